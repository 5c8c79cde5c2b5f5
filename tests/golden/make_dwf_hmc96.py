"""Writes tests/golden/dwf_hmc_96x96x16.npz: two Nf=2 domain-wall HMC trajectories of the
oracle (the reference's FP64 algorithm, domain_wall/generate_configs.cpp:48-207) at
96x96, Ls=16, M5=1, mf=0.05, beta=4, epsilon=0.05, 20 leapfrog steps, from the
reference's cold start theta = 1 (generate_configs.cpp:27-28), host random stream
mt19937_64(1234) in the reference's draw order, force solves to 1e-12 with chronological
guesses.  Stores Delta H, acceptance, acceptance probability, plaquette per trajectory and
the final angles.  Runs on the CPU in this container (about a quarter of an hour):

    python tests/golden/make_dwf_hmc96.py
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402

NT = NX = 96
LS, M5, MF, BETA, EPS, STEPS, NTRAJ, SEED, RTOL = 16, 1.0, 0.05, 4.0, 0.05, 20, 2, 1234, 1e-12


def main():
    t0 = time.time()
    o = oracle.dwf_hmc(np.ones(2 * NT * NX), NT, NX, LS, M5, MF, BETA, EPS, STEPS, NTRAJ, SEED, RTOL, 20000)
    out = os.path.join(HERE, "dwf_hmc_96x96x16.npz")
    np.savez(out, delta_h=o["delta_h"], accepted=o["accepted"], plaquette=o["plaquette"],
             acc_prob=o["acc_prob"], theta=o["theta"],
             params=np.array([NT, NX, LS, M5, MF, BETA, EPS, STEPS, NTRAJ, SEED, RTOL]))
    print(f"wrote {out} in {time.time() - t0:.0f} s: dH {o['delta_h']}, accepted {o['accepted']}, "
          f"plaquette {o['plaquette']}, cg iterations {o.get('cg_iterations')}", flush=True)


if __name__ == "__main__":
    main()
