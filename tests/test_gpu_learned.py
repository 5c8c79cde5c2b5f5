"""GPU parity of the learned preconditioner (csrc/dwf_learned.hip) against the oracle
(oracle/oracle_net.cpp in its bf16-quantized mode, which rounds where the device rounds)
and the flexible CG that uses it.  Tolerances: forward/adjoint 1e-3 relative (FP32
versus FP64 accumulation; a one-ulp bf16 flip of one stored activation, where the two
accumulations straddle a rounding boundary, moves a small output by ~1e-4), far below
the 1e-2 that separates bf16 from FP64 arithmetic; loss 1e-4 and gradient 1e-3
relative; solves certified at the FP64 true residual."""

import numpy as np
import pytest

import oracle
from conftest import print_case
from dense_ref import rand_spinor, rand_theta

import three_part_dwf_study_amd as dwf
from three_part_dwf_study_amd import capi

pytestmark = pytest.mark.gpu

M5, MF = 1.0, 0.05


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def weights32(rng, Ls, scale=0.3):
    """Random weights exactly representable in FP32 (the device's master copy)."""
    return (rng.normal(size=oracle.net_num_weights(Ls)) * scale).astype(np.float32)


@pytest.mark.parametrize("Nt,Nx,Ls", [(8, 6, 4), (16, 12, 8)])
def test_net_apply_matches_oracle(Nt, Nx, Ls):
    print_case(test="net_apply", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(Nt + Ls)
    th = rand_theta(rng, Nt, Nx)
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    u = rand_spinor(rng, Ls * Nt * Nx * 2)
    net = dwf.Network(Ls, seed=1)
    net.set_weights(w)
    assert np.array_equal(net.get_weights(), w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 1)  # the fused kernel in the lattice's own gauge
        z = net.apply(ctx, v)
        za = net.apply(ctx, v, adjoint=True)
        zu = net.apply(ctx, u, adjoint=True)
    w64 = w.astype(np.float64)
    zq = oracle.net_apply(th, w64, v, Nt, Nx, Ls, bf16=True)
    zaq = oracle.net_apply(th, w64, v, Nt, Nx, Ls, adj=True, bf16=True)
    z64 = oracle.net_apply(th, w64, v, Nt, Nx, Ls)
    for r in range(2):
        assert rel(z[r], zq[r]) <= 1e-3, rel(z[r], zq[r])
        assert rel(za[r], zaq[r]) <= 1e-3, rel(za[r], zaq[r])
        assert rel(z[r], z64[r]) <= 2e-2
    # the device's adjoint is the adjoint of its forward to bf16 accuracy: <u, M v> = <M^dag u, v>
    lhs, rhs = np.vdot(u, z[0]), np.vdot(zu, v[0])
    assert abs(lhs - rhs) <= 1e-2 * np.linalg.norm(u) * np.linalg.norm(z[0])


@pytest.mark.parametrize("Nt,Nx,Ls", [(8, 6, 4), (16, 12, 8), (12, 20, 3), (16, 32, 5)])
def test_fused_network_matches_layers(Nt, Nx, Ls):
    """The fused three-layer kernel (LDS-resident hidden layers, output tiles of up to
    8 x 16 sites, wrapped fifth dimension) against the layer-by-layer kernels: the same
    bf16 operands, transports and MFMA order, so bit-identical; and against the oracle."""
    print_case(test="fused_network", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(3 * Nt + Nx + Ls)
    th = rand_theta(rng, Nt, Nx)
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 3 * Ls * Nt * Nx * 2).reshape(3, -1)
    net = dwf.Network(Ls, seed=1)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 1)
        zf = net.apply(ctx, v)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 0)
        zl = net.apply(ctx, v)
    assert np.array_equal(zf, zl)
    zq = oracle.net_apply(th, w.astype(np.float64), v, Nt, Nx, Ls, bf16=True)
    for r in range(3):
        assert rel(zf[r], zq[r]) <= 1e-3, rel(zf[r], zq[r])


def block_gauge(th, Nt, Nx, t_start):
    """Temporal gauge starting at row t_start: Om = 1 on that row, Om(t + 1) = Om(t) T_t(t)
    with T_t the +t transporter (antiperiodic sign on the last row), continued around the
    lattice.  Returns Om [Nt, Nx] and the angles theta' (same layout as theta) whose
    transporters are Om(x) T(x) Om(x + mu)^*: 1 along t except into row t_start."""
    sg = np.ones(Nt)
    sg[-1] = -1.0
    Tt = sg[:, None] * np.exp(1j * th[: Nt * Nx].reshape(Nt, Nx))
    Ux = np.exp(1j * th[Nt * Nx:].reshape(Nt, Nx))
    om = np.ones((Nt, Nx), complex)
    o = np.ones(Nx, complex)
    for k in range(Nt):
        t = (t_start + k) % Nt
        om[t] = o
        o = o * Tt[t]
    Tt2 = om * Tt * np.conj(np.roll(om, -1, axis=0))
    Tx2 = om * Ux * np.conj(np.roll(om, -1, axis=1))
    return om, np.concatenate([np.angle(Tt2 / sg[:, None]).ravel(), np.angle(Tx2).ravel()])


@pytest.mark.parametrize("Nt,Nx,Ls", [(16, 32, 5), (24, 22, 4)])
def test_fused_network_block_gauge_matches_oracle(Nt, Nx, Ls):
    """DWF_TUNE_NET_FUSED 2: each block of the fused kernel computes in a temporal gauge
    of its own (no transport along t inside the block).  Pinned per block against the
    bf16-quantized oracle run on the whole lattice in that block's gauge (Om v in,
    Om^* z' out): the block's output region sees the same transporters and inputs.  The
    shapes are large enough that a block's input region (tile + 3-site halo) does not
    wrap around the lattice."""
    print_case(test="fused_network_block_gauge", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(5 * Nt + Nx + Ls)
    th = rand_theta(rng, Nt, Nx)
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    net = dwf.Network(Ls, seed=1)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 2)
        zg = net.apply(ctx, v)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 1)
        zf = net.apply(ctx, v)
    TY = max(d for d in range(1, 9) if Nt % d == 0)
    TX = max(d for d in range(1, 17) if Nx % d == 0)
    assert Nt >= TY + 6 and Nx >= TX + 6
    w64 = w.astype(np.float64)
    zg4 = zg.reshape(2, Ls, Nt, Nx, 2)
    for t0 in range(0, Nt, TY):
        om, th2 = block_gauge(th, Nt, Nx, (t0 - 3) % Nt)
        vg = (om[None, None, :, :, None] * v.reshape(2, Ls, Nt, Nx, 2)).reshape(2, -1)
        zq = oracle.net_apply(th2, w64, vg, Nt, Nx, Ls, bf16=True).reshape(2, Ls, Nt, Nx, 2)
        zq = np.conj(om)[None, None, :, :, None] * zq
        for x0 in range(0, Nx, TX):
            blk = (slice(None), slice(None), slice(t0, t0 + TY), slice(x0, x0 + TX))
            assert rel(zg4[blk].ravel(), zq[blk].ravel()) <= 1e-3, (t0, x0, rel(zg4[blk].ravel(), zq[blk].ravel()))
    # and the gauge-covariant network gives the same field as without the gauge change,
    # up to the bf16 rounding of the transformed values
    z64 = oracle.net_apply(th, w64, v, Nt, Nx, Ls)
    for r in range(2):
        assert rel(zg[r], zf[r]) <= 2e-2 and rel(zg[r], z64[r]) <= 2e-2


@pytest.mark.parametrize("Nt,Nx,Ls", [(8, 6, 4), (12, 20, 3)])
def test_fused_network_block_gauge_small_lattices(Nt, Nx, Ls):
    """Block gauge on lattices smaller than a block's input region (the region is a patch
    of the covering lattice): same field as the exact fused kernel and the FP64 oracle up
    to bf16 rounding."""
    print_case(test="fused_network_block_gauge_small", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(7 * Nt + Nx + Ls)
    th = rand_theta(rng, Nt, Nx)
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    net = dwf.Network(Ls, seed=1)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 2)
        zg = net.apply(ctx, v)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 1)
        zf = net.apply(ctx, v)
    z64 = oracle.net_apply(th, w.astype(np.float64), v, Nt, Nx, Ls)
    for r in range(2):
        assert rel(zg[r], zf[r]) <= 2e-2 and rel(zg[r], z64[r]) <= 2e-2


def test_net_apply_per_configuration():
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="net_apply_batch")
    rng = np.random.default_rng(21)
    ths = np.stack([rand_theta(rng, Nt, Nx) for _ in range(2)])
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 4 * Ls * Nt * Nx * 2).reshape(4, -1)
    net = dwf.Network(Ls)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, 1)
        z = net.apply(ctx, v)
    for r in range(4):
        want = oracle.net_apply(ths[r // 2], w.astype(np.float64), v[r], Nt, Nx, Ls, bf16=True)
        assert rel(z[r], want) <= 1e-3, r


@pytest.mark.parametrize("loss", [0, 1])
def test_loss_gradient_matches_oracle(loss):
    Nt, Nx, Ls = 8, 6, 4
    print_case(test="net_loss_grad", loss=loss)
    rng = np.random.default_rng(40 + loss)
    ths = np.stack([rand_theta(rng, Nt, Nx) for _ in range(2)])
    w = weights32(rng, Ls)
    n = Ls * Nt * Nx * 2
    b = rand_spinor(rng, 6 * n).reshape(6, n)
    x = rand_spinor(rng, 6 * n).reshape(6, n)
    # FP32-representable pairs, as the device stores them
    b = b.view(np.float64).astype(np.float32).astype(np.float64).view(np.complex128)
    x = x.view(np.float64).astype(np.float32).astype(np.float64).view(np.complex128)
    net = dwf.Network(Ls)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths)
        L, g = net.loss_grad(ctx, b, x, pair_cfg=[0, 0, 0, 1, 1, 1], loss=loss)
    Lw, gw = 0.0, np.zeros(w.size)
    for k in range(2):
        lk, gk = oracle.net_loss_grad(ths[k], w.astype(np.float64), b[3 * k:3 * k + 3], x[3 * k:3 * k + 3], Nt, Nx,
                                      Ls, bf16=True, loss=loss, M5=M5, mf=MF)
        Lw += lk / 2
        gw += gk / 2
    assert abs(L - Lw) <= 1e-4 * Lw, (L, Lw)
    assert rel(g, gw) <= 1e-3, rel(g, gw)
    # and the quantized oracle is within bf16 accuracy of the exact gradient
    g64 = sum(oracle.net_loss_grad(ths[k], w.astype(np.float64), b[3 * k:3 * k + 3], x[3 * k:3 * k + 3], Nt, Nx,
                                   Ls, loss=loss, M5=M5, mf=MF)[1] / 2 for k in range(2))
    assert rel(g, g64) <= 3e-2


def boundary_sources(rng, n, Nt, Nx, Ls):
    eta = rng.normal(size=(n, Nt, Nx, 2)) + 1j * rng.normal(size=(n, Nt, Nx, 2))
    src = np.zeros((n, Ls, Nt, Nx, 2), np.complex128)
    src[:, 0, :, :, 1] = eta[..., 1]
    src[:, Ls - 1, :, :, 0] = eta[..., 0]
    return src.reshape(n, -1)


def test_learned_solve_certified_untrained():
    """FCG with the initial network (about a multiple of the identity) converges like CG
    and is certified on the FP64 true residual of D^dag D x = D^dag src."""
    Nt, Nx, Ls = 16, 16, 8
    print_case(test="learned_solve_untrained")
    rng = np.random.default_rng(2)
    th = rand_theta(rng, Nt, Nx)
    src = boundary_sources(rng, 3, Nt, Nx, Ls)
    net = dwf.Network(Ls, seed=3)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        with pytest.raises(dwf.DwfError, match="no preconditioner attached"):
            ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
        ctx.set_preconditioner(net)
        x, st, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
        _, stc, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_CG64)
    for r in range(3):
        assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10
        # no fallback was needed; M ~ I/4 in bf16 keeps FCG(1) within 2x of CG
        assert st[r]["fp64_iterations"] == 0 and st[r]["reliable_updates"] >= 1
        assert st[r]["iterations"] <= 2 * stc[r]["iterations"], (st[r], stc[r])
        Dx = oracle.dwf_apply(0, th, x[r], Nt, Nx, Ls, M5, MF)
        lhs = oracle.dwf_apply(1, th, src[r] - Dx, Nt, Nx, Ls, M5, MF)
        rhs = oracle.dwf_apply(1, th, src[r], Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= 1.01e-10


def test_learned_solve_fletcher_reeves():
    """The Fletcher-Reeves direction update (no q = A p pass) converges and is certified
    with the untrained network (a multiple of the identity, symmetric), with CG's
    iteration count within 2x."""
    Nt, Nx, Ls = 16, 16, 8
    print_case(test="learned_solve_fr", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(13)
    th = rand_theta(rng, Nt, Nx)
    src = boundary_sources(rng, 3, Nt, Nx, Ls)
    net = dwf.Network(Ls, seed=3)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.set_preconditioner(net)
        ctx.tune(dwf.DWF_TUNE_FCG_BETA, 1)
        x, st, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
        _, stc, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_CG64)
        with pytest.raises(dwf.DwfError):
            ctx.tune(dwf.DWF_TUNE_FCG_BETA, 2)
    for r in range(3):
        assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10 and st[r]["fp64_iterations"] == 0
        assert st[r]["iterations"] <= 2 * stc[r]["iterations"]
        Dx = oracle.dwf_apply(0, th, x[r], Nt, Nx, Ls, M5, MF)
        lhs = oracle.dwf_apply(1, th, src[r] - Dx, Nt, Nx, Ls, M5, MF)
        rhs = oracle.dwf_apply(1, th, src[r], Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= 1.01e-10


def test_learned_solve_fused_and_layers():
    """FCG with the fused network (dot products in its epilogue) and with the layer
    kernels (separate dot-product pass): both certified, iteration counts within 2
    (the two sum the dot products in different orders)."""
    Nt, Nx, Ls = 16, 16, 8
    print_case(test="learned_solve_fused", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(12)
    th = rand_theta(rng, Nt, Nx)
    src = boundary_sources(rng, 3, Nt, Nx, Ls)
    net = dwf.Network(Ls, seed=3)
    res = []
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.set_preconditioner(net)
        for fused in (1, 0, 2):
            ctx.tune(dwf.DWF_TUNE_NET_FUSED, fused)
            res.append(ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED))
    for (x, st, _) in res:
        for r in range(3):
            assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10 and st[r]["fp64_iterations"] == 0
    for r in range(3):
        assert abs(res[0][1][r]["iterations"] - res[1][1][r]["iterations"]) <= 2
        assert rel(res[0][0][r], res[1][0][r]) <= 1e-8
        # the block-gauge network is the same preconditioner up to bf16 rounding
        assert abs(res[2][1][r]["iterations"] - res[0][1][r]["iterations"]) <= 0.1 * res[0][1][r]["iterations"] + 2
        assert rel(res[2][0][r], res[0][0][r]) <= 1e-8


def volume_sources(rng, n, Nt, Nx, Ls):
    shape = (n, Ls * Nt * Nx * 2)
    return rng.normal(size=shape) + 1j * rng.normal(size=shape)


def test_learned_solve_deferred_x_update_staggered():
    """With the LDS-staged stencils the learned solver runs its fused iteration: the D
    stencil forms p = z + beta p_old and applies the previous iteration's x += alpha p_old
    (one iteration late), and the end of a cycle applies each right-hand side's last
    one.  Right-hand sides of different kinds freeze at different iterations of a cycle,
    so some wait with an update pending while others run on.  The solutions must match
    the layer path (separate kernels, x updated at once) to rounding -- and differ from
    it in the last bits, which shows that the two runs took different paths -- and every
    solve must be certified."""
    Nt, Nx, Ls = 16, 24, 8
    print_case(test="learned_deferred_x", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(31)
    th = rand_theta(rng, Nt, Nx)
    src = np.concatenate([boundary_sources(rng, 2, Nt, Nx, Ls), volume_sources(rng, 2, Nt, Nx, Ls)])
    net = dwf.Network(Ls, seed=5)
    res = []
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.set_preconditioner(net)
        ctx.tune(dwf.DWF_TUNE_LDS, 3)  # 64-thread blocks, tiles of 8 x 8 sites
        ctx.tune(dwf.DWF_TUNE_TILE_X, 8)
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA, 1e-3)
        for fr in (1, 0):
            ctx.tune(dwf.DWF_TUNE_FCG_BETA, fr)
            for fused in (1, 0):
                ctx.tune(dwf.DWF_TUNE_NET_FUSED, fused)
                res.append(ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED))
    its = [st["iterations"] for st in res[0][1]]
    assert max(its) - min(its) >= 3, its  # the batch did stagger
    for (x, st, _) in res:
        for r in range(len(src)):
            assert st[r]["converged"] and st[r]["true_relres"] <= 1e-10 and st[r]["fp64_iterations"] == 0
    for f, l in ((0, 1), (2, 3)):
        assert not np.array_equal(res[f][0], res[l][0])
        for r in range(len(src)):
            assert abs(res[f][1][r]["iterations"] - res[l][1][r]["iterations"]) <= 2
            assert rel(res[f][0][r], res[l][0][r]) <= 1e-8


def test_training_cuts_iterations():
    """What the trained network gains: train on (source, CG solution) pairs from 4
    configurations (full-volume Gaussian sources), then solve boundary sources on 2
    held-out configurations.  FCG with the network must need at most 60% of plain FP64
    CG's iterations (measured: about 40%), be certified at 1e-10, and need no fallback."""
    Nt, Nx, Ls, ntrain, per = 16, 16, 8, 4, 16
    print_case(test="training_gain", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(7)
    ths = np.stack([rand_theta(rng, Nt, Nx) for _ in range(ntrain + 2)])
    net = dwf.Network(Ls, seed=11)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths[:ntrain])
        src = volume_sources(rng, ntrain * per, Nt, Nx, Ls)
        b = ctx.apply(dwf.DWF_OP_DDAG, src)
        x, st, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_CG64)
        hist = net.train(ctx, b, x, pair_cfg=np.repeat(np.arange(ntrain), per), steps=300, batch=8, loss=1,
                         lr=3e-3, seed=5)
        assert hist[-20:].mean() < 0.05 * hist[:5].mean(), hist[[0, 50, 100, 200, 299]]
        ctx.set_preconditioner(net)
        for k in range(ntrain, ntrain + 2):
            ctx.set_gauge(ths[k])
            s2 = boundary_sources(rng, 4, Nt, Nx, Ls)
            _, sl, _ = ctx.solve_dirac(s2, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_LEARNED)
            _, sc, _ = ctx.solve_dirac(s2, rtol=1e-10, maxiter=20000, method=dwf.DWF_SOLVE_CG64)
            il = np.array([s["iterations"] for s in sl])
            ic = np.array([s["iterations"] for s in sc])
            print(f"held-out configuration {k}: FCG {il.tolist()} vs CG {ic.tolist()}", flush=True)
            assert all(s["converged"] and s["true_relres"] <= 1e-10 for s in sl)
            assert all(s["fp64_iterations"] == 0 for s in sl)  # no fallback to plain CG
            assert il.mean() <= 0.6 * ic.mean()


def test_learned_dddag_solve_force_and_hmc():
    """DWF_SOLVE_LEARNED on D D^dag (as the rotated D^dag D solve): certified against the
    oracle's D D^dag, with and without an initial guess; the pseudofermion action and
    force agree with FP64 CG's, and a short HMC trajectory with learned force solves has
    FP64 CG's Delta H."""
    Nt, Nx, Ls = 16, 12, 8
    print_case(test="learned_dddag", Nt=Nt, Nx=Nx, Ls=Ls)
    rng = np.random.default_rng(31)
    th = rand_theta(rng, Nt, Nx)
    b = rand_spinor(rng, Ls * Nt * Nx * 2)
    net = dwf.Network(Ls, seed=3)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.set_preconditioner(net)
        x, st, rc = ctx.solve(dwf.DWF_OP_DDDAG, b, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
        r = b - oracle.dwf_apply(dwf.DWF_OP_DDDAG, th, x, Nt, Nx, Ls, M5, MF)
        assert st[0]["converged"] and np.linalg.norm(r) / np.linalg.norm(b) <= 1.01e-10
        x0 = x + 1e-6 * rand_spinor(rng, x.size)
        x2, st2, _ = ctx.solve(dwf.DWF_OP_DDDAG, b, x0=x0, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
        r2 = b - oracle.dwf_apply(dwf.DWF_OP_DDDAG, th, x2, Nt, Nx, Ls, M5, MF)
        assert st2[0]["converged"] and np.linalg.norm(r2) / np.linalg.norm(b) <= 1.01e-10
        assert st2[0]["iterations"] < st[0]["iterations"]
        S, F, *_ = ctx.pf_action_force(b, 1e-12, 5000)
        Sl, Fl, *_ = ctx.pf_action_force(b, 1e-12, 5000, method=dwf.DWF_SOLVE_LEARNED)
        assert abs(Sl - S) <= 1e-10 * abs(S) and rel(Fl, F) <= 1e-9
    hist = []
    for method in (dwf.DWF_SOLVE_CG64, dwf.DWF_SOLVE_LEARNED):
        r = capi.HostRng(77)
        with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
            ctx.set_gauge(np.ones(2 * Nt * Nx))
            ctx.set_preconditioner(net)
            dh = []
            for _ in range(2):
                chi, mom = r.dwf_heatbath(Ls * Nt * Nx * 2, 2 * Nt * Nx)
                res = ctx.hmc_trajectory(chi, mom, r.uniform(), 3.0, 0.05, 6, 1e-11, 10000, method=method)
                dh.append(res["delta_h"])
            hist.append(np.array(dh))
    assert np.max(np.abs(hist[0] - hist[1])) <= 1e-8


def test_learned_solve_full_size():
    """cfg4 at its real size: 96x96, Ls=16.  The network is trained here on the device
    (8 configurations, seeds 1234 + 104729 i for i = 1..8, 8 full-volume Gaussian sources
    each, loss ||D^dag D M b - b||^2 / ||b||^2) and solves 4 boundary sources on the
    held-out configuration (seed 1234) with the bench's settings (LDS stencil, fused
    network, Fletcher-Reeves beta, reliable updates after a 1e-4 drop).  Every solve must
    be certified at 1e-10 against the oracle's operator, need no FP64 fallback, and take
    at most half of plain FP64 CG's iterations."""
    Nt = Nx = 96
    Ls, ntrain, per, rtol = 16, 8, 8, 1e-10
    print_case(test="learned_full_size", Nt=Nt, Nx=Nx, Ls=Ls)
    ths = np.stack([np.random.default_rng(1234 + 104729 * i).normal(0.0, 0.5, size=2 * Nt * Nx)
                    for i in range(ntrain + 1)])
    rng = np.random.default_rng(918273)
    net = dwf.Network(Ls, seed=11)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge_batch(ths[1:])
        src = volume_sources(rng, ntrain * per, Nt, Nx, Ls)
        b = ctx.apply(dwf.DWF_OP_DDAG, src)  # right-hand side k uses configuration k // per
        del src
        hist = net.train(ctx, b, None, pair_cfg=np.repeat(np.arange(ntrain), per), steps=600, batch=8, loss=1,
                         lr=3e-3, seed=5)
        del b
        assert hist[-20:].mean() < 0.1 * hist[:5].mean(), hist[[0, 100, 300, 599]]
        ctx.set_gauge(ths[0])
        ctx.set_preconditioner(net)
        ctx.tune(dwf.DWF_TUNE_LDS, 2)  # 512-thread tiles of 16 x 32 sites
        ctx.tune(dwf.DWF_TUNE_TILE_X, 16)
        ctx.tune(dwf.DWF_TUNE_FCG_BETA, 1)
        ctx.tune(dwf.DWF_TUNE_RELIABLE_DELTA, 1e-4)
        s2 = boundary_sources(rng, 4, Nt, Nx, Ls)
        xl, sl, _ = ctx.solve_dirac(s2, rtol=rtol, maxiter=20000, method=dwf.DWF_SOLVE_LEARNED)
        _, sc, _ = ctx.solve_dirac(s2, rtol=rtol, maxiter=20000, method=dwf.DWF_SOLVE_CG64)
    il = np.array([s["iterations"] for s in sl])
    ic = np.array([s["iterations"] for s in sc])
    print(f"96x96x16 held-out configuration: learned {il.tolist()} vs FP64 CG {ic.tolist()}", flush=True)
    assert all(s["fp64_iterations"] == 0 for s in sl)  # no fallback to plain CG
    assert np.all(il <= 0.5 * ic)
    th = ths[0]
    for r in range(len(s2)):
        assert sl[r]["converged"]
        Dx = oracle.dwf_apply(0, th, xl[r], Nt, Nx, Ls, M5, MF)
        lhs = oracle.dwf_apply(1, th, s2[r] - Dx, Nt, Nx, Ls, M5, MF)
        rhs = oracle.dwf_apply(1, th, s2[r], Nt, Nx, Ls, M5, MF)
        assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= rtol * 1.01, r


@pytest.mark.parametrize("variant", [1, 2])
def test_fused_network_split_fifth_dimension(variant):
    """Splitting each output tile's walk over the fifth dimension into parts of one block
    each (DWF_TUNE_NET_SPLIT) recomputes the parts' edge slices from the same data and
    weights, so every split gives the unsplit kernel's output bit for bit, in both kernel
    variants, down to parts of one slice; an Ls that a split does not divide is refused."""
    Nt, Nx, Ls = 16, 12, 8
    print_case(test="fused_network_split", variant=variant)
    rng = np.random.default_rng(44 + variant)
    th = rand_theta(rng, Nt, Nx)
    w = weights32(rng, Ls)
    v = rand_spinor(rng, 2 * Ls * Nt * Nx * 2).reshape(2, -1)
    net = dwf.Network(Ls, seed=1)
    net.set_weights(w)
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.tune(dwf.DWF_TUNE_NET_FUSED, variant)
        ctx.tune(dwf.DWF_TUNE_NET_SPLIT, 1)
        z1 = net.apply(ctx, v)
        for n in (2, 4, 8, 0):
            ctx.tune(dwf.DWF_TUNE_NET_SPLIT, n)
            assert np.array_equal(net.apply(ctx, v), z1), n
        for bad in (3, 16, -1):
            with pytest.raises(dwf.DwfError):
                ctx.tune(dwf.DWF_TUNE_NET_SPLIT, bad)
    zq = oracle.net_apply(th, w.astype(np.float64), v, Nt, Nx, Ls, bf16=True)
    if variant == 1:
        for r in range(2):
            assert rel(z1[r], zq[r]) <= 1e-3, rel(z1[r], zq[r])


def test_learned_solve_split_network():
    """One right-hand side, the regime the split is for: the learned solve with the
    network's walk split in 4 parts (8 tiles x 4 = 32 blocks, the default there) and
    unsplit both certify at the FP64 true residual with the same iteration count up to the
    last bits of the partial sums' order."""
    Nt, Nx, Ls = 32, 32, 16
    print_case(test="learned_solve_split")
    rng = np.random.default_rng(5)
    th = rand_theta(rng, Nt, Nx)
    src = boundary_sources(rng, 1, Nt, Nx, Ls)
    net = dwf.Network(Ls, seed=3)
    its = {}
    with dwf.DwfContext.dwf(Nt, Nx, Ls, M5, MF) as ctx:
        ctx.set_gauge(th)
        ctx.set_preconditioner(net)
        for n in (1, 4, 0):
            ctx.tune(dwf.DWF_TUNE_NET_SPLIT, n)
            x, st, _ = ctx.solve_dirac(src, rtol=1e-10, maxiter=5000, method=dwf.DWF_SOLVE_LEARNED)
            assert st[0]["converged"] and st[0]["true_relres"] <= 1e-10 and st[0]["fp64_iterations"] == 0
            its[n] = st[0]["iterations"]
            Dx = oracle.dwf_apply(0, th, x[0], Nt, Nx, Ls, M5, MF)
            lhs = oracle.dwf_apply(1, th, src[0] - Dx, Nt, Nx, Ls, M5, MF)
            rhs = oracle.dwf_apply(1, th, src[0], Nt, Nx, Ls, M5, MF)
            assert np.linalg.norm(lhs) / np.linalg.norm(rhs) <= 1.01e-10
    assert its[4] == its[0] and abs(its[4] - its[1]) <= 3, its
