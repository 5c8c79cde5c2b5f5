"""Static check of the gfx950 code of the LDS-tiled kernels (CPU only; needs hipcc).

A block barrier at the top of a loop must be preceded by an LDS wait on every path into
it: a wave whose LDS writes of the previous step are still in flight would otherwise
let other waves pass the barrier and read stale values.  The compiler has dropped the
wait of __syncthreads() on the back-edge path (a CG solve then lost one right-hand side
to a stale halo about once in twenty solves), so the kernels issue it explicitly; this
test compiles the sources to assembly and checks every loop that writes LDS and meets
a barrier before any `s_waitcnt lgkmcnt(0)` on entry."""

import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

CSRC = os.path.join(os.path.dirname(__file__), "..", "three_part_dwf_study_amd", "csrc")
SOURCES = ("dwf_device.hip", "dwf_eo.hip", "dwf_learned.hip")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


def assemble(src, out_dir):
    out = os.path.join(out_dir, src + ".s")
    subprocess.run([HIPCC, "-std=c++17", "-O3", "--offload-arch=gfx950", "-fPIC", "-Wno-unused-function",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--offload-device-only",
                    "-S", "-o", out, os.path.join(CSRC, src)], check=True, capture_output=True)
    return out


def unguarded_loop_barriers(asm):
    """(kernel, loop header) pairs whose loop writes LDS and whose header reaches an
    s_barrier before an lgkmcnt(0) wait."""
    found = []
    for m in re.finditer(r"^(_Z\S+):", asm, re.M):
        end = asm.find(".Lfunc_end", m.end())
        body = [l.strip() for l in asm[m.end():end].splitlines()]
        for h, line in enumerate(body):
            hm = re.match(r"^\.(LBB\d+_\d+):.*Loop Header", line)
            if not hm:
                continue
            lab = hm.group(1)
            inloop, writes = False, False
            for l2 in body:
                bm = re.match(r"^(\.LBB\d+_\d+|; %bb\.\d+):(.*)$", l2)
                if bm:
                    inloop = bm.group(1) == "." + lab or f"Header={lab[1:]}" in bm.group(2)
                    continue
                if inloop and l2.startswith("ds_write"):
                    writes = True
            if not writes:
                continue
            for l2 in body[h + 1:]:
                if re.match(r"s_waitcnt .*lgkmcnt\(0\)", l2):
                    break
                if l2.startswith("s_barrier"):
                    found.append((m.group(1)[:80], lab))
                    break
    return found


def test_unguarded_barrier_detector():
    """The checker itself: a loop that writes LDS and meets a barrier at its header
    without a wait is flagged; with the wait it is not."""
    bad = ("_Zk:\n.LBB0_1:  ; =>This Inner Loop Header: Depth=1\n\ts_barrier\n\tds_write_b128 v1, v[2:5]\n"
           "\ts_cbranch_scc1 .LBB0_1\n.Lfunc_end0:\n")
    good = bad.replace("\ts_barrier", "\ts_waitcnt lgkmcnt(0)\n\ts_barrier")
    assert unguarded_loop_barriers(bad) == [("_Zk", "LBB0_1")]
    assert unguarded_loop_barriers(good) == []


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not available")
def test_loop_barriers_wait_for_lds(tmp_path):
    with ThreadPoolExecutor(len(SOURCES)) as ex:
        outs = list(ex.map(lambda s: assemble(s, str(tmp_path)), SOURCES))
    found = []
    for out in outs:
        found += unguarded_loop_barriers(open(out).read())
    assert not found, found
