"""Graph mode of acfd_step: after one eager step, a step is captured into a
HIP graph and replayed.  The replays run the same kernels in the same order
on the same buffers, so the state must match eager stepping bit for bit.
Replays are separated by eager work (downloads, residual reads, a stage
run through acfd_run_stage) and every chunk is compared."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from tests.test_gpu_parity import two_block_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def start(case, graph, lusgs=1):
    case.cfg.use_graph = 1 if graph else 0
    case.cfg.lusgs_kernel = lusgs
    sol = Solver(case.cfg, case.dims, case.subfaces)
    case.upload(sol)
    sol.run_stage("physical_bc")
    sol.run_stage("getpri")
    sol.prepare()
    return sol


def build(kind):
    if kind == "ns_lusgs_tiles":
        return two_block_case(ns=True, seed=11)
    if kind == "sa_lusgs":
        return two_block_case(ns=True, sa=True, seed=12, itime=0)
    if kind == "euler_bdf2":
        return two_block_case(seed=13, tadv=3)
    if kind == "tgv_blocks":
        return cases.tgv_box(33, blocks=((16,), (), (20,)))
    if kind == "shock_vortex_rk3":
        return cases.shock_vortex(64, 32, 4)
    raise ValueError(kind)


@pytest.mark.parametrize("kind", ["ns_lusgs_tiles", "sa_lusgs", "euler_bdf2", "tgv_blocks",
                                  "shock_vortex_rk3"])
def test_graph_replay_matches_eager(kind):
    say("test_graph_replay_matches_eager", kind)
    chunks = (1, 3, 4, 2)
    a = start(build(kind), graph=False)
    b = start(build(kind), graph=True)
    nblk = len(a._ranges)
    try:
        for c, n in enumerate(chunks):
            a.step(n)
            b.step(n)
            # eager work between replays
            ra, rb = a.residual(), b.residual()
            for m in range(nblk):
                qa, qb = a.download(m + 1, "q"), b.download(m + 1, "q")
                assert np.array_equal(qa, qb), (kind, c, m, np.abs(qa - qb).max())
            assert np.array_equal(ra, rb), (kind, c, ra, rb)
            if c == 1:
                a.run_stage("getpri")
                b.run_stage("getpri")
        say("  %d steps identical, residual %s" % (sum(chunks), np.array2string(rb[:2], precision=4)))
    finally:
        a.close()
        b.close()
