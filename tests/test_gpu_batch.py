"""Batched launches over all blocks.

lusgs_kernel 2: the tile sweeps of all blocks run in one launch per phase,
driven by a table of blocks (tiles by workgroup index when all fit on the
device at once, else by a ticket counter).  batch_blocks 1: every other
stage of the step (primitive variables, inviscid and viscous residuals,
time step, LU-SGS preparation, residual norms) is one launch over all
blocks too.  The arithmetic per node is that of the per-block launches,
which the parity tests hold to the oracle, so the batched state must equal
it bit for bit, for every tile shape the autotuner may pick and with the
step captured in a HIP graph."""
import os
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from tests.test_gpu_parity import two_block_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def build(kind):
    if kind == "vortex_12_uneven":
        return cases.vortex_box(40, blocks=((13, 27), (17,), (9,)))
    if kind == "tgv_8_uneven":
        return cases.tgv_box(33, blocks=((10,), (20,), (16,)))
    if kind == "ns_two_block":
        return two_block_case(ns=True, seed=21)
    if kind == "sa_two_block_local_dt":
        return two_block_case(ns=True, sa=True, seed=22, itime=0)
    if kind == "euler_bdf2_two_block":
        return two_block_case(seed=23, tadv=3)
    if kind == "vortex_12_rk3":
        case = cases.vortex_box(40, blocks=((13, 27), (17,), (9,)))
        case.cfg.iflag_timeadvance = 4
        return case
    raise ValueError(kind)


def run(case, kernel, nsteps, graph=False, shape=None, tickets=False, batch=0):
    case.cfg.lusgs_kernel = kernel
    case.cfg.use_graph = 1 if graph else 0
    case.cfg.batch_blocks = batch
    old = os.environ.get("ACFD_TILE_SHAPE")
    if shape is not None:
        os.environ["ACFD_TILE_SHAPE"] = str(shape)
    if tickets:
        os.environ["ACFD_LU_TICKETS"] = "1"
    try:
        sol = Solver(case.cfg, case.dims, case.subfaces)
        try:
            case.upload(sol)
            sol.run_stage("physical_bc")
            sol.run_stage("getpri")
            sol.prepare()
            sol.step(nsteps)
            q = [sol.download(m + 1, "q") for m in range(len(case.dims))]
            nut = [sol.download(m + 1, "nut") for m in range(len(case.dims))] \
                if case.cfg.iflag_turbulence == 2 else []
            return q, nut, np.concatenate([sol.residual(), sol.stats()])
        finally:
            sol.close()
    finally:
        os.environ.pop("ACFD_LU_TICKETS", None)
        if shape is not None:
            if old is None:
                del os.environ["ACFD_TILE_SHAPE"]
            else:
                os.environ["ACFD_TILE_SHAPE"] = old


def same(a, b, what):
    qa, na, ra = a
    qb, nb, rb = b
    for m, (x, y) in enumerate(zip(qa, qb)):
        assert np.array_equal(x, y), (what, "q", m, np.abs(x - y).max())
    for m, (x, y) in enumerate(zip(na, nb)):
        assert np.array_equal(x, y), (what, "nut", m, np.abs(x - y).max())
    assert np.array_equal(ra, rb), (what, ra, rb)


@pytest.mark.parametrize("kind", ["vortex_12_uneven", "tgv_8_uneven", "ns_two_block", "sa_two_block_local_dt"])
def test_batched_equals_per_block(kind):
    say("test_batched_equals_per_block", kind)
    nsteps = 6
    ref = run(build(kind), 1, nsteps)
    for shape in (None, 1, 3):
        got = run(build(kind), 2, nsteps, shape=shape)
        same(ref, got, "%s shape %s" % (kind, shape))
    got = run(build(kind), 2, nsteps, graph=True)
    same(ref, got, "%s graph" % kind)
    # tiles handed out by ticket (the form used when the launch does not fit
    # on the device at once), eagerly and replayed from a graph
    for graph in (False, True):
        got = run(build(kind), 2, nsteps, graph=graph, tickets=True)
        same(ref, got, "%s tickets graph=%s" % (kind, graph))
    # every stage batched, with either sweep
    for kernel in (1, 2):
        for graph in (False, True):
            got = run(build(kind), kernel, nsteps, graph=graph, batch=1)
            same(ref, got, "%s batch_blocks lusgs %d graph=%s" % (kind, kernel, graph))
    say("  %d blocks, %d steps identical; residual %s" % (len(ref[0]), nsteps, np.array2string(ref[2][:2], precision=4)))


@pytest.mark.parametrize("kind", ["euler_bdf2_two_block", "vortex_12_rk3"])
def test_batched_stages_other_advances(kind):
    """BDF2 and RK3 with every stage batched against block-by-block."""
    say("test_batched_stages_other_advances", kind)
    nsteps = 4
    ref = run(build(kind), 1, nsteps)
    for graph in (False, True):
        got = run(build(kind), 1, nsteps, graph=graph, batch=1)
        same(ref, got, "%s batch_blocks graph=%s" % (kind, graph))
    say("  %d blocks, %d steps identical; residual %s" % (len(ref[0]), nsteps, np.array2string(ref[2][:2], precision=4)))
