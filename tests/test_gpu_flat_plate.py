"""The batched multi-block turbulent flat plate (12 seeded uneven blocks,
tanh wall clustering, SA, wall / inflow / outflow / symmetry / farfield
faces, local time step, first-order implicit) at a reduced size: the GPU
state against the oracle after 10 steps, the batched launches (one launch
per stage and sweep over all blocks, also replayed from a HIP graph)
against the block-by-block launches bit for bit, and the FP32-flux variant's
error against FP64."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from oracle import oracle as O
from tests.helpers import ora_config, ora_states, rel_l2

pytestmark = pytest.mark.gpu

TOLN = 1e-10
SMALL = dict(nblk=(4, 3, 1), lo=(6, 6, 5), hi=(12, 10, 8))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_run(case, nsteps, lusgs=1, batch=0, graph=0, precision=0):
    case.cfg.lusgs_kernel = lusgs
    case.cfg.batch_blocks = batch
    case.cfg.use_graph = graph
    case.cfg.precision = precision
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        case.upload(sol)
        sol.run_stage("physical_bc")
        sol.run_stage("getpri")
        sol.prepare()
        sol.step(nsteps)
        nb = len(case.dims)
        return ([sol.download(m + 1, "q") for m in range(nb)], [sol.download(m + 1, "nut") for m in range(nb)],
                sol.residual())
    finally:
        sol.close()


def test_flat_plate_matches_oracle():
    nsteps = 10
    case = cases.flat_plate(**SMALL)
    say("test_flat_plate_matches_oracle", case.name, "steps", nsteps)
    cfg = ora_config(case.cfg, case.totalnodes)
    states = ora_states(case)
    O.physical_bc(states, cfg)
    for st in states:
        st["pri_v"] = O.getpri(st.ext, st["q"], cfg.gamma, cfg.cv)[0]
    O.prepare(states, cfg)
    for _ in range(nsteps):
        O.step(states, cfg)
    res_o = O.residual(states, case.totalnodes)
    q, nut, res = gpu_run(cases.flat_plate(**SMALL), nsteps)
    worst = 0.0
    for m, st in enumerate(states):
        e = rel_l2(q[m], st["q"])
        en = rel_l2(nut[m], st["nut"])
        worst = max(worst, float(e.max()), float(en))
        assert np.all(e <= TOLN), ("q", m, e)
        assert en <= TOLN, ("nut", m, en)
    say("  worst rel L2 (q, nut) over 12 blocks: %.3e" % worst)
    assert np.allclose(res[:6], res_o[:6], rtol=1e-9, atol=1e-300), (res, res_o)


def test_flat_plate_batched_identical():
    nsteps = 6
    ref = gpu_run(cases.flat_plate(**SMALL), nsteps)
    for graph in (0, 1):
        got = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=2, batch=1, graph=graph)
        for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
            assert np.array_equal(a, b), np.abs(a - b).max()
        assert np.array_equal(ref[2], got[2])
    say("  batched (graph and eager) identical to block-by-block over %d steps" % nsteps)


def test_flat_plate_sa_concurrent_identical():
    """The SA update on its own stream beside the flow update (the default
    with batched sweeps), eager and replayed from a graph, against the two
    updates one after the other."""
    nsteps = 6
    import os
    old = os.environ.get("ACFD_SA_CONCURRENT")
    try:
        os.environ["ACFD_SA_CONCURRENT"] = "0"
        ref = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=2, batch=1)
        os.environ["ACFD_SA_CONCURRENT"] = "1"
        for graph in (0, 1):
            got = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=2, batch=1, graph=graph)
            for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
                assert np.array_equal(a, b), (graph, np.abs(a - b).max())
            assert np.array_equal(ref[2], got[2])
    finally:
        if old is None:
            os.environ.pop("ACFD_SA_CONCURRENT", None)
        else:
            os.environ["ACFD_SA_CONCURRENT"] = old
    say("  concurrent SA update (eager and graph) identical over %d steps" % nsteps)


def test_flat_plate_sweep_variants_identical():
    """Hyperplane launches (lusgs_kernel 0) and tile sweeps, per block (1)
    and batched (2), for both the flow and the SA updates: same bits."""
    nsteps = 6
    ref = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=0)
    for lusgs, batch in ((1, 0), (2, 0), (2, 1)):
        got = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=lusgs, batch=batch)
        for a, b in zip(ref[0] + ref[1], got[0] + got[1]):
            assert np.array_equal(a, b), (lusgs, batch, np.abs(a - b).max())
    say("  hyperplane, tile and batched tile sweeps identical over %d steps" % nsteps)


def test_flat_plate_fp32_flux_error():
    """FP32 reconstruction / splitting / viscous flux with the FP64 update on
    the stretched wall grid: density within 1e-4 relative L2 of FP64."""
    nsteps = 20
    q64 = gpu_run(cases.flat_plate(**SMALL), nsteps)[0]
    q32 = gpu_run(cases.flat_plate(**SMALL), nsteps, lusgs=2, batch=1, precision=1)[0]
    num = sum(float(np.sum((a[..., 0] - b[..., 0]) ** 2)) for a, b in zip(q32, q64))
    den = sum(float(np.sum(b[..., 0] ** 2)) for b in q64)
    l2 = (num / den) ** 0.5
    say("  FP32-flux density rel L2 after %d steps: %.3e" % (nsteps, l2))
    assert 0.0 < l2 <= 1e-4
