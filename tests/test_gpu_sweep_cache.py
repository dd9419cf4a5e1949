"""The tile sweeps' gather leaves the grid operands (dxidx, inv_j) in the
sweep layout between updates of the same block and tiling and moves only
the other fields.  New metrics written between steps must be gathered
again, also by a captured step: a context that ran on one grid, then had
its metrics and state replaced, matches a fresh context on the new grid bit
for bit."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_case(kind, warp):
    case = cases.tgv_box(33) if kind == "one_block" else cases.tgv_box(33, blocks=((16,), (), (20,)))
    if warp:
        # scaled metrics: another valid grid as far as the update goes
        rng = np.random.default_rng(5)
        for fl in case.fields:
            s = 1.0 + 0.02 * rng.random(fl["inv_j"].shape)
            fl["dxidx"] = fl["dxidx"] * s[..., None]
            fl["inv_j"] = fl["inv_j"] * s
    return case


def reset(sol, case):
    case.upload(sol)
    sol.run_stage("physical_bc")
    sol.run_stage("getpri")
    sol.prepare()


@pytest.mark.parametrize("kind,lusgs,graph", [("one_block", 1, False), ("one_block", 1, True),
                                             ("blocks", 2, False), ("blocks", 2, True)])
def test_new_metrics_are_gathered(kind, lusgs, graph):
    say("test_new_metrics_are_gathered", kind, "lusgs", lusgs, "graph", graph)
    old, new = make_case(kind, False), make_case(kind, True)
    for c in (old, new):
        c.cfg.lusgs_kernel = lusgs
        c.cfg.use_graph = 1 if graph else 0
    a = Solver(old.cfg, old.dims, old.subfaces)
    b = Solver(new.cfg, new.dims, new.subfaces)
    try:
        reset(a, old)
        a.step(3)  # the layout now holds the old grid (graph: captured, replayed)
        reset(a, new)  # new metrics and state
        reset(b, new)
        for n in (1, 3):
            a.step(n)
            b.step(n)
            for m in range(len(new.dims)):
                qa, qb = a.download(m + 1, "q"), b.download(m + 1, "q")
                assert np.array_equal(qa, qb), (kind, m, np.abs(qa - qb).max())
        say("  identical after 4 steps on the new grid")
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("kind,lusgs,graph", [("one_block", 1, False), ("one_block", 1, True),
                                             ("blocks", 2, True)])
def test_state_written_between_steps(kind, lusgs, graph):
    """A step starts from the previous step's primitive variables unless q
    was written since: a state uploaded between steps (with no stage run
    after it) is converted again, also under a captured step."""
    say("test_state_written_between_steps", kind, "lusgs", lusgs, "graph", graph)
    case = make_case(kind, False)
    case.cfg.lusgs_kernel = lusgs
    case.cfg.use_graph = 1 if graph else 0
    fresh = make_case(kind, False)
    fresh.cfg.lusgs_kernel = lusgs
    a = Solver(case.cfg, case.dims, case.subfaces)
    b = Solver(fresh.cfg, fresh.dims, fresh.subfaces)
    try:
        reset(a, case)
        a.step(3)
        # b: the state a holds now, perturbed, from scratch (pri_v from getpri)
        rng = np.random.default_rng(9)
        qs = []
        for m in range(len(case.dims)):
            q = a.download(m + 1, "q")
            q[..., 0] *= 1.0 + 1e-3 * rng.random(q.shape[:-1])
            qs.append(q)
        for m, q in enumerate(qs):
            a.upload(m + 1, "q", q)
        fresh.upload(b)
        for m, q in enumerate(qs):
            b.upload(m + 1, "q", q)
        b.run_stage("getpri")
        for n in (1, 2):
            a.step(n)
            b.step(n)
            for m in range(len(case.dims)):
                qa, qb = a.download(m + 1, "q"), b.download(m + 1, "q")
                assert np.array_equal(qa, qb), (kind, m, np.abs(qa - qb).max())
        say("  identical after 3 steps from the uploaded state")
    finally:
        a.close()
        b.close()
