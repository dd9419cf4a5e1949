"""Wall forces and lift/drag on the GPU (acfd_forces) against the
reference's clcd_output: bit for bit on the reference's fixture, and against
the oracle on the state of a flat-plate run."""
import sys
import time

import numpy as np
import pytest

from aerocfd_amd import FlowConfig, Solver, cases
from oracle import oracle as O
from tests.test_golden import load

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_forces_match_reference_fixture():
    say("test_forces_match_reference_fixture")
    g = load("forces")
    nb = len(g["dims"])
    subs, off = [], 0
    for cnt in g["nsub"]:
        subs.append(g["subfaces"][off:off + cnt].tolist())
        off += cnt
    dims = [tuple(int(v) for v in d) for d in g["dims"]]
    pinf, re, aoa, sref = g["params"]
    cfg = FlowConfig(ma=0.3, re=re, aoa_deg=aoa, iflag_solver=0, iflag_turbulence=2, iflag_time=0)
    sol = Solver(cfg, dims, subs)
    try:
        for m in range(nb):
            for k in ("pri_v", "dxidx", "amu", "spci", "spcj", "spck"):
                sol.upload(m + 1, k, g["%s%d" % (k, m)])
        f = sol.forces(sref)
    finally:
        sol.close()
    say("  cl %.17g (reference %.17g), cd %.17g (reference %.17g)" % (f["cl"], g["clcd"][0], f["cd"], g["clcd"][1]))
    assert f["cl"] == g["clcd"][0] and f["cd"] == g["clcd"][1]


def test_forces_on_flat_plate_state():
    """After 10 steps of the small flat plate with device-computed metrics,
    the device forces equal the oracle's on the same downloaded state."""
    case = cases.flat_plate(nblk=(4, 3, 1), lo=(6, 6, 5), hi=(12, 10, 8))
    say("test_forces_on_flat_plate_state", case.name)
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        case.upload(sol)
        sol.grid_metrics(case.coords)
        sol.run_stage("physical_bc")
        sol.run_stage("getpri")
        sol.prepare()
        sol.step(10)
        t0 = time.perf_counter()
        f = sol.forces(1.0)
        dt = time.perf_counter() - t0
        exts = [O.Ext.from_list(r) for r in sol._ranges]
        fields = [{k: sol.download(m + 1, k) for k in ("pri_v", "dxidx", "amu", "spci", "spcj", "spck")}
                  for m in range(len(exts))]
    finally:
        sol.close()
    fp, fv, afp, afv, cl, cd = O.forces(exts, case.subfaces, fields, 1.0 / 1.4 / 0.3 / 0.3, case.cfg.re, 0.0, 1.0)
    say("  cl %.6e cd %.6e (%.2f ms with the copy back)" % (f["cl"], f["cd"], 1e3 * dt))
    assert np.array_equal(f["fp"], fp) and np.array_equal(f["fv"], fv)
    assert f["cl"] == cl and f["cd"] == cd
    assert cd != 0.0
