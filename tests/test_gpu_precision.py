"""Precision of the FP32-flux / FP64-update variant (precision = 1: WENO
reconstruction, Roe splitting and the viscous fluxes in FP32; the
conservative update, metrics, local dt, residual norms and the implicit
solve in FP64) against the FP64 oracle, per configuration, with the FP64
GPU path beside it:

  * density relative L2 and Linf error after N steps;
  * lift and drag (clcd_output on the device, acfd_forces) against the
    oracle's clcd_output on its own state, for the wall-bounded cases;
  * the kinetic-energy history KE(t) = sum(rho |u|^2 / 2) of the
    Taylor-Green vortex, its largest relative error over the steps.

Bounds: FP64 within the path's 1e-10; the FP32 variant within 1e-4
relative L2 in density (the accuracy budget of the variant).  Every figure
is printed for the record."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from oracle import oracle as O
from tests.helpers import ora_config, ora_states
from tests.test_formats_ref import oracle_whole_run

pytestmark = pytest.mark.gpu

FIELDS = ("pri_v", "dxidx", "amu", "spci", "spcj", "spck")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def density_err(qs, states):
    num2 = den2 = 0.0
    linf = ref_max = 0.0
    for q, st in zip(qs, states):
        e = st.ext
        ph = np.s_[e.ks0 - e.ks:e.ke0 - e.ks + 1, e.js0 - e.js:e.je0 - e.js + 1, e.is0 - e.is_:e.ie0 - e.is_ + 1]
        d = q[ph][..., 0] - st["q"][ph][..., 0]
        num2 += float(np.sum(d * d))
        den2 += float(np.sum(st["q"][ph][..., 0] ** 2))
        linf = max(linf, float(np.abs(d).max()))
        ref_max = max(ref_max, float(np.abs(st["q"][ph][..., 0]).max()))
    return np.sqrt(num2 / den2), linf / ref_max


def run_from_coords(case, nsteps, precision):
    """GPU run of a wall-bounded case from its coordinates (metrics, edge
    lengths and wall distance on the device), uniform start; q of every
    block and the device lift / drag."""
    case.cfg.precision = precision
    case.cfg.lusgs_kernel, case.cfg.batch_blocks = 2, 1
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        for m, fl in enumerate(case.fields):
            for k in ("q", "amu", "nut"):
                sol.upload(m + 1, k, fl[k])
            dt = np.full(fl["q"].shape[:-1], case.cfg.dt)
            sol.upload(m + 1, "dt", dt)
        sol.grid_metrics(case.coords)
        sol.run_stage("physical_bc")
        sol.run_stage("getpri")
        sol.prepare()
        sol.step(nsteps)
        qs = [sol.download(m + 1, "q") for m in range(len(case.dims))]
        f = sol.forces(1.0)
    finally:
        sol.close()
    return qs, f


@pytest.mark.parametrize("kind", ["flat_plate", "channel_bump"])
def test_precision_wall_bounded(kind):
    nsteps = 30
    if kind == "flat_plate":
        make = lambda: cases.flat_plate(nblk=(4, 3, 1), lo=(10, 10, 6), hi=(14, 13, 8), first=2e-4)  # noqa: E731
    else:
        make = lambda: cases.channel_bump(lo=(12, 10, 8), hi=(15, 12, 9), first=2.0e-4)  # noqa: E731
    case = make()
    say("test_precision_wall_bounded", case.name, case.totalnodes, "nodes,", nsteps, "steps")
    states, _ = oracle_whole_run(case.cfg, case.dims, case.subfaces, case.coords, nsteps)
    exts = [st.ext for st in states]
    flds = []
    for st, (x, y, z) in zip(states, case.coords):
        sp = O.spacing(st.ext, x, y, z)
        f = {k: st[k] for k in ("pri_v", "dxidx", "amu")}
        f["spci"], f["spcj"], f["spck"] = (np.nan_to_num(a, nan=0.0) for a in sp[:3])
        flds.append(f)
    pinf = 1.0 / 1.4 / case.cfg.ma ** 2
    _, _, _, _, cl_o, cd_o = O.forces(exts, case.subfaces, flds, pinf, case.cfg.re, case.cfg.aoa_deg, 1.0)
    out = {}
    for prec in (0, 1):
        qs, f = run_from_coords(make(), nsteps, prec)
        l2, linf = density_err(qs, states)
        ecl = abs(f["cl"] - cl_o) / max(abs(cl_o), 1e-300)
        ecd = abs(f["cd"] - cd_o) / max(abs(cd_o), 1e-300)
        out[prec] = (l2, linf, ecl, ecd)
        say("  %s: density rel L2 %.3e, rel Linf %.3e; cl %.6e (oracle %.6e) rel err %.3e; cd %.6e (oracle %.6e) "
            "rel err %.3e" % ("FP64" if prec == 0 else "FP32 flux / FP64 update", l2, linf, f["cl"], cl_o, ecl,
                              f["cd"], cd_o, ecd))
    assert out[0][0] <= 1e-10 and out[0][3] <= 1e-8
    assert 0.0 < out[1][0] <= 1e-4
    assert out[1][3] <= 1e-3


def test_precision_vortex_and_tgv_ke():
    """The vortex box (Euler) and the Taylor-Green vortex (laminar NS), 32^3,
    40 steps: density error of both precisions and the TGV's KE(t) error."""
    res = {}
    for kind in ("vortex", "tgv"):
        make = (lambda: cases.vortex_box(32)) if kind == "vortex" else (lambda: cases.tgv_box(32))  # noqa: E731
        base = make()
        cfg = ora_config(base.cfg, base.totalnodes)
        states = ora_states(base)
        O.physical_bc(states, cfg)
        for st in states:
            st["pri_v"] = O.getpri(st.ext, st["q"], cfg.gamma, cfg.cv)[0]
        O.prepare(states, cfg)
        ke_o = []
        for _ in range(40):
            O.step(states, cfg)
            q = states[0]["q"]
            ke_o.append(float(np.sum(0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / q[..., 0])))
        for prec in (0, 1):
            c = make()
            c.cfg.precision = prec
            c.cfg.lusgs_kernel = 1
            sol = Solver(c.cfg, c.dims, c.subfaces)
            ke = []
            try:
                c.upload(sol)
                sol.run_stage("physical_bc")
                sol.run_stage("getpri")
                sol.prepare()
                for _ in range(40):
                    sol.step(1)
                    q = sol.download(1, "q")
                    ke.append(float(np.sum(0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / q[..., 0])))
            finally:
                sol.close()
            l2, linf = density_err([q], states)
            eke = max(abs(a - b) / abs(b) for a, b in zip(ke, ke_o))
            res[(kind, prec)] = (l2, linf, eke)
            say("  %s %s: density rel L2 %.3e, rel Linf %.3e; KE(t) largest rel err %.3e (KE decays %.4f -> %.4f)" % (
                kind, "FP64" if prec == 0 else "FP32 flux / FP64 update", l2, linf, eke, ke_o[0], ke_o[-1]))
    for kind in ("vortex", "tgv"):
        assert res[(kind, 0)][0] <= 1e-10 and res[(kind, 0)][2] <= 1e-10
        assert 0.0 < res[(kind, 1)][0] <= 1e-4 and res[(kind, 1)][2] <= 1e-4
