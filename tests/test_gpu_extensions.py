"""GPU parity of the extensions with no counterpart in the reference
(Steger-Warming + WENO5, WCNS, explicit RK3) against the C oracle, through
the C ABI.  Same bars as the rest of the path: <= 1e-12 norm-wise relative
error per conserved variable for one residual evaluation, <= 1e-10 after
several steps."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from oracle import oracle as O
from tests.helpers import ora_config, ora_states, rel_l2
from tests.test_gpu_parity import two_block_case

pytestmark = pytest.mark.gpu

TOL1 = 1e-12
TOLN = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("scheme", [1, 2])
@pytest.mark.parametrize("ns", [False, True])
def test_stage_inviscid_ext(scheme, ns):
    say("test_stage_inviscid_ext scheme=%d ns=%s two_block 12x9x8" % (scheme, ns))
    case = two_block_case(ns=ns, seed=3)
    case.cfg.inviscid_scheme = scheme
    cfg = ora_config(case.cfg, case.totalnodes)
    states = ora_states(case)
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        case.upload(sol)
        sol.run_stage("inviscid")
        for m, st in enumerate(states):
            ref = O.inviscid_ext(st.ext, st["q"], st["dxidx"], st["inv_j"], cfg.gamma, scheme)
            e = rel_l2(sol.download(m + 1, "rhsi"), ref)
            say("  block", m + 1, "rhsi rel err", e)
            assert np.all(e <= TOL1), e
    finally:
        sol.close()


def run_rk3(case, scheme, nsteps):
    case.cfg.iflag_timeadvance = 4
    case.cfg.inviscid_scheme = scheme
    cfg = ora_config(case.cfg, case.totalnodes)
    states = ora_states(case)
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        case.upload(sol)
        O.physical_bc(states, cfg)
        for st in states:
            st["pri_v"] = O.getpri(st.ext, st["q"], cfg.gamma, cfg.cv)[0]
        sol.run_stage("physical_bc")
        sol.run_stage("getpri")
        O.prepare(states, cfg)
        sol.prepare()
        for _ in range(nsteps):
            O.step_rk3(states, cfg, scheme)
        sol.step(nsteps)
        out = [(sol.download(m + 1, "q"), st["q"]) for m, st in enumerate(states)]
        return out, sol.residual(), O.residual(states, case.totalnodes)
    finally:
        sol.close()


@pytest.mark.parametrize("scheme", [0, 1, 2])
def test_rk3_vortex_steps(scheme):
    say("test_rk3_vortex_steps scheme=%d vortex 20^3 two blocks, 8 steps" % scheme)
    case = cases.vortex_box(20, blocks=((10,), (), ()))
    out, rg, ro = run_rk3(case, scheme, 8)
    for m, (g, o) in enumerate(out):
        e = rel_l2(g, o)
        say("  block", m + 1, "q rel err", e)
        assert np.all(e <= TOLN), e
    assert np.allclose(rg[:5], ro[:5], rtol=1e-9, atol=1e-300)


def test_rk3_tgv_wcns_steps():
    say("test_rk3_tgv_wcns_steps tgv 16^3 two blocks, NS, 5 steps")
    case = cases.tgv_box(16, blocks=((8,), (), ()))
    out, rg, ro = run_rk3(case, 2, 5)
    for m, (g, o) in enumerate(out):
        e = rel_l2(g, o)
        say("  block", m + 1, "q rel err", e)
        assert np.all(e <= TOLN), e


@pytest.mark.parametrize("scheme", [1, 2])
def test_ext_fp32_flux_variant(scheme):
    """FP32 splitting / interpolation / flux with the FP64 update: density
    error against the FP64 path within 1e-4 relative L2 after 10 RK3 steps."""
    say("test_ext_fp32_flux_variant scheme=%d vortex 24^3" % scheme)
    qs = []
    for prec in (0, 1):
        case = cases.vortex_box(24)
        case.cfg.iflag_timeadvance = 4
        case.cfg.inviscid_scheme = scheme
        case.cfg.precision = prec
        sol = Solver(case.cfg, case.dims, case.subfaces)
        try:
            case.upload(sol)
            sol.run_stage("physical_bc")
            sol.run_stage("getpri")
            sol.prepare()
            sol.step(10)
            qs.append(sol.download(1, "q"))
        finally:
            sol.close()
    d = qs[1][..., 0] - qs[0][..., 0]
    l2 = np.linalg.norm(d) / np.linalg.norm(qs[0][..., 0])
    say("  density rel L2 %.3e, rel Linf %.3e" % (l2, np.abs(d).max() / np.abs(qs[0][..., 0]).max()))
    assert 0.0 < l2 <= 1e-4
