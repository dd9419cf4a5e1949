"""Long runs on the HIP path (through the C ABI): 100 whole time steps
against the reference and the oracle, to the path's bar of <= 1e-10
norm-wise relative error per conserved variable after 100 steps.

  * the 100-step fixtures written by the reference's own Fortran
    (tests/golden/make_golden.py: a reduced 12-block SA flat plate, the
    vortex and the Taylor-Green vortex in two blocks), the flat plate on the
    batched path with graph replay (the bench configuration);
  * the 64^3 vortex box of the single-block Euler configuration against the
    oracle: the reference's scheme (WENO3 + Roe, LU-SGS tile sweeps), and
    Steger-Warming + WENO5 with RK3 (the configuration as specified, no
    reference counterpart).
"""
import sys

import numpy as np
import pytest

from aerocfd_amd import FlowConfig, Solver, cases
from oracle import oracle as O
from tests.helpers import ora_config, ora_states, rel_l2
from tests.test_golden import blocks_from, load

pytestmark = pytest.mark.gpu

TOLN = 1e-10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,variant", [("steps100_flat_plate", "batched_graph"),
                                          ("steps100_flat_plate", "per_block"),
                                          ("steps100_vortex", "tiles"), ("steps100_tgv", "tiles")])
def test_100_steps_match_reference(name, variant):
    say("test_100_steps_match_reference", name, variant)
    g = load(name)
    solver, turb, itime, tadv, nsteps, des = [int(x) for x in g["flags"]]
    ma, re, tinf, cfl, dt, aoa = g["params"]
    cfg = FlowConfig(ma=ma, re=re, tinf=tinf, cfl=cfl, dt=dt, aoa_deg=aoa, iflag_solver=solver,
                     iflag_turbulence=turb, iflag_time=itime, iflag_timeadvance=tadv, iflag_des=des)
    if variant == "batched_graph":
        cfg.lusgs_kernel, cfg.batch_blocks, cfg.use_graph = 2, 1, 1
    elif variant == "tiles":
        cfg.lusgs_kernel = 1
    dims = [tuple(int(x) for x in d) for d in g["dims"]]
    sol = Solver(cfg, dims, blocks_from(g))
    try:
        for m in range(len(dims)):
            for k in g:
                pre = "in_"
                suf = str(m)
                if k.startswith(pre) and k.endswith(suf) and k[3:-len(suf)] in O.FIELDS:
                    sol.upload(m + 1, k[3:-len(suf)], g[k])
        sol.step(nsteps)
        worst = 0.0
        for m in range(len(dims)):
            e = rel_l2(sol.download(m + 1, "q"), g["out_q%d" % m])
            worst = max(worst, float(np.max(e)))
            assert np.all(e <= TOLN), "%s block %d q after %d steps: %s" % (name, m + 1, nsteps, e)
            if turb:
                e = rel_l2(sol.download(m + 1, "nut"), g["out_nut%d" % m])
                worst = max(worst, float(e))
                assert e <= TOLN, "%s block %d nut: %s" % (name, m + 1, e)
        say("  worst rel err after %d steps %.3e" % (nsteps, worst))
    finally:
        sol.close()


@pytest.mark.parametrize("scheme", ["weno3_roe_lusgs", "sw_weno5_rk3"])
def test_vortex64_100_steps_against_oracle(scheme):
    """The 64^3 vortex box on [-5,5]^3 (farfield faces, Mach 0.5, strength
    5, dt = 0.5 h / 3), 100 steps, GPU against the oracle."""
    say("test_vortex64_100_steps_against_oracle", scheme)
    case = cases.vortex_box(64)
    if scheme == "weno3_roe_lusgs":
        case.cfg.lusgs_kernel = 1
    else:
        case.cfg.iflag_timeadvance = 4
        case.cfg.inviscid_scheme = 1
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
        sol.step(100)
        q_gpu = sol.download(1, "q")
        res_gpu = sol.residual()
    finally:
        sol.close()
    for n in range(100):
        if scheme == "weno3_roe_lusgs":
            O.step(states, cfg)
        else:
            O.step_rk3(states, cfg, 1)
    e = rel_l2(q_gpu, states[0]["q"])
    say("  q rel err after 100 steps", e)
    assert np.all(e <= TOLN), e
    res_o = O.residual(states, case.totalnodes)
    assert np.allclose(res_gpu[:5], res_o[:5], rtol=1e-8, atol=1e-300), (res_gpu, res_o)
