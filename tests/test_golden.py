"""The C oracle against the golden fixtures written from the reference's own
Fortran (tests/golden/make_golden.py): every routine and whole multi-block
time steps must agree bit for bit.  These run everywhere; the fixtures carry
their inputs."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def ext_of(a):
    return O.Ext.from_list(a)


def test_manifest_lists_every_fixture():
    with open(os.path.join(GOLD, "manifest.json")) as fh:
        man = json.load(fh)
    names = [k for k in man if not k.startswith("_")]
    assert len(names) == 18
    for k in names:
        assert os.path.exists(os.path.join(GOLD, man[k]["file"]))


def test_kernels_match_reference():
    g = load("kernels")
    e = ext_of(g["ext"])
    gamma, cv, re, tinf, prl, prt = g["scalars"]
    q, dx, ij, amu, dt = g["q"], g["dxidx"], g["inv_j"], g["amu"], g["dt"]
    pr = O.getpri(e, q, gamma, cv)[0]
    assert np.array_equal(pr, g["pri_v"])
    for c in (1, 2, 3):
        for d, s in ((1, "l"), (-1, "r")):
            assert np.array_equal(O.weno3(e, q, c, d), g["weno3_c%d_%s" % (c, s)]), (c, s)
    ri = O.inviscid(e, q, dx, ij, gamma)
    assert np.array_equal(ri, g["rhsi"])
    cfg = O.make_config(ma=0.5, re=re, tinf=tinf, solver=0)
    assert cfg.gamma == gamma and cfg.cv == cv and cfg.prl == prl and cfg.prt == prt
    out = O.viscous(e, pr, amu, dx, ij, cfg)
    for a, k in zip(out, ("rhsv", "amu", "vor", "dudx", "tao")):
        assert np.array_equal(a, g["viscous_" + k]), k
    for it in (0, 1):
        d2, cfl2, st = O.find_dt(e, dt, pr, amu, dx, re, 0.5, it)
        assert np.array_equal(d2, g["find_dt%d_dt" % it])
        assert cfl2 == g["find_dt%d_cfl" % it][0]
        assert np.array_equal(np.asarray(st, dtype=np.float64), g["find_dt%d_stats" % it])
    a = O.implicit1st(e, q, dt, pr, ri, g["rhsv_in"], dx, ij, gamma, 0.5)
    assert np.array_equal(a[0], g["implicit1st_q"]) and np.array_equal(a[1], g["implicit1st_rhs"])
    a = O.crank2nd(e, q, g["q0"], dt, pr, ri, g["rhsv_in"], dx, ij, gamma)
    assert np.array_equal(a[0], g["crank2nd_q"]) and np.array_equal(a[1], g["crank2nd_rhs"])
    for v, ref in zip(g["calc_jm_q"], g["calc_jm"]):
        assert np.array_equal(O.calc_jm(v, g["calc_jm_m"], gamma), ref)


@pytest.mark.parametrize("des", [0, 1, 2])
def test_sa_model_matches_reference(des):
    g = load("sa_des%d" % des)
    e = ext_of(g["ext"])
    out = O.sa_model(e, g["nut_in"], g["dt"], g["pri_v"], g["amu"], g["dudx"], g["vor"], g["dxidx"],
                     g["inv_j"], g["aalpha"], g["bbeta"], g["ggamma"], g["dist"], g["spacing"],
                     float(g["re"][0]), float(g["cfl"][0]), des, 0)
    for a, k in zip(out, ("nut", "sa_rhs", "length")):
        assert np.array_equal(a, g[k]), k


def blocks_from(g):
    subs = g["subfaces"]
    out, o = [], 0
    for m, n in enumerate(g["nsub"]):
        out.append(subs[o:o + n].tolist())
        o += n
    return out


@pytest.mark.parametrize("ori", [1, 2, 3, 4])
def test_halo_and_average_match_reference(ori):
    g = load("halo_ori%d" % ori)
    subs = blocks_from(g)
    cfg = O.make_config(solver=0, turb=2, itime=0)
    states = []
    for m in range(2):
        st = O.BlockState(ext_of(g["ext%d" % m]), subs[m])
        st["q"], st["nut"] = g["q_in%d" % m], g["nut_in%d" % m]
        states.append(st)
    O.update_buffer(states, cfg)
    O.average(states, cfg)
    for m in range(2):
        assert np.array_equal(states[m]["q"], g["q%d" % m])
        assert np.array_equal(states[m]["nut"], g["nut%d" % m])


STEP_CASES = ["steps_s1_t0_a2", "steps_s0_t0_a2", "steps_s0_t2_a2", "steps_s0_t0_a3", "steps_s1_t0_a3"]
STEP_INPUTS = ("q", "q0", "dxidx", "inv_j", "amu", "dt", "nut", "dist", "spacing", "aalpha", "bbeta",
               "ggamma", "pri_v")
STEP_OUTPUTS = ("q", "nut", "rhs", "dt", "sa_rhs", "pri_v", "amu")


@pytest.mark.parametrize("name", STEP_CASES)
def test_steps_match_reference(name):
    g = load(name)
    solver, turb, itime, tadv, nsteps = [int(x) for x in g["flags"]]
    ma, re, tinf, cfl, dt, aoa = g["params"]
    subs = blocks_from(g)
    dims = g["dims"]
    cfg = O.make_config(ma=ma, re=re, tinf=tinf, cfl=cfl, aoa_deg=aoa, solver=solver, itime=itime,
                        turb=turb, tadv=tadv, totalnodes=int(sum(int(np.prod(d)) for d in dims)))
    states = []
    for m in range(len(dims)):
        st = O.BlockState(ext_of(g["ext%d" % m]), subs[m])
        for k in STEP_INPUTS:
            st[k] = g["in_%s%d" % (k, m)]
        states.append(st)
    for _ in range(nsteps):
        O.step(states, cfg)
    for m in range(len(dims)):
        for k in STEP_OUTPUTS:
            assert np.array_equal(states[m][k], g["out_%s%d" % (k, m)]), (m, k)


RUN100 = ["steps100_flat_plate", "steps100_vortex", "steps100_tgv"]


def run100_setup(g):
    """Oracle states and config of a 100-step fixture (make_golden.case_run)."""
    solver, turb, itime, tadv, nsteps, des = [int(x) for x in g["flags"]]
    ma, re, tinf, cfl, dt, aoa = g["params"]
    subs = blocks_from(g)
    dims = g["dims"]
    cfg = O.make_config(ma=ma, re=re, tinf=tinf, cfl=cfl, aoa_deg=aoa, solver=solver, itime=itime,
                        turb=turb, des=des, tadv=tadv, totalnodes=int(sum(int(np.prod(d)) for d in dims)))
    states = []
    for m in range(len(dims)):
        st = O.BlockState(ext_of(g["ext%d" % m]), subs[m])
        for k in g:
            if k.startswith("in_") and k.endswith(str(m)) and k[3:-len(str(m))] in O.FIELDS:
                st[k[3:-len(str(m))]] = g[k]
        states.append(st)
    return states, cfg, nsteps


@pytest.mark.parametrize("name", RUN100)
def test_100_steps_match_reference(name):
    """100 whole steps of the reduced flat plate, vortex and Taylor-Green
    cases: the oracle equals the reference bit for bit."""
    g = load(name)
    states, cfg, nsteps = run100_setup(g)
    assert nsteps == 100
    for _ in range(nsteps):
        O.step(states, cfg)
    for m, st in enumerate(states):
        for k in ("q", "nut", "pri_v"):
            assert np.array_equal(st[k], g["out_%s%d" % (k, m)]), (name, m, k)


def same_where_defined(a, ref):
    """Bitwise equality wherever the reference defines the output (the
    fixture holds NaN elsewhere)."""
    m = ~np.isnan(ref)
    assert m.any()
    return np.array_equal(a[m], ref[m])


def test_grid_preprocessing_matches_reference():
    """gridtransformation, init_spacing and init_walldist on the warped
    4-block grid with interface halos and wall faces."""
    g = load("grid")
    exts = [ext_of(g["ext%d" % m]) for m in range(len(g["dims"]))]
    subs, off = [], 0
    for cnt in g["nsub"]:
        subs.append(g["subfaces"][off:off + cnt].tolist())
        off += cnt
    xyz = [(g["x%d" % m], g["y%d" % m], g["z%d" % m]) for m in range(len(exts))]
    wall = O.wall_points(exts, subs, xyz)
    for m, e in enumerate(exts):
        x, y, z = xyz[m]
        for name, a in zip(("dxidx", "inv_j", "aalpha", "bbeta", "ggamma"), O.gridtransformation(e, x, y, z)):
            assert same_where_defined(a, g["%s%d" % (name, m)]), (m, name)
        for name, a in zip(("spci", "spcj", "spck", "spacing", "vol"), O.spacing(e, x, y, z)):
            assert same_where_defined(a, g["%s%d" % (name, m)]), (m, name)
        assert same_where_defined(O.walldist(e, x, y, z, wall), g["dist%d" % m]), m


def test_forces_match_reference():
    g = load("forces")
    exts = [ext_of(g["ext%d" % m]) for m in range(len(g["dims"]))]
    subs, off = [], 0
    for cnt in g["nsub"]:
        subs.append(g["subfaces"][off:off + cnt].tolist())
        off += cnt
    fields = [{k: g["%s%d" % (k, m)] for k in ("pri_v", "dxidx", "amu", "spci", "spcj", "spck")}
              for m in range(len(exts))]
    pinf, re, aoa, sref = g["params"]
    cl, cd = O.forces(exts, subs, fields, pinf, re, aoa, sref)[4:]
    assert cl == g["clcd"][0] and cd == g["clcd"][1]
