"""The reference's file formats on the host (aerocfd_amd.formats): the
control file read back by the reference's own init_readcontrolfile, round
trips of grid / bcinfo / flowfield files, list-directed parsing rules and
the fixed-format residual line."""
import os

import numpy as np
import pytest

from aerocfd_amd import cases
from aerocfd_amd import formats as F
from oracle import refwrap as R

CONTROL = dict(iflag_dimension=3, iflag_solver=0, iflag_time=1, cfl=2.5, dt=1.25e-3, ma=0.3, re=5.0e6,
               tinf=288.15, aoa=2.5, scale_x=1.0, scale_y=2.0, scale_z=0.5, lref=1.5, sref=0.75, fltr=0.2,
               iflag_splittingtype=1, iflag_inviscid=4, iflag_filter=0, iflag_timeadvance=2,
               iflag_turbulence=2, iflag_des=1, writeinterval=100, ntstep=5000, grid="input/plate.grd",
               bcinfo="input/plate.inp", iflag_init=0, restartfile="flowfield", nfile=1, restart_timestep=0,
               nobs=0, ob_dist=0.0, acoustic_length=1.0, iflag_acoustic=0, Start_Acsoutic=0, End_Acsoutic=0,
               iflag_avg=0, Start_Avg=0, End_Avg=0, Avg_interval=1)


def test_list_directed_rules():
    assert F.list_directed("1, 2  3,4") == ["1", "2", "3", "4"]
    assert F.list_directed("3*0.5 'a b' \"c\" / 9") == ["0.5", "0.5", "0.5", "a b", "c"]
    assert F._real("1.5d-3") == 1.5e-3 and F._real("2.0D+2") == 200.0
    # a slash ends the record, also inside an unquoted name (quote paths)
    assert F.list_directed("  input/grid.dat  7") == ["input"]
    assert F.list_directed("  'input/grid.dat'  7") == ["input/grid.dat", "7"]


def test_control_round_trip(tmp_path):
    p = tmp_path / "Control.in"
    F.write_control(str(p), CONTROL)
    assert F.read_control(str(p)) == CONTROL
    cfg = F.flow_config(F.read_control(str(p)))
    assert cfg.ma == 0.3 and cfg.iflag_turbulence == 2 and cfg.aoa_deg == 2.5


@pytest.mark.skipif(not R.available(), reason="compiled reference (oracle/_ref) not built")
def test_control_read_by_the_reference(tmp_path):
    """The file written here, read by the reference's init_readcontrolfile,
    gives the same values (aoa as read, before its conversion to radians)."""
    os.chdir(tmp_path)
    os.makedirs("input")
    R.setup()
    R.begin([(5, 5, 5)], [[[f, f, 1, 5, 1, 5, 1, 5, -10, 0, 0] for f in range(1, 7)]])  # MPI up
    F.write_control("input/Control.in", CONTROL)
    got = R.read_control()
    mine = F.read_control("input/Control.in")
    for k, v in got.items():
        assert mine[k] == v, (k, mine[k], v)


def test_grid_and_bcinfo_round_trip(tmp_path):
    dims, subs, ranges, xyz = cases.warped_grid(seed=2)
    phys = []
    for r, (x, y, z) in zip(ranges, xyz):
        ph = np.s_[r[10] - r[4]:r[11] - r[4] + 1, r[8] - r[2]:r[9] - r[2] + 1, r[6] - r[0]:r[7] - r[0] + 1]
        phys.append((x[ph], y[ph], z[ph]))
    F.write_grid(str(tmp_path / "g.grd"), dims, phys)
    d2, c2 = F.read_grid(str(tmp_path / "g.grd"))
    assert [tuple(d) for d in dims] == d2
    for a, b in zip(phys, c2):
        for u, v in zip(a, b):
            assert np.array_equal(u, v)
    d3, c3 = F.read_grid(str(tmp_path / "g.grd"), scale=(2.0, 1.0, 3.0), lref=4.0)
    assert np.array_equal(c3[0][0], phys[0][0] * 2.0 / 4.0)
    F.write_bcinfo(str(tmp_path / "b.inp"), subs)
    bc = F.read_bcinfo(str(tmp_path / "b.inp"), len(dims))
    assert F.subfaces_of(bc) == [[list(r) for r in s] for s in subs]
    assert all(r[11] == 0 for s in bc for r in s)


def test_flowfield_round_trip(tmp_path):
    case = cases.flat_plate(nblk=(2, 2, 1), lo=(5, 5, 4), hi=(7, 6, 5))
    ma, g = case.cfg.ma, 1.4
    pri, nut = [], []
    rng = np.random.default_rng(0)
    for m, fl in enumerate(case.fields):
        q = fl["q"]
        p = np.zeros(q.shape[:-1] + (7,))
        p[..., 0] = 1.0 + 0.1 * rng.random(q.shape[:-1])
        p[..., 1:4] = rng.standard_normal(q.shape[:-1] + (3,)) * 0.1
        p[..., 4] = 1.0 / g / ma / ma * (1 + 0.05 * rng.random(q.shape[:-1]))
        pri.append(p)
        nut.append(rng.random(q.shape[:-1]))
    F.write_flowfield(str(tmp_path / "f.dat"), case.ranges, case.coords, pri, nut)
    back = F.read_flowfield(str(tmp_path / "f.dat"), case.dims, ma)
    for r, p, nt, b in zip(case.ranges, pri, nut, back):
        ph = np.s_[r[10] - r[4]:r[11] - r[4] + 1, r[8] - r[2]:r[9] - r[2] + 1, r[6] - r[0]:r[7] - r[0] + 1]
        assert np.array_equal(b["nut"], nt[ph]) and np.array_equal(b["q"][..., 0], p[ph + (0,)])
        # momentum and energy from the primitive values as the restart reader forms them
        assert np.allclose(b["q"][..., 1], p[ph + (0,)] * p[ph + (1,)], rtol=1e-15)
        e = p[ph + (4,)] / (g - 1) + 0.5 * p[ph + (0,)] * (p[ph + (1,)] ** 2 + p[ph + (2,)] ** 2 + p[ph + (3,)] ** 2)
        assert np.allclose(b["q"][..., 4], e, rtol=1e-13)


def test_residual_line_format():
    line = F.residual_line(42, [1.23456e-3, -5.0e-7, 0.0, 1.0, 99999.0, 2.5e-12])
    assert line == "000042 0.1235E-02-0.5000E-06 0.0000E+00 0.1000E+01 0.1000E+06 0.2500E-11"
