"""ctypes access to the compiled reference (TEST INFRASTRUCTURE ONLY).

``_ref/libaerocfd_ref.so`` is the reference's own Fortran (jiamin0329/aerocfd,
``src/*.f90`` except the program unit) compiled in place by ``oracle/Makefile``
with amdflang -O2, plus ``ref_harness.f90`` which sets up the module state.
It exists only in the development container, where the reference tree is
mounted; tests that need it skip elsewhere.

Per-block subroutines are called directly with the Fortran calling
convention (every argument by reference).  Module-state routines go through
the ``refh_*`` entry points of the harness.
"""
import ctypes as C
import os

import numpy as np

from .oracle import Ext, ptr

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_DP = C.POINTER(C.c_double)

XFER = {"q": 1, "pri_v": 2, "rhsi": 3, "rhsv": 4, "rhs": 5, "dudx": 6, "tao": 7, "vor": 8,
        "amu": 9, "dt": 10, "dxidx": 11, "inv_j": 12, "aalpha": 13, "bbeta": 14, "ggamma": 15,
        "dist": 16, "spacing": 17, "nut": 18, "sa_rhs": 19, "q0": 20, "length": 21, "x": 22,
        "y": 23, "z": 24, "spci": 25, "spcj": 26, "spck": 27, "vol": 28}
NCOMP = {"q": 5, "pri_v": 7, "rhsi": 5, "rhsv": 5, "rhs": 5, "dudx": 9, "tao": 6, "amu": 3,
         "dxidx": 9, "aalpha": 3, "bbeta": 3, "ggamma": 3, "q0": 10}


def available():
    return os.path.exists(os.path.join(_HERE, "_ref", "libaerocfd_ref.so"))


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(os.path.join(_HERE, "_ref", "libaerocfd_ref.so"), mode=C.RTLD_GLOBAL)
    return _LIB


def _i(x):
    return C.byref(C.c_int(int(x)))


def _d(x):
    return C.byref(C.c_double(float(x)))


def _ext_args(e):
    return [_i(v) for v in e.as_list()]


def setup(dimen=3, solver=1, itime=1, turb=0, des=0, tadv=2, ma=0.5, re=1.0e6, tinf=288.15,
          cfl=0.5, dt=0.01, aoa_deg=0.0):
    lib().refh_setup(C.c_int(dimen), C.c_int(solver), C.c_int(itime), C.c_int(turb), C.c_int(des),
                     C.c_int(tadv), C.c_double(ma), C.c_double(re), C.c_double(tinf),
                     C.c_double(cfl), C.c_double(dt), C.c_double(aoa_deg))


def scalars():
    out = np.zeros(12)
    lib().refh_scalars(ptr(out))
    return dict(zip(["gamma", "cv", "re", "tinf", "prl", "prt", "d0", "u0", "p0", "c0", "aoa",
                     "cfl"], out))


def set_cfl(x):
    lib().refh_set_cfl(C.c_double(x))


def stats():
    out = np.zeros(4)
    lib().refh_stats(ptr(out))
    return out


def begin(dims, subfaces):
    """dims: list of (ni,nj,nk); subfaces: list (per block) of lists of
    11-int records (f_no, face, is, ie, js, je, ks, ke, blk_t, subface_t, ori)."""
    nb = len(dims)
    d = np.ascontiguousarray(np.array(dims, dtype=np.int32).reshape(nb, 3))
    ns = np.array([len(s) for s in subfaces], dtype=np.int32)
    tab = np.ascontiguousarray(np.array([r for s in subfaces for r in s], dtype=np.int32).reshape(-1, 11))
    lib().refh_begin(C.c_int(nb), d.ctypes.data_as(C.POINTER(C.c_int)),
                     ns.ctypes.data_as(C.POINTER(C.c_int)), tab.ctypes.data_as(C.POINTER(C.c_int)))
    return [get_ext(m + 1) for m in range(nb)]


def get_ext(m):
    r = np.zeros(18, dtype=np.int32)
    lib().refh_get_ext(C.c_int(m), r.ctypes.data_as(C.POINTER(C.c_int)))
    return Ext.from_list(r)


def put(m, name, arr):
    a = np.ascontiguousarray(arr, dtype=np.float64)
    lib().refh_xfer(C.c_int(m), C.c_int(XFER[name]), ptr(a), C.c_long(a.size), C.c_int(0))


def get(m, name, ext):
    nk, nj, ni = ext.shape
    nc = NCOMP.get(name, 1)
    shp = (nk, nj, ni) if nc == 1 else (nk, nj, ni, nc)
    a = np.zeros(shp)
    lib().refh_xfer(C.c_int(m), C.c_int(XFER[name]), ptr(a), C.c_long(a.size), C.c_int(1))
    return a


def physical_bc():
    lib().refh_physical_bc()


def update_buffer():
    lib().refh_update_buffer()


def average():
    lib().refh_average()


def update_coordinate():
    lib().refh_update_coordinate()


def residual_call():
    lib().refh_residual()


def step():
    lib().refh_step()


# ------------------------------------------------------ per-block subroutines
def getpri(ext, q, gamma, cv):
    pri = np.zeros(q.shape[:-1] + (7,))
    e = ext.as_list()
    lib().get_primitivevariables_(_i(1), ptr(pri), ptr(q), _d(gamma), _d(cv), *[_i(v) for v in e[:6]])
    return pri


def weno3(ext, f, coord, direction):
    df = np.zeros_like(f)
    e = ext.as_list()
    lib().thirdorder_weno_(ptr(df), ptr(f), *[_i(v) for v in e[:6]], *[_i(v) for v in e[12:18]],
                           _i(coord), _i(direction))
    return df


def roe(ext, ql, qr, dxidx, inv_j, gamma, coord):
    flux = np.zeros_like(ql)
    e = ext.as_list()
    lib().roe_3d_(ptr(ql), ptr(qr), ptr(flux), ptr(dxidx), ptr(inv_j), _d(gamma),
                  *[_i(v) for v in e[:6]], *[_i(v) for v in e[12:18]], _i(coord))
    return flux


def inviscid(ext, q, dxidx, inv_j, gamma):
    rhsi = np.zeros_like(q)
    lib().inviscid_flux_1_(ptr(rhsi), ptr(q), ptr(dxidx), ptr(inv_j), *_ext_args(ext), _d(gamma))
    return rhsi


def viscous(ext, pri_v, amu, dxidx, inv_j, tinf, prl, prt, gamma, cv, re):
    shp = pri_v.shape[:-1]
    rhsv = np.zeros(shp + (5,))
    amu = amu.copy()
    vor = np.zeros(shp)
    dudx = np.zeros(shp + (9,))
    tao = np.zeros(shp + (6,))
    lib().viscous_flux_(ptr(rhsv), ptr(pri_v), ptr(amu), ptr(vor), ptr(dudx), ptr(tao), ptr(dxidx),
                        ptr(inv_j), _d(tinf), _d(prl), _d(prt), _d(gamma), _d(cv), _d(re),
                        *_ext_args(ext), _i(1))
    return rhsv, amu, vor, dudx, tao


def find_dt(ext, dt, pri_v, amu, dxidx, re, cfl, iflag_time):
    dt = dt.copy()
    st = [C.c_double(0) for _ in range(4)]
    c = C.c_double(cfl)
    e = ext.as_list()
    lib().find_dt_(ptr(dt), *[C.byref(x) for x in st], ptr(pri_v), ptr(amu), ptr(dxidx), _d(re),
                   C.byref(c), *[_i(v) for v in e[:6]], _i(iflag_time), _i(0), _i(1))
    return dt, c.value, np.array([x.value for x in st])


def calc_jm(qv, m, gamma):
    a = np.zeros(25)
    lib().calc_jm_(ptr(a), *[_d(x) for x in qv], *[_d(x) for x in m], _d(gamma))
    return a.reshape(5, 5).T.copy()


def implicit1st(ext, q, dt, pri_v, rhsi, rhsv, dxidx, inv_j, gamma, cfl):
    q = q.copy()
    rhs = np.zeros_like(q)
    nk, nj, ni = ext.shape
    xx = np.zeros((nk, nj, ni))
    lib().implicit1st_(ptr(q), ptr(dt), ptr(pri_v), ptr(rhs), ptr(rhsi), ptr(rhsv), ptr(dxidx),
                       ptr(inv_j), *_ext_args(ext), _d(gamma), _d(cfl), ptr(xx), ptr(xx), ptr(xx))
    return q, rhs


def crank2nd(ext, q, q0, dt, pri_v, rhsi, rhsv, dxidx, inv_j, gamma):
    q = q.copy()
    rhs = np.zeros_like(q)
    lib().crank2nd_(ptr(q), ptr(q0), ptr(dt), ptr(pri_v), ptr(rhs), ptr(rhsi), ptr(rhsv),
                    ptr(dxidx), ptr(inv_j), *_ext_args(ext), _d(gamma), _i(1), _i(1))
    return q, rhs


def sa_model(ext, nut, dt, pri_v, amu, dudx, vor, dxidx, inv_j, alpha, beta, gam, dist, spac, re,
             cfl):
    nut = nut.copy()
    sa_rhs = np.zeros_like(nut)
    length = np.zeros_like(nut)
    lib().sa_model_(ptr(nut), ptr(sa_rhs), ptr(dt), ptr(pri_v), ptr(amu), ptr(dudx), ptr(vor),
                    ptr(dxidx), ptr(inv_j), ptr(alpha), ptr(beta), ptr(gam), ptr(length), ptr(dist),
                    ptr(spac), _d(re), _d(cfl), *_ext_args(ext))
    return nut, sa_rhs, length


def gridtransformation(ext, x, y, z):
    nk, nj, ni = ext.shape
    dxidx = np.zeros((nk, nj, ni, 9))
    inv_j = np.zeros((nk, nj, ni))
    al = np.zeros((nk, nj, ni, 3))
    be = np.zeros((nk, nj, ni, 3))
    ga = np.zeros((nk, nj, ni, 3))
    e = ext.as_list()
    lib().gridtransformation_(ptr(dxidx), ptr(inv_j), ptr(al), ptr(be), ptr(ga), ptr(x), ptr(y),
                              ptr(z), *[_i(v) for v in e[:12]], _i(1))
    return dxidx, inv_j, al, be, ga


def init_spacing(ext, x, y, z):
    nk, nj, ni = ext.shape
    out = [np.zeros((nk, nj, ni)) for _ in range(5)]
    e = ext.as_list()
    lib().init_spacing_(*[ptr(a) for a in out], ptr(x), ptr(y), ptr(z), *[_i(v) for v in e[:12]])
    spci, spcj, spck, spacing, vol = out
    return spacing, vol


def init_spacing_all(ext, x, y, z):
    """spci, spcj, spck, spacing, vol of sub_init_spacing.f90:16-110."""
    nk, nj, ni = ext.shape
    out = [np.zeros((nk, nj, ni)) for _ in range(5)]
    e = ext.as_list()
    lib().init_spacing_(*[ptr(a) for a in out], ptr(x), ptr(y), ptr(z), *[_i(v) for v in e[:12]])
    return out


def walldist():
    """init_walldist over the blocks given to begin() (coordinates put as
    x, y, z); writes result/dist_*.dat in the current directory."""
    lib().refh_walldist()


def set_flag(name, value):
    """Set an integer flag of module flag_var directly."""
    C.c_int.in_dll(lib(), "_QMflag_varE" + name).value = int(value)


def clcd(sref):
    """clcd_output (sub_output_forceintegration.f90:8-478) over the blocks
    given to begin(); returns (time, cl, cd) as the reference writes them to
    result/clcd_history.dat (list-directed, every digit of a double)."""
    lib().refh_clcd(C.c_double(sref))
    with open("result/clcd_history.dat") as fh:
        last = fh.read().split()
    return tuple(float(v.replace("D", "E")) for v in last[-3:])


CONTROL_INTS = ("iflag_dimension", "iflag_solver", "iflag_time", "iflag_splittingtype", "iflag_inviscid",
                "iflag_filter", "iflag_timeadvance", "iflag_turbulence", "iflag_des", "writeinterval", "ntstep",
                "iflag_init", "nfile", "restart_timestep", "nobs", "iflag_acoustic", "Start_Acsoutic",
                "End_Acsoutic", "iflag_avg", "Start_Avg", "End_Avg", "Avg_interval")
CONTROL_REALS = ("cfl", "dt", "ma", "re", "tinf", "aoa", "scale_x", "scale_y", "scale_z", "lref", "sref", "fltr",
                 "ob_dist", "acoustic_length")


def read_control():
    """init_readcontrolfile on input/Control.in of the current directory;
    the values it set, by name."""
    iv = np.zeros(22, dtype=np.int32)
    rv = np.zeros(17)
    sv = C.create_string_buffer(540)
    lib().refh_read_control(iv.ctypes.data_as(C.POINTER(C.c_int)), ptr(rv), sv)
    out = {k: int(v) for k, v in zip(CONTROL_INTS, iv)}
    out.update({k: float(v) for k, v in zip(CONTROL_REALS, rv)})
    raw = sv.raw.decode("latin-1")
    for n, k in enumerate(("grid", "bcinfo", "restartfile")):
        out[k] = raw[180 * n:180 * (n + 1)].rstrip()
    return out
