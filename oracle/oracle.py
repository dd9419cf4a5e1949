"""ctypes wrapper of liboracle.so (TEST INFRASTRUCTURE ONLY).

Arrays follow the reference layout: a field with C components on a block of
extended extents (ni, nj, nk) is a C-contiguous float64 numpy array of shape
(nk, nj, ni, C) -- byte-identical to Fortran ``a(C, is:ie, js:je, ks:ke)``.
Scalar fields have shape (nk, nj, ni).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FIELDS = {  # name -> components
    "q": 5, "q0": 10, "pri_v": 7, "rhsi": 5, "rhsv": 5, "rhs": 5, "dudx": 9, "tao": 6,
    "vor": 1, "amu": 3, "dt": 1, "dxidx": 9, "inv_j": 1, "aalpha": 3, "bbeta": 3,
    "ggamma": 3, "dist": 1, "spacing": 1, "length": 1, "nut": 1, "sa_rhs": 1,
}


class Ext(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "is_", "ie", "js", "je", "ks", "ke", "is0", "ie0", "js0", "je0", "ks0", "ke0",
        "is1", "ie1", "js1", "je1", "ks1", "ke1")]

    @classmethod
    def from_list(cls, r):
        return cls(*[int(x) for x in r])

    def as_list(self):
        return [getattr(self, f) for f, _ in self._fields_]

    @property
    def shape(self):
        return (self.ke - self.ks + 1, self.je - self.js + 1, self.ie - self.is_ + 1)


class Subface(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "f_no", "face", "is_", "ie", "js", "je", "ks", "ke", "blk_t", "subface_t", "ori")]


_DP = C.POINTER(C.c_double)


class Block(C.Structure):
    _fields_ = [("e", Ext), ("nsub", C.c_int), ("sub", C.POINTER(Subface))] + [
        (n, _DP) for n in ("q", "q0", "pri_v", "rhsi", "rhsv", "rhs", "dudx", "tao", "vor", "amu",
                           "dt", "dxidx", "inv_j", "aalpha", "bbeta", "ggamma", "dist", "spacing",
                           "length", "nut", "sa_rhs")]


class Config(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "gamma", "cv", "re", "tinf", "prl", "prt", "d0", "u0", "p0", "c0", "aoa", "cfl")] + [
        (n, C.c_int) for n in ("iflag_solver", "iflag_time", "iflag_turbulence", "iflag_des",
                               "iflag_timeadvance")] + [
        ("totalnodes", C.c_long)] + [(n, C.c_double) for n in ("dtmax", "dtmin", "cflmax", "cflmin")]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        _LIB = C.CDLL(path)
        _LIB.ora_getpri.restype = C.c_long
        _LIB.ora_step.restype = C.c_int
        _LIB.ora_ix.restype = C.c_long
    return _LIB


def ptr(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def make_config(ma=0.5, re=1.0e6, tinf=288.15, cfl=0.5, aoa_deg=0.0, solver=1, itime=1, turb=0,
                des=0, tadv=2, totalnodes=1):
    """Non-dimensional constants of sub_init_flowfield.f90:30-39."""
    gamma = 1.4
    c = Config()
    c.gamma = gamma
    c.cv = 1.0 / gamma / (gamma - 1.0) / ma / ma
    c.re, c.tinf, c.prl, c.prt = re, tinf, 0.72, 0.90
    c.d0, c.u0 = 1.0, 1.0
    c.p0 = 1.0 * 1.0 / gamma / ma / ma
    c.c0 = np.sqrt(1.0) / ma
    c.aoa = aoa_deg * 3.1415926535897932 / 180.0
    c.cfl = cfl
    c.iflag_solver, c.iflag_time, c.iflag_turbulence, c.iflag_des = solver, itime, turb, des
    c.iflag_timeadvance = tadv
    c.totalnodes = totalnodes
    return c


class BlockState:
    """Owns the numpy arrays of one block and a ctypes Block view of them."""

    def __init__(self, ext, subfaces=()):
        self.ext = ext if isinstance(ext, Ext) else Ext.from_list(ext)
        nk, nj, ni = self.ext.shape
        self.arrays = {}
        for name, ncomp in FIELDS.items():
            shp = (nk, nj, ni) if ncomp == 1 else (nk, nj, ni, ncomp)
            self.arrays[name] = np.zeros(shp)
        self.subs = (Subface * max(1, len(subfaces)))()
        for i, s in enumerate(subfaces):
            self.subs[i] = Subface(*[int(x) for x in s])
        self.nsub = len(subfaces)

    def __getitem__(self, k):
        return self.arrays[k]

    def __setitem__(self, k, v):
        self.arrays[k][...] = v

    def cblock(self):
        b = Block()
        b.e = self.ext
        b.nsub = self.nsub
        b.sub = C.cast(self.subs, C.POINTER(Subface))
        for name in FIELDS:
            setattr(b, name, ptr(self.arrays[name]))
        return b


def blocks_array(states):
    arr = (Block * len(states))()
    for i, s in enumerate(states):
        arr[i] = s.cblock()
    return arr


# --------------------------------------------------------------- per routine
def getpri(ext, q, gamma, cv):
    pri = np.zeros(q.shape[:-1] + (7,))
    nneg = lib().ora_getpri(C.byref(ext), ptr(q), ptr(pri), C.c_double(gamma), C.c_double(cv))
    return pri, nneg


def weno3(ext, f, coord, direction):
    df = np.zeros_like(f)
    lib().ora_weno3(C.byref(ext), ptr(df), ptr(f), C.c_int(coord), C.c_int(direction))
    return df


def roe(ext, ql, qr, dxidx, inv_j, gamma, coord):
    flux = np.zeros_like(ql)
    lib().ora_roe(C.byref(ext), ptr(ql), ptr(qr), ptr(flux), ptr(dxidx), ptr(inv_j),
                  C.c_double(gamma), C.c_int(coord))
    return flux


def inviscid(ext, q, dxidx, inv_j, gamma):
    rhsi = np.zeros_like(q)
    lib().ora_inviscid(C.byref(ext), ptr(rhsi), ptr(q), ptr(dxidx), ptr(inv_j), C.c_double(gamma))
    return rhsi


def viscous(ext, pri_v, amu, dxidx, inv_j, cfg):
    """Returns rhsv, amu (updated copy), vor, dudx, tao."""
    shp = pri_v.shape[:-1]
    rhsv = np.zeros(shp + (5,))
    amu = amu.copy()
    vor = np.zeros(shp)
    dudx = np.zeros(shp + (9,))
    tao = np.zeros(shp + (6,))
    lib().ora_viscous(C.byref(ext), ptr(rhsv), ptr(pri_v), ptr(amu), ptr(vor), ptr(dudx), ptr(tao),
                      ptr(dxidx), ptr(inv_j), C.c_double(cfg.tinf), C.c_double(cfg.prl),
                      C.c_double(cfg.prt), C.c_double(cfg.gamma), C.c_double(cfg.cv),
                      C.c_double(cfg.re))
    return rhsv, amu, vor, dudx, tao


def find_dt(ext, dt, pri_v, amu, dxidx, re, cfl, iflag_time):
    dt = dt.copy()
    stats = np.zeros(4)
    c = C.c_double(cfl)
    lib().ora_find_dt(C.byref(ext), ptr(dt), ptr(stats), ptr(pri_v), ptr(amu), ptr(dxidx),
                      C.c_double(re), C.byref(c), C.c_int(iflag_time))
    return dt, c.value, stats


def calc_jm(qv, m, gamma):
    a = np.zeros(25)
    lib().ora_calc_jm(ptr(a), *[C.c_double(x) for x in qv], *[C.c_double(x) for x in m],
                      C.c_double(gamma))
    return a.reshape(5, 5).T.copy()  # row-major (r, c)


def implicit1st(ext, q, dt, pri_v, rhsi, rhsv, dxidx, inv_j, gamma, cfl):
    q = q.copy()
    rhs = np.zeros_like(q)
    lib().ora_implicit1st(C.byref(ext), ptr(q), ptr(dt), ptr(pri_v), ptr(rhs), ptr(rhsi), ptr(rhsv),
                          ptr(dxidx), ptr(inv_j), C.c_double(gamma), C.c_double(cfl))
    return q, rhs


def crank2nd(ext, q, q0, dt, pri_v, rhsi, rhsv, dxidx, inv_j, gamma):
    q = q.copy()
    rhs = np.zeros_like(q)
    lib().ora_crank2nd(C.byref(ext), ptr(q), ptr(q0), ptr(dt), ptr(pri_v), ptr(rhs), ptr(rhsi),
                       ptr(rhsv), ptr(dxidx), ptr(inv_j), C.c_double(gamma))
    return q, rhs


def sa_model(ext, nut, dt, pri_v, amu, dudx, vor, dxidx, inv_j, alpha, beta, gam, dist, spac, re,
             cfl, iflag_des, iflag_time):
    nut = nut.copy()
    sa_rhs = np.zeros_like(nut)
    length = np.zeros_like(nut)
    lib().ora_sa_model(C.byref(ext), ptr(nut), ptr(sa_rhs), ptr(dt), ptr(pri_v), ptr(amu), ptr(dudx),
                       ptr(vor), ptr(dxidx), ptr(inv_j), ptr(alpha), ptr(beta), ptr(gam),
                       ptr(length), ptr(dist), ptr(spac), C.c_double(re), C.c_double(cfl),
                       C.c_int(iflag_des), C.c_int(iflag_time))
    return nut, sa_rhs, length


def physical_bc(states, cfg):
    arr = blocks_array(states)
    lib().ora_physical_bc(C.c_int(len(states)), arr, C.byref(cfg))


def update_buffer(states, cfg):
    arr = blocks_array(states)
    lib().ora_update_buffer(C.c_int(len(states)), arr, C.byref(cfg))


def average(states, cfg):
    arr = blocks_array(states)
    lib().ora_average(C.c_int(len(states)), arr, C.byref(cfg))


def residual(states, totalnodes):
    arr = blocks_array(states)
    out = np.zeros(7)
    lib().ora_residual(C.c_int(len(states)), arr, C.c_long(totalnodes), ptr(out))
    return out


def prepare(states, cfg):
    arr = blocks_array(states)
    lib().ora_prepare(C.c_int(len(states)), arr, C.byref(cfg))


def step(states, cfg):
    arr = blocks_array(states)
    rc = lib().ora_step(C.c_int(len(states)), arr, C.byref(cfg))
    if rc != 0:
        raise RuntimeError("ora_step: unsupported time advance %d" % cfg.iflag_timeadvance)


# ------------------------------------------- extensions (no reference form)
def sw_split(qv, M, gamma):
    q = np.ascontiguousarray(qv, dtype=np.float64)
    m = np.ascontiguousarray(M, dtype=np.float64)
    fp = np.zeros(5)
    fm = np.zeros(5)
    lib().ora_sw_split(ptr(q), ptr(m), C.c_double(gamma), ptr(fp), ptr(fm))
    return fp, fm


def weno5_rec(a, b, c, d, e):
    f = lib().ora_weno5_rec
    f.restype = C.c_double
    return f(*[C.c_double(x) for x in (a, b, c, d, e)])


def weno5_interp(a, b, c, d, e):
    f = lib().ora_weno5_interp
    f.restype = C.c_double
    return f(*[C.c_double(x) for x in (a, b, c, d, e)])


def inviscid_ext(ext, q, dxidx, inv_j, gamma, scheme):
    rhsi = np.zeros_like(q)
    lib().ora_inviscid_ext(C.byref(ext), ptr(rhsi), ptr(q), ptr(dxidx), ptr(inv_j), C.c_double(gamma),
                           C.c_int(scheme))
    return rhsi


def step_rk3(states, cfg, scheme):
    arr = blocks_array(states)
    lib().ora_step_rk3(C.c_int(len(states)), arr, C.byref(cfg), C.c_int(scheme))


# ------------------------------------------------------ grid pre-processing
def gridtransformation(ext, x, y, z):
    """sub_init_gridtransformation.f90:27-874 (3-D).  Outputs are NaN at the
    edge and corner halo nodes, which the reference leaves to uninitialised
    temporaries (the nodes with two or three coordinates outside the
    physical range)."""
    nk, nj, ni = ext.shape
    dxidx = np.full((nk, nj, ni, 9), np.nan)
    inv_j = np.full((nk, nj, ni), np.nan)
    al, be, ga = (np.full((nk, nj, ni, 3), np.nan) for _ in range(3))
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    lib().ora_gridtransformation(C.byref(ext), ptr(x), ptr(y), ptr(z), ptr(dxidx), ptr(inv_j), ptr(al),
                                 ptr(be), ptr(ga))
    return dxidx, inv_j, al, be, ga


def spacing(ext, x, y, z):
    """sub_init_spacing.f90:16-110 -> spci, spcj, spck, spacing, vol (NaN
    outside the physical range)."""
    nk, nj, ni = ext.shape
    out = [np.full((nk, nj, ni), np.nan) for _ in range(5)]
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    lib().ora_spacing(C.byref(ext), ptr(x), ptr(y), ptr(z), *[ptr(a) for a in out])
    return out


def walldist(ext, x, y, z, wall_xyz):
    """sub_init_walldist.f90:30-85: distance of the physical nodes to the
    nearest wall point (wall_xyz: (3, nw)); NaN on halos."""
    nk, nj, ni = ext.shape
    dist = np.full((nk, nj, ni), np.nan)
    w = np.ascontiguousarray(wall_xyz, dtype=np.float64).reshape(3, -1)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    lib().ora_walldist(C.byref(ext), ptr(x), ptr(y), ptr(z), C.c_long(w.shape[1]), ptr(w[0]), ptr(w[1]),
                       ptr(w[2]), ptr(dist))
    return dist


def wall_points(exts, subfaces, xyz):
    """Face-node coordinates of every wall subface (blk_t = -20) of every
    block, in block and subface order (sub_init_readbcinfo.f90:294-400);
    xyz[m] = (x, y, z) of block m.  Returns (3, nw)."""
    pts = []
    for e, subs, (x, y, z) in zip(exts, subfaces, xyz):
        for s in subs:
            if s[8] != -20:
                continue
            f = s[1]
            i0, i1, j0, j1, k0, k1 = s[2:8]
            if f == 1:
                i0 = i1 = e.is0
            elif f == 2:
                i0 = i1 = e.ie0
            elif f == 3:
                j0 = j1 = e.js0
            elif f == 4:
                j0 = j1 = e.je0
            elif f == 5:
                k0 = k1 = e.ks0
            else:
                k0 = k1 = e.ke0
            sl = (slice(k0 - e.ks, k1 - e.ks + 1), slice(j0 - e.js, j1 - e.js + 1), slice(i0 - e.is_, i1 - e.is_ + 1))
            pts.append(np.stack([x[sl].ravel(), y[sl].ravel(), z[sl].ravel()]))
    return np.concatenate(pts, axis=1) if pts else np.zeros((3, 0))


def forces(exts, subfaces, fields, pinf, re, aoa_deg=0.0, sref=1.0):
    """clcd_output (sub_output_forceintegration.f90:8-478, 3-D): fields[m]
    holds pri_v, dxidx, amu, spci, spcj, spck of block m.  Returns fp, fv,
    the term magnitudes afp, afv, and cl, cd (div = 1/(0.5 sref) with the
    unit freestream)."""
    fp, fv, afp, afv = (np.zeros(3) for _ in range(4))
    for e, subs, f in zip(exts, subfaces, fields):
        arr = (Subface * max(1, len(subs)))()
        for i, s in enumerate(subs):
            arr[i] = Subface(*[int(v) for v in s])
        a = {k: np.ascontiguousarray(f[k], dtype=np.float64) for k in ("pri_v", "dxidx", "amu", "spci", "spcj",
                                                                        "spck")}
        lib().ora_forces_block(C.byref(e), C.c_int(len(subs)), C.cast(arr, C.POINTER(Subface)), ptr(a["pri_v"]),
                               ptr(a["dxidx"]), ptr(a["amu"]), ptr(a["spci"]), ptr(a["spcj"]), ptr(a["spck"]),
                               C.c_double(pinf), C.c_double(re), ptr(fp), ptr(fv), ptr(afp), ptr(afv))
    aoa = np.deg2rad(aoa_deg)
    div = 1.0 / (0.5 * 1.0 * 1.0 * 1.0 * sref)
    cl = ((fp[1] + fv[1]) * np.cos(aoa) - (fp[0] + fv[0]) * np.sin(aoa)) * div
    cd = ((fp[1] + fv[1]) * np.sin(aoa) + (fp[0] + fv[0]) * np.cos(aoa)) * div
    return fp, fv, afp, afv, cl, cd
