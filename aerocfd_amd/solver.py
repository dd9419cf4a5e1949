"""Host-side mirror of the reference's interface for the residual-and-update path.

Two layers, both thin wrappers of the C ABI in include/aerocfd_amd.h:

* :class:`Solver` -- the resident multi-block context standing in for the
  reference time loop (main.f90:342-613): topology once, fields uploaded once,
  then ``prepare()`` and ``step(n)`` on the GPU, ``residual()`` each step.
* stateless functions named after the reference routines
  (``get_primitivevariables``, ``inviscid_flux_1``, ``viscous_flux``,
  ``find_dt``, ``implicit1st``, ``crank2nd``, ``sa_model``) that run one routine
  on one block given its 18 index ranges.

Arrays use the reference layout: a field with C components on a block with
extended extents (ni, nj, nk) is a C-contiguous float64 array of shape
(nk, nj, ni, C) (scalar fields: (nk, nj, ni)), byte-identical to the Fortran
array ``a(C, is:ie, js:je, ks:ke)``.
"""
import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib

FIELDS = {  # name -> (acfd_field id, components)
    "q": (1, 5), "pri_v": (2, 7), "rhsi": (3, 5), "rhsv": (4, 5), "rhs": (5, 5), "dudx": (6, 9),
    "tao": (7, 6), "vor": (8, 1), "amu": (9, 3), "dt": (10, 1), "dxidx": (11, 9), "inv_j": (12, 1),
    "aalpha": (13, 3), "bbeta": (14, 3), "ggamma": (15, 3), "dist": (16, 1), "spacing": (17, 1),
    "nut": (18, 1), "sa_rhs": (19, 1), "q0": (20, 10), "length": (21, 1),
    # after grid_metrics
    "x": (22, 1), "y": (23, 1), "z": (24, 1), "spci": (25, 1), "spcj": (26, 1), "spck": (27, 1),
}

STAGES = {
    "getpri": 1, "inviscid": 2, "viscous": 3, "find_dt": 4, "implicit1st": 5, "sa": 6,
    "physical_bc": 7, "update_buffer": 8, "average": 9, "crank2nd": 10, "update_q0": 11,
}

# sub-stage timers inside implicit1st: the two Gauss-Seidel sweeps and the
# rest (spectral radii, diagonal, update)
# sub-stage timers of the LU-SGS update: the two sweeps, the gather into the
# sweep layout (tile sweeps), rhs / spectral radii / diag, and q += dq
TIMED_KERNELS = {"lusgs_forward": 12, "lusgs_backward": 13, "lusgs_gather": 14, "lusgs_other": 15,
                 "lusgs_add": 16,
                 # one kernel each, recorded for the batched stages
                 "sa_forward": 17, "sa_backward": 18, "sa_pre": 19, "sa_add": 20, "inviscid_zero": 21,
                 "inviscid_i": 22, "inviscid_j": 23, "inviscid_k": 24, "viscous_node": 25, "viscous_diff": 26,
                 "sa_gather": 27, "sa_pre_b": 28, "viscous_node_dt": 29}
TIME_SLOTS = 32

GAMMA = 1.4


@dataclass
class FlowConfig:
    """Control parameters (sub_init_readcontrolfile.f90:27-47, module_flags.f90)."""
    ma: float = 0.5
    re: float = 1.0e6
    tinf: float = 288.15
    aoa_deg: float = 0.0
    cfl: float = 0.5
    dt: float = 0.01
    prl: float = 0.72
    prt: float = 0.90
    iflag_solver: int = 1        # 0 NS, 1 Euler
    iflag_time: int = 1          # 0 steady, 1 unsteady
    iflag_turbulence: int = 0    # 0 laminar, 2 SA
    iflag_des: int = 0
    iflag_timeadvance: int = 2   # 2 LU-SGS, 3 BDF2, 4 explicit RK3 (no reference counterpart)
    precision: int = 0           # 0 FP64, 1 FP32 flux / FP64 update
    lusgs_kernel: int = 0        # 0 hyperplane launches, 1 pipelined tiles
    inviscid_scheme: int = 0     # 0 WENO3 + Roe (reference); 1 Steger-Warming + WENO5, 2 WCNS
    use_graph: int = 0           # 1 replay one captured HIP graph per step
    batch_blocks: int = 0        # 1 one launch per stage over all blocks

    # non-dimensional freestream, sub_init_flowfield.f90:30-39
    @property
    def cv(self):
        return 1.0 / GAMMA / (GAMMA - 1.0) / self.ma / self.ma

    @property
    def p0(self):
        return 1.0 / GAMMA / self.ma / self.ma

    @property
    def c0(self):
        return 1.0 / self.ma

    def to_c(self):
        c = _lib.Config()
        for f, _ in _lib.Config._fields_:
            if f != "reserved":
                setattr(c, f, getattr(self, f))
        return c


def field_shape(ranges, ncomp):
    r = list(ranges)
    ni, nj, nk = r[1] - r[0] + 1, r[3] - r[2] + 1, r[5] - r[4] + 1
    return (nk, nj, ni) if ncomp == 1 else (nk, nj, ni, ncomp)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _arr(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != tuple(shape):
        raise ValueError("array shape %s, expected %s" % (a.shape, tuple(shape)))
    return a


def _ranges(r):
    r = np.ascontiguousarray(np.asarray(r, dtype=np.int32).reshape(18))
    return r, r.ctypes.data_as(C.POINTER(C.c_int))


class Solver:
    """Resident multi-block state on one GPU (one context, one HIP stream)."""

    def __init__(self, cfg, dims, subfaces):
        """dims: list of (ni, nj, nk); subfaces: per block, list of 11-int
        records (f_no, face, is, ie, js, je, ks, ke, blk_t, subface_t, ori)."""
        self.lib = _lib.load()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = C.c_void_p()
        _lib.check(self.lib.acfd_create(C.byref(self._c), C.byref(h)))
        self.h = h
        nb = len(dims)
        d = np.ascontiguousarray(np.array(dims, dtype=np.int32).reshape(nb, 3))
        ns = np.array([len(s) for s in subfaces], dtype=np.int32)
        recs = [r for s in subfaces for r in s]
        tab = (_lib.Subface * max(1, len(recs)))()
        for x, r in enumerate(recs):
            tab[x] = _lib.Subface(*[int(v) for v in r])
        try:
            _lib.check(self.lib.acfd_set_topology(self.h, C.c_int(nb), d.ctypes.data_as(C.POINTER(C.c_int)),
                                                  ns.ctypes.data_as(C.POINTER(C.c_int)), tab))
        except Exception:
            self.close()
            raise
        self.nblocks = nb
        self.dims = [tuple(x) for x in dims]
        self.subfaces = [list(s) for s in subfaces]
        self._ranges = [self._get_ranges(m + 1) for m in range(nb)]
        self.totalnodes = int(sum(a * b * c for a, b, c in self.dims))

    def _get_ranges(self, m):
        r = np.zeros(18, dtype=np.int32)
        _lib.check(self.lib.acfd_block_ranges(self.h, C.c_int(m), r.ctypes.data_as(C.POINTER(C.c_int))))
        return tuple(int(x) for x in r)

    def ranges(self, m):
        """is,ie,js,je,ks,ke, is0..ke0, is1..ke1 of block m (1-based)."""
        return self._ranges[m - 1]

    def shape(self, m, name):
        return field_shape(self.ranges(m), FIELDS[name][1])

    def upload(self, m, name, arr):
        fid, nc = FIELDS[name]
        a = _arr(arr, self.shape(m, name))
        _lib.check(self.lib.acfd_upload(self.h, C.c_int(m), C.c_int(fid), _dp(a)))

    def download(self, m, name):
        fid, nc = FIELDS[name]
        a = np.zeros(self.shape(m, name))
        _lib.check(self.lib.acfd_download(self.h, C.c_int(m), C.c_int(fid), _dp(a)))
        return a

    def run_stage(self, name):
        _lib.check(self.lib.acfd_run_stage(self.h, C.c_int(STAGES[name])))

    def prepare(self):
        _lib.check(self.lib.acfd_prepare(self.h))

    def step(self, n=1):
        _lib.check(self.lib.acfd_step(self.h, C.c_int(n)))

    def residual(self):
        out = np.zeros(7)
        _lib.check(self.lib.acfd_residual(self.h, _dp(out)))
        return out

    def stats(self):
        out = np.zeros(5)
        _lib.check(self.lib.acfd_stats(self.h, _dp(out)))
        return out

    def sync(self):
        _lib.check(self.lib.acfd_sync(self.h))

    def timing(self, enable=None):
        """Per-stage HIP-event totals {stage: (ms, calls)}; enable=True/False
        restarts recording, None only reads."""
        ms = np.zeros(TIME_SLOTS)
        calls = np.zeros(TIME_SLOTS, dtype=np.int64)
        flag = -1 if enable is None else int(bool(enable))
        _lib.check(self.lib.acfd_timing(self.h, C.c_int(flag), _dp(ms),
                                        calls.ctypes.data_as(C.POINTER(C.c_long))))
        names = dict(STAGES)
        names.update(TIMED_KERNELS)
        return {name: (float(ms[s]), int(calls[s])) for name, s in names.items()}

    # ------------------------------------------------ learned WCNS sensor
    def sensor_weights(self, w=None):
        """Return the 1505 FP32 parameters, or set them when w is given."""
        if w is None:
            out = np.zeros(SENSOR_PARAMS, dtype=np.float32)
            _lib.check(self.lib.acfd_sensor_weights(self.h, out.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(0)))
            return out
        a = np.ascontiguousarray(w, dtype=np.float32)
        _lib.check(self.lib.acfd_sensor_weights(self.h, a.ctypes.data_as(C.POINTER(C.c_float)), C.c_int(1)))

    def sensor_collect(self, stride=1):
        n = C.c_long(0)
        _lib.check(self.lib.acfd_sensor_collect(self.h, C.c_int(stride), C.byref(n)))
        return n.value

    def sensor_train(self, iters, batch=4096, lr=1e-3, seed=0):
        losses = np.zeros(max(iters, 1), dtype=np.float32)
        _lib.check(self.lib.acfd_sensor_train(self.h, C.c_int(iters), C.c_int(batch), C.c_float(lr),
                                              C.c_uint(seed), losses.ctypes.data_as(C.POINTER(C.c_float))))
        return losses[:iters]

    def sensor_evaluate(self, threshold=0.5):
        out = np.zeros(4)
        _lib.check(self.lib.acfd_sensor_evaluate(self.h, C.c_float(threshold), _dp(out)))
        return dict(zip(("accuracy", "linear", "false_linear", "labelled_nonlinear"), out))

    def sensor_enable(self, on=True, threshold=0.5, audit_interval=0, audit_bound=1e-6):
        _lib.check(self.lib.acfd_sensor_enable(self.h, C.c_int(int(on)), C.c_float(threshold),
                                               C.c_int(audit_interval), C.c_double(audit_bound)))

    def sensor_refresh(self, interval=1):
        """Run the network on every interval-th WCNS evaluation of a block
        and reuse its recorded per-face decisions in between."""
        _lib.check(self.lib.acfd_sensor_refresh(self.h, C.c_int(interval)))

    def sensor_stats(self):
        out = np.zeros(6)
        _lib.check(self.lib.acfd_sensor_stats(self.h, _dp(out)))
        return dict(zip(("faces", "linear", "audits", "audit_max_dev", "blocks_fallen_back"), out[:5]))

    def grid_metrics(self, xyz):
        """Grid pre-processing on the device (init_spacing, init_walldist with
        SA, gridtransformation): xyz = list per block of (x, y, z) arrays
        over the extended extents, halos filled.  Writes the metric inputs
        of the resident blocks."""
        ptrs = (C.POINTER(C.c_double) * self.nblocks)()
        keep = []
        for m in range(self.nblocks):
            a = np.ascontiguousarray(np.stack([np.asarray(c, dtype=np.float64) for c in xyz[m]]))
            assert a.shape == (3,) + self.shape(m + 1, "x")
            keep.append(a)
            ptrs[m] = _dp(a)
        _lib.check(self.lib.acfd_grid_metrics(self.h, ptrs))

    def forces(self, sref=1.0):
        """Wall forces of the resident state (clcd_output): dict with the
        pressure force fp, viscous force fv (3 each), cl and cd."""
        out = np.zeros(8)
        _lib.check(self.lib.acfd_forces(self.h, C.c_double(sref), _dp(out)))
        return {"fp": out[:3], "fv": out[3:6], "cl": out[6], "cd": out[7]}

    def exchange_coordinates(self, xyz, shifts=None):
        """Fill the coordinate halos of every block (UpdateBufferCoordinate).
        xyz: list per block of array (3, nk, nj, ni), updated in place;
        shifts: optional per-subface (global order) translation (nsub, 3)."""
        ptrs = (C.POINTER(C.c_double) * self.nblocks)()
        keep = []
        for m in range(self.nblocks):
            a = xyz[m]
            assert a.dtype == np.float64 and a.flags.c_contiguous
            keep.append(a)
            ptrs[m] = _dp(a)
        sh = None
        if shifts is not None:
            sh = np.ascontiguousarray(np.asarray(shifts, dtype=np.float64).reshape(-1, 3))
            keep.append(sh)
        _lib.check(self.lib.acfd_exchange_coordinates(self.h, ptrs, _dp(sh) if sh is not None else None))

    def close(self):
        if getattr(self, "h", None):
            self.lib.acfd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------- stateless routines
def get_primitivevariables(ranges, q, gamma, cv):
    """sub_getpri_v.f90:19-63. Returns pri_v; raises on negative density."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    q = _arr(q, field_shape(r, 5))
    pri = np.zeros(field_shape(r, 7))
    _lib.check(lib.acfd_getpri(rp, C.c_double(gamma), C.c_double(cv), _dp(q), _dp(pri)))
    return pri


def inviscid_flux_1(cfg, ranges, q, dxidx, inv_j):
    """sub_inviscid1.f90:8-189. Returns rhsi."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    out = np.zeros(field_shape(r, 5))
    c = cfg.to_c()
    _lib.check(lib.acfd_inviscid_rhs(C.byref(c), rp, _dp(_arr(q, field_shape(r, 5))),
                                     _dp(_arr(dxidx, field_shape(r, 9))), _dp(_arr(inv_j, field_shape(r, 1))),
                                     _dp(out)))
    return out


def viscous_flux(cfg, ranges, pri_v, amu, dxidx, inv_j):
    """sub_viscous.f90:9-378. Returns rhsv, amu, vor, dudx, tao."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    amu = _arr(amu, field_shape(r, 3)).copy()
    rhsv = np.zeros(field_shape(r, 5))
    vor = np.zeros(field_shape(r, 1))
    dudx = np.zeros(field_shape(r, 9))
    tao = np.zeros(field_shape(r, 6))
    c = cfg.to_c()
    _lib.check(lib.acfd_viscous_rhs(C.byref(c), rp, _dp(_arr(pri_v, field_shape(r, 7))), _dp(amu),
                                    _dp(_arr(dxidx, field_shape(r, 9))), _dp(_arr(inv_j, field_shape(r, 1))),
                                    _dp(rhsv), _dp(vor), _dp(dudx), _dp(tao)))
    return rhsv, amu, vor, dudx, tao


def find_dt(cfg, ranges, pri_v, amu, dxidx, dt, cfl):
    """sub_timeadvancement_finddt.f90:24-130. Returns dt, cfl."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    dt = _arr(dt, field_shape(r, 1)).copy()
    cc = C.c_double(cfl)
    c = cfg.to_c()
    _lib.check(lib.acfd_find_dt(C.byref(c), rp, _dp(_arr(pri_v, field_shape(r, 7))),
                                _dp(_arr(amu, field_shape(r, 3))), _dp(_arr(dxidx, field_shape(r, 9))),
                                _dp(dt), C.byref(cc)))
    return dt, cc.value


def implicit1st(cfg, ranges, q, dt, pri_v, rhsi, rhsv, dxidx, inv_j, cfl):
    """sub_timeadvancement_1stimplicit.f90:7-585. Returns q (updated), rhs."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    q = _arr(q, field_shape(r, 5)).copy()
    rhs = np.zeros(field_shape(r, 5))
    c = cfg.to_c()
    _lib.check(lib.acfd_lusgs(C.byref(c), rp, _dp(q), _dp(_arr(dt, field_shape(r, 1))),
                              _dp(_arr(pri_v, field_shape(r, 7))), _dp(_arr(rhsi, field_shape(r, 5))),
                              _dp(_arr(rhsv, field_shape(r, 5))), _dp(_arr(dxidx, field_shape(r, 9))),
                              _dp(_arr(inv_j, field_shape(r, 1))), C.c_double(cfl), _dp(rhs)))
    return q, rhs


def crank2nd(cfg, ranges, q, q0, dt, pri_v, rhsi, rhsv, dxidx, inv_j):
    """sub_timeadvancement_2ndcrank.f90:7-467. Returns q (updated), rhs."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    q = _arr(q, field_shape(r, 5)).copy()
    rhs = np.zeros(field_shape(r, 5))
    c = cfg.to_c()
    _lib.check(lib.acfd_bdf2(C.byref(c), rp, _dp(q), _dp(_arr(q0, field_shape(r, 10))),
                             _dp(_arr(dt, field_shape(r, 1))), _dp(_arr(pri_v, field_shape(r, 7))),
                             _dp(_arr(rhsi, field_shape(r, 5))), _dp(_arr(rhsv, field_shape(r, 5))),
                             _dp(_arr(dxidx, field_shape(r, 9))), _dp(_arr(inv_j, field_shape(r, 1))),
                             _dp(rhs)))
    return q, rhs


def sa_model(cfg, ranges, nut, dt, pri_v, amu, dudx, vor, dxidx, inv_j, aalpha, bbeta, ggamma, dist,
             spacing, cfl):
    """sub_tur_sa.f90:7-122. Returns nut (updated), sa_rhs, length."""
    lib = _lib.load()
    r, rp = _ranges(ranges)
    s1 = field_shape(r, 1)
    nut = _arr(nut, s1).copy()
    sa_rhs = np.zeros(s1)
    length = np.zeros(s1)
    c = cfg.to_c()
    _lib.check(lib.acfd_sa_step(C.byref(c), rp, _dp(nut), _dp(sa_rhs), _dp(_arr(dt, s1)),
                                _dp(_arr(pri_v, field_shape(r, 7))), _dp(_arr(amu, field_shape(r, 3))),
                                _dp(_arr(dudx, field_shape(r, 9))), _dp(_arr(vor, s1)),
                                _dp(_arr(dxidx, field_shape(r, 9))), _dp(_arr(inv_j, s1)),
                                _dp(_arr(aalpha, field_shape(r, 3))), _dp(_arr(bbeta, field_shape(r, 3))),
                                _dp(_arr(ggamma, field_shape(r, 3))), _dp(length), _dp(_arr(dist, s1)),
                                _dp(_arr(spacing, s1)), C.c_double(cfl)))
    return nut, sa_rhs, length


SENSOR_PARAMS = 1505


def mlp_grad(w, X, Y):
    """Mean loss and gradient of the sensor network over samples X[n][12],
    Y[n] computed by the GPU training kernels."""
    lib = _lib.load()
    w = np.ascontiguousarray(w, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.ascontiguousarray(Y, dtype=np.float32)
    loss = C.c_float(0)
    grad = np.zeros(SENSOR_PARAMS, dtype=np.float32)
    fp = C.POINTER(C.c_float)
    _lib.check(lib.acfd_mlp_grad(w.ctypes.data_as(fp), X.ctypes.data_as(fp), Y.ctypes.data_as(fp),
                                 C.c_long(X.shape[0]), C.byref(loss), grad.ctypes.data_as(fp)))
    return loss.value, grad


def sensor_infer(w, X):
    """Probabilities of the sensor network through its BF16 MFMA routine."""
    lib = _lib.load()
    w = np.ascontiguousarray(w, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    p = np.zeros(X.shape[0], dtype=np.float32)
    fp = C.POINTER(C.c_float)
    _lib.check(lib.acfd_sensor_infer(w.ctypes.data_as(fp), X.ctypes.data_as(fp), C.c_long(X.shape[0]),
                                     p.ctypes.data_as(fp)))
    return p


def mcell_updates_per_s(totalnodes, steps, seconds):
    """Throughput as the reference counts nodes (sub_init_allocatememory.f90:260-263)."""
    return totalnodes * steps / seconds / 1.0e6 if seconds > 0 else math.inf
