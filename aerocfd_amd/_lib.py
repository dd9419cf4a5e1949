"""ctypes binding of libaerocfd_amd.so (the C ABI declared in include/aerocfd_amd.h).

The library is built in-tree by :mod:`aerocfd_amd.build`.  Loading fails with a
RuntimeError when it is missing: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaerocfd_amd.so")

_LIB = None

# Exported entry points (kept in step with include/aerocfd_amd.h).
EXPORTS = (
    "acfd_last_error", "acfd_create", "acfd_destroy", "acfd_set_topology", "acfd_block_ranges",
    "acfd_upload", "acfd_download", "acfd_run_stage", "acfd_prepare", "acfd_step",
    "acfd_residual", "acfd_stats", "acfd_sync", "acfd_timing", "acfd_getpri",
    "acfd_inviscid_rhs", "acfd_viscous_rhs", "acfd_find_dt", "acfd_lusgs", "acfd_bdf2",
    "acfd_sa_step", "acfd_exchange_coordinates", "acfd_sensor_weights", "acfd_sensor_collect",
    "acfd_sensor_train", "acfd_sensor_evaluate", "acfd_sensor_enable", "acfd_sensor_refresh",
    "acfd_sensor_stats", "acfd_mlp_grad", "acfd_sensor_infer", "acfd_grid_metrics", "acfd_forces",
    "acfd_viscous_rhs_noalloc",
)


class Config(C.Structure):
    """acfd_config"""
    _fields_ = [(n, C.c_double) for n in ("ma", "re", "tinf", "aoa_deg", "cfl", "dt", "prl", "prt")] + [
        (n, C.c_int) for n in ("iflag_solver", "iflag_time", "iflag_turbulence", "iflag_des",
                               "iflag_timeadvance", "precision", "lusgs_kernel", "inviscid_scheme",
                               "use_graph", "batch_blocks")] + [
        ("reserved", C.c_int * 2)]


class Subface(C.Structure):
    """acfd_subface"""
    _fields_ = [(n, C.c_int) for n in ("f_no", "face", "is_", "ie", "js", "je", "ks", "ke",
                                       "blk_t", "subface_t", "ori")]


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "aerocfd_amd: %s is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950)" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    lib.acfd_last_error.restype = C.c_char_p
    for name in EXPORTS:
        if name != "acfd_last_error":
            getattr(lib, name).restype = C.c_int
    _LIB = lib
    return lib


class AcfdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("aerocfd_amd error %d: %s" % (code, msg))
        self.code = code


def check(rc):
    if rc != 0:
        raise AcfdError(rc, load().acfd_last_error().decode())
    return rc
