"""Build libaerocfd_amd.so in-tree with hipcc for gfx950.

    python -m aerocfd_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["k_flux.hip", "k_flux5.hip", "k_sensor.hip", "k_lusgs.hip", "k_bdf2.hip", "k_sa.hip", "k_bc.hip",
           "k_grid.hip", "context.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
    # no packed FP32 instructions in any kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    out = os.path.join(HERE, "libaerocfd_amd.so")
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-I" + os.path.join(ROOT, "include")] + \
        [os.path.join(HERE, "csrc", s) for s in SOURCES] + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines()
            if l.strip() and "packed-fp32-ops" not in l]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, out)
    return out


def build_fortran_host():
    """The Fortran host (aerocfd_amd/fortran/acfd_host): the ISO_C_BINDING
    module include/aerocfd_amd.f90 and acfd_host.f90, linked with amdflang
    against the library built above."""
    d = os.path.join(HERE, "fortran")
    r = subprocess.run(["make", "-B", "-C", d], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("Fortran host build failed:\n" + r.stdout[-3000:] + r.stderr[-3000:])
    return os.path.join(d, "acfd_host")


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran_host())
    sys.exit(0)
