"""The Fortran side of the boundary: include/aerocfd_amd.f90 (the
ISO_C_BINDING module) and the Fortran host aerocfd_amd/fortran/acfd_host.f90
compile with amdflang and link against libaerocfd_amd.so, and the module's
bind(C) types have the byte layout of acfd_config / acfd_subface in
include/aerocfd_amd.h field by field (the C compiler's offsetof, the
Fortran compiler's c_loc differences and the ctypes binding all agree).
No GPU is used."""
import ctypes
import os
import shutil
import subprocess

import pytest

from aerocfd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/bin/amdflang"
MODULE = os.path.join(ROOT, "include", "aerocfd_amd.f90")
HOST = os.path.join(ROOT, "aerocfd_amd", "fortran", "acfd_host.f90")

CFG_FIELDS = [n for n, _ in _lib.Config._fields_]
SUB_FIELDS = [n for n, _ in _lib.Subface._fields_]

pytestmark = pytest.mark.skipif(not os.path.exists(FC), reason="amdflang not installed")


def c_layout(tmp_path):
    src = ['#include <stdio.h>', '#include <stddef.h>', '#include "aerocfd_amd.h"', "int main(void) {",
           '  printf("%zu\\n", sizeof(acfd_config));']
    for n in CFG_FIELDS:
        src.append('  printf("%%zu\\n", offsetof(acfd_config, %s));' % n)
    src.append('  printf("%zu\\n", sizeof(acfd_subface));')
    for n in SUB_FIELDS:
        src.append('  printf("%%zu\\n", offsetof(acfd_subface, %s));' % ("is" if n == "is_" else n))
    src += ["  return 0;", "}"]
    (tmp_path / "lay.c").write_text("\n".join(src) + "\n")
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(tmp_path / "lay.c"), "-o",
                    str(tmp_path / "lay_c")], check=True)
    out = subprocess.run([str(tmp_path / "lay_c")], capture_output=True, text=True, check=True).stdout
    return [int(v) for v in out.split()]


def fortran_layout(tmp_path):
    lines = ["program lay", "  use iso_c_binding", "  use aerocfd_amd", "  implicit none",
             "  type(acfd_config), target :: c", "  type(acfd_subface), target :: s",
             "  integer(c_intptr_t) :: b, d", "  b = transfer(c_loc(c), b)", "  print *, c_sizeof(c)"]
    for n in CFG_FIELDS:
        lines.append("  d = transfer(c_loc(c%%%s), d); print *, d - b" % n)
    lines += ["  b = transfer(c_loc(s), b)", "  print *, c_sizeof(s)"]
    for n in SUB_FIELDS:
        lines.append("  d = transfer(c_loc(s%%%s), d); print *, d - b" % ("is" if n == "is_" else n))
    lines.append("end program lay")
    (tmp_path / "lay.f90").write_text("\n".join(lines) + "\n")
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([FC, "-J", str(tmp_path), MODULE, str(tmp_path / "lay.f90"), "-o", str(tmp_path / "lay_f"),
                    "-L" + libdir, "-laerocfd_amd", "-Wl,-rpath," + libdir],
                   check=True, cwd=tmp_path, capture_output=True)
    out = subprocess.run([str(tmp_path / "lay_f")], capture_output=True, text=True, check=True).stdout
    return [int(v) for v in out.split()]


def ctypes_layout():
    out = [ctypes.sizeof(_lib.Config)] + [getattr(_lib.Config, n).offset for n in CFG_FIELDS]
    out += [ctypes.sizeof(_lib.Subface)] + [getattr(_lib.Subface, n).offset for n in SUB_FIELDS]
    return out


def test_bind_c_types_match_the_header(tmp_path):
    c = c_layout(tmp_path)
    f = fortran_layout(tmp_path)
    p = ctypes_layout()
    assert c == p, (c, p)
    assert f == c, (f, c)


def test_module_declares_every_resident_entry_point():
    """Every entry point of the resident context and the grid / force
    calls is bound in the Fortran module under its C name."""
    src = open(MODULE).read()
    for name in ("acfd_last_error", "acfd_create", "acfd_destroy", "acfd_set_topology", "acfd_block_ranges",
                 "acfd_upload", "acfd_download", "acfd_run_stage", "acfd_prepare", "acfd_step", "acfd_residual",
                 "acfd_stats", "acfd_sync", "acfd_exchange_coordinates", "acfd_grid_metrics", "acfd_forces",
                 "acfd_getpri", "acfd_inviscid_rhs", "acfd_lusgs"):
        assert 'bind(C, name="%s")' % name in src, name
        assert name in _lib.EXPORTS


def test_host_compiles_and_links(tmp_path):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libaerocfd_amd.so is not built")
    for f in (MODULE, HOST):
        shutil.copy(f, tmp_path)
    subprocess.run([FC, "-O2", "-J", str(tmp_path), "-c", "aerocfd_amd.f90", "-o", "m.o"], cwd=tmp_path, check=True,
                   capture_output=True)
    subprocess.run([FC, "-O2", "-J", str(tmp_path), "-c", "acfd_host.f90", "-o", "h.o"], cwd=tmp_path, check=True,
                   capture_output=True)
    libdir = os.path.dirname(_lib.LIB_PATH)
    r = subprocess.run([FC, "-o", "acfd_host", "m.o", "h.o", "-L" + libdir, "-laerocfd_amd",
                        "-Wl,-rpath," + libdir], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    # every acfd_ symbol the host needs resolves in the library
    nm = subprocess.run(["nm", "-u", str(tmp_path / "acfd_host")], capture_output=True, text=True).stdout
    need = sorted({ln.split()[-1] for ln in nm.splitlines() if ln.split() and ln.split()[-1].startswith("acfd_")})
    assert "acfd_step" in need and "acfd_grid_metrics" in need
    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert all(hasattr(lib, n) for n in need), need
