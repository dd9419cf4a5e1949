"""Runs the reference's whole program under MPI (TEST INFRASTRUCTURE ONLY).

``_ref/aerocfd_ref.x`` is jiamin0329/aerocfd's own program (``main.f90``
and every other source file) compiled in place by ``oracle/Makefile``
(target ``ref``) with amdflang -O2 and MPICH 3.3.2.  This module writes a
case in the reference's own file formats with :mod:`aerocfd_amd.formats`
(``input/Control.in``, the grid, ``bcinfo`` and, for a restart start, one
``result/flowfield_0000000000_<rank>.dat`` per rank), runs the program with
``mpiexec -n ranks`` in a scratch directory, and reads back what it wrote:

  * the per-step wall time it prints (``main.f90:575-586``: the max and min
    over ranks of each step's MPI_WTIME difference, format f6.3);
  * ``result/resi_history.dat`` (``sub_output_residual.f90:62-64``);
  * ``result/flowfield_<ntstep>_<rank>.dat`` (``sub_output_flowfield.f90:
    30-81``), the physical range of every block after the last step.

Blocks go to ranks as the reference distributes them (``main.f90:99-133``:
myn = ceil(nblock/numprocs) blocks per rank, the remainder on the last).
"""
import os
import re
import resource
import shutil
import subprocess
import tempfile
import time

import numpy as np

from aerocfd_amd import formats

_HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(_HERE, "_ref", "aerocfd_ref.x")
MPIEXEC = "/opt/conda/bin/mpiexec"


def available():
    return os.path.exists(EXE) and os.path.exists(MPIEXEC)


def rank_blocks(nblock, numprocs):
    """Blocks (0-based) of every rank, as main.f90:99-133 assigns them."""
    if nblock % numprocs == 0:
        myn = lastn = nblock // numprocs
    else:
        myn = nblock // numprocs + 1
        lastn = nblock - myn * (numprocs - 1)
        if lastn <= 0:
            myn = nblock // numprocs
            lastn = nblock - myn * (numprocs - 1)
    if lastn <= 0:
        raise ValueError("the reference cannot distribute %d blocks over %d ranks" % (nblock, numprocs))
    out = []
    for r in range(numprocs):
        n = myn if r < numprocs - 1 else lastn
        out.append(list(range(r * myn, r * myn + n)))
    return out


def control_values(cfg, ntstep, iflag_init, writeinterval=None):
    """Control.in values for a FlowConfig (3-D, WENO3 + Roe, no filter,
    no acoustics, no averaging)."""
    return dict(iflag_dimension=3, iflag_solver=cfg.iflag_solver, iflag_time=cfg.iflag_time, cfl=cfg.cfl,
                dt=cfg.dt, ma=cfg.ma, re=cfg.re, tinf=cfg.tinf, aoa=cfg.aoa_deg, scale_x=1.0, scale_y=1.0,
                scale_z=1.0, lref=1.0, sref=1.0, fltr=0.0, iflag_splittingtype=1, iflag_inviscid=4,
                iflag_filter=0, iflag_timeadvance=cfg.iflag_timeadvance, iflag_turbulence=cfg.iflag_turbulence,
                iflag_des=cfg.iflag_des, writeinterval=writeinterval or ntstep, ntstep=ntstep,
                grid="input/grid.dat", bcinfo="input/bcinfo.dat", iflag_init=iflag_init,
                restartfile="none", nfile=1, restart_timestep=0, nobs=0, ob_dist=0.0, acoustic_length=0.0,
                iflag_acoustic=0, Start_Acsoutic=0, End_Acsoutic=0, iflag_avg=0, Start_Avg=0, End_Avg=0,
                Avg_interval=1)


def _physical(r, a):
    """The physical range of an array over a block's extended extents."""
    return a[r[10] - r[4]:r[11] - r[4] + 1, r[8] - r[2]:r[9] - r[2] + 1, r[6] - r[0]:r[7] - r[0] + 1]


def write_case(work, cfg, dims, subfaces, coords, ntstep, ranks, restart=None, writeinterval=None):
    """Write the case into work/ in the reference's formats.  coords[m] =
    (x, y, z) over the extended extents or the physical range; restart =
    per block (ranges, pri_v, nut) over the extended extents for a restart
    start (iflag_init 1), or None for the uniform start (iflag_init 0)."""
    os.makedirs(os.path.join(work, "input"), exist_ok=True)
    os.makedirs(os.path.join(work, "result"), exist_ok=True)
    formats.write_control(os.path.join(work, "input", "Control.in"),
                          control_values(cfg, ntstep, 0 if restart is None else 1, writeinterval))
    phys = []
    for m, (d, c) in enumerate(zip(dims, coords)):
        ni, nj, nk = d
        if c[0].shape == (nk, nj, ni):
            phys.append(c)
        else:
            from aerocfd_amd.cases import ranges_of
            r = ranges_of(d, subfaces[m])
            phys.append(tuple(_physical(r, a) for a in c))
    formats.write_grid(os.path.join(work, "input", "grid.dat"), dims, phys)
    formats.write_bcinfo(os.path.join(work, "input", "bcinfo.dat"), subfaces)
    if restart is not None:
        for rank, blocks in enumerate(rank_blocks(len(dims), ranks)):
            path = os.path.join(work, "result", "flowfield_%010d_%04d.dat" % (0, rank))
            rr = [restart[m][0] for m in blocks]
            from aerocfd_amd.cases import ranges_of
            ext_coords = []
            for m in blocks:
                r = ranges_of(dims[m], subfaces[m])
                x = np.zeros((r[5] - r[4] + 1, r[3] - r[2] + 1, r[1] - r[0] + 1))
                full = []
                for a in range(3):
                    f = x.copy()
                    f[r[10] - r[4]:r[11] - r[4] + 1, r[8] - r[2]:r[9] - r[2] + 1,
                      r[6] - r[0]:r[7] - r[0] + 1] = phys[m][a]
                    full.append(f)
                ext_coords.append(tuple(full))
            formats.write_flowfield(path, rr, ext_coords, [restart[m][1] for m in blocks],
                                    [restart[m][2] for m in blocks])


def _limits():
    # the reference keeps its temporaries on the stack (automatic arrays)
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run(work, ranks, timeout=600, env=None):
    """Run the program in work/; returns (stdout, seconds)."""
    e = dict(os.environ)
    e["OMP_NUM_THREADS"] = "1"
    if env:
        e.update(env)
    t0 = time.perf_counter()
    r = subprocess.run([MPIEXEC, "-n", str(ranks), EXE], cwd=work, capture_output=True, text=True,
                       timeout=timeout, preexec_fn=_limits, env=e)
    dt = time.perf_counter() - t0
    if r.returncode != 0:
        raise RuntimeError("reference program failed (rc %d):\n%s\n%s" % (r.returncode, r.stdout[-3000:],
                                                                          r.stderr[-3000:]))
    return r.stdout, dt


_STEP = re.compile(r"Max CPU Time =\s*([0-9.]+)\s*Min CPU Time =\s*([0-9.]+)")


def step_times(stdout):
    """(max, min) over ranks of every step's wall time, as printed."""
    return [(float(a), float(b)) for a, b in _STEP.findall(stdout)]


def read_output(work, dims, subfaces, ranks, timestep, ma):
    """The flowfield files of a timestep, as one list over the blocks."""
    out = [None] * len(dims)
    for rank, blocks in enumerate(rank_blocks(len(dims), ranks)):
        path = os.path.join(work, "result", "flowfield_%010d_%04d.dat" % (timestep, rank))
        got = formats.read_flowfield(path, [dims[m] for m in blocks], ma)
        for m, g in zip(blocks, got):
            out[m] = g
    return out


def read_residuals(work):
    """Rows of resi_history.dat: timestep and the six residuals."""
    return formats.read_residuals(os.path.join(work, "result", "resi_history.dat"))


def scratch():
    return tempfile.mkdtemp(prefix="aerocfd_ref_")


def cleanup(work):
    shutil.rmtree(work, ignore_errors=True)
