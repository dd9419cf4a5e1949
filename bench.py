"""Throughput of the residual-and-update path on one MI355X.

Default workload (--workload channel_bump): the combined configuration, a
channel with a Gaussian bump (0.05 chord) cut into 40 seeded uneven blocks,
98.1 M nodes in all, tanh wall clustering, Mach 0.6, Re 1e6, Spalart-
Allmaras with DES97, local time step, WENO3 + Roe inviscid flux, central
viscous flux, first-order implicit LU-SGS, FP64; every stage and sweep one
launch over all blocks, one captured HIP graph per step.  Metrics, edge
lengths and wall distance are computed on the device from the coordinates
before the timed region.  A second timed run of the same workload with FP32
reconstruction / splitting / viscous fluxes (FP64 update) is reported under
"fp32_flux_variant".

--workload flat_plate: 12 seeded uneven blocks of a turbulent flat plate
(Mach 0.3, Re_L 5e6, SA, wall / inflow / outflow / symmetry / farfield).
--workload tgv: Taylor-Green vortex, one 256^3 block on [0, 2 pi]^3 with
every face pair self-connected, Mach 0.1, Re 1600, laminar Navier-Stokes.

One step is one pass of main.f90:343-447 plus the residual norm of that
step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload channel_bump|flat_plate|tgv]

Prints one JSON line.  value = Mcell-updates/s = nodes * steps / seconds with
the state resident in HBM; roofline = the dominant single kernel's
algorithmic bytes per launch over its HIP-event-timed average duration;
cpu_baseline = the reference's own Fortran (compiled from its sources into
oracle/_ref) timing the same step on a bounded sample of the workload on one
host core, and its whole program under MPI on several host cores (or the C
oracle when the compiled reference is absent).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# Algorithmic bytes per computed node of each timed kernel/stage (FP64,
# each array element counted once per read and once per write).
BYTES_PER_NODE = {
    "getpri": 40 + 56,                              # q -> pri_v
    "inviscid": 40 + 72 + 8 + 40,                   # q, dxidx, inv_j -> rhsi
    "viscous": 32 + 8 + 72 + 8 + 16 + 40,           # u,v,w,T, mu_t, dxidx, inv_j -> mu, rhsv
    # tile sweeps over the sweep layout: 22 operands in, 5 results out
    "lusgs_forward": 22 * 8 + 40,                   # dxidx, inv_j, q, c, phi/dt, rhs -> dq*
    "lusgs_backward": 22 * 8 + 40,                  # dxidx, inv_j, q, c, phi/dt, dq* -> q + dq
    # 12 operands in (natural) and out (sweep layout): the 10 grid operands
    # stay in the sweep layout from one update of a block to the next
    "lusgs_gather": 2 * 12 * 8,
    "lusgs_add": 40 + 40,                           # q + dq (sweep layout) -> q
    "lusgs_other": 8 + 80 + 40 + 8,                 # dt, rhsi+rhsv -> rhs, phi/dt (tile sweeps)
    # SA scalar sweeps: a, b, c, diag and rhs (forward) / dnut* (backward)
    # in, dnut* / dnut out
    "sa_forward": 5 * 8 + 8,
    "sa_backward": 5 * 8 + 8,
    # the batched SA update's passes: stage A (length, convection, alpha /
    # beta / gamma products): dist, spacing, rho, u, v, w, nut, mu,
    # dxidx, alpha/beta/gamma in; length, sa_rhs, 4 work fields out
    # (SA_PRE_DERIVED: alpha/beta/gamma formed from dxidx and 1/J)
    "sa_pre": 16 + 32 + 8 + 8 + 72 + 72 + 8 + 8 + 32,
    # stage B + sweep operands: 4 work fields, dxidx, inv_j, nut, sa_rhs,
    # rho, u, v, w, mu, vorticity, length, dt in; sa_rhs, a, b, c, diag out
    "sa_pre_b": 32 + 72 + 8 + 8 + 8 + 32 + 8 + 8 + 8 + 8 + 8 + 32,
    "sa_gather": 2 * 5 * 8,                         # a, b, c, diag, rhs -> SA sweep layout
    "sa_add": 8 + 16,                               # dnut (layout), nut -> nut
    # single kernels of the batched inviscid / viscous stages
    "inviscid_zero": 40,                            # rhsi cleared
    "inviscid_i": 40 + 24 + 8 + 40,                 # q, grad xi, inv_j -> rhsi
    "inviscid_j": 40 + 24 + 8 + 80,                 # q, grad eta, inv_j, rhsi -> rhsi
    "inviscid_k": 40 + 24 + 8 + 80,
    "viscous_node": 32 + 8 + 72 + 8 + 96 + 16,      # u,v,w,T, mu_t, dxidx, inv_j -> node fluxes, mu_l, mu
    "viscous_diff": 96 + 40,                        # node fluxes -> rhsv
}
# with SA the viscous node kernel also stores the vorticity magnitude
VISCOUS_NODE_SA = BYTES_PER_NODE["viscous_node"] + 8
# stage A when the metrics come from acfd_grid_metrics: 1/J in place of
# alpha/beta/gamma
SA_PRE_DERIVED = BYTES_PER_NODE["sa_pre"] - 72 + 8
# stage A once the flow-independent length scale (plain SA, DES97) is
# formed: no dist, spacing in, no length out
SA_PRE_LENGTH_KEPT = 24
# the tile sweep backward pass adds dq to q itself on one block's launch
# under 2^22 computed nodes (never in the batched sweeps): q in and q out
# instead of dq out (the add pass is skipped)
FUSED_ADD_MAX_NODES = 1 << 22
FUSED_BACKWARD_BYTES = 24 * 8 + 40 + 40
# scheme -> (inviscid_scheme, iflag_timeadvance, workload suffix)
SCHEMES = {
    "reference": (0, None, "weno3_roe_lusgs"),
    "weno5_rk3": (1, 4, "sw_weno5_rk3"),
    "wcns_rk3": (2, 4, "wcns_rk3"),
}
# stages timed as one call but launching several kernels
LAUNCHES_PER_CALL = {"inviscid": 3, "viscous": 2}
SINGLE_KERNEL = {"getpri", "lusgs_forward", "lusgs_backward", "lusgs_gather", "lusgs_add", "lusgs_other",
                 "sa_forward", "sa_backward", "inviscid_zero", "inviscid_i", "inviscid_j", "inviscid_k",
                 "viscous_node", "viscous_diff", "sa_gather", "sa_pre_b", "sa_pre", "sa_add", "viscous_node_dt"}
# timers that are one kernel only on the batched SA path (per block otherwise)
SA_BATCHED_ONLY = {"sa_pre", "sa_add"}
HYPERPLANE_BYTES = {  # the hyperplane sweeps read the operands in place
    "lusgs_forward": 40 + 72 + 8 + 24 + 8 + 40 + 40,
    "lusgs_backward": 40 + 72 + 8 + 24 + 8 + 40 + 40,
    "lusgs_other": 32 + 72 + 8 + 80 + 40 + 32 + 80 + 80,
    "lusgs_add": 40 + 40 + 40,
}


# HBM bytes per launch of the timed kernels, from two rocprofv3 --pmc runs
# (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB copy read through the same
# counters.  The channel with bump at FP64, batched stages and sweeps
# (profiles/channel_bump_fp64_hbm_traffic.md); the sweeps' figures include
# the refill of the exchange slots that shares their timer.
MEASURED_TRAFFIC_CHANNEL = {
    "getpri": 10.033e9,
    "inviscid_i": 11.021e9,
    "inviscid_j": 17.097e9,
    "inviscid_k": 17.817e9,
    "viscous_node_dt": 30.728e9,
    "viscous_diff": 20.763e9,
    "lusgs_gather": 23.392e9,
    "lusgs_forward": 28.469e9,
    "lusgs_backward": 28.466e9,
    "lusgs_add": 9.256e9,
    "sa_pre": 17.185e9,
    "sa_pre_b": 24.725e9,
    "sa_gather": 9.872e9,
    "sa_forward": 6.167e9,
    "sa_backward": 6.158e9,
    "sa_add": 3.286e9,
}
# The TGV 256^3 block, FP64, tile sweeps (profiles/tgv256_fp64_hbm_traffic.md).
MEASURED_TRAFFIC_256 = {
    "getpri": 1.687e9,
    "lusgs_other": 2.322e9,
    "lusgs_gather": 4.092e9,
    "lusgs_forward": 4.456e9,
    "lusgs_backward": 4.456e9,
    "lusgs_add": 1.576e9,
}


CPU_SAMPLES = {
    "tgv": ("cases.tgv_box(96)", 24, "TGV 96^3 periodic block"),
    "flat_plate": ("cases.flat_plate(nblk=(4, 3, 1), lo=(24, 20, 12), hi=(48, 32, 16))", 150,
                   "flat plate, the same 12-block layout at reduced node counts"),
    "channel_bump": ("cases.channel_bump(lo=(16, 14, 10), hi=(22, 18, 14), first=2.0e-4)", 120,
                     "channel with bump, the same 40-block layout at reduced node counts"),
}


# the reference's whole program under MPI: (case, steps, what)
MPI_SAMPLES = {
    "flat_plate": ("cases.flat_plate(nblk=(4, 3, 1), lo=(40, 30, 20), hi=(64, 44, 28))", 20,
                   "flat plate, the same 12-block layout at reduced node counts"),
    "channel_bump": ("cases.channel_bump(lo=(28, 20, 16), hi=(36, 26, 20), first=2.0e-4)", 20,
                     "channel with bump, the same 40-block layout at reduced node counts"),
}


def cpu_baseline_mpi(workload, ranks, timeout=600):
    """The reference's whole program (main.f90 with every source file,
    compiled into oracle/_ref) on `ranks` MPI ranks of the host, on a
    reduced sample of the multi-block workload written in its own file
    formats: the median over steps 2..K of the per-step wall time it prints
    itself (main.f90:575-586, the max over ranks).  None where the program
    is not built or the workload is the periodic TGV (the reference's
    coordinate exchange has no period shift)."""
    if workload not in MPI_SAMPLES or ranks <= 0:
        return None
    import statistics
    from aerocfd_amd import cases
    from oracle import refrun as RR
    if not RR.available():
        return None
    make, steps, what = MPI_SAMPLES[workload]
    case = eval(make, {"cases": cases})
    work = RR.scratch()
    try:
        RR.write_case(work, case.cfg, case.dims, case.subfaces, case.coords, steps, ranks, writeinterval=10 ** 6)
        out, wall = RR.run(work, ranks, timeout=timeout)
        st = RR.step_times(out)
    except Exception as exc:  # reported, not fatal: the GPU figure stands alone
        return {"value": None, "unit": "Mcell-updates/s", "cores": ranks, "kind": "reference",
                "sample": "failed: %s" % str(exc)[-300:]}
    finally:
        RR.cleanup(work)
    med = statistics.median([a for a, _ in st[1:]])
    return {"value": round(case.totalnodes / med / 1e6, 3), "unit": "Mcell-updates/s", "cores": ranks,
            "kind": "reference",
            "sample": "%s (%d nodes, %d blocks), the reference's whole program (amdflang -O2, MPICH) on %d MPI "
                      "ranks, uniform start, %d steps; median of the per-step max-over-ranks wall time it "
                      "prints, steps 2..%d: %.3f s" % (what, case.totalnodes, len(case.dims), ranks, steps,
                                                       steps, med)}


def cpu_baseline(workload="tgv", ranks=0, timeout=300):
    """The CPU figures beside the GPU line: the reference's whole program
    under MPI on `ranks` host cores (when built; the headline CPU figure),
    and its routines driven in-process on one core (the reference's own
    Fortran through the harness, or the oracle where it is absent)."""
    one = cpu_baseline_1core(workload, timeout)
    mpi = cpu_baseline_mpi(workload, ranks)
    if mpi is None or mpi.get("value") is None:
        if mpi is not None:
            one["mpi_run"] = mpi
        return one
    mpi["single_core"] = one
    return mpi


def cpu_baseline_1core(workload="tgv", timeout=300):
    """Time the reference's own Fortran (or the oracle) on a reduced sample
    of the workload for a few steps in a child process (one core)."""
    make, steps, what = CPU_SAMPLES[workload]
    code = r"""
import json, os, sys, time, tempfile
sys.path.insert(0, %r)
import numpy as np
from aerocfd_amd import cases
from oracle import refwrap as R, oracle as O
from tests.helpers import ora_config, ora_states
steps = %d
case = %s
if "dxidx" not in case.fields[0]:
    # coordinates only: metrics, edge lengths and wall distance by the
    # oracle's restatement of the reference's pre-processing (not timed)
    exts0 = [O.Ext.from_list(r) for r in case.ranges]
    wall = O.wall_points(exts0, case.subfaces, case.coords)
    for m, (e, (x, y, z)) in enumerate(zip(exts0, case.coords)):
        f = dict(zip(("dxidx", "inv_j", "aalpha", "bbeta", "ggamma"), O.gridtransformation(e, x, y, z)))
        f["spacing"] = O.spacing(e, x, y, z)[3]
        f["dist"] = O.walldist(e, x, y, z, wall)
        for k, v in f.items():
            case.fields[m][k] = np.nan_to_num(v, nan=0.0)
kind = "reference" if R.available() else "port"
if kind == "reference":
    os.chdir(tempfile.mkdtemp())
    c = case.cfg
    R.setup(solver=c.iflag_solver, itime=c.iflag_time, turb=c.iflag_turbulence, des=c.iflag_des,
            tadv=c.iflag_timeadvance, ma=c.ma, re=c.re, tinf=c.tinf, cfl=c.cfl, dt=c.dt)
    exts = R.begin(case.dims, case.subfaces)
    sc = R.scalars()
    for m, fl in enumerate(case.fields):
        for k, v in fl.items():
            if k in R.XFER:
                R.put(m + 1, k, v)
        R.put(m + 1, "pri_v", O.getpri(exts[m], fl["q"], sc["gamma"], sc["cv"])[0])
    R.physical_bc(); R.update_buffer(); R.average()
    R.step()
    t0 = time.perf_counter()
    for s in range(steps):
        R.step()
    dt = time.perf_counter() - t0
else:
    cfg = ora_config(case.cfg, case.totalnodes)
    st = ora_states(case)
    for b in st:
        b["pri_v"] = O.getpri(b.ext, b["q"], cfg.gamma, cfg.cv)[0]
    O.prepare(st, cfg)
    O.step(st, cfg)
    t0 = time.perf_counter()
    for s in range(steps):
        O.step(st, cfg)
    dt = time.perf_counter() - t0
print(json.dumps({"kind": kind, "seconds": dt, "nodes": case.totalnodes, "steps": steps}))
""" % (ROOT, steps, make)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    try:
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout,
                           env=env)
        line = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 or not line:
            return {"value": None, "unit": "Mcell-updates/s", "cores": 1, "kind": None,
                    "sample": "failed: " + (r.stderr or r.stdout)[-300:]}
        d = json.loads(line[-1])
    except subprocess.TimeoutExpired:
        return {"value": None, "unit": "Mcell-updates/s", "cores": 1, "kind": None,
                "sample": "timed out after %d s" % timeout}
    return {"value": d["nodes"] * d["steps"] / d["seconds"] / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": d["kind"],
            "sample": "%s (%d nodes), %d steps of the same scheme and flags in FP64, %.2f s on one core"
                      % (what, d["nodes"], d["steps"], d["seconds"])}


def build_case(args):
    from aerocfd_amd import cases
    if args.workload == "channel_bump":
        return cases.channel_bump()
    if args.workload == "flat_plate":
        return cases.flat_plate()
    cut = tuple(round(args.n * c / args.cuts) for c in range(1, args.cuts))
    case = cases.tgv_box(args.n, blocks=(cut, cut, cut))
    if args.scheme != "reference":
        case.cfg.inviscid_scheme, case.cfg.iflag_timeadvance = SCHEMES[args.scheme][:2]
    return case


def timed_run(args, case, fields, coords, precision):
    """Set up one context for the workload at the given precision, run the
    warm-up and timed steps, then (graph runs) two eager steps under the
    per-kernel timers.  Returns wall seconds, the timers, the residual and
    the block ranges."""
    from aerocfd_amd import Solver
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
    cfg = case.cfg
    cfg.precision = precision
    cfg.lusgs_kernel = args.lusgs
    cfg.use_graph = 1 if args.graph else 0
    cfg.batch_blocks = 1 if args.batch else 0
    sol = Solver(cfg, case.dims, case.subfaces)
    try:
        for m, fl in enumerate(fields):
            for k, v in fl.items():
                sol.upload(m + 1, k, v)
        if coords is not None:
            sol.grid_metrics(coords)  # metrics, edge lengths, wall distance on the device
        sol.run_stage("physical_bc")
        sol.run_stage("getpri")
        sol.prepare()
        for _ in range(args.warmup):
            sol.step(1)
        sol.timing(not args.graph)
        sol.sync()
        sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sol.step(1)
        sol.sync()
        sync()
        t1 = time.perf_counter()
        tsteps = args.steps
        if args.graph:
            # the per-kernel timers run eager steps
            tsteps = 2
            sol.timing(True)
            sol.step(tsteps)
        tm = sol.timing(None)
        res = sol.residual()
        rr = [sol.ranges(m + 1) for m in range(len(case.dims))]
    finally:
        sol.close()
    return t1 - t0, tm, tsteps, res, rr


def roofline_of(args, tm, rr, sa, derived=False, length_kept=False):
    """The dominant single-kernel timer (most GPU time per step) and its
    achieved algorithmic GB/s."""
    comp_all = sum((r[13] - r[12] + 1) * (r[15] - r[14] + 1) * (r[17] - r[16] + 1) for r in rr)
    batched = {"lusgs_forward", "lusgs_backward", "lusgs_gather", "lusgs_add", "sa_forward",
               "sa_backward", "sa_gather", "sa_pre", "sa_pre_b", "sa_add"} if args.lusgs == 2 else set()
    if args.batch:
        batched |= {"lusgs_other", "getpri", "inviscid", "viscous", "find_dt", "inviscid_zero", "inviscid_i",
                    "inviscid_j", "inviscid_k", "viscous_node", "viscous_diff", "viscous_node_dt"}
    table = dict(BYTES_PER_NODE)
    if sa:
        table["viscous_node"] = VISCOUS_NODE_SA
    if sa and derived:
        table["sa_pre"] = SA_PRE_DERIVED
    if sa and length_kept and args.lusgs == 2:
        table["sa_pre"] -= SA_PRE_LENGTH_KEPT
    # the viscous node pass with the steady find_dt: + rho, c in, dt out;
    # when it also forms implicit1st's preparation (batched sweeps, no
    # separate preparation timed): + the time term phi/dt out, and the
    # difference pass + rhsi in, rhs out, rhsv not stored
    folded = args.lusgs == 2 and tm.get("lusgs_other", (0, 0))[1] == 0 and tm.get("viscous_node_dt", (0, 0))[1] > 0
    table["viscous_node_dt"] = table["viscous_node"] + 16 + 8 + (8 if folded else 0)
    if folded:
        table["viscous_diff"] = BYTES_PER_NODE["viscous_diff"] + 40
    if args.lusgs == 0:
        table.update(HYPERPLANE_BYTES)
    else:
        env = os.environ.get("ACFD_LU_FUSED_ADD")
        # batched sweeps never fuse; a block's own launch does under 2^22
        # computed nodes
        launch = max((r[13] - r[12] + 1) * (r[15] - r[14] + 1) * (r[17] - r[16] + 1) for r in rr)
        if (int(env) != 0) if env is not None else (args.lusgs != 2 and launch < FUSED_ADD_MAX_NODES):
            table["lusgs_backward"] = FUSED_BACKWARD_BYTES
        if os.environ.get("ACFD_SL_GEO_CACHE", "1") == "0":
            table["lusgs_gather"] = 2 * 22 * 8

    # sa_pre / sa_add are single kernels only on the batched SA path (the
    # per-block path times its launches of every block together)
    sa_batched = "sa_gather" in tm and tm["sa_gather"][1] > 0
    cand = {k: v for k, v in tm.items() if k in table and k in SINGLE_KERNEL and v[1] > 0
            and (sa_batched or k not in SA_BATCHED_ONLY)}
    if not cand:
        # explicit RK3 on one block: no sweep; the stages themselves (one
        # call = all kernels of the stage for one block)
        cand = {k: v for k, v in tm.items() if k in table and v[1] > 0}
    dom = max(cand, key=lambda k: cand[k][0])
    avg_ms = cand[dom][0] / cand[dom][1]
    bytes_launch = table[dom] * (comp_all if dom in batched else comp_all / len(rr))
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    return dom, avg_ms, achieved, table[dom]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", choices=["tgv", "flat_plate", "channel_bump"], default="channel_bump")
    ap.add_argument("--precision", type=int, default=0,
                    help="0 FP64 (default), 1 FP32 flux / FP64 update")
    ap.add_argument("--no-fp32-variant", action="store_true",
                    help="skip the second timed run of a multi-block workload with FP32 fluxes")
    ap.add_argument("--lusgs", type=int, default=None, help="0 hyperplane launches, 1 pipelined tile sweep per block, 2 tile sweeps "
                         "of all blocks in one launch (default: 1 for tgv, 2 for the multi-block workloads)")
    ap.add_argument("--graph", action="store_true",
                    help="replay one captured HIP graph per step in the timed region (per-kernel times then "
                         "come from two eager steps after it); the default for the multi-block workloads")
    ap.add_argument("--no-graph", action="store_true", help="eager launches for the multi-block workloads")
    ap.add_argument("--cuts", type=int, default=1, help="blocks per axis (the TGV split into cuts^3 blocks)")
    ap.add_argument("--batch", type=int, default=None,
                    help="1 = one launch per stage over all blocks (default: 0 for tgv, 1 for the multi-block "
                         "workloads)")
    ap.add_argument("--scheme", choices=list(SCHEMES), default="reference",
                    help="reference: WENO3 + Roe with the first-order implicit LU-SGS update (the reference's "
                         "scheme, the default); weno5_rk3 / wcns_rk3: Steger-Warming + WENO5 or WCNS with "
                         "explicit RK3 (TGV only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-ranks", type=int, default=min(16, os.cpu_count() or 1),
                    help="MPI ranks of the reference's own executable in the CPU baseline (0: skip that run; "
                         "default: 16, the host cores one GPU's share holds, or fewer where the host has fewer)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that "
                         "calibrates HBM byte counters)")
    args = ap.parse_args()
    fp = args.workload in ("flat_plate", "channel_bump")  # multi-block SA workloads
    if args.lusgs is None:
        args.lusgs = 2 if fp else 1
    if args.batch is None:
        args.batch = 1 if fp else 0
    if fp and not args.no_graph:
        args.graph = True
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; every block of the workload is resident on it. "
              "Sharding blocks over several GPUs is not built, so --gpus %d is not supported." % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.scheme != "reference" and fp:
        print("bench.py: --scheme applies to the TGV workload", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mb > 0:
        import torch
        src = torch.ones(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    case = build_case(args)
    fields = case.fields
    coords = getattr(case, "coords", None) if args.workload == "channel_bump" else None
    secs, tm, tsteps, res, rr = timed_run(args, case, fields, coords, args.precision)
    sa = case.cfg.iflag_turbulence == 2 and case.cfg.iflag_solver == 0
    # the reference's totalnodes (sum of ni nj nk); the cut TGV counts the
    # nodes of the single block
    nodes = case.totalnodes if fp else args.n ** 3
    value = nodes * args.steps / secs / 1e6
    dom, avg_ms, achieved, bpn = roofline_of(args, tm, rr, sa, derived=coords is not None,
                                             length_kept=case.cfg.iflag_des != 2)
    default_cfg = (not fp and args.n == 256 and args.lusgs == 1 and args.precision == 0 and args.cuts == 1
                   and args.scheme == "reference")
    traffic = MEASURED_TRAFFIC_256.get(dom) if default_cfg else None
    if args.workload == "channel_bump" and args.precision == 0 and args.lusgs == 2:
        traffic = MEASURED_TRAFFIC_CHANNEL.get(dom)
    if fp:
        wl = case.name + ("_turbulent_sa_des97" if case.cfg.iflag_des == 1 else "_turbulent_sa") + "_weno3_roe_lusgs"
    else:
        wl = "tgv_%d^3_periodic_laminar_ns_%s" % (args.n, SCHEMES[args.scheme][2])
    out = {
        "metric": "Mcell-updates/s",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * secs / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 0 else "f32-flux/f64-update",
        "data": "synthetic",
        "config": {"workload": wl, "nodes": nodes, "blocks": len(case.dims), "graph": bool(args.graph),
                   "batch_blocks": bool(args.batch),
                   "mach": case.cfg.ma, "re": case.cfg.re,
                   "scheme": ("WENO3+Roe, central viscous, %sLU-SGS" % ("SA, " if fp else "")
                              if args.scheme == "reference" else
                              "%s, central viscous, RK3" % ("Steger-Warming+WENO5" if args.scheme == "weno5_rk3"
                                                            else "WCNS")),
                   "lusgs_kernel": {0: "hyperplanes", 1: "pipelined-tiles", 2: "batched-pipelined-tiles"}[args.lusgs]},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 2), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic,
                     "bytes_per_node": bpn, "avg_launch_ms": round(avg_ms, 4)},
        "stage_ms_per_step": {k: round(v[0] / tsteps, 4) for k, v in tm.items() if v[1] > 0},
        "residual": [float("%.6e" % x) for x in res[:5]],
    }
    if fp and args.precision == 0 and not args.no_fp32_variant:
        # the same workload with FP32 reconstruction / splitting / viscous
        # fluxes and the FP64 update (no counterpart in the reference)
        s32, tm32, ts32, _, _ = timed_run(args, case, fields, coords, 1)
        out["fp32_flux_variant"] = {
            "value": round(nodes * args.steps / s32 / 1e6, 3), "ms_per_step": round(1e3 * s32 / args.steps, 4),
            "dtype": "f32-flux/f64-update",
            "stage_ms_per_step": {k: round(v[0] / ts32, 4) for k, v in tm32.items() if v[1] > 0}}
    if not args.no_cpu:
        # the reference has only its own scheme; the variants report none
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_ranks) if args.scheme == "reference" else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
