"""Throughput of the residual-and-update path on one MI355X.

Default workload (--workload tgv): Taylor-Green vortex, one block of 256^3
nodes on [0, 2 pi]^3 with every face pair self-connected (node 256 ==
node 1), Mach 0.1, Re 1600, laminar Navier-Stokes with the reference scheme
(WENO3 + Roe inviscid flux, central viscous flux, first-order implicit
LU-SGS), FP64.  One step is one pass of main.f90:343-447 plus the residual
norm of that step.

--workload flat_plate: 12 seeded uneven blocks of a turbulent flat plate
(tanh wall clustering, Mach 0.3, Re_L 5e6, SA, wall / inflow / outflow /
symmetry / farfield faces, local time step, first-order implicit), by
default with FP32 fluxes and the FP64 update, every stage and sweep one
launch over all blocks.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--workload tgv|flat_plate]

Prints one JSON line.  value = Mcell-updates/s = nodes * steps / seconds with
the state resident in HBM; roofline = the dominant kernel's algorithmic bytes
per launch over its HIP-event-timed average duration in the timed region;
cpu_baseline = the reference's own Fortran (compiled from its sources into
oracle/_ref) timing the same step on a bounded sample of the workload on one
host core (or the C oracle when the compiled reference is absent).
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
    # tile sweeps over the sweep layout: 24 operands in, 5 results out
    "lusgs_forward": 24 * 8 + 40,                   # q, dxidx, inv_j, r_abc, diag, rhs -> dq*
    "lusgs_backward": 24 * 8 + 40,                  # q, dxidx, inv_j, r_abc, diag, dq* -> dq
    # 14 operands in (natural) and out (sweep layout): the 10 grid operands
    # stay in the sweep layout from one update of a block to the next
    "lusgs_gather": 2 * 14 * 8,
    "lusgs_add": 40 + 40,                           # q + dq (sweep layout) -> q
    "lusgs_other": 32 + 72 + 8 + 80 + 40 + 32,      # u,v,w,c, dxidx, dt, rhsi+rhsv -> rhs, r_abc, diag
}
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
SINGLE_KERNEL = {"getpri", "lusgs_forward", "lusgs_backward", "lusgs_gather", "lusgs_add", "lusgs_other"}
HYPERPLANE_BYTES = {  # the hyperplane sweeps read the operands in place
    "lusgs_forward": 40 + 72 + 8 + 24 + 8 + 40 + 40,
    "lusgs_backward": 40 + 72 + 8 + 24 + 8 + 40 + 40,
    "lusgs_other": 32 + 72 + 8 + 80 + 40 + 32 + 80 + 80,
    "lusgs_add": 40 + 40 + 40,
}


# HBM bytes per launch of the timed kernels on the default workload (TGV
# 256^3, FP64, tile sweeps), from two rocprofv3 --pmc runs (FETCH_SIZE,
# WRITE_SIZE) scaled by a 1 GiB copy; profiles/tgv256_fp64_hbm_traffic.md.
MEASURED_TRAFFIC_256 = {
    # each sweep plus the refill of the exchange region (132 MB written)
    # that follows it inside the same timer
    "lusgs_forward": 5.073e9,
    "lusgs_backward": 5.075e9,
    "lusgs_gather": 4.763e9,   # a later gather (the grid operands in place)
    "lusgs_add": 1.572e9,
    "lusgs_other": 4.537e9,
}


CPU_SAMPLES = {
    "tgv": ("cases.tgv_box(96)", 24, "TGV 96^3 periodic block"),
    "flat_plate": ("cases.flat_plate(nblk=(4, 3, 1), lo=(24, 20, 12), hi=(48, 32, 16))", 150,
                   "flat plate, the same 12-block layout at reduced node counts"),
    "channel_bump": ("cases.channel_bump(lo=(16, 14, 10), hi=(22, 18, 14), first=2.0e-4)", 120,
                     "channel with bump, the same 40-block layout at reduced node counts"),
}


def cpu_baseline(workload="tgv", timeout=300):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--workload", choices=["tgv", "flat_plate", "channel_bump"], default="tgv")
    ap.add_argument("--precision", type=int, default=None,
                    help="0 FP64, 1 FP32 flux / FP64 update (default: 0 for tgv, 1 for flat_plate)")
    ap.add_argument("--lusgs", type=int, default=None, help="0 hyperplane launches, 1 pipelined tile sweep per block, 2 tile sweeps "
                         "of all blocks in one launch (default: 1 for tgv, 2 for flat_plate)")
    ap.add_argument("--graph", action="store_true",
                    help="replay one captured HIP graph per step in the timed region (per-kernel times then "
                         "come from two eager steps after it); the default for the multi-block workloads")
    ap.add_argument("--no-graph", action="store_true", help="eager launches for the multi-block workloads")
    ap.add_argument("--cuts", type=int, default=1, help="blocks per axis (the workload split into cuts^3 blocks)")
    ap.add_argument("--batch", type=int, default=None,
                    help="1 = one launch per stage over all blocks (default: 0 for tgv, 1 for flat_plate)")
    ap.add_argument("--scheme", choices=list(SCHEMES), default="reference",
                    help="reference: WENO3 + Roe with the first-order implicit LU-SGS update (the reference's "
                         "scheme, the default); weno5_rk3 / wcns_rk3: Steger-Warming + WENO5 or WCNS with "
                         "explicit RK3 (TGV only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that "
                         "calibrates HBM byte counters)")
    args = ap.parse_args()
    fp = args.workload in ("flat_plate", "channel_bump")  # multi-block SA workloads
    if args.precision is None:
        args.precision = 1 if fp else 0
    if args.lusgs is None:
        args.lusgs = 2 if fp else 1
    if args.batch is None:
        args.batch = 1 if fp else 0
    if fp and not args.no_graph:
        args.graph = True
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; every block of the workload is resident on it. "
              "Sharding blocks over several GPUs is described in DESIGN.md but not built, so --gpus %d "
              "is not supported." % args.gpus, file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from aerocfd_amd import Solver, cases

    if args.copy_probe_mb > 0:
        import torch
        src = torch.ones(args.copy_probe_mb * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    if args.workload == "channel_bump":
        case = cases.channel_bump()
    elif fp:
        case = cases.flat_plate()
    else:
        cut = tuple(round(args.n * c / args.cuts) for c in range(1, args.cuts))
        case = cases.tgv_box(args.n, blocks=(cut, cut, cut))
    if args.scheme != "reference":
        if fp:
            print("bench.py: --scheme applies to the TGV workload", file=sys.stderr)
            sys.exit(2)
        case.cfg.inviscid_scheme, case.cfg.iflag_timeadvance = SCHEMES[args.scheme][:2]
    case.cfg.precision = args.precision
    case.cfg.lusgs_kernel = args.lusgs
    case.cfg.use_graph = 1 if args.graph else 0
    case.cfg.batch_blocks = 1 if args.batch else 0
    sol = Solver(case.cfg, case.dims, case.subfaces)
    case.upload(sol)
    del case.fields
    if args.workload == "channel_bump":
        sol.grid_metrics(case.coords)  # metrics, edge lengths, wall distance on the device
        del case.coords
    sol.run_stage("physical_bc")
    sol.run_stage("getpri")
    sol.prepare()
    for _ in range(args.warmup):
        sol.step(1)
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731
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
        # the per-stage timers run eager steps
        tsteps = 2
        sol.timing(True)
        sol.step(tsteps)
    tm = sol.timing(None)
    res = sol.residual()
    rr = [sol.ranges(m + 1) for m in range(len(case.dims))]
    sol.close()
    secs = t1 - t0
    # the reference's totalnodes (sum of ni nj nk); the cut TGV counts the
    # nodes of the single block
    nodes = case.totalnodes if fp else args.n ** 3
    value = nodes * args.steps / secs / 1e6
    # dominant kernel among the timed ones (per launch averages)
    # computed nodes per launch: is1..ie1 x js1..je1 x ks1..ke1 of one block, or of
    # every block for the batched LU-SGS phases
    comp_all = sum((r[13] - r[12] + 1) * (r[15] - r[14] + 1) * (r[17] - r[16] + 1) for r in rr)
    batched = {"lusgs_forward", "lusgs_backward", "lusgs_gather", "lusgs_add"} if args.lusgs == 2 else set()
    if args.batch:
        batched |= {"lusgs_other", "getpri", "inviscid", "viscous", "find_dt"}
    table = dict(BYTES_PER_NODE)
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
            table["lusgs_gather"] = 2 * 24 * 8
    # timers that time one kernel per launch (the inviscid and viscous
    # stages launch several kernels each)
    cand = {k: v for k, v in tm.items() if k in table and k in SINGLE_KERNEL and k.startswith("lusgs") and v[1] > 0}
    if cand:
        cand = {k: v for k, v in tm.items() if k in table and k in SINGLE_KERNEL and v[1] > 0}
    else:
        # explicit RK3: no sweep; the stages themselves (one call = all
        # kernels of the stage for one block)
        cand = {k: v for k, v in tm.items() if k in table and v[1] > 0}
    dom = max(cand, key=lambda k: cand[k][0])
    avg_ms = cand[dom][0] / cand[dom][1]
    bytes_launch = table[dom] * (comp_all if dom in batched else comp_all / len(rr))
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    default_cfg = (not fp and args.n == 256 and args.lusgs == 1 and args.precision == 0 and args.cuts == 1
                   and args.scheme == "reference")
    traffic = MEASURED_TRAFFIC_256.get(dom) if default_cfg else None
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
        "config": {"workload": (case.name + ("_turbulent_sa_des97" if case.cfg.iflag_des == 1 else "_turbulent_sa")
                                + "_weno3_roe_lusgs" if fp else
                                "tgv_%d^3_periodic_laminar_ns_%s" % (args.n, SCHEMES[args.scheme][2])),
                   "nodes": nodes, "blocks": len(case.dims), "graph": bool(args.graph),
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
                     "bytes_per_node": table[dom], "avg_launch_ms": round(avg_ms, 4),
                     "launches_per_call": LAUNCHES_PER_CALL.get(dom, 1)},
        "stage_ms_per_step": {k: round(v[0] / tsteps, 4) for k, v in tm.items() if v[1] > 0},
        "residual": [float("%.6e" % x) for x in res[:5]],
    }
    if not args.no_cpu:
        # the reference has only its own scheme; the variants report none
        out["cpu_baseline"] = cpu_baseline(args.workload) if args.scheme == "reference" else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
