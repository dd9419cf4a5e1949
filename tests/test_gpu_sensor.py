"""The learned WCNS sensor on the GPU (no counterpart in the reference):
training kernels against the numpy oracle, the BF16 MFMA inference against
the FP32 network, and the trained sensor inside the WCNS residual against
plain WCNS."""
import sys

import numpy as np
import pytest

from aerocfd_amd import Solver, cases
from aerocfd_amd import solver as S
from oracle import mlp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def sample_data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.random((n, mlp.NF)).astype(np.float32) * 1.2
    y = (rng.random(n) > 0.6).astype(np.float32)
    return X, y


def test_training_gradient_matches_oracle():
    """FP32 forward/backward kernels vs the float64 oracle: loss and
    gradient to 1e-4 relative (FP32 sums over 1000 samples)."""
    say("test_training_gradient_matches_oracle n=1000")
    w = mlp.init_params(3)
    X, y = sample_data(1000, 4)
    loss, g = S.mlp_grad(w, X, y)
    lo, go = mlp.loss_and_grad(w, X, y)
    say("  loss gpu %.8f oracle %.8f" % (loss, lo))
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    err = np.linalg.norm(g - go) / np.linalg.norm(go)
    say("  gradient rel err %.3e" % err)
    assert err <= 1e-4


def test_mfma_inference_matches_fp32_network():
    """BF16 MFMA inference (the routine fused into the WCNS kernels) vs the
    FP32 network: probabilities within BF16 rounding."""
    say("test_mfma_inference_matches_fp32_network n=3000")
    rng = np.random.default_rng(5)
    w = mlp.init_params(6) * np.float32(1.5)
    w[mlp.NPARAM - 1] = 0.3
    X = (rng.random((3000, mlp.NF)) * 1.2).astype(np.float32)
    p = S.sensor_infer(w, X)
    po = mlp.predict(w, X)
    d = np.abs(p - po)
    say("  max |p - p_fp32| %.4f, mean %.5f" % (d.max(), d.mean()))
    assert d.max() <= 0.05 and d.mean() <= 0.01
    assert np.corrcoef(p, po)[0, 1] > 0.999


def run(case, nsteps, sensor=None):
    sol = Solver(case.cfg, case.dims, case.subfaces)
    case.upload(sol)
    sol.run_stage("physical_bc")
    sol.run_stage("getpri")
    sol.prepare()
    if sensor is not None:
        sol.sensor_weights(sensor["w"])
        sol.sensor_enable(True, sensor["thr"], sensor["audit"], sensor["bound"])
    sol.step(nsteps)
    q = sol.download(1, "q")
    stats = sol.sensor_stats() if sensor is not None else None
    return sol, q, stats


def test_sensor_trained_on_shock_vortex():
    """Train on interfaces of the shock/vortex/entropy-wave flow, then run
    WCNS with the sensor: classification accuracy, share of faces given
    linear weights, and the density error against plain WCNS."""
    say("test_sensor_trained_on_shock_vortex 96x48x4 Mach 1.5")
    case = cases.shock_vortex(96, 48, 4, mach=1.5, seed=4)
    sol, _, _ = run(case, 10)
    try:
        n = sol.sensor_collect(stride=1)
        say("  samples", n)
        losses = sol.sensor_train(400, batch=4096, lr=3e-3, seed=1)
        say("  loss %.4f -> %.4f" % (losses[0], losses[-1]))
        ev = sol.sensor_evaluate(0.5)
        say("  evaluate", {k: round(v, 4) for k, v in ev.items()})
        w = sol.sensor_weights()
    finally:
        sol.close()
    assert losses[-1] < 0.5 * losses[0]
    assert ev["accuracy"] >= 0.9
    # plain WCNS against WCNS with the sensor, 20 steps from the same start
    base = cases.shock_vortex(96, 48, 4, mach=1.5, seed=4)
    s0, q0, _ = run(base, 20)
    s0.close()
    sv = cases.shock_vortex(96, 48, 4, mach=1.5, seed=4)
    s1, q1, st = run(sv, 20, dict(w=w, thr=0.5, audit=10, bound=5e-2))
    s1.close()
    d = q1[..., 0] - q0[..., 0]
    l2 = np.linalg.norm(d) / np.linalg.norm(q0[..., 0])
    say("  sensor stats", st)
    say("  density rel L2 vs plain WCNS %.3e, rel Linf %.3e" % (l2, np.abs(d).max() / np.abs(q0[..., 0]).max()))
    assert st["faces"] > 0 and st["linear"] > 0
    assert l2 <= 1e-3


def test_sensor_cost_on_smooth_flow():
    """The WCNS residual with and without the sensor on a smooth flow
    (Taylor-Green, 96^3), timed for the record.  With the WENO-JS weights and
    the 1e-3 label bound, smooth but resolved data are labelled nonlinear
    (the weights differ from the linear ones by O(h)), so the sensor gives
    few faces linear weights here, and its evaluation (features, three
    layers through LDS and MFMA) costs more than the weights it saves: the
    measured times are reported, and the residual must stay the plain one
    to rounding where nothing is given linear weights."""
    say("test_sensor_speed_on_smooth_flow tgv 96^3 WCNS")
    case = cases.tgv_box(96)
    case.cfg.iflag_timeadvance = 4
    case.cfg.inviscid_scheme = 2
    case.cfg.iflag_solver = 1
    sol = Solver(case.cfg, case.dims, case.subfaces)
    try:
        case.upload(sol)
        sol.run_stage("getpri")
        sol.prepare()
        sol.sensor_collect(stride=4)
        sol.sensor_train(200, batch=4096, lr=3e-3, seed=2)
        times, rhs = {}, {}
        for on in (False, True):
            sol.sensor_enable(on, 0.5, 0, 1.0)
            sol.run_stage("inviscid")
            sol.sync()
            sol.timing(True)
            for _ in range(5):
                sol.run_stage("inviscid")
            sol.sync()
            times[on] = sol.timing(None)["inviscid"][0] / 5
            rhs[on] = sol.download(1, "rhsi")
        st = sol.sensor_stats()
        share = st["linear"] / max(st["faces"], 1)
        say("  inviscid ms: plain %.3f, sensor %.3f; linear share %.3f" % (times[False], times[True], share))
        dev = np.linalg.norm(rhs[True] - rhs[False]) / np.linalg.norm(rhs[False])
        say("  residual deviation from plain WCNS %.3e" % dev)
        assert st["faces"] > 0
        assert dev <= max(1e-12, 10.0 * share)
    finally:
        sol.close()


def test_sensor_decision_reuse_identical():
    """Decisions recorded by the network and read back from the byte mask
    give the same residual to the last bit: with refresh interval 3 the
    first inviscid evaluation runs the network (and records), the next two
    read the mask, all on the same state; the network run every time
    (interval 1) gives the same bits too."""
    say("test_sensor_decision_reuse_identical 128x64x4 Mach 2.9 WCNS")
    case = cases.shock_vortex(128, 64, 4, mach=2.9, seed=4)
    sol, _, _ = run(case, 5)
    try:
        sol.sensor_collect(stride=1)
        sol.sensor_train(200, batch=4096, lr=3e-3, seed=3)
        rhs = {}
        for interval in (1, 3):
            sol.sensor_enable(True, 0.5, 0, 1.0)
            sol.sensor_refresh(interval)
            for call in range(3):
                sol.run_stage("inviscid")
                rhs[(interval, call)] = sol.download(1, "rhsi")
            st = sol.sensor_stats()
            say("  interval %d: faces %d, linear %d" % (interval, st["faces"], st["linear"]))
            assert 0 < st["linear"] < st["faces"]
    finally:
        sol.close()
    ref = rhs[(1, 0)]
    for key, r in rhs.items():
        assert np.array_equal(r, ref), key
