"""Shared test helpers: oracle states built from a case, perturbations and
error norms."""
import numpy as np

from aerocfd_amd import cases
from oracle import oracle as O

FIELDS_ORA = ("q", "q0", "pri_v", "rhsi", "rhsv", "rhs", "dudx", "tao", "vor", "amu", "dt", "dxidx",
              "inv_j", "aalpha", "bbeta", "ggamma", "dist", "spacing", "length", "nut", "sa_rhs")


def ora_config(cfg, totalnodes):
    return O.make_config(ma=cfg.ma, re=cfg.re, tinf=cfg.tinf, cfl=cfg.cfl, aoa_deg=cfg.aoa_deg,
                         solver=cfg.iflag_solver, itime=cfg.iflag_time, turb=cfg.iflag_turbulence,
                         des=cfg.iflag_des, tadv=cfg.iflag_timeadvance, totalnodes=totalnodes)


def ora_states(case):
    states = []
    for m in range(len(case.dims)):
        st = O.BlockState(case.ranges[m], case.subfaces[m])
        for k, v in case.fields[m].items():
            st[k] = v
        states.append(st)
    return states


def smooth_noise(shape, rng, amp):
    """Low-frequency random field of the given shape (nk,nj,ni)."""
    nk, nj, ni = shape
    k, j, i = np.meshgrid(np.arange(nk), np.arange(nj), np.arange(ni), indexing="ij")
    f = np.zeros(shape)
    for _ in range(4):
        a, b, c = rng.uniform(0.2, 0.8, 3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        f += rng.uniform(-1, 1) * np.sin(a * i + ph[0]) * np.sin(b * j + ph[1]) * np.sin(c * k + ph[2])
    return amp * f / 4.0


def perturb_case(case, seed=0, amp=0.05, sa=False):
    """Add smooth density/velocity/pressure perturbations (keeps the state
    physical) and fill the SA / time-step inputs."""
    rng = np.random.default_rng(seed)
    g = 1.4
    for m in range(len(case.dims)):
        fl = case.fields[m]
        q = fl["q"]
        shp = q.shape[:-1]
        rho = q[..., 0] * (1 + smooth_noise(shp, rng, amp))
        u = q[..., 1] / q[..., 0] + smooth_noise(shp, rng, amp)
        v = q[..., 2] / q[..., 0] + smooth_noise(shp, rng, amp)
        w = q[..., 3] / q[..., 0] + smooth_noise(shp, rng, amp)
        p = (q[..., 4] - 0.5 * (q[..., 1] ** 2 + q[..., 2] ** 2 + q[..., 3] ** 2) / q[..., 0]) * (g - 1)
        p = p * (1 + smooth_noise(shp, rng, amp))
        fl["q"] = cases.conservative(rho, u, v, w, p)
        if sa:
            fl["nut"] = 1.0 + rng.uniform(0, 1, shp)
            fl["dist"] = 0.01 + rng.uniform(0, 1, shp)
            fl["spacing"] = 0.05 + 0.05 * rng.uniform(0, 1, shp)
            amu = fl["amu"].copy()
            amu[..., 1] = 0.01 * rng.uniform(0, 1, shp)
            amu[..., 2] = amu[..., 0] + amu[..., 1]
            fl["amu"] = amu
    return case


def rel_l2(a, b, comp_axis=-1):
    """Norm-wise relative error per component: ||a-b||_2 / ||b||_2."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.ndim == 3 or comp_axis is None:
        nb = np.linalg.norm(b)
        return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
    out = []
    for c in range(a.shape[-1]):
        nb = np.linalg.norm(b[..., c])
        out.append(np.linalg.norm(a[..., c] - b[..., c]) / (nb if nb > 0 else 1.0))
    return np.array(out)


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))
