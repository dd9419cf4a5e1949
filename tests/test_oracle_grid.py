"""Closed-form pins of the oracle's grid pre-processing: metrics of affine
grids, edge lengths of uniform grids, wall distance by brute force and of a
plane wall."""
import numpy as np

from aerocfd_amd import cases
from oracle import oracle as O


def affine_block(A, b, n=(7, 6, 5), codes=(-10,) * 6):
    dims, subs, origins = cases.split_box(*n, phys=codes)
    r = cases.ranges_of(dims[0], subs[0])
    e = O.Ext.from_list(r)
    K, J, I = np.meshgrid(np.arange(r[4], r[5] + 1), np.arange(r[2], r[3] + 1), np.arange(r[0], r[1] + 1),
                          indexing="ij")
    xi = np.stack([I, J, K]).astype(float)
    xyz = np.einsum("ab,b...->a...", A, xi) + np.asarray(b)[:, None, None, None]
    return e, subs[0], [np.ascontiguousarray(c) for c in xyz]


def test_affine_metrics_are_exact():
    rng = np.random.default_rng(0)
    A = np.eye(3) * 0.3 + 0.05 * rng.standard_normal((3, 3))
    e, _, (x, y, z) = affine_block(A, (1.0, -2.0, 0.5))
    dxidx, inv_j, al, be, ga = O.gridtransformation(e, x, y, z)
    Ainv = np.linalg.inv(A)
    det = np.linalg.det(A)
    assert np.allclose(inv_j, det, rtol=1e-13, atol=0)
    for d in range(3):
        for c in range(3):
            assert np.allclose(dxidx[..., 3 * d + c], Ainv[d, c], rtol=1e-12, atol=1e-12)
    G = Ainv @ Ainv.T * det  # (grad xi_d . grad xi_c) / J
    for c in range(3):
        assert np.allclose(al[..., c], G[0, c], rtol=1e-12, atol=1e-12)
        assert np.allclose(be[..., c], G[1, c], rtol=1e-12, atol=1e-12)
        assert np.allclose(ga[..., c], G[2, c], rtol=1e-12, atol=1e-12)


def test_uniform_spacing_and_plane_wall_distance():
    h = (0.2, 0.05, 0.1)
    A = np.diag(h)
    # wall on the j- face, y = 0 at node j = 1
    e, subs, (x, y, z) = affine_block(A, (0.0, -h[1], 0.0), codes=(-10, -10, -20, -10, -30, -30))
    spci, spcj, spck, spacing, vol = O.spacing(e, x, y, z)
    ph = np.s_[e.ks0 - e.ks:e.ke0 - e.ks + 1, e.js0 - e.js:e.je0 - e.js + 1, e.is0 - e.is_:e.ie0 - e.is_ + 1]
    assert np.allclose(spci[ph], h[0], rtol=1e-14) and np.allclose(spcj[ph], h[1], rtol=1e-14)
    assert np.allclose(spck[ph], h[2], rtol=1e-14) and np.allclose(spacing[ph], max(h), rtol=1e-14)
    assert np.allclose(vol[ph], h[0] * h[1] * h[2], rtol=1e-13)
    wall = O.wall_points([e], [subs], [(x, y, z)])
    assert np.allclose(wall[1], 0.0, atol=1e-15)
    d = O.walldist(e, x, y, z, wall)
    assert np.allclose(d[ph], np.abs(y[ph]), rtol=1e-14, atol=1e-15)


def test_walldist_brute_force_on_warped_grid():
    dims, subs, ranges, xyz = cases.warped_grid(seed=5, phys=(-20, -40, -20, -10, -30, -20))
    exts = [O.Ext.from_list(r) for r in ranges]
    wall = O.wall_points(exts, subs, xyz)
    assert wall.shape[1] > 50
    for e, (x, y, z) in zip(exts, xyz):
        d = O.walldist(e, x, y, z, wall)
        ph = np.s_[e.ks0 - e.ks:e.ke0 - e.ks + 1, e.js0 - e.js:e.je0 - e.js + 1, e.is0 - e.is_:e.ie0 - e.is_ + 1]
        p = np.stack([x[ph].ravel(), y[ph].ravel(), z[ph].ravel()], 1)
        ref = np.sqrt(((p[:, None, :] - wall.T[None, :, :]) ** 2).sum(-1)).min(1)
        assert np.allclose(d[ph].ravel(), ref, rtol=1e-14, atol=1e-15)
