"""Synthetic multi-block cases for the residual-and-update path.

Blocks use the reference's overlap convention: neighbouring blocks share their
interface node layer (node ni of one block is node 1 of the next), and each
interface face carries two halo layers filled from the neighbour's nodes 2 and
3 (sub_bc_updatebuffer.f90:57-135).  Subface records are the 11 integers of a
bcinfo line (sub_init_readbcinfo.f90:96-98): f_no, face, is, ie, js, je, ks,
ke, blk_t, subface_t, ori.  The orientation code ``ori`` of every interface is
found by matching node identities through the same face pairing the halo
exchange uses, so a generated topology is geometrically consistent by
construction.

Cases (the configurations the measurements are quoted on):
  * ``vortex_box``  -- single block on [-5,5]^3, farfield on all faces,
    isentropic vortex on a uniform freestream (Euler).
  * ``tgv_box``     -- single block on [0,2 pi]^3 with node n == node 1 on every
    face pair (self-connected interfaces), Taylor-Green vortex (laminar NS).
  * ``split_box``   -- the same kinds of state on a box cut into several
    blocks with interfaces of every orientation.
"""
import math

import numpy as np

GAMMA = 1.4

FACE_AXIS = {1: 0, 2: 0, 3: 1, 4: 1, 5: 2, 6: 2}


# ------------------------------------------------------------ face pairing
def _pair_nodes(subs, dims, recv, send):
    """Sender face node of every receiving face node for the given
    receiving/sending subface records (face_t load, ori permutation, face
    store of sub_bc_updatebuffer.f90:261-539)."""
    f, t = recv, send
    n11, n12, n13 = t[3] - t[2] + 1, t[5] - t[4] + 1, t[7] - t[6] + 1
    l1, m1 = {1: (n13, n12), 2: (n12, n13), 3: (n11, n13), 4: (n13, n11), 5: (n12, n11), 6: (n11, n12)}[t[1]]

    def ua(ll, mm):
        ts, te, js, je, ks, ke = t[2], t[3], t[4], t[5], t[6], t[7]
        return {1: (ts, js + mm - 1, ks + ll - 1), 2: (te, js + ll - 1, ks + mm - 1),
                3: (ts + ll - 1, js, ks + mm - 1), 4: (ts + mm - 1, je, ks + ll - 1),
                5: (ts + mm - 1, js + ll - 1, ks), 6: (ts + ll - 1, js + mm - 1, ke)}[t[1]]

    ori = f[10]
    tr = ori in (2, 4)
    l2, m2 = (m1, l1) if tr else (l1, m1)
    n1, n2, n3 = f[3] - f[2] + 1, f[5] - f[4] + 1, f[7] - f[6] + 1
    ra, rb = {1: (n3, n2), 2: (n2, n3), 3: (n1, n3), 4: (n3, n1), 5: (n2, n1), 6: (n1, n2)}[f[1]]
    if (ra, rb) != (l2, m2):
        return None
    out = []
    for b in range(1, m2 + 1):
        for a in range(1, l2 + 1):
            ll, mm = {1: (a, m1 - b + 1), 2: (b, a), 3: (l1 - a + 1, b), 4: (l1 - b + 1, m1 - a + 1)}[ori]
            s = ua(ll, mm)
            fs, fe, gjs, gje, gks, gke = f[2], f[3], f[4], f[5], f[6], f[7]
            r = {1: (fs, gjs + b - 1, gks + a - 1), 2: (fe, gjs + a - 1, gks + b - 1),
                 3: (fs + a - 1, gjs, gks + b - 1), 4: (fs + b - 1, gje, gks + a - 1),
                 5: (fs + b - 1, gjs + a - 1, gks), 6: (fs + a - 1, gjs + b - 1, gke)}[f[1]]
            out.append((r, s))
    return out


def assign_orientations(dims, subfaces, gid):
    """Choose ori of every interface so that receiving and sending face
    nodes are the same node. gid(m, i, j, k) -> hashable global identity of
    node (i,j,k) of block m (0-based m)."""
    for m, subs in enumerate(subfaces):
        for s in subs:
            if s[8] <= 0:
                continue
            t = subfaces[s[8] - 1][s[9] - 1]
            for ori in (1, 2, 3, 4):
                s[10] = ori
                pn = _pair_nodes(subfaces, dims, s, t)
                if pn is None:
                    continue
                if all(gid(m, *r) == gid(s[8] - 1, *sn) for r, sn in pn):
                    break
            else:
                raise ValueError("no orientation matches interface %s of block %d" % (s, m + 1))
    return subfaces


def _face_range(face, ni, nj, nk):
    return {1: (1, 1, 1, nj, 1, nk), 2: (ni, ni, 1, nj, 1, nk), 3: (1, ni, 1, 1, 1, nk),
            4: (1, ni, nj, nj, 1, nk), 5: (1, ni, 1, nj, 1, 1), 6: (1, ni, 1, nj, nk, nk)}[face]


def box_subfaces(ni, nj, nk, kinds):
    """Six whole-face subfaces; kinds[f-1] = blk_t for face f (physical codes
    < 0) or (blk_t, subface_t) for an interface."""
    subs = []
    for f in range(1, 7):
        kd = kinds[f - 1]
        bt, st = (kd, 0) if isinstance(kd, int) else kd
        subs.append([f, f, *_face_range(f, ni, nj, nk), bt, st, 0])
    return subs


def split_box(NI, NJ, NK, cuts=((), (), ()), phys=(-10,) * 6, periodic=(False, False, False),
              permute=None, flip=None):
    """Cut a box of NI x NJ x NK nodes into blocks at the given node indices
    (1-based global, shared by both sides).  phys: physical code per box face
    where not periodic.  permute: optional per-block axis permutation
    (a tuple of 3: local axis a carries global axis permute[x][a]) applied
    to the block's local axes; flip: optional per-block tuple of 3 booleans,
    local axis a running against its global axis.  Permutations and flips
    whose frame stays right-handed give interfaces of every orientation code.
    Returns dims, subfaces, origins: per block the global index of local
    node (1,1,1) and the axis map, plus the axis signs (+1 / -1) when flip
    is given."""
    edges = []
    for ax, n in enumerate((NI, NJ, NK)):
        e = [1] + sorted(cuts[ax]) + [n]
        edges.append(e)
    nb = [len(e) - 1 for e in edges]
    blocks = []
    for bk in range(nb[2]):
        for bj in range(nb[1]):
            for bi in range(nb[0]):
                lo = (edges[0][bi], edges[1][bj], edges[2][bk])
                hi = (edges[0][bi + 1], edges[1][bj + 1], edges[2][bk + 1])
                blocks.append((bi, bj, bk, lo, hi))
    bid = {(b[0], b[1], b[2]): x for x, b in enumerate(blocks)}
    perms = []
    signs = []
    dims = []
    for x, (bi, bj, bk, lo, hi) in enumerate(blocks):
        p = tuple(permute[x]) if permute else (0, 1, 2)
        sg = tuple(-1 if f else 1 for f in flip[x]) if flip else (1, 1, 1)
        odd = sum(p[a] > p[b] for a in range(3) for b in range(a + 1, 3)) % 2
        if (odd + sum(v < 0 for v in sg)) % 2:
            raise ValueError("block %d: axis map %s with signs %s is left-handed" % (x + 1, p, sg))
        perms.append(p)
        signs.append(sg)
        ext = [hi[a] - lo[a] + 1 for a in range(3)]
        dims.append(tuple(ext[p[a]] for a in range(3)))
    # face of block x (global axis ax, side sd) in local face numbering
    def local_face(x, ax, sd):
        la = perms[x].index(ax)
        return 2 * la + 1 + (sd if signs[x][la] > 0 else 1 - sd)
    subfaces = [[None] * 6 for _ in blocks]
    for x, (bi, bj, bk, lo, hi) in enumerate(blocks):
        idx = (bi, bj, bk)
        for ax in range(3):
            for sd in (0, 1):
                nbr = list(idx)
                nbr[ax] += -1 if sd == 0 else 1
                per = False
                if nbr[ax] < 0 or nbr[ax] >= nb[ax]:
                    if periodic[ax]:
                        nbr[ax] %= nb[ax]
                        per = True
                    else:
                        nbr = None
                f = local_face(x, ax, sd)
                ni, nj, nk = dims[x]
                rng = _face_range(f, ni, nj, nk)
                if nbr is None:
                    subfaces[x][f - 1] = [f, f, *rng, phys[2 * ax + sd], 0, 0]
                else:
                    y = bid[tuple(nbr)]
                    fy = local_face(y, ax, 1 - sd)
                    subfaces[x][f - 1] = [f, f, *rng, y + 1, fy, 0]
    periods = (NI - 1, NJ - 1, NK - 1)

    def first(m):
        # global index of local node (1,1,1)
        lo, hi = blocks[m][3], blocks[m][4]
        return tuple(hi[a] if signs[m][perms[m].index(a)] < 0 else lo[a] for a in range(3))

    def gid(m, i, j, k):
        loc = (i, j, k)
        p = perms[m]
        o = first(m)
        g = [0, 0, 0]
        for a in range(3):
            g[p[a]] = o[p[a]] + signs[m][a] * (loc[a] - 1)
        return tuple(g[a] % periods[a] if periodic[a] else g[a] for a in range(3))

    assign_orientations(dims, subfaces, gid)
    if flip:
        origins = [(first(x), perms[x], signs[x]) for x in range(len(blocks))]
    else:
        origins = [(b[3], perms[x]) for x, b in enumerate(blocks)]
    return dims, subfaces, origins


# ---------------------------------------------------------------- geometry
def ranges_of(dims_m, subs_m):
    """The 18 index ranges of sub_init_allocatememory.f90:73-144."""
    ni, nj, nk = dims_m
    r0 = [1, ni, 1, nj, 1, nk]
    r = list(r0)
    for s in subs_m:
        if s[8] > 0:
            f = s[1]
            if f in (1, 3, 5):
                r[f - 1] -= 2
            else:
                r[f - 1] += 2
    r1 = [r0[a] + (1 if a % 2 == 0 else -1) if r0[a] == r[a] else r0[a] for a in range(6)]
    return tuple(r + r0 + r1)


def node_coords(ranges, origin, spacing, x0, perm=(0, 1, 2), sign=(1, 1, 1)):
    """Cartesian coordinates (3, nk, nj, ni) of a uniform-grid block whose
    local node (1,1,1) is global node ``origin`` (local axis a runs along
    global axis perm[a] with sign[a]); halo nodes extrapolate the lattice
    (the values the coordinate exchange delivers on a uniform grid with
    periodic shifts)."""
    is_, ie, js, je, ks, ke = ranges[:6]
    li = np.arange(is_, ie + 1)
    lj = np.arange(js, je + 1)
    lk = np.arange(ks, ke + 1)
    K, J, I = np.meshgrid(lk, lj, li, indexing="ij")
    loc = (I, J, K)
    xyz = np.zeros((3,) + I.shape)
    for a in range(3):
        ga = perm.index(a)  # local axis that carries global axis a
        xyz[a] = x0[a] + spacing[a] * (origin[a] - 1 + sign[ga] * (loc[ga] - 1))
    return xyz


def uniform_metrics(ranges, spacing, perm=(0, 1, 2), sign=(1, 1, 1)):
    """Metrics of a uniform Cartesian block: dxidx(3(d-1)+c) = d xi_d/d x_c,
    inv_j = 1/J, and alpha/beta/gamma(c) = (grad xi_d . grad xi_c)/J -- the
    values gridtransformation (sub_init_gridtransformation.f90:27-874) returns
    on such a grid up to rounding."""
    is_, ie, js, je, ks, ke = ranges[:6]
    shp = (ke - ks + 1, je - js + 1, ie - is_ + 1)
    h = [spacing[perm[a]] for a in range(3)]  # local-axis spacings
    vol = h[0] * h[1] * h[2]
    dxidx = np.zeros(shp + (9,))
    for d in range(3):
        dxidx[..., 3 * d + perm[d]] = sign[d] / h[d]
    inv_j = np.full(shp, vol)
    abg = []
    for d in range(3):
        a = np.zeros(shp + (3,))
        a[..., d] = vol / (h[d] * h[d])
        abg.append(a)
    return dxidx, inv_j, abg[0], abg[1], abg[2]


# ------------------------------------------------------- initial conditions
def conservative(rho, u, v, w, p):
    q = np.zeros(rho.shape + (5,))
    q[..., 0] = rho
    q[..., 1] = rho * u
    q[..., 2] = rho * v
    q[..., 3] = rho * w
    q[..., 4] = p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return q


def vortex_state(xyz, ma, beta=5.0, center=(0.0, 0.0)):
    """Isentropic vortex (axis along z) on a uniform freestream u = 1 in the
    reference's scaling (rho_inf = u_inf = T_inf = 1, p_inf = 1/(gamma ma^2)):
    velocity perturbation scaled by c_inf/sqrt(gamma)."""
    x = xyz[0] - center[0]
    y = xyz[1] - center[1]
    r2 = x * x + y * y
    cinf = 1.0 / ma
    s = cinf / math.sqrt(GAMMA)
    du = -beta / (2 * math.pi) * np.exp(0.5 * (1 - r2)) * y * s
    dv = beta / (2 * math.pi) * np.exp(0.5 * (1 - r2)) * x * s
    dT = -(GAMMA - 1.0) * beta * beta / (8 * GAMMA * math.pi ** 2) * np.exp(1 - r2)
    rho = (1.0 + dT) ** (1.0 / (GAMMA - 1.0))
    p = 1.0 / (GAMMA * ma * ma) * rho ** GAMMA
    return conservative(rho, 1.0 + du, dv, np.zeros_like(x), p)


def tgv_state(xyz, ma):
    """Taylor-Green vortex: u = sin x cos y cos z, v = -cos x sin y cos z,
    w = 0, p = p_inf + (cos 2x + cos 2y)(cos 2z + 2)/16, isothermal T = 1
    (rho = gamma ma^2 p)."""
    x, y, z = xyz
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(x)
    p = 1.0 / (GAMMA * ma * ma) + (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0) / 16.0
    rho = GAMMA * ma * ma * p
    return conservative(rho, u, v, w, p)


# ------------------------------------------------------------------- cases
class Case:
    """Topology plus per-block input fields, ready for Solver.upload."""

    def __init__(self, name, cfg, dims, subfaces):
        self.name = name
        self.cfg = cfg
        self.dims = dims
        self.subfaces = subfaces
        self.ranges = [ranges_of(d, s) for d, s in zip(dims, subfaces)]
        self.fields = [dict() for _ in dims]

    @property
    def totalnodes(self):
        return int(sum(a * b * c for a, b, c in self.dims))

    def shape(self, m):
        r = self.ranges[m]
        return (r[5] - r[4] + 1, r[3] - r[2] + 1, r[1] - r[0] + 1)

    def upload(self, solver):
        for m, fl in enumerate(self.fields):
            for k, v in fl.items():
                solver.upload(m + 1, k, v)


def _common_fields(case, m, q, dxidx, inv_j, al, be, ga):
    shp = case.shape(m)
    fl = case.fields[m]
    fl["q"] = q
    fl["dxidx"] = dxidx
    fl["inv_j"] = inv_j
    amu = np.zeros(shp + (3,))
    amu[..., 0] = 1.0
    amu[..., 2] = 1.0
    fl["amu"] = amu
    fl["dt"] = np.full(shp, case.cfg.dt)
    if case.cfg.iflag_turbulence == 2:
        fl["aalpha"], fl["bbeta"], fl["ggamma"] = al, be, ga


def vortex_box(n=64, cfg=None, blocks=((), (), ()), permute=None, flip=None):
    """Single (or cut) block n^3 on [-5,5]^3, farfield everywhere, isentropic
    vortex; Euler, first-order implicit, dt = 0.5*(10/(n-1))/(1+2)."""
    from .solver import FlowConfig
    h = 10.0 / (n - 1)
    if cfg is None:
        cfg = FlowConfig(ma=0.5, re=1.0e6, iflag_solver=1, iflag_time=1, iflag_timeadvance=2,
                         dt=0.5 * h / (1.0 + 2.0))
    dims, subs, origins = split_box(n, n, n, blocks, phys=(-10,) * 6, permute=permute, flip=flip)
    case = Case("vortex_box_%d" % n, cfg, dims, subs)
    case.coords = []
    for m in range(len(dims)):
        org, perm, sign = (origins[m] + ((1, 1, 1),))[:3]
        xyz = node_coords(case.ranges[m], org, (h, h, h), (-5.0, -5.0, -5.0), perm, sign)
        case.coords.append((xyz[0], xyz[1], xyz[2]))
        q = vortex_state(xyz, cfg.ma)
        _common_fields(case, m, q, *uniform_metrics(case.ranges[m], (h, h, h), perm, sign))
    return case


def tgv_box(n=256, cfg=None, blocks=((), (), ()), permute=None, flip=None):
    """n^3 nodes on [0,2 pi]^3, node n == node 1 on every face pair (all
    faces self-connected or connected through the cut blocks), Taylor-Green
    vortex at Mach 0.1, Re = 1600, laminar NS, first-order implicit."""
    from .solver import FlowConfig
    h = 2.0 * math.pi / (n - 1)
    if cfg is None:
        cfg = FlowConfig(ma=0.1, re=1600.0, iflag_solver=0, iflag_time=1, iflag_timeadvance=2,
                         dt=0.5 * h / (1.0 + 10.0), prl=0.72)
    dims, subs, origins = split_box(n, n, n, blocks, periodic=(True, True, True), permute=permute, flip=flip)
    case = Case("tgv_box_%d" % n, cfg, dims, subs)
    for m in range(len(dims)):
        org, perm, sign = (origins[m] + ((1, 1, 1),))[:3]
        xyz = node_coords(case.ranges[m], org, (h, h, h), (0.0, 0.0, 0.0), perm, sign)
        q = tgv_state(xyz, cfg.ma)
        _common_fields(case, m, q, *uniform_metrics(case.ranges[m], (h, h, h), perm, sign))
    return case


def shock_vortex(ni=96, nj=48, nk=4, mach=1.5, seed=4, cfg=None):
    """Planar normal shock at x = 1 (Rankine-Hugoniot jump from the Mach
    `mach` freestream moving along +x) hit by an isentropic vortex, with
    entropy waves rho' = 0.2 sin(5 x) upstream; x in [0, 2], uniform spacing,
    extruded in k.  Faces: inflow (i-), outflow (i+), symmetry (j, k).  Euler,
    unsteady, explicit RK3, WCNS.  The vortex position comes from `seed`."""
    from .solver import FlowConfig
    h = 2.0 / (ni - 1)
    if cfg is None:
        cfg = FlowConfig(ma=mach, re=1.0e6, iflag_solver=1, iflag_time=1, iflag_timeadvance=4,
                         inviscid_scheme=2, dt=0.25 * h / (1.0 + 1.0 / mach))
    subs = box_subfaces(ni, nj, nk, [-50, -40, -30, -30, -30, -30])
    case = Case("shock_vortex_%dx%dx%d" % (ni, nj, nk), cfg, [(ni, nj, nk)], [subs])
    rng = case.ranges[0]
    xyz = node_coords(rng, (1, 1, 1), (h, h, h), (0.0, 0.0, 0.0))
    x, y = xyz[0], xyz[1]
    g = GAMMA
    p_inf = 1.0 / (g * mach * mach)
    t_inf = p_inf
    r = np.random.default_rng(seed)
    ly = (nj - 1) * h
    xc, yc, rad, gam = 0.45 + 0.1 * r.random(), ly * (0.4 + 0.2 * r.random()), 0.12, 0.25
    dx, dy = x - xc, y - yc
    rr = (dx * dx + dy * dy) / (rad * rad)
    ex = np.exp(0.5 * (1.0 - rr))
    du = -gam * dy / rad * ex / (2 * math.pi)
    dv = gam * dx / rad * ex / (2 * math.pi)
    dT = -(g - 1.0) * gam * gam * ex * ex / (8.0 * g * math.pi * math.pi)
    T = t_inf + dT
    rho = (T / t_inf) ** (1.0 / (g - 1.0)) * (1.0 + 0.2 * np.sin(5.0 * x))
    p = p_inf * (T / t_inf) ** (g / (g - 1.0))
    u = 1.0 + du
    v = dv
    # post-shock state for x >= 1
    m2 = mach * mach
    r21 = (g + 1.0) * m2 / ((g - 1.0) * m2 + 2.0)
    p21 = 1.0 + 2.0 * g / (g + 1.0) * (m2 - 1.0)
    down = x >= 1.0
    rho = np.where(down, r21, rho)
    u = np.where(down, 1.0 / r21, u)
    v = np.where(down, 0.0, v)
    p = np.where(down, p_inf * p21, p)
    q = conservative(rho, u, v, np.zeros_like(rho), p)
    dxidx, inv_j, al, be, ga = uniform_metrics(rng, (h, h, h))
    _common_fields(case, 0, q, dxidx, inv_j, al, be, ga)
    return case


NUT_INF = 1.341946  # freestream SA working variable (module_tursa.f90:11)


def tanh_stretch(n, height, first):
    """Node heights y_1..y_n of the one-sided tanh clustering
    y = H (1 - tanh(d (1 - s)) / tanh d), s = (j - 1)/(n - 1), with d chosen
    (bisection) so that y_2 - y_1 = first.  Returns (d, y(s), dy/dj(s))."""
    def first_of(d):
        s = 1.0 / (n - 1)
        return height * (1.0 - math.tanh(d * (1.0 - s)) / math.tanh(d))
    lo, hi = 1e-6, 30.0
    if first_of(hi) > first:
        lo = hi
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if first_of(mid) > first:
            lo = mid
        else:
            hi = mid
    d = 0.5 * (lo + hi)

    def y(s):
        return height * (1.0 - np.tanh(d * (1.0 - s)) / math.tanh(d))

    def dyds(s):
        t = np.tanh(d * (1.0 - s))
        return height * d * (1.0 - t * t) / math.tanh(d) / (n - 1)
    return d, y, dyds


def flat_plate(seed=3, nblk=(4, 3, 1), lo=(32, 32, 32), hi=(160, 96, 64), length=(1.0, 0.25, 0.1),
               first=5.4e-6, cfg=None):
    """Batched multi-block turbulent flat plate: a box of nblk[0] x nblk[1] x
    nblk[2] point-matched blocks whose node counts per axis are drawn from a
    seeded generator in [lo, hi]; x uniform over the plate length (inflow at
    x = 0, outflow at x = length[0]), y tanh-clustered to the wall y = 0 with
    first spacing `first` (y+ ~ 1 at Re_L = 5e6), z uniform (symmetry faces).
    Mach 0.3, Re_L = 5e6, SA, local time step, first-order implicit; the
    reference's native uniform start (sub_init_flowfield.f90:69-110).
    Metrics are the exact ones of the separable mapping; wall distance is y;
    spacing follows sub_init_spacing.f90 (largest forward edge length on the
    physical range, last node copied, zero on halos)."""
    from .solver import FlowConfig
    rng = np.random.default_rng(seed)
    sizes = [[int(rng.integers(lo[a], hi[a] + 1)) for _ in range(nblk[a])] for a in range(3)]
    cuts, tot = [], []
    for a in range(3):
        c, pos = [], 1
        for s in sizes[a][:-1]:
            pos += s - 1
            c.append(pos)
        cuts.append(tuple(c))
        tot.append(sum(sizes[a]) - (nblk[a] - 1))
    NI, NJ, NK = tot
    if cfg is None:
        cfg = FlowConfig(ma=0.3, re=5.0e6, iflag_solver=0, iflag_turbulence=2, iflag_des=0, iflag_time=0,
                         iflag_timeadvance=2, cfl=5.0, dt=1.0e-3, prl=0.72)
    dims, subs, origins = split_box(NI, NJ, NK, cuts, phys=(-50, -40, -20, -10, -30, -30))
    case = Case("flat_plate_%dblk_%dx%dx%d" % (len(dims), NI, NJ, NK), cfg, dims, subs)
    case.coords = []  # (x, y, z) per block over the extended extents
    hx, hz = length[0] / (NI - 1), length[2] / (NK - 1)
    _, ymap, dymap = tanh_stretch(NJ, length[1], first)
    g = GAMMA
    cv = 1.0 / g / (g - 1.0) / cfg.ma ** 2
    aoa = math.radians(cfg.aoa_deg)
    for m in range(len(dims)):
        r = case.ranges[m]
        org, _ = origins[m]
        shp = case.shape(m)
        # global node indices of the extended range (halos lie inside the box)
        gi = org[0] + np.arange(r[0], r[1] + 1) - 1
        gj = org[1] + np.arange(r[2], r[3] + 1) - 1
        gk = org[2] + np.arange(r[4], r[5] + 1) - 1
        K, J, I = np.meshgrid(gk, gj, gi, indexing="ij")
        sj = (J - 1.0) / (NJ - 1)
        x = (I - 1.0) * hx
        y = ymap(sj)
        z = (K - 1.0) * hz
        case.coords.append((x, y, z))
        yj = dymap(sj)
        dxidx = np.zeros(shp + (9,))
        dxidx[..., 0] = 1.0 / hx
        dxidx[..., 4] = 1.0 / yj
        dxidx[..., 8] = 1.0 / hz
        inv_j = hx * yj * hz
        al = np.zeros(shp + (3,))
        be = np.zeros(shp + (3,))
        ga = np.zeros(shp + (3,))
        al[..., 0] = inv_j / (hx * hx)
        be[..., 1] = inv_j / (yj * yj)
        ga[..., 2] = inv_j / (hz * hz)
        q = np.zeros(shp + (5,))
        q[..., 0] = 1.0
        q[..., 1] = math.cos(aoa)
        q[..., 2] = math.sin(aoa)
        q[..., 4] = cv + 0.5
        _common_fields(case, m, q, dxidx, inv_j, al, be, ga)
        fl = case.fields[m]
        amu = fl["amu"]
        amu[..., 0] = 1.0
        amu[..., 1] = 0.009
        amu[..., 2] = 1.009
        fl["nut"] = np.full(shp, NUT_INF)
        fl["dist"] = y.copy()
        # sub_init_spacing.f90:39-104 on the physical range
        sp = np.zeros(shp)
        p0 = (r[6] - r[0], r[7] - r[0] + 1, r[8] - r[2], r[9] - r[2] + 1, r[10] - r[4], r[11] - r[4] + 1)
        xs, ys, zs = x[p0[4]:p0[5], p0[2]:p0[3], p0[0]:p0[1]], y[p0[4]:p0[5], p0[2]:p0[3], p0[0]:p0[1]], \
            z[p0[4]:p0[5], p0[2]:p0[3], p0[0]:p0[1]]
        spc = []
        for ax in (2, 1, 0):  # i, j, k edges (arrays are k, j, i)
            d = [np.diff(c, axis=ax) for c in (xs, ys, zs)]
            e = np.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
            last = np.take(e, [-1], axis=ax)
            spc.append(np.concatenate([e, last], axis=ax))
        sp[p0[4]:p0[5], p0[2]:p0[3], p0[0]:p0[1]] = np.maximum(np.maximum(spc[0], spc[1]), spc[2])
        fl["spacing"] = sp
    return case


def channel_bump(seed=5, nblk=(5, 4, 2), lo=(160, 110, 90), hi=(220, 150, 120), length=(4.0, 1.0, 0.5),
                 bump=0.05, first=2.0e-5, cfg=None):
    """The combined workload: a channel whose lower wall carries a Gaussian
    bump of height `bump` (chord 1, y_w = bump exp(-(x - x_c)^2 / 0.06),
    x_c = length[0] / 2), cut into nblk[0] x nblk[1] x nblk[2] point-matched
    blocks whose node counts per axis are drawn from a seeded generator in
    [lo, hi] (40 blocks, ~100 M nodes by default).  x uniform, y tanh-
    clustered from the wall (first spacing `first`) to the flat top at
    y = length[1], z uniform.  Faces: inflow (i-), outflow (i+), no-slip
    wall (j-), symmetry (j+, k).  Mach 0.6, Re = 1e6, SA with DES97, local
    time step, first-order implicit, uniform start.  Only coordinates are
    given (over the extended extents, halos from the same mapping): the
    metrics, edge lengths and wall distance come from acfd_grid_metrics.
    Fields: q, amu, nut, dt."""
    from .solver import FlowConfig
    rng = np.random.default_rng(seed)
    sizes = [[int(rng.integers(lo[a], hi[a] + 1)) for _ in range(nblk[a])] for a in range(3)]
    cuts, tot = [], []
    for a in range(3):
        c, pos = [], 1
        for sz in sizes[a][:-1]:
            pos += sz - 1
            c.append(pos)
        cuts.append(tuple(c))
        tot.append(sum(sizes[a]) - (nblk[a] - 1))
    NI, NJ, NK = tot
    if cfg is None:
        cfg = FlowConfig(ma=0.6, re=1.0e6, iflag_solver=0, iflag_turbulence=2, iflag_des=1, iflag_time=0,
                         iflag_timeadvance=2, cfl=5.0, dt=1.0e-3, prl=0.72)
    dims, subs, origins = split_box(NI, NJ, NK, cuts, phys=(-50, -40, -20, -30, -30, -30))
    case = Case("channel_bump_%dblk_%dx%dx%d" % (len(dims), NI, NJ, NK), cfg, dims, subs)
    case.coords = []
    hx, hz = length[0] / (NI - 1), length[2] / (NK - 1)
    _, ymap, _ = tanh_stretch(NJ, 1.0, first / length[1])
    xc = 0.5 * length[0]
    g = GAMMA
    cv = 1.0 / g / (g - 1.0) / cfg.ma ** 2
    aoa = math.radians(cfg.aoa_deg)
    for m in range(len(dims)):
        r = case.ranges[m]
        org, _ = origins[m]
        shp = case.shape(m)
        gi = org[0] + np.arange(r[0], r[1] + 1) - 1
        gj = org[1] + np.arange(r[2], r[3] + 1) - 1
        gk = org[2] + np.arange(r[4], r[5] + 1) - 1
        K, J, I = np.meshgrid(gk, gj, gi, indexing="ij")
        x = (I - 1.0) * hx
        yw = bump * np.exp(-(x - xc) ** 2 / 0.06)
        y = yw + (length[1] - yw) * ymap((J - 1.0) / (NJ - 1))
        z = (K - 1.0) * hz
        case.coords.append((x, y, z))
        q = np.zeros(shp + (5,))
        q[..., 0] = 1.0
        q[..., 1] = math.cos(aoa)
        q[..., 2] = math.sin(aoa)
        q[..., 4] = cv + 0.5
        fl = case.fields[m]
        fl["q"] = q
        amu = np.zeros(shp + (3,))
        amu[..., 0] = 1.0
        amu[..., 1] = 0.009
        amu[..., 2] = 1.009
        fl["amu"] = amu
        fl["nut"] = np.full(shp, NUT_INF)
        fl["dt"] = np.zeros(shp)
    return case


def warped_grid(NI=14, NJ=12, NK=10, cuts=((7,), (6,), ()), phys=(-10, -40, -20, -10, -30, -30), amp=0.08,
                seed=0):
    """A curvilinear multi-block grid: the box [0,1]^3 (NI x NJ x NK nodes)
    cut into point-matched blocks and warped by a smooth seeded mapping
    x + amp * sum of products of sines, so that every metric term is
    non-trivial.  Halo nodes of interface faces carry the neighbour's
    coordinates (what the coordinate exchange delivers).  Returns dims,
    subfaces, the 18 ranges per block and (x, y, z) per block over the
    extended extents."""
    rng = np.random.default_rng(seed)
    dims, subs, origins = split_box(NI, NJ, NK, cuts, phys=phys)
    ph = rng.uniform(0, 2 * math.pi, (3, 3))
    fr = rng.uniform(0.5, 1.5, (3, 3))

    def mapping(gi, gj, gk):
        s = [(gi - 1.0) / (NI - 1), (gj - 1.0) / (NJ - 1), (gk - 1.0) / (NK - 1)]
        out = []
        for a in range(3):
            w = np.ones_like(s[0])
            for b in range(3):
                if b != a:
                    w = w * np.sin(math.pi * fr[a][b] * s[b] + ph[a][b])
            out.append(s[a] + amp * w * np.sin(math.pi * s[a]) + 0.3 * amp * s[(a + 1) % 3])
        return out

    ranges, coords = [], []
    for m in range(len(dims)):
        r = ranges_of(dims[m], subs[m])
        org, _ = origins[m]
        gi = org[0] + np.arange(r[0], r[1] + 1) - 1
        gj = org[1] + np.arange(r[2], r[3] + 1) - 1
        gk = org[2] + np.arange(r[4], r[5] + 1) - 1
        K, J, I = np.meshgrid(gk, gj, gi, indexing="ij")
        coords.append(tuple(np.ascontiguousarray(c) for c in mapping(I.astype(float), J.astype(float),
                                                                      K.astype(float))))
        ranges.append(r)
    return dims, subs, ranges, coords
