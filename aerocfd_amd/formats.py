"""The reference's on-disk formats, read and written on the host.

  Control.in      sub_init_readcontrolfile.f90:27-47 (six data lines, each
                  after two header lines, Fortran list-directed)
  grid            sub_init_allocatememory.f90:36-37, 64-71, 388-417 (nblock;
                  ni nj nk per block; per block x, y, z, i fastest, scaled
                  by scale_*/lref on reading)
  bcinfo          sub_init_readbcinfo.f90:45-104 (three header lines; per
                  block two lines, the subface count, one line, the
                  subface records f_no face is ie js je ks ke blk_t
                  subface_t ori fwh, one closing line)
  flowfield       sub_output_flowfield.f90:30-81 (writer) and
                  sub_init_flowfield.f90:126-262 (restart reader): per
                  block a variables line, a zone line, then one record per
                  node x y z d u v w p nut over the physical range (a
                  record may wrap over two lines)
  resi_history    sub_output_residual.f90:62-64 (I6.6, 6E11.4)
  clcd_history    sub_output_forceintegration.f90:472-475 (time, cl, cd)

Values are read with the rules of Fortran list-directed input: separators
are blanks and commas, ``r*c`` repeats a value, a slash ends the record,
``d``/``D`` exponents are accepted, strings may be quoted or bare.
"""
import math
import re

import numpy as np

_TOKEN = re.compile(r"""\s*(?:'((?:[^']|'')*)'|"((?:[^"]|"")*)"|([^\s,/]+)|(/))\s*,?""")


def list_directed(line):
    """Tokens of one record of list-directed input (strings unquoted,
    repeat counts expanded; everything after a slash dropped)."""
    out = []
    pos = 0
    line = line.rstrip("\n")
    while pos < len(line):
        m = _TOKEN.match(line, pos)
        if not m or m.end() == pos:
            break
        pos = m.end()
        if m.group(4):
            break
        if m.group(1) is not None:
            out.append(m.group(1).replace("''", "'"))
        elif m.group(2) is not None:
            out.append(m.group(2).replace('""', '"'))
        else:
            tok = m.group(3)
            rep = re.fullmatch(r"(\d+)\*(.*)", tok)
            if rep:
                out.extend([rep.group(2)] * int(rep.group(1)))
            else:
                out.append(tok)
    return out


def _real(tok):
    return float(tok.replace("d", "e").replace("D", "E"))


def _int(tok):
    return int(tok)


# Control.in: (names, types) of the six data lines, in order
CONTROL_LINES = [
    (("iflag_dimension", "i"), ("iflag_solver", "i"), ("iflag_time", "i"), ("cfl", "r"), ("dt", "r")),
    (("ma", "r"), ("re", "r"), ("tinf", "r"), ("aoa", "r")),
    (("scale_x", "r"), ("scale_y", "r"), ("scale_z", "r"), ("lref", "r"), ("sref", "r"), ("fltr", "r")),
    (("iflag_splittingtype", "i"), ("iflag_inviscid", "i"), ("iflag_filter", "i"), ("iflag_timeadvance", "i"),
     ("iflag_turbulence", "i"), ("iflag_des", "i")),
    (("writeinterval", "i"), ("ntstep", "i"), ("grid", "s"), ("bcinfo", "s"), ("iflag_init", "i"),
     ("restartfile", "s"), ("nfile", "i"), ("restart_timestep", "i")),
    (("nobs", "i"), ("ob_dist", "r"), ("acoustic_length", "r"), ("iflag_acoustic", "i"), ("Start_Acsoutic", "i"),
     ("End_Acsoutic", "i"), ("iflag_avg", "i"), ("Start_Avg", "i"), ("End_Avg", "i"), ("Avg_interval", "i")),
]


def read_control(path):
    """Control.in -> dict of the 39 values (sub_init_readcontrolfile.f90:27-47)."""
    with open(path) as fh:
        lines = fh.readlines()
    out = {}
    for n, fields in enumerate(CONTROL_LINES):
        toks = list_directed(lines[3 * n + 2])
        if len(toks) < len(fields):
            raise ValueError("%s: line %d has %d values, %d expected" % (path, 3 * n + 3, len(toks), len(fields)))
        for (name, kind), tok in zip(fields, toks):
            out[name] = _int(tok) if kind == "i" else _real(tok) if kind == "r" else tok
    return out


def write_control(path, values, titles=None):
    """Write a Control.in that read_control and the reference read back
    as `values` (two header lines before every data line)."""
    with open(path, "w") as fh:
        for n, fields in enumerate(CONTROL_LINES):
            t = titles[n] if titles else ("line%d" % (n + 1))
            fh.write("%s\n%s\n" % (t, " ".join(name for name, _ in fields)))
            vals = []
            for name, kind in fields:
                v = values[name]
                vals.append(str(int(v)) if kind == "i" else repr(float(v)) if kind == "r" else "'%s'" % v)
            fh.write(" ".join(vals) + "\n")


def flow_config(ctl, **over):
    """A FlowConfig from Control.in values (the path's flags and constants)."""
    from .solver import FlowConfig
    if ctl["iflag_dimension"] != 3:
        raise ValueError("only 3-D grids are supported")
    cfg = FlowConfig(ma=ctl["ma"], re=ctl["re"], tinf=ctl["tinf"], aoa_deg=ctl["aoa"], cfl=ctl["cfl"],
                     dt=ctl["dt"], iflag_solver=ctl["iflag_solver"], iflag_time=ctl["iflag_time"],
                     iflag_turbulence=ctl["iflag_turbulence"], iflag_des=ctl["iflag_des"],
                     iflag_timeadvance=ctl["iflag_timeadvance"])
    for k, v in over.items():
        setattr(cfg, k, v)
    return cfg


def _numbers(fh):
    """Every value of the rest of a file, in list-directed order."""
    for line in fh:
        for tok in list_directed(line):
            yield tok


def read_grid(path, scale=(1.0, 1.0, 1.0), lref=1.0):
    """3-D grid file -> dims [(ni, nj, nk)], coords [(x, y, z)] with arrays
    (nk, nj, ni) over the physical range, multiplied by scale_* and divided
    by lref as sub_init_allocatememory.f90:408-415 does."""
    with open(path) as fh:
        it = _numbers(fh)
        nb = _int(next(it))
        dims = [tuple(_int(next(it)) for _ in range(3)) for _ in range(nb)]
        coords = []
        for ni, nj, nk in dims:
            c = []
            for a in range(3):
                v = np.array([_real(next(it)) for _ in range(ni * nj * nk)]).reshape(nk, nj, ni)
                c.append(v * scale[a] / lref)
            coords.append(tuple(c))
    return dims, coords


def write_grid(path, dims, coords):
    """Write a 3-D grid file (list-directed, 17 significant digits)."""
    with open(path, "w") as fh:
        fh.write("%d\n" % len(dims))
        for d in dims:
            fh.write("%d %d %d\n" % tuple(d))
        for c in coords:
            for a in range(3):
                flat = np.asarray(c[a], dtype=np.float64).ravel()
                for s in range(0, flat.size, 4):
                    fh.write(" ".join(repr(float(v)) for v in flat[s:s + 4]) + "\n")


def read_bcinfo(path, nblock):
    """bcinfo -> per block a list of 12-int records (f_no, face, is, ie, js,
    je, ks, ke, blk_t, subface_t, ori, fwh)."""
    with open(path) as fh:
        lines = fh.readlines()
    pos = 3
    out = []
    for _ in range(nblock):
        pos += 2
        nsub = _int(list_directed(lines[pos])[0])
        pos += 2
        recs = []
        for _ in range(nsub):
            toks = list_directed(lines[pos])
            pos += 1
            recs.append([_int(t) for t in toks[:12]])
        pos += 1
        out.append(recs)
    return out


def write_bcinfo(path, subfaces):
    """Write a bcinfo file; records of 11 ints get fwh = 0."""
    with open(path, "w") as fh:
        fh.write("bcinfo\n%d blocks\n\n" % len(subfaces))
        for m, recs in enumerate(subfaces):
            fh.write("block %d\nnumber of subfaces\n%d\n" % (m + 1, len(recs)))
            fh.write("f_no face is ie js je ks ke blk_t subface_t ori fwh\n")
            for r in recs:
                r = list(r) + [0] * (12 - len(r))
                fh.write(" ".join("%d" % int(v) for v in r) + "\n")
            fh.write("\n")


def subfaces_of(bcinfo):
    """The 11-int records acfd_set_topology takes (fwh dropped)."""
    return [[r[:11] for r in recs] for recs in bcinfo]


def write_flowfield(path, ranges, coords, pri_v, nut):
    """The reference's flowfield / restart file (sub_output_flowfield.f90:
    52-79): per block the physical range, one node per line.  coords[m],
    pri_v[m], nut[m] over the extended extents (as downloaded)."""
    with open(path, "w") as fh:
        for r, (x, y, z), pv, nt in zip(ranges, coords, pri_v, nut):
            ph = np.s_[r[10] - r[4]:r[11] - r[4] + 1, r[8] - r[2]:r[9] - r[2] + 1, r[6] - r[0]:r[7] - r[0] + 1]
            fh.write(" variables = x,y,z,d,u,v,w,p,nut\n")
            fh.write(" zone i= %d j= %d k= %d datapacking=point\n" % (r[7] - r[6] + 1, r[9] - r[8] + 1,
                                                                       r[11] - r[10] + 1))
            cols = [x[ph], y[ph], z[ph], pv[ph + (0,)], pv[ph + (1,)], pv[ph + (2,)], pv[ph + (3,)],
                    pv[ph + (4,)], nt[ph]]
            a = np.stack([np.asarray(c).ravel() for c in cols], 1)
            for row in a:
                fh.write(" ".join(repr(float(v)) for v in row) + "\n")


def _record(fh, count):
    """The first `count` values of one list-directed READ: the values may
    run on over several lines (flang's list-directed output wraps a long
    record, e.g. a flowfield node of nine doubles), and the rest of the last
    line read is skipped."""
    toks = []
    while len(toks) < count:
        line = fh.readline()
        if not line:
            raise ValueError("end of file inside a record of %d values" % count)
        toks.extend(list_directed(line))
    return toks[:count]


def read_flowfield(path, dims, ma, gamma=1.4):
    """Restart file -> per block dict with x, y, z, q (nk, nj, ni, 5), nut
    and pri_v as sub_init_flowfield.f90:206-258 sets them (pri_v includes
    the reference's q2 = rho^2 + u^2 + v^2; the path recomputes pri_v before
    using it)."""
    cv = 1.0 / gamma / (gamma - 1.0) / ma / ma
    out = []
    with open(path) as fh:
        for ni, nj, nk in dims:
            fh.readline()
            fh.readline()
            n = ni * nj * nk
            a = np.array([[_real(t) for t in _record(fh, 9)] for _ in range(n)])
            x, y, z, d, u, v, w, p, tur = (a[:, c].reshape(nk, nj, ni) for c in range(9))
            t = p / d * gamma * ma * ma
            q = np.stack([d, d * u, d * v, d * w, d * t * cv + 0.5 * d * (u * u + v * v + w * w)], -1)
            pri = np.zeros((nk, nj, ni, 7))
            pri[..., 0] = q[..., 0]
            pri[..., 1] = q[..., 1] / pri[..., 0]
            pri[..., 2] = q[..., 2] / pri[..., 0]
            pri[..., 3] = q[..., 3] / pri[..., 0]
            q2 = pri[..., 0] * pri[..., 0] + pri[..., 1] * pri[..., 1] + pri[..., 2] * pri[..., 2]
            pri[..., 5] = (q[..., 4] / pri[..., 0] - 0.5 * q2) / cv
            pri[..., 4] = pri[..., 0] * pri[..., 5] / gamma / ma / ma
            pri[..., 6] = np.sqrt(pri[..., 4]) / ma
            out.append(dict(x=x, y=y, z=z, q=q, nut=tur, pri_v=pri))
    return out


def _e11_4(v):
    """Fortran E11.4: 0.dddd and a signed two-digit exponent, right-justified
    in 11 columns; the four digits are the correctly rounded ones."""
    v = float(v)
    if v == 0.0:
        return " 0.0000E+00"
    mant, exp = ("%.3e" % abs(v)).split("e")
    digits = mant.replace(".", "")
    s = "%s0.%sE%+03d" % ("-" if v < 0 else "", digits, int(exp) + 1)
    return s.rjust(11)


def residual_line(timestep, res):
    """One line of resi_history.dat: (I6.6, 6E11.4) of timestep, the five
    flow residuals and the SA residual (sub_output_residual.f90:63-64)."""
    return "%06d" % timestep + "".join(_e11_4(float(v)) for v in list(res[:5]) + [res[5]])


def read_residuals(path):
    """resi_history.dat -> array of rows (timestep, five flow residuals, SA
    residual).  The header is list-directed output of one string, which
    the reference's writer may wrap over two lines; data lines are the
    (I6.6, 6E11.4) records."""
    rows = []
    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\n")
            if len(line) < 72 or not line[:6].isdigit():
                continue
            rows.append([int(line[:6])] + [float(line[6 + 11 * n:17 + 11 * n]) for n in range(6)])
    return np.array(rows).reshape(-1, 7)


def append_residual(path, timestep, res):
    with open(path, "a") as fh:
        fh.write(residual_line(timestep, res) + "\n")


def append_clcd(path, time, cl, cd):
    """clcd_history.dat: time, cl, cd in list-directed style (every digit)."""
    with open(path, "a") as fh:
        fh.write(" %r %r %r\n" % (float(time), float(cl), float(cd)))
