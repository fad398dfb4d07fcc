"""Tipsy particle files and ChaNGa's ASCII array outputs.

Tipsy (the format ChaNGa reads, InOutput.cpp:20-60; testdata/tipsydefs.h): a header
{double time; int nbodies, ndim, nsph, ndark, nstar; int pad} followed by
nsph gas, ndark dark and nstar star records of 4-byte floats:
  gas:  mass, pos[3], vel[3], rho, temp, hsmooth, metals, phi  (12)
  dark: mass, pos[3], vel[3], eps, phi                          (9)
  star: mass, pos[3], vel[3], metals, tform, eps, phi           (11)
"Standard" files are big-endian (XDR); native files use the host's order.
The byte order is detected from the header against the file size.

Array outputs (OutputParams with sTipsyExt "acc2" / "pot",
InOutput.h:264-285, 386-410; written by TreePiece::outputASCII,
InOutput.cpp:1735-1990): the particle count on the first line, then one
value per line in "%.14g", particles in input order; a vector writes every
x, then every y, then every z.
"""
import numpy as np

HEADER = [("time", "f8"), ("nbodies", "i4"), ("ndim", "i4"), ("nsph", "i4"), ("ndark", "i4"),
          ("nstar", "i4"), ("pad", "i4")]
GAS = [("mass", "f4"), ("pos", "f4", 3), ("vel", "f4", 3), ("rho", "f4"), ("temp", "f4"),
       ("hsmooth", "f4"), ("metals", "f4"), ("phi", "f4")]
DARK = [("mass", "f4"), ("pos", "f4", 3), ("vel", "f4", 3), ("eps", "f4"), ("phi", "f4")]
STAR = [("mass", "f4"), ("pos", "f4", 3), ("vel", "f4", 3), ("metals", "f4"), ("tform", "f4"),
        ("eps", "f4"), ("phi", "f4")]


def _dtypes(order):
    def d(fields):
        return np.dtype([(f[0], order + f[1], f[2]) if len(f) == 3 else (f[0], order + f[1])
                         for f in fields])
    return d(HEADER), d(GAS), d(DARK), d(STAR)


def _sizes_match(h, gas, dark, star, nbytes):
    ns, nd, nst = int(h["nsph"]), int(h["ndark"]), int(h["nstar"])
    if min(ns, nd, nst) < 0 or ns + nd + nst != int(h["nbodies"]):
        return False
    return h.dtype.itemsize + ns * gas.itemsize + nd * dark.itemsize + nst * star.itemsize == nbytes


def read(path):
    """Returns dict(time, gas, dark, star, big_endian): structured arrays in
    file order (gas, then dark, then star)."""
    raw = np.fromfile(path, dtype=np.uint8)
    for order in (">", "<"):
        hd, gd, dd, sd = _dtypes(order)
        if raw.size < hd.itemsize:
            break
        h = raw[:hd.itemsize].view(hd)[0]
        if _sizes_match(h, gd, dd, sd, raw.size):
            o = hd.itemsize
            out = {"time": float(h["time"]), "big_endian": order == ">"}
            for key, dt, n in (("gas", gd, int(h["nsph"])), ("dark", dd, int(h["ndark"])),
                               ("star", sd, int(h["nstar"]))):
                out[key] = raw[o:o + n * dt.itemsize].view(dt).copy()
                o += n * dt.itemsize
            return out
    raise ValueError(f"{path}: not a tipsy file (header does not match the file size)")


def write(path, time, gas=None, dark=None, star=None, big_endian=True):
    """Write a tipsy file; gas / dark / star are structured arrays (or dicts
    of columns) with the fields above (missing columns are zero)."""
    hd, gd, dd, sd = _dtypes(">" if big_endian else "<")

    def pack(src, dt):
        if src is None:
            return np.zeros(0, dtype=dt)
        names = src.dtype.names if isinstance(src, np.ndarray) else list(src.keys())
        n = len(src[names[0]]) if names else 0
        a = np.zeros(n, dtype=dt)
        for name in dt.names:
            if name in names:
                a[name] = src[name]
        return a

    g, d, s = pack(gas, gd), pack(dark, dd), pack(star, sd)
    h = np.zeros(1, dtype=hd)
    h["time"], h["ndim"] = time, 3
    h["nsph"], h["ndark"], h["nstar"] = len(g), len(d), len(s)
    h["nbodies"] = len(g) + len(d) + len(s)
    with open(path, "wb") as f:
        for a in (h, g, d, s):
            f.write(a.tobytes())


def particles(t):
    """(pos[N,3], mass[N], soft[N]) in file order, FP64; a gas particle's
    softening is its hsmooth (InOutput.cpp:31), stars and dark matter carry
    eps."""
    pos, mass, soft = [], [], []
    for key, eps in (("gas", "hsmooth"), ("dark", "eps"), ("star", "eps")):
        a = t[key]
        if len(a):
            pos.append(a["pos"].astype(np.float64))
            mass.append(a["mass"].astype(np.float64))
            soft.append(a[eps].astype(np.float64))
    if not pos:
        return np.zeros((0, 3)), np.zeros(0), np.zeros(0)
    return np.concatenate(pos), np.concatenate(mass), np.concatenate(soft)


def workload(path, theta=0.7, bucket=16, period=None, n_replicas=1, ew_cut=2.6, ewh_cut=None):
    """A workload dict (as configs.py builds) from a tipsy file; periodic
    when `period` is given: positions wrapped into [0, period), after a shift
    by period / 2 when the file holds negative coordinates (ChaNGa's boxes
    are centred on the origin; a translation leaves periodic forces as they
    are)."""
    t = read(path)
    pos, mass, soft = particles(t)
    w = dict(name=str(path), pos=np.ascontiguousarray(pos), mass=mass, soft=soft, theta=theta,
             bucket=bucket, periodic=False, L=0.0, n_replicas=0, ew_cut=0.0, ewh_cut=0.0,
             active=None)
    if period is not None:
        L = float(period)
        p = np.mod(pos + 0.5 * L, L) if pos.size and pos.min() < 0 else np.mod(pos, L)
        p[p >= L] = 0.0
        w.update(pos=np.ascontiguousarray(p), periodic=True, L=L, n_replicas=n_replicas,
                 ew_cut=ew_cut, ewh_cut=np.sqrt(10.0) if ewh_cut is None else ewh_cut)
    return w


def write_array(path, values):
    """ChaNGa's ASCII array output: N, then the values (a vector column by
    column: every x, every y, every z), "%.14g" one per line."""
    v = np.asarray(values, dtype=np.float64)
    n = v.shape[0]
    cols = [v] if v.ndim == 1 else [v[:, k] for k in range(v.shape[1])]
    with open(path, "w") as f:
        f.write(f"{n}\n")
        for c in cols:
            f.write("".join(f"{x:.14g}\n" for x in c))


def read_array(path, vector=False):
    """Inverse of write_array: [N] or [N, 3]."""
    with open(path) as f:
        n = int(f.readline())
        data = np.loadtxt(f, dtype=np.float64, ndmin=1)
    if vector:
        if data.size != 3 * n:
            raise ValueError(f"{path}: {data.size} values for {n} vectors")
        return np.ascontiguousarray(data.reshape(3, n).T)
    if data.size != n:
        raise ValueError(f"{path}: {data.size} values for {n} particles")
    return data


def write_outputs(prefix, acc, pot):
    """<prefix>.acc2 and <prefix>.pot, particles in input order."""
    write_array(f"{prefix}.acc2", acc)
    write_array(f"{prefix}.pot", pot)


def evaluate(path, prefix, theta=0.7, bucket=16, period=None, device="cuda:0", **kw):
    """One force evaluation of a tipsy file on the GPU (precision form, GPU
    lists), written as <prefix>.acc2 / <prefix>.pot in input order. Returns
    (acc, pot)."""
    import torch

    from changa_amd import engine

    w = workload(path, theta=theta, bucket=bucket, period=period, **kw)
    ev = engine.Evaluation(w, precise=True, device=device, lists="gpu")
    ev.step()
    torch.cuda.synchronize()
    acc_t, pot_t, _ = ev.results()
    n = w["pos"].shape[0]
    acc, pot = np.empty((n, 3)), np.empty(n)
    acc[ev.tree["perm"]] = acc_t
    pot[ev.tree["perm"]] = pot_t
    write_outputs(prefix, acc, pot)
    return acc, pot
