"""Shared helpers for the parity tests: the oracle run on the product's own
tree and interaction lists, and the error metrics."""
import numpy as np

from oracle import oracle as O


def oracle_forces(ev, n_threads=0, buckets=None):
    """FP64 nodeBucketForce/partBucketForce (+ BucketEwald when periodic) on
    the identical tree and lists of an engine.Evaluation, with the oracle's
    own moments. Returns acc[N,3], pot[N], dtgrav[N]."""
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    period = ev.L if ev.w["periodic"] else 0.0
    acc, pot, dtg, _ = O.forces_lists(ev.tree, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists,
                                      period=period, n_threads=n_threads, buckets=buckets)
    if ev.w["periodic"]:
        if buckets is None:
            targets = np.arange(ev.pos.shape[0]) if ev.active is None else np.nonzero(ev.active)[0]
        else:
            tr = ev.tree
            targets = np.concatenate([np.arange(tr["bucket_start"][b],
                                                tr["bucket_start"][b] + tr["bucket_size"][b])
                                      for b in buckets])
        ewt = O.ewald_table(mom[0], geo[0, 3], ev.L, ev.w["ewh_cut"])
        ea, ep, _ = O.bucket_ewald(targets, ev.pos, geo[0, 0:3], geo[0, 5], mom[0], geo[0, 3], ewt,
                                   ev.L, ev.w["n_replicas"], ev.w["ew_cut"], n_threads=n_threads)
        acc += ea
        pot += ep
    return acc, pot, dtg


def rel_err(a, b):
    """Per-particle |a - b| / |b| for vectors (rows) or scalars."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if a.ndim == 2:
        return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    return np.abs(a - b) / np.abs(b)


def summary(e):
    return dict(median=float(np.median(e)), p99=float(np.percentile(e, 99)), max=float(e.max()))


def check_tails(e, med, p99, mx, label=""):
    """Median, 99th percentile and max of a per-particle error against bars
    set at about 1.5x the tails measured for the same case (the cases are
    seeded, so a run reproduces its tails up to the FP32 order of the
    partial sums); None skips a bar."""
    s = summary(e)
    for key, bar in (("median", med), ("p99", p99), ("max", mx)):
        assert bar is None or s[key] <= bar, (label, key, s)


def deep_chain_workload(n=1500):
    """Particles at radii 2^(-k/25) (k < n) around the box centre: a chain of
    ~60 tree levels, ending in one bucket of ~950 particles whose Morton keys
    coincide (beyond 21 bits per axis)."""
    from changa_amd import configs
    w = configs.plummer(n=n)
    rng = np.random.default_rng(3)
    r = 2.0 ** (-np.arange(n) / 25.0)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return dict(w, pos=np.ascontiguousarray(0.5 + 0.5 * r[:, None] * d))
