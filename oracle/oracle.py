"""ctypes wrapper of the FP64 CPU oracle (oracle.c) and of the reference's own
moments.c compiled into oracle/_ref/.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as the
checker and by bench.py's cpu_baseline leg. The product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
NMOM = 22

_lib = None
_ref = None

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def lib():
    """Load liboracle.so (building it with make if absent and gcc exists)."""
    global _lib
    if _lib is not None:
        return _lib
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        import subprocess
        subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    L = ctypes.CDLL(path)
    L.orc_spline.argtypes = [_d, _d, _p, _p]
    L.orc_mom_make.argtypes = [_p, _d, _d, _d, _d, _d]
    L.orc_mom_make.restype = _d
    L.orc_mom_shift.argtypes = [_p, _d, _d, _d, _d]
    L.orc_mom_rescale.argtypes = [_p, _d, _d]
    L.orc_mom_scaled_add.argtypes = [_p, _d, _p, _d]
    L.orc_mom_to_momc.argtypes = [_p, _p]
    L.orc_mom_eval.argtypes = [_p, _d, _d, _d, _d, _p, _p]
    L.orc_build_moments.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p]
    L.orc_forces_lists.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i,
                                   _p, _p, _p, _p, _d, _i, _i, _p, _p, _p, _p]
    L.orc_forces_lists3.argtypes = [_i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _i,
                                    _p, _p, _p, _p, _p, _i, _i, _p, _p, _p, _p]
    L.orc_open_softening.argtypes = [_p, _p, _d, _d, _d]
    L.orc_open_softening.restype = _i
    L.orc_ewald_table.argtypes = [_p, _d, _d, _d, _p, _i]
    L.orc_ewald_table.restype = _i
    L.orc_bucket_ewald.argtypes = [_i, _p, _p, _p, _d, _p, _d, _p, _i, _d, _i, _d, _i,
                                   _p, _p, _p]
    L.orc_direct_sum.argtypes = [_i, _p, _p, _p, _i, _p, _i, _p, _p]
    L.orc_ewald_exact.argtypes = [_i, _p, _p, _p, _i, _p, _d, _d, _d, _i, _i, _p, _p]
    L.orc_adjoint_lists.argtypes = [_i, _p, _i, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p, _p,
                                    _p, _d, _p, _p, _i, _p]
    L.orc_adjoint_ewald.argtypes = [_i, _p, _p, _p, _d, _p, _d, _p, _i, _d, _i, _d, _i, _p, _p]
    L.orc_direct_grad.argtypes = [_i, _p, _p, _p, _p, _i, _p]
    L.orc_rinv_derivs.argtypes = [_p, _i, _p]
    _lib = L
    return L


def ref_lib():
    """The reference's moments.c, compiled from its own source (oracle/_ref).
    Returns None when it was never built (reference tree absent)."""
    global _ref
    if _ref is not None:
        return _ref
    path = os.path.join(HERE, "_ref", "libmoments_ref.so")
    if not os.path.exists(path):
        return None
    R = ctypes.CDLL(path)
    R.momMakeFmomr.argtypes = [_p, _d, _d, _d, _d, _d]
    R.momMakeFmomr.restype = _d
    R.momShiftFmomr.argtypes = [_p, _d, _d, _d, _d]
    R.momRescaleFmomr.argtypes = [_p, _d, _d]
    R.momScaledAddFmomr.argtypes = [_p, _d, _p, _d]
    R.momFmomr2Momc.argtypes = [_p, _p]
    R.momEvalFmomrcm.argtypes = [_p, _d, _d, _d, _d, _d, _p, _p, _p, _p, _p]
    _ref = R
    return R


# ---------------------------------------------------------------- helpers
def spline(r2, twoh):
    a = np.zeros(1)
    b = np.zeros(1)
    lib().orc_spline(r2, twoh, _ptr(a), _ptr(b))
    return a[0], b[0]


def mom_make(m, u, x, y, z):
    q = np.zeros(NMOM)
    lib().orc_mom_make(_ptr(q), m, u, x, y, z)
    return q


def mom_shift(q, u, dx, dy, dz):
    q = np.array(q, dtype=np.float64)
    lib().orc_mom_shift(_ptr(q), u, dx, dy, dz)
    return q


def mom_rescale(q, unew, uold):
    q = np.array(q, dtype=np.float64)
    lib().orc_mom_rescale(_ptr(q), unew, uold)
    return q


def mom_scaled_add(qr, ur, qa, ua):
    qr = np.array(qr, dtype=np.float64)
    qa = np.ascontiguousarray(qa, dtype=np.float64)
    lib().orc_mom_scaled_add(_ptr(qr), ur, _ptr(qa), ua)
    return qr


def mom_to_momc(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    mc = np.zeros(32)
    lib().orc_mom_to_momc(_ptr(q), _ptr(mc))
    return mc


def mom_eval(q, u, x, y, z):
    q = np.ascontiguousarray(q, dtype=np.float64)
    pot = np.zeros(1)
    a = np.zeros(3)
    lib().orc_mom_eval(_ptr(q), u, x, y, z, _ptr(pot), _ptr(a))
    return pot[0], a


def build_moments(tree, pos, mass, soft):
    """FP64 moments of every node of `tree` (a dict of topology arrays as
    produced by changa_amd.host.build_tree). Returns (geo[n,12], mom[n,22])."""
    nn = int(tree["n_nodes"])
    geo = np.zeros((nn, 12))
    mom = np.zeros((nn, NMOM))
    lib().orc_build_moments(nn, _ptr(tree["node_first"]), _ptr(tree["node_count"]),
                            _ptr(tree["node_child"]), _ptr(tree["node_box_radius"]),
                            _ptr(pos), _ptr(mass), _ptr(soft), _ptr(geo), _ptr(mom))
    return geo, mom


def forces_lists(tree, geo, mom, pos, mass, soft, lists, period=0.0, rung=None,
                 active_rung=0, n_threads=0, buckets=None, soft_flags=False):
    """nodeBucketForce/partBucketForce over the interaction lists. `period`:
    the replica period, or three of them (x, y, z) for a box that is not a
    cube. `buckets` restricts the evaluation to a subset of bucket indices
    (the lists are then re-sliced). soft_flags: take the softened-cell decision from bit 31 of
    the list entries instead of openSoftening on `geo` (keeps it fixed when
    positions are perturbed). Returns acc[N,3], pot[N], dtgrav[N], counts[3]."""
    n = pos.shape[0]
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    dtg = np.zeros(n)
    counts = np.zeros(3, dtype=np.int64)
    bnode = tree["bucket_node"]
    cells, cmarks = lists["cells"], lists["cell_marks"]
    parts, pmarks = lists["parts"], lists["part_marks"]
    if buckets is not None:
        buckets = np.asarray(buckets, dtype=np.int64)
        bnode = np.ascontiguousarray(bnode[buckets])
        cells, cmarks = _reslice(cells, cmarks, buckets, 2)
        parts, pmarks = _reslice(parts, pmarks, buckets, 3)
    rg = None if rung is None else np.ascontiguousarray(rung, dtype=np.int32)
    per = np.atleast_1d(np.asarray(period, dtype=np.float64))
    if per.size == 3:
        per3 = np.ascontiguousarray(per)
        lib().orc_forces_lists3(len(bnode), _ptr(bnode), _ptr(tree["node_first"]),
                                _ptr(tree["node_count"]), _ptr(geo), _ptr(mom), _ptr(pos),
                                _ptr(mass), _ptr(soft), _ptr(rg), active_rung,
                                _ptr(cells), _ptr(cmarks), _ptr(parts), _ptr(pmarks),
                                _ptr(per3), int(soft_flags), n_threads, _ptr(acc), _ptr(pot),
                                _ptr(dtg), _ptr(counts))
        return acc, pot, dtg, counts
    period = float(per[0])
    lib().orc_forces_lists(len(bnode), _ptr(bnode), _ptr(tree["node_first"]),
                           _ptr(tree["node_count"]), _ptr(geo), _ptr(mom), _ptr(pos),
                           _ptr(mass), _ptr(soft), _ptr(rg), active_rung,
                           _ptr(cells), _ptr(cmarks), _ptr(parts), _ptr(pmarks),
                           float(period), int(soft_flags), n_threads, _ptr(acc), _ptr(pot),
                           _ptr(dtg), _ptr(counts))
    return acc, pot, dtg, counts


def _reslice(lst, marks, buckets, width):
    lst = lst.reshape(-1, width)
    lens = marks[buckets + 1] - marks[buckets]
    newm = np.zeros(len(buckets) + 1, dtype=np.int64)
    newm[1:] = np.cumsum(lens)
    idx = np.concatenate([np.arange(marks[b], marks[b + 1]) for b in buckets]) \
        if len(buckets) else np.zeros(0, dtype=np.int64)
    return np.ascontiguousarray(lst[idx].reshape(-1)), newm


def open_softening(cell_geo, bucket_geo, off):
    return bool(lib().orc_open_softening(_ptr(np.ascontiguousarray(cell_geo)),
                                         _ptr(np.ascontiguousarray(bucket_geo)),
                                         off[0], off[1], off[2]))


def ewald_table(q_root, u_root, L, hcut, max_n=4096):
    q = np.ascontiguousarray(q_root, dtype=np.float64)
    ewt = np.zeros((max_n, 5))
    n = lib().orc_ewald_table(_ptr(q), u_root, L, hcut, _ptr(ewt), max_n)
    if n < 0:
        raise ValueError("Ewald table larger than max_n")
    return ewt[:n].copy()


def bucket_ewald(targets, pos, root_cm, root_mass, q_root, u_root, ewt, L, n_reps,
                 ew_cut, n_threads=0):
    n = pos.shape[0]
    acc = np.zeros((n, 3))
    pot = np.zeros(n)
    counts = np.zeros(2, dtype=np.int64)
    t = np.ascontiguousarray(targets, dtype=np.int32)
    cm = np.ascontiguousarray(root_cm, dtype=np.float64)
    q = np.ascontiguousarray(q_root, dtype=np.float64)
    e = np.ascontiguousarray(ewt, dtype=np.float64)
    lib().orc_bucket_ewald(len(t), _ptr(t), _ptr(pos), _ptr(cm), root_mass, _ptr(q), u_root,
                           _ptr(e), len(e), L, n_reps, ew_cut, n_threads, _ptr(acc),
                           _ptr(pot), _ptr(counts))
    return acc, pot, counts


def direct_sum(pos, mass, soft, targets, n_threads=0):
    t = np.ascontiguousarray(targets, dtype=np.int32)
    acc = np.zeros((len(t), 3))
    pot = np.zeros(len(t))
    lib().orc_direct_sum(pos.shape[0], _ptr(pos), _ptr(mass), _ptr(soft), len(t), _ptr(t),
                         n_threads, _ptr(acc), _ptr(pot))
    return acc, pot


def ewald_exact(pos, mass, soft, targets, L=1.0, alpha=None, rcut=None, hmax=None,
                n_threads=0):
    """FP64 periodic softened forces on `targets` by a converged Ewald sum."""
    rcut = 0.3 * L if rcut is None else rcut
    alpha = 4.0 / rcut if alpha is None else alpha
    if hmax is None:
        # exp(-k^2/4a^2) < 1e-8 at |k| = 2 pi hmax / L
        hmax = int(np.ceil(2 * alpha * np.sqrt(np.log(1e8)) * L / (2 * np.pi)))
    t = np.ascontiguousarray(targets, dtype=np.int32)
    acc = np.zeros((len(t), 3))
    pot = np.zeros(len(t))
    lib().orc_ewald_exact(pos.shape[0], _ptr(pos), _ptr(mass), _ptr(soft), len(t), _ptr(t),
                          L, alpha, rcut, hmax, n_threads, _ptr(acc), _ptr(pot))
    return acc, pot


def adjoint_lists(tree, geo, mom, pos, mass, soft, lists, w, period=0.0, n_threads=0,
                  root_extra=None):
    """Exact gradient of L = sum_i w_i . a_i for the list evaluation with
    softened cells as flagged in the lists (forces_lists(..., soft_flags=True)).
    root_extra: 35 monomial coefficients added to the root's polynomial (the
    source side of the Ewald term, adjoint_ewald). Returns grad[N, 4]:
    dL/dx, dL/dm."""
    n = pos.shape[0]
    grad = np.zeros((n, 4))
    wv = np.ascontiguousarray(w, dtype=np.float64)
    rx = None if root_extra is None else np.ascontiguousarray(root_extra, dtype=np.float64)
    lib().orc_adjoint_lists(tree["n_nodes"], _ptr(tree["node_child"]), tree["n_buckets"],
                            _ptr(tree["bucket_node"]), _ptr(tree["node_first"]),
                            _ptr(tree["node_count"]), _ptr(geo), _ptr(mom), _ptr(pos),
                            _ptr(mass), _ptr(soft), _ptr(lists["cells"]),
                            _ptr(lists["cell_marks"]), _ptr(lists["parts"]),
                            _ptr(lists["part_marks"]), float(period), _ptr(wv), _ptr(rx),
                            n_threads, _ptr(grad))
    return grad


def adjoint_ewald(pos, w, geo, mom, ewt, L, n_reps, ew_cut, n_threads=0):
    """Adjoint of bucket_ewald over every particle (w zero for non-targets).
    Returns (grad[N, 4] with the target-side terms, hroot[35])."""
    n = pos.shape[0]
    grad = np.zeros((n, 4))
    hroot = np.zeros(35)
    wv = np.ascontiguousarray(w, dtype=np.float64)
    e = np.ascontiguousarray(ewt, dtype=np.float64)
    cm = np.ascontiguousarray(geo[0, 0:3])
    q = np.ascontiguousarray(mom[0])
    lib().orc_adjoint_ewald(n, _ptr(pos), _ptr(wv), _ptr(cm), float(geo[0, 5]), _ptr(q),
                            float(geo[0, 3]), _ptr(e), len(e), float(L), int(n_reps),
                            float(ew_cut), n_threads, _ptr(grad), _ptr(hroot))
    return grad, hroot


def direct_grad(pos, mass, soft, w, n_threads=0):
    """Exact gradient of L = sum_i w_i . a_i for the softened direct sum."""
    n = pos.shape[0]
    grad = np.zeros((n, 4))
    wv = np.ascontiguousarray(w, dtype=np.float64)
    lib().orc_direct_grad(n, _ptr(pos), _ptr(mass), _ptr(soft), _ptr(wv), n_threads, _ptr(grad))
    return grad


def rinv_derivs(r, nmax=6):
    """T[a, b, c] = d^(a+b+c)(1/|r|)/dx^a dy^b dz^c (zero beyond order nmax)."""
    out = np.zeros(343)
    lib().orc_rinv_derivs(_ptr(np.ascontiguousarray(r, dtype=np.float64)), nmax, _ptr(out))
    return out.reshape(7, 7, 7)


# ---------------------------------------------------------------------------
# The local tree walk of ChaNGa's GPU_LOCAL_TREE_WALK build (numpy)
# ---------------------------------------------------------------------------
# CudaNodeType (cuda_typedef.h:87-99)
_LTW_BUCKET, _LTW_INTERNAL, _LTW_BOUNDARY, _LTW_NONLOCAL_BUCKET = 2, 3, 4, 8


def ltw_lists(radius, soft, cm, lo, hi, count, ntype, children, bstart, bsize, bucket_nodes,
              theta, theta_mono, n_replicas=0, period=0.0, root=0, dtype=np.float32):
    """Interaction lists of the reference's gpuLocalTreeWalk (HostCUDA.cu:
    724-942) for each bucket node of `bucket_nodes`: the walk from `root`
    per periodic replica, CUDA_openCriterionNode (CUDAMoments.cu:197-242)
    against the bucket's box, CUDA_openSoftening (:180-195) and the node-type
    actions of CUDA_OptAction (:323-352). The test depends on the particle
    only through its bucket node, so one list per bucket serves all its
    particles.

    dtype float32: the reference's arithmetic (cudatype operands, the opening
    radii formed in double from 2/sqrt(3) and rounded to float); numpy does
    each float32 operation with one rounding and fuses none. dtype float64:
    the same walk in double, which on the same geometry must give the host
    list builder's lists (openCriterionBucket, gravity.h:597-635) -- the
    check that pins this walk's traversal, actions and replica handling.

    Node arrays are per node (cm, lo, hi: [n, 3]; children: [n, 2], -1 none).
    period: one replica period, or three (fperiod, fperiodY, fperiodZ).
    Returns the lists in the layout of forces_lists (flat int32): cells
    (node, offsetID with the replica code in bits 22-30 and bit 31 for a
    softened cell), parts (bucketStart, offsetID, bucketSize), int64 marks
    per bucket; within a bucket the entries are sorted."""
    f = np.dtype(dtype).type
    kgeom = 2.0 / np.sqrt(3.0)
    rad = np.asarray(radius, dtype=f)
    sft = np.asarray(soft, dtype=f)
    cm = np.asarray(cm, dtype=f)
    count = np.asarray(count)
    ntype = np.asarray(ntype)
    children = np.asarray(children).reshape(-1, 2)
    bstart = np.asarray(bstart)
    bsize = np.asarray(bsize)
    bucket_nodes = np.asarray(bucket_nodes, dtype=np.int64)
    nb = len(bucket_nodes)
    bc = cm[bucket_nodes]
    brb = f(2.0) * sft[bucket_nodes]
    blo = np.asarray(lo, dtype=f)[bucket_nodes]
    bhi = np.asarray(hi, dtype=f)[bucket_nodes]
    th, thm = f(theta), f(theta_mono)
    R = int(n_replicas)
    W = 2 * R + 1
    k = np.arange(W ** 3)
    rx, ry, rz = k // (W * W) - R, (k // W) % W - R, k % W - R
    codes = ((rx + 3) | ((ry + 3) << 3) | ((rz + 3) << 6)).astype(np.int64)
    # frontier of (bucket, node, replica code)
    fb = np.repeat(np.arange(nb), len(codes))
    fn = np.full(len(fb), root, dtype=np.int64)
    fc = np.tile(codes, nb)
    per = np.asarray(period, dtype=f)  # one period, or x, y, z

    def box_sphere(b, c, r):
        dsq = np.zeros(len(b), dtype=f)
        for a in range(3):
            d1 = blo[b, a] - c[:, a]
            d2 = c[:, a] - bhi[b, a]
            e = np.where(d1 > 0, d1, np.where(d2 > 0, d2, f(0)))
            dsq = dsq + e * e
        return dsq <= r * r

    def soft_open(b, c, rs):
        d = bc[b] - c
        dist = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
        return (dist <= (brb[b] + rs) * (brb[b] + rs)) | box_sphere(b, c, rs)

    out_c, out_p = [], []
    while len(fb):
        off = np.stack([((fc & 7) - 3), (((fc >> 3) & 7) - 3), (((fc >> 6) & 7) - 3)], axis=1)
        c = cm[fn] + off.astype(f) * per
        r = rad[fn]
        rs = f(2.0) * sft[fn]
        ro = (kgeom * r.astype(np.float64) / np.float64(th)).astype(f)
        ro = np.where(ro < r, r, ro)
        rm = (kgeom * r.astype(np.float64) / np.float64(thm)).astype(f)
        so = soft_open(fb, c, rs)
        opened = (count[fn] <= 6) | box_sphere(fb, c, ro) | (so & box_sphere(fb, c, rm))
        t = ntype[fn]
        compute = ~opened & ((t == _LTW_INTERNAL) | (t == _LTW_BUCKET) | (t == _LTW_BOUNDARY) |
                             (t == _LTW_NONLOCAL_BUCKET))
        oid = (fc << 22).astype(np.int64)
        sel = np.nonzero(compute)[0]
        out_c.append(np.stack([fb[sel], fn[sel],
                               oid[sel] | np.where(so[sel], 1 << 31, 0)], axis=1))
        sel = np.nonzero(opened & (t == _LTW_BUCKET) & (bsize[fn] > 0))[0]
        out_p.append(np.stack([fb[sel], bstart[fn[sel]], oid[sel], bsize[fn[sel]]], axis=1))
        push = opened & ((t == _LTW_INTERNAL) | (t == _LTW_BOUNDARY))
        nb_, nn_, nc_ = [], [], []
        for j in (0, 1):
            ch = children[fn, j]
            s = np.nonzero(push & (ch >= 0))[0]
            nb_.append(fb[s])
            nn_.append(ch[s])
            nc_.append(fc[s])
        fb, fn, fc = np.concatenate(nb_), np.concatenate(nn_).astype(np.int64), np.concatenate(nc_)

    def pack(chunks, width):
        a = np.concatenate(chunks) if chunks else np.zeros((0, width + 1), dtype=np.int64)
        a = a[np.lexsort(a[:, ::-1].T)]  # by bucket, then the entry
        marks = np.zeros(nb + 1, dtype=np.int64)
        np.add.at(marks, a[:, 0] + 1, 1)
        entries = a[:, 1:].astype(np.uint32).view(np.int32).reshape(-1)
        return np.ascontiguousarray(entries), np.cumsum(marks)

    cells, cmarks = pack(out_c, 2)
    parts, pmarks = pack(out_p, 3)
    return dict(cells=cells, cell_marks=cmarks, parts=parts, part_marks=pmarks)
