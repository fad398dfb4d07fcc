"""The oracle's local tree walk (oracle.ltw_lists, the restatement of the
reference's gpuLocalTreeWalk) and the LTW serialization, on the CPU.

Pins: in double arithmetic on the host builder's node geometry the walk
gives the host list builder's lists entry for entry (openCriterionBucket
with the softening re-test, the replicas, particle runs of opened buckets);
the node-type actions of CUDA_OptAction are checked against set operations
on those lists (boundary nodes act as internal ones; non-local, empty and
cached subtrees drop out; a non-local bucket is kept as a cell but never
opened into particles); the FP32 walk of the reference differs from the
double one only where a node sits within rounding of a threshold."""
import numpy as np
import pytest

from changa_amd import configs, host, ltw
from oracle import oracle as O

CASES = {
    "plummer": lambda: (configs.plummer(n=8192), 0),
    "softened": lambda: (configs.plummer(n=2048, eps=0.5), 0),
    "box_r1": lambda: (configs.zeldovich(ng=12, rms=0.05), 1),
    "box_r2": lambda: (configs.zeldovich(ng=10, rms=0.05), 2),
    "box_r3": lambda: (configs.zeldovich(ng=8, rms=0.05), 3),
    "halos": lambda: (configs.halos(n=16384), 0),
}


def _setup(w):
    if w["periodic"]:
        tr = host.build_tree(w["pos"], w["bucket"], box_lo=np.zeros(3), box_size=w["L"])
    else:
        tr = host.build_tree(w["pos"], w["bucket"])
    perm = tr["perm"]
    pos = np.ascontiguousarray(w["pos"][perm])
    m = np.ascontiguousarray(w["mass"][perm])
    s = np.ascontiguousarray(w["soft"][perm])
    geo = host.node_geometry(tr, pos, m, s)
    return tr, pos, m, s, geo


def _types(tr):
    ch = tr["node_child"].reshape(-1, 2)
    return np.where(ch[:, 0] < 0, ltw.BUCKET, ltw.INTERNAL), ch


def _ltw(tr, geo, types, ch, theta, R, period, dtype, theta_mono=None):
    tm = theta ** 4 if theta_mono is None else theta_mono
    return O.ltw_lists(geo[:, 3], geo[:, 4], geo[:, 0:3], geo[:, 6:9], geo[:, 9:12],
                       tr["node_count"], types, np.where(ch[:, :1] < 0, -1, ch),
                       tr["node_first"], tr["node_count"], tr["bucket_node"], theta, tm, R,
                       period, dtype=dtype)


def _entries(lists, with_soft=True):
    """Per bucket: the sets of (node, replica code[, softened]) cells and
    (first, replica code, count) runs."""
    c = lists["cells"].reshape(-1, 2)
    p = lists["parts"].reshape(-1, 3)
    out = []
    for b in range(len(lists["cell_marks"]) - 1):
        cs = c[lists["cell_marks"][b]:lists["cell_marks"][b + 1]]
        ps = p[lists["part_marks"][b]:lists["part_marks"][b + 1]]
        oid = cs[:, 1].astype(np.int64) & 0xffffffff
        cells = {(int(n), int((o >> 22) & 0x1ff)) + ((int(o >> 31),) if with_soft else ())
                 for n, o in zip(cs[:, 0], oid)}
        runs = {(int(a), int((o >> 22) & 0x1ff), int(k)) for a, o, k in ps}
        out.append((cells, runs))
    return out


@pytest.mark.parametrize("case", list(CASES))
def test_double_walk_equals_host_lists(case):
    w, R = CASES[case]()
    tr, pos, m, s, geo = _setup(w)
    period = w["L"] if w["periodic"] else 0.0
    host_lists = host.walk(tr, geo, w["theta"], R, period if R else 1.0)
    types, ch = _types(tr)
    got = _ltw(tr, geo, types, ch, w["theta"], R, period, np.float64)
    a, b = _entries(got), _entries(host_lists)
    assert len(a) == len(b) == tr["n_buckets"]
    for k, (x, y) in enumerate(zip(a, b)):
        assert x == y, f"bucket {k}"
    if case == "softened":
        assert sum(sum(e[2] for e in cells) for cells, _ in a) > 0


def test_node_type_actions():
    """CUDA_OptAction on a serialized tree where some nodes are not local."""
    w = configs.plummer(n=4096)
    tr, pos, m, s, geo = _setup(w)
    types, ch = _types(tr)
    theta = w["theta"]
    base = _entries(_ltw(tr, geo, types, ch, theta, 0, 0.0, np.float64))
    nn = tr["n_nodes"]
    first, count = tr["node_first"], tr["node_count"]

    def subtree(x):
        # every split leaves two non-empty children: the nodes whose particle
        # range lies in x's are x and its descendants
        f, c = first[x], count[x]
        return {n for n in range(nn) if first[n] >= f and first[n] + count[n] <= f + c}

    # the internal node nearest to an eighth of the particles
    inner = np.nonzero(types == ltw.INTERNAL)[0]
    x = int(inner[np.argmin(np.abs(count[inner] - len(pos) // 8))])
    assert count[x] > 64
    sub = subtree(x)
    # boundary nodes act as internal nodes
    t = types.copy()
    t[x] = ltw.BOUNDARY
    assert _entries(_ltw(tr, geo, t, ch, theta, 0, 0.0, np.float64)) == base
    # a non-local (or empty / cached) node and what hangs below it drop out
    for code in (ltw.NONLOCAL, ltw.EMPTY, 9, 10, 11):
        t = types.copy()
        t[x] = code
        got = _entries(_ltw(tr, geo, t, ch, theta, 0, 0.0, np.float64))
        for (gc, gr), (bc, br) in zip(got, base):
            assert gc == {e for e in bc if e[0] not in sub}
            assert gr == {e for e in br if e[0] < first[x] or e[0] >= first[x] + count[x]}
    # a non-local bucket: kept as a cell, never opened into its particles
    bk = int(tr["bucket_node"][len(tr["bucket_node"]) // 3])
    t = types.copy()
    t[bk] = ltw.NONLOCAL_BUCKET
    got = _entries(_ltw(tr, geo, t, ch, theta, 0, 0.0, np.float64))
    for (gc, gr), (bc, br) in zip(got, base):
        assert gc == bc
        assert gr == {e for e in br if e[0] != first[bk]}
    # top and invalid nodes compute nothing (the ERROR action)
    for code in (ltw.TOP, 1):
        t = types.copy()
        t[x] = code
        got = _entries(_ltw(tr, geo, t, ch, theta, 0, 0.0, np.float64))
        assert all(gc == {e for e in bc if e[0] not in sub} for (gc, _), (bc, _) in zip(got, base))


@pytest.mark.parametrize("case", ["plummer", "box_r1", "halos"])
def test_float_walk_differs_only_at_thresholds(case):
    """The reference's FP32 test against the double one on the same nodes
    rounded to FP32: the lists agree up to a few entries per million."""
    w, R = CASES[case]()
    tr, pos, m, s, geo = _setup(w)
    period = w["L"] if w["periodic"] else 0.0
    types, ch = _types(tr)
    g32 = geo.astype(np.float32).astype(np.float64)
    a = _entries(_ltw(tr, g32, types, ch, w["theta"], R, period, np.float32))
    b = _entries(_ltw(tr, g32, types, ch, w["theta"], R, period, np.float64))
    n = sum(len(c) + len(r) for c, r in b)
    diff = sum(len(c1 ^ c2) + len(r1 ^ r2) for (c1, r1), (c2, r2) in zip(a, b))
    print(f"{case}: {diff} of {n} entries differ between FP32 and FP64 tests")
    assert diff <= max(2, n // 200000)


def test_serialize_layout():
    """The LTW node and particle records of a tree: sizes, children, types,
    bucket ranges and each particle's bucket node."""
    w = configs.plummer(n=2048)
    tr, pos, m, s, geo = _setup(w)
    nn = tr["n_nodes"]
    wire = np.zeros((nn, 27), dtype=np.float32)
    wire[:, 0], wire[:, 1], wire[:, 2] = geo[:, 3], geo[:, 4], geo[:, 5]
    wire[:, 3:6] = geo[:, 0:3]
    node64 = np.zeros((nn, 34))
    node64[:, :12] = geo
    nodes, parts = ltw.serialize(tr, wire, node64, pos, m, s)
    assert nodes.itemsize == 160 and parts.itemsize == 24
    assert nodes.dtype.fields["moments"][1] == 76  # offsetof(CudaMultipoleMoments, xx)
    ch = tr["node_child"].reshape(-1, 2)
    leaf = ch[:, 0] < 0
    assert (nodes["type"][leaf] == ltw.BUCKET).all() and (nodes["type"][~leaf] == ltw.INTERNAL).all()
    assert (nodes["children"][leaf] == -1).all() and (nodes["children"][~leaf] == ch[~leaf]).all()
    for b, bn in enumerate(tr["bucket_node"]):
        st, sz = tr["bucket_start"][b], tr["bucket_size"][b]
        assert nodes["bucketStart"][bn] == st and nodes["bucketSize"][bn] == sz
        assert (parts["nodeId"][st:st + sz] == bn).all()
    np.testing.assert_array_equal(nodes["lesser_corner"], geo[:, 6:9].astype(np.float32))


def test_rectangular_periods_against_the_image_sum():
    """A box that is not a cube (gpuLocalTreeWalk's fperiod, fperiodY,
    fperiodZ): the walk's lists with one replica shell and the FP64 list
    forces on them against the direct sum over the same 27 images, to the
    expansion's accuracy; the same lists evaluated with a cubic period
    disagree by orders of magnitude more (the offsets are per axis)."""
    rng = np.random.default_rng(5)
    n, per = 600, np.array([1.0, 1.5, 0.75])
    pos = rng.uniform(size=(n, 3)) * per
    m = np.full(n, 1.0 / n)
    s = np.full(n, 1e-7)
    tr = host.build_tree(pos, 8)
    p = tr["perm"]
    pos, m, s = (np.ascontiguousarray(a[p]) for a in (pos, m, s))
    geo, mom = O.build_moments(tr, pos, m, s)
    types, ch = _types(tr)
    theta = 0.4
    lists = _ltw(tr, geo, types, ch, theta, 1, per, np.float64)
    acc, pot, _, _ = O.forces_lists(tr, geo, mom, pos, m, s, lists, period=per, soft_flags=True)
    ref_a = np.zeros((n, 3))
    ref_p = np.zeros(n)
    for ix in (-1, 0, 1):
        for iy in (-1, 0, 1):
            for iz in (-1, 0, 1):
                d = pos[None, :, :] + np.array([ix, iy, iz]) * per - pos[:, None, :]
                r2 = (d * d).sum(axis=2)
                if (ix, iy, iz) == (0, 0, 0):
                    np.fill_diagonal(r2, np.inf)  # no self-interaction
                ir = 1.0 / np.sqrt(r2)
                ref_a += (d * (m[None, :] * ir ** 3)[:, :, None]).sum(axis=1)
                ref_p -= (m[None, :] * ir).sum(axis=1)
    e = np.linalg.norm(acc - ref_a, axis=1) / np.linalg.norm(ref_a, axis=1)
    ep = np.abs(pot - ref_p) / np.abs(ref_p)
    cube, _, _, _ = O.forces_lists(tr, geo, mom, pos, m, s, lists, period=1.0, soft_flags=True)
    e_cube = np.linalg.norm(cube - ref_a, axis=1) / np.linalg.norm(ref_a, axis=1)
    print(f"rectangular: acc median {np.median(e):.2e} max {e.max():.2e}, pot max {ep.max():.2e};"
          f" cubic period: acc median {np.median(e_cube):.2e}")
    assert np.median(e) < 1e-4 and e.max() < 3e-3 and ep.max() < 1e-4
    assert np.median(e_cube) > 100 * np.median(e)
