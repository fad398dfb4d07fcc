"""Host-side tree and interaction-list invariants (CPU)."""
import numpy as np
import pytest

from changa_amd import configs, engine, host
from oracle import oracle as O


def _build(w, theta=None, R=0, period=1.0):
    if w["periodic"]:
        tr = host.build_tree(w["pos"], w["bucket"], box_lo=np.zeros(3), box_size=w["L"])
    else:
        tr = host.build_tree(w["pos"], w["bucket"])
    perm = tr["perm"]
    pos = np.ascontiguousarray(w["pos"][perm])
    m = np.ascontiguousarray(w["mass"][perm])
    s = np.ascontiguousarray(w["soft"][perm])
    geo = host.node_geometry(tr, pos, m, s)
    lists = host.walk(tr, geo, theta or w["theta"], R, period)
    return tr, pos, m, s, geo, lists


def test_tree_partitions_particles():
    w = configs.halos(n=20000)
    tr, pos, *_ = _build(w)
    assert sorted(tr["perm"]) == list(range(20000))
    starts, sizes = tr["bucket_start"], tr["bucket_size"]
    assert starts[0] == 0 and (starts[1:] == starts[:-1] + sizes[:-1]).all()
    assert starts[-1] + sizes[-1] == 20000
    assert sizes.max() <= w["bucket"] and sizes.min() >= 1
    ch = tr["node_child"].reshape(-1, 2)
    for nd in range(tr["n_nodes"]):
        c0, c1 = ch[nd]
        if c0 < 0:
            continue
        assert c0 > nd and c1 > nd
        f, c = tr["node_first"], tr["node_count"]
        assert f[c0] == f[nd] and f[c1] == f[c0] + c[c0] and c[c0] + c[c1] == c[nd]
        assert c[c0] > 0 and c[c1] > 0
    lo = tr["level_offsets"]
    assert (tr["node_depth"][lo[:-1]] == np.arange(len(lo) - 1)).all()


def test_geometry_bounds_particles():
    w = configs.plummer(n=5000)
    tr, pos, m, s, geo, _ = _build(w)
    for nd in range(0, tr["n_nodes"], 7):
        f, c = tr["node_first"][nd], tr["node_count"][nd]
        p = pos[f:f + c]
        g = geo[nd]
        assert (p >= g[6:9]).all() and (p <= g[9:12]).all()
        assert np.linalg.norm(p - g[0:3], axis=1).max() <= g[3] * (1 + 1e-12)
        assert g[5] == pytest.approx(m[f:f + c].sum(), rel=1e-13)


@pytest.mark.parametrize("R", [0, 1])
def test_every_bucket_sees_the_whole_tree_once(R):
    """Per bucket and per replica, the particles behind the accepted cells and
    the opened buckets cover the whole domain exactly once (ChaNGa's
    CHANGA_REFACTOR_WALKCHECK, TreePiece::checkWalkCorrectness,
    TreePiece.cpp:5709)."""
    w = configs.zeldovich(ng=16, rms=0.05) if R else configs.plummer(n=6000)
    tr, pos, m, s, geo, L = _build(w, R=R, period=1.0)
    n = pos.shape[0]
    cells = L["cells"].reshape(-1, 2)
    parts = L["parts"].reshape(-1, 3)
    for b in range(0, tr["n_buckets"], max(1, tr["n_buckets"] // 40)):
        cover = {}
        for nd, oid in cells[L["cell_marks"][b]:L["cell_marks"][b + 1]]:
            code = (oid >> 22) & 0x1ff
            f, c = tr["node_first"][nd], tr["node_count"][nd]
            cover.setdefault(code, np.zeros(n, np.int32))[f:f + c] += 1
            assert (oid & 0x3fffff) == b
        for f, off, c in parts[L["part_marks"][b]:L["part_marks"][b + 1]]:
            code = (off >> 22) & 0x1ff
            cover.setdefault(code, np.zeros(n, np.int32))[f:f + c] += 1
        assert len(cover) == (2 * R + 1) ** 3
        for v in cover.values():
            assert (v == 1).all()


def test_softened_flag_matches_open_softening():
    w = configs.plummer(n=3000, eps=0.5)
    tr, pos, m, s, geo, L = _build(w)
    cells = L["cells"].reshape(-1, 2)
    nflag = 0
    for b in range(tr["n_buckets"]):
        bg = geo[tr["bucket_node"][b]]
        lo, hi, ss = L["cell_marks"][b], L["cell_marks"][b + 1], L["soft_start"][b]
        for e in range(lo, hi):
            nd, oid = cells[e]
            flagged = bool(oid < 0)  # bit 31 of the int32 offsetID
            assert flagged == (e >= ss)
            assert flagged == O.open_softening(geo[nd], bg, (0.0, 0.0, 0.0))
            nflag += flagged
    assert nflag > 0


def test_accepted_cells_pass_opening_test():
    """Every accepted cell's opening sphere misses the bucket box (gravity.h:
    597-635); every opened bucket's does not (or it has <= 6 particles)."""
    w = configs.plummer(n=4000)
    tr, pos, m, s, geo, L = _build(w)
    k = 2 / np.sqrt(3)
    cells = L["cells"].reshape(-1, 2)

    def box_dist2(lo, hi, c):
        e = np.maximum(0, np.maximum(lo - c, c - hi))
        return (e * e).sum()

    for b in range(0, tr["n_buckets"], 9):
        bg = geo[tr["bucket_node"][b]]
        for nd, oid in cells[L["cell_marks"][b]:L["cell_marks"][b + 1]]:
            g = geo[nd]
            assert tr["node_count"][nd] > 6
            rad = max(k * g[3] / w["theta"], g[3])
            assert box_dist2(bg[6:9], bg[9:12], g[0:3]) > rad * rad


def test_expand_parts_matches_ilpart_runs():
    w = configs.plummer(n=3000)
    tr, pos, m, s, geo, L = _build(w)
    cells, marks = engine.expand_parts(L)
    p = L["parts"].reshape(-1, 3)
    assert marks[-1] == p[:, 2].sum() == L["n_part_sources"]
    c = cells.reshape(-1, 2)
    b = 5
    got = c[marks[b]:marks[b + 1], 0]
    want = np.concatenate([np.arange(f, f + n) for f, _, n in p[L["part_marks"][b]:L["part_marks"][b + 1]]])
    assert (got == want).all()


def test_ewald_real_counts_bruteforce():
    rng = np.random.default_rng(2)
    pos = rng.uniform(size=(50, 3))
    cm = np.array([0.5, 0.5, 0.5])
    c = engine.ewald_real_counts(pos, cm, 1.0, 1, 2.6)
    for i in range(0, 50, 10):
        n = 0
        for ix in range(-3, 4):
            for iy in range(-3, 4):
                for iz in range(-3, 4):
                    hole = max(abs(ix), abs(iy), abs(iz)) <= 1
                    r2 = ((pos[i] - cm + [ix, iy, iz]) ** 2).sum()
                    n += hole or r2 <= 2.6 ** 2
        assert c[i] == n
