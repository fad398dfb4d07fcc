"""The oracle's adjoint (orc_adjoint_lists, orc_direct_grad) pinned by what
fixes it independently of its derivation:
  - central finite differences of the oracle's own forward evaluation on the
    same tree and lists (moments rebuilt from the perturbed particles,
    softened cells as flagged in the lists);
  - translation invariance: sum_k dL/dx_k = 0 (this checks the target-side
    and the cell/particle source-side terms against each other, including the
    centre-of-mass coupling);
  - Euler's theorem for unsoftened gravity: sum_k x_k . dL/dx_k = -2 L;
  - convergence to the exact direct-sum gradient as the opening angle
    shrinks, and the direct-sum gradient itself against finite differences
    of the direct sum;
  - the derivatives of 1/r against closed forms, finite differences and
    Laplace's equation.
"""
import numpy as np
import pytest

from changa_amd import configs, host
from oracle import oracle as O


def _setup(w, theta=None):
    periodic = bool(w["periodic"])
    kw = dict(box_lo=np.zeros(3), box_size=w["L"]) if periodic else {}
    tr = host.build_tree(w["pos"], w["bucket"], **kw)
    perm = tr["perm"]
    pos = np.ascontiguousarray(w["pos"][perm])
    mass = np.ascontiguousarray(w["mass"][perm])
    soft = np.ascontiguousarray(w["soft"][perm])
    geo = host.node_geometry(tr, pos, mass, soft)
    lists = host.walk(tr, geo, w["theta"] if theta is None else theta, w["n_replicas"],
                      w["L"] if periodic else 1.0, 0)
    return tr, pos, mass, soft, lists, (w["L"] if periodic else 0.0)


def _loss(tr, pos, mass, soft, lists, period, wv):
    geo, mom = O.build_moments(tr, pos, mass, soft)
    acc, _, _, _ = O.forces_lists(tr, geo, mom, pos, mass, soft, lists, period=period,
                                  soft_flags=True)
    return float((wv * acc).sum())


def _grad(tr, pos, mass, soft, lists, period, wv):
    geo, mom = O.build_moments(tr, pos, mass, soft)
    return O.adjoint_lists(tr, geo, mom, pos, mass, soft, lists, wv, period=period)


def _fd(tr, pos, mass, soft, lists, period, wv, k, hx, hm):
    out = np.zeros(4)
    for c in range(3):
        p1, p2 = pos.copy(), pos.copy()
        p1[k, c] += hx
        p2[k, c] -= hx
        out[c] = (_loss(tr, p1, mass, soft, lists, period, wv) -
                  _loss(tr, p2, mass, soft, lists, period, wv)) / (2 * hx)
    m1, m2 = mass.copy(), mass.copy()
    m1[k] += hm
    m2[k] -= hm
    out[3] = (_loss(tr, pos, m1, soft, lists, period, wv) -
              _loss(tr, pos, m2, soft, lists, period, wv)) / (2 * hm)
    return out


def _check_fd(w, n_samples, tol, seed=0):
    tr, pos, mass, soft, lists, period = _setup(w)
    rng = np.random.default_rng(seed)
    wv = rng.standard_normal(pos.shape)
    g = _grad(tr, pos, mass, soft, lists, period, wv)
    scale_x = np.abs(g[:, :3]).max()
    scale_m = np.abs(g[:, 3]).max()
    ks = rng.choice(pos.shape[0], size=n_samples, replace=False)
    hx = 1e-5 * (w["L"] if period else 1.0)
    for k in ks:
        fd = _fd(tr, pos, mass, soft, lists, period, wv, k, hx, 1e-4 * mass[k])
        ex = np.abs(fd[:3] - g[k, :3]).max() / scale_x
        em = abs(fd[3] - g[k, 3]) / scale_m
        assert ex < tol and em < tol, (k, fd, g[k], ex, em)
    return g, wv


def test_rinv_derivs_closed_forms_and_laplace():
    r = np.array([0.3, -0.7, 0.45])
    T = O.rinv_derivs(r, 6)
    x, y, z = r
    rr = np.linalg.norm(r)
    assert T[0, 0, 0] == pytest.approx(1 / rr, rel=1e-15)
    assert T[1, 0, 0] == pytest.approx(-x / rr ** 3, rel=1e-14)
    assert T[2, 0, 0] == pytest.approx(3 * x * x / rr ** 5 - 1 / rr ** 3, rel=1e-13)
    assert T[1, 1, 0] == pytest.approx(3 * x * y / rr ** 5, rel=1e-13)
    assert T[1, 1, 1] == pytest.approx(-15 * x * y * z / rr ** 7, rel=1e-13)
    # Laplace: every trace vanishes
    for n in range(2, 7):
        for a in range(n - 1):
            for b in range(n - 1 - a):
                c = n - 2 - a - b
                s = T[a + 2, b, c] + T[a, b + 2, c] + T[a, b, c + 2]
                assert abs(s) < 1e-11 * np.abs(T[a + 2, b, c]) + 1e-12
    # each order is the numerical derivative of the one below
    h = 1e-5
    for (a, b, c) in [(2, 1, 1), (0, 3, 2), (1, 0, 4)]:
        for ax in range(3):
            e = np.zeros(3)
            e[ax] = h
            d = (O.rinv_derivs(r + e, 6)[a, b, c] - O.rinv_derivs(r - e, 6)[a, b, c]) / (2 * h)
            idx = [a, b, c]
            idx[ax] += 1
            assert d == pytest.approx(T[tuple(idx)], rel=1e-6)


def test_direct_grad_matches_finite_differences():
    w = configs.plummer(n=512, eps=0.05)
    pos, mass, soft = w["pos"], w["mass"], w["soft"]
    rng = np.random.default_rng(3)
    wv = rng.standard_normal(pos.shape)
    g = O.direct_grad(pos, mass, soft, wv)
    idx = np.arange(pos.shape[0], dtype=np.int32)

    def loss(p, m):
        a, _ = O.direct_sum(p, m, soft, idx)
        return float((wv * a).sum())

    for k in rng.choice(512, size=6, replace=False):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[k, c] += 1e-5
            p2[k, c] -= 1e-5
            fd = (loss(p1, mass) - loss(p2, mass)) / 2e-5
            assert fd == pytest.approx(g[k, c], rel=1e-6, abs=1e-9 * np.abs(g[:, :3]).max())
        m1, m2 = mass.copy(), mass.copy()
        m1[k] *= 1.0001
        m2[k] *= 0.9999
        fd = (loss(pos, m1) - loss(pos, m2)) / (2e-4 * mass[k])
        assert fd == pytest.approx(g[k, 3], rel=1e-6, abs=1e-9 * np.abs(g[:, 3]).max())


def test_adjoint_finite_differences_plummer():
    _check_fd(configs.plummer(n=2048), n_samples=12, tol=2e-7)


def test_adjoint_finite_differences_softened_cells():
    w = configs.plummer(n=2048, eps=0.5)
    tr, pos, mass, soft, lists, _ = _setup(w)
    assert (lists["cells"].reshape(-1, 2)[:, 1] < 0).any(), "needs softened cells"
    _check_fd(w, n_samples=12, tol=2e-7, seed=1)


def test_adjoint_finite_differences_periodic_replicas():
    w = configs.zeldovich(ng=12, rms=0.05)
    _check_fd(w, n_samples=10, tol=2e-7, seed=2)


def test_adjoint_translation_invariance():
    for w in (configs.plummer(n=4096), configs.plummer(n=2048, eps=0.5),
              configs.halos(n=8192)):
        tr, pos, mass, soft, lists, period = _setup(w)
        wv = np.random.default_rng(4).standard_normal(pos.shape)
        g = _grad(tr, pos, mass, soft, lists, period, wv)
        tot = g[:, :3].sum(axis=0)
        assert np.abs(tot).max() < 1e-11 * np.abs(g[:, :3]).sum(axis=0).max(), tot


def test_adjoint_euler_homogeneity():
    """Unsoftened gravity is homogeneous of degree -2 in the positions, and
    so is the tree expansion on fixed lists: sum_k x_k . dL/dx_k = -2 L."""
    w = configs.plummer(n=4096, eps=1e-9)
    tr, pos, mass, soft, lists, period = _setup(w)
    wv = np.random.default_rng(5).standard_normal(pos.shape)
    g = _grad(tr, pos, mass, soft, lists, period, wv)
    L = _loss(tr, pos, mass, soft, lists, period, wv)
    s = float((pos * g[:, :3]).sum())
    assert s == pytest.approx(-2.0 * L, rel=1e-9)
    # linear in the source masses: a_i is, except through the cell centres;
    # sum_k m_k dL/dm_k = L holds for the direct sum
    gd = O.direct_grad(pos, mass, soft, wv)
    Ld = float((wv * O.direct_sum(pos, mass, soft, np.arange(4096, dtype=np.int32))[0]).sum())
    assert float((mass * gd[:, 3]).sum()) == pytest.approx(Ld, rel=1e-10)


def test_adjoint_converges_to_direct_gradient():
    w = configs.plummer(n=4096)
    errs = []
    for theta in (0.7, 0.4, 0.2):
        tr, pos, mass, soft, lists, period = _setup(w, theta=theta)
        wv = np.random.default_rng(6).standard_normal(pos.shape)
        g = _grad(tr, pos, mass, soft, lists, period, wv)
        gd = O.direct_grad(pos, mass, soft, wv)
        ex = np.linalg.norm(g[:, :3] - gd[:, :3], axis=1) / np.linalg.norm(gd[:, :3], axis=1)
        em = np.abs(g[:, 3] - gd[:, 3]) / np.abs(gd[:, 3])
        errs.append((np.median(ex), np.median(em)))
    assert errs[0][0] < 1e-2 and errs[0][1] < 1e-2, errs
    assert errs[2][0] < errs[1][0] < errs[0][0], errs
    assert errs[2][1] < errs[1][1] < errs[0][1], errs
    assert errs[2][0] < 1e-4 and errs[2][1] < 1e-4, errs


def _periodic_parts(w, pos, mass, soft, tr):
    geo, mom = O.build_moments(tr, pos, mass, soft)
    ewt = O.ewald_table(mom[0], geo[0, 3], w["L"], w["ewh_cut"])
    return geo, mom, ewt


def _loss_periodic(w, tr, pos, mass, soft, lists, wv):
    geo, mom, ewt = _periodic_parts(w, pos, mass, soft, tr)
    acc, _, _, _ = O.forces_lists(tr, geo, mom, pos, mass, soft, lists, period=w["L"],
                                  soft_flags=True)
    ea, _, _ = O.bucket_ewald(np.arange(pos.shape[0]), pos, geo[0, 0:3], geo[0, 5], mom[0],
                              geo[0, 3], ewt, w["L"], w["n_replicas"], w["ew_cut"])
    return float((wv * (acc + ea)).sum())


def _grad_periodic(w, tr, pos, mass, soft, lists, wv):
    geo, mom, ewt = _periodic_parts(w, pos, mass, soft, tr)
    ge, hroot = O.adjoint_ewald(pos, wv, geo, mom, ewt, w["L"], w["n_replicas"], w["ew_cut"])
    g = O.adjoint_lists(tr, geo, mom, pos, mass, soft, lists, wv, period=w["L"],
                        root_extra=hroot)
    return g + ge


def test_adjoint_with_ewald_finite_differences():
    """Tree with replicas plus the Ewald correction: moments, the h-table and
    every term rebuilt from the perturbed particles."""
    w = configs.zeldovich(ng=12, rms=0.05)
    tr, pos, mass, soft, lists, _ = _setup(w)
    rng = np.random.default_rng(9)
    wv = rng.standard_normal(pos.shape)
    g = _grad_periodic(w, tr, pos, mass, soft, lists, wv)
    sx, sm = np.abs(g[:, :3]).max(), np.abs(g[:, 3]).max()
    for k in rng.choice(pos.shape[0], size=8, replace=False):
        fd = np.zeros(4)
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[k, c] += 1e-5
            p2[k, c] -= 1e-5
            fd[c] = (_loss_periodic(w, tr, p1, mass, soft, lists, wv) -
                     _loss_periodic(w, tr, p2, mass, soft, lists, wv)) / 2e-5
        h = 1e-4 * mass[k]
        m1, m2 = mass.copy(), mass.copy()
        m1[k] += h
        m2[k] -= h
        fd[3] = (_loss_periodic(w, tr, pos, m1, soft, lists, wv) -
                 _loss_periodic(w, tr, pos, m2, soft, lists, wv)) / (2 * h)
        ex = np.abs(fd[:3] - g[k, :3]).max() / sx
        em = abs(fd[3] - g[k, 3]) / sm
        assert ex < 2e-7 and em < 2e-7, (k, fd, g[k], ex, em)


def test_adjoint_with_ewald_translation_invariance():
    w = configs.zeldovich(ng=12, rms=0.05)
    tr, pos, mass, soft, lists, _ = _setup(w)
    wv = np.random.default_rng(10).standard_normal(pos.shape)
    g = _grad_periodic(w, tr, pos, mass, soft, lists, wv)
    tot = g[:, :3].sum(axis=0)
    assert np.abs(tot).max() < 1e-11 * np.abs(g[:, :3]).sum(axis=0).max(), tot


def test_adjoint_with_ewald_series_region():
    """A target inside the Ewald series region (|x - cm_root|^2 < 1.2e-3 L^2,
    where the central replica uses the two-term series): its derivative
    follows the series' own dependence on r^2."""
    w = dict(configs.zeldovich(ng=12, rms=0.05))
    pos = w["pos"].copy()
    others = np.delete(np.arange(pos.shape[0]), 0)
    cm = pos[others].mean(axis=0)
    pos[0] = cm + np.array([0.012, -0.018, 0.01]) * w["L"]
    w["pos"] = pos
    tr, tpos, mass, soft, lists, _ = _setup(w)
    geo, mom = O.build_moments(tr, tpos, mass, soft)
    d2 = ((tpos - geo[0, 0:3]) ** 2).sum(axis=1)
    k = int(d2.argmin())
    assert d2[k] < 1.2e-3 * w["L"] ** 2
    wv = np.random.default_rng(11).standard_normal(tpos.shape)
    g = _grad_periodic(w, tr, tpos, mass, soft, lists, wv)
    sx = np.abs(g[:, :3]).max()
    for c in range(3):
        p1, p2 = tpos.copy(), tpos.copy()
        p1[k, c] += 1e-6
        p2[k, c] -= 1e-6
        fd = (_loss_periodic(w, tr, p1, mass, soft, lists, wv) -
              _loss_periodic(w, tr, p2, mass, soft, lists, wv)) / 2e-6
        assert abs(fd - g[k, c]) / sx < 2e-7, (c, fd, g[k, c])
