"""Pins the oracle's force path (spline, PP, PC on lists, Ewald) with closed
forms, invariants and convergence, and with the one numeric anchor the
reference holds (teststep/pkdtest.log, PKDGRAV on teststep/king_soft.bin)."""
import json
import os

import numpy as np
import pytest

from changa_amd import configs, host
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- spline
def test_spline_newtonian_outside():
    for r in [0.2, 0.5, 1.0, 3.0]:
        a, b = O.spline(r * r, 0.2)
        assert a == pytest.approx(1 / r, rel=1e-15)
        assert b == pytest.approx(1 / r ** 3, rel=1e-15)


def test_spline_continuity_and_derivative():
    twoh = 0.3
    h = twoh / 2
    # continuity at u = 1 (r = h) and u = 2 (r = twoh)
    for r0 in [h, twoh]:
        lo = O.spline((r0 * (1 - 1e-9)) ** 2, twoh)
        hi = O.spline((r0 * (1 + 1e-9)) ** 2, twoh)
        assert lo[0] == pytest.approx(hi[0], rel=1e-7)
        assert lo[1] == pytest.approx(hi[1], rel=1e-7)
    # b = -(1/r) da/dr
    for r in np.linspace(0.01, 0.45, 23):
        e = 1e-6 * r
        dadr = (O.spline((r + e) ** 2, twoh)[0] - O.spline((r - e) ** 2, twoh)[0]) / (2 * e)
        assert O.spline(r * r, twoh)[1] == pytest.approx(-dadr / r, rel=1e-6)
    # finite at the centre: a(0) = 7/5 / h
    assert O.spline(1e-30, twoh)[0] == pytest.approx(1.4 / h, rel=1e-12)


# ---------------------------------------------------------------- direct sum
def test_direct_sum_momentum_conservation():
    w = configs.plummer(n=512)
    acc, _ = O.direct_sum(w["pos"], w["mass"], w["soft"], np.arange(512))
    f = (w["mass"][:, None] * acc).sum(0)
    assert np.abs(f).max() < 1e-13 * np.abs(w["mass"][:, None] * acc).sum(0).max()


def test_direct_sum_plummer_analytic():
    """Far outside the core, the mean radial acceleration of a Plummer sphere
    approaches -M(<r)/r^2 with M(<r) = r^3/(r^2+a^2)^(3/2)."""
    w = configs.plummer(n=20000)
    pos = w["pos"]
    r = np.linalg.norm(pos, axis=1)
    sel = np.nonzero((r > 2.0) & (r < 3.0))[0][:400]
    acc, _ = O.direct_sum(pos, w["mass"], w["soft"], sel)
    ar = (acc * pos[sel]).sum(1) / r[sel]
    model = -r[sel] / (r[sel] ** 2 + 1.0) ** 1.5
    assert np.median(np.abs(ar / model - 1)) < 0.05


# ---------------------------------------------------------------- tree forces
def _setup(w, theta=None):
    tr = host.build_tree(w["pos"], w["bucket"])
    perm = tr["perm"]
    pos = np.ascontiguousarray(w["pos"][perm])
    m = np.ascontiguousarray(w["mass"][perm])
    s = np.ascontiguousarray(w["soft"][perm])
    geo = host.node_geometry(tr, pos, m, s)
    lists = host.walk(tr, geo, theta if theta is not None else w["theta"])
    return tr, pos, m, s, geo, lists


def test_tree_converges_to_direct_sum():
    w = configs.plummer(n=3000)
    tr, pos, m, s, geo, _ = _setup(w)
    g, mom = O.build_moments(tr, pos, m, s)
    ad, pd = O.direct_sum(pos, m, s, np.arange(len(pos)))
    med = []
    for th in [0.8, 0.55, 0.35]:
        lists = host.walk(tr, geo, th)
        acc, pot, _, _ = O.forces_lists(tr, g, mom, pos, m, s, lists)
        e = np.linalg.norm(acc - ad, axis=1) / np.linalg.norm(ad, axis=1)
        med.append(np.median(e))
    assert med[0] > med[1] > med[2]
    assert med[2] < 1e-5 and med[0] < 2e-3, med


def test_softened_cells_fall_back_to_monopole():
    """With large softening many accepted cells overlap the bucket's
    softening sphere; the oracle then uses the spline monopole
    (gravity.h:414-421) -- which must still track the direct sum."""
    w = configs.plummer(n=2000, eps=0.5)
    tr, pos, m, s, geo, lists = _setup(w)
    g, mom = O.build_moments(tr, pos, m, s)
    acc, pot, _, cnt = O.forces_lists(tr, g, mom, pos, m, s, lists)
    assert cnt[1] > 0
    ad, pd = O.direct_sum(pos, m, s, np.arange(len(pos)))
    e = np.linalg.norm(acc - ad, axis=1) / np.linalg.norm(ad, axis=1)
    assert np.median(e) < 1e-3


def test_pkdgrav_energy_anchor():
    """teststep/pkdtest.log t=0 (PKDGRAV, theta 0.7, hexadecapole, nBucket 8):
    potential energy 1/2 sum m phi = -64.84776 on king_soft.bin."""
    ref = json.load(open(os.path.join(GOLD, "pkdtest_t0.json")))
    k = np.load(os.path.join(GOLD, "king_soft.npz"))
    w = dict(pos=k["pos"].astype(np.float64), mass=k["mass"].astype(np.float64),
             soft=k["eps"].astype(np.float64), bucket=8, theta=0.7)
    tr, pos, m, s, geo, lists = _setup(w)
    g, mom = O.build_moments(tr, pos, m, s)
    acc, pot, _, _ = O.forces_lists(tr, g, mom, pos, m, s, lists)
    pe = 0.5 * (m * pot).sum()
    assert pe == pytest.approx(ref["potential"], rel=2e-5)
    ke = 0.5 * (w["mass"] * (k["vel"].astype(np.float64) ** 2).sum(1)).sum()
    assert ke == pytest.approx(ref["kinetic"], rel=1e-5)


# ---------------------------------------------------------------- Ewald
MADELUNG_SC = 2.837297479480620  # simple-cubic lattice with neutralizing background


def test_ewald_exact_madelung():
    """One particle in a unit periodic box: its images plus the neutralizing
    background give phi = +2.8372974795 m / L and no force."""
    pos = np.array([[0.3, 0.6, 0.2]])
    acc, pot = O.ewald_exact(pos, np.ones(1), np.full(1, 1e-4), np.array([0]), L=1.0)
    assert pot[0] == pytest.approx(MADELUNG_SC, rel=1e-8)
    assert np.abs(acc).max() < 1e-10


def test_tree_plus_bucket_ewald_madelung():
    """The same constant from ChaNGa's scheme: the tree's 26 explicit replica
    images plus BucketEwald's real-space hole (-erf) and k-space terms."""
    pos = np.array([[0.3, 0.6, 0.2]])
    m, s = np.ones(1), np.full(1, 1e-4)
    tr = host.build_tree(pos, 16, box_lo=np.zeros(3), box_size=1.0)
    geo = host.node_geometry(tr, pos, m, s)
    lists = host.walk(tr, geo, 0.7, n_replicas=1, period=1.0)
    g, mom = O.build_moments(tr, pos, m, s)
    acc, pot, _, _ = O.forces_lists(tr, g, mom, pos, m, s, lists, period=1.0)
    ewt = O.ewald_table(mom[0], g[0, 3], 1.0, np.sqrt(10.0))
    assert len(ewt) == 146
    ea, ep, _ = O.bucket_ewald([0], pos, g[0, 0:3], g[0, 5], mom[0], g[0, 3], ewt, 1.0, 1, 2.6)
    assert pot[0] + ep[0] == pytest.approx(MADELUNG_SC, rel=2e-6)
    assert np.abs(acc + ea).max() < 1e-6


def test_lattice_has_no_force():
    w = configs.zeldovich(ng=8, rms=0.0)
    pos = w["pos"]
    acc, _ = O.ewald_exact(pos, w["mass"], w["soft"], np.arange(0, 512, 37), L=1.0)
    assert np.abs(acc).max() < 1e-9


def test_tree_ewald_matches_exact_periodic():
    """Tree (explicit replicas) + BucketEwald against the converged FP64 Ewald
    sum on a small displaced lattice, at small theta so the tree error is out
    of the way.

    Forces: ChaNGa's correction represents the box by its root hexadecapole,
    whose truncation leaves a floor that shrinks when the explicit replica
    shell grows (nReplicas 1 -> 2).
    Potential: the correction is expanded in trace-free moments, exact only
    for a harmonic kernel; here the kernel's Laplacian is the constant 4 pi/V
    of the neutralizing background, so the trace of the root quadrupole adds
    the constant (2 pi / 3V) sum m |x - cm|^2 to every potential."""
    w = configs.zeldovich(ng=12, rms=0.02)
    pos, m, s = w["pos"], w["mass"], w["soft"]
    tr = host.build_tree(pos, 16, box_lo=np.zeros(3), box_size=1.0)
    perm = tr["perm"]
    pos, m, s = (np.ascontiguousarray(a[perm]) for a in (pos, m, s))
    geo = host.node_geometry(tr, pos, m, s)
    g, mom = O.build_moments(tr, pos, m, s)
    ewt = O.ewald_table(mom[0], g[0, 3], 1.0, np.sqrt(10.0))
    targets = np.arange(0, len(pos), 7)
    xa, xp = O.ewald_exact(pos, m, s, targets, L=1.0)
    offset = 2 * np.pi / 3 * (m * ((pos - g[0, :3]) ** 2).sum(1)).sum()
    med = []
    for R in (1, 2):
        lists = host.walk(tr, geo, 0.2, n_replicas=R, period=1.0)
        acc, pot, _, _ = O.forces_lists(tr, g, mom, pos, m, s, lists, period=1.0)
        ea, ep, _ = O.bucket_ewald(targets, pos, g[0, 0:3], g[0, 5], mom[0], g[0, 3], ewt, 1.0, R,
                                   2.6)
        at = acc[targets] + ea[targets]
        e = np.linalg.norm(at - xa, axis=1) / np.linalg.norm(xa, axis=1)
        med.append(np.median(e))
        dp = pot[targets] + ep[targets] - xp
        assert np.abs(dp - offset).max() < 1e-3 * offset
    assert med[0] < 1e-3 and med[1] < 3e-5 and med[0] / med[1] > 10, med


def test_ewald_table_host_matches_oracle():
    """Two independent statements of EwaldInit (Ewald.cpp:285-383): the
    product's host table and the oracle's."""
    rng = np.random.default_rng(4)
    q = rng.normal(size=22)
    q[0] = 1.0
    a = host.ewald_table(q, 0.8, 1.0, np.sqrt(10.0))
    b = O.ewald_table(q, 0.8, 1.0, np.sqrt(10.0))
    assert a.shape == b.shape == (146, 5)
    assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    # default dEwhCut = 2.8 gives the reference GPU's 80-entry table
    assert len(host.ewald_table(q, 0.8, 1.0, 2.8)) == 80
