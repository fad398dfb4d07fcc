"""Adjoint form on the GPU (changaGravityAdjoint) against the FP64 oracle's
adjoint on the same tree and lists, and, through the test-harness
torch.autograd.Function, against central finite differences of the FP64
oracle's forward evaluation over a kick-drift-kick step.

Bars. The GPU evaluates every pair term in FP32 on bucket-centred
coordinates and accumulates in FP64; per particle, dL/dx and dL/dm are sums
of target-side and source-side terms of either sign. Against the oracle's
adjoint: median relative error <= 2e-5, 99th percentile <= 5e-4 (per
particle, relative to that particle's gradient), at 1.5x the largest tails
measured over the cases: median 1.5e-6 (measured up to 9.4e-7, the softened
sphere's dL/dx) and 99th percentile 6e-5 (4.0e-5, its dL/dm); dL/dm has
zeros, near which its relative error has no useful maximum. Against finite
differences of the FP64 forward: <= 1e-4 relative to the largest gradient
component of the sampled particle.
"""
import numpy as np
import pytest

from changa_amd import configs, engine
from oracle import oracle as O
from tests.helpers import check_tails, summary

pytestmark = pytest.mark.gpu


def _oracle_grad(ev, w, ewald=False):
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    ge, hroot = None, None
    if ewald:
        ge, hroot = O.adjoint_ewald(ev.pos, w, geo, mom, ev.ewt64, ev.L, ev.w["n_replicas"],
                                    ev.w["ew_cut"])
    g = O.adjoint_lists(ev.tree, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists, w,
                        period=ev.L if ev.w["periodic"] else 0.0, root_extra=hroot)
    return g if ge is None else g + ge


def _errors(g, go):
    ex = np.linalg.norm(g[:, :3] - go[:, :3], axis=1) / np.linalg.norm(go[:, :3], axis=1)
    em = np.abs(g[:, 3] - go[:, 3]) / np.abs(go[:, 3])
    return ex, em


CASES = {
    "plummer": lambda: (configs.plummer(n=16384), True),
    "softened": lambda: (configs.plummer(n=4096, eps=0.5), False),
    "periodic": lambda: (configs.zeldovich(ng=16, rms=0.03), True),
    "halos": lambda: (configs.halos(n=32768), True),
}


@pytest.mark.parametrize("case", list(CASES))
def test_gpu_adjoint_matches_oracle(device, case):
    import torch
    w, compact = CASES[case]()
    ev = engine.Evaluation(w, precise=True, pp_compact=compact, device=device)
    if case == "softened":
        assert ev.counts[1] > 0
    wv = np.random.default_rng(21).standard_normal(ev.pos.shape)
    g = ev.adjoint(wv, ewald=False).cpu().numpy()
    torch.cuda.synchronize()
    go = _oracle_grad(ev, wv)
    ex, em = _errors(g, go)
    print(f"adjoint {case}: dL/dx {summary(ex)} dL/dm {summary(em)}", flush=True)
    for name, e in (("dL/dx", ex), ("dL/dm", em)):
        check_tails(e, 1.5e-6, 6e-5, None, label=f"adjoint {case} {name}")
    # translation invariance of the tree force
    tot = np.abs(g[:, :3].sum(axis=0)).max() / np.abs(g[:, :3]).sum(axis=0).max()
    assert tot < 1e-5, tot


@pytest.mark.parametrize("active,reps", [(False, 1), (True, 1), (False, 2)])
def test_gpu_adjoint_ewald_matches_oracle(device, active, reps):
    """Periodic box, L over tree plus Ewald acceleration: the Ewald term's
    target side, and its source side through the root's moments, centre of
    mass and h-table, against the oracle's adjoint_ewald added to the list
    adjoint (root_extra); with one and with two replica shells (BucketEwald's
    hole follows the shells)."""
    import torch
    w = dict(configs.zeldovich(ng=16, rms=0.05, active_frac=0.3 if active else None),
             n_replicas=reps)
    ev = engine.Evaluation(w, precise=True, device=device)
    wt = np.random.default_rng(22).standard_normal((ev.n_targets, 3))
    g = ev.adjoint(wt).cpu().numpy()
    torch.cuda.synchronize()
    wv = np.zeros_like(ev.pos)
    wv[ev.t_index if ev.t_index is not None else slice(None)] = wt
    go = _oracle_grad(ev, wv, ewald=True)
    ex, em = _errors(g, go)
    print(f"adjoint ewald active={active} R={reps}: dL/dx {summary(ex)} dL/dm {summary(em)}", flush=True)
    for name, e in (("dL/dx", ex), ("dL/dm", em)):
        check_tails(e, 1.5e-6, 6e-5, None, label=f"adjoint ewald {name}")
    tot = np.abs(g[:, :3].sum(axis=0)).max() / np.abs(g[:, :3]).sum(axis=0).max()
    assert tot < 1e-5, tot


def test_gpu_adjoint_repeatable(device):
    """The transposed cell lists come from a stable sort (a cell's users in
    bucket order), so every FP32 partial sum is the same from run to run;
    only the order of FP64 atomic additions varies."""
    w = configs.plummer(n=8192)
    ev = engine.Evaluation(w, precise=True, device=device)
    wv = np.random.default_rng(3).standard_normal(ev.pos.shape)
    g1 = ev.adjoint(wv).cpu().numpy()
    g2 = ev.adjoint(wv).cpu().numpy()
    assert np.abs(g1 - g2).max() <= 1e-12 * np.abs(g1).max()


def _oracle_acc(tree, lists, pos, mass, soft):
    geo, mom = O.build_moments(tree, pos, mass, soft)
    acc, _, _, _ = O.forces_lists(tree, geo, mom, pos, mass, soft, lists, soft_flags=True)
    return acc


def test_gpu_adjoint_kdk_finite_differences(device):
    """One kick-drift-kick step of a 4,096-particle Plummer sphere, L = u . v1
    (every term of its gradient goes through the forces), differentiated
    through both force evaluations by the autograd wrapper (positions
    checkpointed, trees and lists rebuilt in backward), against central
    differences (step 1e-4 eps in position, 1e-4 m in mass) of the same step
    evaluated by the FP64 oracle on the frozen nominal trees, on 64 seeded
    particles. Bar: 1e-4 of the largest component, beyond three times the
    differences' own rounding noise (the spread between steps h and 2 h)."""
    import torch

    from tests.gravity_autograd import TreeGravity

    cfg = configs.plummer(n=4096)
    n = cfg["pos"].shape[0]
    rng = np.random.default_rng(8)
    v0 = 0.3 * rng.standard_normal((n, 3))
    wv = rng.standard_normal((n, 3))
    dt = 0.01
    meta = {k: cfg[k] for k in cfg if k not in ("pos", "mass")}

    x0 = torch.tensor(cfg["pos"], dtype=torch.float64, requires_grad=True)
    m = torch.tensor(cfg["mass"], dtype=torch.float64, requires_grad=True)
    a0 = TreeGravity.apply(x0, m, meta)
    vh = torch.tensor(v0) + 0.5 * dt * a0
    x1 = x0 + dt * vh
    a1 = TreeGravity.apply(x1, m, meta)
    v1 = vh + 0.5 * dt * a1
    L = (torch.tensor(wv) * v1).sum()
    L.backward()
    gx, gm = x0.grad.numpy(), m.grad.numpy()

    # the nominal trees and lists, frozen for the differences
    from changa_amd import host

    def tree_of(pos):
        tr = host.build_tree(pos, cfg["bucket"])
        perm = tr["perm"]
        geo = host.node_geometry(tr, pos[perm], cfg["mass"][perm], cfg["soft"][perm])
        lists = host.walk(tr, geo, cfg["theta"], 0, 1.0, 0)
        return tr, perm, lists

    x0n = cfg["pos"]
    t0, p0, l0 = tree_of(x0n)
    a0n = np.empty_like(x0n)
    a0n[p0] = _oracle_acc(t0, l0, x0n[p0], cfg["mass"][p0], cfg["soft"][p0])
    x1n = x0n + dt * (v0 + 0.5 * dt * a0n)
    t1, p1, l1 = tree_of(x1n)

    def loss(xs, ms):
        a0 = np.empty_like(xs)
        a0[p0] = _oracle_acc(t0, l0, xs[p0], ms[p0], cfg["soft"][p0])
        vh = v0 + 0.5 * dt * a0
        x1 = xs + dt * vh
        a1 = np.empty_like(xs)
        a1[p1] = _oracle_acc(t1, l1, x1[p1], ms[p1], cfg["soft"][p1])
        v1 = vh + 0.5 * dt * a1
        return float((wv * v1).sum())

    hx = 1e-4 * cfg["soft"][0]  # step 1e-4 eps
    worst = np.zeros(2)
    for i, k in enumerate(rng.choice(n, size=64, replace=False)):
        fd, fd2 = np.zeros(4), np.zeros(3)
        for c in range(3):
            for step, out in ((hx, fd), (2 * hx, fd2)):
                xp, xm = x0n.copy(), x0n.copy()
                xp[k, c] += step
                xm[k, c] -= step
                out[c] = (loss(xp, cfg["mass"]) - loss(xm, cfg["mass"])) / (2 * step)
        mp, mm = cfg["mass"].copy(), cfg["mass"].copy()
        h = 1e-4 * cfg["mass"][k]
        mp[k] += h
        mm[k] -= h
        fd[3] = (loss(x0n, mp) - loss(x0n, mm)) / (2 * h)
        # the differences' own rounding: L sums 4,096 terms of order one, so
        # at a step of 1e-6 the quotient carries ~1e-9 of absolute noise,
        # measured here as the spread between the steps h and 2 h (whose
        # truncation errors are far below it)
        noise = np.abs(fd[:3] - fd2).max()
        ex = max(0.0, np.abs(fd[:3] - gx[k]).max() - 3 * noise) / np.abs(fd[:3]).max()
        em = abs(fd[3] - gm[k]) / abs(fd[3])
        worst = np.maximum(worst, [ex, em])
        if i % 16 == 15:
            print(f"kdk: {i + 1} of 64 particles, worst dL/dx rel {worst[0]:.2e} "
                  f"dL/dm rel {worst[1]:.2e}", flush=True)
        assert ex <= 1e-4 and em <= 1e-4, (k, fd, gx[k], gm[k])


def test_gpu_adjoint_full_size_plummer(device):
    """The full 2^20-particle Plummer sphere (tree and lists built on the
    GPU), L = sum w . a with seeded w ~ N(0,1)^3: the GPU gradient against
    the FP64 oracle's adjoint on the same tree and lists, on 4,096 seeded
    particles (the oracle computes every particle's gradient: minutes of
    FP64 work on the host cores)."""
    import os
    import time

    import torch
    ev = engine.Evaluation(configs.workload(4), precise=True, device=device, lists="gpu",
                           tree="gpu")
    wv = np.random.default_rng(23).standard_normal(ev.pos.shape)
    g = ev.adjoint(wv).cpu().numpy()
    torch.cuda.synchronize()
    t = time.time()
    threads = min(16, len(os.sched_getaffinity(0)))
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    print("adjoint full size: oracle moments done", flush=True)
    go = O.adjoint_lists(ev.tree, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists, wv, period=0.0,
                         n_threads=threads)
    print(f"adjoint full size: oracle adjoint {time.time() - t:.1f} s on {threads} threads",
          flush=True)
    samp = np.random.default_rng(24).choice(ev.pos.shape[0], size=4096, replace=False)
    ex, em = _errors(g[samp], go[samp])
    print(f"adjoint full size 2^20: dL/dx {summary(ex)} dL/dm {summary(em)}", flush=True)
    check_tails(ex, 5.8e-7, 6.9e-6, 2.1e-5, label="full-size adjoint dL/dx")
    check_tails(em, 5.4e-7, 4.4e-5, None, label="full-size adjoint dL/dm")
