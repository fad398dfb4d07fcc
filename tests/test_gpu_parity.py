"""GPU parity: the HIP kernels through libchanga_gpu.so against the FP64
oracle on identical trees and interaction lists.

Bars (per particle, relative to the oracle's |a|, |phi| and dtGrav): the
median, 99th percentile and max of each case at about 1.5x the measured
tails (check_tails), so a regression of the tail fails. The measured tails:
  - isolated systems, precision form (FP64 bucket-centre subtraction, FP32
    interactions): acceleration median 5e-8, max 7e-6 (the 2^16 Plummer
    sphere); potential max 2.3e-7;
  - isolated systems, plain form (FP32 absolute coordinates, the reference's
    wire layout and what the HostCUDA mirror returns): acceleration median
    7e-8, max 1.0e-5 (close pairs lose digits to absolute FP32 coordinates,
    as in the reference's GPU path);
  - periodic boxes: the tree-with-replicas force and the Ewald correction are
    each ~10x the net force and cancel; FP32 terms are exact to ~1e-7 of
    their own size, so the acceleration has median 2e-6, 99th percentile
    1.1e-5 and a max of 5e-5 to 8e-5 where the net force nearly vanishes;
  - dtGrav (a max of FP32 terms): as the acceleration of the same form.
"""
import numpy as np
import pytest

from changa_amd import configs, engine
from oracle import oracle as O
from tests.helpers import check_tails, deep_chain_workload, oracle_forces, rel_err, summary

pytestmark = pytest.mark.gpu


def _check(ev, label, bars):
    """bars: (median, p99, max) for acc, pot and dtGrav."""
    import torch
    ev.step()
    torch.cuda.synchronize()
    acc, pot, dtg = ev.results()
    oa, op, od = oracle_forces(ev)
    ea, ep, ed = rel_err(acc, oa), rel_err(pot, op), rel_err(dtg, od)
    print(f"{label}: acc {summary(ea)} pot {summary(ep)} dtgrav {summary(ed)}", flush=True)
    for name, e, b in (("acc", ea, bars[0]), ("pot", ep, bars[1]), ("dtgrav", ed, bars[2])):
        check_tails(e, *b, label=f"{label} {name}")


@pytest.mark.parametrize("case", ["plummer", "deep_chain_gpu_tree"])
def test_gpu_moments_match_oracle(device, case):
    """GPU multipole build against the oracle's; the deep chain (61 levels,
    one 950-particle bucket) on the GPU-built tree with device gathers. The
    two builds round differently (the GPU shifts axis by axis, the oracle by
    the binomial sum) and a chain node carries up to 60 shift-and-project
    steps of rounding: per-row bar 1e-9 there, 1e-10 on the sphere."""
    import torch
    if case == "plummer":
        ev = engine.Evaluation(configs.plummer(n=4096), device=device)
    else:
        ev = engine.Evaluation(deep_chain_workload(), device=device, tree="gpu")
    torch.cuda.synchronize()
    g = ev.node64()
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    np.testing.assert_allclose(g[:, 0:12], geo, rtol=1e-12, atol=1e-14)
    scale = np.abs(mom).max(axis=1, keepdims=True)
    assert np.abs(g[:, 12:] - mom).max() / scale.max() < 1e-11
    assert (np.abs(g[:, 12:] - mom) / scale).max() < (1e-10 if case == "plummer" else 1e-9)


@pytest.mark.parametrize("precise", [True, False])
def test_gpu_plummer_65536(device, precise):
    ev = engine.Evaluation(configs.plummer(n=65536), precise=precise, device=device)
    if precise:
        _check(ev, "plummer precise", ((8.3e-8, 5e-7, 1.1e-5), (4.2e-8, 1.7e-7, 3.5e-7),
                                       (2.1e-7, 2.9e-6, 1.1e-5)))
    else:
        _check(ev, "plummer plain", ((1.1e-7, 1.3e-6, 1.5e-5), (4.2e-8, 1.7e-7, 3.5e-7),
                                     (1.1e-6, 9.2e-6, 2.7e-5)))


def test_gpu_softened_cells(device):
    w = configs.plummer(n=4096, eps=0.5)
    ev = engine.Evaluation(w, precise=True, device=device)
    assert ev.counts[1] > 0, "workload must exercise softened cells"
    _check(ev, "softened", ((3.8e-7, 7.3e-7, 8.4e-7), (8.1e-8, 2.2e-7, 2.7e-7),
                            (8.3e-8, 2.2e-6, 4.3e-6)))


def test_gpu_pp_list_formats_agree(device):
    import torch
    w = configs.plummer(n=8192)
    a = engine.Evaluation(w, pp_compact=True, device=device)
    b = engine.Evaluation(w, pp_compact=False, device=device)
    for ev in (a, b):
        ev.zero()
        ev.run_pp()
    torch.cuda.synchronize()
    # same sources and order; only where FP32 tile sums fold into FP64 differs
    va, vb = a.d_vars.cpu().double(), b.d_vars.cpu().double()
    assert ((va - vb).abs() <= 1e-6 * vb.abs().amax(dim=0)).all()


@pytest.mark.parametrize("precise", [True, False])
def test_gpu_periodic_ewald(device, precise):
    w = configs.zeldovich(ng=24, rms=0.02)
    ev = engine.Evaluation(w, precise=precise, device=device)
    if precise:
        _check(ev, "periodic precise", ((3.2e-6, 1.7e-5, 1.3e-4), (2.2e-6, 7.9e-6, 1.25e-5),
                                        (1.9e-7, 8.2e-7, 1.6e-6)))
    else:
        _check(ev, "periodic plain", ((3.3e-6, 1.7e-5, 9.8e-5), (2.2e-6, 7.9e-6, 1.2e-5),
                                      (1.5e-6, 6.9e-6, 1.3e-5)))


@pytest.mark.parametrize("reps", [2, 3])
def test_gpu_periodic_more_replica_shells(device, reps):
    """Two and three explicit replica shells (the tree walks 125 / 343
    replicas; BucketEwald's hole grows to match) through the whole step, PC +
    PP + Ewald, against the FP64 oracle on the same lists, at the periodic
    bars. The potential's tree part grows with the number of shells (the
    replicas' 1/r sums) while the net periodic potential does not, so the
    FP32 partial sums carry more of the net: its median bar is 1e-5 at three
    shells (measured 5.8e-6; 3.0e-6 at two)."""
    import torch
    w = dict(configs.zeldovich(ng=12, rms=0.03), n_replicas=reps)
    ev = engine.Evaluation(w, precise=True, device=device, lists="gpu")
    ev.step()
    torch.cuda.synchronize()
    acc, pot, _ = ev.results()
    oa, op, _ = oracle_forces(ev)
    ea, ep = rel_err(acc, oa), rel_err(pot, op)
    print(f"periodic R={reps}: acc {summary(ea)} pot {summary(ep)}", flush=True)
    bars = {2: ((2.8e-6, 1.6e-5, 5e-5), (4.4e-6, 1.6e-5, 2.1e-5)),
            3: ((3.3e-6, 2.2e-5, 4.6e-5), (8.8e-6, 3e-5, 3.7e-5))}[reps]
    check_tails(ea, *bars[0], label=f"R={reps} acc")
    check_tails(ep, *bars[1], label=f"R={reps} pot")


@pytest.mark.parametrize("cfg", [2, 3, 4, 5])
def test_gpu_full_size_sampled_oracle(device, cfg):
    """The full-size workloads (the 128^3 box, the 2^22-particle halos, the
    2^20 Plummer sphere, the 256^3 box with a 10% active rung), tree and
    lists built on the GPU, through the batched step (with the overlapped
    Ewald pass for the boxes): the targets of 128 seeded buckets against the
    FP64 oracle on the same lists, at 1.5x the measured tails. The periodic
    potential (zero mean over the box) crosses zero among the clustered
    box's particles, so the potential's bars are on its error relative to the
    sample's rms potential; the per-particle error is printed beside it."""
    import torch
    ev = engine.Evaluation(configs.workload(cfg), precise=True, device=device, lists="gpu",
                           tree="gpu")
    ev.batch_plan()
    ev.step_batched()
    torch.cuda.synchronize()
    acc, pot, _ = ev.results()
    tr = ev.tree
    rng = np.random.default_rng(cfg)
    bsel = np.sort(rng.choice(tr["n_buckets"], size=128, replace=False))
    oa, op, _ = oracle_forces(ev, buckets=bsel)
    idx = np.concatenate([np.arange(tr["bucket_start"][b], tr["bucket_start"][b] + tr["bucket_size"][b])
                          for b in bsel])
    if ev.active is not None:
        idx = idx[ev.active[idx]]
    ea = rel_err(acc[idx], oa[idx])
    epp = rel_err(pot[idx], op[idx])
    ep = np.abs(pot[idx] - op[idx]) / np.sqrt(np.mean(op[idx] ** 2))
    print(f"full size {cfg}: {len(idx)} targets, acc {summary(ea)} pot per particle "
          f"{summary(epp)} pot over rms {summary(ep)}", flush=True)
    bars = {2: ((3e-6, 2.2e-5, 9.6e-5), (2e-6, 6.9e-6, 1.03e-5)),
            3: ((1.04e-7, 5.5e-7, 1.3e-6), (3.6e-8, 1.4e-7, 1.85e-7)),
            4: ((7.3e-8, 2.5e-7, 3.8e-7), (2.8e-8, 1.04e-7, 1.3e-7)),
            5: ((4e-7, 2.1e-6, 4.3e-6), (2.6e-6, 8.1e-6, 8.7e-6))}[cfg]
    check_tails(ea, *bars[0], label=f"full size {cfg} acc")
    check_tails(ep, *bars[1], label=f"full size {cfg} pot")
    ev.free_batch()
