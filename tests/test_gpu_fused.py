"""Fused walk and force (changaFusedGravity) against the list path
(changaBuildLists* + changaNodeGravity + changaPartGravity) on the same node
records: the walk accepts the same interactions, so the accelerations and
potentials agree up to the FP32 order of the per-tile sums and dtGrav (a
max over the same per-interaction terms) is identical; and the fused forces
against the FP64 oracle with the bars of the periodic parity test."""
import numpy as np
import pytest

from changa_amd import configs, engine
from tests.helpers import check_tails, deep_chain_workload, oracle_forces, rel_err, summary

pytestmark = pytest.mark.gpu

CASES = {
    "plummer": lambda: configs.plummer(n=32768),
    "softened": lambda: configs.plummer(n=4096, eps=0.5),
    "small_buckets": lambda: dict(configs.plummer(n=8192), bucket=4),
    "periodic": lambda: configs.zeldovich(ng=16, rms=0.05),
    "periodic_r2": lambda: dict(configs.zeldovich(ng=16, rms=0.05), n_replicas=2),
    "periodic_r3": lambda: dict(configs.zeldovich(ng=12, rms=0.05), n_replicas=3),
    "active": lambda: configs.zeldovich(ng=16, rms=0.1, active_frac=0.2),
    "halos": lambda: configs.halos(n=32768),
    "deep_chain": deep_chain_workload,
}


def _both(ev):
    import torch
    ev.zero()
    ev.run_pc()
    ev.run_pp()
    torch.cuda.synchronize()
    ref = ev.d_vars.cpu().numpy().astype(np.float64)
    ev.zero()
    ev.run_fused()
    torch.cuda.synchronize()
    ev.fused_check()
    return ev.d_vars.cpu().numpy().astype(np.float64), ref


def _same(got, ref, case):
    # differences relative to the rms magnitude (periodic forces are small
    # residuals of large cancelling sums), and per particle at the median
    ra = np.linalg.norm(ref[:, :3], axis=1)
    da = np.linalg.norm(got[:, :3] - ref[:, :3], axis=1)
    dp = np.abs(got[:, 3] - ref[:, 3])
    ea, ep = da / np.sqrt(np.mean(ra ** 2)), dp / np.sqrt(np.mean(ref[:, 3] ** 2))
    print(f"fused {case}: acc {summary(ea)} pot {summary(ep)}", flush=True)
    assert ea.max() < 2e-5 and ep.max() < 2e-5
    assert np.median(da / ra) < 1e-6 and np.median(dp / np.abs(ref[:, 3])) < 1e-6
    np.testing.assert_array_equal(got[:, 4], ref[:, 4])


@pytest.mark.parametrize("case", list(CASES))
def test_gpu_fused_matches_list_path(device, case):
    ev = engine.Evaluation(CASES[case](), precise=True, device=device, lists="gpu")
    print(f"fused {case}: {ev.n_targets} targets", flush=True)
    got, ref = _both(ev)
    _same(got, ref, case)
    # a second launch on the same scratch: each bucket's walk and sums are
    # in a fixed order whichever wavefront claims it
    got2, _ = _both(ev)
    np.testing.assert_array_equal(got2, got)


@pytest.mark.parametrize("cfg", [2, 3, 4, 5])
def test_gpu_fused_full_size(device, cfg):
    """The full-size workloads, tree and lists built on the GPU: the 128^3
    box (2,097,152 particles), the 2^22-particle halos, the 2^20 Plummer
    sphere and the 256^3 box with a 10% active rung, at the same bars against
    the list path."""
    ev = engine.Evaluation(configs.workload(cfg), precise=True, device=device, lists="gpu",
                           tree="gpu")
    print(f"fused full size {cfg}: {ev.n_targets} targets", flush=True)
    got, ref = _both(ev)
    _same(got, ref, f"full size {cfg}")


def test_gpu_fused_periodic_oracle(device):
    """A periodic box through the fused launch (plus Ewald) against the FP64
    oracle on the same interaction set."""
    import torch
    ev = engine.Evaluation(configs.zeldovich(ng=24, rms=0.02), precise=True, device=device,
                           lists="gpu")
    ev.step_fused()
    torch.cuda.synchronize()
    ev.fused_check()
    acc, pot, _ = ev.results()
    oa, op, _ = oracle_forces(ev)
    ea, ep = rel_err(acc, oa), rel_err(pot, op)
    print(f"fused periodic acc {summary(ea)} pot {summary(ep)}", flush=True)
    check_tails(ea, 3.1e-6, 1.6e-5, 9.9e-5, label="fused periodic acc")
    check_tails(ep, 2.1e-6, 7.3e-6, 1.3e-5, label="fused periodic pot")


def test_gpu_fused_capacity_error(device):
    """A scratch too small for the walk is reported (error flag), not
    silently truncated."""
    import torch
    ev = engine.Evaluation(configs.plummer(n=8192), precise=True, device=device, lists="gpu")
    ev.fused_setup(max_cells=16, max_runs=16)
    ev.zero()
    ev.run_fused()
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="scratch capacities"):
        ev.fused_check()


def test_gpu_fused_without_lists(device):
    """An evaluation built with lists='none' (tree and moments only, as the
    forward passes of a kick-drift-kick step use it) gives the fused step's
    results bit for bit, and has no lists to hand out."""
    import torch
    w = configs.zeldovich(ng=16, rms=0.05, active_frac=0.3)
    ref = engine.Evaluation(w, precise=True, device=device, lists="gpu", tree="gpu")
    ref.step_fused()
    torch.cuda.synchronize()
    ev = engine.Evaluation(w, precise=True, device=device, lists="none", tree="gpu")
    ev.step_fused()
    torch.cuda.synchronize()
    ev.fused_check()
    np.testing.assert_array_equal(ev.d_vars.cpu().numpy(), ref.d_vars.cpu().numpy())
    with pytest.raises(ValueError, match="no interaction lists"):
        ev.lists


def test_gpu_empty_rung(device):
    """A rung with no active particle: every path (per bucket, batched,
    fused, the Ewald pass, the adjoint) runs over no buckets and leaves the
    results and the gradient at zero."""
    import torch
    w = configs.zeldovich(ng=16, rms=0.05, active_frac=0.3)
    w = dict(w, active=np.zeros_like(w["active"]))
    ev = engine.Evaluation(w, precise=True, device=device, lists="gpu", tree="gpu")
    assert ev.n_targets == 0 and len(ev.t_sizes) == 0 and len(ev.t_starts) == 0
    ev.step()
    ev.batch_plan()
    ev.step_batched()
    ev.step_fused()
    torch.cuda.synchronize()
    ev.fused_check()
    acc, pot, dtg = ev.results()
    assert not acc.any() and not pot.any() and not dtg.any()
    g = ev.adjoint(np.zeros((0, 3)))
    torch.cuda.synchronize()
    assert not g.cpu().numpy().any()
    ev.free_batch()
