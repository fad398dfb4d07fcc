"""Active-target evaluation (a rung): only active particles are targets,
gathered bucket by bucket into their own arrays; every particle is a source.
ChaNGa's counters count exactly these (active target, source) pairs
(Compute.cpp:1643-1651). Checked against the FP64 oracle on the same lists
at the active particles, with the per-form bars of test_gpu_parity."""
import numpy as np
import pytest

from changa_amd import configs, engine
from oracle import oracle as O
from tests.helpers import check_tails, oracle_forces, rel_err, summary

pytestmark = pytest.mark.gpu


def _with_active(w, frac, seed):
    w = dict(w)
    w["active"] = np.random.default_rng(seed).uniform(size=w["pos"].shape[0]) < frac
    return w


def _active_errors(ev, pot_scale=False):
    """Per active particle. pot_scale: the potential's error relative to the
    rms potential (a periodic potential has zeros, where a per-particle
    relative error is undefined)."""
    acc, pot, dtg = ev.results()
    oa, op, od = oracle_forces(ev)
    a = ev.t_index
    ep = (np.abs(pot[a] - op[a]) / np.sqrt(np.mean(op[a] ** 2)) if pot_scale
          else rel_err(pot[a], op[a]))
    return rel_err(acc[a], oa[a]), ep, rel_err(dtg[a], od[a])


def test_gpu_active_isolated(device):
    import torch
    ev = engine.Evaluation(_with_active(configs.plummer(n=32768), 0.1, 5), precise=True,
                           device=device)
    assert ev.n_targets < ev.pos.shape[0]
    ev.step()
    torch.cuda.synchronize()
    bars = {"acc": (6.6e-8, 7.2e-7, 2.3e-6), "pot": (3.9e-8, 1.35e-7, 1.95e-7),
            "dtgrav": (1.5e-7, 2.3e-6, 5.1e-6)}
    for name, e in zip(("acc", "pot", "dtgrav"), _active_errors(ev)):
        print(f"active isolated {name} {summary(e)}", flush=True)
        check_tails(e, *bars[name], label=f"active isolated {name}")


def test_gpu_active_periodic_batched(device):
    import torch
    w = configs.zeldovich(ng=24, rms=0.1, active_frac=0.1)
    ev = engine.Evaluation(w, precise=True, device=device)
    ev.step()
    torch.cuda.synchronize()
    ref = ev.d_vars.cpu().numpy().copy()
    ev.batch_plan()
    ev.step_batched()
    torch.cuda.synchronize()
    v = ev.d_vars.cpu().numpy()
    np.testing.assert_array_equal(v[:, 4], ref[:, 4])
    ea, ep, ed = _active_errors(ev, pot_scale=True)
    print(f"active periodic acc {summary(ea)} pot {summary(ep)}", flush=True)
    check_tails(ea, 4.8e-7, 3.5e-6, 1.7e-5, label="active periodic acc")
    check_tails(ep, 2.5e-6, 7.8e-6, 9.4e-6, label="active periodic pot (over rms)")
    ev.free_batch()


def test_gpu_active_adjoint(device):
    """L over the active targets only: the oracle's adjoint with w zero on
    inactive particles is the same function."""
    import torch
    ev = engine.Evaluation(_with_active(configs.plummer(n=16384), 0.2, 6), precise=True,
                           device=device)
    wt = np.random.default_rng(7).standard_normal((ev.n_targets, 3))
    g = ev.adjoint(wt).cpu().numpy()
    torch.cuda.synchronize()
    wfull = np.zeros_like(ev.pos)
    wfull[ev.t_index] = wt
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    go = O.adjoint_lists(ev.tree, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists, wfull)
    ex = np.linalg.norm(g[:, :3] - go[:, :3], axis=1) / np.linalg.norm(go[:, :3], axis=1)
    em = np.abs(g[:, 3] - go[:, 3]) / np.abs(go[:, 3])
    print(f"active adjoint dL/dx {summary(ex)} dL/dm {summary(em)}", flush=True)
    check_tails(ex, 3.5e-7, 2.2e-6, 1e-5, label="active adjoint dL/dx")
    check_tails(em, 2.4e-7, 1.35e-5, None, label="active adjoint dL/dm")
