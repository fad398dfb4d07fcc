"""A step on moved particles is a complete evaluation: after engine.move()
the multipole build, the Ewald root record and h-loop table
(changaEwaldRoot, TreePiece::EwaldInit, Ewald.cpp:285-383) and the forces
all come from the new positions, with no new Evaluation. Checked against the
FP64 oracle on the new positions with the same tree and lists (its own
moments and its own Ewald table from them), for the batched, per-bucket and
fused steps; and the device root record and table against the host's
EwaldInit restatement."""
import numpy as np
import pytest

from changa_amd import configs, engine, host
from tests.helpers import check_tails, oracle_forces, rel_err, summary

pytestmark = pytest.mark.gpu


def _moved(ev, scale, seed=5):
    rng = np.random.default_rng(seed)
    return ev.pos + scale * rng.standard_normal(ev.pos.shape)


def _results(ev, step):
    import torch
    step()
    torch.cuda.synchronize()
    return ev.results()


def test_gpu_ewald_root_matches_host_table(device):
    ev = engine.Evaluation(configs.zeldovich(ng=16, rms=0.02), precise=True, device=device)
    import torch
    torch.cuda.synchronize()
    root = ev.d_node64[0].cpu().numpy()
    cm, u, M, q = root[0:3], root[3], root[5], root[12:34]
    ref = host.ewald_table(q, u, ev.L, ev.w["ewh_cut"])
    got = ev.d_ewt.cpu().numpy().astype(np.float64)
    assert got.shape == ref.shape == (146, 5)
    np.testing.assert_allclose(got, ref, rtol=2e-7, atol=2e-7 * np.abs(ref[:, 3:]).max())
    r = ev.d_ewald_root.cpu().numpy()
    np.testing.assert_array_equal(r[0:3], cm)
    assert r[3] == M
    np.testing.assert_allclose(r[4:26], host.root_momc(q, u)[:22], rtol=1e-15, atol=0)


def _bars(ea, ep, periodic):
    """At 1.5x the tails measured over the four steps."""
    if periodic:
        check_tails(ea, 3.5e-6, 2e-5, 9.1e-5, label="moved periodic acc")
        check_tails(ep, 2.2e-6, 8e-6, 1.13e-5, label="moved periodic pot")
    else:
        check_tails(ea, 8.4e-8, 9.6e-7, 1.08e-5, label="moved isolated acc")
        check_tails(ep, 4.2e-8, 1.55e-7, 2.7e-7, label="moved isolated pot")


@pytest.mark.parametrize("periodic", [True, False])
def test_gpu_step_on_moved_particles(device, periodic):
    """The list steps (batched, per-bucket) on the lists of the original
    positions; then, with the lists rebuilt from the moved particles' node
    records (build_lists), the batched step again and the fused walk and
    force (whose walk always reads the current node records)."""
    w = configs.zeldovich(ng=16, rms=0.02) if periodic else configs.plummer(n=8192)
    ev = engine.Evaluation(w, precise=True, device=device, lists="gpu")
    ev.batch_plan(64, 256, 1, 1)
    a0, p0, _ = _results(ev, ev.step_batched)
    scale = 2e-3 * (ev.L if periodic else 1.0)
    ev.move(_moved(ev, scale))
    tag = "periodic" if periodic else "isolated"
    oa, op, _ = oracle_forces(ev)
    for name, step in (("batched", ev.step_batched), ("per-bucket", ev.step)):
        acc, pot, _ = _results(ev, step)
        # the forces did change with the positions
        assert np.median(rel_err(acc, a0)) > 1e-3
        ea, ep = rel_err(acc, oa), rel_err(pot, op)
        print(f"moved {tag} {name}: acc {summary(ea)} pot {summary(ep)}", flush=True)
        _bars(ea, ep, periodic)
    # lists of the moved particles
    ev.build_moments()
    ev.build_lists()
    ev.batch_plan(64, 256, 1, 1)
    oa, op, _ = oracle_forces(ev)
    for name, step in (("batched, new lists", ev.step_batched), ("fused", ev.step_fused)):
        acc, pot, _ = _results(ev, step)
        ea, ep = rel_err(acc, oa), rel_err(pot, op)
        print(f"moved {tag} {name}: acc {summary(ea)} pot {summary(ep)}", flush=True)
        _bars(ea, ep, periodic)
    ev.free_batch()
