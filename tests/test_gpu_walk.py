"""Interaction lists built on the GPU (changaBuildListsCount / Fill) against
the host walk (the restatement of ListCompute's walk with
openCriterionBucket and openSoftening) on the same node geometry: the same
lists entry for entry, in the same order; the interaction counts; and forces
through GPU-built lists against the FP64 oracle."""
import numpy as np
import pytest

from changa_amd import configs, engine, host
from tests.helpers import check_tails, deep_chain_workload, oracle_forces, rel_err, summary

pytestmark = pytest.mark.gpu

CASES = {
    "plummer": lambda: configs.plummer(n=32768),
    "softened": lambda: configs.plummer(n=4096, eps=0.5),
    "periodic": lambda: configs.zeldovich(ng=16, rms=0.05),
    "periodic_r2": lambda: dict(configs.zeldovich(ng=16, rms=0.05), n_replicas=2),
    "periodic_r3": lambda: dict(configs.zeldovich(ng=12, rms=0.05), n_replicas=3),
    "halos": lambda: configs.halos(n=32768),
    "deep_chain": deep_chain_workload,
}


def _host_lists(ev):
    """The host walk on the GPU moment build's node geometry (cm, radius,
    softening, mass, bounding box)."""
    w = ev.w
    geo = np.ascontiguousarray(ev.node64()[:, :12])
    return host.walk(ev.tree, geo, w["theta"], w["n_replicas"], ev.L if w["periodic"] else 1.0)


@pytest.mark.parametrize("case", list(CASES))
def test_gpu_lists_match_host_walk(device, case):
    ev = engine.Evaluation(CASES[case](), precise=True, device=device, lists="gpu")
    g, h = ev.lists, _host_lists(ev)
    for k in ("cell_marks", "part_marks", "soft_start"):
        np.testing.assert_array_equal(g[k], h[k], err_msg=k)
    np.testing.assert_array_equal(g["cells"], h["cells"])
    np.testing.assert_array_equal(g["parts"], h["parts"])
    assert g["n_part_sources"] == h["n_part_sources"]
    assert ev.counts == host.interaction_counts(ev.tree, h, ev.active)
    if case == "softened":
        assert (g["cell_marks"][1:] - g["soft_start"]).sum() > 0
    print(f"lists {case}: {len(g['cells']) // 2} cells, {len(g['parts']) // 3} runs", flush=True)


def test_gpu_lists_active_rung(device):
    """Only the buckets holding active particles are walked; their lists are
    the host walk's for those buckets (the bucket field of offsetID counts
    the walked buckets)."""
    w = configs.zeldovich(ng=16, rms=0.1, active_frac=0.2)
    ev = engine.Evaluation(w, precise=True, device=device, lists="gpu")
    h = engine.slice_lists(_host_lists(ev), ev.t_keep)
    np.testing.assert_array_equal(ev.t_lists["cell_marks"], h["cell_marks"])
    np.testing.assert_array_equal(ev.t_lists["part_marks"], h["part_marks"])
    gc = ev.d_cells.cpu().numpy()[:len(h["cells"]) // 2]
    hc = h["cells"].reshape(-1, 2)
    np.testing.assert_array_equal(gc[:, 0], hc[:, 0])
    np.testing.assert_array_equal(gc[:, 1] >> 22, hc[:, 1] >> 22)
    gp = ev.d_plist.cpu().numpy()[:len(h["parts"]) // 3]
    np.testing.assert_array_equal(gp, h["parts"].reshape(-1, 3))


def test_gpu_lists_forces(device):
    """A periodic box evaluated through GPU-built lists, against the FP64
    oracle on the same lists (bars at 1.5x the measured tails)."""
    import torch
    ev = engine.Evaluation(configs.zeldovich(ng=24, rms=0.02), precise=True, device=device,
                           lists="gpu")
    ev.step()
    torch.cuda.synchronize()
    acc, pot, _ = ev.results()
    oa, op, _ = oracle_forces(ev)
    ea, ep = rel_err(acc, oa), rel_err(pot, op)
    print(f"gpu lists periodic acc {summary(ea)} pot {summary(ep)}", flush=True)
    check_tails(ea, 3.1e-6, 1.6e-5, 8.1e-5, label="gpu lists periodic acc")
    check_tails(ep, 2.2e-6, 7.7e-6, 1.25e-5, label="gpu lists periodic pot")


def test_tipsy_evaluate(device, tmp_path):
    """A tipsy file through changa_amd.tipsy.evaluate (GPU lists, precision
    form): the .acc2 / .pot files hold this build's forces in input order."""
    from changa_amd import tipsy
    w = configs.plummer(n=4096)
    path = tmp_path / "plummer.bin"
    tipsy.write(path, 0.0, dark={"mass": w["mass"], "pos": w["pos"], "eps": w["soft"]})
    acc, pot = tipsy.evaluate(path, str(tmp_path / "out"), theta=w["theta"], bucket=w["bucket"],
                              device=device)
    # "%.14g" keeps 14 significant digits
    np.testing.assert_allclose(tipsy.read_array(str(tmp_path / "out.pot")), pot, rtol=1e-13)
    np.testing.assert_allclose(tipsy.read_array(str(tmp_path / "out.acc2"), vector=True), acc,
                               rtol=1e-13, atol=1e-13 * np.abs(acc).max())
    # the same particles (FP32 as the file holds them) through the engine
    import torch
    wf = dict(w, pos=w["pos"].astype(np.float32).astype(np.float64),
              mass=w["mass"].astype(np.float32).astype(np.float64),
              soft=w["soft"].astype(np.float32).astype(np.float64))
    ev = engine.Evaluation(wf, precise=True, device=device, lists="gpu")
    ev.step()
    torch.cuda.synchronize()
    a_t, p_t, _ = ev.results()
    np.testing.assert_array_equal(acc[ev.tree["perm"]], a_t)
    np.testing.assert_array_equal(pot[ev.tree["perm"]], p_t)
