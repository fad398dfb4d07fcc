"""The tree built on the GPU (changaBuildTree: Morton keys, stable radix
sort, one split pass per depth) against the host builder: every array equal
(permutation, topology, breadth-first depths, Morton-cell radii, buckets),
including coincident particles (oversized buckets) and a tree that is a
single bucket; and a force evaluation on the GPU-built tree equal bit for
bit to the host-built one."""
import numpy as np
import pytest

from changa_amd import configs, engine, host

pytestmark = pytest.mark.gpu


def _coincident():
    w = configs.plummer(n=3000)
    pos = w["pos"].copy()
    pos[:100] = pos[0]  # 100 particles on one point: an oversized bucket
    return dict(w, pos=pos)


CASES = {
    "plummer": lambda: configs.plummer(n=32768),
    "halos": lambda: configs.halos(n=32768),
    "periodic": lambda: configs.zeldovich(ng=16, rms=0.05),
    "coincident": _coincident,
    "one_bucket": lambda: configs.plummer(n=5),
}

KEYS = ("perm", "node_first", "node_count", "node_child", "node_depth", "node_box_radius",
        "level_offsets", "bucket_node", "bucket_start", "bucket_size")


def _trees(w, device):
    kw = dict(box_lo=np.zeros(3), box_size=w["L"]) if w["periodic"] else {}
    return (host.build_tree(w["pos"], w["bucket"], **kw),
            engine.gpu_build_tree(w["pos"], w["bucket"], device=device, **kw))


@pytest.mark.parametrize("case", list(CASES))
def test_gpu_tree_matches_host(device, case):
    th, tg = _trees(CASES[case](), device)
    for k in ("n", "n_nodes", "n_buckets", "n_depths"):
        assert tg[k] == th[k], k
    for k in KEYS:
        np.testing.assert_array_equal(tg[k], th[k], err_msg=k)
    if case == "coincident":
        assert th["bucket_size"].max() >= 100
    print(f"tree {case}: {th['n_nodes']} nodes, {th['n_depths']} depths", flush=True)


def test_gpu_tree_evaluation(device):
    import torch
    w = configs.zeldovich(ng=16, rms=0.05)
    out = []
    for tree in ("host", "gpu"):
        ev = engine.Evaluation(w, precise=True, device=device, tree=tree)
        ev.step()
        torch.cuda.synchronize()
        out.append(ev.d_vars.cpu().numpy())
    np.testing.assert_array_equal(out[0], out[1])
