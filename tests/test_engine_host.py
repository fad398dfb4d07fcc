"""Host-side logic of the engine that needs no GPU: list slicing for the
active-target mode and the target set it builds."""
import numpy as np

from changa_amd import configs, engine, host


def test_slice_lists_matches_per_bucket_concatenation():
    rng = np.random.default_rng(0)
    nb = 50
    lens_c, lens_p = rng.integers(0, 7, nb), rng.integers(0, 5, nb)
    cm = np.concatenate([[0], np.cumsum(lens_c)])
    pm = np.concatenate([[0], np.cumsum(lens_p)])
    lists = dict(cells=np.arange(2 * cm[-1], dtype=np.int32), cell_marks=cm,
                 parts=np.arange(3 * pm[-1], dtype=np.int32), part_marks=pm)
    keep = np.sort(rng.choice(nb, 20, replace=False))
    out = engine.slice_lists(lists, keep)
    ref_c = np.concatenate([lists["cells"].reshape(-1, 2)[cm[b]:cm[b + 1]] for b in keep])
    ref_p = np.concatenate([lists["parts"].reshape(-1, 3)[pm[b]:pm[b + 1]] for b in keep])
    np.testing.assert_array_equal(out["cells"], ref_c.reshape(-1))
    np.testing.assert_array_equal(out["parts"], ref_p.reshape(-1))
    np.testing.assert_array_equal(np.diff(out["cell_marks"]), lens_c[keep])
    np.testing.assert_array_equal(np.diff(out["part_marks"]), lens_p[keep])
    empty = engine.slice_lists(lists, np.zeros(0, dtype=np.int64))
    assert empty["cells"].size == 0 and list(empty["cell_marks"]) == [0]


def test_active_target_set_groups_active_particles_by_bucket():
    """The target set of a rung: active particles in tree order, grouped by
    bucket, each active bucket's lists kept."""
    w = configs.zeldovich(ng=16, rms=0.05, active_frac=0.1)

    class Ev(engine.Evaluation):  # the host part only
        def _upload(self):
            pass

    ev = Ev(w, precise=True)
    tr = ev.tree
    owner = np.repeat(np.arange(tr["n_buckets"]), tr["bucket_size"])
    np.testing.assert_array_equal(ev.t_index, np.nonzero(ev.active)[0])
    assert ev.t_sizes.sum() == ev.n_targets == ev.active.sum()
    assert (np.diff(ev.t_index) > 0).all()
    kept = np.unique(owner[ev.t_index])
    np.testing.assert_array_equal(owner[ev.t_index][ev.t_starts], kept)
    assert len(ev.t_lists["cell_marks"]) == len(kept) + 1
    counts = host.interaction_counts(tr, ev.lists, ev.active)
    pc_active = (np.diff(ev.t_lists["cell_marks"]) * ev.t_sizes).sum()
    assert pc_active == counts[0] + counts[1]


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    """Without the built HIP library the product path raises; there is no
    CPU fallback to route through."""
    import pytest

    from changa_amd import gpu
    monkeypatch.setattr(gpu, "_lib", None)
    monkeypatch.setattr(gpu._build, "GPU_LIB", str(tmp_path / "libchanga_gpu.so"))
    with pytest.raises(RuntimeError, match="not built"):
        gpu.lib()
