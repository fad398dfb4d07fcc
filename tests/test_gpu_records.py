"""Cell records (changaCellRecords, ChangaGravityLaunch.cellRecords) on the GPU.

A record holds a cell's FP64 centre of mass, mass, softening and the moments
scaled for the hexadecapole evaluation, one 128-byte line per cell; the
particle-cell loops fetch a tile of them as whole lines. The records are
formed with the same FP32 operations as the per-lane staging of the wire
moments, and the loops evaluate the same staged values in the same order, so:
  - the records equal, bit for bit, a numpy restatement of that staging;
  - forces with records equal forces from the wire moments bit for bit for
    every tile variant of the batched kernel, in the precision and the plain
    form, with softened cells, with one and two periodic replica shells and
    on a 10% active rung; the per-bucket kernel and the fused walk-and-force
    kernel (tiles of 32 cells with records, 64 without) to one FP32 rounding
    of their per-tile sums, dtGrav bit for bit.
"""
import numpy as np
import pytest

from changa_amd import configs, engine

pytestmark = pytest.mark.gpu


def _records_host(mom, cm64):
    """numpy restatement of the record layout (float32 arithmetic)."""
    f32 = np.float32
    n = len(mom)
    r = np.zeros((n, 32), dtype=np.float32)
    r[:, 0:6] = np.ascontiguousarray(cm64, dtype=np.float64).view(np.float32).reshape(n, 6)
    r[:, 6] = mom[:, 2]
    r[:, 7] = mom[:, 1]
    u = mom[:, 0]
    u2 = u * u
    s2, s3, s4 = f32(3.0) * u2, f32(7.5) * u2 * u, f32(17.5) * u2 * u2
    r[:, 8:13] = s2[:, None] * mom[:, 6:11]
    r[:, 13] = -(r[:, 8] + r[:, 11])
    r[:, 14:21] = s3[:, None] * mom[:, 11:18]
    r[:, 21] = -(r[:, 14] + r[:, 15])
    r[:, 22] = -(r[:, 16] + r[:, 17])
    r[:, 23:32] = s4[:, None] * mom[:, 18:27]
    return r


def _pc(ev):
    import torch
    ev.zero()
    ev.run_pc()
    torch.cuda.synchronize()
    return ev.d_vars.cpu().numpy().copy()


def _batched(ev):
    import torch
    ev.run_batched(graph=True)
    torch.cuda.synchronize()
    return ev.d_vars.cpu().numpy().copy()


def _fused(ev):
    import torch
    ev.zero()
    ev.run_fused()
    torch.cuda.synchronize()
    ev.fused_check()
    return ev.d_vars.cpu().numpy().copy()


def test_cell_records_content(device):
    ev = engine.Evaluation(configs.plummer(n=8192), precise=True, device=device)
    import torch
    torch.cuda.synchronize()
    mom = ev.d_moments.cpu().numpy()
    cm = ev.d_cm64.cpu().numpy()
    got = ev.d_cellrec.cpu().numpy()
    want = _records_host(mom, cm)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


CASES = {
    "plummer": lambda: configs.plummer(n=32768),
    "softened": lambda: configs.plummer(n=4096, eps=0.5),
    "box": lambda: configs.zeldovich(ng=20, rms=0.05, cfg=2),
    "box_r2": lambda: dict(configs.zeldovich(ng=16, rms=0.05, cfg=2), n_replicas=2),
    "rung": lambda: configs.zeldovich(ng=24, rms=0.15, cfg=5, active_frac=0.1),
    "halos": lambda: configs.halos(n=65536),
}


@pytest.mark.parametrize("case,precise", [("plummer", True), ("plummer", False),
                                          ("softened", True), ("box", True), ("box", False),
                                          ("box_r2", True), ("rung", True), ("halos", True)])
def test_records_equal_wire_path(device, case, precise):
    w = CASES[case]()
    ev_r = engine.Evaluation(w, precise=precise, device=device, cell_records=True)
    ev_w = engine.Evaluation(w, precise=precise, device=device, cell_records=False)
    if case == "softened":
        assert ev_r.counts[1] > 0, "workload must exercise softened cells"
    # the per-bucket kernel stages records in tiles of 32 cells and wire
    # moments in tiles of 64: the FP32 per-tile sums round in a different
    # order (one rounding of each, 1e-6 of the component's largest
    # magnitude); dtGrav is a max over the same terms
    pr, pw = _pc(ev_r), _pc(ev_w)
    scale = np.abs(pw[:, :4]).max(axis=0)
    assert (np.abs(pr[:, :4].astype(np.float64) - pw[:, :4]) / scale).max() <= 1e-6
    np.testing.assert_array_equal(pr[:, 4], pw[:, 4])
    for pct in (32, 64):
        for kmax in (1, 3):
            o = ev_r.batch_plan(pct, 256, 1, 1, targets_per_lane=kmax, cell_source=1)
            assert o.cellSource == 1
            o = ev_w.batch_plan(pct, 256, 1, 1, targets_per_lane=kmax)
            assert o.cellSource == 2  # no records: the wire moments
            np.testing.assert_array_equal(_batched(ev_r), _batched(ev_w))
    ev_r.free_batch()
    ev_w.free_batch()
    if precise:
        # the fused walk-and-force launch: record tiles of 32 cells, wire
        # tiles of 64 (the per-bucket bar above)
        fr, fw = _fused(ev_r), _fused(ev_w)
        scale = np.abs(fw[:, :4]).max(axis=0)
        assert (np.abs(fr[:, :4].astype(np.float64) - fw[:, :4]) / scale).max() <= 1e-6
        np.testing.assert_array_equal(fr[:, 4], fw[:, 4])


def test_plan_tunes_the_cell_source(device):
    ev = engine.Evaluation(configs.plummer(n=16384), precise=True, device=device)
    o = ev.batch_plan()
    assert o.cellSource in (1, 2)
    ev.free_batch()
