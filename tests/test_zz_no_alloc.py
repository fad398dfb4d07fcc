"""The no-allocation batched launch (changaBatchedGravityNoAlloc,
include/changa_gpu.h Part 3) against the batched plan's launch with the same
options, on two calls in a row, the second on a smaller input.

Both run the same persistent kernel over the same buckets: every bucket is
evaluated by one wavefront in a fixed order and written once, so the results
must be equal bit for bit, whatever the queue split. The static work queue
keeps the first call's contents; the second call must not read them.
"""
import ctypes

import numpy as np
import pytest

from changa_amd import configs, engine, gpu

pytestmark = pytest.mark.gpu

OPTS = (64, 256, 1, 1)


def _plan_result(ev, kmax):
    import torch
    ev.batch_plan(*OPTS, zero_vars=True, targets_per_lane=kmax)
    ev.d_vars.fill_(5.0)
    ev.run_batched(graph=False)
    torch.cuda.synchronize()
    out = ev.d_vars.cpu().numpy().copy()
    ev.free_batch()
    return out


def _noalloc_result(ev, kmax):
    import torch
    pc, pp = ev._launch(True), ev._launch(False)
    o = gpu.ChangaBatchOptions(*OPTS, 1, 0.0, kmax)
    ev.d_vars.fill_(float("nan"))  # zeroVars must clear every particle of the buckets
    gpu.check(gpu.lib().changaBatchedGravityNoAlloc(ctypes.byref(pc), ctypes.byref(pp),
                                                    ctypes.byref(o), ev._sp()),
              "changaBatchedGravityNoAlloc")
    torch.cuda.synchronize()
    return ev.d_vars.cpu().numpy().copy()


def _covered(ev):
    idx = np.zeros(ev.n_targets, dtype=bool)
    for s, n in zip(ev.t_starts, ev.t_sizes):
        idx[s:s + n] = True
    return idx


@pytest.mark.parametrize("precise,compact,kmax", [(True, True, 3), (False, False, 1)])
def test_no_alloc_equals_plan_two_calls(device, precise, compact, kmax):
    import torch
    big = engine.Evaluation(configs.halos(n=1 << 16, cfg=3), precise=precise, pp_compact=compact,
                            device=device)
    small = engine.Evaluation(configs.plummer(n=6000), precise=precise, pp_compact=compact,
                              device=device)
    assert len(small.t_sizes) < len(big.t_sizes)
    ref_big, ref_small = _plan_result(big, kmax), _plan_result(small, kmax)
    got_big = _noalloc_result(big, kmax)
    free0 = torch.cuda.mem_get_info()[0]
    got_small = _noalloc_result(small, kmax)
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 == free0, "the second call changed the device's free memory"
    for ev, got, ref in ((big, got_big, ref_big), (small, got_small, ref_small)):
        m = _covered(ev)
        assert m.any()
        assert np.isfinite(got[m]).all()
        np.testing.assert_array_equal(got[m], ref[m])


def test_no_alloc_rejects_costliest_first_and_runs_empty(device):
    ev = engine.Evaluation(configs.plummer(n=4096), precise=True, device=device)
    pc, pp = ev._launch(True), ev._launch(False)
    o = gpu.ChangaBatchOptions(64, 256, 1, 2, 1, 0.0)
    rc = gpu.lib().changaBatchedGravityNoAlloc(ctypes.byref(pc), ctypes.byref(pp), ctypes.byref(o),
                                               ev._sp())
    assert rc != 0
    pc.numBuckets = pp.numBuckets = 0
    assert gpu.lib().changaBatchedGravityNoAlloc(ctypes.byref(pc), ctypes.byref(pp), None,
                                                 ev._sp()) == 0
