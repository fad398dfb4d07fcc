"""The HostCUDA/EwaldCUDA mirror (include/changa_gpu.h Part 1) on the GPU.

Drives the entry points the way ChaNGa's DataManager / ListCompute /
TreePiece::EwaldGPU do: upload the local tree, send the cell and particle
interaction lists as several requests with host-side lists and per-request
bucket markers, run EwaldHost, copy the accelerations back, and check that
each request's callback fires exactly once. The result must equal the
device-resident path on the wire moments bit for bit (same kernels, same
per-bucket lists)."""
import ctypes

import numpy as np
import pytest

from changa_amd import configs, engine, gpu

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def _requests(ev, lst, marks, width, n_req):
    """Split the buckets into n_req requests with request-local markers."""
    nb = ev.tree["n_buckets"]
    cuts = np.linspace(0, nb, n_req + 1).astype(int)
    out = []
    for r in range(n_req):
        b0, b1 = cuts[r], cuts[r + 1]
        m = marks[b0:b1 + 1].astype(np.int64)
        seg = np.ascontiguousarray(lst.reshape(-1, width)[m[0]:m[-1]].reshape(-1))
        out.append(dict(list=seg, marks=np.ascontiguousarray((m - m[0]).astype(np.int32)),
                        starts=np.ascontiguousarray(ev.tree["bucket_start"][b0:b1]),
                        sizes=np.ascontiguousarray(ev.tree["bucket_size"][b0:b1]),
                        n=int(m[-1] - m[0]), nb=int(b1 - b0)))
    return out


def _run_mirror(ev, pp_expanded):
    import torch

    L = gpu.lib()
    fired = []
    handler = gpu.CALLBACK(lambda cb: fired.append(cb))
    L.changaGpuSetCallbackHandler(handler)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = ev.pos.shape[0]
    moments = ev.d_moments.cpu().numpy()
    parts = ev.d_parts.cpu().numpy()
    varp = np.zeros((n, 5), dtype=np.float32)
    dm, dp, dv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    L.DataManagerTransferLocalTree(moments.ctypes.data, moments.nbytes, parts.ctypes.data,
                                   parts.nbytes, varp.ctypes.data, varp.nbytes, ctypes.byref(dm),
                                   ctypes.byref(dp), ctypes.byref(dv), stream, n,
                                   ctypes.c_void_p(1001))
    keep = []
    cb = 2000

    def send(fn, reqs, sources_remote=False):
        nonlocal cb
        for r in reqs:
            q = gpu.CudaRequest()
            q.stream = stream
            q.d_localMoments = dm
            q.d_localParts = dp
            q.d_localVars = dv
            q.list = r["list"].ctypes.data
            q.bucketMarkers = r["marks"].ctypes.data
            q.bucketStarts = r["starts"].ctypes.data
            q.bucketSizes = r["sizes"].ctypes.data
            q.numInteractions = r["n"]
            q.numBucketsPlusOne = r["nb"] + 1
            q.fperiod = ev.L
            q.cb = cb
            cb += 1
            keep.append((q, r))
            fn(ctypes.byref(q))

    send(L.TreePieceCellListDataTransferLocal, _requests(ev, ev.lists["cells"],
                                                         ev.lists["cell_marks"], 2, 3))
    if pp_expanded:
        cells, marks = engine.expand_parts(ev.lists)
        send(L.TreePiecePartListDataTransferLocal, _requests(ev, cells, marks, 2, 3))
    else:
        send(L.TreePiecePartListDataTransferLocalCompact,
             _requests(ev, ev.lists["parts"], ev.lists["part_marks"], 3, 3))
    ncb = cb - 2000
    if ev.w["periodic"]:
        h = gpu.EwaldData()
        L.EwaldHostMemorySetup(ctypes.byref(h), n, len(ev.ewt64), 1)
        for i in range(n):
            h.EwaldMarkers[i] = i
        for i, e in enumerate(ev.ewt64.astype(np.float32)):
            h.ewt[i] = gpu.EwtData(*[float(x) for x in e])
        ctypes.memmove(h.cachedData, ctypes.byref(ev.ewald_ro), ctypes.sizeof(ev.ewald_ro))
        h.EwaldRange[0], h.EwaldRange[1] = 0, n - 1
        L.EwaldHost(dp, dv, ctypes.byref(h), stream, ctypes.c_void_p(3001), 0, 1)
    host_out = ctypes.c_void_p()
    L.allocatePinnedHostMemory(ctypes.byref(host_out), n * 20)
    L.TransferParticleVarsBack(host_out, n * 20, dv, stream, ctypes.c_void_p(4001))
    torch.cuda.synchronize()
    out = np.ctypeslib.as_array(ctypes.cast(host_out, ctypes.POINTER(ctypes.c_float)),
                                shape=(n, 5)).copy()
    if ev.w["periodic"]:
        L.EwaldHostMemoryFree(ctypes.byref(h), 1)
    L.freePinnedHostMemory(host_out)
    hip = _hip()
    for p in (dm, dp, dv):
        assert hip.hipFree(p) == 0
    L.changaGpuSetCallbackHandler(gpu.CALLBACK(0))
    expect = [1001] + list(range(2000, 2000 + ncb)) + ([3001] if ev.w["periodic"] else []) + [4001]
    return out, fired, expect


@pytest.mark.parametrize("case", ["plummer_soft", "periodic", "periodic_r2"])
def test_gpu_mirror_matches_device_path(device, case):
    import torch
    if case == "plummer_soft":
        w = configs.plummer(n=3000, eps=0.5)
    else:
        w = dict(configs.zeldovich(ng=16, rms=0.03), n_replicas=2 if case == "periodic_r2" else 1)
    ev = engine.Evaluation(w, precise=False, pp_compact=False, device=device,
                           cell_records=False)
    assert case != "plummer_soft" or ev.counts[1] > 0
    out, fired, expect = _run_mirror(ev, pp_expanded=True)
    assert sorted(fired) == sorted(expect) and len(fired) == len(set(fired))
    ev.zero()
    ev.run_pc()
    ev.run_pp()
    ev.run_ewald(host_record=True)  # the record EwaldHost was given
    torch.cuda.synchronize()
    ref = ev.d_vars.cpu().numpy()
    np.testing.assert_array_equal(out, ref)


def test_gpu_mirror_compact_part_lists(device):
    import torch
    w = configs.plummer(n=3000)
    ev = engine.Evaluation(w, precise=False, pp_compact=True, device=device,
                           cell_records=False)
    out, fired, expect = _run_mirror(ev, pp_expanded=False)
    assert sorted(fired) == sorted(expect)
    ev.zero()
    ev.run_pc()
    ev.run_pp()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out, ev.d_vars.cpu().numpy())


def test_gpu_error_mode_returns(device):
    """With abort-on-error off, a bad launch returns an error code and the
    message names the call (no fault: the library rejects it on the host)."""
    L = gpu.lib()  # error-return mode (the binding switches abort-on-error off)
    a = gpu.ChangaGravityLaunch()
    a.numBuckets = 5  # no arrays
    rc = L.changaNodeGravity(ctypes.byref(a), None)
    assert rc != 0
    msg = ctypes.c_char_p()
    assert L.changaGpuLastError(ctypes.byref(msg)) != 0
    assert b"missing" in msg.value


def test_gpu_ewald_host_small_phase(device):
    """EwaldHost's two phases (HostCUDA.cu:1917-1946, EwaldKernel :1958-1971,
    filled as Ewald.cpp:416-441 fills them): the large phase evaluates the
    particles markers[0 .. EwaldRange[1]], the small phase the contiguous
    particles EwaldRange[0] .. EwaldRange[1]. The same particles through
    either phase get the same result bit for bit; no other particle is
    touched."""
    import torch
    L = gpu.lib()
    ev = engine.Evaluation(configs.zeldovich(ng=12, rms=0.05), precise=False, device=device,
                           cell_records=False)
    torch.cuda.synchronize()
    n = ev.pos.shape[0]
    a, b = 211, n - 307
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dp = ctypes.c_void_p(ev.d_parts.data_ptr())

    def run(large):
        h = gpu.EwaldData()
        L.EwaldHostMemorySetup(ctypes.byref(h), n, len(ev.ewt64), int(large))
        for i, e in enumerate(ev.ewt64.astype(np.float32)):
            h.ewt[i] = gpu.EwtData(*[float(x) for x in e])
        ro = gpu.EwaldReadOnlyData.from_buffer_copy(ev.ewald_ro)
        if large:
            for i in range(b - a + 1):
                h.EwaldMarkers[i] = a + i
            h.EwaldRange[0], h.EwaldRange[1] = 0, b - a
            ro.n = b - a + 1
        else:
            assert not h.EwaldMarkers
            h.EwaldRange[0], h.EwaldRange[1] = a, b
            ro.n = n
        ctypes.memmove(h.cachedData, ctypes.byref(ro), ctypes.sizeof(ro))
        out = torch.zeros((n, 5), dtype=torch.float32, device=device)
        L.EwaldHost(dp, ctypes.c_void_p(out.data_ptr()), ctypes.byref(h), stream, None, 0,
                    int(large))
        torch.cuda.synchronize()
        L.EwaldHostMemoryFree(ctypes.byref(h), int(large))
        msg = ctypes.c_char_p()
        assert L.changaGpuLastError(ctypes.byref(msg)) == 0, msg.value
        return out.cpu().numpy()

    big, small = run(True), run(False)
    np.testing.assert_array_equal(small, big)
    assert np.abs(big[a:b + 1, 0:4]).sum(axis=1).min() > 0
    assert not big[:a].any() and not big[b + 1:].any()
