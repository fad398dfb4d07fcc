"""Argument validation of the device-resident entry points, on the CPU: with
abort-on-error switched off, a malformed call returns a non-zero code and
records a message naming the problem, before any device work is issued; a
call over no buckets (an empty rung) succeeds without arrays."""
import ctypes

import pytest

from changa_amd import gpu

DUMMY = 0x1000  # never dereferenced: validation fails before any use


@pytest.fixture
def lib():
    L = gpu.lib()  # the binding switches abort-on-error off
    L.changaGpuLastError(ctypes.byref(ctypes.c_char_p()))  # clear
    yield L


def _last(L):
    msg = ctypes.c_char_p()
    rc = L.changaGpuLastError(ctypes.byref(msg))
    return rc, (msg.value or b"").decode()


def _fused(nb=4):
    f = gpu.ChangaFusedLaunch()
    f.cells.numBuckets = f.parts.numBuckets = nb
    for name in ("bucketNode", "first", "count", "child", "node64", "work"):
        setattr(f, name, DUMMY)
    for side in (f.cells, f.parts):
        for name in ("targets", "vars", "moments", "sources", "starts", "sizes"):
            setattr(side, name, DUMMY)
    f.numNodes, f.theta, f.maxCells, f.maxRuns = 7, 0.7, 16, 16
    return f


def test_fused_rejects_the_plain_form(lib):
    f = _fused()  # no FP64 positions: not the precision form
    assert lib.changaFusedGravity(ctypes.byref(f), None) != 0
    rc, msg = _last(lib)
    assert rc != 0 and "precision form" in msg, msg


def test_fused_rejects_replicas_beyond_three(lib):
    f = _fused()
    for side in (f.cells, f.parts):
        side.tpos64 = side.spos64 = side.cm64 = DUMMY
    f.nReplicas, f.period = 4, 1.0
    assert lib.changaFusedGravity(ctypes.byref(f), None) != 0
    assert "nReplicas" in _last(lib)[1]


def test_fused_over_no_buckets_needs_no_arrays(lib):
    f = gpu.ChangaFusedLaunch()
    assert lib.changaFusedGravity(ctypes.byref(f), None) == 0
    assert _last(lib)[0] == 0


def test_list_build_rejects_a_zero_opening_angle(lib):
    b = gpu.ChangaListBuild()
    b.nBuckets, b.numNodes = 4, 7
    for name in ("bucketNode", "first", "count", "child", "node64", "cellMarks", "partMarks",
                 "softStart", "work"):
        setattr(b, name, DUMMY)
    b.theta = 0.0
    n1, n2 = ctypes.c_longlong(), ctypes.c_longlong()
    assert lib.changaBuildListsCount(ctypes.byref(b), ctypes.byref(n1), ctypes.byref(n2),
                                     None) != 0
    assert "theta" in _last(lib)[1]


def test_tree_build_rejects_an_empty_bucket_size(lib):
    t = gpu.ChangaTreeBuild()
    t.n, t.bucketSize, t.boxSize = 100, 0, 1.0
    for name in ("pos", "perm", "first", "count", "child", "boxRadius", "bucketNode",
                 "levelOffsets", "work"):
        setattr(t, name, DUMMY)
    assert lib.changaBuildTree(ctypes.byref(t), None) != 0
    assert "changaBuildTree" in _last(lib)[1]


def test_adjoint_rejects_missing_weights(lib):
    a = gpu.ChangaAdjointLaunch()
    a.cells.numBuckets = a.parts.numBuckets = 4
    a.grad, a.nParticles = DUMMY, 10
    assert lib.changaGravityAdjoint(ctypes.byref(a), None) != 0
    assert "missing argument" in _last(lib)[1]


def test_add_vars(lib):
    assert lib.changaAddVars(None, None, 0, None) == 0
    assert lib.changaAddVars(None, None, 5, None) != 0
    assert "changaAddVars" in _last(lib)[1]


def test_node_gravity_rejects_a_nonzero_variant(lib):
    a = gpu.ChangaGravityLaunch()
    a.numBuckets = 4
    for name in ("targets", "vars", "moments", "list", "marks", "starts", "sizes"):
        setattr(a, name, DUMMY)
    a.variant = 1  # reserved field: only the one particle-cell kernel exists
    assert lib.changaNodeGravity(ctypes.byref(a), None) != 0
    assert "variant" in _last(lib)[1]



def _noalloc_launches(nb):
    pc, pp = gpu.ChangaGravityLaunch(), gpu.ChangaGravityLaunch()
    for side in (pc, pp):
        side.numBuckets = nb
        for name in ("targets", "vars", "moments", "sources", "list", "marks", "starts",
                     "sizes"):
            setattr(side, name, DUMMY)
    pp.listKind = 1
    return pc, pp


def test_no_alloc_rejects_more_buckets_than_its_queue_holds(lib):
    pc, pp = _noalloc_launches(gpu.NOALLOC_MAX_BUCKETS + 1)
    assert lib.changaBatchedGravityNoAlloc(ctypes.byref(pc), ctypes.byref(pp), None, None) != 0
    assert "CHANGA_NOALLOC_MAX_BUCKETS" in _last(lib)[1]


def test_no_alloc_rejects_the_costliest_first_order(lib):
    pc, pp = _noalloc_launches(4)
    o = gpu.ChangaBatchOptions(64, 256, 1, 2, 0, 0.0)
    assert lib.changaBatchedGravityNoAlloc(ctypes.byref(pc), ctypes.byref(pp), ctypes.byref(o),
                                           None) != 0
    assert "tree order" in _last(lib)[1]


def test_cell_records_reject_misalignment(lib):
    rc = lib.changaCellRecords(ctypes.c_void_p(DUMMY), None, 2, ctypes.c_void_p(DUMMY + 4), None)
    assert rc != 0
    assert "aligned" in _last(lib)[1]
    rc = lib.changaCellRecords(None, None, 2, ctypes.c_void_p(DUMMY), None)
    assert rc != 0 and "missing" in _last(lib)[1]
    assert lib.changaCellRecords(None, None, 0, None, None) == 0


def test_node_gravity_rejects_misaligned_records(lib):
    a = gpu.ChangaGravityLaunch()
    a.numBuckets = 1
    for name in ("targets", "vars", "moments", "list", "marks", "starts", "sizes"):
        setattr(a, name, DUMMY)
    a.cellRecords = DUMMY + 64
    assert lib.changaNodeGravity(ctypes.byref(a), None) != 0
    assert "aligned" in _last(lib)[1]
