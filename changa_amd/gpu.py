"""ctypes view of libchanga_gpu.so (include/changa_gpu.h).

Loading fails loudly when the library is not built: there is no CPU fallback
for the GPU path.
"""
import ctypes
import os

from . import build as _build

c_int = ctypes.c_int
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t


class CudaVector3D(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


class MultipoleMomentsData(ctypes.Structure):
    _fields_ = [("totalMass", c_float), ("cmx", c_float), ("cmy", c_float), ("cmz", c_float)]


MOMC_FIELDS = ["m", "xx", "yy", "xy", "xz", "yz", "xxx", "xyy", "xxy", "yyy", "xxz", "yyz", "xyz",
               "xxxx", "xyyy", "xxxy", "yyyy", "xxxz", "yyyz", "xxyy", "xxyz", "xyyz", "zz", "xzz",
               "yzz", "zzz", "xxzz", "xyzz", "xzzz", "yyzz", "yzzz", "zzzz"]


class MomcData(ctypes.Structure):
    _fields_ = [(f, c_float) for f in MOMC_FIELDS]


class EwaldReadOnlyData(ctypes.Structure):
    _fields_ = [("mm", MultipoleMomentsData), ("momcRoot", MomcData), ("n", c_int),
                ("nReps", c_int), ("nEwReps", c_int), ("nEwhLoop", c_int), ("L", c_float),
                ("fEwCut", c_float), ("alpha", c_float), ("alpha2", c_float), ("k1", c_float),
                ("ka", c_float), ("fEwCut2", c_float), ("fInner2", c_float)]


class EwtData(ctypes.Structure):
    _fields_ = [("hx", c_float), ("hy", c_float), ("hz", c_float), ("hCfac", c_float),
                ("hSfac", c_float)]


class EwaldData(ctypes.Structure):
    _fields_ = [("EwaldRange", c_int * 2), ("EwaldMarkers", ctypes.POINTER(c_int)),
                ("ewt", ctypes.POINTER(EwtData)),
                ("cachedData", ctypes.POINTER(EwaldReadOnlyData))]


class CudaRequest(ctypes.Structure):
    _fields_ = [("stream", c_void_p), ("d_localMoments", c_void_p), ("d_remoteMoments", c_void_p),
                ("d_localParts", c_void_p), ("d_remoteParts", c_void_p), ("d_localVars", c_void_p),
                ("sMoments", c_size_t), ("sCompactParts", c_size_t), ("sVarParts", c_size_t),
                ("list", c_void_p), ("bucketMarkers", c_void_p), ("bucketStarts", c_void_p),
                ("bucketSizes", c_void_p), ("numInteractions", c_int),
                ("numBucketsPlusOne", c_int), ("tp", c_void_p), ("missedNodes", c_void_p),
                ("missedParts", c_void_p), ("sMissed", c_size_t), ("affectedBuckets", c_void_p),
                ("cb", c_void_p), ("state", c_void_p), ("fperiod", c_float),
                ("node", ctypes.c_bool), ("remote", ctypes.c_bool)]


class CudaRequestLTW(ctypes.Structure):
    """CudaRequest of the GPU_LOCAL_TREE_WALK layout (the tail of HostCUDA.h:79-88)."""
    _fields_ = CudaRequest._fields_ + [
        ("firstParticle", c_int), ("lastParticle", c_int), ("rootIdx", c_int),
        ("theta", ctypes.c_double), ("thetaMono", ctypes.c_double), ("nReplicas", c_int),
        ("fperiodY", c_float), ("fperiodZ", c_float)]


class ChangaGravityLaunch(ctypes.Structure):
    _fields_ = [("targets", c_void_p), ("vars", c_void_p), ("moments", c_void_p),
                ("sources", c_void_p), ("list", c_void_p), ("marks", c_void_p),
                ("starts", c_void_p), ("sizes", c_void_p), ("numBuckets", c_int),
                ("listKind", c_int), ("fperiod", c_float), ("variant", c_int),
                ("tpos64", c_void_p), ("spos64", c_void_p), ("cm64", c_void_p),
                ("cellRecords", c_void_p), ("fperiodY", c_float), ("fperiodZ", c_float)]


class ChangaBatchOptions(ctypes.Structure):
    _fields_ = [("pcTile", c_int), ("ppTile", c_int), ("listsPerGrab", c_int), ("order", c_int),
                ("zeroVars", c_int), ("tunedMs", c_float), ("targetsPerLane", c_int),
                ("cellSource", c_int)]


class ChangaAdjointLaunch(ctypes.Structure):
    _fields_ = [("cells", ChangaGravityLaunch), ("parts", ChangaGravityLaunch), ("w", c_void_p),
                ("targetIndex", c_void_p), ("grad", c_void_p), ("nodeWork", c_void_p), ("listWork", c_void_p),
                ("listWorkInts", c_size_t), ("numCellEntries", c_int), ("numPartEntries", c_int),
                ("nNodes", c_int),
                ("nDepths", c_int), ("nParticles", c_int), ("levelOffsets", c_void_p),
                ("first", c_void_p), ("count", c_void_p), ("child", c_void_p),
                ("node64", c_void_p), ("pos64", c_void_p), ("mass64", c_void_p),
                ("ewaldRO", c_void_p), ("ewt", c_void_p), ("rootCm64", c_void_p),
                ("nTargets", c_int), ("ewaldWork", c_void_p), ("ewaldWorkDoubles", c_size_t)]


class ChangaListBuild(ctypes.Structure):
    _fields_ = [("nBuckets", c_int), ("numNodes", c_int), ("bucketNode", c_void_p),
                ("first", c_void_p),
                ("count", c_void_p), ("child", c_void_p), ("node64", c_void_p),
                ("theta", ctypes.c_double), ("nReplicas", c_int), ("period", ctypes.c_double),
                ("cellMarks", c_void_p), ("partMarks", c_void_p), ("softStart", c_void_p),
                ("sourceCounts", c_void_p), ("cells", c_void_p), ("parts", c_void_p),
                ("work", c_void_p), ("workBytes", c_size_t)]


class ChangaTreeBuild(ctypes.Structure):
    _fields_ = [("n", c_int), ("pos", c_void_p), ("boxLo", ctypes.c_double * 3),
                ("boxSize", ctypes.c_double), ("bucketSize", c_int), ("perm", c_void_p),
                ("first", c_void_p), ("count", c_void_p), ("child", c_void_p),
                ("boxRadius", c_void_p), ("bucketNode", c_void_p), ("maxNodes", c_int),
                ("levelOffsets", c_void_p), ("maxDepths", c_int), ("work", c_void_p),
                ("workBytes", c_size_t), ("nNodes", c_int), ("nBuckets", c_int),
                ("nDepths", c_int)]


class ChangaFusedLaunch(ctypes.Structure):
    _fields_ = [("cells", ChangaGravityLaunch), ("parts", ChangaGravityLaunch),
                ("bucketNode", c_void_p), ("first", c_void_p), ("count", c_void_p),
                ("child", c_void_p), ("node64", c_void_p), ("numNodes", c_int),
                ("theta", ctypes.c_double), ("nReplicas", c_int), ("period", ctypes.c_double),
                ("maxCells", c_int), ("maxRuns", c_int), ("work", c_void_p),
                ("workBytes", c_size_t), ("errorFlag", c_void_p)]


CALLBACK = ctypes.CFUNCTYPE(None, c_void_p)

# every extern "C" symbol include/changa_gpu.h declares
EXPORTS = [
    "allocatePinnedHostMemory", "freePinnedHostMemory", "DataManagerTransferLocalTree",
    "DataManagerTransferRemoteChunk", "TransferParticleVarsBack",
    "TreePieceCellListDataTransferLocal", "TreePieceCellListDataTransferRemote",
    "TreePieceCellListDataTransferRemoteResume", "TreePiecePartListDataTransferLocal",
    "TreePiecePartListDataTransferLocalSmallPhase", "TreePiecePartListDataTransferRemote",
    "TreePiecePartListDataTransferRemoteResume", "EwaldHostMemorySetup", "EwaldHostMemoryFree",
    "EwaldHost", "changaGpuSetCallbackHandler", "changaGpuSetAbortOnError", "changaGpuLastError",
    "changaGpuReleasePools", "TreePiecePartListDataTransferLocalCompact", "changaNodeGravity",
    "changaPartGravity", "changaZeroVars", "changaAddVars", "changaEwald", "changaBuildMoments",
    "changaBatchPlanCreate", "changaBatchPlanLaunch", "changaBatchPlanRun", "changaBatchPlanDestroy",
    "changaGravityAdjoint", "changaGravityAdjointWorkInts", "changaGravityAdjointEwaldWorkDoubles",
    "changaBuildListsWorkBytes", "changaBuildListsCount", "changaBuildListsFill",
    "changaBuildTreeWorkBytes", "changaBuildTree", "changaFusedGravityWorkBytes",
    "changaFusedGravity", "changaBatchedGravityNoAlloc", "changaEwaldRoot", "changaEwaldResident",
    "changaCellRecords",
]

# the symbols of libchanga_gpu_ltw.so: the HostCUDA / EwaldCUDA boundary and
# the per-list kernels it calls, on the GPU_LOCAL_TREE_WALK wire layouts
LTW_EXPORTS = [
    "allocatePinnedHostMemory", "freePinnedHostMemory", "DataManagerTransferLocalTree",
    "DataManagerTransferRemoteChunk", "TransferParticleVarsBack",
    "TreePieceCellListDataTransferLocal", "TreePieceCellListDataTransferRemote",
    "TreePieceCellListDataTransferRemoteResume", "TreePiecePartListDataTransferLocal",
    "TreePiecePartListDataTransferLocalSmallPhase", "TreePiecePartListDataTransferRemote",
    "TreePiecePartListDataTransferRemoteResume", "TreePiecePartListDataTransferLocalCompact",
    "EwaldHostMemorySetup", "EwaldHostMemoryFree", "EwaldHost", "changaGpuSetCallbackHandler",
    "changaGpuSetAbortOnError", "changaGpuLastError", "changaGpuReleasePools",
    "changaNodeGravity", "changaPartGravity", "changaZeroVars", "changaAddVars", "changaEwald",
    "changaEwaldRoot", "changaEwaldResident", "changaCellRecords",
]

# cells changaCellRecords accepts (CHANGA_CELL_RECORD_MAX)
CELL_RECORD_MAX = 1 << 25

# doubles of the device Ewald root record of changaEwaldRoot
EWALD_ROOT_DOUBLES = 26

# buckets the no-allocation batched launch accepts (its static queue's size)
NOALLOC_MAX_BUCKETS = 1 << 22

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_build.GPU_LIB):
        raise RuntimeError("libchanga_gpu.so is not built: run `python -m changa_amd.build`")
    L = ctypes.CDLL(_build.GPU_LIB)
    P = c_void_p
    L.allocatePinnedHostMemory.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
    L.freePinnedHostMemory.argtypes = [P]
    L.DataManagerTransferLocalTree.argtypes = [P, c_size_t, P, c_size_t, P, c_size_t,
                                               ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                               ctypes.POINTER(c_void_p), P, c_int, P]
    L.DataManagerTransferRemoteChunk.argtypes = [P, c_size_t, P, c_size_t,
                                                 ctypes.POINTER(c_void_p),
                                                 ctypes.POINTER(c_void_p), P, P]
    L.TransferParticleVarsBack.argtypes = [P, c_size_t, P, P, P]
    for name in ["TreePieceCellListDataTransferLocal", "TreePieceCellListDataTransferRemote",
                 "TreePieceCellListDataTransferRemoteResume", "TreePiecePartListDataTransferLocal",
                 "TreePiecePartListDataTransferRemote", "TreePiecePartListDataTransferRemoteResume",
                 "TreePiecePartListDataTransferLocalCompact"]:
        getattr(L, name).argtypes = [ctypes.POINTER(CudaRequest)]
    L.TreePiecePartListDataTransferLocalSmallPhase.argtypes = [ctypes.POINTER(CudaRequest), P, c_int]
    L.EwaldHostMemorySetup.argtypes = [ctypes.POINTER(EwaldData), c_int, c_int, c_int]
    L.EwaldHostMemoryFree.argtypes = [ctypes.POINTER(EwaldData), c_int]
    L.EwaldHost.argtypes = [P, P, ctypes.POINTER(EwaldData), P, P, c_int, c_int]
    L.changaGpuSetCallbackHandler.argtypes = [CALLBACK]
    L.changaGpuSetAbortOnError.argtypes = [c_int]
    L.changaGpuLastError.argtypes = [ctypes.POINTER(ctypes.c_char_p)]
    L.changaGpuLastError.restype = c_int
    L.changaNodeGravity.argtypes = [ctypes.POINTER(ChangaGravityLaunch), P]
    L.changaNodeGravity.restype = c_int
    L.changaCellRecords.argtypes = [P, P, c_int, P, P]
    L.changaCellRecords.restype = c_int
    L.changaPartGravity.argtypes = [ctypes.POINTER(ChangaGravityLaunch), P]
    L.changaPartGravity.restype = c_int
    L.changaZeroVars.argtypes = [P, c_int, P]
    L.changaZeroVars.restype = c_int
    L.changaAddVars.argtypes = [P, P, c_int, P]
    L.changaAddVars.restype = c_int
    L.changaEwald.argtypes = [P, P, P, c_int, c_int, ctypes.POINTER(EwaldReadOnlyData), P, P, P, P]
    L.changaEwald.restype = c_int
    L.changaEwaldRoot.argtypes = [P, ctypes.POINTER(EwaldReadOnlyData), ctypes.c_double, P, P, P]
    L.changaEwaldRoot.restype = c_int
    L.changaEwaldResident.argtypes = [P, P, P, c_int, c_int, ctypes.POINTER(EwaldReadOnlyData), P,
                                      P, P, P]
    L.changaEwaldResident.restype = c_int
    L.changaBuildMoments.argtypes = [c_int, c_int, P, P, P, P, P, P, P, P, P, P, P, P]
    L.changaBuildMoments.restype = c_int
    L.changaBatchPlanCreate.argtypes = [ctypes.POINTER(ChangaGravityLaunch),
                                        ctypes.POINTER(ChangaGravityLaunch),
                                        ctypes.POINTER(ChangaBatchOptions),
                                        ctypes.POINTER(c_void_p), P]
    L.changaBatchPlanCreate.restype = c_int
    for name in ["changaBatchPlanLaunch", "changaBatchPlanRun"]:
        getattr(L, name).argtypes = [P, P]
        getattr(L, name).restype = c_int
    L.changaBatchPlanDestroy.argtypes = [P]
    L.changaBatchPlanDestroy.restype = c_int
    L.changaBatchedGravityNoAlloc.argtypes = [ctypes.POINTER(ChangaGravityLaunch),
                                              ctypes.POINTER(ChangaGravityLaunch),
                                              ctypes.POINTER(ChangaBatchOptions), P]
    L.changaBatchedGravityNoAlloc.restype = c_int
    L.changaGravityAdjoint.argtypes = [ctypes.POINTER(ChangaAdjointLaunch), P]
    L.changaGravityAdjoint.restype = c_int
    L.changaGravityAdjointWorkInts.argtypes = [c_int, c_int, c_int, c_int]
    L.changaGravityAdjointWorkInts.restype = c_size_t
    L.changaGravityAdjointEwaldWorkDoubles.argtypes = [c_int, c_int]
    L.changaGravityAdjointEwaldWorkDoubles.restype = c_size_t
    L.changaBuildListsWorkBytes.argtypes = [c_int, c_int]
    L.changaBuildListsWorkBytes.restype = c_size_t
    L.changaBuildListsCount.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
    L.changaBuildListsFill.argtypes = [c_void_p, c_void_p]
    L.changaBuildTreeWorkBytes.argtypes = [c_int]
    L.changaBuildTreeWorkBytes.restype = c_size_t
    L.changaBuildTree.argtypes = [c_void_p, c_void_p]
    L.changaFusedGravityWorkBytes.argtypes = [c_int, c_int, c_int]
    L.changaFusedGravityWorkBytes.restype = c_size_t
    L.changaFusedGravity.argtypes = [c_void_p, c_void_p]
    # Python callers get a RuntimeError from check() instead of the library's
    # default abort (the cudaChk convention of HostCUDA.cu:39-47)
    L.changaGpuSetAbortOnError(0)
    _lib = L
    return L


_ltw = None


def lib_ltw():
    """libchanga_gpu_ltw.so: the boundary built on the GPU_LOCAL_TREE_WALK
    layouts (160-byte nodes, 24-byte particles, CudaRequestLTW), whose
    TreePieceCellListDataTransferLocal runs the local tree walk. Its errors
    are reported through its own changaGpuLastError (check_ltw)."""
    global _ltw
    if _ltw is not None:
        return _ltw
    if not os.path.exists(_build.LTW_LIB):
        raise RuntimeError("libchanga_gpu_ltw.so is not built: run `python -m changa_amd.build`")
    L = ctypes.CDLL(_build.LTW_LIB)
    P = c_void_p
    for name in ["TreePieceCellListDataTransferLocal", "TreePieceCellListDataTransferRemote",
                 "TreePieceCellListDataTransferRemoteResume", "TreePiecePartListDataTransferLocal",
                 "TreePiecePartListDataTransferRemote", "TreePiecePartListDataTransferRemoteResume",
                 "TreePiecePartListDataTransferLocalCompact"]:
        getattr(L, name).argtypes = [ctypes.POINTER(CudaRequestLTW)]
    L.TreePiecePartListDataTransferLocalSmallPhase.argtypes = [ctypes.POINTER(CudaRequestLTW), P,
                                                               c_int]
    L.allocatePinnedHostMemory.argtypes = [ctypes.POINTER(c_void_p), c_size_t]
    L.freePinnedHostMemory.argtypes = [P]
    L.EwaldHostMemorySetup.argtypes = [ctypes.POINTER(EwaldData), c_int, c_int, c_int]
    L.EwaldHostMemoryFree.argtypes = [ctypes.POINTER(EwaldData), c_int]
    L.EwaldHost.argtypes = [P, P, ctypes.POINTER(EwaldData), P, P, c_int, c_int]
    L.changaGpuSetCallbackHandler.argtypes = [CALLBACK]
    L.TransferParticleVarsBack.argtypes = [P, c_size_t, P, P, P]
    L.changaGpuSetAbortOnError.argtypes = [c_int]
    L.changaGpuLastError.argtypes = [ctypes.POINTER(ctypes.c_char_p)]
    L.changaGpuLastError.restype = c_int
    L.changaGpuReleasePools.argtypes = []
    L.changaZeroVars.argtypes = [P, c_int, P]
    L.changaZeroVars.restype = c_int
    L.changaGpuSetAbortOnError(0)
    _ltw = L
    return L


def check_ltw(what):
    """Raise when the LTW library recorded an error (its calls return void)."""
    msg = ctypes.c_char_p()
    rc = lib_ltw().changaGpuLastError(ctypes.byref(msg))
    if rc != 0:
        raise RuntimeError(f"{what}: {msg.value.decode() if msg.value else rc}")


def check(rc, what):
    if rc != 0:
        msg = ctypes.c_char_p()
        lib().changaGpuLastError(ctypes.byref(msg))
        raise RuntimeError(f"{what}: {msg.value.decode() if msg.value else rc}")
