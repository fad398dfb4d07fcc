// extern "C" work-request surface: ChaNGa's HostCUDA.h / EwaldCUDA.h entry
// points (HostCUDA.cu:66-630, 1887-1956) over the MI355X kernels.
//
// Per-request device copies (interaction list, bucket markers / starts /
// sizes, missed nodes or particles, Ewald markers and h-loop table) go into
// persistent per-stream pools: a slot grows (after synchronizing its stream)
// but is never freed while a kernel of that stream may read it. Completion is
// one host function enqueued behind the request's device work
// (hipLaunchHostFunc), which hands the request's `cb` pointer to the handler
// installed by changaGpuSetCallbackHandler.
#include <atomic>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "common.h"

namespace cgpu {

int launch_ewald(const CompactPartData* parts, VariablePartData* vars, const int* markers,
                 int first, int n, const EwaldReadOnlyData* ro, const EwtData* d_ewt,
                 const double* pos64, const double* rootCm64, const double* d_root,
                 hipStream_t s);
int launch_ewald_root(const double* node64, const EwaldReadOnlyData* ro, double hcut,
                      double* root, EwtData* ewt, hipStream_t s);
#ifdef GPU_LOCAL_TREE_WALK
int ltw_work_bytes(int n, int maxCells, int maxRuns, size_t* bytes);
int ltw_launch(const CudaMultipoleMoments* nodes, int numNodes, const CompactPartData* parts,
               int numParts, VariablePartData* vars, int first, int last, int root, float theta,
               float thetaMono, int nReplicas, float fperiod, float fperiodY, float fperiodZ,
               void* work, size_t workBytes, int maxCells, int maxRuns, int* errOut,
               hipStream_t s);
// per-wavefront scratch of a local-tree-walk request: cells and particle runs
// (229 KB a wavefront; a request runs at most one wavefront per 64
// particles, up to the device's resident count)
constexpr int kLtwMaxCells = 16384;
constexpr int kLtwMaxRuns = 8192;
#endif

static std::atomic<int> g_abort{1};
static std::atomic<changa_gpu_callback_fn> g_handler{nullptr};
static std::mutex g_err_mu;
static char g_err_msg[512];
static int g_err_code = 0;

int fail(const char* what, hipError_t e, const char* file, int line) {
  {
    std::lock_guard<std::mutex> lk(g_err_mu);
    snprintf(g_err_msg, sizeof(g_err_msg), "%s failed at %s:%d: %s (%d)", what, file, line,
             hipGetErrorString(e), (int)e);
    g_err_code = e == hipSuccess ? -1 : (int)e;
  }
  if (g_abort.load()) {
    fprintf(stderr, "Fatal HIP error: %s\n", g_err_msg);
    abort();
  }
  return g_err_code;
}

enum Slot { S_LIST = 0, S_MARKS, S_STARTS, S_SIZES, S_MISSED, S_EWT, S_EWMARK, S_LTW, S_NSLOT };

struct Pool {
  void* buf[S_NSLOT] = {};
  size_t cap[S_NSLOT] = {};
  int* flag = nullptr;  // pinned: the error flag of the stream's last local tree walk
};

static std::mutex g_pool_mu;
static std::unordered_map<hipStream_t, Pool>* g_pools = nullptr;

static int pool_get(hipStream_t s, int slot, size_t bytes, void** out) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pools) g_pools = new std::unordered_map<hipStream_t, Pool>();
  Pool& p = (*g_pools)[s];
  if (bytes == 0) bytes = 16;
  if (p.cap[slot] < bytes) {
    if (p.buf[slot]) {
      // earlier work on this stream may still read the old buffer
      CGPU_CHECK(hipStreamSynchronize(s));
      CGPU_CHECK(hipFree(p.buf[slot]));
      p.buf[slot] = nullptr;
    }
    size_t cap = bytes + bytes / 4;
    CGPU_CHECK(hipMalloc(&p.buf[slot], cap));
    p.cap[slot] = cap;
  }
  *out = p.buf[slot];
  return 0;
}

#ifdef GPU_LOCAL_TREE_WALK
static int pool_flag(hipStream_t s, int** out) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pools) g_pools = new std::unordered_map<hipStream_t, Pool>();
  Pool& p = (*g_pools)[s];
  if (!p.flag) {
    CGPU_CHECK(hipHostMalloc((void**)&p.flag, sizeof(int), hipHostMallocDefault));
    *p.flag = 0;
  }
  *out = p.flag;
  return 0;
}
#endif

static void host_trampoline(void* cb) {
  changa_gpu_callback_fn fn = g_handler.load();
  if (fn) fn(cb);
}

static int add_callback(hipStream_t s, void* cb) {
  if (!cb || !g_handler.load()) return 0;
  CGPU_CHECK(hipLaunchHostFunc(s, host_trampoline, cb));
  return 0;
}

static int upload(hipStream_t s, int slot, const void* host, size_t bytes, void** dev) {
  if (int e = pool_get(s, slot, bytes, dev)) return e;
  if (bytes) CGPU_CHECK(hipMemcpyAsync(*dev, host, bytes, hipMemcpyHostToDevice, s));
  return 0;
}

// Common path of the TreePiece*ListDataTransfer* calls.
static int list_request(CudaRequest* d, bool cells, size_t entryBytes, int listKind,
                        const CudaMultipoleMoments* moments, const CompactPartData* sources) {
  if (!d) return fail("request: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)d->stream;
  const int nb = d->numBucketsPlusOne - 1;
  if (nb > 0 && d->numInteractions > 0) {
    void *dl, *dm, *ds, *dz;
    if (int e = upload(s, S_LIST, d->list, (size_t)d->numInteractions * entryBytes, &dl)) return e;
    if (int e = upload(s, S_MARKS, d->bucketMarkers, sizeof(int) * (size_t)(nb + 1), &dm)) return e;
    if (int e = upload(s, S_STARTS, d->bucketStarts, sizeof(int) * (size_t)nb, &ds)) return e;
    if (int e = upload(s, S_SIZES, d->bucketSizes, sizeof(int) * (size_t)nb, &dz)) return e;
    ChangaGravityLaunch a;
    std::memset(&a, 0, sizeof(a));
    a.targets = d->d_localParts;
    a.vars = d->d_localVars;
    a.moments = moments;
    a.sources = sources;
    a.list = dl;
    a.marks = (const int*)dm;
    a.starts = (const int*)ds;
    a.sizes = (const int*)dz;
    a.numBuckets = nb;
    a.listKind = listKind;
    a.fperiod = d->fperiod;
    int e = cells ? changaNodeGravity(&a, s) : changaPartGravity(&a, s);
    if (e) return e;
  }
  return add_callback(s, d->cb);
}

}  // namespace cgpu

using namespace cgpu;

extern "C" {

void changaGpuSetCallbackHandler(changa_gpu_callback_fn fn) { g_handler.store(fn); }
void changaGpuSetAbortOnError(int a) { g_abort.store(a ? 1 : 0); }

int changaGpuLastError(const char** message) {
  std::lock_guard<std::mutex> lk(g_err_mu);
  int c = g_err_code;
  if (message) *message = g_err_msg;
  g_err_code = 0;
  return c;
}

void changaGpuReleasePools(void) {
  std::lock_guard<std::mutex> lk(g_pool_mu);
  if (!g_pools) return;
  for (auto& kv : *g_pools) {
    for (int k = 0; k < S_NSLOT; ++k)
      if (kv.second.buf[k]) (void)hipFree(kv.second.buf[k]);
    if (kv.second.flag) (void)hipHostFree(kv.second.flag);
  }
  g_pools->clear();
}

void allocatePinnedHostMemory(void** ptr, size_t size) {
  if (!ptr) return;
  *ptr = nullptr;
  if (size == 0) {
    fail("allocatePinnedHostMemory: 0 size", hipErrorInvalidValue, __FILE__, __LINE__);
    return;
  }
  hipError_t e = hipHostMalloc(ptr, size, hipHostMallocDefault);
  if (e != hipSuccess) fail("hipHostMalloc", e, __FILE__, __LINE__);
}

void freePinnedHostMemory(void* ptr) {
  if (!ptr) {
    fail("freePinnedHostMemory: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
    return;
  }
  hipError_t e = hipHostFree(ptr);
  if (e != hipSuccess) fail("hipHostFree", e, __FILE__, __LINE__);
}

static int transfer_local_tree(void* moments, size_t sMoments, void* compactParts,
                               size_t sCompactParts, void* varParts, size_t sVarParts,
                               void** d_localMoments, void** d_compactParts, void** d_varParts,
                               void* stream, int numParticles, void* callback) {
  hipStream_t s = (hipStream_t)stream;
  CGPU_CHECK(hipMalloc(d_localMoments, sMoments ? sMoments : 16));
  CGPU_CHECK(hipMalloc(d_compactParts, sCompactParts ? sCompactParts : 16));
  CGPU_CHECK(hipMalloc(d_varParts, sVarParts ? sVarParts : 16));
  if (sMoments)
    CGPU_CHECK(hipMemcpyAsync(*d_localMoments, moments, sMoments, hipMemcpyHostToDevice, s));
  if (sCompactParts)
    CGPU_CHECK(hipMemcpyAsync(*d_compactParts, compactParts, sCompactParts,
                              hipMemcpyHostToDevice, s));
  if (sVarParts)
    CGPU_CHECK(hipMemcpyAsync(*d_varParts, varParts, sVarParts, hipMemcpyHostToDevice, s));
  if (int e = changaZeroVars((VariablePartData*)*d_varParts, numParticles, s)) return e;
  return add_callback(s, callback);
}

void DataManagerTransferLocalTree(void* moments, size_t sMoments, void* compactParts,
                                  size_t sCompactParts, void* varParts, size_t sVarParts,
                                  void** d_localMoments, void** d_compactParts, void** d_varParts,
                                  void* stream, int numParticles, void* callback) {
  (void)transfer_local_tree(moments, sMoments, compactParts, sCompactParts, varParts, sVarParts,
                            d_localMoments, d_compactParts, d_varParts, stream, numParticles,
                            callback);
}

static int transfer_remote(void* moments, size_t sMoments, void* parts, size_t sParts,
                           void** d_remoteMoments, void** d_remoteParts, void* stream,
                           void* callback) {
  hipStream_t s = (hipStream_t)stream;
  CGPU_CHECK(hipMalloc(d_remoteMoments, sMoments ? sMoments : 16));
  CGPU_CHECK(hipMalloc(d_remoteParts, sParts ? sParts : 16));
  if (sMoments)
    CGPU_CHECK(hipMemcpyAsync(*d_remoteMoments, moments, sMoments, hipMemcpyHostToDevice, s));
  if (sParts)
    CGPU_CHECK(hipMemcpyAsync(*d_remoteParts, parts, sParts, hipMemcpyHostToDevice, s));
  return add_callback(s, callback);
}

void DataManagerTransferRemoteChunk(void* moments, size_t sMoments, void* compactParts,
                                    size_t sCompactParts, void** d_remoteMoments,
                                    void** d_remoteParts, void* stream, void* callback) {
  (void)transfer_remote(moments, sMoments, compactParts, sCompactParts, d_remoteMoments,
                        d_remoteParts, stream, callback);
}

static int vars_back(VariablePartData* hostBuffer, size_t size, void* d_varParts, void* stream,
                     void* cb) {
  hipStream_t s = (hipStream_t)stream;
  if (size) CGPU_CHECK(hipMemcpyAsync(hostBuffer, d_varParts, size, hipMemcpyDeviceToHost, s));
  return add_callback(s, cb);
}

void TransferParticleVarsBack(VariablePartData* hostBuffer, size_t size, void* d_varParts,
                              void* stream, void* cb) {
  (void)vars_back(hostBuffer, size, d_varParts, stream, cb);
}

#ifdef GPU_LOCAL_TREE_WALK
// runs in stream order after the walk's flag has been copied to host memory
static void ltw_check(void* p) {
  const int f = *static_cast<const int*>(p);
  if (f & 1)
    fail("TreePieceCellListDataTransferLocal: a local tree walk exceeded its scratch",
         hipErrorInvalidValue, __FILE__, __LINE__);
  if (f & 2)
    fail("TreePieceCellListDataTransferLocal: a node or particle index outside its array",
         hipErrorInvalidValue, __FILE__, __LINE__);
}

// The local tree walk of particles [firstParticle, lastParticle]
// (HostCUDA.cu:190-224): the request's interaction list is a placeholder
// (Compute.cpp:1437-1530) and is not read.
static int local_tree_walk(CudaRequest* d) {
  if (!d) return fail("request: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)d->stream;
  const long long nn = (long long)(d->sMoments / sizeof(CudaMultipoleMoments));
  const long long np = (long long)(d->sCompactParts / sizeof(CompactPartData));
  if (nn > 0x7fffffffll || np > 0x7fffffffll)
    return fail("TreePieceCellListDataTransferLocal: sMoments / sCompactParts too large",
                hipErrorInvalidValue, __FILE__, __LINE__);
  if (d->lastParticle < d->firstParticle) return add_callback(s, d->cb);  // no particles
  size_t bytes = 0;
  const int n = d->lastParticle - d->firstParticle + 1;
  if (int e = ltw_work_bytes(n, kLtwMaxCells, kLtwMaxRuns, &bytes)) return e;
  void* work = nullptr;
  int* flag = nullptr;
  if (int e = pool_get(s, S_LTW, bytes, &work)) return e;
  if (int e = pool_flag(s, &flag)) return e;
  if (int e = ltw_launch(d->d_localMoments, (int)nn, d->d_localParts, (int)np, d->d_localVars,
                         d->firstParticle, d->lastParticle, d->rootIdx, (float)d->theta,
                         (float)d->thetaMono, d->nReplicas, d->fperiod, d->fperiodY,
                         d->fperiodZ, work, bytes, kLtwMaxCells, kLtwMaxRuns, flag, s))
    return e;
  CGPU_CHECK(hipLaunchHostFunc(s, ltw_check, flag));
  return add_callback(s, d->cb);
}

void TreePieceCellListDataTransferLocal(CudaRequest* d) { (void)local_tree_walk(d); }
#else
void TreePieceCellListDataTransferLocal(CudaRequest* d) {
  (void)list_request(d, true, sizeof(ILCell), 0, d ? d->d_localMoments : nullptr, nullptr);
}
#endif

void TreePieceCellListDataTransferRemote(CudaRequest* d) {
  (void)list_request(d, true, sizeof(ILCell), 0, d ? d->d_remoteMoments : nullptr, nullptr);
}

static int cell_remote_resume(CudaRequest* d) {
  if (!d) return fail("request: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
  void* dm;
  if (int e = upload((hipStream_t)d->stream, S_MISSED, d->missedNodes, d->sMissed, &dm)) return e;
  return list_request(d, true, sizeof(ILCell), 0, (const CudaMultipoleMoments*)dm, nullptr);
}

void TreePieceCellListDataTransferRemoteResume(CudaRequest* d) { (void)cell_remote_resume(d); }

void TreePiecePartListDataTransferLocal(CudaRequest* d) {
  (void)list_request(d, false, sizeof(ILCell), 0, nullptr, d ? d->d_localParts : nullptr);
}

void TreePiecePartListDataTransferLocalCompact(CudaRequest* d) {
  (void)list_request(d, false, sizeof(ILPart), 1, nullptr, d ? d->d_localParts : nullptr);
}

static int part_small_phase(CudaRequest* d, CompactPartData* parts, int len) {
  if (!d) return fail("request: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
  void* dp;
  if (int e = upload((hipStream_t)d->stream, S_MISSED, parts,
                     sizeof(CompactPartData) * (size_t)(len > 0 ? len : 0), &dp))
    return e;
  return list_request(d, false, sizeof(ILCell), 0, nullptr, (const CompactPartData*)dp);
}

void TreePiecePartListDataTransferLocalSmallPhase(CudaRequest* d, CompactPartData* parts,
                                                  int len) {
  (void)part_small_phase(d, parts, len);
}

void TreePiecePartListDataTransferRemote(CudaRequest* d) {
  (void)list_request(d, false, sizeof(ILCell), 0, nullptr, d ? d->d_remoteParts : nullptr);
}

static int part_remote_resume(CudaRequest* d) {
  if (!d) return fail("request: NULL", hipErrorInvalidValue, __FILE__, __LINE__);
  void* dp;
  if (int e = upload((hipStream_t)d->stream, S_MISSED, d->missedParts, d->sMissed, &dp)) return e;
  return list_request(d, false, sizeof(ILCell), 0, nullptr, (const CompactPartData*)dp);
}

void TreePiecePartListDataTransferRemoteResume(CudaRequest* d) { (void)part_remote_resume(d); }

void EwaldHostMemorySetup(EwaldData* h, int size, int nEwhLoop, int largephase) {
  if (!h) return;
  if (largephase)
    allocatePinnedHostMemory((void**)&h->EwaldMarkers, (size_t)(size > 0 ? size : 1) * sizeof(int));
  else
    h->EwaldMarkers = nullptr;
  allocatePinnedHostMemory((void**)&h->ewt, (size_t)(nEwhLoop > 0 ? nEwhLoop : 1) * sizeof(EwtData));
  allocatePinnedHostMemory((void**)&h->cachedData, sizeof(EwaldReadOnlyData));
}

void EwaldHostMemoryFree(EwaldData* h, int largephase) {
  if (!h) return;
  if (largephase && h->EwaldMarkers) freePinnedHostMemory(h->EwaldMarkers);
  if (h->ewt) freePinnedHostMemory(h->ewt);
  if (h->cachedData) freePinnedHostMemory(h->cachedData);
}

static int ewald_host(CompactPartData* parts, VariablePartData* vars, EwaldData* h, void* stream,
                      void* cb, int largephase) {
  if (!h || !h->cachedData)
    return fail("EwaldHost: missing data", hipErrorInvalidValue, __FILE__, __LINE__);
  hipStream_t s = (hipStream_t)stream;
  const EwaldReadOnlyData* ro = h->cachedData;
  int n = largephase ? ro->n : h->EwaldRange[1] - h->EwaldRange[0] + 1;
  void *dmark = nullptr, *dewt = nullptr;
  if (largephase)
    if (int e = upload(s, S_EWMARK, h->EwaldMarkers, sizeof(int) * (size_t)(n > 0 ? n : 0), &dmark))
      return e;
  if (int e = upload(s, S_EWT, h->ewt, sizeof(EwtData) * (size_t)ro->nEwhLoop, &dewt)) return e;
  if (int e = launch_ewald(parts, vars, (const int*)dmark, largephase ? 0 : h->EwaldRange[0], n, ro,
                           (const EwtData*)dewt, nullptr, nullptr, nullptr, s))
    return e;
  return add_callback(s, cb);
}

void EwaldHost(CompactPartData* d_localParts, VariablePartData* d_localVars, EwaldData* h_idata,
               void* stream, void* cb, int myIndex, int largephase) {
  (void)myIndex;
  (void)ewald_host(d_localParts, d_localVars, h_idata, stream, cb, largephase);
}

int changaEwald(const CompactPartData* parts, VariablePartData* vars, const int* markers,
                int first, int n, const EwaldReadOnlyData* ro, const EwtData* ewt,
                const double* pos64, const double* rootCm64, void* stream) {
  if (n == 0) return 0;  // no active particles
  if (!ro || !ewt || !vars || (!parts && !pos64))
    return fail("changaEwald: missing argument", hipErrorInvalidValue, __FILE__, __LINE__);
  return launch_ewald(parts, vars, markers, first, n, ro, ewt, pos64, rootCm64, nullptr,
                      (hipStream_t)stream);
}

int changaEwaldRoot(const double* node64, const EwaldReadOnlyData* ro, double hCut, double* root,
                    EwtData* ewt, void* stream) {
  return launch_ewald_root(node64, ro, hCut, root, ewt, (hipStream_t)stream);
}

int changaEwaldResident(const CompactPartData* parts, VariablePartData* vars, const int* markers,
                        int first, int n, const EwaldReadOnlyData* ro, const EwtData* ewt,
                        const double* pos64, const double* root, void* stream) {
  if (n == 0) return 0;  // no active particles
  if (!ro || !ewt || !vars || !root || (!parts && !pos64))
    return fail("changaEwaldResident: missing argument", hipErrorInvalidValue, __FILE__,
                __LINE__);
  return launch_ewald(parts, vars, markers, first, n, ro, ewt, pos64, nullptr, root,
                      (hipStream_t)stream);
}

}  // extern "C"
