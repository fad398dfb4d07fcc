/*
 * changa_gpu.h -- C-ABI of the MI355X gravity force engine (libchanga_gpu.so).
 *
 * Part 1 mirrors ChaNGa's GPU work-request surface, HostCUDA.h:99-126 and
 * EwaldCUDA.h:59-64, with the same names, argument meaning and wire layouts
 * (changa_gpu_types.h). Differences a caller sees:
 *   - every function has C linkage; streams are hipStream_t passed as void*;
 *   - completion: instead of hapiAddCallback(stream, cb), each call enqueues
 *     one host function on its stream after all of its device work; that
 *     function calls the handler installed with changaGpuSetCallbackHandler
 *     with the request's `cb` pointer (nothing happens when either is NULL);
 *   - per-request device copies live in persistent per-stream pools: nothing
 *     is freed while a kernel may still read it;
 *   - cells flagged CHANGA_ILCELL_SOFTENED are evaluated on the GPU as a
 *     spline-softened monopole;
 *   - the Ewald h-loop table has no 80-entry limit;
 *   - errors abort the process with a message (the reference's cudaChk)
 *     unless changaGpuSetAbortOnError(0) was called; then the call returns,
 *     changaGpuLastError() reports the failure and the callback is not run.
 *
 * Part 2 is the device-resident interface the mirror is built on: the same
 * kernels on lists already in device memory, the GPU multipole build, and the
 * precision and batched forms. It has no counterpart in the reference.
 */
#ifndef CHANGA_GPU_H
#define CHANGA_GPU_H

#include "changa_gpu_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================
 * Part 1: HostCUDA.h / EwaldCUDA.h mirror
 * ====================================================================== */

/* HostCUDA.cu:66-83 (hapiMallocHost): pinned host memory. */
void allocatePinnedHostMemory(void **ptr, size_t size);
/* HostCUDA.cu:85-101 (hapiFreeHost). */
void freePinnedHostMemory(void *ptr);

/* HostCUDA.cu:115-149: allocate the device tree (moments, particles,
   accelerations), copy it host->device on `stream` and zero the first
   numParticles accelerations (ZeroVars, HostCUDA.cu:2195-2205). The three
   device pointers are returned through the out-parameters; the caller
   releases them with hipFree after its last request completed. */
void DataManagerTransferLocalTree(void *moments, size_t sMoments, void *compactParts,
                                  size_t sCompactParts, void *varParts, size_t sVarParts,
                                  void **d_localMoments, void **d_compactParts,
                                  void **d_varParts, void *stream, int numParticles,
                                  void *callback);

/* HostCUDA.cu:159-183: remote moments and particles to the device. */
void DataManagerTransferRemoteChunk(void *moments, size_t sMoments, void *compactParts,
                                    size_t sCompactParts, void **d_remoteMoments,
                                    void **d_remoteParts, void *stream, void *callback);

/* HostCUDA.cu:623-630: accelerations device->host. */
void TransferParticleVarsBack(VariablePartData *hostBuffer, size_t size, void *d_varParts,
                              void *stream, void *cb);

/* HostCUDA.cu:190-248 / 252-292 / 296-341: particle-cell (hexadecapole)
   forces. data->list is ILCell[numInteractions] (host memory);
   bucketMarkers[numBucketsPlusOne], bucketStarts/bucketSizes[numBuckets]
   index d_localParts; results are added into d_localVars (+= for a, phi;
   max for dtGrav). Sources: d_localMoments, d_remoteMoments, or
   data->missedNodes (CudaMultipoleMoments, data->sMissed bytes).
   In the GPU_LOCAL_TREE_WALK build (libchanga_gpu_ltw.so, this header with
   that macro defined) TreePieceCellListDataTransferLocal is the local tree
   walk of HostCUDA.cu:210-224 instead: the list fields are not read; every
   particle in [firstParticle, lastParticle] of d_localParts walks the
   serialized local tree d_localMoments (sMoments bytes) from rootIdx, per
   periodic replica, with theta / thetaMono, and gets the PC, softened-cell
   and PP forces of what the walk accepts (+= / max as above), with the
   replica periods fperiod, fperiodY, fperiodZ per axis; a walk beyond the
   per-wavefront scratch (16,384 cells, 8,192 particle runs) or a child or
   bucket index outside its array is reported as an error. */
void TreePieceCellListDataTransferLocal(CudaRequest *data);
void TreePieceCellListDataTransferRemote(CudaRequest *data);
void TreePieceCellListDataTransferRemoteResume(CudaRequest *data);

/* HostCUDA.cu:412-460 / 345-408 / 464-513 / 517-572: particle-particle
   forces. As ChaNGa's GenericList<ILPart>::serialize (Compute.cpp:1118-1218)
   produces it, data->list is ILCell[numInteractions], one entry per source
   particle {particle index, offsetID}. Sources: d_localParts, the `parts`
   array (small phase), d_remoteParts, or data->missedParts. */
void TreePiecePartListDataTransferLocal(CudaRequest *data);
void TreePiecePartListDataTransferLocalSmallPhase(CudaRequest *data, CompactPartData *parts,
                                                  int len);
void TreePiecePartListDataTransferRemote(CudaRequest *data);
void TreePiecePartListDataTransferRemoteResume(CudaRequest *data);

/* HostCUDA.cu:1887-1901: pinned Ewald buffers (markers only in large phase). */
void EwaldHostMemorySetup(EwaldData *h_idata, int size, int nEwhLoop, int largephase);
void EwaldHostMemoryFree(EwaldData *h_idata, int largephase);
/* HostCUDA.cu:1917-1956: Ewald correction (real-space replicas + k-space
   h-loop) for the active particles: markers[0..n) in large phase, the range
   EwaldRange otherwise. Results are added into d_localVars. */
void EwaldHost(CompactPartData *d_localParts, VariablePartData *d_localVars,
               EwaldData *h_idata, void *stream, void *cb, int myIndex, int largephase);

/* ======================================================================
 * Part 1b: runtime controls (no reference counterpart)
 * ====================================================================== */
typedef void (*changa_gpu_callback_fn)(void *cb);
void changaGpuSetCallbackHandler(changa_gpu_callback_fn fn);
void changaGpuSetAbortOnError(int abort_on_error);
/* 0 when no error happened since the last call; the message is kept in a
   static buffer. */
int changaGpuLastError(const char **message);
/* Releases the per-stream pools (after the caller synchronized). */
void changaGpuReleasePools(void);

/* ILPart-list variant of TreePiecePartListDataTransferLocal: data->list is
   ILPart[numInteractions] (one entry per source bucket, 12 B instead of 8 B
   per source particle) and is expanded on the GPU. */
void TreePiecePartListDataTransferLocalCompact(CudaRequest *data);

/* ======================================================================
 * Part 2: device-resident interface
 * ====================================================================== */

/* One particle-cell or particle-particle launch over lists in device memory.
   Precision form: when pos64 is non-NULL the kernels take target and source
   positions (and cell centres, cm64) in FP64, subtract the bucket centre
   (the first target particle's position) in FP64 and evaluate the
   interaction in FP32 on the bucket-centred coordinates. */
typedef struct {
  const CompactPartData *targets;        /* particles indexed by starts[] */
  VariablePartData *vars;                /* results (+=, max) */
  const CudaMultipoleMoments *moments;   /* PC sources */
  const CompactPartData *sources;        /* PP sources */
  const void *list;                      /* ILCell* (PC, PP) or ILPart* (PP compact) */
  const int *marks;                      /* numBuckets + 1 */
  const int *starts;                     /* numBuckets */
  const int *sizes;                      /* numBuckets */
  int numBuckets;
  int listKind;                          /* 0 ILCell per source, 1 ILPart runs */
  float fperiod;
  int variant;                           /* reserved: must be 0 */
  const double *tpos64;                  /* 3 per target particle, or NULL */
  const double *spos64;                  /* 3 per source particle, or NULL */
  const double *cm64;                    /* 3 per cell, or NULL */
  const void *cellRecords;               /* 128-byte cell records of changaCellRecords (built
                                            from these moments and cm64), or NULL: with them a
                                            tile of cells is fetched as whole cache lines */
  float fperiodY, fperiodZ;              /* periods of the y and z replica offsets (a box
                                            that is not a cube); 0: fperiod */
} ChangaGravityLaunch;

int changaNodeGravity(const ChangaGravityLaunch *a, void *stream);
/* No counterpart in the reference (ChangaGravityLaunch.cellRecords): one
   128-byte record per cell, built from the wire moments and the FP64 centres
   of mass (cm64, or NULL: the wire centres), with the moments pre-scaled for
   the hexadecapole evaluation. `records` holds CHANGA_CELL_RECORD_BYTES x
   numCells bytes, 128-byte aligned (hipMalloc gives 256), for at most
   CHANGA_CELL_RECORD_MAX cells (the loops address them with 32-bit byte
   offsets). Rebuild after every moment build. */
#define CHANGA_CELL_RECORD_BYTES 128
#define CHANGA_CELL_RECORD_MAX (1 << 25)
int changaCellRecords(const CudaMultipoleMoments *moments, const double *cm64, int numCells,
                      void *records, void *stream);
int changaPartGravity(const ChangaGravityLaunch *a, void *stream);
int changaZeroVars(VariablePartData *vars, int n, void *stream);
/* dst += src (a and potential; dtGrav = max): merges a result array filled on
   another stream, e.g. an Ewald pass overlapped with the tree forces. */
int changaAddVars(VariablePartData *dst, const VariablePartData *src, int n, void *stream);

/* Ewald on device-resident data. markers: n active particle indices into
   parts (device memory; NULL: particles first..first+n-1). ewt: the
   ro->nEwhLoop table entries in device memory. ro: host struct (its
   momcRoot, mm, L, fEwCut, nReps and nEwReps are used). Precision form:
   pos64 (device, 3 per particle) and rootCm64 (host, 3 doubles) or NULL. */
int changaEwald(const CompactPartData *parts, VariablePartData *vars, const int *markers,
                int first, int n, const EwaldReadOnlyData *ro, const EwtData *ewt,
                const double *pos64, const double *rootCm64, void *stream);

/* TreePiece::EwaldInit (Ewald.cpp:285-383) on the device, so that an
   evaluation on moved particles needs no host round trip between the
   multipole build and the Ewald pass: from the root's node record
   (node64[0] of changaBuildMoments) the root record `root` (device,
   CHANGA_EWALD_ROOT_DOUBLES doubles: centre of mass x, y, z, mass, then the
   22 reduced moments at unit scale in FMOMR order, as momRescaleFmomr +
   momFmomr2Momc give them) and the h-loop table `ewt` (device, the
   ro->nEwhLoop entries with 0 < |h|^2 <= hCut^2, hx outermost, hz innermost;
   a different ro->nEwhLoop is an error). Uses ro->L. */
#define CHANGA_EWALD_ROOT_DOUBLES 26
int changaEwaldRoot(const double *node64, const EwaldReadOnlyData *ro, double hCut, double *root,
                    EwtData *ewt, void *stream);
/* changaEwald with the root of changaEwaldRoot (device) in place of
   ro->mm, ro->momcRoot and rootCm64. */
int changaEwaldResident(const CompactPartData *parts, VariablePartData *vars, const int *markers,
                        int first, int n, const EwaldReadOnlyData *ro, const EwtData *ewt,
                        const double *pos64, const double *root, void *stream);

/* GPU multipole build (the CPU path is GenericTreeNode::makeBucket,
   GenericTreeNode.h:219-257, and LocalTreeBuilder::doneChildren,
   Compute.cpp:2534-2580). Node topology (breadth-first numbering, children
   after parents, depth d occupying nodes [levelOffsets[d],
   levelOffsets[d+1])): first/count (particle range), child (2 per node, -1
   for a bucket), boxRadius (initial expansion radius of a bucket: half the
   diagonal of its spatial cell). Particles: pos (FP64 x,y,z), mass, soft, in
   tree order. Outputs: moments (wire, FP32), node64 (34 doubles per node:
   cm.xyz, radius, soft, mass, lo.xyz, hi.xyz, 22 radius-scaled reduced
   moments) and cm64 (3 doubles per node, may be NULL). All pointers are
   device pointers. */
int changaBuildMoments(int nNodes, int nDepths, const int *levelOffsets /* host */,
                       const int *first, const int *count, const int *child,
                       const double *boxRadius, const double *pos, const double *mass,
                       const double *soft, CudaMultipoleMoments *moments, double *node64,
                       double *cm64, void *stream);

/* ======================================================================
 * Part 3: batched form
 *
 * Every bucket's cell list and particle list in one persistent launch: a
 * wavefront claims buckets from per-XCD work queues and evaluates the cell
 * list, then the particle list, of each into one set of accumulators (one
 * += per particle and bucket instead of one per list). The plan is bound to
 * the buffers of the two launch descriptions it was made from; make a new
 * plan when they change. The graph it holds contains kernels only.
 * ====================================================================== */
typedef struct ChangaBatchPlan ChangaBatchPlan;

typedef struct {
  int pcTile;        /* cells per LDS tile: 32 or 64; 0 = autotune */
  int ppTile;        /* source particles per LDS tile: 128 or 256; 0 = autotune */
  int listsPerGrab;  /* bucket lists a wavefront claims per queue access, 1..64; 0 = autotune */
  int order;         /* 1 tree order, 2 costliest bucket first (per XCD range); 0 = autotune */
  int zeroVars;      /* 1: each run first zeroes vars (particles covered by the buckets) */
  float tunedMs;     /* out: time of one evaluation with the chosen options, or 0 */
  int targetsPerLane; /* 1: one target per lane; 3: one to three, by bucket size (more slots
                         busy, but a kernel three chunk loops long); 0 = autotune */
  int cellSource;     /* 1: the launch's cellRecords (whole-line fetches: faster when buckets
                         have few targets); 2: the wire moments (per-lane fetches); 0 =
                         autotune between the two when the launch has cellRecords, else 2 */
} ChangaBatchOptions;

/* cells: the particle-cell launch (listKind 0); parts: the particle-particle
   launch (listKind 0 or 1) over the same buckets, targets, vars and
   precision-form positions. Options left at 0 are chosen by timing the
   candidates on these inputs (results go to a scratch array; vars is not
   touched): the default (64, 256, 1, tree order, 3) is timed first and
   another candidate is taken only when it is more than 3% faster. *opts
   returns the choice. Synchronizes the stream. */
int changaBatchPlanCreate(const ChangaGravityLaunch *cells, const ChangaGravityLaunch *parts,
                          ChangaBatchOptions *opts, ChangaBatchPlan **plan, void *stream);
/* One evaluation, enqueued as kernels on `stream` (usable inside the
   caller's own stream capture). */
int changaBatchPlanLaunch(ChangaBatchPlan *plan, void *stream);
/* One evaluation, replaying the plan's graph on `stream`. */
int changaBatchPlanRun(ChangaBatchPlan *plan, void *stream);
int changaBatchPlanDestroy(ChangaBatchPlan *plan);

/* No counterpart in the reference. For callers that cannot allocate device
   memory (code under stream capture, tight serving loops, harnesses that
   forbid hipMalloc): one batched evaluation like changaBatchPlanLaunch that
   calls no allocator. Its work queue (range starts and bucket ids in tree
   order), queue counters and cost sums live in a static device array of
   4 x (CHANGA_NOALLOC_MAX_BUCKETS + 1,296) bytes (about 16 MiB), which each
   GPU the library runs on holds until the process ends; every call rewrites
   what it reads of it. Calls that overlap on one GPU (two streams, two host
   threads) race on that array: serialize them. opts may be NULL (cell tile
   64, particle tile 256, one bucket per claim, up to three targets per
   lane, the cell records when the launch has them); fields left at 0 take those values;
   order must be 1 (tree order: the costliest-first order needs a sort); zeroVars zeroes the results of the particles the buckets cover.
   More than CHANGA_NOALLOC_MAX_BUCKETS buckets is an error. Results equal
   changaBatchPlanLaunch's with the same options. */
#define CHANGA_NOALLOC_MAX_BUCKETS (1 << 22)
int changaBatchedGravityNoAlloc(const ChangaGravityLaunch *cells, const ChangaGravityLaunch *parts,
                                const ChangaBatchOptions *opts, void *stream);

/* ======================================================================
 * Part 4: adjoint form
 *
 * For L = sum_i w_i . a_i, with a the tree acceleration of the cell and
 * particle lists (precision form), the exact gradient dL/dx and dL/dm of
 * every particle on the same lists: target-side tidal terms, particle-source
 * terms, and cell-source terms carried down the tree as polynomials about
 * each node's centre of mass. Softening lengths are constants. Periodic
 * replicas in the lists are handled. With ewaldRO set, L also holds the
 * Ewald correction of changaEwald at the targets (its target-side terms, and
 * its dependence on the root's centre of mass, moments and h-table).
 * ====================================================================== */
typedef struct {
  ChangaGravityLaunch cells;      /* particle-cell launch (listKind 0, tpos64 and cm64 set) */
  ChangaGravityLaunch parts;      /* particle-particle launch (spos64 set), same buckets */
  const float *w;                 /* dL/da, 3 per target, in target order (device) */
  const int *targetIndex;         /* particle index of each target (device), or NULL when the
                                     targets are the particles themselves */
  double *grad;                   /* out, overwritten: dL/dx (3), dL/dm per particle (device) */
  double *nodeWork;               /* scratch: 112 doubles per node (device) */
  int *listWork;                  /* scratch: changaGravityAdjointWorkInts(nNodes,
                                     numCellEntries, numPartEntries, nParticles) ints (device) */
  size_t listWorkInts;            /* its size in ints */
  int numCellEntries;             /* entries of the cell list (cells.marks[numBuckets]) */
  int numPartEntries;             /* entries of the particle list (parts.marks[numBuckets]) */
  int nNodes, nDepths, nParticles;
  const int *levelOffsets;        /* host, nDepths + 1 (as for changaBuildMoments) */
  const int *first, *count, *child;  /* device, as for changaBuildMoments */
  const double *node64;           /* device: node64 of changaBuildMoments */
  const double *pos64, *mass64;   /* device: particles in tree order */
  /* Ewald term (optional; all NULL / 0 for none) */
  const EwaldReadOnlyData *ewaldRO;  /* host: the record given to changaEwald */
  const EwtData *ewt;                /* device: its h-table */
  const double *rootCm64;            /* host: root centre of mass (FP64), as for changaEwald */
  int nTargets;                      /* targets (rows of w), positions cells.tpos64 */
  double *ewaldWork;                 /* scratch: changaGravityAdjointEwaldWorkDoubles(nTargets,
                                        ewaldRO->nEwhLoop) doubles (device) */
  size_t ewaldWorkDoubles;           /* its size in doubles */
} ChangaAdjointLaunch;

/* Ints of listWork for a tree of nNodes nodes, a cell list of numCellEntries
   entries and a particle list of numPartEntries entries over nParticles
   source particles (the cell list transposed per cell; an ILPart particle
   list transposed per source run). */
size_t changaGravityAdjointWorkInts(int nNodes, int numCellEntries, int numPartEntries,
                                    int nParticles);
/* Doubles of ewaldWork for nTargets targets and an h-table of nH entries. */
size_t changaGravityAdjointEwaldWorkDoubles(int nTargets, int nH);
int changaGravityAdjoint(const ChangaAdjointLaunch *a, void *stream);

/* ======================================================================
 * Part 5: interaction-list construction on the GPU
 *
 * The caller side of the cell and particle lists (ListCompute::doWork /
 * stateReady, Compute.cpp:690-884, 1543-1922, with openCriterionBucket,
 * gravity.h:597-635, and openSoftening, gravity.h:251-260) as a GPU pass
 * over the node records changaBuildMoments writes: for each target bucket,
 * the ILCell list (hexadecapole cells, then softened cells flagged in bit
 * 31 of offsetID) and the ILPart list, in the layout the launches above
 * take. Two calls: Count (per-bucket sizes into cellMarks / partMarks /
 * softStart, synchronizes the stream and returns the totals), then, once
 * the caller has allocated cells and parts, Fill.
 * ====================================================================== */
typedef struct {
  int nBuckets;                   /* target buckets */
  int numNodes;                   /* nodes of the tree */
  const int *bucketNode;          /* device: node of each target bucket */
  const int *first, *count, *child;  /* device: topology, as for changaBuildMoments */
  const double *node64;           /* device: node records of changaBuildMoments */
  double theta;                   /* opening angle */
  int nReplicas;                  /* periodic replicas per side (0: isolated; at most 3) */
  double period;                  /* box size when periodic */
  int *cellMarks, *partMarks;     /* device out: nBuckets + 1 list offsets each */
  int *softStart;                 /* device out: first softened cell of each bucket */
  int *sourceCounts;              /* device out, optional: source particles of each bucket */
  ILCell *cells;                  /* device out (Fill): the cell list */
  ILPart *parts;                  /* device out (Fill): the particle runs */
  void *work;                     /* device scratch: changaBuildListsWorkBytes(nBuckets,
                                     numNodes); Count leaves the walk records there for Fill */
  size_t workBytes;
} ChangaListBuild;

size_t changaBuildListsWorkBytes(int nBuckets, int numNodes);
int changaBuildListsCount(const ChangaListBuild *b, long long *numCells, long long *numParts,
                          void *stream);
int changaBuildListsFill(const ChangaListBuild *b, void *stream);

/* ======================================================================
 * Part 6: tree build on the GPU
 *
 * The caller side's tree (TreePiece::buildTree, TreePiece.cpp:2619, and
 * LocalTreeBuilder, Compute.cpp:2440-2580): particles in Morton order (21
 * bits per axis, top bit x, then y, then z: the axis cycle of
 * makeOctChildren, GenericTreeNode.h:473-575), binary nodes split on the
 * highest key bit on which their particles differ until they hold at most
 * bucketSize particles (Compute.cpp:2476), numbered breadth first; buckets
 * in particle order. The outputs are the topology changaBuildMoments,
 * changaBuildLists* and changaGravityAdjoint take. Synchronizes the stream
 * (one count per depth is read back).
 * ====================================================================== */
typedef struct {
  int n;                          /* particles */
  const double *pos;              /* device: FP64 x, y, z per particle, input order */
  double boxLo[3], boxSize;       /* the cube the Morton keys cover */
  int bucketSize;                 /* nBucket */
  int *perm;                      /* device out, n: input index of each tree-order particle */
  int *first, *count, *child;     /* device out: node arrays (child: 2 per node, -1 for buckets) */
  double *boxRadius;              /* device out: half diagonal of each node's Morton cell */
  int *bucketNode;                /* device out: bucket nodes in particle order (n entries) */
  int maxNodes;                   /* capacity of the node arrays: at least 2n - 1 */
  int *levelOffsets;              /* host out: depth d holds nodes [levelOffsets[d],
                                     levelOffsets[d + 1]) */
  int maxDepths;                  /* capacity of levelOffsets */
  void *work;                     /* device scratch: changaBuildTreeWorkBytes(n) bytes */
  size_t workBytes;
  int nNodes, nBuckets, nDepths;  /* out */
} ChangaTreeBuild;

size_t changaBuildTreeWorkBytes(int n);
int changaBuildTree(ChangaTreeBuild *t, void *stream);

/* ======================================================================
 * Part 7: fused walk and force
 *
 * The device-resident counterpart of gpuLocalTreeWalk (CudaFunctions.h:10-23,
 * HostCUDA.cu:724-942), which the reference launches from
 * TreePieceCellListDataTransferLocal (HostCUDA.cu:210-224) with the
 * CudaRequest fields of HostCUDA.h:80-89; that request route itself is the
 * GPU_LOCAL_TREE_WALK build of Part 1.
 * One persistent launch in which each wavefront walks the tree for a target
 * bucket (the opening tests of changaBuildLists*, so the same interaction
 * set; the reference kernel walks per particle instead) and evaluates what
 * it accepts at once, with no interaction list in
 * device memory beyond a per-wavefront scratch of maxCells cells and maxRuns
 * particle runs. Precision form only (pos64 in both launches). Results are
 * added into cells.vars like changaNodeGravity + changaPartGravity; they
 * equal the list path's up to the FP32 order of the per-tile sums.
 * errorFlag (host, optional) receives 1, once the stream reaches it, when
 * a bucket's walk exceeded a capacity (hexadecapole cells + 2 x softened
 * cells > maxCells, or runs > maxRuns); the results are then incomplete.
 * ====================================================================== */
typedef struct {
  ChangaGravityLaunch cells;      /* targets, vars, moments, cm64, tpos64, starts, sizes,
                                     numBuckets, fperiod; list and marks unused */
  ChangaGravityLaunch parts;      /* sources, spos64 (same targets and buckets) */
  const int *bucketNode;          /* device: node of each target bucket */
  const int *first, *count, *child;  /* device: topology, as for changaBuildMoments */
  const double *node64;           /* device: node records of changaBuildMoments */
  int numNodes;
  double theta;
  int nReplicas;                  /* 0..3 */
  double period;
  int maxCells, maxRuns;          /* per-wavefront scratch capacities */
  void *work;                     /* device scratch: changaFusedGravityWorkBytes(numNodes,
                                     maxCells, maxRuns) */
  size_t workBytes;
  int *errorFlag;                 /* host out, optional */
} ChangaFusedLaunch;

size_t changaFusedGravityWorkBytes(int numNodes, int maxCells, int maxRuns);
int changaFusedGravity(const ChangaFusedLaunch *f, void *stream);

#ifdef __cplusplus
}
#endif
#endif
