/*
 * changa_gpu_types.h -- wire types of the GPU gravity boundary.
 *
 * Byte-for-byte the layouts ChaNGa's host code fills and its GPU code reads
 * (HEXADECAPOLE build): cuda_typedef.h:104-272 and EwaldCUDA.h:11-57. With
 * GPU_LOCAL_TREE_WALK defined (ChaNGa's --enable-gpu-local-tree-walk) the
 * node and particle records carry the tree-walk fields and the request its
 * walk parameters, as in that build of ChaNGa. The static asserts pin the
 * sizes of both layouts.
 */
#ifndef CHANGA_GPU_TYPES_H
#define CHANGA_GPU_TYPES_H

#include <stddef.h>
#include <stdint.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef float cudatype;

/* cuda_typedef.h:53-70 */
typedef struct CudaVector3D {
  cudatype x, y, z;
} CudaVector3D;

#ifdef GPU_LOCAL_TREE_WALK
/* cuda_typedef.h:87-99: NodeType of a serialized local-tree node */
enum CudaNodeType {
  CudaInvalid = 1,
  CudaBucket = 2,
  CudaInternal = 3,
  CudaBoundary = 4,
  CudaNonLocal = 5,
  CudaEmpty = 6,
  CudaTop = 7,
  CudaNonLocalBucket = 8,
  CudaCached = 9,
  CudaCachedBucket = 10,
  CudaCachedEmpty = 11
};
#endif

/* cuda_typedef.h:104-131: 27 floats = 108 B (160 B with the tree-walk
   fields). Moments are radius-scaled, reduced and trace-free (component
   order xx, xy, xz, yy, yz, ...). */
typedef struct CudaMultipoleMoments {
  cudatype radius;
  cudatype soft;
  cudatype totalMass;
  CudaVector3D cm;
#ifdef GPU_LOCAL_TREE_WALK
  /* cuda_typedef.h:114-121: bounding box, the node's particles in the
     particle array (bucketStart, bucketSize; INT_MAX and 0 where none are
     local), its particle count, its own index, its two children (-1: none)
     and its CudaNodeType */
  CudaVector3D lesser_corner;
  CudaVector3D greater_corner;
  int bucketStart;
  int bucketSize;
  int particleCount;
  int nodeArrayIndex;
  int children[2];
  int type;
#endif
  cudatype xx, xy, xz, yy, yz;
  cudatype xxx, xyy, xxy, yyy, xxz, yyz, xyz;
  cudatype xxxx, xyyy, xxxy, yyyy, xxxz, yyyz, xxyy, xxyz, xyyz;
} CudaMultipoleMoments;

/* cuda_typedef.h:208-221: a run of `num` source particles starting at
   `index`; `off` carries the periodic replica in bits 22-30. */
typedef struct ILPart {
  int index;
  int off;
  int num;
} ILPart;

/* cuda_typedef.h:225-235: a source cell. offsetID = encodeOffset(bucket, x,
   y, z) (TreePiece.cpp:3631-3645): bits 0-21 bucket, 22-24 / 25-27 / 28-30
   replica x / y / z + 3. Bit 31, unused by the reference, marks a cell whose
   softening sphere overlaps the target bucket (openSoftening,
   gravity.h:251-260); such a cell is evaluated as a spline-softened
   monopole instead of being sent back to the CPU (Compute.cpp:1683-1698). */
typedef struct ILCell {
  int index;
  int offsetID;
} ILCell;

#define CHANGA_ILCELL_SOFTENED ((int)0x80000000u)

/* cuda_typedef.h:240-248: 20 B (24 B with the index of the particle's
   bucket node, cuda_typedef.h:245-247) */
typedef struct CompactPartData {
  cudatype mass;
  cudatype soft;
  CudaVector3D position;
#ifdef GPU_LOCAL_TREE_WALK
  int nodeId;
#endif
} CompactPartData;

/* cuda_typedef.h:268-272: 20 B */
typedef struct VariablePartData {
  CudaVector3D a;
  cudatype potential;
  cudatype dtGrav;
} VariablePartData;

/* EwaldCUDA.h:11-14 */
typedef struct {
  cudatype hx, hy, hz;
  cudatype hCfac, hSfac;
} EwtData;

/* EwaldCUDA.h:18-25 (HEXADECAPOLE: no quadrupole fields) */
typedef struct {
  cudatype totalMass;
  cudatype cmx, cmy, cmz;
} MultipoleMomentsData;

/* EwaldCUDA.h:29-37: complete moments, MOMC order (moments.h:42-50) */
typedef struct {
  cudatype m;
  cudatype xx, yy, xy, xz, yz;
  cudatype xxx, xyy, xxy, yyy, xxz, yyz, xyz;
  cudatype xxxx, xyyy, xxxy, yyyy, xxxz, yyyz, xxyy, xxyz, xyyz;
  cudatype zz;
  cudatype xzz, yzz, zzz;
  cudatype xxzz, xyzz, xzzz, yyzz, yzzz, zzzz;
} MomcData;

/* EwaldCUDA.h:41-47 */
typedef struct {
  MultipoleMomentsData mm;
  MomcData momcRoot;
  int n, nReps, nEwReps, nEwhLoop;
  cudatype L, fEwCut, alpha, alpha2, k1, ka, fEwCut2, fInner2;
} EwaldReadOnlyData;

/* EwaldCUDA.h:50-56 */
typedef struct {
  int EwaldRange[2];
  int *EwaldMarkers;
  EwtData *ewt;
  EwaldReadOnlyData *cachedData;
} EwaldData;

/* HostCUDA.h:31-89; the stream is an opaque hipStream_t. */
typedef struct _CudaRequest {
  void *stream;
  CudaMultipoleMoments *d_localMoments;
  CudaMultipoleMoments *d_remoteMoments;
  CompactPartData *d_localParts;
  CompactPartData *d_remoteParts;
  VariablePartData *d_localVars;
  size_t sMoments;
  size_t sCompactParts;
  size_t sVarParts;
  void *list;
  int *bucketMarkers;
  int *bucketStarts;
  int *bucketSizes;
  int numInteractions;
  int numBucketsPlusOne;
  void *tp;
  void *missedNodes;
  void *missedParts;
  size_t sMissed;
  int *affectedBuckets;
  void *cb;
  void *state;
  cudatype fperiod;
  bool node;
  bool remote;
#ifdef GPU_LOCAL_TREE_WALK
  /* HostCUDA.h:79-88: the local walk of particles [firstParticle,
     lastParticle] from node rootIdx; theta and thetaMono are cosmoType
     (double unless COSMO_FLOAT) */
  int firstParticle;
  int lastParticle;
  int rootIdx;
  double theta;
  double thetaMono;
  int nReplicas;
  cudatype fperiodY;
  cudatype fperiodZ;
#endif
} CudaRequest;

#ifdef __cplusplus
#ifdef GPU_LOCAL_TREE_WALK
static_assert(sizeof(CudaMultipoleMoments) == 160, "CudaMultipoleMoments layout");
static_assert(offsetof(CudaMultipoleMoments, xx) == 76, "CudaMultipoleMoments layout");
static_assert(sizeof(CompactPartData) == 24, "CompactPartData layout");
#else
static_assert(sizeof(CudaMultipoleMoments) == 108, "CudaMultipoleMoments layout");
static_assert(offsetof(CudaMultipoleMoments, xx) == 24, "CudaMultipoleMoments layout");
static_assert(sizeof(CompactPartData) == 20, "CompactPartData layout");
#endif
static_assert(sizeof(VariablePartData) == 20, "VariablePartData layout");
static_assert(sizeof(ILCell) == 8, "ILCell layout");
static_assert(sizeof(ILPart) == 12, "ILPart layout");
static_assert(sizeof(EwtData) == 20, "EwtData layout");
static_assert(sizeof(MomcData) == 128, "MomcData layout");
}
#endif

#endif
