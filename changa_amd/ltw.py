"""Host side of the GPU_LOCAL_TREE_WALK request route (libchanga_gpu_ltw.so).

ChaNGa built with --enable-gpu-local-tree-walk serializes its local tree into
CudaMultipoleMoments records that carry the walk fields (DataManager.cpp:
469-500, addTreeNodeToList :811-822, transformLocalTreeRecursive :912-953),
tags every particle with its bucket node (TreePiece::fillGPUBuffer,
TreePiece.cpp:3864-3892) and sends one request for all of a TreePiece's
particles (ListCompute::sendLocalTreeWalkTriggerToGpu, Compute.cpp:
1437-1530). `serialize` builds those arrays from this package's tree (node
index = tree index, root 0; every node of a single local tree is a bucket or
an internal node) and `walk` runs one request through the C ABI.
"""
import ctypes

import numpy as np

from . import gpu

# CudaMultipoleMoments with the walk fields (cuda_typedef.h:104-131): 160 B
NODE_DTYPE = np.dtype([
    ("radius", "<f4"), ("soft", "<f4"), ("totalMass", "<f4"), ("cm", "<f4", 3),
    ("lesser_corner", "<f4", 3), ("greater_corner", "<f4", 3), ("bucketStart", "<i4"),
    ("bucketSize", "<i4"), ("particleCount", "<i4"), ("nodeArrayIndex", "<i4"),
    ("children", "<i4", 2), ("type", "<i4"), ("moments", "<f4", 21)])
# CompactPartData with the bucket node (cuda_typedef.h:240-248): 24 B
PART_DTYPE = np.dtype([("mass", "<f4"), ("soft", "<f4"), ("position", "<f4", 3),
                       ("nodeId", "<i4")])
assert NODE_DTYPE.itemsize == 160 and PART_DTYPE.itemsize == 24

# CudaNodeType (cuda_typedef.h:87-99)
BUCKET, INTERNAL, BOUNDARY, NONLOCAL, EMPTY, TOP, NONLOCAL_BUCKET = 2, 3, 4, 5, 6, 7, 8


def serialize(tree, wire_moments, node64, pos, mass, soft):
    """(nodes[NODE_DTYPE], parts[PART_DTYPE]) of a tree: the wire moments
    (radius, soft, mass, cm, 21 moments; float32 [n_nodes, 27]), the FP64
    node records for the bounding boxes (lo at 6-8, hi at 9-11), and the
    particles in tree order."""
    nn = int(tree["n_nodes"])
    child = np.asarray(tree["node_child"], dtype=np.int32).reshape(nn, 2)
    first = np.asarray(tree["node_first"], dtype=np.int32)
    count = np.asarray(tree["node_count"], dtype=np.int32)
    w = np.asarray(wire_moments, dtype=np.float32).reshape(nn, 27)
    g = np.asarray(node64, dtype=np.float64)
    nodes = np.zeros(nn, dtype=NODE_DTYPE)
    nodes["radius"], nodes["soft"], nodes["totalMass"] = w[:, 0], w[:, 1], w[:, 2]
    nodes["cm"] = w[:, 3:6]
    nodes["moments"] = w[:, 6:27]
    # OrientedBox<double> corners stored into CudaVector3D (float)
    nodes["lesser_corner"] = g[:, 6:9].astype(np.float32)
    nodes["greater_corner"] = g[:, 9:12].astype(np.float32)
    leaf = child[:, 0] < 0
    nodes["type"] = np.where(leaf, BUCKET, INTERNAL)
    nodes["children"] = np.where(leaf[:, None], -1, child)
    nodes["nodeArrayIndex"] = np.arange(nn, dtype=np.int32)
    nodes["particleCount"] = count
    # buckets: their particles; internal nodes: the union of their
    # children's (contiguous in tree order)
    nodes["bucketStart"] = first
    nodes["bucketSize"] = count
    n = pos.shape[0]
    parts = np.zeros(n, dtype=PART_DTYPE)
    parts["mass"], parts["soft"] = mass, soft
    parts["position"] = np.asarray(pos, dtype=np.float32)
    bn = np.asarray(tree["bucket_node"])
    parts["nodeId"] = np.repeat(bn, count[bn])
    return nodes, parts


def serialize_evaluation(ev):
    """serialize() of an engine.Evaluation's tree, moments and particles."""
    return serialize(ev.tree, ev.d_moments.cpu().numpy(), ev.node64(), ev.pos, ev.mass, ev.soft)


class LocalWalk:
    """A TreePieceCellListDataTransferLocal request of the local tree walk on
    device copies of `nodes` and `parts` that stay resident, for repeated
    runs; `period` is one replica period or three (fperiod, fperiodY,
    fperiodZ); theta_mono defaults to theta^4 (TreePiece.cpp:5036)."""

    def __init__(self, nodes, parts, theta, n_replicas=0, period=0.0, first=0, last=None,
                 root=0, device="cuda:0", theta_mono=None):
        import torch
        self.L = gpu.lib_ltw()
        self.n = len(parts)
        dev = torch.device(device)
        self.d_nodes = torch.from_numpy(nodes.view(np.uint8).copy()).to(dev)
        self.d_parts = torch.from_numpy(parts.view(np.uint8).copy()).to(dev)
        self.d_vars = torch.zeros((self.n, 5), dtype=torch.float32, device=dev)
        self.device = dev
        self.stream = torch.cuda.current_stream(dev)
        r = gpu.CudaRequestLTW()
        r.stream = self.stream.cuda_stream
        r.d_localMoments = self.d_nodes.data_ptr()
        r.d_localParts = self.d_parts.data_ptr()
        r.d_localVars = self.d_vars.data_ptr()
        r.sMoments = nodes.nbytes
        r.sCompactParts = parts.nbytes
        r.sVarParts = self.n * 20
        r.numBucketsPlusOne = 1
        r.firstParticle, r.rootIdx = first, root
        r.lastParticle = self.n - 1 if last is None else last
        r.theta = theta
        r.thetaMono = theta * theta * theta * theta if theta_mono is None else theta_mono
        r.nReplicas = n_replicas
        px, py, pz = (float(x) for x in np.broadcast_to(np.asarray(period, dtype=float), (3,)))
        r.fperiod, r.fperiodY, r.fperiodZ = px, py, pz
        r.node, r.remote = True, False
        self.request = r

    def run(self, zero=True):
        """Enqueue one request (after zeroing the results unless zero=False)."""
        if zero:
            rc = self.L.changaZeroVars(ctypes.c_void_p(self.d_vars.data_ptr()), self.n,
                                       ctypes.c_void_p(self.stream.cuda_stream))
            if rc:
                gpu.check_ltw("changaZeroVars")
        self.L.TreePieceCellListDataTransferLocal(ctypes.byref(self.request))

    def results(self):
        """Synchronize, raise on a recorded error, and return the
        VariablePartData of every particle (float32 [n, 5])."""
        import torch
        torch.cuda.synchronize(self.device)
        gpu.check_ltw("TreePieceCellListDataTransferLocal")
        return self.d_vars.cpu().numpy()


def walk(nodes, parts, theta, n_replicas=0, period=0.0, first=0, last=None, root=0,
         device="cuda:0", theta_mono=None):
    """One request of the local tree walk on fresh device copies; returns
    the VariablePartData of every particle (a[3], potential, dtGrav)."""
    w = LocalWalk(nodes, parts, theta, n_replicas, period, first, last, root, device, theta_mono)
    w.run()
    return w.results()
