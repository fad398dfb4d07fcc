"""Python view of the host-side library (libchanga_host.so): tree build, node
geometry, interaction lists and the Ewald h-loop table. This is the caller
side of the GPU boundary (the part ChaNGa's TreePiece / ListCompute play).
"""
import ctypes
import os

import numpy as np

from . import build as _build

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p
_lib = None


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_build.HOST_LIB):
        raise RuntimeError("libchanga_host.so is not built: run `python -m changa_amd.build`")
    L = ctypes.CDLL(_build.HOST_LIB)
    L.cgh_tree_build.argtypes = [_i, _p, _p, _d, _i]
    L.cgh_tree_build.restype = _p
    L.cgh_tree_sizes.argtypes = [_p, _p]
    L.cgh_tree_export.argtypes = [_p] + [_p] * 8
    L.cgh_tree_free.argtypes = [_p]
    L.cgh_node_geometry.argtypes = [_i] + [_p] * 8
    L.cgh_walk.argtypes = [_i, _p, _p, _p, _p, _p, _d, _i, _d, _i]
    L.cgh_walk.restype = _p
    L.cgh_lists_sizes.argtypes = [_p, _p]
    L.cgh_lists_export.argtypes = [_p] * 6
    L.cgh_lists_free.argtypes = [_p]
    L.cgh_ewald_table.argtypes = [_p, _d, _d, _d, _p, _i]
    L.cgh_ewald_table.restype = _i
    L.cgh_root_momc.argtypes = [_p, _d, _p]
    _lib = L
    return L


def build_tree(pos, bucket_size, box_lo=None, box_size=None):
    """Morton-ordered binary tree. Returns a dict of topology arrays; `perm`
    maps tree position -> input particle index."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    if box_lo is None:
        lo, hi = pos.min(axis=0), pos.max(axis=0)
        box_size = float((hi - lo).max()) * (1 + 1e-12) + 1e-300
        box_lo = lo
    box_lo = np.ascontiguousarray(box_lo, dtype=np.float64)
    L = lib()
    h = L.cgh_tree_build(n, _ptr(pos), _ptr(box_lo), float(box_size), int(bucket_size))
    sz = np.zeros(3, dtype=np.int32)
    L.cgh_tree_sizes(h, _ptr(sz))
    nn, nb, nd = (int(x) for x in sz)
    t = dict(
        n=n, n_nodes=nn, n_buckets=nb, n_depths=nd,
        perm=np.zeros(n, dtype=np.int32),
        node_first=np.zeros(nn, dtype=np.int32),
        node_count=np.zeros(nn, dtype=np.int32),
        node_child=np.zeros(2 * nn, dtype=np.int32),
        node_depth=np.zeros(nn, dtype=np.int32),
        node_box_radius=np.zeros(nn, dtype=np.float64),
        level_offsets=np.zeros(nd + 1, dtype=np.int32),
        bucket_node=np.zeros(nb, dtype=np.int32),
    )
    L.cgh_tree_export(h, _ptr(t["perm"]), _ptr(t["node_first"]), _ptr(t["node_count"]),
                      _ptr(t["node_child"]), _ptr(t["node_depth"]), _ptr(t["node_box_radius"]),
                      _ptr(t["level_offsets"]), _ptr(t["bucket_node"]))
    L.cgh_tree_free(h)
    t["bucket_start"] = np.ascontiguousarray(t["node_first"][t["bucket_node"]])
    t["bucket_size"] = np.ascontiguousarray(t["node_count"][t["bucket_node"]])
    return t


def node_geometry(tree, pos, mass, soft):
    """FP64 node geometry [n_nodes, 12]: cm.xyz, radius, soft, mass, lo.xyz, hi.xyz.
    pos/mass/soft are in tree order."""
    geo = np.zeros((tree["n_nodes"], 12))
    lib().cgh_node_geometry(tree["n_nodes"], _ptr(tree["node_first"]), _ptr(tree["node_count"]),
                            _ptr(tree["node_child"]), _ptr(tree["node_box_radius"]), _ptr(pos),
                            _ptr(mass), _ptr(soft), _ptr(geo))
    return geo


def walk(tree, geo, theta, n_replicas=0, period=1.0, n_threads=0):
    """Per-bucket interaction lists in ChaNGa's wire layout."""
    L = lib()
    h = L.cgh_walk(tree["n_buckets"], _ptr(tree["bucket_node"]), _ptr(tree["node_count"]),
                   _ptr(tree["node_first"]), _ptr(tree["node_child"]), _ptr(geo),
                   float(theta), int(n_replicas), float(period), int(n_threads))
    if not h:
        raise ValueError("too many buckets for the 22-bit encodeOffset bucket field")
    sz = np.zeros(3, dtype=np.int64)
    L.cgh_lists_sizes(h, _ptr(sz))
    nc, npt, nsrc = (int(x) for x in sz)
    nb = tree["n_buckets"]
    lists = dict(
        cells=np.zeros(2 * nc, dtype=np.int32),
        cell_marks=np.zeros(nb + 1, dtype=np.int64),
        soft_start=np.zeros(nb, dtype=np.int64),
        parts=np.zeros(3 * npt, dtype=np.int32),
        part_marks=np.zeros(nb + 1, dtype=np.int64),
        n_part_sources=nsrc,
    )
    L.cgh_lists_export(h, _ptr(lists["cells"]), _ptr(lists["cell_marks"]),
                       _ptr(lists["soft_start"]), _ptr(lists["parts"]), _ptr(lists["part_marks"]))
    L.cgh_lists_free(h)
    return lists


def interaction_counts(tree, lists, active=None):
    """(N_PC, N_softened, N_PP) as ChaNGa counts them (Compute.cpp:1636-1651):
    sources x active targets of each bucket."""
    nb = tree["n_buckets"]
    if active is None:
        act = tree["bucket_size"].astype(np.int64)
    else:
        starts = tree["bucket_start"]
        cs = np.concatenate([[0], np.cumsum(active.astype(np.int64))])
        act = cs[starts + tree["bucket_size"]] - cs[starts]
    ncell = np.diff(lists["cell_marks"])
    nsoft = lists["cell_marks"][1:] - lists["soft_start"]
    nhex = ncell - nsoft
    p = lists["parts"].reshape(-1, 3)
    per_entry = p[:, 2].astype(np.int64)
    pm = lists["part_marks"]
    cs = np.concatenate([[0], np.cumsum(per_entry)])
    nsrc = cs[pm[1:]] - cs[pm[:-1]]
    assert len(nsrc) == nb
    return int((nhex * act).sum()), int((nsoft * act).sum()), int((nsrc * act).sum())


def ewald_table(q_root, u_root, L, hcut, max_n=8192):
    q = np.ascontiguousarray(q_root, dtype=np.float64)
    ewt = np.zeros((max_n, 5))
    n = lib().cgh_ewald_table(_ptr(q), float(u_root), float(L), float(hcut), _ptr(ewt), max_n)
    if n < 0:
        raise ValueError("Ewald table exceeds max_n")
    return ewt[:n].copy()


def root_momc(q_root, u_root):
    """Complete root moments at unit scale (MOMC order, 32 values)."""
    q = np.ascontiguousarray(q_root, dtype=np.float64)
    mc = np.zeros(32)
    lib().cgh_root_momc(_ptr(q), float(u_root), _ptr(mc))
    return mc
