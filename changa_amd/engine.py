"""One gravity force evaluation on cuda:0 through the C-ABI.

The sequence mirrors ChaNGa's GPU path (TreePiece::startGravity ->
DataManager::transferLocalToGPU -> TreePiece{Cell,Part}ListDataTransferLocal
-> EwaldHost -> TransferParticleVarsBack): the tree and the interaction
lists come from the host (libchanga_host.so) or the GPU (changaBuildTree,
changaBuildLists*); the GPU builds the multipole moments, then evaluates PC,
PP and Ewald into the VariablePartData array (libchanga_gpu.so), from the
lists or with the walk fused into the force launch (changaFusedGravity).
PyTorch only provides device memory and the stream.
"""
import ctypes
import functools
import math

import numpy as np
import torch

from . import gpu, host


def _dp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def compact_parts(pos, mass, soft):
    """CompactPartData[N] {mass, soft, x, y, z} (cuda_typedef.h:240-248)."""
    a = np.empty((pos.shape[0], 5), dtype=np.float32)
    a[:, 0] = mass
    a[:, 1] = soft
    a[:, 2:] = pos
    return a


def expand_parts(lists):
    """ILPart runs -> one ILCell per source particle, as ChaNGa's
    GenericList<ILPart>::serialize sends them (Compute.cpp:1118-1218)."""
    p = lists["parts"].reshape(-1, 3)
    num = p[:, 2].astype(np.int64)
    total = int(num.sum())
    run = np.repeat(np.arange(len(p)), num)
    start = np.concatenate([[0], np.cumsum(num)[:-1]])
    k = np.arange(total) - start[run]
    cells = np.empty((total, 2), dtype=np.int32)
    cells[:, 0] = p[run, 0] + k
    cells[:, 1] = p[run, 1]
    cs = np.concatenate([[0], np.cumsum(num)])
    marks = cs[lists["part_marks"]]
    return cells.reshape(-1), marks


def ewald_real_counts(pos, cm, L, n_reps, ew_cut, chunk=1 << 18):
    """Evaluated real-space Ewald terms per particle (BucketEwald's test,
    Ewald.cpp:126-138): |d + nL|^2 <= (fEwCut L)^2 or n in the hole."""
    R = max(int(math.ceil(ew_cut)), n_reps)
    g = np.arange(-R, R + 1)
    nv = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    hole = np.all(np.abs(nv) <= n_reps, axis=1)
    cut2 = (ew_cut * L) ** 2
    out = np.zeros(pos.shape[0], dtype=np.int64)
    for i in range(0, pos.shape[0], chunk):
        d = pos[i:i + chunk] - cm
        r2 = ((d[:, None, :] + nv[None, :, :] * L) ** 2).sum(axis=2)
        out[i:i + chunk] = ((r2 <= cut2) | hole[None, :]).sum(axis=1)
    return out


def slice_lists(lists, buckets):
    """The interaction lists of a subset of buckets (in the given order),
    with re-based markers."""
    out = {}
    buckets = np.asarray(buckets, dtype=np.int64)
    for key, mkey, width in (("cells", "cell_marks", 2), ("parts", "part_marks", 3)):
        lst, marks = lists[key].reshape(-1, width), np.asarray(lists[mkey], dtype=np.int64)
        lens = marks[buckets + 1] - marks[buckets]
        new_marks = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        idx = np.repeat(marks[buckets] - new_marks[:-1], lens) + np.arange(new_marks[-1])
        out[key] = np.ascontiguousarray(lst[idx].reshape(-1))
        out[mkey] = new_marks
    return out


def gpu_build_tree(pos, bucket_size, box_lo=None, box_size=None, device="cuda:0",
                   keep_device=False):
    """The tree host.build_tree makes, built on the GPU (changaBuildTree);
    returns the same dict of host arrays (with keep_device, also the device
    arrays and the uploaded positions under "_dev")."""
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = pos.shape[0]
    dev = torch.device(device)
    i32 = dict(dtype=torch.int32, device=dev)
    d_pos = torch.from_numpy(pos).to(dev)
    if box_lo is None:
        # the host builder's box; min and max are exact, so the device gives the same values
        lo, hi = d_pos.amin(dim=0).cpu().numpy(), d_pos.amax(dim=0).cpu().numpy()
        box_size = float((hi - lo).max()) * (1 + 1e-12) + 1e-300
        box_lo = lo
    nn_cap = max(2 * n - 1, 1)
    d_perm = torch.empty(n, **i32)
    d_first, d_count = torch.empty(nn_cap, **i32), torch.empty(nn_cap, **i32)
    d_child = torch.empty(2 * nn_cap, **i32)
    d_boxr = torch.empty(nn_cap, dtype=torch.float64, device=dev)
    d_bnode = torch.empty(n, **i32)
    nbytes = gpu.lib().changaBuildTreeWorkBytes(n)
    if nbytes == 0:
        raise RuntimeError("changaBuildTreeWorkBytes failed")
    work = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    levels = np.zeros(130, dtype=np.int32)
    t = gpu.ChangaTreeBuild()
    t.n = n
    t.pos = d_pos.data_ptr()
    for k in range(3):
        t.boxLo[k] = float(box_lo[k])
    t.boxSize = float(box_size)
    t.bucketSize = int(bucket_size)
    t.perm, t.first, t.count = d_perm.data_ptr(), d_first.data_ptr(), d_count.data_ptr()
    t.child, t.boxRadius, t.bucketNode = d_child.data_ptr(), d_boxr.data_ptr(), d_bnode.data_ptr()
    t.maxNodes = nn_cap
    t.levelOffsets = levels.ctypes.data
    t.maxDepths = len(levels)
    t.work, t.workBytes = work.data_ptr(), work.numel()
    s = torch.cuda.current_stream(dev)
    gpu.check(gpu.lib().changaBuildTree(ctypes.byref(t), ctypes.c_void_p(s.cuda_stream)),
              "changaBuildTree")
    nn, nb, nd = t.nNodes, t.nBuckets, t.nDepths
    lo_off = levels[:nd + 1].copy()
    depth = np.repeat(np.arange(nd, dtype=np.int32), np.diff(lo_off))
    tr = dict(n=n, n_nodes=nn, n_buckets=nb, n_depths=nd,
              perm=d_perm.cpu().numpy(), node_first=d_first[:nn].cpu().numpy(),
              node_count=d_count[:nn].cpu().numpy(), node_child=d_child[:2 * nn].cpu().numpy(),
              node_depth=depth, node_box_radius=d_boxr[:nn].cpu().numpy(),
              level_offsets=lo_off, bucket_node=d_bnode[:nb].cpu().numpy())
    tr["bucket_start"] = np.ascontiguousarray(tr["node_first"][tr["bucket_node"]])
    tr["bucket_size"] = np.ascontiguousarray(tr["node_count"][tr["bucket_node"]])
    if keep_device:
        # the device copies, for an evaluation that gathers on the device
        tr["_dev"] = dict(pos=d_pos, perm=d_perm, first=d_first[:nn], count=d_count[:nn],
                          child=d_child[:2 * nn], box_radius=d_boxr[:nn], bucket_node=d_bnode[:nb])
    return tr


class Evaluation:
    """Tree, lists and device state for one workload (see configs.py).

    precise: FP64 bucket-centre subtraction form; pp_compact: ILPart lists
    expanded on the GPU (else ChaNGa's per-particle ILCell lists).

    Targets: every particle, or, when the workload carries an `active` mask
    (a rung), only the active ones: they are gathered bucket by bucket into
    their own target arrays (CompactPartData, FP64 positions, results), and
    only buckets holding an active particle keep their lists. Sources are
    always every particle."""

    def __init__(self, w, precise=False, pp_compact=True, n_threads=0, device="cuda:0",
                 lists="host", tree="host", cell_records=True):
        if lists not in ("host", "gpu", "none") or tree not in ("host", "gpu"):
            raise ValueError("lists: 'host', 'gpu' or 'none' (fused form only); tree: 'host' or "
                             "'gpu'")
        if lists == "gpu" and not pp_compact:
            raise ValueError("GPU-built lists are ILPart runs (pp_compact=True)")
        self.w = w
        self.precise = precise
        self.pp_compact = pp_compact
        # 128-byte cell records (changaCellRecords) for the particle-cell
        # launches, rebuilt with every moment build; False: wire moments
        self.cell_records = bool(cell_records)
        self.list_source = lists
        self.device = torch.device(device)
        periodic = bool(w["periodic"])
        self.L = float(w["L"]) if periodic else 0.0
        build = host.build_tree if tree == "host" else functools.partial(
            gpu_build_tree, device=device, keep_device=True)
        if periodic:
            tr = build(w["pos"], w["bucket"], box_lo=np.zeros(3), box_size=self.L)
        else:
            tr = build(w["pos"], w["bucket"])
        self._dev = tr.pop("_dev", None)
        self.tree = tr
        self._host = {}
        if self._dev is None:
            perm = tr["perm"]
            self._host = dict(pos=np.ascontiguousarray(w["pos"][perm]),
                              mass=np.ascontiguousarray(w["mass"][perm]),
                              soft=np.ascontiguousarray(w["soft"][perm]),
                              active=None if w.get("active") is None
                              else np.ascontiguousarray(w["active"][perm]))
        else:
            # tree order gathered on the device from the device permutation;
            # host copies are made on first use
            d_perm = self._dev["perm"].long()
            self._dev["pos"] = self._dev["pos"][d_perm].contiguous()
            for key in ("mass", "soft"):
                self._dev[key] = torch.from_numpy(np.ascontiguousarray(
                    w[key], dtype=np.float64)).to(self.device)[d_perm].contiguous()
            self._dev["active"] = None if w.get("active") is None else torch.from_numpy(
                np.ascontiguousarray(w["active"], dtype=np.bool_)).to(self.device)[d_perm]
        self._lists = None
        if lists == "host":
            self.geo = host.node_geometry(tr, self.pos, self.mass, self.soft)
            self._lists = host.walk(tr, self.geo, w["theta"], w["n_replicas"],
                                    self.L if periodic else 1.0, n_threads)
            self.counts = host.interaction_counts(tr, self._lists, self.active)
        self._targets()
        self._upload()

    def _host_array(self, key):
        if key not in self._host:
            d = self._dev[key]
            self._host[key] = None if d is None else d.cpu().numpy()
        return self._host[key]

    @property
    def pos(self):
        """Positions in tree order (host FP64)."""
        return self._host_array("pos")

    @property
    def mass(self):
        return self._host_array("mass")

    @property
    def soft(self):
        return self._host_array("soft")

    @property
    def active(self):
        """Rung mask in tree order, or None when every particle is a target."""
        return self._host_array("active")

    @property
    def lists(self):
        """Interaction lists of every bucket of the tree (host arrays, the
        host walk's layout); with GPU-built lists, downloaded on first use
        (buckets without targets have empty lists)."""
        if self._lists is None:
            if self.list_source == "none":
                raise ValueError("this evaluation has no interaction lists (lists='none')")
            self._lists = self._download_lists()
        return self._lists

    def _targets(self):
        tr = self.tree
        self.n = int(tr["n"])
        if self._dev is not None and self._dev["active"] is not None:
            # the rung's targets, bucket by bucket, selected on the device
            nb = tr["n_buckets"]
            d_size = torch.from_numpy(tr["bucket_size"].astype(np.int64)).to(self.device)
            owner = torch.repeat_interleave(torch.arange(nb, device=self.device), d_size)
            tidx = torch.nonzero(self._dev["active"]).flatten()
            counts = torch.bincount(owner[tidx], minlength=nb)
            keep = torch.nonzero(counts > 0).flatten()
            sizes = counts[keep].to(torch.int32)
            self._dev["tindex"] = tidx.to(torch.int32)
            self.t_index = self._dev["tindex"].cpu().numpy()
            self.n_targets = len(self.t_index)
            self.t_sizes = np.ascontiguousarray(sizes.cpu().numpy())
            self.t_starts = np.ascontiguousarray(
                (np.cumsum(self.t_sizes) - self.t_sizes).astype(np.int32))
            self.t_keep = keep.cpu().numpy()
            return
        if self.active is None:
            self.t_index = None
            self.n_targets = self.n
            self.t_starts, self.t_sizes = tr["bucket_start"], tr["bucket_size"]
            self.t_keep = np.arange(tr["n_buckets"])
            if self._lists is not None:
                self.t_lists = self._lists
            return
        nb = tr["n_buckets"]
        owner = np.repeat(np.arange(nb), tr["bucket_size"])
        tidx = np.nonzero(self.active)[0]
        counts = np.bincount(owner[tidx], minlength=nb)
        keep = np.nonzero(counts > 0)[0]
        self.t_index = tidx.astype(np.int32)
        self.n_targets = len(tidx)
        self.t_sizes = np.ascontiguousarray(counts[keep].astype(np.int32))
        self.t_starts = np.ascontiguousarray(
            (np.cumsum(self.t_sizes) - self.t_sizes).astype(np.int32))
        self.t_keep = keep
        if self._lists is not None:
            self.t_lists = slice_lists(self._lists, keep)

    # ------------------------------------------------------------------ device
    def _t(self, a, dtype=None):
        return torch.from_numpy(np.ascontiguousarray(a)).to(self.device)

    def _upload(self):
        tr = self.tree
        dv = self._dev
        if dv is None:
            self.d_parts = self._t(compact_parts(self.pos, self.mass, self.soft))
            self.d_pos64 = self._t(self.pos)
            self.d_mass64 = self._t(self.mass)
            self.d_soft64 = self._t(self.soft)
        else:
            self.d_pos64, self.d_mass64, self.d_soft64 = dv["pos"], dv["mass"], dv["soft"]
            # CompactPartData {mass, soft, x, y, z}, each FP64 value rounded to FP32
            self.d_parts = torch.cat([self.d_mass64[:, None], self.d_soft64[:, None],
                                      self.d_pos64], dim=1).to(torch.float32).contiguous()
        if self.t_index is None:
            self.d_targets, self.d_tpos64, self.d_tindex = self.d_parts, self.d_pos64, None
        elif dv is not None:
            self.d_tindex = dv["tindex"]
            ti = self.d_tindex.long()
            self.d_targets = self.d_parts[ti].contiguous()
            self.d_tpos64 = self.d_pos64[ti].contiguous()
        else:
            ti = self.t_index
            self.d_targets = self._t(compact_parts(self.pos[ti], self.mass[ti], self.soft[ti]))
            self.d_tpos64 = self._t(self.pos[ti])
            self.d_tindex = self._t(ti)
        self.d_vars = torch.zeros((self.n_targets, 5), dtype=torch.float32, device=self.device)
        if dv is None:
            self.d_first = self._t(tr["node_first"])
            self.d_count = self._t(tr["node_count"])
            self.d_child = self._t(tr["node_child"])
            self.d_boxr = self._t(tr["node_box_radius"])
        else:
            self.d_first, self.d_count = dv["first"], dv["count"]
            self.d_child, self.d_boxr = dv["child"], dv["box_radius"]
        nn = tr["n_nodes"]
        self.d_moments = torch.zeros((nn, 27), dtype=torch.float32, device=self.device)
        self.d_node64 = torch.zeros((nn, 34), dtype=torch.float64, device=self.device)
        self.d_cm64 = torch.zeros((nn, 3), dtype=torch.float64, device=self.device)
        # (the loops address at most gpu.CELL_RECORD_MAX records; larger
        # trees use the wire moments)
        self.d_cellrec = (torch.zeros((nn, 32), dtype=torch.float32, device=self.device)
                          if self.cell_records and nn <= gpu.CELL_RECORD_MAX else None)
        self.d_starts = self._t(self.t_starts)
        self.d_sizes = self._t(self.t_sizes)
        self.stream = torch.cuda.current_stream(self.device)
        if self.list_source == "gpu":
            self.build_moments()
            self._gpu_lists()
        elif self.list_source == "none":
            # no interaction lists: the fused walk-and-force form only
            self.build_moments()
            self.counts = None
            self.t_lists = None
        else:
            self._upload_host_lists()
            self.build_moments()
        torch.cuda.synchronize(self.device)
        if self.w["periodic"]:
            self._setup_ewald()

    def move(self, pos):
        """New particle positions (tree order, [n, 3] FP64) on the same tree
        topology, interaction lists and device buffers (every address a plan
        or launch record holds stays valid). The next step's multipole build,
        Ewald root record and forces use them; the lists are not rebuilt
        (build_lists does that for GPU-built lists)."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        if pos.shape != (self.n, 3):
            raise ValueError(f"pos must be [{self.n}, 3] in tree order")
        d = torch.from_numpy(pos).to(self.device)
        self.d_pos64.copy_(d)
        self.d_parts[:, 2:5].copy_(d.to(torch.float32))  # CompactPartData {mass, soft, x, y, z}
        if self.t_index is not None:
            ti = self.d_tindex.long()
            self.d_tpos64.copy_(self.d_pos64[ti])
            self.d_targets.copy_(self.d_parts[ti])
        if self._dev is not None:
            self._dev["pos"] = self.d_pos64
        self._host["pos"] = pos
        self._host_root_stale = True
        self.ewald_count_real = None

    def _refresh_host_root(self):
        """The host copies of the Ewald root (centre of mass, mass, MOMC) from
        the device root record, after moves (the adjoint's Ewald term reads
        them)."""
        if not (self.w["periodic"] and getattr(self, "_host_root_stale", False)):
            return
        r = self.d_ewald_root.cpu().numpy()
        self.root_cm64 = np.ascontiguousarray(r[0:3])
        ro = self.ewald_ro
        ro.mm.totalMass = r[3]
        ro.mm.cmx, ro.mm.cmy, ro.mm.cmz = r[0:3]
        q = r[4:26]
        mc = host.root_momc(q, 1.0)
        for f, v in zip(gpu.MOMC_FIELDS, mc):
            setattr(ro.momcRoot, f, float(v))
        self._host_root_stale = False

    def _upload_host_lists(self):
        L = self.t_lists
        if L["cell_marks"][-1] >= 2 ** 31:
            raise ValueError("cell list exceeds the int32 marker range of one request")
        self.d_cells = self._t(L["cells"])
        self.d_cmarks = self._t(L["cell_marks"].astype(np.int32))
        if self.pp_compact:
            self.d_plist = self._t(L["parts"])
            self.d_pmarks = self._t(L["part_marks"].astype(np.int32))
        else:
            cells, marks = expand_parts(L)
            if marks[-1] >= 2 ** 31:
                raise ValueError("expanded particle list exceeds the int32 marker range")
            self.d_plist = self._t(cells)
            self.d_pmarks = self._t(marks.astype(np.int32))

    def _bucket_nodes(self):
        """Device int32: the node of each target bucket."""
        if self._dev is not None:
            bn = self._dev["bucket_node"]
            if len(self.t_keep) == self.tree["n_buckets"]:
                return bn
            return bn[torch.from_numpy(self.t_keep).to(self.device)].contiguous()
        return self._t(self.tree["bucket_node"][self.t_keep].astype(np.int32))

    def build_lists(self):
        """Rebuild the GPU-built lists from the current node records (after
        build_moments, e.g. for new positions on the same tree topology).

        The new lists live in new device arrays, so everything that holds the
        old arrays' addresses is dropped: the batched plan (its graph keeps
        the launch records by value), the fused launch record and the
        adjoint's list work buffer (sized for the old list lengths)."""
        if self.list_source != "gpu":
            raise ValueError("the lists of this evaluation come from the host walk")
        self.free_batch()
        self._fused = None
        self.d_list_work = None
        self._lists = None
        self._gpu_lists()

    def _gpu_lists(self):
        """The target buckets' lists built on the GPU from the node records
        of the moment build (changaBuildListsCount / Fill); the interaction
        counts come from its per-bucket sizes."""
        tr, w = self.tree, self.w
        keep = self.t_keep
        nb = len(keep)
        i32 = dict(dtype=torch.int32, device=self.device)
        self.d_bucket_node = self._bucket_nodes()
        self.d_cmarks = torch.empty(nb + 1, **i32)
        self.d_pmarks = torch.empty(nb + 1, **i32)
        self.d_soft_start = torch.empty(max(nb, 1), **i32)
        d_src = torch.empty(max(nb, 1), **i32)
        nbytes = gpu.lib().changaBuildListsWorkBytes(nb, tr["n_nodes"])
        if nbytes == 0:
            raise RuntimeError("changaBuildListsWorkBytes failed")
        work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        b = gpu.ChangaListBuild()
        b.nBuckets = nb
        b.numNodes = tr["n_nodes"]
        b.bucketNode = self.d_bucket_node.data_ptr()
        b.first = self.d_first.data_ptr()
        b.count = self.d_count.data_ptr()
        b.child = self.d_child.data_ptr()
        b.node64 = self.d_node64.data_ptr()
        b.theta = float(w["theta"])
        b.nReplicas = int(w["n_replicas"]) if w["periodic"] else 0
        b.period = self.L if w["periodic"] else 0.0
        b.cellMarks = self.d_cmarks.data_ptr()
        b.partMarks = self.d_pmarks.data_ptr()
        b.softStart = self.d_soft_start.data_ptr()
        b.sourceCounts = d_src.data_ptr()
        b.work = work.data_ptr()
        b.workBytes = work.numel()
        nc, npr = ctypes.c_longlong(0), ctypes.c_longlong(0)
        gpu.check(gpu.lib().changaBuildListsCount(ctypes.byref(b), ctypes.byref(nc),
                                                  ctypes.byref(npr), self._sp()),
                  "changaBuildListsCount")
        self.d_cells = torch.empty((max(nc.value, 1), 2), **i32)
        self.d_plist = torch.empty((max(npr.value, 1), 3), **i32)
        b.cells = self.d_cells.data_ptr()
        b.parts = self.d_plist.data_ptr()
        gpu.check(gpu.lib().changaBuildListsFill(ctypes.byref(b), self._sp()),
                  "changaBuildListsFill")
        cm = self.d_cmarks.cpu().numpy().astype(np.int64)
        pm = self.d_pmarks.cpu().numpy().astype(np.int64)
        ss = self.d_soft_start.cpu().numpy()[:nb].astype(np.int64)
        src = d_src.cpu().numpy()[:nb].astype(np.int64)
        self.t_lists = dict(cell_marks=cm, part_marks=pm, soft_start=ss)
        act = self.t_sizes.astype(np.int64)
        nhex, nsoft = ss - cm[:-1], cm[1:] - ss
        self.counts = (int((nhex * act).sum()), int((nsoft * act).sum()), int((src * act).sum()))

    def _download_lists(self):
        """GPU-built lists as host arrays over every bucket of the tree."""
        tr = self.tree
        nbt, keep = tr["n_buckets"], self.t_keep
        L = self.t_lists
        cells = self.d_cells.cpu().numpy().reshape(-1)[:2 * int(L["cell_marks"][-1])]
        parts = self.d_plist.cpu().numpy().reshape(-1)[:3 * int(L["part_marks"][-1])]
        out = dict(cells=np.ascontiguousarray(cells), parts=np.ascontiguousarray(parts))
        for key in ("cell_marks", "part_marks"):
            lens = np.zeros(nbt, dtype=np.int64)
            lens[keep] = np.diff(L[key])
            out[key] = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
        ss = out["cell_marks"][:-1].copy()
        ss[keep] += L["soft_start"] - L["cell_marks"][:-1]
        out["soft_start"] = ss
        p = parts.reshape(-1, 3)
        out["n_part_sources"] = int(p[:, 2].astype(np.int64).sum())
        return out

    def _sp(self):
        return ctypes.c_void_p(self.stream.cuda_stream)

    def build_moments(self):
        tr = self.tree
        lo = np.ascontiguousarray(tr["level_offsets"], dtype=np.int32)
        rc = gpu.lib().changaBuildMoments(
            tr["n_nodes"], tr["n_depths"], lo.ctypes.data_as(ctypes.c_void_p), _dp(self.d_first),
            _dp(self.d_count), _dp(self.d_child), _dp(self.d_boxr), _dp(self.d_pos64),
            _dp(self.d_mass64), _dp(self.d_soft64), _dp(self.d_moments), _dp(self.d_node64),
            _dp(self.d_cm64), self._sp())
        gpu.check(rc, "changaBuildMoments")
        if self.d_cellrec is not None:
            gpu.check(gpu.lib().changaCellRecords(_dp(self.d_moments), _dp(self.d_cm64),
                                                  tr["n_nodes"], _dp(self.d_cellrec),
                                                  self._sp()), "changaCellRecords")

    def _setup_ewald(self):
        w = self.w
        root = self.d_node64[0].cpu().numpy()
        cm, u, M, q = root[0:3], root[3], root[5], root[12:34]
        self.root_cm64 = np.ascontiguousarray(cm)
        self.ewt64 = host.ewald_table(q, u, self.L, w["ewh_cut"])
        mc = host.root_momc(q, u)
        ro = gpu.EwaldReadOnlyData()
        ro.mm.totalMass = M
        ro.mm.cmx, ro.mm.cmy, ro.mm.cmz = cm
        for f, v in zip(gpu.MOMC_FIELDS, mc):
            setattr(ro.momcRoot, f, float(v))
        L = self.L
        alpha = 2.0 / L
        self.ew_markers = None
        ro.n = self.n_targets
        ro.nReps = w["n_replicas"]
        ro.nEwReps = int(math.ceil(w["ew_cut"]))
        ro.nEwhLoop = len(self.ewt64)
        ro.L = L
        ro.fEwCut = w["ew_cut"]
        ro.alpha = alpha
        ro.alpha2 = alpha * alpha
        ro.k1 = math.pi / (alpha * alpha * L ** 3)
        ro.ka = 2.0 * alpha / math.sqrt(math.pi)
        ro.fEwCut2 = w["ew_cut"] ** 2 * L * L
        ro.fInner2 = 1.1e-2 * L * L
        self.ewald_ro = ro
        self.d_ewt = self._t(self.ewt64.astype(np.float32))
        self.d_ewald_root = torch.empty(gpu.EWALD_ROOT_DOUBLES, dtype=torch.float64,
                                        device=self.device)
        self.ewald_root()
        self.ewald_count_real = None

    def ewald_root(self):
        """The Ewald root record and h-loop table rebuilt on the device from
        the last moment build's root node (changaEwaldRoot): the Ewald pass of
        a step on moved particles uses the step's own root multipole."""
        if not self.w["periodic"]:
            return
        gpu.check(gpu.lib().changaEwaldRoot(_dp(self.d_node64), ctypes.byref(self.ewald_ro),
                                            float(self.w["ewh_cut"]), _dp(self.d_ewald_root),
                                            _dp(self.d_ewt), self._sp()), "changaEwaldRoot")

    def ewald_counts(self):
        if not self.w["periodic"]:
            return 0, 0
        if self.ewald_count_real is None:
            sel = slice(None) if self.active is None else self.active
            c = ewald_real_counts(self.pos[sel], self.root_cm64, self.L, self.w["n_replicas"],
                                  self.w["ew_cut"])
            self.ewald_count_real = int(c.sum())
        nact = self.n if self.active is None else int(self.active.sum())
        return self.ewald_count_real, nact * len(self.ewt64)

    # ------------------------------------------------------------------ launches
    def _launch(self, pc):
        a = gpu.ChangaGravityLaunch()
        a.targets = self.d_targets.data_ptr()
        a.vars = self.d_vars.data_ptr()
        a.moments = self.d_moments.data_ptr()
        a.sources = self.d_parts.data_ptr()
        if self.list_source != "none":
            a.list = (self.d_cells if pc else self.d_plist).data_ptr()
            a.marks = (self.d_cmarks if pc else self.d_pmarks).data_ptr()
        a.starts = self.d_starts.data_ptr()
        a.sizes = self.d_sizes.data_ptr()
        a.numBuckets = len(self.t_sizes)
        a.listKind = 0 if pc else (1 if self.pp_compact else 0)
        a.fperiod = self.L
        if self.precise:
            a.tpos64 = self.d_tpos64.data_ptr()
            a.spos64 = self.d_pos64.data_ptr()
            a.cm64 = self.d_cm64.data_ptr()
        if pc and self.d_cellrec is not None:
            a.cellRecords = self.d_cellrec.data_ptr()
        return a

    def zero(self):
        gpu.check(gpu.lib().changaZeroVars(_dp(self.d_vars), self.n_targets, self._sp()),
                  "changaZeroVars")

    def _need_lists(self):
        if self.list_source == "none":
            raise ValueError("this evaluation has no interaction lists (lists='none'): "
                             "use the fused form")

    def run_pc(self):
        self._need_lists()
        a = self._launch(True)
        gpu.check(gpu.lib().changaNodeGravity(ctypes.byref(a), self._sp()), "changaNodeGravity")

    def run_pp(self):
        self._need_lists()
        a = self._launch(False)
        gpu.check(gpu.lib().changaPartGravity(ctypes.byref(a), self._sp()), "changaPartGravity")

    def run_ewald(self, vars_=None, stream=None, host_record=False):
        """The Ewald correction added into vars_ (default: the results) on
        stream (default: the evaluation's), with the device root record of
        the last moment build; host_record: with the host EwaldReadOnlyData
        record instead (changaEwald, the record EwaldHost takes)."""
        if not self.w["periodic"]:
            return
        pos64 = _dp(self.d_tpos64) if self.precise else None
        sp = self._sp() if stream is None else ctypes.c_void_p(stream.cuda_stream)
        if host_record:
            self._refresh_host_root()
            cm = self.root_cm64.ctypes.data_as(ctypes.c_void_p) if self.precise else None
            rc = gpu.lib().changaEwald(
                _dp(self.d_targets), _dp(self.d_vars if vars_ is None else vars_),
                _dp(self.ew_markers) if self.ew_markers is not None else None, 0,
                self.ewald_ro.n, ctypes.byref(self.ewald_ro), _dp(self.d_ewt), pos64, cm, sp)
            gpu.check(rc, "changaEwald")
            return
        rc = gpu.lib().changaEwaldResident(
            _dp(self.d_targets), _dp(self.d_vars if vars_ is None else vars_),
            _dp(self.ew_markers) if self.ew_markers is not None else None, 0, self.ewald_ro.n,
            ctypes.byref(self.ewald_ro), _dp(self.d_ewt), pos64, _dp(self.d_ewald_root), sp)
        gpu.check(rc, "changaEwaldResident")

    def _ewald_begin(self):
        """Start the Ewald correction on a second stream into its own result
        array, so that it overlaps the tree forces; _ewald_end merges it. It
        waits for the work enqueued so far on the evaluation's stream (the
        moment build and the device root record it reads)."""
        if getattr(self, "_ew_stream", None) is None:
            self._ew_stream = torch.cuda.Stream(self.device)  # default priority
            self._ew_vars = torch.empty_like(self.d_vars)
            self._ew_go = torch.cuda.Event()
            self._ew_done = torch.cuda.Event()
        # the previous merge must have read the array before it is cleared
        self._ew_go.record(self.stream)
        self._ew_stream.wait_event(self._ew_go)
        sp = ctypes.c_void_p(self._ew_stream.cuda_stream)
        gpu.check(gpu.lib().changaZeroVars(_dp(self._ew_vars), self.n_targets, sp),
                  "changaZeroVars")
        self.run_ewald(self._ew_vars, self._ew_stream)
        self._ew_done.record(self._ew_stream)

    def _ewald_end(self):
        self.stream.wait_event(self._ew_done)
        gpu.check(gpu.lib().changaAddVars(_dp(self.d_vars), _dp(self._ew_vars), self.n_targets,
                                          self._sp()), "changaAddVars")

    def step(self, moments=True):
        """One force evaluation: multipole build (and the Ewald root record
        from it), zero, PC, PP, Ewald."""
        if moments:
            self.build_moments()
            self.ewald_root()
        self.zero()
        self.run_pc()
        self.run_pp()
        self.run_ewald()

    # ------------------------------------------------------------------ batched form
    def batch_plan(self, pc_tile=0, pp_tile=0, lists_per_grab=0, order=0, zero_vars=True,
                   targets_per_lane=0, cell_source=0):
        """Make (replacing any previous) the batched-form plan over this
        evaluation's lists; options left at 0 are autotuned. cell_source 1:
        the cell records, 2: the wire moments. Returns the chosen
        ChangaBatchOptions."""
        self._need_lists()
        self.free_batch()
        o = gpu.ChangaBatchOptions(pc_tile, pp_tile, lists_per_grab, order, int(zero_vars), 0.0,
                                   targets_per_lane, cell_source)
        self._batch_launches = (self._launch(True), self._launch(False))
        plan = ctypes.c_void_p()
        gpu.check(gpu.lib().changaBatchPlanCreate(ctypes.byref(self._batch_launches[0]),
                                                  ctypes.byref(self._batch_launches[1]),
                                                  ctypes.byref(o), ctypes.byref(plan), self._sp()),
                  "changaBatchPlanCreate")
        self._batch_plan = plan
        self.batch_options = o
        return o

    def run_batched(self, graph=True):
        """PC + PP of every bucket in one persistent launch (graph replay, or
        the same kernels enqueued directly)."""
        fn = gpu.lib().changaBatchPlanRun if graph else gpu.lib().changaBatchPlanLaunch
        gpu.check(fn(self._batch_plan, self._sp()), "changaBatchPlan")

    def step_batched(self, moments=True, overlap_ewald=True):
        """One force evaluation in the batched form: multipole build (and the
        Ewald root record from it), the plan's graph (zeroing, PC + PP),
        Ewald. With overlap_ewald (periodic boxes) the Ewald pass runs on a
        second stream beside the graph and is merged after it; the results
        are bit-identical (the merge adds the same FP32 values the direct
        pass adds)."""
        overlap = overlap_ewald and self.w["periodic"]
        if moments:
            self.build_moments()
            self.ewald_root()
        if overlap:
            self._ewald_begin()
        self.run_batched()
        if overlap:
            self._ewald_end()
        else:
            self.run_ewald()

    # ------------------------------------------------------------------ fused form
    def fused_setup(self, max_cells=16384, max_runs=8192):
        """Scratch of the fused walk-and-force launch (per-wavefront cell and
        run capacities; a bucket whose walk exceeds them fails the launch)."""
        if not self.precise:
            raise ValueError("the fused form runs on the precision form (FP64 positions)")
        nbytes = gpu.lib().changaFusedGravityWorkBytes(self.tree["n_nodes"], int(max_cells),
                                                       int(max_runs))
        if nbytes == 0:
            raise RuntimeError("changaFusedGravityWorkBytes failed")
        self.d_fused_work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self._fused_err = torch.zeros(1, dtype=torch.int32).pin_memory()
        if getattr(self, "d_bucket_node", None) is None:
            self.d_bucket_node = self._bucket_nodes()
        f = gpu.ChangaFusedLaunch()
        f.cells = self._launch(True)
        f.parts = self._launch(False)
        f.bucketNode = self.d_bucket_node.data_ptr()
        f.first = self.d_first.data_ptr()
        f.count = self.d_count.data_ptr()
        f.child = self.d_child.data_ptr()
        f.node64 = self.d_node64.data_ptr()
        f.numNodes = self.tree["n_nodes"]
        f.theta = float(self.w["theta"])
        f.nReplicas = int(self.w["n_replicas"]) if self.w["periodic"] else 0
        f.period = self.L
        f.maxCells, f.maxRuns = int(max_cells), int(max_runs)
        f.work = self.d_fused_work.data_ptr()
        f.workBytes = nbytes
        f.errorFlag = self._fused_err.data_ptr()
        self._fused = f

    def run_fused(self):
        """PC + PP of every target bucket by the fused walk-and-force launch."""
        if getattr(self, "_fused", None) is None:
            self.fused_setup()
        gpu.check(gpu.lib().changaFusedGravity(ctypes.byref(self._fused), self._sp()),
                  "changaFusedGravity")

    def fused_check(self):
        """Raise when the last fused launch (after a synchronize) overflowed
        its scratch."""
        if int(self._fused_err[0]):
            raise RuntimeError("changaFusedGravity: a walk exceeded the scratch capacities "
                               "(raise max_cells / max_runs in fused_setup)")

    def step_fused(self, moments=True, overlap_ewald=True):
        """One force evaluation in the fused form: multipole build, zero, the
        fused walk and force, Ewald (overlapped on a second stream as in
        step_batched)."""
        overlap = overlap_ewald and self.w["periodic"]
        if moments:
            self.build_moments()
            self.ewald_root()
        if overlap:
            self._ewald_begin()
        self.zero()
        self.run_fused()
        if overlap:
            self._ewald_end()
        else:
            self.run_ewald()

    # ------------------------------------------------------------------ adjoint form
    def adjoint(self, w, ewald=None):
        """dL/dx and dL/dm (tree order, [N, 4] FP64 on the device) of
        L = sum_i w_i . a_i over the targets, for this evaluation's lists
        (precision form), with the moments of the last build. w: [n_targets, 3]
        in target order (tree order when every particle is a target).
        ewald: include the Ewald correction in a (default: when periodic)."""
        if not self.precise:
            raise ValueError("the adjoint runs on the precision form (FP64 positions)")
        self._need_lists()
        n = self.n
        wt = torch.as_tensor(w, dtype=torch.float32).to(self.device).contiguous()
        if tuple(wt.shape) != (self.n_targets, 3):
            raise ValueError(f"w must be [{self.n_targets}, 3]")
        tr = self.tree
        grad = torch.empty((n, 4), dtype=torch.float64, device=self.device)
        entries = int(self.t_lists["cell_marks"][-1])
        part_entries = int(self.t_lists["part_marks"][-1])
        if getattr(self, "d_node_work", None) is None:
            self.d_node_work = torch.empty((tr["n_nodes"], 112), dtype=torch.float64,
                                           device=self.device)
        ints = gpu.lib().changaGravityAdjointWorkInts(tr["n_nodes"], entries, part_entries, n)
        if ints == 0:
            raise RuntimeError("changaGravityAdjointWorkInts failed")
        work = getattr(self, "d_list_work", None)
        if work is None or work.numel() < ints:
            # sized for the current lists (they change with build_lists)
            self.d_list_work = torch.empty(ints, dtype=torch.int32, device=self.device)
        lo = np.ascontiguousarray(tr["level_offsets"], dtype=np.int32)
        a = gpu.ChangaAdjointLaunch()
        a.cells = self._launch(True)
        a.parts = self._launch(False)
        a.w = wt.data_ptr()
        a.targetIndex = self.d_tindex.data_ptr() if self.d_tindex is not None else None
        a.grad = grad.data_ptr()
        a.nodeWork = self.d_node_work.data_ptr()
        a.listWork = self.d_list_work.data_ptr()
        a.listWorkInts = self.d_list_work.numel()
        a.numCellEntries = entries
        a.numPartEntries = part_entries
        a.nNodes = tr["n_nodes"]
        a.nDepths = tr["n_depths"]
        a.nParticles = n
        a.levelOffsets = lo.ctypes.data
        a.first = self.d_first.data_ptr()
        a.count = self.d_count.data_ptr()
        a.child = self.d_child.data_ptr()
        a.node64 = self.d_node64.data_ptr()
        a.pos64 = self.d_pos64.data_ptr()
        a.mass64 = self.d_mass64.data_ptr()
        if ewald is None:
            ewald = bool(self.w["periodic"])
        if ewald:
            self._refresh_host_root()
            if not self.w["periodic"]:
                raise ValueError("the Ewald term needs a periodic box")
            nd = gpu.lib().changaGravityAdjointEwaldWorkDoubles(self.n_targets,
                                                                 self.ewald_ro.nEwhLoop)
            work = getattr(self, "d_ewald_work", None)
            if work is None or work.numel() < nd:
                self.d_ewald_work = torch.empty(nd, dtype=torch.float64, device=self.device)
            a.ewaldRO = ctypes.addressof(self.ewald_ro)
            a.ewt = self.d_ewt.data_ptr()
            a.rootCm64 = self.root_cm64.ctypes.data
            a.nTargets = self.n_targets
            a.ewaldWork = self.d_ewald_work.data_ptr()
            a.ewaldWorkDoubles = self.d_ewald_work.numel()
        gpu.check(gpu.lib().changaGravityAdjoint(ctypes.byref(a), self._sp()),
                  "changaGravityAdjoint")
        return grad

    def free_batch(self):
        plan = getattr(self, "_batch_plan", None)
        if plan is not None and plan.value:
            torch.cuda.synchronize(self.device)
            gpu.lib().changaBatchPlanDestroy(plan)
        self._batch_plan = None

    def __del__(self):
        try:
            self.free_batch()
        except Exception:
            pass

    def results(self):
        """(acc[N,3], pot[N], dtgrav[N]) in tree order, FP32 as the wire;
        zero for particles that are not targets."""
        v = self.d_vars.cpu().numpy().astype(np.float64)
        if self.t_index is not None:
            full = np.zeros((self.n, 5))
            full[self.t_index] = v
            v = full
        return v[:, 0:3], v[:, 3], v[:, 4]

    def node64(self):
        return self.d_node64.cpu().numpy()

    def n_interactions(self):
        pc, soft, pp = self.counts
        er, ek = self.ewald_counts()
        return dict(pc=pc, soft=soft, pp=pp, ewald_real=er, ewald_k=ek)
