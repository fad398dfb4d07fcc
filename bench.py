"""Gravity force-evaluation benchmark on one MI355X.

One step = one pass of the hot path over the resident workload: GPU
multipole build, zeroing of the accelerations, particle-cell (hexadecapole)
and particle-particle evaluation, and the Ewald correction, on interaction
lists already in HBM. The cell and particle lists run either in the batched
form (one persistent launch replayed from a hipGraph, autotuned; the
default) or per bucket (one launch per list kind, ChaNGa's pattern); both are
timed and reported.

Workloads: --config 3 (default), 2^22 particles in 512 Hernquist halos plus
a 20% uniform background (theta = 0.6, bucket 32, isolated), the clustered
configuration the metric is quoted on; --config 2, the 128^3 periodic
Zel'dovich box (2,097,152 particles, Ewald nReplicas = 1, fEwCut = 2.6,
|h|^2 <= 10, theta = 0.7, bucket 16); --config 4,
a Plummer sphere of 2^20 particles (eps 0.01, theta 0.7, bucket 16);
--config 5, the 256^3 box (rms 0.15) with a rung of 10% active particles,
the only targets. The metric is G interactions/s, an interaction being one
(active target, source cell / particle / Ewald term) pair as ChaNGa counts
them.

Adjoint (configs 4 and 5, or --adjoint): the time of changaGravityAdjoint
for L = sum w . a (w ~ N(0,1)^3 per target) on the same lists; --kdk adds one
kick-drift-kick step of forward forces and adjoint with only the positions
kept between the passes (trees, lists and moments rebuilt in the backward
pass); --grad-accuracy adds the gradient error against the FP64 oracle's
adjoint on the same lists (and, for config 4, against the FP64 direct-sum
gradient of a 4,096-particle Plummer sphere).

The tree and the interaction lists are built once, before the timed steps,
on the GPU (--tree gpu --lists gpu, the defaults; changaBuildTree,
changaBuildLists*) or on the host (--tree host, --lists host); both give the
same tree and lists. config.tree_build_ms (including the positions' upload
and the download of the topology) and config.list_build_ms report the GPU
builds' times.

--pcie adds the PCIe-inclusive rate of the request route (HostCUDA.h's
cell and particle list requests on pageable host lists, uploaded per step)
beside the resident one; --plan fixes the batched kernel's configuration
instead of autotuning it; --path picks the step behind `value` (batched,
per-bucket, fused walk and force, or the GPU_LOCAL_TREE_WALK request).

Prints one JSON line. --gpus N with N > 1 exits: this build runs on one GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12  # MI355X FP32 vector peak (FLOP/s), packed issue
PEAK_FP32_NONPACKED = 78.6e12
FLOP = dict(pc=183, soft=23, pp=23, ewald_real=333, ewald_k=58)
FORCE_ERR_BAR = 1e-4  # the metric's bar: median relative force error vs FP64 exact forces
BOX_CPU_SHARE = 16    # host threads one GPU's share of the host may use


def cpu_threads():
    """(threads used by the CPU legs, host CPUs, CPUs this process may run
    on): one thread per CPU this process may use, at most the host share of
    one GPU."""
    host = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = host
    return max(1, min(allowed, BOX_CPU_SHARE)), host, allowed


def log(msg):
    """Progress of the long legs, on stderr (stdout carries the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--form", choices=["precise", "plain"], default="precise")
    ap.add_argument("--plan", default=None,
                    help="pin the batched plan instead of autotuning it: PCTILE,PPTILE,GRAB,ORDER,"
                         "K,SOURCE (cell tile, particle tile, buckets per claim, order 1 tree / 2 "
                         "costliest first, targets per lane, cell source 1 records / 2 wire); "
                         "profiles pin the plan the untraced bench chooses")
    ap.add_argument("--path", choices=["batched", "per-bucket", "fused", "ltw"], default="batched",
                    help="ltw: the GPU_LOCAL_TREE_WALK request route (plain form)")
    ap.add_argument("--config", type=int, choices=[2, 3, 4, 5], default=3)
    ap.add_argument("--ng", type=int, default=128, help="lattice size of the periodic box")
    ap.add_argument("--n", type=int, default=1 << 22, help="particles of the halo workload")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-accuracy", action="store_true", help="skip the accuracy legs")
    ap.add_argument("--pcie", action="store_true",
                    help="also time the request route on host lists (PCIe-inclusive rate)")
    ap.add_argument("--cpu-frac", type=float, default=None,
                    help="fraction of the buckets the CPU baseline evaluates "
                         "(default 1 for config 2, 1/4 for 3, 1/10 for 5: about 10-30 s of "
                         "CPU work)")
    ap.add_argument("--samples", type=int, default=4096,
                    help="particles sampled for the error against FP64 exact forces")
    ap.add_argument("--lists", choices=["gpu", "host"], default="gpu",
                    help="interaction lists built on the GPU (default) or by the host walk")
    ap.add_argument("--tree", choices=["gpu", "host"], default="gpu",
                    help="tree built on the GPU (default) or on the host")
    ap.add_argument("--adjoint", action="store_true",
                    help="time the adjoint (default for configs 4 and 5)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the adjoint leg that configs 4 and 5 run by default")
    ap.add_argument("--kdk", action="store_true",
                    help="one kick-drift-kick step of forward + adjoint, checkpointed")
    ap.add_argument("--grad-accuracy", action="store_true",
                    help="gradient error against the FP64 oracle (minutes of CPU work)")
    return ap.parse_args()


def parse_plan(text):
    """--plan PCTILE,PPTILE,GRAB,ORDER,K,SOURCE -> the six ChangaBatchOptions
    values, each checked against what changaBatchPlanCreate accepts."""
    try:
        v = [int(x) for x in text.split(",")]
    except ValueError:
        v = []
    if len(v) != 6:
        raise SystemExit("--plan: six integers PCTILE,PPTILE,GRAB,ORDER,K,SOURCE")
    pct, ppt, grab, order, k, src = v
    if (pct not in (32, 64) or ppt not in (128, 256) or not 1 <= grab <= 64
            or order not in (1, 2) or k not in (1, 3) or src not in (1, 2)):
        raise SystemExit("--plan: cell tile 32|64, particle tile 128|256, 1..64 per claim, "
                         "order 1|2, targets per lane 1|3, cell source 1|2")
    return pct, ppt, grab, order, k, src


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(multi-GPU sharding is a design, not a measured path)"}))
        sys.exit(2)
    import torch

    from changa_amd import configs, engine

    t0 = time.time()
    if args.config == 2:
        w = configs.zeldovich(ng=args.ng, rms=0.02, cfg=2)
        workload = (f"periodic Zel'dovich box {args.ng}^3 ({args.ng ** 3} particles), "
                    "P(k)~k^-2.5, rms 0.02, Ewald nReplicas=1 fEwCut=2.6 |h|^2<=10, "
                    "theta=0.7, bucket 16, hexadecapole")
        data = "synthetic (seeded Zel'dovich lattice, numpy PCG64 seed 1002)"
        tag = f"zeldovich{args.ng}"
    elif args.config == 5:
        ng = 256 if args.ng == 128 else args.ng
        w = configs.zeldovich(ng=ng, rms=0.15, cfg=5, active_frac=0.1)
        workload = (f"periodic Zel'dovich box {ng}^3 ({ng ** 3} particles), P(k)~k^-2.5, "
                    "rms 0.15, rung mask with 10% of particles active (targets), Ewald "
                    "nReplicas=1 fEwCut=2.6 |h|^2<=10, theta=0.7, bucket 16, hexadecapole")
        data = "synthetic (seeded Zel'dovich lattice and Bernoulli(0.1) rung, numpy PCG64 seed 1005)"
        tag = f"zeldovich{ng}_active"
    elif args.config == 4:
        w = configs.workload(4)
        workload = ("Plummer sphere, 1048576 equal-mass particles (a=1, total mass 1), eps=0.01 "
                    "spline softening, theta=0.7, bucket 16, hexadecapole, isolated")
        data = "synthetic (seeded Plummer sphere, numpy PCG64 seed 1004)"
        tag = "plummer1048576"
    else:
        w = configs.halos(n=args.n, cfg=3)
        workload = (f"{args.n} particles: 512 Hernquist halos (dN/dM~M^-1.9 over 3 decades, "
                    "a~M^(1/3)) + 20% uniform background, isolated, theta=0.6, bucket 32, "
                    "hexadecapole")
        data = "synthetic (seeded halos, numpy PCG64 seed 1003)"
        tag = f"halos{args.n}"
    ev = engine.Evaluation(w, precise=(args.form == "precise"), lists=args.lists, tree=args.tree)
    setup_s = time.time() - t0
    tree_ms = None
    if args.tree == "gpu":
        kw = dict(box_lo=np.zeros(3), box_size=w["L"]) if w["periodic"] else {}
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        engine.gpu_build_tree(w["pos"], w["bucket"], **kw)
        tree_ms = (time.perf_counter() - t1) * 1e3
    list_ms = None
    if args.lists == "gpu":
        ev.build_lists()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ev.build_lists()
        torch.cuda.synchronize()
        list_ms = (time.perf_counter() - t1) * 1e3
    log(f"workload set up in {setup_s:.1f} s")
    counts = ev.n_interactions()
    n_int = sum(counts.values())
    flops = sum(FLOP[k] * v for k, v in counts.items())
    t0 = time.time()
    if args.plan:
        pct, ppt, grab, order, kmax, src = parse_plan(args.plan)
        opts = ev.batch_plan(pct, ppt, grab, order, targets_per_lane=kmax, cell_source=src)
    else:
        opts = ev.batch_plan()
    tune_s = time.time() - t0
    stream = torch.cuda.current_stream()

    def time_steps(step):
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(args.steps):
            step()
        stop.record(stream)
        torch.cuda.synchronize()
        return start.elapsed_time(stop) / args.steps

    step_ms = {"batched": time_steps(ev.step_batched), "per-bucket": time_steps(ev.step)}
    if ev.precise:
        # walk and force in one launch: the list build is inside this step
        step_ms["fused"] = time_steps(ev.step_fused)
        ev.fused_check()
    elif ev.active is None:
        # the local-tree-walk request route (libchanga_gpu_ltw.so): the
        # serialized local tree walked per bucket and evaluated, no lists
        # (every particle is a target of that route)
        from changa_amd import ltw
        nodes, parts = ltw.serialize_evaluation(ev)
        lw = ltw.LocalWalk(nodes, parts, float(w["theta"]),
                           int(w["n_replicas"]) if w["periodic"] else 0, ev.L, device=ev.device)
        step_ms["ltw"] = time_steps(lw.run)
        lw.results()
    if args.path not in step_ms:
        raise SystemExit("the fused path runs on the precision form, the ltw path on the plain "
                         "form of a workload without an active rung")
    ms = step_ms[args.path]
    value = n_int / (ms * 1e-3) / 1e9

    # per-kernel durations (HIP events on the launch stream, the median of
    # the passes), separate pass
    def time_fn(fn):
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * args.steps)]
        for i in range(args.steps):
            evs[2 * i].record(stream)
            fn()
            evs[2 * i + 1].record(stream)
        torch.cuda.synchronize()
        return float(np.median([evs[2 * i].elapsed_time(evs[2 * i + 1])
                                for i in range(args.steps)]))

    kern = {"moments": time_fn(ev.build_moments), "zero": time_fn(ev.zero),
            "pc": time_fn(ev.run_pc), "pp": time_fn(ev.run_pp),
            "batched_graph": time_fn(ev.run_batched)}
    if ev.precise:
        kern["fused"] = time_fn(ev.run_fused)
    elif "ltw" in step_ms:
        kern["ltw"] = time_fn(lw.run)
        lw.results()
    if w["periodic"]:
        kern["ewald"] = time_fn(ev.run_ewald)
    pc_flop = FLOP["pc"] * counts["pc"] + FLOP["soft"] * counts["soft"]
    if args.path == "batched":
        # the graph holds the zeroing kernel, the queue reset and batch_kernel
        dom, dom_ms = "batch_kernel", kern["batched_graph"] - kern["zero"]
        dom_flop = pc_flop + FLOP["pp"] * counts["pp"]
    elif args.path == "fused":
        dom, dom_ms = "fused_kernel", kern["fused"]
        dom_flop = pc_flop + FLOP["pp"] * counts["pp"]
    elif args.path == "ltw":
        # the zeroing kernel, the counter reset and ltw_kernel
        dom, dom_ms = "ltw_kernel", kern["ltw"] - kern["zero"]
        dom_flop = pc_flop + FLOP["pp"] * counts["pp"]
    else:
        dom, dom_ms, dom_flop = "pc_kernel", kern["pc"], pc_flop
    achieved = dom_flop / (dom_ms * 1e-3) / 1e12

    out = {
        "metric": "G interactions/s (PC+PP+Ewald) & % FP32 VALU peak, 1 GPU, at <=1e-4 force err",
        "value": round(value, 2),
        "unit": "G interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": data,
        "config": {
            "workload": workload,
            "form": args.form,
            "path": args.path,
            "step": ("zero + one local-tree-walk request (walk + PC + PP, no lists) on the "
                     "serialized tree; the multipole build and Ewald are not in this step"
                     if args.path == "ltw" else
                     "GPU multipole build + zero + PC + PP" + (" + Ewald" if w["periodic"] else "")
                     + (" with the tree walk fused in (no lists)" if args.path == "fused"
                        else " on resident lists")),
            "interactions_per_step": counts,
            "n_buckets": int(ev.tree["n_buckets"]),
            "n_nodes": int(ev.tree["n_nodes"]),
            "host_setup_s": round(setup_s, 2),
            "lists": args.lists,
            "list_build_ms": None if list_ms is None else round(list_ms, 2),
            "tree": args.tree,
            "tree_build_ms": None if tree_ms is None else round(tree_ms, 1),
        },
        "step_ms": {k: round(v, 4) for k, v in step_ms.items()},
        "batched_options": {"pcTile": opts.pcTile, "ppTile": opts.ppTile,
                            "listsPerGrab": opts.listsPerGrab,
                            "order": {1: "tree", 2: "costliest first"}[opts.order],
                            "targetsPerLane": opts.targetsPerLane,
                            "cellSource": {1: "records", 2: "wire"}[opts.cellSource],
                            "tuned_ms": round(opts.tunedMs, 4), "tune_s": round(tune_s, 2)},
        "fp32_valu_frac": round(flops / (ms * 1e-3) / PEAK_FP32, 4),
        "algorithmic_tflops": round(flops / (ms * 1e-3) / 1e12, 2),
        "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
        "roofline": {
            "bound": "valu",
            "pipe": "FP32 VALU (no matrix work; packed FP32 disabled, so the non-packed "
                    "ceiling is 78.6 TFLOP/s)",
            "kernel": dom,
            "achieved": round(achieved, 3),
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": round(achieved * 1e12 / PEAK_FP32, 4),
        },
    }
    # the last template arguments of the batched kernel: cell tile, particle
    # tile, targets per lane, cell records (true) or wire moments (false)
    rec = "true" if opts.cellSource == 1 else "false"
    variant = {"batch_kernel": f"batch_kernel<{opts.pcTile},{opts.ppTile},{opts.targetsPerLane},"
                               f"{rec}>",
               "fused_kernel": "fused_kernel<64,256>"}.get(dom, dom)
    traffic, traffic_kernel = traffic_from_profile(
        tag, variant + (" graph replays" if dom == "batch_kernel" else ""))
    out["roofline"]["kernel"] = variant
    out["roofline"]["kernel_ms"] = round(dom_ms, 4)
    out["roofline"]["traffic"] = traffic
    out["roofline"]["traffic_kernel"] = traffic_kernel
    # the same kernel's median duration in the committed rocprofv3 trace of
    # this workload, and the fraction it gives
    prof_ms, prof_kernel = duration_from_profile(tag, variant)
    if prof_ms:
        out["roofline"]["profile_kernel"] = prof_kernel
        out["roofline"]["profile_kernel_ms"] = prof_ms
        out["roofline"]["profile_frac"] = round(dom_flop / (prof_ms * 1e-3) / PEAK_FP32, 4)

    log(f"timed: {value:.1f} G interactions/s")
    if args.pcie:
        if ev.active is not None:
            raise SystemExit("--pcie times every bucket's lists: a workload without an active rung")
        out["pcie_inclusive"] = pcie_leg(ev, w, args, n_int)
        log("PCIe-inclusive leg done")
    if not args.no_accuracy:
        def step_ltw():
            # the request's results into the evaluation's, then the Ewald pass
            lw.run()
            ev.d_vars.copy_(lw.d_vars)
            if w["periodic"]:
                ev.run_ewald()

        out["accuracy"] = accuracy(ev, {"batched": ev.step_batched, "fused": ev.step_fused,
                                        "ltw": step_ltw}.get(args.path, ev.step), args.samples)
        log("accuracy leg done")
    if args.adjoint or args.kdk or args.grad_accuracy or (args.config in (4, 5)
                                                          and not args.no_adjoint):
        out["adjoint"] = adjoint_leg(ev, w, args)
        log("adjoint leg done")
    if not args.no_cpu:
        frac = args.cpu_frac if args.cpu_frac is not None else {2: 1.0, 3: 0.25, 4: 0.25,
                                                                5: 0.1}[args.config]
        out["cpu_baseline"] = cpu_baseline(ev, frac)
        log("cpu baseline done")
    print(json.dumps(out))


def pcie_leg(ev, w, args, n_int):
    """The HostCUDA.h request route on host lists: per step, zero, one
    TreePieceCellListDataTransferLocal and one
    TreePiecePartListDataTransferLocalCompact over all buckets, with the
    lists in pageable host memory (as Compute.cpp:1158 serializes them) that
    the library uploads on the stream, plus the Ewald pass of a periodic box;
    tree, moments and particles stay resident, and the forces are the plain
    form's. Wall-clock time per step (a pageable upload holds the host),
    and the step's interactions over it: the rate with the lists' PCIe
    transfers in it."""
    import ctypes

    import torch

    from changa_amd import gpu
    L = gpu.lib()
    stream = torch.cuda.current_stream()
    nb = int(ev.tree["n_buckets"])
    starts = np.ascontiguousarray(ev.tree["bucket_start"], dtype=np.int32)
    sizes = np.ascontiguousarray(ev.tree["bucket_size"], dtype=np.int32)
    keep, reqs, nbytes = [starts, sizes], [], 0
    for name, d_list, d_marks in (("TreePieceCellListDataTransferLocal", ev.d_cells, ev.d_cmarks),
                                  ("TreePiecePartListDataTransferLocalCompact", ev.d_plist,
                                   ev.d_pmarks)):
        lst = np.ascontiguousarray(d_list.cpu().numpy())
        marks = np.ascontiguousarray(d_marks.cpu().numpy().astype(np.int32))
        nbytes += lst.nbytes + marks.nbytes
        q = gpu.CudaRequest()
        q.stream = stream.cuda_stream
        q.d_localMoments = ev.d_moments.data_ptr()
        q.d_localParts = ev.d_parts.data_ptr()
        q.d_localVars = ev.d_vars.data_ptr()
        q.list = lst.ctypes.data
        q.bucketMarkers = marks.ctypes.data
        q.bucketStarts = starts.ctypes.data
        q.bucketSizes = sizes.ctypes.data
        q.numInteractions = int(marks[-1])
        q.numBucketsPlusOne = nb + 1
        q.fperiod = ev.L
        keep += [lst, marks]
        reqs.append((name, getattr(L, name), q))

    def step():
        ev.zero()
        for name, fn, q in reqs:
            gpu.check(fn(ctypes.byref(q)), name)
        if w["periodic"]:
            ev.run_ewald()

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    return {"value": round(n_int / (ms * 1e-3) / 1e9, 2), "unit": "G interactions/s",
            "ms_per_step": round(ms, 3), "host_list_bytes": int(nbytes),
            "route": "TreePieceCellListDataTransferLocal + "
                     "TreePiecePartListDataTransferLocalCompact on pageable host lists"
                     + (" + Ewald" if w["periodic"] else "") + ", plain form"}


def _events_ms(fn, steps):
    import torch
    stream = torch.cuda.current_stream()
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def _target_original_index(ev):
    """Original (workload) index of each target, in target order."""
    perm = ev.tree["perm"]
    return perm if ev.t_index is None else perm[ev.t_index]


def adjoint_leg(ev, w, args):
    """Adjoint time on the evaluation's lists, optionally one checkpointed
    kick-drift-kick step and the gradient error against the FP64 oracle."""
    import torch

    rng = np.random.default_rng(12)
    wt = rng.standard_normal((ev.n_targets, 3))
    wt_dev = torch.tensor(wt, dtype=torch.float32, device=ev.device)
    steps = max(1, min(args.steps, 10))
    fwd_ms = _events_ms(ev.step_batched, steps)
    adj_ms = _events_ms(lambda: ev.adjoint(wt_dev), steps)
    out = {"loss": "L = sum_targets w . a, w ~ N(0,1)^3 (seed 12)",
           "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
           "adjoint_over_forward": round(adj_ms / fwd_ms, 2),
           "includes_ewald": bool(w["periodic"])}
    if w["periodic"]:
        out["adjoint_tree_only_ms"] = round(
            _events_ms(lambda: ev.adjoint(wt_dev, ewald=False), steps), 4)
    if args.grad_accuracy:
        out["grad_accuracy"] = grad_accuracy(ev, wt, w)
        log("gradient accuracy done")
    if args.kdk:
        out["kdk"] = kdk_step(w)
        log("kdk step done")
    return out


def kdk_step(w, dt=0.01, return_grad=False):
    """One kick-drift-kick step x0 -> x1 (active particles kicked, every
    particle drifted; periodic positions wrapped), L = sum_active w . v1,
    forward then backward. The forward passes use the fused walk-and-force
    step (no lists); between the passes only the positions are kept: the
    backward pass rebuilds each tree, its lists and moments (checkpoint and
    recompute), then runs the adjoint. GPU time of the force and adjoint
    passes by HIP events; set-up time (tree, moments, lists, upload) by the
    wall clock."""
    import torch

    from changa_amd import engine

    n = w["pos"].shape[0]
    act = w.get("active")
    aidx = np.arange(n) if act is None else np.nonzero(act)[0]
    rng = np.random.default_rng(14)
    v0 = 0.01 * rng.standard_normal((n, 3))
    wv = rng.standard_normal((len(aidx), 3))
    host_s, gpu_ms = 0.0, 0.0
    stream = torch.cuda.current_stream()

    def timed_gpu(fn):
        nonlocal gpu_ms
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(stream)
        r = fn()
        b.record(stream)
        torch.cuda.synchronize()
        gpu_ms += a.elapsed_time(b)
        return r

    def build(x, lists):
        nonlocal host_s
        t = time.time()
        ev = engine.Evaluation(dict(w, pos=x), precise=True, lists=lists, tree="gpu")
        torch.cuda.synchronize()
        host_s += time.time() - t
        log(f"kdk: tree, moments{' and lists' if lists != 'none' else ''} in "
            f"{time.time() - t:.1f} s")
        return ev

    def forces(x):
        # the forward passes walk and evaluate in one launch; only the
        # adjoint needs the interaction lists
        ev = build(x, "none")
        timed_gpu(ev.step_fused)
        acc, _, _ = ev.results()
        a = np.empty((n, 3))
        a[ev.tree["perm"]] = acc
        del ev
        return a[aidx]

    def adjoint(x, w_act):
        ev = build(x, "gpu")
        # w_act is per active particle (original order); targets in target order
        orig = _target_original_index(ev)
        pos_in_act = np.empty(n, dtype=np.int64)
        pos_in_act[aidx] = np.arange(len(aidx))
        wt = w_act[pos_in_act[orig]]
        g = timed_gpu(lambda: ev.adjoint(wt)).cpu().numpy()
        out = np.empty((n, 4))
        out[ev.tree["perm"]] = g
        del ev
        return out

    x0 = w["pos"]
    box = float(w["L"]) if w["periodic"] else None
    a0 = forces(x0)
    vh = v0.copy()
    vh[aidx] += 0.5 * dt * a0
    x1 = x0 + dt * vh
    if box is not None:
        x1 = np.mod(x1, box)
        x1[x1 >= box] = 0.0
    a1 = forces(x1)
    # backward: L = sum wv . (vh + dt/2 a1)
    g1 = adjoint(x1, 0.5 * dt * wv)
    dl_da0 = 0.5 * dt * wv + 0.5 * dt * dt * g1[aidx, :3]
    g0 = adjoint(x0, dl_da0)
    grad = g0 + g1
    info = {"steps": "forward F(x0), F(x1); backward adjoint at x1, then at x0",
            "dt": dt, "gpu_ms": round(gpu_ms, 3), "host_s": round(host_s, 2),
            "total_s": round(host_s + gpu_ms * 1e-3, 3),
            "forward": "fused walk-and-force step (no lists)",
            "grad_norm_x": float(np.linalg.norm(grad[:, :3])),
            "grad_norm_m": float(np.linalg.norm(grad[:, 3]))}
    return (info, grad) if return_grad else info


def grad_accuracy(ev, wt, w, n_samples=4096):
    """GPU gradient (L = sum w . a on this evaluation's lists) against the
    FP64 oracle's exact adjoint on the same lists, on seeded particles; for
    the isolated Plummer sphere also the tree adjoint of a 4,096-particle
    sphere against the FP64 direct-sum gradient."""
    import torch

    from changa_amd import configs, engine
    from oracle import oracle as O

    threads = cpu_threads()[0]
    g = ev.adjoint(torch.tensor(wt, dtype=torch.float32, device=ev.device)).cpu().numpy()
    wfull = np.zeros_like(ev.pos)
    wfull[ev.t_index if ev.t_index is not None else slice(None)] = wt
    t = time.time()
    log("gradient accuracy: FP64 oracle adjoint")
    geo, mom = O.build_moments(ev.tree, ev.pos, ev.mass, ev.soft)
    hroot, ge = None, None
    if w["periodic"]:
        ge, hroot = O.adjoint_ewald(ev.pos, wfull, geo, mom, ev.ewt64, ev.L, w["n_replicas"],
                                    w["ew_cut"], n_threads=threads)
        log(f"gradient accuracy: oracle Ewald adjoint done ({time.time() - t:.1f} s)")
    go = O.adjoint_lists(ev.tree, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists, wfull,
                         period=ev.L, n_threads=threads, root_extra=hroot)
    if ge is not None:
        go = go + ge
    oracle_s = time.time() - t
    samp = np.random.default_rng(15).choice(ev.pos.shape[0], size=min(n_samples, ev.pos.shape[0]),
                                            replace=False)
    ex = (np.linalg.norm(g[samp, :3] - go[samp, :3], axis=1) /
          np.linalg.norm(go[samp, :3], axis=1))
    em = np.abs(g[samp, 3] - go[samp, 3]) / np.abs(go[samp, 3])

    def st(e):
        return {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                "max": float(e.max())}

    def st_rms(d, ref):
        """|error| relative to the rms of the reference over the samples
        (the gradient has zeros, where a per-particle ratio is undefined)"""
        return st(np.abs(d) / np.sqrt(np.mean(ref ** 2)))

    out = {"vs_oracle_same_lists": {"samples": int(len(samp)), "dL_dx": st(ex), "dL_dm": st(em),
                                    "oracle_s": round(oracle_s, 1), "cores": threads}}
    if not w["periodic"]:
        ws = configs.plummer(n=4096, cfg=4)
        es = engine.Evaluation(ws, precise=True, device=ev.device)
        wv = np.random.default_rng(16).standard_normal((4096, 3))
        gs = es.adjoint(wv).cpu().numpy()
        gd = O.direct_grad(es.pos, es.mass, es.soft, wv, n_threads=threads)
        ex = np.linalg.norm(gs[:, :3] - gd[:, :3], axis=1) / np.linalg.norm(gd[:, :3], axis=1)
        em = np.abs(gs[:, 3] - gd[:, 3]) / np.abs(gd[:, 3])
        ex_r = st_rms(np.linalg.norm(gs[:, :3] - gd[:, :3], axis=1),
                      np.linalg.norm(gd[:, :3], axis=1))
        em_r = st_rms(gs[:, 3] - gd[:, 3], gd[:, 3])
        out["vs_fp64_direct_4096"] = {"theta": ws["theta"], "dL_dx": st(ex), "dL_dm": st(em),
                                      "dL_dx_over_rms": ex_r, "dL_dm_over_rms": em_r}
    return out


def traffic_from_profile(tag, kernel):
    """(HBM bytes per launch, the profiled kernel) of `kernel` (for the
    batched kernel, the graph replays of its variant, e.g.
    "batch_kernel<64,256,3,false> graph replays": cell tile, particle tile,
    targets per lane, cell records), from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE and WRITE_SIZE, copy-calibrated).
    (None, None) when the summary holds no figure for exactly that kernel
    (the profiled plan chose another variant)."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{tag}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path):
        if " traffic_bytes_per_launch:" in line:
            name, v = line.split(" traffic_bytes_per_launch:")
            if name.strip() == kernel:
                return int(v), kernel
    return None, None


def duration_from_profile(tag, kernel):
    """(median ms, profiled kernel) of `kernel` in the committed rocprofv3
    kernel trace summary of this workload: for the batched kernel the median
    of the graph replays of that tile variant (the summary's line for the
    variant the plan chose in the profiled run), else the table's median
    column."""
    import re
    path = os.path.join(ROOT, "profiles", f"kernel_stats_{tag}.md")
    if not os.path.exists(path):
        return None, None
    base = kernel.split("<")[0]
    want = kernel[len(base) + 1:-1].split(",") if "<" in kernel else None
    for line in open(path):
        m = re.match(r"^(.*\)): \d+ graph replays.* median ([0-9.]+) ms", line)
        if m and base == "batch_kernel" and "batch_kernel<" in m.group(1):
            targs = m.group(1).split("batch_kernel<")[1].split(">")[0].split(", ")
            if want and targs[-len(want):] == want:
                return float(m.group(2)), m.group(1)
        elif line.startswith("| ") and base != "batch_kernel" and f"::{base}" in line:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) >= 6:
                return float(cells[5]), cells[0]
    return None, None


def accuracy(ev, step, n_samples):
    """Force error of this GPU step against (a) the FP64 oracle on the same
    lists and (b) FP64 exact forces (a converged Ewald sum for the periodic
    box, direct summation otherwise), on seeded samples."""
    import torch

    from oracle import oracle as O
    from tests.helpers import oracle_forces, rel_err

    step()
    torch.cuda.synchronize()
    acc, pot, _ = ev.results()
    rng = np.random.default_rng(7)
    tr = ev.tree
    bsel = np.sort(rng.choice(tr["n_buckets"], size=max(1, n_samples // 8), replace=False))
    oa, op, _ = oracle_forces(ev, buckets=bsel)
    idx = np.concatenate([np.arange(tr["bucket_start"][b], tr["bucket_start"][b] + tr["bucket_size"][b])
                          for b in bsel])
    targets = np.arange(ev.pos.shape[0]) if ev.active is None else np.nonzero(ev.active)[0]
    if ev.active is not None:
        idx = idx[ev.active[idx]]
    e_or = rel_err(acc[idx], oa[idx])
    n_samples = min(n_samples, len(targets))
    samp = np.sort(rng.choice(targets, size=n_samples, replace=False)).astype(np.int32)
    threads = cpu_threads()[0]
    t = time.time()
    if ev.w["periodic"]:
        xa, xp = O.ewald_exact(ev.pos, ev.mass, ev.soft, samp, L=ev.L, rcut=0.45 * ev.L,
                               n_threads=threads)
        exact = "vs_fp64_ewald"
    else:
        xa, xp = O.direct_sum(ev.pos, ev.mass, ev.soft, samp, n_threads=threads)
        exact = "vs_fp64_direct"
    e_ex = rel_err(acc[samp], xa)
    med = float(np.median(e_ex))
    return {
        "vs_oracle_same_lists": {"particles": int(len(idx)), "median": float(np.median(e_or)),
                                 "p99": float(np.percentile(e_or, 99)), "max": float(e_or.max())},
        exact: {"samples": int(n_samples), "median": med,
                "p99": float(np.percentile(e_ex, 99)), "max": float(e_ex.max()),
                "oracle_s": round(time.time() - t, 1)},
        "median_within_bar": bool(med <= FORCE_ERR_BAR),
        "bar": FORCE_ERR_BAR,
    }


def cpu_baseline(ev, frac):
    """The FP64 oracle (nodeBucketForce/partBucketForce/BucketEwald
    restatement) on a bounded sample of the same buckets and lists, on the
    host cores."""
    from oracle import oracle as O

    threads, host, allowed = cpu_threads()
    rng = np.random.default_rng(11)
    tr = ev.tree
    nb = max(1, int(frac * tr["n_buckets"]))
    bsel = np.sort(rng.choice(tr["n_buckets"], size=nb, replace=False))
    geo, mom = O.build_moments(tr, ev.pos, ev.mass, ev.soft)
    rung = None if ev.active is None else ev.active.astype(np.int32)
    t = time.time()
    _, _, _, c = O.forces_lists(tr, geo, mom, ev.pos, ev.mass, ev.soft, ev.lists,
                                period=ev.L, rung=rung, active_rung=1, n_threads=threads,
                                buckets=bsel)
    targets = np.concatenate([np.arange(tr["bucket_start"][b], tr["bucket_start"][b] + tr["bucket_size"][b])
                              for b in bsel])
    if ev.active is not None:
        targets = targets[ev.active[targets]]
    ce = np.zeros(1, dtype=np.int64)
    if ev.w["periodic"]:
        ewt = O.ewald_table(mom[0], geo[0, 3], ev.L, ev.w["ewh_cut"])
        _, _, ce = O.bucket_ewald(targets, ev.pos, geo[0, 0:3], geo[0, 5], mom[0], geo[0, 3], ewt,
                                  ev.L, ev.w["n_replicas"], ev.w["ew_cut"], n_threads=threads)
    dt = time.time() - t
    n = int(c.sum() + ce.sum())
    return {"value": round(n / dt / 1e9, 4), "unit": "G interactions/s", "cores": threads,
            "host_cpus": host, "cpus_allowed": allowed,
            "threads_note": f"one OpenMP thread per CPU this process may use, capped at the "
                            f"{BOX_CPU_SHARE}-thread host share of one GPU",
            "kind": "port",
            "sample": f"{nb} of {tr['n_buckets']} buckets ({len(targets)} particles), "
                      f"{n} interactions, {dt:.1f} s"}


if __name__ == "__main__":
    main()
