"""Benchmark of the collision path on one MI355X.

Default workload: 16 stacked cloth layers of 180x180 quads (1,036,800
triangles, 524,176 vertices), per-layer rigid tilt + z-twist + noise, ACCD
thickness 1e-4 — the "1M triangles" configuration of the headline metric.
One step = update (leaf boxes, Morton codes, radix sort, LBVH) + traverse
(self-traversal, candidate generation, FP64 ACCD, discrete tri-tri) with the
inputs already resident in HBM; the step ends when every result is on the
device and its counts are known on the host.

Prints ONE JSON line.  The roofline figure is for the dominant kernel, timed
with HIP events recorded on the library's stream around its launches; the
update kernels' HBM rates are under kernels.build.  The CPU baseline is the
oracle (a C++ restatement of the reference path, OpenMP over the same
decomposition) timed on this host's cores on the same full workload (one warm
frame at 1M triangles), with the CPU model, the cold first update and the
8,192-triangle cloth at all cores and at 1 thread beside it.  Other workloads:
--workload cloth | disk | disk-boundary (the 2D edge paths).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ms per update+traverse at 1M tris; batched scenes/sec vs plain GPU at 1e-9 TOI tol"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector, FMA counted as 2 (39.3 T non-FMA ops/s)

# Algorithmic flops per unit (add/sub/mul/div/sqrt = 1; compare, min/max,
# negation = 0), from NarrowPhase.cpp:634-750 and Projection.hpp:60-145,193-204,241-302.
# Setup: displacements, centring, l_p (one norm per vertex), per stencil.
FLOPS_ACCD_SETUP = {3: 73, 2: 37}
# One ACCD iteration: step x += t_l p (3D 24, 2D 12), the distance (VF 68, EE 67,
# 2D 21) and eps / t_l (13).
FLOPS_PER_EVAL = {"vf3": 105, "ee3": 104, "vf2": 46}


# HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE) on
# this workload, scaled by the same counters on a 1 GiB device copy; the
# summary (with the commands) is the file below.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "stacked_layers_1M_pmc_traffic.md")


def profiled_traffic(kernel):
    """'traffic_bytes_per_launch <kernel>: <bytes>' from the committed PMC summary, else None."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                line = line.strip().strip("`")
                if line.startswith("traffic_bytes_per_launch ") and ":" in line:
                    k, v = line[len("traffic_bytes_per_launch "):].rsplit(":", 1)
                    if k.strip() == kernel:
                        return float(v)
    except OSError:
        pass
    return None


# FP64 VALU counters per launch (rocprofv3 --pmc SQ_INSTS_VALU_*_F64 ...) on
# this workload; the summary (with the command) is the file below.
FP64_FILE = os.path.join(ROOT, "profiles", "stacked_layers_1M_pmc_fp64_valu.md")
CUS, CLOCK_HZ = 256, 2.4e9  # MI355X: one wave64 FP64 VALU instruction per CU per cycle at peak


def profiled_fp64(kernel, ms):
    """FP64 VALU use of one ACCD launch from the committed PMC summary, scaled by
    this run's kernel time: VALU issue fraction, FP64 TFLOP/s (FMA = 2), lane use."""
    try:
        with open(FP64_FILE) as f:
            for line in f:
                line = line.strip().strip("`")
                if not line.startswith("fp64_valu ") or ":" not in line:
                    continue
                k, rest = line[len("fp64_valu "):].rsplit(":", 1)
                if k.strip() != kernel:
                    continue
                v = dict(kv.split("=") for kv in rest.split())
                v = {a: float(b) for a, b in v.items()}
                t = ms * 1e-3
                tf = v["f64_flop_insts"] * 64 * v["lane_use"] / t / 1e12
                return {"valu_issue_frac": round(v["valu_insts"] / (CUS * CLOCK_HZ * t), 3),
                        "fp64_share_of_valu": round(v["f64_insts"] / v["valu_insts"], 3),
                        "lane_use": v["lane_use"], "fp64_tflops": round(tf, 2),
                        "fp64_frac": round(tf / FP64_PEAK_TFLOPS, 4),
                        "note": "hardware FP64 instruction counts (division and sqrt sequences included) x 64 x "
                                "lane use / this run's kernel time; issue fraction against one wave64 VALU "
                                "instruction per CU per cycle"}
    except (OSError, ValueError, KeyError):
        pass
    return None


def accd_flops(c, dim):
    """Algorithmic ACCD flops of one traverse: (VF, EE)."""
    vf_cand = c["vf_candidates"]
    ee_cand = c["candidates"] - vf_cand
    vf_evals = c["dist_evals"] - c["ee_dist_evals"]
    su = FLOPS_ACCD_SETUP[dim]
    if dim == 2:
        return su * c["candidates"] + FLOPS_PER_EVAL["vf2"] * c["dist_evals"], 0
    return (su * vf_cand + FLOPS_PER_EVAL["vf3"] * vf_evals,
            su * ee_cand + FLOPS_PER_EVAL["ee3"] * c["ee_dist_evals"])


def roofline(c, avg, n, nv, dim):
    """Roofline of the longest single kernel of the step (HIP-event times)."""
    kern = {"traverse": avg["traverse"], "accd_vf": avg["accd_vf"], "accd_ee": avg["accd_ee"],
            "expand": avg["expand"]}
    dom = max(kern, key=kern.get)
    node_b, leaf_b = (64, 144) if dim == 3 else (48, 96)
    if dom == "traverse":
        # compulsory bytes: sorted nodes and FP64 leaf boxes once, 8 B per emitted pair
        byts = node_b * (2 * n - 1) + leaf_b * n + 8 * c["leaf_pairs"]
        achieved = byts / (avg["traverse"] * 1e-3) / 1e9
        kernel = f"k_traverse<{dim}>"
        return {"kernel": kernel, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": profiled_traffic(kernel) if dim == 3 else None,
                "algorithmic_bytes": byts, "ms": round(avg["traverse"], 3),
                "note": f"algorithmic bytes = {node_b}*(2n-1) nodes + {leaf_b}*n FP64 leaf boxes + 8*leaf_pairs; "
                        "traffic = rocprof FETCH_SIZE+WRITE_SIZE per launch (calibrated on a 1 GiB copy)"}
    if dom == "expand":
        head_b, vert_b = (24, 32) if dim == 3 else (16, 16)
        byts = 16 * c["leaf_pairs"] + head_b * n + vert_b * nv + 16 * c["candidates"] + 8 * c["discrete"]
        achieved = byts / (avg["expand"] * 1e-3) / 1e9
        kernel = f"k_expand<{dim}, false>"
        return {"kernel": kernel + " (+ count pass and scans)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": profiled_traffic(kernel) if dim == 3 else None, "algorithmic_bytes": byts,
                "note": f"algorithmic bytes = pairs read twice (8+8) + {head_b} B face arrays per leaf + {vert_b} B V1 "
                        "per vertex + 16 B per stencil + 8 B per discrete hit"}
    # one ACCD kernel: the vertex-face or the edge-edge stencils
    vf = dom == "accd_vf"
    fv, fe = accd_flops(c, dim)
    flops = fv if vf else fe
    cand = c["vf_candidates"] if vf else c["candidates"] - c["vf_candidates"]
    evals = c["dist_evals"] - c["ee_dist_evals"] if vf else c["ee_dist_evals"]
    per_eval = FLOPS_PER_EVAL["vf2" if dim == 2 else ("vf3" if vf else "ee3")]
    ms = avg[dom]
    achieved = flops / (ms * 1e-3) / 1e12
    kernel = f"k_accd<{dim}, true, double>" if vf else f"k_accd<{dim}, false, double>"
    return {"kernel": f"{kernel} (FP64 ACCD, {'VF' if vf else 'EE'} stencils)", "bound": "valu",
            "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": profiled_traffic(kernel) if dim == 3 else None,
            "algorithmic_flops": flops, "stencils": cand, "dist_evals": evals, "ms": round(ms, 3),
            "pmc": profiled_fp64(kernel, ms) if dim == 3 else None,
            "note": f"algorithmic flops = {FLOPS_ACCD_SETUP[dim]}*stencils + {per_eval}*distance evaluations, "
                    "div/sqrt = 1 flop; FP64 vector peak 78.6 TFLOP/s counts FMA as 2 and parity forbids "
                    "contraction, so the usable rate is 39.3 T ops/s"}


def build_rates(avg, n, nv, dim, pdim):
    """HBM rate of each update kernel group against its compulsory bytes."""
    rec_b = 64 if dim == 3 else 32        # packed vertex record (t0 and t1 positions)
    leaf_b = 144 if dim == 3 else 96      # FP64 leaf box (swept, t0, t1)
    node_b = 64 if dim == 3 else 48       # FP32 node
    cp_b = 2 * node_b                     # child-pair record per internal node
    phases = {
        "pack": (4 * pdim * n + 16 * dim * nv + rec_b * nv,
                 "index range check reads P; read V0 and V1, write the packed records"),
        "leaf_boxes": (4 * pdim * n + rec_b * nv + (leaf_b + 1) * n,
                       "read P and each vertex record once, write FP64 leaf boxes and active flags"),
        "morton": ((leaf_b // 3) * n + 8 * n, "read swept boxes, write keys and indices"),
        "sort": (4 * 20 * n, "4 passes of 8-bit digits: histogram reads keys (4 B), scatter reads and writes "
                             "key + value (16 B)"),
        "lbvh": (8 * n + 2 * leaf_b * n + node_b * (2 * n - 1) + (cp_b + 16) * (n - 1),
                 "read keys, indices and leaf boxes; write sorted leaf boxes, nodes, child pairs, child and range "
                 "links"),
    }
    out = {}
    total_b, total_ms = 0, 0.0
    for k, (b, what) in phases.items():
        ms = avg["update_" + k]
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else None
        out[k] = {"ms": round(ms, 4), "bytes": int(b), "gb_s": round(gbs, 1) if gbs else None,
                  "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None, "bytes_are": what}
        total_b += b
        total_ms += ms
    out["total"] = {"ms": round(total_ms, 4), "bytes": int(total_b),
                    "gb_s": round(total_b / (total_ms * 1e-3) / 1e9, 1) if total_ms > 0 else None}
    return out


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


def make_workload(name):
    from mclccd_amd import scenes

    if name == "layers":
        V0, V1, F, th = scenes.stacked_layers()
        desc = "16 stacked cloth layers, 180x180 quads each, rigid tilt U(0,0.3) + z-twist U(0,0.5) + noise 0.005, seed 2"
    elif name == "cloth":
        V0, V1, F, th = scenes.cloth_fold()
        desc = "64x64-quad cloth sheet folded by 0.3 sin(3 pi x) sin(2 pi y) + noise, seed 0"
    elif name in ("disk", "disk-boundary"):
        edges = "boundary" if name == "disk-boundary" else "all"
        V0, V1, F, th = scenes.disk2d(edges=edges)
        desc = (f"2D triangulated disk, 150k uniform points (seed 1), {edges} edges, swirl 0.4 r exp(-r^2) + "
                "noise 1e-3, theta 1e-3")
    else:
        raise SystemExit(f"unknown workload {name}")
    return V0, V1, F, th, desc


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform

    return platform.processor() or "unknown"


def _oracle_frames(V0, V1, F, th, threads, runs, deadline_s):
    """Cold update (reptri) once, then warm update+traverse frames; callbacks
    append to per-thread lists (the oracle keeps every hit).  Returns (cold
    update ms, sorted warm ms, oracle)."""
    from oracle.oracle import Oracle

    o = Oracle(V0.shape[1])
    o.set_options(box_eta=th, vf_ccd_eta=th, ee_ccd_eta=th, nthreads=threads)
    t0 = time.perf_counter()
    o.update(V0, V1, F)
    cold = (time.perf_counter() - t0) * 1e3
    times = []
    t_end = time.time() + deadline_s
    for _ in range(runs):
        t0 = time.perf_counter()
        o.update(V0, V1, F)
        o.traverse(leaf_pairs=False, candidates=False)
        times.append((time.perf_counter() - t0) * 1e3)
        if time.time() > t_end:
            break
    return cold, sorted(times), o


def cpu_baseline(name, V0, V1, F, th, gpu_counts):
    """Times the oracle (the reference path restated in C++, OpenMP over the
    same decomposition: leaves, then the frontlist pairs) on this host's cores,
    on the full workload: one warm frame for the 1M-triangle scene (tens of
    seconds), the median of up to 5 otherwise; plus the 8,192-triangle cloth
    at all cores and at 1 thread (median of 5)."""
    from mclccd_amd import scenes

    cores = _cores()
    runs = 1 if name == "layers" else 5
    cold, times, o = _oracle_frames(V0, V1, F, th, cores, runs, 60.0)
    ms = times[len(times) // 2]
    same = {k: o.count(i) == gpu_counts[k] for i, k in ((2, "hits"), (3, "discrete"))}
    del o
    cV0, cV1, cF, cth = scenes.cloth_fold()
    _, c_all, _ = _oracle_frames(cV0, cV1, cF, cth, cores, 5, 10.0)
    _, c_one, _ = _oracle_frames(cV0, cV1, cF, cth, 1, 5, 20.0)
    what = "one warm update+traverse" if runs == 1 else f"median of {len(times)} warm update+traverse runs"
    return {"value": round(ms, 3), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"full workload, {what} (callbacks append every hit and discrete pair to per-thread lists); "
                      f"cold first update (reptri) {cold:.1f} ms",
            "cpu_model": cpu_model(), "cold_update_ms": round(cold, 1),
            "same_counts_as_gpu": same,
            "cloth_8192_tris_ms": {"all_cores": round(c_all[len(c_all) // 2], 3),
                                   "one_thread": round(c_one[len(c_one) // 2], 3)}}


def batch_bench(steps, autotune=True):
    """512 independent folded sheets (n log-uniform in [16,180]): one batched
    frame (graph replay) vs the plain path (one update+traverse per scene on one
    context, inputs resident in HBM).  Returns the numbers for the JSON line."""
    import torch

    from mclccd_amd import scenes
    from mclccd_amd.batch import from_scenes
    from mclccd_amd.tree import BVHTree

    sc = scenes.batch_sheets()
    b, V0, V1 = from_scenes(sc)
    dev = torch.device("cuda:0")
    dV0 = torch.from_numpy(V0).to(dev)
    dV1 = torch.from_numpy(V1).to(dev)
    torch.cuda.synchronize()
    b.run(dV0, dV1)
    if autotune:
        b.autotune(None, reps=2)
    b.run(dV0, dV1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run(dV0, dV1)
    torch.cuda.synchronize()
    batch_ms = (time.perf_counter() - t0) * 1e3 / steps
    cb = b.counts()
    # plain GPU path: one BVHTree, scene after scene
    t = BVHTree(3)
    per = []
    for a0, a1, F, th in sc:
        per.append((torch.from_numpy(a0).to(dev), torch.from_numpy(a1).to(dev), torch.from_numpy(F).to(dev), th))
    torch.cuda.synchronize()

    def plain_pass():
        hits = 0
        for a0, a1, F, th in per:
            t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
            t.update(a0, a1, F)
            t.run(a0, a1, F)
            hits += t.counts()["hits"]
        return hits

    plain_pass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, steps // 2)):
        hits = plain_pass()
    torch.cuda.synchronize()
    plain_ms = (time.perf_counter() - t0) * 1e3 / max(1, steps // 2)
    assert hits == cb["hits"], (hits, cb["hits"])
    n = len(sc)
    return {"scenes": n, "tris": int(sum(len(s[2]) for s in sc)), "batched_ms_per_frame": round(batch_ms, 3),
            "batched_scenes_per_s": round(n / batch_ms * 1e3, 1), "plain_ms_per_frame": round(plain_ms, 3),
            "plain_scenes_per_s": round(n / plain_ms * 1e3, 1), "speedup": round(plain_ms / batch_ms, 2),
            "hits": cb["hits"], "params": b.params(),
            "note": "hit total asserted equal to the plain path's here; every scene of this exact mix is checked "
                    "against the oracle in tests/test_gpu_batch_c4.py (VF bit-identical, EE within 1e-9); "
                    "plain = one update+traverse per scene on one context"}


def cpu_batch():
    """The 512-scene mix on the oracle, scene after scene at all host cores:
    one warm pass (cold updates first, outside the timing)."""
    from mclccd_amd import scenes
    from oracle.oracle import Oracle

    cores = _cores()
    sc = scenes.batch_sheets()
    trees = []
    for a0, a1, F, th in sc:
        o = Oracle(3)
        o.set_options(box_eta=th, vf_ccd_eta=th, ee_ccd_eta=th, nthreads=cores)
        o.update(a0, a1, F)
        trees.append(o)
    hits = 0
    t0 = time.perf_counter()
    for (a0, a1, F, th), o in zip(sc, trees):
        o.update(a0, a1, F)
        o.traverse(leaf_pairs=False, candidates=False)
        hits += o.count(2)
    ms = (time.perf_counter() - t0) * 1e3
    return {"ms_per_frame": round(ms, 1), "scenes_per_s": round(len(sc) / ms * 1e3, 2), "cores": cores,
            "hits": hits, "sample": "one warm pass over all 512 scenes"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="layers", choices=["layers", "cloth", "disk", "disk-boundary"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the TOI adjoint")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes comparison")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one 1 GiB device-to-device copy (reference kernel for PMC byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (the path runs on one GPU; "
              "sharding by ranges of sorted query leaves is not implemented)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mclccd_amd.tree import BVHTree

    V0, V1, F, th, desc = make_workload(args.workload)
    dev = torch.device("cuda:0")
    if args.calibrate_copy:  # reads and writes exactly 1 GiB
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    dV0 = torch.from_numpy(V0).to(dev)
    dV1 = torch.from_numpy(V1).to(dev)
    dF = torch.from_numpy(F).to(dev)
    torch.cuda.synchronize()

    dim = V0.shape[1]
    t = BVHTree(dim)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    print(f"[bench] workload: {desc}; {len(F)} primitives", file=sys.stderr, flush=True)
    for i in range(args.warmup):
        t.update(dV0, dV1, dF)
        t.run(dV0, dV1, dF)
        print(f"[bench] warmup {i}: {t.counts()} {t.timings()}", file=sys.stderr, flush=True)

    phases = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        t.update(dV0, dV1, dF)
        upd = t.timings()
        t.run(dV0, dV1, dF)
        tr = t.timings()
        ph = {k: upd[k] for k in upd if k.startswith("update")}
        ph.update({k: tr[k] for k in tr if not k.startswith("update")})
        phases.append(ph)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / args.steps
    c = t.counts()
    avg = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    print(f"[bench] counts {c}", file=sys.stderr, flush=True)
    print(f"[bench] phase ms {avg}", file=sys.stderr, flush=True)

    roof = roofline(c, avg, len(F), len(V0), dim)
    # the other large kernels, for context
    fv, fe = accd_flops(c, dim)
    others = {"traverse_ms": round(avg["traverse"], 3), "expand_ms": round(avg["expand"], 3),
              "accd_vf_ms": round(avg["accd_vf"], 3), "accd_ee_ms": round(avg["accd_ee"], 3),
              "accd_tflops": round((fv + fe) / ((avg["accd_vf"] + avg["accd_ee"]) * 1e-3) / 1e12, 3)
              if avg["accd_vf"] + avg["accd_ee"] > 0 else None,
              "build": build_rates(avg, len(F), len(V0), dim, F.shape[1])}

    # TOI adjoint over every hit of the last traverse, w_i ~ U(0,1) (seed 4)
    import ctypes as C

    from mclccd_amd._lib import lib

    adjoint = None
    if not args.no_adjoint:
        gen = torch.Generator(device=dev).manual_seed(4)
        w = torch.rand(c["hits"], dtype=torch.float64, device=dev, generator=gen)
        gV0 = torch.empty_like(dV0)
        gV1 = torch.empty_like(dV0)
        adj = []
        for i in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t._check(lib.mclg_toi_grad(t._h, C.c_void_p(w.data_ptr()), C.c_void_p(gV0.data_ptr()),
                                       C.c_void_p(gV1.data_ptr()), 1))
            if i >= args.warmup:
                adj.append(t.timings()["adjoint"])
        del w, gV0, gV1
        adjoint = {"ms": round(float(np.median(adj)), 3), "hits": c["hits"], "weights": "U(0,1), seed 4",
                   "note": "dL/dV0, dL/dV1 of L = sum w_i toi_i over every CCD hit (FP64, HIP events)"}
        print(f"[bench] adjoint {adjoint}", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batch:
        batched = batch_bench(max(args.steps, 3))
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        del t
        torch.cuda.empty_cache()
        print("[bench] CPU baseline (oracle, all host cores) ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.workload, V0, V1, F, th, c)
        print(f"[bench] cpu {cpu}", file=sys.stderr, flush=True)
        if batched is not None:
            batched["cpu"] = cpu_batch()
            batched["cpu"]["same_hits_as_gpu"] = batched["cpu"]["hits"] == batched["hits"]
            print(f"[bench] cpu batched {batched['cpu']}", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": round(wall_ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "prims": int(len(F)), "verts": int(len(V0)), "theta": th,
                   "results": "left on the device (hit, discrete and pair lists in HBM; the host reads the counts); "
                              "no copy to the host and no callback replay inside the timed region",
                   "leaf_pairs": c["leaf_pairs"], "candidates": c["candidates"], "hits": c["hits"],
                   "discrete": c["discrete"], "dist_evals": c["dist_evals"],
                   "traverse_wave_steps": c["traverse_steps"], "accd_suspended": c["suspended"],
                   "phase_ms": {k: round(v, 4) for k, v in avg.items()}},
        "roofline": roof,
        "kernels": others,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "batched": batched,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
