"""Benchmark of the collision path on one MI355X.

Default workload: 16 stacked cloth layers of 180x180 quads (1,036,800
triangles, 524,176 vertices), per-layer rigid tilt + z-twist + noise, ACCD
thickness 1e-4 — the "1M triangles" configuration of the headline metric.
One step = update (leaf boxes, Morton codes, radix sort, LBVH) + traverse
(self-traversal, candidate generation, FP64 ACCD, discrete tri-tri) with the
inputs already resident in HBM; the step ends when every result is on the
device and its counts are known on the host.

Prints ONE JSON line.  The roofline figure is for the dominant kernel, timed
with HIP events recorded on the library's stream around its launches.  The CPU
baseline is the oracle (a C++ restatement of the reference path, OpenMP over
the same decomposition) timed on this host on a bounded sample.
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
# negation = 0), from NarrowPhase.cpp:634-750 and Projection.hpp:60-145,241-302:
FLOPS_ACCD_SETUP = 73  # displacements, centring, l_p (4 norms)
FLOPS_PER_EVAL = 105   # one ACCD iteration: step x += t_l p (24), distance (~68), eps and t_l (~13)


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


def roofline(c, avg, n, nv):
    """Roofline of the longest single kernel of the step (HIP-event times)."""
    kern = {"traverse": avg["traverse"], "accd_vf": avg["accd_vf"], "accd_ee": avg["accd_ee"],
            "expand": avg["expand"]}
    dom = max(kern, key=kern.get)
    if dom == "traverse":
        # compulsory bytes: sorted nodes and FP64 leaf boxes once, 8 B per emitted pair
        byts = 64 * (2 * n - 1) + 144 * n + 8 * c["leaf_pairs"]
        achieved = byts / (avg["traverse"] * 1e-3) / 1e9
        return {"kernel": "k_traverse<3>", "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": profiled_traffic("k_traverse<3>"),
                "algorithmic_bytes": byts, "ms": round(avg["traverse"], 3),
                "note": "algorithmic bytes = 64*(2n-1) nodes + 144*n FP64 leaf boxes + 8*leaf_pairs; traffic = "
                        "rocprof FETCH_SIZE+WRITE_SIZE per launch (calibrated on a 1 GiB copy)"}
    if dom == "expand":
        byts = 16 * c["leaf_pairs"] + 24 * n + 32 * nv + 16 * c["candidates"] + 8 * c["discrete"]
        achieved = byts / (avg["expand"] * 1e-3) / 1e9
        return {"kernel": "k_expand<3> (+ count pass and scans)", "bound": "hbm", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                "traffic": profiled_traffic("k_expand<3>"), "algorithmic_bytes": byts,
                "note": "algorithmic bytes = pairs read twice (8+8) + 24 B face arrays per leaf + 32 B V1 per vertex + 16 B per "
                        "stencil + 8 B per discrete hit"}
    # one ACCD kernel: the vertex-face or the edge-edge stencils
    vf = dom == "accd_vf"
    cand = c["vf_candidates"] if vf else c["candidates"] - c["vf_candidates"]
    evals = c["dist_evals"] - c["ee_dist_evals"] if vf else c["ee_dist_evals"]
    flops = FLOPS_ACCD_SETUP * cand + FLOPS_PER_EVAL * evals
    ms = avg[dom]
    achieved = flops / (ms * 1e-3) / 1e12
    kernel = "k_accd<3, true>" if vf else "k_accd<3, false>"
    return {"kernel": f"{kernel} (FP64 ACCD, {'VF' if vf else 'EE'} stencils)", "bound": "valu",
            "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": profiled_traffic(kernel),
            "algorithmic_flops": flops, "stencils": cand, "dist_evals": evals, "ms": round(ms, 3),
            "pmc": profiled_fp64(kernel, ms),
            "note": f"algorithmic flops = {FLOPS_ACCD_SETUP}*stencils + {FLOPS_PER_EVAL}*distance evaluations, "
                    "div/sqrt = 1 flop; FP64 vector peak 78.6 TFLOP/s counts FMA as 2 and parity forbids "
                    "contraction, so the usable rate is 39.3 T ops/s"}


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
    else:
        raise SystemExit(f"unknown workload {name}")
    return V0, V1, F, th, desc


def cpu_baseline(name, V0, V1, F, th, gpu_leaf_pairs):
    """Times the oracle on this host.  For the 1M-triangle workload a CPU run
    takes tens of seconds per frame, so the sample is the same generator at
    90x90 quads per layer (a quarter of the triangles; about 10-20 s of host time
    for the timed frames) and the time is scaled by the ratio of broadphase leaf
    pairs (the CPU path's work is linear in them)."""
    from oracle.oracle import Oracle
    from mclccd_amd import scenes

    cores = _cores()
    if name == "layers":
        sV0, sV1, sF, sth = scenes.stacked_layers(n=90)
        sample = f"stacked layers at 90x90 quads per layer ({len(sF):,} tris), scaled by leaf pairs"
    else:
        sV0, sV1, sF, sth = V0, V1, F, th
        sample = "full workload"
    o = Oracle(3)
    o.set_options(box_eta=sth, vf_ccd_eta=sth, ee_ccd_eta=sth, nthreads=cores)
    o.update(sV0, sV1, sF)  # cold update (reptri) outside the timing
    times = []
    t_end = time.time() + 20.0
    for _ in range(5):
        t0 = time.perf_counter()
        o.update(sV0, sV1, sF)
        o.traverse(leaf_pairs=False, candidates=False)
        times.append((time.perf_counter() - t0) * 1e3)
        if time.time() > t_end:
            break
    times.sort()
    ms = times[len(times) // 2]
    lp = o.count(0)
    # leaf pairs are counted by the brute-force-checked traversal: rerun once with them kept
    o.traverse(leaf_pairs=True, candidates=False)
    lp = o.count(0)
    scale = gpu_leaf_pairs / lp if name == "layers" and lp > 0 else 1.0
    return {"value": round(ms * scale, 3), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"{sample}: median of {len(times)} warm update+traverse runs = {ms:.1f} ms for {lp} leaf pairs"
                      + (f", x{scale:.1f}" if scale != 1.0 else "")}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="layers", choices=["layers", "cloth"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
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

    t = BVHTree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    print(f"[bench] workload: {desc}; {len(F)} tris", file=sys.stderr, flush=True)
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

    roof = roofline(c, avg, len(F), len(V0))
    # the other large kernels, for context
    others = {"traverse_ms": round(avg["traverse"], 3), "expand_ms": round(avg["expand"], 3),
              "accd_vf_ms": round(avg["accd_vf"], 3), "accd_ee_ms": round(avg["accd_ee"], 3),
              "accd_tflops": round((FLOPS_ACCD_SETUP * c["candidates"] + FLOPS_PER_EVAL * c["dist_evals"]) /
                                   ((avg["accd_vf"] + avg["accd_ee"]) * 1e-3) / 1e12, 3)
              if avg["accd_vf"] + avg["accd_ee"] > 0 else None}

    # TOI adjoint over every hit of the last traverse, w_i ~ U(0,1) (seed 4)
    import ctypes as C

    from mclccd_amd._lib import lib

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
    del w
    adjoint = {"ms": round(float(np.median(adj)), 3), "hits": c["hits"], "weights": "U(0,1), seed 4",
               "note": "dL/dV0, dL/dV1 of L = sum w_i toi_i over every CCD hit (FP64, HIP events)"}
    print(f"[bench] adjoint {adjoint}", file=sys.stderr, flush=True)

    batched = None
    if not args.no_batch:
        batched = batch_bench(max(args.steps, 3))
        print(f"[bench] batched {batched}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.workload, V0, V1, F, th, c["leaf_pairs"])

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
        "config": {"workload": desc, "tris": int(len(F)), "verts": int(len(V0)), "theta": th,
                   "leaf_pairs": c["leaf_pairs"], "candidates": c["candidates"], "hits": c["hits"],
                   "discrete": c["discrete"], "dist_evals": c["dist_evals"],
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
