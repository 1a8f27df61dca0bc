"""Benchmark: Msamples/s of the plain GPU path tracer on the mixed-material 1080p / 256-spp workload.

One step = one complete render job of the workload (every pixel, every sample) through libmrg's
plain (one job at a time) schedule. Inputs (the scene) are resident in HBM before the timed region;
the timed region is submit -> job complete on the device.

Prints one JSON line with the throughput, a roofline entry for the dominant kernel (the bounce traversal
launch; VALU-issue bound, quoted against the FP32 peak with the useful box / triangle test operations per
ray of the FP64 oracle on the reference BVH, and with its measured HBM and L2 fractions), the served
request stream at several arrival rates (plain vs batched, p50/p99 latencies) and the CPU baseline (the
FP64 oracle built -march=native, on the host's CPU share and at the reference's default thread count).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload mixed|cornell|intersect|batch|...]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak (packed FP32 issue)
FP32_UNPACKED_TFLOPS = 78.6       # the same without packed FP32 (the kernels are built without it)
B_BOX, B_TRI, B_RAY_IN, B_HIT_OUT, B_OCCLUDED_OUT, B_FILM = 24, 36, 24, 8, 4, 12
FLOP_BOX, FLOP_TRI = 15, 30       # useful FP32 operations of one slab box test / one watertight triangle test

WORKLOADS = {
    # 2M-triangle fBm terrain + 64 spheres, sun + sky, thin lens f/2.8
    "mixed": dict(width=1920, height=1080, spp=256, max_depth=16, seed=0, cpu_rows=(516, 564)),
    # Cornell box 512^2, 64 spp, depth 8
    "cornell": dict(width=512, height=512, spp=64, max_depth=8, seed=0, cpu_rows=(192, 320)),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_scene(name):
    from minirender_amd import scenegen
    if name == "mixed":
        return scenegen.mixed_terrain()
    if name == "cornell":
        return scenegen.cornell_box()
    if name == "intersect":
        return scenegen.icosphere(level=8)
    raise SystemExit(f"unknown workload {name}")


def read_profile_value(name, key):
    """A `key = value` line of a profiles/ summary (figures measured by rocprofv3 or the cache probe), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key):
            return float(line.split("=")[1])
    return None


def read_traffic(kernel, workload):
    """HBM bytes per launch measured by rocprofv3 --pmc (see profiles/), or None."""
    return read_profile_value(f"{kernel}_{workload}_traffic.txt", "traffic_bytes_per_launch")


def host_cpu():
    """CPU model and core counts of the host this process runs on: std::thread::hardware_concurrency
    (logical CPUs), the CPUs this process may run on (affinity) and the cgroup CPU quota, if any."""
    import subprocess
    model = None
    try:
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith("Model name:"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if not model and os.path.exists("/proc/cpuinfo"):
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    hw = os.cpu_count() or 1
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else hw
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    share = max(1, min(usable, int(quota))) if quota else usable
    return {"model": model, "hardware_concurrency": hw, "affinity_cpus": usable, "cgroup_cpu_quota": quota,
            "cpu_share": share, "reference_default_threads": max(1, min(hw - 1, 32))}


def tests_per_ray(st):
    """Box and triangle tests per extension / shadow ray of the FP64 oracle (reference BVH topology)."""
    return ((st["bounds_tests"] - st["shadow_bounds_tests"]) / max(st["extension_rays"], 1),
            (st["triangle_tests"] - st["shadow_triangle_tests"]) / max(st["extension_rays"], 1),
            st["shadow_bounds_tests"] / max(st["shadow_rays"], 1), st["shadow_triangle_tests"] / max(st["shadow_rays"], 1))


def roofline_entry(kernel, workload, avg_s, ext_rays, sh_rays, eb, et, sb, stt, hit_bytes=B_HIT_OUT):
    """Roofline of a traversal launch. The kernel is bound by VALU issue (its rocprofv3 counters show the VALU
    busy in nearly every cycle, see profiles/), so the entry is quoted against the FP32 peak: achieved =
    useful FP32 operations per launch (15 per box test, 30 per triangle test, the FP64 oracle's tests per ray on
    the reference BVH) / the launch's average duration. The HBM and L2 fractions use the bytes the counters
    measured; the reference-BVH gather bytes of the algorithmic model are kept as their own fields."""
    flops = ext_rays * (FLOP_BOX * eb + FLOP_TRI * et) + sh_rays * (FLOP_BOX * sb + FLOP_TRI * stt)
    alg = ext_rays * (B_RAY_IN + hit_bytes + B_BOX * eb + B_TRI * et) + sh_rays * (B_RAY_IN + B_OCCLUDED_OUT + B_BOX * sb + B_TRI * stt)
    achieved = flops / avg_s / 1e12
    e = {"kernel": kernel, "bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
         "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
         "frac_of_unpacked_fp32_peak": round(achieved / FP32_UNPACKED_TFLOPS, 4),
         "traffic": read_traffic(kernel, workload), "avg_launch_ms": round(avg_s * 1e3, 4),
         "rays_per_launch": round(ext_rays + sh_rays, 1), "useful_flops_per_launch": round(flops),
         "bounds_tests_per_extension_ray": round(eb, 2), "triangle_tests_per_extension_ray": round(et, 2),
         "bounds_tests_per_shadow_ray": round(sb, 2), "triangle_tests_per_shadow_ray": round(stt, 2),
         "algorithmic_bytes_per_launch": round(alg), "algorithmic_gb_per_s": round(alg / avg_s / 1e9, 1)}
    prof = "path_kernel_counters_mixed1080p.txt"
    for key in ("valu_busy", "valu_lane_utilisation", "occupancy"):
        v = read_profile_value(prof, f"{kernel}_{key}")
        if v is not None:
            e[key] = v
    if e["traffic"]:
        hbm = e["traffic"] / avg_s / 1e9
        e.update({"hbm_achieved_gb_per_s": round(hbm, 1), "hbm_peak_gb_per_s": HBM_PEAK_GBS, "hbm_frac": round(hbm / HBM_PEAK_GBS, 4)})
    l2_bytes = read_profile_value(f"{kernel}_{workload}_traffic.txt", "l2_bytes_per_launch")
    l2_peak = read_profile_value("memory_hierarchy_probe.txt", "l2_read_gb_per_s")
    if l2_bytes and l2_peak:
        l2 = l2_bytes / avg_s / 1e9
        e.update({"l2_bytes_per_launch": round(l2_bytes), "l2_achieved_gb_per_s": round(l2, 1), "l2_peak_measured_gb_per_s": l2_peak,
                  "l2_frac": round(l2 / l2_peak, 4)})
    return e


def cpu_baseline(scene, w, threads):
    """FP64 oracle (restatement of the reference CPU path tracer) on a bounded crop at full spp."""
    from oracle import oracle as O
    o = O.OracleScene(scene)
    y0, y1 = w["cpu_rows"]
    t = time.perf_counter()
    _, st = o.render(scene.camera, w["width"], w["height"], w["spp"], seed=w["seed"], max_depth=w["max_depth"],
                     precision=64, threads=threads, crop=(0, y0, w["width"], y1))
    dt = time.perf_counter() - t
    return st, dt, (y0, y1)


def bench_render(args):
    from minirender_amd import api
    w = WORKLOADS[args.workload]
    log(f"[bench] generating workload {args.workload}")
    scene = make_scene(args.workload)
    dev = api.Device(0, schedule="plain_fill", persistent_blocks_per_cu=args.blocks_per_cu,
                     path_capacity=args.path_capacity, traversal=args.traversal, use_graph=args.graph,
                     batches_in_flight=args.batches_in_flight, hbm_budget_bytes=args.hbm_budget_gib << 30)
    info = dev.query()
    t0 = time.perf_counter()
    bvh = dev.upload(scene)
    log(f"[bench] device {info['name']} CUs={info['compute_units']}; scene {scene.triangle_count} tris uploaded in "
        f"{time.perf_counter() - t0:.2f}s; {bvh.info()['nodes']} nodes")
    rs = api.RenderSettings(w["width"], w["height"], w["spp"], w["seed"])
    ins = api.IntegratorSettings("mis", w["max_depth"], True, 3)

    def run_job():
        job = api.Job(bvh, scene.camera, rs, ins)
        job.wait(-1)
        p = job.progress()
        job.release()
        if p.state != "completed":
            raise SystemExit(f"render ended {p.state}: {p.error}")
        return p

    for i in range(args.warmup):
        t = time.perf_counter()
        run_job()
        log(f"[bench] warmup {i} {time.perf_counter() - t:.3f}s")
    dev.kernel_times(reset=True)
    times, progs = [], []
    for i in range(args.steps):
        t = time.perf_counter()
        progs.append(run_job())
        times.append(time.perf_counter() - t)
        log(f"[bench] step {i} {times[-1]:.3f}s")
    kt = dev.kernel_times()
    samples = w["width"] * w["height"] * w["spp"]
    total = sum(times)
    value = samples * args.steps / total / 1e6
    rays = kt["extension_rays"] + kt["shadow_rays"]

    result = {
        "metric": "Msamples/s at 1080p 256spp; p50/p99 job latency, served vs plain, 1 GPU",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} {w['width']}x{w['height']} {w['spp']}spp depth {w['max_depth']} "
                               f"MIS RR3 ({scene.triangle_count} triangles, plain one-job schedule, pool-filling batches)",
                   "width": w["width"], "height": w["height"], "spp": w["spp"], "max_depth": w["max_depth"],
                   "triangles": scene.triangle_count},
        "mrays_per_s": round(rays / total / 1e6, 1),
        "kernel_ms_per_step": {k: round(kt[k] / args.steps, 2) for k in
                               ("generate_ms", "extend_ms", "shade_ms", "shadow_ms", "accumulate_ms")},
    }

    bvh.release()
    if not args.no_cpu and not args.graph:
        result["config1_gpu_msamples_per_s"] = round(cornell_gpu_rate(dev), 1)
    dev.close()
    if not args.no_serving:
        result["serving"] = bench_serving(args)

    cpu = None
    if not args.no_cpu and args.graph:
        log("[bench] --graph: kernel times are not measured, so no roofline entry; CPU baseline skipped")
    elif not args.no_cpu:
        cpu, st = cpu_baselines(scene, w, value, result.get("config1_gpu_msamples_per_s"))
        fused = kt["shadow_launches"] == 0  # one traversal kernel per bounce traces both kinds of ray
        refill = fused and kt["refill_launches"] == kt["extend_launches"]
        kernel = ("k_trace_refill" if refill else "k_trace") if fused else "k_extend"
        launches = max(kt["extend_launches"], 1)
        avg_s = kt["extend_ms"] / launches / 1e3
        eb, et, sb, stt = tests_per_ray(st)
        if avg_s > 0:
            result["roofline"] = roofline_entry(kernel, args.workload, avg_s, kt["extension_rays"] / launches,
                                                (kt["shadow_rays"] if fused else 0) / launches, eb, et, sb, stt)
    result["cpu_baseline"] = cpu
    return result


def cornell_gpu_rate(dev):
    """GPU Msamples/s of the Cornell box workload (512^2, 64 spp, depth 8), one warm and one timed render."""
    from minirender_amd import api
    w = WORKLOADS["cornell"]
    scene = make_scene("cornell")
    rs, ins = api.RenderSettings(w["width"], w["height"], w["spp"], w["seed"]), api.IntegratorSettings("mis", w["max_depth"], True, 3)
    api.render(dev, scene, rs, ins)
    t = time.perf_counter()
    api.render(dev, scene, rs, ins)
    return w["width"] * w["height"] * w["spp"] / (time.perf_counter() - t) / 1e6


def cpu_baselines(scene, w, gpu_value, cornell_gpu):
    """The FP64 oracle (the reference CPU path tracer restated: same RNG streams, 16^2 tiles, 4-sample
    batches), built -march=native on this host, timed at the reference's default thread count
    min(hardware_concurrency - 1, 32) (render_session.cpp:389-392) and at the host's CPU share, on a
    fixed band of image rows at full spp; plus the whole Cornell-box image at both thread counts."""
    from oracle import oracle as O
    t = time.perf_counter()
    O.use_native()
    build_s = time.perf_counter() - t
    host = host_cpu()
    counts = sorted({host["reference_default_threads"], host["cpu_share"]})
    log(f"[bench] cpu baseline: {host['model']}, hardware_concurrency {host['hardware_concurrency']}, "
        f"share {host['cpu_share']} CPUs; threads {counts}")
    runs, st = {}, None
    for threads in counts:
        st, dt, rows = cpu_baseline(scene, w, threads)
        runs[threads] = (st["paths"] / dt / 1e6, dt)
        log(f"[bench] cpu {threads} threads: {runs[threads][0]:.3f} Msamples/s ({dt:.1f}s)")
    cornell = make_scene("cornell")
    cw = WORKLOADS["cornell"]
    oc = O.OracleScene(cornell)
    c_runs = {}
    for threads in counts:
        t = time.perf_counter()
        _, cst = oc.render(cornell.camera, cw["width"], cw["height"], cw["spp"], seed=cw["seed"], max_depth=cw["max_depth"],
                           precision=64, threads=threads)
        dt = time.perf_counter() - t
        c_runs[threads] = (cst["paths"] / dt / 1e6, dt)
    share = host["cpu_share"]
    value = runs[share][0]
    cpu = {"value": round(value, 4), "unit": "Msamples/s", "cores": share, "kind": "port",
           "sample": f"rows {rows[0]}-{rows[1] - 1} ({(rows[1] - rows[0]) * w['width']} px) of the {w['width']}x{w['height']} "
                     f"workload at {w['spp']} spp, FP64 oracle built -march=native ({build_s:.0f}s build), {runs[share][1]:.1f}s",
           "gpu_over_cpu": round(gpu_value / value, 1),
           "host": host,
           "by_threads": {str(k): {"msamples_per_s": round(v[0], 4), "seconds": round(v[1], 2)} for k, v in runs.items()},
           "gpu_over_cpu_reference_default_threads": round(gpu_value / runs[host["reference_default_threads"]][0], 1),
           "config1_cornell_512x512_64spp": {
               "by_threads": {str(k): {"msamples_per_s": round(v[0], 3), "seconds": round(v[1], 2)} for k, v in c_runs.items()},
               "gpu_msamples_per_s": cornell_gpu,
               "gpu_over_cpu": round(cornell_gpu / c_runs[share][0], 1) if cornell_gpu else None}}
    return cpu, st


def bench_serving(args):
    """Served request stream (config 5): 64 uneven jobs over 8 resident scenes, 64 GB HBM budget, 16-spp
    progressive passes, 5% of the jobs cancelled after 1-3 passes, replayed at each arrival rate of
    --serve-rates (Poisson; inf = every request at once, the saturated stream) on the plain (one job at a
    time), plain-fill and batched schedules: latency against throughput."""
    import math
    from minirender_amd import serve
    rates = [math.inf if r == "inf" else float(r) for r in args.serve_rates.split(",")]
    log(f"[bench] serving: building the 8 resident scenes; rates {rates}")
    rows = serve.sweep(rates)
    keep = ("completed", "cancelled", "failed", "samples", "makespan_s", "msamples_per_s", "first_pass_p50_s",
            "first_pass_p99_s", "completion_p50_s", "completion_p99_s")
    out = []
    for row in rows:
        r = {"rate_jobs_per_s": row["rate"] if math.isfinite(row["rate"]) else "inf"}
        for m in serve.SCHEDULES:
            r[m] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in row[m].items() if k in keep}
        r["speedup"] = round(row["speedup"], 3)
        r["speedup_vs_plain_fill"] = round(row["speedup_vs_plain_fill"], 3)
        r["last_arrival_s"] = round(row["last_arrival_s"], 4)
        r["speedup_ceiling"] = round(row["speedup_ceiling"], 3) if math.isfinite(row["rate"]) else None
        log(f"[bench] serving rate {r['rate_jobs_per_s']}: plain {r['plain']['msamples_per_s']:.1f}, plain_fill "
            f"{r['plain_fill']['msamples_per_s']:.1f}, batched {r['batched']['msamples_per_s']:.1f} Msamples/s")
        out.append(r)
    return {"stream": "64 jobs, Poisson arrivals, side 128-1024, spp 16-1024, depth 8-16, 8 scenes (36-410k tris), "
                      "5% cancelled after 1-3 passes, 16-spp passes, 2^27-path pool, 64 GB budget",
            "rates": out,
            "saturated_speedup": out[-1]["speedup"] if out and out[-1]["rate_jobs_per_s"] == "inf" else None}


def bench_intersect(args):
    """Intersection microbench: 16,777,216 seeded rays against the 1,310,720-triangle displaced icosphere, closest
    hit, rays and hits resident in HBM. Parity and per-ray algorithmic bytes come from the FP64 oracle
    on every k-th ray of the same set (--cpu-stride, default 1: all 16M rays, about 5 s on 16 cores)."""
    import ctypes

    from minirender_amd import api, scenegen
    scene = make_scene("intersect")
    n = 1 << 24
    log("[bench] intersect: generating 16M rays")
    rays = scenegen.intersection_rays(scene, n)
    dev = api.Device(0, persistent_blocks_per_cu=args.blocks_per_cu, query_precision=args.query_precision,
                     path_capacity=1 << 16)  # ray queries only: a small render pool
    bvh = dev.upload(scene)
    hip = ctypes.CDLL("libamdhip64.so")
    d_rays, d_hits = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_rays), ctypes.c_size_t(rays.nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(d_hits), ctypes.c_size_t(n * 8)) == 0
    assert hip.hipMemcpy(d_rays, rays.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(rays.nbytes), 1) == 0
    for _ in range(args.warmup):
        bvh.intersect_device(d_rays.value, n, d_hits.value)
    ms = [bvh.intersect_device(d_rays.value, n, d_hits.value) for _ in range(args.steps)]
    # host-buffer entry point (mrg_intersect): rays in and hits out over PCIe, for the PCIe-inclusive rate
    t = time.perf_counter()
    bvh.intersect(rays)
    pcie_s = time.perf_counter() - t
    hits = np.empty(n, np.dtype([("t", "<f4"), ("primitive", "<i4")]))
    assert hip.hipMemcpy(hits.ctypes.data_as(ctypes.c_void_p), d_hits, ctypes.c_size_t(n * 8), 2) == 0
    hip.hipFree(d_rays)
    hip.hipFree(d_hits)
    avg = sum(ms) / len(ms)
    value = n / avg / 1e3
    result = {"metric": "Mrays/s closest hit (intersection microbench)", "value": round(value, 1), "unit": "Mrays/s",
              "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(avg, 3),
              "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
              "dtype": "f32 traversal, f64 hit decisions" if args.query_precision == 64 else "f32", "data": "synthetic",
              "config": {"workload": f"icosphere level 8 ({scene.triangle_count} tris, 5% value-noise displacement), "
                                     f"{n} rays, closest hit", "triangles": scene.triangle_count, "rays": n,
                         "query_precision": args.query_precision},
              "pcie_inclusive_mrays_per_s": round(n / pcie_s / 1e6, 1)}
    if not args.no_cpu:
        from oracle import oracle as O
        O.use_native()
        host = host_cpu()
        threads = host["cpu_share"]
        o = O.OracleScene(scene)
        idx = np.arange(0, n, args.cpu_stride)
        sample = rays[idx].astype(np.float64)
        t = time.perf_counter()
        ref = o.intersect(sample, precision=64, second=True, threads=threads)
        dt = time.perf_counter() - t
        b = ref["stats"]["bounds_tests"] / len(idx)
        tt = ref["stats"]["triangle_tests"] / len(idx)
        with np.errstate(invalid="ignore"):
            near = np.abs(ref["t2"] - ref["t"]) <= 1e-6 * np.abs(ref["t"])
        mism = hits["primitive"][idx] != ref["prim"]
        gp, gt = hits["primitive"][idx], hits["t"][idx]
        edge = np.zeros(len(idx), bool)
        for i in np.where(mism & ~near & (gp >= 0) & (ref["prim"] >= 0))[0]:
            if abs(float(gt[i]) - ref["t"][i]) <= 1e-6 * ref["t"][i]:
                va = {tuple(v) for v in scene.positions[gp[i]].tolist()}
                edge[i] = bool(va & {tuple(v) for v in scene.positions[ref["prim"][i]].tolist()})
        both = (~mism) & (ref["prim"] >= 0)
        rel = np.abs(hits["t"][idx][both] - ref["t"][both]) / ref["t"][both]
        hit64 = ref["prim"] >= 0
        result["parity"] = {"sample_rays": int(len(idx)), "id_mismatch": int(mism.sum()),
                            "near_ties": int(near.sum()), "id_mismatch_outside_near_ties": int((mism & ~near).sum()),
                            "of_which_shared_edge_crossings": int(edge.sum()),
                            "max_t_rel": float(rel.max(initial=0)),
                            "t_equal_to_fp64_rounded": int((hits["t"][idx][hit64] == ref["t"][hit64].astype(np.float32)).sum()),
                            "hits": int(hit64.sum())}
        kernel = "k_query_exact" if args.query_precision == 64 else "k_query"
        result["roofline"] = roofline_entry(kernel, "intersect", avg / 1e3, n, 0, b, tt, 0, 0)
        result["cpu_baseline"] = {"value": round(len(idx) / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads,
                                  "kind": "port", "host": host,
                                  "sample": (f"all {len(idx)} rays" if args.cpu_stride == 1 else
                                             f"every {args.cpu_stride}-th ray ({len(idx)} rays)") + ", FP64 oracle (-march=native) "
                                  f"closest + second hit, {dt:.1f}s"}
    bvh.release()
    dev.close()
    return result


def bench_convergence(args):
    """RMSE of 64-spp Cornell renders (GPU FP32, and the FP64 oracle) against an FP64 render at
    65,536 spp of a 32x32 crop (Cornell 512^2 camera, depth 8, RR 3, seed 0)."""
    from minirender_amd import api, scenegen
    from oracle import oracle as O
    w = WORKLOADS["cornell"]
    scene = scenegen.cornell_box()
    crop = (240, 200, 272, 232)
    threads = host_cpu()["cpu_share"]
    o = O.OracleScene(scene)
    t = time.perf_counter()
    ref, _ = o.render(scene.camera, w["width"], w["height"], 65536, seed=99, max_depth=8, precision=64,
                      threads=threads, crop=crop)
    t_ref = time.perf_counter() - t
    cpu64, _ = o.render(scene.camera, w["width"], w["height"], w["spp"], seed=0, max_depth=8, precision=64,
                        threads=threads, crop=crop)
    with api.Device(0) as dev:
        res = {}
        for spp in (w["spp"], 1024, 16384):
            img, _ = api.render(dev, scene, api.RenderSettings(w["width"], w["height"], spp, 0),
                                api.IntegratorSettings("mis", 8, True, 3))
            res[spp] = img[crop[1]:crop[3], crop[0]:crop[2]]
    rmse = lambda a: float(np.sqrt(np.mean((a.astype(np.float64) - ref.astype(np.float64)) ** 2)))
    out = {"metric": "RMSE vs FP64 65536-spp reference (linear RGB)", "crop": crop,
           "reference_seconds": round(t_ref, 1), "reference_threads": threads,
           "rmse_gpu_fp32_64spp": rmse(res[w["spp"]]), "rmse_cpu_fp64_64spp": rmse(cpu64),
           "rmse_gpu_fp32_1024spp": rmse(res[1024]), "rmse_gpu_fp32_16384spp": rmse(res[16384]),
           "mean_reference": float(ref.mean())}
    out["gpu_over_cpu_rmse_64spp"] = out["rmse_gpu_fp32_64spp"] / out["rmse_cpu_fp64_64spp"]
    return out


def bench_batch(args):
    """Batched uneven jobs: 64 jobs (sides log-uniform 128-1024, spp log-uniform 16-1024, depth 8-16)
    over 8 resident scenes (36 to ~410k triangles, seed 4), all submitted at once. The autotuner
    first picks each schedule's device options on a calibration copy of the job set (spp / 8), then
    the full set is rendered on the plain (one job at a time), plain-fill and batched schedules."""
    from minirender_amd import autotune, serve
    scenes = serve.scene_pool(seed=4)
    jobs = serve.batch_jobs(seed=4)
    calib = serve.batch_jobs(seed=4, spp_divisor=8)
    knobs = {"path_capacity": [1 << 23, 1 << 24, 1 << 25, 1 << 26, 1 << 27, 1 << 28], "persistent_blocks_per_cu": [0, 4, 6],
             "traversal": ["split", "fused", "refill", "auto"], "use_graph": [False, True], "max_jobs_in_flight": [16, 64],
             "batches_in_flight": [1, 2]}
    tuned = {}
    for mode in ("plain", "plain_fill", "batched"):
        t = time.perf_counter()
        tuned[mode] = autotune.tune(scenes, calib, schedule=mode, knobs=knobs, repeats=1, log=log)
        log(f"[bench] autotune {mode}: {tuned[mode]['options']} in {time.perf_counter() - t:.0f}s")
    out = {}
    for mode in ("plain", "plain_fill", "batched"):
        value, secs = autotune.run_workload(tuned[mode]["options"], mode, scenes, jobs, repeats=args.steps)
        out[mode] = {"msamples_per_s": round(value / 1e6, 1), "seconds": [round(x, 3) for x in secs],
                     "options": tuned[mode]["options"],
                     "calibration_trials": len(tuned[mode]["trials"])}
        log(f"[bench] batch {mode}: {value / 1e6:.1f} Msamples/s")
    total = sum(r.width * r.height * r.spp for r in jobs)
    return {"metric": "Msamples/s, 64 uneven jobs submitted at once (batched vs plain)",
            "value": out["batched"]["msamples_per_s"], "unit": "Msamples/s", "n_gpus": 1, "steps": args.steps,
            "higher_is_better": True, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"64 jobs, sides 128-1024, spp 16-1024, depth 8-16, 8 scenes, {total} samples"},
            "schedules": out,
            "speedup_vs_plain": round(out["batched"]["msamples_per_s"] / out["plain"]["msamples_per_s"], 3),
            "speedup_vs_plain_fill": round(out["batched"]["msamples_per_s"] / out["plain_fill"]["msamples_per_s"], 3)}


def bench_copy(args):
    """Plain 1 GiB device-to-device copy: the known-size kernel that scales rocprofv3's FETCH_SIZE and
    WRITE_SIZE readings (KiB per dispatch) into bytes for the traffic figures of the path kernels."""
    import torch
    n = 1 << 28
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(max(args.steps, 1)):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / max(args.steps, 1)
    return {"metric": "device copy GB/s (read + write)", "value": round(2 * a.numel() * 4 / dt / 1e9, 1), "unit": "GB/s",
            "bytes_read_per_dispatch": a.numel() * 4, "bytes_written_per_dispatch": a.numel() * 4}


def bench_cache(args):
    """Read bandwidth of each level of the memory hierarchy (libmrg_probe.so): a grid of 8 256-thread
    blocks per CU streams a resident buffer whose size picks the level serving the reads (a few MB:
    the per-XCD L2; tens of MB: the Infinity Cache; GBs: HBM). The roofline of the traversal kernel
    is quoted against these measured figures as well as against the 8 TB/s HBM peak."""
    import ctypes as C
    from minirender_amd import build as B
    lib = C.CDLL(B.PROBE_LIB)
    lib.mrg_probe_read_bandwidth.restype = C.c_int
    lib.mrg_probe_read_bandwidth.argtypes = [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]
    per_cu, iters = 8, 1024  # 8.6 GB read per launch on 256 CUs
    levels = []
    for label, size in (("l2", 2 << 20), ("l2", 3 << 20), ("infinity_cache", 64 << 20), ("infinity_cache", 128 << 20),
                        ("hbm", 4 << 30)):
        gbps, ms, nbytes = C.c_double(), C.c_double(), C.c_uint64()
        rc = lib.mrg_probe_read_bandwidth(0, size, per_cu, iters, max(args.steps, 3), C.byref(gbps), C.byref(ms),
                                          C.byref(nbytes))
        if rc != 0:
            raise SystemExit(f"probe failed with HIP error {rc}")
        log(f"[bench] cache probe {size >> 20} MiB: {gbps.value:.0f} GB/s ({ms.value:.3f} ms per launch)")
        levels.append({"level": label, "bytes": size, "gb_per_s": round(gbps.value, 1), "ms_per_launch": round(ms.value, 4),
                       "bytes_read_per_launch": nbytes.value})
    return {"metric": "read bandwidth by memory level (GB/s)", "value": max(l["gb_per_s"] for l in levels), "unit": "GB/s",
            "blocks_per_cu": per_cu, "threads_per_block": 256, "float4_loads_per_thread": iters, "levels": levels}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed", choices=["mixed", "cornell", "intersect", "convergence", "copy", "batch",
                                                                 "cache"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--serve-rates", default="40,80,160,320,inf",
                    help="arrival rates (jobs/s) of the served stream, comma-separated; inf = all at once")
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--path-capacity", type=int, default=0, help="paths in flight per batch (0 = library default)")
    ap.add_argument("--batches-in-flight", type=int, default=0, help="batches rendering at once (0 = library default)")
    ap.add_argument("--hbm-budget-gib", type=int, default=0, help="device admission budget (0 = library default)")
    ap.add_argument("--query-precision", type=int, default=64, choices=[32, 64],
                    help="intersect workload: hit decisions in FP64 (the reference intersector's) or the render kernels' FP32")
    ap.add_argument("--traversal", default="auto", choices=["auto", "split", "fused", "refill"])
    ap.add_argument("--cpu-stride", type=int, default=1,
                    help="intersect workload: the FP64 oracle checks and times every k-th ray")
    ap.add_argument("--graph", action="store_true", help="capture the bounce loop in a hipGraph (no per-kernel times)")
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel (calibration runs only)")
    args = ap.parse_args()
    if args.spp:
        WORKLOADS[args.workload]["spp"] = args.spp
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs the path on a single MI355X; samples "
              "shard by pass across GPUs with one film reduce, which this build does not run; use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "intersect":
        res = bench_intersect(args)
    elif args.workload == "batch":
        res = bench_batch(args)
    elif args.workload == "copy":
        res = bench_copy(args)
    elif args.workload == "cache":
        res = bench_cache(args)
    elif args.workload == "convergence":
        res = bench_convergence(args)
    else:
        res = bench_render(args)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
