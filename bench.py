"""Benchmark: path-traced Msamples/s of the stream path tracer on one MI355X.

Workload (configuration 1 of the stream renderer's benchmark set): 1920x1080, 64 spp, max
depth 8, thin-lens camera, 4096 icospheres at subdivision 4 (20,971,522 triangles) with
diffuse / rough-conductor / rough-dielectric materials, lit by a sun/sky plus a distant light;
Sobol-Burley seed 0.  One step = one sample pass (one sample for every pixel, one
Renderer::render call, render/path_renderer_stream.h:150-187); --steps 64 renders the full
frame.  Scene generation, upload and the BVH build happen before the timed region (the BVH
build time is reported separately, as Embree's buildMs is).

Output: one JSON line with the metric, the roofline of the dominant kernel (algorithmic bytes
from the kernels' own counters over the HIP-event time of that kernel), a CPU baseline (the
oracle, the project's C++ restatement of the reference path, timed on the host cores on a
bounded sample of the same frame) and the time-to-relMSE comparison of plain rendering against
the learned denoiser on the same frame ("relmse"; --no-relmse skips it).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("oidn-protoray_amd", os.path.join("oidn-protoray_amd", "scenes")):
    sys.path.insert(0, os.path.join(ROOT, sub))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# Algorithmic bytes per unit of work (fixed by the queue records the kernels read and write).
NODE_BYTES = 128       # one 4-wide BVH node: four child boxes + child references
TRI_BYTES = 48         # one triangle record: v0, e1, e2 (+ primitive ID)
RAY_BYTES_CLOSEST = 4 + 32 + 16   # queue entry, ray (origin+tnear, direction), hit (t, u, v, prim)
RAY_BYTES_SHADOW = 4 + 32 + 4 + 16 + 16  # queue entry, shadow ray, flags, radiance read + direct light


def _profile_value(fname, key):
    path = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + " ="):
            return float(line.split("=")[1])
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary (FETCH_SIZE and
    WRITE_SIZE, calibrated against a copy kernel of known size), or None."""
    return _profile_value("traffic_spheres_sunsky_1080p.txt",
                          "traffic_bytes_per_launch.k" + kernel[0].upper() + kernel[1:])


def measured_l2_hit_rate(kernel):
    """L2 hit rate of `kernel` (TCC_HIT / (TCC_HIT + TCC_MISS)) from the committed rocprofv3 PMC
    summary, or None."""
    return _profile_value("l2_spheres_sunsky_1080p.txt", "l2_hit_rate.k" + kernel[0].upper() + kernel[1:])


def host_cpu():
    """(threads to use, CPUs in this process's affinity mask, logical CPUs of the machine, model).
    The threads are every CPU the process may run on, capped by OMP_NUM_THREADS when the host sets
    it (the share of the machine a job on a shared GPU host gets)."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    threads = min(affinity, int(share)) if share and share.isdigit() and int(share) > 0 else affinity
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return threads, affinity, os.cpu_count(), model


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=64)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample time")
    p.add_argument("--small", action="store_true", help="tiny scene, for profiling smoke runs only")
    p.add_argument("--no-relmse", action="store_true", help="skip the time-to-relMSE comparison (plain vs denoised)")
    p.add_argument("--relmse-ref-spp", type=int, default=4096, help="samples of the relMSE reference render")
    p.add_argument("--copy-calibration-mib", type=int, default=0,
                   help="before the timed region, run one device-to-device copy of this many MiB (a copy "
                        "kernel of known traffic that calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def cpu_baseline(scene, target_s):
    """Times the oracle (its AVX2 + FMA build, as the reference's host build) on a bounded band of
    rows of the same frame, on every host core the process may use."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.select_library("liboracle_bench.so")
    threads, affinity, logical, model = host_cpu()
    W, H = scene["width"], scene["height"]
    t0 = time.time()
    osc = O.Scene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"])
    osc.set_lights(scene["lights"], scene.get("env"))
    build_s = time.time() - t0
    rows = (H // 2 - 16, H // 2 + 16)
    # calibrate with one pass, then size the sample to about target_s seconds
    t0 = time.time()
    osc.render(scene["camera"], W, H, spp_start=0, spp_count=1, max_depth=scene["max_depth"], threads=threads, rows=rows)
    t1 = time.time() - t0
    # budget in units of one sample over the 32-row band: all of the workload's samples first,
    # then a taller band (rows in steps of 32 around the middle of the frame)
    budget = target_s / max(t1, 1e-3)
    spp = max(1, min(scene["spp"], int(budget)))
    if budget >= 2 * scene["spp"]:
        band = min(H, 32 * int(budget // scene["spp"]))
        rows = (max(0, H // 2 - band // 2), max(0, H // 2 - band // 2) + band)
    t0 = time.time()
    _, rays = osc.render(scene["camera"], W, H, spp_start=0, spp_count=spp, max_depth=scene["max_depth"],
                         threads=threads, rows=rows)
    dt = time.time() - t0
    samples = (rows[1] - rows[0]) * W * spp
    return {"value": samples / dt / 1e6, "unit": "Msamples/s", "cores": threads, "kind": "port",
            "cores_in_affinity": affinity, "logical_cpus": logical, "cpu_model": model,
            "build": "oracle/liboracle_bench.so (-O3 -march=x86-64-v3, FMA contraction)",
            "sample": f"rows {rows[0]}-{rows[1]} of the {W}x{H} frame x {spp} spp ({samples} samples, "
                      f"{dt:.1f} s, {rays / dt / 1e6:.2f} Mray/s, {threads} threads); CPU BVH build "
                      f"{build_s:.1f} s not timed"}


def relmse_leg(ref_spp):
    """Second half of the metric: relMSE against a ref_spp render of the same frame (a different
    seed) for plain rendering at 4..1024 samples and for the learned denoiser (committed weights,
    trained on 4-sample / 4096-sample pairs the batch renderer produced) on 4, 8 and 16 samples,
    and for the batched form (the 4-sample frame from the persistent batch renderer, then the
    denoiser), with the render (+ denoise) time of each point — "ms" wall time with the denoiser
    reading the frame in HBM, "ms_device" with its kernels only — and the time to relMSE 1e-3
    where reached.  The reference's own noise term (from its two independent halves) is reported,
    and relmse_corrected subtracts it."""
    import protoray_hip as P
    import denoise_train as T
    dn = P.DenoiserHip(0)
    dn.setWeights(np.load(os.path.join(ROOT, "oidn-protoray_amd", "weights", "unet_sunsky.npy")))
    ev = T.evaluate(dn, ref_spp=ref_spp, log=lambda m: print(m, file=sys.stderr, flush=True))
    dn.close()
    best_dn = min(ev["denoised"], key=lambda r: r["relmse"])
    first_dn = ev["denoised"][0]
    return {
        "definition": "mean over pixels and channels of (x - y)^2 / (y^2 + 1e-2), y = reference",
        "reference_spp": ref_spp,
        "ms_per_spp": round(ev["ms_per_spp"], 3),
        "reference_noise_relmse": round(ev["reference_noise_relmse"], 5),
        "plain": [{"spp": r["spp"], "ms": round(r["ms"], 1), "relmse": round(r["relmse"], 5),
                   "relmse_corrected": round(r["relmse_corrected"], 5)} for r in ev["plain"]],
        "denoised": [{"spp": r["spp"], "ms": round(r["ms"], 1), "ms_device": round(r["ms_device"], 1),
                      "relmse": round(r["relmse"], 5), "relmse_corrected": round(r["relmse_corrected"], 5),
                      "denoise_kernel_ms": round(r["denoise_kernel_ms"], 3)}
                     for r in ev["denoised"]],
        "denoised_batched": {k: (round(v, 5) if k.startswith("relmse") else round(v, 3) if isinstance(v, float) else v)
                             for k, v in ev["denoised_batched"].items()},
        "time_to_relmse_1e-3_ms": ev["time_to_relmse_1e-3_ms"],
        "time_to_relmse_1e-3_ms_corrected": ev["time_to_relmse_1e-3_ms_corrected"],
        "time_to_denoised_4spp_quality_ms": {"denoised": round(first_dn["ms"], 1),
                                             "plain": T.time_to(ev["plain"], first_dn["relmse"])},
        "best_denoised_relmse": round(best_dn["relmse"], 5),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only: --gpus {args.gpus} is not supported (the path shards by "
              f"image tile across GPUs by design, but no multi-GPU run is part of this build)", file=sys.stderr)
        return 2
    import protoray_hip as P
    import scenegen as G

    if args.small:
        scene = G.spheres_sunsky(n_spheres=64, subdiv=2, width=256, height=128)
    else:
        scene = G.spheres_sunsky()
    W, H = scene["width"], scene["height"]
    ntris = int(scene["indices"].shape[0])
    dev = P.DeviceHip(0)
    info = dev.getInfo()
    build_ms = dev.initScene(scene["positions"], scene["indices"], scene["tri_material"], scene["materials"])
    dev.initLights(scene["lights"], scene["env"])
    dev.initRenderer(maxDepth=scene["max_depth"])
    dev.initCamera(scene["camera"])
    dev.initFrame(W, H)
    bvh = dev.bvhStats()

    if args.copy_calibration_mib > 0:
        copy_ms = dev.copyCalibration(args.copy_calibration_mib << 20)
        print(f"copy calibration: {args.copy_calibration_mib} MiB in {copy_ms:.3f} ms "
              f"({2 * (args.copy_calibration_mib << 20) / copy_ms / 1e6:.0f} GB/s read+write)", file=sys.stderr)
    dev.initSampler(0)
    for _ in range(args.warmup):
        dev.render(1)
    # timed region: sample indices 0 .. steps-1 of seed 0 into a cleared frame
    dev.initSampler(0)
    dev.clearFrame()
    dev.synchronize()
    t0 = time.perf_counter()
    st = dev.render(args.steps)
    dev.synchronize()
    elapsed = time.perf_counter() - t0
    kt = dev.kernelTimes()
    img = dev.blitFrame()
    # SIMD utilisation of the traversal: two more passes with the counting builds of the kernels
    dev.countSteps(True)
    dev.render(2)
    steps = dev.traceSteps()
    dev.countSteps(False)

    samples = args.steps * W * H
    value = samples / elapsed / 1e6
    # roofline of the dominant kernel (bytes per launch / mean launch time == total / total)
    kernels = {
        "intersect": (kt["intersect_ms"], kt["rays"] * RAY_BYTES_CLOSEST + kt["node_visits"] * NODE_BYTES
                      + kt["tri_tests"] * TRI_BYTES, kt["intersect_launches"]),
        "occluded": (kt["occluded_ms"], kt["shadow_rays"] * RAY_BYTES_SHADOW + kt["shadow_node_visits"] * NODE_BYTES
                     + kt["shadow_tri_tests"] * TRI_BYTES, kt["occluded_launches"]),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    ms, nbytes, launches = kernels[dom]
    achieved = nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    traffic = measured_traffic(dom)
    mean_launch_ms = ms / max(launches, 1)
    # HBM bandwidth the kernel's measured traffic takes; below half the peak the kernel waits on
    # dependent loads rather than on bytes
    hbm_gbs = traffic / (mean_launch_ms * 1e-3) / 1e9 if traffic and mean_launch_ms > 0 else None
    bound = "hbm" if hbm_gbs is not None and hbm_gbs >= 0.5 * HBM_PEAK_GBS else "latency" if hbm_gbs else "hbm"
    lanes = {k: round(steps[k + "_lane_steps"] / (64.0 * steps[k + "_wave_steps"]), 4)
             for k in ("closest", "shadow") if steps[k + "_wave_steps"] > 0}
    result = {
        "metric": "path-traced Msamples/s (1080p, depth 8)",
        "value": round(value, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"spheres_sunsky {W}x{H}, {args.steps} spp of 64, max depth {scene['max_depth']}, "
                        f"{ntris} triangles, thin lens r=0.05 f=15, sun/sky (35 deg, T=3) + distant light, seed 0",
            "device": info,
            "bvh_build_ms": round(build_ms, 1),
            "bvh_max_depth": bvh["max_depth"],
            "mray_per_s": round(st["rays"] / elapsed / 1e6, 1),
            "rays": int(st["rays"]),
            "kernel_ms": {k: round(kt[k + "_ms"], 2) for k in ("generate", "intersect", "shade", "occluded")},
            "nodes_per_ray": round(kt["node_visits"] / max(kt["rays"], 1), 2),
            "tris_per_ray": round(kt["tri_tests"] / max(kt["rays"], 1), 2),
            "traversal_active_lane_fraction": lanes,
            "image_mean": [round(float(x), 5) for x in img.reshape(-1, 3).mean(axis=0)],
        },
        "roofline": {
            "kernel": dom,
            "bound": bound,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "hbm_gbs": None if hbm_gbs is None else round(hbm_gbs, 1),
            "hbm_frac": None if hbm_gbs is None else round(hbm_gbs / HBM_PEAK_GBS, 4),
            "l2_hit_rate": measured_l2_hit_rate(dom),
            "active_lane_fraction": lanes.get("closest" if dom == "intersect" else "shadow"),
            "algorithmic_bytes_per_launch": round(nbytes / max(launches, 1)),
            "launches": int(launches),
            "mean_launch_ms": round(mean_launch_ms, 4),
        },
    }
    dev.close()
    if not args.no_relmse and not args.small:
        result["relmse"] = relmse_leg(args.relmse_ref_spp)
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(scene, args.cpu_seconds)
        ratio = value / result["cpu_baseline"]["value"]
        result["cpu_baseline"]["gpu_over_cpu"] = round(ratio, 1)
        result["cpu_baseline"]["target_gpu_over_cpu"] = 20.0
        result["cpu_baseline"]["meets_target"] = bool(ratio >= 20.0)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
