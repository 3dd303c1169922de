"""Training and evaluation of the learned denoiser on data the renderer produces.

Training set: frames of the batched dataset configuration's scenes (sun/sky spheres, diffuse /
conductor / dielectric / coated materials), each rendered twice in one persistent batch launch:
a 4-sample noisy frame with its 4-sample first-hit albedo and normal, and a reference frame of
the same camera and seed with ref_spp samples.  Training: random 256x256 crops, batch 16,
Adam, L1 loss of log(1 + colour) (the kernels of csrc/denoise.hip).

Evaluation: relMSE = mean over pixels and channels of (x - y)^2 / (y^2 + 1e-2) against a
high-sample reference of the 1920x1080 sun/sky workload (20,971,522 triangles, depth 8), for
plain rendering at increasing sample counts and for the denoised 4-, 8- and 16-sample images;
the time to reach relMSE 1e-3 is interpolated in log-log space from those points.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
for _p in (HERE, os.path.join(HERE, "scenes")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np

import dataset as D
import protoray_hip as P
import scenegen as G


def training_set(n_frames=24, size=(384, 384), ref_spp=4096, n_scenes=4, max_tris=1_000_000, seed=11, device=0, log=print):
    scenes = D.dataset_scenes(n_scenes, max_tris=max_tris)
    devs = []
    for sc in scenes:
        d = P.DeviceHip(device)
        P.load_scene(d, sc)
        devs.append(d)
    rng = np.random.default_rng(seed)
    batch = P.BatchHip(device)
    W, H = size
    for i in range(n_frames):
        pos = (rng.uniform(-6, 6), rng.uniform(0.5, 5), rng.uniform(-4, 2))
        tgt = (rng.uniform(-6, 6), rng.uniform(0, 4), rng.uniform(12, 24))
        cam = D.look_at(pos, tgt, fov_deg=float(rng.uniform(35, 60)), lens_radius=float(rng.choice([0.0, 0.05])),
                        focal=float(np.linalg.norm(np.subtract(tgt, pos))))
        d = devs[i % n_scenes]
        batch.addFrame(d, cam, W, H, 1000 + i, 4, aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1)  # noisy input
        batch.addFrame(d, cam, W, H, 1000 + i, ref_spp)                                   # reference
    t0 = time.perf_counter()
    r = batch.render()
    log(f"training set: {n_frames} frames {W}x{H}, 4 spp + {ref_spp} spp references, {r['samples'] / 1e9:.2f} G samples "
        f"in {r['ms'] / 1e3:.1f} s ({time.perf_counter() - t0:.1f} s wall)")
    inputs = np.zeros((n_frames, H, W, 9), np.float32)
    targets = np.zeros((n_frames, H, W, 3), np.float32)
    for i in range(n_frames):
        inputs[i, ..., 0:3] = batch.get(2 * i)
        inputs[i, ..., 3:6] = batch.get(2 * i, P.BATCH_ALBEDO1)
        inputs[i, ..., 6:9] = batch.get(2 * i, P.BATCH_NORMAL1)
        targets[i] = batch.get(2 * i + 1)
    batch.close()
    for d in devs:
        d.close()
    return inputs, targets, r


# Input sets of the filter (Denoiser::Params with and without the auxiliary images): the channels of
# the 9-channel input each one keeps, and the weights file trained for it.
INPUT_SETS = {"full": 9, "color_albedo": 6, "color": 3}


def weights_path(inputs="full", directory=os.path.join(HERE, "weights")):
    suffix = "" if inputs == "full" else "_" + inputs
    return os.path.join(directory, "unet_sunsky%s.npy" % suffix)


def train(dn, inputs, targets, steps=1500, batch=16, crop=256, lr=1e-3, seed=5, log=print, input_set="full"):
    """input_set: the auxiliary channels beyond it are zero in every batch (the filter's input
    when the caller gives no albedo / normal image)."""
    keep = INPUT_SETS[input_set]
    rng = np.random.default_rng(seed)
    n, H, W, _ = inputs.shape
    losses, dev_ms = [], 0.0
    t0 = time.perf_counter()
    for step in range(steps):
        idx = rng.integers(0, n, batch)
        ys = rng.integers(0, H - crop + 1, batch)
        xs = rng.integers(0, W - crop + 1, batch)
        bi = np.stack([inputs[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)])
        bi[..., keep:] = 0.0
        bt = np.stack([targets[i, y:y + crop, x:x + crop] for i, y, x in zip(idx, ys, xs)])
        rate = lr * 0.5 * (1 + math.cos(math.pi * step / steps))  # cosine decay
        loss, ms = dn.trainStep(bi, bt, rate)
        losses.append(loss)
        dev_ms += ms
        if step % 250 == 0 or step == steps - 1:
            log(f"step {step}: loss {np.mean(losses[-50:]):.4f}, {dev_ms / (step + 1):.1f} ms per step on the device")
    return {"steps": steps, "final_loss": float(np.mean(losses[-50:])), "first_loss": float(losses[0]),
            "device_ms_per_step": dev_ms / steps, "wall_s": time.perf_counter() - t0}


def relmse(x, ref):
    return G.relmse(x, ref)


def evaluate(dn, ref_spp=4096, plain_spp=(4, 16, 64, 256, 1024), denoised_spp=(4, 8, 16), device=0, log=print):
    """relMSE against a ref_spp render of the 1080p sun/sky workload for plain and denoised images.

    The reference is rendered in two halves of ref_spp / 2 samples (A: sample indices 0 .. n/2-1,
    then the frame continues to n samples); B = 2 * full - A is the mean of the second half.  Each
    half is an independent estimate, so mean((A - B)^2 / (ref^2 + 1e-2)) / 4 estimates the
    reference's own noise term, which every relMSE against it contains in expectation:
    E[relMSE(x, ref)] = relMSE(x, truth) + that term.  relmse_corrected subtracts it.

    Denoised points run the denoiser on the frame where the renderer left it (prt_denoiser_
    execute_frame / execute_batch): "ms" is render + denoise wall time, "ms_device" render +
    denoiser kernels."""
    sc = G.spheres_sunsky()
    dev = P.DeviceHip(device)
    P.load_scene(dev, sc)
    W, H = sc["width"], sc["height"]
    dev.initFrame(W, H, P.AOV_ALBEDO1 | P.AOV_NORMAL1)
    # reference: sample indices of a different seed than the evaluated images
    dev.initSampler(7)
    dev.clearFrame()
    t0 = time.perf_counter()
    half = ref_spp // 2
    for _ in range(half // 64):
        dev.render(64)
    ref_a = dev.blitFrame()
    for _ in range((ref_spp - half) // 64):
        dev.render(64)
    ref = dev.blitFrame()
    ref_b = 2.0 * ref - ref_a
    ref_noise = float(np.mean((ref_a - ref_b) ** 2 / (ref ** 2 + 1e-2)) / 4.0)
    log(f"reference {ref_spp} spp in {time.perf_counter() - t0:.1f} s; its noise term {ref_noise:.5f}")
    res = {"plain": [], "denoised": []}
    dev.initSampler(0)
    dev.clearFrame()
    done = 0
    want = sorted(set(plain_spp) | set(denoised_spp))
    for spp in want:
        dev.synchronize()
        dev.render(spp - done)
        dev.synchronize()
        done = spp
        img = dev.blitFrame()
        if spp in plain_spp:
            res["plain"].append({"spp": spp, "relmse": relmse(img, ref)})
        if spp in denoised_spp:
            secs = dn.executeFrame(dev)
            out = dn.output()
            res["denoised"].append({"spp": spp, "relmse": relmse(out, ref), "denoise_s": secs, "denoise_kernel_ms": dn.kernelMs(),
                                    "noisy_relmse": relmse(img, ref)})
    # the batched form: the same 4-sample frame (seed 0) from the persistent batch renderer in one
    # launch (it equals the wavefront renderer's image), then the denoiser on the batch's output
    batch = P.BatchHip(device)
    batch.addFrame(dev, sc["camera"], W, H, 0, 4, aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1)
    batch.render()  # first launch builds the graph
    br = batch.render()
    bsecs = dn.executeBatch(batch, 0)
    bout = dn.output()
    batch.close()
    res["denoised_batched"] = {"spp": 4, "render_ms": br["ms"], "denoise_s": bsecs, "denoise_kernel_ms": dn.kernelMs(),
                               "relmse": relmse(bout, ref), "ms": br["ms"] + bsecs * 1e3,
                               "ms_device": br["ms"] + dn.kernelMs()}
    # time per sample pass of this frame
    dev.synchronize()
    t0 = time.perf_counter()
    dev.render(16)
    dev.synchronize()
    ms_per_spp = (time.perf_counter() - t0) * 1e3 / 16
    dev.close()
    for r in res["plain"]:
        r["ms"] = r["spp"] * ms_per_spp
    for r in res["denoised"]:
        r["ms"] = r["spp"] * ms_per_spp + r["denoise_s"] * 1e3
        r["ms_device"] = r["spp"] * ms_per_spp + r["denoise_kernel_ms"]
    for r in res["plain"] + res["denoised"] + [res["denoised_batched"]]:
        r["relmse_corrected"] = r["relmse"] - ref_noise
    res["ms_per_spp"] = ms_per_spp
    res["ref_spp"] = ref_spp
    res["reference_noise_relmse"] = ref_noise
    res["time_to_relmse_1e-3_ms"] = {"plain": time_to(res["plain"], 1e-3), "denoised": time_to(res["denoised"], 1e-3)}
    corrected = lambda pts: [{"ms": r["ms"], "relmse": r["relmse_corrected"]} for r in pts if r["relmse_corrected"] > 0]
    res["time_to_relmse_1e-3_ms_corrected"] = {"plain": time_to(corrected(res["plain"]), 1e-3),
                                               "denoised": time_to(corrected(res["denoised"]), 1e-3)}
    log(json.dumps(res))
    return res


def time_to(points, target):
    """Time at which relMSE reaches target, interpolated in log-log space between the measured
    points; beyond them, extrapolated from the last two (marked) while the curve still falls at
    least as t^-1/2, else marked not reached."""
    pts = sorted((p["ms"], p["relmse"]) for p in points)
    for (t0, e0), (t1, e1) in zip(pts, pts[1:]):
        if e0 > target >= e1:
            a = (math.log(target) - math.log(e0)) / (math.log(e1) - math.log(e0))
            return math.exp(math.log(t0) + a * (math.log(t1) - math.log(t0)))
    if pts and pts[0][1] <= target:
        return pts[0][0]
    if len(pts) >= 2 and pts[-1][1] < pts[-2][1]:
        (t0, e0), (t1, e1) = pts[-2], pts[-1]
        slope = (math.log(e1) - math.log(e0)) / (math.log(t1) - math.log(t0))
        # extrapolate only a curve still falling at least as fast as t^-1/2 (Monte Carlo noise
        # falls as t^-1); a flattening curve (bias, or the reference's own noise) is not extrapolated
        if slope <= -0.5:
            return {"extrapolated": math.exp(math.log(t1) + (math.log(target) - math.log(e1)) / slope)}
        return {"not_reached": "the curve flattens (slope %.2f in log-log)" % slope}
    return None


def kernel_bench(inferences=4, train_steps=6, width=1920, height=1080, batch=16, crop=256, log=print):
    """Device time of the denoiser's kernels on synthetic inputs: 1080p inferences with the
    committed weights, then training steps on batch x crop x crop (the training configuration)."""
    dn = P.DenoiserHip(0)
    dn.setWeights(np.load(os.path.join(HERE, "weights", "unet_sunsky.npy")))
    rng = np.random.default_rng(0)
    col, alb, nrm = (rng.random((height, width, 3), np.float32) for _ in range(3))
    res = {"inference_ms": [], "train_step_ms": []}
    for _ in range(inferences):
        dn.denoise(col, alb, nrm)
        res["inference_ms"].append(dn.kernelMs())
    inp = rng.random((batch, crop, crop, 9), np.float32)
    tgt = rng.random((batch, crop, crop, 3), np.float32)
    for _ in range(train_steps):
        res["train_step_ms"].append(dn.trainStep(inp, tgt, 1e-4)[1])
    dn.close()
    log(json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser(description="train the denoiser on rendered data and evaluate relMSE at 1080p")
    ap.add_argument("--kernel-bench", action="store_true",
                    help="only time the kernels on synthetic inputs (1080p inference, 16 x 256^2 training steps)")
    ap.add_argument("--frames", type=int, default=24)
    ap.add_argument("--train-ref-spp", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--eval-ref-spp", type=int, default=4096)
    ap.add_argument("--inputs", default="full",
                    help="comma-separated input sets to train for (full, color_albedo, color; all: every set), "
                         "each from the same rendered training set")
    ap.add_argument("--weights-dir", default=os.path.join(HERE, "weights"))
    ap.add_argument("--no-eval", action="store_true", help="skip the 1080p relMSE evaluation")
    ap.add_argument("--json-out", default=None)
    args = ap.parse_args()
    log = lambda m: print(m, flush=True)
    if args.kernel_bench:
        kernel_bench(log=log)
        return
    inputs, targets, r = training_set(args.frames, ref_spp=args.train_ref_spp, log=log)
    sets = sorted(INPUT_SETS) if args.inputs == "all" else args.inputs.split(",")
    for name in sets:
        if name not in INPUT_SETS:
            raise SystemExit(f"unknown input set {name}")
    out = {"training_data": {"frames": args.frames, "ref_spp": args.train_ref_spp, "render_ms": r["ms"], "samples": r["samples"]},
           "training": {}}
    dn = P.DenoiserHip(0)
    for name in sets:
        dn.setWeights(dn.initialWeights)
        out["training"][name] = train(dn, inputs, targets, steps=args.steps, log=log, input_set=name)
        os.makedirs(args.weights_dir, exist_ok=True)
        np.save(weights_path(name, args.weights_dir), dn.getWeights())
    if not args.no_eval:
        dn.setWeights(np.load(weights_path("full", args.weights_dir)))
        out["evaluation"] = evaluate(dn, ref_spp=args.eval_ref_spp, log=log)
    if args.json_out:
        with open(args.json_out, "w") as f:
            json.dump(out, f, indent=1)
    log(json.dumps(out))
    dn.close()


if __name__ == "__main__":
    main()
