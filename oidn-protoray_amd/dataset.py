"""Dataset writer and batch driver for denoiser training data.

Files per frame follow the reference's batch output (main/render/render_window.cpp:1291-1444,
image/image.cpp:81-115, image/image_io.cpp:225-250): OpenEXR images with ZIP compression,
colour and AOVs as FP16 RGB, depth as FP32 Y, and a JSON file of camera and exposure metadata
with the reference's keys.  Names follow the OIDN training layout:
    <name>_0001.hdr.exr, <name>_0004.hdr.exr   noisy colour after 1 and 4 samples
    <name>_ref.hdr.exr                          reference colour
    <name>.alb.exr, <name>.nrm.exr              first-hit albedo and normal
    <name>.z.exr                                first-hit depth along the view axis (FP32 Y)
    <name>.json                                 metadata

The OpenEXR encoder below is written from the file-format specification (single-part scanline
file: magic and version, attribute header, offset table, 16-scanline ZIP chunks whose bytes are
split into even/odd halves and delta-predicted before zlib); `read_exr` decodes the same subset.
"""
import json
import math
import os
import struct
import zlib

import numpy as np

EXR_MAGIC = 20000630
HALF, FLOAT = 1, 2
ZIP_COMPRESSION = 3
ZIP_LINES = 16


def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _zip_pack(raw):
    """ZIP chunk encoding: interleave (even bytes, then odd bytes), delta predictor, zlib."""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]])
    d = t.astype(np.int16)
    d[1:] = (t[1:].astype(np.int16) - t[:-1].astype(np.int16) + 128) & 0xFF
    return zlib.compress(d.astype(np.uint8).tobytes(), 9)


def _zip_unpack(data, size):
    t = np.frombuffer(zlib.decompress(data), np.uint8).astype(np.int32)
    t = np.cumsum(np.concatenate([t[:1], t[1:] - 128])) & 0xFF
    t = t.astype(np.uint8)
    half = (size + 1) // 2
    out = np.empty(size, np.uint8)
    out[0::2] = t[:half]
    out[1::2] = t[half:]
    return out.tobytes()


def write_exr(path, img, half=True, channels=None):
    """Writes an H x W x C float image (C = 3: R, G, B; C = 1: Y) as a ZIP-compressed OpenEXR file."""
    img = np.asarray(img, np.float32)
    if img.ndim == 2:
        img = img[:, :, None]
    H, W, C = img.shape
    names = channels or (("R", "G", "B") if C == 3 else ("Y",) if C == 1 else tuple("C%d" % i for i in range(C)))
    ptype = HALF if half else FLOAT
    # channels are stored in alphabetical order of their names
    order = sorted(range(C), key=lambda i: names[i])
    chlist = b"".join(names[i].encode() + b"\0" + struct.pack("<iB3xii", ptype, 0, 1, 1) for i in order) + b"\0"
    header = b"".join([
        _attr("channels", "chlist", chlist),
        _attr("compression", "compression", bytes([ZIP_COMPRESSION])),
        _attr("dataWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1)),
        _attr("displayWindow", "box2i", struct.pack("<iiii", 0, 0, W - 1, H - 1)),
        _attr("lineOrder", "lineOrder", bytes([0])),
        _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0)),
        _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0)),
        _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)),
    ]) + b"\0"
    dt = np.float16 if half else np.float32
    planes = img.astype(dt)
    chunks = []
    for y0 in range(0, H, ZIP_LINES):
        rows = planes[y0:y0 + ZIP_LINES]
        # per scanline: every channel's W values in turn
        raw = np.ascontiguousarray(rows[:, :, order].transpose(0, 2, 1)).astype(dt).tobytes()
        packed = _zip_pack(raw)
        data = packed if len(packed) < len(raw) else raw
        chunks.append(struct.pack("<ii", y0, len(data)) + data)
    prefix = struct.pack("<ii", EXR_MAGIC, 2) + header
    table_end = len(prefix) + 8 * len(chunks)
    offsets, pos = [], table_end
    for c in chunks:
        offsets.append(pos)
        pos += len(c)
    with open(path, "wb") as f:
        f.write(prefix)
        f.write(struct.pack("<%dQ" % len(offsets), *offsets))
        for c in chunks:
            f.write(c)


def read_exr(path):
    """Reads a single-part scanline OpenEXR file (ZIP or uncompressed, HALF/FLOAT channels) written
    as above: returns (H x W x C float32 array in the file's channel order, channel names)."""
    buf = open(path, "rb").read()
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != EXR_MAGIC or (version & 0xFF) != 2:
        raise ValueError("not a single-part OpenEXR 2 file")
    pos, attrs = 8, {}
    while buf[pos] != 0:
        e = buf.index(b"\0", pos)
        name = buf[pos:e].decode()
        e2 = buf.index(b"\0", e + 1)
        typ = buf[e + 1:e2].decode()
        (size,) = struct.unpack_from("<i", buf, e2 + 1)
        attrs[name] = (typ, buf[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    chl = attrs["channels"][1]
    chans, p = [], 0
    while chl[p] != 0:
        e = chl.index(b"\0", p)
        ptype = struct.unpack_from("<i", chl, e + 1)[0]
        chans.append((chl[p:e].decode(), ptype))
        p = e + 1 + 16
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    comp = attrs["compression"][1][0]
    lines = ZIP_LINES if comp == ZIP_COMPRESSION else 1
    nchunks = (H + lines - 1) // lines
    offsets = struct.unpack_from("<%dQ" % nchunks, buf, pos)
    out = np.zeros((H, W, len(chans)), np.float32)
    for off in offsets:
        yy, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8:off + 8 + size]
        nl = min(lines, H - (yy - y0))
        row_bytes = sum(W * (2 if t == HALF else 4) for _, t in chans)
        raw_size = nl * row_bytes
        raw = data if size == raw_size else _zip_unpack(data, raw_size)
        q = 0
        for l in range(nl):
            for ci, (_, t) in enumerate(chans):
                dt, nb = (np.float16, 2) if t == HALF else (np.float32, 4)
                out[yy - y0 + l, :, ci] = np.frombuffer(raw, dt, W, q)
                q += W * nb
    return out, [c for c, _ in chans]


def scene_scale(positions):
    """RenderWindow's automatic sceneScale (main/render/render_window.cpp:153-154): a tenth of the
    smallest extent of the scene's bounding box."""
    p = np.asarray(positions, np.float64).reshape(-1, 3)
    return float((p.max(axis=0) - p.min(axis=0)).min() * 0.1)


def camera_metadata(camera, width, height, spp, max_depth, exposure=1.0, scene_scale_=None, jitter=None,
                    pixel_filter="box", pixel_filter_width=0.0):
    """The reference's per-image JSON keys (main/render/render_window.cpp:1406-1444): the pixel
    filter (or the jitter offset), the sampler, the camera frame and lens, worldToViewMatrix and
    viewToClipMatrix (Device::getWorldToViewD3D / getViewToClipD3D, rows as toJSON(Mat4f) writes
    them), sceneScale, exposure, path depth and sample count."""
    import protoray_hip as P
    U, V, N = (np.asarray(camera[k], float) for k in ("U", "V", "N"))
    world_to_view, view_to_clip, _ = P.camera_matrices(camera, width, height)
    meta = {}
    if jitter is not None:
        meta["jitterOffset"] = [0.5 - float(jitter[0]), 0.5 - float(jitter[1])]
    else:
        meta["pixelFilter"] = pixel_filter
        meta["pixelFilterWidth"] = float(pixel_filter_width)
    meta.update({
        "sampler": "sobolBurley",
        "cameraPosition": [float(x) for x in camera["pos"]],
        "cameraForward": [float(-x) for x in N],
        "cameraUp": [float(x) for x in V],
        "cameraRight": [float(x) for x in U],
        "cameraAspectRatio": width / height,
        "cameraFovVertical": float(camera["fov"]),
        "cameraLensRadius": float(camera["lensRadius"]),
        "cameraFocalDistance": float(camera["focalDistance"]),
        "cameraNear": float(camera["nearClip"]),
        "worldToViewMatrix": [[float(x) for x in row] for row in world_to_view],
        "viewToClipMatrix": [[float(x) for x in row] for row in view_to_clip],
    })
    if scene_scale_ is not None:
        meta["sceneScale"] = scene_scale_
    meta.update({"exposure": exposure, "maxPathDepth": max_depth, "samplesPerPixel": spp})
    return meta


def write_frame(outdir, name, noisy, ref, albedo, normal, meta, depth=None):
    """noisy: {spp: H x W x 3}; depth: H x W or None; writes the files listed in the module
    docstring (the depth as FP32, as the reference saves it: saveImageSafe(..., 1, 32))."""
    os.makedirs(outdir, exist_ok=True)
    paths = []
    for spp, img in sorted(noisy.items()):
        paths.append(os.path.join(outdir, "%s_%04d.hdr.exr" % (name, spp)))
        write_exr(paths[-1], img)
    for suffix, img in (("_ref.hdr", ref), (".alb", albedo), (".nrm", normal)):
        if img is not None:
            paths.append(os.path.join(outdir, name + suffix + ".exr"))
            write_exr(paths[-1], img)
    if depth is not None:
        paths.append(os.path.join(outdir, name + ".z.exr"))
        write_exr(paths[-1], depth, half=False)
    paths.append(os.path.join(outdir, name + ".json"))
    with open(paths[-1], "w") as f:
        json.dump(meta, f, indent=2)
    return paths


# ---------------------------------------------------------------------------------------
# The batched dataset configuration: 64 frames from 8 seeded scenes of the sun/sky sphere
# generator (about 100 k to 4 M triangles), sizes between 256x256 and 2048x1152, max depth 8,
# Sobol-Burley seeds 0..63; each frame keeps the means after 1 and 4 samples, the first-hit
# albedo and normal, and the reference mean after ref_spp samples.
# ---------------------------------------------------------------------------------------
def dataset_scenes(n_scenes=8, min_tris=100_000, max_tris=4_000_000, palette=256):
    import scenegen as G
    scenes = []
    for i in range(n_scenes):
        target = min_tris * (max_tris / min_tris) ** (i / max(n_scenes - 1, 1))
        subdiv = 3 if target < 1_000_000 else 4
        per = 20 * 4 ** subdiv
        n = max(8, int(round(target / per)))
        sc = G.spheres_sunsky(seed=i, n_spheres=n, subdiv=subdiv, palette=palette, coated=0.3)
        scenes.append(sc)
    return scenes


def look_at(pos, target, fov_deg=45.0, lens_radius=0.0, focal=1.0):
    import scenegen as G
    f = np.asarray(target, float) - np.asarray(pos, float)
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0])
    N = -f
    U = np.cross(up, N)
    U /= np.linalg.norm(U)
    V = np.cross(N, U)
    c = np.zeros((), G.CAMERA_DTYPE)
    c["pos"] = pos
    c["U"], c["V"], c["N"] = U, V, N
    c["fov"] = math.radians(fov_deg)
    c["nearClip"] = 0.01
    c["lensRadius"] = lens_radius
    c["focalDistance"] = focal
    return c


def dataset_frames(n_frames=64, n_scenes=8, seed=7):
    """(scene index, camera, width, height, seed) of every frame."""
    rng = np.random.default_rng(seed)
    frames = []
    for i in range(n_frames):
        W = int(rng.integers(32, 257)) * 8
        H = int(rng.integers(32, 145)) * 8
        pos = (rng.uniform(-6, 6), rng.uniform(0.5, 5), rng.uniform(-4, 2))
        tgt = (rng.uniform(-6, 6), rng.uniform(0, 4), rng.uniform(12, 24))
        cam = look_at(pos, tgt, fov_deg=float(rng.uniform(35, 60)), lens_radius=float(rng.choice([0.0, 0.05])),
                      focal=float(np.linalg.norm(np.subtract(tgt, pos))))
        frames.append((i % n_scenes, cam, W, H, i))
    return frames


def render_dataset(outdir=None, ref_spp=4096, n_frames=64, n_scenes=8, max_tris=4_000_000, device=0, tune=True, log=print):
    """Renders the batched dataset in one persistent launch; writes the files when outdir is given.
    Returns a summary (timings, sample counts, tuned launch shape)."""
    import protoray_hip as P
    scenes = dataset_scenes(n_scenes, max_tris=max_tris)
    devs = []
    for sc in scenes:
        d = P.DeviceHip(device)
        P.load_scene(d, sc)
        devs.append(d)
    batch = P.BatchHip(device)
    frames = dataset_frames(n_frames, n_scenes)
    for si, cam, W, H, seed in frames:
        batch.addFrame(devs[si], cam, W, H, seed, ref_spp, snapshots=(1, 4),
                       aov_mask=P.AOV_ALBEDO1 | P.AOV_NORMAL1 | P.AOV_DEPTH)
    tuned = batch.tune(1) if tune else None
    r = batch.render()
    summary = {"frames": n_frames, "scenes": n_scenes, "triangles": [int(sc["indices"].shape[0]) for sc in scenes],
               "pixels": int(sum(W * H for _, _, W, H, _ in frames)), "ref_spp": ref_spp, "launch_ms": r["ms"],
               "samples": r["samples"], "msamples_per_s": r["samples"] / (r["ms"] * 1e3), "tuned": tuned}
    if outdir:
        for i, (si, cam, W, H, seed) in enumerate(frames):
            meta = camera_metadata(cam, W, H, ref_spp, scenes[si]["max_depth"],
                                   scene_scale_=scene_scale(scenes[si]["positions"]))
            meta.update({"scene": si, "seed": seed})
            write_frame(outdir, "frame%02d" % i, {1: batch.get(i, P.BATCH_SNAPSHOT), 4: batch.get(i, P.BATCH_SNAPSHOT + 1)},
                        batch.get(i, P.BATCH_COLOR), batch.get(i, P.BATCH_ALBEDO1), batch.get(i, P.BATCH_NORMAL1), meta,
                        depth=batch.get(i, P.BATCH_DEPTH))
    batch.close()
    for d in devs:
        d.close()
    log(json.dumps(summary))
    return summary
