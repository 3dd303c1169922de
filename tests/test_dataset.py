"""Dataset writer: OpenEXR encoding (ZIP scanline chunks) and the frame layout.

No OpenEXR library is installed here, so the encoder is pinned by the specification's fixed
bytes (magic number, version, attribute and channel-list layout), a hand-computed ZIP
pre-processing example, and exact round trips through the decoder (half and float, ragged
last chunk, incompressible data stored raw); agreement with OpenEXR's own reader is unpinned.
"""
import json
import struct
import zlib

import numpy as np

import dataset as D


def test_zip_preprocessing_by_hand():
    raw = bytes([10, 200, 12, 201, 15, 199])
    # even bytes then odd bytes: 10 12 15 200 201 199; deltas + 128 (mod 256) after the first
    expect = bytes([10, (12 - 10 + 128) & 255, (15 - 12 + 128) & 255, (200 - 15 + 128) & 255, (201 - 200 + 128) & 255,
                    (199 - 201 + 128) & 255])
    assert zlib.decompress(D._zip_pack(raw)) == expect
    assert D._zip_unpack(D._zip_pack(raw), len(raw)) == raw


def test_header_layout(tmp_path):
    p = tmp_path / "a.exr"
    D.write_exr(p, np.zeros((5, 3, 3), np.float32))
    b = p.read_bytes()
    assert b[:4] == bytes([0x76, 0x2F, 0x31, 0x01]) and struct.unpack("<i", b[4:8])[0] == 2
    assert b"channels\0chlist\0" in b and b"compression\0compression\0" in b
    i = b.index(b"chlist\0") + 7
    (size,) = struct.unpack("<i", b[i:i + 4])
    chl = b[i + 4:i + 4 + size]
    # B, G, R in alphabetical order, HALF (1), linear 0, sampling 1 x 1
    assert chl == b"".join(c + b"\0" + struct.pack("<iB3xii", 1, 0, 1, 1) for c in (b"B", b"G", b"R")) + b"\0"


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for H, W, C, half in ((37, 29, 3, True), (16, 8, 1, False), (1, 1, 3, True), (40, 33, 3, False)):
        img = (rng.standard_normal((H, W, C)) * 10).astype(np.float32)
        p = tmp_path / f"r{H}_{C}_{half}.exr"
        D.write_exr(p, img, half=half)
        back, names = D.read_exr(p)
        assert names == (["B", "G", "R"] if C == 3 else ["Y"])
        ref = img.astype(np.float16).astype(np.float32) if half else img
        if C == 3:
            ref = ref[:, :, ::-1]  # file order B, G, R
        assert np.array_equal(back, ref)
    # white noise bytes do not compress: the chunk is stored raw and still decodes
    img = rng.random((16, 64, 3)).astype(np.float32)
    D.write_exr(tmp_path / "n.exr", img, half=False)
    back, _ = D.read_exr(tmp_path / "n.exr")
    assert np.array_equal(back[:, :, ::-1], img)


def test_write_frame_files(tmp_path):
    import scenegen as G
    cam = D.look_at((0, 1, 0), (0, 1, 10))
    img = np.ones((8, 16, 3), np.float32)
    meta = D.camera_metadata(cam, 16, 8, 64, 8, scene_scale_=D.scene_scale([[0, 0, 0], [3, 2, 5]]))
    depth = np.linspace(1, 9, 8 * 16, dtype=np.float32).reshape(8, 16)
    depth[0, 0] = np.inf  # a pixel whose samples all missed
    paths = D.write_frame(tmp_path, "f", {1: img, 4: img}, img, img, img, meta, depth=depth)
    names = sorted(p.split("/")[-1] for p in paths)
    assert names == ["f.alb.exr", "f.json", "f.nrm.exr", "f.z.exr", "f_0001.hdr.exr", "f_0004.hdr.exr",
                     "f_ref.hdr.exr"]
    z, zn = D.read_exr(tmp_path / "f.z.exr")
    assert zn == ["Y"] and np.array_equal(z[:, :, 0], depth)  # FP32, exact
    m = json.load(open(tmp_path / "f.json"))
    np.testing.assert_allclose(m["cameraForward"], [0, 0, 1], atol=1e-7)
    assert m["cameraAspectRatio"] == 2.0 and m["samplesPerPixel"] == 64 and m["maxPathDepth"] == 8
    # the keys of main/render/render_window.cpp:1408-1441, in its order
    assert list(m) == ["pixelFilter", "pixelFilterWidth", "sampler", "cameraPosition", "cameraForward", "cameraUp",
                       "cameraRight", "cameraAspectRatio", "cameraFovVertical", "cameraLensRadius",
                       "cameraFocalDistance", "cameraNear", "worldToViewMatrix", "viewToClipMatrix", "sceneScale",
                       "exposure", "maxPathDepth", "samplesPerPixel"]
    assert m["pixelFilter"] == "box" and m["pixelFilterWidth"] == 0.0 and abs(m["sceneScale"] - 0.2) < 1e-12
    # worldToViewD3D maps the camera position to the origin and a point ahead to +z (row vectors,
    # D3D view space looks along +z)
    wv = np.array(m["worldToViewMatrix"])
    assert wv.shape == (4, 4)
    np.testing.assert_allclose(np.array([0, 1, 0, 1]) @ wv, [0, 0, 0, 1], atol=1e-6)
    np.testing.assert_allclose(np.array([0, 1, 5, 1]) @ wv, [0, 0, 5, 1], atol=1e-5)
    assert np.array(m["viewToClipMatrix"]).shape == (4, 4)
    jm = D.camera_metadata(cam, 16, 8, 64, 8, jitter=(0.25, 0.5))
    assert jm["jitterOffset"] == [0.25, 0.0] and "pixelFilter" not in jm


def test_dataset_configuration_shape():
    frames = D.dataset_frames()
    assert len(frames) == 64 and {f[0] for f in frames} == set(range(8))
    assert all(256 <= W <= 2048 and 256 <= H <= 1152 and W % 8 == 0 and H % 8 == 0 for _, _, W, H, _ in frames)
    assert [f[4] for f in frames] == list(range(64))
