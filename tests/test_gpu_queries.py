"""Debug queries on the GPU: Renderer::queryPixel (render/renderer.h:82-121 with the hit
properties of render/renderer_stream.h:16-62) and queryOcclusionRay with the transparent-shadow
filter (renderer_stream.h:64-97), on a scene with transparent and thin-walled materials.

* The first hit of a pixel query is the closest hit of the same camera ray found by brute force
  (the primitive, the distance to 1e-5), and its point, offset and geometric normal equal the
  oracle's post-intersection of that hit.
* Later hits step past transparent surfaces; each hit reports the material's transparency
  (0 for opaque materials, the opacity complement for non-thin ones) and its depth along the
  view axis.
* An occlusion query along the same ray reports occlusion exactly when the pixel query stops at
  an opaque surface or the transmittance product falls below 1e-5.
"""
import numpy as np
import pytest

import oracle as O
import protoray_hip as P
import scenegen as G

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = P.DeviceHip(0)
    yield d
    d.close()


def test_query_pixel_and_filtered_occlusion(dev):
    sc = G.spheres_lobes()
    P.load_scene(dev, sc)
    W, H = 96, 64
    dev.initFrame(W, H)
    mats = sc["materials"]
    osc = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], mats)
    cam = sc["camera"]
    assert float(cam["lensRadius"]) == 0.0
    cfg = [W, H, float(cam["fov"]), float(cam["nearClip"]), 0.0, float(cam["focalDistance"]), *map(float, cam["pos"])]
    pix = [(x, y) for y in range(0, H, 3) for x in range(0, W, 3)]
    s = np.array([[(x + 0.5) / W, (y + 0.5) / H, 0.0, 0.0] for x, y in pix], np.float32)
    r = O.camera_rays(cfg, s)[:, 0]  # pinhole: origin, direction, near
    rays = np.concatenate([r[:, :7], np.full((len(pix), 1), 3.4028235e38, np.float32)], 1)
    dc = O.camera_rays(cfg, np.array([[0.5, 0.5, 0.0, 0.0]], np.float32))[0, 0, 3:6]
    bprim, bt, bu, bv = osc.intersect(rays, mode=1)
    occ = dev.queryOcclusion(rays)
    hit_rows, first_prims = [], []
    multi = transparent_hits = edge = 0
    for k, (x, y) in enumerate(pix):
        hits = dev.queryPixel(x, y)
        if bprim[k] < 0 and len(hits) == 0:
            assert occ[k] == 0
            continue
        # the device forms the camera ray with fused multiply-adds: its direction can differ from
        # the oracle's in the last bits, which moves t by a few ulp and, on a triangle edge, the
        # primitive
        if len(hits) == 0 or hits[0]["prim"] != bprim[k]:
            edge += 1
            continue
        np.testing.assert_allclose(hits[0]["dist"], bt[k], rtol=1e-5)
        hit_rows.append(k)
        first_prims.append(hits[0]["prim"])
        assert np.all(np.diff(hits["dist"]) > 0)
        multi += len(hits) > 1
        Tprod = np.ones(3)
        for h in hits:
            m = mats[h["material"]]
            transparent = m["transparency"] != 0 or m["kind"] == G.KIND_THIN_DIELECTRIC
            assert bool(h["material_type"] & 1) == transparent
            if not transparent:
                assert np.all(h["T"] == 0)
            elif m["kind"] != G.KIND_THIN_DIELECTRIC:
                transparent_hits += 1
                np.testing.assert_allclose(h["T"], 1.0 - min(max(1.0 - m["transparency"], 0.0), 1.0), rtol=1e-6)
            else:
                assert np.all((h["T"] >= 0) & (h["T"] <= 1))
            np.testing.assert_allclose(h["depth"], h["dist"] * float(np.dot(rays[k, 3:6], dc)), rtol=1e-5)
            assert np.dot(h["Ng"], rays[k, 3:6]) <= 0  # oriented towards the viewer
            Tprod = Tprod * h["T"]
        stops_opaque = hits["T"][-1].max() < 1e-5
        assert occ[k] == (1 if stops_opaque or Tprod.max() < 1e-5 else 0), (x, y, hits)
    assert len(hit_rows) > 200 and multi > 20 and transparent_hits > 20 and edge <= 2
    # the first hits' geometry against the oracle's post-intersection (before orientation)
    hr = np.array(hit_rows)
    r9 = np.concatenate([rays[hr, :6], bt[hr, None], bu[hr, None], bv[hr, None]], 1)  # [o d t u v]
    ctx, _ = O.post_intersect(sc["positions"], None, sc["indices"], r9, np.array(first_prims, np.int32))
    first = np.array([dev.queryPixel(*pix[k], max_hits=1)[0] for k in hit_rows], P.PIXEL_HIT_DTYPE)
    # p = o + t d may be contracted to a fused multiply-add on the device: a few ulp
    np.testing.assert_allclose(first["p"], ctx[:, 0:3], rtol=1e-5, atol=1e-5)
    np.testing.assert_allclose(first["eps"], ctx[:, 3], rtol=1e-5)
    np.testing.assert_allclose(np.abs(np.sum(first["Ng"] * ctx[:, 4:7], 1)), 1.0, atol=1e-5)


def test_query_pixel_needs_a_frame_and_valid_counts(dev):
    sc = G.spheres_ambient(n_spheres=8, subdiv=1)
    P.load_scene(dev, sc)
    dev.initFrame(32, 8)
    assert len(dev.queryPixel(16, 4, max_hits=1)) <= 1
    with pytest.raises(ValueError):
        dev.queryPixel(16, 4, max_hits=33)
