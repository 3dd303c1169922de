"""Volume transmission of the oracle against a closed form (path_renderer_stream.h:652-657,
material/standard.cpp:169-173, material/material_factory.cpp:159-175).

Two slabs of a smooth dielectric with an IOR of 1.0001 (Fresnel 2.5e-9 at normal incidence, so rays
cross them almost straight and unreflected) under an ambient light: a camera ray through slab k
returns ambient x (colour_k ^ depth_k) ^ (thickness / cos theta).  The slabs share their IOR and
differ in transmission colour and depth, so the two halves of the image also pin the media's
deduplication by (colour, IOR); a wrong exponent (colour ^ (1 / depth), the path length before
the refraction instead of inside the slab, the attenuation applied on the segment after the exit)
fails the comparison."""
import numpy as np

import estimator_closed_form as C
import oracle as O
import scenegen as G


def render_slabs(sc, spp=4):
    W, H = sc["width"], sc["height"]
    s = O.Scene(sc["positions"], sc["indices"], sc["tri_material"], sc["materials"])
    s.set_lights(sc["lights"])
    s.set_pixel_sample(jitter=(0.5, 0.5))
    acc, _ = s.render(sc["camera"], W, H, spp_count=spp, max_depth=sc["max_depth"], rr_depth=sc["rr_depth"], threads=4)
    return acc["color"] / acc["weight"][None]


def slab_reference(sc):
    d = C.pinhole_dirs(sc["camera"], sc["width"], sc["height"], 0.5, 0.5)
    ref, k = G.glass_slab_radiance(sc, d)
    x_hit = d[:, 0] / d[:, 2] * sc["slab_z0"]
    keep = np.abs(x_hit) > 0.2  # away from the seam between the slabs
    return ref, k, keep


def test_oracle_slab_absorption_closed_form():
    sc = G.glass_slabs()
    mean = render_slabs(sc)
    ref, k, keep = slab_reference(sc)
    assert keep.mean() > 0.9 and (k[keep] == 0).any() and (k[keep] == 1).any()
    np.testing.assert_allclose(mean[:, keep].T, ref[keep], rtol=2e-3)
    # the two media are told apart (same IOR, other colour): slab 0 is blue-tinted, slab 1 grey
    left, right = mean[:, keep & (k == 0)], mean[:, keep & (k == 1)]
    assert (left[2] > 2 * left[0]).all() and np.allclose(right[0], right[2], rtol=1e-3)


def test_oracle_clear_media():
    # transmission depth 0 or colour 1 leave the medium clear (colour ^ 0 = 1, 1 ^ d = 1)
    sc = G.glass_slabs(slabs=(((0.3, 0.6, 0.9), 0.0), ((1.0, 1.0, 1.0), 3.0)))
    mean = render_slabs(sc, spp=2)
    _, _, keep = slab_reference(sc)
    np.testing.assert_allclose(mean[:, keep], 1.0, rtol=2e-3)
