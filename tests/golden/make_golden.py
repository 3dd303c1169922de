"""Writes the golden fixtures from the reference's own code (run in the build container only).

Builds oracle/_ref/ from the sources under /root/reference with oracle/Makefile.ref, runs the
harness oracle/ref_golden.cpp, and packs its raw outputs into tests/golden/ref_vectors.npz
(inputs and reference outputs; see the harness header for the reference code exercised).
Also bakes the sun/sky environment of the 1080p benchmark scene with oracle/ref_sunsky.cpp
(the reference's vendored Hosek-Wilkie model) into
oidn-protoray_amd/scenes/sunsky_e35_t3_az30.npz, a scene input asset.

Usage: python tests/golden/make_golden.py
"""
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
ORACLE = os.path.join(ROOT, "oracle")

# Sun of the 1080p scene: elevation 35 deg, azimuth 30 deg in the reference's convention
# phi = atan2(x, -z) (light/sun_sky_light.cpp:42-43); turbidity 3, ground albedo 0.2.
SUN_ELEVATION_DEG = 35.0
SUN_AZIMUTH_DEG = 30.0
TURBIDITY = 3.0
GROUND_ALBEDO = 0.2


def sun_direction():
    el, az = math.radians(SUN_ELEVATION_DEG), math.radians(SUN_AZIMUTH_DEG)
    return (math.sin(az) * math.cos(el), math.sin(el), -math.cos(az) * math.cos(el))


def main():
    subprocess.check_call(["make", "-f", "Makefile.ref", "-j8"], cwd=ORACLE)
    with tempfile.TemporaryDirectory() as tmp:
        subprocess.check_call([os.path.join(ORACLE, "_ref", "ref_golden"), tmp])
        with open(os.path.join(tmp, "manifest.json")) as f:
            manifest = json.load(f)
        arrays = {}
        for name, meta in manifest.items():
            a = np.fromfile(os.path.join(tmp, name + ".bin"), dtype=np.dtype(meta["dtype"]))
            arrays[name] = a.reshape(meta["shape"])
        out = os.path.join(ROOT, "tests", "golden", "ref_vectors.npz")
        np.savez_compressed(out, **arrays)
        print("wrote", out, len(arrays), "arrays")

        d = sun_direction()
        subprocess.check_call([os.path.join(ORACLE, "_ref", "ref_sunsky"), tmp,
                               repr(d[0]), repr(d[1]), repr(d[2]), repr(TURBIDITY), repr(GROUND_ALBEDO)])
        sky = np.fromfile(os.path.join(tmp, "sky.bin"), dtype=np.float32).reshape(256, 512, 3)
        sun = np.array([float(x) for x in open(os.path.join(tmp, "sun.txt")).read().split()], dtype=np.float32)
        asset = os.path.join(ROOT, "oidn-protoray_amd", "scenes", "sunsky_e35_t3_az30.npz")
        np.savez_compressed(asset, sky=sky, sun_radiance=sun[0:3], sun_radius=sun[3:4], sun_dir=sun[4:7],
                            elevation_deg=np.float32(SUN_ELEVATION_DEG), azimuth_deg=np.float32(SUN_AZIMUTH_DEG),
                            turbidity=np.float32(TURBIDITY), ground_albedo=np.float32(GROUND_ALBEDO))
        print("wrote", asset)
    return 0


if __name__ == "__main__":
    sys.exit(main())
