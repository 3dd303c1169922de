"""CPU-side checks of the product library: it loads and exports every entry point of
include/prt_hip.h (no compute calls without a GPU)."""
import ctypes
import os
import re

import protoray_hip as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "prt_hip.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(prt_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_and_binding_agree():
    assert declared_entry_points() == sorted(P.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = P.load_library()
    for name in declared_entry_points():
        assert hasattr(lib, name), name


def test_last_error_is_callable_without_a_device():
    lib = P.load_library()
    assert isinstance(lib.prt_last_error(), bytes)
    # a null context is rejected with PRT_ERR_INVALID, no GPU needed
    assert lib.prt_frame_clear(None) == P.PRT_ERR_INVALID
    assert b"null context" in lib.prt_last_error()


def test_camera_matrices_match_reference(golden):
    # camera/pinhole_camera.h:41-68 through the C ABI (host only): worldToViewD3D, viewToClipD3D,
    # worldToRaster against the reference's PinholeCamera for axis-aligned and rotated bases
    import numpy as np
    import protoray_hip as P
    cin, cout = golden["cam_mat_in"], golden["cam_mat_out"]
    for row, ref in zip(cin, cout):
        cam = np.zeros((), P.CAMERA_DTYPE)
        cam["fov"], cam["nearClip"] = row[2], row[3]
        cam["pos"], cam["U"], cam["V"], cam["N"] = row[4:7], row[7:10], row[10:13], row[13:16]
        mats = P.camera_matrices(cam, int(row[0]), int(row[1]))
        for m, r in zip(mats, ref):
            np.testing.assert_allclose(m.ravel(), r, rtol=1e-6, atol=1e-5 * np.abs(r).max())
        # a world point projects to the raster position of the ray through it
        W, H = int(row[0]), int(row[1])
        p = row[4:7] - 5.0 * row[13:16] + 0.3 * row[7:10] - 0.2 * row[10:13]
        q = mats[2] @ np.append(p, 1.0)
        px, py = q[0] / q[3], q[1] / q[3]
        assert 0 < px < W and 0 < py < H
