"""Writers of small scene files for the scene-file tests (.rkscene + .rkbin, .gltf with external
or data-URI buffers, the text mesh description the reference harness turns into a .mesh blob,
.mtl).  They follow the published layouts: rkscene/rkscene.h (typed buffer views into one .rkbin
file), glTF 2.0 (accessors over buffer views, node TRS), the harness's DESC text."""
import base64
import json
import os
import struct

import numpy as np


def icosphere(subdiv=1):
    """Unit icosphere: (positions, indices), outward vertex normals = positions."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, float) / np.linalg.norm(p) for p in v]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[k] = len(verts) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(verts, np.float32), np.array(f, np.uint32)


def quaternion(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    s = np.sin(angle / 2)
    return [float(axis[0] * s), float(axis[1] * s), float(axis[2] * s), float(np.cos(angle / 2))]


def trs_matrix(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    """4x4 row-major matrix of translation * rotation(q = x, y, z, w) * scale."""
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(s, float)[None, :]
    M[:3, 3] = t
    return M


class RkWriter:
    """Builds an .rkscene document and its .rkbin buffer."""

    def __init__(self):
        self.bin = bytearray()
        self.doc = {"materials": [], "geometries": [], "objects": [], "lights": []}

    def view(self, arr, typ):
        off = len(self.bin)
        self.bin += np.ascontiguousarray(arr).tobytes()
        count = {"float3": len(arr) // 1 if np.ndim(arr) == 2 else len(arr) // 3, "uint": np.size(arr),
                 "float3x4": len(arr)}[typ]
        return {"type": typ, "count": int(count), "byteOffset": off}

    def geometry(self, pos, idx, material=None, normals=None):
        g = {"type": "triangle", "positions": self.view(np.asarray(pos, np.float32).reshape(-1, 3), "float3"),
             "indices": self.view(np.asarray(idx, np.uint32).ravel(), "uint")}
        if normals is not None:
            g["normals"] = self.view(np.asarray(normals, np.float32).reshape(-1, 3), "float3")
        if material is not None:
            g["material"] = material
        self.doc["geometries"].append(g)
        return len(self.doc["geometries"]) - 1

    def object(self, geoms, instances=None):
        o = {"geometries": list(geoms)}
        if instances is not None:
            # float3x4 = 4 columns (linear part, translation) from 4x4 row-major matrices
            cols = np.stack([np.concatenate([M[:3, 0], M[:3, 1], M[:3, 2], M[:3, 3]]) for M in instances]).astype(np.float32)
            o["instances"] = self.view(cols, "float3x4")
        self.doc["objects"].append(o)

    def save(self, path):
        base = os.path.splitext(path)[0]
        with open(base + ".rkbin", "wb") as f:
            f.write(bytes(self.bin))
        doc = dict(self.doc)
        doc["buffers"] = [{"uri": os.path.basename(base) + ".rkbin", "byteLength": len(self.bin)}]
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)


class GltfWriter:
    """Builds a glTF 2.0 document with one binary buffer (external .bin or a data URI)."""

    def __init__(self):
        self.bin = bytearray()
        self.doc = {"asset": {"version": "2.0"}, "buffers": [], "bufferViews": [], "accessors": [], "meshes": [],
                    "materials": [], "nodes": [], "scenes": [{"nodes": []}], "scene": 0}

    def _align(self, n=4):
        while len(self.bin) % n:
            self.bin.append(0)

    def accessor(self, arr, ctype, typ, stride=None):
        """ctype: 5126 float, 5125 uint32, 5123 uint16, 5121 uint8.  stride > element size
        interleaves padding bytes after each element."""
        self._align()
        dt = {5126: np.float32, 5125: np.uint32, 5123: np.uint16, 5121: np.uint8}[ctype]
        a = np.ascontiguousarray(arr, dt)
        comps = {"SCALAR": 1, "VEC2": 2, "VEC3": 3}[typ]
        count = a.size // comps
        elem = a.itemsize * comps
        off = len(self.bin)
        if stride and stride > elem:
            rows = a.reshape(count, comps)
            for r in rows:
                self.bin += r.tobytes() + bytes(stride - elem)
            view = {"buffer": 0, "byteOffset": off, "byteLength": count * stride, "byteStride": stride}
        else:
            self.bin += a.tobytes()
            view = {"buffer": 0, "byteOffset": off, "byteLength": a.nbytes}
        self.doc["bufferViews"].append(view)
        self.doc["accessors"].append({"bufferView": len(self.doc["bufferViews"]) - 1, "componentType": ctype,
                                      "count": int(count), "type": typ})
        return len(self.doc["accessors"]) - 1

    def mesh(self, prims):
        self.doc["meshes"].append({"primitives": prims})
        return len(self.doc["meshes"]) - 1

    def node(self, mesh=None, children=None, root=False, **trs):
        n = {}
        if mesh is not None:
            n["mesh"] = mesh
        if children:
            n["children"] = list(children)
        n.update(trs)
        self.doc["nodes"].append(n)
        i = len(self.doc["nodes"]) - 1
        if root:
            self.doc["scenes"][0]["nodes"].append(i)
        return i

    def save(self, path, data_uri=False):
        doc = dict(self.doc)
        if data_uri:
            doc["buffers"] = [{"byteLength": len(self.bin),
                               "uri": "data:application/octet-stream;base64," + base64.b64encode(bytes(self.bin)).decode()}]
        else:
            binname = os.path.splitext(os.path.basename(path))[0] + ".bin"
            with open(os.path.join(os.path.dirname(path), binname), "wb") as f:
                f.write(bytes(self.bin))
            doc["buffers"] = [{"byteLength": len(self.bin), "uri": binname}]
        with open(path, "w") as f:
            json.dump(doc, f, indent=1)


def write_mesh_desc(path, pos, idx, mat, names, normals=None):
    """The text mesh description read by the reference harness (ref_scene mesh / clone)."""
    with open(path, "w") as f:
        f.write(f"{len(pos)} {len(idx)} {1 if normals is not None else 0} {len(names)}\n")
        for i, p in enumerate(pos):
            row = list(p) + (list(normals[i]) if normals is not None else [])
            f.write(" ".join("%.9g" % v for v in row) + "\n")
        for t, m in zip(idx, mat):
            f.write(f"{t[0]} {t[1]} {t[2]} {m}\n")
        for n in names:
            f.write((n if n else "-") + "\n")


def write_mesh_blob(path, pos, idx, mat, names, normals=None):
    """A .mesh blob in the TriangleMesh stream layout (geometry/triangle_mesh.cpp:382-397): blob ID,
    bounds, counts, flags, indices, material IDs, interleaved vertex attributes, material names."""
    pos = np.asarray(pos, np.float32)
    out = bytearray(struct.pack("<i", 0x04df8666))
    out += struct.pack("<6f", *pos.min(0), *pos.max(0))
    out += struct.pack("<iiBb", len(idx), len(pos), 1 if normals is not None else 0, 0)
    out += np.asarray(idx, np.int32).tobytes() + np.asarray(mat, np.int32).tobytes()
    attr = pos if normals is None else np.concatenate([pos, np.asarray(normals, np.float32)], 1)
    out += np.ascontiguousarray(attr, np.float32).tobytes()
    out += struct.pack("<i", len(names))
    for n in names:
        b = n.encode()
        out += struct.pack("<i", len(b)) + b
    with open(path, "wb") as f:
        f.write(bytes(out))
