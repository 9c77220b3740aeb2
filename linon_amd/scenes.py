"""Scene inputs for the ray-cast path: the Cornell box, linon's vertex normalisation, and a
seeded displaced icosphere.

The reference host owns its meshes and hands them to the compute pass as two storage buffers
(vertices as f32 x,y,z with a 12-byte stride, faces as u32 a,b,c with a 12-byte stride;
src/main_view.wgsl:33-53).  These helpers build the same arrays on the host side of this
library: float32 ``(nv, 3)`` vertices and uint32 ``(nf, 3)`` faces.
"""
from __future__ import annotations

import numpy as np

# The Cornell box (Cornell University Program of Computer Graphics measurements), as the
# 18 quads of src/cornell_box.rs:2-99, each quad split into two triangles (q0,q1,q2) and
# (q2,q3,q0) in quad order, which reproduces the face table of src/cornell_box.rs:102-163.
_CORNELL_QUADS = (
    # floor
    ((552.8, 0.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, 559.2), (549.6, 0.0, 559.2)),
    ((130.0, 0.0, 65.0), (82.0, 0.0, 225.0), (240.0, 0.0, 272.0), (290.0, 0.0, 114.0)),
    ((423.0, 0.0, 247.0), (265.0, 0.0, 296.0), (314.0, 0.0, 456.0), (472.0, 0.0, 406.0)),
    # ceiling
    ((556.0, 548.8, 0.0), (556.0, 548.8, 559.2), (0.0, 548.8, 559.2), (0.0, 548.8, 0.0)),
    ((343.0, 548.8, 227.0), (343.0, 548.8, 332.0), (213.0, 548.8, 332.0), (213.0, 548.8, 227.0)),
    # back wall
    ((549.6, 0.0, 559.2), (0.0, 0.0, 559.2), (0.0, 548.8, 559.2), (556.0, 548.8, 559.2)),
    # right wall
    ((0.0, 0.0, 559.2), (0.0, 0.0, 0.0), (0.0, 548.8, 0.0), (0.0, 548.8, 559.2)),
    # left wall
    ((552.8, 0.0, 0.0), (549.6, 0.0, 559.2), (556.0, 548.8, 559.2), (556.0, 548.8, 0.0)),
    # short block: top, then four sides
    ((130.0, 165.0, 65.0), (82.0, 165.0, 225.0), (240.0, 165.0, 272.0), (290.0, 165.0, 114.0)),
    ((290.0, 0.0, 114.0), (290.0, 165.0, 114.0), (240.0, 165.0, 272.0), (240.0, 0.0, 272.0)),
    ((130.0, 0.0, 65.0), (130.0, 165.0, 65.0), (290.0, 165.0, 114.0), (290.0, 0.0, 114.0)),
    ((82.0, 0.0, 225.0), (82.0, 165.0, 225.0), (130.0, 165.0, 65.0), (130.0, 0.0, 65.0)),
    ((240.0, 0.0, 272.0), (240.0, 165.0, 272.0), (82.0, 165.0, 225.0), (82.0, 0.0, 225.0)),
    # tall block: top, then four sides
    ((423.0, 330.0, 247.0), (265.0, 330.0, 296.0), (314.0, 330.0, 456.0), (472.0, 330.0, 406.0)),
    ((423.0, 0.0, 247.0), (423.0, 330.0, 247.0), (472.0, 330.0, 406.0), (472.0, 0.0, 406.0)),
    ((472.0, 0.0, 406.0), (472.0, 330.0, 406.0), (314.0, 330.0, 456.0), (314.0, 0.0, 456.0)),
    ((314.0, 0.0, 456.0), (314.0, 330.0, 456.0), (265.0, 330.0, 296.0), (265.0, 0.0, 296.0)),
    ((265.0, 0.0, 296.0), (265.0, 330.0, 296.0), (423.0, 330.0, 247.0), (423.0, 0.0, 247.0)),
)


def cornell_box_raw():
    """Raw Cornell box in millimetres: (72, 3) float32 vertices, (36, 3) uint32 faces."""
    verts = np.asarray(_CORNELL_QUADS, dtype=np.float32).reshape(-1, 3)
    q = np.arange(len(_CORNELL_QUADS), dtype=np.uint32)[:, None] * 4
    faces = np.concatenate([q + [0, 1, 2], q + [2, 3, 0]], axis=1).reshape(-1, 3).astype(np.uint32)
    return verts, faces


def normalize_vertices(verts):
    """One global (not per-axis) min/max rescale of every coordinate to [-1, 1], in f32
    (src/vertices.rs:3-17)."""
    v = np.asarray(verts, dtype=np.float32)
    flat = v.reshape(-1)
    lo = np.float32(flat.min())
    hi = np.float32(flat.max())
    out = (flat - lo) / (hi - lo) * np.float32(2.0) - np.float32(1.0)
    return out.astype(np.float32).reshape(v.shape)


def get_center(verts):
    """Centre of the axis-aligned bounding box, in f32 (src/vertices.rs:19-57)."""
    v = np.asarray(verts, dtype=np.float32).reshape(-1, 3)
    return ((v.min(axis=0) + v.max(axis=0)) / np.float32(2.0)).astype(np.float32)


def reference_startup_scene():
    """The scene linon starts with: Cornell box with x and z negated, then normalised
    (src/application.rs:105-111).  Returns (verts, faces, center)."""
    v, f = cornell_box_raw()
    v = v.copy()
    v[:, 0] = -v[:, 0]
    v[:, 2] = -v[:, 2]
    v = normalize_vertices(v)
    return v, f, get_center(v)


def parse_obj(text: str):
    """The first model of a Wavefront OBJ text as (verts f32 [n, 3], faces u32 [m, 3]), the
    arrays linon's drop handler passes on (src/main.rs:139-157: tobj::load_obj with
    triangulate, then models[0].mesh.positions / indices).

    * A new model starts at an `o` or `g` line once the current one has faces.
    * Face corners may be `p`, `p/t`, `p//n` or `p/t/n`; indices are 1-based, and negative ones
      count back from the last vertex read so far.
    * Polygons become triangle fans (0, i, i + 1).
    * A model's vertices are the positions its faces use, renumbered in order of first use.

    tobj itself (a crate the reference pins but does not vendor) cannot run here.  Its exact
    vertex order within a model changes no rendered output: normalize_vertices uses only the
    extent, and faces keep the file's order, which is the hit id.
    """
    positions = []
    models, cur = [], []
    for raw in text.splitlines():
        line = raw.split("#", 1)[0].strip()
        if not line:
            continue
        tok = line.split()
        if tok[0] == "v":
            positions.append([float(x) for x in tok[1:4]])
        elif tok[0] == "f":
            corners = []
            for c in tok[1:]:
                i = int(c.split("/")[0])
                if i == 0:
                    raise ValueError(f"OBJ: index 0 in {raw!r}")
                i = i - 1 if i > 0 else len(positions) + i
                if not 0 <= i < len(positions):
                    raise ValueError(f"OBJ: vertex index out of range in {raw!r}")
                corners.append(i)
            if len(corners) < 3:
                raise ValueError(f"OBJ: face with fewer than 3 corners: {raw!r}")
            for k in range(1, len(corners) - 1):
                cur.append((corners[0], corners[k], corners[k + 1]))
        elif tok[0] in ("o", "g") and cur:
            models.append(cur)
            cur = []
    if cur:
        models.append(cur)
    if not models:
        raise ValueError("OBJ: no faces")
    first = models[0]
    remap = {}
    for tri in first:
        for i in tri:
            remap.setdefault(i, len(remap))
    verts = np.zeros((len(remap), 3), np.float32)
    for g, l in remap.items():
        verts[l] = positions[g]
    faces = np.array([[remap[i] for i in tri] for tri in first], dtype=np.uint32)
    return verts, faces


def load_obj_scene(path):
    """A dropped OBJ file as linon loads it (src/application.rs:199-224): the first model,
    normalised by normalize_vertices.  Returns (verts, faces, center)."""
    with open(path, "r", encoding="utf-8", errors="replace") as fh:
        v, f = parse_obj(fh.read())
    v = normalize_vertices(v)
    return v, f, get_center(v)


def cornell_box_centered():
    """The start-up scene translated so its bounding-box centre is the origin (the camera
    configurations look at the origin)."""
    v, f, c = reference_startup_scene()
    return (v - c).astype(np.float32), f


# ---------------------------------------------------------------------------------------
# Seeded displaced icosphere
# ---------------------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 (Steele, Lea, Flood 2014) started from ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def splitmix_uniform(seed: int, n: int, lo: float, hi: float) -> np.ndarray:
    """n doubles uniform in [lo, hi) from the top 53 bits of SplitMix64."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array(
        [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0),
         (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    f = np.array(
        [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11),
         (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6), (7, 1, 8),
         (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9),
         (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere_unit(level: int):
    """Unit icosphere after ``level`` 4-way subdivisions: 10*4^level+2 vertices,
    20*4^level faces (level 5: 20,480 faces; level 8: 1,310,720)."""
    v, f = _icosahedron()
    for _ in range(level):
        nv = len(v)
        e = np.stack([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=1).reshape(-1, 2)
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = (inv.reshape(-1, 3) + nv)
        v = np.concatenate([v, mid], axis=0)
        f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.stack([
            np.stack([f0, m01, m20], 1), np.stack([f1, m12, m01], 1),
            np.stack([f2, m20, m12], 1), np.stack([m01, m12, m20], 1)], axis=1).reshape(-1, 3)
    return v, f


def displaced_icosphere(level: int = 5, radius: float = 0.4, displacement: float = 0.05,
                        seed: int = 0, center=(0.0, 0.0, 0.0)):
    """Icosphere whose vertex i sits at radius + U(-displacement, displacement) along its
    direction, U drawn from SplitMix64(seed) in vertex order."""
    v, f = icosphere_unit(level)
    r = radius + splitmix_uniform(seed, len(v), -displacement, displacement)
    pos = v * r[:, None] + np.asarray(center, dtype=np.float64)
    return pos.astype(np.float32), f.astype(np.uint32)


def concat_meshes(*meshes):
    """Concatenate (verts, faces) pairs; face indices of later meshes are offset, face order
    is kept (face index is the hit-ID and the tie-break key)."""
    vs, fs, off = [], [], 0
    for v, f in meshes:
        vs.append(np.asarray(v, dtype=np.float32).reshape(-1, 3))
        fs.append(np.asarray(f, dtype=np.uint32).reshape(-1, 3) + np.uint32(off))
        off += len(vs[-1])
    return np.concatenate(vs).astype(np.float32), np.concatenate(fs).astype(np.uint32)


def benchmark_scene(level: int = 5, seed: int = 0):
    """Centred Cornell box plus the displaced icosphere (radius 0.4, U(-0.05,0.05)) at the
    origin: 36 + 20*4^level triangles."""
    return concat_meshes(cornell_box_centered(), displaced_icosphere(level, 0.4, 0.05, seed))
