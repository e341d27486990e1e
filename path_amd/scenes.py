"""Procedural benchmark scenes written in the reference's on-disk formats
(CS123 XML scene file + OBJ + MTL), so that every render goes through the
scene loader.

- cornell_box(): 36 triangles. Five diffuse walls (white 0.73, red
  0.65/0.05/0.05, green 0.12/0.45/0.15), two axis-aligned white boxes and a
  0.25 x 0.25 quad emitter of radiance (17, 12, 4) just below the ceiling.
- cornell_icosphere(): the box plus a level-6 icosphere (81,920 triangles)
  with seeded per-vertex radius jitter of +-5%; materials by triangle index
  mod 3: diffuse, glossy (Phong exponent 50), dielectric (IOR 1.5).

The camera follows the reference scenes (scenes/cornell-diffuse.xml): position
(0, 1, 3.6), focus (0, 1, 0), up +y, height angle 45 degrees.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

CAMERA = {"pos": (0.0, 1.0, 3.6), "focus": (0.0, 1.0, 0.0), "up": (0.0, 1.0, 0.0),
          "heightangle": 45.0}


def _quad(a, b, c, d):
    """Quad (a, b, c, d) counter-clockwise seen from the side its normal points to."""
    return [(a, b, c, d)]


def _box_quads(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    # outward normals
    return [
        ((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)),  # top +y
        ((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # bottom -y
        ((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # front +z
        ((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)),  # back -z
        ((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)),  # right +x
        ((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # left -x
    ]


def cornell_quads():
    """[(material_name, quad)] with inward-facing walls."""
    q = []
    q.append(("floor", ((-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1))))
    q.append(("ceiling", ((-1, 2, -1), (1, 2, -1), (1, 2, 1), (-1, 2, 1))))
    q.append(("backWall", ((-1, 0, -1), (1, 0, -1), (1, 2, -1), (-1, 2, -1))))
    q.append(("leftWall", ((-1, 0, -1), (-1, 2, -1), (-1, 2, 1), (-1, 0, 1))))
    q.append(("rightWall", ((1, 0, -1), (1, 0, 1), (1, 2, 1), (1, 2, -1))))
    for quad in _box_quads((0.15, 0.0, 0.0), (0.75, 0.6, 0.6)):
        q.append(("shortBox", quad))
    for quad in _box_quads((-0.75, 0.0, -0.75), (-0.15, 1.2, -0.15)):
        q.append(("tallBox", quad))
    h = 0.125
    # light faces down (-y)
    q.append(("light", ((-h, 1.98, -h), (h, 1.98, -h), (h, 1.98, h), (-h, 1.98, h))))
    return q


CORNELL_MTL = """newmtl floor
  illum 2
  Kd 0.73 0.73 0.73
newmtl ceiling
  illum 2
  Kd 0.73 0.73 0.73
newmtl backWall
  illum 2
  Kd 0.73 0.73 0.73
newmtl leftWall
  illum 2
  Kd 0.65 0.05 0.05
newmtl rightWall
  illum 2
  Kd 0.12 0.45 0.15
newmtl shortBox
  illum 2
  Kd 0.73 0.73 0.73
newmtl tallBox
  illum 2
  Kd 0.73 0.73 0.73
newmtl light
  illum 2
  Kd 0.78 0.78 0.78
  Ke 17 12 4
"""


def _write_cornell_obj(path: Path):
    lines = ["mtllib cornell.mtl"]
    cur = None
    for name, quad in cornell_quads():
        for v in quad:
            lines.append("v %r %r %r" % tuple(float(c) for c in v))
        if name != cur:
            lines.append(f"usemtl {name}")
            cur = name
        lines.append("f -4 -3 -2 -1")
    path.write_text("\n".join(lines) + "\n")


def icosphere(level: int, radius: float, center, jitter: float, seed: int):
    """Subdivided icosahedron; each vertex radius scaled by U[1-jitter, 1+jitter]."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
             (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
             (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
             (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    v = np.array(verts)
    rng = np.random.default_rng(seed)
    scale = rng.uniform(1.0 - jitter, 1.0 + jitter, size=len(v))
    v = v * (radius * scale)[:, None] + np.asarray(center, np.float64)[None, :]
    return v.astype(np.float32), np.array(faces, np.int64)


ICOSPHERE_MTL = """newmtl sphereDiffuse
  illum 2
  Kd 0.73 0.73 0.73
newmtl sphereGlossy
  illum 2
  Ks 0.73 0.73 0.73
  Ns 50
newmtl sphereGlass
  illum 5
  Ni 1.5
"""


def _write_icosphere_obj(path: Path, level: int, seed: int):
    v, f = icosphere(level, 0.35, (0.3, 1.0, 0.15), 0.05, seed)
    names = ["sphereDiffuse", "sphereGlossy", "sphereGlass"]
    out = ["mtllib icosphere.mtl"]
    out += ["v %.9g %.9g %.9g" % tuple(p) for p in v]
    for i, (a, b, c) in enumerate(f):
        out.append("usemtl " + names[i % 3])
        out.append(f"f {a + 1} {b + 1} {c + 1}")
    path.write_text("\n".join(out) + "\n")


def _write_xml(path: Path, objs):
    c = CAMERA
    objects = "\n".join(
        f'      <object type="primitive" name="mesh" filename="{o}"> </object>' for o in objs)
    path.write_text(f"""<scenefile>
  <cameradata>
    <pos x="{c['pos'][0]}" y="{c['pos'][1]}" z="{c['pos'][2]}"/>
    <up x="{c['up'][0]}" y="{c['up'][1]}" z="{c['up'][2]}"/>
    <focus x="{c['focus'][0]}" y="{c['focus'][1]}" z="{c['focus'][2]}"/>
    <heightangle v="{c['heightangle']}"/>
  </cameradata>
  <object type="tree" name="root">
    <transblock>
      <translate x="0" y="0" z="0"/>
{objects}
    </transblock>
  </object>
</scenefile>
""")


def cornell_box(directory) -> Path:
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    (d / "cornell.mtl").write_text(CORNELL_MTL)
    _write_cornell_obj(d / "cornell.obj")
    _write_xml(d / "cornell.xml", ["cornell.obj"])
    return d / "cornell.xml"


def cornell_icosphere(directory, level: int = 6, seed: int = 0) -> Path:
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    (d / "cornell.mtl").write_text(CORNELL_MTL)
    _write_cornell_obj(d / "cornell.obj")
    (d / "icosphere.mtl").write_text(ICOSPHERE_MTL)
    _write_icosphere_obj(d / "icosphere.obj", level, seed)
    name = f"cornell_icosphere_l{level}_s{seed}.xml"
    _write_xml(d / name, ["cornell.obj", "icosphere.obj"])
    return d / name
