"""Scene loader (CS123 XML + OBJ/MTL), MTL -> BSDF conversion, camera,
YAML settings and the C ABI surface. CPU only."""
import re
from pathlib import Path

import numpy as np
import pytest

from path_amd import _lib as L
from path_amd.renderer import Scene, Settings, camera_from_look

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_cornell_diffuse_loads_like_the_reference():
    sc, ftw = Scene.load(GOLDEN / "scenes" / "cornell-diffuse.xml", 1.0)
    n_tris, n_mats, n_light = sc.counts()
    # 18 faces (17 quads + 1 quad light), all triangulated into 2 triangles
    assert n_tris == 36 and n_mats == 8 and n_light == 2
    flat = sc.flat()
    mats = flat["materials"]
    light = [m for m in mats[:n_mats] if m.emission[0] > 0]
    assert len(light) == 1 and list(light[0].emission) == [51, 36, 12]
    assert all(m.type == L.BSDF["Diffuse"] for m in mats[:n_mats])
    # the light quad (-0.24..0.23 x -0.22..0.16 at y=1.98): two equal-area triangles
    assert np.allclose(flat["light_cdf"], [0.5, 1.0])
    # quad (0,1,2,3) -> (0,1,2), (0,2,3): first floor triangle's vertices
    t = flat["tris"][0]
    v0, v1, v2 = t[:3], t[:3] + t[3:6], t[:3] + t[6:9]
    assert np.allclose([v0, v1, v2], [[-1.01, 0, 0.99], [1.0, 0, 0.99], [1.0, 0, -1.04]], atol=1e-6)
    # face normals for OBJ files without vn; floor normal +y
    assert np.allclose(flat["normals"][0, :3], [0, 1, 0], atol=1e-6)


def test_glass_sphere_materials():
    sc, _ = Scene.load(GOLDEN / "scenes" / "cornell-glass-sphere.xml", 1.0)
    flat = sc.flat()
    types = {m.type for m in flat["materials"][:flat["n_materials"]]}
    # leftSphere Ks 0.95 Ns 1000 -> Mirror; rightSphere illum 5 Ni 2.5 -> glass
    assert L.BSDF["Mirror"] in types and L.BSDF["DielectricRefractive"] in types
    glass = [m for m in flat["materials"][:flat["n_materials"]]
             if m.type == L.BSDF["DielectricRefractive"]][0]
    assert abs(glass.ior - 2.5) < 1e-6 and list(glass.specular) == [1, 1, 1]


def test_glossy_floor_material():
    sc, _ = Scene.load(GOLDEN / "scenes" / "cornell-glossy.xml", 1.0)
    flat = sc.flat()
    glossy = [m for m in flat["materials"][:flat["n_materials"]] if m.type == L.BSDF["Glossy"]]
    assert len(glossy) == 1 and glossy[0].shininess == 40


def test_camera_transform_closed_form():
    # pos (0,1,3.6) looking at (0,1,0), 45 degree height angle, square image
    ftw = camera_from_look([0, 0, -3.6], [0, 1, 0], [0, 1, 3.6], 45.0, 1.0).reshape(3, 4)
    assert np.allclose(ftw[:, 3], [0, 1, 3.6])
    t = np.tan(np.pi * 45 / 360)
    # film point (1, 1, -1) maps to pos + (t, t, -1) in world space
    p = ftw[:, :3] @ np.array([1, 1, -1]) + ftw[:, 3]
    assert np.allclose(p, [t, 1 + t, 2.6], atol=1e-6)


def test_missing_scene_is_an_error():
    with pytest.raises(L.PathError) as e:
        Scene.load("/nonexistent/scene.xml")
    assert e.value.code == L.PT_ERR_IO


def test_unsupported_illum_is_an_error(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl a\n  illum 7\n  Kd 0.5 0.5 0.5\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    sc = Scene()
    with pytest.raises(L.PathError) as e:
        sc.add_obj(tmp_path / "m.obj")
    assert "illum" in str(e.value)


@pytest.mark.parametrize("mtl,kind,diffuse,specular,weights", [
    # illum 2, Kd only -> Diffuse
    ("illum 2\nKd 0.2 0.4 0.6", "Diffuse", (0.2, 0.4, 0.6), None, None),
    # Ks only, Ns <= 100 -> Glossy (Phong exponent Ns)
    ("illum 2\nKs 0.3 0.3 0.3\nNs 40", "Glossy", None, (0.3, 0.3, 0.3), None),
    # Ks only, Ns > 100 -> Mirror
    ("illum 2\nKs 0.9 0.8 0.7\nNs 1000", "Mirror", None, (0.9, 0.8, 0.7), None),
    # Kd + Ks: weights by channel sums, lobes divided by their weight
    ("illum 2\nKd 0.3 0.3 0.3\nKs 0.1 0.1 0.1\nNs 20", "DiffuseGlossy",
     (0.4, 0.4, 0.4), (0.4, 0.4, 0.4), (0.75, 0.25)),
    ("illum 2\nKd 0.3 0.3 0.3\nKs 0.1 0.1 0.1\nNs 500", "DiffuseMirror",
     (0.4, 0.4, 0.4), (0.4, 0.4, 0.4), (0.75, 0.25)),
    # illum 5 -> dielectric, specular 1, ior Ni (Kd ignored)
    ("illum 5\nKd 0.5 0.5 0.5\nNi 1.33", "DielectricRefractive", None, (1, 1, 1), None),
])
def test_mtl_conversion_branches(tmp_path, mtl, kind, diffuse, specular, weights):
    """tinyobj_material_conversion.h:13-101, branch by branch."""
    (tmp_path / "m.mtl").write_text("newmtl a\n" + mtl + "\nKe 1 2 3\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    sc = Scene()
    sc.add_obj(tmp_path / "m.obj")
    m = sc.materials()[0]
    assert m.type == L.BSDF[kind]
    assert np.allclose(list(m.emission), [1, 2, 3])
    if diffuse is not None:
        assert np.allclose(list(m.diffuse), diffuse, rtol=1e-6)
    if specular is not None:
        assert np.allclose(list(m.specular), specular, rtol=1e-6)
    if weights is not None:
        assert np.allclose([m.diffuse_weight, m.glossy_weight], weights, rtol=1e-6)
    if kind == "Glossy":
        assert m.shininess == 40
    if kind == "DielectricRefractive":
        assert abs(m.ior - 1.33) < 1e-6


@pytest.mark.parametrize("mtl,msg", [
    ("illum 2\nKd 0.7 0.7 0.7\nKs 0.5 0.5 0.5", "exceeds 1"),   # always_assert
    ("illum 2\nNs 1000", "mirror"),                             # mirror without Ks
    ("illum 2", "unhandled material"),                          # neither Kd nor Ks
    ("Kd 0.5 0.5 0.5", "illum"),                                # tinyobj default illum 0
])
def test_mtl_conversion_errors(tmp_path, mtl, msg):
    (tmp_path / "m.mtl").write_text("newmtl a\n" + mtl + "\n")
    (tmp_path / "m.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    with pytest.raises(L.PathError, match=msg):
        Scene().add_obj(tmp_path / "m.obj")


def test_ear_clipping_of_a_concave_polygon(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl a\n  illum 2\n  Kd 0.5 0.5 0.5\n")
    # L-shaped hexagon in the z=0 plane
    pts = [(0, 0), (2, 0), (2, 1), (1, 1), (1, 2), (0, 2)]
    obj = "mtllib m.mtl\n" + "".join(f"v {x} {y} 0\n" for x, y in pts) + "usemtl a\nf 1 2 3 4 5 6\n"
    (tmp_path / "m.obj").write_text(obj)
    sc = Scene()
    sc.add_obj(tmp_path / "m.obj")
    flat = sc.flat()
    assert flat["tris"].shape[0] == 4
    area = sum(0.5 * abs(np.cross(t[3:6], t[6:9])[2]) for t in flat["tris"])
    assert abs(area - 3.0) < 1e-6


def _affine(translate=(0, 0, 0), axis=None, angle_deg=0.0, scale=(1, 1, 1)):
    """4x4 matrices of the CS123 transformations (Eigen Translation,
    AngleAxis with the axis as given, Scaling)."""
    m = np.eye(4)
    if axis is not None:
        a = np.asarray(axis, float)
        th = np.radians(angle_deg)
        k = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
        m[:3, :3] = np.cos(th) * np.eye(3) + np.sin(th) * k + (1 - np.cos(th)) * np.outer(a, a)
    m[:3, :3] = m[:3, :3] @ np.diag(scale)
    m[:3, 3] = translate
    return m


def test_scene_graph_transforms_matrix_and_instancing(tmp_path):
    """Nested transblocks compose on the right (scenefile_loader.cpp:70-102);
    the k-th element of <matrix> is column k (CS123XmlSceneParser.cpp:236-262);
    a node's primitives come before its subtrees; a mesh file used twice is
    loaded once and instanced; master objects are reusable subtrees."""
    (tmp_path / "m.mtl").write_text("newmtl a\n  illum 2\n  Kd 0.5 0.5 0.5\n")
    (tmp_path / "a.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    (tmp_path / "b.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 0 0 1\nv 1 1 0\nusemtl a\nf 1 2 3\n")
    xml = """<scenefile>
  <object type="tree" name="m">
    <transblock>
      <translate x="0" y="0" z="-1"/>
      <object type="primitive" name="mesh" filename="b.obj"/>
    </transblock>
  </object>
  <object type="tree" name="root">
    <transblock>
      <translate x="1" y="2" z="3"/>
      <rotate x="0" y="0" z="1" angle="90"/>
      <scale x="2" y="1" z="1"/>
      <object type="tree">
        <transblock>
          <matrix>
            <row a="0" b="1" c="0" d="0"/>
            <row a="-1" b="0" c="0" d="0"/>
            <row v1="0" v2="0" v3="1" v4="0"/>
            <row a="5" b="6" c="7" d="1"/>
          </matrix>
          <object type="primitive" name="mesh" filename="a.obj"/>
        </transblock>
      </object>
      <object type="primitive" name="mesh" filename="a.obj"/>
    </transblock>
    <transblock>
      <scale x="1" y="1" z="2"/>
      <object type="master" name="m"/>
    </transblock>
  </object>
</scenefile>
"""
    (tmp_path / "g.xml").write_text(xml)
    sc, _ = Scene.load(tmp_path / "g.xml", 1.0)
    tris = sc.flat()["tris"].astype(np.float64)
    world = np.stack([tris[:, 0:3], tris[:, 0:3] + tris[:, 3:6], tris[:, 0:3] + tris[:, 6:9]], 1)
    a = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    b = np.array([[0, 0, 0], [0, 0, 1], [1, 1, 0]], float)
    t1 = _affine(translate=(1, 2, 3)) @ _affine(axis=(0, 0, 1), angle_deg=90) @ _affine(scale=(2, 1, 1))
    mat = np.array([[0, -1, 0, 5], [1, 0, 0, 6], [0, 0, 1, 7], [0, 0, 0, 1]], float)
    t_master = _affine(scale=(1, 1, 2)) @ _affine(translate=(0, 0, -1))

    def apply(m, v):
        return v @ m[:3, :3].T + m[:3, 3]

    # primitives of the first transblock, then its subtree, then the master
    expected = np.stack([apply(t1, a), apply(t1 @ mat, a), apply(t_master, b)])
    assert world.shape == expected.shape
    assert np.allclose(world, expected, atol=1e-5)


def test_matrix_with_three_rows_is_an_error(tmp_path):
    (tmp_path / "m.mtl").write_text("newmtl a\n  illum 2\n  Kd 0.5 0.5 0.5\n")
    (tmp_path / "a.obj").write_text("mtllib m.mtl\nv 0 0 0\nv 1 0 0\nv 0 1 0\nusemtl a\nf 1 2 3\n")
    rows = "".join(f'<row a="{i}" b="0" c="0" d="0"/>' for i in range(3))
    (tmp_path / "g.xml").write_text(
        '<scenefile><object type="tree" name="root"><transblock><matrix>' + rows +
        '</matrix><object type="primitive" name="mesh" filename="a.obj"/></transblock>'
        "</object></scenefile>")
    with pytest.raises(L.PathError, match="four"):
        Scene.load(tmp_path / "g.xml", 1.0)


def test_yaml_reference_config():
    s = Settings.from_yaml(GOLDEN / "configs" / "benchmark.yaml")
    assert s.kernel_approach == L.KERNEL_APPROACH["MegaKernel"]
    assert s.max_blocks_per_launch_gpu == 4096 and s.max_blocks_per_launch_cpu == 128
    assert s.reduce_work_division.target_x_block_size == 256
    assert s.accel == L.ACCEL["SBVH"] and s.target_objects == 4
    assert s.light_sampler == L.LIGHT_SAMPLER["RandomTriangle"]
    assert s.binary_search_threshold == 1073741823
    assert s.term_prob == L.TERM_PROB["MultiplierFunc"] and s.multiplier_func_exp == 50.0
    assert s.rng == L.RNG["Uniform"] and s.max_sample_size == 256
    assert s.back_cull_emission == 1


def test_yaml_round_trip(tmp_path):
    s = Settings()
    s.kernel_approach = "Streaming"
    s.term_prob = "NIters"
    s.n_iters = 7
    s.dir_sampler = "Uniform"
    s.max_num_samples_per_launch = 12345
    p = tmp_path / "c.yaml"
    p.write_text(s.to_yaml())
    t = Settings.from_yaml(p)
    for f, _ in L.PtSettings._fields_:
        a, b = getattr(s.raw, f), getattr(t.raw, f)
        if isinstance(a, L.PtWorkDivisionSettings):
            a, b = bytes(a), bytes(b)
        assert a == b, f


def test_yaml_unknown_enum_is_an_error(tmp_path):
    p = tmp_path / "bad.yaml"
    p.write_text("settings:\n  term_prob:\n    type: Bogus\n    value: {}\n")
    with pytest.raises(L.PathError):
        Settings.from_yaml(p)


def test_host_fma_is_an_instruction():
    """The host sources use explicit fused multiply-adds (scene flattening,
    shading-record packing, the SAH builder); built for baseline x86-64 each
    one was a call into libm's fmaf, which made packing a 1024-scene batch
    about 3x slower. The library must import neither fmaf nor fma."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--undefined-only", str(L.LIB_PATH)], check=True,
                         capture_output=True, text=True).stdout
    imported = {line.split()[-1].split("@")[0] for line in out.splitlines() if line.strip()}
    assert not imported & {"fmaf", "fma"}, sorted(imported & {"fmaf", "fma"})


def test_library_exports_every_declared_entry_point():
    header = (Path(__file__).resolve().parent.parent / "include" / "path_amd.h").read_text()
    declared = set(re.findall(r"^(?:int|void|const char \*)\s*(pt_\w+)\(", header, re.M))
    assert len(declared) >= 20
    lib = L.lib()
    for name in declared:
        assert hasattr(lib, name), name
    assert declared <= set(L.SIGNATURES)
