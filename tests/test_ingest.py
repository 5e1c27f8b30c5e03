"""Mesh ingest from pbrt files (gonzales_amd/ingest.py) on the reference's own regression-test
scene (tests/golden/regression-test.pbrt, the file of Tests/reference) and on small files written
here: shapes, materials, area lights, transforms and render settings."""
from pathlib import Path

import numpy as np
import pytest

from gonzales_amd import abi
from gonzales_amd.camera import pbrt_transform
from gonzales_amd.ingest import ingest, tokenize, transform_points

from test_ply import LAYOUTS, write_ply

GOLDEN = Path(__file__).resolve().parent / "golden" / "regression-test.pbrt"


def test_regression_scene_shapes_materials_and_light():
    s = ingest(str(GOLDEN))
    meshes = s.scene.meshes
    assert len(meshes) == 8  # floor, ceiling, back, right, left walls, short box, tall box, light
    tris = [m.number_triangles for m in meshes]
    assert tris == [2, 2, 2, 2, 2, 12, 12, 2]
    # the floor's points as the file writes them (identity transform inside WorldBegin)
    floor = meshes[0].points[:, :3]
    assert np.array_equal(floor, np.array([[-1, 1.74846e-7, -1], [-1, 1.74846e-7, 1], [1, -1.74846e-7, 1],
                                           [1, -1.74846e-7, -1]], np.float32))
    assert meshes[0].vertex_indices.tolist() == [0, 1, 2, 0, 2, 3]
    assert meshes[0].normals is not None and meshes[0].uvs is not None
    # one material per non-light shape, in shape order: Floor, Ceiling, BackWall, RightWall,
    # LeftWall, ShortBox, TallBox
    m = s.materials
    assert len(m) == 7 and np.all(m["type"] == abi.GZ_MAT_DIFFUSE)
    assert np.allclose(m["reflectance"][3], (0.14, 0.45, 0.091))
    assert np.allclose(m["reflectance"][4], (0.63, 0.065, 0.05))
    assert np.allclose(m["reflectance"][0], (0.725, 0.71, 0.68))
    # the light quad: two area-light triangles of radiance (17, 12, 4)
    kinds = s.prims.kind
    assert (kinds == 3).sum() == 2 and np.array_equal(s.scene.light_emission, [[17, 12, 4], [17, 12, 4]])
    assert np.array_equal(s.prims.material[kinds == 1], np.repeat(np.arange(7), tris[:7]))
    # render settings
    assert s.fov == 19.5 and s.resolution == (128, 128) and s.spp == 64 and s.max_depth == 4
    assert s.filter == {"type": "gaussian", "radius": (5.0, 5.0), "sigma": 2.5}
    assert np.array_equal(s.world_to_camera, pbrt_transform([1, -0, -0, -0, -0, 1, -0, -0, -0, -0, -1, -0, -0, -1,
                                                             6.8, 1]))
    cam = s.camera()
    assert cam.c.resX == 128 and cam.spp == 64


def test_transform_points_float32_order_and_weight():
    m = np.array([[2, 0, 0, 1], [0, 3, 0, -2], [0, 0, 1, 0.5], [0, 0, 0, 1]], np.float64)
    p = np.array([[1, 2, 3], [0.1, 0.2, 0.3]], np.float32)
    got = transform_points(m, p)
    f = np.float32
    want0 = ((f(2) * f(1) + f(0) * f(2)) + f(0) * f(3)) + f(1)
    assert got[0, 0] == want0 and got[0, 1] == f(4.0) and got[0, 2] == f(3.5)
    assert np.array_equal(transform_points(np.eye(4), p), p)
    # a projective row divides by the weight
    proj = np.eye(4)
    proj[3] = (0, 0, 1, 0)
    assert np.allclose(transform_points(proj, np.array([[2, 4, 2]], np.float32)), [[1, 2, 1]])


def test_tokenizer_and_transform_stack(tmp_path):
    assert tokenize('Shape "trianglemesh" # comment\n "point3 P" [ 0 0 0 ]') == \
        ['Shape', '"trianglemesh"', '"point3 P"', '[', '0', '0', '0', ']']
    data = write_ply(LAYOUTS["xyz"], np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), [[0, 1, 2]])
    (tmp_path / "tri.ply").write_bytes(data)
    (tmp_path / "s.pbrt").write_text('''
WorldBegin
Material "conductor" "float roughness" [ 0.2 ]
AttributeBegin
  Translate 1 2 3
  Scale 2 2 2
  Shape "plymesh" "string filename" "tri.ply"
AttributeEnd
Shape "trianglemesh" "point3 P" [ 0 0 0 1 0 0 0 1 0 ] "integer indices" [ 0 1 2 ]
Shape "sphere" "float radius" 1
LightSource "infinite" "rgb L" [ 0.1 0.1 0.1 ]
AttributeBegin
  AreaLightSource "diffuse" "rgb L" [ 5 5 5 ]
  Rotate 90 0 0 1
  Shape "trianglemesh" "point3 P" [ 0 0 0 1 0 0 0 1 0 ] "integer indices" [ 0 1 2 ]
AttributeEnd
''')
    s = ingest(str(tmp_path / "s.pbrt"))
    assert len(s.scene.meshes) == 3 and s.skipped_shapes == ["sphere"]
    assert np.allclose(s.scene.meshes[0].points[:, :3], [[1, 2, 3], [3, 2, 3], [1, 4, 3]])
    assert np.allclose(s.scene.meshes[1].points[:, :3], [[0, 0, 0], [1, 0, 0], [0, 1, 0]])
    assert np.allclose(s.scene.meshes[2].points[:, :3], [[0, 0, 0], [0, 1, 0], [-1, 0, 0]], atol=1e-6)
    assert len(s.materials) == 2 and np.all(s.materials["type"] == abi.GZ_MAT_CONDUCTOR)
    assert np.allclose(s.materials["roughness"][:, 0], 0.2)
    assert s.scene.light_emission.tolist() == [[5, 5, 5]]
    assert s.world_to_camera is None


@pytest.mark.parametrize("text,msg", [
    ('WorldBegin\nMaterial "conductor" "spectrum eta" "metal-Au-eta"\n', "named spectrum"),
    ('WorldBegin\nMaterial "hair"\n', "not supported"),
    ('WorldBegin\nObjectBegin "x"\n', "instancing"),
    ('WorldBegin\nNamedMaterial "nope"\n', "unknown named material"),
])
def test_unsupported_input_is_rejected(tmp_path, text, msg):
    (tmp_path / "bad.pbrt").write_text(text)
    with pytest.raises(ValueError, match=msg):
        ingest(str(tmp_path / "bad.pbrt"))
