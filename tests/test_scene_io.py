"""Scene / configuration front end in the reference's JSON schema
(simulator/base.py:31-45, simulator/scene.py:23-50, affine_body.py:59-67)."""
import json
import os

import numpy as np
import pytest

from abd_hip import scene_io, scenes


def test_obj_reader_handles_corner_forms_polygons_and_negative_indices(tmp_path):
    p = tmp_path / "quad.obj"
    p.write_text("# a unit square as one quad and a triangle with negative indices\n"
                 "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nvt 0 0\nvn 0 0 1\n"
                 "f 1/1/1 2//1 3/1 4\n"
                 "f -4 -3 -1\n")
    V, F = scene_io.read_obj(str(p))
    assert V.shape == (4, 3) and np.array_equal(V[2], [1, 1, 0])
    # fan from the first corner, then the triangle (1, 2, 4) in 0-based ids
    assert F.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 3]]


def test_obj_reader_rejects_out_of_range_faces(tmp_path):
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nf 1 2 3\n")
    with pytest.raises(ValueError):
        scene_io.read_obj(str(p))


def test_reference_schema_scene_with_procedural_and_obj_meshes(tmp_path):
    (tmp_path / "assets").mkdir()
    V, F = scenes.cube_mesh()
    scene_io.write_obj(str(tmp_path / "assets" / "cube.obj"), V, F)
    (tmp_path / "scenes").mkdir()
    cfg = {"objects": [
        {"file": "cube.obj", "folder": "assets/", "p": [0.0, 0.1, 0.0],
         "Adot": [[0.0, 1.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]},
        {"file": "box", "p": [0.6, 3.0, 0.15], "mass": 500.0, "I0": (np.eye(3) / 12).tolist()},
        {"file": "sphere", "p": [3.0, 1.0, 0.0], "pdot": [1.0, 0.0, 0.0]},
    ]}
    (tmp_path / "scenes" / "s.json").write_text(json.dumps(cfg))
    spec = scene_io.load_scene(str(tmp_path / "scenes" / "s.json"))
    assert spec.n_bodies == 3 and spec.n_scenes == 1
    assert np.array_equal(spec.x0[:8], V) and np.array_equal(spec.x0[8:16], V)
    assert spec.voff[3] - spec.voff[2] == 162                    # icosphere-2
    assert np.array_equal(spec.Adot[0], cfg["objects"][0]["Adot"])
    assert np.array_equal(spec.pdot[2], [1.0, 0.0, 0.0])
    assert spec.mass.tolist() == [1e3, 500.0, 1e3]               # D13: reference constants by default
    assert np.isclose(spec.I0[1], 1.0 / 12) and spec.I0[0] == 1e2


def test_bgeo_and_usd_meshes_are_refused_with_a_clear_error(tmp_path):
    (tmp_path / "scenes").mkdir()
    (tmp_path / "scenes" / "b.json").write_text(json.dumps({"objects": [{"file": "box.bgeo", "folder": "assets/"}]}))
    with pytest.raises(NotImplementedError):
        scene_io.load_scene(str(tmp_path / "scenes" / "b.json"))


def test_save_load_round_trip_and_config(tmp_path):
    spec = scenes.config_small()
    (tmp_path / "scenes").mkdir()
    path = str(tmp_path / "scenes" / "small.json")
    scene_io.save_scene(spec, path)
    back = scene_io.load_scene(path)
    for f in ("x0", "tris", "edges", "voff", "foff", "eoff", "A", "p", "Adot", "pdot", "mass", "I0"):
        assert np.array_equal(getattr(back, f), getattr(spec, f)), f
    assert back.ground == spec.ground
    (tmp_path / "config.json").write_text(json.dumps({"scene": "scenes/small.json", "frame_out": 7, "dt": 0.005,
                                                      "kappa": 1e6, "max_iter": 20, "verbose": 1}))
    cfg = scene_io.load_config(str(tmp_path / "config.json"))
    assert cfg.frame_out == 7 and cfg.params.dt == 0.005 and cfg.params.kappa == 1e6 and cfg.params.max_iter == 20
    assert cfg.params.ground == spec.ground and cfg.extra == {"verbose": 1}
    assert cfg.scene.n_bodies == 2
