"""Scene and configuration files in the reference's JSON schema.

config.json (simulator/base.py:31-45):
    {"scene": "scenes/case1.json", "frame_in": 0, "frame_out": 100, ...,
     "dt": 0.01, "dhat": 1e-3, "kappa": 1e7, "stiffness": 1e8, "max_iter": 10, "tol": 1e-4}
  The simulator arguments of abd.py:107-121 are honoured here (the reference
  shadows them with module constants, D7); unknown keys are kept in `extra`.

scenes/*.json (simulator/scene.py:23-50, affine_body.py:59-67):
    {"objects": [{"file": "box.obj", "folder": "assets/", "p": [..], "A": [[..]],
                  "Adot": [[..]], "pdot": [..], "mass": m, "I0": i}]}
  Meshes: Wavefront OBJ files (read here; polygons fan-triangulated as
  igl.read_triangle_mesh does) or procedural names: "box" / "cube" (unit cube
  [-0.5, 0.5]^3), "sphere" (icosphere-2, radius 0.5), "icosphere-N" (radius 0.5),
  "box-N" (N x N x N subdivided unit box).  The reference's bgeo and USD readers
  depend on Houdini / pxr modules that are not part of this build: such files
  raise NotImplementedError.  "mass" and "I0" (a scalar, or a 3x3 matrix whose
  diagonal mean is used) set the per-body values (D13); absent, the reference's
  kernel constants m = 1e3, I0 = 1e2 apply.
"""
from __future__ import annotations

import dataclasses
import json
import os
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import scenes
from .const_params import ContactParams
from .scenes import SceneSpec


def read_obj(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """Vertices [V,3] float64 and triangles [F,3] int32 of a Wavefront OBJ file
    (`v` and `f` records; `f` corners may be `v`, `v/vt`, `v//vn`, `v/vt/vn`,
    negative indices count from the end; polygons are fanned from the first corner)."""
    V: List[List[float]] = []
    F: List[List[int]] = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            if parts[0] == "v":
                V.append([float(parts[1]), float(parts[2]), float(parts[3])])
            elif parts[0] == "f":
                idx = []
                for c in parts[1:]:
                    k = int(c.split("/")[0])
                    idx.append(k - 1 if k > 0 else len(V) + k)
                for j in range(1, len(idx) - 1):
                    F.append([idx[0], idx[j], idx[j + 1]])
    Vn = np.asarray(V, np.float64).reshape(-1, 3)
    Fn = np.asarray(F, np.int32).reshape(-1, 3)
    if Fn.size and (Fn.min() < 0 or Fn.max() >= Vn.shape[0]):
        raise ValueError(f"{path}: face index out of range")
    return Vn, Fn


def write_obj(path: str, V: np.ndarray, F: np.ndarray) -> None:
    with open(path, "w") as fh:
        for v in V:
            fh.write(f"v {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
        for f in F:
            fh.write(f"f {f[0] + 1} {f[1] + 1} {f[2] + 1}\n")


def _mesh_of(obj: Dict, base_dir: str) -> Tuple[np.ndarray, np.ndarray]:
    file = obj.get("file", "assets/stanford-bunny.obj")
    folder = obj.get("folder", "")
    name = file.lower()
    if name in ("box", "cube"):
        return scenes.cube_mesh()
    if name == "sphere":
        return scenes.icosphere_mesh(2, 0.5)
    if name.startswith("icosphere-"):
        return scenes.icosphere_mesh(int(name.split("-")[1]), 0.5)
    if name.startswith("box-"):
        return scenes.subdivided_box_mesh(int(name.split("-")[1]))
    if name.endswith((".bgeo", ".usd", ".usda", ".usdz")):
        raise NotImplementedError(f"{file}: bgeo / USD meshes need the Houdini / pxr readers, not part of this "
                                  "build; convert the mesh to OBJ")
    path = os.path.join(base_dir, folder + file)
    return read_obj(path)


def _scalar_inertia(v) -> float:
    a = np.asarray(v, np.float64)
    return float(a) if a.ndim == 0 else float(np.mean(np.diag(a.reshape(3, 3))))


def load_scene(path: str, ground: Optional[float] = None, base_dir: Optional[str] = None) -> SceneSpec:
    """A scenes/*.json file as a SceneSpec (one independent scene).  `base_dir`
    (default: the directory holding the JSON file's parent, where the reference
    keeps assets/) resolves relative mesh paths."""
    with open(path) as fh:
        cfg = json.load(fh)
    if base_dir is None:
        base_dir = os.path.dirname(os.path.dirname(os.path.abspath(path)))
    meshes, A, p, Adot, pdot, mass, I0 = [], [], [], [], [], [], []
    for obj in cfg["objects"]:
        meshes.append(_mesh_of(obj, base_dir))
        A.append(np.asarray(obj.get("A", np.eye(3)), np.float64).reshape(3, 3))
        p.append(np.asarray(obj.get("p", np.zeros(3)), np.float64).reshape(3))
        Adot.append(np.asarray(obj.get("Adot", np.zeros((3, 3))), np.float64).reshape(3, 3))
        pdot.append(np.asarray(obj.get("pdot", np.zeros(3)), np.float64).reshape(3))
        mass.append(float(obj["mass"]) if "mass" in obj else 1e3)
        I0.append(_scalar_inertia(obj["I0"]) if "I0" in obj else 1e2)
    g = float(cfg.get("ground", -0.5 - 0.5 * scenes.DHAT)) if ground is None else ground
    name = os.path.splitext(os.path.basename(path))[0]
    return scenes.build_scene(meshes, np.stack(A), np.stack(p), np.stack(Adot), np.stack(pdot), g,
                              mass=np.array(mass), I0=np.array(I0), name=name)


def save_scene(spec: SceneSpec, path: str, mesh_dir: Optional[str] = None) -> None:
    """Writes a SceneSpec in the same schema: one OBJ per distinct mesh under
    `mesh_dir` (default: assets/ beside the JSON file's directory)."""
    out_dir = os.path.dirname(os.path.abspath(path))
    base = os.path.dirname(out_dir)
    mesh_dir = mesh_dir or os.path.join(base, "assets")
    os.makedirs(mesh_dir, exist_ok=True)
    seen: Dict[bytes, str] = {}
    objs = []
    for b in range(spec.n_bodies):
        V = spec.x0[spec.voff[b]:spec.voff[b + 1]]
        F = spec.tris[spec.foff[b]:spec.foff[b + 1]]
        key = V.tobytes() + F.tobytes()
        if key not in seen:
            fname = f"{spec.name}_mesh{len(seen)}.obj"
            write_obj(os.path.join(mesh_dir, fname), V, F)
            seen[key] = fname
        objs.append({"file": seen[key], "folder": os.path.relpath(mesh_dir, base) + "/",
                     "p": spec.p[b].tolist(), "A": spec.A[b].tolist(), "Adot": spec.Adot[b].tolist(),
                     "pdot": spec.pdot[b].tolist(), "mass": float(spec.mass[b]), "I0": float(spec.I0[b])})
    with open(path, "w") as fh:
        json.dump({"objects": objs, "ground": spec.ground}, fh, indent=1)


@dataclasses.dataclass
class SimConfig:
    scene: SceneSpec
    params: ContactParams
    frame_in: int = 0
    frame_out: int = 100
    extra: Dict = dataclasses.field(default_factory=dict)


_PARAM_KEYS = ("dt", "dhat", "kappa", "stiffness", "gravity", "ground", "mass", "I0", "c1", "alpha_min", "tol",
               "max_iter", "solver_tol", "solver_maxiter", "pt", "ee", "vg", "solver", "psd")


def load_config(path: str) -> SimConfig:
    """config.json -> scene, runtime parameters and frame range.  The scene path is
    resolved against the config file's directory."""
    with open(path) as fh:
        cfg = json.load(fh)
    root = os.path.dirname(os.path.abspath(path))
    kw = {k: cfg[k] for k in _PARAM_KEYS if k in cfg}
    spec = load_scene(os.path.join(root, cfg["scene"]), ground=kw.get("ground"), base_dir=root)
    kw.setdefault("ground", spec.ground)
    extra = {k: v for k, v in cfg.items() if k not in _PARAM_KEYS and k not in ("scene", "frame_in", "frame_out")}
    return SimConfig(spec, ContactParams(**kw), int(cfg.get("frame_in", 0)), int(cfg.get("frame_out", 100)), extra)
