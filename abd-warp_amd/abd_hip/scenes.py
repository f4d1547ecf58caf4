"""Procedural scenes for the affine-body contact path.

The reference loads meshes from .bgeo/.obj/.usd files through a JSON scene
(simulator/scene.py:23-60, simulator/mesh.py:8-54, affine_body.py:59-71); those
readers and assets are absent, so scenes here are generated procedurally and
seeded.  A scene is plain numpy data (`SceneSpec`) that both the GPU path and the
CPU oracle consume.

Edge ordering: the reference takes edges from ``igl.edges(F)`` (affine_body.py:114),
whose order is not pinned offline.  Here edges are the unique undirected
triangle edges ``(min, max)`` sorted lexicographically.
"""
from __future__ import annotations

import dataclasses
from typing import List, Optional, Sequence, Tuple

import numpy as np


# --------------------------------------------------------------------- meshes
def edges_of(F: np.ndarray) -> np.ndarray:
    """Unique undirected edges of a triangle list, (min, max), lexicographic."""
    e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]], axis=0)
    e = np.sort(e, axis=1)
    e = np.unique(e, axis=0)
    return e.astype(np.int32)


def cube_mesh() -> Tuple[np.ndarray, np.ndarray]:
    """Unit cube [-0.5, 0.5]^3: 8 vertices, 12 outward triangles, 18 edges.

    The bottom (y = -0.5) face is split along the (-x,-z)-(+x,+z) diagonal and
    the top (y = +0.5) face along the other one, so that two stacked copies
    have crossing face diagonals (a valid edge-edge contact at the face centre).
    """
    V = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], dtype=np.float64)
    # index = 4*ix + 2*iy + iz
    def idx(ix, iy, iz):
        return 4 * ix + 2 * iy + iz
    F = []
    # -x face (ix = 0), outward -x
    F += [[idx(0, 0, 0), idx(0, 0, 1), idx(0, 1, 1)], [idx(0, 0, 0), idx(0, 1, 1), idx(0, 1, 0)]]
    # +x face
    F += [[idx(1, 0, 0), idx(1, 1, 1), idx(1, 0, 1)], [idx(1, 0, 0), idx(1, 1, 0), idx(1, 1, 1)]]
    # -y face (bottom), diagonal (0,0,0)-(1,0,1)
    F += [[idx(0, 0, 0), idx(1, 0, 1), idx(0, 0, 1)], [idx(0, 0, 0), idx(1, 0, 0), idx(1, 0, 1)]]
    # +y face (top), diagonal (0,1,1)-(1,1,0)
    F += [[idx(0, 1, 1), idx(1, 1, 0), idx(0, 1, 0)], [idx(0, 1, 1), idx(1, 1, 1), idx(1, 1, 0)]]
    # -z face
    F += [[idx(0, 0, 0), idx(0, 1, 0), idx(1, 1, 0)], [idx(0, 0, 0), idx(1, 1, 0), idx(1, 0, 0)]]
    # +z face
    F += [[idx(0, 0, 1), idx(1, 1, 1), idx(0, 1, 1)], [idx(0, 0, 1), idx(1, 0, 1), idx(1, 1, 1)]]
    F = np.array(F, dtype=np.int32)
    return V, F


def icosphere_mesh(subdiv: int, radius: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Icosahedron subdivided `subdiv` times, vertices projected to `radius`.

    Counts: V = 10*4^n + 2, F = 20*4^n, E = 30*4^n (n=2: 162/320/480, n=3: 642/1280/1920).
    """
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in V]
    for _ in range(subdiv):
        mid = {}
        nF = []

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nF += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = nF
    return np.array(V) * radius, np.array(F, dtype=np.int32)


def subdivided_box_mesh(n: int = 6, size: float = 1.0) -> Tuple[np.ndarray, np.ndarray]:
    """Surface of an n x n x n subdivided box of edge `size` centred at 0.

    n = 6: 218 vertices, 432 triangles, 648 edges.
    """
    idx = {}
    V = []

    def vid(i, j, k):
        key = (i, j, k)
        if key not in idx:
            idx[key] = len(V)
            V.append([i / n - 0.5, j / n - 0.5, k / n - 0.5])
        return idx[key]
    F = []
    for axis in range(3):
        for side in (0, n):
            for a in range(n):
                for b in range(n):
                    def P(u, v):
                        c = [0, 0, 0]
                        c[axis] = side
                        c[(axis + 1) % 3] = u
                        c[(axis + 2) % 3] = v
                        return vid(*c)
                    q = [P(a, b), P(a + 1, b), P(a + 1, b + 1), P(a, b + 1)]
                    if side == n:
                        F += [[q[0], q[1], q[2]], [q[0], q[2], q[3]]]
                    else:
                        F += [[q[0], q[2], q[1]], [q[0], q[3], q[2]]]
    return np.array(V, dtype=np.float64) * size, np.array(F, dtype=np.int32)


def ellipsoid_mesh(rng: np.random.Generator, subdiv: int = 4) -> Tuple[np.ndarray, np.ndarray]:
    """Icosphere-4 ellipsoid with radii U[0.3, 0.8]^3 and a seeded radial perturbation (+-5%)."""
    V, F = icosphere_mesh(subdiv, 1.0)
    radii = rng.uniform(0.3, 0.8, size=3)
    pert = 1.0 + 0.05 * rng.uniform(-1.0, 1.0, size=V.shape[0])
    return V * pert[:, None] * radii[None, :], F


def random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Uniform random rotation: QR of a Gaussian 3x3 with the sign fix, det +1."""
    M = rng.normal(size=(3, 3))
    Q, R = np.linalg.qr(M)
    Q = Q * np.sign(np.diag(R))[None, :]
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    return Q


# ---------------------------------------------------------------------- scene
@dataclasses.dataclass
class SceneSpec:
    """Flattened multi-body scene (host numpy).

    Per body b: vertices x0[voff[b]:voff[b+1]] (rest shape), triangles
    tris[foff[b]:foff[b+1]] and edges edges[eoff[b]:eoff[b+1]] in body-local
    vertex indices.  Initial state per body: A (rows a1, a2, a3; x = A^T x0 + p),
    p, Adot, pdot.  body_scene[b] is the independent scene a body belongs to
    (bodies of different scenes never interact).
    """
    x0: np.ndarray
    tris: np.ndarray
    edges: np.ndarray
    voff: np.ndarray
    foff: np.ndarray
    eoff: np.ndarray
    A: np.ndarray
    p: np.ndarray
    Adot: np.ndarray
    pdot: np.ndarray
    mass: np.ndarray
    I0: np.ndarray
    body_scene: np.ndarray
    ground: float
    name: str = "scene"

    @property
    def n_bodies(self) -> int:
        return int(self.voff.shape[0] - 1)

    @property
    def n_scenes(self) -> int:
        return int(self.body_scene.max()) + 1 if self.n_bodies else 0

    @property
    def n_verts(self) -> int:
        return int(self.voff[-1])

    @property
    def n_tris(self) -> int:
        return int(self.foff[-1])

    @property
    def n_edges(self) -> int:
        return int(self.eoff[-1])


def build_scene(meshes: Sequence[Tuple[np.ndarray, np.ndarray]], A: np.ndarray, p: np.ndarray, Adot: np.ndarray,
                pdot: np.ndarray, ground: float, mass: Optional[np.ndarray] = None, I0: Optional[np.ndarray] = None,
                body_scene: Optional[np.ndarray] = None, name: str = "scene",
                mass_const: float = 1e3, I0_const: float = 1e2) -> SceneSpec:
    nb = len(meshes)
    voff = np.zeros(nb + 1, np.int32)
    foff = np.zeros(nb + 1, np.int32)
    eoff = np.zeros(nb + 1, np.int32)
    xs, fs, es = [], [], []
    for b, (V, F) in enumerate(meshes):
        E = edges_of(F)
        voff[b + 1] = voff[b] + V.shape[0]
        foff[b + 1] = foff[b] + F.shape[0]
        eoff[b + 1] = eoff[b] + E.shape[0]
        xs.append(np.asarray(V, np.float64))
        fs.append(np.asarray(F, np.int32))
        es.append(E)
    return SceneSpec(
        x0=np.ascontiguousarray(np.concatenate(xs, 0)),
        tris=np.ascontiguousarray(np.concatenate(fs, 0)),
        edges=np.ascontiguousarray(np.concatenate(es, 0)),
        voff=voff, foff=foff, eoff=eoff,
        A=np.ascontiguousarray(A, np.float64), p=np.ascontiguousarray(p, np.float64),
        Adot=np.ascontiguousarray(Adot, np.float64), pdot=np.ascontiguousarray(pdot, np.float64),
        mass=np.full(nb, mass_const) if mass is None else np.asarray(mass, np.float64),
        I0=np.full(nb, I0_const) if I0 is None else np.asarray(I0, np.float64),
        body_scene=np.zeros(nb, np.int32) if body_scene is None else np.asarray(body_scene, np.int32),
        ground=float(ground), name=name)


# -------------------------------------------------------------------- configs
DHAT = 1e-3


def config_small(dhat: float = DHAT) -> SceneSpec:
    """Plain small: two unit cubes, B at the origin, A at height 1.05, zero velocity.

    The ground sits at -0.5 - dhat/2 so that B's bottom face starts at distance
    dhat/2 (a ground exactly at -0.5 would put B at d = 0, where the barrier is
    infinite).
    """
    V, F = cube_mesh()
    A = np.stack([np.eye(3)] * 2)
    p = np.array([[0.0, 0.0, 0.0], [0.0, 1.05, 0.0]])
    z = np.zeros((2, 3))
    return build_scene([(V, F), (V, F)], A, p, np.zeros((2, 3, 3)), z, ground=-0.5 - dhat / 2, name="small")


def config_medium(seed: int = 0, n_side: int = 4, subdiv: int = 3) -> SceneSpec:
    """Plain medium: 64 icosphere-3 bodies (r 0.5) on a 4x4x4 grid, spacing 1.1.

    Velocities U[-1,1]^3, uniformly random rotations, zero angular rate.  The
    floor is y = 0 and the lowest layer's centres sit at 0.5 + 0.1.
    """
    rng = np.random.default_rng(seed)
    V, F = icosphere_mesh(subdiv, 0.5)
    n = n_side ** 3
    ps, As = [], []
    for i in range(n_side):
        for j in range(n_side):
            for k in range(n_side):
                ps.append([1.1 * i, 0.6 + 1.1 * j, 1.1 * k])
    for _ in range(n):
        As.append(random_rotation(rng))
    pdot = rng.uniform(-1.0, 1.0, size=(n, 3))
    return build_scene([(V, F)] * n, np.array(As), np.array(ps), np.zeros((n, 3, 3)), pdot, ground=0.0,
                       name="medium")


def _stack_scene(rng: np.random.Generator, n_bodies: int, meshes_lib, gap: float = 0.02):
    """Random non-overlapping stacks: bodies drop into random (x, z) cells of a
    4 x 4 grid and are stacked in y with `gap` between bounding boxes."""
    cells = {}
    meshes, ps = [], []
    for _ in range(n_bodies):
        kind = int(rng.integers(0, len(meshes_lib)))
        V, F = meshes_lib[kind]
        cx, cz = int(rng.integers(0, 4)), int(rng.integers(0, 4))
        ext = V.max(0) - V.min(0)
        y0 = cells.get((cx, cz), 0.0) + gap
        p = np.array([1.2 * cx, y0 - V[:, 1].min(), 1.2 * cz])
        cells[(cx, cz)] = y0 + ext[1]
        meshes.append((V, F))
        ps.append(p)
    return meshes, np.array(ps)


def config_batched(n_scenes: int = 256, seed0: int = 1000, nb_range=(2, 64), speed: float = 2.0) -> SceneSpec:
    """Batched: n_scenes independent scenes, scene s seeded 1000 + s; each draws
    N_b ~ U{2..64} bodies, each a cube, icosphere-2 or 6x6x6 subdivided box,
    stacked without overlap, speed U[0, speed] (2 by default) in a uniform direction."""
    lib = [cube_mesh(), icosphere_mesh(2, 0.5), subdivided_box_mesh(6)]
    meshes, ps, pd, bs = [], [], [], []
    for s in range(n_scenes):
        rng = np.random.default_rng(seed0 + s)
        nb = int(rng.integers(nb_range[0], nb_range[1] + 1))
        m, p = _stack_scene(rng, nb, lib)
        # scenes are independent; offset them in x so boxes of different scenes never touch
        p = p + np.array([10.0 * s, 0.0, 0.0])
        d = rng.normal(size=(nb, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        v = d * rng.uniform(0.0, speed, size=(nb, 1))
        meshes += m
        ps.append(p)
        pd.append(v)
        bs += [s] * nb
    nb = len(meshes)
    return build_scene(meshes, np.stack([np.eye(3)] * nb), np.concatenate(ps), np.zeros((nb, 3, 3)),
                       np.concatenate(pd), ground=0.0 - 0.5 * DHAT, body_scene=np.array(bs), name="batched")


def config_diff(seed: int = 0, n_side: int = 4) -> SceneSpec:
    """Differentiation: 16 icosphere-2 bodies (162 V, 320 F, radius 0.5) on a
    4 x 4 grid with spacing 1.5, centres 0.5 + U[0.02, 0.1] above the floor
    (y = 0), identity orientation, seeded velocities U[-0.5, 0.5]^3."""
    rng = np.random.default_rng(seed)
    V, F = icosphere_mesh(2, 0.5)
    n = n_side * n_side
    ps = []
    for i in range(n):
        ps.append([1.5 * (i % n_side), 0.5 - V[:, 1].min() - 0.5 + rng.uniform(0.02, 0.1), 1.5 * (i // n_side)])
    pdot = rng.uniform(-0.5, 0.5, size=(n, 3))
    return build_scene([(V, F)] * n, np.stack([np.eye(3)] * n), np.array(ps), np.zeros((n, 3, 3)), pdot,
                       ground=0.0, name="diff")


def diff_targets(spec: SceneSpec, seed: int = 1) -> np.ndarray:
    """Seeded targets for the differentiation loss: each body's start position
    moved by U[-0.3, 0.3]^3."""
    rng = np.random.default_rng(seed)
    return spec.p + rng.uniform(-0.3, 0.3, size=spec.p.shape)


def _packed_lattice(rng: np.random.Generator, meshes, nx: int, nz: int, gap: float = 0.02,
                    jitter: float = 0.005) -> np.ndarray:
    """Positions packing `meshes` (axis-aligned, identity orientation) into a lattice of
    rows of nx bodies along x, nz rows per layer along z and as many layers along y as
    needed: along x each body's box starts `gap` after its neighbour's; rows and layers
    are spaced by their largest box plus `gap`; the lowest layer's lowest vertex sits
    `gap` above y = 0.  Each position is jittered by U[-jitter, jitter]^3 (jitter <
    gap / 2 keeps every box apart)."""
    lo = np.array([V.min(0) for V, _ in meshes])
    hi = np.array([V.max(0) for V, _ in meshes])
    n = len(meshes)
    per_layer = nx * nz
    p = np.zeros((n, 3))
    y = None
    for k0 in range(0, n, per_layer):
        layer = np.arange(k0, min(k0 + per_layer, n))
        y = gap - lo[layer, 1].min() if y is None else y + gap - lo[layer, 1].min()
        z = None
        for r0 in range(0, len(layer), nx):
            row = layer[r0:r0 + nx]
            z = -lo[row, 2].min() if z is None else z + gap - lo[row, 2].min()
            x = 0.0
            for t, b in enumerate(row):
                x = -lo[b, 0] if t == 0 else x + gap - lo[b, 0]
                p[b] = (x, y, z)
                x += hi[b, 0]
            z += hi[row, 2].max()
        y += hi[layer, 1].max()
    return p + rng.uniform(-jitter, jitter, size=p.shape)


def _stress_meshes(rng: np.random.Generator, n_bodies: int):
    ico3 = icosphere_mesh(3, 0.5)
    cube = cube_mesh()
    return [cube if b % 3 == 0 else ico3 if b % 3 == 1 else ellipsoid_mesh(rng) for b in range(n_bodies)]


def config_stress(seed: int = 0, n_bodies: int = 1024, gap: float = 0.02, jitter: float = 0.005) -> SceneSpec:
    """Stress: 1024 bodies in equal thirds cube / icosphere-3 / icosphere-4 ellipsoid
    (radii U[0.3, 0.8]^3, radial perturbation +-5%), packed in a 10 x 10 x 10.24
    jittered lattice with `gap` between neighbouring boxes (_packed_lattice), at
    rest above a floor at y = 0; they collapse under gravity into a pile.  With a gap
    below dhat (config_stress_compressed) the bodies start in contact with their
    neighbours and the floor."""
    rng = np.random.default_rng(seed)
    meshes = _stress_meshes(rng, n_bodies)
    p = _packed_lattice(rng, meshes, 10, 10, gap, jitter)
    A = np.stack([np.eye(3)] * n_bodies)
    return build_scene(meshes, A, p, np.zeros((n_bodies, 3, 3)), np.zeros((n_bodies, 3)), ground=0.0,
                       name="stress")


def config_stress_compressed(seed: int = 0, n_bodies: int = 1024) -> SceneSpec:
    """config_stress packed with a gap of 2 dhat (jitter 2e-4): the lowest layer
    lands on the floor within two frames and each layer on the one below right
    after, so the pile is in contact within the first frames (a gap inside dhat
    would start every contact deep in the barrier)."""
    return config_stress(seed, n_bodies, gap=2 * DHAT, jitter=2e-4)


def config_stress_batched(n_scenes: int = 8, n_bodies: int = 128, seed0: int = 2000, gap: float = 0.02,
                          jitter: float = 0.005) -> SceneSpec:
    """Batched stress: n_scenes independent scenes (scene s seeded seed0 + s), each
    n_bodies of the stress mix packed in an 8 x 8 x 2 lattice with config_stress's
    layout and gap, offset 40 apart in x."""
    specs = []
    for s in range(n_scenes):
        rng = np.random.default_rng(seed0 + s)
        meshes = _stress_meshes(rng, n_bodies)
        p = _packed_lattice(rng, meshes, 8, 8, gap, jitter) + np.array([40.0 * s, 0.0, 0.0])
        specs.append(build_scene(meshes, np.stack([np.eye(3)] * n_bodies), p, np.zeros((n_bodies, 3, 3)),
                                 np.zeros((n_bodies, 3)), ground=0.0, name=f"stress-{s}"))
    return concat_scenes(specs, name="stress-batched")


def subset_scene(spec: SceneSpec, bodies: Sequence[int]) -> SceneSpec:
    """A scene made of some bodies of another (keeps their meshes and states)."""
    meshes = []
    for b in bodies:
        V = spec.x0[spec.voff[b]:spec.voff[b + 1]]
        F = spec.tris[spec.foff[b]:spec.foff[b + 1]]
        meshes.append((V, F))
    idx = np.asarray(bodies)
    # scene ids renumbered 0, 1, ... in order (bodies of a scene stay contiguous)
    _, scene = np.unique(spec.body_scene[idx], return_inverse=True)
    return build_scene(meshes, spec.A[idx], spec.p[idx], spec.Adot[idx], spec.pdot[idx], spec.ground,
                       mass=spec.mass[idx], I0=spec.I0[idx], body_scene=scene.astype(np.int32),
                       name=spec.name + "-sub")


def concat_scenes(specs: Sequence[SceneSpec], name: str = "batch") -> SceneSpec:
    """Independent scenes side by side in one SceneSpec (scene k = specs[k]; the
    scenes never interact, whatever their positions).  All must share the ground."""
    meshes, A, p, Adot, pdot, mass, I0, bs = [], [], [], [], [], [], [], []
    for k, sp in enumerate(specs):
        if sp.ground != specs[0].ground:
            raise ValueError("scenes of a batch share the ground height")
        for b in range(sp.n_bodies):
            meshes.append((sp.x0[sp.voff[b]:sp.voff[b + 1]], sp.tris[sp.foff[b]:sp.foff[b + 1]]))
        A.append(sp.A)
        p.append(sp.p)
        Adot.append(sp.Adot)
        pdot.append(sp.pdot)
        mass.append(sp.mass)
        I0.append(sp.I0)
        bs += [k] * sp.n_bodies
    return build_scene(meshes, np.concatenate(A), np.concatenate(p), np.concatenate(Adot), np.concatenate(pdot),
                       specs[0].ground, mass=np.concatenate(mass), I0=np.concatenate(I0),
                       body_scene=np.array(bs, np.int32), name=name)


def scene_of(spec: SceneSpec, s: int) -> SceneSpec:
    """Independent scene `s` of a batched SceneSpec, as a plain one-scene spec."""
    return subset_scene(spec, np.nonzero(spec.body_scene == s)[0])
