"""Independent numpy statements of reference formulas used as pins for the oracle.

Each function restates a DIFFERENT formula of the reference than the one the
oracle implements for the same quantity (e.g. the block formula of d2Psi in
psd/hl.py:20-70 versus the analytic eigensystem of psd/hl.py:156-195), or a
textbook closed form, so that agreement is evidence, not a tautology."""
from __future__ import annotations

import numpy as np


def signed_distance(e0, e1, e2):
    n = np.cross(e0, e1)
    return float(np.dot(e2, n) / np.linalg.norm(n))


def d2psi(e0o, e1o, e2o):
    """psd/hl.py:20-34 Hl blocks and :62-70 d2Psi = 2 l Hl + 2 gl gl^T (9x9)."""
    l = signed_distance(e0o, e1o, e2o)
    n0, n1, n2 = e0o @ e0o, e1o @ e1o, e2o @ e2o
    l02 = -l * np.outer(e2o, e0o) / (n2 * n0)
    l12 = -l * np.outer(e2o, e1o) / (n2 * n1)
    u0, u1 = e0o / np.sqrt(n0), e1o / np.sqrt(n1)
    block = np.eye(3) - np.outer(u0, u0) - np.outer(u1, u1)
    l00 = -l * block / n0
    l11 = -l * block / n1
    gl = [np.zeros(3), np.zeros(3), e2o * l / n2]
    H = np.zeros((9, 9))
    B = [[l00, None, l02], [None, l11, l12], [l02.T, l12.T, None]]
    for a in range(3):
        for b in range(3):
            blk = 2.0 * np.outer(gl[a], gl[b])
            if B[a][b] is not None:
                blk = blk + 2.0 * l * B[a][b]
            H[3 * a:3 * a + 3, 3 * b:3 * b + 3] = blk
    return H


def point_plane_distance(p, t0, t1, t2):
    n = np.cross(t1 - t0, t2 - t0)
    return float(abs(np.dot(p - t0, n)) / np.linalg.norm(n))


def line_line_distance(a0, a1, b0, b1):
    n = np.cross(a1 - a0, b1 - b0)
    return float(abs(np.dot(b0 - a0, n)) / np.linalg.norm(n))


def barycentric(p, t0, t1, t2):
    """Coordinates (u, v) of the projection of p: p ~ t0 + u (t1 - t0) + v (t2 - t0) (least squares)."""
    M = np.stack([t1 - t0, t2 - t0], 1)
    uv, *_ = np.linalg.lstsq(M, p - t0, rcond=None)
    return uv


def affine_map(q24, r4, nI):
    """Vertex positions from the 24 affine DOFs of bodies I (first 12) and J:
    x = p + x0_0 a1 + x0_1 a2 + x0_2 a3 with q = (p, a1, a2, a3) per body."""
    x = np.zeros((4, 3))
    for v in range(4):
        q = q24[:12] if v < nI else q24[12:]
        p, a = q[:3], q[3:].reshape(3, 3)
        x[v] = p + r4[v] @ a
    return x
