"""Separation checks of a simulated state, for the long property runs (test
infrastructure, CPU numpy).

Textbook geometry, independent of the oracle and of the library: point-triangle
distance (the plane distance when the projection falls inside the triangle,
else the nearest of its three edges), segment-segment distance (the clamped
closest-point parameters of Ericson, Real-Time Collision Detection, 2005,
section 5.1.9) and segment-triangle crossing (Moller and Trumbore, 1997).

`separation(spec, x, ground)` reports, over every pair of bodies whose boxes
come within `pad`, the smallest point-triangle and edge-edge distance between
the two bodies, the number of edges of one body crossing a triangle of the
other (an interpenetration), and the smallest vertex height above the ground.
Only primitives inside the other body's padded box take part: a primitive
outside it is farther than `pad` from every primitive of that body.
"""
from __future__ import annotations

import numpy as np


def point_segment_dist(p, a, b):
    ab = b - a
    t = np.einsum("...k,...k->...", p - a, ab) / np.maximum(np.einsum("...k,...k->...", ab, ab), 1e-300)
    t = np.clip(t, 0.0, 1.0)
    return np.linalg.norm(p - (a + t[..., None] * ab), axis=-1)


def point_triangle_dist(p, t0, t1, t2):
    """Distance of points p [..., 3] to triangles (t0, t1, t2) [..., 3] (broadcast)."""
    e0, e1 = t1 - t0, t2 - t0
    n = np.cross(e0, e1)
    nn = np.einsum("...k,...k->...", n, n)
    w = p - t0
    # barycentric coordinates of the projection
    d00 = np.einsum("...k,...k->...", e0, e0)
    d01 = np.einsum("...k,...k->...", e0, e1)
    d11 = np.einsum("...k,...k->...", e1, e1)
    d20 = np.einsum("...k,...k->...", w, e0)
    d21 = np.einsum("...k,...k->...", w, e1)
    den = d00 * d11 - d01 * d01
    ok = den > 0
    v = np.where(ok, (d11 * d20 - d01 * d21) / np.where(ok, den, 1.0), -1.0)
    u = np.where(ok, (d00 * d21 - d01 * d20) / np.where(ok, den, 1.0), -1.0)
    inside = (v >= 0) & (u >= 0) & (v + u <= 1)
    plane = np.abs(np.einsum("...k,...k->...", w, n)) / np.sqrt(np.maximum(nn, 1e-300))
    edge = np.minimum(np.minimum(point_segment_dist(p, t0, t1), point_segment_dist(p, t1, t2)),
                      point_segment_dist(p, t2, t0))
    return np.where(inside, plane, edge)


def segment_segment_dist(p1, q1, p2, q2):
    d1, d2, r = q1 - p1, q2 - p2, p1 - p2
    a = np.einsum("...k,...k->...", d1, d1)
    e = np.einsum("...k,...k->...", d2, d2)
    f = np.einsum("...k,...k->...", d2, r)
    c = np.einsum("...k,...k->...", d1, r)
    b = np.einsum("...k,...k->...", d1, d2)
    den = a * e - b * b
    s = np.where(den > 1e-20 * a * e, np.clip((b * f - c * e) / np.where(den > 0, den, 1.0), 0.0, 1.0), 0.0)
    t = (b * s + f) / e
    lo, hi = t < 0, t > 1
    s = np.where(lo, np.clip(-c / a, 0.0, 1.0), np.where(hi, np.clip((b - c) / a, 0.0, 1.0), s))
    t = np.clip(t, 0.0, 1.0)
    return np.linalg.norm(p1 + s[..., None] * d1 - (p2 + t[..., None] * d2), axis=-1)


def segment_crosses_triangle(p, q, t0, t1, t2):
    """True where segment pq passes through the interior of triangle (t0, t1, t2)."""
    e1, e2, d = t1 - t0, t2 - t0, q - p
    h = np.cross(d, e2)
    a = np.einsum("...k,...k->...", e1, h)
    ok = np.abs(a) > 1e-300
    f = 1.0 / np.where(ok, a, 1.0)
    s = p - t0
    u = f * np.einsum("...k,...k->...", s, h)
    qv = np.cross(s, e1)
    v = f * np.einsum("...k,...k->...", d, qv)
    t = f * np.einsum("...k,...k->...", e2, qv)
    return ok & (u >= 0) & (v >= 0) & (u + v <= 1) & (t >= 0) & (t <= 1)


def _in_box(pts_lo, pts_hi, lo, hi):
    return np.all(pts_hi >= lo, axis=-1) & np.all(pts_lo <= hi, axis=-1)


def separation(spec, x, ground, pad=0.01, max_pairs=None):
    """min point-triangle / edge-edge distance between bodies, edge-triangle crossings
    and min vertex height over the ground at positions x [V, 3]."""
    nb = spec.n_bodies
    voff, foff, eoff = spec.voff, spec.foff, spec.eoff
    blo = np.array([x[voff[b]:voff[b + 1]].min(0) for b in range(nb)]) - pad
    bhi = np.array([x[voff[b]:voff[b + 1]].max(0) for b in range(nb)]) + pad
    scene = np.asarray(spec.body_scene)
    # candidate body pairs: boxes overlapping after padding, same scene
    pairs = []
    order = np.argsort(blo[:, 0])
    for k, i in enumerate(order):
        for j in order[k + 1:]:
            if blo[j, 0] > bhi[i, 0]:
                break
            if scene[i] == scene[j] and np.all(blo[i] <= bhi[j]) and np.all(blo[j] <= bhi[i]):
                pairs.append((min(i, j), max(i, j)))
    if max_pairs is not None:
        pairs = pairs[:max_pairs]
    min_pt, min_ee, cross = np.inf, np.inf, 0

    def tris_of(b):
        return spec.tris[foff[b]:foff[b + 1]] + voff[b]

    def edges_of(b):
        return spec.edges[eoff[b]:eoff[b + 1]] + voff[b]

    for (i, j) in pairs:
        for (I, J) in ((i, j), (j, i)):
            V = np.arange(voff[I], voff[I + 1])
            V = V[_in_box(x[V], x[V], blo[J], bhi[J])]
            F = tris_of(J)
            tf = x[F]
            F = F[_in_box(tf.min(1), tf.max(1), blo[I], bhi[I])]
            E = edges_of(I)
            te = x[E]
            E = E[_in_box(te.min(1), te.max(1), blo[J], bhi[J])]
            if len(V) and len(F):
                d = point_triangle_dist(x[V][:, None], x[F[:, 0]][None], x[F[:, 1]][None], x[F[:, 2]][None])
                min_pt = min(min_pt, float(d.min()))
            if len(E) and len(F):
                c = segment_crosses_triangle(x[E[:, 0]][:, None], x[E[:, 1]][:, None], x[F[:, 0]][None],
                                             x[F[:, 1]][None], x[F[:, 2]][None])
                cross += int(c.sum())
        Ei, Ej = edges_of(i), edges_of(j)
        ti, tj = x[Ei], x[Ej]
        Ei = Ei[_in_box(ti.min(1), ti.max(1), blo[j], bhi[j])]
        Ej = Ej[_in_box(tj.min(1), tj.max(1), blo[i], bhi[i])]
        if len(Ei) and len(Ej):
            d = segment_segment_dist(x[Ei[:, 0]][:, None], x[Ei[:, 1]][:, None], x[Ej[:, 0]][None],
                                     x[Ej[:, 1]][None])
            min_ee = min(min_ee, float(d.min()))
    return dict(pairs=len(pairs), min_pt=min_pt, min_ee=min_ee, crossings=cross,
                min_vg=float((x[:, 1] - ground).min()))
