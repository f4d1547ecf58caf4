"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

  fixed_pair.npz   the reference's only fixed contact input (ipc.py:98-103,
                   hl_entry.py:8-13) evaluated as a PT and as an EE contact with
                   dhat = 0.5 (so that it is active): geometry, eigensystem,
                   dcdx, gradient and the three 12x12 blocks;
  cubic_kat.npz    known-root cubics (cubic_roots.py:225-246 construction,
                   seeded) with the roots found on [-5, 5];
  small_frames.npz body states (A, p) after each of the first 20 frames of the
                   two-cube scene (solver tolerance 1e-12).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "abd-warp_amd")]

from abd_hip import scenes  # noqa: E402
from abd_hip.const_params import ContactParams  # noqa: E402
from oracle import oracle as O  # noqa: E402

FIXED = np.array([[0.4360, 0.0259, 0.5497], [0.4353, 0.4204, 0.3303], [0.2046, 0.6193, 0.2997],
                  [0.2668, 0.6211, 0.5291]])


def fixed_pair():
    """The fixed input is an inactive pair for both PT and EE (the closest points fall
    outside the triangle / edges), so it fixes the validity gate and, as in
    hl_entry.py:86-147, the eigensystem; seeded near-contacts fix the assembly."""
    cp = ContactParams(dhat=1e-3)
    prm = O.params(cp)
    out = {"x": FIXED, "dhat": cp.dhat, "kappa": cp.kappa}
    for kind, tag in ((0, "pt"), (1, "ee")):
        G = O.pair_geometry(kind, FIXED)
        act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, FIXED, FIXED)
        assert not act
        out.update({f"{tag}_c": np.concatenate([G["e0p"], G["e1p"], G["e2p"]]),
                    f"{tag}_bgl": np.array([G["beta"], G["gamma"], G["l"]]), f"{tag}_valid": G["valid"],
                    f"{tag}_Q": G["Q"], f"{tag}_lam": G["lam"], f"{tag}_D": G["D"], f"{tag}_s": s})
    # seeded active contacts: positions, rest positions, g24 and the 24x24 block
    rng = np.random.default_rng(7)
    for kind, tag in ((0, "pt"), (1, "ee")):
        X, R, Gs, Hs = [], [], [], []
        while len(X) < 16:
            t = rng.uniform(-1, 1, size=(4, 3))
            if kind == 0:
                e0, e1 = t[2] - t[1], t[3] - t[1]
                n = np.cross(e0, e1)
                if np.linalg.norm(n) < 0.3:
                    continue
                n /= np.linalg.norm(n)
                t[0] = t[1] + 0.3 * e0 + 0.3 * e1 + rng.uniform(-9e-4, 9e-4) * n
            else:
                d0, d1 = t[1] - t[0], t[3] - t[2]
                n = np.cross(d0, d1)
                if np.linalg.norm(n) < 0.3:
                    continue
                n /= np.linalg.norm(n)
                c = t[0] + 0.45 * d0 + rng.uniform(-9e-4, 9e-4) * n
                t[2], t[3] = c - 0.55 * d1, c + 0.45 * d1
            r = t + 0.2 * rng.uniform(-1, 1, size=(4, 3))
            act, s, g24, Hi, Hj, Hij = O.pair_gh(kind, prm, t, r)
            if not act:
                continue
            X.append(t)
            R.append(r)
            Gs.append(g24)
            Hs.append(np.block([[Hi, Hij], [Hij.T, Hj]]))
        out.update({f"{tag}_act_x": np.array(X), f"{tag}_act_r": np.array(R), f"{tag}_act_g24": np.array(Gs),
                    f"{tag}_act_H24": np.array(Hs)})
    np.savez(os.path.join(HERE, "fixed_pair.npz"), **out)


def cubic_kat():
    rng = np.random.default_rng(2024)
    coefs, refs, n_out, r_out = [], [], [], []
    for _ in range(64):
        abc = np.sort(rng.uniform(0, 5.0, 3))
        a0 = rng.uniform(0.1, 5.0)
        c = np.array([-abc[0] * abc[1] * abc[2], abc[0] * abc[1] + abc[0] * abc[2] + abc[1] * abc[2], -abc.sum(),
                      1.0]) * a0
        n, r = O.cubic_roots(c, -5.0, 5.0)
        coefs.append(c)
        refs.append(abc)
        n_out.append(n)
        r_out.append(r)
    np.savez(os.path.join(HERE, "cubic_kat.npz"), coef=np.array(coefs), known=np.array(refs), n=np.array(n_out),
             roots=np.array(r_out))


def small_frames():
    spec = scenes.config_small()
    cp = ContactParams(ground=spec.ground, solver_tol=1e-12)
    sim = O.OracleSim(spec, cp)
    A, p, its = [], [], []
    for _ in range(20):
        its.append(sim.step())
        A.append(sim.A.copy())
        p.append(sim.p.copy())
    np.savez(os.path.join(HERE, "small_frames.npz"), A=np.array(A), p=np.array(p), iters=np.array(its))


if __name__ == "__main__":
    fixed_pair()
    cubic_kat()
    small_frames()
    print("golden fixtures written")
