"""Writes tests/golden/kat.json: known-answer vectors computed here in pure
Python/numpy from the mathematics each reference function implements, with
no use of the oracle or the product library.

  * Morton keys by explicit bit interleaving (morton.cpp:12-25 is a magic-mask
    form of the same map: x in even bits, y in odd bits).
  * scale_to_grid (morton.cpp:3-10) evaluated with numpy float32 arithmetic.
  * bounding boxes (globalLogic.cpp:1083-1099) in float32.
  * SPH kernel constants (SPH.h:38-56): closed forms.
  * a hand-checked 2D tree over 8 particles (quadtree.cpp:5-107).

Run: python tests/golden/make_golden.py
"""
import json
import math
import os

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kat.json")


def interleave(ix, iy):
    k = 0
    for b in range(18):
        k |= ((ix >> b) & 1) << (2 * b)
        k |= ((iy >> b) & 1) << (2 * b + 1)
    return k


def scale_to_grid(p, lo, hi):
    p, lo, hi = np.float32(p), np.float32(lo), np.float32(hi)
    if hi <= lo:
        return 0
    c = min(max(p, lo), hi)
    norm = np.float32(np.float32(c - lo) / np.float32(hi - lo))
    s = int(np.float32(norm * np.float32(262144.0)))
    return min(s, 262143)


def bbox(x, y):
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    mnx, mxx, mny, mxy = x.min(), x.max(), y.min(), y.max()
    size = max(np.float32(mxx - mnx), np.float32(mxy - mny))
    cx = np.float32(np.float32(mnx + mxx) * np.float32(0.5))
    cy = np.float32(np.float32(mny + mxy) * np.float32(0.5))
    half = np.float32(size * np.float32(0.5))
    return [float(np.float32(cx - half)), float(np.float32(cy - half)), float(size)]


def main():
    rng = np.random.default_rng(12345)
    morton = [[0, 0, 0], [1, 0, 1], [0, 1, 2], [1, 1, 3], [2, 0, 4], [3, 5, interleave(3, 5)],
              [262143, 262143, (1 << 36) - 1], [262143, 0, interleave(262143, 0)],
              [0, 262143, interleave(0, 262143)]]
    for _ in range(64):
        ix, iy = int(rng.integers(0, 262144)), int(rng.integers(0, 262144))
        morton.append([ix, iy, interleave(ix, iy)])
    grid = [[0.0, 0.0, 10.0, 0], [10.0, 0.0, 10.0, 262143], [-5.0, 0.0, 10.0, 0],
            [15.0, 0.0, 10.0, 262143], [5.0, 0.0, 10.0, 131072], [1.0, 1.0, 1.0, 0]]
    for _ in range(64):
        lo = float(np.float32(rng.uniform(-1000, 1000)))
        hi = float(np.float32(lo + rng.uniform(1e-3, 3000)))
        p = float(np.float32(rng.uniform(lo - 10, hi + 10)))
        grid.append([p, lo, hi, scale_to_grid(p, lo, hi)])
    boxes = []
    for n in (1, 2, 7, 100):
        x = rng.uniform(-500, 500, n).astype(np.float32)
        y = rng.uniform(-50, 300, n).astype(np.float32)
        boxes.append({"x": x.tolist(), "y": y.tolist(), "bb": bbox(x, y)})
    # SPH constants (SPH.h:38-56): W volume pi h^4 / 6, spiky -45 / (pi h^6),
    # laplacian 45 / (pi h^6), each operation in float with the float pi (the
    # reference's pow(float, float) is the float overload; h^4 and h^6 are
    # exact floats for these h).  At h = 2.5, 3 and 3.5 the float form differs
    # in the last bit from the same expression evaluated in double and rounded.
    F = np.float32
    pi_f = F(math.pi)
    sph = []
    for h in (1.0, 3.0, 4.0, 2.5, 3.5):
        h4, h6 = F(h ** 4), F(h ** 6)
        sph.append({"h": h, "volume": float((pi_f * h4) / F(6.0)),
                    "spiky": float(F(-45.0) / (pi_f * h6)),
                    "lap": float(F(45.0) / (pi_f * h6))})
    # Hand-checked tree: bbox (0,0,size 8); 8 particles placed so that the
    # root splits into four quadrants, the lower-left quadrant (size 4, 5
    # particles) splits again and the others are leaves by the count <= 4 &&
    # size <= 2 rule or the single-particle rule.  Particles are listed in
    # Morton order.  Expected pre-order nodes: (start, end, next, size).
    tree_pts = [(0.5, 0.5), (1.5, 0.5), (0.5, 1.5), (1.5, 1.5), (3.0, 3.0),
                (6.0, 1.0), (1.0, 6.0), (6.0, 6.0)]
    tree_nodes = [
        # root [0,8): children q0 [0,5), q1 [5,6), q2 [6,7), q3 [7,8)
        (0, 8, 8, 8.0),
        # q0 size 4: mid (2,2); [0,4) lower-left (size 2, 4 particles -> leaf), [4,5) upper-right
        (0, 5, 2, 4.0),
        (0, 4, 0, 2.0),
        (4, 5, 0, 2.0),
        (5, 6, 0, 4.0),
        (6, 7, 0, 4.0),
        (7, 8, 0, 4.0),
    ]
    # total nodes: root + q0 + its 2 children + 3 leaves = 7; root.next = 6
    tree_nodes[0] = (0, 8, 6, 8.0)
    data = {"morton": morton, "scale_to_grid": grid, "bbox": boxes, "sph_constants": sph,
            "tree": {"points": tree_pts, "bb": [0.0, 0.0, 8.0], "nodes": tree_nodes}}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
