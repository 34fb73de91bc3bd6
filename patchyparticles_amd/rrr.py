"""The reference's ``.rrr`` configuration format in numpy.

Header ``step N Lx Ly Lz``; then three lines per particle: orientation row 0,
orientation row 1, position (output.c:302-317).  Loading normalises rows 0
and 1, sets row 2 = row0 x row1 and Gram-Schmidts the three rows
(system.c:213-245); every operation below is elementwise in the reference's
order, so the doubles match the reference loader bit for bit.
"""
from __future__ import annotations

import numpy as np


def _dot(a, b):
    return a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1] + a[..., 2] * b[..., 2]


def _normalize(v):
    n = np.sqrt(_dot(v, v))
    return np.stack([v[..., 0] / n, v[..., 1] / n, v[..., 2] / n], axis=-1)


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def gram_schmidt(O):
    """Rows of O (..., 3, 3): utils.c:91-113."""
    v1, v2, v3 = O[..., 0, :].copy(), O[..., 1, :].copy(), O[..., 2, :].copy()
    n1 = _dot(v1, v1)
    c21 = _dot(v2, v1) / n1
    v2 = np.stack([v2[..., k] - c21 * v1[..., k] for k in range(3)], axis=-1)
    s31 = _dot(v3, v1)
    s32 = _dot(v3, v2)
    n2 = _dot(v2, v2)
    c31 = s31 / n1
    c32 = s32 / n2
    v3 = np.stack([v3[..., k] - (c31 * v1[..., k] + c32 * v2[..., k]) for k in range(3)], axis=-1)
    return np.stack([_normalize(v1), _normalize(v2), _normalize(v3)], axis=-2)


def read_rrr(path):
    """Returns (step, box[3], r[N,3], O[N,3,3])."""
    with open(path) as f:
        head = f.readline().split()
        if len(head) < 5:
            raise ValueError(f"{path}: malformed header")
        step, n = int(head[0]), int(head[1])
        box = np.array([float(x) for x in head[2:5]])
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != 9 * n:
        raise ValueError(f"{path}: found {vals.size // 9} particles, header says {n}")
    v = vals.reshape(n, 3, 3)
    a = _normalize(v[:, 0])
    b = _normalize(v[:, 1])
    O = gram_schmidt(np.stack([a, b, _cross(a, b)], axis=1))
    return step, box, np.ascontiguousarray(v[:, 2]), O


def write_rrr(path, step, box, r, O):
    """output_save (output.c:302-317): %lld %d %.12lf x3, rows %lf, positions %.12lf."""
    r = np.asarray(r).reshape(-1, 3)
    O = np.asarray(O).reshape(-1, 3, 3)
    lines = ["%d %d %.12f %.12f %.12f" % (step, len(r), box[0], box[1], box[2])]
    for i in range(len(r)):
        lines.append("%f %f %f" % tuple(O[i, 0]))
        lines.append("%f %f %f" % tuple(O[i, 1]))
        lines.append("%.12f %.12f %.12f" % tuple(r[i]))
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
