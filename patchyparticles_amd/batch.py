"""Batched workloads: many independent replicas in one context.

``uneven_replicas`` draws the batched-replica workload of many independent
NVT systems with uneven size, temperature and density: N log-uniform in
[n_min, n_max], T uniform in [t_min, t_max], rho uniform in [rho_min,
rho_max], each started by random sequential insertion with uniform random
orientations (its own seed), with the reference's default move sizes and
patch geometry.
"""
from __future__ import annotations

import numpy as np

from ._lib import generate


def uneven_replicas(count=256, seed=0, n_range=(500, 8000), t_range=(0.08, 0.20), rho_range=(0.05, 0.60)):
    """Returns (rs, Os, boxes, params) for System.batch."""
    rng = np.random.default_rng(seed)
    n = np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), count)).astype(int)
    T = rng.uniform(*t_range, count)
    rho = rng.uniform(*rho_range, count)
    rs, Os, boxes, params = [], [], [], []
    for k in range(count):
        r, O, box = generate(int(n[k]), float(rho[k]), int(seed * 100003 + k))
        rs.append(r)
        Os.append(O)
        boxes.append(box)
        params.append(dict(T=float(T[k])))
    return rs, Os, boxes, params
