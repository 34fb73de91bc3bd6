"""Python face of the GPU hot path: a ``System`` holds one or more replicas in a
device context and advances them with ``do_ensemble`` (the reference's
``syst.do_ensemble`` seam, main.c:119).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import NET_BIASES, NET_INPUTS, NET_OUTPUTS, NET_WEIGHTS, Counters, Options, Params, PPError, _arr, _ptr, check, lib

NVT, GC, NPT = 0, 1, 5  # defs.h:21-25


class System:
    """Replicas of hard spheres with 4 Kern-Frenkel patches, on one GPU.

    Positions ``r`` (N, 3) and orientation matrices ``O`` (N, 3, 3) by rows, as
    the reference's PatchyParticle (defs.h:47-56).  For several replicas pass
    lists of per-replica arrays and parameters (see :meth:`batch`).
    """

    def __init__(self, r, O, box, *, T, disp_max=0.05, theta_max=0.1, kf_delta=0.12, kf_cosmax=0.92,
                 ensemble=NVT, P=0.0, rescale_factor_max=0.0, activity=0.0, gc_n_max=0, seed=1, threads_per_cell=0,
                 block_size=0, cell_occupancy=0.0, device=0, launch=0, trials_per_cell=0):
        self._h = None
        self._make([r], [O], [box], [dict(T=T, disp_max=disp_max, theta_max=theta_max, kf_delta=kf_delta,
                                          kf_cosmax=kf_cosmax, ensemble=ensemble, P=P,
                                          rescale_factor_max=rescale_factor_max, activity=activity,
                                          gc_n_max=gc_n_max)],
                   seed, threads_per_cell, block_size, cell_occupancy, device, launch, trials_per_cell)

    @classmethod
    def batch(cls, rs, Os, boxes, params, *, seed=1, threads_per_cell=0, block_size=0, cell_occupancy=0.0,
              device=0, launch=0, trials_per_cell=0):
        """Many independent replicas in one context; ``params`` is a list of
        dicts with the keyword arguments of the constructor (T, P, ...)."""
        self = cls.__new__(cls)
        self._h = None
        self._make(rs, Os, boxes, params, seed, threads_per_cell, block_size, cell_occupancy, device, launch,
                   trials_per_cell)
        return self

    def _make(self, rs, Os, boxes, params, seed, tpc, block, occ, device, launch, tpc_trials):
        R = len(rs)
        self.n = [int(np.asarray(r).size // 3) for r in rs]
        P = (Params * R)()
        for k, p in enumerate(params):
            P[k].ensemble = int(p.get("ensemble", NVT))
            P[k].temperature = float(p["T"])
            P[k].disp_max = float(p.get("disp_max", 0.05))
            P[k].theta_max = float(p.get("theta_max", 0.1))
            P[k].kf_delta = float(p.get("kf_delta", 0.12))
            P[k].kf_cosmax = float(p.get("kf_cosmax", 0.92))
            P[k].pressure = float(p.get("P", 0.0))
            P[k].rescale_factor_max = float(p.get("rescale_factor_max", 0.0))
            P[k].activity = float(p.get("activity", 0.0))
            P[k].gc_n_max = int(p.get("gc_n_max", 0))
        o = Options(int(seed) & 0xFFFFFFFFFFFFFFFF, int(tpc), int(block), float(occ), int(device), int(launch),
                    int(tpc_trials))
        r_all = np.concatenate([_arr(r) for r in rs]) if R else np.zeros(0)
        O_all = np.concatenate([_arr(x) for x in Os]) if R else np.zeros(0)
        b_all = np.concatenate([_arr(b) for b in boxes])
        h = C.c_void_p()
        check(lib().pp_create_batch(R, P, C.byref(o), np.asarray(self.n, np.int32), np.ascontiguousarray(r_all),
                                    np.ascontiguousarray(O_all), np.ascontiguousarray(b_all), C.byref(h)))
        self._h = h
        self.R = R
        self._gc = [P[k].ensemble == GC for k in range(R)]
        self._theta_max0 = float(P[0].theta_max) if R else 0.1
        self._net = None

    def _n(self, rep):
        """Particles of replica rep now (a grand-canonical replica's count
        changes); an index out of range is refused as the C ABI refuses it
        (PP_ERR_ARG), not wrapped around."""
        if not 0 <= int(rep) < self.R:
            raise PPError(1, f"replica {rep} out of range")
        if self._gc[rep]:
            return self.counters(rep)["n"]
        return self.n[rep]

    def close(self):
        if self._h is not None:
            lib().pp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- the hot path -------------------------------------------------------
    def do_ensemble(self, sweeps=1):
        """Advance every replica by ``sweeps`` sweeps (N attempts each)."""
        check(lib().pp_sweeps(self._h, int(sweeps)))

    def energy(self, rep=0):
        """(bonds[N], overlap[N], bonded pairs, any overlap) recomputed from
        scratch, as MC_energy over every particle (MC.c:272-309)."""
        n = self._n(rep)
        bonds = np.zeros(max(n, 1), np.int32)
        ov = np.zeros(max(n, 1), np.uint8)
        tot = C.c_longlong(0)
        anyo = C.c_int(0)
        check(lib().pp_energy(self._h, rep, _ptr(bonds), _ptr(ov), C.byref(tot), C.byref(anyo)))
        return bonds[:n], ov[:n].astype(bool), tot.value, bool(anyo.value)

    def check_energy(self):
        """MC_check_energy (MC.c:766-784) on every replica; returns energies."""
        e = np.zeros(self.R)
        check(lib().pp_check_energy(self._h, _ptr(e)))
        return e

    def state(self, rep=0):
        n = self._n(rep)
        r = np.zeros(3 * max(n, 1))
        O = np.zeros(9 * max(n, 1))
        box = np.zeros(3)
        check(lib().pp_download(self._h, rep, _ptr(r), _ptr(O), _ptr(box)))
        return r[:3 * n].reshape(n, 3), O[:9 * n].reshape(n, 3, 3), box

    def snapshot_start(self, rep=0):
        """Start an asynchronous copy of replica rep's state; later sweeps run
        while it proceeds.  snapshot_finish() returns that state."""
        self._snap_n = self._n(rep)
        check(lib().pp_snapshot_start(self._h, int(rep)))

    def snapshot_finish(self):
        """(r, O, box) of the replica as it was at snapshot_start."""
        n = getattr(self, "_snap_n", 0)
        r = np.zeros(3 * max(n, 1))
        O = np.zeros(9 * max(n, 1))
        box = np.zeros(3)
        check(lib().pp_snapshot_finish(self._h, _ptr(r), _ptr(O), _ptr(box)))
        return r[:3 * n].reshape(n, 3), O[:9 * n].reshape(n, 3, 3), box

    def upload(self, rep, r, O, box=None):
        """Replace replica rep's particles (a grand-canonical replica takes
        any number up to its gc_n_max)."""
        r = _arr(r)
        O = _arr(O)
        b = None if box is None else _arr(box)
        check(lib().pp_upload_n(self._h, rep, r.size // 3, _ptr(r), _ptr(O), _ptr(b)))

    def counters(self, rep=0):
        c = Counters()
        check(lib().pp_get_counters(self._h, rep, C.byref(c)))
        return dict(tries=list(c.tries), accepted=list(c.accepted), bonds=c.bonds, sweeps=c.sweeps,
                    box=list(c.box), volume=c.volume, n=c.n, gc_tries=list(c.gc_tries),
                    gc_accepted=list(c.gc_accepted))

    def counters_all(self):
        """Counters of every replica: dict of numpy arrays (one entry per replica)."""
        cs = (Counters * self.R)()
        check(lib().pp_get_counters_all(self._h, cs))
        return dict(tries=np.array([list(c.tries) for c in cs]), accepted=np.array([list(c.accepted) for c in cs]),
                    bonds=np.array([c.bonds for c in cs]), box=np.array([list(c.box) for c in cs]),
                    volume=np.array([c.volume for c in cs]), n=np.array([c.n for c in cs]),
                    gc_tries=np.array([list(c.gc_tries) for c in cs]),
                    gc_accepted=np.array([list(c.gc_accepted) for c in cs]))

    def reset_counters(self):
        check(lib().pp_reset_counters(self._h))

    def grid(self, rep=0):
        n = (C.c_int * 3)()
        s = (C.c_double * 3)()
        check(lib().pp_get_grid(self._h, rep, n, s))
        return list(n), list(s)

    def layout(self):
        """(cells reserved over all replicas, cells per scan tile, scan tiles)."""
        cap, tile, tiles = C.c_longlong(0), C.c_int(0), C.c_int(0)
        check(lib().pp_get_layout(self._h, C.byref(cap), C.byref(tile), C.byref(tiles)))
        return cap.value, tile.value, tiles.value

    def colour_conflicts(self, rep=0, nside=None):
        """Particles of replica rep within the bond range of a particle in
        another cell of the same colour: on the grid of the current cell lists
        (nside None) or on nside = (nx, ny, nz) cells over the box (a debug
        check of the checkerboard invariant; 0 on every grid the sweeps use)."""
        n = C.c_longlong(0)
        arg = None if nside is None else (C.c_int * 3)(*[int(v) for v in nside])
        check(lib().pp_colour_conflicts(self._h, int(rep), arg, C.byref(n)))
        return n.value

    def colour_check(self, on=None):
        """Live checkerboard checks after the rebuild and every colour pass of
        each later sweep: on=True starts them (and resets the totals),
        on=False stops them, None only reads.  Returns (conflicts, replica
        checks) since the last reset."""
        t, k = C.c_longlong(0), C.c_longlong(0)
        check(lib().pp_colour_check(self._h, -1 if on is None else (1 if on else 0), C.byref(t), C.byref(k)))
        return t.value, k.value

    def set_trace(self, capacity):
        """Record the bonded-pair count after each of the next `capacity` sweeps."""
        self._trace_cap = int(capacity)
        check(lib().pp_set_trace(self._h, int(capacity)))

    def trace(self, rep=0):
        """Energy E = -bonds after each traced sweep (numpy int64, oldest first)."""
        cap = getattr(self, "_trace_cap", 0)
        out = np.zeros(max(cap, 1), np.int64)
        cnt = C.c_int(0)
        check(lib().pp_get_trace(self._h, rep, _ptr(out), cap, C.byref(cnt)))
        return -out[:cnt.value]

    def autotune(self, sweeps=20):
        """Pick the fastest launch shape (threads per cell, workgroup size) by
        timing `sweeps` real sweeps with each; returns (threads_per_cell,
        block_size, microseconds per sweep)."""
        t, b, us = C.c_int(0), C.c_int(0), C.c_double(0)
        check(lib().pp_autotune(self._h, int(sweeps), C.byref(t), C.byref(b), C.byref(us)))
        return t.value, b.value, us.value

    def set_timing(self, every=1):
        """Bracket the launches of every ``every``-th sweep with HIP events
        (0 or False: off).  Resets the accumulated times."""
        check(lib().pp_set_timing(self._h, int(every)))

    def timing(self, name=None):
        nl = C.c_longlong(0)
        ms = C.c_double(0)
        check(lib().pp_get_timing(self._h, None if name is None else name.encode(), C.byref(nl), C.byref(ms)))
        return nl.value, ms.value

    def set_lx_move(self, lx_change_max, lyz_min, lyz_max, rep=-1, reference_rule=False):
        """Box-shape moves at constant volume (NVT): Lx += U(-1/2, 1/2)
        lx_change_max, Ly = Lz = sqrt(V / Lx) within [lyz_min, lyz_max],
        accepted with min(1, exp(-dE/T)), or with the reference's exp(+dE/T)
        (MC.c:562) when ``reference_rule``."""
        check(lib().pp_set_lx_move(self._h, int(rep), float(lx_change_max), float(lyz_min), float(lyz_max)))
        check(lib().pp_set_lx_acceptance(self._h, int(rep), 1 if reference_rule else 0))

    def bond_list(self, rep=0, max_per=16):
        """Bonded partners of every particle: a list of sorted index arrays."""
        n = self._n(rep)
        cnt = np.zeros(max(n, 1), np.int32)
        part = np.zeros(max(n, 1) * max_per, np.int32)
        check(lib().pp_bond_list(self._h, rep, max_per, _ptr(cnt), _ptr(part)))
        if (cnt[:n] > max_per).any():
            raise RuntimeError("more bonded partners than max_per")
        return [np.sort(part[i * max_per:i * max_per + cnt[i]]) for i in range(n)]

    def checkpoint_save(self, path):
        """Exact state, counters and sweep index of every replica (bit-exact resume)."""
        check(lib().pp_checkpoint_save(self._h, str(path).encode()))

    def checkpoint_load(self, path):
        """Resume from pp_checkpoint_save's file (same parameters, sizes, seed)."""
        check(lib().pp_checkpoint_load(self._h, str(path).encode()))

    # ---- learned proposal ------------------------------------------------

    def attach(self, net, learned_fraction=0.8, theta_max=None):
        """Use ``net`` (a :class:`ProposalNet`) for a share ``learned_fraction``
        of the colour passes of later sweeps; ``net=None`` detaches.  The
        rotation angle of a learned move is uniform in [0, theta_max)
        (default: the replicas' own theta_max, of replica 0)."""
        if theta_max is None:
            theta_max = self._theta_max0
        check(lib().pp_net_attach(self._h, None if net is None else net._h, float(learned_fraction),
                                  float(theta_max)))
        self._net = net  # keep it alive while attached

    def train(self, net, steps, batch=4096, lr=1e-3, pos_noise=0.05, theta_max=None, sweeps_between=1):
        """Train ``net`` by denoising moves drawn from this context's replicas
        (one Adam step per batch, ``sweeps_between`` sweeps between steps);
        returns the mean loss of the first and last batch."""
        if theta_max is None:
            theta_max = self._theta_max0
        a, b = C.c_double(0), C.c_double(0)
        check(lib().pp_net_train(self._h, net._h, int(steps), int(batch), float(lr), float(pos_noise),
                                 float(theta_max), int(sweeps_between), C.byref(a), C.byref(b)))
        return a.value, b.value

    def learn_counters(self, rep=0):
        """(tried, accepted) learned moves of replica ``rep`` since the last reset."""
        t, a = C.c_longlong(0), C.c_longlong(0)
        check(lib().pp_get_learn_counters(self._h, rep, C.byref(t), C.byref(a)))
        return t.value, a.value


class ProposalNet:
    """The proposal network of the learned move (device weights; FP32 master
    copy, BF16 operands).  Weights are laid out as include/pp.h describes."""

    def __init__(self, seed=1, device=0):
        self._h = C.c_void_p()
        check(lib().pp_net_create(int(device), int(seed), C.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None):
            lib().pp_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def get(self):
        w = np.zeros(NET_WEIGHTS, np.float32)
        b = np.zeros(NET_BIASES, np.float32)
        check(lib().pp_net_get(self._h, _ptr(w), _ptr(b)))
        return w, b

    def set(self, w, b):
        check(lib().pp_net_set(self._h, np.ascontiguousarray(w, np.float32), np.ascontiguousarray(b, np.float32)))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, NET_INPUTS)
        y = np.zeros((x.shape[0], NET_OUTPUTS), np.float32)
        check(lib().pp_net_forward(self._h, x.shape[0], _ptr(x), _ptr(y)))
        return y

    def loss_grad(self, x, target):
        x = np.ascontiguousarray(x, np.float32).reshape(-1, NET_INPUTS)
        t = np.ascontiguousarray(target, np.float32).reshape(-1, 6)
        gw = np.zeros(NET_WEIGHTS, np.float32)
        gb = np.zeros(NET_BIASES, np.float32)
        loss = C.c_double(0)
        check(lib().pp_net_loss_grad(self._h, x.shape[0], x, t, C.byref(loss), _ptr(gw), _ptr(gb)))
        return loss.value, gw, gb
