"""ctypes bindings for the CPU oracle (``oracle/liboracle.so``) and, where it
has been built, the reference itself (``oracle/_ref/libppref.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, the smoke check and the
cpu_baseline leg of ``bench.py`` as the checker.  The product package never
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_LIB_PATH = os.path.join(REF_DIR, "libppref.so")

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")

OVERLAP, NO_BOND, BOND = -1, 0, 1


class OrcParams(C.Structure):
    _fields_ = [
        ("ensemble", C.c_int),
        ("temperature", C.c_double),
        ("disp_max", C.c_double),
        ("theta_max", C.c_double),
        ("kf_delta", C.c_double),
        ("kf_cosmax", C.c_double),
        ("pressure", C.c_double),
        ("rescale_factor_max", C.c_double),
        ("seed", C.c_long),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.orc_base_patches.argtypes = [_dp]
        L.orc_patches.argtypes = [_dp, _dp, _dp]
        L.orc_pair.argtypes = [_dp, C.c_double, C.c_double, _dp, _dp, _dp, _dp, C.c_void_p, C.c_void_p]
        L.orc_pair.restype = C.c_int
        L.orc_gram_schmidt.argtypes = [_dp]
        L.orc_cell_index.argtypes = [_dp, _dp, _ip, _ip]
        L.orc_cell_index.restype = C.c_int
        L.orc_grid.argtypes = [_dp, C.c_double, _ip]
        for f in (L.orc_energy_cells, L.orc_energy_brute):
            f.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _ip, _u8p]
            f.restype = C.c_longlong
        L.orc_sim_create.argtypes = [C.POINTER(OrcParams), C.c_int, _dp, _dp, _dp]
        L.orc_sim_create.restype = C.c_void_p
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_init_energy.argtypes = [C.c_void_p]
        L.orc_sim_init_energy.restype = C.c_int
        L.orc_sim_sweeps.argtypes = [C.c_void_p, C.c_longlong]
        L.orc_sim_check_energy.argtypes = [C.c_void_p]
        L.orc_sim_check_energy.restype = C.c_int
        L.orc_sim_cells_count.argtypes = [C.c_void_p]
        L.orc_sim_cells_count.restype = C.c_int
        L.orc_sim_energy.argtypes = [C.c_void_p]
        L.orc_sim_energy.restype = C.c_double
        L.orc_sim_get.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_sim_counters.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.orc_sim_reset_counters.argtypes = [C.c_void_p]
        L.orc_generate.argtypes = [C.c_int, C.c_double, C.c_long, _dp, _dp, _dp]
        L.orc_generate.restype = C.c_int
        L.orc_rrr_load.argtypes = [C.c_char_p, C.POINTER(C.c_longlong), _dp, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_rrr_load.restype = C.c_int
        L.orc_rrr_save.argtypes = [C.c_char_p, C.c_longlong, C.c_int, _dp, _dp, _dp]
        L.orc_rrr_save.restype = C.c_int
        L.orc_rrr_save_generated.argtypes = [C.c_char_p, C.c_int, _dp, _dp, _dp]
        L.orc_rrr_save_generated.restype = C.c_int
        L.orc_philox4x32_10.argtypes = [_u32p, _u32p, _u32p]
        L.orc_philox4x32_10_n.argtypes = [C.c_int, _u32p, _u32p, _u32p]
        _lib = L
    return _lib


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1) if n is None else a.reshape(n)


def base_patches():
    out = np.zeros(12)
    lib().orc_base_patches(out)
    return out.reshape(4, 3)


def patches(O):
    out = np.zeros(12)
    lib().orc_patches(_f64(O, 9), _f64(base_patches()), out)
    return out.reshape(4, 3)


def pair(rp, Op, rq, Oq, box, delta=0.12, cosmax=0.92):
    rc = 1.0 + delta
    Pp, Pq = patches(Op), patches(Oq)
    return lib().orc_pair(_f64(box, 3), rc * rc, cosmax, _f64(rp, 3), _f64(Pp), _f64(rq, 3), _f64(Pq), None, None)


def gram_schmidt(O):
    O = _f64(O, 9).copy()
    lib().orc_gram_schmidt(O)
    return O.reshape(3, 3)


def energy(r, O, box, delta=0.12, cosmax=0.92, brute=False):
    """Per-particle (bonds, overlap) and total bonded pairs (-1 on overlap)."""
    r = _f64(r)
    O = _f64(O)
    n = r.size // 3
    bonds = np.zeros(n, np.int32)
    ov = np.zeros(n, np.uint8)
    f = lib().orc_energy_brute if brute else lib().orc_energy_cells
    tot = f(n, r, O, _f64(box, 3), delta, cosmax, bonds, ov)
    return bonds, ov, tot


def generate(n, rho, seed=-1):
    r = np.zeros(3 * n)
    O = np.zeros(9 * n)
    box = np.zeros(3)
    if lib().orc_generate(n, rho, seed, r, O, box) != 0:
        raise ValueError("density above 0.7")
    return r.reshape(n, 3), O.reshape(n, 3, 3), box


def rrr_load(path):
    step = C.c_longlong(0)
    box = np.zeros(3)
    n = lib().orc_rrr_load(path.encode(), C.byref(step), box, None, None, 0)
    if n < 0:
        raise ValueError(f"bad rrr {path}")
    r = np.zeros(3 * n)
    O = np.zeros(9 * n)
    m = lib().orc_rrr_load(path.encode(), C.byref(step), box, r.ctypes.data, O.ctypes.data, n)
    if m != n:
        raise ValueError(f"bad rrr body {path}")
    return step.value, box, r.reshape(n, 3), O.reshape(n, 3, 3)


def rrr_save(path, step, box, r, O, generator_header=False):
    r = _f64(r)
    O = _f64(O)
    n = r.size // 3
    if generator_header:
        return lib().orc_rrr_save_generated(path.encode(), n, _f64(box, 3), r, O)
    return lib().orc_rrr_save(path.encode(), step, n, _f64(box, 3), r, O)


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def philox_n(ctr, key):
    """Philox4x32-10 of n blocks: ctr (n, 4), key (n, 2) uint32 -> (n, 4)."""
    ctr = np.ascontiguousarray(ctr, np.uint32).reshape(-1, 4)
    key = np.ascontiguousarray(key, np.uint32).reshape(-1, 2)
    out = np.zeros_like(ctr)
    lib().orc_philox4x32_10_n(len(ctr), ctr, key, out)
    return out


class Sim:
    """The serial reference-algorithm sampler (NVT Ensemble=0 / NPT Ensemble=5)."""

    def __init__(self, r, O, box, *, T, disp_max=0.05, theta_max=0.1, delta=0.12, cosmax=0.92,
                 ensemble=0, pressure=0.0, rescale_factor_max=0.0, seed=1):
        self.n = np.asarray(r).size // 3
        p = OrcParams(ensemble, T, disp_max, theta_max, delta, cosmax, pressure, rescale_factor_max, seed)
        self._h = lib().orc_sim_create(C.byref(p), self.n, _f64(r), _f64(O), _f64(box, 3))
        if lib().orc_sim_init_energy(self._h):
            raise ValueError("initial configuration contains an overlap")

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_sim_destroy(self._h)
            self._h = None

    def sweeps(self, k):
        lib().orc_sim_sweeps(self._h, k)

    def check_energy(self):
        return lib().orc_sim_check_energy(self._h)

    def run(self, steps, print_every):
        """main.c:78-120 cadence without output: check energy at print steps."""
        for st in range(steps):
            if print_every > 0 and st % print_every == 0:
                lib().orc_sim_reset_counters(self._h)
                if st != 0:
                    rc = self.check_energy()
                    if rc:
                        raise RuntimeError(f"energy check failed ({rc}) at step {st}")
            self.sweeps(1)

    @property
    def energy(self):
        return lib().orc_sim_energy(self._h)

    def state(self):
        r = np.zeros(3 * self.n)
        O = np.zeros(9 * self.n)
        box = np.zeros(3)
        lib().orc_sim_get(self._h, r, O, box)
        return r.reshape(self.n, 3), O.reshape(self.n, 3, 3), box

    def counters(self):
        t = (C.c_longlong * 2)()
        a = (C.c_longlong * 2)()
        lib().orc_sim_counters(self._h, t, a)
        return list(t), list(a)


# ---------------------------------------------------------------------------
# the reference itself (only where oracle/_ref has been built)

def ref_available():
    return os.path.exists(REF_LIB_PATH)


_ref = None


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB_PATH)
        L.ref_load.argtypes = [C.c_char_p]
        L.ref_load.restype = C.c_int
        L.ref_set_state.argtypes = [_dp, _dp, C.c_void_p]
        L.ref_get_state.argtypes = [_dp, _dp, _dp]
        L.ref_particle_energies.argtypes = [_ip, _u8p]
        L.ref_run.argtypes = [C.c_longlong, C.c_longlong]
        L.ref_energy.restype = C.c_double
        L.ref_volume.restype = C.c_double
        L.ref_counters.argtypes = [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.ref_pair.argtypes = [_dp, _dp, _dp, _dp, _dp]
        L.ref_pair.restype = C.c_int
        L.ref_base_patches.argtypes = [_dp]
        _ref = L
    return _ref


def write_input(path, conf, *, T, seed, ensemble=0, disp_max=0.05, theta_max=0.1, delta=0.12,
                cosmax=0.92, pressure=None, rescale_factor_max=None, steps=0, print_every=1000,
                save_every=1000000, folder=".", extra=None):
    lines = [
        "Dynamics = 0",
        f"Ensemble = {ensemble}",
        f"Seed = {seed}",
        f"Temperature = {T!r}",
        f"Disp_max = {disp_max!r}",
        f"Theta_max = {theta_max!r}",
        f"KF_delta = {delta!r}",
        f"KF_cosmax = {cosmax!r}",
        f"Initial_conditions_file = {conf}",
        f"Steps = {steps}",
        f"Print_every = {print_every}",
        f"Save_every = {save_every}",
        f"Configuration_folder = {folder}",
    ]
    if pressure is not None:
        lines.append(f"P = {pressure!r}")
    if rescale_factor_max is not None:
        lines.append(f"rescale_factor_max = {rescale_factor_max!r}")
    for k, v in (extra or {}).items():
        lines.append(f"{k} = {v}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


class Ref:
    """The reference's own code (system_init / MC_energy / do_ensemble) via
    oracle/_ref/libppref.so.  One live instance per process (the reference
    keeps global state)."""

    def __init__(self, conf_path, **kw):
        self.tmp = tempfile.mkdtemp(prefix="ppref_")
        inp = os.path.join(self.tmp, "input")
        write_input(inp, conf_path, folder=self.tmp, **kw)
        self.n = ref_lib().ref_load(inp.encode())
        if self.n < 0:
            raise ValueError(f"reference load failed ({self.n})")

    def set_state(self, r, O, box=None):
        b = None if box is None else _f64(box, 3)
        ref_lib().ref_set_state(_f64(r), _f64(O), None if b is None else b.ctypes.data)

    def state(self):
        r = np.zeros(3 * self.n)
        O = np.zeros(9 * self.n)
        box = np.zeros(3)
        ref_lib().ref_get_state(r, O, box)
        return r.reshape(self.n, 3), O.reshape(self.n, 3, 3), box

    def energies(self):
        b = np.zeros(self.n, np.int32)
        ov = np.zeros(self.n, np.uint8)
        ref_lib().ref_particle_energies(b, ov)
        return b, ov

    def run(self, steps, print_every):
        ref_lib().ref_run(steps, print_every)

    @property
    def energy(self):
        return ref_lib().ref_energy()

    def counters(self):
        t = (C.c_longlong * 2)()
        a = (C.c_longlong * 2)()
        ref_lib().ref_counters(t, a)
        return list(t), list(a)

    def pair(self, rp, Op, rq, Oq, box):
        return ref_lib().ref_pair(_f64(rp, 3), _f64(Op, 9), _f64(rq, 3), _f64(Oq, 9), _f64(box, 3))
