"""ctypes binding of ``libpp_hip.so`` (the C ABI declared in ``include/pp.h``).

Loading fails loudly when the library has not been built: there is no CPU
fallback on this path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libpp_hip.so")

# every symbol include/pp.h declares
EXPORTS = (
    "pp_create", "pp_create_batch", "pp_sweeps", "pp_energy", "pp_check_energy", "pp_download",
    "pp_upload", "pp_get_counters", "pp_get_counters_all", "pp_reset_counters", "pp_num_replicas", "pp_num_particles",
    "pp_energy_snapshot", "pp_generate", "pp_set_trace", "pp_get_trace", "pp_set_timing", "pp_get_timing", "pp_get_grid", "pp_destroy",
    "pp_strerror", "pp_last_error", "pp_autotune",
    "pp_net_create", "pp_net_set", "pp_net_get", "pp_net_forward", "pp_net_loss_grad", "pp_net_destroy",
    "pp_net_attach", "pp_net_train", "pp_get_learn_counters", "pp_net_descriptor",
    "pp_checkpoint_save", "pp_checkpoint_load", "pp_bond_list", "pp_set_lx_move",
    "pp_energy_snapshot_noalloc", "pp_set_lx_acceptance", "pp_get_layout", "pp_upload_n",
    "pp_colour_conflicts", "pp_colour_check", "pp_snapshot_start", "pp_snapshot_finish",
)

# sizes of the proposal network (PP_NET_* in include/pp.h)
NET_INPUTS = 84
NET_OUTPUTS = 26
NET_WEIGHTS = 49152
NET_BIASES = 416

# PP_NOALLOC_MAX_N in include/pp.h
NOALLOC_MAX_N = 65536

PP_OK = 0
STATUS = {
    0: "PP_OK", 1: "PP_ERR_ARG", 2: "PP_ERR_HIP", 3: "PP_ERR_OVERLAP", 4: "PP_ERR_BOX",
    5: "PP_ERR_CELL_CAP", 6: "PP_ERR_ENERGY", 7: "PP_ERR_NOMEM", 8: "PP_ERR_UNSUPPORTED", 9: "PP_ERR_CELLS",
}


class PPError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("ensemble", C.c_int),
        ("temperature", C.c_double),
        ("disp_max", C.c_double),
        ("theta_max", C.c_double),
        ("kf_delta", C.c_double),
        ("kf_cosmax", C.c_double),
        ("pressure", C.c_double),
        ("rescale_factor_max", C.c_double),
        ("activity", C.c_double),
        ("gc_n_max", C.c_int),
    ]


class Options(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("threads_per_cell", C.c_int),
        ("block_size", C.c_int),
        ("cell_occupancy", C.c_double),
        ("device", C.c_int),
        ("launch", C.c_int),
        ("trials_per_cell", C.c_int),
    ]


class Counters(C.Structure):
    _fields_ = [
        ("tries", C.c_longlong * 2),
        ("accepted", C.c_longlong * 2),
        ("bonds", C.c_longlong),
        ("sweeps", C.c_longlong),
        ("box", C.c_double * 3),
        ("volume", C.c_double),
        ("gc_tries", C.c_longlong * 2),
        ("gc_accepted", C.c_longlong * 2),
        ("n", C.c_int),
        ("pad", C.c_int),
    ]


_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_vp = C.c_void_p

_lib = None


def lib():
    """Load libpp_hip.so (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                          "python -m patchyparticles_amd.build")
    L = C.CDLL(LIB_PATH)
    st = C.c_int
    L.pp_create.argtypes = [C.POINTER(Params), C.POINTER(Options), C.c_int, _dp, _dp, _dp, C.POINTER(_vp)]
    L.pp_create.restype = st
    L.pp_create_batch.argtypes = [C.c_int, C.POINTER(Params), C.POINTER(Options), _ip, _dp, _dp, _dp,
                                  C.POINTER(_vp)]
    L.pp_create_batch.restype = st
    L.pp_sweeps.argtypes = [_vp, C.c_longlong]
    L.pp_sweeps.restype = st
    L.pp_energy.argtypes = [_vp, C.c_int, _vp, _vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    L.pp_energy.restype = st
    L.pp_check_energy.argtypes = [_vp, _vp]
    L.pp_check_energy.restype = st
    L.pp_download.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.pp_download.restype = st
    L.pp_upload.argtypes = [_vp, C.c_int, _vp, _vp, _vp]
    L.pp_upload.restype = st
    L.pp_upload_n.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp, _vp]
    L.pp_upload_n.restype = st
    L.pp_get_counters.argtypes = [_vp, C.c_int, C.POINTER(Counters)]
    L.pp_get_counters.restype = st
    L.pp_get_counters_all.argtypes = [_vp, C.POINTER(Counters)]
    L.pp_get_counters_all.restype = st
    L.pp_reset_counters.argtypes = [_vp]
    L.pp_reset_counters.restype = st
    L.pp_num_replicas.argtypes = [_vp]
    L.pp_num_replicas.restype = C.c_int
    L.pp_num_particles.argtypes = [_vp, C.c_int]
    L.pp_num_particles.restype = C.c_int
    L.pp_energy_snapshot.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, C.c_double, _vp, _vp,
                                     C.POINTER(C.c_longlong), C.POINTER(C.c_int)]
    L.pp_energy_snapshot.restype = st
    L.pp_energy_snapshot_noalloc.argtypes = L.pp_energy_snapshot.argtypes
    L.pp_energy_snapshot_noalloc.restype = st
    L.pp_generate.argtypes = [C.c_int, C.c_double, C.c_long, _dp, _dp, _dp]
    L.pp_generate.restype = st
    L.pp_set_trace.argtypes = [_vp, C.c_int]
    L.pp_set_trace.restype = st
    L.pp_get_trace.argtypes = [_vp, C.c_int, _vp, C.c_int, C.POINTER(C.c_int)]
    L.pp_get_trace.restype = st
    L.pp_autotune.argtypes = [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.pp_autotune.restype = st
    L.pp_set_timing.argtypes = [_vp, C.c_int]
    L.pp_set_timing.restype = st
    L.pp_get_timing.argtypes = [_vp, C.c_char_p, C.POINTER(C.c_longlong), C.POINTER(C.c_double)]
    L.pp_get_timing.restype = st
    L.pp_get_grid.argtypes = [_vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_double)]
    L.pp_get_grid.restype = st
    L.pp_get_layout.argtypes = [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pp_get_layout.restype = st
    L.pp_colour_conflicts.argtypes = [_vp, C.c_int, _vp, C.POINTER(C.c_longlong)]
    L.pp_colour_conflicts.restype = st
    L.pp_colour_check.argtypes = [_vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.pp_colour_check.restype = st
    L.pp_snapshot_start.argtypes = [_vp, C.c_int]
    L.pp_snapshot_start.restype = st
    L.pp_snapshot_finish.argtypes = [_vp, _vp, _vp, _vp]
    L.pp_snapshot_finish.restype = st
    fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    L.pp_net_create.argtypes = [C.c_int, C.c_uint64, C.POINTER(_vp)]
    L.pp_net_create.restype = st
    L.pp_net_set.argtypes = [_vp, fp, fp]
    L.pp_net_set.restype = st
    L.pp_net_get.argtypes = [_vp, _vp, _vp]
    L.pp_net_get.restype = st
    L.pp_net_forward.argtypes = [_vp, C.c_int, _vp, _vp]
    L.pp_net_forward.restype = st
    L.pp_net_loss_grad.argtypes = [_vp, C.c_int, fp, fp, C.POINTER(C.c_double), _vp, _vp]
    L.pp_net_loss_grad.restype = st
    L.pp_net_destroy.argtypes = [_vp]
    L.pp_net_destroy.restype = None
    L.pp_net_attach.argtypes = [_vp, _vp, C.c_double, C.c_double]
    L.pp_net_attach.restype = st
    L.pp_net_train.argtypes = [_vp, _vp, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.pp_net_train.restype = st
    L.pp_get_learn_counters.argtypes = [_vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.pp_get_learn_counters.restype = st
    L.pp_net_descriptor.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, C.c_int, _ip, _vp]
    L.pp_net_descriptor.restype = st
    L.pp_checkpoint_save.argtypes = [_vp, C.c_char_p]
    L.pp_checkpoint_save.restype = st
    L.pp_checkpoint_load.argtypes = [_vp, C.c_char_p]
    L.pp_checkpoint_load.restype = st
    L.pp_bond_list.argtypes = [_vp, C.c_int, C.c_int, _vp, _vp]
    L.pp_bond_list.restype = st
    L.pp_set_lx_move.argtypes = [_vp, C.c_int, C.c_double, C.c_double, C.c_double]
    L.pp_set_lx_move.restype = st
    L.pp_set_lx_acceptance.argtypes = [_vp, C.c_int, C.c_int]
    L.pp_set_lx_acceptance.restype = st
    L.pp_destroy.argtypes = [_vp]
    L.pp_destroy.restype = None
    L.pp_strerror.argtypes = [st]
    L.pp_strerror.restype = C.c_char_p
    L.pp_last_error.argtypes = []
    L.pp_last_error.restype = C.c_char_p
    _lib = L
    return L


def check(status):
    if status != PP_OK:
        raise PPError(status, lib().pp_last_error().decode())


def _arr(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.reshape(-1)


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None else None


def energy_snapshot(r, O, box, kf_delta=0.12, kf_cosmax=0.92):
    """Per-particle (bonds, overlap), bonded pairs and any-overlap of a snapshot
    computed on the GPU (MC_energy of every particle, MC.c:272-309)."""
    r = _arr(r)
    O = _arr(O)
    n = r.size // 3
    bonds = np.zeros(max(n, 1), np.int32)
    ov = np.zeros(max(n, 1), np.uint8)
    tot = C.c_longlong(0)
    anyo = C.c_int(0)
    check(lib().pp_energy_snapshot(n, r, O, _arr(box), kf_delta, kf_cosmax, _ptr(bonds), _ptr(ov),
                                   C.byref(tot), C.byref(anyo)))
    return bonds[:n], ov[:n].astype(bool), tot.value, bool(anyo.value)


def energy_snapshot_noalloc(r, O, box, kf_delta=0.12, kf_cosmax=0.92):
    """energy_snapshot through the allocation-free entry point (at most
    NOALLOC_MAX_N particles; the library allocates nothing for the call)."""
    r = _arr(r)
    O = _arr(O)
    n = r.size // 3
    bonds = np.zeros(max(n, 1), np.int32)
    ov = np.zeros(max(n, 1), np.uint8)
    tot = C.c_longlong(0)
    anyo = C.c_int(0)
    check(lib().pp_energy_snapshot_noalloc(n, r, O, _arr(box), kf_delta, kf_cosmax, _ptr(bonds), _ptr(ov),
                                           C.byref(tot), C.byref(anyo)))
    return bonds[:n], ov[:n].astype(bool), tot.value, bool(anyo.value)


def generate(n, rho, seed=-1):
    """Random sequential insertion start (the reference generator's algorithm):
    returns r (n, 3), O (n, 3, 3), box (3,)."""
    r = np.zeros(3 * max(n, 1))
    O = np.zeros(9 * max(n, 1))
    box = np.zeros(3)
    check(lib().pp_generate(int(n), float(rho), int(seed), r, O, box))
    return r[:3 * n].reshape(n, 3), O[:9 * n].reshape(n, 3, 3), box


def net_descriptor(r, O, box, idx, device=0):
    """Descriptors of particles ``idx`` (every other particle a candidate), as
    the learned move computes them: (len(idx), 84) float32 holding BF16 values."""
    r = _arr(r)
    O = _arr(O)
    idx = np.ascontiguousarray(idx, np.int32)
    x = np.zeros((len(idx), NET_INPUTS), np.float32)
    check(lib().pp_net_descriptor(int(device), r.size // 3, r, O, _arr(box), len(idx), idx, _ptr(x)))
    return x
