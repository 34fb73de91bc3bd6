"""MI355X (gfx950) implementation of the PatchyParticles rototranslation Monte
Carlo hot path: Kern-Frenkel pair test, per-particle energy with cell lists,
cell rebuild, as checkerboard sweeps in hand-written HIP kernels behind the C
ABI of include/pp.h (libpp_hip.so).
"""
from ._lib import EXPORTS, LIB_PATH, NOALLOC_MAX_N, PPError, energy_snapshot, energy_snapshot_noalloc, generate, lib, net_descriptor
from .rrr import read_rrr, write_rrr
from .system import GC, NPT, NVT, ProposalNet, System
from .batch import uneven_replicas

__all__ = ["System", "NVT", "NPT", "GC", "energy_snapshot", "generate", "read_rrr", "write_rrr", "PPError", "lib", "LIB_PATH",
           "EXPORTS", "uneven_replicas", "ProposalNet", "net_descriptor", "energy_snapshot_noalloc", "NOALLOC_MAX_N"]
