"""pfc_amd -- MI355X-native engine for pfc-relax's semi-implicit spectral
relaxation loop (phase field crystal models of graphene, hBN and their
heterostructures).

The compute path is libpfcgpu.so (hand-written HIP for gfx950, C ABI in
include/pfc_relax.h); the pfc-relax binary in pfc_amd/bin is the drop-in
command-line front end.  This package holds the ctypes binding used by
tests and benchmarks, parameter helpers and synthetic input generators.
"""
from .params import (FP64, MIXED, ENERGY_MODEL, ENERGY_NU2, PfcParams, make_params,
                     graphene_params, hetero_params, LATTICE_A)

__all__ = ["FP64", "MIXED", "ENERGY_MODEL", "ENERGY_NU2", "PfcParams", "make_params",
           "graphene_params", "hetero_params", "LATTICE_A"]
