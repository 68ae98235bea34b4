"""MI355X-native (gfx950) particle-DEM step with FullC0ntact's ParticleSystem API.

The compute path is libfc_dem.so (hand-written HIP kernels behind the C ABI in
include/fc_dem.h).  Submodules load it on first use; `python -m
fullc0ntact_amd.build` compiles it.
"""
from __future__ import annotations

import importlib

__all__ = ["ParticleSystem", "SimParams", "Context", "DeviceArray", "FcError", "make_params"]

_LAZY = {
    "ParticleSystem": "particle_system",
    "make_params": "particle_system",
    "SimParams": "_lib",
    "Context": "_lib",
    "DeviceArray": "_lib",
    "FcError": "_lib",
}


def __getattr__(name: str):
    mod = _LAZY.get(name)
    if mod is None:
        raise AttributeError(name)
    return getattr(importlib.import_module(f"{__name__}.{mod}"), name)
