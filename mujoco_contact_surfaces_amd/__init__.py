"""MI355X-native hydroelastic contact-surface engine (drop-in for the
per-step contact-surface and wrench call of ubi-agni/mujoco_contact_surfaces).

The engine is libhcs.so (C-ABI in include/hcs.h, HIP kernels for gfx950);
this package is its host-side mirror: ctypes binding (_abi), the plugin-shaped
wrapper (engine) and seeded workload generators (scenes).
"""
from .engine import (KPOLYGON, KTRIANGLE, Batch, ContactSurfacesPlugin, Engine, HcsError,  # noqa: F401
                     props_from_cs)
