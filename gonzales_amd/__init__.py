"""gonzales_amd — the MI355X (gfx950) hot path of the Gonzales path tracer.

BVH2 closest-hit / any-hit traversal, the Z-Sobol sampler and the wavefront diffuse shade as
hand-written HIP kernels behind a C-ABI (include/gonzales_amd.h) that mirrors the reference's
mojo_kernel.h; this package holds the ctypes binding (``abi``), the host-side mirror of the
reference's Accelerator API (``accelerator``) and the seeded synthetic workloads (``scenes``).
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
