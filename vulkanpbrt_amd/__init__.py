"""MI355X (gfx950) frame path of VulkanPBRT: BVH build, wavefront path tracing,
temporal accumulation, BMFR/BFR denoising and TAA behind a C ABI.

The compute path is ``libpbrt.so`` (hand-written HIP kernels for gfx950);
:mod:`vulkanpbrt_amd.capi` is its ctypes view.  Importing the package does not
touch the GPU.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
