"""MI355X (gfx950) splat rasterizer and optimizer loop of nprush.

Hot path in hand-written HIP kernels (nprush_amd/csrc) behind a C ABI
(include/nprush_hip.h, libnprush_hip.so); Python classes mirroring the
reference's GpuSplatOptimizerManager / GpuSplatForwardPipelineManager.
"""
from . import camera, meshes, runner, scenes  # noqa: F401
from .managers import (  # noqa: F401
    DensifierMLP,
    GpuBezierForwardPipelineManager,
    GpuBezierOptimizerManager,
    GpuBlurPipelineManager,
    GpuMeshRenderPipelineManager,
    GpuMultiViewStep,
    GpuPathTracePipelineManager,
    GpuSplatForwardPipelineManager,
    GpuSplatOptimizerManager,
    compose_display,
    debug_math,
    reference_bezier_init,
    reference_init,
)
