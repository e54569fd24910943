"""The splat part of the reference's per-frame schedule.

GpuRunner (src/lib/viewer/GpuRunner.svelte.ts) renders the model into
optimisation-resolution textures, blurs them, trains the splats against them
and renders the splats back at that resolution, once per animation frame:

* optimisation resolution: the short side is OPTIM_SHORT = 256 pixels, the
  other matches the panel's aspect ratio, rounded (GpuRunner.svelte.ts:15,
  rebuildOptimTextures :284-296);
* target preparation: with compareBlurred (default true,
  viewer/ViewerState.svelte.ts:26) the colour target is blurred in sRGB and the
  encoded-depth target linearly, both with radius blurRadius (16, :28) and
  sigma radius / 2 (GpuRunner.svelte.ts:644-667);
* the camera's VP and inverse are written with blurEnabled = compareBlurred
  whenever the camera changes (:249);
* setBackwardTarget(blurred colour, blurred depth) -> dispatch (backward, Adam
  step, densification every 25th) -> forward render of the current splats at
  the optimisation resolution with clear = true (:689-709).

The mesh and path-trace passes that produce the target (:578-642) need mesh
assets the reference does not ship; a caller passes the rendered target
texture instead (``frame(target_rgba, target_depth_rgba)``).  Everything runs
through the C ABI (nps_*, npf_*, npb_*).
"""
from __future__ import annotations

import numpy as np

from .managers import GpuBlurPipelineManager, GpuSplatForwardPipelineManager, GpuSplatOptimizerManager

OPTIM_SHORT = 256


def optim_size(panel_aspect: float) -> tuple[int, int]:
    """(width, height) of the optimisation textures for a panel aspect ratio
    (GpuRunner.svelte.ts:284-296).  JavaScript's Math.round rounds halves up."""
    if panel_aspect >= 1.0:
        return int(np.floor(OPTIM_SHORT * panel_aspect + 0.5)), OPTIM_SHORT
    return OPTIM_SHORT, int(np.floor(OPTIM_SHORT / panel_aspect + 0.5))


class GpuRunner:
    """Splat optimisation frames of the reference's GpuRunner, REF semantics."""

    def __init__(self, num_splats: int, panel_aspect: float, *, splats: np.ndarray | None = None,
                 blur_radius: int = 16, compare_blurred: bool = True, adc_period: int = 25, hip_device: int = 0):
        self.optim_width, self.optim_height = optim_size(panel_aspect)
        self.blur_radius = int(blur_radius)
        self.compare_blurred = bool(compare_blurred)
        self.optimizer = GpuSplatOptimizerManager(numSplats=num_splats, splats=splats, adc_period=adc_period,
                                                  hip_device=hip_device)
        self.forward = GpuSplatForwardPipelineManager(splatBuffer=self.optimizer)
        self.blur = GpuBlurPipelineManager(hip_device=hip_device)
        self.forward.setTarget(None, None, self.optim_width, self.optim_height)

    def set_camera(self, vp: np.ndarray, vp_inv: np.ndarray):
        """The camera effect (:249) and the forward pass's VP."""
        self.optimizer.writeSplatVPMatrix(vp, vp_inv, self.compare_blurred)
        self.forward.writeVPMatrix(vp)

    def prepare_target(self, target_rgba: np.ndarray, target_depth_rgba: np.ndarray):
        """The textures setBackwardTarget receives: blurred (colour in sRGB, depth
        linear) when compare_blurred, else the inputs."""
        W, H = self.optim_width, self.optim_height
        if not self.compare_blurred:
            return target_rgba, target_depth_rgba
        r = self.blur_radius
        c = self.blur.blur(None, target_rgba, None, None, W, H, r, r / 2, True)
        d = self.blur.blur(None, target_depth_rgba, None, None, W, H, r, r / 2, False)
        return c, d

    def frame(self, target_rgba: np.ndarray, target_depth_rgba: np.ndarray):
        """One frame: target preparation, optimiser dispatch, forward render at the
        optimisation resolution.  Returns the rendered (colour, depth) textures."""
        W, H = self.optim_width, self.optim_height
        if target_rgba.shape[:2] != (H, W) or target_depth_rgba.shape[:2] != (H, W):
            raise ValueError(f"targets must be {H} x {W} (the optimisation resolution)")
        c, d = self.prepare_target(target_rgba, target_depth_rgba)
        self.optimizer.setBackwardTarget(c, d[..., 0], W, H)
        self.optimizer.dispatch()
        self.forward.setTarget(None, None, W, H)
        self.forward.dispatch(None, clear=True)
        return self.forward.read_target()

    def close(self):
        self.forward.close()
        self.blur.close()
        self.optimizer.close()
