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

The target itself (:578-642): the mesh is rendered over the environment into
the optimisation-resolution colour and encoded-depth textures (mesh.wgsl,
envmap.wgsl), the path tracer adds one sample per pixel to its accumulator
(reset whenever the camera changes, :237-242) and its resolved output is the
colour target; the depth target is the raster's.  ``attach_scene`` sets this
up on generated meshes (the reference's .glb/.hdr assets are not shipped) and
``scene_frame`` runs it; ``frame(target_rgba, target_depth_rgba)`` takes the
textures from the caller instead.  Everything runs through the C ABI (nps_*,
npf_*, npb_*, npp_*).
"""
from __future__ import annotations

import numpy as np

from .managers import (GpuBlurPipelineManager, GpuMeshRenderPipelineManager, GpuPathTracePipelineManager,
                       GpuSplatForwardPipelineManager, GpuSplatOptimizerManager)

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

        self.mesh = None
        self.pt = None
        self._view = None

    def attach_scene(self, meshes, env: np.ndarray, shading_mode: float = 1.0, hip_device: int = 0):
        """The target generators of the frame: meshes is a list of (vertices [n, 12],
        indices [m, 3]) (the model first, then the ground), env the HDR panorama used
        as environment and matcap (GpuRunner.svelte.ts:148-177)."""
        self.mesh = GpuMeshRenderPipelineManager(mesh=meshes[0], matcapTexture=env, hip_device=hip_device)
        if len(meshes) > 1:
            self.mesh.setGroundMesh(meshes[1])
        self.pt = GpuPathTracePipelineManager(envTexture=env, hip_device=hip_device)
        self.pt.setMeshes(meshes)
        self.pt.setOutputSize(self.optim_width, self.optim_height)
        self.shading_mode = float(shading_mode)

    def set_camera(self, vp: np.ndarray, vp_inv: np.ndarray, view: np.ndarray | None = None):
        """The camera effect (:237-249): the splat VP, the forward pass's VP, and
        with a scene attached the raster uniforms and the path tracer's inverse VP
        with its accumulator reset."""
        self.optimizer.writeSplatVPMatrix(vp, vp_inv, self.compare_blurred)
        self.forward.writeVPMatrix(vp)
        if self.pt is not None:
            if view is None:
                raise ValueError("a scene is attached: the view matrix is needed for the raster")
            self.mesh.writeUniforms(vp, view, vp_inv, self.shading_mode)
            self.pt.writeInvViewProjMat(vp_inv)
            self.pt.reset()

    def scene_textures(self):
        """The frame's target textures from the attached scene: the raster pass
        (colour, encoded depth), then one path-trace sample; returns (path-traced
        colour, raster colour, raster depth)."""
        W, H = self.optim_width, self.optim_height
        col, dep = self.mesh.addDraw(None, None, W, H)
        self.pt.dispatch()
        return self.pt.read_output(), col, dep

    def scene_frame(self):
        """One frame of the reference with the target rendered from the attached
        scene: raster, path-trace sample, then frame(path-traced colour, raster
        depth).  Returns (the forward render, the path-traced target)."""
        pt_col, _, dep = self.scene_textures()
        return self.frame(pt_col, dep), pt_col

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
        if self.pt is not None:
            self.pt.close()
            self.mesh.close()
        self.forward.close()
        self.blur.close()
        self.optimizer.close()
