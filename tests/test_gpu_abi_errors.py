"""Error paths of the C ABI on the GPU: calls in the wrong state return
NPS_ERR_STATE (-3) with a message, counts that do not match the handle
NPS_ERR_ARG (-1); handles that borrow another handle's buffers (npf, nbf,
npt) refuse every call once the owner is destroyed and can still be freed;
a display-raster target of a new size starts cleared."""
import ctypes as C

import numpy as np
import pytest

from nprush_amd import (GpuBezierForwardPipelineManager, GpuBezierOptimizerManager, GpuMultiViewStep,
                        GpuSplatForwardPipelineManager, GpuSplatOptimizerManager, _lib, camera, scenes)

pytestmark = pytest.mark.gpu

ERR_ARG, ERR_STATE = -1, -3


def test_state_and_size_errors():
    lib = _lib.load()
    m = GpuSplatOptimizerManager(numSplats=64, splats=scenes.make_splats(64, 0))
    h = m.handle
    # no target, no camera
    assert lib.nps_dispatch(h) == ERR_STATE
    assert b"target" in lib.nps_last_error(h)
    assert lib.nps_backward(h) == ERR_STATE
    # tiles before any backward
    nt = C.c_int32()
    nid = C.c_int64()
    assert lib.nps_read_tiles(h, None, None, None, C.byref(nt), C.byref(nid)) == ERR_STATE
    # a count that is not the handle's
    buf = np.zeros((65, 16), np.float32)
    assert lib.nps_write_splats(h, buf.ctypes.data_as(C.POINTER(C.c_float)), 65) == ERR_ARG
    assert b"num_splats" in lib.nps_last_error(h)
    # EXT-only calls on a REF handle
    sh = np.zeros((64, 48), np.float32)
    assert lib.nps_write_sh(h, sh.ctypes.data_as(C.POINTER(C.c_float)), 64) == ERR_STATE
    assert lib.nps_forward(h) == ERR_STATE
    # a bad target size
    assert lib.nps_set_target(h, buf.ctypes.data_as(C.POINTER(C.c_float)),
                              buf.ctypes.data_as(C.POINTER(C.c_float)), 0, 4) == ERR_ARG
    m.close()
    # blur: bad size
    b = C.c_void_p()
    assert lib.npb_create(0, C.byref(b)) == 0
    img = np.zeros(64, np.float32)
    p = img.ctypes.data_as(C.POINTER(C.c_float))
    assert lib.npb_blur(b, p, p, 0, 4, 1, 1.0, 1) == ERR_ARG
    lib.npb_destroy(b)


def test_borrowing_handles_after_owner_destroyed():
    lib = _lib.load()
    # splat render handle
    m = GpuSplatOptimizerManager(numSplats=32, splats=scenes.make_splats(32, 1))
    f = GpuSplatForwardPipelineManager(splatBuffer=m)
    fh = f._f
    m.close()
    vp = np.eye(4, dtype=np.float32).reshape(16)
    assert lib.npf_write_vp(fh, vp.ctypes.data_as(C.POINTER(C.c_float))) == ERR_STATE
    assert lib.npf_dispatch(fh, 1, 1) == ERR_STATE
    f.close()   # still frees it
    # Bezier display raster
    z = GpuBezierOptimizerManager(numBeziers=16)
    d = GpuBezierForwardPipelineManager(bezierBuffer=z, brushTexture=scenes.make_brush(16))
    dh = d._f
    assert lib.nbf_set_target(dh, 8, 6) == 0
    out = np.full((6, 8, 4), 7.0, np.float32)
    assert lib.nbf_read_target(dh, out.ctypes.data_as(C.POINTER(C.c_float))) == 0
    assert not out.any(), "a new target starts cleared"
    z.close()
    assert lib.nbf_dispatch(dh, 1, 1) == ERR_STATE
    assert lib.nbf_read_target(dh, out.ctypes.data_as(C.POINTER(C.c_float))) == ERR_STATE
    d.close()
    # multi-view step over two scenes, one destroyed
    s0, h0 = scenes.make_ext_splats(500, 0, sh_degree=1)
    s1, h1 = scenes.make_ext_splats(500, 1, sh_degree=1)
    a = GpuSplatOptimizerManager(numSplats=500, mode="ext", sh_degree=1, splats=s0, sh=h0, adc_enabled=False)
    b = GpuSplatOptimizerManager(numSplats=500, mode="ext", sh_degree=1, splats=s1, sh=h1, adc_enabled=False)
    st = GpuMultiViewStep([a, b], lanes=2)
    vp, vi = camera.view_proj((0.0, 0.0, 3.0), aspect=1.0)
    st.add_view(0, vp, vi, np.ones((32, 32, 4), np.float32))
    st.add_view(1, vp, vi, np.ones((32, 32, 4), np.float32))
    st.step(graph=True)
    th = st._t
    b.close()
    assert lib.npt_step(th, 1) == ERR_STATE
    assert b"destroyed" in lib.npt_last_error(th)
    st.close()
    a.step()   # the surviving scene is untouched by the detached step
    a.close()
