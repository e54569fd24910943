"""The C-ABI library loads and exports every entry point the header declares
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "nprush_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:nps|npf|npb|npt|nbz|nbf|nml|npd|npp)_[a-z0-9_]+)\s*\(", src)))


def test_header_symbols_exported():
    from nprush_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from nprush_amd import _lib

    assert set(_declared()) == set(_lib.SIGNATURES)


def test_manager_method_names_match_reference():
    """Method names of the reference managers (GpuSplatOptimizerManager.ts:281,303,365;
    GpuSplatForwardPipelineManager.ts:94,104,114; GpuBezierForwardPipelineManager.ts:100,110,131)."""
    from nprush_amd import GpuBezierForwardPipelineManager, GpuSplatForwardPipelineManager, GpuSplatOptimizerManager

    for m in ("writeSplatVPMatrix", "setBackwardTarget", "dispatch"):
        assert callable(getattr(GpuSplatOptimizerManager, m))
    for m in ("writeVPMatrix", "setTarget", "dispatch"):
        assert callable(getattr(GpuSplatForwardPipelineManager, m))
        assert callable(getattr(GpuBezierForwardPipelineManager, m))
    from nprush_amd import GpuBlurPipelineManager   # GpuBlurPipelineManager.ts:51

    assert callable(GpuBlurPipelineManager.blur)


def test_null_handles_are_refused_without_a_gpu():
    """Every family of entry points refuses a null handle with NPS_ERR_ARG (-1)
    before touching the GPU, and the error-text calls answer for a null handle."""
    from nprush_amd import _lib

    lib = _lib.load()
    null = None
    buf = (ctypes.c_float * 64)()
    ibuf = (ctypes.c_int32 * 64)()
    d2 = (ctypes.c_double * 2)()
    calls = [
        ("nps_write_splats", (null, buf, 1)), ("nps_read_splats", (null, buf, 1)),
        ("nps_set_camera", (null, buf, buf, 0)), ("nps_set_target", (null, buf, buf, 1, 1)),
        ("nps_backward", (null,)), ("nps_forward", (null,)), ("nps_step", (null,)), ("nps_adc", (null,)),
        ("nps_dispatch", (null,)), ("nps_read_loss", (null, d2)), ("nps_read_grads", (null, ibuf, 1)),
        ("npf_write_vp", (null, buf)), ("npf_set_target", (null, 4, 4)), ("npf_dispatch", (null, 1, 1)),
        ("npt_step", (null, 1)), ("npt_autotune", (null,)), ("npt_read_losses", (null, d2)),
        ("npb_blur", (null, buf, buf, 4, 4, 1, 1.0, 1)),
        ("nbz_backward", (null,)), ("nbz_dispatch", (null,)), ("nbz_read_loss", (null, d2)),
        ("nbf_write_vp", (null, buf)), ("nbf_dispatch", (null, 1, 1)),
        ("nml_forward", (null, buf, 1, buf)), ("nml_read_params", (null, buf)),
    ]
    for name, args in calls:
        assert getattr(lib, name)(*args) == -1, name
    for name in ("nps_last_error", "npt_last_error", "npb_last_error", "nbz_last_error", "nml_last_error"):
        assert getattr(lib, name)(null), name
    # creation with bad arguments: refused before any device call
    out = ctypes.c_void_p()
    assert lib.nps_create(None, -1, ctypes.byref(out)) == -1
    assert lib.npt_create(None, 0, 4, ctypes.byref(out)) == -1
    assert lib.nbz_create(0, -1, ctypes.byref(out)) == -1
    assert lib.nml_create(0, 0, ctypes.byref(out)) == -1
    assert lib.npb_create(0, None) == -1
