"""The C-ABI library loads and exports every entry point the header declares
(no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "nprush_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b((?:nps|npf|npb|npt|nbz|nbf|nml)_[a-z0-9_]+)\s*\(", src)))


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
