"""CPU-side checks of the C ABI: the built library loads and exports every
function declared in include/exahype_amd.h (no GPU calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "exahype_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(exa_[a-z0-9_]+)\s*\(", txt)))


def test_header_and_python_binding_agree():
    import exahype_amd
    assert sorted(exahype_amd.EXPORTS) == declared_symbols()


def test_library_exports_every_declared_symbol():
    import exahype_amd
    if not os.path.exists(exahype_amd.LIB_PATH):
        pytest.skip("library not built (python -m exahype_amd.build)")
    lib = ctypes.CDLL(exahype_amd.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_no_gpu_fails_loudly():
    """Without a GPU the product refuses to run instead of falling back."""
    import exahype_amd
    if os.path.exists("/dev/kfd"):
        pytest.skip("GPU driver present")
    if not os.path.exists(exahype_amd.LIB_PATH):
        pytest.skip("library not built")
    with pytest.raises(RuntimeError):
        exahype_amd.ADERDGSolver(3)
