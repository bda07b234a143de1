"""MI355X gPb hot path: HIP kernels for gfx950 behind a C ABI (libgpb.so).

The package directory holds the HIP/C++ sources (``csrc/``), the build recipe
(``build.py``) and a ctypes binding (``gpb.py``).  Load it with
``load()``; the binding raises ImportError when the library is not built.
"""
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent


def load():
    """Import and return the ctypes binding module (requires libgpb.so)."""
    import importlib.util
    import sys

    name = "gpb_amd_binding"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, PKG_DIR / "gpb.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod
