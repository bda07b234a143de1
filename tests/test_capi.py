"""The C-ABI library loads and exports every symbol include/*.h declares (no
compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(gpb_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_declared_symbols(gpb):
    lib = ctypes.CDLL(str(ROOT / "ucm-mpi-gpu_amd" / "libgpb.so"))
    names = _declared()
    assert len(names) >= 6
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_synth_deterministic(gpb):
    a = gpb.synth_voronoi(50, 40, 9, 6.0, 42)
    b = gpb.synth_voronoi(50, 40, 9, 6.0, 42)
    c = gpb.synth_voronoi(50, 40, 9, 6.0, 43)
    assert a.shape == (40, 50, 3) and a.dtype.name == "uint8"
    assert (a == b).all() and not (a == c).all()
    _, reg = gpb.synth_voronoi(80, 60, 12, 0.0, 1, with_regions=True)
    assert set(np.unique(reg)) <= set(range(12)) and len(np.unique(reg)) >= 8


import numpy as np  # noqa: E402
