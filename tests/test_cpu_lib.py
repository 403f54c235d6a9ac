"""The C-ABI library builds, loads and exports every entry point declared in include/*.h (CPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    names = []
    for h in (ROOT / "include").glob("*.h"):
        names += re.findall(r"^\s*(?:int|const char \*)\s*(phast_\w+)\s*\(", h.read_text(), flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared()
    assert "phast_q4_explicit_step" in names and "phast_hex8_mech_action" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from phast_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run python -m phast_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert _lib.lib().phast_version().startswith(b"phast-mi355x")


def test_product_has_no_cpu_fallback():
    """Compute entry points refuse CPU tensors instead of computing on the host."""
    import torch
    from phast_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    with pytest.raises(ValueError):
        _lib.ptr(torch.zeros(4, dtype=torch.float64))
    if not torch.cuda.is_available():
        with pytest.raises(_lib.PhastLibraryError):
            _lib.require_device()


def test_product_does_not_import_oracle():
    for p in (ROOT / "phast_amd").rglob("*.py"):
        text = p.read_text()
        assert "import oracle" not in text and "from oracle" not in text, p


def test_t3_plan_validates_connectivity_before_any_device_call():
    """Host-side checks of phast_t3_plan_create reject bad meshes with PHAST_ERR_ARG."""
    import numpy as np
    from phast_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built")
    L = _lib.lib()
    nodes = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    h = ctypes.c_void_p()
    for els, split in ((np.array([[0, 1, 3]], dtype=np.int32), 0), (np.array([[0, 1, -1]], dtype=np.int32), 0),
                       (np.array([[0, 1, 2]], dtype=np.int32), 7)):
        rc = L.phast_t3_plan_create(3, 1, nodes.ctypes.data, els.ctypes.data, 1.0, 0.2, 1.0, 1.0, 1e-7, 0, split,
                                    split, 0, 0.0, ctypes.byref(h))
        assert rc == _lib.PHAST_ERR_ARG and not h.value


def _declared_arity():
    """Parameter count of every ``int phast_*(...)`` declaration in include/*.h."""
    out = {}
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        for m in re.finditer(r"^\s*(?:int|const char \*)\s*(phast_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.M | re.S):
            params = m.group(2).strip()
            out[m.group(1)] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_bound_argtypes_match_the_header():
    """Every ctypes binding the host modules set up passes as many arguments as the C
    declaration takes (a wrong count only shows at the first call, on the GPU)."""
    from phast_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("library not built (run python -m phast_amd.build)")
    import phast_amd.solid3d, phast_amd.q4, phast_amd.t3, phast_amd.batch, phast_amd.adjoint3d  # noqa: F401,E401
    arity = _declared_arity()
    lib = _lib.lib()
    checked, wrong = 0, []
    for name, n in arity.items():
        fn = getattr(lib, name, None)
        if fn is None or fn.argtypes is None:
            continue
        checked += 1
        if len(fn.argtypes) != n:
            wrong.append((name, len(fn.argtypes), n))
    assert not wrong, wrong
    assert checked >= 60
