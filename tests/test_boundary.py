"""The C-ABI boundary on the CPU: the library loads, exports every function
the headers declare, the Python binding covers them, and the host-side error
paths report errors instead of crashing (no kernel is launched here)."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADERS = sorted((ROOT / "include").glob("*.h"))


def declared_functions():
    names = set()
    for h in HEADERS:
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:wg_status|const char\s*\*)\s+(wg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_functions():
    names = declared_functions()
    assert "wg_agg_consume" in names and "wg_join_probe" in names and "wg_gen" in names
    assert len(names) > 30


def test_library_exports_every_declared_symbol():
    from wadjet_amd._lib import LIB_PATH
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB_PATH)], capture_output=True, text=True, check=True)
    exported = set(re.findall(r"\sT\s+(wg_\w+)$", out.stdout, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"


def test_binding_covers_every_declared_symbol():
    import wadjet_amd._lib as L
    L.lib()
    missing = [n for n in declared_functions() if n not in L.SIGNATURES and n not in L.STRING_FNS]
    assert not missing, f"no ctypes signature for: {missing}"
    for n in L.SIGNATURES:
        assert hasattr(L.lib(), n), n


def test_version_and_error_string_without_gpu():
    import wadjet_amd._lib as L
    assert b"gfx950" in L.lib().wg_version()
    assert isinstance(L.lib().wg_last_error(), bytes)


def test_null_arguments_are_rejected_on_the_host():
    import ctypes as C

    import wadjet_amd._lib as L
    with pytest.raises(L.WgError) as e:
        L.call("wg_ctx_create", 0, None)
    assert e.value.status == L.WG_ERR_INVALID
    with pytest.raises(L.WgError) as e:
        L.call("wg_agg_consume", None, None, 0, None, 0, None, 0, 1)
    assert e.value.status == L.WG_ERR_INVALID
    with pytest.raises(L.WgError):
        L.call("wg_agg_num_groups", None, C.byref(C.c_uint64()))


def test_product_fails_loudly_without_the_library(tmp_path, monkeypatch):
    import wadjet_amd._lib as L
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", tmp_path / "missing.so")
    with pytest.raises(ImportError):
        L.lib()
