"""The C-ABI library loads and exports every function include/verion_hip.h declares (no GPU calls)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "verion_hip.h")
LIB = os.path.join(ROOT, "verion_amd", "libverion_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vh_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_contract():
    names = declared()
    for must in ["vh_context_create", "vh_renderer_create", "vh_accumulate", "vh_acc_read_film", "vh_acc_mean",
                 "vh_set_adaptive", "vh_set_restir", "vh_set_integrator", "vh_render_normals", "vh_render_aovs",
                 "vh_primary_hits", "vh_last_error"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libverion_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = C.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_backend_refuses_to_run_without_the_library(tmp_path, monkeypatch):
    import verion_amd.backend as b

    monkeypatch.setattr(b, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(b, "_lib", None)
    with pytest.raises(ImportError):
        b.Context()
