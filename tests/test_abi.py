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


def test_struct_layouts_match_the_header(tmp_path):
    # sizes and field offsets of every struct of the header, from gcc, against the ctypes mirror
    import subprocess

    from verion_amd import abi

    pairs = {"vh_scene_info": abi.SceneInfo, "vh_view": abi.ViewS, "vh_scene_tables": abi.SceneTables,
             "vh_adaptive": abi.Adaptive, "vh_restir": abi.Restir, "vh_memory_usage": abi.MemoryUsage,
             "vh_update_report": abi.UpdateReport, "vh_nrc_config": abi.NrcConfig, "vh_batch_job": abi.BatchJob,
             "vh_batch_tuning": abi.BatchTuning, "vh_update_plan": abi.UpdatePlanS}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'printf("{cname}.{field} %zu\\n", offsetof({cname}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines())
    for cname, py in pairs.items():
        assert int(got[cname]) == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert int(got[f"{cname}.{field}"]) == getattr(py, field).offset, (cname, field)
