"""The C-ABI library loads without a GPU and exports every entry point that
include/happyray.h declares; the Python binding covers each of them."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "happyray.h")


def declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:hr_status|const char\*)\s+(hr_\w+)\s*\(", src, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "hr_create" in names and "hr_render" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(hr):
    lib = ctypes.CDLL(hr.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from happyray_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared()


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs have the sizes and field offsets that a C
    compiler gives the header."""
    import shutil
    import subprocess
    from happyray_amd import _lib
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    structs = {"hr_options": _lib.Options, "hr_grid_info": _lib.GridInfo,
               "hr_render_params": _lib.RenderParams, "hr_timing": _lib.Timing,
               "hr_work_stats": _lib.WorkStats, "hr_tuning": _lib.Tuning,
               "hr_scene_info": _lib.SceneInfo, "hr_mesh_desc": _lib.MeshDesc,
               "hr_ray_batch": _lib.RayBatch, "hr_batch_tuning": _lib.BatchTuning,
               "hr_autotune_result": _lib.AutotuneResult, "hr_server_options": _lib.ServerOptions,
               "hr_request": _lib.Request, "hr_request_result": _lib.RequestResult,
               "hr_server_scene": _lib.ServerScene, "hr_server_stats": _lib.ServerStats}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "happyray.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    out = dict(l.rsplit(" ", 1) for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for cname, py in structs.items():
        assert int(out[cname]) == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(py, fname).offset, (cname, fname)


def test_no_device_is_an_error_not_a_crash(hr):
    """Without a GPU, hr_create reports HR_ERR_DEVICE (the reference exits)."""
    import subprocess
    import sys
    code = ("import sys; sys.path.insert(0, %r); import ctypes; from happyray_amd import _lib;"
            "L = _lib.load(); h = ctypes.c_void_p(); o = _lib.Options(0, 1, 2);"
            "s = L.hr_create(ctypes.byref(h), ctypes.byref(o)); print(s, L.hr_last_error().decode())"
            % ROOT)
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    status = int(r.stdout.split()[0])
    assert status == 2, r.stdout
