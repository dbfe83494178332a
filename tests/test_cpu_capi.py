"""The C-ABI library: loads, exports every declared symbol, fails cleanly."""
import ctypes
import os
import re
import subprocess

import pytest

import maboss.gpu_amd as mg
from maboss.gpu_amd import _capi
from conftest import LIB, ROOT, model_paths

HEADER = os.path.join(ROOT, "include", "maboss_gpu.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mbg_\w+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = ctypes.CDLL(LIB)
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_capi.EXPORTED) == syms


def test_nm_exports_match():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (mbg_\w+)", out))
    assert set(declared_symbols()) <= exported


def test_null_arguments():
    lib = _capi.lib()
    assert lib.mbg_model_load(None, None, None) == 1
    assert lib.mbg_last_error().decode()
    assert lib.mbg_model_get_info(None, None) == 1
    assert lib.mbg_simulate(None, None, None, None) == 1
    assert lib.mbg_result_window_size(None, 0) == -1
    lib.mbg_model_free(None)
    lib.mbg_program_free(None)
    lib.mbg_result_free(None)
    assert lib.mbg_version().decode().startswith("maboss-gpu-amd")


def test_compile_without_gpu_fails_cleanly(gpu_available):
    if gpu_available:
        pytest.skip("a GPU is present")
    m = mg.Model.load(*model_paths("cellcycle"))
    with pytest.raises(mg.MbgError) as e:
        m.compile()
    assert e.value.code == 7  # MBG_ERR_HIP


def test_cli_usage():
    p = mg.run_cli()
    assert p.returncode == 1 and "Usage: MaBoSS.GPU [-o prefix] bnd_file cfg_file" in p.stdout
    p = mg.run_cli("/nonexistent.bnd", "/nonexistent.cfg")
    assert p.returncode == 1 and "cannot open" in p.stderr


STRUCTS = {
    "mbg_model_info": _capi.ModelInfo, "mbg_compile_opts": _capi.CompileOpts,
    "mbg_sim_params": _capi.SimParams, "mbg_result_info": _capi.ResultInfo,
    "mbg_variant": _capi.Variant, "mbg_batch_params": _capi.BatchParams,
    "mbg_batch_info": _capi.BatchInfo, "mbg_grad_params": _capi.GradParams,
    "mbg_grad_info": _capi.GradInfo,
}


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the C structs have the header's size and field offsets."""
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "maboss_gpu.h"', "int main(void){"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    for line in out.splitlines():
        cname, field, val = line.split()
        py = STRUCTS[cname]
        want = ctypes.sizeof(py) if field == "sizeof" else getattr(py, field).offset
        assert int(val) == want, (cname, field)
