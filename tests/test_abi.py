"""The C-ABI library loads without a GPU and exports every entry point the header declares."""

import os
import re

import pytest

from jqmc_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jqmc_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(jqmc_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "jqmc_vmc_sweep" in names and "jqmc_local_energy" in names
    assert len(names) >= 12


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
        assert name in _lib.EXPORTED_SYMBOLS, f"{name} has no ctypes signature"
    assert lib.jqmc_abi_version() == _lib.ABI_VERSION == 6


@pytest.mark.skipif(not os.path.exists(_lib.LIB_PATH), reason="library not built")
def test_error_path_reports_message():
    lib = _lib.load()
    import ctypes

    rc = lib.jqmc_system_create(None, ctypes.byref(ctypes.c_void_p()))
    assert rc == 1
    assert b"NULL" in lib.jqmc_last_error()


def test_struct_layout_matches_header():
    """Field order of the ctypes mirrors equals the order in the C structs."""
    text = open(HEADER).read()
    desc = text[text.index("typedef struct {", text.index("jqmc_system_desc") - 4000):text.index("} jqmc_system_desc;")]
    ordered = []
    for line in desc.splitlines():
        line = line.split("/*")[0]
        for m in re.finditer(r"(?:int32_t|double|const \w+\*)\s+([\w\s,*]+);", line):
            for name in m.group(1).split(","):
                ordered.append(name.strip().lstrip("*"))
    assert [f[0] for f in _lib.SystemDesc._fields_] == ordered
    wk = text[text.index("typedef struct {", text.index("jqmc_system;")):text.index("} jqmc_walkers;")]
    wnames = re.findall(r"\*\s*(\w+)\s*;", wk)
    assert [f[0] for f in _lib.Walkers._fields_] == wnames


def test_lrdmc_params_layout_matches_header():
    text = open(HEADER).read()
    body = text[text.index("typedef struct {", text.index("JQMC_LRDMC_FIXED_TAU")):text.index("} jqmc_lrdmc_params;")]
    names = []
    for line in body.splitlines():
        m = re.match(r"\s*(?:int32_t|double)\s+(\w+);", line.split("/*")[0])
        if m:
            names.append(m.group(1))
    assert [f[0] for f in _lib.LrdmcParams._fields_] == names


def test_packed_fp32_flag_is_honoured_by_the_device_pass(tmp_path):
    """The target feature build() passes removes the packed FP32 FMA that hipcc -O3 otherwise makes
    from two independent float FMAs (the check build() runs on the library's own assembly)."""
    import shutil
    import subprocess

    import __graft_entry__ as ge

    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "k.hip"
    src.write_text("#include <hip/hip_runtime.h>\n__global__ void k(float2* a, float2 b) {\n"
                   "  float2 x = a[threadIdx.x]; x.x = x.x * b.x + b.y; x.y = x.y * b.x + b.y; a[threadIdx.x] = x;\n}\n")
    out = {}
    for name, flags in (("plain", []), ("guarded", ge.NO_PACKED_FP32)):
        asm = tmp_path / f"{name}.s"
        subprocess.run([hipcc, "--offload-arch=gfx950", "-O3", *flags, "--cuda-device-only", "-S", "-o", str(asm),
                        str(src)], check=True, capture_output=True)
        out[name] = asm.read_text()
    assert "v_pk_fma_f32" in out["plain"]
    assert not any(op in out["guarded"] for op in ge.PACKED_FP32)
