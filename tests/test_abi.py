"""CPU tests of the C-ABI boundary: the library loads, exports every declared symbol, links
under the reference's own test program, and rejects bad arguments the reference's way."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

import knnhip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "knncuda.h")
REF_TEST = os.path.join(ROOT, "oracle", "_ref", "test_ref")


def declared_c_functions():
    text = open(HEADER).read()
    body = text.split('extern "C" {', 1)[1]
    return sorted(set(re.findall(r"\b(knnhip_\w+)\s*\(", body)))


def test_library_builds_and_loads():
    lib = knnhip.load()
    assert knnhip.load() is lib
    assert lib.knnhip_version().decode().startswith("knnhip")


def test_exports_every_declared_symbol():
    declared = declared_c_functions()
    assert set(declared) == set(knnhip.C_SYMBOLS)
    lib = knnhip.load()
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", knnhip.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    for sym in knnhip.CXX_SYMBOLS + tuple(declared):
        assert re.search(rf"\bT {re.escape(sym)}\b", out), sym


def test_cxx_linkage_matches_reference_header():
    # The reference header declares the three functions with C++ linkage (no extern "C",
    # code/knncuda.h:14-21,37-44,61-68); g++ mangles them to these names.
    src = "bool knn_cuda_global(const float*,int,const float*,int,int,int,float*,int*);\n" \
          "bool knn_cuda_texture(const float*,int,const float*,int,int,int,float*,int*);\n" \
          "bool knn_cublas(const float*,int,const float*,int,int,int,float*,int*);\n" \
          "void* p[] = {(void*)knn_cuda_global, (void*)knn_cuda_texture, (void*)knn_cublas};\n"
    out = subprocess.run(["g++", "-x", "c++", "-S", "-o", "-", "-"], input=src, capture_output=True,
                         text=True, check=True).stdout
    for sym in knnhip.CXX_SYMBOLS:
        assert sym in out


@pytest.mark.skipif(not os.path.exists(REF_TEST), reason="reference test program not built")
def test_reference_test_program_links_against_library():
    out = subprocess.run(["ldd", REF_TEST], capture_output=True, text=True, check=True).stdout
    line = [l for l in out.splitlines() if "libknnhip.so" in l]
    assert line and "not found" not in line[0]


@pytest.mark.parametrize("entry", ["knn_cuda_global", "knn_cuda_texture", "knn_cublas"])
@pytest.mark.parametrize("case", ["k_too_large", "k_zero", "no_points"])
def test_bad_arguments_return_false_with_error_line(entry, case, capfd):
    ref = np.zeros((4, 10), dtype=np.float32)
    qry = np.zeros((4, 3), dtype=np.float32)
    k = {"k_too_large": 11, "k_zero": 0, "no_points": 2}[case]
    if case == "no_points":
        ref = np.zeros((4, 0), dtype=np.float32)
    with pytest.raises(knnhip.KnnError):
        getattr(knnhip, entry)(ref, qry, k)
    out = capfd.readouterr().out
    assert out.startswith("ERROR:")


def _has_gpu():
    try:
        import torch

        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present")
def test_no_device_is_reported_like_the_reference(capfd):
    # code/knncuda.cu:289-293: no device -> "ERROR: ..." and false
    ref = np.zeros((4, 10), dtype=np.float32)
    with pytest.raises(knnhip.KnnError, match="No HIP device"):
        knnhip.knn_cuda_global(ref, ref[:, :3].copy(), 2)
    assert "ERROR: No HIP device found" in capfd.readouterr().out


def test_options_struct_matches_header():
    text = open(HEADER).read()
    fields = re.search(r"typedef struct knnhip_options \{(.*?)\} knnhip_options;", text, re.S).group(1)
    names = re.findall(r"int\s+(\w+)(?:\[\d+\])?;", fields)
    assert names == [f[0] for f in knnhip.Options._fields_]
    arrays = [int(n) for n in re.findall(r"int\s+\w+\[(\d+)\];", fields)]
    ints = len(names) - len(arrays) + sum(arrays)
    assert ctypes.sizeof(knnhip.Options) == 4 * ints
    assert ints == 13  # the layout callers compiled against: 13 ints
