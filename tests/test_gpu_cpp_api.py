"""The reference-compatible C++ API (include/pbig.h) through a small driver
compiled against libpbig.so (float 2-D records) and libpbig_u32.so (the
reference's unsigned 3-D records): pbig_check's report count, and the
compressed stream's round trip, against the oracle."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def build_driver(tmp_path, lib, defines):
    exe = str(tmp_path / f"driver_{lib}")
    subprocess.run(["g++", "-O1", "-std=c++17", *defines, f"-I{ROOT}/include",
                    os.path.join(ROOT, "tests", "cpp", "api_driver.cpp"), f"-L{ROOT}/pbig_amd", f"-l{lib}",
                    f"-Wl,-rpath,{ROOT}/pbig_amd", "-o", exe], check=True)
    return exe


def run(exe, data, n, range_, cs):
    out = subprocess.run([exe, str(data), str(n), str(range_), str(cs)], check=True, capture_output=True,
                         text=True, timeout=120).stdout.split()
    return int(out[1]), int(out[3])


def test_float_api_and_stream(torch_cuda, tmp_path, orc):
    from pbig_amd import workloads

    boxes, L = workloads.c1_small_uniform()
    f = tmp_path / "boxes.f32"
    boxes.astype("<f4").tofile(f)
    exe = build_driver(tmp_path, "pbig", [])
    pairs, stream_pairs = run(exe, f, boxes.shape[0], int(L), 0)
    want = len(orc.brute_pairs(boxes))
    assert pairs == want and stream_pairs == want


def test_u32_api_and_stream(torch_cuda, tmp_path, orc):
    b = orc.ref_generate(200000, 10000, 100, "UN", 3)
    f = tmp_path / "boxes.u32"
    b.astype("<u4").tofile(f)
    exe = build_driver(tmp_path, "pbig_u32", ["-DPBIG_COMPAT_U32", "-DDIMS=3"])
    pairs, stream_pairs = run(exe, f, b.shape[0], 10000, 0)
    want = orc.grid_u32(b, 3, 10000, orc.decide_cell_size_u32(b, 3, 10000))["pairs"]
    assert pairs == want and stream_pairs == want
