"""The C-ABI library loads and exports every entry point its header declares
(no compute: no GPU needed); host-only functions agree with the oracle."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(pbig_hip_\w+)\s*\(", text)))


def dynsyms(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_header_declares_exports():
    import pbig_amd

    assert declared("pbig_hip.h") == sorted(pbig_amd.EXPORTS)


def test_library_exports_every_declared_symbol():
    import pbig_amd

    lib = pbig_amd.load_library()
    syms = dynsyms(pbig_amd.library_path())
    for name in declared("pbig_hip.h"):
        assert name in syms, name
        assert hasattr(lib, name)


def test_cpp_api_library_exports_reference_symbols():
    """libpbig.so (float 2-D records) and libpbig_u32.so (the reference's
    unsigned 3-D records) export the reference's C++ API (src/pbig/pbig.h:9-11)
    with the same mangled names, plus report's out-of-line members."""
    for lib in ("libpbig.so", "libpbig_u32.so"):
        syms = dynsyms(os.path.join(ROOT, "pbig_amd", lib))
        for name in ("_Z9pbig_initjjjjj", "_Z10pbig_checkP8RectInfoP6report", "_Z13pbig_finalizev",
                     "_ZN6reportC1Ev", "_ZN6report5printEv", "_ZN6report5foundEPK8RectInfoS2_"):
            assert name in syms, (lib, name)


def test_reference_cell_size_u32_matches_oracle(orc):
    """The unsigned cell-size rule against the oracle's (pinned by the
    reference's trace: cs 128 at 10^7 boxes)."""
    import pbig_amd

    for sim, n, rng_, S, dims in (("UU", 1000, 10000, 100, 3), ("UN", 5000, 10000, 100, 3),
                                  ("UU", 20000, 4000, 30, 2), ("NU", 300, 100000, 500, 3)):
        b = orc.ref_generate(n, rng_, S, sim, dims)
        assert pbig_amd.reference_cell_size_u32(b, dims, rng_) == orc.decide_cell_size_u32(b, dims, rng_)


def test_reference_cell_size_matches_oracle(orc):
    import pbig_amd
    from pbig_amd import workloads

    rng = np.random.default_rng(5)
    cases = [workloads.c1_small_uniform(), (workloads.uniform(300000, 20000.0, 1, 50, 9), 20000.0)]
    for _ in range(5):
        n = int(rng.integers(1, 5000))
        L = float(rng.uniform(10, 5000))
        cases.append((workloads.uniform(n, L, 0.5, float(rng.uniform(1, 100)), int(rng.integers(1 << 30))), L))
    for b, L in cases:
        assert pbig_amd.reference_cell_size(b, L) == orc.decide_cell_size(b, L)


def test_rect_layout_matches_reference_record():
    import pbig_amd

    assert pbig_amd.RECT_DTYPE.itemsize == 20  # RectInfo, src/pbig/rect.h:15-18 (2-D)
    r = pbig_amd.as_rects(np.arange(8, dtype=np.float32).reshape(2, 4))
    assert r["rid"].tolist() == [0, 1]


def test_report_layout_per_build(tmp_path):
    """libpbig_u32.so is the reference's shipped build: its report object
    keeps the reference's layout (one int, src/pbig/report.h:15), so a
    caller compiled against the reference's header can link to it.  The
    float build's report holds a 64-bit counter.  (The header checks both
    with static_assert; this compiles it each way.)"""
    import subprocess

    src = tmp_path / "r.cpp"
    src.write_text('#include "report.h"\nint size_of_report() { return (int)sizeof(report); }\n')
    inc = os.path.join(ROOT, "include")
    for flags in (["-DPBIG_COMPAT_U32", "-DDIMS=3"], ["-DPBIG_COMPAT_U32", "-DDIMS=2"], []):
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", f"-I{inc}", *flags, str(src)], check=True)
