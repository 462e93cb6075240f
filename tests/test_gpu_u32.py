"""GPU tests of the reference-compatible unsigned mode (the reference's
shipped build: CType = unsigned int, DIMS = 3 or 2) against the reference's
own logged answers and the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


@pytest.fixture(scope="module")
def eng(torch_cuda):
    import pbig_amd

    e = pbig_amd.Engine(0)
    yield e
    e.close()


def dev(torch, b):
    return torch.from_numpy(np.ascontiguousarray(b, dtype=np.uint32).view(np.int32)).to("cuda:0")


def test_reference_logged_answer_test1(torch_cuda, eng, orc):
    """benchmarks/test1/filter_file:7: 10^7 boxes, UU, S=100, range 10^4,
    srand(100), 3-D unsigned: 49,514,252 pairs, with the reference's rule
    choosing the cell size (128)."""
    b = orc.ref_generate(10**7, 10000, 100, "UU", 3)
    n = eng.check_u32(dev(torch_cuda, b), 3, 10000, 0)
    st = eng.stats()
    assert st["cell_size"] == 128
    assert n == 49514252


def test_reference_logged_answer_test3(torch_cuda, eng, orc):
    """benchmarks/test3/filter_file:17-49: 10^7 boxes, UN (normal centres):
    cs 128, 493,039 cells, 26,668,901 memberships, max 3,543 per cell,
    1,121,815,231 pairs."""
    b = orc.ref_generate(10**7, 10000, 100, "UN", 3)
    n = eng.check_u32(dev(torch_cuda, b), 3, 10000, 0)
    st = eng.stats()
    assert st["cell_size"] == 128
    assert st["n_cells"] == 493039
    assert st["memberships"] == 26668901
    assert st["max_per_cell"] == 3543
    assert n == 1121815231


@pytest.mark.parametrize("dims", [3, 2])
@pytest.mark.parametrize("sim", ["UU", "UN", "NU"])
@pytest.mark.parametrize("cs", [16, 64, 512])
def test_pairs_match_brute_force(torch_cuda, eng, orc, dims, sim, cs):
    b = orc.ref_generate(20000, 3000, 150, sim, dims)
    want = orc.sort_pairs(orc.brute_u32(b, dims))
    out = torch_cuda.empty((len(want) + 1024, 2), dtype=torch_cuda.int32, device="cuda:0")
    n = eng.check_u32(dev(torch_cuda, b), dims, 3000, cs, out)
    assert n == len(want)
    got = orc.sort_pairs(out[:n].cpu().numpy().view(np.uint32))
    assert np.array_equal(got, want)
    # count-only mode gives the same number
    assert eng.check_u32(dev(torch_cuda, b), dims, 3000, cs) == n


def test_host_records_entry_point(torch_cuda, eng, orc):
    b = orc.ref_generate(5000, 2000, 100, "UN", 3)
    want = orc.sort_pairs(orc.brute_u32(b, 3))
    got = orc.sort_pairs(eng.check_host_u32(b, 3, 2000, 0))
    assert np.array_equal(got, want)
    # the pairs handed to a callback chunk by chunk
    chunks = []
    n = eng.check_host_u32_each(b, 3, 2000, 0, lambda p: chunks.append(p.copy()))
    assert n == want.shape[0]
    assert np.array_equal(orc.sort_pairs(np.concatenate(chunks)), want)


def test_out_of_range_coordinates_follow_the_reference(torch_cuda, eng, orc):
    """Coordinates at or past the range alias into other cells through the
    reference's 32-bit cell ids (and ids past the grid are skipped); the
    count must be the literal restatement's."""
    rng = np.random.default_rng(3)
    b = orc.ref_generate(4000, 1000, 80, "UU", 3)
    k = rng.integers(0, 4000, 300)
    b[k, 0::2] += 900  # push some boxes past range 1000
    b[k, 1::2] += 900
    for cs in (32, 128):
        want = orc.grid_u32(b, 3, 1000, cs)["pairs"]
        assert eng.check_u32(dev(torch_cuda, b), 3, 1000, cs) == want


def test_rejects_bad_arguments(torch_cuda, eng, orc):
    import pbig_amd

    b = orc.ref_generate(100, 1000, 50, "UU", 3)
    with pytest.raises(pbig_amd.PbigError):
        eng.check_u32(dev(torch_cuda, b), 3, 1000, 48)  # not a power of two
    with pytest.raises(pbig_amd.PbigError):
        eng.check_u32(dev(torch_cuda, b), 4, 1000, 64)
