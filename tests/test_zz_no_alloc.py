"""GPU test of pbig_hip_check_device_noalloc (the entry point that allocates
nothing): its pairs against pbig_hip_check_device_async's on the same input,
sorted and bit-identical, on two calls in a row, the second with a smaller
input (the shared scratch array keeps the first call's contents); and the
size limit refused before anything is launched."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _pairs(torch, pairs, res):
    import oracle

    n, err = (int(v) for v in res.cpu().tolist())
    assert err == 0, f"error word {err}"
    return oracle.sort_pairs(pairs[:n].cpu().numpy().view(np.uint32))


def test_noalloc_matches_async_twice(torch_cuda):
    import oracle
    import pbig_amd
    from pbig_amd import workloads

    torch = torch_cuda
    cases = [(workloads.uniform(600_000, 65536.0 * (600_000 / 4194304) ** 0.5, 1.0, 32.0, 901),
              65536.0 * (600_000 / 4194304) ** 0.5, 32.0),
             (workloads.uniform(70_000, 1500.0, 1.0, 8.0, 902), 1500.0, 6.0)]
    with pbig_amd.Engine(0) as eng:
        for k, (boxes, L, cell) in enumerate(cases):
            print(f"case {k}: {boxes.shape[0]} boxes, L {L:.1f}, cell {cell}", flush=True)
            d = torch.from_numpy(boxes).to("cuda:0")
            cap = 2 * boxes.shape[0]
            p_async = torch.full((cap, 2), -1, dtype=torch.int32, device="cuda:0")
            p_none = torch.full((cap, 2), -1, dtype=torch.int32, device="cuda:0")
            r_async = torch.zeros(2, dtype=torch.int64, device="cuda:0")
            r_none = torch.zeros(2, dtype=torch.int64, device="cuda:0")
            eng.check_device(d, L, cell, p_async)  # sizes the allocating path's workspace
            eng.check_device_async(d, L, cell, p_async, r_async)
            eng.check_device_noalloc(d, L, cell, p_none, r_none)
            torch.cuda.synchronize()
            want = _pairs(torch, p_async, r_async)
            got = _pairs(torch, p_none, r_none)
            assert got.shape[0] > 0
            assert np.array_equal(got, want), (k, got.shape[0], want.shape[0])
            if k == 1:  # the small case against the oracle as well
                assert np.array_equal(got, oracle.sort_pairs(oracle.sweep_pairs(boxes)))


def test_noalloc_rejects_oversized_input(torch_cuda):
    import pbig_amd

    torch = torch_cuda
    with pbig_amd.Engine(0) as eng:
        n = 3 << 20
        d = torch.zeros((n, 4), dtype=torch.float32, device="cuda:0")
        pairs = torch.empty((16, 2), dtype=torch.int32, device="cuda:0")
        res = torch.zeros(2, dtype=torch.int64, device="cuda:0")
        with pytest.raises(pbig_amd.PbigError) as e:
            eng.check_device_noalloc(d, 65536.0, 32.0, pairs, res)
        assert e.value.status == pbig_amd.PBIG_E_ARG
        assert "scratch" in eng.last_error()
