"""awp_misfit_noalloc (partial sums in the library's own device array, no
allocation) against awp_misfit (partial sums in a caller-given work array):
residuals and misfit bit for bit, on two calls in a row, the second with a
smaller input, so the second call runs over the first call's scratch."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _misfit(torch, lib, fn, d, obs, work=None):
    n = d.numel()
    res = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    J = torch.full((1,), np.nan, dtype=torch.float64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    if work is None:
        fn(d.data_ptr(), obs.data_ptr(), n, res.data_ptr(), J.data_ptr(), st)
    else:
        fn(d.data_ptr(), obs.data_ptr(), n, res.data_ptr(), J.data_ptr(), work.data_ptr(), st)
    torch.cuda.synchronize()
    return res.cpu().numpy(), float(J.item())


def test_misfit_noalloc_matches_misfit(torch_gpu):
    torch = torch_gpu
    from awp_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(11)
    # large first, then smaller: 300,001 samples spread over all 256 blocks,
    # then 777 samples, where most blocks sum nothing
    for n in (300_001, 777):
        _say(f"misfit_noalloc: n = {n}")
        d = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
        obs = torch.from_numpy(rng.standard_normal(n).astype(np.float32)).cuda()
        work = torch.full((lib_parts(),), np.nan, dtype=torch.float64, device="cuda")
        r_ref, J_ref = _misfit(torch, lib, lib.awp_misfit, d, obs, work)
        _lib.check()
        r_na, J_na = _misfit(torch, lib, lib.awp_misfit_noalloc, d, obs)
        _lib.check()
        assert np.array_equal(r_na, r_ref)
        assert J_na == J_ref, (J_na, J_ref)
        # the sum itself, in binary64 on the host
        r64 = d.cpu().numpy().astype(np.float64) - obs.cpu().numpy().astype(np.float64)
        assert np.array_equal(r_na, r64)
        assert abs(J_na - 0.5 * np.sum(r64 * r64)) <= 1e-12 * abs(J_na)


def test_misfit_noalloc_rejects_negative_n(torch_gpu):
    from awp_amd import _lib
    lib = _lib.load()
    lib.awp_clear_error()
    lib.awp_misfit_noalloc(None, None, -1, None, None, None)
    assert lib.awp_last_error() == 1
    assert b"negative" in lib.awp_last_error_string()
    lib.awp_clear_error()


def lib_parts():
    # AWP_MISFIT_PARTS in include/awp_hip.h
    return 256
