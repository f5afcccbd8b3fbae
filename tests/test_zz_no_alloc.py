"""cm_ihc_lambda_run_noalloc (the IHC/lambda chain with no device allocation)
against cm_ihc_lambda_run on the same inputs, bit for bit, on two calls in a
row, the second with a smaller input, and its refusals of requests beyond the
staging array."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
FS = 44100.0


def _cuda():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _noalloc(bi, d_bm, rows, ivs, d_lambda, d_stages=None, n_ac=None):
    from cochlearmodel_amd import _lib, stages
    from cochlearmodel_amd._lib import ptr
    L = _lib.lib()
    ivs = np.ascontiguousarray(ivs, _lib.IHC_INTERVAL_DTYPE)
    return L.cm_ihc_lambda_run_noalloc(ptr(d_bm), rows, ptr(ivs), len(ivs), ptr(bi.b),
                                       bi.n_ac if n_ac is None else n_ac, ptr(bi.dc), bi.n_dc, ptr(bi.damage),
                                       ptr(bi.A), bi.spont, stages.LAMBDA_SAT, stages.ETA_AC, stages.ETA_DC, 1.0,
                                       ptr(d_lambda), ptr(d_stages))


def test_noalloc_matches_the_allocating_entry_point():
    torch = _cuda()
    from cochlearmodel_amd import _lib, stages
    bi = stages.BatchedIHC(FS, ihc_vectors=[8.0, 6.0])
    rng = np.random.default_rng(31)
    L = _lib.lib()
    L.cm_set_stream(C.c_void_p(torch.cuda.current_stream().cuda_stream))
    # two calls in a row, the second smaller: the staging array keeps the first
    # call's tables, which the second must not read
    for rows, lens in ((9000, [4000, 3100, 1900]), (2500, [2500])):
        print(f"case: no-alloc IHC/lambda, {rows} rows in {len(lens)} intervals", flush=True)
        bm = (1e-3 * rng.standard_normal((rows, 256))).astype(np.float32)
        d_bm = torch.from_numpy(bm).cuda()
        starts = np.cumsum([0] + lens[:-1])
        ivs = stages.signal_intervals(starts, lens, sets=[k % 2 for k in range(len(lens))])
        lam_ref, st_ref = bi.run(d_bm, rows, ivs, stages=True)
        d_lam = torch.full((3, rows, 256), float("nan"), dtype=torch.float32, device="cuda")
        d_st = torch.full((5, rows, 256), float("nan"), dtype=torch.float32, device="cuda")
        _lib.check(_noalloc(bi, d_bm, rows, ivs, d_lam, d_st), "cm_ihc_lambda_run_noalloc")
        torch.cuda.synchronize()
        np.testing.assert_array_equal(d_lam.cpu().numpy(), lam_ref.cpu().numpy())
        np.testing.assert_array_equal(d_st.cpu().numpy(), st_ref.cpu().numpy())


def test_noalloc_refuses_requests_beyond_its_array():
    torch = _cuda()
    from cochlearmodel_amd import _lib, stages
    bi = stages.BatchedIHC(FS)
    rows = 64
    d_bm = torch.zeros((rows, 256), dtype=torch.float32, device="cuda")
    d_lam = torch.empty((3, rows, 256), dtype=torch.float32, device="cuda")
    L = _lib.lib()
    many = stages.signal_intervals(np.zeros(4097, np.int64), np.ones(4097, np.int64))
    assert _noalloc(bi, d_bm, rows, many, d_lam) == -3 and "at most" in L.cm_last_error().decode()
    assert _noalloc(bi, d_bm, rows, stages.signal_intervals([0], [rows]), d_lam, n_ac=513) == -3
    cont = stages.signal_intervals([0, 32], [32, 32])
    cont["kbase"] = [0, 32]
    assert _noalloc(bi, d_bm, rows, cont, d_lam) == -3 and "kbase" in L.cm_last_error().decode()
    out = stages.signal_intervals([0], [rows + 1])
    assert _noalloc(bi, d_bm, rows, out, d_lam) == -3
