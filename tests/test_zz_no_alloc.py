"""The allocation-free metric entry point (gsea_metric_noalloc) against
gsea_metric_f32/f64 through the C ABI: bit-identical metric rows on two
calls in a row (the second smaller), for every metric, both precisions, the
label swap and a permutation shift; oversized calls rejected; no device
memory allocated by the call."""
import ctypes

import numpy as np
import pytest

from cudagsea_amd import workloads

pytestmark = pytest.mark.gpu


def _dp(t):
    return ctypes.c_void_p(t.data_ptr())


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _metric_pair(L, X, lab, P, shift, mid, flags, dbl):
    import torch
    S, N = X.shape
    tdt = torch.float64 if dbl else torch.float32
    ref = torch.full((P, N), -7.0, dtype=tdt, device="cuda")
    got = torch.full((P, N), 7.0, dtype=tdt, device="cuda")
    fn = L.gsea_metric_f64 if dbl else L.gsea_metric_f32
    assert fn(_dp(X), N, S, _ptr(lab), P, shift, mid, flags, _dp(ref), None) == 0
    rc = L.gsea_metric_noalloc(int(dbl), _dp(X), N, S, _ptr(lab), P, shift, mid, flags,
                               _dp(got), None)
    assert rc == 0, L.gsea_last_error()
    return got.cpu().numpy(), ref.cpu().numpy()


def _same_bits(a, b):
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    return ((a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))).all()


@pytest.mark.parametrize("dbl", [False, True])
def test_noalloc_metric_equals_allocating_entry_point(gpu, dbl):
    import torch
    L = gpu.lib()
    dt = np.float64 if dbl else np.float32
    big = workloads.make("c3", num_genes=2000, num_samples=120, num_perms=96)
    small = workloads.make("c1", num_genes=700, num_samples=30, num_perms=20)
    for mid in range(14):
        for w, P, shift, flags in ((big, 96, 3, 0), (small, 20, 0, 2)):
            print("noalloc", "fp64" if dbl else "fp32", "metric", mid, "N", w["N"],
                  "P", P, flush=True)
            X = torch.from_numpy(w["X"].astype(dt)).cuda()
            lab = np.ascontiguousarray(w["labels"], dtype=np.uint8)
            got, ref = _metric_pair(L, X, lab, P, shift, mid, flags, dbl)
            assert _same_bits(got, ref), (mid, w["N"], P)


def test_noalloc_rejects_oversized_calls_and_allocates_nothing(gpu):
    import torch
    L = gpu.lib()
    w = workloads.make("c1", num_genes=500, num_samples=500, num_perms=8)
    X = torch.from_numpy(w["X"]).cuda()
    lab = np.ascontiguousarray(w["labels"], dtype=np.uint8)
    r = torch.empty((30000, 500), dtype=torch.float32, device="cuda")
    # 30,000 permutations x 500 samples need ~45 MB of scratch: rejected
    rc = L.gsea_metric_noalloc(0, _dp(X), 500, 500, _ptr(lab), 30000, 0, 10, 0, _dp(r), None)
    assert rc == 31 and b"scratch" in L.gsea_last_error()
    # a call that fits leaves the device's free memory unchanged (the
    # static scratch is resident from the first call on)
    assert L.gsea_metric_noalloc(0, _dp(X), 500, 500, _ptr(lab), 64, 0, 10, 0, _dp(r),
                                 None) == 0
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for P in (8000, 100, 5000):
        assert L.gsea_metric_noalloc(0, _dp(X), 500, 500, _ptr(lab), P, 1, 10, 0, _dp(r),
                                     None) == 0
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0
