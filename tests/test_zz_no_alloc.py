"""The allocation-free LARS entry point (spx_lasso_dtx_dev_noalloc_f32),
called through the C ABI, against the allocating one
(spx_lasso_dtx_dev_f32): the (value, index) lists and counts equal bit for
bit on two calls in a row, the second with fewer signals, at the headline
shape with per-signal lambda small enough that active sets outgrow the LDS
pass (hand-over states and inverse slots in the static scratch); device
memory in use does not change across the allocation-free calls; a call
whose scratch exceeds the static array fails with a message."""
import ctypes

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def _lists(torch, n, Lout):
    V = torch.full((n, Lout), -7.0, dtype=torch.float32, device="cuda")
    R = torch.full((n, Lout), -7, dtype=torch.int32, device="cuda")
    c = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    return V, R, c


def _equal(a, b):
    (Va, Ra, ca), (Vb, Rb, cb) = a, b
    ca, cb = ca.cpu().numpy(), cb.cpu().numpy()
    assert np.array_equal(ca, cb)
    Va, Ra, Vb, Rb = (t.cpu().numpy() for t in (Va, Ra, Vb, Rb))
    for i in range(len(ca)):
        k = ca[i]
        assert np.array_equal(Ra[i, :k], Rb[i, :k]), i
        assert np.array_equal(Va[i, :k].view(np.uint32), Vb[i, :k].view(np.uint32)), i


def test_noalloc_lasso_equals_allocating_entry_point():
    import torch

    import spx
    print("case noalloc lasso m=256 p=1024", flush=True)
    ctx = spx.Context(0)
    ctx.set_stream(torch.cuda.current_stream().cuda_stream)
    X, D, lam = W.lars_headline(seed=91, n=4000)
    lam = (lam * 0.5).astype(np.float32)          # longer paths: the hand-over runs
    d = spx.Dictionary(ctx, D)
    m, p = D.shape
    Xt = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    lt = torch.from_numpy(lam).cuda()
    ldg = (p + 15) // 16 * 16
    DtX = torch.zeros((4000, ldg), dtype=torch.float32, device="cuda")
    norms = torch.zeros((4000,), dtype=torch.float32, device="cuda")
    spx.dtx_dev(ctx, d, Xt, DtX)
    spx.sqnorms_dev(ctx, Xt, norms)
    Lout = spx.lasso_L(m, p, -1, spx.PENALTY, float(lam.max()))
    runs = []
    for n in (4000, 1500):                          # two calls, the second smaller
        a = _lists(torch, n, Lout)
        b = _lists(torch, n, Lout)
        spx.lasso_dtx_dev(ctx, d, DtX[:n], norms[:n], lt[:n], *a)
        torch.cuda.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        spx.lasso_dtx_dev(ctx, d, DtX[:n], norms[:n], lt[:n], *b, noalloc=True)
        torch.cuda.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        assert free1 >= free0 - (2 << 20), (free0, free1)
        _equal(a, b)
        assert int(a[2].max().item()) > 24          # the full-capacity pass ran
        runs.append(n)
    # the same call straight through ctypes, with lambda2 > 1e-10 (a ridge
    # Gram formed in the static scratch) and the plain form
    f = spx.lib().spx_lasso_dtx_dev_noalloc_f32
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                  ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_float,
                  ctypes.POINTER(spx.spx_lasso_opts), ctypes.POINTER(spx.spx_lists_dev)]
    n = 800
    for form in ("batched", "plain"):
        a = _lists(torch, n, Lout)
        b = _lists(torch, n, Lout)
        spx.lasso_dtx_dev(ctx, d, DtX[:n], norms[:n], lt[:n], *a, lambda2=0.01, form=form)
        o = spx.spx_lasso_opts(-1, spx.PENALTY, 0, -1, 0, spx.FORM[form], 0)
        lists = spx.spx_lists_dev(b[0].data_ptr(), b[1].data_ptr(), b[2].data_ptr(), Lout, None)
        rc = f(ctx.h, d.h, DtX.data_ptr(), ldg, norms.data_ptr(), n, lt.data_ptr(), n, 0.01,
               ctypes.byref(o), ctypes.byref(lists))
        assert rc == 0, spx.lib().spx_last_error()
        torch.cuda.synchronize()
        _equal(a, b)
    # too many signals for the static scratch: refused with a message
    big = 40000
    DtXb = torch.zeros((big, ldg), dtype=torch.float32, device="cuda")
    nb = torch.ones((big,), dtype=torch.float32, device="cuda")
    lb = torch.full((1,), 0.05, dtype=torch.float32, device="cuda")
    Vb, Rb, cb = _lists(torch, big, Lout)
    with pytest.raises(RuntimeError, match="static scratch"):
        spx.lasso_dtx_dev(ctx, d, DtXb, nb, lb, Vb, Rb, cb, noalloc=True)
    d.close()
    ctx.close()
