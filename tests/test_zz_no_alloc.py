"""The allocation-free preconditioned apply, heomq_apply_packed_precond_noalloc,
against the allocating entry point it varies, heomq_apply_packed_precond
(form 1), through the C ABI.

Both run the same kernels on the same inputs, so the outputs must be
bit-identical.  Two calls in a row, the second on a smaller problem: the
static device scratch keeps the first call's contents, which the second call
must overwrite wherever it reads.
"""
import ctypes as C

import pytest

import heomq
from heomq import _lib

pytestmark = pytest.mark.gpu


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def test_noalloc_matches_the_allocating_entry_point():
    import torch
    L = _lib.lib()
    big = heomq.Context(heomq.Model(heomq.two_impurity(nk=2, ntop=2)))
    small = heomq.Context(heomq.Model(heomq.siam(nk=2, ntop=2, V=0.2)))
    assert big.packed_info()["packed"] > small.packed_info()["packed"]
    for ctx in (big, small):
        n = ctx.packed_info()["packed"]
        print(f"packed elements {n}", flush=True)
        g = torch.Generator(device="cuda:0").manual_seed(7)
        xp = torch.randn(n, dtype=torch.complex128, device="cuda:0", generator=g)
        y_ref = torch.zeros_like(xp)
        y_na = torch.full_like(xp, float("nan"))
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(L.heomq_apply_packed_precond(ctx._h, _ptr(xp), _ptr(y_ref), heomq.TRACE_RDO, 1, st))
        _lib.check(L.heomq_apply_packed_precond_noalloc(ctx._h, _ptr(xp), _ptr(y_na), heomq.TRACE_RDO, st))
        torch.cuda.synchronize()
        assert torch.equal(y_ref, y_na)
    with pytest.raises(heomq.HeomqError):   # x and y must differ, as in the allocating form
        _lib.check(L.heomq_apply_packed_precond_noalloc(small._h, _ptr(xp), _ptr(xp), heomq.TRACE_RDO, st))
    big.close()
    small.close()
