"""The allocation-free permutation entry point
(bihip_cumulative_offspring_to_ancestors_permute_noalloc) against the
allocating one and the oracle, through the C ABI: bit-exact ancestors on
consecutive calls that share the static scratch, the second call smaller than
the first, and the size limit rejected before any kernel runs."""
import ctypes

import numpy as np
import pytest

from libbi_amd import capi
from gpu_helpers import dev, empty, host

pytestmark = pytest.mark.gpu


def _c(x):
    return ctypes.c_void_p(x.data_ptr())


def _cases(P, rng):
    yield "multinomial", rng.multinomial(P, np.ones(P) / P)
    o = np.zeros(P, dtype=np.int64)
    o[P // 3] = P
    yield "one survivor", o
    yield "dirichlet 0.05", rng.multinomial(P, rng.dirichlet(np.full(P, 0.05)))


def _both(ctx, torch, O):
    P = len(O)
    Od = dev(O.astype(np.int32))
    a_alloc = empty(P, torch.int32)
    a_free = empty(P, torch.int32)
    a_free.fill_(-7)
    L = capi.lib()
    capi.check(L.bihip_cumulative_offspring_to_ancestors_permute(ctx.h, _c(Od), P, _c(a_alloc)))
    capi.check(L.bihip_cumulative_offspring_to_ancestors_permute_noalloc(ctx.h, _c(Od), P,
                                                                         _c(a_free)))
    return host(a_alloc), host(a_free)


def test_noalloc_permute_equals_allocating_entry_point(ctx, orc, torch_gpu):
    rng = np.random.default_rng(2026)
    # the largest size first, then smaller ones: the later calls run on the
    # scratch the earlier ones left behind
    for P in (capi.NOALLOC_MAX_P, 100003, 2049, 1):
        for name, o in _cases(P, rng):
            print(f"noalloc permute: P={P} case={name}", flush=True)
            O = np.cumsum(o)
            a_alloc, a_free = _both(ctx, torch_gpu, O)
            np.testing.assert_array_equal(a_free, a_alloc)
            np.testing.assert_array_equal(a_free, orc.cumulative_offspring_to_ancestors_permute(
                O.astype(np.int32)))


def test_noalloc_permute_rejects_sizes_beyond_the_scratch(ctx, torch_gpu):
    L = capi.lib()
    P = capi.NOALLOC_MAX_P + 1
    Od = dev(np.arange(1, P + 1, dtype=np.int32))
    a = empty(P, torch_gpu.int32)
    a.fill_(-7)
    rc = L.bihip_cumulative_offspring_to_ancestors_permute_noalloc(ctx.h, _c(Od), P, _c(a))
    assert rc == capi.EINVAL
    assert b"static scratch" in L.bihip_last_error()
    assert rc == L.bihip_cumulative_offspring_to_ancestors_permute_noalloc(ctx.h, _c(Od), 0, _c(a))
    assert (host(a) == -7).all()
