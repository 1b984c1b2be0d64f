"""bsg_sigma_apply_noalloc (partial tiles in the library's static device
scratch, no allocation) against bsg_sigma_apply through the C-ABI: bitwise
the same sigma, FP64 and split precision, on two calls in a row with the
second superblock smaller than the first; and the FP64 result against the
oracle."""
import ctypes

import numpy as np
import pytest

import _oracle
import block_amd
from capi import check, lib, vp
from helpers import config_fcidump, random_fcidump, rng_vec

pytestmark = pytest.mark.gpu


def _run(sb, seed):
    h = block_amd.multiply_h(sb)
    ctx, plan, arena, dc, _ = h.raw()
    n = sb.wave_size
    need = lib.bsg_plan_partial_doubles(plan)
    print(f"[gpu] noalloc n={n} partial tiles {need} doubles", flush=True)
    c = rng_vec(n, seed)
    bufs = [vp(), vp()]
    for b in bufs:
        check(lib.bsg_malloc(ctx, 8 * n, ctypes.byref(b)))
    out = {}
    try:
        check(lib.bsg_memcpy_h2d(ctx, dc, c.ctypes.data, 8 * n))
        for prec in (block_amd.FP64, block_amd.SPLIT_BF16):
            h.precision = prec  # split: sets up the plan's slices (an allocating call, before)
            got = []
            for b, f in zip(bufs, (lib.bsg_sigma_apply, lib.bsg_sigma_apply_noalloc)):
                check(lib.bsg_memset_zero(ctx, b, 8 * n))
                check(f(ctx, plan, arena, dc, b, 1, prec))
                y = np.empty(n)
                check(lib.bsg_memcpy_d2h(ctx, y.ctypes.data, b, 8 * n))
                got.append(y)
            assert np.array_equal(got[0], got[1]), prec
            out[prec] = got[1]
    finally:
        for b in bufs:
            if b.value:
                lib.bsg_mfree(ctx, b)
    h.precision = block_amd.FP64
    ref = _oracle.sigma(sb, c, nthreads=8)
    rel = np.linalg.norm(out[block_amd.FP64] - ref) / np.linalg.norm(ref)
    assert rel <= 1e-12
    rel_s = np.linalg.norm(out[block_amd.SPLIT_BF16] - ref) / np.linalg.norm(ref)
    assert 0.0 < rel_s <= 2e-6
    return need


def test_noalloc_sigma_matches_allocating_entry():
    big = block_amd.SuperBlock(config_fcidump("hubbard12"), 12, 0, 0, 5, 100)
    small = block_amd.SuperBlock(random_fcidump(16, 21), 16, 0, 0, 7, 80)
    assert small.wave_size < big.wave_size
    need_big = _run(big, 3)
    assert need_big > 0  # the static scratch really holds partial tiles
    _run(small, 4)
