"""GPU multiply_h through the C-ABI against the CPU oracle (direct
restatement of multiplyH / TensorMultiply): <= 1e-12 relative per sigma."""
import os
import sys

import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump, random_fcidump, rng_vec

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _cases():
    return [
        ("exact6", lambda: _oracle.exact_superblock(random_fcidump(6, 3), 6, 0, 0, 3)),
        ("exact7d2h", lambda: _oracle.exact_superblock(random_fcidump(7, 5, d2h=True, nelec=6), 6, 0, 0, 4)),
        ("hubbard12_M100", lambda: block_amd.SuperBlock(config_fcidump("hubbard12"), 12, 0, 0, 5, 100)),
        ("syn16_M80", lambda: block_amd.SuperBlock(random_fcidump(16, 21), 16, 0, 0, 7, 80)),
        ("syn16d2h_M80", lambda: block_amd.SuperBlock(random_fcidump(16, 22, d2h=True), 16, 0, 0, 7, 80)),
        # total spin S = 1: the St-dependent 9j factors and signs of the work list
        ("exact6_S1", lambda: _oracle.exact_superblock(random_fcidump(6, 3), 6, 2, 0, 3)),
        ("syn16_M80_S1", lambda: block_amd.SuperBlock(random_fcidump(16, 21), 16, 2, 0, 7, 80)),
        # mid-chain superblock of the 24-orbital DMRG warm-up at M = 500: V
        # blocks up to ~350 wide, the wide-tile kernel classes
        ("syn24_M500_dmrg", lambda: block_amd.SuperBlock.from_dmrg_warmup(config_fcidump("syn24"), 24, 0, 0, 500)),
        # the same at M = 700: a V block 479 wide, on the 768-wide FP64 tiles
        # (eight waves in one wave row)
        ("syn24_M700_dmrg", lambda: block_amd.SuperBlock.from_dmrg_warmup(config_fcidump("syn24"), 24, 0, 0, 700)),
    ]


# cases that must contain a V block wider than the 384-wide tiles
WIDE = {"syn24_M700_dmrg": 385}


@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
def test_sigma_parity(name, make):
    print(f"[gpu] sigma parity {name}", flush=True)
    sb = make()
    if name in WIDE:
        assert max(c for _, c in sb.block_shapes()) >= WIDE[name]
    h = block_amd.multiply_h(sb)
    c = rng_vec(sb.wave_size, 7)
    ref = _oracle.sigma(sb, c, nthreads=8)
    got = h(c)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"[gpu] {name} n={sb.wave_size} rel={rel:.3e}", flush=True)
    assert rel <= TOL
    # bitwise reproducible
    assert np.array_equal(h(c), got)
    # Hermiticity <x,Hy> = <Hx,y>
    x = rng_vec(sb.wave_size, 8)
    hx = h(x)
    assert abs(x @ got - hx @ c) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(got)
    # linearity
    assert np.linalg.norm(h(2.0 * c + x) - (2.0 * got + hx)) <= 1e-12 * np.linalg.norm(got)
    # diagonal
    d = h.diagonal()
    np.testing.assert_allclose(d, _oracle.diagonal(sb), rtol=0, atol=1e-12 * max(1.0, np.abs(d).max()))


SPLIT_TOL = 2e-6  # three BF16 slices: ~2^-24 per element, FP32 sums of 32-deep steps


@pytest.mark.parametrize("name,make", _cases(), ids=[c[0] for c in _cases()])
def test_split_sigma(name, make):
    """Split-precision sigma against the FP64 oracle; switching back restores
    the FP64 result bit for bit."""
    print(f"[gpu] split sigma {name}", flush=True)
    sb = make()
    h = block_amd.multiply_h(sb)
    c = rng_vec(sb.wave_size, 11)
    ref = _oracle.sigma(sb, c, nthreads=8)
    f64 = h(c)
    h.precision = block_amd.SPLIT_BF16
    got = h(c)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"[gpu] {name} split rel={rel:.3e}", flush=True)
    assert rel <= SPLIT_TOL
    assert rel > 0.0  # the BF16 path really ran
    assert np.array_equal(h(c), got)
    # Rayleigh quotient error of the split sigma
    cn = c / np.linalg.norm(c)
    assert abs(cn @ h(cn) - cn @ (ref / np.linalg.norm(c))) <= SPLIT_TOL * np.linalg.norm(ref) / np.linalg.norm(c)
    h.precision = block_amd.FP64
    assert np.array_equal(h(c), f64)


def test_copy_segments():
    """bsg_copy_segments: many device copies (segments longer than one
    chunk, empty ones, reversed order) in one launch, through the C-ABI."""
    import ctypes
    print("[gpu] bsg_copy_segments", flush=True)
    lib = block_amd.sigma_lib
    vp = ctypes.c_void_p

    class Seg(ctypes.Structure):
        _fields_ = [("src", vp), ("dst", vp), ("n", ctypes.c_int64)]

    lib.bsg_init.argtypes = [ctypes.POINTER(vp), ctypes.c_int]
    lib.bsg_malloc.argtypes = [vp, ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.bsg_mfree.argtypes = [vp, vp]
    lib.bsg_free.argtypes = [vp]
    lib.bsg_memcpy_h2d.argtypes = [vp, vp, vp, ctypes.c_size_t]
    lib.bsg_memcpy_d2h.argtypes = [vp, vp, vp, ctypes.c_size_t]
    lib.bsg_memset_zero.argtypes = [vp, vp, ctypes.c_size_t]
    lib.bsg_copy_segments.argtypes = [vp, ctypes.POINTER(Seg), ctypes.c_int32]
    n = 100_000
    x = rng_vec(n, 5)
    ctx, dx, dy = vp(), vp(), vp()
    assert lib.bsg_init(ctypes.byref(ctx), 0) == 0
    try:
        assert lib.bsg_malloc(ctx, 8 * n, ctypes.byref(dx)) == 0
        assert lib.bsg_malloc(ctx, 8 * n, ctypes.byref(dy)) == 0
        assert lib.bsg_memcpy_h2d(ctx, dx, x.ctypes.data, 8 * n) == 0
        assert lib.bsg_memset_zero(ctx, dy, 8 * n) == 0
        # (src offset, dst offset, length): 40000 spans three 16384-double chunks
        cuts = [(0, 50_000, 40_000), (40_000, 0, 7), (60_000, 10, 0), (90_000, 20, 10_000), (40_007, 95_000, 5_000)]
        segs = (Seg * len(cuts))(*[Seg(dx.value + 8 * s, dy.value + 8 * d, m) for s, d, m in reversed(cuts)])
        assert lib.bsg_copy_segments(ctx, segs, len(cuts)) == 0
        y = np.empty(n)
        assert lib.bsg_memcpy_d2h(ctx, y.ctypes.data, dy, 8 * n) == 0
        want = np.zeros(n)
        for s, d, m in cuts:
            want[d:d + m] = x[s:s + m]
        assert np.array_equal(y, want)
    finally:
        for p in (dx, dy):
            if p.value:
                lib.bsg_mfree(ctx, p)
        lib.bsg_free(ctx)


BATCH_CASES = ["exact7d2h", "hubbard12_M100", "syn16_M80", "syn24_M500_dmrg"]


@pytest.mark.parametrize("name,make", [c for c in _cases() if c[0] in BATCH_CASES], ids=BATCH_CASES)
def test_batched_roots(name, make):
    """Several roots in one set of launches (linear.C:234-258): every root's
    sigma is bitwise the one-vector sigma, for a batch of four, a batch of
    three over four vectors (3 + 1) and in split precision."""
    print(f"[gpu] batched roots {name}", flush=True)
    sb = make()
    h = block_amd.multiply_h(sb, max_vec=4)
    c = np.stack([rng_vec(sb.wave_size, 20 + i) for i in range(4)])
    for prec in (block_amd.FP64, block_amd.SPLIT_BF16):
        h.precision = prec
        h.set_batch(1)
        one = np.stack([h(c[i]) for i in range(4)])
        for nb in (4, 3):
            h.set_batch(nb)
            got = h(c)
            assert np.array_equal(got, one), (prec, nb)
    h.precision = block_amd.FP64
    ref = _oracle.sigma(sb, c[2], nthreads=8)
    assert np.linalg.norm(got[2] - one[2]) == 0.0
    h.set_batch(4)
    assert np.linalg.norm(h(c)[2] - ref) <= TOL * np.linalg.norm(ref)


@pytest.mark.parametrize("name,make", [c for c in _cases() if c[0] in ("exact7d2h", "syn16_M80", "syn24_M500_dmrg")],
                         ids=["exact7d2h", "syn16_M80", "syn24_M500_dmrg"])
def test_graph_replays_match_eager(name, make):
    """Applications through a captured HIP graph: every replay, with eager
    applications and host round trips between them, FP64 and split
    precision, is bitwise the eagerly launched sigma."""
    print(f"[gpu] graph replays {name}", flush=True)
    sb = make()
    h = block_amd.multiply_h(sb)
    c = rng_vec(sb.wave_size, 31)
    x = rng_vec(sb.wave_size, 32)
    for prec in (block_amd.FP64, block_amd.SPLIT_BF16):
        h.precision = prec
        h.set_graph(False)
        want, want_x = h(c), h(x)
        h.set_graph(True)
        for rep in range(4):
            assert np.array_equal(h(c), want), (prec, rep)
            h.set_graph(False)
            assert np.array_equal(h(x), want_x)  # eager work between replays
            h.set_graph(True)
            assert np.array_equal(h(x), want_x)  # a replay after a new capture
    h.precision = block_amd.FP64


@pytest.mark.timeout(480)
def test_syn30_four_root_superblock():
    """The 30-orbital / 30-electron C1 synthetic Hamiltonian at M = 800 with
    four state-averaged roots (the batched configuration): the mid-chain
    superblock of its DMRG warm-up, sigma against the oracle at full size
    (FP64 and split), and the four roots batched in one launch set bitwise
    equal to four single applications."""
    print("[gpu] syn30 M=800 four-root superblock", flush=True)
    sb = block_amd.SuperBlock.from_dmrg_warmup(config_fcidump("syn30"), 30, 0, 0, 800, nroots=4)
    print(f"[gpu] syn30 wave {sb.wave_size}, {sb.flops:.3e} flop", flush=True)
    assert sb.wave_size > 100_000
    h = block_amd.multiply_h(sb, max_vec=4)
    c = np.stack([rng_vec(sb.wave_size, 40 + i) for i in range(4)])
    ref = _oracle.sigma(sb, c[0], nthreads=16)
    h.set_batch(4)
    got = h(c)
    rel = np.linalg.norm(got[0] - ref) / np.linalg.norm(ref)
    print(f"[gpu] syn30 FP64 rel={rel:.3e}", flush=True)
    assert rel <= TOL
    h.set_batch(1)
    for i in range(4):
        assert np.array_equal(h(c[i]), got[i])
    h.precision = block_amd.SPLIT_BF16
    h.set_batch(4)
    gs = h(c)
    rel_s = np.linalg.norm(gs[0] - ref) / np.linalg.norm(ref)
    print(f"[gpu] syn30 split rel={rel_s:.3e}", flush=True)
    assert 0.0 < rel_s <= SPLIT_TOL


@pytest.mark.parametrize("name,make", [c for c in _cases() if c[0] in ("exact6", "syn24_M500_dmrg")],
                         ids=["exact6", "syn24_M500_dmrg"])
def test_split_share(name, make):
    """Share of the flops split mode runs on BF16 slices: a small one when
    every V block is at most 32 x 32 (their pairs stay FP64 except those
    with an identity A and an operator B, which run in the grouped kernel),
    most of them when the V blocks are hundreds wide."""
    print(f"[gpu] split share {name}", flush=True)
    sb = make()
    share = block_amd.multiply_h(sb).split_share
    big = max(max(m, n) for m, n in sb.block_shapes())
    print(f"[gpu] {name} largest V block side {big}, split share {share:.3f}", flush=True)
    assert 0.0 <= share <= 1.0
    if big <= 32:
        assert share < 0.1
    else:
        assert share > 0.5


@pytest.mark.timeout(480)
def test_syn50_superblock():
    """The 50-orbital / 50-electron C1 synthetic Hamiltonian (the large
    configuration's system) at M = 300: the mid-chain superblock of its
    infinite-system warm-up, sigma against the oracle at full size, FP64 and
    split."""
    print("[gpu] syn50 M=300 superblock", flush=True)
    sb = block_amd.SuperBlock.from_dmrg_warmup(config_fcidump("syn50"), 50, 0, 0, 300)
    print(f"[gpu] syn50 wave {sb.wave_size}, {sb.flops:.3e} flop", flush=True)
    assert sb.wave_size > 50_000
    h = block_amd.multiply_h(sb)
    c = rng_vec(sb.wave_size, 50)
    ref = _oracle.sigma(sb, c, nthreads=16)
    got = h(c)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"[gpu] syn50 FP64 rel={rel:.3e}", flush=True)
    assert rel <= TOL
    h.precision = block_amd.SPLIT_BF16
    gs = h(c)
    rel_s = np.linalg.norm(gs - ref) / np.linalg.norm(ref)
    print(f"[gpu] syn50 split rel={rel_s:.3e}", flush=True)
    assert 0.0 < rel_s <= SPLIT_TOL


def _wave_cases():
    # superblocks with V blocks above 32 rows or columns (S = 0 and S = 1)
    return [c for c in _cases() if c[0] in ("hubbard12_M100", "syn24_M500_dmrg")] + [
        ("hubbard12_M100_S1", lambda: block_amd.SuperBlock(config_fcidump("hubbard12"), 12, 2, 0, 5, 100)),
        ("syn16_M200_S1", lambda: block_amd.SuperBlock(random_fcidump(16, 21), 16, 2, 0, 7, 200)),
    ]


@pytest.mark.parametrize("name,make", _wave_cases(), ids=[c[0] for c in _wave_cases()])
def test_wave_tiles(name, make):
    """The per-wave tile kernel (the FP64 tile-table candidate after the
    geometries) on every V block above 32 rows or columns: sigma against the
    oracle at 1e-12, a different summation order from the grouped kernel
    (so the per-wave kernel ran), bitwise reproducible, four roots batched
    bitwise equal to single applications, every graph replay equal to the
    eager sigma, and split precision (the grouped BF16 kernel keeps these
    blocks, their pairs in the per-wave order) within its tolerance."""
    print(f"[gpu] wave tiles {name}", flush=True)
    sb = make()
    big = max(max(m, n) for m, n in sb.block_shapes())
    assert big > 32
    c = np.stack([rng_vec(sb.wave_size, 50 + i) for i in range(4)])
    ref = _oracle.sigma(sb, c[0], nthreads=8)
    h = block_amd.multiply_h(sb, max_vec=4)
    wave = block_amd.tile_candidates(block_amd.FP64) - 1
    try:
        block_amd.set_tile_table([])  # the grouped kernel for every block above 32 x 32
        h.replan()
        grouped = h(c[0])
        h.precision = block_amd.SPLIT_BF16
        split = h(c[0])
        h.precision = block_amd.FP64
        block_amd.set_tile_table([(1, 1 << 20, 1, 1 << 20, wave, -1)])
        h.replan()
        got = h(c[0])
        rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"[gpu] {name} largest V block side {big}, wave-tile rel={rel:.3e}", flush=True)
        assert rel <= TOL
        assert not np.array_equal(got, grouped)
        assert np.array_equal(h(c[0]), got)
        h.set_batch(4)
        batch = h(c)
        h.set_batch(1)
        for i in range(4):
            assert np.array_equal(batch[i], h(c[i])), i
        h.set_graph(True)
        for rep in range(3):
            assert np.array_equal(h(c[0]), got), rep
            h.set_graph(False)
            assert np.array_equal(h(c[1]), batch[1])
            h.set_graph(True)
        h.set_graph(False)
        h.precision = block_amd.SPLIT_BF16
        got_s = h(c[0])
        rel_s = np.linalg.norm(got_s - ref) / np.linalg.norm(ref)
        assert 0.0 < rel_s <= SPLIT_TOL
        assert np.linalg.norm(got_s - split) <= SPLIT_TOL * np.linalg.norm(ref)
        h.precision = block_amd.FP64
    finally:
        block_amd.set_tile_table(block_amd.read_tile_table() if os.path.exists(block_amd.TILE_TABLE) else [])
