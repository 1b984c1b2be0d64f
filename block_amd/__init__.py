"""MI355X-native Davidson sigma step (H.C) of spin-adapted two-site DMRG.

The compute path is native: ``libblock_sigma.so`` (hand-written HIP kernels
for gfx950 behind the C-ABI of ``include/block_sigma.h``) driven by
``libblock_host.so`` (the C++ restatement of Block's superblock, operator
algebra, Davidson solver and sweep).  This module only loads them with ctypes;
there is no Python or CPU fallback: if the libraries are missing the import
fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def _load(name):
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run __graft_entry__.build() (make -C block_amd)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)


sigma_lib = _load("libblock_sigma.so")
host_lib = _load("libblock_host.so")

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_h = host_lib
_sig(_h, "blkh_last_error", ctypes.c_char_p, [])
_sig(_h, "blkh_load_fcidump", _vp, [ctypes.c_char_p])
_sig(_h, "blkh_free_integrals", None, [_vp])
for _n in ("blkh_norb", "blkh_nelec", "blkh_isym", "blkh_ms2"):
    _sig(_h, _n, ctypes.c_int, [_vp])
_sig(_h, "blkh_ecore", ctypes.c_double, [_vp])
_sig(_h, "blkh_snapshot", _vp, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int])
_sig(_h, "blkh_snapshot_dmrg", _vp, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int])
_sig(_h, "blkh_free_superblock", None, [_vp])
_sig(_h, "blkh_superblock_from_state", _vp, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int])
_sig(_h, "blkh_wave_size", ctypes.c_long, [_vp])
_sig(_h, "blkh_flops", ctypes.c_double, [_vp])
_sig(_h, "blkh_nterms", ctypes.c_int, [_vp])
_sig(_h, "blkh_nsectors", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_int)])
_sig(_h, "blkh_operator_bytes", ctypes.c_long, [_vp])
_sig(_h, "blkh_wave_blocks", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int])
_sig(_h, "blkh_diag_cpu", ctypes.c_int, [_vp, _dp])
_sig(_h, "blkh_gpu_sigma_create", _vp, [_vp, ctypes.c_int, ctypes.c_long])
_sig(_h, "blkh_gpu_sigma_free", None, [_vp])
_sig(_h, "blkh_gpu_sigma_apply", ctypes.c_int, [_vp, _dp, _dp])
_sig(_h, "blkh_gpu_sigma_diag", ctypes.c_int, [_vp, _dp])
_sig(_h, "blkh_gpu_sigma_flops", ctypes.c_double, [_vp])
_sig(_h, "blkh_gpu_sigma_split_share", ctypes.c_double, [_vp])
_sig(_h, "blkh_gpu_sigma_pairs", ctypes.c_long, [_vp])
_sig(_h, "blkh_gpu_sigma_phase_flops", ctypes.c_int, [_vp, _dp, _dp])
_sig(_h, "blkh_gpu_sigma_arena_bytes", ctypes.c_long, [_vp])
_sig(_h, "blkh_gpu_sigma_timed", ctypes.c_int, [_vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _lp])
_sig(_h, "blkh_gpu_sigma_set_batch", ctypes.c_int, [_vp, ctypes.c_int])
_sig(_h, "blkh_gpu_sigma_set_graph", ctypes.c_int, [_vp, ctypes.c_int])
_sig(_h, "blkh_gpu_sigma_replan", ctypes.c_int, [_vp])
_sig(sigma_lib, "bsg_set_tile_table", ctypes.c_int, [ctypes.POINTER(ctypes.c_int32), ctypes.c_int32])
_sig(sigma_lib, "bsg_tile_candidates", ctypes.c_int, [ctypes.c_int32])
_sig(sigma_lib, "bsg_last_error", ctypes.c_char_p, [])

_sig(_h, "blkh_gpu_sigma_apply_n", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int])
_sig(_h, "blkh_gpu_sigma_set_precision", ctypes.c_int, [_vp, ctypes.c_int])
_sig(_h, "blkh_gpu_sigma_raw", ctypes.c_int, [_vp] + [ctypes.POINTER(_vp)] * 5)
_sig(_h, "blkh_dmrg_gpu", ctypes.c_int,
     [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_int,
      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int,
      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp])
_sig(_h, "blkh_build_check", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _lp])
_sig(_h, "blkh_build_check_cpu", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _lp])
_sig(_h, "blkh_build_times", ctypes.c_int, [_dp])
_sig(_h, "blkh_state_block_dims", ctypes.c_int, [_vp, ctypes.c_char_p] + [ctypes.POINTER(ctypes.c_int)] * 4)

_sig(_h, "blkh_davidson", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp,
                                        ctypes.c_double, ctypes.c_double, _dp, _dp, ctypes.POINTER(ctypes.c_int)])
_sig(_h, "blkh_rotation_check", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _lp])
_ip = ctypes.POINTER(ctypes.c_int)
_sig(_h, "blkh_fiedler_order", ctypes.c_int, [_vp, _ip])
_sig(_h, "blkh_fiedler_order_matrix", ctypes.c_int, [ctypes.c_int, _dp, _ip])
_sig(_h, "blkh_reorder_integrals", _vp, [_vp, _ip])

FP64, SPLIT_BF16 = 0, 1
SWEEP_COLS = ("sweep", "forward", "M", "E0", "discarded", "seconds", "sigma_seconds", "sigma_flops",
              "sigma_calls", "split_calls", "build_seconds", "davidson_seconds", "renorm_seconds", "guessed_steps")


class BlockError(RuntimeError):
    pass


def _err():
    return BlockError(host_lib.blkh_last_error().decode())


TILE_TABLE = os.path.join(_HERE, "tile_table.txt")


def set_tile_table(rules):
    """Measured tile choices for the plans created from now on: rules of
    (m_lo, m_hi, n_lo, n_hi, fp64 geometry, split geometry), geometry numbers
    into the kernels' candidate lists, -1 for the model's choice
    (bsg_set_tile_table)."""
    flat = [int(x) for r in rules for x in r]
    arr = (ctypes.c_int32 * max(1, len(flat)))(*flat)
    if sigma_lib.bsg_set_tile_table(arr, len(rules)) != 0:
        raise BlockError(sigma_lib.bsg_last_error().decode())


def read_tile_table(path=TILE_TABLE):
    """Rules of a tile table file: one rule of six integers per line, '#'
    comments."""
    rules = []
    for line in open(path):
        line = line.split("#", 1)[0].split()
        if line:
            rules.append(tuple(int(x) for x in line))
    return rules


def tile_candidates(precision):
    return sigma_lib.bsg_tile_candidates(int(precision))


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Integrals:
    """One- and two-electron integrals read from an FCIDUMP (input.C:1407-1560)."""

    def __init__(self, path):
        self._h = host_lib.blkh_load_fcidump(str(path).encode())
        if not self._h:
            raise _err()
        self.path = str(path)

    norb = property(lambda s: host_lib.blkh_norb(s._h))
    nelec = property(lambda s: host_lib.blkh_nelec(s._h))
    isym = property(lambda s: host_lib.blkh_isym(s._h))
    ms2 = property(lambda s: host_lib.blkh_ms2(s._h))
    ecore = property(lambda s: host_lib.blkh_ecore(s._h))

    def fiedler_order(self):
        """Orbital order from the Fiedler vector of the exchange-integral graph
        (fiedler.C:47-81): order[a] is the orbital placed at position a."""
        out = (ctypes.c_int * self.norb)()
        if host_lib.blkh_fiedler_order(self._h, out) != 0:
            raise _err()
        return list(out)

    def reordered(self, order):
        """The integrals with orbital order[a] moved to position a."""
        arr = (ctypes.c_int * self.norb)(*order)
        h = host_lib.blkh_reorder_integrals(self._h, arr)
        if not h:
            raise _err()
        obj = object.__new__(Integrals)
        obj._h = h
        obj.path = self.path
        return obj

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib.blkh_free_integrals(self._h)
            self._h = None


class SuperBlock:
    """Mid-chain two-site superblock L (x) R (sweep.C:91-178): sys = orbitals
    [0, nsys), dots nsys and nsys+1, environment the rest, blocks grown with a
    Wilson-type truncation to M multiplets."""

    def __init__(self, ints, nelec, twoS, irrep, nsys, M, _handle=None):
        self.ints = ints
        self._h = _handle if _handle is not None else host_lib.blkh_snapshot(ints._h, nelec, twoS, irrep, nsys, M)
        if not self._h:
            raise _err()

    wave_size = property(lambda s: host_lib.blkh_wave_size(s._h))
    flops = property(lambda s: host_lib.blkh_flops(s._h))
    nterms = property(lambda s: host_lib.blkh_nterms(s._h))
    operator_bytes = property(lambda s: host_lib.blkh_operator_bytes(s._h))

    @classmethod
    def from_dmrg_warmup(cls, ints, nelec, twoS, irrep, M, nroots=1, device=0, verbose=0):
        """The mid-chain superblock reached by the infinite-system DMRG warm-up:
        both halves truncated to M by density matrices of the partial-chain
        ground states (Davidson sigma on the GPU)."""
        h = host_lib.blkh_snapshot_dmrg(ints._h, nelec, twoS, irrep, nroots, M, device, verbose)
        if not h:
            raise _err()
        return cls(ints, nelec, twoS, irrep, 0, M, _handle=h)

    @classmethod
    def from_sweep_state(cls, ints, path, nelec=None, twoS=0, irrep=None, step=-1):
        """The two-site superblock of forward step `step` (sys = orbitals
        0..step, dots step+1 and step+2, environment the rest; -1: the middle
        of the chain) rebuilt from the blocks of a saved sweep state (a
        checkpoint file of dmrg)."""
        nelec = ints.nelec if nelec is None else nelec
        irrep = ints.isym if irrep is None else irrep
        h = host_lib.blkh_superblock_from_state(ints._h, str(path).encode(), nelec, twoS, irrep, step)
        if not h:
            raise _err()
        return cls(ints, nelec, twoS, irrep, 0, 0, _handle=h)

    def sectors(self):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        host_lib.blkh_nsectors(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def block_shapes(self):
        """(rows, cols) of every wavefunction (V) block, in storage order."""
        n = self.sectors()[2]
        r, c = (ctypes.c_int * max(1, n))(), (ctypes.c_int * max(1, n))()
        host_lib.blkh_wave_blocks(self._h, r, c, n)
        return [(r[i], c[i]) for i in range(n)]

    def diagonal_cpu(self):
        out = np.zeros(self.wave_size)
        if host_lib.blkh_diag_cpu(self._h, _ptr(out)) != 0:
            raise _err()
        return out

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib.blkh_free_superblock(self._h)
            self._h = None


class multiply_h:
    """GPU sigma functor for one superblock: ``v = H c`` (davidson.h:22-29).

    Construction flattens the superblock into the GEMM-pair work list and makes
    every operator block resident in HBM; calls run the fused MFMA kernels in
    FP64 or, with ``precision=SPLIT_BF16``, as split BF16 slice products."""

    def __init__(self, sb, device=0, precision=FP64, max_vec=1):
        self.sb = sb
        self._h = host_lib.blkh_gpu_sigma_create(sb._h, device, int(max_vec))
        if not self._h:
            raise _err()
        self.n = sb.wave_size
        self.max_vec = int(max_vec)
        self.precision = precision

    @property
    def precision(self):
        return self._precision

    @precision.setter
    def precision(self, mode):
        if host_lib.blkh_gpu_sigma_set_precision(self._h, int(mode)) != 0:
            raise _err()
        self._precision = int(mode)

    flops = property(lambda s: host_lib.blkh_gpu_sigma_flops(s._h))
    # share of the flops split mode runs on BF16 slices (the rest is FP64)
    split_share = property(lambda s: host_lib.blkh_gpu_sigma_split_share(s._h))
    pairs = property(lambda s: host_lib.blkh_gpu_sigma_pairs(s._h))
    arena_bytes = property(lambda s: host_lib.blkh_gpu_sigma_arena_bytes(s._h))

    def phase_flops(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        host_lib.blkh_gpu_sigma_phase_flops(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def __call__(self, c):
        """H c for one vector (shape (n,)) or for several (shape (k, n), k <=
        max_vec: the roots run together in launches of the plan's batch)."""
        c = np.ascontiguousarray(c, dtype=np.float64)
        if c.shape == (self.n,):
            v = np.zeros(self.n)
            if host_lib.blkh_gpu_sigma_apply(self._h, _ptr(c), _ptr(v)) != 0:
                raise _err()
            return v
        if c.ndim != 2 or c.shape[1] != self.n or not 1 <= c.shape[0] <= self.max_vec:
            raise ValueError(f"wavefunctions must have shape ({self.n},) or (k <= {self.max_vec}, {self.n})")
        v = np.zeros_like(c)
        if host_lib.blkh_gpu_sigma_apply_n(self._h, _ptr(c), _ptr(v), c.shape[0]) != 0:
            raise _err()
        return v

    def set_batch(self, nroot):
        """Roots per launch (bsg_plan_set_batch)."""
        if host_lib.blkh_gpu_sigma_set_batch(self._h, int(nroot)) != 0:
            raise _err()

    def replan(self):
        """Build the plan again from the work list (tile choices set since
        with set_tile_table apply)."""
        if host_lib.blkh_gpu_sigma_replan(self._h) != 0:
            raise _err()

    def set_graph(self, on):
        """Applications as one captured HIP graph launch each (bsg_plan_capture
        / bsg_plan_replay) instead of their kernel launches."""
        if host_lib.blkh_gpu_sigma_set_graph(self._h, int(bool(on))) != 0:
            raise _err()

    def raw(self):
        """The C-ABI handles (bsg_context*, bsg_plan*, arena, device c, device v
        of wave_size doubles each) for calling libblock_sigma.so directly."""
        out = [_vp() for _ in range(5)]
        host_lib.blkh_gpu_sigma_raw(self._h, *[ctypes.byref(x) for x in out])
        return tuple(x.value for x in out)

    def diagonal(self):
        out = np.zeros(self.n)
        if host_lib.blkh_gpu_sigma_diag(self._h, _ptr(out)) != 0:
            raise _err()
        return out

    def timed(self, c, warmup, reps):
        """Device-resident repeated sigma in the current precision on one
        vector (n,) or several (k, n): returns (wall_ms, [fused, reduce, side]
        kernel ms, kernel dispatches) over the `reps` timed applications,
        measured with HIP events on the plan stream."""
        c = np.ascontiguousarray(c, dtype=np.float64)
        nvec = 1 if c.ndim == 1 else c.shape[0]
        wall = ctypes.c_double()
        km = np.zeros(3)
        la = (ctypes.c_long * 3)()
        if host_lib.blkh_gpu_sigma_timed(self._h, _ptr(c), nvec, warmup, reps, ctypes.byref(wall), _ptr(km), la) != 0:
            raise _err()
        return wall.value, km.tolist(), la[0]

    def __del__(self):
        if getattr(self, "_h", None):
            host_lib.blkh_gpu_sigma_free(self._h)
            self._h = None


def _schedule_array(schedule):
    rows = [(float(s), float(m), float(t), float(n)) for (s, m, t, n) in schedule]
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1))


def dmrg(ints, nelec=None, twoS=0, irrep=None, nroots=1, weights=None, schedule=None, max_sweeps=20,
         sweep_tol=1e-8, device=0, verbose=0, ladder=0.0, guess=True, ladder_from=0, checkpoint=None, resume=None,
         host_ahead=True, device_blocks=True, ladder_min_share=0.5):
    """Two-site spin-adapted DMRG (Sweep::do_one, sweep.C:341) with the Davidson
    sigma on the GPU.  schedule: rows (start_sweep, M, davidson_tol, noise).
    device_blocks: build the composite operators in HBM from device-resident
    renormalised blocks (no operator crosses PCIe); False: host builds and
    uploads (then host_ahead applies).
    ladder > 0: split-precision sigma until the Davidson residual norm falls
    below it, FP64 after (convergence is always decided on FP64 sigmas), from
    sweep ladder_from on (0: also in the warm-up), on superblocks whose flops
    run at least ladder_min_share on BF16 slices (multiply_h.split_share; the
    rest of a superblock stays FP64 in split mode).
    guess: start every step from the previous step's wavefunction carried into
    the new basis (guess_wavefunction.C:524-636), else from the diagonal.
    checkpoint: file the sweep state is written to after every sweep; resume:
    such a file to continue from (blocks, last wavefunction, sweep records)
    instead of the warm-up.
    host_ahead: build the host-side operators that do not depend on the
    current solve (the new system block's complementary operators, the next
    environment block) on a host thread while the GPU runs the Davidson solve.
    Returns (energies, sweeps) with sweeps rows named by SWEEP_COLS."""
    nelec = ints.nelec if nelec is None else nelec
    irrep = ints.isym if irrep is None else irrep
    schedule = schedule or [(0, 50, 1e-6, 1e-5), (2, 100, 1e-8, 1e-6), (6, 100, 1e-12, 0.0)]
    w = np.ascontiguousarray(weights if weights is not None else np.full(nroots, 1.0 / nroots), dtype=np.float64)
    sched = _schedule_array(schedule)
    e = np.zeros(nroots)
    nc = len(SWEEP_COLS)
    sw = np.zeros(nc * max_sweeps)
    n = host_lib.blkh_dmrg_gpu(ints._h, nelec, twoS, irrep, nroots, _ptr(w), len(schedule), _ptr(sched), max_sweeps,
                               sweep_tol, float(ladder), int(ladder_from), int(guess),
                               None if checkpoint is None else str(checkpoint).encode(),
                               None if resume is None else str(resume).encode(), int(bool(host_ahead)),
                               int(bool(device_blocks)), device, verbose, float(ladder_min_share), _ptr(e),
                               _ptr(sw))
    if n < 0:
        raise _err()
    return e, sw[: nc * n].reshape(n, nc)


def state_block_dims(ints, path):
    """Kept states and sectors of every block of a saved sweep state:
    (fwd_states, bwd_states, fwd_sectors, bwd_sectors), lists over the
    orbitals (fwd[l]: orbitals 0..l, bwd[l]: orbitals l..norb-1; 0 where
    the state holds no block)."""
    k = ints.norb
    arrs = [(ctypes.c_int * k)() for _ in range(4)]
    if host_lib.blkh_state_block_dims(ints._h, str(path).encode(), *arrs) != 0:
        raise _err()
    return tuple(list(a) for a in arrs)


def build_times():
    """Accumulated recorded-build phases of this process: dict of seconds
    (record, execute, grouping, device), algorithmic bytes and items."""
    out = np.zeros(6)
    host_lib.blkh_build_times(_ptr(out))
    return dict(zip(("record", "execute", "grouping", "device", "bytes", "items"), out.tolist()))


def build_check_cpu(ints, nsys, M):
    """The recorded composite-operator build (the items and groups the device
    executes) run on the host against the host build: (largest absolute
    element difference, operators compared)."""
    worst = ctypes.c_double()
    n = ctypes.c_long()
    if host_lib.blkh_build_check_cpu(ints._h, nsys, M, ctypes.byref(worst), ctypes.byref(n)) != 0:
        raise _err()
    return worst.value, n.value


def build_check(ints, nsys, M, device=0):
    """Composite operators built in HBM against the host build (a Wilson-grown
    block of nsys sites and the next dot, normal and complementary operator
    sets, from a host block and from a device-only one); returns (largest
    absolute element difference, operators compared)."""
    worst = ctypes.c_double()
    n = ctypes.c_long()
    if host_lib.blkh_build_check(ints._h, nsys, M, device, ctypes.byref(worst), ctypes.byref(n)) != 0:
        raise _err()
    return worst.value, n.value


def davidson(sb, guesses, tol=1e-12, lower=None, ladder=0.0, device=0, host_loop=False):
    """One Davidson solve on superblock sb with the GPU sigma
    (Linear::block_davidson, linear.C:179-385), from the given guesses (one
    row per root).  lower: states to project off (the reference's
    lowerStates).  host_loop: run the host control flow with the GPU sigma
    instead of the device-resident solve.  Returns (energies, vectors,
    (iterations, sigma calls, split-precision sigma calls), device_solve)
    where device_solve says whether the device-resident solve ran."""
    g = np.ascontiguousarray(np.atleast_2d(guesses), dtype=np.float64)
    nroots = g.shape[0]
    low = None if lower is None or len(lower) == 0 else np.ascontiguousarray(np.atleast_2d(lower), dtype=np.float64)
    e = np.zeros(nroots)
    vecs = np.zeros_like(g)
    st = (ctypes.c_int * 3)()
    rc = host_lib.blkh_davidson(sb._h, device, int(bool(host_loop)), nroots, _ptr(g),
                                0 if low is None else low.shape[0], None if low is None else _ptr(low), tol,
                                float(ladder), _ptr(e), _ptr(vecs), st)
    if rc < 0:
        raise _err()
    return e, vecs, (st[0], st[1], st[2]), (rc == 0 and not host_loop)


def rotation_check(ints, nsys, M, device=0):
    """Renormalise a Wilson-grown block of nsys sites plus one dot with a random
    density matrix truncated to M, with the host rotations and with the GPU
    rotator (transform_operators, save_load_block.C:267-319); returns (largest
    relative difference over all rotated operators, operators compared)."""
    worst = ctypes.c_double()
    n = ctypes.c_long()
    if host_lib.blkh_rotation_check(ints._h, nsys, M, device, ctypes.byref(worst), ctypes.byref(n)) != 0:
        raise _err()
    return worst.value, n.value


def fiedler_order_matrix(m):
    """fiedler_reorder (fiedler.C:47-81) on a symmetric interaction matrix."""
    m = np.ascontiguousarray(m, dtype=np.float64)
    n = m.shape[0]
    out = (ctypes.c_int * n)()
    if host_lib.blkh_fiedler_order_matrix(n, _ptr(m), out) != 0:
        raise _err()
    return list(out)


# the committed tile table (bench.py --tune) applies to every plan this
# process creates
if os.path.exists(TILE_TABLE):
    set_tile_table(read_tile_table())
