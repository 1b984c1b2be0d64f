"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- test
infrastructure: tests use it only as the reference their results are
checked against."""
import ctypes
import os

import numpy as np

import block_amd

_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_path = os.path.join(_ROOT, "oracle", "liboracle.so")
if not os.path.exists(_path):
    raise ImportError(f"{_path} not built; run __graft_entry__.build()")
lib = ctypes.CDLL(_path)
_vp, _dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
for name, res, args in [
    ("orc_last_error", ctypes.c_char_p, []),
    ("orc_sigma", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int]),
    ("orc_sigma_timed", ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int, _dp]),
    ("orc_diag", ctypes.c_int, [_vp, _dp]),
    ("orc_sigma_flops", ctypes.c_double, [_vp]),
    ("orc_exec_worklist", ctypes.c_int, [_vp, _dp, _dp]),
    ("orc_fci", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp, _dp]),
    ("orc_exact_superblock", _vp, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    ("orc_davidson", ctypes.c_int, [_vp, ctypes.c_int, _dp, ctypes.c_int, _dp, ctypes.c_double, ctypes.c_int, _dp,
                                    _dp, ctypes.POINTER(ctypes.c_int)]),
    ("orc_dmrg_cpu", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                    _dp, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_char_p,
                                    ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
]:
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args


def _p(a):
    return a.ctypes.data_as(_dp)


def _chk(rc):
    if rc < 0:
        raise RuntimeError(lib.orc_last_error().decode())
    return rc


def sigma(sb, c, nthreads=1):
    c = np.ascontiguousarray(c, dtype=np.float64)
    v = np.zeros(sb.wave_size)
    _chk(lib.orc_sigma(sb._h, _p(c), _p(v), nthreads))
    return v


def sigma_timed(sb, c, nthreads):
    c = np.ascontiguousarray(c, dtype=np.float64)
    v = np.zeros(sb.wave_size)
    t = ctypes.c_double()
    _chk(lib.orc_sigma_timed(sb._h, _p(c), _p(v), nthreads, ctypes.byref(t)))
    return v, t.value


def sigma_flops(sb):
    return lib.orc_sigma_flops(sb._h)


def diagonal(sb):
    d = np.zeros(sb.wave_size)
    _chk(lib.orc_diag(sb._h, _p(d)))
    return d


def exec_worklist(sb, c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    v = np.zeros(sb.wave_size)
    _chk(lib.orc_exec_worklist(sb._h, _p(c), _p(v)))
    return v


def fci(ints, nelec, nroots=1, penalty=3.0, irrep=-1):
    e = np.zeros(nroots)
    s2 = np.zeros(nroots)
    n = _chk(lib.orc_fci(ints._h, nelec, nroots, penalty, irrep, _p(e), _p(s2)))
    return e[:n], s2[:n]


def exact_superblock(ints, nelec, twoS, irrep, nL):
    h = lib.orc_exact_superblock(ints._h, nelec, twoS, irrep, nL)
    if not h:
        raise RuntimeError(lib.orc_last_error().decode())
    return block_amd.SuperBlock(ints, nelec, twoS, irrep, 0, 0, _handle=h)


def dense_h(sb, nthreads=4):
    n = sb.wave_size
    H = np.zeros((n, n))
    e = np.zeros(n)
    for i in range(n):
        e[:] = 0.0
        e[i] = 1.0
        H[:, i] = sigma(sb, e, nthreads)
    return H


def davidson(sb, guesses, tol=1e-12, lower=None, nthreads=4):
    """block_davidson with the oracle sigma from the given guesses (rows);
    returns (energies, vectors, (iterations, sigma calls, split calls))."""
    g = np.ascontiguousarray(np.atleast_2d(guesses), dtype=np.float64)
    nroots = g.shape[0]
    low = None if lower is None or len(lower) == 0 else np.ascontiguousarray(np.atleast_2d(lower), dtype=np.float64)
    e = np.zeros(nroots)
    vecs = np.zeros_like(g)
    st = (ctypes.c_int * 3)()
    _chk(lib.orc_davidson(sb._h, nroots, _p(g), 0 if low is None else low.shape[0], None if low is None else _p(low),
                          tol, nthreads, _p(e), _p(vecs), st))
    return e, vecs, (st[0], st[1], st[2])


def dmrg_cpu(ints, nelec=None, twoS=0, irrep=None, nroots=1, weights=None, schedule=None, max_sweeps=20,
             sweep_tol=1e-8, nthreads=4, verbose=0, guess=True, checkpoint=None, resume=None, ahead=-1):
    nelec = ints.nelec if nelec is None else nelec
    irrep = ints.isym if irrep is None else irrep
    w = np.ascontiguousarray(weights if weights is not None else np.full(nroots, 1.0 / nroots), dtype=np.float64)
    sched = block_amd._schedule_array(schedule)
    e = np.zeros(nroots)
    nc = len(block_amd.SWEEP_COLS)
    sw = np.zeros(nc * max_sweeps)
    n = _chk(lib.orc_dmrg_cpu(ints._h, nelec, twoS, irrep, nroots, _p(w), len(schedule), _p(sched), max_sweeps,
                              sweep_tol, int(guess), None if checkpoint is None else str(checkpoint).encode(),
                              None if resume is None else str(resume).encode(), int(ahead), nthreads, verbose, _p(e),
                              _p(sw)))
    return e, sw[: nc * n].reshape(n, nc)
