"""Block's scalar factors of a TensorMultiply sector triple, restated from the
reference for Abelian point groups (spatial Clebsch-Gordan and 9j factors 1,
Symmetry.C:534-538), and the packer that turns them into the alpha_a and
alpha of a bsg_pair_t -- the adapter arithmetic a Block maintainer writes
(INTEGRATION.md's pack_superblock):

* Transposeview::get_scaling (BaseOperator.C:56-91): the factor of a
  transposed operator view between sectors (l, r) of spins lspin, rspin for
  an operator of spin cspin (all doubled);
* getCommuteParity (BaseOperator.C:20-53): the factor of swapping two
  coupled operators;
* the normalised Ninej (couplingCoeffs.h:88-95) with TensorMultiply's
  argument order and fermion parity (operatorfunctions.C:520-530).
Test infrastructure: used by tests/test_block_conventions.py only."""
import ctypes

import block_amd

_h = block_amd.host_lib
for _n in ("blkh_cg",):
    f = getattr(_h, _n)
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_int] * 6
_h.blkh_ninej.restype = ctypes.c_double
_h.blkh_ninej.argtypes = [ctypes.c_int] * 9
NUMERICAL_ZERO = 1e-15


def clebsch(j1, m1, j2, m2, J, M):
    return _h.blkh_cg(j1, m1, j2, m2, J, M)


def nine_j(*a):
    return _h.blkh_ninej(*a)


def Ninej(ja, jb, jc, jd, je, jf, jg, jh, ji):
    """couplingCoeffs.h:88-95"""
    return ((jg + 1) * (jh + 1) * (jc + 1) * (jf + 1)) ** 0.5 * nine_j(ja, jb, jc, jd, je, jf, jg, jh, ji)


def get_scaling(lspin, rspin, cspin):
    """Transposeview::get_scaling for a transposed view (conjugacy 't'),
    BaseOperator.C:56-91: the first non-vanishing Clebsch-Gordan pair."""
    for lsz in range(-lspin, lspin + 1, 2):
        for rsz in range(-rspin, rspin + 1, 2):
            cleb = clebsch(lspin, lsz, cspin, -cspin, rspin, rsz)
            if abs(cleb) <= NUMERICAL_ZERO:
                continue
            return (-1.0) ** cspin * cleb / clebsch(rspin, rsz, cspin, cspin, lspin, lsz)
    raise ValueError("get_scaling: inappropriate arguments")


def get_commute_parity(aspin, bspin, cspin, a_fermion, b_fermion):
    """getCommuteParity, BaseOperator.C:20-53 (spin part)."""
    parity = -1.0 if (a_fermion and b_fermion) else 1.0
    for asz in range(-aspin, aspin + 1, 2):
        for bsz in range(-bspin, bspin + 1, 2):
            cleb = clebsch(aspin, asz, bspin, bsz, cspin, cspin)
            if abs(cleb) <= NUMERICAL_ZERO:
                continue
            return parity * cleb / clebsch(bspin, bsz, aspin, asz, cspin, cspin)
    raise ValueError("getCommuteParity: inappropriate arguments")


def adjoint_convention(k2):
    """This library's adjoint tensor (T^dag)_q = (-1)^(k-q) (T_-q)^dag differs
    from the one Block's get_scaling encodes by (-1)^(2k): the factor a
    packer applies to every transposed view of spin k (k2 = 2k; -1 for the
    half-integer-spin, fermionic operators)."""
    return -1.0 if k2 & 1 else 1.0


def pack_pair(coef, ka, kb, a_dag, b_dag, b_fermion, a_ident, b_ident, lq, lqp, rq, rqp, St):
    """alpha_a and alpha of one GEMM pair from Block's factors
    (operatorfunctions.C:508-530): W = get_scaling(A) op(A) C, then
    V += scale Ninej parity get_scaling(B) W op(B)^T, with every transposed
    view's get_scaling carried into this library's adjoint convention
    (adjoint_convention).  lq... are (N, 2S) of the sectors l, l', r, r'; St
    the target's 2S; coef the term coefficient in this library's operator
    normalisation."""
    alpha_a = 1.0
    if a_dag and not a_ident:
        alpha_a = get_scaling(lq[1], lqp[1], ka) * adjoint_convention(ka)
    f = coef * Ninej(lqp[1], rqp[1], St, ka, kb, 0, lq[1], rq[1], St)
    if b_fermion and (lqp[0] & 1):
        f = -f
    if b_dag and not b_ident:
        f *= get_scaling(rq[1], rqp[1], kb) * adjoint_convention(kb)
    return alpha_a, f


def worklist_provenance(sb):
    """(pairs [(term, l, l', r, r', ident flags, alpha_a, alpha)], terms
    [(ka, kb, a_dag, b_dag, b_fermion, dn_a, dn_b, coef)], left sectors,
    right sectors, target 2S) of a superblock's compiled work list."""
    f = _h.blkh_worklist_provenance
    ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
    f.restype = ctypes.c_long
    f.argtypes = [ctypes.c_void_p, ctypes.c_long, ip, dp, ip, dp, ip, ip, ip]
    n = f(sb._h, 0, None, None, None, None, None, None, None)
    if n < 0:
        raise block_amd._err()
    nterm = max(1, sb.nterms)
    nl, nr, _ = sb.sectors()
    pi, pd = (ctypes.c_int * (6 * max(1, n)))(), (ctypes.c_double * (2 * max(1, n)))()
    ti, td = (ctypes.c_int * (7 * nterm))(), (ctypes.c_double * nterm)()
    lq, rq, meta = (ctypes.c_int * (3 * nl))(), (ctypes.c_int * (3 * nr))(), (ctypes.c_int * 4)()
    if f(sb._h, n, pi, pd, ti, td, lq, rq, meta) != n:
        raise block_amd._err()
    pairs = [tuple(pi[6 * p:6 * p + 6]) + (pd[2 * p], pd[2 * p + 1]) for p in range(n)]
    terms = [tuple(ti[7 * t:7 * t + 7]) + (td[t],) for t in range(meta[0])]
    L = [tuple(lq[3 * i:3 * i + 3]) for i in range(meta[1])]
    R = [tuple(rq[3 * i:3 * i + 3]) for i in range(meta[2])]
    return pairs, terms, L, R, meta[3]
