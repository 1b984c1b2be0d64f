"""The scalar factors of the compiled GEMM pairs against Block's conventions:
every pair's alpha_a and alpha recomputed with Block's get_scaling, Ninej
argument order and fermion parity (operatorfunctions.C:508-530,
BaseOperator.C:56-91, couplingCoeffs.h:88-95) from the pair's term
coefficient and sectors must equal what compile_worklist emits; and
getCommuteParity (BaseOperator.C:20-53) equals the Clebsch-Gordan exchange
symmetry (-1)^((a+b-c)/2) times the fermion sign.  Block's get_scaling and
this library's adjoint phase differ by (-1)^(2k) for an operator of spin k
(block_conventions.adjoint_convention), which the check pins for every
transposed view it meets.  The term coefficients themselves are this
build's (own operator normalisation); the check pins the per-sector factors
a Block-side packer derives."""
import pytest

import _oracle
import block_amd
import block_conventions as bc
from helpers import config_fcidump, random_fcidump


@pytest.mark.parametrize("name", ["hubbard12_S0", "syn10_S0", "syn10d2h_S1", "exact6_S2"])
def test_pair_factors_match_block_conventions(name):
    sb = {"hubbard12_S0": lambda: block_amd.SuperBlock(config_fcidump("hubbard12"), 12, 0, 0, 5, 30),
          "syn10_S0": lambda: block_amd.SuperBlock(random_fcidump(10, 51), 10, 0, 0, 4, 40),
          "syn10d2h_S1": lambda: block_amd.SuperBlock(random_fcidump(10, 52, d2h=True), 10, 2, 0, 4, 40),
          "exact6_S2": lambda: _oracle.exact_superblock(random_fcidump(6, 3), 6, 4, 0, 3)}[name]()
    pairs, terms, L, R, St = bc.worklist_provenance(sb)
    assert len(pairs) > 50
    transposed = 0
    for (t, l, lp, r, rp, idf, alpha_a, alpha) in pairs:
        ka, kb, a_dag, b_dag, b_ferm, _, _, coef = terms[t]
        ea, e = bc.pack_pair(coef, ka, kb, a_dag, b_dag, b_ferm, idf & 1, idf & 2, L[l], L[lp], R[r], R[rp], St)
        assert alpha_a == pytest.approx(ea, rel=1e-13, abs=1e-15)
        assert alpha == pytest.approx(e, rel=1e-13, abs=1e-15)
        transposed += bool(a_dag or b_dag)
    assert transposed > 0  # the adjoint views' scalings were exercised


@pytest.mark.parametrize("a,b,c", [(1, 1, 0), (1, 1, 2), (2, 2, 0), (2, 2, 2), (2, 2, 4), (1, 2, 1), (1, 2, 3),
                                   (2, 1, 3), (3, 1, 2), (3, 1, 4), (2, 4, 2)])
def test_commute_parity_closed_form(a, b, c):
    for fa in (False, True):
        for fb in (False, True):
            sign = -1.0 if (fa and fb) else 1.0
            want = sign * (-1.0) ** ((a + b - c) // 2)
            assert bc.get_commute_parity(a, b, c, fa, fb) == pytest.approx(want, abs=1e-14)
