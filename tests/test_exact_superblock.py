"""An untruncated superblock built with the spin-adapted operator algebra must
reproduce the determinant FCI spectrum of its (N, S=0, irrep) sector; this
pins every coupling coefficient, fermion sign and complementary-operator
recursion that feeds multiplyH (spinblock.C:722-789)."""
import numpy as np
import pytest

import _oracle
from helpers import random_fcidump


@pytest.mark.parametrize("n,nL,nelec,d2h,seed", [(4, 2, 4, False, 1), (5, 2, 4, False, 2), (6, 3, 6, False, 3),
                                                 (6, 2, 6, True, 4), (7, 4, 6, True, 5)])
def test_exact_superblock_spectrum_matches_fci(n, nL, nelec, d2h, seed):
    ints = random_fcidump(n, seed, d2h=d2h, nelec=nelec)
    sb = _oracle.exact_superblock(ints, nelec, 0, 0, nL)
    H = _oracle.dense_h(sb)
    assert np.abs(H - H.T).max() < 1e-12
    w = np.linalg.eigvalsh(H)
    e, s2 = _oracle.fci(ints, nelec, nroots=2, penalty=3.0, irrep=0)
    assert np.all(np.abs(s2) < 1e-6)
    assert w[0] == pytest.approx(e[0], abs=1e-9)
    assert w[1] == pytest.approx(e[1], abs=1e-9)
    # the diagonal restatements agree with the dense matrix
    np.testing.assert_allclose(_oracle.diagonal(sb), np.diag(H), atol=1e-12)
    np.testing.assert_allclose(sb.diagonal_cpu(), np.diag(H), atol=1e-12)


@pytest.mark.parametrize("n,nL,nelec,d2h,seed,twoS", [(5, 2, 4, False, 2, 2), (6, 3, 6, False, 3, 2),
                                                      (6, 3, 6, False, 3, 4), (6, 2, 6, True, 4, 2)])
def test_exact_superblock_spectrum_matches_fci_spin(n, nL, nelec, d2h, seed, twoS):
    """The same for S > 0 targets (the St-dependent 9j factors and fermion
    signs of the work list): the superblock of total spin S reproduces the
    FCI roots whose <S^2> is S(S+1) (tiny penalty, roots filtered by S^2)."""
    ints = random_fcidump(n, seed, d2h=d2h, nelec=nelec)
    sb = _oracle.exact_superblock(ints, nelec, twoS, 0, nL)
    H = _oracle.dense_h(sb)
    assert np.abs(H - H.T).max() < 1e-12
    w = np.linalg.eigvalsh(H)
    e, s2 = _oracle.fci(ints, nelec, nroots=60, penalty=1e-4, irrep=0)
    S = twoS / 2
    sel = e[np.abs(s2 - S * (S + 1)) < 1e-6]
    assert len(sel) >= 2
    assert w[0] == pytest.approx(sel[0], abs=1e-9)
    assert w[1] == pytest.approx(sel[1], abs=1e-9)
    np.testing.assert_allclose(_oracle.diagonal(sb), np.diag(H), atol=1e-12)
    np.testing.assert_allclose(sb.diagonal_cpu(), np.diag(H), atol=1e-12)
