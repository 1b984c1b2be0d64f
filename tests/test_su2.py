"""SU(2) coefficients pinned by closed forms and orthogonality (the values
Block takes from new_anglib.C:19-333 and couplingCoeffs.C:54-68)."""
import itertools
import math

import numpy as np
import pytest

from block_amd import host_lib as L

cg, sixj, ninej, dph = L.blkh_cg, L.blkh_sixj, L.blkh_ninej, L.blkh_dagger_phase
for f in (cg, sixj):
    f.restype = __import__("ctypes").c_double
ninej.restype = __import__("ctypes").c_double
dph.restype = __import__("ctypes").c_double


def test_cg_known_values():
    assert cg(1, 1, 1, -1, 0, 0) == pytest.approx(1 / math.sqrt(2), abs=1e-15)
    assert cg(1, -1, 1, 1, 0, 0) == pytest.approx(-1 / math.sqrt(2), abs=1e-15)
    assert cg(2, 2, 1, -1, 1, 1) == pytest.approx(math.sqrt(2 / 3), abs=1e-15)
    assert cg(2, 0, 1, 1, 1, 1) == pytest.approx(-math.sqrt(1 / 3), abs=1e-15)
    assert cg(2, 0, 2, 0, 0, 0) == pytest.approx(-1 / math.sqrt(3), abs=1e-15)


@pytest.mark.parametrize("j1,j2", [(1, 1), (2, 1), (3, 2), (4, 4), (5, 3)])
def test_cg_orthogonality(j1, j2):
    Js = range(abs(j1 - j2), j1 + j2 + 1, 2)
    for J, Jp in itertools.product(Js, Js):
        for M in range(-min(J, Jp), min(J, Jp) + 1, 2):
            s = sum(cg(j1, m1, j2, M - m1, J, M) * cg(j1, m1, j2, M - m1, Jp, M) for m1 in range(-j1, j1 + 1, 2))
            assert s == pytest.approx(1.0 if J == Jp else 0.0, abs=1e-13)


def test_sixj_values_and_orthogonality():
    assert sixj(2, 2, 2, 2, 2, 2) == pytest.approx(1 / 6, abs=1e-15)
    assert sixj(1, 1, 2, 1, 1, 0) == pytest.approx(1 / 2, abs=1e-15)
    # sum_x (2x+1)(2p+1) {a b x; c d p}{a b x; c d q} = delta_pq
    a, b, c, d = 2, 3, 3, 2
    ps = [p for p in range(0, 12) if sixj(a, b, 1 if False else 1, 0, 0, 0) is not None]
    for p, q in itertools.product(range(0, 7, 2), range(0, 7, 2)):
        s = sum((x + 1) * (p + 1) * sixj(a, b, x, c, d, p) * sixj(a, b, x, c, d, q) for x in range(1, 6, 2))
        if not (abs(a - d) <= p <= a + d and abs(b - c) <= p <= b + c):
            continue
        assert s == pytest.approx(1.0 if p == q else 0.0, abs=1e-13)


@pytest.mark.parametrize("a,b,c,d,e,f", [(1, 1, 2, 1, 1, 0), (2, 1, 1, 2, 3, 1), (2, 2, 2, 2, 2, 2), (3, 1, 2, 1, 3, 2)])
def test_ninej_reduces_to_sixj(a, b, c, d, e, f):
    # {a b e; c d e; f f 0} = (-1)^{b+c+e+f} {a b e; d c f} / sqrt((2e+1)(2f+1))
    lhs = ninej(a, b, e, c, d, e, f, f, 0)
    ph = (-1) ** ((b + c + e + f) // 2)
    rhs = ph * sixj(a, b, e, d, c, f) / math.sqrt((e + 1) * (f + 1))
    assert lhs == pytest.approx(rhs, abs=1e-14)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_dagger_twice_is_parity(k):
    # (T^dag)^dag = (-1)^{2k} T, so phase(J,J') phase(J',J) = (-1)^{2k}
    for J in range(0, 7):
        for Jp in range(0, 7):
            if abs(J - Jp) <= k <= J + Jp and (J + Jp + k) % 2 == 0:
                assert dph(J, Jp, k) * dph(Jp, J, k) == pytest.approx((-1) ** k, abs=1e-14)
