"""The determinant FCI oracle pinned on the reference's own fixture: the
8-site Hubbard ring of dmrg_tests/hubbard (FCIDUMP copied to tests/golden)
whose ground state the reference asserts at -6.56819216 (runtest:30, 1e-6)."""
import os

import pytest

import block_amd
import _oracle


def test_hubbard_ring_fci(golden):
    ints = block_amd.Integrals(os.path.join(golden, "hubbard8_ring.FCIDUMP"))
    e, s2 = _oracle.fci(ints, 8, nroots=1, penalty=3.0)
    assert e[0] == pytest.approx(-6.56819216, abs=1e-6)
    assert e[0] == pytest.approx(-6.568192162874748, abs=1e-9)
    assert abs(s2[0]) < 1e-8
