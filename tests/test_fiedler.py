"""Fiedler-vector orbital ordering (fiedler.C:47-81, input.C:2366-2376)."""
import os

import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump


def test_fiedler_reference_unit_case():
    # the reference's own check (fiedler.C, BOOST_AUTO_TEST_CASE(Fiedler)): an
    # 8-site Hueckel chain permuted by {1,0,5,7,4,2,3,6} must come back in
    # chain order; the eigenvector's sign is a convention, so either direction
    idx = [1, 0, 5, 7, 4, 2, 3, 6]
    n = 8
    h = np.zeros((n, n))
    for i in range(n - 1):
        h[i, i + 1] = h[i + 1, i] = -1.0
    hp = np.zeros((n, n))
    for i in range(n):
        for j in range(n):
            hp[idx[i], idx[j]] = h[i, j]
    got = block_amd.fiedler_order_matrix(hp)
    expected = [idx[n - 1 - i] for i in range(n)]
    assert got == expected or got == expected[::-1]


def test_fiedler_recovers_permuted_chain():
    # hopping-only couplings of the 12-site open chain (the exchange integrals
    # vanish): the Fiedler vector of a path graph is monotone along the path
    ints = config_fcidump("hubbard12")
    perm = list(np.random.default_rng(3).permutation(ints.norb))
    shuffled = ints.reordered(perm)
    inv = [0] * ints.norb
    for a, x in enumerate(perm):
        inv[x] = a
    got = shuffled.fiedler_order()
    assert got == inv or got == inv[::-1]
    # identity on the unshuffled chain
    o = ints.fiedler_order()
    assert o == list(range(ints.norb)) or o == list(range(ints.norb))[::-1]


def test_reordered_energy_invariant(golden):
    # the ground state does not depend on the orbital order (untruncated ring;
    # noise in the first sweeps so that states the warm-up's partial chains
    # left at zero weight can come back)
    ints = block_amd.Integrals(os.path.join(golden, "hubbard8_ring.FCIDUMP"))
    r = ints.reordered(ints.fiedler_order())
    sched = [(0, 1000, 1e-8, 1e-5), (2, 1000, 1e-12, 0.0)]
    e, _ = _oracle.dmrg_cpu(r, schedule=sched, max_sweeps=10, sweep_tol=1e-11)
    assert e[0] == pytest.approx(-6.568192162874748, abs=1e-9)
