"""Linear::block_davidson (linear.C:179-385) on its own: the host restatement
with the oracle sigma (CPU) and the device-resident solve with the GPU sigma.

Pinned by the dense spectrum of untruncated superblocks (numpy eigh of the
oracle's H, which itself reproduces determinant FCI): the lowest roots, many
roots at once (a subspace beyond 32 vectors, as the reference's nroots+20
deflation size gives for 12 or more roots, input.C:885), and the lowerStates
projection (linear.C:203-208,311-316,369-375), which must converge to the
lowest state orthogonal to the given ones."""
import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump, random_fcidump


def _exact(n=7, nelec=6, seed=5, twoS=0):
    ints = random_fcidump(n, seed, nelec=nelec)
    sb = _oracle.exact_superblock(ints, nelec, twoS, 0, n // 2)
    H = _oracle.dense_h(sb)
    w, V = np.linalg.eigh(H)
    return sb, w, V


def _guesses(sb, nroots, seed=3):
    """Unit vectors at the lowest diagonal elements with a little seeded noise
    (the sweep's guess without a carried wavefunction)."""
    d = _oracle.diagonal(sb)
    rng = np.random.default_rng(seed)
    g = 1e-3 * rng.standard_normal((nroots, sb.wave_size))
    for r, i in enumerate(np.argsort(d, kind="stable")[:nroots]):
        g[r, i] += 1.0
    return g


def test_lowest_roots_cpu():
    sb, w, _ = _exact()
    e, vecs, st = _oracle.davidson(sb, _guesses(sb, 3), tol=1e-16)
    np.testing.assert_allclose(e, w[:3], rtol=0, atol=1e-10)
    assert st[1] >= 3


def test_many_roots_cpu():
    """14 roots: a subspace of up to 35 vectors (nroots + 20 + 1)."""
    sb, w, _ = _exact()
    e, _, st = _oracle.davidson(sb, _guesses(sb, 14), tol=1e-14)
    assert st[0] < 200  # converged, not stopped at the iteration limit
    np.testing.assert_allclose(e, w[:14], rtol=0, atol=1e-9)


def test_lower_states_cpu():
    """With the two lowest eigenvectors as lower states the solve finds the
    third eigenvalue, and its vector is orthogonal to them."""
    sb, w, V = _exact()
    lower = V[:, :2].T.copy()
    e, vecs, _ = _oracle.davidson(sb, _guesses(sb, 1), tol=1e-16, lower=lower)
    assert e[0] == pytest.approx(w[2], abs=1e-10)
    assert np.abs(lower @ vecs[0]).max() < 1e-8


@pytest.mark.gpu
def test_device_davidson_matches_host_solve():
    """Same guesses, same control flow: energies to 1e-10 and the same number
    of iterations and sigma calls as the host restatement with the oracle
    sigma; the host control flow with the GPU sigma agrees as well."""
    ints = config_fcidump("hubbard12")
    sb = block_amd.SuperBlock(ints, 12, 0, 0, 5, 60)
    g = _guesses(sb, 1)
    e_ref, v_ref, st_ref = _oracle.davidson(sb, g, tol=1e-12)
    e, v, st, on_dev = block_amd.davidson(sb, g, tol=1e-12)
    assert on_dev
    assert abs(e[0] - e_ref[0]) < 1e-10
    assert st[:2] == st_ref[:2]
    assert abs(abs(v[0] @ v_ref[0]) - 1.0) < 1e-10
    e_h, _, st_h, _ = block_amd.davidson(sb, g, tol=1e-12, host_loop=True)
    assert abs(e_h[0] - e_ref[0]) < 1e-10
    assert st_h[:2] == st_ref[:2]


@pytest.mark.gpu
def test_device_davidson_many_roots_and_lower_states():
    sb, w, V = _exact()
    e, _, st, on_dev = block_amd.davidson(sb, _guesses(sb, 14), tol=1e-14)
    assert on_dev and st[0] < 200
    np.testing.assert_allclose(e, w[:14], rtol=0, atol=1e-9)
    lower = V[:, :2].T.copy()
    e, vecs, _, on_dev = block_amd.davidson(sb, _guesses(sb, 1), tol=1e-16, lower=lower)
    assert on_dev
    assert e[0] == pytest.approx(w[2], abs=1e-10)
    assert np.abs(lower @ vecs[0]).max() < 1e-8
