"""Davidson vector kernels of the C-ABI against numpy restatements of the
reference's vector algebra, on random vectors: Linear::precondition and
olsenPrecondition (linear.C:27-60), the subspace matrix and Ritz rotation of
block_davidson (linear.C:265-294) beyond one 32 x 32 tile, and the
Gram-Schmidt of the new vector with its Normalise (linear.C:358-366,
BaseOperator.C:282-287)."""
import ctypes

import numpy as np
import pytest

from capi import Device, check, lib

pytestmark = pytest.mark.gpu
N = 12_345


def _precond(x, d, e, shift):
    den = e - d
    out = x.copy()
    m = np.abs(den) > 1e-12
    out[m] = x[m] / (den[m] + shift)
    return out


def test_precondition_and_olsen():
    print("[gpu] precondition / olsen", flush=True)
    rng = np.random.default_rng(1)
    x, c, d = rng.standard_normal(N), rng.standard_normal(N), rng.standard_normal(N)
    e = 0.3
    d[::97] = e  # elements where |e - H_ii| <= 1e-12 are left alone
    with Device() as g:
        dx, dc, dd, dw = g.upload(x), g.upload(c), g.upload(d), g.zeros(N)
        check(lib.bsg_precondition(g.ctx, dx, dd, e, 0.0, N))
        np.testing.assert_allclose(g.download(dx, N), _precond(x, d, e, 0.0), rtol=1e-14, atol=0)
        dx = g.upload(x)
        check(lib.bsg_olsen_precondition(g.ctx, dx, dc, dd, e, 0.0, dw, N))
        ct = _precond(c, d, e, 0.0)
        want = _precond(x - (ct @ x) / (c @ ct) * c, d, e, 0.0)
        np.testing.assert_allclose(g.download(dx, N), want, rtol=1e-11, atol=1e-11 * np.abs(want).max())


@pytest.mark.parametrize("m", [1, 7, 32, 41])
def test_gram_and_lincomb(m):
    print(f"[gpu] gram / lincomb m={m}", flush=True)
    rng = np.random.default_rng(m)
    X, Y = rng.standard_normal((m, N)), rng.standard_normal((m, N))
    with Device() as g:
        dX, dY = g.upload(X), g.upload(Y)
        G = np.zeros((m, m))
        check(lib.bsg_gram(g.ctx, dX, dY, m, N, N, G.ctypes.data))
        np.testing.assert_allclose(G, X @ Y.T, rtol=1e-12, atol=1e-12 * N)
        mo = max(1, m - 4)
        coef = rng.standard_normal((m, mo))
        dZ = g.zeros(mo * N)
        check(lib.bsg_lincomb(g.ctx, dX, m, N, N, coef.ctypes.data, mo, dZ, N))
        np.testing.assert_allclose(g.download(dZ, (mo, N)), coef.T @ X, rtol=1e-12, atol=1e-12 * m)


@pytest.mark.parametrize("m,normalise", [(5, 1), (41, 1), (41, 0)])
def test_gs_orthonormalise(m, normalise):
    print(f"[gpu] gs_orthonormalise m={m}", flush=True)
    rng = np.random.default_rng(10 + m)
    B = np.linalg.qr(rng.standard_normal((N, m)))[0].T.copy()
    r = rng.standard_normal(N)
    want = r.copy()
    for j in range(m):  # linear.C:358-366
        want /= np.sqrt(want @ want)
        want -= (want @ B[j]) * B[j]
    if normalise:
        want /= np.sqrt(want @ want)
    with Device() as g:
        dr, dB = g.upload(r), g.upload(B)
        check(lib.bsg_gs_orthonormalise(g.ctx, dr, dB, m, N, N, normalise))
        got = g.download(dr, N)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    assert np.abs(B @ got).max() < 1e-12
