"""Guess wavefunction of the next sweep step (guess_wavefunction.C:524-636 and
the half swap at a turn, :55-139), pinned by exactness: when no block is
truncated the carried wavefunction is the next superblock's eigenvector, so
every step converges on its first sigma; with truncation it must change no
converged energy."""
import os

import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump

COLS = block_amd.SWEEP_COLS
SCHED = [(0, 50, 1e-6, 1e-5), (2, 100, 1e-8, 1e-6), (6, 100, 1e-12, 0.0)]


@pytest.mark.parametrize("nroots", [1, 2])
def test_guess_exact_without_truncation(golden, nroots):
    # 8-site ring, M far above every block dimension: both rotations keep the
    # wavefunction's support, so the 6j recoupling, the fermion signs and the
    # turn's transpose must reproduce the eigenvector exactly
    ints = block_amd.Integrals(os.path.join(golden, "hubbard8_ring.FCIDUMP"))
    sched = [(0, 1000, 1e-12, 0.0)]
    e, sw = _oracle.dmrg_cpu(ints, nroots=nroots, schedule=sched, max_sweeps=4, sweep_tol=1e-12)
    e0, sw0 = _oracle.dmrg_cpu(ints, nroots=nroots, schedule=sched, max_sweeps=4, sweep_tol=1e-12, guess=False)
    steps = ints.norb - 3
    guessed = sw[:, COLS.index("guessed_steps")]
    calls = sw[:, COLS.index("sigma_calls")]
    # sweep 0 starts where the warm-up ended; every later sweep has all steps
    assert guessed[0] >= 1 and np.all(guessed[1:] == steps)
    # one sigma per root and step (the converged guess is accepted at once)
    assert np.all(calls[1:] == nroots * steps), calls
    assert calls[1:].sum() * 5 < sw0[1:, COLS.index("sigma_calls")].sum()
    assert np.abs(e - e0).max() <= 1e-10
    assert e[0] == pytest.approx(-6.568192162874748, abs=1e-9)


def test_guess_truncated_chain_same_energy():
    # 12-site chain at M = 50 / 100 with noise: the transformed guess is
    # approximate; it must cut the sigma count and leave the energy unchanged
    ints = config_fcidump("hubbard12")
    e, sw = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=14, sweep_tol=1e-9)
    e0, sw0 = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=14, sweep_tol=1e-9, guess=False)
    assert abs(e[0] - e0[0]) <= 1e-10
    assert e[0] == pytest.approx(-6.526243384454524, abs=1e-6)
    assert sw[:, COLS.index("sigma_calls")].sum() * 3 < sw0[:, COLS.index("sigma_calls")].sum()

