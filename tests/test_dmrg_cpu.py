"""Sweep driver + Davidson restatement with the oracle sigma, pinned by the
reference's regression energies (dmrg_tests/runtest) and exact diagonalisation."""
import os

import pytest

import _oracle
import block_amd
from helpers import config_fcidump

SCHED = [(0, 50, 1e-6, 1e-5), (2, 100, 1e-8, 1e-6), (6, 100, 1e-12, 0.0)]


def test_hubbard_ring_two_roots(golden):
    # dmrg_tests/hubbard: nroots 2, weights 0.5 0.5; runtest:30 asserts -6.56819216
    ints = block_amd.Integrals(os.path.join(golden, "hubbard8_ring.FCIDUMP"))
    e, sweeps = _oracle.dmrg_cpu(ints, nroots=2, schedule=SCHED, max_sweeps=12, sweep_tol=1e-9)
    assert e[0] == pytest.approx(-6.56819216, abs=1e-6)
    assert e[0] == pytest.approx(-6.568192162874748, abs=1e-9)


def test_hubbard_chain_12_sites_M100():
    # open chain t=1 U=4, N=12, S=0: exact ground state -6.526243384454524
    ints = config_fcidump("hubbard12")
    e, sweeps = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=14, sweep_tol=1e-9)
    assert e[0] >= -6.526243384454524 - 1e-9
    assert e[0] == pytest.approx(-6.526243384454524, abs=1e-6)


@pytest.mark.parametrize("nroots", [1, 2])
def test_host_ahead_same_sweeps(golden, nroots):
    # complementary operators and the next environment block built on a host
    # thread beside the Davidson solve (the GPU runs' default) are the same
    # operators: bitwise the same energies, truncation and Davidson work (one
    # sigma thread, so that the sigma's own sums are in a fixed order)
    ints = block_amd.Integrals(os.path.join(golden, "hubbard8_ring.FCIDUMP"))
    sched = [(0, 20, 1e-7, 1e-5), (2, 30, 1e-10, 0.0)]
    (e0, s0), (e1, s1) = (_oracle.dmrg_cpu(ints, nroots=nroots, schedule=sched, max_sweeps=4, sweep_tol=0.0,
                                           nthreads=1, ahead=a) for a in (0, 1))
    assert (e0 == e1).all()
    cols = block_amd.SWEEP_COLS
    for c in ("E0", "discarded", "sigma_calls", "guessed_steps"):
        i = cols.index(c)
        assert (s0[:, i] == s1[:, i]).all(), c


def test_state_averaged_four_roots():
    # four state-averaged roots (weights 1/4, the batched configuration's
    # Davidson): untruncated halves of an 8-orbital synthetic system, so the
    # sweep energies are the four lowest S=0 determinant-FCI energies.  With
    # the subspace deflated below nroots vectors (deflation_min_size must be
    # raised to nroots, input.C:882-885) roots 3 and 4 come out wrong.
    from helpers import random_fcidump
    ints = random_fcidump(8, 7)
    e_fci, s2 = _oracle.fci(ints, 8, nroots=10, penalty=3.0)
    singlets = sorted(e for e, s in zip(e_fci, s2) if s < 0.5)[:4]
    sched = [(0, 200, 1e-9, 1e-5), (2, 200, 1e-11, 0.0)]
    e, sw = _oracle.dmrg_cpu(ints, nroots=4, schedule=sched, max_sweeps=6, sweep_tol=0.0)
    assert e == pytest.approx(singlets, abs=1e-8)
