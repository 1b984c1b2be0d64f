"""Checkpoint and resume of the sweep (the part of Block's block, rotation
and sweep-state files, save_load_block.C:23-91, rotationmat.C:21-60,
sweep_params.C:156-185): a run interrupted after some sweeps and resumed
from its saved state continues exactly as the uninterrupted run."""
import os

import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump

COLS = block_amd.SWEEP_COLS
SCHED = [(0, 50, 1e-6, 1e-5), (2, 100, 1e-10, 0.0)]


def test_resume_continues_the_sweeps(tmp_path):
    ints = config_fcidump("hubbard12")
    e_all, sw_all = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=6, sweep_tol=0.0)
    ck = str(tmp_path / "state.bin")
    e3, sw3 = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=3, sweep_tol=0.0, checkpoint=ck)
    assert os.path.getsize(ck) > 0
    e, sw = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=6, sweep_tol=0.0, resume=ck)
    # the resumed run carries the first three sweeps' records and runs 3..5
    assert sw.shape == sw_all.shape
    assert np.array_equal(sw[:3, COLS.index("sweep")], sw3[:, COLS.index("sweep")])
    np.testing.assert_allclose(sw[:, COLS.index("E0")], sw_all[:, COLS.index("E0")], rtol=0, atol=1e-10)
    assert abs(e[0] - e_all[0]) <= 1e-10
    # every resumed step started from a carried wavefunction (the turn too)
    assert np.all(sw[3:, COLS.index("guessed_steps")] == ints.norb - 3)


def test_resume_rejects_a_mismatched_run(tmp_path):
    """A state saved after three sweeps cannot continue a run of fewer sweeps
    (its records would not fit the caller's buffer), of another number of
    roots or of another target."""
    ints = config_fcidump("hubbard12")
    ck = str(tmp_path / "state.bin")
    _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=3, sweep_tol=0.0, checkpoint=ck)
    with pytest.raises(RuntimeError, match="max_sweeps"):
        _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=1, sweep_tol=0.0, resume=ck)
    with pytest.raises(RuntimeError, match="roots"):
        _oracle.dmrg_cpu(ints, nroots=2, schedule=SCHED, max_sweeps=6, sweep_tol=0.0, resume=ck)
    with pytest.raises(RuntimeError, match="target"):
        _oracle.dmrg_cpu(ints, twoS=2, schedule=SCHED, max_sweeps=6, sweep_tol=0.0, resume=ck)
    # exactly as many sweeps as saved: nothing left to run, the saved records come back
    e, sw = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=3, sweep_tol=0.0, resume=ck)
    assert sw.shape[0] == 3


def test_superblock_from_saved_state(tmp_path):
    """The mid-chain superblock rebuilt from a saved state's blocks solves to
    the energy the sweeps reached (the state's blocks are converged ones)."""
    ints = config_fcidump("hubbard12")
    ck = str(tmp_path / "state.bin")
    e, sw = _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=6, sweep_tol=0.0, checkpoint=ck)
    sb = block_amd.SuperBlock.from_sweep_state(ints, ck)
    d = _oracle.diagonal(sb)
    g = np.zeros((1, sb.wave_size))
    g[0, np.argmin(d)] = 1.0
    e_sb, _, _ = _oracle.davidson(sb, g, tol=1e-14)
    assert abs(e_sb[0] - e[0]) < 1e-7
    with pytest.raises(block_amd.BlockError):
        block_amd.SuperBlock.from_sweep_state(ints, ck, step=ints.norb - 2)


def test_resume_rejects_garbage(tmp_path):
    ints = config_fcidump("hubbard12")
    bad = tmp_path / "bad.bin"
    bad.write_bytes(b"not a state file at all")
    with pytest.raises(RuntimeError):
        _oracle.dmrg_cpu(ints, schedule=SCHED, max_sweeps=2, resume=str(bad))
