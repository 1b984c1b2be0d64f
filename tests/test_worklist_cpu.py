"""The GEMM-pair work list the GPU executes, run on the CPU, equals the direct
restatement of multiplyH (the contract of the device path)."""
import numpy as np
import pytest

import _oracle
import block_amd
from helpers import config_fcidump, random_fcidump, rng_vec


@pytest.mark.parametrize("case", ["hubbard12", "syn10", "syn10d2h"])
def test_worklist_equals_direct(case):
    if case == "hubbard12":
        ints = config_fcidump("hubbard12")
        sb = block_amd.SuperBlock(ints, 12, 0, 0, 5, 40)
    else:
        ints = random_fcidump(10, 11, d2h=case.endswith("d2h"))
        sb = block_amd.SuperBlock(ints, 10, 0, 0, 4, 40)
    c = rng_vec(sb.wave_size, 5)
    ref = _oracle.sigma(sb, c, nthreads=4)
    got = _oracle.exec_worklist(sb, c)
    assert np.linalg.norm(got - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("case", ["hubbard12", "syn10", "syn10d2h", "syn10_S1"])
def test_sigma_flops_equals_brute_force(case):
    """The flop count the sweep records (walks the operators' allowed blocks)
    equals the brute-force count over every sector quadruple."""
    twoS = 2 if case.endswith("S1") else 0
    if case == "hubbard12":
        sb = block_amd.SuperBlock(config_fcidump("hubbard12"), 12, 0, 0, 5, 40)
    else:
        sb = block_amd.SuperBlock(random_fcidump(10, 11, d2h="d2h" in case), 10, twoS, 0, 4, 40)
    want = _oracle.sigma_flops(sb)
    assert want > 0
    assert sb.flops == pytest.approx(want, rel=1e-15)
