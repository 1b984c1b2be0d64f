"""Operator rotations of the renormalisation on the GPU (R_r^T O R_c of every
stored and complementary operator, run through the sigma kernels) against
the host rotations of the same block."""
import pytest

import block_amd
from helpers import config_fcidump

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,nsys,M,nmin", [("hubbard12", 5, 60, 10), ("syn24", 9, 150, 200)])
def test_gpu_rotation_matches_host(name, nsys, M, nmin):
    print(f"[gpu] rotation check {name} nsys={nsys} M={M}", flush=True)
    ints = config_fcidump(name)
    worst, n = block_amd.rotation_check(ints, nsys, M)
    print(f"[gpu] {name}: {n} operators, max rel diff {worst:.3e}", flush=True)
    assert n >= nmin
    assert worst <= 1e-12
