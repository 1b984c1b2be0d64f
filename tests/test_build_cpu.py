"""The composite-operator build as recorded scaled block copies (what the
device executes) against the host build, on the CPU: the recording of
tensor_product / SparseOp::axpy, the grouping by destination rectangle and
the per-element summation order (k_build's arithmetic run on the host) must
reproduce the host build bit for bit (operatorfunctions.C:146-254,
Operators.C:624-2395 through the block (x) dot recursions)."""
import pytest

import block_amd
from helpers import config_fcidump, random_fcidump


@pytest.mark.parametrize("name,nsys,M", [("hubbard12", 4, 30), ("syn12", 5, 40), ("syn12d2h", 5, 40)])
def test_recorded_build_matches_host_build(name, nsys, M):
    ints = {"hubbard12": lambda: config_fcidump("hubbard12"),
            "syn12": lambda: random_fcidump(12, 41),
            "syn12d2h": lambda: random_fcidump(12, 42, d2h=True)}[name]()
    worst, n = block_amd.build_check_cpu(ints, nsys, M)
    assert n > 20
    assert worst == 0.0
