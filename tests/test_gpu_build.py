"""Composite operators built in HBM (the block (x) dot recursions of
Operators.C:624-691,1059-1196,1472-1616,1905-1966,2322-2395 through
TensorProduct, operatorfunctions.C:146-254, as recorded scaled block copies
executed by bsg_build_apply) against the host build of the same operators,
and sweeps whose blocks never leave HBM against host-built sweeps."""
import numpy as np
import pytest

import block_amd
from helpers import config_fcidump, random_fcidump

pytestmark = pytest.mark.gpu
COLS = block_amd.SWEEP_COLS


@pytest.mark.parametrize("name,nsys,M", [("hubbard12", 4, 40), ("syn16", 6, 60), ("syn16d2h", 6, 60)])
def test_device_build_matches_host_build(name, nsys, M):
    """Normal and complementary operator sets of sys (x) dot and sys (x)
    dot's complementary operators in its own basis, from a host block and
    from a device-only one: the same elements as the host build (the device
    sums every element's terms in the host's order, so they agree bitwise)."""
    print(f"[gpu] device build {name}", flush=True)
    ints = {"hubbard12": lambda: config_fcidump("hubbard12"),
            "syn16": lambda: random_fcidump(16, 21),
            "syn16d2h": lambda: random_fcidump(16, 22, d2h=True)}[name]()
    worst, n = block_amd.build_check(ints, nsys, M)
    print(f"[gpu] {name}: {n} operators, max |host - device| = {worst:.3e}", flush=True)
    assert n > 20
    assert worst <= 1e-14


def test_device_blocks_sweep_matches_host_blocks():
    """The same sweeps with blocks kept in HBM and with host-built blocks:
    the same energies (to 1e-12) and sigma counts per sweep."""
    print("[gpu] device-block sweep", flush=True)
    ints = random_fcidump(12, 31)
    sched = [(0, 40, 1e-8, 1e-5), (2, 80, 1e-11, 0.0)]
    e_d, sw_d = block_amd.dmrg(ints, schedule=sched, max_sweeps=4, sweep_tol=0.0, device_blocks=True)
    e_h, sw_h = block_amd.dmrg(ints, schedule=sched, max_sweeps=4, sweep_tol=0.0, device_blocks=False)
    np.testing.assert_allclose(sw_d[:, COLS.index("E0")], sw_h[:, COLS.index("E0")], rtol=0, atol=1e-12)
    assert np.array_equal(sw_d[:, COLS.index("sigma_calls")], sw_h[:, COLS.index("sigma_calls")])
    assert abs(e_d[0] - e_h[0]) <= 1e-12
