"""libblock_sigma.so loads and exports every entry point include/block_sigma.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import block_amd


def test_exports_every_declared_symbol():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(root, "include", "block_sigma.h")).read()
    names = sorted(set(re.findall(r"^\s*(?:int|int64_t|double|const char\*)\s+(bsg_\w+)\s*\(", hdr, re.M)))
    assert len(names) >= 18
    lib = ctypes.CDLL(os.path.join(root, "block_amd", "libblock_sigma.so"))
    for n in names:
        assert hasattr(lib, n), n
    assert block_amd.sigma_lib is not None
