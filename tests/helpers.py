import os
import tempfile

import numpy as np

import block_amd
from block_amd import integrals as gen

_TMP = tempfile.mkdtemp(prefix="blk_tests_")


def random_fcidump(n, seed, d2h=False, nelec=None):
    """Small synthetic system from the benchmark generator (8-fold symmetric,
    PSD two-electron part), optionally with D2h irreps round-robin."""
    h1, h2, orbsym = gen.synthetic(n, seed, d2h=d2h, nchol=2 * n)
    h1 = h1 * 10.0  # stronger one-body part for a gapped spectrum
    h2 = h2 * 10.0
    path = os.path.join(_TMP, f"rnd_{n}_{seed}_{int(d2h)}.FCIDUMP")
    gen.write_fcidump(path, h1, h2, nelec if nelec is not None else n, orbsym=orbsym, ecore=0.25)
    return block_amd.Integrals(path)


def config_fcidump(name):
    path = os.path.join(_TMP, f"{name}.FCIDUMP")
    if not os.path.exists(path):
        gen.make_config_fcidump(name, path)
    return block_amd.Integrals(path)


def rng_vec(n, seed):
    return np.random.default_rng(seed).standard_normal(n)
