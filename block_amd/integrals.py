"""FCIDUMP inputs of the benchmark configurations (written, then read by the
native FCIDUMP reader, input.C:1407-1560).

* ``hubbard_chain``: open 1D Fermi-Hubbard chain, h_{i,i+1} = -t, (ii|ii) = U.
* ``synthetic``: the seeded ab-initio-like generator
    h_ij = d_ij eps_i + (1 - d_ij) 0.1 N(0,1) exp(-|i-j|/4)   (symmetrised),
    eps_i linear from -2 to +1 Eh,
    (ij|kl) = sum_{L=1..72} B^L_ij B^L_kl,  B^L symmetric, 0.1 N(0,1) exp(-|i-j|/3),
  optionally with D2h irreps assigned round-robin (i mod 8) and every integral
  whose irrep product is not totally symmetric set to zero.
"""
import numpy as np


def write_fcidump(path, h1, h2, nelec, ms2=0, orbsym=None, isym=1, ecore=0.0, tol=0.0):
    n = h1.shape[0]
    orbsym = list(orbsym) if orbsym is not None else [1] * n
    with open(path, "w") as f:
        f.write(f" &FCI NORB={n},NELEC={nelec},MS2={ms2},\n  ORBSYM={','.join(str(x) for x in orbsym)},\n")
        f.write(f"  ISYM={isym},\n &END\n")
        lines = []
        for i in range(n):
            for j in range(i + 1):
                ij = i * (i + 1) // 2 + j
                for k in range(n):
                    for l in range(k + 1):
                        if k * (k + 1) // 2 + l > ij:
                            continue
                        v = h2[i, j, k, l]
                        if abs(v) > tol:
                            lines.append(f"{v:23.16e} {i + 1:4d} {j + 1:4d} {k + 1:4d} {l + 1:4d}\n")
        for i in range(n):
            for j in range(i + 1):
                if abs(h1[i, j]) > tol:
                    lines.append(f"{h1[i, j]:23.16e} {i + 1:4d} {j + 1:4d} {0:4d} {0:4d}\n")
        lines.append(f"{ecore:23.16e} {0:4d} {0:4d} {0:4d} {0:4d}\n")
        f.writelines(lines)


def hubbard_chain(n, t=1.0, U=4.0):
    h1 = np.zeros((n, n))
    for i in range(n - 1):
        h1[i, i + 1] = h1[i + 1, i] = -t
    h2 = np.zeros((n, n, n, n))
    for i in range(n):
        h2[i, i, i, i] = U
    return h1, h2


def synthetic(n, seed, d2h=False, nchol=72):
    rng = np.random.default_rng(seed)
    idx = np.arange(n)
    dist = np.abs(idx[:, None] - idx[None, :])
    eps = np.linspace(-2.0, 1.0, n)
    off = 0.1 * rng.standard_normal((n, n)) * np.exp(-dist / 4.0)
    off = 0.5 * (off + off.T)
    h1 = np.where(dist == 0, np.diag(eps), off)
    B = 0.1 * rng.standard_normal((nchol, n, n)) * np.exp(-dist / 3.0)[None]
    B = 0.5 * (B + B.transpose(0, 2, 1))
    Bf = B.reshape(nchol, n * n)
    h2 = (Bf.T @ Bf).reshape(n, n, n, n)
    orbsym = [1] * n
    if d2h:
        irr = idx % 8
        orbsym = [int(g) + 1 for g in irr]
        sym1 = (irr[:, None] ^ irr[None, :]) == 0
        h1 = np.where(sym1, h1, 0.0)
        x = irr[:, None, None, None] ^ irr[None, :, None, None] ^ irr[None, None, :, None] ^ irr[None, None, None, :]
        h2 = np.where(x == 0, h2, 0.0)
    return h1, h2, orbsym


# Benchmark configurations (seeds are part of the definition).  ramp: the
# sweep schedule (start sweep, M, Davidson tolerance on the squared residual
# norm, density-matrix noise) that grows the blocks to M, ramp_sweeps sweeps
# long; settle: noise-free sweeps at M (Davidson tolerance settle_tol) that
# converge the state after the ramp; timed_tol: the Davidson tolerance of the
# timed sweeps that follow.
_RAMP = {
    500: [(0, 250, 1e-7, 1e-5), (2, 500, 1e-8, 1e-6)],
    800: [(0, 200, 1e-6, 1e-4), (2, 400, 1e-7, 1e-5), (4, 800, 1e-8, 1e-6)],
    1000: [(0, 250, 1e-6, 1e-4), (1, 500, 1e-7, 1e-4), (2, 1000, 1e-8, 1e-5)],
    2000: [(0, 250, 1e-6, 1e-4), (2, 500, 1e-7, 1e-4), (4, 1000, 1e-8, 1e-5), (6, 2000, 1e-8, 1e-6)],
}
CONFIGS = {
    "hubbard12": dict(kind="hubbard", norb=12, nelec=12, M=100, ramp=[(0, 50, 1e-7, 1e-5), (2, 100, 1e-9, 1e-6)],
                      ramp_sweeps=4, timed_tol=1e-12),
    "syn24": dict(kind="synthetic", norb=24, nelec=24, M=500, seed=2024, d2h=False, ramp=_RAMP[500], ramp_sweeps=4,
                  timed_tol=1e-12),
    "syn36d2h": dict(kind="synthetic", norb=36, nelec=36, M=1000, seed=2036, d2h=True, ramp=_RAMP[1000],
                     ramp_sweeps=4, settle=4, settle_tol=1e-10, timed_tol=1e-12),
    "syn30": dict(kind="synthetic", norb=30, nelec=30, M=800, seed=2030, d2h=False, nroots=4, ramp=_RAMP[800],
                  ramp_sweeps=6, timed_tol=1e-12),
    "syn50": dict(kind="synthetic", norb=50, nelec=50, M=2000, seed=2050, d2h=False, ramp=_RAMP[2000],
                  ramp_sweeps=8, timed_tol=1e-12),
}


def make_config_fcidump(name, path):
    cfg = CONFIGS[name]
    if cfg["kind"] == "hubbard":
        h1, h2 = hubbard_chain(cfg["norb"])
        write_fcidump(path, h1, h2, cfg["nelec"])
    else:
        h1, h2, orbsym = synthetic(cfg["norb"], cfg["seed"], cfg["d2h"])
        write_fcidump(path, h1, h2, cfg["nelec"], orbsym=orbsym, tol=1e-14)
    return cfg
