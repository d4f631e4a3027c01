"""Readers for the reference's committed test fixtures (data only)."""
import numpy as np


def read_raw_fixture(path):
    """Raw f2 fixture (tests/reference/test_qpadm_parity.cu:143-167):
    int32 P, int32 n_block, int32[n_block] block_sizes, float64[P*P*n_block] at
    i + P*j + P*P*b.  Returns (f2 as (P, P, nb) F-order, block_sizes)."""
    raw = open(path, "rb").read()
    P, nb = np.frombuffer(raw, np.int32, 2)
    bs = np.frombuffer(raw, np.int32, nb, offset=8).copy()
    off = 8 + 4 * int(nb)
    f2 = np.frombuffer(raw, np.float64, int(P) * int(P) * int(nb), offset=off)
    return f2.reshape((int(P), int(P), int(nb)), order="F").copy(), bs


def random_tensor(P, nb, seed, scale=0.05, bmin=20, bmax=400):
    """A symmetric random f2 tensor (P, P, nb) with zero diagonal, correlated across
    populations through a shared base matrix, and ragged block sizes in [bmin, bmax)."""
    rng = np.random.default_rng(seed)
    base = rng.uniform(0.2, 1.0, (P, P))
    base = base + base.T
    f2 = np.empty((P, P, nb))
    for b in range(nb):
        noise = rng.normal(0.0, 0.15, (P, P))
        m = scale * base * (1.0 + noise + noise.T)
        np.fill_diagonal(m, 0.0)
        f2[:, :, b] = m
    bs = rng.integers(bmin, bmax, nb).astype(np.int32)
    return f2, bs
