"""Synthetic 3DGS PLY records for the codec tests."""
import numpy as np


def random_records(n, seed=0, with_index=False, dup_every=7):
    """n records of 62 (or 63) floats: positions in a box with every
    dup_every-th position repeated (Morton ties), unnormalised quaternions,
    log-scales, opacity logits, SH of distinct values."""
    rng = np.random.default_rng(seed)
    r = np.zeros((n, 63 if with_index else 62), np.float32)
    r[:, 0:3] = rng.uniform(-3.0, 5.0, size=(n, 3))
    if dup_every and n > dup_every:
        dst = np.arange(dup_every, n, dup_every)
        r[dst, 0:3] = r[rng.integers(0, dup_every, size=dst.shape[0]), 0:3]
    r[:, 3:6] = rng.normal(size=(n, 3))
    r[:, 6:54] = rng.normal(size=(n, 48)) * 0.3
    r[:, 54] = rng.normal(size=n) * 2.0
    r[:, 55:58] = rng.uniform(-6.0, -1.0, size=(n, 3))
    r[:, 58:62] = rng.normal(size=(n, 4)) * rng.uniform(0.2, 3.0, size=(n, 1))
    if with_index:
        r[:, 62] = rng.permutation(n).astype(np.float32) + rng.uniform(-0.4, 0.4, size=n)
    return r
