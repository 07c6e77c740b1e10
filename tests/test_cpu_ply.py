"""The PLY codec's restatement (oracle.py: morton_codes, ply_load,
ply_save_records, ply_write / ply_read) pinned before it checks the GPU codec,
and the C-ABI header check (arapgs_ply_info, host only).

* Morton codes agree with an independent bit-spreading construction (the
  shift-and-mask "part1by2" spread of a 21-bit integer) and with the closed
  forms min -> 0, max -> 2^63 - 1; de-interleaving returns the integers;
* loading permutes the records by a stable sort of the codes, applies the
  activations (unit quaternions, exp, sigmoid) and interleaves the SH so
  that coefficient j of channel c is f_dc_c (j = 0) or f_rest_{15c + j - 1};
* save(load(records)) gives the records back up to the normals (written 0),
  the quaternion normalisation and the exp/log and sigmoid/logit round trips;
  the crop box is inclusive and too-small Gaussians are dropped;
* arapgs_ply_info reads the count and the index flag and rejects other layouts.
"""
import os

import numpy as np
import pytest

import oracle as O
from arapgs_amd import arapgs
from plyhelp import random_records


def part1by2(x):
    x = x.astype(np.uint64) & np.uint64(0x1FFFFF)
    x = (x | (x << np.uint64(32))) & np.uint64(0x1F00000000FFFF)
    x = (x | (x << np.uint64(16))) & np.uint64(0x1F0000FF0000FF)
    x = (x | (x << np.uint64(8))) & np.uint64(0x100F00F00F00F00F)
    x = (x | (x << np.uint64(4))) & np.uint64(0x10C30C30C30C30C3)
    x = (x | (x << np.uint64(2))) & np.uint64(0x1249249249249249)
    return x


def test_morton_codes_match_bit_spreading_and_closed_forms():
    rng = np.random.default_rng(1)
    pos = rng.uniform(-2, 7, size=(5000, 3)).astype(np.float32)
    code, mn, mx = O.morton_codes(pos)
    scaled = np.float32((1 << 21) - 1) * ((pos - mn) / (mx - mn))
    xyz = scaled.astype(np.int64)
    ref = part1by2(xyz[:, 0]) | (part1by2(xyz[:, 1]) << np.uint64(1)) | (part1by2(xyz[:, 2]) << np.uint64(2))
    assert np.array_equal(code, ref)
    # closed forms: a point at the box minimum / maximum
    corner = np.concatenate([pos, mn[None], mx[None]])
    c2, _, _ = O.morton_codes(corner)
    assert c2[-2] == 0 and c2[-1] == (1 << 63) - 1
    # de-interleave
    for a in range(3):
        back = np.zeros(len(code), np.int64)
        for b in range(21):
            back |= ((code >> np.uint64(3 * b + a)) & np.uint64(1)).astype(np.int64) << b
        assert np.array_equal(back, xyz[:, a])


def test_degenerate_axis_gives_zero_bits():
    pos = np.array([[0, 1, 2], [1, 1, 3], [2, 1, 2.5]], np.float32)
    code, _, _ = O.morton_codes(pos)
    assert not np.any(code & part1by2(np.full(3, (1 << 21) - 1)) << np.uint64(1))


def test_load_permutes_activates_and_interleaves(tmp_path):
    recs = random_records(3000, seed=2)
    p = str(tmp_path / "a.ply")
    O.ply_write(p, recs)
    back, stride = O.ply_read(p)
    assert stride == 62 and np.array_equal(back, recs)
    g = O.ply_load(back)
    code, _, _ = O.morton_codes(recs[:, 0:3])
    order = g["index"]
    assert np.array_equal(np.sort(order), np.arange(3000))
    c = code[order]
    assert np.all(c[1:] >= c[:-1])
    ties = c[1:] == c[:-1]
    assert ties.sum() > 100 and np.all(order[1:][ties] > order[:-1][ties])
    assert np.array_equal(g["pos"], recs[order, 0:3])
    assert np.abs(np.linalg.norm(g["rot"], axis=1) - 1).max() < 1e-6
    assert np.allclose(g["scale"], np.exp(recs[order, 55:58].astype(np.float64)), rtol=1e-6)
    assert np.allclose(g["opacity"], 1 / (1 + np.exp(-recs[order, 54].astype(np.float64))), rtol=1e-6)
    sh = g["sh"].reshape(-1, 16, 3)
    r = recs[order]
    for ch in range(3):
        assert np.array_equal(sh[:, 0, ch], r[:, 6 + ch])
        for j in range(1, 16):
            assert np.array_equal(sh[:, j, ch], r[:, 9 + 15 * ch + j - 1])


def test_index_property_rounds_half_away_from_zero():
    recs = random_records(6, seed=3, with_index=True)
    recs[:, 62] = [0.5, 1.5, 2.5, -0.5, 3.49, 7.0]
    g = O.ply_load(recs, morton=False)
    assert list(g["index"]) == [1, 2, 3, -1, 3, 7]


def test_save_inverts_load_with_box_and_too_small():
    recs = random_records(2000, seed=4)
    g = O.ply_load(recs)
    out = O.ply_save_records(g["pos"], g["rot"], g["scale"], g["opacity"], g["sh"])
    r = recs[g["index"]]
    assert np.array_equal(out[:, 0:3], r[:, 0:3]) and not out[:, 3:6].any()
    assert np.array_equal(out[:, 6:54], r[:, 6:54])
    assert np.allclose(out[:, 55:58], r[:, 55:58], atol=2e-6)
    assert np.allclose(out[:, 54], r[:, 54], atol=1e-4 * np.maximum(1, np.abs(r[:, 54])).max())
    qn = r[:, 58:62] / np.linalg.norm(r[:, 58:62].astype(np.float64), axis=1, keepdims=True)
    assert np.allclose(out[:, 58:62], qn, atol=1e-6)
    lo, hi = np.array([-1, -1, -1], np.float32), np.array([2, 2, 2], np.float32)
    pos = g["pos"].copy()
    pos[0] = lo  # on the boundary: kept
    small = np.zeros(2000, bool)
    small[::3] = True
    small[0] = False
    out = O.ply_save_records(pos, g["rot"], g["scale"], g["opacity"], g["sh"], box=(lo, hi), too_small=small)
    keep = ((pos >= lo) & (pos <= hi)).all(1) & ~small
    assert out.shape[0] == keep.sum() > 0
    assert np.array_equal(out[:, 0:3], pos[keep])


def test_ply_info_reads_header_and_rejects_other_layouts(tmp_path):
    p = str(tmp_path / "a.ply")
    O.ply_write(p, random_records(17, seed=5))
    assert arapgs.ply_info(p) == (17, False)
    O.ply_write(p, random_records(9, seed=5, with_index=True))
    assert arapgs.ply_info(p) == (9, True)
    O.ply_write(p, np.zeros((0, 62), np.float32))
    assert arapgs.ply_info(p) == (0, False)
    with open(p, "rb") as f:
        data = f.read()
    bad = data.replace(b"f_rest_7\n", b"f_rest_X\n")
    with open(p, "wb") as f:
        f.write(bad)
    with pytest.raises(arapgs.ArapgsError):
        arapgs.ply_info(p)
    with open(p, "wb") as f:
        f.write(data.replace(b"binary_little_endian", b"ascii"))
    with pytest.raises(arapgs.ArapgsError):
        arapgs.ply_info(p)
    with pytest.raises(arapgs.ArapgsError):
        arapgs.ply_info(str(tmp_path / "missing.ply"))
