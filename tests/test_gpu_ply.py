"""GPU PLY codec (arapgs_ply_load / arapgs_ply_save) against the oracle's
restatement of loadPly / savePly (oracle.py, pinned by tests/test_cpu_ply.py).

* load, Morton order and file order, with and without the `index` property:
  positions, SH, index (the permutation) and the bounding box bit-exact; unit
  quaternions within 1 ulp-scale (2e-7); exp(scale) and sigmoid(opacity)
  within 2e-6 relative (the device and host exp differ in the last bit);
* save: the records written (crop box, too-small mask, compaction order) are
  the restatement's: every field bit-exact except log(scale) and the opacity
  logit, within 2e-6 absolute (1e-5 for the logit of opacities near 1);
* round trip save -> load returns the saved Gaussians;
* an empty PLY loads and saves; 1M Gaussians load into a permutation with
  non-decreasing Morton codes.
"""
import numpy as np
import pytest

import oracle as O
from plyhelp import random_records

pytestmark = pytest.mark.gpu


def H(t):
    return t.detach().cpu().numpy()


def _check_load(g, ref):
    assert np.array_equal(H(g["index"]), ref["index"])
    assert np.array_equal(H(g["pos"]), ref["pos"])
    assert np.array_equal(H(g["sh"]), ref["sh"])
    assert np.array_equal(g["box"], ref["box"])
    assert np.abs(H(g["rot"]) - ref["rot"]).max() <= 2e-7
    assert np.abs(H(g["scale"]) / ref["scale"] - 1).max() <= 2e-6
    assert np.abs(H(g["opacity"]) / ref["opacity"] - 1).max() <= 2e-6


@pytest.mark.parametrize("morton", [True, False])
@pytest.mark.parametrize("with_index", [False, True])
def test_load_matches_restatement(ctx, tmp_path, morton, with_index):
    recs = random_records(40000, seed=7, with_index=with_index)
    p = str(tmp_path / "a.ply")
    O.ply_write(p, recs)
    print(f"ply_load n=40000 morton={morton} index={with_index}", flush=True)
    g = ctx.ply_load(p, morton=morton)
    _check_load(g, O.ply_load(recs, morton=morton))


def test_save_matches_restatement_and_round_trips(ctx, tmp_path):
    import torch
    recs = random_records(30000, seed=8)
    p = str(tmp_path / "a.ply")
    O.ply_write(p, recs)
    g = ctx.ply_load(p)
    n = g["pos"].shape[0]
    small = torch.zeros(n, dtype=torch.bool, device="cuda")
    small[::5] = True
    lo, hi = np.array([-2, -1, -2.5], np.float32), np.array([4, 3.5, 4], np.float32)
    q = str(tmp_path / "b.ply")
    print("ply_save n=30000 box+too_small", flush=True)
    written = ctx.ply_save(q, g["pos"], g["rot"], g["scale"], g["opacity"], g["sh"], box=(lo, hi), too_small=small)
    got, stride = O.ply_read(q)
    ref = O.ply_save_records(H(g["pos"]), H(g["rot"]), H(g["scale"]), H(g["opacity"]), H(g["sh"]), box=(lo, hi),
                             too_small=H(small))
    assert stride == 62 and written == ref.shape[0] == got.shape[0] > 0
    exact = np.r_[0:54, 58:62]
    assert np.array_equal(got[:, exact], ref[:, exact])
    assert np.abs(got[:, 55:58] - ref[:, 55:58]).max() <= 2e-6
    assert np.abs(got[:, 54] - ref[:, 54]).max() <= 1e-5 * max(1.0, np.abs(ref[:, 54]).max())
    # round trip: the saved file loads back to the kept Gaussians (file order)
    back = ctx.ply_load(q, morton=False)
    keep = H(small) == 0
    keep &= ((H(g["pos"]) >= lo) & (H(g["pos"]) <= hi)).all(1)
    assert np.array_equal(H(back["pos"]), H(g["pos"])[keep])
    assert np.array_equal(H(back["sh"]), H(g["sh"])[keep])
    assert np.abs(H(back["scale"]) / H(g["scale"])[keep] - 1).max() <= 4e-6
    assert np.abs(H(back["opacity"]) - H(g["opacity"])[keep]).max() <= 4e-6
    assert np.abs(H(back["rot"]) - H(g["rot"])[keep]).max() <= 2e-7


def test_save_everything_without_box(ctx, tmp_path):
    recs = random_records(5000, seed=9)
    p = str(tmp_path / "a.ply")
    O.ply_write(p, recs)
    g = ctx.ply_load(p)
    q = str(tmp_path / "b.ply")
    assert ctx.ply_save(q, g["pos"], g["rot"], g["scale"], g["opacity"], g["sh"]) == 5000
    got, _ = O.ply_read(q)
    assert np.array_equal(got[:, 0:3], H(g["pos"]))


def test_empty_ply(ctx, tmp_path):
    import torch
    p = str(tmp_path / "e.ply")
    O.ply_write(p, np.zeros((0, 62), np.float32))
    g = ctx.ply_load(p)
    assert g["pos"].shape == (0, 3)
    e = torch.zeros((0, 3), device="cuda")
    q = str(tmp_path / "f.ply")
    assert ctx.ply_save(q, e, torch.zeros((0, 4), device="cuda"), e, torch.zeros(0, device="cuda"),
                        torch.zeros((0, 48), device="cuda")) == 0
    assert O.ply_read(q)[0].shape == (0, 62)


def test_million_gaussians_load_in_morton_order(ctx, tmp_path):
    n = 1 << 20
    recs = random_records(n, seed=10, dup_every=0)
    p = str(tmp_path / "m.ply")
    O.ply_write(p, recs)
    print(f"ply_load n={n}", flush=True)
    g = ctx.ply_load(p)
    idx = H(g["index"])
    assert np.array_equal(np.sort(idx), np.arange(n))
    code, _, _ = O.morton_codes(recs[:, 0:3])
    c = code[idx]
    assert np.all(c[1:] >= c[:-1])
    assert np.array_equal(H(g["pos"]), recs[idx, 0:3])
