"""arapgs_knn_noalloc (the KNN entry point that allocates nothing; its scratch
is a static device array) against arapgs_knn, bit for bit, through the C ABI:
two calls in a row on caller-owned outputs, the second with fewer queries and
nodes, and queries with duplicated nodes so that the wave fallback (the part
that uses the scratch list) runs; sizes beyond the static scratch are
rejected."""
import numpy as np
import pytest

from arapgs_amd import arapgs

pytestmark = pytest.mark.gpu


def _case(rng, m, q, dup):
    nodes = rng.uniform(-1, 1, size=(m, 3)).astype(np.float32)
    nodes[1:1 + dup] = nodes[0]  # equal distances: many queries need the full selection-sort replay
    queries = rng.uniform(-1.2, 1.2, size=(q, 3)).astype(np.float32)
    queries[: q // 4] = nodes[0] + 1e-3 * rng.normal(size=(q // 4, 3)).astype(np.float32)
    return nodes, queries


def test_knn_noalloc_matches_knn(ctx):
    import torch
    rng = np.random.default_rng(3)
    for m, q, k, dup in ((2048, 300_000, 8, 24), (300, 5_000, 6, 12)):
        print(f"knn_noalloc: m={m} q={q} k={k}", flush=True)
        nodes, queries = _case(rng, m, q, dup)
        tn, tq = torch.from_numpy(nodes).cuda(), torch.from_numpy(queries).cuda()
        ref_idx, ref_w = ctx.knn(tn, tq, k)
        idx = torch.full((q, k + 1), -7, dtype=torch.int32, device="cuda")
        w = torch.full((q, k), -7.0, dtype=torch.float32, device="cuda")
        ctx.knn_noalloc(tn, tq, k, idx, w)
        torch.cuda.synchronize()
        assert torch.equal(idx, ref_idx)
        # bit patterns: rows whose k+1 nearest nodes are all duplicates at one
        # distance have 0/0 weights (NaN) in both
        assert torch.equal(w.view(torch.int32), ref_w.view(torch.int32))
    big = torch.zeros(((1 << 23) + 1, 3), dtype=torch.float32, device="cuda")
    out = torch.empty((1, 9), dtype=torch.int32, device="cuda")
    with pytest.raises(arapgs.ArapgsError, match="static scratch"):
        ctx.knn_noalloc(tn, big, 8, out, None)
