"""Oracle grid build (getGridSamples, GaussianView.cpp:3870-4149; GetBoxesGsGrid,
cudakdtree.cu:425-451) pinned on the CPU by integer closed forms and invariants:

  scene AABB     contains every centre, 1.1x expansion about the mid point, clamped
                 to at least +-0.75, step = largest extent / G
  reorder        a permutation; cell keys non-decreasing; stable within a cell
  Gaussian AABB  half extent per axis = max_d |R[c][d]| sqrt(-2 ln(alpha_c / o)) s_d
                 (the box of the six axis end points); a point when o <= 1/255
  cell lists     brute force per cell: exactly the Gaussians whose padded cell
                 range contains it, ascending; prefix = cumulative counts; valid =
                 non-empty cells ascending
  samples        a 4x4x4 lattice inside every valid cell, centred on the cell

The density sample itself is this build's own definition (the reference kernel is
un-vendored): its test here checks the implementation against that definition by
brute force, which pins nothing about the reference (parity unpinned).
"""
import numpy as np

import oracle as O


def scene(seed=0, N=400):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-1.2, 1.0, size=(N, 3)).astype(np.float32)
    pos[:, 2] *= 0.3
    rot = rng.normal(size=(N, 4)).astype(np.float32)
    scale = np.exp(rng.normal(-3.5, 0.5, size=(N, 3))).astype(np.float32)
    opac = rng.uniform(0.0, 1.0, size=N).astype(np.float32)
    opac[:10] = 1.0 / 255.0   # at the cutoff: point boxes
    return pos, rot, scale, opac


def test_scene_aabb_and_step():
    pos, *_ = scene()
    for G in (16, 64):
        mn, mx, step = O.scene_aabb(pos, G)
        assert np.all(mn <= pos.min(0)) and np.all(mx >= pos.max(0))
        assert np.all(mn <= -0.75) and np.all(mx >= 0.75)
        c = (pos.max(0) + pos.min(0)) / 2
        ext = pos.max(0) - pos.min(0)
        # 1.1x about the mid point (and a little more on the max side), or the +-0.75 clamp
        lo_exp = c - 0.55 * ext
        np.testing.assert_allclose(mn, np.minimum(lo_exp, -0.75), atol=1e-5)
        assert np.all(mx >= np.maximum(c + 0.55 * ext, 0.75) - 1e-5)
        np.testing.assert_allclose(step, (mx - mn).max() / G, rtol=1e-6)


def cells(pos, mn, step, G):
    ijk = np.floor((pos - mn) / step).astype(np.int64)
    return ijk[:, 0] * G * G + ijk[:, 1] * G + ijk[:, 2]


def test_reorder_is_stable_cell_sort():
    pos, *_ = scene(1)
    G = 16
    mn, _, step = O.scene_aabb(pos, G)
    order = O.grid_reorder(pos, mn, step, G)
    assert sorted(order.tolist()) == list(range(pos.shape[0]))
    key = cells(pos, mn, step, G)[order]
    assert np.all(np.diff(key) >= 0)
    same = np.diff(key) == 0
    assert np.all(np.diff(order)[same] > 0)


def quat_mat(q):
    q = q / np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def test_gaussian_aabbs_closed_form():
    pos, rot, scale, opac = scene(2)
    box = O.gaussian_aabbs(pos, rot, scale, opac)
    for g in range(pos.shape[0]):
        if opac[g] <= 1.0 / 255.0:
            np.testing.assert_array_equal(box[g, :3], pos[g])
            np.testing.assert_array_equal(box[g, 3:], pos[g])
            continue
        f = np.sqrt(-2.0 * np.log((1.0 / 255.0) / float(opac[g])))
        R = quat_mat(rot[g].astype(np.float64))
        half = np.abs(R * (f * scale[g].astype(np.float64))[None, :]).max(1)
        np.testing.assert_allclose(box[g, :3], pos[g] - half, atol=1e-6)
        np.testing.assert_allclose(box[g, 3:], pos[g] + half, atol=1e-6)


def test_cell_lists_brute_force_and_samples():
    pos, rot, scale, opac = scene(3, N=300)
    G = 16
    mn, _, step = O.scene_aabb(pos, G)
    box = O.gaussian_aabbs(pos, rot, scale, opac)
    prefix, lst, valid, samples = O.grid_lists(box, mn, step, G, padding=1)
    lo = np.maximum(np.floor((box[:, :3] - mn) / step).astype(np.int64) - 1, 0)
    hi = np.minimum(np.floor((box[:, 3:] - mn) / step).astype(np.int64) + 1, G - 1)
    counts = np.zeros(G ** 3, np.int64)
    start = 0
    nonempty = []
    for c in range(G ** 3):
        ijk = np.array([c // (G * G), (c // G) % G, c % G])
        members = np.nonzero(np.all((lo <= ijk) & (ijk <= hi), axis=1))[0]
        counts[c] = members.size
        assert list(lst[start:start + members.size]) == members.tolist()
        start += members.size
        if members.size:
            nonempty.append(c)
    np.testing.assert_array_equal(prefix, np.cumsum(counts))
    np.testing.assert_array_equal(valid, nonempty)
    # every centre is listed in its own cell
    own = cells(pos, mn, step, G)
    for g in range(0, pos.shape[0], 7):
        c = own[g]
        s = 0 if c == 0 else prefix[c - 1]
        assert g in lst[s:prefix[c]]
    # 4^3 samples inside each valid cell, centred on it
    smp = samples.reshape(len(valid), 64, 3)
    for v, c in enumerate(valid[:50]):
        ijk = np.array([c // (G * G), (c // G) % G, c % G])
        cmin = mn + ijk * step
        assert np.all(smp[v] > cmin - 1e-6) and np.all(smp[v] < cmin + step + 1e-6)
        np.testing.assert_allclose(smp[v].mean(0), cmin + step / 2, atol=1e-6)
        assert len({tuple(p) for p in np.round(smp[v], 6)}) == 64


def test_density_sample_matches_its_definition():
    """Own definition (parity unpinned): a_g(x) = o exp(-1/2 d^T (Sigma + lpf I)^-1 d)."""
    pos, rot, scale, opac = scene(4, N=120)
    sh = np.random.default_rng(4).normal(size=(pos.shape[0], 48)).astype(np.float32)
    G = 8
    mn, _, step = O.scene_aabb(pos, G)
    box = O.gaussian_aabbs(pos, rot, scale, opac)
    prefix, lst, valid, samples = O.grid_lists(box, mn, step, G, padding=1)
    lpf = float((step / 4) ** 2 * 0.2)
    feat, op = O.density_sample(valid, prefix, lst, samples, pos, rot, scale, opac, sh, lpf)
    for v in range(0, len(valid), max(1, len(valid) // 10)):
        c = valid[v]
        s = 0 if c == 0 else prefix[c - 1]
        gs = lst[s:prefix[c]]
        x = samples[v * 64:(v + 1) * 64].astype(np.float64)
        a = np.zeros((64, len(gs)))
        for j, g in enumerate(gs):
            R = quat_mat(rot[g].astype(np.float64))
            Sig = R @ np.diag(scale[g].astype(np.float64) ** 2) @ R.T + lpf * np.eye(3)
            d = x - pos[g]
            a[:, j] = opac[g] * np.exp(-0.5 * np.einsum("ni,ij,nj->n", d, np.linalg.inv(Sig), d))
        np.testing.assert_allclose(op[v * 64:(v + 1) * 64], a.sum(1), rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(feat[v * 64:(v + 1) * 64], a @ sh[gs].astype(np.float64), rtol=1e-4, atol=1e-6)


def _lattice(rng, nv, interval):
    """nv cells of 4x4x4 samples in the grid's order (sample x*16 + y*4 + z)."""
    ijk = np.array([(x, y, z) for x in range(4) for y in range(4) for z in range(4)], np.float64) * interval
    base = rng.uniform(-1, 1, size=(nv, 1, 3))
    return ijk[None], base


def test_ada_lpf_recovers_affine_maps():
    """GetAdaLpfRatio (GaussianView.cpp:4670-4749): the cell's filter is
    lpf_param * M M^T with M the linear map carrying the unit cube's corners to
    the corner samples over the 3 sample intervals. For samples that are an
    exact affine image T u + c of the lattice u, M = interval * T (closed form);
    at rest (the grid's own samples) the filter is lpf_param * interval^2 * I."""
    rng = np.random.default_rng(7)
    interval, lpf_param = 0.05, 0.2
    u, base = _lattice(rng, 12, interval)
    T = np.eye(3)[None] + 0.3 * rng.normal(size=(12, 3, 3))
    x = (np.einsum("vij,vnj->vni", T, np.broadcast_to(u, (12, 64, 3))) + base).astype(np.float32)
    got = O.ada_lpf(x, lpf_param)
    for v in range(12):
        Mv = interval * T[v]
        L = lpf_param * Mv @ Mv.T
        want = np.array([L[0, 0], L[1, 1], L[2, 2], L[0, 1], L[0, 2], L[1, 2]])
        # the corner differences are float32 samples: relative rounding of a few 1e-7 per entry
        assert np.abs(got[v] - want).max() <= 2e-5 * np.abs(want).max(), (v, got[v], want)
    pos, rot, scale, opac = scene(5, N=200)
    G = 8
    mn, _, step = O.scene_aabb(pos, G)
    _, _, valid, samples = O.grid_lists(O.gaussian_aabbs(pos, rot, scale, opac), mn, step, G, padding=1)
    rest = O.ada_lpf(samples, lpf_param)
    spacing = np.float64(np.float32(step) / np.float32(4.0))
    want = lpf_param * spacing ** 2
    assert np.abs(rest[:, :3] - want).max() <= 1e-5 * want
    assert np.abs(rest[:, 3:]).max() <= 1e-6 * want


def test_ada_lpf_is_the_per_row_least_squares_fit():
    """For corners that no linear map fits, each row m_j of M minimises
    sum_c (m_j . q_c - p_jc)^2 over the 8 corners: numpy's lstsq on the 8 x 3
    corner matrix (a different route than the reference's 24 x 9 normal
    equations) gives the same M."""
    rng = np.random.default_rng(8)
    x = rng.normal(size=(6, 64, 3)).astype(np.float32)
    got = O.ada_lpf(x, 1.0)
    Qc = np.array([[c % 2, (c // 2) % 2, (c // 4) % 2] for c in range(8)], np.float64)
    for v in range(6):
        idx = [48 * q[0] + 12 * q[1] + 3 * q[2] for q in Qc.astype(int)]
        P = ((x[v, idx] - x[v, 0]) / np.float32(3.0)).astype(np.float64)  # 8 x 3
        M = np.linalg.lstsq(Qc, P, rcond=None)[0].T  # rows m_j
        L = M @ M.T
        want = np.array([L[0, 0], L[1, 1], L[2, 2], L[0, 1], L[0, 2], L[1, 2]])
        assert np.abs(got[v] - want).max() <= 1e-9 * np.abs(want).max()


def test_density_with_isotropic_ada_equals_scalar_lpf():
    """A per-cell filter of lpf I is the scalar filter: identical outputs; a
    per-cell anisotropic filter changes each cell by its own matrix only
    (brute force on a few cells)."""
    pos, rot, scale, opac = scene(6, N=120)
    sh = np.random.default_rng(6).normal(size=(pos.shape[0], 48)).astype(np.float32)
    G = 8
    mn, _, step = O.scene_aabb(pos, G)
    box = O.gaussian_aabbs(pos, rot, scale, opac)
    prefix, lst, valid, samples = O.grid_lists(box, mn, step, G, padding=1)
    lpf = float(np.float32((step / 4) ** 2 * 0.2))
    iso = np.zeros((len(valid), 6), np.float32)
    iso[:, :3] = lpf
    f0, o0 = O.density_sample(valid, prefix, lst, samples, pos, rot, scale, opac, sh, lpf)
    f1, o1 = O.density_sample(valid, prefix, lst, samples, pos, rot, scale, opac, sh, 123.0, ada=iso)
    np.testing.assert_array_equal(f0, f1)
    np.testing.assert_array_equal(o0, o1)
    rng = np.random.default_rng(9)
    A = rng.normal(size=(len(valid), 3, 3)) * np.sqrt(lpf)
    Lm = np.einsum("vij,vkj->vik", A, A) + lpf * np.eye(3)[None]
    ada = np.stack([Lm[:, 0, 0], Lm[:, 1, 1], Lm[:, 2, 2], Lm[:, 0, 1], Lm[:, 0, 2], Lm[:, 1, 2]], 1)
    ada = ada.astype(np.float32)
    feat, op = O.density_sample(valid, prefix, lst, samples, pos, rot, scale, opac, sh, 0.0, ada=ada)
    for v in range(0, len(valid), max(1, len(valid) // 8)):
        c = valid[v]
        s = 0 if c == 0 else prefix[c - 1]
        gs = lst[s:prefix[c]]
        x = samples[v * 64:(v + 1) * 64].astype(np.float64)
        L = ada[v].astype(np.float64)
        Lv = np.array([[L[0], L[3], L[4]], [L[3], L[1], L[5]], [L[4], L[5], L[2]]])
        a = np.zeros((64, len(gs)))
        for j, g in enumerate(gs):
            R = quat_mat(rot[g].astype(np.float64))
            Sig = R @ np.diag(scale[g].astype(np.float64) ** 2) @ R.T + Lv
            d = x - pos[g]
            a[:, j] = opac[g] * np.exp(-0.5 * np.einsum("ni,ij,nj->n", d, np.linalg.inv(Sig), d))
        np.testing.assert_allclose(op[v * 64:(v + 1) * 64], a.sum(1), rtol=1e-5, atol=1e-7)
        np.testing.assert_allclose(feat[v * 64:(v + 1) * 64], a @ sh[gs].astype(np.float64), rtol=1e-4, atol=1e-6)
