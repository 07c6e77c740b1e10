"""Pins the oracle's FP64 differentiable chain (oracle/chain.cpp) before it is
used to check the HIP adjoint.

* identity node transforms leave every Gaussian and sample at rest (the
  six-point closed form: M = R(q0) diag(len), polar -> R(q0), S_ii = len_i,
  so s' = s0, q' = +-q0, c' = c0, mu' = mu0);
* its FP64 render and density agree with the float restatements
  (oracle/raster.cpp, oracle/grid.cpp) at the same deformed state;
* its forward-mode directional derivatives agree with FP64 central
  differences of its own forward: the density term with eps 1e-4, the render
  term with eps 1e-8 (the render is piecewise smooth: the 1/255 alpha cut, the
  transmittance stop and the tile rectangles switch contributors, and eps
  1e-4 steps cross such switches in this scene).
"""
import numpy as np
import pytest

import oracle as O
from helpers import chain_scene, oracle_chain


@pytest.fixture(scope="module")
def sc():
    return chain_scene()


def _params(sc):
    g = sc["g"]
    return sc["R"], sc["t"], g.rot.astype(np.float64), g.scale.astype(np.float64)


def test_identity_transforms_keep_rest_state(sc):
    ch = oracle_chain(sc)
    g = sc["g"]
    M = sc["nodes"].shape[0]
    out = ch.jvp(np.tile(np.eye(3).reshape(9), (M, 1)), np.zeros((M, 3)), g.rot.astype(np.float64),
                 g.scale.astype(np.float64))
    assert np.abs(out["pos"] - g.pos).max() < 1e-6
    assert np.abs(out["scale"] - g.scale).max() < 1e-6 * g.scale.max()
    assert np.abs(out["sh"] - g.sh).max() < 1e-5
    qn = g.rot / np.linalg.norm(g.rot, axis=1, keepdims=True)
    err = np.minimum(np.abs(out["rot"] - qn).max(1), np.abs(out["rot"] + qn).max(1))
    assert err.max() < 1e-6
    assert np.abs(out["samples"] - sc["samples"]).max() < 1e-6


def test_fp64_render_and_density_match_float_restatements(sc):
    ch = oracle_chain(sc)
    out = ch.jvp(*_params(sc))
    pos, rot, scale, sh = (out[k].astype(np.float32) for k in ("pos", "rot", "scale", "sh"))
    for v, cam in enumerate(sc["cams"]):
        ref = O.raster(pos, sh, sc["g"].opacity, scale, rot, cam.view, cam.proj, cam.campos, cam.tanfovx,
                       cam.tanfovy, sc["W"], sc["H"], sc["bg"], use_rects=False)
        assert np.abs(out["images"][v] - ref["image"]).max() < 1e-3
        assert np.mean(np.abs(out["images"][v] - ref["image"])) < 1e-5
    feat, dens = O.density_sample(sc["valid"], sc["prefix"], sc["list"], out["samples"].astype(np.float32), pos,
                                  rot, scale, sc["g"].opacity, sh, sc["lpf"])
    assert np.abs(out["dens"] - dens).max() <= 1e-4 * max(1.0, np.abs(dens).max())
    assert np.abs(out["feat"] - feat).max() <= 1e-4 * max(1.0, np.abs(feat).max())


@pytest.mark.parametrize("term,eps", [("density", 1e-4), ("render", 1e-8)])
def test_directional_derivative_matches_central_differences(sc, term, eps):
    ch = oracle_chain(sc)
    R, t, q0, s0 = _params(sc)
    rng = np.random.default_rng(5)
    base = ch.forward(R, t, q0, s0)
    gI, gF, gD = rng.normal(size=base[0].shape), rng.normal(size=base[1].shape), rng.normal(size=base[2].shape)
    wI, wF = (1.0, 0.0) if term == "render" else (0.0, 1.0)

    def loss(o):
        return wI * np.sum(gI * o[0]) + wF * (np.sum(gF * o[1]) + np.sum(gD * o[2]))

    for trial in range(4):
        d = [rng.normal(size=R.shape), rng.normal(size=t.shape), rng.normal(size=q0.shape),
             rng.normal(size=s0.shape) * s0]
        d = [x if (trial == 3 or i == trial) else None for i, x in enumerate(d)]
        out = ch.jvp(R, t, q0, s0, *d)
        jv = wI * np.sum(gI * out["d_images"]) + wF * (np.sum(gF * out["d_feat"]) + np.sum(gD * out["d_dens"]))
        step = [(x if dx is None else x + eps * dx, x if dx is None else x - eps * dx)
                for x, dx in zip((R, t, q0, s0), d)]
        fd = (loss(ch.forward(*[s[0] for s in step])) - loss(ch.forward(*[s[1] for s in step]))) / (2 * eps)
        assert abs(jv - fd) <= 1e-5 * max(abs(fd), 1.0), (trial, jv, fd)


def test_density_jvp_is_linear_in_opacity_and_sh(sc):
    """The density sample is linear in the opacities and the SH: the forward-mode
    derivative along a direction equals the difference of two evaluations."""
    g = sc["g"]
    rng = np.random.default_rng(3)
    args = (sc["valid"], sc["prefix"], sc["list"], sc["samples"], g.pos, g.rot, g.scale)
    dop = (0.01 * rng.normal(size=g.opacity.shape)).astype(np.float32)
    dsh = (0.01 * rng.normal(size=g.sh.shape)).astype(np.float32)
    f0, d0, jf, jd = O.density_jvp(*args, g.opacity, g.sh, sc["lpf"], dopac=dop)
    f1, d1, _, _ = O.density_jvp(*args, g.opacity + dop, g.sh, sc["lpf"])
    assert np.abs((d1 - d0) - jd).max() <= 1e-5 * np.abs(jd).max()
    assert np.abs((f1 - f0) - jf).max() <= 1e-5 * np.abs(jf).max()
    f0, d0, jf, jd = O.density_jvp(*args, g.opacity, g.sh, sc["lpf"], dsh=dsh)
    f1, _, _, _ = O.density_jvp(*args, g.opacity, g.sh + dsh, sc["lpf"])
    assert np.abs(jd).max() == 0.0
    assert np.abs((f1 - f0) - jf).max() <= 1e-5 * np.abs(jf).max()


def test_density_jvp_with_adaptive_filter_matches_central_differences(sc):
    """With per-cell adaptive filters the forward-mode derivative of the density
    (rotations, scales, positions, samples) equals FP64 central differences.
    The density takes float32 inputs, so the differences are taken between the
    float32-rounded points x +- eps d and the derivative along their exact
    difference."""
    g = sc["g"]
    rng = np.random.default_rng(12)
    nv = len(sc["valid"])
    ada = np.tile(np.array([sc["lpf"], 1.5 * sc["lpf"], 0.7 * sc["lpf"], 0.2 * sc["lpf"], -0.1 * sc["lpf"],
                            0.05 * sc["lpf"]], np.float32), (nv, 1))
    args = (sc["valid"], sc["prefix"], sc["list"])
    base = dict(samples=sc["samples"], pos=g.pos, rot=g.rot, scale=g.scale)
    raw = dict(samples=rng.normal(size=sc["samples"].shape) * 1e-2, pos=rng.normal(size=g.pos.shape) * 1e-2,
               rot=rng.normal(size=g.rot.shape), scale=rng.normal(size=g.scale.shape) * g.scale)
    eps = 1e-3
    plus = {k: (base[k] + eps * raw[k]).astype(np.float32) for k in base}
    minus = {k: (base[k] - eps * raw[k]).astype(np.float32) for k in base}
    mid = {k: ((plus[k].astype(np.float64) + minus[k]) / 2).astype(np.float32) for k in base}
    dirs = {k: (plus[k].astype(np.float64) - minus[k]) / (2 * eps) for k in base}
    gF = rng.normal(size=(64 * nv, 48))
    gD = rng.normal(size=64 * nv)
    _, _, jf, jd = O.density_jvp(*args, mid["samples"], mid["pos"], mid["rot"], mid["scale"], g.opacity, g.sh,
                                 sc["lpf"], ada=ada, dpos=dirs["pos"], drot=dirs["rot"], dscale=dirs["scale"],
                                 dsamples=dirs["samples"])
    jv = np.sum(gF * jf) + np.sum(gD * jd)
    outs = []
    for pt in (plus, minus):
        f, dn, _, _ = O.density_jvp(*args, pt["samples"], pt["pos"], pt["rot"], pt["scale"], g.opacity, g.sh,
                                    sc["lpf"], ada=ada)
        outs.append(np.sum(gF * f) + np.sum(gD * dn))
    fd = (outs[0] - outs[1]) / (2 * eps)
    assert abs(jv - fd) <= 1e-4 * max(abs(fd), 1.0), (jv, fd)
