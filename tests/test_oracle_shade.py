"""Pins of the FP64 shading oracle (oracle/oracle_shade.c) that need no GPU.

* closed forms of Cook-Torrance at normal incidence (F = F0, G = 1, D = 1/(pi r^4));
* the hand-derived adjoint against central finite differences on random
  records (the loss is polynomial in albedo and smooth in roughness and
  intensity; FP64 central differences with h = 1e-6 agree to ~1e-9, and the
  gap shrinks as h^2, as a correct derivative must);
* linearity in the light intensities and a zero gradient at zero residual.
"""
import math

import numpy as np


def random_records(n=4000, ntri=25, nl=8, seed=0):
    import oracle
    rng = np.random.default_rng(seed)
    rec = np.zeros(n, oracle.SHADE_RECORD_DTYPE)
    p = rng.uniform(-3, 3, size=(n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    eye = p + nrm * rng.uniform(1, 5, size=(n, 1)) + rng.normal(scale=0.5, size=(n, 3))
    q = p + rng.normal(scale=2.0, size=(n, 3))
    nq = rng.normal(size=(n, 3))
    nq /= np.linalg.norm(nq, axis=1, keepdims=True)
    rec["p"], rec["n"], rec["eye"], rec["q"], rec["nq"] = p, nrm, eye, q, nq
    rec["tri0"] = rng.integers(-1, ntri, n)            # -1: primary miss
    rec["tri1"] = rng.integers(-1, ntri, n)
    rec["vis"] = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    albedo = rng.uniform(0.05, 0.95, size=3 * ntri)
    rough = rng.uniform(0.05, 1.0, size=ntri)
    lpos = rng.uniform(-6, 6, size=(nl, 3)) + np.array([0, 8, 0])
    inten = rng.uniform(1, 10, size=nl)
    return rec, albedo, rough, lpos.ravel(), inten


def test_cook_torrance_normal_incidence(orc):
    for r in (0.1, 0.35, 0.8, 1.0):
        A, kd, spec, dspec = orc.cook_torrance([0, 0, 0], [0, 0, 1], [0, 0, 1], [0, 0, 2], 4.0, r)
        assert A == 1.0                              # radiance 4/2^2, NL = 1
        assert abs(kd - 0.96) < 1e-15               # 1 - F0
        D = 1.0 / (3.14159265359 * r ** 4)
        assert abs(spec - 0.04 * D / 4.0) <= 1e-12 * spec
        assert abs(dspec - (-0.04 / (3.14159265359 * r ** 5))) <= 1e-12 * abs(dspec)


def test_cook_torrance_grazing_light_is_dark(orc):
    # light below the horizon: NL clamps to 0, so the whole term vanishes
    A, kd, spec, dspec = orc.cook_torrance([0, 0, 0], [0, 0, 1], [0, 0, 1], [0, 3, -1], 5.0, 0.5)
    assert A == 0.0


def test_adjoint_matches_central_differences(orc):
    rec, albedo, rough, lpos, inten = random_records()
    rng = np.random.default_rng(1)
    target = (orc.shade_forward(rec, albedo, rough, lpos, inten) *
              rng.uniform(0.5, 1.5, size=(rec.size, 3))).astype(np.float32)
    loss, ga, gr, gi = orc.shade_adjoint(rec, albedo, rough, lpos, inten, target)
    kind = np.concatenate([np.zeros(ga.size), np.ones(gr.size), np.full(gi.size, 2)]).astype(np.int32)
    index = np.concatenate([np.arange(ga.size), np.arange(gr.size), np.arange(gi.size)]).astype(np.int32)
    g = np.concatenate([ga, gr, gi])
    scale = np.abs(g).max()
    errs = []
    for h in (1e-4, 1e-5, 1e-6):
        fd = orc.shade_fd(rec, albedo, rough, lpos, inten, target, kind, index, h=h)
        errs.append(np.abs(fd - g) / np.maximum(np.abs(g), 1e-6 * scale))
    best = np.minimum(errs[1], errs[2])   # h small enough, rounding not yet dominant
    assert best.max() < 1e-6, (best.max(), np.argmax(best))
    # second-order convergence where the truncation error dominates rounding
    big = errs[0] > 1e-6
    assert (errs[1][big] < errs[0][big] / 20).all()
    assert loss > 0


def test_zero_residual_has_zero_gradient(orc):
    rec, albedo, rough, lpos, inten = random_records(n=500, seed=3)
    target = orc.shade_forward(rec, albedo, rough, lpos, inten).astype(np.float32)
    # round-trip through float32 leaves a residual of one float32 ulp at most
    loss, ga, gr, gi = orc.shade_adjoint(rec, albedo, rough, lpos, inten, target)
    c = orc.shade_forward(rec, albedo, rough, lpos, inten)
    assert loss <= 1e-10 * max(1.0, float((c ** 2).sum()))
    assert np.abs(np.concatenate([ga, gr, gi])).max() < 1e-5 * max(1.0, np.abs(c).max())


def test_color_is_linear_in_intensity(orc):
    rec, albedo, rough, lpos, inten = random_records(n=800, seed=4)
    c1 = orc.shade_forward(rec, albedo, rough, lpos, inten)
    c2 = orc.shade_forward(rec, albedo, rough, lpos, 2.0 * inten)
    assert np.allclose(c2, 2.0 * c1, rtol=1e-13, atol=0)
    miss = rec["tri0"] < 0
    assert (c1[miss] == 0).all()
    only0 = inten.copy()
    only0[1:] = 0
    # with all but light 0 switched off, pixels that cannot see light 0 are black
    c0 = orc.shade_forward(rec, albedo, rough, lpos, only0)
    dark = ((rec["vis"] & 1) == 0) & (((rec["vis"] >> 16) & 1) == 0)
    assert (c0[dark] == 0).all()
    assert math.isfinite(float(c0.sum()))
