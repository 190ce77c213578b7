"""Pins of the host GetData restatement (oracle/oracle_attr.c) that need no GPU.

* unpackHalf2x16: every one of the 65,536 halves against numpy's float16
  (an independent exact conversion);
* the sRGB decode table against the sRGB formula evaluated here in float64;
* the sampler: texel centres return the texel exactly, a half-texel offset
  returns the mean of two texels, repeat wrap, agreement with a float64
  bilinear filter;
* GetData: normals and UVs at the vertices, the normalized mean at the
  centroid, misses, model-colour fallback, out-of-table ids, entity data.
"""
import numpy as np
import pytest


def test_half_unpack_all_values(orc):
    h = np.arange(65536, dtype=np.uint32)
    ref = h.astype(np.uint16).view(np.float16).astype(np.float32)
    got = np.array([orc.half_to_float(int(x)) for x in h], np.float32)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    assert same.all()
    for bits, val in ((0x3C00, 1.0), (0xC000, -2.0), (0x3800, 0.5), (0x0001, 2.0 ** -24), (0x7BFF, 65504.0)):
        assert orc.half_to_float(bits) == val


def test_srgb_table(orc):
    t = orc.srgb_table()
    c = np.arange(256, dtype=np.float64) / 255.0
    ref = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4).astype(np.float32)
    assert np.array_equal(t, ref)
    assert t[0] == 0.0 and t[255] == 1.0
    assert (np.diff(t) > 0).all()


def _tex(rng, w, h):
    return rng.integers(0, 2 ** 32, size=(h, w), dtype=np.uint64).astype(np.uint32)


def test_sampler_closed_forms(orc):
    rng = np.random.default_rng(0)
    w, h = 8, 4
    tex = _tex(rng, w, h)
    lut = orc.srgb_table()
    dec = lambda v: np.array([lut[v & 255], lut[(v >> 8) & 255], lut[(v >> 16) & 255]], np.float32)
    for j in range(h):
        for i in range(w):
            got = orc.sample(tex, w, h, (i + 0.5) / w, (j + 0.5) / h)
            assert np.array_equal(got, dec(int(tex[j, i])))                  # texel centre: the texel
            mid = orc.sample(tex, w, h, (i + 1.0) / w, (j + 0.5) / h)        # halfway to the right neighbour
            ref = 0.5 * (dec(int(tex[j, i])).astype(np.float64) + dec(int(tex[j, (i + 1) % w])))
            assert np.abs(mid - ref).max() <= 1e-7
    # repeat wrap: shifting UV by whole periods changes nothing (UVs on a 2^-10 grid stay exact)
    for _ in range(200):
        u, v = rng.integers(0, 1024, 2) / 1024.0
        a = orc.sample(tex, w, h, u, v)
        assert np.array_equal(a, orc.sample(tex, w, h, u + 3.0, v - 2.0))


def test_sampler_vs_float64_bilinear(orc):
    rng = np.random.default_rng(1)
    w, h = 16, 8
    tex = _tex(rng, w, h)
    lut = orc.srgb_table().astype(np.float64)
    rgb = np.stack([lut[tex & 255], lut[(tex >> 8) & 255], lut[(tex >> 16) & 255]], axis=-1)
    for _ in range(500):
        u, v = rng.uniform(-3, 3, 2)
        x = (u - np.floor(u)) * w - 0.5
        y = (v - np.floor(v)) * h - 0.5
        i0, j0 = int(np.floor(x)), int(np.floor(y))
        a, b = x - i0, y - j0
        i0w, i1w, j0w, j1w = i0 % w, (i0 + 1) % w, j0 % h, (j0 + 1) % h
        ref = ((rgb[j0w, i0w] * (1 - a) + rgb[j0w, i1w] * a) * (1 - b) +
               (rgb[j1w, i0w] * (1 - a) + rgb[j1w, i1w] * a) * b)
        assert np.abs(orc.sample(tex, w, h, u, v) - ref).max() < 2e-6


def _one_triangle():
    import oracle
    from candela_amd import scenes
    pos = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]], np.float32)
    uv = np.array([[0.25, 0.5], [0.75, 0.5], [0.25, 1.0]], np.float32)
    v = scenes.make_vertices(pos, nrm, uv)
    tris = np.zeros(1, oracle.TRIANGLE_DTYPE)
    tris["packed"][0] = [0, 1, 2, 3]
    ent = oracle.identity_entity(1)
    ent["data"][0, 0] = np.float32(2.5).view(np.int32)
    ent["data"][0, 1] = np.float32(0.75).view(np.int32)
    return v, tris, ent


def test_get_data_closed_forms(orc):
    import oracle
    v, tris, ent = _one_triangle()
    refs = np.zeros(4, oracle.TEXTURE_REF_DTYPE)
    refs["albedo"] = -1
    refs["model_color"][3] = [0.1, 0.2, 0.3, 1.0]
    hits = np.zeros(5, oracle.HIT_DTYPE)
    hits["t"] = 1.0
    hits["mesh"] = 3
    hits["tri"] = 0
    hits["entity"] = 0
    bary = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1 / 3, 1 / 3, 1 / 3)]
    for k, (a, b, c) in enumerate(bary):
        hits["u"][k], hits["v"][k], hits["w"][k] = a, b, c
    hits["t"][4] = -1.0                                     # a miss
    out = orc.hit_attributes(tris, v, ent, refs, [], hits)
    assert np.array_equal(out["normal"][0], [0, 0, 1]) and np.array_equal(out["normal"][1], [0, 1, 0])
    assert np.array_equal(out["normal"][2], [1, 0, 0])
    assert np.abs(out["normal"][3] - 1 / np.sqrt(3)).max() < 1e-6
    assert np.array_equal(out["albedo"][0], np.float32([0.1, 0.2, 0.3]))  # model colour
    assert (out["emissivity"][:4] == 2.5).all() and (out["alpha"][:4] == 0.75).all()
    assert np.array_equal(out["normal"][4], [-1, -1, -1]) and (out["albedo"][4] == 0).all()
    assert out["alpha"][4] == 0 and out["emissivity"][4] == 0
    # a textured mesh: at a vertex the UV is the vertex UV, so the sample is the sampler's at that UV
    rng = np.random.default_rng(3)
    tex = _tex(rng, 4, 4)
    refs["albedo"][3] = 0
    out = orc.hit_attributes(tris, v, ent, refs, [(tex, 4, 4)], hits)
    assert np.array_equal(out["albedo"][0], orc.sample(tex, 4, 4, 0.25, 0.5))
    assert np.array_equal(out["albedo"][2], orc.sample(tex, 4, 4, 0.25, 1.0))
    # texture index outside the table -> model colour; mesh outside the table -> albedo 0
    refs["albedo"][3] = 7
    out = orc.hit_attributes(tris, v, ent, refs, [(tex, 4, 4)], hits)
    assert np.array_equal(out["albedo"][0], np.float32([0.1, 0.2, 0.3]))
    hits["mesh"] = 9
    out = orc.hit_attributes(tris, v, ent, refs, [(tex, 4, 4)], hits)
    assert (out["albedo"][:4] == 0).all()
