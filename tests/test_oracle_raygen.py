"""Pins of the ray-generation host twin (oracle/oracle_raygen.c) that need no GPU.

* diffuse / AO: against the float32 numpy generator (scenes.secondary_rays,
  which uses numpy's sin / cos and norms): same rays in the same order, same
  origins, directions within 1e-6;
* specular: roughness below 0.01 is the exact mirror direction; for rough
  surfaces the half vector recovered from each ray, H = normalize(R - I),
  follows the sampling law of SampleGGXVNDF: its polar angle maps back to a
  uniform on [0, 0.7) (the tail-control scale) and its azimuth to a uniform
  on [0, 0.8); with zero draws accepted it would equal the normal;
* probes: origins equal the grid formula in float32, directions are unit
  vectors with mean 0 and E[z^2] = 1/3, and z = 2 u1 - 1;
* transformed entities: the normal is perpendicular to the world-space edges.
"""
import numpy as np
import pytest
from scipy import stats

from oracle_cases import random_rays


def _scene_hits(crt, orc, meshes, model=None, w=96, h=64):
    from candela_amd import scenes
    r = crt.RayIntersector(0)
    r.add_object(0, meshes)
    hb = r.host_buffers()
    import oracle
    ent = oracle.identity_entity(hb["nodes"].size)
    if model is not None:
        m = np.asarray(model, np.float32)
        ent["model"][0] = m.T.reshape(16)
        ent["inverse"][0] = np.linalg.inv(m.astype(np.float64)).astype(np.float32).T.reshape(16)
    iv, ip = scenes.camera_inverses(aspect=w / h, **scenes.BOXES_CAMERA)
    rays = orc.primary_rays(w, h, iv.T.reshape(16), ip.T.reshape(16))
    hits = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], ent, 0, rays)
    return hb, ent, rays, hits


def test_diffuse_matches_numpy_generator(crt, orc):
    from candela_amd import scenes
    hb, ent, rays, hits = _scene_hits(crt, orc, scenes.boxes_scene())
    p = crt.raygen_params(seed=2, offset=1e-4)
    ao, gi = orc.gen_diffuse_rays(p, rays, hits, ent, tris=hb["triangles"], verts=hb["vertices"])
    pos = hb["vertices"]["position"][:, :3]
    rao, rgi = scenes.secondary_rays(rays, hits, pos, hb["triangles"], seed=2)
    assert ao.size == gi.size == rao.size == int((hits["t"] > 0).sum()) > 1000
    for got, ref in ((ao, rao), (gi, rgi)):
        assert np.array_equal(got["id"], ref["id"])
        assert np.array_equal(got["tmax"], ref["tmax"])
        assert np.allclose(got["org"], ref["org"], rtol=0, atol=2e-6 * (1 + np.abs(ref["org"]).max()))
        assert np.abs(got["dir"] - ref["dir"]).max() < 1e-6
    n = np.linalg.norm(gi["dir"].astype(np.float64), axis=1)
    assert np.abs(n - 1).max() < 1e-6


def test_diffuse_transformed_entity(crt, orc):
    from candela_amd import scenes
    c, s = np.cos(0.7), np.sin(0.7)
    model = np.array([[2 * c, 0, 2 * s, 1.0], [0, 0.5, 0, 0.3], [-s, 0, c, -2.0], [0, 0, 0, 1]], np.float32)
    meshes = scenes.boxes_scene()
    hb, ent, rays, hits = _scene_hits(crt, orc, meshes, model)
    p = crt.raygen_params(seed=3, offset=0.5)          # a long offset makes N recoverable from the origin
    ao, gi = orc.gen_diffuse_rays(p, rays, hits, ent, tris=hb["triangles"], verts=hb["vertices"])
    sel = hits[hits["t"] > 0]
    pos = hb["vertices"]["position"][:, :3].astype(np.float64)
    tv = hb["triangles"]["packed"][sel["tri"], :3]
    a, b, cc = (pos[tv[:, k]] @ model[:3, :3].T.astype(np.float64) for k in range(3))
    # the bounce origin sits on the +normal side; recover N from (org - P) / offset
    r = rays[hits["t"] > 0]
    P = r["org"].astype(np.float64) + r["dir"].astype(np.float64) * sel["t"][:, None]
    N = (gi["org"].astype(np.float64) - P) / 0.5
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    for e in (b - a, cc - a):
        e /= np.linalg.norm(e, axis=1, keepdims=True)
        assert np.abs(np.einsum("ij,ij->i", N, e)).max() < 1e-4
    assert (np.einsum("ij,ij->i", N, r["dir"]) < 0).all()


def _floor_case(n=20000, seed=7):
    """Primary hits on the y = 0 floor of a two-triangle quad, seen from above at random angles."""
    from candela_amd import scenes
    # edges of 128 make |n| = 2^14, so the normalized normal is exactly (0, 1, 0)
    pos = np.array([[-64, 0, -64], [64, 0, -64], [64, 0, 64], [-64, 0, 64]], np.float32)
    meshes = [(scenes.make_vertices(pos), np.array([0, 2, 1, 0, 3, 2], np.uint32), 0)]
    rays = random_rays(n, [-20, 5, -20], [20, 30, 20], seed=seed, toward=[0, 0, 0])
    return meshes, rays


def test_specular_mirror_and_ggx_law(crt, orc):
    import oracle
    meshes, rays = _floor_case()
    r = crt.RayIntersector(0)
    r.add_object(0, meshes)
    hb = r.host_buffers()
    ent = oracle.identity_entity(hb["nodes"].size)
    hits = orc.trace(0, hb["nodes"], hb["triangles"], hb["vertices"], ent, 0, rays)
    live = hits["t"] > 0
    assert live.mean() > 0.5
    inc = rays["dir"][live].astype(np.float32)
    dd = (inc[:, 0] * inc[:, 0] + inc[:, 1] * inc[:, 1]) + inc[:, 2] * inc[:, 2]
    inc = inc * (np.float32(1) / np.sqrt(dd))[:, None]    # normalize as v * (1 / sqrt(v.v))

    # roughness 0: the mirror direction, exactly (N = +y)
    p = crt.raygen_params(seed=4, roughness_scale=0.9)
    out = orc.gen_specular_rays(p, rays, hits, ent, tris=hb["triangles"], verts=hb["vertices"],
                                roughness=np.array([0.0], np.float32))
    want = inc.copy()
    want[:, 1] = -want[:, 1]
    assert np.array_equal(out["dir"], want)

    # roughness 0.8 (0.72 after the 0.9 scale): the half vector's law
    rough = 0.8
    out = orc.gen_specular_rays(p, rays, hits, ent, tris=hb["triangles"], verts=hb["vertices"],
                                roughness=np.array([rough], np.float32))
    R = out["dir"].astype(np.float64)
    H = R - inc.astype(np.float64)
    H /= np.linalg.norm(H, axis=1, keepdims=True)
    H *= np.sign(H[:, 1])[:, None]      # R - I = -2 (H.I) H: the sign flips where I lies in front of H
    c = H[:, 1]
    assert (c > 0.001).all()
    alpha2 = (rough * 0.9) ** 4
    y = (1 - c * c) / (1 + (alpha2 - 1) * c * c)        # inverse of cos^2 = (1 - y) / (1 + (alpha2 - 1) y)
    assert stats.kstest(y / 0.7, "uniform").pvalue > 1e-3
    # azimuth about N in the sampler's basis: up = (0,0,1), tangent = normalize(up x N), bitangent = N x tangent
    t = np.cross([0.0, 0.0, 1.0], [0.0, 1.0, 0.0])
    b = np.cross([0.0, 1.0, 0.0], t)
    phi = np.mod(np.arctan2(H @ b, H @ t), 2 * np.pi) / (2 * np.pi)
    assert phi.max() < 0.8 + 1e-5
    assert stats.kstest(phi / 0.8, "uniform").pvalue > 1e-3
    # origin offset: mix(0.05, 0.1, clamp(r * 1.4, 0, 1)) = 0.1 along +y
    P = rays["org"][live].astype(np.float64) + rays["dir"][live].astype(np.float64) * hits["t"][live][:, None]
    assert np.abs(out["org"][:, 1] - (P[:, 1] + 0.1)).max() < 1e-5


def test_probe_rays():
    import oracle
    g = np.zeros(1, oracle.PROBE_GRID_DTYPE)
    g["res"], g["box_origin"], g["size"], g["seed"], g["tmax"] = (24, 16, 20), (1.5, -2.0, 3.0), (40, 20, 30), 9, 500
    out = oracle.gen_probe_rays(g)
    assert out.size == 24 * 16 * 20
    idx = np.arange(out.size)
    x, y, z = idx % 24, (idx // 24) % 16, idx // (24 * 16)
    for k, (v, res) in enumerate(((x, 24), (y, 16), (z, 20))):
        c = (v.astype(np.float32) / np.float32(res)) * np.float32(2) - np.float32(1)
        want = np.float32(g["box_origin"][0][k]) + c * np.float32(g["size"][0][k])
        assert np.array_equal(out["org"][:, k], want)
    assert np.array_equal(out["id"], idx.astype(np.uint32))
    assert (out["tmax"] == 500).all()
    d = out["dir"].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
    se = 1 / np.sqrt(out.size)
    assert np.abs(d.mean(0)).max() < 5 * se
    assert abs((d[:, 2] ** 2).mean() - 1 / 3) < 5 * se
    from candela_amd import scenes
    u = scenes.hash_uniforms(idx, 9, 2)
    assert np.abs(d[:, 2] - (2 * u[:, 1] - 1)).max() < 1e-6
