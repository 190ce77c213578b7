"""Pins of the host path twin (oracle/oracle_path.c) that need no GPU.

* the azimuth sampler's sine / cosine polynomial against math.sin / cos;
* closed forms: over one large floor plane every camera path hits the floor
  and its bounce escapes, so radiance = albedo x sky exactly in 2 segments; a
  zero bounce budget ends every path after its camera segment;
* a white scene (albedo 1, no roulette) gives radiance 0 or sky exactly;
* roulette with p = 0.8 leaves the mean radiance of a furnace unchanged
  within its statistical error, and the same seed reproduces every path.
"""
import math

import numpy as np
import pytest


def test_sincos_polynomial(orc):
    for u in np.linspace(0.0, 1.0, 4001, endpoint=False):
        s, c = orc.sincos_turn(u)
        assert abs(s - math.sin(2 * math.pi * u)) < 1e-6 and abs(c - math.cos(2 * math.pi * u)) < 1e-6


def _plane_scene(crt):
    import oracle
    from candela_amd import scenes
    pos = np.array([[-1e4, 0, -1e4], [1e4, 0, -1e4], [1e4, 0, 1e4], [-1e4, 0, 1e4]], np.float32)
    meshes = [(scenes.make_vertices(pos), np.array([0, 2, 1, 0, 3, 2], np.uint32), 0)]
    r = crt.RayIntersector(0)
    r.add_object(0, meshes)
    hb = r.host_buffers()
    hb["entities"] = oracle.identity_entity(hb["nodes"].size)
    return crt.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 4, 2)


def _params(crt, w, h, spp, eye, center, **kw):
    from candela_amd import scenes
    iv, ip = scenes.camera_inverses(eye=eye, center=center, fovy_deg=60.0, aspect=w / h)
    return crt.path_params(w, h, spp, iv, ip, **kw)


def test_floor_plane_closed_form(crt, orc):
    w = _plane_scene(crt)
    p = _params(crt, 64, 32, 2, (0.0, 5.0, 0.0), (0.0, 0.0, 20.0), sky=(1.0, 2.0, 0.5))
    albedo = np.array([[0.25, 0.5, 0.75]], np.float32)
    out = orc.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, np.arange(64 * 32 * 2))
    down = out["segments"] == 2
    # the upper half of the image sees the sky directly
    assert down.mean() > 0.4 and (out["segments"][~down] == 1).all()
    assert np.array_equal(out["radiance"][down], np.broadcast_to(np.float32([0.25, 1.0, 0.375]), (down.sum(), 3)))
    assert np.array_equal(out["radiance"][~down], np.broadcast_to(np.float32([1.0, 2.0, 0.5]), ((~down).sum(), 3)))
    p["max_bounces"] = 0
    out0 = orc.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, np.arange(64 * 32 * 2))
    assert (out0["segments"] == 1).all() and (out0["radiance"][down] == 0).all()


def test_white_scene_and_roulette(crt, orc):
    import oracle
    from candela_amd import scenes
    meshes = scenes.large_scene(n=64, spheres=12, subdivisions=2, seed=7, size=256.0, height=40.0)
    r = crt.RayIntersector(0)
    r.add_object(0, meshes)
    hb = r.host_buffers()
    hb["entities"] = oracle.identity_entity(hb["nodes"].size)
    w = crt.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 4, 2)
    nmesh = 13
    cam = dict(eye=(128.0, 120.0, -60.0), center=(128.0, 10.0, 128.0))
    paths = np.arange(32 * 18 * 4)
    p = _params(crt, 32, 18, 4, **cam, rr=1.0)
    white = orc.trace_paths(p, 4, w.groups, w.records, w.entities, np.ones((nmesh, 3), np.float32), paths)
    assert set(np.unique(white["radiance"][:, 0])) <= {0.0, 1.0}
    assert (white["segments"] >= 1).all() and (white["segments"] <= 3).all()
    assert (white["status"] == 0).all()
    # Russian roulette keeps the mean: albedo 0.6, survive 0.8 from the second bounce on
    grey = np.full((nmesh, 3), 0.6, np.float32)
    many = np.arange(32 * 18 * 64)
    p = _params(crt, 32, 18, 64, **cam, rr=1.0)
    a = orc.trace_paths(p, 4, w.groups, w.records, w.entities, grey, many)["radiance"][:, 0].astype(np.float64)
    p = _params(crt, 32, 18, 64, **cam, rr=0.8)
    b = orc.trace_paths(p, 4, w.groups, w.records, w.entities, grey, many)["radiance"][:, 0].astype(np.float64)
    err = 4.0 * math.sqrt(a.var() / a.size + b.var() / b.size)
    assert abs(a.mean() - b.mean()) < err
    again = orc.trace_paths(p, 4, w.groups, w.records, w.entities, grey, many[:500])
    assert np.array_equal(again["radiance"][:, 0], b[:500].astype(np.float32))
