"""GPU parity of GetData (crt_hit_attributes) against the host restatement.

Closest hits of primary and random rays over two objects (the boxes scene and
an icosphere cloud, every vertex with a random half-packed normal and UV),
three entities with distinct emissive / translucency values, a material table
with model colours, three textures (16x16, 64x32, 1x1), an out-of-table
texture index and a mesh id beyond the table.  The attribute records must be
bit-identical; the hits come from the BVH2 and from the BVH4 layout.
"""
import sys

import numpy as np
import pytest

from oracle_cases import random_rays

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _attributed(meshes, rng):
    out = []
    for v, idx, mid in meshes:
        v = v.copy()
        n = rng.normal(size=(v.size, 3)).astype(np.float32)
        uv = rng.uniform(-2, 3, size=(v.size, 2)).astype(np.float32)
        from candela_amd.scenes import _pack_half2
        v["normal_tangent_data"][:, 0] = _pack_half2(n[:, 0], n[:, 1])
        v["normal_tangent_data"][:, 1] = _pack_half2(n[:, 2], np.zeros(v.size, np.float32))
        v["texcoords"] = _pack_half2(uv[:, 0], uv[:, 1])
        out.append((v, idx, mid))
    return out


@pytest.fixture(scope="module")
def setup(crt):
    from candela_amd import scenes
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] attribute scene")
    rng = np.random.default_rng(31)
    r = crt.RayIntersector(0)
    r.add_object(0, _attributed(scenes.boxes_scene(), rng))
    r.add_object(1, _attributed([(v, i, 2 + m) for v, i, m in
                                 scenes.icosphere_cloud(count=30, subdivisions=2, seed=9, box=12.0)], rng))
    up = np.eye(4, dtype=np.float32)
    up[1, 3] = 5.0
    r.push_entity(0, None, emissive=0.5, translucency=0.0)
    r.push_entity(1, up, emissive=3.0, translucency=0.25)
    r.push_entity(1, None, emissive=0.0, translucency=0.5)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(4, 2)
    nmesh = 2 + 30
    refs = np.zeros(nmesh - 1, crt.TEXTURE_REF_DTYPE)      # the last mesh id is beyond the table
    refs["model_color"] = rng.uniform(0, 1, size=(refs.size, 4)).astype(np.float32)
    refs["albedo"] = rng.choice([-1, 0, 1, 2, 99], size=refs.size)
    refs["normal"] = -1
    textures = [rng.integers(0, 2 ** 32, size=s, dtype=np.uint64).astype(np.uint32) for s in ((16, 16), (32, 64), (1, 1))]
    mats = crt.MaterialTable(refs, textures)
    rays = np.concatenate([random_rays(40000, [-4, 0.5, -4], [20, 12, 20], seed=32),
                           random_rays(20000, [-10, 6, -10], [-5, 12, -5], seed=33, toward=[8, 2, 8])])
    return r, refs, textures, mats, rays


def _check(crt, orc, r, scene, refs, textures, mats, rays, label):
    dr = crt.DeviceArray.from_host(rays)
    dh = crt.DeviceArray(crt.HIT_DTYPE, rays.size)
    crt.trace(scene, crt.QUERY_CLOSEST, dr, dh)
    da = crt.DeviceArray(crt.HIT_ATTR_DTYPE, rays.size)
    crt.hit_attributes(scene, mats, dh, da)
    crt.lib().crt_device_sync()
    hits = dh.download()
    got = da.download()
    hb = r.host_buffers()
    ref = orc.hit_attributes(hb["triangles"], hb["vertices"], hb["entities"], refs,
                             [(t, t.shape[1], t.shape[0]) for t in textures], hits)
    same = got.tobytes() == ref.tobytes()
    hit = hits["t"] > 0
    textured = hit & (hits["mesh"] < refs.size)
    textured &= np.isin(refs["albedo"][np.minimum(hits["mesh"], refs.size - 1)], [0, 1, 2])
    log(f"[parity] attributes {label}: {rays.size} hits ({hit.mean():.3f} hit, {textured.mean():.3f} textured), "
        f"bit-identical: {same}")
    assert same
    assert textured.sum() > 1000 and (hit & (hits["mesh"] >= 2)).sum() > 1000


def test_attributes_bvh2(crt, orc, setup):
    r, refs, textures, mats, rays = setup
    _check(crt, orc, r, r.scene(), refs, textures, mats, rays, "BVH2")


def test_attributes_wide(crt, orc, setup):
    r, refs, textures, mats, rays = setup
    _check(crt, orc, r, r.wide_scene(), refs, textures, mats, rays, "BVH4")


def test_attributes_errors(crt, setup):
    r, refs, textures, mats, rays = setup
    dh = crt.DeviceArray(crt.HIT_DTYPE, 4)
    da = crt.DeviceArray(crt.HIT_ATTR_DTYPE, 2)
    with pytest.raises(ValueError):
        crt.hit_attributes(r.scene(), mats, dh, da)
