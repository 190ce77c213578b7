"""GPU parity of the multi-bounce path kernel (crt_trace_paths) against its host twin.

A reduced version of the large-scene workload (a 256 x 256 terrain plus 60
icospheres, 256 x 144 pixels x 2 samples, 2 bounces, roulette 0.8 after the
first bounce): every preset and launch shape must give the same per-path
records, and 5,000 sampled paths must equal the host twin bit for bit.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scene(crt):
    from candela_amd import scenes
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] reduced large scene")
    meshes = scenes.large_scene(n=256, spheres=60, subdivisions=3, seed=11, size=512.0, height=60.0)
    r = crt.RayIntersector(0)
    r.add_object(0, meshes)
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    r.build_wide(4, 2)
    hb = r.host_buffers()
    w = crt.wide_build(0, hb["nodes"], hb["triangles"], hb["vertices"], hb["entities"], 4, 2)
    rng = np.random.default_rng(12)
    albedo = rng.uniform(0.1, 0.9, size=(61, 3)).astype(np.float32)
    return r, w, albedo


def _run(crt, r, albedo, w_img, h_img, spp, variant=0, grid=0, **kw):
    from candela_amd import scenes
    iv, ip = scenes.camera_inverses(eye=(256.0, 140.0, -80.0), center=(256.0, 20.0, 256.0), fovy_deg=60.0,
                                    aspect=w_img / h_img)
    dalb = crt.DeviceArray.from_host(albedo.ravel())
    p = crt.path_params(w_img, h_img, spp, iv, ip, albedo=dalb, **kw)
    n = w_img * h_img * spp
    out = crt.DeviceArray(crt.PATH_SAMPLE_DTYPE, n)
    out.fill_bytes(0xAB)
    crt.trace_paths(r.wide_scene(), p, out, n, variant=variant, grid=grid)
    crt.lib().crt_device_sync()
    return p, out.download()


def test_paths_match_host_twin(crt, orc, scene):
    r, w, albedo = scene
    p, got = _run(crt, r, albedo, 256, 144, 2)
    idx = np.random.default_rng(13).choice(got.size, 5000, replace=False)
    ref = orc.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, idx)
    same = got[idx].tobytes() == ref.tobytes()
    segs = got["segments"]
    log(f"[parity] paths: {got.size} paths, {segs.sum()} segments (1/2/3: {np.bincount(segs, minlength=4)[1:]}), "
        f"5000 sampled paths bit-identical to the host twin: {same}")
    assert same
    assert (got["status"] == 0).all() and (segs >= 1).all() and (segs <= 3).all()
    assert (segs == 3).mean() > 0.03 and (segs == 1).mean() > 0.01


def test_presets_and_grids_agree(crt, scene):
    r, w, albedo = scene
    _, base = _run(crt, r, albedo, 128, 72, 2)
    for variant, grid in ((1, 0), (2, 3), (3, 0), (0, 1), (4, 0), (5, 2)):
        _, got = _run(crt, r, albedo, 128, 72, 2, variant=variant, grid=grid)
        assert got.tobytes() == base.tobytes(), f"variant {variant} grid {grid}"
    log(f"[parity] paths: presets 0..3 and grids 1, 3 give identical records for {base.size} paths")


def _path_fields(a):
    """Radiance bits, segment counts and status: what every preset must agree on (visit and test counts of the
    speculative presets follow their wave schedule)."""
    return a["radiance"].tobytes(), a["segments"].tobytes(), a["status"].tobytes()


def test_speculative_presets(crt, orc, scene):
    """Presets 6..9 walk each segment with the speculative walk: radiance, segments and status of every path
    equal the sequential presets', sampled paths equal the host twin's, and the counters are live."""
    r, w, albedo = scene
    p, base = _run(crt, r, albedo, 256, 144, 2)
    idx = np.random.default_rng(14).choice(base.size, 3000, replace=False)
    ref = orc.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, idx)
    for variant, grid in ((6, 0), (7, 0), (8, 3), (9, 0)):
        log(f"[case] speculative path preset {variant} grid {grid}")
        _, got = _run(crt, r, albedo, 256, 144, 2, variant=variant, grid=grid)
        assert _path_fields(got) == _path_fields(base), f"speculative preset {variant} grid {grid}"
        assert _path_fields(got[idx]) == _path_fields(ref)
        assert (got["visits"] > 0).all() and (got["tests"] > 0).mean() > 0.5
    log(f"[parity] paths: speculative presets 6..9 give the sequential presets' radiance, segments and status "
        f"for {base.size} paths, and the host twin's on 3000 sampled paths")


def test_errors(crt, scene):
    r, w, albedo = scene
    from candela_amd import scenes
    iv, ip = scenes.camera_inverses(eye=(0, 1, 0), center=(0, 0, 1), fovy_deg=60.0, aspect=1.0)
    out = crt.DeviceArray(crt.PATH_SAMPLE_DTYPE, 16)
    with pytest.raises(crt.CrtError):
        crt.trace_paths(r.scene(), crt.path_params(4, 4, 1, iv, ip), out)          # BVH2 scene: wide only
    with pytest.raises(crt.CrtError):
        crt.trace_paths(r.wide_scene(), crt.path_params(4, 4, 1, iv, ip, rr=0.0), out)
    with pytest.raises(crt.CrtError):
        crt.trace_paths(r.wide_scene(), crt.path_params(4, 4, 1, iv, ip), out, variant=10)
