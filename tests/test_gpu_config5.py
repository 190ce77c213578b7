"""GPU parity on the large-scene multi-bounce workload at its full size.

A 2048 x 2048 fBm terrain (8,388,608 triangles) plus 1,400 subdivision-4
icospheres (7,168,000), about 15.6M triangles in one SAH BVH and its BVH4
layout; 3840 x 2160 pixels x 4 samples (33,177,600 paths), 2 diffuse
bounces, Russian roulette p = 0.8 from the second bounce, all in one
persistent crt_trace_paths launch.  2,000 sampled paths must equal the host
twin bit for bit, every path must end without a guard flag, and the segment
counts must follow the roulette (camera segment always, a third segment on
about 0.8 of the paths that reach it).  The speculative walk (preset 7) must
give every path the same radiance, segments and status.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_config5_paths_match_host_twin(crt, orc):
    from candela_amd import scenes
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    log("[case] config 5: building the 15.6M-triangle scene")
    meshes = scenes.large_scene()
    r = crt.RayIntersector(crt.FORMAT_STACKLESS)
    r.add_object(0, meshes)
    del meshes
    r.push_entity(0)
    r.buffer_entities()
    hb = r.host_buffers()
    ntri = int(hb["triangles"].size)
    assert ntri == 8_388_608 + 7_168_000
    dw = crt.upload_wide(crt.FORMAT_STACKLESS, hb, 4, 2)
    nmesh = int(hb["triangles"]["packed"][:, 3].max()) + 1
    rng = np.random.default_rng(5)
    albedo = rng.uniform(0.2, 0.8, size=(nmesh, 3)).astype(np.float32)
    albedo[0] = (0.45, 0.4, 0.3)
    dalb = crt.DeviceArray.from_host(albedo.ravel())
    W, H, spp = 3840, 2160, 4
    iv, ip = scenes.camera_inverses(aspect=W / H, **scenes.LARGE_CAMERA)
    p = crt.path_params(W, H, spp, iv, ip, albedo=dalb, seed=5, max_bounces=2, rr=0.8, sky=(1.0, 1.0, 1.0))
    n = W * H * spp
    out = crt.DeviceArray(crt.PATH_SAMPLE_DTYPE, n)
    out.fill_bytes(0xAB)
    log(f"[case] config 5: {ntri} triangles, {n} paths, preset 4")
    crt.trace_paths(dw.scene, p, out, n, variant=4)
    crt.lib().crt_device_sync()
    got = out.download()
    idx = np.random.default_rng(6).choice(n, 2000, replace=False)
    w = dw.host
    ref = orc.trace_paths(p, 4, w.groups, w.records, w.entities, albedo, idx)
    diff = int(np.count_nonzero(got[idx].view(np.uint32).reshape(-1, 8) != ref.view(np.uint32).reshape(-1, 8)))
    segs = got["segments"]
    counts = np.bincount(segs, minlength=4)
    log(f"[parity] config 5: {n} paths, {int(segs.sum())} segments (1/2/3: {counts[1:4]}); 2000 sampled paths "
        f"differing words: {diff}")
    assert diff == 0, "sampled paths differ from the host twin"
    assert (got["status"] == 0).all()
    assert (segs >= 1).all() and (segs <= 3).all()
    # roulette: of the paths that reach a second bounce (3 segments possible), about 0.8 continue
    assert counts[3] > 0 and counts[2] > 0
    # the speculative walk (preset 7): radiance, segments and status of every path equal preset 4's
    log("[case] config 5: speculative preset 7")
    out.fill_bytes(0xAB)
    crt.trace_paths(dw.scene, p, out, n, variant=7)
    crt.lib().crt_device_sync()
    spec = out.download()
    assert spec["radiance"].tobytes() == got["radiance"].tobytes(), "speculative radiance differs"
    assert np.array_equal(spec["segments"], got["segments"]) and np.array_equal(spec["status"], got["status"])
    log(f"[parity] config 5: speculative preset 7 gives every path's radiance, segments and status of preset 4")
