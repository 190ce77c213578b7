"""The entity record PushEntity builds (Intersector.h:207-213), without a GPU.

The inverse must be glm::inverse(model) bit for bit: tests/golden/glm_inverse.npz
holds seeded models and the inverses glm 0.9.8.5 computes for them (written by
tests/golden/make_glm_inverse.py from a program compiled against the
reference's vendored glm headers).
"""
import ctypes as C
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def make_entity(crt, model, emissive=0.0, translucency=0.0, node_offset=0, node_count=0):
    e = np.zeros(1, crt.ENTITY_DTYPE)
    m = np.ascontiguousarray(model, np.float32)
    crt.check(crt.lib().crt_make_entity(m.ctypes.data, emissive, translucency, node_offset, node_count, e.ctypes.data))
    return e[0]


def test_inverse_matches_glm(crt):
    d = np.load(os.path.join(HERE, "golden", "glm_inverse.npz"))
    models, inverses = d["model"], d["inverse"]
    assert len(models) >= 100
    bad = []
    for k, (m, inv) in enumerate(zip(models, inverses)):
        e = make_entity(crt, m)
        if e["inverse"].tobytes() != inv.tobytes():
            bad.append(k)
        assert e["model"].tobytes() == m.tobytes()
    assert not bad, f"{len(bad)} of {len(models)} inverses differ from glm::inverse (first {bad[:5]})"


def test_record_fields(crt):
    m = np.eye(4, dtype=np.float32)
    m[3, :3] = (1.0, 2.0, 3.0)                  # column-major: translation in the last column
    e = make_entity(crt, m.reshape(16), emissive=2.5, translucency=0.25, node_offset=7, node_count=99)
    assert (e["node_offset"], e["node_count"]) == (7, 99)
    assert e["data"][0] == np.float32(2.5).view(np.int32)
    assert e["data"][1] == np.float32(0.75).view(np.int32)
    assert (e["data"][2:] == 0).all()
    inv = e["inverse"].reshape(4, 4)
    assert np.array_equal(inv[3, :3], np.float32([-1.0, -2.0, -3.0]))


def test_singular_is_rejected(crt):
    with pytest.raises(crt.CrtError, match="singular"):
        make_entity(crt, np.zeros(16, np.float32))


@pytest.mark.gpu
def test_push_entity_uses_the_same_record(crt):
    """crt_push_entity's record equals crt_make_entity's (checked through BufferEntities' host copy)."""
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    from candela_amd import scenes
    d = np.load(os.path.join(HERE, "golden", "glm_inverse.npz"))
    r = crt.RayIntersector(crt.FORMAT_STACKLESS)
    r.add_object(0, scenes.boxes_scene())
    for m in d["model"][:10]:
        r.push_entity(0, m.reshape(4, 4).T)
    r.buffer_entities()
    ents = r.host_buffers()["entities"]
    assert ents["inverse"].tobytes() == d["inverse"][:10].tobytes()
