"""Two independent ray batches on two streams (crt_stream_wait_event), through the C ABI.

A fork event on the first stream, a wait on the second, one batch traced on each, and a join event back: every
record must equal those of the same launches run one after the other on one stream, for the refilling speculative
GI launch with the speculative grid AO launch (the config-2 pairing) and for a wave-walk pairing, over several steps
in a row.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle_cases import random_rays

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _event(gpu):
    h = C.c_void_p()
    gpu.check(gpu.lib().crt_event_create(C.byref(h)))
    return h.value


@pytest.mark.parametrize("gi,ao", [((5, 4), (4, 5)), ((3, 0), (3, 1))])
def test_two_streams_equal_one(crt, gi, ao):
    from candela_amd import scenes
    if crt.device_count() < 1:
        pytest.fail("no HIP device visible")
    r = crt.RayIntersector(crt.FORMAT_STACKLESS)
    r.add_object(0, scenes.icosphere_cloud(count=300, subdivisions=3, seed=31, box=40.0))
    r.push_entity(0)
    r.buffer_entities()
    r.buffer_data()
    dw = crt.upload_wide(crt.FORMAT_STACKLESS, r.host_buffers(), 4, 2)
    d_gi = crt.DeviceArray.from_host(random_rays(60000, [-5, -5, -5], [45, 45, 45], seed=71))
    d_ao = crt.DeviceArray.from_host(random_rays(50000, [-5, -5, -5], [45, 45, 45], seed=72, tmax=3.0))
    h_gi, h_ao = crt.DeviceArray(crt.HIT_DTYPE, 60000), crt.DeviceArray(crt.HIT_DTYPE, 50000)
    s1, s2 = crt.Stream(), crt.Stream()
    crt.trace(dw.scene, crt.QUERY_ANY, d_ao, h_ao, kernel=ao[0], variant=ao[1], stream=s1)
    crt.trace(dw.scene, crt.QUERY_CLOSEST, d_gi, h_gi, kernel=gi[0], variant=gi[1], stream=s1)
    s1.sync()
    seq = (h_ao.download(), h_gi.download())
    fork, join = _event(crt), _event(crt)
    lib = crt.lib()
    for step in range(3):
        h_ao.fill_bytes(0xAB)
        h_gi.fill_bytes(0xAB)
        lib.crt_device_sync()
        log(f"[case] two streams gi={gi} ao={ao} step {step}")
        crt.check(lib.crt_event_record(fork, s1.handle))
        crt.check(lib.crt_stream_wait_event(s2.handle, fork))
        crt.trace(dw.scene, crt.QUERY_ANY, d_ao, h_ao, kernel=ao[0], variant=ao[1], stream=s2)
        crt.trace(dw.scene, crt.QUERY_CLOSEST, d_gi, h_gi, kernel=gi[0], variant=gi[1], stream=s1)
        crt.check(lib.crt_event_record(join, s2.handle))
        crt.check(lib.crt_stream_wait_event(s1.handle, join))
        s1.sync()
        assert h_ao.download().tobytes() == seq[0].tobytes(), f"AO records differ (step {step})"
        assert h_gi.download().tobytes() == seq[1].tobytes(), f"GI records differ (step {step})"
    for e in (fork, join):
        lib.crt_event_destroy(e)
    with pytest.raises(crt.CrtError, match="null event"):
        crt.check(lib.crt_stream_wait_event(s1.handle, None))
    s1.close()
    s2.close()
    log(f"[parity] two streams gi={gi} ao={ao}: records equal the one-stream launches over 3 steps")
