"""The allocation-free contact record (vx3mi_batch_record_contacts_static, the
library's static per-GPU queue) against the allocating one
(vx3mi_batch_record_contacts), through the C ABI: two calls in a row, the
second on a smaller robot (fewer pairs than the queue still holds from the
first), each against a batch that records with its own allocation.  The pairs
and the states must be identical, and the device's free memory must not change
across the call and the step it records."""
import ctypes as C

import numpy as np
import pytest

import vx3mi
from vx3mi import configs

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime the library runs on (the ROCm one mapped into this
    process; another copy, e.g. one a torch wheel brings, is a separate runtime)."""
    vx3mi.lib()
    paths = sorted({line.split()[-1] for line in open("/proc/self/maps") if "libamdhip64" in line})
    rocm = [p for p in paths if p.startswith("/opt/rocm")]
    return C.CDLL((rocm or paths)[0])


def free_bytes(batch):
    batch.sync()
    free, total = C.c_size_t(), C.c_size_t()
    assert hip_runtime().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def jostled(dims, seed):
    prob = configs.config1(seed=seed, dims=dims)
    m = vx3mi.Model.from_problem(prob)
    b = vx3mi.Batch([m], vx3mi.FP32)
    pos = b.state()["pos"] + np.random.default_rng(seed).normal(scale=1.6e-3, size=(b.nv, 3))
    b.close()
    return m, pos


@pytest.fixture(scope="module", autouse=True)
def setup_queue():
    """The first call on a device takes the queue's device address (the runtime
    loads the library's code object holding the array then); later calls must
    not change the device's free memory."""
    m, _ = jostled((4, 4, 3), 1)
    b = vx3mi.Batch([m], vx3mi.FP32)
    b.record_contacts_static(16)
    b.step(1)
    b.close()


@pytest.mark.parametrize("mode", [vx3mi.MODE_PERSISTENT, vx3mi.MODE_FLAT])
def test_static_record_equals_allocating_record(mode):
    counts = []
    for dims, seed in (((14, 12, 6), 2), ((6, 5, 4), 3)):  # the second call smaller
        m, pos = jostled(dims, seed)
        out = []
        for static in (False, True):
            b = vx3mi.Batch([m, m], vx3mi.FP32)  # two robots: the persistent kernel runs
            b.set_state(0, pos=pos)
            b.set_state(1, pos=pos)
            b.step(5, mode=mode)
            before = free_bytes(b)
            if static:
                b.record_contacts_static(1 << 16)
            else:
                b.record_contacts(1 << 16)
            b.step(1, mode=mode)
            after = free_bytes(b)
            if static and mode == vx3mi.MODE_PERSISTENT:  # (flat steps capture their graph again)
                assert after == before, (before, after)
            out.append((b.contact_pairs(-1), b.state()))
            b.close()
        (pa, sa), (pb, sb) = out
        assert len(pa) > 0
        assert np.array_equal(pa, pb)
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(sa[k], sb[k]), k
        counts.append(len(pa))
    assert counts[1] < counts[0]


def test_static_record_rejects_too_large_capacity():
    m, _ = jostled((6, 5, 4), 3)
    b = vx3mi.Batch([m], vx3mi.FP32)
    with pytest.raises(vx3mi.VX3MIError):
        b.record_contacts_static((1 << 20) + 1)
    b.record_contacts_static(0)  # stops recording
    b.close()
