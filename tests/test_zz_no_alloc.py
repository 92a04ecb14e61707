"""plq_chain_correlator_noalloc through the C ABI (GPU).

The entry point that allocates nothing must return exactly what
plq_chain_correlator returns on the same chain, call after call: its scratch
is one device array shared by every call, so a second call on a smaller
configuration must not see the first call's contents.  It must leave the
device's free memory unchanged and refuse configurations beyond its fixed
scratch with PLQ_EINVAL.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_P = C.POINTER


def _lens(shape):
    return [int(l) // 2 + 1 for l in shape]


def _reference(gpu, chain, model):
    n = sum(_lens(chain.shape))
    sums = np.zeros(n, np.int64)
    vals = np.zeros(n, np.float64)
    rc = gpu.lib().plq_chain_correlator(chain._h, int(model), sums.ctypes.data_as(_P(C.c_int64)),
                                        vals.ctypes.data_as(_P(C.c_double)), n)
    assert rc == 0, gpu.lib().plq_last_error()
    return sums, vals


def _noalloc(gpu, chain, model, want_sums=True, want_vals=True):
    n = sum(_lens(chain.shape))
    sums = np.full(n, -7, np.int64)
    vals = np.full(n, np.nan, np.float64)
    rc = gpu.lib().plq_chain_correlator_noalloc(chain._h, int(model),
                                                sums.ctypes.data_as(_P(C.c_int64)) if want_sums else None,
                                                vals.ctypes.data_as(_P(C.c_double)) if want_vals else None, n)
    return rc, sums, vals


def _free_bytes():
    try:
        hip = C.CDLL("libamdhip64.so")
    except OSError:
        hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _hot(n, q, seed):
    return np.random.default_rng(seed).integers(0, q, n).astype(np.uint8)


def test_noalloc_correlator_equals_the_allocating_one_call_after_call(gpu):
    # a large configuration first, then a smaller one on another chain: the
    # second call must not read the first call's planes or counts
    big = gpu.potts_gpu_chain([256, 128, 16], 5, 1.0, None, 0.6, 3, _hot(256 * 128 * 16, 5, 1))
    big.advance(3)
    small = gpu.ising_gpu_chain([64, 48], 1.0, 0.0, 0.44, 4, _hot(64 * 48, 2, 2))
    small.advance(5)
    for chain, model in ((big, gpu.Model.POTTS), (small, gpu.Model.ISING), (big, gpu.Model.POTTS)):
        want_s, want_v = _reference(gpu, chain, model)
        rc, s, v = _noalloc(gpu, chain, model)
        assert rc == 0, gpu.lib().plq_last_error()
        assert np.array_equal(s, want_s)
        assert np.array_equal(v, want_v)  # bit for bit: the same integer sums, the same final division
        rc, _, v = _noalloc(gpu, chain, model, want_sums=False)
        assert rc == 0 and np.array_equal(v, want_v)
        rc, s, _ = _noalloc(gpu, chain, model, want_vals=False)
        assert rc == 0 and np.array_equal(s, want_s)
    big.close()
    small.close()


def test_noalloc_correlator_on_a_cluster_chain_and_after_more_sweeps(gpu):
    c = gpu.GpuClusterChain.swendsen_wang([128, 96], 3, 1.0, 1.0050525, 9, _hot(128 * 96, 3, 3))
    for _ in range(3):
        c.advance(2)
        want_s, want_v = _reference(gpu, c, gpu.Model.POTTS)
        rc, s, v = _noalloc(gpu, c, gpu.Model.POTTS)
        assert rc == 0 and np.array_equal(s, want_s) and np.array_equal(v, want_v)
    c.close()


def test_noalloc_correlator_leaves_free_memory_unchanged(gpu):
    c = gpu.ising_gpu_chain([128, 128, 64], 1.0, 0.0, 0.2216546, 5, _hot(128 * 128 * 64, 2, 4))
    c.advance(2)
    assert _noalloc(gpu, c, gpu.Model.ISING)[0] == 0  # warm-up: code objects loaded
    before = _free_bytes()
    for _ in range(3):
        assert _noalloc(gpu, c, gpu.Model.ISING)[0] == 0
    assert _free_bytes() == before
    c.close()


def test_noalloc_correlator_refuses_what_its_scratch_cannot_hold(gpu):
    # more than 2^21 plane words (rows of two sites take one word each)
    c = gpu.ising_gpu_chain([2, 4194306], 1.0, 0.0, 0.3, 6, np.zeros(2 * 4194306, np.uint8))
    rc, _, _ = _noalloc(gpu, c, gpu.Model.ISING)
    assert rc == gpu.PLQ_EINVAL and b"no-allocation" in gpu.lib().plq_last_error()
    c.close()
    # more than 8,192 entries (axis 1 alone has 8,201)
    c = gpu.ising_gpu_chain([2, 16400], 1.0, 0.0, 0.3, 6, np.zeros(2 * 16400, np.uint8))
    rc, _, _ = _noalloc(gpu, c, gpu.Model.ISING)
    assert rc == gpu.PLQ_EINVAL
    # nothing to return into
    assert gpu.lib().plq_chain_correlator_noalloc(c._h, int(gpu.Model.ISING), None, None, 8203) == gpu.PLQ_EINVAL
    # and the allocating entry point still serves it
    s, _ = _reference(gpu, c, gpu.Model.ISING)
    assert s[0] == 2 * 16400
    c.close()
