"""krr_accel_trace_closest_noalloc against krr_accel_trace_closest.

The no-allocation variant runs the same traversal kernel with its work
counter and stack overflow in a fixed per-GPU device array, on a grid sized
to fit that array.  Closest hits do not depend on the grid or on where the
stack lives, so its hits must equal the allocating entry point's bit for
bit.  Two calls in a row, the second with fewer rays, check that the shared
array's leftovers (the counter, old stack entries) do not leak into a call,
and that nothing is written past the n hits asked for.
"""
import numpy as np
import pytest

from conftest import icosphere_scene

pytestmark = pytest.mark.gpu


def _rays(pos, idx, n, seed):
    rng = np.random.default_rng(seed)
    tri = rng.integers(0, len(idx), n)
    b = rng.random((n, 3)).astype(np.float32)
    b /= b.sum(1, keepdims=True)
    target = pos[idx[tri, 0]] * b[:, :1] + pos[idx[tri, 1]] * b[:, 1:2] + pos[idx[tri, 2]] * b[:, 2:3]
    org = (np.array([-0.5, 0.0, -0.5]) + np.array([2.0, 1.7, 2.0]) * rng.random((n, 3))).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 3] = np.inf
    rays[::5, 3] = rng.random(len(rays[::5])).astype(np.float32)  # some finite tmax
    rays[:, 4:7] = target - org + 0.02 * (rng.random((n, 3)) - 0.5)
    return rays


@pytest.mark.parametrize("stack_lds", [16, 0])
def test_noalloc_trace_equals_trace_closest(krr, gpu, stack_lds):
    print(f"case: trace_closest_noalloc vs trace_closest, stack_lds_entries={stack_lds}", flush=True)
    lib = krr.load_library()
    assert lib.krr_accel_noalloc_scratch_bytes() == 16 << 20
    s = krr.Scene(icosphere_scene(count=256, level=2, res=(64, 64)))
    pos, idx, _ = s.arrays()
    acc = krr.Accel(s.desc, {"stack_lds_entries": stack_lds})
    n1, n2 = 400_000, 7_001
    big = gpu.DeviceBuffer(n1 * 32)
    want = gpu.DeviceBuffer(n1 * 16)
    got = gpu.DeviceBuffer(n1 * 16)
    sentinel = np.full((n1, 4), -7.0, np.float32)
    for call, (n, seed) in enumerate([(n1, 11), (n2, 12)]):
        rays = _rays(pos, idx, n, seed)
        big.copy_from_numpy(rays)
        got.copy_from_numpy(sentinel)
        acc.trace_closest(big.ptr, n, want.ptr, 0)
        acc.trace_closest_noalloc(big.ptr, n, got.ptr, 0)
        gpu.synchronize()
        w = want.to_numpy(np.float32, (n1, 4))[:n]
        g = got.to_numpy(np.float32, (n1, 4))
        hit = w[:, 1].view(np.int32) >= 0
        print(f"call {call + 1}: {n} rays, {int(hit.sum())} hits, "
              f"equal {int((g[:n].view(np.int32) == w.view(np.int32)).all(1).sum())}", flush=True)
        assert hit.mean() > 0.3
        assert np.array_equal(g[:n].view(np.int32), w.view(np.int32))
        assert np.array_equal(g[n:], sentinel[n:])
    acc.close()
