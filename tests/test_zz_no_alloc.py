"""wf_ray_schedule_noalloc, the allocation-free variant of wf_ray_schedule (it has no counterpart in
the reference; it is for callers that cannot allocate device memory). Called through the C-ABI
and checked against wf_ray_schedule and the oracle on two calls in a row, the second smaller, at
the scratch capacity, and for its guards (too many rays, kernel timing on)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib

pytestmark = pytest.mark.gpu
abi = oracle_lib.abi
f32 = np.float32
BMIN, BMAX = (-10.0, -10.0, -10.0), (10.0, 10.0, 10.0)


def make_rays(rng, n):
    o = rng.uniform(-12, 12, (n, 3)).astype(f32)
    o[:, 0] = np.round(o[:, 0])              # equal keys: ties keep ascending index
    d = rng.normal(size=(n, 3)).astype(f32)
    d[::7, 1] = f32(-0.0)                     # -0.0 sets no octant bit
    return {"ox": o[:, 0].copy(), "oy": o[:, 1].copy(), "oz": o[:, 2].copy(),
            "dx": d[:, 0].copy(), "dy": d[:, 1].copy(), "dz": d[:, 2].copy()}


def schedule_noalloc(wf, ctx, q, n):
    idx = wf.DeviceBuffer(ctx, max(n, 1), np.int32)
    keys = wf.DeviceBuffer(ctx, max(n, 1), np.uint64)
    bmin = (C.c_float * 3)(*BMIN)
    bmax = (C.c_float * 3)(*BMAX)
    soa = q.soa()
    abi.check(ctx._lib.wf_ray_schedule_noalloc(ctx.handle, C.byref(soa), n, bmin, bmax,
                                               C.c_void_p(idx.ptr), C.c_void_p(keys.ptr)))
    ctx.synchronize()
    return idx.download(n), keys.download(n)


def test_noalloc_equals_ray_schedule_twice(wf, ctx, orc):
    print("noalloc: two calls, 300001 then 70003 rays", flush=True)
    rng = np.random.default_rng(41)
    for n in (300_001, 70_003):
        rays = make_rays(rng, n)
        q = wf.RayQueue(ctx, n)
        q.upload(rays)
        got_idx, got_keys = schedule_noalloc(wf, ctx, q, n)
        want_idx, want_keys = wf.ray_schedule(ctx, q, n, BMIN, BMAX)
        assert np.array_equal(got_keys, want_keys)
        assert np.array_equal(got_idx, want_idx)
        assert np.array_equal(got_idx, orc.ray_schedule(rays, BMIN, BMAX))


def test_noalloc_at_capacity(wf, ctx):
    n = abi.NOALLOC_MAX_RAYS
    print(f"noalloc: {n} rays (the capacity)", flush=True)
    rays = make_rays(np.random.default_rng(5), n)
    q = wf.RayQueue(ctx, n)
    q.upload(rays)
    got_idx, _ = schedule_noalloc(wf, ctx, q, n)
    want_idx, _ = wf.ray_schedule(ctx, q, n, BMIN, BMAX)
    assert np.array_equal(got_idx, want_idx)
    small = make_rays(np.random.default_rng(6), 1)
    q1 = wf.RayQueue(ctx, 1)
    q1.upload(small)
    assert schedule_noalloc(wf, ctx, q1, 1)[0].tolist() == [0]


def test_noalloc_guards(wf, ctx):
    q = wf.RayQueue(ctx, 4)
    with pytest.raises(wf.WavefrontError, match="exceed"):
        schedule_noalloc(wf, ctx, q, abi.NOALLOC_MAX_RAYS + 1)
    ctx.profile(True)
    try:
        with pytest.raises(wf.WavefrontError, match="timing"):
            schedule_noalloc(wf, ctx, q, 4)
    finally:
        ctx.profile(False)
    assert schedule_noalloc(wf, ctx, q, 0)[0].size == 0
