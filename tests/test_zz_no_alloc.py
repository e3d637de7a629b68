"""The allocation-free draw-list entry point (vx_frustum_draw_list_noalloc),
called through the C ABI on device buffers: two calls in a row, the second
with fewer meshes, each equal byte for byte to vx_frustum_draw_list on the
same inputs; the context's device memory does not change across the calls;
n = 0 clears the count; n above the scratch limit is refused."""
import ctypes as C

import numpy as np
import pytest

import camera
import voxel_amd as V

pytestmark = pytest.mark.gpu

NOALLOC_MAX = 4194304  # VX_DRAW_LIST_NOALLOC_MAX in include/voxel.h


def draw_inputs(rng, n):
    side = int(np.ceil(np.sqrt(n)))
    coords = np.array([(x - side // 2, int(rng.integers(-1, 2)), z - side // 2)
                       for x in range(side) for z in range(side)][:n], np.int32)
    coords = coords[rng.permutation(n)]  # every prefix holds some visible chunks
    faces = rng.integers(0, 5000, n).astype(np.uint32)
    first = np.concatenate([[0], np.cumsum(faces[:-1], dtype=np.uint64)]).astype(np.uint64)
    return coords, faces, first


def call(ctx, fn, f, bufs, n, mode, args, cnt):
    return getattr(ctx.L, fn)(ctx.h, C.byref(f.c), bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, n, 32,
                              mode, args.ptr, cnt.ptr)


@pytest.mark.parametrize("mode", [V.DRAW_INSTANCED, V.DRAW_EXPANDED])
def test_noalloc_draw_list_matches_allocating_entry_point(ctx, mode):
    print("[case] draw list without allocation, mode", mode, flush=True)
    rng = np.random.default_rng(17 + mode)
    eye = np.array([10.0, 30.0, -20.0])
    f = V.Frustum(camera.colmajor(camera.perspective(np.radians(70), 1.6, 0.1, 1024) @
                                  camera.look_at(eye, eye + np.array([0.7, -0.2, 0.4]))))
    n_max = 70000
    coords, faces, first = draw_inputs(rng, n_max)
    bufs = [V.DeviceBuffer.of(ctx, coords), V.DeviceBuffer.of(ctx, faces),
            V.DeviceBuffer.of(ctx, first)]
    a_ref, c_ref = V.DeviceBuffer(ctx, n_max * 20), V.DeviceBuffer(ctx, 8)
    a_got, c_got = V.DeviceBuffer(ctx, n_max * 20), V.DeviceBuffer(ctx, 8)
    try:
        for n in (n_max, 1500):  # a second call with fewer meshes
            # poison the outputs so stale contents cannot pass
            a_got.upload(np.full(n_max * 5, 0xA5A5A5A5, np.uint32))
            c_got.upload(np.array([0xFFFFFFFF, 0xFFFFFFFF], np.uint32))
            ctx.check(call(ctx, "vx_frustum_draw_list", f, bufs, n, mode, a_ref, c_ref),
                      "vx_frustum_draw_list")
            ctx.synchronize()
            used0 = ctx.memory()[0]
            ctx.check(call(ctx, "vx_frustum_draw_list_noalloc", f, bufs, n, mode, a_got, c_got),
                      "vx_frustum_draw_list_noalloc")
            ctx.synchronize()
            assert ctx.memory()[0] == used0  # nothing allocated
            cr, cg = c_ref.download(np.uint32, 2), c_got.download(np.uint32, 2)
            assert (cr == cg).all() and cr[1] == 0
            k = int(cr[0])
            assert 0 < k < n
            assert (a_ref.download(np.uint32, 5 * k) == a_got.download(np.uint32, 5 * k)).all()
            # the oracle of the cull: the host test of each AABB, in mesh order
            vis = [i for i in range(n) if f.contains(tuple(coords[i]), 32)]
            got = a_got.download(np.uint32, 5 * k).reshape(k, 5)
            col = 4 if mode == V.DRAW_INSTANCED else 2
            div = 1 if mode == V.DRAW_INSTANCED else 6
            assert (got[:, col] // div == first[vis].astype(np.uint32)).all()
        # n = 0 clears the count with a kernel; n beyond the scratch is refused
        ctx.check(call(ctx, "vx_frustum_draw_list_noalloc", f, bufs, 0, mode, a_got, c_got),
                  "vx_frustum_draw_list_noalloc")
        ctx.synchronize()
        assert (c_got.download(np.uint32, 2) == 0).all()
        st = call(ctx, "vx_frustum_draw_list_noalloc", f, bufs, NOALLOC_MAX + 1, mode, a_got, c_got)
        assert st == V.ERR_INVALID
    finally:
        for b in bufs + [a_ref, c_ref, a_got, c_got]:
            b.free()
