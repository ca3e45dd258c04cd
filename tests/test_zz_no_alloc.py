"""The allocation-free standalone narrowphase (kzg_collide_batch_device):
device pointers in and out, its scratch in the library's per-GPU array.
Checked bit for bit against kzg_collide_batch on the same pairs, on two
calls in a row (the second smaller, so stale scratch from the first would
show), at the size limit, and on its error paths."""
import ctypes as C

import numpy as np
import pytest

from kizuriphysics3d_amd import device as D
from kizuriphysics3d_amd import scenes as S
from kizuriphysics3d_amd.world import build_hulls

pytestmark = pytest.mark.gpu

MAX_PAIRS = 16384       # KZG_NO_ALLOC_MAX_PAIRS
ERR_ARGUMENT = -1       # KZG_ERR_ARGUMENT


def _err(ctx):
    return D.library().kzg_last_error(ctx.h).decode()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime the library itself
    is linked against (the caller owns the device buffers; the entry point
    allocates nothing)."""

    def __init__(self):
        D.library()  # loads libamdhip64.so.7 as the library's dependency
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a: np.ndarray) -> C.c_void_p:
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.live.append(p)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p: C.c_void_p, out: np.ndarray) -> np.ndarray:
        assert self.L.hipDeviceSynchronize() == 0
        if out.nbytes:
            assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def _call(ctx, sa, sb, xa, xb, max_sep=0.02):
    """Run kzg_collide_batch_device on device copies of the inputs."""
    hip = _Hip()
    try:
        n = len(sa)
        d_sa, d_sb = hip.put(sa.astype(np.uint32)), hip.put(sb.astype(np.uint32))
        d_xa, d_xb = hip.put(xa.astype(np.float32)), hip.put(xb.astype(np.float32))
        # outputs pre-filled with junk: every slot must be written
        d_counts = hip.put(np.full(max(n, 1), 0x7F7F7F7F, np.uint32))
        d_pts = hip.put(np.full(max(n, 1) * 4 * 11, 12345, np.uint32))
        rc = D.library().kzg_collide_batch_device(ctx.h, d_sa, d_sb, d_xa, d_xb, n, float(max_sep), d_counts, d_pts)
        counts = hip.get(d_counts, np.zeros(max(n, 1), np.uint32))[:n]
        pts = hip.get(d_pts, np.zeros((max(n, 1), 4), S.CONTACT_DTYPE))[:n]
        return rc, counts, pts
    finally:
        hip.free()


def _host(ctx, sa, sb, xa, xb):
    counts, pts, _ = ctx.collide_batch(sa, sb, xa, xb)
    return counts, pts[:, :4]


def _same(a_counts, a_pts, b_counts, b_pts):
    assert np.array_equal(a_counts, b_counts), np.nonzero(a_counts != b_counts)[0][:10]
    assert a_pts.tobytes() == b_pts.tobytes()


def test_no_alloc_narrowphase_matches_the_allocating_entry_point():
    pb = S.pair_batch(MAX_PAIRS, seed=0x4B5A0109, num_shapes=512)
    ctx = D.DeviceContext(S.world_settings())
    ctx.upload_shapes(pb.shapes, build_hulls(pb.shapes, pb.hull_points))
    # first call: 6,000 mixed pairs (GJK, EPA, clipping paths)
    n1 = 6000
    args1 = (pb.shape_a[:n1], pb.shape_b[:n1], pb.xf_a[:n1], pb.xf_b[:n1])
    rc, c1, p1 = _call(ctx, *args1)
    assert rc == 0, _err(ctx)
    hc1, hp1 = _host(ctx, *args1)
    _same(c1, p1, hc1, hp1)
    assert (c1 == 4).sum() > 100 and (c1 == 1).sum() > 100
    # second call, smaller, other pairs at every index (reversed order)
    sel = np.arange(n1)[::-1][:1500] + 3000
    args2 = (pb.shape_a[sel], pb.shape_b[sel], pb.xf_a[sel], pb.xf_b[sel])
    rc, c2, p2 = _call(ctx, *args2)
    assert rc == 0, _err(ctx)
    hc2, hp2 = _host(ctx, *args2)
    _same(c2, p2, hc2, hp2)
    print(f"no-alloc narrowphase: {n1} then {len(sel)} pairs bit-identical to kzg_collide_batch; "
          f"points/pair {np.bincount(c1, minlength=5).tolist()}", flush=True)


def test_no_alloc_narrowphase_at_the_limit_and_errors():
    pb = S.pair_batch(MAX_PAIRS, seed=0x4B5A010A, num_shapes=256)
    ctx = D.DeviceContext(S.world_settings())
    ctx.upload_shapes(pb.shapes, build_hulls(pb.shapes, pb.hull_points))
    args = (pb.shape_a, pb.shape_b, pb.xf_a, pb.xf_b)
    rc, c, p = _call(ctx, *args)
    assert rc == 0, _err(ctx)
    hc, hp = _host(ctx, *args)
    _same(c, p, hc, hp)
    # one pair more than the limit: rejected before any launch
    big = tuple(np.concatenate([a, a[:1]]) for a in args)
    rc, _, _ = _call(ctx, *big)
    assert rc == ERR_ARGUMENT and "limit" in _err(ctx)
    # a shape index outside the table: that pair reports no contact, the others are exact
    sa = pb.shape_a[:2000].copy()
    sa[7] = len(pb.shapes) + 5
    args = (sa, pb.shape_b[:2000], pb.xf_a[:2000], pb.xf_b[:2000])
    rc, c, p = _call(ctx, *args)
    assert rc == ERR_ARGUMENT and "shape index" in _err(ctx)
    assert c[7] == 0
    keep = np.arange(2000) != 7
    hc, hp = _host(ctx, pb.shape_a[:2000][keep], pb.shape_b[:2000][keep], pb.xf_a[:2000][keep], pb.xf_b[:2000][keep])
    _same(c[keep], p[keep], hc, hp)
