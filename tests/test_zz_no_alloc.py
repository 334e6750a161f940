"""The no-allocation checker (sw_subset_check_noalloc) against the checker
it mirrors (sw_subset_check), called through the C ABI with device
buffers: two calls in a row, the second on smaller lists, so the scratch
the two share must be cleared between them."""
import ctypes as C

import numpy as np
import pytest

import synchrowords_amd as sw
from helpers import random_sets

pytestmark = pytest.mark.gpu


def _noalloc(ctx, n, small, large, thresh, want_rows=True, want_count=True, want_any=True):
    ds = sw.DeviceArray.from_host(ctx, small)
    dl = sw.DeviceArray.from_host(ctx, large)
    rows = sw.DeviceArray(ctx, max(len(small), 1)) if want_rows else None
    cnt = sw.DeviceArray(ctx, 8) if want_count else None
    anyf = sw.DeviceArray(ctx, 4) if want_any else None
    try:
        ctx.subset_check_noalloc(n, ds, len(small), dl, len(large), thresh, rows, cnt, anyf)
        ctx.sync()
        r = rows.to_host(np.uint8, (len(small),)).astype(bool) if want_rows else None
        c = int(cnt.to_host(np.uint64, (1,))[0]) if want_count else None
        a = bool(anyf.to_host(np.int32, (1,))[0]) if want_any else None
        return r, c, a
    finally:
        for d in (ds, dl, rows, cnt, anyf):
            if d is not None:
                d.free()


@pytest.mark.parametrize("n", [100, 300])
def test_noalloc_matches_subset_check_twice(ctx, n):
    rng = np.random.default_rng(77 + n)
    for count_s, count_l in ((4000, 6000), (500, 700)):
        small = random_sets(rng, count_s, n, 0.06)
        large = random_sets(rng, count_l, n, 0.8)
        small[3] = large[5]
        for thresh in (1, 2):
            want_rows, want_cnt, want_any = ctx.subset_check(n, small, large, thresh=thresh, per_row=True)
            rows, cnt, anyf = _noalloc(ctx, n, small, large, thresh)
            np.testing.assert_array_equal(rows, want_rows)
            assert cnt == want_cnt and anyf == want_any
        # the count alone, then existence alone (the early-exit kernel)
        _, cnt, anyf = _noalloc(ctx, n, small, large, 1, want_rows=False)
        assert cnt == want_cnt and anyf == want_any
        _, _, anyf = _noalloc(ctx, n, small, large, 1, want_rows=False, want_count=False)
        assert anyf == want_any
    # no pair at all: empty large sets cannot hold a nonempty small set
    small = random_sets(rng, 300, n, 0.2)
    small[:, 0] |= np.uint64(1)
    large = np.zeros((400, (n + 63) // 64), dtype=np.uint64)
    rows, cnt, anyf = _noalloc(ctx, n, small, large, 1)
    assert not rows.any() and cnt == 0 and not anyf


def test_noalloc_rejects_rows_beyond_scratch(ctx):
    ds = sw.DeviceArray(ctx, 64)
    out = sw.DeviceArray(ctx, 64)
    try:
        rc = sw.load().sw_subset_check_noalloc(ctx.handle, 64, ds.ptr, (1 << 22) + 1, ds.ptr, 1, 1, out.ptr, None,
                                               None)
        assert rc == -1  # SW_ERR_INVALID
        assert b"4194304" in sw.load().sw_last_error()
    finally:
        ds.free()
        out.free()
