"""Host-side checks of the work-item arithmetic the re-sort kernels use
(mls-mpm_amd/csrc/mpm_tiled_dev.h), restated in numpy float32 / integers.

* div_small(r, d) = trunc((float(r) + 0.5) * rcp(d)) must equal r // d for
  every r < 2^13 and 1 <= d <= 16, with the hardware reciprocal anywhere
  within one ulp of 1/d (v_rcp_f32's bound).
* item_slot(r, c) must map the ranks 0..c-1 of an item of c particles
  one-to-one onto its slots (rows of 256, column-major), for every c up to
  4096 (16 particles per thread).
* Size classes: items ordered largest first by rows of 256 cover every item
  exactly once, whatever the brick counts (the scan's dispatch positions).
"""
import numpy as np


def div_small(r, d, ulp=0):
    rc = np.float32(1.0) / np.float32(d)
    if ulp:
        rc = np.nextafter(rc, np.float32(np.inf if ulp > 0 else -np.inf), dtype=np.float32)
    x = (r.astype(np.float32) + np.float32(0.5)) * rc
    assert x.dtype == np.float32
    return x.astype(np.uint32)


def test_div_small_is_floor_division_for_item_ranks():
    r = np.arange(1 << 13, dtype=np.uint32)
    for d in range(1, 17):
        for ulp in (-1, 0, 1):
            np.testing.assert_array_equal(div_small(r, d, ulp), r // d, err_msg=f"d={d} ulp={ulp}")


def item_slot(r, c):
    K = (c + 255) >> 8
    L = c - 256 * (K - 1)
    if r < L * K:
        t = int(div_small(np.array([r], np.uint32), K)[0])
        k = r - t * K
    else:
        r2 = r - L * K
        q = int(div_small(np.array([r2], np.uint32), K - 1)[0])
        t = L + q
        k = r2 - q * (K - 1)
    return k * 256 + t


def test_item_slot_is_a_bijection_onto_the_item():
    for c in list(range(1, 600)) + list(range(600, 4097, 37)) + [1024, 2048, 2047, 4095, 4096]:
        slots = [item_slot(r, c) for r in range(c)]
        assert sorted(slots) == list(range(c)), c
        # consecutive ranks go down a thread's column: rank 1 is the next row
        if c > 256:
            assert slots[0] == 0 and slots[1] == 256


def test_size_classes_cover_every_item_once():
    rng = np.random.default_rng(0)
    for chs in (10, 11, 12):
        ch = 1 << chs
        R = ch // 256
        for trial in range(20):
            counts = rng.integers(0, 3 * ch, 300)
            counts[rng.random(300) < 0.3] = 0
            full = counts >> chs
            rest = counts & (ch - 1)
            cls = np.where(rest > 0, R - (rest + 255) // 256, -1)
            ncls = np.zeros(R, np.int64)
            ncls[0] = full.sum() + (cls == 0).sum()
            for k in range(1, R):
                ncls[k] = (cls == k).sum()
            base = np.concatenate([[0], np.cumsum(ncls)[:-1]])
            cur = base.copy()
            pos = []
            sizes = []
            for b in range(len(counts)):
                for p in range(full[b]):
                    pos.append(cur[0]); cur[0] += 1; sizes.append(ch)
                if rest[b]:
                    k = cls[b]
                    pos.append(cur[k]); cur[k] += 1; sizes.append(rest[b])
            total = int(full.sum() + (rest > 0).sum())
            assert sorted(pos) == list(range(total))
            order = np.empty(total, np.int64)
            order[np.array(pos)] = np.array(sizes)
            rows = (order + 255) // 256
            assert np.all(np.diff(rows) <= 0), "dispatch order is by rows, largest first"
