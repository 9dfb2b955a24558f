"""CPU model of the cast kernel's wave-spread leaf round (kernels.hip,
flat_leaf_round / keyed_leaf_round): the mapping of a wave's (ray, triangle)
positions to owners and records, and the keyed fold, checked against the
reference's sequential fold `r > 0 && (best < 0 || r < best)`
(kernel.cl:170-185) on random queues, exact ties, infinities and misses."""
from __future__ import annotations

import numpy as np

WAVE = 64


def owners_by_chunk(cnt):
    """Owner of every position, chunk by chunk, as the kernel finds it: the
    lane whose exclusive prefix starts in the chunk marks its start, a
    running maximum carries the mark forward, and a chunk without a start in
    its first lane continues the previous chunk's last owner."""
    incl = np.cumsum(cnt)
    excl = incl - cnt
    total = int(incl[-1])
    out = np.empty(total, np.int64)
    carry = 0
    for base in range(0, total, WAVE):
        mark = np.full(WAVE, -1)
        for lane in range(WAVE):
            if cnt[lane] > 0 and base <= excl[lane] < base + WAVE:
                mark[excl[lane] - base] = lane
        own = np.maximum.accumulate(mark)
        own[own < 0] = carry
        carry = own[-1]
        n = min(WAVE, total - base)
        out[base:base + n] = own[:n]
    return out, excl


def queue_record(off, cnt_q, nq, k):
    """Record of position k in a queue of nq leaves (offsets off, counts
    cnt_q), clamped into the selected leaf (queue_record in kernels.hip)."""
    c = cnt_q[0]
    at = off[0] + min(max(k, 0), c - 1)
    for q in range(1, len(off)):
        if q < nq:
            k -= c
            c = cnt_q[q]
            if k >= 0:
                at = off[q] + min(k, c - 1)
    return at


def seq_fold(best, best_id, rs, ids):
    for r, i in zip(rs, ids):
        if r > 0 and (best < 0 or r < best):
            best, best_id = r, i
    return best, best_id


def keyed_fold(best, best_id, rs, ids):
    keys = [(int(np.float32(r).view(np.uint32)) << 32) | k for k, r in enumerate(rs) if r > 0]
    if not keys:
        return best, best_id
    kv = min(keys)
    r = np.uint32(kv >> 32).view(np.float32)
    if best < 0 or r < best:
        return r, ids[kv & 0xFFFFFFFF]
    return best, best_id


def test_positions_map_to_their_owner_and_record():
    rng = np.random.default_rng(1)
    Q = 4
    for trial in range(300):
        nq = rng.integers(0, Q + 1, WAVE)
        nq[rng.random(WAVE) < 0.3] = 0
        cnt_q = rng.integers(1, 40, (WAVE, Q))
        off = rng.integers(0, 10**6, (WAVE, Q))
        cnt = np.array([cnt_q[l, :nq[l]].sum() for l in range(WAVE)])
        if cnt.sum() == 0:
            continue
        own, excl = owners_by_chunk(cnt)
        for j, o in enumerate(own):
            k = j - excl[o]
            assert 0 <= k < cnt[o], (trial, j)
            at = queue_record(off[o], cnt_q[o], nq[o], k)
            # the k-th record of the owner's queued leaves, in queue order
            flat = np.concatenate([off[o, q] + np.arange(cnt_q[o, q]) for q in range(nq[o])])
            assert at == flat[k]


def test_keyed_fold_equals_sequential_fold():
    rng = np.random.default_rng(2)
    pool = np.array([-1.0, 0.5, 0.25, 0.5, np.inf, 3.0, 1e-30, 2.0], np.float32)
    for trial in range(20000):
        n = int(rng.integers(0, 12))
        rs = rng.choice(pool, n) if trial % 2 else rng.uniform(-1, 4, n).astype(np.float32)
        ids = rng.integers(0, 50, n)
        best = np.float32(rng.choice([-42.0, 0.5, 2.0, np.inf]))
        best_id = -1 if best < 0 else 7
        a = seq_fold(best, best_id, rs, ids)
        b = keyed_fold(best, best_id, rs, ids)
        assert a[0] == b[0] or (np.isinf(a[0]) and np.isinf(b[0])), (rs, ids, best)
        assert a[1] == b[1], (rs, ids, best)


def test_positive_float_bits_order_as_values():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(0, 10, 10000), 10.0 ** rng.uniform(-45, 38, 10000),
                        [np.inf, 1e-45, 3.4e38]]).astype(np.float32)
    x = x[x > 0]
    assert np.array_equal(np.argsort(x, kind="stable"),
                          np.argsort(x.view(np.uint32), kind="stable"))
