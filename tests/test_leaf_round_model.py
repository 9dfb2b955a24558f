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


def queue_record(base, start, nq, k, n_entries):
    """Leaf entry of position k of a lane's work (queue_record in
    kernels.hip): slot q holds start[q], the position of its leaf's first
    triangle, and base[q], its first entry minus start[q]; slots from nq on
    hold stale values and do not take part.  The leaf holding k is the
    number of later queued slots starting at or before k; the entry is
    clamped into the scene's leaf lists."""
    q = sum(1 for s in range(1, len(base)) if s < nq and start[s] <= k)
    return min(max(int(base[q]) + k, 0), n_entries - 1)


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
    n_entries = 10**6 + 64
    for Q in (4, 8):
        for trial in range(300):
            nq = rng.integers(0, Q + 1, WAVE)
            nq[rng.random(WAVE) < 0.3] = 0
            cnt_q = rng.integers(1, 40, (WAVE, Q))
            off = rng.integers(0, 10**6, (WAVE, Q))
            # what the lanes wrote when queueing; the slots from nq on keep
            # whatever an earlier round left there
            start = rng.integers(-50, 400, (WAVE, Q))
            base = rng.integers(-10**6, 2 * 10**6, (WAVE, Q))
            for l in range(WAVE):
                s0 = 0
                for q in range(nq[l]):
                    start[l, q] = s0
                    base[l, q] = off[l, q] - s0
                    s0 += cnt_q[l, q]
            cnt = np.array([cnt_q[l, :nq[l]].sum() for l in range(WAVE)])
            if cnt.sum() == 0:
                continue
            own, excl = owners_by_chunk(cnt)
            for j, o in enumerate(own):
                k = j - excl[o]
                assert 0 <= k < cnt[o], (trial, j)
                at = queue_record(base[o], start[o], nq[o], k, n_entries)
                # the k-th record of the owner's queued leaves, in queue order
                flat = np.concatenate([off[o, q] + np.arange(cnt_q[o, q]) for q in range(nq[o])])
                assert at == flat[k]


def test_queue_record_stays_in_the_leaf_lists():
    """Whatever the queue holds, the entry read is inside the leaf lists."""
    rng = np.random.default_rng(4)
    for _ in range(20000):
        Q = 8
        base = rng.integers(-2**31, 2**31 - 2**20, Q)
        start = rng.integers(-2**20, 2**20, Q)
        nq = int(rng.integers(0, Q + 1))
        k = int(rng.integers(0, 2**20))
        n_entries = int(rng.integers(1, 2**30))
        assert 0 <= queue_record(base, start, nq, k, n_entries) < n_entries


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


def dpp_fold(r, ids, seg_start):
    """The leaf round's segmented fold (wave_seg_fold in kernels.hip) lane
    by lane: steps 1, 2, 4, 8 take lane - step inside each 16-lane row,
    then rows 1 and 3 take the last lane of the row before, then lanes
    32..63 take lane 31; a lane takes a source only if the source lies in
    its segment (at or after seg_start).  Every lane reads the values the
    step started with."""
    r = np.array(r, np.float32)
    ids = np.array(ids)
    lanes = np.arange(WAVE)

    def step(src, take):
        nr, ni = r.copy(), ids.copy()
        for l in lanes:
            if take[l]:
                ru, iu = r[src[l]], ids[src[l]]
                if ru > 0 and not (r[l] > 0 and r[l] < ru):
                    nr[l], ni[l] = ru, iu
        return nr, ni

    rel = lanes - seg_start
    for s in (1, 2, 4, 8):
        r, ids = step(np.maximum(lanes - s, 0), ((lanes & 15) >= s) & (rel >= s))
    src16 = (lanes & ~15) - 1
    r, ids = step(np.maximum(src16, 0), ((lanes & 31) >= 16) & (src16 >= seg_start))
    r, ids = step(np.full(WAVE, 31), (lanes >= 32) & (31 >= seg_start))
    return r, ids


def test_segmented_fold_is_the_sequential_fold_per_segment():
    """Each lane ends with the sequential fold of its segment from the
    segment's start to itself, for random segmentations of a chunk (owners'
    runs of positions), with misses, ties and infinities."""
    rng = np.random.default_rng(5)
    pool = np.array([-1.0, 0.5, 0.25, 0.5, np.inf, 3.0, 2.0], np.float32)
    for trial in range(400):
        cuts = np.sort(rng.choice(np.arange(1, WAVE), int(rng.integers(0, 20)), replace=False))
        starts = np.concatenate([[0], cuts])
        seg_start = np.zeros(WAVE, np.int64)
        for a in starts:
            seg_start[a:] = a
        r = rng.choice(pool, WAVE) if trial % 2 else rng.uniform(-1, 4, WAVE).astype(np.float32)
        ids = rng.integers(0, 1000, WAVE)
        fr, fi = dpp_fold(r, ids, seg_start)
        for l in range(WAVE):
            b, bi = seq_fold(np.float32(-42.0), -1, r[seg_start[l]:l + 1], ids[seg_start[l]:l + 1])
            if bi == -1:     # nothing positive: the lane keeps its own entry
                assert not (fr[l] > 0), (trial, l)
                continue
            assert fr[l] == b and fi[l] == bi, (trial, l)


MAILBOX = 8


def mailbox_rounds(work, dist, rays_of, clear_on_new_ray=True, mark_before_test=False):
    """The leaf rounds of the buffer forms with the per-ray mailbox
    (flat_leaf_round in kernels.hip): `work[r]` is round r's list of
    (lane, record) positions in position order, `rays_of[r][lane]` the ray
    the lane holds in round r and `dist(ray, record)` the test's distance.
    Chunk by chunk, every position first reads its owner's mailbox slot
    record & 7; positions whose slot holds their record are skipped, the
    others are tested and then write their record there.  The mailbox is
    cleared when a lane takes a new ray.  Returns each ray's (best, id) and
    the number of skipped tests."""
    mb = {}
    best = {}
    skipped = 0
    held = {}
    for r, positions in enumerate(work):
        for lane, ray in rays_of[r].items():
            if held.get(lane) != ray:
                held[lane] = ray
                if clear_on_new_ray:
                    mb[lane] = [-1] * MAILBOX
                else:
                    mb.setdefault(lane, [-1] * MAILBOX)
                best[ray] = (-42.0, -1)
        for base in range(0, len(positions), WAVE):
            chunk = positions[base:base + WAVE]
            if mark_before_test:  # a planted mistake: slots written before they are read
                for lane, k in chunk:
                    mb[lane][k % MAILBOX] = k
            seen = [mb[lane][k % MAILBOX] == k for lane, k in chunk]
            rs, ids, owners = [], [], []
            for (lane, k), skip in zip(chunk, seen):
                ray = rays_of[r][lane]
                owners.append(ray)
                if skip:
                    skipped += 1
                    rs.append(-1.0)
                else:
                    rs.append(dist(ray, k))
                    mb[lane][k % MAILBOX] = k
                ids.append(k)
            for ray in dict.fromkeys(owners):
                sel = [i for i, o in enumerate(owners) if o == ray]
                b, bi = seq_fold(*best[ray], [rs[i] for i in sel], [ids[i] for i in sel])
                best[ray] = (b, bi)
    return best, skipped


def test_mailbox_skips_only_repeats_of_the_same_ray():
    """Random rounds of a wave: lanes keep rays for several rounds and then
    take new ones, each ray's records repeat often (a triangle listed in
    many leaves), distances are drawn from a few values so that different
    records tie exactly, and some are misses (<= 0) or infinite.  With the
    mailbox, every ray's (best, id) equals the reference's sequential fold
    over all its tests; many tests are skipped.  Planted mistakes fail: a
    mailbox kept across rays skips a new ray's first tests, and one written
    before it is read skips first copies."""
    rng = np.random.default_rng(23)
    n_rounds, lanes = 40, 64
    values = np.float32([0.5, 0.75, 1.0, 1.0, 2.5, -1.0, 0.0, np.inf])
    table = {}

    def dist(ray, k):
        if (ray, k) not in table:
            table[(ray, k)] = float(values[rng.integers(0, len(values))])
        return table[(ray, k)]

    rays_of, work, all_tests = [], [], {}
    next_ray, cur = 0, {}
    for r in range(n_rounds):
        for lane in range(lanes):
            if lane not in cur or rng.random() < 0.15:
                cur[lane] = next_ray
                next_ray += 1
        rays_of.append(dict(cur))
        positions = []
        for lane in range(lanes):
            if rng.random() < 0.3:
                continue
            pool = rng.integers(0, 12, 3)  # a few records a ray meets again and again
            n = int(rng.integers(1, 40))
            recs = np.where(rng.random(n) < 0.6, rng.choice(pool, n), rng.integers(0, 5000, n))
            for k in recs:
                positions.append((lane, int(k)))
                all_tests.setdefault(cur[lane], []).append(int(k))
        work.append(positions)
    got, skipped = mailbox_rounds(work, dist, rays_of)
    for ray, recs in all_tests.items():
        want = seq_fold(-42.0, -1, [dist(ray, k) for k in recs], recs)
        assert got[ray] == want, (ray, got[ray], want)
    assert skipped > 0.05 * sum(len(v) for v in all_tests.values())
    kept, _ = mailbox_rounds(work, dist, rays_of, clear_on_new_ray=False)
    assert any(kept[ray] != seq_fold(-42.0, -1, [dist(ray, k) for k in recs], recs)
               for ray, recs in all_tests.items())
    early, _ = mailbox_rounds(work, dist, rays_of, mark_before_test=True)
    assert any(early[ray] != seq_fold(-42.0, -1, [dist(ray, k) for k in recs], recs)
               for ray, recs in all_tests.items())
