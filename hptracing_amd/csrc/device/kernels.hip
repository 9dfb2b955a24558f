// HIP kernels of the wavefront path tracer for gfx950 (CDNA4, wave64).
//
//   k_cast     KD traversal + triangle intersection   (kdtree_intersect, kernel.cl:111-208)
//   k_camera   pinhole / equirectangular camera rays  (set_viewdirs[_vr], kernel.cl:559-613)
//   k_primary  primary-hit cache and background       (kernel.cl:187-207, trace_runner.cpp:192-213)
//   k_expand   per-pass restore of the cached hits    (trace_runner.cpp:223-229)
//   k_shade    emission, interpolation, texture, branch choice and the four
//              bounce rules fused, with stream compaction of the survivors
//              (s1_run + s2_*_run, kernel.cl:258-557)
//   k_resolve  box filter, tone map, UNORM8           (get_image, kernel.cl:615-652)
//
// The cast and shade kernels are persistent: a grid sized to the machine
// pulls work until the device-side live count is exhausted, so the depth
// loop needs no host round trip and can be captured in a hipGraph.
#include "dmath.hpp"
#include "kernels.hpp"
#include "launch.hpp"

namespace hpt {

// ------------------------------------------------------------------ cast
// One ray of kdtree_intersect.  The node walk is the reference's stackless
// pre-order walk expressed with escape indices: a box hit on an interior node
// descends to idx + 1, anything else continues at the node's escape index
// (idx + 1 for a leaf).  Leaves are tested the moment they are reached,
// which is the order the reference tests its collected list in; the walk
// stops after the 512th non-empty leaf because the reference collects no
// more (kernel.cl:146) and so can change nothing after it.
// FAST forms the slab quotients through the ray's approximate reciprocal
// direction and certifies each decision against the rounding error, falling
// back to the division when it cannot (box_hit_rcp): same decisions, almost
// no divisions.
template <bool COUNT, bool FAST>
__device__ __forceinline__ int trace_ray(const DevScene &sc, V3 sp, V3 d, float &t_out,
                                         int *cnt) {
    int idx = 0, leaves = 0, best_id = -1;
    float best = -42.f;
    int visits = 0, tested = 0, dropped = 0, visits_cap = -1;
    const int nn = sc.n_nodes;
    const bool rcp = FAST && sc.boxes_fast && fast_box_ray(sp, d);
    const V3 inv = approx_rcp(d);
    while (idx < nn) {
        const float4 lo = sc.nodes[2 * idx];
        const float4 hi = sc.nodes[2 * idx + 1];
        const int a = __float_as_int(lo.w);
        const int b = __float_as_int(hi.w);
        if (COUNT) ++visits;
        const bool hit = rcp ? box_hit_rcp(xyz(lo), xyz(hi), sp, d, inv)
                             : box_hit(xyz(lo), xyz(hi), sp, d);
        if (b < 0) {
            idx = hit ? idx + 1 : a;
            continue;
        }
        if (hit && a > 0) {
            if (!COUNT || leaves < kMaxMatchData) {
                const float4 *rec = sc.tris + 3 * (int64_t)b;
                for (int k = 0; k < a; ++k) {
                    const float4 r0 = rec[3 * k], r1 = rec[3 * k + 1], r2 = rec[3 * k + 2];
                    const float r = sc.arith_ocl ? tri_hit_ocl(sp, d, xyz(r0), xyz(r1), xyz(r2))
                                              : tri_hit(sp, d, xyz(r0), xyz(r1), xyz(r2));
                    if (r > 0 && (best < 0 || r < best)) {
                        best = r;
                        best_id = __float_as_int(r0.w);
                    }
                }
                if (COUNT) tested += a;
                ++leaves;
                if (leaves == kMaxMatchData) {
                    if (!COUNT) break;
                    visits_cap = visits;
                }
            } else {
                ++dropped;
            }
        }
        ++idx;
    }
    if (COUNT) {
        cnt[0] = visits_cap >= 0 ? visits_cap : visits;
        cnt[1] = leaves;
        cnt[2] = tested;
        cnt[3] = dropped;
        cnt[4] = visits;
    }
    t_out = best;
    return best_id;
}

// The same search organised as alternating phases for the whole wave: every
// lane walks nodes until it reaches a non-empty leaf (or the end of the
// tree), then all lanes test their leaf's triangles together.  A lane tests
// its leaves in the order it reaches them, as above, so the result is the
// same; the triangle loop runs once per leaf round instead of on almost
// every node step of some lane.
template <bool FAST>
__device__ __forceinline__ int trace_ray_ww(const DevScene &sc, V3 sp, V3 d, float &t_out) {
    int idx = 0, leaves = 0, best_id = -1;
    float best = -42.f;
    const int nn = sc.n_nodes;
    const bool rcp = FAST && sc.boxes_fast && fast_box_ray(sp, d);
    const V3 inv = approx_rcp(d);
    bool done = nn <= 0;
    int pcnt = 0, poff = 0;
    for (;;) {
        while (!done && pcnt == 0) {
            const float4 lo = sc.nodes[2 * idx];
            const float4 hi = sc.nodes[2 * idx + 1];
            const int a = __float_as_int(lo.w);
            const int b = __float_as_int(hi.w);
            const bool hit = rcp ? box_hit_rcp(xyz(lo), xyz(hi), sp, d, inv)
                                 : box_hit(xyz(lo), xyz(hi), sp, d);
            if (b < 0) {
                idx = hit ? idx + 1 : a;
            } else {
                if (hit && a > 0) {
                    pcnt = a;
                    poff = b;
                }
                ++idx;
            }
            if (idx >= nn) done = true;
        }
        if (pcnt == 0) break;
        const float4 *rec = sc.tris + 3 * (int64_t)poff;
        for (int k = 0; k < pcnt; ++k) {
            const float4 r0 = rec[3 * k], r1 = rec[3 * k + 1], r2 = rec[3 * k + 2];
            const float r = sc.arith_ocl ? tri_hit_ocl(sp, d, xyz(r0), xyz(r1), xyz(r2))
                                              : tri_hit(sp, d, xyz(r0), xyz(r1), xyz(r2));
            if (r > 0 && (best < 0 || r < best)) {
                best = r;
                best_id = __float_as_int(r0.w);
            }
        }
        pcnt = 0;
        if (++leaves == kMaxMatchData) done = true;
    }
    t_out = best;
    return best_id;
}

__device__ __forceinline__ void wave_done(WorkCounter *wc) {
    if ((threadIdx.x & 63) == 0) {
        const int total_waves = (int)(gridDim.x * (blockDim.x / 64));
        __threadfence();
        const int prev = atomicAdd(&wc->done, 1);
        if (prev == total_waves - 1) {
            atomicExch(&wc->next, 0);
            atomicExch(&wc->done, 0);
        }
    }
}

// Wave-wide inclusive scans over lanes 0..63 on DPP moves (VALU, no LDS
// round trip): row_shr 1, 2, 4, 8 inside each 16-lane row, then the last
// lane of each even row into the next row (row_bcast:15), then lane 31 into
// the upper half (row_bcast:31).  A lane combines only where its source lane
// exists, so each step extends its range by the range just before it.
template <int CTRL>
__device__ __forceinline__ int dpp_mov(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xf, 0xf, false);
}
__device__ __forceinline__ bool dpp_step_valid(int step, int lane) {
    return step < 16 ? (lane & 15) >= step : (step == 16 ? (lane & 31) >= 16 : lane >= 32);
}
template <class Op>
__device__ __forceinline__ int wave_scan(int v, int lane, Op op) {
    int t;
    t = dpp_mov<0x111>(v); if (dpp_step_valid(1, lane)) v = op(t, v);
    t = dpp_mov<0x112>(v); if (dpp_step_valid(2, lane)) v = op(t, v);
    t = dpp_mov<0x114>(v); if (dpp_step_valid(4, lane)) v = op(t, v);
    t = dpp_mov<0x118>(v); if (dpp_step_valid(8, lane)) v = op(t, v);
    t = dpp_mov<0x142>(v); if (dpp_step_valid(16, lane)) v = op(t, v);
    t = dpp_mov<0x143>(v); if (dpp_step_valid(32, lane)) v = op(t, v);
    return v;
}
__device__ __forceinline__ int wave_scan_add(int v, int lane) {
    return wave_scan(v, lane, [](int a, int b) { return a + b; });
}
__device__ __forceinline__ int wave_scan_max(int v, int lane) {
    return wave_scan(v, lane, [](int a, int b) { return a > b ? a : b; });
}

// Segmented fold of (r, id) pairs in lane order.  A lane's segment is the
// run of lanes from `seg_start` up to it (positions of one owner are
// contiguous).  The operator keeps the earlier entry unless only the later
// has r > 0 or the later's r is smaller -- the reference's
// `r > 0 && (best < 0 || r < best)` applied in order -- and is associative,
// so the scan (steps 1, 2, 4, 8 inside each 16-lane row, then the last lane
// of rows 0 and 2 into rows 1 and 3, then lane 31 into the upper half)
// leaves each lane with the sequential fold from its segment's start to
// itself: a lane takes a step's source only where the source lies in its
// segment, and the source's range then ends where the lane's begins.
template <int CTRL>
__device__ __forceinline__ void fold_step(bool take, float &r, int &id) {
    const float r_up = __int_as_float(dpp_mov<CTRL>(__float_as_int(r)));
    const int id_up = dpp_mov<CTRL>(id);
    if (take && r_up > 0 && !(r > 0 && r < r_up)) {
        r = r_up;
        id = id_up;
    }
}
__device__ __forceinline__ void wave_seg_fold(int lane, int seg_start, float &r, int &id) {
    const int rel = lane - seg_start;  // lanes of the segment before this one
    fold_step<0x111>((lane & 15) >= 1 && rel >= 1, r, id);
    fold_step<0x112>((lane & 15) >= 2 && rel >= 2, r, id);
    fold_step<0x114>((lane & 15) >= 4 && rel >= 4, r, id);
    fold_step<0x118>((lane & 15) >= 8 && rel >= 8, r, id);
    fold_step<0x142>((lane & 31) >= 16 && (lane & ~15) - 1 >= seg_start, r, id);
    fold_step<0x143>(lane >= 32 && 31 >= seg_start, r, id);
}

// Per-lane queue of pending leaves in LDS (FLAT form): up to Q entries per
// lane, lane-major so that a wave's accesses to one slot fall in distinct
// banks.  A lane's work is the triangles of its queued leaves in queue
// order; for the leaf in slot q, start[q] is the position of its first
// triangle in that work (the counts of the leaves before it) and base[q]
// its first leaf entry minus start[q], so that position k of the lane's
// work is leaf entry base[q] + k for the last q with start[q] <= k.
template <int Q, int ROWS = Q + 1>
struct LeafQueue {
    // With ROWS = Q + 1, slot Q takes the writes of lanes that queue nothing
    // (the single-node reciprocal visit); the window forms need Q rows.
    int base[ROWS][kBlock];
    int start[ROWS][kBlock];
};

// The record of leaf entry `at`: in leaf order (direct form), or through
// the entry's record number in the per-triangle records (IND, indirect
// form); both hold the same values, the triangle's index in w.
template <bool IND>
__device__ __forceinline__ const float4 *tri_record(const DevScene &sc, int at) {
    return IND ? sc.tri_recs + 3 * (int64_t)sc.leaf_ids[at] : sc.tris + 3 * (int64_t)at;
}

// Buffer loads: a resource over `bytes` bytes from p (under 2 GiB) takes
// 32-bit byte offsets, with no 64-bit address arithmetic per load, and an
// offset at or past the end reads zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_over(const void *p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, 0x00020000);
}

__device__ __forceinline__ float4 buffer_float4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    return make_float4(__int_as_float(v[0]), __int_as_float(v[1]), __int_as_float(v[2]),
                       __int_as_float(v[3]));
}

// Leaf entry of position k (0 <= k < the lane's count) of the work of the
// lane whose queue is in column `col`, with `nq` leaves queued: the slots
// from nq on hold stale or never-written values and do not take part.  The
// starts of the queued leaves increase (every queued leaf has triangles),
// so the leaf holding k is the number of later slots starting at or before
// k.  The entry is clamped into the scene's leaf lists, so that no error in
// the bookkeeping can read outside them.
template <int Q, int R>
__device__ __forceinline__ int queue_record(const LeafQueue<Q, R> &lq, int col, int nq, int k,
                                            int n_entries) {
    int q = 0;
#pragma unroll
    for (int s = 1; s < Q; ++s) q += (s < nq && lq.start[s][col] <= k) ? 1 : 0;
    const int at = lq.base[q][col] + k;
    return min(max(at, 0), n_entries - 1);
}

// Each lane's mailbox of the record numbers its ray has tested (IND and BUF
// forms): slot k & 7 holds the last record number k tested there, and -1
// when the lane takes a ray.  A test of a record its ray has already tested
// gives the same distance again and cannot change the result under the
// strict `r < best` fold (the earlier test set best <= r, or lost), so it
// is skipped: no record loads, no test.  Skipping only later copies keeps
// the first smallest distance, so ties come out as the reference's.
constexpr int kMailbox = 8;

// One leaf round of the FLAT form (see cast_refill).  A lane's pending work
// is the triangles of its `nq` queued leaves in queue order, `cnt` in all
// (0 and 0 if none); best / best_id are folded in place.  All 64 lanes must
// call it.  `mb` is the block's mailboxes (IND and BUF forms).
template <int Q, bool IND, bool BUF, int R>
__device__ __forceinline__ void flat_leaf_round(const DevScene &sc, const LeafQueue<Q, R> &lq,
                                                int lane, int nq, int cnt, V3 sp, V3 d,
                                                float &best, int &best_id,
                                                int (*mb)[kBlock]) {
    __shared__ int s_mark[kBlock];
    const int w = threadIdx.x & ~63;
    const int incl = wave_scan_add(cnt, lane);
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - cnt;
    int carry = 0;  // owner of the position just before the current chunk
    for (int base = 0; base < total; base += 64) {
        // Owner of each position of the chunk: segment starts are marked
        // with the owning lane and carried forward by a running maximum
        // (owners are in lane order, their segments in position order).
        s_mark[w + lane] = -1;
        __builtin_amdgcn_wave_barrier();
        if (cnt > 0 && excl >= base && excl < base + 64) s_mark[w + excl - base] = lane;
        __builtin_amdgcn_wave_barrier();
        int own = wave_scan_max(s_mark[w + lane], lane);
        if (own < 0) own = carry;
        carry = __shfl(own, 63, 64);
        const int o_excl = __shfl(excl, own, 64);
        const int o_nq = Q > 1 ? __shfl(nq, own, 64) : 1;
        const V3 o_sp = mk(__shfl(sp.x, own, 64), __shfl(sp.y, own, 64), __shfl(sp.z, own, 64));
        const V3 o_d = mk(__shfl(d.x, own, 64), __shfl(d.y, own, 64), __shfl(d.z, own, 64));
        const int j = base + lane;
        float r = -1.f;
        int id = -1;
        if (j < total) {
            // Position j - o_excl of the owner's work (0 <= it < the owner's
            // count, the sum of the counts of its o_nq queued leaves).
            const int at = queue_record<Q>(lq, w + own, o_nq, j - o_excl, sc.n_entries);
            if constexpr (IND && BUF) {
                // The owner's mailbox is read by every position of the chunk
                // before any of them writes it, so a repeat inside one chunk
                // is tested again (harmless); repeats of earlier chunks and
                // rounds are skipped.
                const int k = __builtin_amdgcn_raw_buffer_load_b32(
                    buffer_over(sc.leaf_ids, (uint32_t)sc.n_entries * 4u),
                    (int)((uint32_t)at << 2), 0, 0);
                int &slot = mb[k & (kMailbox - 1)][w + own];
                if (slot != k) {
                    const __amdgpu_buffer_rsrc_t recs =
                        buffer_over(sc.tri_recs, (uint32_t)sc.n_recs * 48u);
                    const uint32_t off = (uint32_t)k * 48u;
                    const float4 r0 = buffer_float4(recs, off);
                    const float4 r1 = buffer_float4(recs, off + 16u);
                    const float4 r2 = buffer_float4(recs, off + 32u);
                    r = sc.arith_ocl ? tri_hit_ocl(o_sp, o_d, xyz(r0), xyz(r1), xyz(r2))
                                     : tri_hit(o_sp, o_d, xyz(r0), xyz(r1), xyz(r2));
                    id = __float_as_int(r0.w);
                    slot = k;
                }
            } else {
                const float4 *rec = tri_record<IND>(sc, at);
                const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
                r = sc.arith_ocl ? tri_hit_ocl(o_sp, o_d, xyz(r0), xyz(r1), xyz(r2))
                                 : tri_hit(o_sp, o_d, xyz(r0), xyz(r1), xyz(r2));
                id = __float_as_int(r0.w);
            }
        }
        // Fold the chunk per owner: a segment is one owner's run of
        // positions, starting at its first position or at the chunk start.
        // (Every lane executes each cross-lane move: a lane that skipped one
        // would supply no value to its neighbour.)
        wave_seg_fold(lane, max(o_excl - base, 0), r, id);
        // Each owner with positions in this chunk takes its segment's fold,
        // which follows everything it has folded before.
        const int last = (excl + cnt < base + 64 ? excl + cnt : base + 64) - 1 - base;
        const bool mine = cnt > 0 && excl < base + 64 && excl + cnt > base;
        const int src = mine ? last : lane;
        const float rr = __shfl(r, src, 64);
        const int rid = __shfl(id, src, 64);
        if (mine && rr > 0 && (best < 0 || rr < best)) {
            best = rr;
            best_id = rid;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// Rays of a block's lanes in LDS, written when a lane takes a ray, so that
// a tester reads its owner's ray with two loads instead of six shuffles.
struct RaySlots {
    float4 org[kBlock];
    float4 dir[kBlock];
};

// The leaf round with a keyed fold (KEYED form).  The reference keeps, of a
// ray's tests in order, the first with the smallest r > 0.  For r > 0
// (+inf included; NaN never passes r > 0) the order of the float bits is
// the order of the values, so the key (bits of r) << 32 | position is
// smallest exactly for that test: every tester with r > 0 takes the 64-bit
// minimum of its key and the owner's LDS slot, and after the round each
// owner reads the winner of all its positions at once and combines it with
// its running best by the reference's rule (the round's tests come after
// every earlier one, so a tie keeps the earlier best).  The winner's id is
// read back from its record.  Same contract as flat_leaf_round.
template <int Q, bool IND>
__device__ __forceinline__ void keyed_leaf_round(const DevScene &sc, const LeafQueue<Q> &lq,
                                                 const RaySlots &rs,
                                                 unsigned long long *key, int lane, int nq,
                                                 int cnt, float &best, int &best_id) {
    __shared__ int s_mark[kBlock];
    const int w = threadIdx.x & ~63;
    const int incl = wave_scan_add(cnt, lane);
    const int total = __shfl(incl, 63, 64);
    const int excl = incl - cnt;
    key[threadIdx.x] = ~0ULL;
    int carry = 0;
    for (int base = 0; base < total; base += 64) {
        s_mark[w + lane] = -1;
        __builtin_amdgcn_wave_barrier();
        if (cnt > 0 && excl >= base && excl < base + 64) s_mark[w + excl - base] = lane;
        __builtin_amdgcn_wave_barrier();
        int own = wave_scan_max(s_mark[w + lane], lane);
        if (own < 0) own = carry;
        carry = __shfl(own, 63, 64);
        const int o_excl = __shfl(excl, own, 64);
        const int o_nq = Q > 1 ? __shfl(nq, own, 64) : 1;
        const int j = base + lane;
        if (j < total) {
            const int k = j - o_excl;
            const int at = queue_record<Q>(lq, w + own, o_nq, k, sc.n_entries);
            const float4 o4 = rs.org[w + own], d4 = rs.dir[w + own];
            const float4 *rec = tri_record<IND>(sc, at);
            const float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
            const float r = sc.arith_ocl ? tri_hit_ocl(xyz(o4), xyz(d4), xyz(r0), xyz(r1), xyz(r2))
                                         : tri_hit(xyz(o4), xyz(d4), xyz(r0), xyz(r1), xyz(r2));
            if (r > 0.f)
                atomicMin(key + w + own,
                          ((unsigned long long)__float_as_uint(r) << 32) | (unsigned)k);
        }
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_wave_barrier();
    if (cnt > 0) {
        const unsigned long long kv = key[threadIdx.x];
        if (kv != ~0ULL) {
            const float r = __uint_as_float((unsigned)(kv >> 32));
            if (best < 0 || r < best) {
                best = r;
                const int at = queue_record<Q>(lq, threadIdx.x, nq, (int)(unsigned)kv,
                                               sc.n_entries);
                best_id = __float_as_int(tri_record<IND>(sc, at)[0].w);
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
}

// The search with per-lane refill.  Per-ray costs differ by orders of
// magnitude (a ray skimming the water grid crosses thousands of nodes), so a
// wave that waits for its slowest lane idles most of its lanes.  Here every
// lane carries its own ray through a sequence of rounds -- at most
// `node_steps` node visits until it holds a non-empty leaf, then that
// leaf's triangles -- and a lane whose ray is finished takes the next ray
// from the shared counter as soon as `refill` lanes of the wave are idle.
// A round's node phase also ends once fewer than `walk_min` lanes can walk.
// Each ray's leaves are still tested in the order its walk reaches them, so
// results are those of trace_ray.
//
// Refinements, none of which changes a result:
//  WIN    nodes are fetched WIN at a time (idx .. idx + WIN - 1, contiguous;
//         WIN = 2 is the node pair).  After a leaf, and after an interior
//         node whose box is hit, the walk moves to idx + 1, so the next
//         nodes are usually the next ones visited and their box tests need
//         no further memory trip; only a missed interior node (a jump to its
//         escape) waits, unless its escape still lies in the window.
//  PREF   a leaf's triangle records are fetched one triangle ahead, so the
//         loads of the next record are in flight during the current test.
//  Q > 0  FLAT form: the leaf phase is spread over the whole wave.  A lane
//         walks until it has queued Q non-empty leaves (in LDS) or its walk
//         ends; then the (ray, triangle) pairs of every lane's queue are
//         numbered by a wave prefix sum and tested 64 at a time, one pair
//         per lane, each tester reading its ray from the owning lane; a
//         segmented wave scan folds each owner's distances in position
//         order, which is its queue order and the reference's order, so
//         best and ties come out the same.  A leaf
//         round costs ceil(sum of counts / 64) tests instead of the
//         largest count, and lanes keep walking while their queue fills.
//  KEYED  (FLAT only) the leaf round folds through 64-bit keys in LDS
//         (keyed_leaf_round) and testers read rays from LDS.
//  IND    (FLAT only) leaf rounds read the records once per triangle
//         through the leaf entries' triangle ids (tri_record).
//  LINE   the window ends at the end of the 128-byte line (four nodes) that
//         holds idx: a trip reads only nodes of one line, so it touches one
//         line and one or two 64-byte sectors instead of up to two lines.
//  BUF    nodes, and (IND) the leaf entries' record numbers and the
//         records, are read with buffer loads: 32-bit byte offsets (no
//         64-bit address arithmetic) and a range check that returns zeros
//         past the last node (such window slots are never visited), so the
//         window needs no clamp.  Each array must be under 2 GiB (fetch_fits).
//         With FLAT and IND, leaf rounds skip records the ray has already
//         tested (the mailbox, kMailbox).
template <bool PATHS, bool FAST, int WIN, bool PREF, int Q, bool KEYED, bool IND, bool LINE,
          bool BUF>
__device__ __forceinline__ void cast_refill(const DevScene &sc, const float *org3,
                                            const float *dir3, const float4 *org4,
                                            const float4 *dir4, int64_t n, WorkCounter *wc,
                                            int *out_id, float *out_t) {
    constexpr bool FLAT = Q > 0;
    constexpr int QN = Q > 0 ? Q : 1;
    // The window forms (WIN > 1) never use the reciprocal single-node
    // visit, so their queue has no spare row.
    constexpr int kSpare = WIN < 2 ? QN : 0;
    __shared__ LeafQueue<QN, WIN < 2 ? QN + 1 : QN> lq;
    __shared__ int s_mb[(FLAT && IND && BUF) ? kMailbox : 1][kBlock];
    __shared__ RaySlots rs;
    __shared__ unsigned long long s_key[kBlock];
    const int lane = threadIdx.x & 63;
    const int tid = threadIdx.x;
    const int nn = sc.n_nodes;
    const int threshold = sc.refill < 1 ? 1 : (sc.refill > 64 ? 64 : sc.refill);
    const int node_steps = sc.node_steps < 1 ? 1 : sc.node_steps;
    const int walk_min = sc.walk_min;
    int64_t ray = -1;
    V3 sp = mk(0.f, 0.f, 0.f), d = sp, inv = sp;
    bool rcp = false, done = true;
    // Node q's two halves (box_start + a, box_end + b).
    auto fetch = [&](int q, float4 &lo, float4 &hi) {
        if constexpr (BUF) {
            const __amdgpu_buffer_rsrc_t nodes = buffer_over(sc.nodes, (uint32_t)nn * 32u);
            const uint32_t off = (uint32_t)q << 5;
            lo = buffer_float4(nodes, off);
            hi = buffer_float4(nodes, off + 16u);
        } else {
            lo = sc.nodes[2 * q];
            hi = sc.nodes[2 * q + 1];
        }
    };
    // pending: leaves held for the next leaf round (FLAT: queued in LDS,
    // pcnt = their triangle total; otherwise pcnt / poff of the one leaf).
    int idx = 0, leaves = 0, best_id = -1, pcnt = 0, poff = 0, pending = 0;
    float best = -42.f;
    bool exhausted = false;
    // One node visit: box test, then the next index (idx + 1 after a hit
    // interior node or any leaf, the escape after a missed interior node),
    // and a pending leaf when a non-empty leaf's box is hit.  The reference
    // collects no leaf after its 512th, so the walk ends there.
    auto visit = [&](float4 lo, float4 hi) {
        const int a = __float_as_int(lo.w);
        const int b = __float_as_int(hi.w);
        const bool hit = rcp ? box_hit_rcp(xyz(lo), xyz(hi), sp, d, inv)
                             : box_hit(xyz(lo), xyz(hi), sp, d);
        if (b < 0) {
            idx = hit ? idx + 1 : a;
        } else {
            if (hit && a > 0) {
                if (FLAT) {
                    lq.base[pending][tid] = b - pcnt;
                    lq.start[pending][tid] = pcnt;
                    pcnt += a;
                } else {
                    pcnt = a;
                    poff = b;
                }
                ++pending;
                if (leaves + pending == kMaxMatchData) done = true;
            }
            ++idx;
        }
        if (idx >= nn) done = true;
    };
    // The same visit for a wave whose lanes all use the reciprocal slab test
    // (FLAT form), written without lane branches: the queue write always
    // happens, into slot QN when nothing is queued, and the rare
    // uncertified box test is decided by the division form under a
    // wave-uniform check.
    auto visit_rcp = [&](float4 lo, float4 hi) {
        const int a = __float_as_int(lo.w);
        const int b = __float_as_int(hi.w);
        bool open;
        bool hit = box_hit_rcp_part(xyz(lo), xyz(hi), sp, inv, open);
        if (__ballot(open)) {
            if (open) hit = box_hit(xyz(lo), xyz(hi), sp, d);
        }
        const bool leaf = b >= 0;
        const bool take = leaf && hit && a > 0;
        idx = (!leaf && !hit) ? a : idx + 1;
        const int slot = take ? pending : kSpare;
        lq.base[slot][tid] = b - pcnt;
        lq.start[slot][tid] = pcnt;
        pcnt += take ? a : 0;
        pending += take ? 1 : 0;
        done = done || (take && leaves + pending == kMaxMatchData) || idx >= nn;
    };
    auto test = [&](float4 r0, float4 r1, float4 r2) {
        const float r = sc.arith_ocl ? tri_hit_ocl(sp, d, xyz(r0), xyz(r1), xyz(r2))
                                     : tri_hit(sp, d, xyz(r0), xyz(r1), xyz(r2));
        if (r > 0 && (best < 0 || r < best)) {
            best = r;
            best_id = __float_as_int(r0.w);
        }
    };
    for (;;) {
        const bool idle = ray < 0;
        const uint64_t m = __ballot(idle);
        const int n_idle = __popcll(m);
        if (!exhausted && n_idle > 0 && (n_idle >= threshold || n_idle == 64)) {
            int base = 0;
            if (lane == 0) base = atomicAdd(&wc->next, n_idle);
            base = __shfl(base, 0, 64);
            if ((int64_t)base + n_idle >= n) exhausted = true;
            if (idle) {
                const int64_t my = (int64_t)base + __popcll(m & ((1ULL << lane) - 1ULL));
                if (my < n) {
                    ray = my;
                    if (PATHS) {
                        sp = xyz(org4[my]);
                        d = xyz(dir4[my]);
                    } else {
                        sp = mk(org3[3 * my], org3[3 * my + 1], org3[3 * my + 2]);
                        d = mk(dir3[3 * my], dir3[3 * my + 1], dir3[3 * my + 2]);
                    }
                    rcp = FAST && sc.boxes_fast && fast_box_ray(sp, d);
                    inv = approx_rcp(d);
                    if (KEYED) {
                        rs.org[tid] = make_float4(sp.x, sp.y, sp.z, 0.f);
                        rs.dir[tid] = make_float4(d.x, d.y, d.z, 0.f);
                    }
                    idx = 0;
                    leaves = 0;
                    best = -42.f;
                    best_id = -1;
                    pcnt = 0;
                    pending = 0;
                    done = nn <= 0;
                    if constexpr (FLAT && IND && BUF) {
#pragma unroll
                        for (int m = 0; m < kMailbox; ++m) s_mb[m][tid] = -1;
                    }
                }
            }
        }
        if (__ballot(ray >= 0) == 0) {
            if (exhausted) break;
            continue;
        }
        const bool all_rcp = FLAT && FAST && __ballot(ray >= 0 && !rcp) == 0;
        // Node steps until no lane can walk, `node_steps` steps, or (after the
        // first step, so that every round makes progress) fewer than
        // `walk_min` lanes walk.  The loop is written once per visit form
        // (the form is fixed for the round), so that the loop-carried
        // walk state needs no merge of two code paths on every step.
        if (all_rcp && WIN < 2) {
            for (int k = 0; k < node_steps; ++k) {
                const bool walk = ray >= 0 && !done && pending < QN;
                const uint64_t wm = __ballot(walk);
                if (wm == 0 || (k > 0 && (int)__popcll(wm) < walk_min)) break;
                if (walk) {
                    float4 lo, hi;
                    fetch(idx, lo, hi);
                    visit_rcp(lo, hi);
                }
            }
        } else {
            for (int k = 0; k < node_steps; ++k) {
                const bool walk = ray >= 0 && !done && pending < QN;
                const uint64_t wm = __ballot(walk);
                if (wm == 0 || (k > 0 && (int)__popcll(wm) < walk_min)) break;
                if (walk) {
                    if (WIN > 1) {
                        // The walk only moves forward, so every node of the
                        // window that it reaches is visited in order; the
                        // loads past the last node repeat that node.
                        float4 lo[WIN], hi[WIN];
                        // nodes of the window: all WIN, or (LINE) those up
                        // to the end of idx's line
                        const int len = LINE ? min(WIN, 4 - (idx & 3)) : WIN;
#pragma unroll
                        for (int j = 0; j < WIN; ++j)
                            if (!LINE || j < len)
                                fetch(BUF || idx + j < nn ? idx + j : nn - 1, lo[j], hi[j]);
                        const int first = idx;
                        visit(lo[0], hi[0]);
#pragma unroll
                        for (int j = 1; j < WIN; ++j)
                            if (!done && pending < QN && idx == first + j && (!LINE || j < len))
                                visit(lo[j], hi[j]);
                    } else {
                        float4 lo, hi;
                        fetch(idx, lo, hi);
                        visit(lo, hi);
                    }
                }
            }
        }
        if (FLAT) {
            __builtin_amdgcn_wave_barrier();
            const bool own = ray >= 0 && pending > 0;
            if constexpr (KEYED)
                keyed_leaf_round<QN, IND>(sc, lq, rs, s_key, lane, own ? pending : 0,
                                     own ? pcnt : 0, best, best_id);
            else
                flat_leaf_round<QN, IND, BUF>(sc, lq, lane, own ? pending : 0, own ? pcnt : 0, sp, d,
                                              best, best_id, s_mb);
            if (own) {
                leaves += pending;
                if (leaves >= kMaxMatchData) done = true;
                pcnt = 0;
                pending = 0;
            }
        } else if (ray >= 0 && pcnt > 0) {
            const float4 *rec = sc.tris + 3 * (int64_t)poff;
            if (PREF) {
                float4 r0 = rec[0], r1 = rec[1], r2 = rec[2];
                for (int k = 0; k < pcnt; ++k) {
                    const int kn = k + 1 < pcnt ? k + 1 : k;
                    const float4 n0 = rec[3 * kn], n1 = rec[3 * kn + 1], n2 = rec[3 * kn + 2];
                    test(r0, r1, r2);
                    r0 = n0;
                    r1 = n1;
                    r2 = n2;
                }
            } else {
                for (int k = 0; k < pcnt; ++k) test(rec[3 * k], rec[3 * k + 1], rec[3 * k + 2]);
            }
            pcnt = 0;
            pending = 0;
            if (++leaves == kMaxMatchData) done = true;
        }
        if (ray >= 0 && done && pending == 0) {
            out_id[ray] = best_id;
            out_t[ray] = best;
            ray = -1;
        }
    }
}

// The lane-refill search (traversal 2) as a kernel of its own per fetch
// form, so that each form is compiled with the registers it needs.  FETCH
// bits: 1 node pairs, 2 triangle prefetch, 4 FLAT, and with 4 the leaf
// queue depth 1 << ((FETCH >> 3) & 3); bit 32 compiles the form for six
// waves per SIMD (80 VGPRs) instead of eight (64); bit 64 selects the keyed
// leaf-round fold, bit 128 the indirect (per-triangle) records; with bit 1,
// bits 8-9 widen the node window from 2 to 2 + ((FETCH >> 8) & 3); bit 10
// ends the window at the end of idx's 128-byte line; bit 11 reads the nodes
// with buffer loads.
// The keyed forms hold 21 KB (queue of 4) or 29 KB (queue of 8) of LDS per
// block, so at most 7 or 5 blocks fit in a CU's 160 KB.
constexpr int cast_waves_per_eu(int fetch) {
    const int regs = (fetch & 32) ? 6 : 8;
    const int lds = (fetch & 64) ? (((fetch >> 3) & 3) == 3 ? 5 : 7) : 8;
    return regs < lds ? regs : lds;
}
template <bool PATHS, bool FAST, int FETCH>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    cast_waves_per_eu(FETCH), cast_waves_per_eu(FETCH)))) void k_cast_refill(
    DevScene sc, const float *org3, const float *dir3, const float4 *org4, const float4 *dir4,
    CountRef n_dev, int64_t n_host, WorkCounter *wc, int *out_id, float *out_t) {
    const int64_t n = n_dev.base ? (int64_t)*n_dev.ptr() : n_host;
    constexpr int Q = (FETCH & 4) ? (1 << ((FETCH >> 3) & 3)) : 0;
    constexpr int WIN = (FETCH & 1) ? 2 + ((FETCH >> 8) & 3) : 1;
    cast_refill<PATHS, FAST, WIN, (FETCH & 6) == 2, Q, (FETCH & 64) != 0, (FETCH & 128) != 0,
                (FETCH & 1024) != 0, (FETCH & 2048) != 0>(sc, org3, dir3, org4, dir4, n, wc,
                                                          out_id, out_t);
    wave_done(wc);
}

// Rays given as xyz triples (hpt_cast) or as path state (render).
// With COUNT, per-ray counters go to `counts` (HPT_CAST_COUNTS ints per ray)
// when it is given, and their wave sums (nodes to the cap, leaves,
// triangles) to the 64-bit `totals` when that is given.
template <bool COUNT, bool PATHS, bool FAST>
__global__ __launch_bounds__(kBlock) void k_cast(DevScene sc, const float *org3, const float *dir3,
                                                 const float4 *org4, const float4 *dir4,
                                                 CountRef n_dev, int64_t n_host,
                                                 WorkCounter *wc, int *out_id, float *out_t,
                                                 int *counts, unsigned long long *totals) {
    const int64_t n = n_dev.base ? (int64_t)*n_dev.ptr() : n_host;
    const int lane = threadIdx.x & 63;
    unsigned long long tot[3] = {0, 0, 0};
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(&wc->next, 64);
        base = __shfl(base, 0, 64);
        if ((int64_t)base >= n) break;
        const int64_t i = (int64_t)base + lane;
        if (i < n) {
            V3 sp, d;
            if (PATHS) {
                sp = xyz(org4[i]);
                d = xyz(dir4[i]);
            } else {
                sp = mk(org3[3 * i], org3[3 * i + 1], org3[3 * i + 2]);
                d = mk(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2]);
            }
            float t;
            int c[HPT_CAST_COUNTS];
            const int id = (!COUNT && sc.traversal == 1) ? trace_ray_ww<FAST>(sc, sp, d, t)
                                                         : trace_ray<COUNT, FAST>(sc, sp, d, t, c);
            out_id[i] = id;
            out_t[i] = t;
            if (COUNT) {
                if (counts)
                    for (int k = 0; k < HPT_CAST_COUNTS; ++k) counts[HPT_CAST_COUNTS * i + k] = c[k];
                tot[0] += (unsigned)c[0];
                tot[1] += (unsigned)c[1];
                tot[2] += (unsigned)c[2];
            }
        }
    }
    if (COUNT && totals) {
        for (int k = 0; k < 3; ++k) {
            unsigned long long v = tot[k];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) atomicAdd(totals + k, v);
        }
    }
    wave_done(wc);
}

// --------------------------------------------------------------- camera
__global__ __launch_bounds__(kBlock) void k_camera(const ViewParams *views, int view_id,
                                                   Paths out) {
    const ViewParams &vp = views[view_id];
    const int64_t n = (int64_t)vp.sx * vp.sy;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int gx = (int)(r % vp.sx), gy = (int)(r / vp.sx);
        V3 d;
        int64_t index;
        if (vp.vr) {
            const float lon = (float)(gx + vp.vr_shift) / (float)(vp.vr_sample_y * 2) * kPi * 2.f - kPi;
            const float lat = (float)gy / (float)vp.vr_sample_y * kPi - kPi / 2.0f;
            d = normalize3(mk(cosf(lat) * cosf(lon), sinf(lat), cosf(lat) * sinf(lon)));
            index = (int64_t)vp.vr_sample_x * gy + gx;
        } else {
            const V3 up = mk(vp.up[0], vp.up[1], vp.up[2]);
            const V3 right = mk(vp.right[0], vp.right[1], vp.right[2]);
            V3 nd = cross(up, right);
            nd = add(nd, scl((float)gy / (float)vp.sy * vp.height - vp.height * 0.5f, up));
            nd = add(nd, scl((float)gx / (float)vp.sx * vp.width - vp.width * 0.5f, right));
            d = normalize3(nd);
            index = (int64_t)vp.sx * gy + gx;
        }
        const int64_t g = vp.ray_base + index;
        out.org[g] = make_float4(vp.view[0], vp.view[1], vp.view[2], __int_as_float(view_id));
        out.dir[g] = make_float4(d.x, d.y, d.z, 0.f);
        out.thr[g] = make_float4(1.f, 1.f, 1.f, __int_as_float((int)g));
    }
}

// ----------------------------------------------------- block compaction
// Exclusive position of each surviving lane in the block's output run and
// the run's base in the destination, from wave64 ballots and an LDS scan of
// the per-wave totals; one atomic per block.
__device__ __forceinline__ int64_t block_compact(bool alive, int *counter, int &pos_in_run) {
    __shared__ int wave_tot[kBlock / 64];
    __shared__ int run_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t mask = __ballot(alive);
    const int before = __popcll(mask & ((1ULL << lane) - 1ULL));
    if (lane == 0) wave_tot[wave] = __popcll(mask);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < kBlock / 64; ++w) {
            const int c = wave_tot[w];
            wave_tot[w] = acc;
            acc += c;
        }
        run_base = acc ? atomicAdd(counter, acc) : 0;
    }
    __syncthreads();
    pos_in_run = wave_tot[wave] + before;
    const int64_t base = run_base;
    __syncthreads();  // LDS reused by the next call
    return base;
}

// -------------------------------------------------------------- primary
// After the primary cast of one view: misses write background * spp once
// (kernel.cl:193-207 with sample_n = spp), hits are cached and listed.
__global__ __launch_bounds__(kBlock) void k_primary(const ViewParams *views, int view_id,
                                                    const int *hit_id, const float *hit_t,
                                                    float *accum, int2 *prim_hit,
                                                    int *prim_list, int *prim_count) {
    const ViewParams &vp = views[view_id];
    const int64_t n = (int64_t)vp.sx * vp.sy;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t r = base + threadIdx.x;
        bool alive = false;
        int64_t g = 0;
        if (r < n) {
            g = vp.ray_base + r;
            const int id = hit_id[g];
            float *acc = accum + 3 * g;
            if (id < 0) {
                const float s = (float)vp.sample;
                acc[0] = s * (vp.background[0] * 1.f);
                acc[1] = s * (vp.background[1] * 1.f);
                acc[2] = s * (vp.background[2] * 1.f);
            } else {
                acc[0] = acc[1] = acc[2] = 0.f;
                prim_hit[g] = make_int2(id, __float_as_int(hit_t[g]));
                alive = true;
            }
            for (int k = 1; k < kPassSlices; ++k) {
                float *sl = accum + 3 * (g + k * vp.accum_stride);
                sl[0] = sl[1] = sl[2] = 0.f;
            }
        }
        int pos;
        const int64_t run = block_compact(alive, prim_count + view_id, pos);
        if (alive) prim_list[vp.ray_base + run + pos] = (int)g;
    }
}

// --------------------------------------------------------------- expand
// Restore the cached primary hits of every view taking part in a pass
// into the path buffer (the reference copies data_initial_mem and the S1
// cache, trace_runner.cpp:223-229), seeding each path's own LCG stream.
// The pass comes from device memory (`pass`), its segments from `passes`.
__global__ __launch_bounds__(kBlock) void k_expand(const ViewParams *views, const Segment *segs_all,
                                                   const PassDesc *passes, const int *pass,
                                                   const int *prim_list, const int2 *prim_hit,
                                                   const float4 *prim_dir, Paths out, int *out_id,
                                                   float *out_t, CountRef live) {
    const PassDesc pd = passes[*pass];
    const Segment *segs = segs_all + pd.seg_begin;
    const int n_segs = pd.n_segs;
    const int64_t total = pd.total;
    if (blockIdx.x == 0 && threadIdx.x == 0) *live.ptr() = (int)total;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < total;
         k += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = n_segs - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (segs[mid].out_offset <= k) lo = mid;
            else hi = mid - 1;
        }
        const Segment &s = segs[lo];
        const ViewParams &vp = views[s.view];
        const int g = prim_list[s.prim_offset + (k - s.out_offset)];
        const int2 h = prim_hit[g];
        const float4 d = prim_dir[g];
        out.org[k] = make_float4(vp.view[0], vp.view[1], vp.view[2], __int_as_float(s.view));
        out.dir[k] = d;
        // thr.w: the accumulator entry of this path, in its pass's slice.
        const int64_t entry = g + (int64_t)(s.pass % kPassSlices) * vp.accum_stride;
        out.thr[k] = make_float4(1.f, 1.f, 1.f, __int_as_float((int)entry));
        out.seed[k] = path_seed(vp.seed, (int64_t)g - vp.ray_base, s.pass);
        out_id[k] = h.x;
        out_t[k] = __int_as_float(h.y);
    }
}

// ---------------------------------------------------------------- shade
// read_imagef of s1_run with CLK_NORMALIZED_COORDS_TRUE | ADDRESS_REPEAT |
// FILTER_LINEAR on a 512x512 RGBA UNORM8 layer (OpenCL 1.2, section 8.2).
__device__ __forceinline__ V3 texel(const DevScene &sc, int layer, int i, int j) {
    const uchar4 p = sc.textures[((int64_t)layer * 512 + j) * 512 + i];
    return mk(p.x / 255.f, p.y / 255.f, p.z / 255.f);
}

__device__ __forceinline__ V3 read_texture(const DevScene &sc, int layer, float su, float sv) {
    if (layer < 0) layer = 0;
    if (layer > sc.n_textures - 1) layer = sc.n_textures - 1;
    const float u = (su - floorf(su)) * 512.f;
    const float v = (sv - floorf(sv)) * 512.f;
    const float fu = floorf(u - 0.5f), fv = floorf(v - 0.5f);
    int i0 = (int)fu, j0 = (int)fv;
    int i1 = i0 + 1, j1 = j0 + 1;
    if (i0 < 0) i0 = 512 + i0;
    if (j0 < 0) j0 = 512 + j0;
    if (i1 > 511) i1 = i1 - 512;
    if (j1 > 511) j1 = j1 - 512;
    if (i0 < 0 || i0 > 511 || i1 < 0 || i1 > 511) i0 = i1 = 0;
    if (j0 < 0 || j0 > 511 || j1 < 0 || j1 > 511) j0 = j1 = 0;
    const float a = (u - 0.5f) - fu, b = (v - 0.5f) - fv;
    const V3 t00 = texel(sc, layer, i0, j0), t10 = texel(sc, layer, i1, j0);
    const V3 t01 = texel(sc, layer, i0, j1), t11 = texel(sc, layer, i1, j1);
    const float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    return mk(((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x,
              ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y,
              ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z);
}

__device__ __forceinline__ V3 ld_point(const DevScene &sc, int i) { return xyz(sc.points[i]); }

// One path: s1_run (kernel.cl:283-408) then, unless this is the last depth
// level, the chosen s2 rule (kernel.cl:413-557).  Returns true when the
// path continues; its next ray is left in o/d/strength/seed.
__device__ __forceinline__ bool shade_path(const DevScene &sc, const ViewParams &vp, int id,
                                           float t, V3 &o, V3 &d, V3 &strength, uint64_t &seed,
                                           float *acc, bool last) {
    const int4 g = sc.mesh[id];
    const V3 A = ld_point(sc, g.x), B = ld_point(sc, g.y), Cc = ld_point(sc, g.z);
    const hpt_material &mat = sc.materials[g.w];
    const V3 P = add(o, scl(t, d));
    const V3 f1 = sub(A, P), f2 = sub(B, P), f3 = sub(Cc, P);
    const float area_all = length3(cross(sub(B, A), sub(Cc, A)));
    const float a1 = length3(cross(f2, f3)) / area_all;
    const float a2 = length3(cross(f1, f3)) / area_all;
    const float a3 = length3(cross(f1, f2)) / area_all;
    V3 normal;
    if (sc.n_normals == 0) {
        normal = normalize3(cross(sub(B, A), sub(Cc, A)));
    } else {
        normal = add(add(scl(a1, xyz(sc.normals[g.x])), scl(a2, xyz(sc.normals[g.y]))),
                     scl(a3, xyz(sc.normals[g.z])));
    }
    float uvx = 0.f, uvy = 0.f;
    const bool textured = mat.ambient_texture_id >= 0 || mat.diffuse_texture_id >= 0;
    if (textured) {
        if (sc.n_texcoords == 0) {
            const float l = sqrtf(P.x * P.x + P.y * P.y);
            uvx = P.x / l;
            uvy = P.y / l;
        } else {
            const float2 ta = sc.texcoords[g.x], tb = sc.texcoords[g.y], tc = sc.texcoords[g.z];
            uvx = (a1 * ta.x + a2 * tb.x) + a3 * tc.x;
            uvy = (a1 * ta.y + a2 * tb.y) + a3 * tc.y;
        }
        uvy = 1.f - uvy;
    }
    V3 ambient = mk(mat.ambient[0], mat.ambient[1], mat.ambient[2]);
    if (mat.ambient_texture_id >= 0) ambient = read_texture(sc, mat.ambient_texture_id, uvx, uvy);
    V3 result = mulv(strength, add(ambient, mk(mat.emission[0], mat.emission[1], mat.emission[2])));
    result = scl(fabsf(dot(normal, d)), result);
    acc[0] += result.x;
    acc[1] += result.y;
    acc[2] += result.z;

    if (last || length3(strength) < kGeneralThreshold) return false;
    const float rn = lcg_randf(seed);
    int branch;
    if (rn < mat.specular_possibility) {
        strength = mulv(strength, mk(mat.specular[0], mat.specular[1], mat.specular[2]));
        branch = 1;
    } else if (rn < mat.refract_possibility) {
        strength = scl(1.0f - mat.dissolve, strength);
        branch = 0;
    } else {
        V3 diffuse = mk(mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]);
        if (mat.diffuse_texture_id >= 0) diffuse = read_texture(sc, mat.diffuse_texture_id, uvx, uvy);
        strength = mulv(strength, diffuse);
        branch = (rn < mat.diffuse_possibility && vp.disable_diffuse == 0) ? 2 : 3;
    }

    if (branch == 0) {  // s2_refract_run
        const float cos_alpha = dot(d, neg(normal));
        float reverse = 1.0f;
        if (cos_alpha < 0) reverse = -1;
        const float alpha = acosf(cos_alpha);
        const V3 p = scl(cos_alpha, normal);
        const V3 q = normalize3(add(d, p));
        const float sin_beta = sinf(alpha) * powf(mat.optical_density, -reverse);
        V3 fd;
        if (sin_beta <= 1.f) {
            const float beta = asinf(sin_beta);
            fd = add(scl(-reverse * cosf(beta), normal), scl(sin_beta, q));
        } else {
            fd = add(add(d, p), p);
        }
        o = add(P, scl(kStepEpsilon, fd));
        d = fd;
        return true;
    }
    if (branch == 1) {  // s2_specular_run
        V3 rv = lcg_randf3(seed);
        const float sh = mat.shininess;
        rv = mk(rv.x / 2.0f / sh, rv.y / 2.0f / sh, rv.z / 2.0f / sh);
        const float dd = dot(d, normal);
        const V3 projection = scl(dd, normal);
        V3 refl = sub(d, scl(2.0f, projection));
        refl = add(refl, rv);
        o = add(P, scl(kStepEpsilon, refl));
        d = refl;
        return true;
    }
    if (branch == 2) {  // s2_diffuse_run
        const bool front = dot(d, normal) < 0;
        V3 p = lcg_randf3(seed);
        float dn = dot(p, normal);
        if (dn < 0) {
            dn = -dn;
            if (front) p = neg(p);
        }
        strength = scl(dn, strength);
        o = add(P, scl(kStepEpsilon, p));
        d = p;
        return true;
    }
    // s2_light_run
    if (sc.n_lights == 0) return false;
    const bool front = dot(d, normal) < 0;
    const int li = (int)(lcg_rand(seed) % (uint32_t)sc.n_lights);
    const int4 l = sc.lights[li];
    const V3 pa = ld_point(sc, l.x), pb = ld_point(sc, l.y), pc = ld_point(sc, l.z);
    float rx = lcg_randf(seed), ry = lcg_randf(seed);
    if (rx + ry > 1) {
        rx = 1 - rx;
        ry = 1 - ry;
    }
    const V3 point = add(add(pa, scl(rx, sub(pb, pa))), scl(ry, sub(pc, pa)));
    const V3 p = normalize3(sub(point, P));
    const float dl = dot(p, normal);
    if ((dl > 0) == front) {
        strength = scl(fabsf(dl), strength);
        o = add(P, scl(kStepEpsilon, p));
        d = p;
        return true;
    }
    return false;
}

__global__ __launch_bounds__(kBlock) void k_shade(DevScene sc, const ViewParams *views,
                                                  int depth_index, CountRef n_in, Paths in,
                                                  const int *hit_id, const float *hit_t,
                                                  Paths out, CountRef n_out_ref, float *accum) {
    const int64_t n = *n_in.ptr();
    int *n_out = n_out_ref.ptr();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n; base += stride) {
        const int64_t k = base + threadIdx.x;
        bool alive = false;
        float4 o4 = make_float4(0, 0, 0, 0), d4 = o4, th = o4;
        uint64_t seed = 0;
        V3 o, d, strength;
        if (k < n) {
            o4 = in.org[k];
            d4 = in.dir[k];
            th = in.thr[k];
            seed = in.seed[k];
            const int view = __float_as_int(o4.w);
            const int gpix = __float_as_int(th.w);
            const ViewParams &vp = views[view];
            float *acc = accum + 3 * (int64_t)gpix;
            o = xyz(o4);
            d = xyz(d4);
            strength = xyz(th);
            const int id = hit_id[k];
            if (id < 0) {
                // Miss: background * strength * sample_n with sample_n = 1.
                acc[0] += 1.f * (vp.background[0] * strength.x);
                acc[1] += 1.f * (vp.background[1] * strength.y);
                acc[2] += 1.f * (vp.background[2] * strength.z);
            } else {
                alive = shade_path(sc, vp, id, hit_t[k], o, d, strength, seed, acc,
                                   depth_index >= vp.depth - 1);
            }
        }
        int pos;
        const int64_t run = block_compact(alive, n_out, pos);
        if (alive) {
            const int64_t w = run + pos;
            out.org[w] = make_float4(o.x, o.y, o.z, o4.w);
            out.dir[w] = make_float4(d.x, d.y, d.z, 0.f);
            out.thr[w] = make_float4(strength.x, strength.y, strength.z, th.w);
            out.seed[w] = seed;
        }
    }
}

// -------------------------------------------------------------- resolve
__device__ __forceinline__ unsigned char unorm8(float f) {
    const float x = f * 255.0f;
    if (!(x > 0.f)) return 0;  // also NaN
    if (x >= 255.f) return 255;
    return (unsigned char)rintf(x);
}

__global__ __launch_bounds__(kBlock) void k_resolve(const ViewParams *views, int view_id,
                                                    int width, int height, int ssx, int ssy,
                                                    int samples, float brightness,
                                                    const float *accum, uchar4 *rgba) {
    const ViewParams &vp = views[view_id];
    const int64_t n = (int64_t)width * height;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int gx = (int)(p % width), gy = (int)(p / width);
        float px = 0.f, py = 0.f, pz = 0.f;
        for (int i = 0; i < ssx; ++i)
            for (int j = 0; j < ssy; ++j) {
                const int64_t row = (int64_t)gy * ssy + j, col = (int64_t)gx * ssx + i;
                const int64_t index = vp.ray_base + row * width * ssx + col;
                px += accum[index * 3];
                py += accum[index * 3 + 1];
                pz += accum[index * 3 + 2];
            }
        const float nss = (float)(ssx * ssy), sp = (float)samples;
        px /= nss; py /= nss; pz /= nss;
        px /= sp; py /= sp; pz /= sp;
        px *= brightness; py *= brightness; pz *= brightness;
        px = 1.f - expf(-px);
        py = 1.f - expf(-py);
        pz = 1.f - expf(-pz);
        rgba[p] = make_uchar4(unorm8(px), unorm8(py), unorm8(pz), 0);
    }
}

// Sum a view's accumulator slices in slice order into `out` (same camera-
// ray indexing as slice 0; `out` may be `accum` itself: each thread reads
// all slices of its entry before writing it).
__global__ __launch_bounds__(kBlock) void k_fold(const ViewParams *views, int view_id,
                                                 const float *accum, float *out) {
    const ViewParams &vp = views[view_id];
    const int64_t n = (int64_t)vp.sx * vp.sy;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const float *a = accum + 3 * (vp.ray_base + r);
        float x = a[0], y = a[1], z = a[2];
        for (int k = 1; k < kPassSlices; ++k) {
            const float *b = a + 3 * k * vp.accum_stride;
            x += b[0];
            y += b[1];
            z += b[2];
        }
        float *o = out + 3 * (vp.ray_base + r);
        o[0] = x;
        o[1] = y;
        o[2] = z;
    }
}

// Background of the camera rays whose primary ray missed, for an image of
// `samples` passes: what k_primary writes for a render of that many
// samples (slice 0 = samples x background; the paths never touch these
// entries).  Lets a progressive request resolve after any pass count.
__global__ __launch_bounds__(kBlock) void k_misses(const ViewParams *views, int view_id,
                                                   const int *hit_id, int samples,
                                                   float *accum) {
    const ViewParams &vp = views[view_id];
    const int64_t n = (int64_t)vp.sx * vp.sy;
    const float s = (float)samples;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = vp.ray_base + r;
        if (hit_id[g] >= 0) continue;
        float *acc = accum + 3 * g;
        acc[0] = s * (vp.background[0] * 1.f);
        acc[1] = s * (vp.background[1] * 1.f);
        acc[2] = s * (vp.background[2] * 1.f);
    }
}

// ------------------------------------------------------------ launchers
static int grid_for(int64_t n, int cap) {
    int64_t g = (n + kBlock - 1) / kBlock;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (int)g;
}

void launch_cast(const DevScene &sc, const float *org3, const float *dir3, const float4 *org4,
                 const float4 *dir4, CountRef n_dev, int64_t n_host, WorkCounter *wc,
                 int *out_id, float *out_t, int *counts, unsigned long long *totals, int grid,
                 hipStream_t st) {
    const bool paths = org4 != nullptr;
    const bool count = counts || totals;
    const bool fast = sc.fast_div != 0;
    if (!count && sc.traversal == 2) {
#define HPT_REFILL(P, F, M)                                                                      \
    k_cast_refill<P, F, M><<<grid, kBlock, 0, st>>>(sc, org3, dir3, org4, dir4, n_dev, n_host, wc, \
                                                    out_id, out_t)
#define HPT_REFILL_M(P, F)                                                                       \
    switch (sc.fetch) {                                                                          \
    case 1: HPT_REFILL(P, F, 1); break;                                                          \
    case 2: HPT_REFILL(P, F, 2); break;                                                          \
    case 3: HPT_REFILL(P, F, 3); break;                                                          \
    case 4: HPT_REFILL(P, F, 4); break;                                                          \
    case 5: HPT_REFILL(P, F, 5); break;                                                          \
    case 12: HPT_REFILL(P, F, 12); break;                                                        \
    case 13: HPT_REFILL(P, F, 13); break;                                                        \
    case 20: HPT_REFILL(P, F, 20); break;                                                        \
    case 21: HPT_REFILL(P, F, 21); break;                                                        \
    case 28: HPT_REFILL(P, F, 28); break;                                                        \
    case 29: HPT_REFILL(P, F, 29); break;                                                        \
    case 52: HPT_REFILL(P, F, 52); break;                                                        \
    case 60: HPT_REFILL(P, F, 60); break;                                                        \
    case 84: HPT_REFILL(P, F, 84); break;                                                        \
    case 92: HPT_REFILL(P, F, 92); break;                                                        \
    case 116: HPT_REFILL(P, F, 116); break;                                                      \
    case 124: HPT_REFILL(P, F, 124); break;                                                      \
    case 148: HPT_REFILL(P, F, 148); break;                                                      \
    case 149: HPT_REFILL(P, F, 149); break;                                                      \
    case 156: HPT_REFILL(P, F, 156); break;                                                      \
    case 157: HPT_REFILL(P, F, 157); break;                                                      \
    case 189: HPT_REFILL(P, F, 189); break;                                                      \
    case 220: HPT_REFILL(P, F, 220); break;                                                      \
    case 252: HPT_REFILL(P, F, 252); break;                                                      \
    case 413: HPT_REFILL(P, F, 413); break;                                                      \
    case 445: HPT_REFILL(P, F, 445); break;                                                      \
    case 669: HPT_REFILL(P, F, 669); break;                                                      \
    case 701: HPT_REFILL(P, F, 701); break;                                                      \
    case 1469: HPT_REFILL(P, F, 1469); break;                                                    \
    case 1725: HPT_REFILL(P, F, 1725); break;                                                    \
    case 3517: HPT_REFILL(P, F, 3517); break;                                                    \
    case 3773: HPT_REFILL(P, F, 3773); break;                                                    \
    default: HPT_REFILL(P, F, 0); break;                                                         \
    }
        if (paths) { if (fast) { HPT_REFILL_M(true, true) } else { HPT_REFILL_M(true, false) } }
        else { if (fast) { HPT_REFILL_M(false, true) } else { HPT_REFILL_M(false, false) } }
#undef HPT_REFILL_M
#undef HPT_REFILL
        return;
    }
#define HPT_CAST(C, P, F)                                                                       \
    k_cast<C, P, F><<<grid, kBlock, 0, st>>>(sc, org3, dir3, org4, dir4, n_dev, n_host, wc, out_id, \
                                             out_t, counts, totals)
    if (count) {
        if (paths) { if (fast) HPT_CAST(true, true, true); else HPT_CAST(true, true, false); }
        else { if (fast) HPT_CAST(true, false, true); else HPT_CAST(true, false, false); }
    } else {
        if (paths) { if (fast) HPT_CAST(false, true, true); else HPT_CAST(false, true, false); }
        else { if (fast) HPT_CAST(false, false, true); else HPT_CAST(false, false, false); }
    }
#undef HPT_CAST
}

void launch_fold(const ViewParams *views, int view_id, int64_t n, const float *accum, float *out,
                 hipStream_t st) {
    k_fold<<<grid_for(n, 8192), kBlock, 0, st>>>(views, view_id, accum, out);
}

void launch_misses(const ViewParams *views, int view_id, int64_t n, const int *hit_id,
                   int samples, float *accum, hipStream_t st) {
    k_misses<<<grid_for(n, 8192), kBlock, 0, st>>>(views, view_id, hit_id, samples, accum);
}

void launch_camera(const ViewParams *views, int view_id, int64_t n, Paths out, hipStream_t st) {
    k_camera<<<grid_for(n, 8192), kBlock, 0, st>>>(views, view_id, out);
}

void launch_primary(const ViewParams *views, int view_id, int64_t n, const int *hit_id,
                    const float *hit_t, float *accum, int2 *prim_hit, int *prim_list,
                    int *prim_count, hipStream_t st) {
    k_primary<<<grid_for(n, 8192), kBlock, 0, st>>>(views, view_id, hit_id, hit_t, accum,
                                                     prim_hit, prim_list, prim_count);
}

void launch_expand(const ViewParams *views, const Segment *segs_all, const PassDesc *passes,
                   const int *pass, int64_t max_total, const int *prim_list, const int2 *prim_hit,
                   const float4 *prim_dir, Paths out, int *out_id, float *out_t, CountRef live,
                   hipStream_t st) {
    k_expand<<<grid_for(max_total, 8192), kBlock, 0, st>>>(views, segs_all, passes, pass,
                                                            prim_list, prim_hit, prim_dir, out,
                                                            out_id, out_t, live);
}

void launch_shade(const DevScene &sc, const ViewParams *views, int depth_index, CountRef n_in,
                  Paths in, const int *hit_id, const float *hit_t, Paths out, CountRef n_out,
                  float *accum, int grid, hipStream_t st) {
    k_shade<<<grid, kBlock, 0, st>>>(sc, views, depth_index, n_in, in, hit_id, hit_t, out, n_out,
                                     accum);
}

__global__ void k_next_pass(int *pass) { *pass += 1; }

void launch_next_pass(int *pass, hipStream_t st) { k_next_pass<<<1, 1, 0, st>>>(pass); }

__global__ void k_zero_counter(WorkCounter *wc) {
    wc->next = 0;
    wc->done = 0;
}

void launch_zero_counter(WorkCounter *wc, hipStream_t st) {
    k_zero_counter<<<1, 1, 0, st>>>(wc);
}

void launch_resolve(const ViewParams *views, int view_id, int width, int height, int ssx, int ssy,
                    int samples, float brightness, const float *accum, uchar4 *rgba,
                    hipStream_t st) {
    k_resolve<<<grid_for((int64_t)width * height, 8192), kBlock, 0, st>>>(
        views, view_id, width, height, ssx, ssy, samples, brightness, accum, rgba);
}

int occupancy_blocks_cast() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_cast_refill<true, true, 0>, kBlock, 0) != hipSuccess) b = 4;
    return b;
}

int occupancy_blocks_shade() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_shade, kBlock, 0) != hipSuccess) b = 4;
    return b;
}

}  // namespace hpt
