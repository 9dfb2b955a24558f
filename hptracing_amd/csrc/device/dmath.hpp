// Device arithmetic of the path tracer.  Every operation keeps the order
// written in src/hp/cl_src/kernel.cl, with the OpenCL builtins spelled out:
//   dot(a,b)   = (a.x*b.x + a.y*b.y) + a.z*b.z
//   cross(a,b) = (a.y*b.z - a.z*b.y, a.z*b.x - a.x*b.z, a.x*b.y - a.y*b.x)
//   length(a)  = sqrt(dot(a,a)), normalize(a) = a / length(a) per component
// The library is compiled with -ffp-contract=off and correctly rounded
// division and square root, so +,-,*,/ and sqrt round exactly as IEEE
// float32 does on the host.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hpt {

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 mk(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 xyz(float4 a) { return V3{a.x, a.y, a.z}; }
__device__ __forceinline__ V3 add(V3 a, V3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 mulv(V3 a, V3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ V3 scl(float s, V3 a) { return mk(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ V3 neg(V3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length3(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize3(V3 a) {
    const float l = length3(a);
    return mk(a.x / l, a.y / l, a.z / l);
}

constexpr float kPi = 3.1415926535f;            // PI, kernel.cl:27
constexpr float kGeneralThreshold = 1e-4f;      // kernel.cl:49
constexpr float kStepEpsilon = 1e-5f;           // kernel.cl:411
constexpr int kMaxMatchData = 512;              // kernel.cl:133

// Slab test `_box_intersect` (kernel.cl:51-83): a start point inside the
// (closed) box hits; otherwise per-axis quotients (division, not a
// reciprocal), IEEE minNum/maxNum, reject when the nearest exit is <= 0.
__device__ __forceinline__ bool box_hit(V3 bs, V3 be, V3 sp, V3 d) {
    if (sp.x >= bs.x && sp.x <= be.x && sp.y >= bs.y && sp.y <= be.y && sp.z >= bs.z &&
        sp.z <= be.z)
        return true;
    const float tx0 = (bs.x - sp.x) / d.x, tx1 = (be.x - sp.x) / d.x;
    const float ty0 = (bs.y - sp.y) / d.y, ty1 = (be.y - sp.y) / d.y;
    const float tz0 = (bs.z - sp.z) / d.z, tz1 = (be.z - sp.z) / d.z;
    const float max_of_mins = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    const float min_of_maxs = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    if (min_of_maxs <= 0) return false;
    return max_of_mins <= min_of_maxs;
}

// Slab test with the six quotients formed as (s - p) * y, y = v_rcp_f32(d)
// (within 1 ulp of 1/d), certified against the rounding error so that the
// answer is box_hit's for every input it is used on.
//
// Preconditions (checked by fast_box_ray per ray; the scene's box
// coordinates are checked on upload): every direction component has
// magnitude in [2^-40, 2^40] (none is 0); every origin and box coordinate
// is 0 or has magnitude in [2^-60, 2^80].  Then every nonzero difference
// a = p - s has |a| >= 2^-84 (both operands are multiples of 2^-84), so
// each quotient is 0 or a normal number no larger than 2^121 (never
// infinite or NaN), and the approximate quotient q' = fl(a * y) stays
// within relative 2^-21.9 of the correctly rounded q = fl(a / d), with the
// same sign class (negative, zero, positive).
//
//  * The exit (min of maxs) formed from q' has the sign class of the one
//    formed from q: fminf/fmaxf pick one of their inputs, and the inputs'
//    sign classes agree one for one.  So "nearest exit <= 0" is exact.
//  * A start point inside the closed box has lo - p <= 0 <= hi - p on every
//    axis, so each axis's quotients straddle 0 and entry <= 0 <= exit.  With
//    exit > 0 the box is then hit by the slab rule as well, and exit < 0
//    rules the start point out; only exit == 0 needs the start-inside test,
//    and that rare case is left to the division form.
//  * x -> x -+ e|x| is monotone for e < 1, so the entry (max of mins) and
//    exit formed from q' lie within relative e of those formed from q.
//    When they differ by at least kCertify = 2^-20 of the sum of their
//    magnitudes (3.7 times 2^-21.9, the rounding of the check itself and a
//    subnormal bound included), their order is the exact order.
//    Otherwise the division form decides.
__device__ __forceinline__ bool fast_box_ray(V3 sp, V3 d) {
    auto dir_ok = [](float c) {
        const float a = fabsf(c);
        return a >= 0x1p-40f && a <= 0x1p40f;
    };
    auto org_ok = [](float c) {
        const float a = fabsf(c);
        return a == 0.f || (a >= 0x1p-60f && a <= 0x1p80f);
    };
    return dir_ok(d.x) && dir_ok(d.y) && dir_ok(d.z) && org_ok(sp.x) && org_ok(sp.y) &&
           org_ok(sp.z);
}

__device__ __forceinline__ V3 approx_rcp(V3 d) {
    return mk(__builtin_amdgcn_rcpf(d.x), __builtin_amdgcn_rcpf(d.y), __builtin_amdgcn_rcpf(d.z));
}

// Relative margin of the certified comparison (see above).
constexpr float kCertify = 0x1p-20f;

//
// The certified part alone: the reciprocal form's decision, and in `open`
// whether it is not certified (the caller then asks box_hit).
__device__ __forceinline__ bool box_hit_rcp_part(V3 bs, V3 be, V3 sp, V3 y, bool &open) {
    const float tx0 = (bs.x - sp.x) * y.x, tx1 = (be.x - sp.x) * y.x;
    const float ty0 = (bs.y - sp.y) * y.y, ty1 = (be.y - sp.y) * y.y;
    const float tz0 = (bs.z - sp.z) * y.z, tz1 = (be.z - sp.z) * y.z;
    const float entry = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fminf(tz0, tz1));
    const float exit = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fmaxf(tz0, tz1));
    const float bound = (fabsf(entry) + fabsf(exit)) * kCertify;
    const bool sure = fabsf(entry - exit) >= bound;
    const bool ahead = exit > 0.f;
    open = exit == 0.f || (ahead && !sure);
    return ahead && entry <= exit;
}

__device__ __forceinline__ bool box_hit_rcp(V3 bs, V3 be, V3 sp, V3 d, V3 y) {
    bool open;
    const bool hit = box_hit_rcp_part(bs, be, sp, y, open);
    if (!open) return hit;
    return box_hit(bs, be, sp, d);
}

// Correctly rounded 1/d (the value of the reference's `1.f / det` under
// -cl-fp32-correctly-rounded-divide-sqrt) without the division sequence:
// v_rcp_f32 (within 1 ulp) refined by one Newton step in fused
// multiply-adds.  Checked on the MI355X for every float with a biased
// exponent in [2, 252] (|d| in [2^-125, 2^126)), both signs: equal to the
// division bit for bit (tests/cpp/rcp_exhaustive.hip).  Zero, subnormal,
// huge, infinite and NaN inputs take the division.
__device__ __forceinline__ float rcp_rn(float d) {
    const float y = __builtin_amdgcn_rcpf(d);
    float r = fmaf(fmaf(-d, y, 1.f), y, y);
    const uint32_t ex = (__float_as_uint(d) >> 23) & 0xFFu;
    if (__builtin_expect(ex - 2u > 250u, 0)) r = 1.f / d;
    return r;
}

// Moller-Trumbore `_single_intersect` (kernel.cl:85-109) on a triangle
// record that stores pa and the edges e1 = pb - pa, e2 = pc - pa (the
// same float subtractions the kernel performs).  The degenerate guard of
// the reference can never fire, so none is applied; 1/det is rcp_rn's
// correctly rounded reciprocal.
__device__ __forceinline__ float tri_hit(V3 sp, V3 d, V3 pa, V3 e1, V3 e2) {
    const V3 P = cross(d, e2);
    const float det = dot(e1, P);
    const float inv_det = rcp_rn(det);
    const V3 T = sub(sp, pa);
    const float u = dot(T, P) * inv_det;
    if (u < 0.f || u > 1.f) return -1.f;
    const V3 Q = cross(T, e1);
    const float v = dot(d, Q) * inv_det;
    if (v < 0.f || u + v > 1.f) return -1.f;
    const float t = dot(e2, Q) * inv_det;
    if (t > 0) return t;
    return -1.f;
}

// The same test with dot and cross evaluated the way ROCm's OpenCL device
// library evaluates its builtins (measured on gfx950 against the library,
// bit for bit on 2^18 random inputs):
//   dot(a,b)   = fma(a.z, b.z, fma(a.y, b.y, a.x*b.x))
//   cross(a,b) = (fma(a.y, b.z, -(a.z*b.y)), fma(a.z, b.x, -(a.x*b.z)),
//                 fma(a.x, b.y, -(a.y*b.x)))
// With it the cast reproduces the reference kernel compiled by that library
// bit for bit, t included.
__device__ __forceinline__ float dot_ocl(V3 a, V3 b) {
    return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x));
}
__device__ __forceinline__ V3 cross_ocl(V3 a, V3 b) {
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
              fmaf(a.x, b.y, -(a.y * b.x)));
}
__device__ __forceinline__ float tri_hit_ocl(V3 sp, V3 d, V3 pa, V3 e1, V3 e2) {
    const V3 P = cross_ocl(d, e2);
    const float det = dot_ocl(e1, P);
    const float inv_det = rcp_rn(det);
    const V3 T = sub(sp, pa);
    const float u = dot_ocl(T, P) * inv_det;
    if (u < 0.f || u > 1.f) return -1.f;
    const V3 Q = cross_ocl(T, e1);
    const float v = dot_ocl(d, Q) * inv_det;
    if (v < 0.f || u + v > 1.f) return -1.f;
    const float t = dot_ocl(e2, Q) * inv_det;
    if (t > 0) return t;
    return -1.f;
}

// ------------------------------------------------------------------ RNG
// 48-bit LCG of kernel.cl:29-47.
__device__ __forceinline__ uint32_t lcg_rand(uint64_t &s) {
    s = (s * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
    return (uint32_t)(s >> 16);
}
__device__ __forceinline__ float lcg_randf(uint64_t &s) {
    return (float)lcg_rand(s) / 4294967296.0f;  // float(0xFFFFFFFF) rounds to 2^32
}
__device__ __forceinline__ V3 lcg_randf3(uint64_t &s) {
    const float alpha = lcg_randf(s) * kPi;
    const float beta = lcg_randf(s) * kPi;
    const float cos_beta = cosf(beta);
    return mk(cos_beta * cosf(alpha), cos_beta * sinf(alpha), sinf(beta));
}

__host__ __device__ __forceinline__ uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// Initial LCG state of camera ray `ray` in pass `pass` of a render seeded
// with `seed`: every path owns its random stream, so a pixel's result does
// not depend on which other rays share the launch.
__host__ __device__ __forceinline__ uint64_t path_seed(uint64_t seed, int64_t ray, int32_t pass) {
    return splitmix(splitmix(seed ^ splitmix((uint64_t)ray)) + (uint64_t)(uint32_t)pass) &
           ((1ULL << 48) - 1);
}

}  // namespace hpt
