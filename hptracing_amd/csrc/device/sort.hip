// Coherent ordering of hpt_cast batches (hpt_tuning.sort_rays).
//
// A batch of incoherent rays (origins spread over the scene, directions over
// the sphere) walks the whole tree in every wave, so the rays resident on
// the machine at any moment touch nodes and records all over it.  Sorted by
// the Morton code of their origins in the scene box, neighbouring rays start
// in the same region and share the upper walk and many leaves, which the
// caches then serve.  Each ray's result does not depend on the order, so
// the batch is cast in sorted order and the results are written back to the
// caller's positions: same bits as the unsorted cast.
//
//   k_ray_keys      30-bit Morton code (10 bits per axis) of each origin,
//                   clamped to the scene box, and the identity permutation
//   (radix sort of the (key, index) pairs, hipcub)
//   k_gather_rays   the rays in sorted order as float4 origin/direction
//   (the cast, path form)
//   k_scatter_hits  hit id and t back to the caller's positions
//
// hpt_cast takes the buffers from the stream-ordered allocator and releases
// them on the same stream, so concurrent casts on other streams share
// nothing; hpt_cast_noalloc carves them out of its static scratch array.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"
#include "launch.hpp"

namespace hpt {

namespace {

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t quantise(float x, float lo, float scale) {
    // fmaxf / fminf return the other operand for a NaN, so every origin maps
    // into [0, 1023].
    return (uint32_t)fminf(fmaxf((x - lo) * scale, 0.f), 1023.f);
}

__global__ __launch_bounds__(kBlock) void k_ray_keys(const float *org3, int64_t n, float3 lo,
                                                     float3 scale, uint32_t *keys, int32_t *idx) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t qx = quantise(org3[3 * i], lo.x, scale.x);
        const uint32_t qy = quantise(org3[3 * i + 1], lo.y, scale.y);
        const uint32_t qz = quantise(org3[3 * i + 2], lo.z, scale.z);
        keys[i] = spread10(qx) | (spread10(qy) << 1) | (spread10(qz) << 2);
        idx[i] = (int32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_gather_rays(const float *org3, const float *dir3,
                                                        const int32_t *idx, int64_t n,
                                                        float4 *org4, float4 *dir4) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx[j];
        org4[j] = make_float4(org3[3 * i], org3[3 * i + 1], org3[3 * i + 2], 0.f);
        dir4[j] = make_float4(dir3[3 * i], dir3[3 * i + 1], dir3[3 * i + 2], 0.f);
    }
}

__global__ __launch_bounds__(kBlock) void k_scatter_hits(const int32_t *idx, const int *ids_s,
                                                         const float *t_s, int64_t n, int *out_id,
                                                         float *out_t) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx[j];
        out_id[i] = ids_s[j];
        out_t[i] = t_s[j];
    }
}

// Stream-ordered buffers released on the stream when the cast is done.
struct AsyncBuffers {
    hipStream_t st;
    std::vector<void *> ptrs;
    explicit AsyncBuffers(hipStream_t s) : st(s) {}
    template <class T>
    T *get(size_t count) {
        void *p = nullptr;
        hip_check(hipMallocAsync(&p, count * sizeof(T) + 256, st), "hipMallocAsync");
        ptrs.push_back(p);
        return static_cast<T *>(p);
    }
    ~AsyncBuffers() {
        for (void *p : ptrs) (void)hipFreeAsync(p, st);
    }
};

}  // namespace

size_t sort_temp_bytes(int64_t n, hipStream_t st) {
    size_t bytes = 0;
    hipcub::DoubleBuffer<uint32_t> keys(nullptr, nullptr);
    hipcub::DoubleBuffer<int32_t> idx(nullptr, nullptr);
    hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys, idx, (int)n, 0, 30, st),
              "radix sort size");
    return bytes;
}

void cast_sorted_into(const hpt_scene *s, const LaunchForm &f, const float *org3,
                      const float *dir3, int64_t n, int *out_id, float *out_t,
                      const SortBuffers &b, hipStream_t st) {
    float3 lo, scale;
    float *lo_p = &lo.x, *scale_p = &scale.x;
    for (int k = 0; k < 3; ++k) {
        lo_p[k] = s->box_lo[k];
        const float ext = s->box_hi[k] - s->box_lo[k];
        scale_p[k] = ext > 0.f ? 1023.f / ext : 0.f;
    }
    const int grid = std::max(1, s->n_cus) * 8;
    k_ray_keys<<<grid, kBlock, 0, st>>>(org3, n, lo, scale, b.keys[0], b.idx[0]);
    hip_check(hipGetLastError(), "k_ray_keys");
    // The double-buffer form of the sort needs temporary storage for its
    // digit counts only; which buffer holds the result is known on the
    // host when the call returns.
    hipcub::DoubleBuffer<uint32_t> keys(b.keys[0], b.keys[1]);
    hipcub::DoubleBuffer<int32_t> idx(b.idx[0], b.idx[1]);
    size_t temp_bytes = b.temp_bytes;
    hip_check(hipcub::DeviceRadixSort::SortPairs(b.temp, temp_bytes, keys, idx, (int)n, 0, 30, st),
              "radix sort");
    const int32_t *order = idx.Current();
    k_gather_rays<<<grid, kBlock, 0, st>>>(org3, dir3, order, n, b.org4, b.dir4);
    hip_check(hipGetLastError(), "k_gather_rays");
    launch_cast(f.dev, nullptr, nullptr, b.org4, b.dir4, CountRef{}, n, b.wc, b.ids, b.t, nullptr,
                nullptr, f.cast_grid, st);
    hip_check(hipGetLastError(), "k_cast launch");
    k_scatter_hits<<<grid, kBlock, 0, st>>>(order, b.ids, b.t, n, out_id, out_t);
    hip_check(hipGetLastError(), "k_scatter_hits");
}

void cast_sorted(const hpt_scene *s, const LaunchForm &f, const float *org3, const float *dir3,
                 int64_t n, int *out_id, float *out_t, hipStream_t st) {
    AsyncBuffers buf(st);
    SortBuffers b{};
    for (int k = 0; k < 2; ++k) {
        b.keys[k] = buf.get<uint32_t>((size_t)n);
        b.idx[k] = buf.get<int32_t>((size_t)n);
    }
    b.org4 = buf.get<float4>((size_t)n);
    b.dir4 = buf.get<float4>((size_t)n);
    b.ids = buf.get<int>((size_t)n);
    b.t = buf.get<float>((size_t)n);
    b.temp_bytes = sort_temp_bytes(n, st);
    b.temp = buf.get<uint8_t>(b.temp_bytes);
    StreamCounter wc(st);
    b.wc = wc.p;
    cast_sorted_into(s, f, org3, dir3, n, out_id, out_t, b, st);
}

}  // namespace hpt
