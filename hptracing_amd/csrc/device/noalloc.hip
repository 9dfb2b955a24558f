// hpt_cast_noalloc: hpt_cast without any device allocation.  No counterpart
// in the reference; it is for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  hpt_cast takes a stream-ordered work counter per call and,
// when the scene sorts rays, stream-ordered sort buffers; here all of them
// are carved out of `scratch`, a file-scope device array of fixed size that
// each GPU holds from the library's load until the process ends.
//
// Layout of scratch (HPT_NOALLOC_MAX_RAYS = N rays):
//   [0, 256)            the work counter (cleared by a fill kernel per call)
//   then, N entries each: keys x2 (u32), idx x2 (i32), org4 and dir4
//   (float4), hit id (i32) and t (f32) of the sorted order
//   then kSortTempBytes for the radix sort's digit counts.
// Every call on a GPU uses that GPU's one array, so calls that overlap on a
// GPU (other streams, other threads) race; a caller serialises them.
#include <map>

#include "internal.hpp"
#include "launch.hpp"

namespace hpt {
namespace {

constexpr size_t kMaxRays = HPT_NOALLOC_MAX_RAYS;
constexpr size_t kCounterBytes = 256;
constexpr size_t kSortTempBytes = (size_t)8 << 20;
constexpr size_t kPerRayBytes = 4 * 2 + 4 * 2 + 16 * 2 + 4 + 4;
constexpr size_t kScratchBytes = kCounterBytes + kMaxRays * kPerRayBytes + kSortTempBytes;
static_assert(kScratchBytes == ((size_t)64 << 20) + 256, "64 MiB of buffers and a counter");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> its copy of `scratch`

// The address of `scratch` on `device` (the current device), looked up once
// per device.  In host code `scratch` names a host-side stand-in, so every
// kernel argument below is derived from this address.
unsigned char *scratch_on(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) return it->second;
    unsigned char *ptr = nullptr;
    hip_check(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)),
              "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = ptr;
    return ptr;
}

}  // namespace
}  // namespace hpt

extern "C" {

int hpt_cast_noalloc(hpt_scene *s, const float *org, const float *dir, int64_t n,
                     int32_t *tri_id, float *t, void *stream) {
    using namespace hpt;
    return guarded([&] {
        if (!s) fail(HPT_ERR_INVALID, "null scene");
        if (n < 0 || n > (int64_t)kMaxRays)
            fail(HPT_ERR_INVALID, "hpt_cast_noalloc: " + std::to_string(n) +
                                      " rays; its static scratch holds at most "
                                      "HPT_NOALLOC_MAX_RAYS = " + std::to_string(kMaxRays) +
                                      " (hpt_cast takes larger batches)");
        if (n == 0) return;
        if (!org || !dir || !tri_id || !t) fail(HPT_ERR_INVALID, "null buffer");
        int device = -1;
        hip_check(hipGetDevice(&device), "hipGetDevice");
        if (device != s->device)
            fail(HPT_ERR_INVALID, "hpt_cast_noalloc: the current device is not the scene's");
        unsigned char *base = scratch_on(device);
        const LaunchForm f = s->form();
        hipStream_t st = (hipStream_t)stream;
        WorkCounter *wc = reinterpret_cast<WorkCounter *>(base);
        // The array keeps the previous call's contents: the counter starts
        // at zero for this launch.
        launch_zero_counter(wc, st);
        hip_check(hipGetLastError(), "k_zero_counter");
        if (f.sort_rays && n >= kSortMinRays) {
            unsigned char *q = base + kCounterBytes;
            SortBuffers b{};
            for (int k = 0; k < 2; ++k) {
                b.keys[k] = reinterpret_cast<uint32_t *>(q);
                q += 4 * kMaxRays;
                b.idx[k] = reinterpret_cast<int32_t *>(q);
                q += 4 * kMaxRays;
            }
            b.org4 = reinterpret_cast<float4 *>(q);
            q += 16 * kMaxRays;
            b.dir4 = reinterpret_cast<float4 *>(q);
            q += 16 * kMaxRays;
            b.ids = reinterpret_cast<int *>(q);
            q += 4 * kMaxRays;
            b.t = reinterpret_cast<float *>(q);
            q += 4 * kMaxRays;
            b.temp = q;
            b.temp_bytes = sort_temp_bytes(n, st);
            if (b.temp_bytes > kSortTempBytes)
                fail(HPT_ERR_INVALID, "hpt_cast_noalloc: the radix sort asks for " +
                                          std::to_string(b.temp_bytes) +
                                          " bytes of temporary storage, the scratch holds " +
                                          std::to_string(kSortTempBytes));
            b.wc = wc;
            cast_sorted_into(s, f, org, dir, n, tri_id, t, b, st);
            return;
        }
        launch_cast(f.dev, org, dir, nullptr, nullptr, CountRef{}, n, wc, tri_id, t, nullptr,
                    nullptr, f.cast_grid, st);
        hip_check(hipGetLastError(), "k_cast launch");
    });
}

}  // extern "C"
