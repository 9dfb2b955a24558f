// Exhaustive check of hpt::rcp_rn (hptracing_amd/csrc/device/dmath.hpp), the
// reciprocal the intersection test uses for the reference's `1.f / det`:
// for every one of the 2^32 float bit patterns, rcp_rn(d) must equal the
// correctly rounded division 1.f / d bit for bit (built with the library's
// flags: correctly rounded division, no contraction, denormals kept).
// Prints "checked N mismatches M first 0x..." and exits 0 only when M == 0.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#include "dmath.hpp"

constexpr uint32_t kThreads = 1u << 23;           // each thread takes 512 patterns
constexpr uint32_t kPerThread = 1u << 9;          // kThreads * kPerThread = 2^32

__global__ void check(unsigned long long *mismatches, unsigned *first) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= kThreads) return;
    unsigned long long bad = 0;
    for (uint32_t k = 0; k < kPerThread; ++k) {
        const uint32_t bits = t * kPerThread + k;
        const float d = __uint_as_float(bits);
        const float a = hpt::rcp_rn(d);
        const float b = 1.f / d;
        if (__float_as_uint(a) != __float_as_uint(b)) {
            if (bad == 0) atomicMin(first, bits);       // one atomic per thread at most
            ++bad;
        }
    }
    if (bad) atomicAdd(mismatches, bad);
}

int main() {
    unsigned long long *d_bad = nullptr;
    unsigned *d_first = nullptr;
    if (hipMalloc(&d_bad, sizeof(*d_bad)) != hipSuccess ||
        hipMalloc(&d_first, sizeof(*d_first)) != hipSuccess) {
        std::fprintf(stderr, "hipMalloc failed\n");
        return 2;
    }
    const unsigned long long zero = 0;
    const unsigned none = 0xFFFFFFFFu;
    (void)hipMemcpy(d_bad, &zero, sizeof(zero), hipMemcpyHostToDevice);
    (void)hipMemcpy(d_first, &none, sizeof(none), hipMemcpyHostToDevice);
    constexpr int block = 256;
    check<<<kThreads / block, block>>>(d_bad, d_first);
    const hipError_t e = hipDeviceSynchronize();
    unsigned long long bad = 0;
    unsigned first = 0;
    (void)hipMemcpy(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost);
    (void)hipMemcpy(&first, d_first, sizeof(first), hipMemcpyDeviceToHost);
    std::printf("checked %llu mismatches %llu first 0x%08x status %s\n",
                (unsigned long long)kThreads * kPerThread, bad, first, hipGetErrorString(e));
    return (e == hipSuccess && bad == 0) ? 0 : 1;
}
