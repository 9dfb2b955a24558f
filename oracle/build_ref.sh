#!/bin/bash
# Builds the reference's own device code for the GPU box -- TEST INFRASTRUCTURE.
#
#   oracle/_ref/ref_kernels.co    src/hp/cl_src/types.h.cl + kernel.cl, unmodified,
#                                 compiled in place from /root/reference with ROCm's
#                                 clang and its OpenCL device library for gfx950
#   oracle/_ref/ref_kernels_notex.co  the same sources and options, with the texture
#                                 argument of s1_run given as an int: HIP launches
#                                 a code object's kernels from raw argument bytes and
#                                 cannot pass an OpenCL image object, so for the
#                                 whole-render harness the image type becomes `int`
#                                 and read_imagef a zero vector (-D options, the source
#                                 is untouched).  Valid for untextured scenes only,
#                                 where s1_run never reaches a texture read.
#   oracle/_ref/ref_kernels_notex_contract.co  the untextured build with multiply-adds
#                                 contracted into fused ones (-ffp-contract=fast, as
#                                 NVIDIA's OpenCL compiler contracts by default), to
#                                 show which image features depend on contraction
#   oracle/_ref/libref_harness.so oracle/ref_harness.hip: loads those code objects and
#                                 launches kdtree_intersect the way TraceRunner::run
#                                 does (trace_runner.cpp:192-206), or the whole render
#                                 loop (trace_runner.cpp:116-327)
#
# Build options: -cl-std=CL1.2 (the reference's sampler branch), -include types.h.cl
# (the host prepends it, trace_runner.cpp:47-52), -cl-single-precision-constant
# (double literals such as `/ 2.0` at kernel.cl:463 are float on a device
# without fp64; clang rejects them otherwise), -cl-fp32-correctly-rounded-divide-sqrt
# and -ffp-contract=off (IEEE division and no contraction, the arithmetic the
# parity comparison is defined on; the _contract build excepted), and no packed FP32 instructions
# (v_pk_add/mul/fma_f32 have given wrong results on these machines now and
# then; the library's own device code is built the same way).  Only the
# outputs above are written; no reference source is copied.
set -euo pipefail
REF=${HPT_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
CLANG=/opt/rocm/lib/llvm/bin/clang
if [ ! -f "$REF/src/hp/cl_src/kernel.cl" ]; then
    echo "build_ref: $REF not present, skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
"$CLANG" -x cl -cl-std=CL1.2 -target amdgcn-amd-amdhsa -mcpu=gfx950 \
    --rocm-path=/opt/rocm -Xclang -finclude-default-header \
    -include "$REF/src/hp/cl_src/types.h.cl" \
    -cl-single-precision-constant -cl-fp32-correctly-rounded-divide-sqrt -ffp-contract=off -O3 \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    "$REF/src/hp/cl_src/kernel.cl" -o "$OUT/ref_kernels.co.tmp"
mv "$OUT/ref_kernels.co.tmp" "$OUT/ref_kernels.co"
"$CLANG" -x cl -cl-std=CL1.2 -target amdgcn-amd-amdhsa -mcpu=gfx950 \
    --rocm-path=/opt/rocm -Xclang -finclude-default-header \
    -include "$REF/src/hp/cl_src/types.h.cl" \
    -cl-single-precision-constant -cl-fp32-correctly-rounded-divide-sqrt -ffp-contract=off -O3 \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -D__read_only= -Dimage2d_array_t=int '-Dread_imagef(i,s,c)=((float4)(0.0f))' \
    "$REF/src/hp/cl_src/kernel.cl" -o "$OUT/ref_kernels_notex.co.tmp"
mv "$OUT/ref_kernels_notex.co.tmp" "$OUT/ref_kernels_notex.co"
"$CLANG" -x cl -cl-std=CL1.2 -target amdgcn-amd-amdhsa -mcpu=gfx950 \
    --rocm-path=/opt/rocm -Xclang -finclude-default-header \
    -include "$REF/src/hp/cl_src/types.h.cl" \
    -cl-single-precision-constant -cl-fp32-correctly-rounded-divide-sqrt -ffp-contract=fast -O3 \
    -Xclang -target-feature -Xclang -packed-fp32-ops \
    -D__read_only= -Dimage2d_array_t=int '-Dread_imagef(i,s,c)=((float4)(0.0f))' \
    "$REF/src/hp/cl_src/kernel.cl" -o "$OUT/ref_kernels_notex_contract.co.tmp"
mv "$OUT/ref_kernels_notex_contract.co.tmp" "$OUT/ref_kernels_notex_contract.co"
/opt/rocm/bin/hipcc -std=c++17 -O2 -fPIC -shared "$HERE/ref_harness.hip" \
    -o "$OUT/libref_harness.so.tmp" -Wno-unused-command-line-argument
mv "$OUT/libref_harness.so.tmp" "$OUT/libref_harness.so"
echo "build_ref: wrote $OUT/ref_kernels.co, $OUT/ref_kernels_notex.co, $OUT/ref_kernels_notex_contract.co and $OUT/libref_harness.so" >&2
