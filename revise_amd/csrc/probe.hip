// Measurement helpers: kernels that measure the chip's own peaks for the
// roofline of the path's kernels (no reference counterpart).  Each runs on
// every CU with a grid of many waves and reports its work count; the caller
// times the launch (HIP events on the same stream) and divides.
//
//  * valu: non-packed fp32 fma issue -- 8 independent fma chains per lane
//    (no dependency stalls), as many waves as fit;
//  * lds: ds_read_b128 of distinct, conflict-free addresses per lane;
//  * gather: dword loads from an L1-resident buffer, either coalesced (a
//    wave reads 256 consecutive bytes: 4 64-byte sectors) or scattered (each
//    lane a different 64-byte sector: 64 sectors per load), the two ends of
//    the L1 tag-lookup cost the ray-casts' corner loads pay;
//  * stream: a non-temporal copy of a device buffer (read + write), all CUs.

#include "rv_internal.hpp"

namespace rv {
namespace {

__global__ __launch_bounds__(256) void probe_valu_kernel(float* out, uint32_t iters, float a, float b)
{
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
        x[k] = (float)(threadIdx.x + k);
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            x[k] = fmaf(x[k], a, b);
    }
    float s = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
        s += x[k];
    if (s == 1.2345f) // never true for the arguments used; keeps the chains live
        out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void probe_lds_kernel(float* out, uint32_t iters, uint32_t step)
{
    __shared__ float4 buf[1024]; // 16 KB
    for (uint32_t i = threadIdx.x; i < 1024u; i += 256u)
        buf[i] = make_float4((float)i, 1.f, 2.f, 3.f);
    __syncthreads();
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t j = threadIdx.x;
    for (uint32_t i = 0; i < iters; ++i) {
        // consecutive lanes, consecutive 16-byte entries: conflict-free; the
        // step comes from the host, so no two iterations' loads can be
        // proven to repeat
        const float4 v = buf[j & 1023u];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        j += step;
    }
    if (acc.x + acc.y + acc.z + acc.w == -1.f)
        out[blockIdx.x * blockDim.x + threadIdx.x] = acc.x;
}

// Dword loads of one 16 KB region that every workgroup reads (so it stays
// in each CU's L1): lane l of wave w in iteration i reads float
// (l * stride + w * 1024 + i * 16) mod 4096.
__global__ __launch_bounds__(256) void probe_gather_kernel(const float* __restrict__ src, float* out,
                                                           uint32_t iters, uint32_t stride, uint32_t step)
{
    const uint32_t lane = threadIdx.x & 63u;
    const float* base = src;
    float acc = 0.f;
    uint32_t o = lane * stride + (threadIdx.x >> 6) * 1024u;
    for (uint32_t i = 0; i < iters; ++i) {
        acc += base[o & 4095u];
        o += step; // 16 floats (one 64-byte sector), from the host
    }
    if (acc == -1.f)
        out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void probe_stream_kernel(const f32x4* __restrict__ src, f32x4* __restrict__ dst,
                                                           uint64_t n4)
{
    for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * 256u) {
        const f32x4 v = __builtin_nontemporal_load(src + i);
        __builtin_nontemporal_store(v, dst + i);
    }
}

} // namespace
} // namespace rv

using namespace rv;

extern "C" int rv_probe_valu(float* d_sink, uint32_t blocks, uint32_t iters, uint64_t* fma_count,
                             rv_stream_t stream)
{
    RV_REQUIRE(d_sink && fma_count && blocks > 0, "rv_probe_valu: bad arguments");
    hipLaunchKernelGGL(probe_valu_kernel, dim3(blocks), dim3(256), 0, S(stream), d_sink, iters, 0.999f, 0.001f);
    RV_LAUNCH_CHECK();
    *fma_count = (uint64_t)blocks * 256u * iters * 8u;
    return RV_OK;
}

extern "C" int rv_probe_lds(float* d_sink, uint32_t blocks, uint32_t iters, uint64_t* bytes, rv_stream_t stream)
{
    RV_REQUIRE(d_sink && bytes && blocks > 0, "rv_probe_lds: bad arguments");
    hipLaunchKernelGGL(probe_lds_kernel, dim3(blocks), dim3(256), 0, S(stream), d_sink, iters, 320u);
    RV_LAUNCH_CHECK();
    *bytes = (uint64_t)blocks * 256u * iters * 16u;
    return RV_OK;
}

extern "C" int rv_probe_gather(const float* d_src, float* d_sink, uint32_t blocks, uint32_t iters, int scattered,
                               uint64_t* loads, rv_stream_t stream)
{
    RV_REQUIRE(d_src && d_sink && loads && blocks > 0, "rv_probe_gather: bad arguments");
    hipLaunchKernelGGL(probe_gather_kernel, dim3(blocks), dim3(256), 0, S(stream), d_src, d_sink, iters,
                       scattered ? 16u : 1u, 16u);
    RV_LAUNCH_CHECK();
    *loads = (uint64_t)blocks * 4u * iters; // wave-level load instructions
    return RV_OK;
}

extern "C" int rv_probe_stream(const void* d_src, void* d_dst, uint64_t bytes, uint32_t blocks, rv_stream_t stream)
{
    RV_REQUIRE(d_src && d_dst && blocks > 0 && (bytes & 15u) == 0u, "rv_probe_stream: bad arguments");
    hipLaunchKernelGGL(probe_stream_kernel, dim3(blocks), dim3(256), 0, S(stream),
                       static_cast<const f32x4*>(d_src), static_cast<f32x4*>(d_dst), bytes >> 4);
    RV_LAUNCH_CHECK();
    return RV_OK;
}
