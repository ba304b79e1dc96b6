// gb_probe.hip -- VALU issue-rate probes for the benchmark's roofline (not part of the
// verifier's C ABI).  Each kernel runs a fixed number of vector ALU instructions per wave
// (written as inline asm so the count is exact), so the rate a launch reaches is
// instructions / time, and rocprofv3's SQ_INSTS_VALU / SQ_BUSY_CYCLES of the same launch give
// the count and the clock the hardware saw.
//
//   kind 0: eight independent chains of v_add_u32 (the issue ceiling of simple 32-bit ops)
//   kind 1: eight independent chains of v_mad_u64_u32 (the 32x32->64 product the Montgomery
//           arithmetic is built from)
//   kind 2: one dependent chain of v_add_u32 (latency-bound issue)
//
// LDS probes (gb_probe_lds): two 512-thread workgroups per CU, each with 64 KB of LDS, every
// thread issuing a fixed count of 32-bit LDS operations at addresses stepped by a per-lane odd
// stride (two VALU instructions per operation, masked into the array):
//
//   kind 0: ds_and_b32 at scattered words (the sieve's marks)
//   kind 1: ds_and_b32 at words 64x + lane (no two lanes of a wave on one bank)
//   kind 2: ds_read_b32 at scattered words
//   kind 3: ds_and_b32 at scattered words from the even lanes only (half the exec mask)
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

constexpr int BS = 256;

template <int KIND>
__global__ __launch_bounds__(BS) void valu_probe(uint32_t iters, uint32_t* out) {
    uint32_t a[8];
    uint64_t w[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = threadIdx.x + k;
        w[k] = threadIdx.x * 3u + k;
    }
    const uint32_t b = blockIdx.x | 1u;
    for (uint32_t i = 0; i < iters; ++i) {
        if constexpr (KIND == 0) {
#pragma unroll
            for (int k = 0; k < 8; ++k) asm volatile("v_add_u32 %0, %1, %0" : "+v"(a[k]) : "v"(b));
        } else if constexpr (KIND == 1) {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(w[k]) : "v"(a[k]), "v"(b) : "vcc");
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) asm volatile("v_add_u32 %0, %1, %0" : "+v"(a[0]) : "v"(b));
        }
    }
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) r += a[k] + (uint32_t)w[k] + (uint32_t)(w[k] >> 32);
    out[blockIdx.x * BS + threadIdx.x] = r;
}

constexpr int LDS_BS = 512;
constexpr uint32_t LDS_WORDS = 16384;   // 64 KB per workgroup

template <int KIND>
__global__ __launch_bounds__(LDS_BS) void lds_probe(uint32_t iters, uint32_t* out) {
    __shared__ uint32_t w[LDS_WORDS];
    const uint32_t tid = threadIdx.x, lane = tid & 63;
    for (uint32_t i = tid; i < LDS_WORDS; i += LDS_BS) w[i] = ~0u;
    __syncthreads();
    constexpr uint32_t AMASK = (LDS_WORDS - 1) * 4;   // byte offsets of whole words
    uint32_t b = ((tid * 2654435761u) >> 8) & AMASK;
    const uint32_t s4 = (((tid * 40503u + blockIdx.x * 97u) | 1u) * 4u) & AMASK;
    const uint32_t m = ~(1u << (tid & 31));
    uint32_t acc = 0;
    char* base = reinterpret_cast<char*>(w);
    for (uint32_t k = 0; k < iters; ++k) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            b = (b + s4) & AMASK;
            const uint32_t a = KIND == 1 ? ((b & ~255u) | (lane * 4)) : b;
            uint32_t* pw = reinterpret_cast<uint32_t*>(base + a);
            if constexpr (KIND == 2) acc += *pw;
            else if constexpr (KIND == 3) { if (!(lane & 1)) atomicAnd(pw, m); }
            else atomicAnd(pw, m);
        }
    }
    __syncthreads();
    uint32_t x = acc;
    for (uint32_t i = tid; i < LDS_WORDS; i += LDS_BS) x ^= w[i];
    out[blockIdx.x * LDS_BS + tid] = x;
}

}  // namespace

extern "C" {

// Runs probe `kind` over `blocks` workgroups of 256 threads, `iters` iterations of 8 probe
// instructions per thread; *ms = the launch's device time, *wave_insts = probe instructions
// issued (per wave, 8 x iters).  Returns a hipError_t value (0 = success).
int gb_probe_valu(int kind, uint32_t blocks, uint32_t iters, double* ms, double* wave_insts) {
    if (kind < 0 || kind > 2 || blocks == 0 || !ms || !wave_insts) return (int)hipErrorInvalidValue;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&out), (size_t)blocks * BS * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
        if (kind == 0) hipLaunchKernelGGL(valu_probe<0>, dim3(blocks), dim3(BS), 0, nullptr, iters, out);
        else if (kind == 1) hipLaunchKernelGGL(valu_probe<1>, dim3(blocks), dim3(BS), 0, nullptr, iters, out);
        else hipLaunchKernelGGL(valu_probe<2>, dim3(blocks), dim3(BS), 0, nullptr, iters, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    *ms = f;
    *wave_insts = (double)blocks * (BS / 64) * 8.0 * iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return (int)e;
}

// Runs LDS probe `kind` over `blocks` workgroups of 512 threads (two per CU hold 128 KB of
// its LDS), `iters` iterations of 8 LDS operations per thread; *ms = the launch's device time,
// *wave_insts = the probe's wave-level LDS instructions (8 x iters per wave).  Returns a
// hipError_t value (0 = success).
int gb_probe_lds(int kind, uint32_t blocks, uint32_t iters, double* ms, double* wave_insts) {
    if (kind < 0 || kind > 3 || blocks == 0 || !ms || !wave_insts) return (int)hipErrorInvalidValue;
    uint32_t* out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&out), (size_t)blocks * LDS_BS * 4);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
        if (kind == 0) hipLaunchKernelGGL(lds_probe<0>, dim3(blocks), dim3(LDS_BS), 0, nullptr, iters, out);
        else if (kind == 1) hipLaunchKernelGGL(lds_probe<1>, dim3(blocks), dim3(LDS_BS), 0, nullptr, iters, out);
        else if (kind == 2) hipLaunchKernelGGL(lds_probe<2>, dim3(blocks), dim3(LDS_BS), 0, nullptr, iters, out);
        else hipLaunchKernelGGL(lds_probe<3>, dim3(blocks), dim3(LDS_BS), 0, nullptr, iters, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float f = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&f, e0, e1);
    *ms = f;
    *wave_insts = (double)blocks * (LDS_BS / 64) * 8.0 * iters;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out) (void)hipFree(out);
    return (int)e;
}

}  // extern "C"
