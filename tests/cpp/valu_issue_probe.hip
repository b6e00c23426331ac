// Issue rate of single VALU instruction types on this GPU, in wavefront
// instructions per CU per 2.4 GHz clock.  Each lane runs 8 independent
// chains of one instruction (inline asm, so the count is exact), 16 or 32
// wavefronts per CU, 4,096 iterations; five timed launches after one warm-up.
// Prints one JSON object per (instruction, wavefronts per CU).
#include <hip/hip_runtime.h>
#include <cstdio>

constexpr int kIters = 4096;
constexpr int kChains = 8;

#define PROBE(NAME, BODY)                                                        \
    __global__ __launch_bounds__(256) void NAME(float *out, float seed) {         \
        float f[kChains];                                                         \
        double d[kChains];                                                        \
        unsigned u[kChains];                                                      \
        for (int i = 0; i < kChains; ++i) {                                       \
            f[i] = seed + threadIdx.x + i;                                        \
            d[i] = seed * 0.5 + threadIdx.x + i;                                  \
            u[i] = threadIdx.x * 7u + i + 1u;                                     \
        }                                                                         \
        for (int it = 0; it < kIters; ++it) {                                     \
            _Pragma("unroll") for (int i = 0; i < kChains; ++i) { BODY; }         \
        }                                                                         \
        float s = 0.0f;                                                           \
        for (int i = 0; i < kChains; ++i) s += f[i] + (float)d[i] + (float)u[i];  \
        out[blockIdx.x * blockDim.x + threadIdx.x] = s;                           \
    }

PROBE(p_add_f32, asm volatile("v_add_f32 %0, %0, 1.0" : "+v"(f[i])))
PROBE(p_fma_f32, asm volatile("v_fma_f32 %0, %0, %0, 1.0" : "+v"(f[i])))
PROBE(p_max_f32, asm volatile("v_max_f32 %0, %0, %1" : "+v"(f[i]) : "v"(f[(i + 1) % kChains])))
PROBE(p_add_u32, asm volatile("v_add_u32 %0, %0, 7" : "+v"(u[i])))
PROBE(p_and_b32, asm volatile("v_and_b32 %0, %0, %1" : "+v"(u[i]) : "v"(u[(i + 1) % kChains])))
PROBE(p_lshl_add_u32,
      asm volatile("v_lshl_add_u32 %0, %0, 2, %1" : "+v"(u[i]) : "v"(u[(i + 1) % kChains])))
PROBE(p_ffbl_b32, asm volatile("v_ffbl_b32 %0, %0" : "+v"(u[i])))
PROBE(p_cvt_f32_u32, asm volatile("v_cvt_f32_u32 %0, %1" : "=v"(f[i]) : "v"(u[i])))
PROBE(p_cvt_f64_f32, asm volatile("v_cvt_f64_f32 %0, %1" : "=v"(d[i]) : "v"(f[i])))
PROBE(p_fma_f64, asm volatile("v_fma_f64 %0, %0, %0, 1.0" : "+v"(d[i])))
PROBE(p_min_f64, asm volatile("v_min_f64 %0, %0, %1" : "+v"(d[i]) : "v"(d[(i + 1) % kChains])))
PROBE(p_pk_min_u16,
      asm volatile("v_pk_min_u16 %0, %0, %1" : "+v"(u[i]) : "v"(u[(i + 1) % kChains])))
PROBE(p_mov_dpp, asm volatile("v_mov_b32_dpp %0, %1 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf"
                              : "=v"(u[i]) : "v"(u[(i + 1) % kChains])))

typedef void (*Kern)(float *, float);

int main() {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 1;
    const int cus = p.multiProcessorCount;
    float *out = nullptr;
    if (hipMalloc(&out, sizeof(float) * (size_t)cus * 8 * 256) != hipSuccess) return 1;
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 1;
    const struct { const char *name; Kern k; } probes[] = {
        {"v_add_f32", p_add_f32},           {"v_fma_f32", p_fma_f32},
        {"v_max_f32", p_max_f32},           {"v_add_u32", p_add_u32},
        {"v_and_b32", p_and_b32},           {"v_lshl_add_u32", p_lshl_add_u32},
        {"v_ffbl_b32", p_ffbl_b32},         {"v_cvt_f32_u32", p_cvt_f32_u32},
        {"v_cvt_f64_f32", p_cvt_f64_f32},   {"v_fma_f64", p_fma_f64},
        {"v_min_f64", p_min_f64},           {"v_pk_min_u16", p_pk_min_u16},
        {"v_mov_b32_dpp", p_mov_dpp}};
    for (const auto &pr : probes) {
        for (int waves : {16, 32}) {
            const int blocks = cus * waves / 4;   // 4 wavefronts per 256-thread block
            hipLaunchKernelGGL(pr.k, dim3(blocks), dim3(256), 0, 0, out, 1.0f);
            if (hipDeviceSynchronize() != hipSuccess) return 1;
            (void)hipEventRecord(a);
            for (int r = 0; r < 5; ++r)
                hipLaunchKernelGGL(pr.k, dim3(blocks), dim3(256), 0, 0, out, 1.0f);
            (void)hipEventRecord(b);
            if (hipEventSynchronize(b) != hipSuccess) return 1;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, a, b);
            const double winst = 5.0 * blocks * 4 * (double)kIters * kChains;
            printf("{\"instruction\": \"%s\", \"waves_per_cu\": %d, "
                   "\"winst_per_cu_clock\": %.3f}\n",
                   pr.name, waves, winst / (ms * 1e-3) / cus / 2.4e9);
        }
    }
    (void)hipFree(out);
    return 0;
}
