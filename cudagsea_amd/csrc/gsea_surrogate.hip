// gsea_surrogate.hip -- the learned part of the engine: an MLP that predicts
// a gene set's null enrichment-score moments from set features, used to
// spend permutations adaptively.  Everything here is HIP written for this
// project (the reference has no counterpart):
//   k_set_features     16 features per set from the observed metric and ranks
//   k_null_moments     targets: mean/std of the positive and the negative null
//                      ES of a set (permutations 1..P-1)
//   k_mlp_fwd_bwd      16 -> 64 -> 64 -> 4 MLP with GELU, forward (+ backward
//                      to the layer inputs) per sample, fp32
//   k_mlp_wgrad        weight gradients, one thread per parameter, summed over
//                      the batch in a fixed order (deterministic)
//   k_adam             Adam update
//   k_mlp_infer_bf16   inference with bf16 weights and bf16 activations
//                      between layers, fp32 accumulation
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "gsea_device.hpp"
#include "gsea_hip.h"

namespace gsea {
int set_error(int code, const std::string &msg);
}

namespace gsea {

constexpr int kIn = GSEA_MLP_IN, kH = GSEA_MLP_HIDDEN, kOut = GSEA_MLP_OUT;
// packed parameter offsets: W1[kH][kIn] b1[kH] W2[kH][kH] b2[kH] W3[kOut][kH] b3[kOut]
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH,
              oW3 = oB2 + kH, oB3 = oW3 + kOut * kH, kNumParams = oB3 + kOut;
static_assert(kNumParams == GSEA_MLP_NUM_PARAMS, "parameter count");

__device__ __forceinline__ float gelu(float z) {
    return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float z) {
    const float cdf = 0.5f * (1.0f + erff(z * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * z * z);
    return cdf + z * pdf;
}

// ---------------------------------------------------------------------------
// features: one wavefront per set.  r0 = observed metric [N], pos0 = observed
// rank of each gene [N], glob = {std r, mean |r|, max |r|, skew r} over genes.
// ---------------------------------------------------------------------------
__global__ void k_set_features(const float *__restrict__ r0,
                               const int32_t *__restrict__ pos0, int N,
                               const int32_t *__restrict__ off,
                               const int32_t *__restrict__ genes, int G,
                               const float *__restrict__ glob,
                               float *__restrict__ feat) {
    const int s = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (s >= G) return;
    const int lane = lane_id();
    const int o0 = off[s], k = off[s + 1] - o0;
    double sr = 0, sr2 = 0, sr3 = 0, sa = 0, sa2 = 0, sq = 0, sq2 = 0, amax = 0;
    double c10 = 0, c25 = 0, b25 = 0, b10 = 0;
    const double n = (double)N;
    for (int j = lane; j < k; j += kWave) {
        const int g = genes[o0 + j];
        const double r = r0[g];
        const double a = fabs(r);
        const double q = (double)pos0[g] / n;
        sr += r; sr2 += r * r; sr3 += r * r * r;
        sa += a; sa2 += a * a; sq += q; sq2 += q * q;
        amax = a > amax ? a : amax;
        c10 += q < 0.1; c25 += q < 0.25; b25 += q >= 0.75; b10 += q >= 0.9;
    }
    sr = wave_sum(sr); sr2 = wave_sum(sr2); sr3 = wave_sum(sr3);
    sa = wave_sum(sa); sa2 = wave_sum(sa2); sq = wave_sum(sq); sq2 = wave_sum(sq2);
    c10 = wave_sum(c10); c25 = wave_sum(c25); b25 = wave_sum(b25); b10 = wave_sum(b10);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double t = __shfl_xor(amax, o, kWave);
        amax = t > amax ? t : amax;
    }
    if (lane != 0) return;
    const double kk = k > 0 ? (double)k : 1.0;
    const double gstd = glob[0] > 0 ? glob[0] : 1.0;
    const double gabs = glob[1] > 0 ? glob[1] : 1.0;
    const double gmax = glob[2] > 0 ? glob[2] : 1.0;
    const double mr = sr / kk, vr = fmax(sr2 / kk - mr * mr, 0.0);
    const double sdr = sqrt(vr);
    const double skew = sdr > 0 ? (sr3 / kk - 3 * mr * vr - mr * mr * mr) / (sdr * sdr * sdr) : 0.0;
    const double ma = sa / kk, va = fmax(sa2 / kk - ma * ma, 0.0);
    const double mq = sq / kk, vq = fmax(sq2 / kk - mq * mq, 0.0);
    float *f = feat + (size_t)s * kIn;
    f[0] = (float)(log(kk) / log(n > 1 ? n : 2.0));
    f[1] = (float)(kk / n);
    f[2] = (float)(mr / gstd);
    f[3] = (float)(sdr / gstd);
    f[4] = (float)fmax(fmin(skew, 10.0), -10.0);
    f[5] = (float)(ma / gabs);
    f[6] = (float)(sqrt(va) / gabs);
    f[7] = (float)mq;
    f[8] = (float)sqrt(vq);
    f[9] = (float)(c10 / kk);
    f[10] = (float)(c25 / kk);
    f[11] = (float)(b25 / kk);
    f[12] = (float)(b10 / kk);
    f[13] = (float)(amax / gmax);
    f[14] = (float)(gstd / gabs);
    f[15] = (float)fmax(fmin((double)glob[3], 10.0), -10.0);
}

// rank of each gene from the gene at each rank
__global__ void k_invert_ranks(const int32_t *__restrict__ rank_genes, int N,
                               int32_t *__restrict__ pos) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) pos[rank_genes[i]] = i;
}

// global summaries of the observed metric (one workgroup)
__global__ void k_metric_summary(const float *__restrict__ r0, int N,
                                 float *__restrict__ glob) {
    __shared__ double part[5][16];
    double s1 = 0, s2 = 0, s3 = 0, sa = 0, mx = 0;
    for (int g = threadIdx.x; g < N; g += blockDim.x) {
        const double r = r0[g];
        s1 += r; s2 += r * r; s3 += r * r * r;
        const double a = fabs(r);
        sa += a;
        mx = a > mx ? a : mx;
    }
    s1 = wave_sum(s1); s2 = wave_sum(s2); s3 = wave_sum(s3); sa = wave_sum(sa);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const double t = __shfl_xor(mx, o, kWave);
        mx = t > mx ? t : mx;
    }
    const int w = threadIdx.x / kWave;
    if (lane_id() == 0) {
        part[0][w] = s1; part[1][w] = s2; part[2][w] = s3; part[3][w] = sa; part[4][w] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t1 = 0, t2 = 0, t3 = 0, ta = 0, tm = 0;
        for (int i = 0; i < (int)(blockDim.x / kWave); ++i) {
            t1 += part[0][i]; t2 += part[1][i]; t3 += part[2][i]; ta += part[3][i];
            tm = part[4][i] > tm ? part[4][i] : tm;
        }
        const double n = (double)N;
        const double m = t1 / n, v = fmax(t2 / n - m * m, 0.0), sd = sqrt(v);
        glob[0] = (float)sd;
        glob[1] = (float)(ta / n);
        glob[2] = (float)tm;
        glob[3] = (float)(sd > 0 ? (t3 / n - 3 * m * v - m * m * m) / (sd * sd * sd) : 0.0);
    }
}

// ---------------------------------------------------------------------------
// null moments per set from es[P][G] (permutations 1..P-1):
// out[s] = {mean ES+, std ES+, mean |ES-|, std |ES-|, fraction positive}
// ---------------------------------------------------------------------------
__global__ void k_null_moments(const float *__restrict__ es, int64_t P, int G,
                               float *__restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= G) return;
    double np = 0, sp = 0, sp2 = 0, nn = 0, sn = 0, sn2 = 0;
    for (int64_t p = 1; p < P; ++p) {
        const double v = es[p * G + s];
        if (v != v) continue;
        if (v > 0) { np += 1; sp += v; sp2 += v * v; }
        else { nn += 1; sn -= v; sn2 += v * v; }
    }
    const double mp = np > 0 ? sp / np : 0.0, mn = nn > 0 ? sn / nn : 0.0;
    float *o = out + (size_t)s * 5;
    o[0] = (float)mp;
    o[1] = (float)(np > 1 ? sqrt(fmax(sp2 / np - mp * mp, 0.0)) : 0.0);
    o[2] = (float)mn;
    o[3] = (float)(nn > 1 ? sqrt(fmax(sn2 / nn - mn * mn, 0.0)) : 0.0);
    o[4] = (float)(np + nn > 0 ? np / (np + nn) : 0.5);
}

// ---------------------------------------------------------------------------
// MLP forward (+ backward to layer inputs), one thread per sample.  Weights in
// LDS.  Scratch per sample: z1, h1, z2, h2 (forward) and d1, d2, d3 (backward)
// kept for the weight-gradient kernel.
// ---------------------------------------------------------------------------
struct MlpScratch {
    float *z1, *h1, *z2, *h2, *d1, *d2, *d3;   // [B][kH] except d3 [B][kOut]
};

// norm (optional, inference): x_mean[16] x_inv_std[16] y_scale[4] y_shift[4]
constexpr int kNormLen = 2 * kIn + 2 * kOut;

template <bool BACKWARD>
__global__ __launch_bounds__(256) void k_mlp_fwd_bwd(
        const float *__restrict__ params, const float *__restrict__ x,
        const float *__restrict__ t, int B, float *__restrict__ y,
        MlpScratch sc, float *__restrict__ loss_part,
        const float *__restrict__ norm) {
    __shared__ float W[kNumParams];
    __shared__ float red[256 / kWave];
    for (int i = threadIdx.x; i < kNumParams; i += blockDim.x) W[i] = params[i];
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    float sq = 0.0f;
    if (b < B) {
        float xi[kIn];
#pragma unroll
        for (int j = 0; j < kIn; ++j) {
            xi[j] = x[(size_t)b * kIn + j];
            if (norm) xi[j] = (xi[j] - norm[j]) * norm[kIn + j];
        }
        float h1[kH];
        for (int i = 0; i < kH; ++i) {
            float z = W[oB1 + i];
#pragma unroll
            for (int j = 0; j < kIn; ++j) z += W[oW1 + i * kIn + j] * xi[j];
            h1[i] = gelu(z);
            if (BACKWARD) {
                sc.z1[(size_t)b * kH + i] = z;
                sc.h1[(size_t)b * kH + i] = h1[i];
            }
        }
        float h2[kH];
        for (int i = 0; i < kH; ++i) {
            float z = W[oB2 + i];
#pragma unroll 8
            for (int j = 0; j < kH; ++j) z += W[oW2 + i * kH + j] * h1[j];
            h2[i] = gelu(z);
            if (BACKWARD) {
                sc.z2[(size_t)b * kH + i] = z;
                sc.h2[(size_t)b * kH + i] = h2[i];
            }
        }
        float out[kOut];
#pragma unroll
        for (int o = 0; o < kOut; ++o) {
            float z = W[oB3 + o];
#pragma unroll 8
            for (int j = 0; j < kH; ++j) z += W[oW3 + o * kH + j] * h2[j];
            out[o] = z;
            if (y) y[(size_t)b * kOut + o] =
                norm ? z * norm[2 * kIn + o] + norm[2 * kIn + kOut + o] : z;
        }
        if (BACKWARD) {
            // mean squared error over B x kOut values
            const float scale = 2.0f / ((float)B * (float)kOut);
            float d3[kOut];
#pragma unroll
            for (int o = 0; o < kOut; ++o) {
                const float e = out[o] - t[(size_t)b * kOut + o];
                sq += e * e;
                d3[o] = scale * e;
                sc.d3[(size_t)b * kOut + o] = d3[o];
            }
            float d2[kH];
            for (int i = 0; i < kH; ++i) {
                float a = 0.0f;
#pragma unroll
                for (int o = 0; o < kOut; ++o) a += W[oW3 + o * kH + i] * d3[o];
                d2[i] = a * gelu_grad(sc.z2[(size_t)b * kH + i]);
                sc.d2[(size_t)b * kH + i] = d2[i];
            }
            for (int j = 0; j < kH; ++j) {
                float a = 0.0f;
#pragma unroll 8
                for (int i = 0; i < kH; ++i) a += W[oW2 + i * kH + j] * d2[i];
                sc.d1[(size_t)b * kH + j] = a * gelu_grad(sc.z1[(size_t)b * kH + j]);
            }
        }
    }
    if (BACKWARD && loss_part) {
        sq = wave_sum(sq);
        if (lane_id() == 0) red[threadIdx.x / kWave] = sq;
        __syncthreads();
        if (threadIdx.x == 0) {
            float tot = 0.0f;
            for (int i = 0; i < (int)(blockDim.x / kWave); ++i) tot += red[i];
            loss_part[blockIdx.x] = tot;
        }
    }
}

// weight gradients: thread = parameter, fixed-order sum over the batch
__global__ void k_mlp_wgrad(const float *__restrict__ x, MlpScratch sc, int B,
                            float *__restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    const float *dv, *av;
    int dstride, astride, di, ai;
    bool bias = false;
    if (p < oB1) { di = (p - oW1) / kIn; ai = (p - oW1) % kIn; dv = sc.d1; dstride = kH; av = x; astride = kIn; }
    else if (p < oW2) { di = p - oB1; ai = 0; dv = sc.d1; dstride = kH; av = nullptr; astride = 0; bias = true; }
    else if (p < oB2) { di = (p - oW2) / kH; ai = (p - oW2) % kH; dv = sc.d2; dstride = kH; av = sc.h1; astride = kH; }
    else if (p < oW3) { di = p - oB2; ai = 0; dv = sc.d2; dstride = kH; av = nullptr; astride = 0; bias = true; }
    else if (p < oB3) { di = (p - oW3) / kH; ai = (p - oW3) % kH; dv = sc.d3; dstride = kOut; av = sc.h2; astride = kH; }
    else { di = p - oB3; ai = 0; dv = sc.d3; dstride = kOut; av = nullptr; astride = 0; bias = true; }
    float acc = 0.0f;
    for (int b = 0; b < B; ++b) {
        const float d = dv[(size_t)b * dstride + di];
        acc += bias ? d : d * av[(size_t)b * astride + ai];
    }
    grad[p] = acc;
}

__global__ void k_adam(float *__restrict__ params, float *__restrict__ m,
                       float *__restrict__ v, const float *__restrict__ grad,
                       float lr, float b1, float b2, float eps, float wd,
                       float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNumParams) return;
    const float g = grad[p] + wd * params[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// bf16 inference: weights and inter-layer activations rounded to bf16,
// products accumulated in fp32
__global__ __launch_bounds__(256) void k_mlp_infer_bf16(
        const float *__restrict__ params, const float *__restrict__ x, int B,
        float *__restrict__ y, const float *__restrict__ norm) {
    __shared__ __hip_bfloat16 W[kNumParams];
    for (int i = threadIdx.x; i < kNumParams; i += blockDim.x)
        W[i] = __float2bfloat16(params[i]);
    __syncthreads();
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    float xi[kIn];
#pragma unroll
    for (int j = 0; j < kIn; ++j) {
        float v = x[(size_t)b * kIn + j];
        if (norm) v = (v - norm[j]) * norm[kIn + j];
        xi[j] = __bfloat162float(__float2bfloat16(v));
    }
    float h1[kH];
    for (int i = 0; i < kH; ++i) {
        float z = __bfloat162float(W[oB1 + i]);
#pragma unroll
        for (int j = 0; j < kIn; ++j) z += __bfloat162float(W[oW1 + i * kIn + j]) * xi[j];
        h1[i] = __bfloat162float(__float2bfloat16(gelu(z)));
    }
    float h2[kH];
    for (int i = 0; i < kH; ++i) {
        float z = __bfloat162float(W[oB2 + i]);
#pragma unroll 8
        for (int j = 0; j < kH; ++j) z += __bfloat162float(W[oW2 + i * kH + j]) * h1[j];
        h2[i] = __bfloat162float(__float2bfloat16(gelu(z)));
    }
#pragma unroll
    for (int o = 0; o < kOut; ++o) {
        float z = __bfloat162float(W[oB3 + o]);
#pragma unroll 8
        for (int j = 0; j < kH; ++j) z += __bfloat162float(W[oW3 + o * kH + j]) * h2[j];
        y[(size_t)b * kOut + o] = norm ? z * norm[2 * kIn + o] + norm[2 * kIn + kOut + o] : z;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int hip_fail(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? GSEA_ERR_OUT_OF_MEMORY : GSEA_ERR_HIP,
                     std::string(what) + ": " + hipGetErrorString(e));
}

#define SG_TRY(expr)                                   \
    do {                                               \
        hipError_t e_ = (expr);                        \
        if (e_ != hipSuccess) return hip_fail(e_, #expr); \
    } while (0)

struct MlpWork {
    float *buf = nullptr;
    size_t cap = 0;   // floats
};
// Scratch of gsea_mlp_grad (g_work) and gsea_set_features (g_pos), shared
// by every caller in the process.  Both entry points hold g_sg_mu from the
// first use of their scratch until their stream has drained, so a buffer is
// never regrown or reused while a kernel of another call still reads it,
// whichever threads or streams the calls come from.
static MlpWork g_work, g_pos;
static std::mutex g_sg_mu;

static int scratch(int B, MlpScratch *sc, float **loss_part, float **grad) {
    const size_t need = (size_t)B * (6 * kH + kOut) + (size_t)(B / 256 + 2) + kNumParams;
    if (g_work.cap < need) {
        if (g_work.buf) SG_TRY(hipFree(g_work.buf));
        g_work.buf = nullptr;
        g_work.cap = 0;
        SG_TRY(hipMalloc(&g_work.buf, need * sizeof(float)));
        g_work.cap = need;
    }
    float *p = g_work.buf;
    sc->z1 = p; p += (size_t)B * kH;
    sc->h1 = p; p += (size_t)B * kH;
    sc->z2 = p; p += (size_t)B * kH;
    sc->h2 = p; p += (size_t)B * kH;
    sc->d1 = p; p += (size_t)B * kH;
    sc->d2 = p; p += (size_t)B * kH;
    sc->d3 = p; p += (size_t)B * kOut;
    *loss_part = p; p += (size_t)(B / 256 + 2);
    *grad = p;
    return GSEA_OK;
}

}  // namespace gsea

using namespace gsea;

extern "C" {

int gsea_mlp_num_params(void) { return kNumParams; }

int gsea_mlp_forward(const float *d_params, const float *d_x, int64_t B,
                     int bf16, const float *d_norm, float *d_y, void *stream) {
    if (B <= 0 || B > (1LL << 30) || !d_params || !d_x || !d_y)
        return set_error(GSEA_ERR_INVALID_ARGUMENT, "bad mlp forward arguments");
    hipStream_t st = (hipStream_t)stream;
    const int blocks = (int)((B + 255) / 256);
    if (bf16) {
        hipLaunchKernelGGL(k_mlp_infer_bf16, dim3(blocks), dim3(256), 0, st, d_params,
                           d_x, (int)B, d_y, d_norm);
    } else {
        MlpScratch sc{};
        hipLaunchKernelGGL(k_mlp_fwd_bwd<false>, dim3(blocks), dim3(256), 0, st,
                           d_params, d_x, (const float *)nullptr, (int)B, d_y, sc,
                           (float *)nullptr, d_norm);
    }
    SG_TRY(hipGetLastError());
    SG_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

int gsea_mlp_grad(const float *d_params, const float *d_x, const float *d_t,
                  int64_t B, float *d_grad, float *loss, void *stream) {
    if (B <= 0 || B > (1LL << 26) || !d_params || !d_x || !d_t || !d_grad)
        return set_error(GSEA_ERR_INVALID_ARGUMENT, "bad mlp grad arguments");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_sg_mu);
    MlpScratch sc;
    float *loss_part, *grad;
    int rc = scratch((int)B, &sc, &loss_part, &grad);
    if (rc) return rc;
    const int blocks = (int)((B + 255) / 256);
    hipLaunchKernelGGL(k_mlp_fwd_bwd<true>, dim3(blocks), dim3(256), 0, st, d_params,
                       d_x, d_t, (int)B, (float *)nullptr, sc, loss_part,
                       (const float *)nullptr);
    SG_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_wgrad, dim3((kNumParams + 255) / 256), dim3(256), 0, st,
                       d_x, sc, (int)B, d_grad);
    SG_TRY(hipGetLastError());
    if (loss) {
        std::vector<float> parts(blocks);
        SG_TRY(hipMemcpyAsync(parts.data(), loss_part, sizeof(float) * blocks,
                              hipMemcpyDeviceToHost, st));
        SG_TRY(hipStreamSynchronize(st));
        double tot = 0;
        for (float v : parts) tot += v;
        *loss = (float)(tot / ((double)B * kOut));
    } else {
        SG_TRY(hipStreamSynchronize(st));
    }
    return GSEA_OK;
}

int gsea_mlp_adam(float *d_params, float *d_m, float *d_v, const float *d_grad,
                  int64_t step, float lr, float weight_decay, void *stream) {
    if (step < 1) return set_error(GSEA_ERR_INVALID_ARGUMENT, "adam step starts at 1");
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((kNumParams + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, d_params, d_m, d_v, d_grad, lr, b1, b2,
                       1e-8f, weight_decay, c1, c2);
    SG_TRY(hipGetLastError());
    return GSEA_OK;
}

int gsea_set_features(const float *d_r0, const int32_t *d_rank_genes, int64_t N,
                      const int32_t *d_set_offsets, const int32_t *d_set_genes,
                      int64_t G, float *d_glob, float *d_feat, void *stream) {
    if (N <= 0 || N > GSEA_MAX_GENES || G <= 0 || !d_r0 || !d_rank_genes || !d_feat ||
        !d_glob)
        return set_error(GSEA_ERR_INVALID_ARGUMENT, "bad feature arguments");
    hipStream_t st = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_sg_mu);
    if (g_pos.cap < (size_t)N) {
        if (g_pos.buf) SG_TRY(hipFree(g_pos.buf));
        g_pos.buf = nullptr;
        g_pos.cap = 0;
        SG_TRY(hipMalloc(&g_pos.buf, sizeof(int32_t) * GSEA_MAX_GENES));
        g_pos.cap = GSEA_MAX_GENES;
    }
    int32_t *d_pos0 = reinterpret_cast<int32_t *>(g_pos.buf);
    hipLaunchKernelGGL(k_invert_ranks, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                       d_rank_genes, (int)N, d_pos0);
    SG_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_metric_summary, dim3(1), dim3(1024), 0, st, d_r0, (int)N, d_glob);
    SG_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_set_features, dim3((unsigned)((G + 3) / 4)), dim3(256), 0, st,
                       d_r0, d_pos0, (int)N, d_set_offsets, d_set_genes, (int)G,
                       (const float *)d_glob, d_feat);
    SG_TRY(hipGetLastError());
    SG_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

int gsea_null_moments(const float *d_es_pg, int64_t P, int64_t G, float *d_out,
                      void *stream) {
    if (P <= 1 || G <= 0 || !d_es_pg || !d_out)
        return set_error(GSEA_ERR_INVALID_ARGUMENT, "bad null-moment arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_null_moments, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, st,
                       d_es_pg, P, (int)G, d_out);
    SG_TRY(hipGetLastError());
    SG_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

}  // extern "C"
