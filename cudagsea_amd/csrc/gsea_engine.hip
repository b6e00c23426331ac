// gsea_engine.hip -- host orchestration of the permutation-GSEA path on one
// MI355X and the extern "C" entry points declared in include/gsea_hip.h.
//
// Pipeline per permutation batch (the reference's gsea_proxy loop,
// cudaGSEA_wrapper.cu:295-372, without its per-tile host work):
//   k_class_lists -> k_metric -> k_rank -> k_es     (ES[P][G] in HBM)
// then once: k_set_stats -> k_perm_extrema -> k_fwer -> FDR kernels, and
// k_transpose when the caller wants the [G][P] matrix.
//
// Device memory comes from a grow-only per-device workspace, so repeated
// calls do not allocate.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gsea_hip.h"
#include "gsea_kernels.hpp"
#include "gsea_launch.hpp"

namespace gsea {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

static int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int set_error(int code, const std::string &msg) { return fail(code, msg); }

#define HIP_TRY(expr)                                                        \
    do {                                                                     \
        hipError_t e_ = (expr);                                              \
        if (e_ != hipSuccess)                                                \
            return fail(e_ == hipErrorOutOfMemory ? GSEA_ERR_OUT_OF_MEMORY   \
                                                  : GSEA_ERR_HIP,            \
                        std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

static const char *kMetricNames[14] = {
    "naive_diff_of_classes",  "naive_ratio_of_classes",
    "naive_log2_ratio_of_classes", "stable_diff_of_classes",
    "stable_ratio_of_classes", "stable_log2_ratio_of_classes",
    "onepass_signal2noise",   "onepass_t_test",
    "twopass_signal2noise",   "twopass_t_test",
    "stable_signal2noise",    "stable_t_test",
    "overkill_signal2noise",  "overkill_t_test"};

// ---------------------------------------------------------------------------
// per-device workspace
// ---------------------------------------------------------------------------
struct Buffer {
    void *ptr = nullptr;
    size_t cap = 0;
};

enum BufId {
    kBufLabels, kBufSig, kBufLists, kBufR, kBufES, kBufAGlob, kBufFactor,
    kBufMax, kBufMin, kBufThr, kBufHist, kBufX, kBufOff, kBufGenes,
    kBufStats, kBufEsOut, kBufPos, kBufBudget, kBufCv, kBufRcp,
    kBufFdrTable, kBufFdrInfo, kBufCheck, kBufStamps, kBufRankScratch, kNumBufs
};

// set tiers of one set list (slot 0: all sets; slots 1..: budget levels)
constexpr int kTierSlots = 24;
enum TierBuf { kTierIds, kTierGm, kTierDecay, kTierGenes, kNumTierBufs };

// the set tiers of the last set list (slot 0), reused while the list's
// pointers, sizes, split and contents (checksum) are unchanged
struct TierKey {
    const int32_t *off = nullptr, *genes = nullptr;
    int64_t G = -1, N = -1;
    int lane_max = -1, pair_max = -1;
    size_t tsize = 0;
    unsigned long long sum = 0, total = 0;
};

struct DeviceState {
    bool init = false;
    int lds_max = 65536;
    int cu_count = 256;
    Buffer bufs[kNumBufs];
    Buffer tier_bufs[kTierSlots][kNumTierBufs];
    int64_t rcp_len = 0;   // entries of the 1/(i+1) table in kBufRcp
    TierKey tier_key;
    bool tier_valid = false;
    hipStream_t own_stream = nullptr;   // capture stream for callers on the null stream
    std::vector<hipEvent_t> events;
};

static std::mutex g_mu;
static DeviceState g_dev[64];
static gsea_tuning g_tune = {};

static int current_device(int *dev) {
    HIP_TRY(hipGetDevice(dev));
    if (*dev < 0 || *dev >= 64) return fail(GSEA_ERR_HIP, "device id out of range");
    return GSEA_OK;
}

static int device_state(DeviceState **out) {
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    DeviceState &d = g_dev[dev];
    if (!d.init) {
        int v = 0;
        HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        d.lds_max = v;
        // the opt-in limit (dynamic LDS above the default) where reported
        int optin = 0;
        if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, dev) ==
                hipSuccess && optin > d.lds_max)
            d.lds_max = optin;
        (void)hipGetLastError();
        HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev));
        d.cu_count = v;
        d.init = true;
    }
    *out = &d;
    return GSEA_OK;
}

// bumped whenever a workspace buffer moves: captured graphs hold addresses
static uint64_t g_ws_generation = 0;

static int ensure_buf(Buffer &b, size_t bytes, void **out) {
    bytes = std::max<size_t>(bytes, 256);
    if (b.cap < bytes) {
        ++g_ws_generation;
        if (b.ptr) HIP_TRY(hipFree(b.ptr));
        b.ptr = nullptr;
        b.cap = 0;
        const size_t want = bytes + bytes / 8;
        HIP_TRY(hipMalloc(&b.ptr, want));
        b.cap = want;
    }
    *out = b.ptr;
    return GSEA_OK;
}

static int ensure(DeviceState *d, BufId id, size_t bytes, void **out) {
    return ensure_buf(d->bufs[id], bytes, out);
}

// events recorded around launches; durations summed per stage after sync
// device wall clock at a point of a stream (one thread, after the previous
// launch on the stream completes and before the next starts)
__global__ void k_stamp(unsigned long long *slot) { *slot = (unsigned long long)wall_clock64(); }

// Stage boundaries of a run: HIP events, or -- for sequences captured into a
// graph, where event nodes do not time -- device wall-clock stamps written
// by k_stamp into `stamps` (read after the stream is synchronised).
struct StageClock {
    DeviceState *d;
    hipStream_t st;
    std::vector<std::pair<int, std::pair<int, int>>> marks;  // stage, ev0, ev1
    int next = 0;
    int open_stage = -1, open_ev = -1;
    unsigned long long *stamps = nullptr;   // device; stamp mode when set
    int stamp_cap = 0;

    int event(int *idx) {
        if (stamps) {
            if (next >= stamp_cap) return fail(GSEA_ERR_INVALID_ARGUMENT, "too many stages");
            *idx = next++;
            hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, st, stamps + *idx);
            HIP_TRY(hipGetLastError());
            return GSEA_OK;
        }
        if (next >= (int)d->events.size()) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            d->events.push_back(e);
        }
        *idx = next++;
        HIP_TRY(hipEventRecord(d->events[*idx], st));
        return GSEA_OK;
    }
    int begin(int stage) {
        open_stage = stage;
        return event(&open_ev);
    }
    int end() {
        int e1;
        int rc = event(&e1);
        if (rc) return rc;
        marks.push_back({open_stage, {open_ev, e1}});
        return GSEA_OK;
    }
};

static int finish_timings(StageClock &clk, gsea_timings *t, int64_t batches) {
    if (!t) return GSEA_OK;
    std::memset(t, 0, sizeof(*t));
    t->perm_batches = batches;
    if (clk.marks.empty()) return GSEA_OK;
    std::vector<unsigned long long> ts;
    double ms_per_tick = 0.0;
    if (clk.stamps) {
        int dev = 0, khz = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        if (khz <= 0) return fail(GSEA_ERR_HIP, "no wall clock rate");
        ms_per_tick = 1.0 / (double)khz;
        ts.resize((size_t)clk.next);
        HIP_TRY(hipMemcpy(ts.data(), clk.stamps, sizeof(unsigned long long) * ts.size(),
                          hipMemcpyDeviceToHost));
    }
    auto span = [&](int a, int b, float *ms) -> int {
        if (clk.stamps) {
            *ms = (float)((double)(long long)(ts[(size_t)b] - ts[(size_t)a]) * ms_per_tick);
            return GSEA_OK;
        }
        HIP_TRY(hipEventElapsedTime(ms, clk.d->events[a], clk.d->events[b]));
        return GSEA_OK;
    };
    // stages: 0 lists, 1 metric, 2 rank, 3 statistics, 4 transpose, 5 ES
    double *slots[6] = {&t->ms_lists, &t->ms_metric, &t->ms_rank,
                        &t->ms_stats, &t->ms_transpose, &t->ms_es};
    int rc;
    for (auto &m : clk.marks) {
        float ms = 0.f;
        if ((rc = span(m.second.first, m.second.second, &ms))) return rc;
        *slots[m.first] += ms;
        if (m.first == 5) ++t->es_launches;
    }
    t->ms_rank_es = t->ms_rank + t->ms_es;
    float tot = 0.f;
    if ((rc = span(clk.marks.front().second.first, clk.marks.back().second.second, &tot)))
        return rc;
    t->ms_total = tot;
    return GSEA_OK;
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------
static size_t align16(size_t x) { return (x + 15) / 16 * 16; }

struct Labels {
    std::vector<uint8_t> lab;
    int nA = 0, nB = 0;          // divisors: label-0 / label-1 counts
    int lenA = 0;                // samples with label 0 after any swap
    int offB = 0, lstride = 0;
};

// class sizes as gsea_proxy counts them (cudaGSEA_wrapper.cu:239-251)
static int prepare_labels(const uint8_t *labels, int64_t S, int flags,
                          Labels *L) {
    L->lab.assign(labels, labels + S);
    int nA = 0, nB = 0;
    for (auto v : L->lab) { nA += v == 0; nB += v == 1; }
    if (flags & GSEA_FLAG_SWAP_LABELS) {
        for (auto &v : L->lab) v = (uint8_t)((v + 1) & 1);
        std::swap(nA, nB);
    }
    if (nA == 0 || nB == 0)
        return fail(GSEA_ERR_INVALID_ARGUMENT,
                    "both phenotype classes need at least one sample");
    int n0 = 0;
    for (auto v : L->lab) n0 += v == 0;
    L->nA = nA;
    L->nB = nB;
    L->lenA = n0;
    L->offB = (n0 + 7) / 8 * 8;
    L->lstride = L->offB + ((int)S - n0 + 7) / 8 * 8;
    return GSEA_OK;
}

// X tile width: the tuned width if its [S][GT] tile fits LDS, else the
// widest of 64/32/16 genes whose tile fits 64 KiB (two workgroups per CU);
// otherwise X is read from global memory.
template <class T>
static int metric_tile(int64_t S, int kind, bool *xlds, int lds_max) {
    const int tuned = g_tune.metric_gene_tile_kind[kind] ? g_tune.metric_gene_tile_kind[kind]
                                                         : g_tune.metric_gene_tile;
    if ((tuned == 16 || tuned == 32 || tuned == 64) &&
        (size_t)S * tuned * sizeof(T) <= (size_t)lds_max) {
        *xlds = true;
        return tuned;
    }
    for (int gt : {64, 32, 16}) {
        if ((size_t)S * gt * sizeof(T) <= 65536) { *xlds = true; return gt; }
    }
    *xlds = false;
    return 64;
}

// k_metric over Pb permutations' class lists; rcp[i] = fl64(1/(i+1)), i < S
template <class T>
static int launch_metric_rcp(const T *dX, int64_t N, int64_t S, const Labels &L,
                             const uint16_t *lists, int64_t Pb, int metric, T *r,
                             const double *rcp, const DeviceState *d, hipStream_t st) {
    MetricArgs<T> a;
    a.X = dX; a.N = (int)N; a.S = (int)S; a.lists = lists;
    a.lstride = L.lstride; a.nA = L.nA; a.nB = L.nB; a.offB = L.offB;
    a.lenA = L.lenA;
    a.Pb = (int)Pb; a.metric = metric; a.r = r;
    a.rcp = rcp;
    bool xlds = false;
    const int kind = acc_kind(metric);
    const int gt = metric_tile<T>(S, kind, &xlds, d->lds_max);
    const int64_t gx = (N + gt - 1) / gt;
    // enough workgroups to fill the device several times over, each streaming
    // as many permutations as possible over its X tile
    const int64_t target = 8LL * d->cu_count;
    int64_t ppb = (Pb * gx + target - 1) / target;
    ppb = std::max<int64_t>(32, std::min<int64_t>(512, (ppb + 31) / 32 * 32));
    if (g_tune.metric_perms_per_block_kind[kind] > 0)
        ppb = g_tune.metric_perms_per_block_kind[kind];
    else if (g_tune.metric_perms_per_block > 0)
        ppb = g_tune.metric_perms_per_block;
    a.ppb = (int)std::min<int64_t>(ppb, std::max<int64_t>(Pb, 1));
    HIP_TRY(launch_metric_kernel<T>(a, kind, gt, xlds, st));
    return GSEA_OK;
}

template <class T>
static int launch_metric(const T *dX, int64_t N, int64_t S, const Labels &L,
                         const uint16_t *lists, int64_t Pb, int metric,
                         T *r, DeviceState *d, hipStream_t st) {
    // 1/(i+1) for the Welford running counts, correctly rounded doubles
    if (d->rcp_len < S) {
        std::vector<double> rcp((size_t)S);
        for (int64_t i = 0; i < S; ++i) rcp[(size_t)i] = 1.0 / (double)(i + 1);
        void *pr;
        int rc = ensure(d, kBufRcp, sizeof(double) * S, &pr);
        if (rc) return rc;
        HIP_TRY(hipMemcpy(pr, rcp.data(), sizeof(double) * S, hipMemcpyHostToDevice));
        d->rcp_len = S;
    }
    return launch_metric_rcp<T>(dX, N, S, L, lists, Pb, metric, r,
                                (const double *)d->bufs[kBufRcp].ptr, d, st);
}

// LDS layouts.  k_rank: [keys (fp32 only)] | ids | 16 x 256 counters | misc.
// k_es: [|metric| by rank (fp32 when it fits)] | pos | per-wavefront bitmaps
// of the wavefronts that score sets | misc.
struct LdsPlan {
    bool alds = false;
    size_t lds = 0;
    int waves = 16;
    size_t off_ids = 0, off_cnt = 0, off_misc = 0;
};

struct RankPlan {
    bool big = false;  // k_rank_big: more genes than k_rank's LDS sort holds
    int npad = 0;      // k_rank_big: ids per wavefront segment x 16, multiple of 1024
    int maxch = 1;   // k_rank: chunks of 64 keys per wavefront
    int W = 1;       // k_es: bitmap words per lane
    int nw = 0;      // k_es: bitmap words per wavefront (64 W)
    LdsPlan rank, es;
};

static bool fit(size_t region0, size_t npad, size_t region2, int lds_max, LdsPlan *o) {
    size_t off = align16(region0);
    o->off_ids = off;
    off += align16(npad * 2);
    o->off_cnt = off;
    off += align16(region2);
    o->off_misc = off;
    off += 64 * 4;
    o->lds = off;
    return off <= (size_t)lds_max;
}

template <class T>
static int plan_rank(int64_t N, const DeviceState *d, RankPlan *pl) {
    pl->maxch = pick_maxch(N);
    pl->big = pl->maxch == 0;
    pl->npad = (int)((N + kRankThreads - 1) / kRankThreads * kRankThreads);
    if (pl->big) pl->maxch = 32;   // k_rank's LDS fields below are unused
    const size_t npad = (size_t)kRankThreads * pl->maxch;
    // k_es's bitmap holds ranks 0..N: rank N is the sentinel's (|metric| 0,
    // bit cleared before the walk), so N + 1 bits, W words per lane
    pl->W = (int)((N + 1 + 32 * kWave - 1) / (32 * kWave));
    pl->nw = kWave * pl->W;
    const size_t counters = (size_t)kRankWaves * 256 * 4;
    // k_rank keeps fp32 keys in LDS; fp64 keys are recomputed from r
    pl->rank.alds = sizeof(T) == 4 && fit(npad * 4, npad, counters, d->lds_max, &pl->rank);
    if (!pl->big && !pl->rank.alds && !fit(0, npad, counters, d->lds_max, &pl->rank))
        return fail(GSEA_ERR_UNSUPPORTED_SIZE, "gene count does not fit the rank kernel");
    // k_es: |metric| by rank and pos for genes/ranks 0..N (the sentinel
    // included), not padded to the rank kernel's chunks
    const size_t nes = (size_t)N + 1;
    const int ew_max = g_tune.es_waves > 0 ? std::min(g_tune.es_waves, kRankWaves) : kRankWaves;
    for (int alds = (sizeof(T) == 4) ? 1 : 0; alds >= 0; --alds) {
        for (int ew = ew_max; ew >= 1; --ew) {
            if (fit(alds ? nes * 4 : 0, nes, (size_t)ew * pl->nw * 4, d->lds_max, &pl->es)) {
                pl->es.alds = alds != 0;
                pl->es.waves = ew;
                return GSEA_OK;
            }
        }
    }
    return fail(GSEA_ERR_UNSUPPORTED_SIZE, "gene count does not fit the ES kernel");
}

// Work split of the ES phase: sets above the pair class are scored by one
// wavefront each (in set order), sets of at most kSmallSet genes by one lane
// each in chunks of 64 sets of one size class (<= 8, <= 16, <= 32, <= 64,
// <= 128 genes), and sets of 129..2*kSmallSet genes by a pair of lanes each
// in chunks of 32 sets (lane 2i holds entries 0..127, lane 2i+1 the rest).
struct SetTiers {
    int n_large = 0, n_chunks = 0, chunk_end8 = 0, chunk_end16 = 0, chunk_end32 = 0;
    int chunk_end64 = 0, chunk_end128 = 0;
    int stride = 0;
    const int4 *large_desc = nullptr;
    const int32_t *small_ids = nullptr, *small_k = nullptr, *small_kmax = nullptr;
    const uint16_t *small_gm = nullptr;
    const uint16_t *large_genes = nullptr;   // large sets' gene ids + 1
    const void *decay = nullptr;         // [G] of T
    const void *large_decay = nullptr;   // [n_large] of T
};

template <class T>
static int build_tiers(const int32_t *d_off, const int32_t *d_genes, int64_t G,
                       int64_t N, DeviceState *d, hipStream_t st, SetTiers *t,
                       int slot = 0, const std::vector<int32_t> *subset = nullptr) {
    if (slot < 0 || slot >= kTierSlots)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "too many set lists");
    std::vector<int32_t> off((size_t)G + 1);
    HIP_TRY(hipMemcpyAsync(off.data(), d_off, sizeof(int32_t) * (G + 1),
                           hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<int32_t> all;
    if (!subset) {
        all.resize((size_t)G);
        for (int64_t s = 0; s < G; ++s) all[(size_t)s] = (int32_t)s;
        subset = &all;
    }
    // sets up to lane_max genes: one lane each (tuned: 8..128, or none);
    // up to pair_max: a pair of lanes each (with lane_max 128 only)
    const int lane_max = g_tune.es_lane_set_max < 0 ? 0
                       : g_tune.es_lane_set_max > 0 ? std::min(g_tune.es_lane_set_max, kSmallSet)
                       : kSmallSet;
    const int pair_max = lane_max != kSmallSet || g_tune.es_pair_set_max < 0 ? 0
                       : g_tune.es_pair_set_max > 0 ? g_tune.es_pair_set_max
                       : 2 * kSmallSet;
    std::vector<int32_t> large, cls[6];
    for (int32_t s : *subset) {
        const int k = off[(size_t)s + 1] - off[(size_t)s];
        if (k < 0) return fail(GSEA_ERR_INVALID_ARGUMENT, "set_offsets must be non-decreasing");
        if (k > lane_max && k <= pair_max) cls[5].push_back((int32_t)s);
        else if (k > lane_max) large.push_back((int32_t)s);
        else cls[k <= 8 ? 0 : k <= 16 ? 1 : k <= 32 ? 2 : k <= 64 ? 3 : 4].push_back((int32_t)s);
    }
    // large sets stay in set order: a workgroup's wavefronts claim items in
    // order, so the ES stores of one permutation row fill cache lines
    // together.  k_es claims the chunks first, longest class first, and the
    // large sets last (each a few microseconds), so the wavefronts of a
    // workgroup finish a permutation close together
    // within a class, sets of similar size share a chunk of 64 lanes
    for (auto &c : cls)
        std::stable_sort(c.begin(), c.end(), [&](int32_t x, int32_t y) {
            return off[x + 1] - off[x] < off[y + 1] - off[y];
        });
    std::vector<int32_t> sid, sk;
    int ends[6];
    for (int c = 0; c < 6; ++c) {
        for (int32_t s : cls[c]) {
            // a pair set takes two adjacent slots, both with its id and size
            for (int h = 0; h < (c == 5 ? 2 : 1); ++h) {
                sid.push_back(s);
                sk.push_back(off[s + 1] - off[s]);
            }
        }
        while (sid.size() % kWave) { sid.push_back(-1); sk.push_back(0); }
        ends[c] = (int)(sid.size() / kWave);
    }
    t->n_large = (int)large.size();
    t->n_chunks = ends[5];
    t->chunk_end128 = ends[4];
    t->chunk_end8 = ends[0];
    t->chunk_end16 = ends[1];
    t->chunk_end32 = ends[2];
    t->chunk_end64 = ends[3];
    t->stride = (int)sid.size();
    const size_t nl = large.size(), ns = sid.size();
    // int32 words: large descriptors {set, offset, size, 0} | small ids |
    // small sizes.  offset: the set's first entry in this list's 16-bit copy
    // of the large sets' gene ids (made on every call: the caller's CSR may
    // change between calls at the same address)
    const size_t nch = ns / kWave;
    std::vector<int32_t> host(4 * nl + 2 * ns + nch);
    int64_t nlg = 0;
    for (size_t i = 0; i < nl; ++i) {
        const int32_t s = large[i];
        host[4 * i] = s;
        host[4 * i + 1] = (int32_t)nlg;
        host[4 * i + 2] = off[s + 1] - off[s];
        host[4 * i + 3] = 0;
        nlg += off[s + 1] - off[s];
    }
    std::copy(sid.begin(), sid.end(), host.begin() + 4 * nl);
    std::copy(sk.begin(), sk.end(), host.begin() + 4 * nl + ns);
    // the largest set of each chunk bounds its lanes' gather and walk
    for (size_t c = 0; c < nch; ++c)
        host[4 * nl + 2 * ns + c] =
            *std::max_element(sk.begin() + c * kWave, sk.begin() + (c + 1) * kWave);
    // per-set decay of the running sum, enrichment_scores.cuh:311, by set id
    // and again in large-item order
    std::vector<T> decay((size_t)G + nl);
    for (int64_t s = 0; s < G; ++s)
        decay[(size_t)s] = (T)(-1.0 / (double)(N - (off[(size_t)s + 1] - off[(size_t)s])));
    for (size_t i = 0; i < nl; ++i) decay[(size_t)G + i] = decay[(size_t)large[i]];
    void *pt, *pg, *pd, *pl;
    Buffer *tb = d->tier_bufs[slot];
    if (slot == 0) d->tier_valid = false;   // slot 0's buffers change below
    int rc = ensure_buf(tb[kTierIds], sizeof(int32_t) * std::max<size_t>(host.size(), 1), &pt);
    // 1 KiB of slack after the last set: a phase-A load past a set's end is
    // range-checked against that set, never outside the allocation
    if (!rc) rc = ensure_buf(tb[kTierGenes], sizeof(uint16_t) * (size_t)nlg + 1024, &pl);
    if (!rc) rc = ensure_buf(tb[kTierDecay], sizeof(T) * std::max<size_t>(decay.size(), 1), &pd);
    if (!rc) rc = ensure_buf(tb[kTierGm], sizeof(uint16_t) * std::max<size_t>(ns * kSmallSet, 1), &pg);
    if (rc) return rc;
    if (!host.empty())
        HIP_TRY(hipMemcpyAsync(pt, host.data(), sizeof(int32_t) * host.size(),
                               hipMemcpyHostToDevice, st));
    if (!decay.empty())
        HIP_TRY(hipMemcpyAsync(pd, decay.data(), sizeof(T) * decay.size(),
                               hipMemcpyHostToDevice, st));
    t->decay = pd;
    t->large_decay = (const T *)pd + G;
    t->large_desc = (const int4 *)pt;
    t->small_ids = (const int32_t *)pt + 4 * nl;
    t->small_k = (const int32_t *)pt + 4 * nl + ns;
    t->small_kmax = (const int32_t *)pt + 4 * nl + 2 * ns;
    t->small_gm = (const uint16_t *)pg;
    t->large_genes = (const uint16_t *)pl;
    if (nl) {
        hipLaunchKernelGGL(k_large_genes, dim3((unsigned)nl), dim3(256), 0, st, d_off, d_genes,
                           t->large_desc, (uint16_t *)pl);
        HIP_TRY(hipGetLastError());
    }
    if (ns) {
        hipLaunchKernelGGL(k_small_gm, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, st,
                           d_off, d_genes, t->small_ids, (int)ns, ends[4] * kWave,
                           (uint16_t *)pg);
        HIP_TRY(hipGetLastError());
    }
    // the host vectors above must outlive the asynchronous copy
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

// slot-0 tiers of the last set list, per device
static SetTiers g_tier0[64];

// tiers of every set of a list, reused across calls while the list is the
// same (pointers, sizes, lane split, value type and contents by checksum)
template <class T>
static int cached_tiers(const int32_t *d_off, const int32_t *d_genes, int64_t G, int64_t N,
                        DeviceState *d, hipStream_t st, SetTiers *out) {
    int dev = 0;
    int rc = current_device(&dev);
    if (rc) return rc;
    void *pc;
    rc = ensure(d, kBufCheck, 2 * sizeof(unsigned long long), &pc);
    if (rc) return rc;
    hipLaunchKernelGGL(k_csr_checksum, dim3(1), dim3(1024), 0, st, d_off, d_genes, G,
                       (unsigned long long *)pc);
    HIP_TRY(hipGetLastError());
    unsigned long long sum[2];
    HIP_TRY(hipMemcpyAsync(sum, pc, sizeof(sum), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    TierKey k;
    k.off = d_off; k.genes = d_genes; k.G = G; k.N = N;
    k.lane_max = g_tune.es_lane_set_max;
    k.pair_max = g_tune.es_pair_set_max;
    k.tsize = sizeof(T);
    k.sum = sum[0];
    k.total = sum[1];
    const TierKey &c = d->tier_key;
    if (d->tier_valid && c.off == k.off && c.genes == k.genes && c.G == k.G && c.N == k.N &&
        c.lane_max == k.lane_max && c.pair_max == k.pair_max && c.tsize == k.tsize &&
        c.sum == k.sum && c.total == k.total) {
        *out = g_tier0[dev];
        return GSEA_OK;
    }
    d->tier_valid = false;
    rc = build_tiers<T>(d_off, d_genes, G, N, d, st, out);
    if (rc) return rc;
    g_tier0[dev] = *out;
    d->tier_key = k;
    d->tier_valid = true;
    return GSEA_OK;
}

template <class T>
static void fill_rank_args(RankArgs<T> &a, const RankPlan &pl, int64_t N, int64_t Pb,
                           int64_t perm_base, const int32_t *d_off,
                           const int32_t *d_genes, int64_t G, int flags, T *es_pg,
                           const SetTiers *tiers) {
    a.r = nullptr; a.N = (int)N; a.Pb = (int)Pb; a.perm_base = perm_base;
    a.G = (int)G; a.set_off = d_off; a.set_genes = d_genes; a.es = es_pg;
    a.asc = (flags & GSEA_FLAG_ASCENDING) ? 1 : 0;
    a.nw = pl.nw;
    a.W = pl.W;
    a.a_glob = nullptr;
    a.pos_out = nullptr;
    a.p_lo = 0;
    a.idx_out = nullptr;
    a.n_large = tiers ? tiers->n_large : 0;
    a.n_items = tiers ? tiers->n_large + tiers->n_chunks : 0;
    a.large_desc = tiers ? tiers->large_desc : nullptr;
    a.large_decay = tiers ? (const T *)tiers->large_decay : nullptr;
    a.large_genes = tiers ? tiers->large_genes : nullptr;
    a.small_ids = tiers ? tiers->small_ids : nullptr;
    a.small_k = tiers ? tiers->small_k : nullptr;
    a.small_gm = tiers ? tiers->small_gm : nullptr;
    a.small_kmax = tiers ? tiers->small_kmax : nullptr;
    a.small_stride = tiers ? tiers->stride : 0;
    a.chunk_end8 = tiers ? tiers->chunk_end8 : 0;
    a.chunk_end16 = tiers ? tiers->chunk_end16 : 0;
    a.chunk_end32 = tiers ? tiers->chunk_end32 : 0;
    a.chunk_end64 = tiers ? tiers->chunk_end64 : 0;
    a.chunk_end128 = tiers ? tiers->chunk_end128 : 0;
    a.decay = tiers ? (const T *)tiers->decay : nullptr;
}

// k_es over rank rows [row_lo, row_hi) of the current batch (permutations
// perm_base + row) for the sets of `tiers`; the rows were written by k_rank
// into the rank-row buffers, which hold rows_cap rows
template <class T>
static int launch_es_rows(int64_t N, int64_t rows_cap, int64_t perm_base,
                          const int32_t *d_off, const int32_t *d_genes, int64_t G,
                          int flags, T *es_pg, DeviceState *d, hipStream_t st,
                          const SetTiers *tiers, int64_t row_lo, int64_t row_hi) {
    if (row_hi <= row_lo || !tiers || tiers->n_large + tiers->n_chunks == 0)
        return GSEA_OK;
    RankPlan pl;
    int rc = plan_rank<T>(N, d, &pl);
    if (rc) return rc;
    RankArgs<T> a;
    fill_rank_args<T>(a, pl, N, row_hi - row_lo, perm_base, d_off, d_genes, G, flags,
                      es_pg, tiers);
    void *pa = nullptr, *pp = nullptr;
    rc = ensure(d, kBufAGlob, (size_t)rows_cap * N * sizeof(T), &pa);
    if (!rc) rc = ensure(d, kBufPos, (size_t)rows_cap * N * sizeof(uint16_t), &pp);
    if (rc) return rc;
    a.a_glob = (T *)pa;
    a.pos_out = (uint16_t *)pp;
    a.p_lo = (int)row_lo;
    a.es_waves = pl.es.waves;
    a.off_ids = pl.es.off_ids; a.off_cnt = pl.es.off_cnt; a.off_misc = pl.es.off_misc;
    const int grid = (int)(row_hi - row_lo);
    hipError_t e = launch_es_kernel<T>(a, pl.es.alds, pl.es.lds, grid, st);
    HIP_TRY(e);
    return GSEA_OK;
}

// workgroups of k_rank_big: two per CU, each sorting its rows in turn
static int64_t rank_big_grid(const DeviceState *d) { return 2LL * d->cu_count; }

// k_rank over the Pb rows of a batch (rank rows kept for k_es), then k_es of
// every row for the sets of `tiers` (none: ranks only)
template <class T>
static int launch_rank(const T *r, int64_t N, int64_t Pb, int64_t perm_base,
                       const int32_t *d_off, const int32_t *d_genes, int64_t G,
                       int flags, T *es_pg, DeviceState *d, hipStream_t st,
                       const SetTiers *tiers, uint32_t *idx_out = nullptr) {
    RankPlan pl;
    int rc = plan_rank<T>(N, d, &pl);
    if (rc) return rc;
    RankArgs<T> a;
    fill_rank_args<T>(a, pl, N, Pb, perm_base, d_off, d_genes, G, flags, es_pg, tiers);
    a.r = r;
    a.idx_out = idx_out;
    if (G > 0) {
        // rank rows handed from k_rank to k_es: |metric| by rank, rank by gene
        void *pa = nullptr, *pp = nullptr;
        rc = ensure(d, kBufAGlob, (size_t)Pb * N * sizeof(T), &pa);
        if (!rc) rc = ensure(d, kBufPos, (size_t)Pb * N * sizeof(uint16_t), &pp);
        if (rc) return rc;
        a.a_glob = (T *)pa;
        a.pos_out = (uint16_t *)pp;
    }
    a.es_waves = kRankWaves;
    if (pl.big) {
        // ids ping-pong through 2 x npad entries per workgroup
        const int grid = (int)std::min<int64_t>(Pb, rank_big_grid(d));
        void *ps;
        rc = ensure(d, kBufRankScratch, (size_t)grid * 2 * pl.npad * sizeof(uint16_t), &ps);
        if (rc) return rc;
        HIP_TRY(launch_rank_big_kernel<T>(a, (uint16_t *)ps, pl.npad, grid, st));
    } else {
        const int grid = (int)Pb;
        a.off_ids = pl.rank.off_ids; a.off_cnt = pl.rank.off_cnt; a.off_misc = pl.rank.off_misc;
        HIP_TRY(launch_rank_kernel<T>(a, pl.rank.alds, pl.maxch, pl.rank.lds, grid, st));
    }
    if (G > 0)
        return launch_es_rows<T>(N, Pb, perm_base, d_off, d_genes, G, flags, es_pg, d, st,
                                 tiers, 0, Pb);
    return GSEA_OK;
}

// ---------------------------------------------------------------------------
// statistics
// ---------------------------------------------------------------------------
template <class T>
struct StatsBufs {
    T *factor = nullptr, *maxv = nullptr, *minv = nullptr, *thr = nullptr;
    unsigned long long *hist = nullptr;   // hub | hlb | signs[2] | hmax | hmin
    int32_t *table = nullptr;
    FdrInfo<T> *info = nullptr;
    int npow2 = 0;        // keys of the device threshold sort (0: host sort)
    size_t fdr_lds = 0;   // k_fdr_hist's LDS (0: histograms in global memory)
};

static int64_t hist_len(int64_t G) { return 4 * (G + 1) + 2; }

// every buffer the statistics of P x G scores use, and the kernels' LDS
// limits (before any capture: nothing here is a stream operation)
template <class T>
static int prepare_stats(int64_t P, int64_t G, DeviceState *d, StatsBufs<T> *b) {
    void *pf, *pmax, *pmin, *pthr, *phist, *ptab, *pinfo;
    int rc = ensure(d, kBufFactor, 2 * (size_t)G * sizeof(T), &pf);
    if (!rc) rc = ensure(d, kBufMax, (size_t)P * sizeof(T), &pmax);
    if (!rc) rc = ensure(d, kBufMin, (size_t)P * sizeof(T), &pmin);
    if (!rc) rc = ensure(d, kBufThr, (size_t)G * sizeof(T), &pthr);
    if (!rc) rc = ensure(d, kBufHist, sizeof(unsigned long long) * hist_len(G), &phist);
    if (!rc) rc = ensure(d, kBufFdrTable, sizeof(int32_t) * (kFdrBuckets + 1), &ptab);
    if (!rc) rc = ensure(d, kBufFdrInfo, sizeof(FdrInfo<T>), &pinfo);
    if (rc) return rc;
    b->factor = (T *)pf; b->maxv = (T *)pmax; b->minv = (T *)pmin; b->thr = (T *)pthr;
    b->hist = (unsigned long long *)phist; b->table = (int32_t *)ptab;
    b->info = (FdrInfo<T> *)pinfo;
    typedef typename KeyOf<T>::type K;
    int np = 2;
    while (np < G) np <<= 1;
    b->npow2 = (size_t)np * sizeof(K) <= (size_t)d->lds_max && np <= 32768 ? np : 0;
    if (b->npow2) {
        HIP_TRY(hipFuncSetAttribute((const void *)k_sort_thresholds<T>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)(b->npow2 * sizeof(K))));
    }
    const size_t lds = align16(sizeof(T) * G) + align16(sizeof(int32_t) * (kFdrBuckets + 1)) +
                       2 * sizeof(uint32_t) * (G + 1);
    b->fdr_lds = lds <= (size_t)d->lds_max ? lds : 0;
    if (b->fdr_lds) {
        HIP_TRY(hipFuncSetAttribute((const void *)k_fdr_hist<T, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    return GSEA_OK;
}

// the sorted thresholds and bucket table computed on the host (more sets
// than the device sort holds): the same values k_sort_thresholds writes
template <class T>
static int host_thresholds(const T *d_stats, int64_t G, const StatsBufs<T> &b,
                           hipStream_t st) {
    std::vector<T> nes((size_t)G);
    HIP_TRY(hipMemcpyAsync(nes.data(), d_stats + G, sizeof(T) * G, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<T> thr;
    thr.reserve((size_t)G);
    for (T v : nes)
        if (v == v) thr.push_back(v);
    std::sort(thr.begin(), thr.end());
    FdrInfo<T> f{};
    f.n = (int)thr.size();
    std::vector<int32_t> table;
    if (f.n > 1 && thr.back() > thr.front()) {
        f.nb = std::min(kFdrBuckets, f.n);
        const double w = ((double)thr.back() - (double)thr.front()) / f.nb;
        f.lo = thr.front();
        f.inv_w = (T)(1.0 / w);
        table.resize((size_t)f.nb + 1);
        for (int bk = 0; bk <= f.nb; ++bk) {
            const double edge = (double)thr.front() + bk * w;
            table[(size_t)bk] = (int32_t)(std::lower_bound(thr.begin(), thr.end(), edge,
                                                           [](T x, double e) {
                                                               return (double)x < e;
                                                           }) - thr.begin());
        }
        table[0] = 0;
        table[(size_t)f.nb] = f.n;
    } else if (f.n) {
        f.lo = thr.front();
    }
    if (f.n)
        HIP_TRY(hipMemcpyAsync(b.thr, thr.data(), sizeof(T) * f.n, hipMemcpyHostToDevice, st));
    if (f.nb)
        HIP_TRY(hipMemcpyAsync(b.table, table.data(), sizeof(int32_t) * (f.nb + 1),
                               hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b.info, &f, sizeof(f), hipMemcpyHostToDevice, st));
    // the host vectors above feed asynchronous copies
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

// ES, NES, NP, FWER and FDR of es[P][G] into d_stats [5][G]
// (statistics.cuh:46-117; FDR defined by this project), queued on `st`.
// With the device threshold sort (b.npow2 > 0) nothing here waits for the
// device or reads device memory on the host, so the sequence can be captured
// into a graph.  budget (optional): set s uses permutations 0..budget[s]-1.
// flags: GSEA_FLAG_SIGNED_NES selects the per-sign normalisation.
template <class T>
static int enqueue_stats(const T *es_pg, int64_t P, int64_t G, T *d_stats,
                         const StatsBufs<T> &b, const DeviceState *d, hipStream_t st,
                         int flags, const int32_t *d_budget = nullptr,
                         float *d_cv = nullptr) {
    const int g = (int)G;
    // one wavefront per 64 sets: the sequential per-set sums are latency
    // bound, so spread them over as many CUs as possible
    if (flags & GSEA_FLAG_SIGNED_NES)
        hipLaunchKernelGGL((k_set_stats<T, true>), dim3((g + 63) / 64), dim3(64), 0, st,
                           es_pg, P, g, d_stats, b.factor, d_budget, d_cv);
    else
        hipLaunchKernelGGL((k_set_stats<T, false>), dim3((g + 63) / 64), dim3(64), 0, st,
                           es_pg, P, g, d_stats, b.factor, d_budget, d_cv);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_perm_extrema<T>, dim3((unsigned)((P + 3) / 4)), dim3(256), 0,
                       st, es_pg, P, g, b.factor, d_budget, b.maxv, b.minv);
    HIP_TRY(hipGetLastError());
    // sorted observed NES (FDR thresholds, FWER through histograms)
    if (b.npow2) {
        typedef typename KeyOf<T>::type K;
        hipLaunchKernelGGL(k_sort_thresholds<T>, dim3(1), dim3(1024), b.npow2 * sizeof(K), st,
                           (const T *)d_stats, g, b.npow2, b.thr, b.info, b.table);
        HIP_TRY(hipGetLastError());
    } else {
        int rc = host_thresholds<T>(d_stats, G, b, st);
        if (rc) return rc;
    }
    unsigned long long *hub = b.hist, *hlb = hub + (G + 1), *signs = hlb + (G + 1);
    unsigned long long *hmax = signs + 2, *hmin = hmax + (G + 1);
    const int64_t hl = hist_len(G);
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)std::min<int64_t>((hl + 255) / 256, 1024)),
                       dim3(256), 0, st, b.hist, hl);
    HIP_TRY(hipGetLastError());
    if (d_budget) {
        // each set over its own permutations
        hipLaunchKernelGGL(k_fwer<T>, dim3((g + 255) / 256), dim3(256), 0, st, d_stats,
                           P, g, b.maxv, b.minv, d_budget);
        HIP_TRY(hipGetLastError());
    } else {
        const unsigned blocks = (unsigned)std::min<int64_t>((P + 255) / 256, 4LL * d->cu_count);
        hipLaunchKernelGGL(k_fwer_hist<T>, dim3(blocks), dim3(256), 0, st, (const T *)b.maxv,
                           (const T *)b.minv, P, (const T *)b.thr,
                           (const FdrInfo<T> *)b.info, hmax, hmin);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_prefix_u64, dim3(1), dim3(64), 0, st, hmax, hmin,
                           (const int *)&b.info->n);
        HIP_TRY(hipGetLastError());
        hipLaunchKernelGGL(k_fwer_final<T>, dim3((g + 255) / 256), dim3(256), 0, st, d_stats,
                           P, g, (const T *)b.thr, (const FdrInfo<T> *)b.info,
                           (const unsigned long long *)hmax, (const unsigned long long *)hmin);
        HIP_TRY(hipGetLastError());
    }
    if (P > 1) {
        if (b.fdr_lds) {
            // the thresholds, bucket table and two histograms fill most of a
            // CU's LDS, so one 16-wavefront workgroup per CU hides the ES
            // row loads' latency
            const int blocks = (int)std::min<int64_t>(P - 1, d->cu_count);
            hipLaunchKernelGGL((k_fdr_hist<T, true>), dim3(blocks), dim3(1024), b.fdr_lds, st,
                               es_pg, P, g, (const T *)b.factor, d_budget, (const T *)b.thr,
                               (const FdrInfo<T> *)b.info, (const int32_t *)b.table, hub, hlb,
                               signs);
        } else {
            const int blocks = (int)std::min<int64_t>(P - 1, 4LL * d->cu_count);
            hipLaunchKernelGGL((k_fdr_hist<T, false>), dim3(blocks), dim3(256), 0, st,
                               es_pg, P, g, (const T *)b.factor, d_budget, (const T *)b.thr,
                               (const FdrInfo<T> *)b.info, (const int32_t *)b.table, hub, hlb,
                               signs);
        }
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_prefix_u64, dim3(1), dim3(64), 0, st, hub, hlb,
                       (const int *)&b.info->n);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_fdr_final<T>, dim3((g + 255) / 256), dim3(256), 0, st, d_stats,
                       g, (const T *)b.thr, (const FdrInfo<T> *)b.info,
                       (const unsigned long long *)hub, (const unsigned long long *)hlb,
                       (const unsigned long long *)signs);
    HIP_TRY(hipGetLastError());
    return GSEA_OK;
}

template <class T>
static int run_stats(const T *es_pg, int64_t P, int64_t G, T *d_stats,
                     DeviceState *d, hipStream_t st, int flags,
                     const int32_t *d_budget = nullptr, float *d_cv = nullptr) {
    StatsBufs<T> b;
    int rc = prepare_stats<T>(P, G, d, &b);
    if (!rc) rc = enqueue_stats<T>(es_pg, P, G, d_stats, b, d, st, flags, d_budget, d_cv);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

// ---------------------------------------------------------------------------
// the whole path with device-resident inputs
// ---------------------------------------------------------------------------
static int check_sizes(int64_t N, int64_t S, int64_t G, int64_t P, int metric) {
    if (metric < 0 || metric > 13)
        return fail(GSEA_ERR_UNKNOWN_METRIC, "no known metric specified");
    if (N <= 0 || S <= 0 || P <= 0)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "empty expression data or no permutations");
    if (G <= 0) return fail(GSEA_ERR_NO_PATHS, "no gene sets specified");
    if (N > GSEA_MAX_GENES)
        return fail(GSEA_ERR_UNSUPPORTED_SIZE, "more genes than GSEA_MAX_GENES");
    if (S > GSEA_MAX_SAMPLES)
        return fail(GSEA_ERR_UNSUPPORTED_SIZE, "more samples than GSEA_MAX_SAMPLES");
    if (G > (int64_t)INT32_MAX / 2 || P > (int64_t)1 << 40)
        return fail(GSEA_ERR_UNSUPPORTED_SIZE, "too many sets or permutations");
    return GSEA_OK;
}

// permutations per batch: the metric matrix of one batch stays below 2 GiB
// (or the tuned count)
static int64_t batch_perms(int64_t N, int64_t P, size_t w) {
    int64_t cap = std::max<int64_t>(1, ((int64_t)2 << 30) / (N * (int64_t)w));
    if (g_tune.batch_perms > 0) cap = std::min(cap, g_tune.batch_perms);
    return std::min<int64_t>(P, cap);
}

// Everything a run needs before its first launch: buffers, the labels on
// the device, the reciprocal table, the set tiers and the statistics'
// buffers.  May wait for the device; enqueue_run then queues the whole path
// without waiting and without reading device memory on the host.
template <class T>
struct RunCtx {
    DeviceState *d = nullptr;
    const T *dX = nullptr;
    int64_t N = 0, S = 0, G = 0, P = 0, Pb = 0;
    int flags = 0;
    const int32_t *d_off = nullptr, *d_genes = nullptr;
    Labels L;
    uint8_t *plab = nullptr, *psig = nullptr;
    uint16_t *plists = nullptr;
    T *pr = nullptr, *pes = nullptr;
    SetTiers tiers;
    StatsBufs<T> sb;
};

template <class T>
static int prepare_run(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                       const int32_t *d_off, const int32_t *d_genes, int64_t G, int64_t P,
                       int flags, hipStream_t st, RunCtx<T> *c) {
    if (!dX || !labels || !d_off || !d_genes)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    int rc = device_state(&c->d);
    if (rc) return rc;
    DeviceState *d = c->d;
    rc = prepare_labels(labels, S, flags, &c->L);
    if (rc) return rc;
    c->dX = dX; c->N = N; c->S = S; c->G = G; c->P = P; c->flags = flags;
    c->d_off = d_off; c->d_genes = d_genes;
    c->Pb = batch_perms(N, P, sizeof(T));
    void *plab, *psig, *plists, *pr, *pes;
    rc = ensure(d, kBufLabels, (size_t)S, &plab);
    if (!rc) rc = ensure(d, kBufSig, (size_t)c->Pb * S, &psig);
    if (!rc) rc = ensure(d, kBufLists, (size_t)c->Pb * c->L.lstride * 2, &plists);
    if (!rc) rc = ensure(d, kBufR, (size_t)c->Pb * N * sizeof(T), &pr);
    if (!rc) rc = ensure(d, kBufES, (size_t)P * G * sizeof(T), &pes);
    if (rc) return rc;
    c->plab = (uint8_t *)plab; c->psig = (uint8_t *)psig; c->plists = (uint16_t *)plists;
    c->pr = (T *)pr; c->pes = (T *)pes;
    HIP_TRY(hipMemcpyAsync(plab, c->L.lab.data(), (size_t)S, hipMemcpyHostToDevice, st));
    // 1/(i+1) for the Welford running counts, correctly rounded doubles
    if (d->rcp_len < S) {
        std::vector<double> rcp((size_t)S);
        for (int64_t i = 0; i < S; ++i) rcp[(size_t)i] = 1.0 / (double)(i + 1);
        void *prc;
        rc = ensure(d, kBufRcp, sizeof(double) * S, &prc);
        if (rc) return rc;
        HIP_TRY(hipMemcpy(prc, rcp.data(), sizeof(double) * S, hipMemcpyHostToDevice));
        d->rcp_len = S;
    }
    rc = cached_tiers<T>(d_off, d_genes, G, N, d, st, &c->tiers);
    if (!rc) rc = prepare_stats<T>(P, G, d, &c->sb);
    if (rc) return rc;
    // rank rows handed from k_rank to k_es, and k_rank_big's id scratch
    void *pa, *pp;
    rc = ensure(d, kBufAGlob, (size_t)c->Pb * N * sizeof(T), &pa);
    if (!rc) rc = ensure(d, kBufPos, (size_t)c->Pb * N * sizeof(uint16_t), &pp);
    if (rc) return rc;
    RankPlan rp;
    if ((rc = plan_rank<T>(N, d, &rp))) return rc;
    if (rp.big) {
        void *ps;
        const int64_t grid = std::min<int64_t>(c->Pb, rank_big_grid(d));
        rc = ensure(d, kBufRankScratch, (size_t)grid * 2 * rp.npad * sizeof(uint16_t), &ps);
        if (rc) return rc;
    }
    HIP_TRY(hipStreamSynchronize(st));   // the labels' host copy
    return GSEA_OK;
}

// class lists of permutations [base, base + pb)
template <class T>
static int enqueue_lists(const RunCtx<T> &c, int64_t base, int64_t pb, hipStream_t st) {
    hipLaunchKernelGGL(k_class_lists, dim3((unsigned)((pb + 63) / 64)), dim3(64), 0, st,
                       (const uint8_t *)c.plab, (int)c.S, base, (int)pb, c.psig, c.L.offB,
                       c.plists, c.L.lstride);
    HIP_TRY(hipGetLastError());
    return GSEA_OK;
}

// The whole path for one metric: per permutation batch class lists, metric,
// rank rows and enrichment scores (es[P][G]), then the statistics and, when
// asked, the [G][P] matrix.  lists_ready: the class lists of a single batch
// are already in place (a sweep over metrics computes them once).  clk
// (optional) brackets the stages with events.
template <class T>
static int enqueue_run(const RunCtx<T> &c, int metric, T *d_stats, T *d_es_out,
                       hipStream_t st, StageClock *clk, bool lists_ready, int64_t *batches) {
    int rc;
    int64_t nb = 0;
    const double *rcp = (const double *)c.d->bufs[kBufRcp].ptr;
    for (int64_t base = 0; base < c.P; base += c.Pb, ++nb) {
        const int64_t pb = std::min(c.Pb, c.P - base);
        if (!(lists_ready && c.Pb >= c.P)) {
            if (clk && (rc = clk->begin(0))) return rc;
            if ((rc = enqueue_lists<T>(c, base, pb, st))) return rc;
            if (clk && (rc = clk->end())) return rc;
        }
        if (clk && (rc = clk->begin(1))) return rc;
        rc = launch_metric_rcp<T>(c.dX, c.N, c.S, c.L, c.plists, pb, metric, c.pr, rcp, c.d, st);
        if (rc) return rc;
        if (clk && (rc = clk->end())) return rc;
        if (clk && (rc = clk->begin(2))) return rc;
        rc = launch_rank<T>((const T *)c.pr, c.N, pb, base, c.d_off, c.d_genes, c.G, c.flags,
                            c.pes, c.d, st, nullptr);
        if (rc) return rc;
        if (clk && (rc = clk->end())) return rc;
        if (clk && (rc = clk->begin(5))) return rc;
        rc = launch_es_rows<T>(c.N, pb, base, c.d_off, c.d_genes, c.G, c.flags, c.pes, c.d, st,
                               &c.tiers, 0, pb);
        if (rc) return rc;
        if (clk && (rc = clk->end())) return rc;
    }
    if (batches) *batches = nb;
    if (clk && (rc = clk->begin(3))) return rc;
    rc = enqueue_stats<T>((const T *)c.pes, c.P, c.G, d_stats, c.sb, c.d, st, c.flags);
    if (rc) return rc;
    if (clk && (rc = clk->end())) return rc;
    if (d_es_out) {
        if (clk && (rc = clk->begin(4))) return rc;
        hipLaunchKernelGGL(k_transpose<T>,
                           dim3((unsigned)((c.G + 31) / 32), (unsigned)((c.P + 31) / 32)),
                           dim3(32, 8), 0, st, (const T *)c.pes, c.P, (int)c.G, d_es_out);
        HIP_TRY(hipGetLastError());
        if (clk && (rc = clk->end())) return rc;
    }
    return GSEA_OK;
}

template <class T>
static int run_device(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                      const int32_t *d_off, const int32_t *d_genes, int64_t G,
                      int64_t P, int metric, int flags, T *d_es_out, T *d_stats,
                      hipStream_t st, gsea_timings *t) {
    int rc = check_sizes(N, S, G, P, metric);
    if (rc) return rc;
    if (!d_stats) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    RunCtx<T> c;
    rc = prepare_run<T>(dX, N, S, labels, d_off, d_genes, G, P, flags, st, &c);
    if (rc) return rc;
    StageClock clk{c.d, st};
    int64_t batches = 0;
    rc = enqueue_run<T>(c, metric, d_stats, d_es_out, st, &clk, false, &batches);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return finish_timings(clk, t, batches);
}

// ---------------------------------------------------------------------------
// sweep over metrics: one HIP graph per shape
// ---------------------------------------------------------------------------
struct SweepKey {
    const void *dX = nullptr;
    int64_t N = 0, S = 0, G = 0, P = 0;
    std::vector<uint8_t> labels;
    const int32_t *off = nullptr, *genes = nullptr;
    std::vector<int> metrics;
    int flags = 0;
    const void *stats = nullptr;
    hipStream_t st = nullptr;
    size_t tsize = 0;
    gsea_tuning tune = {};
    TierKey tiers;
    uint64_t generation = 0;

    bool operator==(const SweepKey &o) const {
        const TierKey &a = tiers, &b = o.tiers;
        return dX == o.dX && N == o.N && S == o.S && G == o.G && P == o.P &&
               labels == o.labels && off == o.off && genes == o.genes &&
               metrics == o.metrics && flags == o.flags && stats == o.stats && st == o.st &&
               tsize == o.tsize && std::memcmp(&tune, &o.tune, sizeof(tune)) == 0 &&
               a.sum == b.sum && a.total == b.total && a.lane_max == b.lane_max &&
               a.pair_max == b.pair_max &&
               generation == o.generation;
    }
};

struct SweepGraph {
    bool valid = false;
    SweepKey key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // stage events recorded by the graph's event nodes (stage, ev0, ev1)
    std::vector<std::pair<int, std::pair<int, int>>> marks;
    int nstamps = 0;
    int64_t batches = 0;
};

static SweepGraph g_sweep[64];

static void drop_graph(SweepGraph &sg) {
    if (sg.exec) (void)hipGraphExecDestroy(sg.exec);
    if (sg.graph) (void)hipGraphDestroy(sg.graph);
    sg.exec = nullptr;
    sg.graph = nullptr;
    sg.valid = false;
}

// The whole path for several metrics over the same inputs: class lists once
// (a single permutation batch), then per metric the metric rows, rank rows,
// enrichment scores and statistics into d_stats[m][5][G].  With use_graph
// the sequence is captured into one HIP graph on the first call and replayed
// on later calls with the same key; otherwise (or when the statistics need
// the host: more sets than the device sort holds) it is queued directly.
template <class T>
static int run_sweep(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                     const int32_t *d_off, const int32_t *d_genes, int64_t G, int64_t P,
                     const int *metrics, int nm, int flags, T *d_stats, hipStream_t st,
                     bool use_graph, gsea_timings *t) {
    if (!metrics || nm < 1 || !d_stats)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "no metrics or null pointer argument");
    for (int m = 0; m < nm; ++m) {
        int rc = check_sizes(N, S, G, P, metrics[m]);
        if (rc) return rc;
    }
    DeviceState *d = nullptr;
    int rc = device_state(&d);
    if (rc) return rc;
    int dev = 0;
    if ((rc = current_device(&dev))) return rc;
    if (!st) {
        // a blocking stream: ordered after the null stream's earlier work and
        // before its later work, and capturable
        if (!d->own_stream) HIP_TRY(hipStreamCreate(&d->own_stream));
        st = d->own_stream;
    }
    RunCtx<T> c;
    rc = prepare_run<T>(dX, N, S, labels, d_off, d_genes, G, P, flags, st, &c);
    if (rc) return rc;
    const bool one_batch = c.Pb >= P;
    const bool capturable = use_graph && c.sb.npow2 > 0;
    // stage events bracket every launch; in a graph they become event nodes
    StageClock clk{d, st};
    int64_t batches = 0;
    auto enqueue_all = [&]() -> int {
        int r;
        if (one_batch) {
            if ((r = clk.begin(0))) return r;
            if ((r = enqueue_lists<T>(c, 0, P, st))) return r;
            if ((r = clk.end())) return r;
        }
        for (int m = 0; m < nm; ++m) {
            int64_t nb = 0;
            r = enqueue_run<T>(c, metrics[m], d_stats + (size_t)m * 5 * G, nullptr, st, &clk,
                               one_batch, &nb);
            if (r) return r;
            batches += nb;
        }
        return GSEA_OK;
    };
    if (!capturable) {
        if ((rc = enqueue_all())) return rc;
        HIP_TRY(hipStreamSynchronize(st));
        return finish_timings(clk, t, batches);
    }
    SweepKey k;
    k.dX = dX; k.N = N; k.S = S; k.G = G; k.P = P;
    k.labels.assign(labels, labels + S);
    k.off = d_off; k.genes = d_genes;
    k.metrics.assign(metrics, metrics + nm);
    k.flags = flags; k.stats = d_stats; k.st = st; k.tsize = sizeof(T);
    k.tune = g_tune;
    k.tiers = d->tier_key;
    SweepGraph &sg = g_sweep[dev];
    // stage stamps of the graph: two per stage launch
    const int cap = 2 * (1 + nm * (int)(4 * ((P + c.Pb - 1) / c.Pb) + 2));
    void *pstamp;
    if ((rc = ensure(d, kBufStamps, sizeof(unsigned long long) * cap, &pstamp))) return rc;
    k.generation = g_ws_generation;
    if (!(sg.valid && sg.key == k)) {
        drop_graph(sg);
        clk.stamps = (unsigned long long *)pstamp;
        clk.stamp_cap = cap;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        rc = enqueue_all();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(st, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess)
            return fail(GSEA_ERR_HIP, std::string("hipStreamEndCapture: ") +
                                          hipGetErrorString(ec));
        sg.graph = g;
        HIP_TRY(hipGraphInstantiate(&sg.exec, g, nullptr, nullptr, 0));
        sg.key = k;
        sg.marks = clk.marks;
        sg.nstamps = clk.next;
        sg.batches = batches;
        sg.valid = g_ws_generation == k.generation;
    }
    HIP_TRY(hipGraphLaunch(sg.exec, st));
    HIP_TRY(hipStreamSynchronize(st));
    StageClock done{d, st};
    done.marks = sg.marks;
    done.stamps = (unsigned long long *)pstamp;
    done.stamp_cap = cap;
    done.next = sg.nstamps;
    return finish_timings(done, t, sg.batches);
}

// ---------------------------------------------------------------------------
// budgeted run: set s is scored for permutations 0..budget[s]-1 only
// ---------------------------------------------------------------------------

// Sets (budget < P) to score for all permutations.  se = cv*|NES|/sqrt(B) is
// a set's NES standard error.  Call rule (ext.fdr > 0): among the sets of
// the same NES sign whose NES lies within z*se of this set's, FDR values on
// both sides of ext.fdr (or equal to it) mean the set's call is not settled.
// Precision rule (ext.nes_tol > 0): z*se > nes_tol.
template <class T>
static std::vector<int32_t> extension_sets(const T *st5, const float *cv,
                                           const int32_t *B, int64_t G, int64_t P,
                                           const gsea_extension &ext) {
    const T *nes = st5 + G, *fdr = st5 + 4 * G;
    std::vector<int32_t> pos, neg;
    for (int64_t s = 0; s < G; ++s) {
        if (!(nes[s] == nes[s]) || !(fdr[s] == fdr[s])) continue;
        (nes[s] > T(0) ? pos : neg).push_back((int32_t)s);
    }
    auto by_nes = [&](int32_t a, int32_t b) { return nes[a] < nes[b]; };
    std::sort(pos.begin(), pos.end(), by_nes);
    std::sort(neg.begin(), neg.end(), by_nes);
    std::vector<int32_t> out;
    for (int64_t s = 0; s < G; ++s) {
        if (B[s] >= P) continue;
        const double n = (double)nes[s];
        const double se = (double)cv[s] * std::fabs(n) / std::sqrt((double)B[s]);
        bool grow = false;
        if (ext.nes_tol > 0.0) grow = !(ext.z * se <= ext.nes_tol);
        if (!grow && ext.fdr > 0.0 && n == n) {
            const std::vector<int32_t> &v = n > 0.0 ? pos : neg;
            const double lo = n - ext.z * se, hi = n + ext.z * se;
            auto a = std::lower_bound(v.begin(), v.end(), lo, [&](int32_t x, double y) {
                return (double)nes[x] < y;
            });
            bool below = false, above = false;
            for (auto it = a; it != v.end() && (double)nes[*it] <= hi; ++it) {
                const double f = (double)fdr[*it];
                below = below || f <= ext.fdr;
                above = above || f >= ext.fdr;
            }
            const double f0 = (double)fdr[s];
            below = below || f0 <= ext.fdr;
            above = above || f0 >= ext.fdr;
            grow = below && above;
        }
        if (grow) out.push_back((int32_t)s);
    }
    return out;
}

template <class T>
static int run_device_budget(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                             const int32_t *d_off, const int32_t *d_genes, int64_t G,
                             int64_t P, int metric, int flags, const int32_t *budget,
                             const gsea_extension *ext, int32_t *budget_out,
                             T *d_es_pg, T *d_stats, hipStream_t st, gsea_timings *t) {
    int rc = check_sizes(N, S, G, P, metric);
    if (rc) return rc;
    if (!dX || !labels || !d_off || !d_genes || !d_stats || !budget)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    if (G < 1 || P < 2) return fail(GSEA_ERR_INVALID_ARGUMENT, "budgeted run needs sets and P >= 2");
    std::vector<int32_t> B(budget, budget + G);
    for (int64_t s = 0; s < G; ++s)
        if (B[(size_t)s] < 2 || B[(size_t)s] > P || P % B[(size_t)s] != 0)
            return fail(GSEA_ERR_INVALID_ARGUMENT,
                        "each budget must divide num_perms and be at least 2");
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    Labels L;
    rc = prepare_labels(labels, S, flags, &L);
    if (rc) return rc;

    const int64_t Pb = batch_perms(N, P, sizeof(T));
    void *plab, *psig, *plists, *pr, *pes, *pbud, *pcv;
    rc = ensure(d, kBufLabels, (size_t)S, &plab);
    if (!rc) rc = ensure(d, kBufSig, (size_t)Pb * S, &psig);
    if (!rc) rc = ensure(d, kBufLists, (size_t)Pb * L.lstride * 2, &plists);
    if (!rc) rc = ensure(d, kBufR, (size_t)Pb * N * sizeof(T), &pr);
    if (!rc) rc = ensure(d, kBufES, (size_t)P * G * sizeof(T), &pes);
    if (!rc) rc = ensure(d, kBufBudget, sizeof(int32_t) * G, &pbud);
    if (!rc) rc = ensure(d, kBufCv, sizeof(float) * G, &pcv);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(plab, L.lab.data(), (size_t)S, hipMemcpyHostToDevice, st));
    StageClock clk{d, st};

    // one set list per distinct budget (largest first), tier slots 1..
    auto make_levels = [&](const std::vector<int32_t> &ids, const std::vector<int32_t> &bud,
                           int slot0, std::vector<int64_t> *lv,
                           std::vector<SetTiers> *tiers) -> int {
        std::vector<int64_t> vals;
        for (int32_t s : ids) vals.push_back(bud[(size_t)s]);
        std::sort(vals.begin(), vals.end(), std::greater<int64_t>());
        vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
        if (slot0 + (int)vals.size() > kTierSlots)
            return fail(GSEA_ERR_INVALID_ARGUMENT, "too many distinct budgets");
        lv->assign(vals.begin(), vals.end());
        tiers->assign(vals.size(), SetTiers());
        for (size_t l = 0; l < vals.size(); ++l) {
            std::vector<int32_t> sub;
            for (int32_t s : ids)
                if (bud[(size_t)s] == vals[l]) sub.push_back(s);
            int r = build_tiers<T>(d_off, d_genes, G, N, d, st, &(*tiers)[l],
                                   slot0 + (int)l, &sub);
            if (r) return r;
        }
        return GSEA_OK;
    };
    std::vector<int32_t> ids((size_t)G);
    for (int64_t s = 0; s < G; ++s) ids[(size_t)s] = (int32_t)s;
    std::vector<int64_t> lv;
    std::vector<SetTiers> tiers;
    if ((rc = make_levels(ids, B, 1, &lv, &tiers))) return rc;

    // metric and rank rows of batch [base, base + pb); k_es of the given
    // levels over rows [lo_l - base, hi_l - base) of it
    auto batch = [&](int64_t base, int64_t pb) -> int {
        int r;
        if ((r = clk.begin(0))) return r;
        hipLaunchKernelGGL(k_class_lists, dim3((unsigned)((pb + 63) / 64)), dim3(64), 0, st,
                           (const uint8_t *)plab, (int)S, base, (int)pb, (uint8_t *)psig,
                           L.offB, (uint16_t *)plists, L.lstride);
        HIP_TRY(hipGetLastError());
        if ((r = clk.end())) return r;
        if ((r = clk.begin(1))) return r;
        r = launch_metric<T>(dX, N, S, L, (const uint16_t *)plists, pb, metric, (T *)pr, d, st);
        if (r) return r;
        if ((r = clk.end())) return r;
        if ((r = clk.begin(2))) return r;
        r = launch_rank<T>((const T *)pr, N, pb, base, d_off, d_genes, G, flags, (T *)pes,
                           d, st, nullptr);
        if (r) return r;
        return clk.end();
    };
    auto score = [&](int64_t base, int64_t pb, const std::vector<int64_t> &lo,
                     const std::vector<int64_t> &hi, std::vector<SetTiers> &tv) -> int {
        int r;
        for (size_t l = 0; l < tv.size(); ++l) {
            const int64_t a = std::max(lo[l], base) - base;
            const int64_t b = std::min(hi[l], base + pb) - base;
            if (b <= a) continue;
            if ((r = clk.begin(5))) return r;
            r = launch_es_rows<T>(N, Pb, base, d_off, d_genes, G, flags, (T *)pes, d, st,
                                  &tv[l], a, b);
            if (r) return r;
            if ((r = clk.end())) return r;
        }
        return GSEA_OK;
    };

    int64_t batches = 0, resident = -1;
    const std::vector<int64_t> zeros(lv.size(), 0);
    for (int64_t base = 0; base < lv[0]; base += Pb, ++batches) {
        const int64_t pb = std::min(Pb, P - base);
        if ((rc = batch(base, pb))) return rc;
        if ((rc = score(base, pb, zeros, lv, tiers))) return rc;
        resident = base;
    }
    HIP_TRY(hipMemcpyAsync(pbud, B.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, st));
    if ((rc = clk.begin(3))) return rc;
    rc = run_stats<T>((const T *)pes, P, G, d_stats, d, st, flags, (const int32_t *)pbud,
                      (float *)pcv);
    if (rc) return rc;
    if ((rc = clk.end())) return rc;

    // extension: sets whose significance call could change within the
    // uncertainty of their NES, or whose NES is imprecise, get all P
    // permutations
    if (ext) {
        std::vector<T> st5((size_t)5 * G);
        std::vector<float> cv((size_t)G);
        HIP_TRY(hipMemcpyAsync(st5.data(), d_stats, sizeof(T) * 5 * G, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipMemcpyAsync(cv.data(), pcv, sizeof(float) * G, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const std::vector<int32_t> grow =
            extension_sets(st5.data(), cv.data(), B.data(), G, P, *ext);
        std::vector<int32_t> ext_ids, oldB = B;
        for (int32_t s : grow) { ext_ids.push_back(s); B[(size_t)s] = (int32_t)P; }
        if (!ext_ids.empty()) {
            // groups by previous budget: score permutations [old, P)
            std::vector<int64_t> elv;
            std::vector<SetTiers> etiers;
            if ((rc = make_levels(ext_ids, oldB, 1 + (int)lv.size(), &elv, &etiers))) return rc;
            const std::vector<int64_t> ehi(elv.size(), P);
            const int64_t first = elv.back();   // smallest previous budget
            for (int64_t base = (first / Pb) * Pb; base < P; base += Pb) {
                const int64_t pb = std::min(Pb, P - base);
                if (base != resident) {
                    if ((rc = batch(base, pb))) return rc;
                    resident = base;
                }
                if ((rc = score(base, pb, elv, ehi, etiers))) return rc;
            }
            HIP_TRY(hipMemcpyAsync(pbud, B.data(), sizeof(int32_t) * G, hipMemcpyHostToDevice, st));
            if ((rc = clk.begin(3))) return rc;
            rc = run_stats<T>((const T *)pes, P, G, d_stats, d, st, flags, (const int32_t *)pbud,
                              (float *)pcv);
            if (rc) return rc;
            if ((rc = clk.end())) return rc;
        }
    }
    if (d_es_pg)
        HIP_TRY(hipMemcpyAsync(d_es_pg, pes, sizeof(T) * P * G, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (budget_out) std::copy(B.begin(), B.end(), budget_out);
    return finish_timings(clk, t, batches);
}

// host CSR check: offsets monotone from 0, genes strictly increasing per set
static int check_sets(const int32_t *off, const int32_t *genes, int64_t G,
                      int64_t N) {
    if (off[0] != 0) return fail(GSEA_ERR_INVALID_ARGUMENT, "set_offsets[0] must be 0");
    for (int64_t s = 0; s < G; ++s) {
        if (off[s + 1] < off[s])
            return fail(GSEA_ERR_INVALID_ARGUMENT, "set_offsets must be non-decreasing");
        for (int64_t i = off[s]; i < off[s + 1]; ++i) {
            if (genes[i] < 0 || genes[i] >= N)
                return fail(GSEA_ERR_INVALID_ARGUMENT, "gene index out of range");
            if (i > off[s] && genes[i] <= genes[i - 1])
                return fail(GSEA_ERR_INVALID_ARGUMENT,
                            "gene indices of a set must be strictly increasing");
        }
    }
    return GSEA_OK;
}

template <class T>
static int run_host(const T *exprs, int64_t N, int64_t S, const uint8_t *labels,
                    const int32_t *off, const int32_t *genes, int64_t G, int64_t P,
                    int metric, int flags, T *es_out, T *stats_out,
                    gsea_timings *t) {
    int rc = check_sizes(N, S, G, P, metric);
    if (rc) return rc;
    if (!exprs || !labels || !off || !genes || !stats_out)
        return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    rc = check_sets(off, genes, G, N);
    if (rc) return rc;
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    const int64_t total = off[G];
    void *px, *poff, *pgen, *pst, *pes = nullptr;
    rc = ensure(d, kBufX, (size_t)N * S * sizeof(T), &px);
    if (!rc) rc = ensure(d, kBufOff, (size_t)(G + 1) * 4, &poff);
    if (!rc) rc = ensure(d, kBufGenes, (size_t)std::max<int64_t>(total, 1) * 4, &pgen);
    if (!rc) rc = ensure(d, kBufStats, (size_t)5 * G * sizeof(T), &pst);
    if (!rc && es_out) rc = ensure(d, kBufEsOut, (size_t)G * P * sizeof(T), &pes);
    if (rc) return rc;
    hipStream_t st = nullptr;
    HIP_TRY(hipMemcpy(px, exprs, (size_t)N * S * sizeof(T), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(poff, off, (size_t)(G + 1) * 4, hipMemcpyHostToDevice));
    if (total) HIP_TRY(hipMemcpy(pgen, genes, (size_t)total * 4, hipMemcpyHostToDevice));
    rc = run_device<T>((const T *)px, N, S, labels, (const int32_t *)poff,
                       (const int32_t *)pgen, G, P, metric, flags, (T *)pes,
                       (T *)pst, st, t);
    if (rc) return rc;
    HIP_TRY(hipMemcpy(stats_out, pst, (size_t)5 * G * sizeof(T), hipMemcpyDeviceToHost));
    if (es_out)
        HIP_TRY(hipMemcpy(es_out, pes, (size_t)G * P * sizeof(T), hipMemcpyDeviceToHost));
    return GSEA_OK;
}

template <class T>
static int metric_stage(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                        int64_t P, int64_t shift, int metric, int flags, T *d_r,
                        hipStream_t st) {
    int rc = check_sizes(N, S, 1, P, metric);
    if (rc) return rc;
    if (shift < 0) return fail(GSEA_ERR_INVALID_ARGUMENT, "negative permutation shift");
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    Labels L;
    rc = prepare_labels(labels, S, flags, &L);
    if (rc) return rc;
    void *plab, *psig, *plists;
    rc = ensure(d, kBufLabels, (size_t)S, &plab);
    if (!rc) rc = ensure(d, kBufSig, (size_t)P * S, &psig);
    if (!rc) rc = ensure(d, kBufLists, (size_t)P * L.lstride * 2, &plists);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(plab, L.lab.data(), (size_t)S, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_class_lists, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, st,
                       (const uint8_t *)plab, (int)S, shift, (int)P, (uint8_t *)psig,
                       L.offB, (uint16_t *)plists, L.lstride);
    HIP_TRY(hipGetLastError());
    rc = launch_metric<T>(dX, N, S, L, (const uint16_t *)plists, P, metric, d_r, d, st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

// ---------------------------------------------------------------------------
// allocation-free metric stage.  Its scratch -- the labels, the permuted
// labels and class lists of every permutation, and the Welford reciprocal
// table -- lives in `scratch`, a static device array that every GPU the
// library runs on holds until the process ends.  Host code reaches it only
// through the per-device address hipGetSymbolAddress returns (the array's
// name is a host-side stand-in, not a device address).  All calls on one GPU
// share it: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static void *g_scratch[64];   // per device: device address of `scratch`

static int scratch_ptr(int dev, void **out) {
    if (!g_scratch[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(GSEA_ERR_HIP, "no device address for the static scratch");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return GSEA_OK;
}

// labels (0/1 bytes) swapped in place
__global__ void k_swap_labels(uint8_t *lab, int S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) lab[i] = (uint8_t)((lab[i] + 1) & 1);
}

// rcp[i] = 1/(i+1), correctly rounded (IEEE double division)
__global__ void k_recip_table(double *rcp, int S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) rcp[i] = 1.0 / (double)(i + 1);
}

// class sizes exactly as prepare_labels counts them, without a host copy
static int label_counts(const uint8_t *labels, int64_t S, int flags, Labels *L) {
    int n0 = 0, n1 = 0, odd = 0;
    for (int64_t i = 0; i < S; ++i) {
        n0 += labels[i] == 0;
        n1 += labels[i] == 1;
        odd += labels[i] & 1;
    }
    const bool swap = (flags & GSEA_FLAG_SWAP_LABELS) != 0;
    L->nA = swap ? n1 : n0;
    L->nB = swap ? n0 : n1;
    if (L->nA == 0 || L->nB == 0)
        return fail(GSEA_ERR_INVALID_ARGUMENT,
                    "both phenotype classes need at least one sample");
    // label-0 samples after the swap (v -> (v + 1) & 1 maps odd v to 0)
    L->lenA = swap ? odd : n0;
    L->offB = (L->lenA + 7) / 8 * 8;
    L->lstride = L->offB + ((int)S - L->lenA + 7) / 8 * 8;
    return GSEA_OK;
}

// bytes of scratch a metric call of P permutations over S samples needs
static size_t noalloc_layout(int64_t S, int64_t P, int lstride, size_t off[4]) {
    auto up = [](size_t x) { return (x + 255) / 256 * 256; };
    off[0] = 0;                                        // labels [S]
    off[1] = off[0] + up((size_t)S);                   // rcp [S] doubles
    off[2] = off[1] + up(sizeof(double) * (size_t)S);  // permuted labels [P][S]
    off[3] = off[2] + up((size_t)P * (size_t)S);       // class lists [P][lstride]
    return off[3] + (size_t)P * (size_t)lstride * 2;
}

template <class T>
static int metric_stage_noalloc(const T *dX, int64_t N, int64_t S, const uint8_t *labels,
                                int64_t P, int64_t shift, int metric, int flags, T *d_r,
                                hipStream_t st) {
    int rc = check_sizes(N, S, 1, P, metric);
    if (rc) return rc;
    if (shift < 0) return fail(GSEA_ERR_INVALID_ARGUMENT, "negative permutation shift");
    if (!dX || !labels || !d_r) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    Labels L;
    rc = label_counts(labels, S, flags, &L);
    if (rc) return rc;
    size_t off[4];
    const size_t need = noalloc_layout(S, P, L.lstride, off);
    if (need > kNoAllocScratchBytes)
        return fail(GSEA_ERR_UNSUPPORTED_SIZE,
                    "allocation-free metric call needs " + std::to_string(need) +
                        " bytes of scratch, more than its " +
                        std::to_string(kNoAllocScratchBytes) + " (fewer permutations per call)");
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    int dev = 0;
    if ((rc = current_device(&dev))) return rc;
    void *base = nullptr;
    if ((rc = scratch_ptr(dev, &base))) return rc;
    unsigned char *sp = (unsigned char *)base;
    uint8_t *plab = sp + off[0];
    double *prcp = (double *)(sp + off[1]);
    uint8_t *psig = sp + off[2];
    uint16_t *plists = (uint16_t *)(sp + off[3]);
    // every scratch byte a kernel reads is written first in this call: the
    // labels and reciprocals here, permuted labels and class lists by
    // k_class_lists (k_metric reads list entries [0, nA) and
    // [offB, offB + S - nA) only)
    HIP_TRY(hipMemcpyAsync(plab, labels, (size_t)S, hipMemcpyHostToDevice, st));
    const unsigned sb = (unsigned)((S + 255) / 256);
    if (flags & GSEA_FLAG_SWAP_LABELS) {
        hipLaunchKernelGGL(k_swap_labels, dim3(sb), dim3(256), 0, st, plab, (int)S);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_recip_table, dim3(sb), dim3(256), 0, st, prcp, (int)S);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_class_lists, dim3((unsigned)((P + 63) / 64)), dim3(64), 0, st,
                       (const uint8_t *)plab, (int)S, shift, (int)P, psig, L.offB, plists,
                       L.lstride);
    HIP_TRY(hipGetLastError());
    rc = launch_metric_rcp<T>(dX, N, S, L, plists, P, metric, d_r, prcp, d, st);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

template <class T>
static int rank_es_stage(const T *d_r, int64_t N, int64_t P, const int32_t *d_off,
                         const int32_t *d_genes, int64_t G, int flags, T *d_es,
                         hipStream_t st) {
    int rc = check_sizes(N, 1, G, P, 0);
    if (rc) return rc;
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    SetTiers tiers;
    rc = build_tiers<T>(d_off, d_genes, G, N, d, st, &tiers);
    if (rc) return rc;
    rc = launch_rank<T>(d_r, N, P, 0, d_off, d_genes, G, flags, d_es, d, st, &tiers);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

template <class T>
static int argsort_stage(const T *d_keys, int64_t N, int64_t P, int flags,
                         uint32_t *d_index, hipStream_t st) {
    int rc = check_sizes(N, 1, 1, P, 0);
    if (rc) return rc;
    if (!d_keys || !d_index) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    DeviceState *d = nullptr;
    rc = device_state(&d);
    if (rc) return rc;
    rc = launch_rank<T>(d_keys, N, P, 0, nullptr, nullptr, 0, flags, nullptr, d, st,
                        nullptr, d_index);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return GSEA_OK;
}

}  // namespace gsea

// ===========================================================================
// extern "C"
// ===========================================================================
using namespace gsea;

extern "C" {

int gsea_metric_id(const char *name) {
    if (!name) return -1;
    for (int i = 0; i < 14; ++i)
        if (std::strcmp(name, kMetricNames[i]) == 0) return i;
    return -1;
}

const char *gsea_metric_name(int id) {
    return (id >= 0 && id < 14) ? kMetricNames[id] : "";
}

const char *gsea_last_error(void) { return g_last_error.c_str(); }

const char *gsea_version(void) { return "cudagsea_amd 0.2.0 (gfx950)"; }

int gsea_set_tuning(const gsea_tuning *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (t) {
        const int gt = t->metric_gene_tile, lm = t->es_lane_set_max;
        const int pm = t->es_pair_set_max;
        bool bad = t->batch_perms < 0 || t->metric_perms_per_block < 0 || t->es_waves < 0 ||
                   t->es_waves > kRankWaves || (gt && gt != 16 && gt != 32 && gt != 64) ||
                   (lm > 0 && lm != 8 && lm != 16 && lm != 32 && lm != 64 && lm != 128) ||
                   lm < -1 || pm < -1 || (pm > 0 && (pm <= kSmallSet || pm > 2 * kSmallSet));
        for (int k = 0; k < 6; ++k) {
            const int g = t->metric_gene_tile_kind[k];
            bad |= (g && g != 16 && g != 32 && g != 64) || t->metric_perms_per_block_kind[k] < 0;
        }
        if (bad) return fail(GSEA_ERR_INVALID_ARGUMENT, "tuning value out of range");
    }
    g_tune = t ? *t : gsea_tuning{};
    return GSEA_OK;
}

int gsea_get_tuning(gsea_tuning *t) {
    if (!t) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    std::lock_guard<std::mutex> lk(g_mu);
    *t = g_tune;
    return GSEA_OK;
}

int gsea_device_count(int *count) {
    if (!count) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    HIP_TRY(hipGetDeviceCount(count));
    return GSEA_OK;
}

int gsea_get_device(int *device) {
    if (!device) return fail(GSEA_ERR_INVALID_ARGUMENT, "null pointer argument");
    HIP_TRY(hipGetDevice(device));
    return GSEA_OK;
}

int gsea_set_device(int device) {
    HIP_TRY(hipSetDevice(device));
    return GSEA_OK;
}

int gsea_device_name(int device, char *name, size_t name_len, size_t *total_bytes) {
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (name && name_len) {
        // some runtimes leave the marketing name empty: fall back to the ISA
        std::snprintf(name, name_len, "%s", prop.name[0] ? prop.name : prop.gcnArchName);
    }
    if (total_bytes) *total_bytes = prop.totalGlobalMem;
    return GSEA_OK;
}

int gsea_run_f32(const float *exprs, int64_t N, int64_t S, const uint8_t *labels,
                 const int32_t *off, const int32_t *genes, int64_t G, int64_t P,
                 int metric, int flags, float *es_out, float *stats_out,
                 gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_host<float>(exprs, N, S, labels, off, genes, G, P, metric, flags,
                           es_out, stats_out, t);
}

int gsea_run_f64(const double *exprs, int64_t N, int64_t S, const uint8_t *labels,
                 const int32_t *off, const int32_t *genes, int64_t G, int64_t P,
                 int metric, int flags, double *es_out, double *stats_out,
                 gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_host<double>(exprs, N, S, labels, off, genes, G, P, metric, flags,
                            es_out, stats_out, t);
}

int gsea_run_device_f32(const float *dX, int64_t N, int64_t S, const uint8_t *labels,
                        const int32_t *d_off, const int32_t *d_genes, int64_t G,
                        int64_t P, int metric, int flags, float *d_es_out,
                        float *d_stats, void *stream, gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_device<float>(dX, N, S, labels, d_off, d_genes, G, P, metric, flags,
                             d_es_out, d_stats, (hipStream_t)stream, t);
}

int gsea_run_device_f64(const double *dX, int64_t N, int64_t S, const uint8_t *labels,
                        const int32_t *d_off, const int32_t *d_genes, int64_t G,
                        int64_t P, int metric, int flags, double *d_es_out,
                        double *d_stats, void *stream, gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_device<double>(dX, N, S, labels, d_off, d_genes, G, P, metric, flags,
                              d_es_out, d_stats, (hipStream_t)stream, t);
}

int gsea_sweep_device_f32(const float *dX, int64_t N, int64_t S, const uint8_t *labels,
                          const int32_t *d_off, const int32_t *d_genes, int64_t G, int64_t P,
                          const int *metric_ids, int num_metrics, int flags, int use_graph,
                          float *d_stats, void *stream, gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_sweep<float>(dX, N, S, labels, d_off, d_genes, G, P, metric_ids, num_metrics,
                            flags, d_stats, (hipStream_t)stream, use_graph != 0, t);
}

int gsea_sweep_device_f64(const double *dX, int64_t N, int64_t S, const uint8_t *labels,
                          const int32_t *d_off, const int32_t *d_genes, int64_t G, int64_t P,
                          const int *metric_ids, int num_metrics, int flags, int use_graph,
                          double *d_stats, void *stream, gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_sweep<double>(dX, N, S, labels, d_off, d_genes, G, P, metric_ids, num_metrics,
                             flags, d_stats, (hipStream_t)stream, use_graph != 0, t);
}

int gsea_run_budget_device_f32(const float *dX, int64_t N, int64_t S,
                               const uint8_t *labels, const int32_t *d_off,
                               const int32_t *d_genes, int64_t G, int64_t P, int metric,
                               int flags, const int32_t *budget,
                               const gsea_extension *ext, int32_t *budget_out,
                               float *d_es_pg, float *d_stats, void *stream,
                               gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_device_budget<float>(dX, N, S, labels, d_off, d_genes, G, P, metric, flags,
                                    budget, ext, budget_out, d_es_pg, d_stats,
                                    (hipStream_t)stream, t);
}

int gsea_run_budget_device_f64(const double *dX, int64_t N, int64_t S,
                               const uint8_t *labels, const int32_t *d_off,
                               const int32_t *d_genes, int64_t G, int64_t P, int metric,
                               int flags, const int32_t *budget,
                               const gsea_extension *ext, int32_t *budget_out,
                               double *d_es_pg, double *d_stats, void *stream,
                               gsea_timings *t) {
    std::lock_guard<std::mutex> lk(g_mu);
    return run_device_budget<double>(dX, N, S, labels, d_off, d_genes, G, P, metric, flags,
                                     budget, ext, budget_out, d_es_pg, d_stats,
                                     (hipStream_t)stream, t);
}

int gsea_metric_f32(const float *dX, int64_t N, int64_t S, const uint8_t *labels,
                    int64_t P, int64_t shift, int metric, int flags, float *d_r,
                    void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return metric_stage<float>(dX, N, S, labels, P, shift, metric, flags, d_r,
                               (hipStream_t)stream);
}

int gsea_metric_f64(const double *dX, int64_t N, int64_t S, const uint8_t *labels,
                    int64_t P, int64_t shift, int metric, int flags, double *d_r,
                    void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return metric_stage<double>(dX, N, S, labels, P, shift, metric, flags, d_r,
                                (hipStream_t)stream);
}

int gsea_metric_noalloc(int double_precision, const void *d_exprs, int64_t N, int64_t S,
                        const uint8_t *labels, int64_t P, int64_t shift, int metric,
                        int flags, void *d_r, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (double_precision)
        return metric_stage_noalloc<double>((const double *)d_exprs, N, S, labels, P, shift,
                                            metric, flags, (double *)d_r, (hipStream_t)stream);
    return metric_stage_noalloc<float>((const float *)d_exprs, N, S, labels, P, shift, metric,
                                       flags, (float *)d_r, (hipStream_t)stream);
}

int gsea_rank_es_f32(const float *d_r, int64_t N, int64_t P, const int32_t *d_off,
                     const int32_t *d_genes, int64_t G, int flags, float *d_es,
                     void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return rank_es_stage<float>(d_r, N, P, d_off, d_genes, G, flags, d_es,
                                (hipStream_t)stream);
}

int gsea_rank_es_f64(const double *d_r, int64_t N, int64_t P, const int32_t *d_off,
                     const int32_t *d_genes, int64_t G, int flags, double *d_es,
                     void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return rank_es_stage<double>(d_r, N, P, d_off, d_genes, G, flags, d_es,
                                 (hipStream_t)stream);
}

int gsea_argsort_f32(const float *d_keys, int64_t N, int64_t P, int flags,
                     uint32_t *d_index, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return argsort_stage<float>(d_keys, N, P, flags, d_index, (hipStream_t)stream);
}

int gsea_argsort_f64(const double *d_keys, int64_t N, int64_t P, int flags,
                     uint32_t *d_index, void *stream) {
    std::lock_guard<std::mutex> lk(g_mu);
    return argsort_stage<double>(d_keys, N, P, flags, d_index, (hipStream_t)stream);
}

}  // extern "C"
