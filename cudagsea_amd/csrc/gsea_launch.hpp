// gsea_launch.hpp -- launchers of the templated hot-path kernels.  Each
// family is instantiated in its own translation unit (launch_*.hip), so the
// library's kernels compile in parallel.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include "gsea_metric.hpp"
#include "gsea_rank.hpp"

namespace gsea {

// k_metric for accumulator `kind`, X tile width gt (64/32/16 genes) staged in
// LDS when xlds; grid (ceil(N/gt), ceil(Pb/ppb)) of 256 threads
template <class T>
hipError_t launch_metric_kernel(const MetricArgs<T> &a, int kind, int gt, bool xlds,
                                hipStream_t st);

// k_rank with maxch chunks of 64 keys per wavefront (1, 2, 4, 8, 12, ..., 32),
// fp32 keys in LDS when alds; one 1024-thread workgroup per row, grid rows
template <class T>
hipError_t launch_rank_kernel(const RankArgs<T> &a, bool alds, int maxch, size_t lds,
                              int grid, hipStream_t st);

// k_rank_big (gene counts beyond k_rank's LDS sort): scratch holds
// 2 * npad 16-bit ids per workgroup of the grid
template <class T>
hipError_t launch_rank_big_kernel(const RankArgs<T> &a, uint16_t *scratch, int npad, int grid,
                                  hipStream_t st);

// k_es, |metric| by rank in LDS when alds (fp32 only)
template <class T>
hipError_t launch_es_kernel(const RankArgs<T> &a, bool alds, size_t lds, int grid,
                            hipStream_t st);

template <>
hipError_t launch_metric_kernel<float>(const MetricArgs<float> &, int, int, bool, hipStream_t);
template <>
hipError_t launch_metric_kernel<double>(const MetricArgs<double> &, int, int, bool,
                                        hipStream_t);
template <>
hipError_t launch_es_kernel<float>(const RankArgs<float> &, bool, size_t, int, hipStream_t);
template <>
hipError_t launch_es_kernel<double>(const RankArgs<double> &, bool, size_t, int, hipStream_t);
extern template hipError_t launch_rank_big_kernel<float>(const RankArgs<float> &, uint16_t *,
                                                         int, int, hipStream_t);
extern template hipError_t launch_rank_big_kernel<double>(const RankArgs<double> &, uint16_t *,
                                                          int, int, hipStream_t);
extern template hipError_t launch_rank_kernel<float>(const RankArgs<float> &, bool, int,
                                                     size_t, int, hipStream_t);
extern template hipError_t launch_rank_kernel<double>(const RankArgs<double> &, bool, int,
                                                      size_t, int, hipStream_t);

// the chunk counts k_rank is instantiated for
inline int pick_maxch(long long n) {
    const int need = (int)((n + kRankThreads - 1) / kRankThreads);
    for (int c : {1, 2, 4, 8, 12, 16, 20, 24, 28, 32})
        if (c >= need) return c;
    return 0;   // more keys than one workgroup sorts
}

}  // namespace gsea
