// k_rank and k_es instantiations (fp32 and fp64)
#include "gsea_launch.hpp"

namespace gsea {

// launches with more than the default 64 KiB of dynamic LDS opt in first
template <class T, bool ALDS, int MAXCH>
static hipError_t launch_rank_mc(const RankArgs<T> &a, size_t lds, int grid,
                                 hipStream_t st) {
    auto kern = k_rank<T, ALDS, MAXCH>;
    hipError_t e = hipFuncSetAttribute((const void *)kern,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kRankThreads), lds, st, a);
    return hipGetLastError();
}

template <class T, bool ALDS>
static hipError_t launch_rank_alds(const RankArgs<T> &a, int maxch, size_t lds, int grid,
                                   hipStream_t st) {
    switch (maxch) {
    case 1: return launch_rank_mc<T, ALDS, 1>(a, lds, grid, st);
    case 2: return launch_rank_mc<T, ALDS, 2>(a, lds, grid, st);
    case 4: return launch_rank_mc<T, ALDS, 4>(a, lds, grid, st);
    case 8: return launch_rank_mc<T, ALDS, 8>(a, lds, grid, st);
    case 12: return launch_rank_mc<T, ALDS, 12>(a, lds, grid, st);
    case 16: return launch_rank_mc<T, ALDS, 16>(a, lds, grid, st);
    case 20: return launch_rank_mc<T, ALDS, 20>(a, lds, grid, st);
    case 24: return launch_rank_mc<T, ALDS, 24>(a, lds, grid, st);
    case 28: return launch_rank_mc<T, ALDS, 28>(a, lds, grid, st);
    case 32: return launch_rank_mc<T, ALDS, 32>(a, lds, grid, st);
    default: return hipErrorInvalidValue;
    }
}

template <class T>
hipError_t launch_rank_kernel(const RankArgs<T> &a, bool alds, int maxch, size_t lds,
                              int grid, hipStream_t st) {
    return alds ? launch_rank_alds<T, true>(a, maxch, lds, grid, st)
                : launch_rank_alds<T, false>(a, maxch, lds, grid, st);
}

template <class T>
hipError_t launch_rank_big_kernel(const RankArgs<T> &a, uint16_t *scratch, int npad, int grid,
                                  hipStream_t st) {
    hipLaunchKernelGGL(k_rank_big<T>, dim3(grid), dim3(kRankThreads), 0, st, a, scratch, npad);
    return hipGetLastError();
}

template hipError_t launch_rank_big_kernel<float>(const RankArgs<float> &, uint16_t *, int, int,
                                                  hipStream_t);
template hipError_t launch_rank_big_kernel<double>(const RankArgs<double> &, uint16_t *, int,
                                                   int, hipStream_t);
template hipError_t launch_rank_kernel<float>(const RankArgs<float> &, bool, int, size_t,
                                              int, hipStream_t);
template hipError_t launch_rank_kernel<double>(const RankArgs<double> &, bool, int, size_t,
                                               int, hipStream_t);

}  // namespace gsea
