// ref_rng.cpp -- exposes the reference's permutation generator through a C
// ABI so the tests can pin the oracle's RNG against it.  TEST INFRASTRUCTURE
// ONLY.
//
// It #includes cudaGSEA/src/include/rngpu/rngpu.hpp where it lies under
// /root/reference (oracle/ref/Makefile); nothing of the reference is copied.
// That header is plain C++ when __CUDA_ARCH__ is undefined (rngpu.hpp:4-9)
// and compiles as shipped with g++; it only expects the fixed-width integer
// types to be declared before it, which the standard headers below do.  The
// other reference headers on the path (functors.cuh, statistics.cuh,
// enrichment_scores.cuh, ...) need the CUDA compiler's __host__/__device__
// keywords or the CUDA runtime and are not built here.
#include <cstddef>
#include <cstdint>
#include <cstring>

#include "rngpu/rngpu.hpp"

extern "C" {

// labels of global permutation `perm`: the observed labels for perm == 0,
// otherwise their Fisher-Yates shuffle seeded with perm + 42, as
// correlate_genes.cuh:94-103 (CPU) and :44-55 (GPU) do
void ref_permute_labels(const uint8_t *labels, int64_t S, uint64_t perm,
                        uint8_t *out) {
    std::memcpy(out, labels, (size_t)S);
    if (perm == 0) return;
    auto state = get_initial_fast_kiss_state32((uint32_t)(perm + 42));
    fisher_yates_shuffle(fast_kiss32, &state, out, (size_t)S);
}

// the hash the generator is seeded with (rngpu.hpp:17,27-38)
uint32_t ref_hash32(uint32_t x) { return hash32(x); }

// the first n draws of the fast KISS generator seeded with `seed`
// (rngpu.hpp:145-178)
void ref_kiss_stream(uint32_t seed, uint32_t *out, int64_t n) {
    auto state = get_initial_fast_kiss_state32(seed);
    for (int64_t i = 0; i < n; ++i) out[i] = fast_kiss32(&state);
}

}  // extern "C"
