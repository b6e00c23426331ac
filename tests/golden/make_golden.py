"""Write tests/golden/ref_vectors.npz from the reference's own permutation
generator.

Runs only where the reference sources exist: oracle/ref/Makefile compiles
oracle/ref/ref_rng.cpp with plain g++ against the reference header
cudaGSEA/src/include/rngpu/rngpu.hpp (as shipped, nothing forced in or
stubbed) into oracle/_ref/libgsea_ref.so; this script calls that library and
stores inputs and outputs (data only) so the tests can pin the oracle without
the reference present.

Vectors:
  perm_ids, perm_S<S>   label permutations of arange(S) (mod 251 above 255)
                        for the global permutation ids perm_ids (seed
                        perm + 42, fast KISS + Fisher-Yates, rngpu.hpp)
  kiss_seeds, kiss      the first 64 draws of the generator for each seed
  hash_in, hash_out     the seeding hash (rngpu.hpp:27-38)

The reference's metric functors, enrichment scores and statistics are not
built here (they need nvcc's keywords, the CUDA runtime or CUB); the tests
pin those oracle functions by closed forms, float64 brute force and
invariants instead.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle  # noqa: E402

PERMS = [1, 2, 3, 17, 255, 4096, 65535, 1000003]
SIZES = [2, 3, 10, 40, 200, 500]
SEEDS = [0, 1, 43, 44, 4138, 4294967295]


def main():
    oracle.build(ref=True)
    out = {"perm_ids": np.array(PERMS, dtype=np.uint64)}
    for S in SIZES:
        base = np.arange(S, dtype=np.uint8) if S <= 255 else (
            np.arange(S) % 251).astype(np.uint8)
        out["perm_S%d" % S] = np.stack(
            [oracle.permute_labels(base, p, use_ref=True) for p in PERMS])
    out["kiss_seeds"] = np.array(SEEDS, dtype=np.uint32)
    out["kiss"] = np.stack([oracle.kiss_stream(s, 64, use_ref=True) for s in SEEDS])
    hin = np.array([0, 1, 42, 43, 12345, 0x7fffffff, 0xffffffff], dtype=np.uint32)
    out["hash_in"] = hin
    out["hash_out"] = np.array([oracle.ref().ref_hash32(int(x)) for x in hin],
                               dtype=np.uint32)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
