#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference generator.

The values come from oracle/_ref/libmg_ref.so, the reference's own
matrix_generator.cpp compiled by oracle/Makefile (`make -C oracle ref`), so
the fixtures pin the reference's behaviour, not this library's.  float -> half
rounding uses numpy's float16 conversion (round to nearest even, as
__float2half) and the BCSR arrays follow formats/BCSRMatrix.cu:36-71
(stored iff a half is nonzero, ascending block-columns, row-major blocks).

Outputs
  golden.json       hdr/idx of the small case (1024^2, 10% of 16x16 blocks,
                    seed 123) in full, SHA-256 of every array of that case
                    and of the 16384^2 cases, and the 8x8 checkerboard of
                    matrix_generator_test.cpp.
  tiny_patterns.npz float32 values of every reference pattern on small,
                    ragged shapes (and int values for the int instantiation).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402

PATTERNS = ["random", "checkerboard", "diagonal", "blockdiagonal", "blockrandom",
            "pattern64by16", "pattern32by16"]
TINY_SHAPES = [(40, 48, 8, 0.7, 42), (37, 70, 16, 0.5, 7), (64, 64, 16, 0.9, 123)]

# (name, M, K, sparsity, pattern, generator blocksize, bh, bw)
LARGE_CASES = [
    ("c1_16x16_d1", 16384, 16384, 0.99, "blockrandom", 16, 16, 16),
    ("c1_16x16_d5", 16384, 16384, 0.95, "blockrandom", 16, 16, 16),
    ("c1_16x16_d20", 16384, 16384, 0.80, "blockrandom", 16, 16, 16),
    ("c1_32x16_d5", 16384, 16384, 0.95, "pattern32by16", 32, 32, 16),
    ("c1_64x16_d5", 16384, 16384, 0.95, "pattern64by16", 64, 64, 16),
    ("c1_blockdiag256", 16384, 16384, 0.0, "blockdiagonal", 256, 16, 16),
]
B_CASES = [("B_16384x64", 16384, 64), ("B_16384x256", 16384, 256), ("B_16384x512", 16384, 512),
           ("B_1024x256", 1024, 256)]


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def to_half_bits(a32):
    return np.ascontiguousarray(a32, dtype=np.float32).astype(np.float16).view(np.uint16)


def bcsr_case(M, K, sp, pattern, gbs, bh, bw, seed=123):
    A = oracle.ref_generate_f32(M, K, sp, pattern, gbs, seed)
    A16 = to_half_bits(A)
    del A
    hdr, idx, data = oracle.bcsr_from_dense(A16, bh, bw)
    return A16, hdr, idx, data


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libmg_ref.so missing: run `make -C oracle ref` where the reference "
                 "tree is present")
    out = {"source": "reference matrix_generator.cpp via oracle/_ref; fp16 by numpy RNE"}

    A16, hdr, idx, data = bcsr_case(1024, 1024, 0.9, "blockrandom", 16, 16, 16)
    out["blockrandom_1024x1024_bs16"] = {"M": 1024, "K": 1024, "sparsity": 0.9, "pattern": "blockrandom",
                      "blocksize": 16, "seed": 123, "nnzb": int(hdr[-1]),
                      "hdr": hdr.tolist(), "idx": idx.tolist(), "sha_data": sha(data),
                      "sha_dense_f16": sha(A16)}
    large = {}
    for name, M, K, sp, pat, gbs, bh, bw in LARGE_CASES:
        A16, hdr, idx, data = bcsr_case(M, K, sp, pat, gbs, bh, bw)
        large[name] = {"M": M, "K": K, "sparsity": sp, "pattern": pat, "blocksize": gbs,
                       "bh": bh, "bw": bw, "seed": 123, "nnzb": int(hdr[-1]),
                       "sha_hdr": sha(hdr), "sha_idx": sha(idx), "sha_data": sha(data)}
        print(name, int(hdr[-1]), flush=True)
        del A16, data
    out["large"] = large
    bs = {}
    for name, K, N in B_CASES:
        B = oracle.ref_generate_f32(K, N, 0.0, "random", 16, 456)
        bs[name] = {"K": K, "N": N, "seed": 456, "sha_f16": sha(to_half_bits(B))}
    out["B"] = bs
    chk = oracle.ref_generate_matrix(8, 8, 0.7, "checkerboard", 2, 42)
    out["generator_test_checkerboard"] = chk.astype(np.float32).view(np.uint32).tolist()
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1)

    arrays = {}
    for pat in PATTERNS:
        for (r, c, bs_, sp, seed) in TINY_SHAPES:
            key = f"{pat}_{r}x{c}_b{bs_}_s{sp}_seed{seed}"
            arrays[key + "_float"] = oracle.ref_generate_matrix(r, c, sp, pat, bs_, seed,
                                                                "float").astype(np.float32)
            arrays[key + "_int"] = oracle.ref_generate_matrix(r, c, sp, pat, bs_, seed,
                                                              "int").astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "tiny_patterns.npz"), **arrays)
    print("wrote golden.json and tiny_patterns.npz")


if __name__ == "__main__":
    main()
