"""Write tests/golden/reference_factors.npz from the reference itself.

Runs the reference library compiled from its own sources
(oracle/_ref/libbkref.so, built by ``make -C oracle``) on the matrices its
own tests use -- the fixed 4x4 cases of
tests/block_ldlt_decomposition_test.cc:421-508, RandomSymmetricMatrix
(:148-161) seeds of TestRandomFactorizations (:510-522) and of the default
stress list (:40-99) -- and stores, per case, the upper triangle of the
compact factors, Permutation(), PivotTypes(), the block geometry of
CompactFactors() and the full reconstruction residual.  The inputs A_* are
drawn by the reference's own RandomSymmetricMatrix (compiled from its test
source by oracle/ref_tests_decomposition.cc).  Also stores, from the large
stress test's own code (oracle/ref_tests_stress.cc): the 4096
residual-sample positions of ResidualSamples
(tests/block_ldlt_large_stress_test.cc:204-218) for one (size, seed), the
RandomSymmetricBlockMatrix input of two small cases (:116-136), and one
whole stress case (its compact factors and SampledRelativeResidual,
:220-265).

Usage (in a container that has the reference tree):
    make -C oracle && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import oracle as O  # noqa: E402

FIXED = [
    [6.0, 12.0, 3.0, -6.0, 12.0, -8.0, -13.0, 4.0, 3.0, -13.0, -1.0, 1.0, -6.0, 4.0, 1.0, 6.0],
    [0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0],
    [2.0, 0.0, 0.0, 0.0, 0.0, -3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 5.0],
]
SCALES = [1.0e-12, 1.0e-9, 1.0e-6, 1.0e-3, 1.0, 1.0e3, 1.0e6, 1.0e9, 1.0e12]


def cases():
    out = []
    for fi, vals in enumerate(FIXED):
        for bs in range(1, 5):
            out.append(("fixed%d" % fi, np.array(vals).reshape(4, 4), bs, 0, 1.0))
    # TestRandomFactorizations seeds: 10000*n + 100*bs + seed
    for n in (5, 8, 13, 16):
        for bs in sorted({1, 2, 3, n}):
            for seed in range(2):
                sd = 10000 * n + 100 * bs + seed
                out.append(("random", O.ref_random_symmetric_matrix(n, sd), bs, sd, 1.0))
    # default stress seeds: 910000 + 10000*n + 100*bs + 10*scale_index
    for n, bs, si in ((20, 3, 0), (20, 5, 8), (50, 1, 4), (50, 8, 2), (50, 21, 6),
                      (50, 32, 4)):
        sd = 910000 + 10000 * n + 100 * bs + 10 * si
        out.append(("stress", O.ref_random_symmetric_matrix(n, sd, SCALES[si]), bs, sd,
                    SCALES[si]))
    return out


def main():
    O.ref()
    store = {}
    meta = []
    for idx, (kind, A, bs, seed, scale) in enumerate(cases()):
        n = A.shape[0]
        r = O.ref_dsytrf(A, bs, full=True)
        iu = np.triu_indices(n)
        store["A_%d" % idx] = A[iu]
        store["U_%d" % idx] = r.compact[iu]
        store["perm_%d" % idx] = r.perm
        store["piv_%d" % idx] = r.piv
        store["geom_%d" % idx] = r.geometry
        store["res_%d" % idx] = np.array([r.residual])
        meta.append((n, bs, seed, scale))
    store["meta"] = np.array(meta, dtype=np.float64)
    store["kinds"] = np.array([c[0] for c in cases()])
    rows, cols = O.ref_residual_samples(5000, 931300)
    store["samples_5000_931300"] = np.stack([rows, cols])
    # RandomSymmetricBlockMatrix inputs: (n, bs, seed, scale)
    blk = [(37, 8, 930100, 1.0), (64, 21, 930200, 1.0e6)]
    store["block_meta"] = np.array(blk, dtype=np.float64)
    for t, (n, bs, sd, sc) in enumerate(blk):
        A = O.ref_random_symmetric_block_matrix(n, bs, sd, sc)
        store["block_A_%d" % t] = A[np.triu_indices(n)]
    # one whole stress case: compact factors and the sampled residual
    n, bs, sd, sc = 200, 32, 931300, 1.0e-3
    f, sres = O.ref_stress_case(n, bs, sd, sc)
    store["stress_meta"] = np.array([n, bs, sd, sc], dtype=np.float64)
    store["stress_U"] = f.compact[np.triu_indices(n)]
    store["stress_perm"] = f.perm
    store["stress_piv"] = f.piv
    store["stress_sampled_residual"] = np.array([sres])
    path = os.path.join(HERE, "reference_factors.npz")
    np.savez_compressed(path, **store)
    print("wrote", path, os.path.getsize(path), "bytes,", len(meta), "cases")


if __name__ == "__main__":
    main()
