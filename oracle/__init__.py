"""CPU oracle for the brute-force L2 kNN path -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py use
this package, as the checker / the reported CPU baseline. The product
(curplsh_amd, libcurplsh_amd.so, libcurplsh.so) never imports or links it.

Functions and the reference code they restate (paths relative to the
reference repository):

  norms_canonical   computeL2Norm, src/internal/norm.cu:191-202 (squared)
  search_canonical  searchL2Distance, src/internal/search.cu:79-216, with the
                    GEMM of src/internal/matrix_ops.cu:56-58 (alpha=-2), the
                    base-norm add + select of src/internal/select.cu:189-224,
                    the +||q||^2 of src/internal/broadcast.cu:15-38 and the
                    exact-distance choice of src/index_bf.cpp:68-69; empty
                    index / padding per search.cu:93-100 and index.h:42-44
  search_fp64       FP64 brute force (no reference counterpart: the judge of
                    recall and distance error for every mode)
  recall            Dataset::evaluate, include/dataset.h:135-152
  bf16 / e4m3       encoders from IEEE-754 and the OCP OFP8 spec
  bf16_tree_dot     the LSH pre-rank's BF16 dot product (piece/tree order)
  lsh_hash /        the library's random-projection LSH front end (IndexRPLSH);
  lsh_search        the reference has no LSH step (only the Gaussian
                    initialiser of src/internal/matrix_ops.cu:71-76), so this
                    restates the library's own specification

Pinning (tests/test_oracle.py): closed-form fixtures of test/test_search.cpp:17-36
and test/test_norm.cpp:15-22, exact rational arithmetic on tiny inputs, numpy
float64 brute force, torch's bfloat16 / float8_e4m3fn conversions, and the
exhaustive e4m3 code table. LSH (tests/test_oracle_lsh.py): float64 signs of
the projections, the all-buckets case against search_canonical, and the
candidate sets rebuilt by set membership and by an explicit bucket
concatenation in numpy.
"""
from .oracle import (  # noqa: F401
    LIB_PATH, bf16_rne, bf16_to_float, bf16_tree_dot, e4m3_rne_sat, e4m3_to_float, load, lsh_hash, lsh_search,
    norms_canonical, pair_dist_fp64, recall, search_canonical, search_fp64,
)
