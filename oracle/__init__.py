"""CPU oracle for the matrix-free REML path — TEST INFRASTRUCTURE ONLY.

This package is a NumPy/SciPy restatement of the reference algorithm
(GPU_REML: ``geno_stream``, ``kv_impl``, ``pcg``, ``precond``, ``slq``,
``reml``), written from the reference's source text; every function cites
the reference file:line it follows.  It exists to check the HIP product
path, and only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it.  The product package ``coherit_amd``
never imports it and has no CPU fallback.

Pinning (what checks each module against something other than itself):

* ``geno``   — PLINK BED spec closed form, hand-built golden vectors of every
  code in every in-byte position, the reference's own genotype fixtures
  (``tests/test_sparse_prediction.py:146-174``) and its stats formula.
* ``kv``     — brute-force dense ``Z Z^T / eff`` in FP64; symmetry/PSD; the
  identity ``sum_g eff_g K_g = Z Z^T`` for contiguous partitions.
* ``pcg``    — the eight behaviour tests of ``tests/test_pcg_stability.py``
  (closed-form diagonal/dense answers, operator-call counts, restart).
* ``precond``— closed forms: ``M^{-1}(M v) = v``, ``log|M|`` = dense slogdet.
* ``slq``    — the dense probe oracle ``sum z^T log(H) z / S`` at full depth,
  finite differences of its own value in FP64, batching invariance.
* ``reml``   — exact dense REML (explicit K, Cholesky) on tiny n, finite
  differences of the objective, ascent of the AI direction, SciPy NNLS KKT.

The reference itself cannot run here (it needs jax/jaxlib/numba, absent from
this image), so there is no ``oracle/_ref`` build; probes are generated by
this project's own counter-based generator and fed identically to the
oracle and to the GPU path.
"""
