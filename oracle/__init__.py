"""CPU oracle of the blocks_3d hot path -- TEST INFRASTRUCTURE ONLY.

Nothing in the product (`blocksvibe_amd`) imports this package.  Only `tests/`, `__graft_entry__.smoke()`
and the `cpu_baseline` leg of `bench.py` use it, and only as the checker.

Modules:
  mpf_ref        bit-level specification of the multiprecision number format and of every rounding the
                 HIP kernels perform, restated with Python integers (pinned by exact-rational error bounds,
                 brute force over small inputs and special cases: tests/test_oracle_mpf.py)
  structure_ref  pole/spin bookkeeping (reference blocksvibe/structure.py:16-83), scalar form (pinned against
                 vectors written from the reference's own structure.py, tests/golden/)
  recursion_ref  the radial recursion in exact arithmetic (reference blocksvibe/recursion_ref.py:11-44)
  plan_ref       per-task index plan (reference blocksvibe/plan_ref.py:9-132; pinned against the reference
                 module's own plans, tests/golden/plan.json)
  recursion_bits the GPU recursion's exact bit-level semantics (Cauchy and pole steps) on top of mpf_ref
  derivs_ref     residues -> numerators, mpmath (reference tests/test_blocks_derivatives.py:13-80)
  output_ref     write_output maps, mpmath (reference blocksvibe/output.py:35-95, tests/test_blocks_output.py:13-47)
  c/             a C restatement of the recursion kernels (CPU baseline and large-size bit parity)
"""
