"""CPU oracle for the multifrontal LU path -- TEST INFRASTRUCTURE ONLY.

This package restates, on the CPU, what the reference PureUMFPACK.jl computes on
the `splu(A; method=:multifrontal)` -> `F \\ b` / `solve(F, b; refine=k)` path, so
that the HIP library can be checked against it.  It is imported only by
`tests/`, by `__graft_entry__.smoke()` (as the checker) and by the
`cpu_baseline` leg of `bench.py`.  The product path (`pureumfpack_hip`) never
imports, links or runs anything in here.

Modules
-------
matrices   generators for the test matrices (bench/matrices.jl:5-99 plus the
           perturbed / batched / convection-diffusion workloads).
scaling    row_scaling / apply_row_scaling (src/scaling.jl:92-126).
amd        approximate minimum degree on A+A^T and on A^T A
           (src/amd.jl:84-188, the CSparse cs_amd algorithm).
symbolic   etree, postorder, column structure, fundamental supernodes
           (src/symbolic.jl:15-237).
numeric    ctypes front end of mf_oracle.c: multifrontal numeric factor
           (src/multifrontal.jl:73-375), triangular solves (src/solve.jl:10-79),
           refinement (src/interface.jl:123-177); plus a dense brute-force
           restricted-pivoting elimination used to pin the C code.

Pinning (what each function is checked against, in tests/test_oracle_*.py):
  scaling   the doctest at src/scaling.jl:86-89 and closed forms.
  amd       published fill counts (results/SUMMARY.md:32-39,52,56), permutation
            validity; brute-force minimum-degree agreement is NOT claimed.
  symbolic  parent[j] > j, predicted_fill == nnz of a dense symbolic
            elimination (brute force on small n), published nsuper / max front
            (results/SCALING_DIAGNOSIS.md:24-30).
  numeric   dense restricted-pivot Gaussian elimination on small n (brute
            force), reconstruction (Rs.*A)[p,q] = L*U, OpenBLAS and loop
            backends agreeing, np=1 case reducing to unpivoted Doolittle.
  solve     the 2x2 doctests (src/interface.jl:85-90,161-167) and dense solves.
"""
