"""CPU oracle for the dense DMRG hot path — TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy, what the reference (ITensorsGPU.jl on
top of ITensors.jl 0.2 / NDTensors) computes on the DMRG two-site-update path.
It is the checker for the HIP library: only `tests/`, `__graft_entry__.smoke()`
and the `cpu_baseline` leg of `bench.py` may import it. The product path
(`itensorsgpu.jl_amd/`) never imports, calls or falls back to anything here.

Modules and the reference code each follows
  tensor_ops  contraction-label bookkeeping, contraction, permute, permuted add
              (src/tensor/cudense.jl:83-110,238-331,333-445,447-500;
               src/tensor/dense.jl:1-48)
  truncate    truncate! with both the NDTensors CPU semantics and the
              reference GPU kernel's semantics (src/tensor/cutruncate.jl:1-93)
  linalg      eigen/svd/qr of order-2 tensors (src/tensor/culinearalgebra.jl)
  philox      Philox4x32-10 + Box-Muller Gaussian stream (replaces CURAND
              randn!, src/cuitensor.jl:37-49, src/mps/cumps.jl:44-53)
  models      closed-form MPOs and exact-diagonalisation Hamiltonians
  dmrg        ProjMPO, davidson, replacebond!/factorize, two-site dmrg
              (ITensors 0.2 algorithms reached via test/dmrg.jl:27,75)
  mlp         forward/backward/Adam of the learned Davidson preconditioner
              (no reference counterpart)

How each part is pinned (details in the tests):
  tensor_ops  brute-force index loops on tiny inputs; closed-form
              transpose/add vectors of test/test_cudense.jl:16-42
  truncate    the three golden vectors of test/test_cutruncate.jl:9-17
  linalg      reconstruction/orthogonality invariants
              (test/test_cuitensor.jl:105-130) and scipy LAPACK
  philox      Random123 known-answer vectors for philox4x32-10
  models      exact diagonalisation (Kronecker-product Hamiltonians)
  dmrg        ED ground-state energies (N=10,12,20 Heisenberg), the critical
              TFIM closed form of test/dmrg.jl:79, the variational bound
  mlp         central finite differences
"""
