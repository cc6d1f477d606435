"""CPU oracle for the partial Green-Jacobi homogenization hot path.

TEST INFRASTRUCTURE ONLY. This package is a plain numpy/scipy FP64 restatement
of the reference algorithm. It is imported by ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` as the
checker and as the timed CPU baseline. The product path (``pgj``) never imports
it and has no CPU fallback.

Modules and the reference files they restate (paths relative to the
reference checkout):

- ``hex8``: exact 2x2x2-Gauss Q1/HEX8 element stiffness, the periodic
  matrix-free operator K(a), rhs, element energies, mean stress and the nodal
  diagonal. Follows ``exp/asg_candidates_20261103/candidates.py:30-54``,
  ``exp/asg_scenarios_20261105/verify.py:8-32``,
  ``solver/experiments/ladecky_comparison/study.py:147-174``,
  ``solver/src/materials/hybrid_kernels.cu:25-106``,
  ``exp/asg_scenarios_20261105/energy.cu:7-24``,
  ``solver/experiments/smooth_contrast/study.py:56-67``.
- ``green``: closed-form FI Green symbol, the FFT Green apply, the projection
  P and the partial Green-Jacobi preconditioner M_alpha^-1. Follows
  ``solver/src/reference_medium/formula_adapter.py:32-68``,
  ``solver/src/reference_medium/formula_kernels.cu:104-126``,
  ``solver/src/spectral/precision_kernels.cu:57-105``,
  ``solver/experiments/adaptive_green_scaling/study.py:61-101``.
- ``pcg``: the resident multi-load PCG recurrence with masks, latches, the
  4000-iteration cap, restarts and true-residual certification, and the
  generic best-iterate ``solve_pcg_batched``. Follows
  ``exp/asg_gpu_resident_20261103/resident.py:28-209``,
  ``exp/asg_sctop_reproduction_20261105/reproduction.py:144-151``,
  ``exp/asg_topopt_20261103/backend.py:123-144``,
  ``solver/src/runtime/fi/source/femlib/batched/pcg.py:20-252``.
- ``design``: the seeded filtered-noise density, SIMP material, C^H by
  energy polarization, sensitivities, and both OC density updates. Follows
  ``exp/asg_sctop_reproduction_20261105/{reproduction.py,outer.cu,verify.py}``,
  ``exp/asg_gpu_resident_20261103/{outer.py,outer.cu}``,
  ``exp/asg_topopt_20261103/study.py:61-100``.

Every function's docstring names the reference lines it follows and the
test in ``tests/test_oracle_*.py`` that pins it.
"""
