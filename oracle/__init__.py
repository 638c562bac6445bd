"""CPU oracle for the micromagnetic hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain NumPy/SciPy, what the reference computes on the
path the HIP library replaces: geometry precompute, the assembled Poisson /
divergence / exchange+anisotropy operators, energy and gradient, the PCG and
AMG V-cycle, and the minimiser pieces (Cayley update, Armijo line search,
Steihaug-PCG preconditioner, pcohen_hs and L-BFGS steps, convergence test).
Every function cites the reference file:line it follows.

``cpu.py`` / ``cpu_minimizer.cpp`` restate the pcohen_hs field step of ``minimize.py`` and
``solvers.py`` in C++ with OpenMP (the multi-core CPU baseline); ``tests/test_oracle_cpu.py``
checks them against the NumPy code.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker. The product package
``mumag_matrixfree2_amd`` never imports it and has no CPU fallback.

How each part is pinned (the tests under ``tests/test_oracle_*.py``):
  geometry           closed forms (regular tets, sum of grad phi = 0, sum Ve = box volume),
                     the reference fixture ``single_solid.npz``
  assembled A, D, G  matrix-free element loops computed independently, symmetry,
                     A.1 = reg on interior rows, exact A.x for linear x
  energy / gradient  analytic KATs of ``tests/test_energy.py`` (helix exchange, Zeeman,
                     45 deg anisotropy, cube demag ~ 1/3), central differences of
                     ``tests/test_gradients.py``, E = 1/2 m.(g + g_z)
  PCG / solve_U      scipy ``spsolve`` of the same masked matrix; residual <= tol
  V-cycle            one level = the exact inverse; on two and three levels the error
                     propagation equals (I - MA)(I - P B_c R A)(I - MA) with the coarse
                     V-cycle B_c (dense random Galerkin hierarchies, both smoothers);
                     symmetric positive definite on the mesh hierarchy; planted mistakes
                     (no post-smoothing, subtracted coarse correction) fail the check
  Cayley / tangent   |m| = 1, rotation by 2 atan|k| about k, orthogonality
  Steihaug PCG       the stagnation exit after one step returns the Cauchy point (2 iterations),
                     negative curvature and zero iterations fall back to z = M g_tan, the 10x
                     clip gives y = 10 z on a near-singular P, exact solve on an SPD system
  line search        a one-node Cayley path with a prescribed ratio d(s): hand-derived accepted
                     step and evaluation count for expansion, contraction (towards s_min),
                     failure (tau = 0 after 16 evaluations), the 15-evaluation expansion cap,
                     non-finite energies and the warm-start chain of the solves; planted
                     changes of the rule fail
  pcohen_hs          restart period 1 reproduces preconditioned steepest descent iterate for
                     iterate; the HS direction is conjugate to the gradient difference up to
                     the projection defect (the Polak-Ribiere beta fails this); restart period
  lbfgs_two_loop     the explicit product of BFGS inverse updates from H0 = gam diag(invMr),
                     with the ring of pairs wrapped, and the secant condition H y = s.
  minimisers         energy decrease per accepted step, |m| = 1, S,Y tangent, s.y > 0, and
                     pcohen_hs and L-BFGS converge to the same minimum.
Parity of the L-BFGS iterate trajectory against the reference JAX code rests on the two-loop
check and these invariants: JAX is not installed here. The SA hierarchy (pyamg's role) is
the build's own and has no reference to be pinned against.
"""
