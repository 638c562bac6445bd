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
  V-cycle            SPD check (z.r > 0) and a dense reference on small levels
  Cayley / tangent   |m| = 1, orthogonality of the rotation, closed-form small rotations
  Steihaug PCG       one iteration equals the closed-form Cauchy step; exact solve
                     on an SPD-projected system
  line search        closed-form quadratic model (known accepted step)
  pcohen_hs / lbfgs  energy decrease per accepted step, |m| = 1, S,Y tangent, s.y > 0,
                     and both methods converge to the same minimum.
  lbfgs_two_loop     the explicit product of BFGS inverse updates from H0 = gam diag(invMr),
                     with the ring of pairs wrapped, and the secant condition H y = s.
Parity of the L-BFGS iterate trajectory against the reference JAX code is
unpinned: JAX is not installed here, so only the checks above cover it.
"""
