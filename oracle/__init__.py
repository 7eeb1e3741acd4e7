"""FP64 CPU oracle for the SISL dynamical-core step.

TEST INFRASTRUCTURE ONLY. This package restates, in plain NumPy, the
reference algorithm of Suetes' regional 3-D semi-implicit semi-Lagrangian
step (``suetes/regional3d/steppers.py:720-1073`` and everything it calls).
It is the checker the HIP product path is compared against: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import it.
The product package ``suetes_amd`` never imports it and has no CPU fallback.

Every function cites the reference ``file:line`` it follows. Parity pins
(closed forms, invariants, textbook routines) live in ``tests/test_oracle_*.py``.
Parity status per function:

* tricubic, trilinear, departure points, C-grid operators, tendencies,
  linear operator, preconditioner, FFSL, sponge, full step: pinned
  (closed forms / invariants / SciPy cross-checks, see tests).
* GMRES (restatement of ``jax.scipy.sparse.linalg.gmres``, batched mode):
  the solution is pinned by residual tests and by SciPy's GMRES; the iterate
  sequence of JAX (Gram-Schmidt variant, rounding) is parity unpinned because
  JAX is not installed and its source is not vendored.
* DaviesSponge float32 masks: the cos() rounding of XLA is parity unpinned
  (the oracle and the product both use NumPy float32).
* ``split_explicit``: the split-explicit RK3 core (steppers.py:1076-1737),
  pinned by the reference's own regressions (theta' Exner flux, closed-box
  budgets), the upwind flux on cell averages of polynomials, the coupled
  off-centred column equations and a resting atmosphere.
"""
