"""CPU FP64 oracle for the 2D unsplit wave-propagation step.

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, as the checker.  The product (``clawpack_amd``) never does.

Two independent restatements of the classic Clawpack 2D step live here:

* ``claw2d_oracle.c`` (bound by :mod:`oracle.corc`): slice-by-slice, in the loop order of the
  upstream Fortran ``step2``/``flux2``/``limiter`` and the ``rpn2``/``rpt2`` solvers
  ``shallow_roe_with_efix`` and ``euler_4wave``.
* :mod:`oracle.torch_step`: a whole-array float64 PyTorch restatement of the same step, written
  in the gather form (per-interface quantities, then per-cell sums).  It is differentiable and
  serves as the autograd reference for the adjoint.

Parity status: the upstream Fortran is absent from the reference tree (its pyclaw/riemann
submodules are empty, .gitmodules:1-12), so both restatements are pinned by closed forms and
invariants in tests/test_oracle_*.py (Roe property, consistency, conservation, symmetry,
fixed points, 1D reduction, convergence) and by their agreement with each other; they are
"parity unpinned" with respect to bitwise upstream Fortran output.
"""
