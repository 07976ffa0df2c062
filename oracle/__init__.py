"""Oracle: CPU restatements of Mahakala's per-ray hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in ``mahakala_amd`` may import, link or
run anything in this directory; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and only to check results
or to time the CPU baseline.

Two restatements live here:

* ``mahakala_np``  -- NumPy FP64, written to follow the reference line by
  line (same expression graph for the metric, forward-mode derivatives of that
  expression via dual numbers in place of ``jax.jacfwd``, an LU inverse in
  place of ``jnp.linalg.inv``, the fixed-N scan semantics and trim rule,
  explicit-Euler transfer scanned backward).  This is the parity oracle.
* ``c/mk_oracle.c`` -- plain C FP64 with closed-form metric derivatives and
  the closed-form inverse, OpenMP over rays.  It is pinned against
  ``mahakala_np`` on small inputs and used for large parity cases and as the
  CPU baseline that ``bench.py`` times.

The reference itself (Python + JAX) cannot run here: JAX is not installed.
"""
