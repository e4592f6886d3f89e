"""CPU oracle for the range-analysis hot path of neural-implicit-queries.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package may import, call or
execute anything under ``oracle/``; only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it, and there only as the
checker (or, for ``cpu_baseline``, as the CPU restatement that is timed).

This is a NumPy restatement of the reference's JAX algorithm, written from the
reference source read as text (JAX is not installed in this container, so the
reference cannot be run).  Every function cites the reference ``file:line`` it
follows.  All functions are batched over a leading box/ray/point axis and are
generic over the floating-point dtype (float32 restates the reference; float64
is the high-precision oracle used for the soundness checks).

Canonical accumulation order (where the reference leaves it to XLA):
  * radius sums over error-term rows are sequential in row order;
  * dense products use NumPy/BLAS ``matmul`` in the given dtype.

How each module is pinned (no reference golden vectors exist for this path):
  * ``mlp``      - key grammar checked against the reference's sample ``.npz``
                   files; point evaluation checked against closed forms.
  * ``affine``   - exactness on activation-free MLPs (closed form), Chebyshev
                   ReLU special cases, sampled soundness, FP32 vs FP64.
  * ``slope``    - exactness of the slope bound on activation-free MLPs,
                   sampled soundness.
  * ``tree``     - brute-force ancestor walk on the finest grid.
  * ``rays``     - closed-form plane hits; no sign change before the hit.
  * ``closest``  - plane closed form (up to the reference's node-centre rule).
  * ``mc``       - table invariants; closed-form area of a plane cut.
The TanH rule has no reference counterpart (the reference implements relu,
elu and sin only); it is pinned by sampled soundness and FP64 invariants only.
"""
