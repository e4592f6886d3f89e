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
  * ``frustum``  - single-pixel frusta equal the single-ray oracle; a plane's
                   hits; the subdivision partitions the image.
  * ``intersect``- half-spaces (overlapping, disjoint, touching); grid brute
                   force.
  * ``sampling`` - Philox4x32-10 against the Random123 known-answer vectors;
                   round-size independence against a one-at-a-time loop;
                   closed-form slab, volume and centroid of half-spaces.
  * ``funcs``    - the sdf mode against the exact ball-misses-plane test of a
                   unit-normal plane; sign-constancy of determinate boxes.
  * ``render``   - plane normals exactly; float64 gradients of a smooth net;
                   the matcap lookup against scipy.ndimage.map_coordinates.
The golden fixtures under tests/golden/ (written by make_golden.py) freeze the
outputs of these modules for one seeded model, so later changes are caught.
The TanH rule has no reference counterpart (the reference implements relu,
elu and sin only); it is pinned by sampled soundness and FP64 invariants only.
Parity unpinned (no available check fixes them): the FP32 bits of XLA's dot
and reduction order, XLA's transcendental ulps, and the tie order of argsort /
top_k in affine_truncate / affine_append (ties are broken lowest index first).
"""
