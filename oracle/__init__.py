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
                   ReLU special cases, sampled soundness, FP32 vs FP64; the
                   sin slope against its definition (midrange of cos over the
                   interval, dense sampling) and closed forms on [0, pi/2],
                   [-pi/2, pi/2], [0, pi]; truncate_affine's kept rows and
                   affine_append's kept deltas against exhaustive subset
                   search, and both against affine_all when nothing is
                   dropped.
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
  * ``margins``  - not a restatement: it evaluates each decision of the
                   queries in FP32 and FP64 and marks the ones within the
                   FP32 rounding band of their threshold (1e-6 relative plus
                   4x the FP32-FP64 difference), so tests can require every
                   GPU result that differs from the oracle to carry a mark.
The golden fixtures under tests/golden/ (written by make_golden.py) freeze the
outputs of these modules for one seeded model, so later changes are caught.
The TanH rule has no reference counterpart (the reference implements relu,
elu and sin only); it is pinned by sampled soundness, FP64 invariants and its
tangent limit only.
Parity unpinned (no available check fixes them): the FP32 bits of XLA's dot
and reduction order, XLA's transcendental ulps, and the order argsort / top_k
give exact ties in affine_truncate / affine_append (broken lowest index first
here; the kept set itself is pinned).
"""
