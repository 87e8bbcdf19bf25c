"""CPU oracle for the SIMP finite-element hot path (test infrastructure only).

This package is a numpy/scipy restatement of topoptlab's per-iteration SIMP
pipeline on structured Q4/H8 grids.  It exists to *check* the HIP product path
and to provide the timed CPU baseline; nothing in the product
(``topoptlab_amd``) may import it.  Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` use it.

Every function cites the reference file:line it restates (paths relative to
the topoptlab repository root).

Pinning status (what checks each part independently of this restatement):

* ``edofmat_2d`` / ``edofmat_3d``: closed form + assembled K of the committed
  mbb2d fixture (tests/golden) -> bit-exact.
* ``ke_elast_2d``: Andreassen/Aage top88 closed form (``ke_elast_2d_top88``)
  and the fixture.
* ``ke_elast_3d``: hand-evaluated entries of the reference closed form
  (K[0,0], K[0,1] at l=1, g=0), symmetry, 6 rigid-body null vectors, patch
  energy of affine fields.
* ``ke_poisson_2d``: closed form [[4,-1,-2,-1],...]/6.
* ``assemble_matrix_filter``: normalisation / matrix == convolution properties
  and brute-force weights.
* BC generators: closed forms checked index by index in the tests.
* ``oc_top88`` / ``topopt_loop``: end-to-end compliance goldens of the
  reference test-suite (233.71457, 216.81366, 208.44294, 473.73370).
* ``pcg`` / ``smoothed_jacobi``: the 3x3 known-answer system of the reference
  tests + scipy's ``spsolve``.
* GMG (``prolongator``, ``galerkin_hierarchy``, ``vcycle``): invariants
  (exact reproduction of coarse multilinear fields, injection at coincident
  nodes, row sums 1, SPD coarse operators) and convergence to the direct
  solution.  The reference's own ``gmg.create_interpolator`` is never called
  by the reference and raises ValueError on every grid (its corner columns
  index past the coarse grid; recorded in tests/golden/interpolator_vectors.npz
  from the reference checkout), so there is no reference P to reproduce; P is
  the intended shape-function prolongation, pinned by the invariants above
  and by the reference's shape-function weights at its relative coordinates.
"""
