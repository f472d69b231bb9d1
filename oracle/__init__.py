"""Complex128 CPU oracle for the slab Foldy-Lax solve -- TEST INFRASTRUCTURE ONLY.

This package restates, in NumPy/SciPy, the reference algorithm of
tmnestor/MultipleScatteringCalculations for the FFT-accelerated slab solve
``(I - G0 T0) psi = psi0``.  Every function cites the reference file:line it
follows (paths relative to the reference repository root).

It is the checker, never the product:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it;
* the HIP product path in ``multiplescatteringcalculations_amd`` never imports,
  links or calls anything here.

Pinning status of each module (how a plausible mistake would be caught):

* ``tmatrix``     -- pinned by the Mathematica golden A,B,C,T1,T2,T3 values
  (cubic_scattering/tests/test_cubic_tmatrix.py:463-485), by quadrature of the
  geometric constants g0, j1, j2, k1, by a Duffy-transformed cubature of
  Gamma0 = int_cube G_11 and by the Born limit.
* ``propagator``  -- pinned by finite differences of the Kupradze tensor
  (G -> Gd -> Gdd), by the engineering-strain finite-difference checks of the
  C, H and S Voigt blocks, and by reflection/rotation covariance.
* ``slab``        -- matvec pinned by a brute-force dense O(N^2) sum on tiny
  slabs (and an explicit circular sum in periodic mode); incident field by its
  closed form; R_PP by the Kennett single-layer closed form at the 5%/15%
  physics level and by golden vectors written from the reference
  (tests/golden/).
* ``gmres``       -- pinned against numpy.linalg.solve on dense systems.
"""
