"""FP64 CPU oracle for the somax QG / shallow-water time step.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/SciPy restatement of
the reference algorithm (jejjohnson/somax v0.0.6 plus the finitevolX v0.0.39
operators it calls) used as the correctness checker.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker or the timed CPU baseline.  The product
(``somax_amd``) never imports, calls or links anything here; it fails loudly
when its HIP library is missing instead of falling back to this code.

Parity pins (each function's docstring names its own):

* ``dst``       -- pinned by the executed outputs of the reference tutorials
                   step09/step10/step11 (ten published numbers), by
                   ``scipy.fft.dst(type=1)`` and by a dense brute-force inverse
                   of the 5-point operator on tiny grids.
* ``modes``     -- pinned by closed forms (``Cl2m @ A @ Cm2l`` diagonal,
                   ``Cl2m @ Cm2l = I``) and, for the rigid-lid variant, by the
                   homogeneous-solution statistics printed in
                   ``notebooks/dev/QG_Multilayer/mqg_demo_retangular.ipynb``.
* ``stencils``  -- Arakawa Jacobian pinned by invariants (J(x, y) = 1,
                   antisymmetry, sum(psi J) = sum(q J) = 0) and by a
                   brute-force continuous-limit check; Laplacian exact on
                   quadratics; WENO-Z exact on constants, 5th order on smooth
                   data, conservative.  The finitevolX index conventions
                   themselves are *parity unpinned* (no importable reference);
                   each assumption is stated where it is made.
* ``qg``        -- zero state -> zero tendency, forcing locality, BC ghost
                   semantics (reference tests), SSP-RK3 order 3, homsol
                   known answers.
* ``swm``       -- rest state -> zero tendency, forcing locality, wall BC,
                   exact mass conservation of the flux-form continuity
                   equation, C-grid dispersion relation.
* ``rk``        -- Tsit5 tableau pinned by the 17 order conditions of orders
                   1-5 and by its dt^5 convergence on y' = lambda y.
* ``reparam``   -- ReparameterizedQG: Q(G(psi)) = (lap - f0^2 A) psi where the
                   stencils are inside the array, divergence-free geostrophic
                   velocity, and the reference's qualitative tests (rest state
                   fixed point, approximate idempotence).
* ``run``       -- the runner's per-field reduction (NaN-skipping min / mean /
                   max, non-finite count): the reference's own cases and a
                   brute-force loop over IEEE-754 bit classes.
"""
