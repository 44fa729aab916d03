"""cnpc_ref -- FP64 CPU restatement of CosmoNPC's mesh-estimator hot path.

TEST INFRASTRUCTURE ONLY.  This package is the oracle the HIP path is checked
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may import it, and only as the checker / the timed CPU
baseline.  The product package ``cosmonpc_amd`` never imports it, and nothing
in the product falls back to it.

It restates, on full complex128 c2c grids exactly as the reference does with
``pmesh.ParticleMesh(dtype="complex128")`` (``mesh_generator.py:189-191``):

* ``conventions`` -- the pmesh conventions the reference relies on (paint
  windows, FFT normalisation, k / x / circular axes).  pmesh is not vendored
  in the reference and cannot be imported here; these conventions are pinned
  by closed forms and invariants in ``tests/test_oracle_*.py``.
* ``angular``     -- Ylm / Legendre / Wigner-3j / associated Legendre / q_ell
  (``math_evaluator.py``), k-bin counting and radial binning, compensations.
* ``box``         -- ``get_mesh_box`` + ``calculate_power_spectrum_box`` +
  ``calculate_bk_sugi_box`` (``mesh_generator.py``, ``clustering_estimator.py``).
* ``survey``      -- the survey meshes, particle sums, N-field, pk (with G_l
  and ``high_order_mode``) and bk_sugi estimators with their shot noise.
* ``catalog``     -- comoving distances and the survey catalog processing
  (``catalog_processor.py``).

Every function is pinned (closed forms, brute-force sums, invariants,
Poisson-ensemble expectations, distant-observer reductions) by
``tests/test_oracle_*.py``; none is left unpinned.

Every function cites the reference file:line it follows.
"""
