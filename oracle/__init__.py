"""CPU oracle for the EasyFsi pressure projection and neo-Hookean MPM transfers.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``easyfsi_amd``) imports,
links or runs anything under ``oracle/``.  The only callers are ``tests/``,
``__graft_entry__.smoke()`` (as the checker) and ``bench.py``'s CPU-baseline
leg.

Contents
--------
* ``oracle.fluid`` -- NumPy float64 restatement of the FV pressure operator,
  divergence, gradient subtraction, smoothers, V-cycle and flexible PCG of
  ``simulation_core/fluid.py`` (each function cites the lines it follows).
* ``oracle.mpm``   -- NumPy restatement of ``NeoHookeanMpmState._step_kernel``
  (``simulation_core/neo_hookean_mpm.py:695-959``).
* ``oracle/c/``    -- plain C + OpenMP restatement of the same plain PCG +
  V-cycle in float64, used as the CPU baseline and as a fast oracle at sizes
  NumPy is too slow for.

How the oracle is pinned (the reference is Taichi, which is not installed and
whose runtime rejects CPU execution, so nothing of it can be run here):
the reference's own known-answer tests re-expressed in ``tests/``, closed
forms (Richardson convergence of the Laplacian, B-spline partition of unity),
invariants (V-weighted self-adjointness, div-grad adjointness, global
conservation, momentum conservation) and library reductions
(``scipy.sparse.linalg.spsolve`` for the operator's solution,
``numpy.linalg.svd`` for the 3x3 SVD).  Functions not pinned by any of these
say "parity unpinned" in their docstring.
"""
