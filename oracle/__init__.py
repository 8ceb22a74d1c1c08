"""CPU oracle for the assembly + condensation + Krylov path (test infrastructure only).

This package is a float64 CPU restatement of the reference algorithm and is
used only as the checker: by ``tests/``, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py``.  The product package
``tensormesh_amd`` never imports it.

How each function is pinned (beyond restating the reference):

* ``pattern.pattern_and_slots`` - SciPy canonical CSR exactly as
  ``element_assembler.py:838-861``; pinned by brute-force pair sets on tiny
  meshes, by the closed-form Kuhn-grid nnz count and by 15 entries per
  interior row (tests/test_oracle.py).
* ``assemble.*`` - einsum / inverse / det sequence of
  ``element.py:958-966`` and ``transformation.py:1284-1290``, quadrature sum
  ``element_assembler.py:196-221``, ``index_add_`` reduce
  ``projector.py:95-130``; pinned by single-element known answers (unit
  right triangle, Kuhn tetrahedron), ``K 1 = 0``, symmetry / PSD,
  ``1^T M 1 = |Omega|``, ``b = M f`` and the O(h^2) convergence to the exact
  solution of the 2D quickstart problem.  ``element_matrices_general``
  (constant ``U^T D V`` integrands) is pinned by reducing to the Laplace and
  mass matrices for ``D = diag(0, I)`` and ``D = e_0 e_0^T``, by
  ``K(D^T) = K(D)^T`` (the u basis is the row) and by the exact integral of
  a linear field's convection term (tests/test_general_form.py).
* ``condense.*`` - ``condense.py:182-309, 388-458``; pinned by the 4x4
  tridiagonal fixture of ``tests/operator/test_condenser.py:48-99`` and
  dense block elimination.
* ``solve.*`` - textbook Jacobi-PCG (Saad, Iterative Methods, Alg. 9.1) with
  float64 reductions; pinned against ``scipy.sparse.linalg.spsolve`` and an
  independently recomputed residual.
* ``hyperelastic.*`` - neo-Hookean density of ``builtin.py:285-315`` with
  closed-form P and dP/dF; pinned by central finite differences (h = 1e-6)
  and by torch autograd of the energy, and psi(F = I) = 0.
* ``torch_cpu.*`` - the same Map / Reduce / condensation / Jacobi-PCG in
  float64 torch on the CPU, used only as the timed CPU baseline of
  ``bench.py``; its assembly is checked against ``assemble.element_matrices``
  + ``assemble.reduce`` (tests/test_oracle.py).
* ``facet.*`` - boundary-facet loads of ``facet_assembler.py:123-277,
  368-515``; pinned by closed forms on the sides of the unit square and cube
  (constant loads, ``int x_0``, tractions, zero-sum gradient fluxes, nodal
  values ``c h``) and by the reference's selection rule (tests/test_facet.py).
* ``assemble.element_matrices_quadrature`` / ``load_vector_quadrature`` -
  the quadrature-level Map for spatially varying integrands; pinned by
  ``kappa(centroid) |T| grad . grad`` for linear ``kappa``, the exact
  barycentric integrals of ``x_0 phi_a phi_b`` (order-3 rule) and of
  ``x_0 phi_a`` (tests/test_spatial_forms.py).
* ``elasticity.*`` - Voigt ``B^T C B`` of ``builtin.py:173-179`` and
  ``functional/elasticity.py:132-255``; pinned by the rigid-body modes
  (translations and rotations, which a transposed block fails for
  lam != mu), the constant-strain patch energy
  ``u^T K u = |Omega| (lam tr(eps)^2 + 2 mu eps:eps)`` and equality with the
  neo-Hookean tangent at u = 0 (tests/test_elasticity.py).
"""
