"""CPU oracle for the crystal-plasticity hot path — TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 NumPy/SciPy restatement of the reference algorithm
(SuperkakaSCU/JAX-CPFEM, built on the external jax-fem library).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it,
and only as the reference those compare against / the CPU baseline: the product path (``cpfem_amd``) never
imports, links or executes anything from here, and fails loudly when its HIP
library is missing.

Modules
-------
``cp_local``  per-quadrature-point Kalidindi update (rotations, Schmid tensors,
              latent matrix, cubic stiffness, local Newton with backtracking,
              IFT consistent tangent, history update, Cauchy stress).
``fe``        hex8 2x2x2 finite elements, element residual / stiffness,
              COO->CSR assembly, row-elimination Dirichlet BCs, Jacobi-BiCGSTAB
              with jax.scipy semantics, the global Newton loop, load-step driver.

Pinning (what checks each function beyond retyping its formula) is listed in
each module header; every function cites the reference file:line it follows.
"""
