"""CPU oracle: a NumPy float64 restatement of tackandr/CFD-demo's hot path.

TEST INFRASTRUCTURE ONLY. Nothing in the product package imports, calls or
links this package. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as the
timed CPU baseline), never as the thing measured on the GPU or shipped.

The reference is a JAX program (``src/cfd_demo/*.py``) that cannot run here
(no jax, Python 3.10 < 3.12). Each function below restates one reference
function from its source text, op for op, and cites ``file:line``. Array
layouts follow the reference: states are ``(Np, K, 4)`` with conserved
variables ``(rho, rho*u, rho*v, E)`` on the last axis.

Pinning (each function is checked by something other than a retyped formula):

* reference element: GLL nodes, exact differentiation of degree-4 monomials,
  exact monomial integrals, ``w . LIFT`` = face weights, P1 projection and
  basis identities, face-node index sets (``tests/test_oracle_reference.py``);
* mesh / connectivity / geometry: counts, orientation, symmetric
  connectivity, opposite normals, coincident paired nodes, closed-form
  inverse metric (``tests/test_oracle_mesh.py``);
* pointwise physics, fluxes and limiter functions: round trips, closed forms,
  consistency, conservation, rotational invariance, contact exactness,
  supersonic upwinding, the limiter known-answer table
  (``tests/test_oracle_physics.py``);
* RHS: free-stream preservation on a distorted mesh, closed box at rest, the
  slip-wall closed form, the discrete conservation identity, fourth-order
  consistency against finite differences, and an independent per-element
  loop on tiny meshes (``tests/test_oracle_dg.py``);
* limiting and time integration: average preservation, linear data untouched,
  single-element flattening, positivity repair, the CFL formula, exact end
  time, run conservation, third-order convergence, rejection paths, and the
  Sod exact Riemann solution (``tests/test_oracle_time.py``).

Nothing here is "parity unpinned" except the bits of XLA's reduction order,
which no reference test fixes either.
"""
