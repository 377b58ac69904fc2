"""Learned coarse space / initial guess of the BEM solve (no reference
counterpart): native MLP kernels (``mlp``), the preconditioner built on them
(``precond``) and its training on solver-produced solutions (``train``)."""
