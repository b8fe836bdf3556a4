"""CPU oracle for the microfacet render / per-pixel SVBRDF optimisation path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
directory.  It is used by ``tests/`` (as the checker), by
``__graft_entry__.smoke()`` (as the checker) and by the CPU-baseline leg of
``bench.py`` (timed on the host cores as the reference CPU path).

Modules
-------
microfacet_np
    NumPy restatement (float32 or float64) of ``src/microfacet.py:10-120``,
    the MSE loss of ``src/optimization.py:14,28-29``, the log-space L2 loss
    this project defines, the hand-derived adjoint of the whole chain and the
    Adam update of ``torch.optim.Adam`` (single-tensor order).
microfacet_torch
    Eager PyTorch restatement at the reference's op granularity (one torch
    op per term of ``src/microfacet.py:84-120``), used for autograd
    gradients, ``gradcheck`` and the CPU-baseline timing.
synthetic_np
    Seeded synthetic maps and flash geometry used by the tests.

How each function is pinned (checked by something other than itself) is
listed in the module docstrings; every function here is checked against the
golden vectors in ``tests/golden/`` that were written by the reference's own
``Microfacet`` class, and against closed forms / finite differences.
"""
