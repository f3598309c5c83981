"""CPU oracle for the spectral RK4 time step (TEST INFRASTRUCTURE ONLY).

This package restates, in plain NumPy FP64, the reference algorithm of the
spherical-backend hot path: basis construction, dense point-set and Gauss
transforms, spectral operators, the pseudospectral Jacobian, the
barotropic-vorticity and shallow-water tendencies and the classical RK4
combination.  Each function cites the reference file:line it follows
(paths relative to the reference repository root; ``num/`` stands for
``src/planetary_sandbox/numerics/``, ``phy/`` for
``src/planetary_sandbox/physics/``, ``run/`` for
``src/planetary_sandbox/run/``, ``pla/`` for ``src/planetary_sandbox/planet/``).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker.  The product
path (``aeolus_amd``) never imports it and fails loudly when its HIP library
is missing.

Pinning status (which check pins each function) is listed in each module's
header.  Functions without an independent check say "parity unpinned".
"""
