"""aeolus_amd: MI355X-native spectral time step behind Aeolus's spherical-backend interface.

The numerical classes mirror the reference interface (transform objects,
backends, ``SpectralOperators``, ``BarotropicVorticity``,
``ShallowWaterModel``, ``PrimitiveEquationsModel`` with ``SigmaGrid``,
``rk4_step`` / ``rk4_step_array`` and the run drivers) with PyTorch
device tensors in place of CuPy arrays; every computation runs in the HIP
library ``aeolus_amd/lib/libaeolus_hip.so`` (see ``include/aeolus_hip.h``).
"""
__version__ = "0.1.0"
