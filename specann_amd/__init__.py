"""MI355X-native SpecAnn alpha-sweep: HIP kernels behind a C ABI, driven from Python host code.

The solver classes mirror the reference's plugin API (models/spectral_annealing.py:172-180,
models/spec_adj.py:44-48, models/spec_signed_lap.py:46-50); the compute runs in the HIP library
``libspecann.so`` built from ``specann_amd/csrc``.
"""
__version__ = "0.1.0"
