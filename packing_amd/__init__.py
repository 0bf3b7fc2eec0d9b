"""MI355X-native relaxation core of the Santa-2025 packing solver.

Drop-in modules named after the reference's: pack_cuda, pack_cuda_lut,
lbfgs_torch_parallel, lap_batch, pack_cost, pack_dynamics.  All compute runs in
hand-written HIP kernels (libpacking.so, C-ABI in include/packing.h).
"""
__version__ = "0.1.0"
