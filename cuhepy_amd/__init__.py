"""cuhepy_amd: MI355X-native (gfx950, HIP) Paillier batched-modexp path.

Drop-in for the GPU extensions of XTraceAI/cuhepy
(``cuhepy.paillier._gpu_ext`` and ``cuhepy.paillier._lookup_gpu_ext``):
Python host code over hand-written HIP kernels behind the C ABI in
``include/chx.h`` (loaded with ctypes from ``cuhepy_amd/libchx.so``).
"""

__version__ = "0.1.0"
