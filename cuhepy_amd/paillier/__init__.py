"""Drop-in GPU clients for the reference's Paillier extensions.

``cuhepy_amd.paillier._gpu_ext.PaillierGPUClient`` replaces
``cuhepy.paillier._gpu_ext.PaillierGPUClient`` and
``cuhepy_amd.paillier._lookup_gpu_ext.PaillierGPULookupClient`` replaces
``cuhepy.paillier._lookup_gpu_ext.PaillierGPULookupClient``.
"""
