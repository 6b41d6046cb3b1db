"""ctypes binding of ``libchx.so`` (C ABI declared in ``include/chx.h``).

The library is loaded from the package directory. There is no fallback: if
the library is missing or cannot be loaded, importing this module raises
ImportError, and every device entry point fails loudly.
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_size_t, c_uint8, c_uint32, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CHX_LIBRARY", os.path.join(_HERE, "libchx.so"))

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"HIP library not built: {LIB_PATH} is missing. Build it with "
        "`python -c 'import __graft_entry__ as g; g.build()'` or `python -m cuhepy_amd._build`."
    )

# One HIP runtime per process. PyTorch bundles its own libamdhip64 /
# libhsa-runtime64 with the same sonames as the system ones; whichever is
# loaded first is the one both sides bind to. If the system runtime were loaded
# first, PyTorch would still load its own copy and the second runtime to open
# the device finds no GPU. So when PyTorch is installed it is imported first
# and the library binds to PyTorch's runtime.
try:
    import torch  # noqa: F401
except ImportError:  # pragma: no cover - plain C/ctypes use without PyTorch
    pass

try:
    lib = ctypes.CDLL(LIB_PATH)
except OSError as e:  # pragma: no cover - depends on the ROCm runtime
    raise ImportError(f"cannot load {LIB_PATH}: {e}") from e

u32p = POINTER(c_uint32)
u8p = POINTER(c_uint8)
vp = c_void_p

CHX_OK, CHX_EINVAL, CHX_EDEVICE, CHX_ENOMEM, CHX_EUNSUPPORTED = 0, 1, 2, 3, 4
OP_MODEXP, OP_MULMOD, OP_LOOKUP, OP_RAGGED = 0, 1, 2, 3



class RaggedSubStruct(ctypes.Structure):
    """chx_ragged_sub"""
    _fields_ = [("mod", c_void_p), ("base", c_void_p), ("exps", c_void_p), ("out", c_void_p),
                ("count", c_uint32), ("stride", c_uint32), ("e_limbs", c_uint32), ("e_bits", c_uint32)]


# name -> (restype, argtypes); the list is the exported surface of chx.h
SIGNATURES = {
    "chx_version": (c_char_p, []),
    "chx_last_error": (c_int, [c_char_p, c_size_t]),
    "chx_ctx_create": (c_int, [c_int, POINTER(vp)]),
    "chx_ctx_destroy": (c_int, [vp]),
    "chx_ctx_set_stream": (c_int, [vp, vp]),
    "chx_ctx_get_stream": (vp, [vp]),
    "chx_ctx_sync": (c_int, [vp]),
    "chx_ctx_set_tuning": (c_int, [vp, c_int, c_int, c_int, c_int]),
    "chx_supported_tpi": (c_int, [c_int, POINTER(c_int), c_int]),
    "chx_supported_tpi_op": (c_int, [c_int, c_int, POINTER(c_int), c_int]),
    "chx_ctx_scratch_bytes": (c_size_t, [vp]),
    "chx_mod_create": (c_int, [vp, u32p, c_int, POINTER(vp)]),
    "chx_mod_destroy": (c_int, [vp]),
    "chx_mod_limbs": (c_int, [vp]),
    "chx_mulmod": (c_int, [vp, vp, u32p, u32p, u32p, c_size_t]),
    "chx_mulmod_dev": (c_int, [vp, vp, vp, vp, vp, c_size_t, c_size_t]),
    "chx_mulmod_bcast_dev": (c_int, [vp, vp, vp, u32p, vp, c_size_t, c_size_t]),
    "chx_mulmod_operand_words": (c_int, [vp]),
    "chx_mulmod_operands_dev": (c_int, [vp, vp, vp, vp, c_size_t]),
    "chx_mulmod_operands": (c_int, [vp, vp, u32p, vp, c_size_t]),
    "chx_mulmod_multi_dev": (c_int, [vp, vp, vp, c_size_t, c_size_t, c_size_t, c_int, POINTER(vp), POINTER(vp),
                                     c_size_t]),
    "chx_mulmod_multi_stamped_dev": (c_int, [vp, vp, vp, c_size_t, c_size_t, c_size_t, c_int, POINTER(vp),
                                             POINTER(vp), c_size_t, vp, vp]),
    "chx_ctx_wall_clock_hz": (c_int, [vp, POINTER(c_double)]),
    "chx_modexp": (c_int, [vp, vp, u32p, c_size_t, u32p, c_int, u32p]),
    "chx_modexp_dev": (c_int, [vp, vp, vp, vp, c_size_t, c_size_t, u32p, c_int]),
    "chx_modexp_each": (c_int, [vp, vp, u32p, c_size_t, u32p, c_int, u32p]),
    "chx_modexp_each_dev": (c_int, [vp, vp, vp, vp, c_int, c_int, vp, c_size_t, c_size_t]),
    "chx_ragged_create": (c_int, [vp, POINTER(RaggedSubStruct), c_int, c_int, c_int, c_int, POINTER(vp)]),
    "chx_ragged_run": (c_int, [vp, c_int]),
    "chx_ragged_run_plain": (c_int, [vp]),
    "chx_ragged_destroy": (c_int, [vp]),
    "chx_paillier_key_create": (c_int, [vp, u32p, c_int, u32p, u32p, u32p, u32p, c_int, u32p, u32p, u32p,
                                        POINTER(vp)]),
    "chx_paillier_key_destroy": (c_int, [vp]),
    "chx_paillier_width": (c_int, [vp]),
    "chx_paillier_encrypt": (c_int, [vp, vp, u32p, u32p, u32p, c_size_t]),
    "chx_paillier_encrypt_dev": (c_int, [vp, vp, vp, vp, vp, c_size_t, c_size_t]),
    "chx_paillier_decrypt": (c_int, [vp, vp, u32p, u32p, c_size_t]),
    "chx_paillier_decrypt_dev": (c_int, [vp, vp, vp, vp, c_size_t, c_size_t]),
    "chx_paillier_decrypt_crt": (c_int, [vp, vp, u32p, u32p, c_size_t]),
    "chx_paillier_decrypt_crt_dev": (c_int, [vp, vp, vp, vp, c_size_t, c_size_t]),
    "chx_paillier_decrypt_crt_noalloc": (c_int, [vp, vp, vp, vp, c_size_t, c_size_t]),
    "chx_paillier_decode_hamming": (c_int, [vp, vp, u32p, u32p, u32p, c_size_t]),
    "chx_lookup_key_create": (c_int, [vp, u32p, c_int, u32p, c_int, u32p, u32p, c_int, u32p, c_int,
                                      POINTER(vp)]),
    "chx_lookup_key_destroy": (c_int, [vp]),
    "chx_lookup_width": (c_int, [vp]),
    "chx_paillier_key_set_idempotents": (c_int, [vp, vp, u32p, u32p]),
    "chx_paillier_key_has_idempotents": (c_int, [vp]),
    "chx_lookup_key_set_crt": (c_int, [vp, vp, u32p, u32p, c_int, u32p, u32p]),
    "chx_lookup_key_has_crt": (c_int, [vp]),
    "chx_lookup_key_set_table_rows": (c_int, [vp, vp, c_int]),
    "chx_lookup_encrypt": (c_int, [vp, vp, u8p, u8p, c_int, u32p, c_size_t]),
    "chx_lookup_encrypt_dev": (c_int, [vp, vp, vp, vp, c_int, vp, c_size_t, c_size_t]),
    "chx_lookup_decrypt": (c_int, [vp, vp, u32p, u32p, c_size_t]),
    "chx_lookup_decode_hamming": (c_int, [vp, vp, u32p, u32p, u32p, c_size_t]),
    "chx_lookup_reencrypt": (c_int, [vp, vp, vp, u32p, u8p, c_int, u32p, c_size_t]),
    "chx_lookup_reencrypt_dev": (c_int, [vp, vp, vp, vp, vp, c_int, vp, c_size_t, c_size_t]),
    "chx_aos_to_il_dev": (c_int, [vp, vp, vp, c_size_t, c_int, c_size_t]),
    "chx_il_to_aos_dev": (c_int, [vp, vp, vp, c_size_t, c_int, c_size_t]),
    "chx_host_schedule": (c_int, [u32p, c_int, c_int, u32p, c_int]),
    "chx_host_mont_consts": (c_int, [u32p, c_int, u32p, u32p]),
    "chx_host_paillier_check": (c_int, [u32p, c_int, u32p, u32p, u32p, u32p, c_int, u32p, u32p, u32p]),
    "chx_probe_mad": (c_int, [vp, c_int, c_int, POINTER(c_double), POINTER(c_double)]),
    "chx_probe_rate": (c_int, [vp, c_int, c_int, c_int, POINTER(c_double), POINTER(c_double)]),
    "chx_probe_copy": (c_int, [vp, vp, vp, c_size_t, c_size_t, c_int, c_int]),
    "chx_ctx_set_timing": (c_int, [vp, c_int]),
    "chx_ctx_kernel_time": (c_int, [vp, c_char_p, POINTER(c_double), POINTER(c_int)]),
    "chx_ctx_reset_timing": (c_int, [vp]),
}

for _name, (_res, _args) in SIGNATURES.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib.chx_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


class ChxError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"chx error {code}: {message}")
        self.code = code


def check(rc: int) -> None:
    """Map a C-ABI status to a Python exception (ValueError for bad input)."""
    if rc == CHX_OK:
        return
    msg = last_error()
    if rc == CHX_EINVAL:
        raise ValueError(msg)
    raise ChxError(rc, msg)
