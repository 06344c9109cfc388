"""ctypes binding of ``libmi355df.so`` (the HIP library's C ABI).

The library is built ahead of time (``python -m electrolites_amd.build``)
and loaded from the package directory.  There is no fallback: if the library
is missing or its ABI version differs, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, 'libmi355df.so')
ABI_VERSION = 12
DTYPE_F64, DTYPE_F32 = 0, 1
PREC_FP64 = 0
PREC_FP32 = 1

P = ctypes.c_void_p
L = ctypes.c_long
I = ctypes.c_int
SZ = ctypes.c_size_t
D = ctypes.c_double

# name -> (restype, argtypes)
_SIGS = {
    'mdf_abi_version': (I, []),
    'mdf_last_error': (ctypes.c_char_p, []),
    'mdf_device_info': (I, [ctypes.POINTER(I), ctypes.POINTER(ctypes.c_longlong),
                            ctypes.c_char_p, I]),
    'mdf_get_j_workspace': (SZ, [L, L, I]),
    'mdf_get_j': (I, [P, I, L, L, L, P, I, L, P, L, P, L, I, P, SZ, P]),
    'mdf_get_j_noalloc': (I, [P, I, L, L, L, P, I, L, P, L, P, L, I, P]),
    'mdf_get_k_workspace': (SZ, [L, I, I, I, I, L]),
    'mdf_get_k': (I, [P, I, L, L, L, P, I, P, P, I, L, I, P, L, I, P, SZ, P]),
    'mdf_pack_dm': (I, [P, I, I, L, P, L, I, P, L, P]),
    'mdf_unpack_pairs': (I, [P, I, L, P, L, I, P, L, P]),
    'mdf_build_pair_map': (I, [P, L, I, P, P]),
    'mdf_fill_zero': (I, [P, L, P]),
    'mdf_cast_f32': (I, [P, L, L, L, P, L, P]),
    'mdf_widen_f32': (I, [P, L, L, L, P, L, P]),
    'mdf_copy_2d': (I, [P, L, P, L, L, L, P]),
    'mdf_transpose': (I, [P, I, L, L, L, P, L, P]),
    'mdf_j_rho': (I, [P, I, L, L, L, P, I, L, P, L, I, P, SZ, P]),
    'mdf_j_vj': (I, [P, I, L, L, L, P, I, L, P, L, I, P, SZ, P]),
    'mdf_int3c2e': (I, [P, P, P, P, P, P, L, P, P, L, P]),
    'mdf_int2c2e': (I, [P, P, P, P, L, P]),
    'mdf_k_step1': (I, [P, I, L, L, L, P, I, P, L, I, P, L, L, P]),
    'mdf_k_step1_lower': (I, [P, I, L, L, L, P, I, P, L, I, P, L, L, P]),
    'mdf_int3c2e_ip': (I, [P, P, P, P, P, P, P, L, P, P, P, I, P, P, P]),
    'mdf_int2c2e_ip': (I, [P, P, P, P, L, P, I, P, P]),
    'mdf_int1e': (I, [P, I, P, P, I, P, P, P, P, P, P]),
    'mdf_gemm': (I, [I, I, I, L, L, L, D, P, L, P, L, D, P, L, P]),
    'mdf_gemm_batched': (I, [I, I, I, L, L, L, D, P, L, L, P, L, L, D, P, L, L, L, P]),
    'mdf_gemm_pairs_batched': (I, [I, L, L, D, P, L, L, P, L, L, P, P, L, P, P, L, P]),
    'mdf_sym_pack': (I, [L, P, L, P, L, D, D, L, P]),
    'mdf_sym_unpack': (I, [L, P, L, P, L, D, D, L, P]),
    'mdf_sp2_tiles': (L, [L]),
    'mdf_sp2_begin': (I, [L, P, L, P, P]),
    'mdf_sp2_step': (I, [L, D, P, L, P, P, L, P, P]),
    'mdf_potrf_workspace': (SZ, [L]),
    'mdf_potrf': (I, [P, L, L, P, P, SZ, P]),
    'mdf_trtri_workspace': (SZ, [L]),
    'mdf_trtri': (I, [P, L, L, P, P, L, P, SZ, P]),
    'mdf_trmm': (I, [P, L, L, P, L, L, P, L, P]),
    'mdf_grouped_plan_create': (I, [P, I, I, I, I, I, SZ, ctypes.POINTER(P), ctypes.POINTER(SZ)]),
    'mdf_grouped_jk': (I, [P, P, SZ, P]),
    'mdf_grouped_plan_destroy': (I, [P]),
    'mdf_grouped_plan_info': (I, [P, ctypes.POINTER(L), I]),
}


class MDFError(RuntimeError):
    """A call into libmi355df failed; the message is the library's."""


_lib = None


def load():
    """Load and validate the library (once per process)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise MDFError(f'{LIBPATH} is missing: build it with '
                       '`python -m electrolites_amd.build` (there is no '
                       'non-HIP fallback for this path)')
    lib = ctypes.CDLL(LIBPATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    ver = lib.mdf_abi_version()
    if ver != ABI_VERSION:
        raise MDFError(f'libmi355df ABI {ver}, expected {ABI_VERSION}; rebuild')
    _lib = lib
    return lib


def exported():
    """Names of the ``mdf_*`` entry points this binding declares."""
    return sorted(_SIGS)


def call(name, *args):
    """Call ``name``; raise :class:`MDFError` with the library's message on a
    non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.mdf_last_error().decode(errors='replace')
        raise MDFError(f'{name} failed ({rc}): {msg}')
    return rc


def query(name, *args):
    return getattr(load(), name)(*args)
