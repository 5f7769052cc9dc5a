"""Loader and ctypes signatures for libvkhel.so."""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libvkhel.so")
REPO = os.path.dirname(_HERE)
HEADERS = [os.path.join(REPO, "include", "vkhel", h)
           for h in ("vkhel.h", "vkhel_ext.h")]

_LIB = None


class VkhelError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"vkhel error {code}: {message}")
        self.code = code


_p = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int
_pu64 = ctypes.POINTER(ctypes.c_uint64)

# name: (restype, argtypes)
SIGNATURES = {
    "vkhel_ctx_create": (_p, []),
    "vkhel_ctx_destroy": (None, [_p]),
    "vkhel_ntt_tables_create": (_p, [_u64, _u64, _u64]),
    "vkhel_ntt_tables_destroy": (None, [_p]),
    "vkhel_vector_create": (_p, [_p, _u64]),
    "vkhel_vector_create2": (_p, [_p, _u64, ctypes.c_bool]),
    "vkhel_vector_destroy": (None, [_p]),
    "vkhel_vector_dup": (_p, [_p]),
    "vkhel_vector_copy_from_host": (None, [_p, _pu64]),
    "vkhel_vector_map": (None, [_p, ctypes.POINTER(_p), ctypes.c_size_t]),
    "vkhel_vector_unmap": (None, [_p]),
    "vkhel_vector_elemfma": (None, [_p, _p, _p, _u64, _u64]),
    "vkhel_vector_elemmod": (None, [_p, _p, _u64, _u64]),
    "vkhel_vector_elemmul": (None, [_p, _p, _p, _u64]),
    "vkhel_vector_elemgtadd": (None, [_p, _p, _u64, _u64]),
    "vkhel_vector_elemgtsub": (None, [_p, _p, _u64, _u64, _u64]),
    "vkhel_vector_forward_transform": (None, [_p, _p, _p]),
    "vkhel_vector_inverse_transform": (None, [_p, _p, _p]),
    "vkhel_ext_set_abort_on_error": (None, [_int]),
    "vkhel_ext_last_error": (_int, []),
    "vkhel_ext_last_error_message": (ctypes.c_char_p, []),
    "vkhel_ext_clear_error": (None, []),
    "vkhel_ext_ctx_stream": (_p, [_p]),
    "vkhel_ext_ctx_synchronize": (_int, [_p]),
    "vkhel_ext_ctx_device_name": (ctypes.c_char_p, [_p]),
    "vkhel_ext_vector_device_ptr": (_p, [_p]),
    "vkhel_ext_vector_length": (_u64, [_p]),
    "vkhel_ext_vector_view": (_p, [_p, _u64, _u64]),
    "vkhel_ext_vector_read": (_int, [_p, _u64, _u64, _pu64]),
    "vkhel_ext_vector_write": (_int, [_p, _u64, _u64, _pu64]),
    "vkhel_ext_set_staging_pool": (_int, [_p, _int]),
    "vkhel_ext_staging_bytes": (_u64, [_p]),
    "vkhel_ext_host_alloc": (_p, [ctypes.c_size_t]),
    "vkhel_ext_host_free": (None, [_p]),
    "vkhel_ext_vector_write_async": (_int, [_p, _u64, _u64, _pu64, _p]),
    "vkhel_ext_vector_read_async": (_int, [_p, _u64, _u64, _pu64, _p]),
    "vkhel_ext_tables_info": (_int, [_p, _pu64, _pu64, _pu64]),
    "vkhel_ext_tables_copy": (_int, [_p, _pu64, _pu64, _pu64, _pu64]),
    "vkhel_ext_tables_upload": (_int, [_p, _p]),
    "vkhel_ext_tables_create_device": (_p, [_p, _u64, _u64, _u64]),
    "vkhel_ext_tables_root_ok": (_int, [_p]),
    "vkhel_ext_tables_read_device": (_int, [_p, _p, _pu64, _pu64, _pu64, _pu64]),
    "vkhel_ext_set_ntt_algo": (_int, [_p, _int]),
    "vkhel_ext_autotune": (_int, [_p, _u32, _u32, _u32]),
    "vkhel_ext_tile_bits": (_int, [_p, _u32]),
    "vkhel_ext_set_tile_bits": (_int, [_p, _u32, _int]),
    "vkhel_ext_pass_mode": (_int, [_p, _u32]),
    "vkhel_ext_set_pass_mode": (_int, [_p, _u32, _int]),
    "vkhel_ext_set_handoff_ahead": (_int, [_p, _int]),
    "vkhel_ext_split": (_int, [_p, _u32]),
    "vkhel_ext_set_split": (_int, [_p, _u32, _int]),
    "vkhel_ext_tile_bits2": (_int, [_p, _u32]),
    "vkhel_ext_set_tile_bits2": (_int, [_p, _u32, _int]),
    "vkhel_ext_autotune_save": (_int, [_p, ctypes.c_char_p]),
    "vkhel_ext_autotune_load": (_int, [_p, ctypes.c_char_p]),
    "vkhel_server_create": (_p, [_p, _p]),
    "vkhel_server_destroy": (None, [_p]),
    "vkhel_server_submit": (_u64, [_p, _int, _p, _p, _p, ctypes.POINTER(_p), _u32]),
    "vkhel_server_submit_rns": (_u64, [_p, _int, _p, _p, ctypes.POINTER(_p), _u32,
                                       _p]),
    "vkhel_server_wait": (_int, [_p, _u64, _u64, _p]),
    "vkhel_server_cancel": (_int, [_p, _u64]),
    "vkhel_server_drain": (_int, [_p]),
    "vkhel_server_get_stats": (_int, [_p, _p]),
    "vkhel_server_replay": (_int, [_p, _p, _u64, ctypes.c_double, _u64, _p]),
    "vkhel_ext_vector_fill_uniform": (_int, [_p, _u64, _u64, _u64, _u64]),
    "vkhel_batch_create": (_p, [_p]),
    "vkhel_batch_destroy": (None, [_p]),
    "vkhel_batch_clear": (None, [_p]),
    "vkhel_batch_add": (_int, [_p, _int, _p, _p, _p, ctypes.POINTER(_p), _u32]),
    "vkhel_batch_add_rns": (_int, [_p, _int, _p, _p, ctypes.POINTER(_p), _u32,
                                   _pu64]),
    "vkhel_batch_launch": (_int, [_p]),
    "vkhel_batch_wait": (_int, [_p]),
    "vkhel_batch_query": (_int, [_p]),
    "vkhel_batch_run": (_int, [_p]),
    "vkhel_batch_num_jobs": (_u64, [_p]),
    "vkhel_ext_polymul_noalloc": (_int, [_p, _p, _p, _p, ctypes.POINTER(_p), _u32,
                                         _p]),
}


def header_symbols():
    """Every function name declared in include/vkhel/*.h."""
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(vkhel_\w+)\s*\(", text))
    return sorted(names)


def load():
    """Loads libvkhel.so (no GPU needed to load); raises if it is missing."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with python -c "
            "'import __graft_entry__ as g; g.build()' or make -C vkhel_amd")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L.vkhel_ext_set_abort_on_error(0)
    _LIB = L
    return L


def symbols():
    return sorted(SIGNATURES)


def check(rc=None):
    """Raises VkhelError if the last call on this thread recorded an error."""
    L = load()
    code = L.vkhel_ext_last_error()
    if code:
        msg = L.vkhel_ext_last_error_message().decode(errors="replace")
        L.vkhel_ext_clear_error()
        raise VkhelError(code, msg)
    if rc:
        raise VkhelError(rc, "call failed")
