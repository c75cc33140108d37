"""ctypes binding of libwotk.so (the C ABI declared in include/wotk.h).

The library is built in-tree by ``python -m wot_amd.build``.  Loading fails
loudly if it is missing: there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libwotk.so')

WOTK_OK = 0
WOTK_EINVAL = 1
WOTK_EHIP = 2
WOTK_ENAN = 3
WOTK_EWORKSPACE = 4

COST_F64 = 0
COST_BF16X3 = 1
COST_BF16X6 = 2
COST_MODES = {'f64': COST_F64, 'bf16x3': COST_BF16X3, 'bf16x6': COST_BF16X6}

KERNEL_ROW_SWEEP = 0
KERNEL_BATCH_SWEEP = 1
KERNEL_BATCH_SHARED = 2
BATCH_SHARE_EXP = 1
BATCH_SHARE_EPS = 2

DTYPE_F64 = 0
DTYPE_F32 = 1
PUSH_FORWARD = 0
PULL_BACK = 1


class Tile(ctypes.Structure):
    _fields_ = [('rows', ctypes.c_int32), ('threads', ctypes.c_int32)]

    def as_tuple(self):
        return (int(self.rows), int(self.threads))


class TMap(ctypes.Structure):
    """wotk_tmap: one stored transport map (device pointer, row-major)."""
    _fields_ = [('data', ctypes.c_void_p), ('dtype', ctypes.c_int32), ('reserved', ctypes.c_int32),
                ('ld', ctypes.c_int64), ('rows', ctypes.c_int64), ('cols', ctypes.c_int64),
                ('transposed', ctypes.c_void_p), ('ld_t', ctypes.c_int64)]


class Stats(ctypes.Structure):
    _fields_ = [('iterations', ctypes.c_int64),
                ('stabilisations', ctypes.c_int64),
                ('gap_checks', ctypes.c_int64),
                ('gap', ctypes.c_double),
                ('epsilon_final', ctypes.c_double),
                ('max_iter_reached', ctypes.c_int32),
                ('nonfinite', ctypes.c_int32),
                ('primal', ctypes.c_double),
                ('dual', ctypes.c_double),
                ('sweeps', ctypes.c_int64),
                ('sweep_ms', ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_D = ctypes.c_double
_SZ = ctypes.c_size_t

# name -> (restype, argtypes); every symbol include/wotk.h declares
SIGNATURES = {
    'wotk_last_error': (ctypes.c_char_p, []),
    'wotk_abi_version': (ctypes.c_int, []),
    'wotk_cost_workspace_size': (_SZ, [_I64, _I64, _I64, _I32]),
    'wotk_cost_sqeuclid': (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _I32, _P, _I64, _P, _I64, _P, _P, _P,
                                          _SZ, _P]),
    'wotk_sqeuclid_f64': (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _I64, _P]),
    'wotk_median_workspace_size': (_SZ, [_I64]),
    'wotk_median_f64': (ctypes.c_int, [_P, _I64, _P, _P, _P, _SZ, _P]),
    'wotk_cost_import': (ctypes.c_int, [_P, _I64, _I64, _I64, _D, _P, _I64, _P, _I64, _P]),
    'wotk_solver_workspace_size': (_SZ, [_I64, _I64]),
    'wotk_transport_stablev2': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _I64, _P, _D, _D, _D, _I64, _D,
                                               _D, _I64, _I64, _P, _I64, _P, _P, _P, _P, _SZ,
                                               ctypes.POINTER(Stats), _P]),
    'wotk_scratch_capacity': (_SZ, []),
    'wotk_transport_stablev2_noalloc': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _I64, _P, _D, _D, _D, _I64, _D,
                                                       _D, _I64, _I64, _P, _I64, _P, _P, _P,
                                                       ctypes.POINTER(Stats), _P]),
    'wotk_transport_duality_gap': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _I64, _P, _D, _D, _D, _I64,
                                                  _D, _D, _D, _I64, _P, _I64, _P, _P, _P, _P, _SZ,
                                                  ctypes.POINTER(Stats), _P]),
    'wotk_plan': (ctypes.c_int, [_P, _I64, _I64, _I64, _P, _P, _D, _D, _P, _I64, _P, _P]),
    'wotk_batch_workspace_size': (_SZ, [_P, _I32, _I32, _I64]),
    'wotk_batch_stablev2': (ctypes.c_int, [_P, _P, _I32, _P, _I32, _P, _P, _D, _I64, _D, _D, _I64, _I64, _P, _P,
                                           _I32, _I32, _P, _SZ, ctypes.POINTER(Stats), _P]),
    'wotk_batch_begin': (ctypes.c_int, [_P, _P, _I32, _P, _I32, _I32, _P, _P, _P, _I64, _D, _I64, _I64, _I32, _I32,
                                        _P, _SZ, ctypes.POINTER(ctypes.c_void_p), _P]),
    'wotk_batch_advance': (ctypes.c_int, [_P, _I64, ctypes.POINTER(ctypes.c_int64), _P]),
    'wotk_batch_potentials': (ctypes.c_int, [_P, _P, _P, _P, ctypes.POINTER(Stats), _P]),
    'wotk_batch_modes': (ctypes.c_int, [_P, _P, _I64]),
    'wotk_batch_end': (None, [_P]),
    'wotk_fused_workspace_size': (_SZ, [_I64, _I64, _I64]),
    'wotk_transport_stablev2_coords': (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _D, _D, _D, _I64, _D, _D, _I64,
                                                      _I64, _I32, _P, _P, _P, _P, _P, _SZ, ctypes.POINTER(Stats),
                                                      _P]),
    'wotk_duality_gap_coords_workspace_size': (_SZ, [_I64, _I64, _I64]),
    'wotk_transport_duality_gap_coords': (ctypes.c_int, [_P, _P, _I64, _I64, _I64, _P, _D, _D, _D, _I64, _D, _D, _D,
                                                         _I64, _I32, _P, _P, _P, _P, _P, _SZ, ctypes.POINTER(Stats),
                                                         _P]),
    'wotk_gemm_f32': (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _I64, _I64, _I64, _P]),
    'wotk_adjoint_workspace_size': (_SZ, [_I64, _I64, _I64]),
    'wotk_ot_adjoint': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _I64, _P, _P, _I64, _P, _D, _P, _P, _D, _D, _D,
                                       _P, _P, _P, _P, _P, _SZ, _P]),
    'wotk_primal_dual': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _I64, _P, _P, _P, _D, _D, _D,
                                        ctypes.POINTER(ctypes.c_double), _P, _SZ, _P]),
    'wotk_tmap_workspace_size': (_SZ, [ctypes.POINTER(TMap), _I32, _I32, _I32]),
    'wotk_tmap_propagate': (ctypes.c_int, [ctypes.POINTER(TMap), _I32, _I32, _P, _I64, _I32, _I32,
                                           ctypes.POINTER(_P), _I64, _P, _SZ, ctypes.POINTER(Stats), _P]),
    'wotk_rows_normalize_f64': (ctypes.c_int, [_P, _I64, _I32, _I64, _P]),
    'wotk_tmap_transpose': (ctypes.c_int, [_P, _I32, _I64, _I64, _I64, _P, _I64, _P]),
    'wotk_pca_center_workspace_size': (_SZ, [_I64, _I64]),
    'wotk_interp_weights': (ctypes.c_int, [_P, _I64, _I64, _I64, _D, _P, _P, _P]),
    'wotk_interp_sample': (ctypes.c_int, [_P, _I64, _P, _I64, _P, _P]),
    'wotk_interp_gather': (ctypes.c_int, [_P, _I64, _P, _I64, _I64, _P, _I64, _I64, _I64, _D, _P, _I64, _P]),
    'wotk_emd': (ctypes.c_int, [_P, _I64, _P, _I64, _P, _I64, _I64, ctypes.POINTER(ctypes.c_double), _P,
                                ctypes.POINTER(_I64), ctypes.POINTER(_I32)]),
    'wotk_pca_center': (ctypes.c_int, [_P, _I32, _I64, _I64, _I64, _P, _I64, _P, _P, _P, _SZ, _P]),
    'wotk_probe_sink_bytes': (_SZ, []),
    'wotk_probe_valu_rates': (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), _I32, _P, _SZ, _P]),
    'wotk_probe_issue_mix': (ctypes.c_int, [ctypes.POINTER(ctypes.c_double), _I32, _P, _SZ, _P]),
    'wotk_tile_variants': (_I32, [_I32, ctypes.POINTER(Tile), _I32]),
    'wotk_shape_bucket': (_I32, [_I32, _I64, _I64, _I32]),
    'wotk_num_buckets': (_I32, [_I32]),
    'wotk_set_tile': (ctypes.c_int, [_I32, _I32, ctypes.POINTER(Tile)]),
    'wotk_get_tile': (ctypes.c_int, [_I32, _I32, ctypes.POINTER(Tile)]),
}


class BatchPair(ctypes.Structure):
    _fields_ = [('c_offset', ctypes.c_int64), ('ct_offset', ctypes.c_int64), ('ldc', ctypes.c_int64),
                ('ldct', ctypes.c_int64), ('n_i', ctypes.c_int64), ('n_j', ctypes.c_int64),
                ('g_offset', ctypes.c_int64)]


class NativeError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libwotk.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libwotk.so not found at %s; build it with `python -m wot_amd.build`'
                              % LIB_PATH)
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    """Map a status code to the reference's exception types."""
    if rc == WOTK_OK:
        return
    msg = lib().wotk_last_error().decode(errors='replace')
    if rc == WOTK_EINVAL:
        raise ValueError(msg)
    if rc == WOTK_ENAN:
        raise RuntimeError(msg)
    raise NativeError('libwotk error %d: %s' % (rc, msg))
