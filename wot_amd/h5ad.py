"""h5ad transport maps without h5py / anndata: a ctypes binding of the HDF5 C
library writing and reading the AnnData on-disk layout that wot's transport
maps use (wot/io/io.py:439-452 ``ds.write(path)``; read back by
wot/io/io.py:403-404 ``anndata.read(path)`` and, for the cell ids only, by
wot/tmap/transport_map_model.py:700-712).

Layout written (AnnData's h5ad encoding 0.1.0, readable by anndata >= 0.7):

    /                attrs encoding-type='anndata', encoding-version='0.1.0'
    /X               dense float64 (or float32) dataset, cells x cells
    /obs, /var       groups, attrs encoding-type='dataframe', encoding-version='0.1.0',
                     _index='_index', column-order=[column names]
                     /_index          variable-length UTF-8 strings (cell ids)
                     /<column>        one dataset per column (float, int, bool, strings)
    /uns /obsm /varm /obsp /varp /layers   empty groups (encoding-type='dict')

The library is ``$WOTK_HDF5_LIB`` or the first libhdf5 found on the system.
"""
import ctypes
import ctypes.util
import os
import threading

import numpy as np
import pandas as pd

_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_ulonglong
_P = ctypes.c_void_p
_S = ctypes.c_char_p

H5P_DEFAULT = 0
H5S_ALL = 0
H5S_SCALAR = 0
H5F_ACC_RDONLY = 0x0000
H5F_ACC_TRUNC = 0x0002
H5T_VARIABLE = ctypes.c_size_t(-1).value
H5T_CSET_UTF8 = 1
H5T_STR_NULLTERM = 0
H5T_INTEGER, H5T_FLOAT, H5T_STRING, H5T_ENUM = 0, 1, 3, 8
H5T_SGN_NONE = 0
H5I_GROUP, H5I_DATASET = 2, 5

_SIGS = {
    'H5open': (_herr, []),
    'H5get_libversion': (_herr, [ctypes.POINTER(ctypes.c_uint)] * 3),
    'H5Fcreate': (_hid, [_S, ctypes.c_uint, _hid, _hid]),
    'H5Fopen': (_hid, [_S, ctypes.c_uint, _hid]),
    'H5Fclose': (_herr, [_hid]),
    'H5Gcreate2': (_hid, [_hid, _S, _hid, _hid, _hid]),
    'H5Gopen2': (_hid, [_hid, _S, _hid]),
    'H5Gclose': (_herr, [_hid]),
    'H5Screate': (_hid, [ctypes.c_int]),
    'H5Screate_simple': (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    'H5Sget_simple_extent_ndims': (ctypes.c_int, [_hid]),
    'H5Sget_simple_extent_dims': (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
    'H5Sclose': (_herr, [_hid]),
    'H5Tcopy': (_hid, [_hid]),
    'H5Tset_size': (_herr, [_hid, ctypes.c_size_t]),
    'H5Tset_cset': (_herr, [_hid, ctypes.c_int]),
    'H5Tset_strpad': (_herr, [_hid, ctypes.c_int]),
    'H5Tget_class': (ctypes.c_int, [_hid]),
    'H5Tget_size': (ctypes.c_size_t, [_hid]),
    'H5Tget_sign': (ctypes.c_int, [_hid]),
    'H5Tis_variable_str': (ctypes.c_int, [_hid]),
    'H5Tclose': (_herr, [_hid]),
    'H5Tenum_create': (_hid, [_hid]),
    'H5Tenum_insert': (_herr, [_hid, _S, _P]),
    'H5Dcreate2': (_hid, [_hid, _S, _hid, _hid, _hid, _hid, _hid]),
    'H5Dopen2': (_hid, [_hid, _S, _hid]),
    'H5Dwrite': (_herr, [_hid, _hid, _hid, _hid, _hid, _P]),
    'H5Dread': (_herr, [_hid, _hid, _hid, _hid, _hid, _P]),
    'H5Dget_type': (_hid, [_hid]),
    'H5Dget_space': (_hid, [_hid]),
    'H5Dvlen_reclaim': (_herr, [_hid, _hid, _hid, _P]),
    'H5Dclose': (_herr, [_hid]),
    'H5Acreate2': (_hid, [_hid, _S, _hid, _hid, _hid, _hid]),
    'H5Aopen': (_hid, [_hid, _S, _hid]),
    'H5Aexists': (ctypes.c_int, [_hid, _S]),
    'H5Awrite': (_herr, [_hid, _hid, _P]),
    'H5Aread': (_herr, [_hid, _hid, _P]),
    'H5Aget_type': (_hid, [_hid]),
    'H5Aget_space': (_hid, [_hid]),
    'H5Aclose': (_herr, [_hid]),
    'H5Lexists': (ctypes.c_int, [_hid, _S, _hid]),
    'H5Oopen': (_hid, [_hid, _S, _hid]),
    'H5Oclose': (_herr, [_hid]),
    'H5Iget_type': (ctypes.c_int, [_hid]),
}

_lock = threading.Lock()
_lib = None
_T = {}


def _candidates():
    env = os.environ.get('WOTK_HDF5_LIB')
    if env:
        yield env
    found = ctypes.util.find_library('hdf5')
    if found:
        yield found
    for d in ('/opt/conda/lib', '/usr/lib/x86_64-linux-gnu/hdf5/serial', '/usr/lib/x86_64-linux-gnu',
              '/usr/local/lib', '/usr/lib64'):
        yield os.path.join(d, 'libhdf5.so')


def lib():
    """The HDF5 C library (loaded once); ImportError if none is installed."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        err = []
        for path in _candidates():
            try:
                h = ctypes.CDLL(path)
            except OSError as e:
                err.append(str(e))
                continue
            for name, (res, args) in _SIGS.items():
                fn = getattr(h, name)
                fn.restype, fn.argtypes = res, args
            if h.H5open() < 0:
                continue
            for key, sym in (('f64', 'H5T_NATIVE_DOUBLE_g'), ('f32', 'H5T_NATIVE_FLOAT_g'),
                             ('i64', 'H5T_NATIVE_LLONG_g'), ('i32', 'H5T_NATIVE_INT_g'),
                             ('i8', 'H5T_NATIVE_SCHAR_g'), ('c_s1', 'H5T_C_S1_g')):
                _T[key] = _hid.in_dll(h, sym).value
            _lib = h
            return h
        raise ImportError('h5ad needs the HDF5 C library (libhdf5.so); set WOTK_HDF5_LIB. Tried: %s' % err)


def version():
    L = lib()
    v = [ctypes.c_uint() for _ in range(3)]
    L.H5get_libversion(*[ctypes.byref(x) for x in v])
    return tuple(x.value for x in v)


def _ok(rc, what):
    if rc < 0:
        raise OSError('HDF5 call failed: %s' % what)
    return rc


# ---------------------------------------------------------------------------
# writing
# ---------------------------------------------------------------------------
def _str_type():
    L = lib()
    t = _ok(L.H5Tcopy(_T['c_s1']), 'H5Tcopy')
    _ok(L.H5Tset_size(t, H5T_VARIABLE), 'H5Tset_size')
    _ok(L.H5Tset_cset(t, H5T_CSET_UTF8), 'H5Tset_cset')
    _ok(L.H5Tset_strpad(t, H5T_STR_NULLTERM), 'H5Tset_strpad')
    return t


def _space(shape):
    L = lib()
    if shape == ():
        return _ok(L.H5Screate(H5S_SCALAR), 'H5Screate')
    dims = (_hsize * len(shape))(*shape)
    return _ok(L.H5Screate_simple(len(shape), dims, None), 'H5Screate_simple')


def _str_buffer(values):
    enc = [str(v).encode('utf-8') for v in values]
    buf = (ctypes.c_char_p * max(len(enc), 1))(*enc)
    return buf, enc  # keep enc alive


def _write_attr(obj, name, value):
    L = lib()
    if isinstance(value, str):
        t, sp = _str_type(), _space(())
        buf, keep = _str_buffer([value])
    elif isinstance(value, (list, tuple, np.ndarray)) and (len(value) == 0 and isinstance(value, np.ndarray)
                                                            and value.dtype.kind == 'f'):
        t, sp = L.H5Tcopy(_T['f64']), _space((0,))
        buf, keep = (ctypes.c_double * 1)(), None
    elif isinstance(value, (list, tuple, np.ndarray)):
        vals = list(value)
        t, sp = _str_type(), _space((len(vals),))
        buf, keep = _str_buffer(vals)
    else:
        raise TypeError('unsupported attribute %r' % (value,))
    a = _ok(L.H5Acreate2(obj, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT), 'H5Acreate2 ' + name)
    try:
        _ok(L.H5Awrite(a, t, ctypes.cast(buf, _P)), 'H5Awrite ' + name)
    finally:
        L.H5Aclose(a)
        L.H5Sclose(sp)
        L.H5Tclose(t)
    del keep


def _write_array(loc, name, arr):
    L = lib()
    arr = np.asarray(arr)
    if arr.dtype.kind in 'OUS':
        t, sp = _str_type(), _space(arr.shape)
        buf, keep = _str_buffer(arr.ravel())
        mem = t
    else:
        if arr.dtype == np.bool_:  # h5py's encoding: an int8 enum FALSE = 0, TRUE = 1
            arr = np.ascontiguousarray(arr.astype(np.int8))
            t = _ok(L.H5Tenum_create(_T['i8']), 'H5Tenum_create')
            for nm, v in ((b'FALSE', 0), (b'TRUE', 1)):
                c = ctypes.c_int8(v)
                _ok(L.H5Tenum_insert(t, nm, ctypes.byref(c)), 'H5Tenum_insert')
            mem, sp, keep = t, _space(arr.shape), arr
            buf = arr.ctypes.data_as(_P)
            key = None
        elif arr.dtype.kind == 'f':
            arr = arr.astype(np.float32 if arr.dtype == np.float32 else np.float64, copy=False)
            key = 'f32' if arr.dtype == np.float32 else 'f64'
        elif arr.dtype.kind in 'iu':
            wide = arr.dtype.itemsize > 4 or arr.dtype.kind == 'u'
            arr = arr.astype(np.int64 if wide else np.int32, copy=False)
            key = 'i64' if arr.dtype == np.int64 else 'i32'
        else:
            raise TypeError('unsupported column dtype %s' % arr.dtype)
        if key is not None:
            arr = np.ascontiguousarray(arr)
            t = L.H5Tcopy(_T[key])
            mem, sp, keep = t, _space(arr.shape), arr
            buf = arr.ctypes.data_as(_P)
    d = _ok(L.H5Dcreate2(loc, name.encode(), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), 'H5Dcreate2 ' + name)
    try:
        if arr.size:
            _ok(L.H5Dwrite(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, ctypes.cast(buf, _P)), 'H5Dwrite ' + name)
    finally:
        L.H5Dclose(d)
        L.H5Sclose(sp)
        L.H5Tclose(t)
    del keep


def _write_dataframe(f, key, df):
    L = lib()
    g = _ok(L.H5Gcreate2(f, key.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), 'H5Gcreate2 ' + key)
    try:
        cols = [str(c) for c in df.columns]
        if '_index' in cols:
            raise ValueError("a column may not be named '_index'")
        _write_attr(g, '_index', '_index')
        _write_attr(g, 'column-order', cols if cols else np.zeros(0))
        _write_attr(g, 'encoding-type', 'dataframe')
        _write_attr(g, 'encoding-version', '0.1.0')
        _write_array(g, '_index', np.asarray([str(v) for v in df.index], dtype=object))
        for c in df.columns:
            v = df[c]
            vals = v.astype(str).values.astype(object) if v.dtype.kind in 'OUSc' or str(v.dtype) == 'category' \
                else v.values
            _write_array(g, str(c), vals)
    finally:
        L.H5Gclose(g)


def write_h5ad(path, X, obs, var):
    """Write a transport map (or any dense AnnData) as h5ad."""
    L = lib()
    X = np.asarray(X)
    if X.ndim != 2 or len(obs) != X.shape[0] or len(var) != X.shape[1]:
        raise ValueError('X %s does not match obs (%d) / var (%d)' % (X.shape, len(obs), len(var)))
    f = _ok(L.H5Fcreate(str(path).encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), 'H5Fcreate ' + str(path))
    try:
        _write_attr(f, 'encoding-type', 'anndata')
        _write_attr(f, 'encoding-version', '0.1.0')
        _write_array(f, 'X', X if X.dtype in (np.float32, np.float64) else X.astype(np.float64))
        _write_dataframe(f, 'obs', obs)
        _write_dataframe(f, 'var', var)
        for k in ('uns', 'obsm', 'varm', 'obsp', 'varp', 'layers'):
            g = _ok(L.H5Gcreate2(f, k.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), 'H5Gcreate2 ' + k)
            _write_attr(g, 'encoding-type', 'dict')
            _write_attr(g, 'encoding-version', '0.1.0')
            L.H5Gclose(g)
    finally:
        L.H5Fclose(f)
    return str(path)


# ---------------------------------------------------------------------------
# reading
# ---------------------------------------------------------------------------
def _read_strings(L, obj, t, sp, reader):
    n = _n_elements(L, sp)
    mem = _str_type()
    try:
        if L.H5Tis_variable_str(t) > 0:
            buf = (ctypes.c_char_p * max(n, 1))()
            _ok(reader(obj, mem, ctypes.cast(buf, _P)), 'read strings')
            out = [buf[i].decode('utf-8') if buf[i] is not None else '' for i in range(n)]
            L.H5Dvlen_reclaim(mem, sp, H5P_DEFAULT, ctypes.cast(buf, _P))
            return out
        size = L.H5Tget_size(t)
        raw = ctypes.create_string_buffer(size * max(n, 1))
        _ok(reader(obj, t, ctypes.cast(raw, _P)), 'read fixed strings')
        return [raw.raw[i * size:(i + 1) * size].split(b'\0', 1)[0].decode('utf-8') for i in range(n)]
    finally:
        L.H5Tclose(mem)


def _n_elements(L, sp):
    nd = L.H5Sget_simple_extent_ndims(sp)
    if nd <= 0:
        return 1
    dims = (_hsize * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    return int(np.prod([dims[i] for i in range(nd)]))


def _shape(L, sp):
    nd = L.H5Sget_simple_extent_ndims(sp)
    if nd <= 0:
        return ()
    dims = (_hsize * nd)()
    L.H5Sget_simple_extent_dims(sp, dims, None)
    return tuple(int(dims[i]) for i in range(nd))


def _dread(d, mem, buf):
    return lib().H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf)


def _aread(a, mem, buf):
    return lib().H5Aread(a, mem, buf)


def _numeric(L, t):
    cls = L.H5Tget_class(t)
    size = L.H5Tget_size(t)
    if cls == H5T_FLOAT:
        return ('f32', np.float32) if size == 4 else ('f64', np.float64)
    if cls in (H5T_INTEGER, H5T_ENUM):
        if size == 1:
            return 'i8', np.int8
        return ('i32', np.int32) if size <= 4 else ('i64', np.int64)
    return None


def _read_dataset(L, loc, name):
    d = _ok(L.H5Dopen2(loc, name.encode(), H5P_DEFAULT), 'H5Dopen2 ' + name)
    t = L.H5Dget_type(d)
    sp = L.H5Dget_space(d)
    try:
        shape = _shape(L, sp)
        if L.H5Tget_class(t) == H5T_STRING:
            return np.asarray(_read_strings(L, d, t, sp, _dread), dtype=object).reshape(shape)
        num = _numeric(L, t)
        if num is None:
            raise TypeError('unsupported HDF5 type in %s' % name)
        key, dt = num
        out = np.empty(shape, dtype=dt)
        if out.size:
            _ok(L.H5Dread(d, _T[key], H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data_as(_P)), 'H5Dread ' + name)
        if L.H5Tget_class(t) == H5T_ENUM and size_is_bool(L, t):
            return out.astype(bool)
        return out
    finally:
        L.H5Sclose(sp)
        L.H5Tclose(t)
        L.H5Dclose(d)


def size_is_bool(L, t):
    return L.H5Tget_size(t) == 1


def _read_attr(L, obj, name):
    if L.H5Aexists(obj, name.encode()) <= 0:
        return None
    a = _ok(L.H5Aopen(obj, name.encode(), H5P_DEFAULT), 'H5Aopen ' + name)
    t = L.H5Aget_type(a)
    sp = L.H5Aget_space(a)
    try:
        shape = _shape(L, sp)
        if L.H5Tget_class(t) == H5T_STRING:
            vals = _read_strings(L, a, t, sp, _aread)
            return vals[0] if shape == () else vals
        num = _numeric(L, t)
        if num is None:
            return None
        key, dt = num
        out = np.empty(shape, dtype=dt)
        if out.size:
            _ok(L.H5Aread(a, _T[key], out.ctypes.data_as(_P)), 'H5Aread ' + name)
        return out.tolist() if shape else out.item()
    finally:
        L.H5Sclose(sp)
        L.H5Tclose(t)
        L.H5Aclose(a)


def _exists(L, loc, name):
    return L.H5Lexists(loc, name.encode(), H5P_DEFAULT) > 0


def _read_index(L, f, key):
    g = _ok(L.H5Gopen2(f, key.encode(), H5P_DEFAULT), 'H5Gopen2 ' + key)
    try:
        idx_key = _read_attr(L, g, '_index') or 'index'
        return g, idx_key, pd.Index([str(v) for v in _read_dataset(L, g, idx_key)])
    except Exception:
        L.H5Gclose(g)
        raise


def read_index(path, key='obs'):
    """Cell ids of /obs or /var (the '_index' attribute names the dataset,
    'index' when absent), without reading X (transport_map_model.py:705-710)."""
    L = lib()
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), 'H5Fopen ' + str(path))
    try:
        g, _, idx = _read_index(L, f, key)
        L.H5Gclose(g)
        return idx
    finally:
        L.H5Fclose(f)


def _read_dataframe(L, f, key):
    g, idx_key, index = _read_index(L, f, key)
    try:
        order = _read_attr(L, g, 'column-order')
        cols = [c for c in (order or []) if isinstance(c, str)]
        data = {}
        for c in cols:
            if _exists(L, g, c):
                obj = L.H5Oopen(g, c.encode(), H5P_DEFAULT)
                kind = L.H5Iget_type(obj)
                L.H5Oclose(obj)
                if kind == H5I_DATASET:
                    data[c] = _read_dataset(L, g, c)
        return pd.DataFrame(data, index=index, columns=cols if data else None)
    finally:
        L.H5Gclose(g)


def _read_sparse(L, f):
    """X stored as a CSR / CSC group (data, indices, indptr, attr shape)."""
    import scipy.sparse
    g = _ok(L.H5Gopen2(f, b'X', H5P_DEFAULT), 'H5Gopen2 X')
    try:
        fmt = _read_attr(L, g, 'encoding-type') or _read_attr(L, g, 'h5sparse_format') or 'csr_matrix'
        shape = _read_attr(L, g, 'shape') or _read_attr(L, g, 'h5sparse_shape')
        data = _read_dataset(L, g, 'data')
        indices = _read_dataset(L, g, 'indices')
        indptr = _read_dataset(L, g, 'indptr')
    finally:
        L.H5Gclose(g)
    cls = scipy.sparse.csc_matrix if str(fmt).startswith('csc') else scipy.sparse.csr_matrix
    return cls((data, indices, indptr), shape=tuple(int(s) for s in shape)).toarray()


def read_h5ad(path):
    """Dense X (float64 as stored, sparse groups densified), obs and var."""
    from . import _anndata
    L = lib()
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), 'H5Fopen ' + str(path))
    try:
        obj = _ok(L.H5Oopen(f, b'X', H5P_DEFAULT), 'H5Oopen X')
        kind = L.H5Iget_type(obj)
        L.H5Oclose(obj)
        X = _read_dataset(L, f, 'X') if kind == H5I_DATASET else _read_sparse(L, f)
        obs = _read_dataframe(L, f, 'obs')
        var = _read_dataframe(L, f, 'var')
    finally:
        L.H5Fclose(f)
    return _anndata.AnnData(X, obs, var)


def read_attrs(path, key=''):
    """Attributes of the root ('') or of a group, as a dict (format checks)."""
    L = lib()
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), 'H5Fopen ' + str(path))
    try:
        g = f if not key else _ok(L.H5Gopen2(f, key.encode(), H5P_DEFAULT), 'H5Gopen2 ' + key)
        names = ('encoding-type', 'encoding-version', '_index', 'column-order', 'shape')
        out = {n: _read_attr(L, g, n) for n in names}
        if key:
            L.H5Gclose(g)
        return {k: v for k, v in out.items() if v is not None}
    finally:
        L.H5Fclose(f)


def has_key(path, key):
    L = lib()
    f = _ok(L.H5Fopen(str(path).encode(), H5F_ACC_RDONLY, H5P_DEFAULT), 'H5Fopen ' + str(path))
    try:
        return _exists(L, f, key)
    finally:
        L.H5Fclose(f)
