"""ctypes binding of libqradient_amd.so (the C ABI in include/qradient_amd.h).

The library is required: importing the circuit classes works without it, but
the first GPU call raises ``RuntimeError`` if ``libqradient_amd.so`` was not
built (``python -m qradient_amd._build``).  There is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libqradient_amd.so')

QG_OK, QG_ERR_INVALID, QG_ERR_UNSUPPORTED, QG_ERR_HIP, QG_ERR_NOMEM, QG_ERR_PENDING = 0, 1, 2, 3, 4, 5
QG_SERVER_SIMULATE = 1
QG_SERVER_SHORTEST_FIRST = 2
QG_C128, QG_C64 = 0, 1
QG_ENT_CNOT, QG_ENT_CZ = 0, 1
QG_HISTORY = -1
TERM_X, TERM_Y, TERM_Z, TERM_ZZ = 0, 1, 2, 3

DTYPES = {'complex128': QG_C128, 'c128': QG_C128, np.complex128: QG_C128,
          'complex64': QG_C64, 'c64': QG_C64, np.complex64: QG_C64}
ENTANGLERS = {'cnot': QG_ENT_CNOT, 'cz': QG_ENT_CZ}


class QgTerm(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int32), ('i', ctypes.c_int32), ('j', ctypes.c_int32),
                ('pad', ctypes.c_int32), ('w', ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [('queue_ms', ctypes.c_double), ('run_ms', ctypes.c_double), ('latency_ms', ctypes.c_double),
                ('device_start_ms', ctypes.c_double), ('device_end_ms', ctypes.c_double),
                ('slot', ctypes.c_int32), ('ckpt_every', ctypes.c_int32), ('footprint', ctypes.c_uint64)]


class ServerCounters(ctypes.Structure):
    _fields_ = [('queued', ctypes.c_int64), ('running', ctypes.c_int64), ('completed', ctypes.c_int64),
                ('arena_bytes', ctypes.c_uint64), ('arena_in_use', ctypes.c_uint64), ('arena_peak', ctypes.c_uint64)]


_vp = ctypes.c_void_p
_pp = ctypes.POINTER(ctypes.c_void_p)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_tp = ctypes.POINTER(QgTerm)
_i = ctypes.c_int
_d = ctypes.c_double

# name -> argtypes; every function returns int status unless listed in _RESTYPES
SIGNATURES = {
    'qg_last_error': [],
    'qg_version': [],
    'qg_ctx_create': [_i, _vp, _pp],
    'qg_ctx_destroy': [_vp],
    'qg_ctx_synchronize': [_vp],
    'qg_stream_copy': [_vp, _vp, _vp, ctypes.c_size_t, _i, _i, _i, _dp],
    'qg_mcclean_plan': [_vp, _i, _i, _i, _i, _i, _pp],
    'qg_mcclean_grad_run': [_vp, _ip, _dp, _tp, _i, _dp, _dp, _dp],
    'qg_mcclean_workspace_bytes': [_vp, _i, _i, _i, _i, ctypes.POINTER(ctypes.c_size_t), _ip],
    'qg_mcclean_expec_val': [_vp, _ip, _dp, _tp, _i, _dp, _dp],
    'qg_mcclean_grad_run_noalloc': [_vp, _i, _i, _i, _ip, _dp, _tp, _i, _dp, _dp],
    'qg_mcclean_shift_term_prob': [_vp, _ip, _dp, _dp, _i, _ip, _dp, _tp, _i, _dp],
    'qg_mcclean_shift_sample': [_vp, _ip, _dp, _dp, _i, _ip, _dp, _ip, _i, _dp, _ip],
    'qg_state_sample_basis': [_vp, _ip, _i, _dp, _ip],
    'qg_state_copy': [_vp, _vp],
    'qg_mcclean_batch_plan': [_vp, _i, _i, _i, _i, _i, _pp],
    'qg_mcclean_batch_grad_run': [_vp, _i, _ip, _dp, _tp, _i, _dp, _dp],
    'qg_qaoa_plan': [_vp, _i, _i, _i, _i, _pp],
    'qg_qaoa_tune': [_vp, _i, _i, _i, _i],
    'qg_qaoa_grad_run': [_vp, _i, _tp, _ip, _dp, _dp, _dp, _dp],
    'qg_qaoa_expec_val': [_vp, _i, _tp, _ip, _dp, _dp, _dp],
    'qg_qaoa_grad_run_ini': [_vp, _i, _tp, _ip, _dp, _dp, _dp, _dp, _dp],
    'qg_qaoa_expec_val_ini': [_vp, _i, _tp, _ip, _dp, _dp, _dp, _dp],
    'qg_plan_reshape': [_vp, _i, _i, _i],
    'qg_plan_workspace_size': [_vp, ctypes.POINTER(ctypes.c_size_t)],
    'qg_plan_set_workspace': [_vp, _vp, ctypes.c_size_t],
    'qg_plan_get_state': [_vp, _dp],
    'qg_plan_last_ms': [_vp, _dp],
    'qg_plan_set_profiling': [_vp, _i],
    'qg_plan_pass_times': [_vp, _i, ctypes.POINTER(ctypes.c_float), _ip, _ip],
    'qg_plan_tune': [_vp, _i, _i, _i, _i],
    'qg_plan_set_fused': [_vp, _i],
    'qg_plan_ckpt_every': [_vp, _ip],
    'qg_plan_set_lds64': [_vp, _i],
    'qg_plan_set_graph': [_vp, _i],
    'qg_plan_lds64': [_vp, _ip],
    'qg_plan_destroy': [_vp],
    'qg_state_create': [_vp, _i, _i, ctypes.c_char, _pp],
    'qg_state_destroy': [_vp],
    'qg_state_reset': [_vp, ctypes.c_char],
    'qg_state_set': [_vp, _dp],
    'qg_state_get': [_vp, _dp],
    'qg_state_rot': [_vp, _i, _d, _i, _i],
    'qg_state_rot_all': [_vp, _i, _d],
    'qg_state_cnot_ladder': [_vp, _i],
    'qg_state_cz_ladder': [_vp],
    'qg_state_cnot': [_vp, _i, _i],
    'qg_state_exp_diag': [_vp, _tp, _i, _d],
    'qg_state_multiply_observable': [_vp, _tp, _i],
    'qg_state_expec_val': [_vp, _tp, _i, _dp],
    'qg_state_norm_error': [_vp, _dp],
    'qg_state_term_prob': [_vp, _tp, _i, _dp],
    'qg_plan_term_prob': [_vp, _tp, _i, _dp],
    'qg_plan_expec_val': [_vp, _tp, _i, _dp],
    'qg_state_scale': [_vp, _d, _d],
    'qg_server_create': [_i, _i, _i, ctypes.c_size_t, ctypes.c_size_t, _i, _i, _i, _pp],
    'qg_server_footprint': [_vp, _i, _i, ctypes.POINTER(ctypes.c_size_t), _ip],
    'qg_submit': [_vp, _i, _i, _ip, _dp, _tp, _i, ctypes.POINTER(ctypes.c_int64)],
    'qg_poll': [_vp, ctypes.c_int64, _i, _dp, _dp, ctypes.POINTER(RequestInfo)],
    'qg_wait_any': [_vp, _i, ctypes.POINTER(ctypes.c_int64)],
    'qg_server_stats': [_vp, ctypes.POINTER(ServerCounters)],
    'qg_server_destroy': [_vp],
}
_RESTYPES = {'qg_last_error': ctypes.c_char_p}

_lib = None
_lock = threading.Lock()


def load():
    """Load the library (once) and declare every signature; raises if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('libqradient_amd.so is not built: run `python -m qradient_amd._build` '
                               '(there is no CPU fallback)')
        lib = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = lib
        return lib


def check(status):
    """Map a QG_* status to the reference's exception types."""
    if status == QG_OK:
        return
    msg = load().qg_last_error().decode(errors='replace')
    if status == QG_ERR_INVALID:
        raise ValueError(msg)
    if status == QG_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == QG_ERR_NOMEM:
        raise MemoryError(msg)
    raise RuntimeError(msg)


def call(name, *args):
    check(getattr(load(), name)(*args))


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


# numpy mirror of struct qg_term (24 bytes: 4 x int32, float64)
TERM_DTYPE = np.dtype([('kind', '<i4'), ('i', '<i4'), ('j', '<i4'), ('pad', '<i4'), ('w', '<f8')])
assert TERM_DTYPE.itemsize == ctypes.sizeof(QgTerm)


def terms_array(terms):
    """[(kind, i, j, w)] -> (ctypes array, count)."""
    arr = (QgTerm * max(1, len(terms)))()
    if terms:
        a = np.frombuffer(arr, dtype=TERM_DTYPE)
        t = np.asarray([(int(k), int(i), int(j), 0, float(w)) for (k, i, j, w) in terms], dtype=TERM_DTYPE)
        a[:len(terms)] = t
    return arr, len(terms)


class Context:
    """One HIP device + stream; shared by the plans and states made from it."""

    _default = {}

    def __init__(self, device=0, stream=None):
        h = ctypes.c_void_p()
        call('qg_ctx_create', int(device), stream, ctypes.byref(h))
        self.handle = h
        self.device = device

    @classmethod
    def default(cls, device=0):
        if device not in cls._default:
            cls._default[device] = cls(device)
        return cls._default[device]

    def synchronize(self):
        call('qg_ctx_synchronize', self.handle)

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h and _lib is not None:
            _lib.qg_ctx_destroy(h)
            self.handle = None
