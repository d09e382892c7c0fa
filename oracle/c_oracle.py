"""Build and bind the C restatement in oracle/c/oracle_ref.c.

TEST INFRASTRUCTURE ONLY: used by tests/ as a checker and by bench.py as the
timed single-core CPU baseline.  Built with gcc by ``build()`` (called from
``__graft_entry__.build()``) twice: in double precision (liboracle_ref.so) and
in x87 extended precision (liboracle_ref_ld.so, ``extended=True``), the
arbiter that pins FP64 results at depth.  The shared objects are git-ignored
and travel with the working tree.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'c', 'oracle_ref.c')
LIB = os.path.join(HERE, 'c', 'liboracle_ref.so')
LIB_LD = os.path.join(HERE, 'c', 'liboracle_ref_ld.so')
_libs = {}


def build():
    for out, extra in ((LIB, []), (LIB_LD, ['-DORACLE_REAL=long double'])):
        # a temporary name per process: parallel test workers build at once
        tmp = '%s.tmp%d' % (out, os.getpid())
        cmd = ['gcc', '-O3', '-march=x86-64-v2', '-fPIC', '-shared', '-std=gnu11'] + extra + [SRC, '-o', tmp, '-lm']
        subprocess.run(cmd, check=True, capture_output=True)
        os.replace(tmp, out)
    return LIB


def load(extended=False):
    path = LIB_LD if extended else LIB
    if path not in _libs:
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        ip, dp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)
        lib.oracle_mcclean_grad_run.argtypes = [ctypes.c_int] * 3 + [ip, dp, ctypes.c_int, ip, ip, ip, dp, dp, dp]
        lib.oracle_qaoa_grad_run.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.c_int, ip, ip, ip, dp, dp, dp]
        _libs[path] = lib
    return _libs[path]


def _terms(obs):
    from .statevector import observable_terms
    _, terms = observable_terms(obs)
    km = {'x': 0, 'y': 1, 'z': 2, 'zz': 3}
    kind = np.array([km[t[0]] for t in terms], dtype=np.int32)
    ii = np.array([t[1] for t in terms], dtype=np.int32)
    jj = np.array([t[2] for t in terms], dtype=np.int32)
    w = np.array([t[3] for t in terms], dtype=np.float64)
    return len(terms), kind, ii, jj, w


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def mcclean_grad_run(n, axes, angles, obs, layers=None, extended=False):
    """(E, grad) of the reference algorithm; ``layers`` < L runs a bounded sample;
    ``extended``: in x87 extended precision."""
    lib = load(extended)
    L = axes.shape[0]
    Ls = L if layers is None else int(layers)
    ax = np.ascontiguousarray(axes, dtype=np.int32)
    an = np.ascontiguousarray(angles, dtype=np.float64)
    nt, kind, ii, jj, w = _terms(obs)
    E = np.zeros(1)
    grad = np.zeros((Ls, n))
    ci, cd = ctypes.c_int, ctypes.c_double
    rc = lib.oracle_mcclean_grad_run(n, L, Ls, _p(ax, ci), _p(an, cd), nt, _p(kind, ci), _p(ii, ci), _p(jj, ci),
                                     _p(w, cd), _p(E, cd), _p(grad, cd))
    if rc != 0:
        raise MemoryError('oracle_mcclean_grad_run: allocation failed')
    return float(E[0]), grad


def qaoa_grad_run(n, obs, betas, gammas, extended=False):
    lib = load(extended)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    nt, kind, ii, jj, w = _terms(obs)
    E = np.zeros(1)
    grad = np.zeros((b.size, 2))
    ci, cd = ctypes.c_int, ctypes.c_double
    rc = lib.oracle_qaoa_grad_run(n, b.size, _p(b, cd), _p(g, cd), nt, _p(kind, ci), _p(ii, ci), _p(jj, ci),
                                  _p(w, cd), _p(E, cd), _p(grad, cd))
    if rc != 0:
        raise MemoryError('oracle_qaoa_grad_run: allocation failed')
    return float(E[0]), grad
