"""Loading helpers for the golden fixtures (see make_golden.py)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def obs_from(d):
    """The observable dict stored as obs_z / obs_zz arrays (NaN for None)."""
    obs = {}
    for key in ('z', 'zz'):
        k = 'obs_' + key
        if k in d.files:
            a = d[k]
            o = np.empty(a.shape, dtype=object)
            for idx in np.ndindex(a.shape):
                o[idx] = None if np.isnan(a[idx]) else float(a[idx])
            obs[key] = o
    return obs


def maxcut_obs(n, edges):
    zz = np.full((n, n), None)
    for a, b in edges:
        zz[int(a), int(b)] = 1.0
    return {'zz': zz}
