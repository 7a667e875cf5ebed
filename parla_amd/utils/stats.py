"""Log-linear / log-log fits for convergence checks (parla/utils/stats.py:6-56).

Host-side statistics on short error logs; not part of any GPU computation.
"""
import warnings

import numpy as np
import scipy.linalg as la


def loglinear_fit(x, y):
    """Fit log(y) = a + b x; return ([a, b], R^2) (stats.py:6-28)."""
    x = np.asarray(x, dtype=float).ravel()
    y = np.asarray(y, dtype=float).ravel()
    assert x.size == y.size
    if np.any(y <= 0):
        warnings.warn('Dropping samples "i" where y[i] <= 0.')
        x, y = x[y > 0], y[y > 0]
    ly = np.log(y)
    X = np.column_stack([np.ones(x.size), x])
    fit = la.lstsq(X, ly)[0]
    ss_tot = np.sum((ly - ly.mean()) ** 2)
    return fit, 1 - np.sum((ly - X @ fit) ** 2) / ss_tot


def loglog_fit(x, y):
    """Fit log(y) = a + b log(x) (stats.py:31-56)."""
    x = np.asarray(x, dtype=float).ravel()
    if np.any(x <= 0):
        raise ValueError('Input x must be positive.')
    return loglinear_fit(np.log(x), y)
