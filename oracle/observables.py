"""Local-value (E_loc) estimators restated from the reference
(observables/monte_carlo/operator.py, observables/monte_carlo/observable.py)."""
import numpy as np


def flat_local_connections_log_values(wave_function, conn, use):
    """``get_flat_local_connections_log_values`` (operator.py:6-10): used
    connections in sample-major, connection-minor order."""
    nc, B = use.shape
    flat = np.moveaxis(conn, 1, 0).reshape((-1,) + conn.shape[2:])
    return wave_function(flat[use.astype(bool).T.reshape(-1)])


def local_values_unbalanced(wave_function, conn, mel, use):
    """operator.py:13-33: per sample exp(logpsi' - logpsi[0]) (complex64
    differences, :31) times the used matrix elements, summed."""
    B = use.shape[1]
    out = np.zeros(B, dtype=np.complex128)
    logv = flat_local_connections_log_values(wave_function, conn, use)
    counts = use.sum(axis=0).astype(np.int64)
    idx = 0
    for i in range(B):
        lv = logv[idx:idx + counts[i]]
        idx += counts[i]
        div = np.exp(lv - lv[0])
        out[i] = (mel[use[:, i].astype(bool), i] * div).sum()
    return out


def local_values_balanced(wave_function, conn, mel):
    """operator.py:35-42: every one of the Nc*B configurations evaluated."""
    nc, B = mel.shape
    flat = conn.reshape((-1,) + conn.shape[2:])
    logv = wave_function(flat).reshape(nc, B)
    return (mel * np.exp(logv - logv[0])).sum(axis=0)


def local_values(wave_function, conn, mel, use):
    """operator.py:44-54: path switch on use.mean() < 0.95."""
    if use.mean() < 0.95:
        return local_values_unbalanced(wave_function, conn, mel, use)
    return local_values_balanced(wave_function, conn, mel)


def estimate(local_vals):
    """observable.py:10-14: (mean, population variance of the real part,
    local values)."""
    return np.mean(local_vals), np.var(np.real(local_vals)), local_vals
