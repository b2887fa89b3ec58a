"""Exact enumeration helpers (exact/utils.py:39-58) and exact diagonalisation
of the find_conn matrix elements."""
import numpy as np
import scipy.sparse
import scipy.sparse.linalg


def binary_array_to_decimal_array(digits):
    """exact/utils.py:39-46: little-endian, +1 -> bit 1."""
    digits = np.asarray(digits)
    out = np.zeros(digits.shape[:-1], dtype=np.int64)
    power = 1
    for i in range(digits.shape[-1]):
        out += (digits[..., i] == 1) * power
        power *= 2
    return out


def decimal_array_to_binary_array(decimal, num_of_bits):
    """exact/utils.py:49-58 with zero_one_base=False."""
    decimal = np.asarray(decimal, dtype=np.int64).copy()
    out = np.zeros((decimal.shape[0], num_of_bits))
    for i in range(num_of_bits):
        out[:, i] = 2 * (decimal % 2) - 1
        decimal //= 2
    return out


def all_states(shape):
    n = int(np.prod(shape))
    return decimal_array_to_binary_array(np.arange(2 ** n), n).reshape((2 ** n,) + tuple(shape))


def sparse_hamiltonian(find_conn, shape, chunk=4096):
    """H[s, s'] summed over used connections from ``find_conn``."""
    n = int(np.prod(shape))
    N = 2 ** n
    rows, cols, vals = [], [], []
    for start in range(0, N, chunk):
        idx = np.arange(start, min(N, start + chunk))
        st = decimal_array_to_binary_array(idx, n).reshape((len(idx),) + tuple(shape))
        conn, mel, use = find_conn(st)
        nc = conn.shape[0]
        tgt = binary_array_to_decimal_array(conn.reshape(nc, len(idx), n))
        for c in range(nc):
            m = use[c] & (mel[c] != 0)
            rows.append(idx[m])
            cols.append(tgt[c][m])
            vals.append(mel[c][m])
    rows, cols, vals = map(np.concatenate, (rows, cols, vals))
    return scipy.sparse.csr_matrix((vals.real.astype(np.float64), (rows, cols)), shape=(N, N))


def ground_energy(find_conn, shape):
    H = sparse_hamiltonian(find_conn, shape)
    if H.shape[0] <= 64:
        return float(np.linalg.eigvalsh(H.toarray())[0])
    return float(scipy.sparse.linalg.eigsh(H, k=1, which='SA', tol=1e-12)[0][0])


def exact_energy(find_conn, shape, log_psi_all):
    """<psi|H|psi>/<psi|psi> for an enumerated log psi over all states."""
    H = sparse_hamiltonian(find_conn, shape)
    m = np.real(log_psi_all).max()
    psi = np.exp(log_psi_all - m)
    return np.vdot(psi, H @ psi) / np.vdot(psi, psi)
