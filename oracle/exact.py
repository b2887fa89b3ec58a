"""Exact enumeration helpers (exact/utils.py:39-58), exact diagonalisation
of the find_conn matrix elements, and the quantities of the exact
enumeration path -- ExactVariational / ExactObservable
(optimization/exact_variational.py:10-161) and ExactSampler
(samplers/exact_sampler.py:14-20) -- restated in FP64 numpy.  Test
infrastructure only."""
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


def complex_norm_log_fsum_exp(arr):
    """exact/utils.py:79-82 (numpy.sum in place of accupy's fsum)."""
    real_arr = np.real(arr)
    m = np.max(real_arr)
    return np.log(np.sum(np.exp(real_arr - m))) + m


def exact_variational(log_psi_all, find_conn, shape):
    """ExactVariational._update_wave_function_arrays (exact_variational.py:
    127-133), ExactObservable.update_local_energy (:80-100) with the variance
    of the local operator, and _update_local_energy (:135-138), following the
    reference's formulas term by term: energies_k = sum_j conj(mel_j)
    exp(conj(log psi(k'_j)) + log psi(k)) / |psi|^2, current_energy = sum_k
    energies_k, naive energies E_loc(k) = sum_j mel_j exp(log psi(k'_j) -
    log psi(k)), variance = sum_k p_k Re(E_loc(k) - current_energy)^2,
    coefficients = energies - p * current_energy."""
    n = int(np.prod(shape))
    N = 2 ** n
    lp = np.asarray(log_psi_all, dtype=np.complex128).reshape(N)
    psi_squared = 2.0 * lp
    log_norm2 = complex_norm_log_fsum_exp(psi_squared)
    norm2 = np.exp(log_norm2)
    log_probs = np.real(psi_squared) - log_norm2
    probs = np.exp(log_probs)
    states = all_states(shape)
    conn, mel, use = find_conn(states)
    nc = conn.shape[0]
    idx = binary_array_to_decimal_array(np.asarray(conn).reshape(nc, N, n))
    mel = np.asarray(mel).astype(np.complex128) * np.asarray(use)
    lpc = lp[idx]                                   # (nc, N)
    energies = np.sum(np.conj(mel) * np.exp(np.conj(lpc) + lp[None, :]), axis=0) / norm2
    current_energy = np.sum(energies)
    naive = np.sum(mel * np.exp(lpc - lp[None, :]), axis=0)
    d = np.real(naive - current_energy)
    variance = float(np.sum(d * d * probs))
    coefficients = energies - probs * current_energy
    return dict(probs=probs, log_norm2=log_norm2, eloc=naive, energies=energies, current_energy=current_energy,
                variance=variance, coefficients=coefficients, states=states)


def exact_sample(probs, uniforms):
    """samplers/exact_sampler.py:14-20 on given uniforms: numpy's choice is
    cdf = cumsum(p), cdf /= cdf[-1], index = searchsorted(cdf, u, 'right')."""
    cdf = np.cumsum(np.asarray(probs, dtype=np.float64))
    cdf /= cdf[-1]
    return np.searchsorted(cdf, np.asarray(uniforms), side='right'), cdf
