"""Stochastic reconfiguration restated
(optimizers/stochastic_reconfiguration/optimizer.py:14-124,
optimizers/stochastic_reconfiguration/linear_equations.py:34-137)."""
import numpy as np


def centered(O):
    """``get_wave_function_jacobian_minus_mean`` (optimizer.py:111-114)."""
    return O - O.mean(axis=0, keepdims=True)


def force(O, eloc):
    """``get_energy_grad`` (optimizer.py:94-109) with targets
    conj(E_loc - E)/B (optimization/variational_monte_carlo.py:38-40,50):
    F = (O - Obar)^H conj(targets)."""
    B = O.shape[0]
    t = np.conj(eloc - eloc.mean()) / B
    return centered(O).conj().T @ np.conj(t)


def s_matvec(O, v, diag_shift):
    """optimizer.py:78-81: (O-Obar)^H ((O-Obar) v) / B + shift * v."""
    Oc = centered(O)
    return Oc.conj().T @ (Oc @ v) / O.shape[0] + diag_shift * v


def s_matrix(O, diag_shift):
    Oc = centered(O)
    return Oc.conj().T @ Oc / O.shape[0] + diag_shift * np.eye(O.shape[1])


def pinv_solve(A, b, reltol=1e-6):
    """tests/test_stochastic_reconfiguration.py:19-29."""
    u, s, vh = np.linalg.svd(A)
    keep = s > s.max() * reltol
    sinv = np.where(keep, 1.0 / np.where(keep, s, 1.0), 0.0)
    return vh.conj().T @ (sinv * (u.conj().T @ b))


def conjugate_gradient(apply, rhs, tol=1e-4, max_iter=20):
    """linear_equations.py:34-137 without preconditioner: x0 = 0, r0 = rhs,
    stop when i >= max_iter or ||r|| <= tol*||r0|| (norms as float32, :87-93),
    dot(x, y) = conj(x).y.  Returns (x, iterations, residual)."""
    x = np.zeros_like(rhs)
    r = rhs.copy()
    p = r.copy()
    gamma = np.vdot(r, r)
    tol_abs = np.float32(tol) * np.float32(np.linalg.norm(r))
    i = 0
    while (max_iter is None or i < max_iter) and np.float32(np.linalg.norm(r)) > tol_abs:
        z = apply(p)
        alpha = gamma / np.vdot(p, z)
        x = x + alpha * p
        r = r - alpha * z
        g_new = np.vdot(r, r)
        p = r + (g_new / gamma) * p
        gamma = g_new
        i += 1
    return x, i, r


def real_parameter_sr(O, eloc, diag_shift):
    """SR for real parameters with a complex log-amplitude (this build's
    definition for ConvNetAutoregressive2D, whose reference optimizer only
    accepts complex layers, optimizers/complex_values_optimizer.py:12-25):
    S = Re[(O-Obar)^H (O-Obar)]/B + shift*I, F = Re[(O-Obar)^H (E_loc - E)]/B,
    returns S^{-1} F by dense solve."""
    B, P = O.shape
    Oc = centered(O)
    S = np.real(Oc.conj().T @ Oc) / B + diag_shift * np.eye(P)
    F = np.real(Oc.conj().T @ (eloc - eloc.mean())) / B
    return np.linalg.solve(S, F), S, F
