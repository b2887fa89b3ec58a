"""find_conn restated from the reference operators (numpy, CPU).

Layout contract (reference ``operators/heisenberg.py:63-82``): for a batch of
``B`` samples of shape ``(H, W)`` with entries +-1,
    conn  (Nc, B, H, W)  -- conn[0] is the sample itself
    mel   (Nc, B)        -- float64 matrix elements, mel[0] the diagonal
    use   (Nc, B)        -- bool, use[0] is True
"""
import numpy as np


def _shape2d(shape):
    shape = tuple(int(s) for s in shape)
    if len(shape) == 1:
        return (shape[0], 1), 1
    assert len(shape) == 2
    return shape, 2


def heisenberg_find_conn(sample, pbc=True, unitary_rotation=True):
    """Restates ``HeisenbergFindConn.find_conn/calc_conn_and_mel``
    (operators/heisenberg.py:47-121).

    Connection 1 + (i*W + j)*dim + d is the bond from (i, j) to its neighbour
    along direction d (d counts only dimensions of size > 1, :88-106; ``dim``
    is the rank of the *original* shape, :18).  Off-diagonal element is
    ``off_diag`` (-2 with the Marshall rotation, :11-14) when the two spins
    differ, else 0 and unused (:93,99,110,116).  The swapped configuration is
    written whether used or not (:94-95,111-112).  Open-boundary bonds are
    unused (:102-103,119-120).  Diagonal = sum over bonds of s_a*s_b (:91,
    :98,108,115,78-79).
    """
    sample = np.asarray(sample)
    orig_rank = sample.ndim - 1
    if orig_rank == 1:
        sample = sample[..., None]
    (H, W), _ = _shape2d(sample.shape[1:])
    dim = orig_rank
    off_diag = -2.0 if unitary_rotation else 2.0
    B = sample.shape[0]
    nc = H * W * dim + 1
    conn = np.repeat(sample[None].astype(np.float64), nc, axis=0)
    mel = np.zeros((nc, B))
    use = np.zeros((nc, B), dtype=bool)
    use[0] = True
    self_mel = np.zeros((B, H, W))
    for i in range(H):
        for j in range(W):
            d = 0
            for axis, size in ((0, H), (1, W)):
                if size <= 1:
                    continue
                ni, nj = (i + 1, j) if axis == 0 else (i, j + 1)
                wrap = (ni >= H) or (nj >= W)
                c = 1 + (i * W + j) * dim + d
                if wrap and not pbc:
                    use[c] = False
                else:
                    ni, nj = ni % H, nj % W
                    a, b = sample[:, i, j], sample[:, ni, nj]
                    self_mel[:, i, j] += a * b
                    use[c] = a != b
                    conn[c, :, i, j] = b
                    conn[c, :, ni, nj] = a
                mel[c] = np.where(use[c], off_diag, 0.0)
                d += 1
    mel[0] = self_mel.sum(axis=(1, 2))
    if orig_rank == 1:
        conn = conn[..., 0]
    return conn, mel, use


def ising_find_conn(sample, h=1.0, j=1.0, pbc=True):
    """Restates ``Ising.find_conn`` (operators/ising.py:17-46).

    conn[1 + i*W + j] flips site (i, j) (:18,31,36,42-43); off-diagonal -h for
    every connection, all used (:44-46).  Diagonal: per-site float32
    accumulation of -j*s_ij*s_right/down with wrap (pbc) or zero padding (obc)
    over dimensions of size > 1 (:19-23,37-41), summed over sites (:41).
    """
    sample = np.asarray(sample)
    shape = sample.shape[1:]
    B = sample.shape[0]
    n = int(np.prod(shape))
    flat = sample.reshape(B, n)
    # the arithmetic below is the reference's for float64 +-1 samples (what
    # its samplers return): every -j*s*s' term in float64, subtracted in
    # place from the float32 per-site array, then numpy's float32 sum
    s64 = sample.astype(np.float64)
    conn = np.repeat(flat[None], n + 1, axis=0).astype(sample.dtype)
    for s in range(n):
        conn[1 + s, :, s] *= -1
    conn = conn.reshape((n + 1, B) + shape)
    if len(shape) == 1:
        s2 = s64[:, :, None]
    else:
        s2 = s64
    H, W = s2.shape[1:]
    mode = 'wrap' if pbc else 'constant'
    padded = np.pad(s2, ((0, 0), (0, 1), (0, 1)), mode=mode)
    per_site = np.zeros((B, H, W), dtype=np.float32)
    for i in range(H):
        for jj in range(W):
            if len(shape) == 1:
                per_site[:, i, jj] -= j * s2[:, i, jj] * padded[:, i + 1, jj]
                continue
            if H > 1:
                per_site[:, i, jj] -= j * s2[:, i, jj] * padded[:, i + 1, jj]
            if W > 1:
                per_site[:, i, jj] -= j * s2[:, i, jj] * padded[:, i, jj + 1]
    mel = np.full((n + 1, B), -float(h))
    mel[0] = per_site.sum(axis=(1, 2))
    return conn, mel, np.ones_like(mel, dtype=bool)


def j1j2_edges(L1, L2, pbc):
    """Restates the bond list of ``j1j2_two_dim_netket_operator``
    (operators/j1j2.py:7-35): (site_a, site_b, color) with color 1 = J1
    (nearest neighbour) and color 2 = J2 (next-nearest neighbour), in the
    order the reference appends them."""
    e = []
    for hh in range(L1):
        for w in range(L2 - 1):
            e.append((w + L2 * hh, w + 1 + L2 * hh, 1))
            if hh < L1 - 1:
                e.append((w + L2 * hh, w + L2 * (hh + 1), 1))
                e.append((w + L2 * hh, w + 1 + L2 * (hh + 1), 2))
            elif pbc:
                e.append((w + L2 * hh, w, 1))
                e.append((w + L2 * hh, w + 1, 2))
            if hh > 0:
                e.append((w + L2 * hh, w + 1 + L2 * (hh - 1), 2))
            elif pbc:
                e.append((w + L2 * hh, w + 1 + L2 * (L1 - 1), 2))
        w = L2 - 1
        if pbc:
            e.append((L2 - 1 + L2 * hh, L2 * hh, 1))
            e.append((w + L2 * hh, L2 * ((hh + 1) % L1), 2))
            e.append((w + L2 * hh, L2 * ((L1 + hh - 1) % L1), 2))
        if hh < L1 - 1:
            e.append((w + L2 * hh, w + L2 * (hh + 1), 1))
        elif pbc:
            e.append((w + L2 * hh, w, 1))
    return e


def j1j2_find_conn(sample, j2=0.5, pbc=False):
    """J1J2 connections from the bond operator J_b*(sz sz + sx sx + sy sy)
    (operators/j1j2.py:37-55) on the in-repo edge list.

    Matrix elements: diagonal sum_b J_b s_a s_b; each antiparallel bond gives
    the swapped configuration with element +2*J_b (no Marshall sign).  The
    padding/use rule follows ``NetketOperatorWrapper.new_netket_find_conn``
    (operators/netket_operator.py:46-66): use = (mel != 0), use[0] = True,
    unused slots hold all-zero configurations with mel 0.

    Slot order: this build's canonical one-slot-per-bond order, slot 1 + k for
    bond k of ``j1j2_edges``; netket's own order cannot be pinned here (netket
    is not installed), so parity on J1J2 is defined on this layout and, for
    netket, as a set.  Nc = 2*dim*N + 1 (operators/j1j2.py:59).
    """
    sample = np.asarray(sample)
    B = sample.shape[0]
    L1, L2 = sample.shape[1:]
    n = L1 * L2
    edges = j1j2_edges(L1, L2, pbc)
    J = {1: 1.0, 2: float(j2)}
    nc = n * 2 * 2 + 1
    flat = sample.reshape(B, n).astype(np.float64)
    conn = np.zeros((nc, B, n))
    mel = np.zeros((nc, B))
    conn[0] = flat
    for k, (a, b, col) in enumerate(edges):
        sa, sb = flat[:, a], flat[:, b]
        mel[0] += J[col] * sa * sb
        anti = sa != sb
        c = 1 + k
        mel[c] = np.where(anti, 2.0 * J[col], 0.0)
        swapped = flat.copy()
        swapped[:, a], swapped[:, b] = sb, sa
        # a connection exists only where its element is non-zero; every other
        # slot is padding and holds the all-zero configuration (:56-60)
        conn[c] = np.where((mel[c] != 0.0)[:, None], swapped, 0.0)
    use = mel != 0.0
    use[0] = True
    return conn.reshape((nc, B, L1, L2)), mel, use


def compact_connections(use):
    """Ragged index map of the unbalanced E_loc path
    (observables/monte_carlo/operator.py:6-10,26-30): per-sample used counts,
    exclusive prefix-sum offsets, and the flat (sample, connection) list in
    sample-major / connection-minor order."""
    use = np.asarray(use, dtype=bool)
    counts = use.sum(axis=0).astype(np.int64)
    offsets = np.concatenate([[0], np.cumsum(counts)[:-1]])
    b_idx, c_idx = np.nonzero(use.T)
    return counts, offsets, b_idx, c_idx
