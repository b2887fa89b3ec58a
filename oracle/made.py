"""Masked autoregressive MLP (MADE) used as the learned independent proposal
of the Markov-chain sampler -- float64 numpy restatement, TEST
INFRASTRUCTURE ONLY.  This proposal has no counterpart in the reference; it
is pinned by the mathematics: sum_x q(x) == 1 over all 2^N states (a wrong
mask breaks it), the analytic gradient against central finite differences,
and exact-distribution z-tests of the sampler.

Network: x (N spins in {-1,+1}) -> h1 = relu(W1 x + b1) -> h2 = relu(W2 h1 + b2)
-> logits l = W3 h2 + b3; log q(x) = sum_i log sigmoid(x_i * l_i).
Degrees: input i -> i+1, hidden unit h -> 1 + h mod (N-1); W1[h, i] is live
iff i+1 <= deg(h), W2[t, k] iff deg(k) <= deg(t), W3[o, h] iff deg(h) <= o.
"""
import numpy as np


def degrees(N, Hd):
    return 1 + np.arange(Hd) % (N - 1)


def masks(N, Hd):
    dh = degrees(N, Hd)
    m1 = (np.arange(N)[None, :] + 1 <= dh[:, None]).astype(np.float64)   # (Hd, N)
    m2 = (dh[None, :] <= dh[:, None]).astype(np.float64)                 # (Hd, Hd) [t, k]
    m3 = (dh[None, :] <= np.arange(N)[:, None]).astype(np.float64)       # (N, Hd)
    return m1, m2, m3


def unpack(master, N, Hd):
    """[w1 (Hd,N) | b1 | w2t (Hd,Hd) with w2t[k,t] = W2[t,k] | b2 | w3 (N,Hd) | b3]
    -> (W1, b1, W2, b2, W3, b3) float64 with the masks applied."""
    v = np.asarray(master, dtype=np.float64)
    sizes = [Hd * N, Hd, Hd * Hd, Hd, N * Hd, N]
    parts, off = [], 0
    for n in sizes:
        parts.append(v[off:off + n])
        off += n
    m1, m2, m3 = masks(N, Hd)
    W1 = parts[0].reshape(Hd, N) * m1
    W2 = parts[2].reshape(Hd, Hd).T * m2
    W3 = parts[4].reshape(N, Hd) * m3
    return W1, parts[1], W2, parts[3], W3, parts[5]


def bf16_round(x):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    x32 = np.asarray(x, dtype=np.float32)
    bits = x32.view(np.uint32).astype(np.uint64)
    bits = (bits + 0x7FFF + ((bits >> 16) & 1)) & 0xFFFF0000
    return bits.astype(np.uint32).view(np.float32).astype(np.float64)


def _log_sigmoid(z):
    return np.where(z >= 0, -np.log1p(np.exp(-np.abs(z))), z - np.log1p(np.exp(-np.abs(z))))


def logits(params, x, hidden_bf16=False):
    """Logits of every site; with hidden_bf16 the hidden activations are
    rounded to bfloat16 (the proposal kernels hold them in BF16)."""
    W1, b1, W2, b2, W3, b3 = params
    rnd = bf16_round if hidden_bf16 else (lambda v: v)
    x = np.asarray(x, dtype=np.float64)
    h1 = rnd(np.maximum(x @ W1.T + b1, 0))
    h2 = rnd(np.maximum(h1 @ W2.T + b2, 0))
    return h2 @ W3.T + b3


def log_q(params, x, hidden_bf16=False):
    x = np.asarray(x, dtype=np.float64).reshape(len(x), -1)
    return _log_sigmoid(x * logits(params, x, hidden_bf16)).sum(axis=1)


def sample(params, uniforms, hidden_bf16=False):
    """Site by site: x_i = +1 iff sigmoid(l_i(x_<i)) > u_i."""
    B, N = uniforms.shape
    x = np.zeros((B, N))
    lq = np.zeros(B)
    for i in range(N):
        z = logits(params, x, hidden_bf16)[:, i]
        up = 1.0 / (1.0 + np.exp(-z)) > uniforms[:, i]
        x[:, i] = np.where(up, 1.0, -1.0)
        lq += _log_sigmoid(np.where(up, z, -z))
    return x, lq


def grad_mean_nll(params, x):
    """Gradient of -mean_b log q(x_b) w.r.t. (W1, b1, W2, b2, W3, b3) by
    backpropagation (masked entries zero)."""
    W1, b1, W2, b2, W3, b3 = params
    x = np.asarray(x, dtype=np.float64).reshape(len(x), -1)
    B, N = x.shape
    a1 = x @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    l = h2 @ W3.T + b3
    dl = 1.0 / (1.0 + np.exp(-l)) - (x > 0)               # d(-log sigmoid(x l))/dl
    dh2 = (dl @ W3) * (a2 > 0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    m1, m2, m3 = masks(N, W1.shape[0])
    return ((dh1.T @ x) / B * m1, dh1.mean(0), (dh2.T @ h1) / B * m2, dh2.mean(0), (dl.T @ h2) / B * m3, dl.mean(0))


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.T.ravel(), b2, W3.ravel(), b3])
