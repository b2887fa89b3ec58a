"""ConvNetAutoregressive2D restated in float64 numpy
(machines/conv_net_autoregressive_2D.py:8-78, machines/abstract_machine.py:31-71).

Parameters are a list of (kernel, bias) pairs in network order:
    for each causal block b = 0 .. 2*depth-3:
        vconv  kernel (k, k, Cin_b, C)     bias (C,)        :38
        hconv  kernel (1, k, Cin_b, C)     bias (C,)        :39
        x1x1   kernel (1, 1, C, C//2)      bias (C//2,)     :43
        y1x1   kernel (1, 1, C, C//2)      bias (C//2,)     :45
        h3conv kernel (k, k, C, C)         bias (C,)        :50
    head       kernel (1, 1, C, 4)         bias (4,)        :73
with Cin_0 = 1 and Cin_b = C otherwise.  Weight normalisation is applied by the
caller (``weight_norm`` below restates deepar/layers/wrappers.py:123-135).
"""
import numpy as np


def layer_shapes(depth, channels, kernel_size=3):
    k, C = kernel_size, channels
    shapes = []
    n_blocks = 2 * depth - 2 if depth >= 2 else 2
    for b in range(n_blocks):
        cin = 1 if b == 0 else C
        shapes += [((k, k, cin, C), (C,)), ((1, k, cin, C), (C,)),
                   ((1, 1, C, C // 2), (C // 2,)), ((1, 1, C, C // 2), (C // 2,)),
                   ((k, k, C, C), (C,))]
    shapes.append(((1, 1, C, 4), (4,)))
    return shapes


def num_causal_blocks(depth):
    # block 0, (depth-2) residual blocks of 2 causal blocks, final block
    # (conv_net_autoregressive_2D.py:68-72); range(depth-2) is empty for depth<2.
    return 2 + 2 * max(depth - 2, 0)


def glorot_params(rng, depth, channels, kernel_size=3, bias_scale=0.0):
    """Glorot-uniform kernels, zero biases (Keras Conv2D defaults)."""
    out = []
    for kshape, bshape in layer_shapes(depth, channels, kernel_size):
        fan_in = kshape[0] * kshape[1] * kshape[2]
        fan_out = kshape[0] * kshape[1] * kshape[3]
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        kern = rng.uniform(-lim, lim, size=kshape)
        bias = bias_scale * rng.standard_normal(bshape)
        out.append((kern, bias))
    return out


def weight_norm(v, g, exponential_norm=True):
    """deepar/layers/wrappers.py:123-135: l2-normalise v over all axes but the
    last, scale per output channel by g (exp(g) when exponential_norm)."""
    gg = np.exp(g) if exponential_norm else g
    # tf.nn.l2_normalize(v, axes): v / sqrt(max(sum v^2, 1e-12))
    norm = np.sqrt(np.maximum((v * v).sum(axis=(0, 1, 2), keepdims=True), 1e-12))
    return v / norm * gg


def _act(x, activation):
    if activation == 'relu':
        return np.maximum(x, 0.0)
    if activation == 'elu':
        return np.where(x > 0, x, np.expm1(np.minimum(x, 0.0)))
    if activation == 'linear':
        return x
    raise ValueError(activation)


def _conv(x, kern, bias, pad):
    """Keras Conv2D 'valid' after ZeroPadding2D(pad=((top,bottom),(left,right)))."""
    (pt, pb), (pl, pr) = pad
    xp = np.pad(x, ((0, 0), (pt, pb), (pl, pr), (0, 0)))
    kh, kw = kern.shape[:2]
    B, Hp, Wp, _ = xp.shape
    H, W = Hp - kh + 1, Wp - kw + 1
    out = np.broadcast_to(bias, (B, H, W, kern.shape[3])).copy()
    for di in range(kh):
        for dj in range(kw):
            out += xp[:, di:di + H, dj:dj + W, :] @ kern[di, dj]
    return out


def _down_shift(x):
    # deepar/layers/masking.py:8-15 with axis=1
    out = np.zeros_like(x)
    out[:, 1:] = x[:, :-1]
    return out


def _right_shift(x):
    # deepar/layers/masking.py:8-15 with axis=2
    out = np.zeros_like(x)
    out[:, :, 1:] = x[:, :, :-1]
    return out


def _causal_block(v, h, layers, k, activation, use_horizontal_mask=False):
    """``_causal_conv_2d`` (conv_net_autoregressive_2D.py:33-51); returns the
    pre-activation (v', h')."""
    p = k - 1
    (vk, vb), (hk, hb), (xk, xb), (yk, yb), (ck, cb) = layers
    v_out = _conv(v, vk, vb, ((p, 0), (p // 2, p // 2)))
    x = _act(_conv(h, hk, hb, ((0, 0), (p, 0))), activation)
    if use_horizontal_mask:
        x = _right_shift(x)
    x = _conv(x, xk, xb, ((0, 0), (0, 0)))
    y = _down_shift(_act(v_out, activation))
    y = _conv(y, yk, yb, ((0, 0), (0, 0)))
    x, y = _act(x, activation), _act(y, activation)
    cat = np.concatenate([x, y], axis=-1)
    h_out = _conv(cat, ck, cb, ((p, 0), (p, 0)))
    return v_out, h_out


def unnormalized_logits(params, spins, depth, kernel_size=3, activation='relu'):
    """conv_net_autoregressive_2D.py:64-74 -> (B, H, W, 4) pre-complex logits."""
    k = kernel_size
    x = np.asarray(spins, dtype=np.float64)[..., None]          # ExpandInputDim+ToFloat32
    v, h = x, x
    it = iter(range(0, 5 * num_causal_blocks(depth), 5))

    def blk(v, h, mask=False):
        s = next(it)
        return _causal_block(v, h, params[s:s + 5], k, activation, mask)

    v, h = blk(v, h)
    v, h = _act(v, activation), _act(h, activation)
    for _ in range(depth - 2):                                   # _resiual_block :53-62
        v_in, h_in = v, h
        v, h = blk(v, h)
        v, h = _act(v, activation), _act(h, activation)
        v, h = blk(v, h)
        v, h = _act(v_in + v, activation), _act(h_in + h, activation)
    _, h = blk(v, h, mask=True)
    hk, hb = params[-1]
    return _conv(_act(h, activation), hk, hb, ((0, 0), (0, 0)))


def conditional_log_wave_function(logits):
    """VectorToComplexNumber (layers/complex/casting.py:16-24) then
    normalize_in_log_space(norm_type=2) (deepar/layers/autoregressive.py:7-15):
    Re -= 0.5*logsumexp(2*Re) over the two spin values."""
    re, im = logits[..., 0:2], logits[..., 2:4]
    m = re.max(axis=-1, keepdims=True)
    lse = m + 0.5 * np.log(np.exp(2 * (re - m)).sum(axis=-1, keepdims=True))
    return re - lse, im


def conditional_log_probs(params, spins, depth, kernel_size=3, activation='relu'):
    """AutoNormalizedAutoregressiveMachine.conditional_log_probs
    (abstract_machine.py:56-57): 2*Re of the normalised conditionals,
    shape (B, H, W, 2); index 0 is spin +1 (deepar/layers/one_hot.py:7-9)."""
    re, _ = conditional_log_wave_function(
        unnormalized_logits(params, spins, depth, kernel_size, activation))
    return 2.0 * re


def log_psi(params, spins, depth, kernel_size=3, activation='relu'):
    """combine_autoregressive_conditionals with PlusMinusOneToOneHot
    (deepar/layers/autoregressive.py:18-22, one_hot.py:7-9,
    abstract_machine.py:34-37): sum over sites of the conditional at the
    actual spin, idx = (1 - s) // 2.  Returns complex128 (B,)."""
    spins = np.asarray(spins)
    re, im = conditional_log_wave_function(
        unnormalized_logits(params, spins, depth, kernel_size, activation))
    idx = ((1 - spins.astype(np.int64)) // 2)[..., None]
    r = np.take_along_axis(re, idx, axis=-1)[..., 0].sum(axis=(1, 2))
    i = np.take_along_axis(im, idx, axis=-1)[..., 0].sum(axis=(1, 2))
    return r + 1j * i


def num_params(depth, channels, kernel_size=3):
    return sum(int(np.prod(k)) + int(np.prod(b))
               for k, b in layer_shapes(depth, channels, kernel_size))


def flops_per_forward(H, W, depth, channels):
    """Multiply-adds x2 of one full forward (k=3): first block 10C^2+12C MACs
    per site, each further causal block 22C^2, head 4C."""
    C = channels
    nb = num_causal_blocks(depth)
    macs = (10 * C * C + 12 * C) + (nb - 1) * 22 * C * C + 4 * C
    return 2 * H * W * macs
