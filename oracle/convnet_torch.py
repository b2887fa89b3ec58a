"""ConvNetAutoregressive2D restated with torch.nn.functional.conv2d on the CPU
-- TEST INFRASTRUCTURE ONLY (and the timed CPU baseline of bench.py).

The same network as oracle/convnet.py (machines/conv_net_autoregressive_2D.py:
33-74, machines/abstract_machine.py:31-71), written a second way: NCHW
tensors, Keras ZeroPadding2D + 'valid' Conv2D as explicit F.pad + F.conv2d,
DownShift / RightShift (deepar/layers/masking.py:8-29) as pad-and-crop.  In
float64 it is the fast checker for large batches (it equals the numpy oracle
to 1e-12, tests/test_oracle.py); in float32 it is the CPU baseline's forward
(the arithmetic the reference's TensorFlow CPU path uses).
"""
import numpy as np
import torch
import torch.nn.functional as F

from .convnet import num_causal_blocks


def _act(x, activation):
    if activation == 'relu':
        return F.relu(x)
    if activation == 'elu':
        return F.elu(x)
    raise ValueError(activation)


def torch_params(params, dtype=torch.float64):
    """(kernel (kh, kw, cin, cout), bias) pairs -> (weight (cout, cin, kh, kw),
    bias) CPU tensors of ``dtype``."""
    return [(torch.as_tensor(np.ascontiguousarray(k)).permute(3, 2, 0, 1).contiguous().to(dtype),
             torch.as_tensor(np.asarray(b)).to(dtype)) for k, b in params]


def _conv(x, wb, pad):
    (pt, pb), (pl, pr) = pad
    return F.conv2d(F.pad(x, (pl, pr, pt, pb)), wb[0], wb[1])


def _block(v, h, layers, activation, mask=False):
    """_causal_conv_2d (conv_net_autoregressive_2D.py:33-51), pre-activation."""
    v_out = _conv(v, layers[0], ((2, 0), (1, 1)))
    x = _act(_conv(h, layers[1], ((0, 0), (2, 0))), activation)
    if mask:
        x = F.pad(x, (1, 0, 0, 0))[..., :-1]                    # RightShift
    x = _conv(x, layers[2], ((0, 0), (0, 0)))
    y = F.pad(_act(v_out, activation), (0, 0, 1, 0))[..., :-1, :]  # DownShift
    y = _conv(y, layers[3], ((0, 0), (0, 0)))
    cat = torch.cat([_act(x, activation), _act(y, activation)], dim=1)
    return v_out, _conv(cat, layers[4], ((2, 0), (2, 0)))


@torch.no_grad()
def unnormalized_logits(tparams, spins, depth, activation='relu'):
    """(B, 4, H, W) logits of conv_net_autoregressive_2D.py:64-74 for spins
    (B, H, W), with the parameters' dtype."""
    dtype = tparams[0][0].dtype
    x = torch.as_tensor(np.asarray(spins)).to(dtype)[:, None]
    it = iter(range(0, 5 * num_causal_blocks(depth), 5))

    def blk(v, h, mask=False):
        i = next(it)
        return _block(v, h, tparams[i:i + 5], activation, mask)

    v, h = blk(x, x)
    v, h = _act(v, activation), _act(h, activation)
    for _ in range(depth - 2):
        vi, hi = v, h
        v, h = blk(v, h)
        v, h = _act(v, activation), _act(h, activation)
        v, h = blk(v, h)
        v, h = _act(vi + v, activation), _act(hi + h, activation)
    _, h = blk(v, h, mask=True)
    return _conv(_act(h, activation), tparams[-1], ((0, 0), (0, 0)))


@torch.no_grad()
def log_psi(tparams, spins, depth, activation='relu', chunk=4096):
    """log psi (B,) complex128 (abstract_machine.py:34-37,49-57): normalised
    conditionals summed at the actual spins, in the parameters' dtype."""
    spins = np.asarray(spins)
    out = []
    for lo in range(0, spins.shape[0], chunk):
        s = spins[lo:lo + chunk]
        z = unnormalized_logits(tparams, s, depth, activation)
        re, im = z[:, 0:2], z[:, 2:4]
        re = re - 0.5 * torch.logsumexp(2 * re, dim=1, keepdim=True)
        idx = torch.as_tensor((1 - s.astype(np.int64)) // 2)[:, None]
        r = torch.gather(re, 1, idx)[:, 0].sum(dim=(1, 2))
        i = torch.gather(im, 1, idx)[:, 0].sum(dim=(1, 2))
        out.append(r.double().numpy() + 1j * i.double().numpy())
    return np.concatenate(out) if out else np.zeros(0, np.complex128)
