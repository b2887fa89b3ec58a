"""Wave functions evaluated by the HIP kernels.

``ConvNetAutoregressive2D`` mirrors machines/conv_net_autoregressive_2D.py
(with AutoNormalizedAutoregressiveMachine, machines/abstract_machine.py:49-71):
same constructor arguments, same architecture, same outputs -- log psi (B, 1)
complex64 and conditional log-probabilities (B, H, W, 2).  Parameters are one
flat float32 device tensor in the canonical (Keras) order; weight
normalisation (deepar/layers/wrappers.py:123-135) is a per-step transform of
that vector.  Every evaluation runs in the fused forward kernel
(fk_convar_forward).
"""
import numpy as np
import torch

from . import _lib

_ACT = {'relu': _lib.FK_ACT_RELU, 'elu': _lib.FK_ACT_ELU}
_PREC = {'fp32': _lib.FK_PREC_FP32, 'float32': _lib.FK_PREC_FP32, 'bf16': _lib.FK_PREC_BF16,
         'bfloat16': _lib.FK_PREC_BF16, 'fp16': _lib.FK_PREC_FP16, 'float16': _lib.FK_PREC_FP16}


def num_causal_blocks(depth):
    """Block 0, (depth-2) residual blocks of two causal blocks, and the masked
    final block (conv_net_autoregressive_2D.py:68-72)."""
    return 2 + 2 * max(depth - 2, 0)


def layer_shapes(depth, channels, kernel_size=3):
    """(kernel shape, bias shape) of every convolution in network order:
    per causal block vconv, hconv, x1x1, y1x1, h3conv
    (conv_net_autoregressive_2D.py:38-50), then the 1x1 C->4 head (:73)."""
    k, C = kernel_size, channels
    shapes = []
    for b in range(num_causal_blocks(depth)):
        cin = 1 if b == 0 else C
        shapes += [((k, k, cin, C), (C,)), ((1, k, cin, C), (C,)), ((1, 1, C, C // 2), (C // 2,)),
                   ((1, 1, C, C // 2), (C // 2,)), ((k, k, C, C), (C,))]
    shapes.append(((1, 1, C, 4), (4,)))
    return shapes


class _ConditionalLogProbsModel:
    """Stand-in for ``Model(inputs, machine.conditional_log_probs)``."""

    def __init__(self, machine):
        self.machine = machine
        self.input_shape = (None,) + machine.input_size

    def predict(self, x, batch_size=None):
        return self.machine.conditional_log_probs(x)


class ConvNetAutoregressive2D:
    """machines/conv_net_autoregressive_2D.py:8-78.

    ``keras_input_layer`` is the lattice shape (H, W) (or any object with a
    ``shape`` attribute ending in (H, W)).  ``precision`` selects the kernel
    arithmetic: 'fp32' (FP32 MFMA), 'bf16' (BF16 MFMA, FP32 accumulate) or
    'fp16' (FP16 MFMA, FP32 accumulate: same speed as bf16, 8x finer
    rounding of weights and activations, range limited to |x| < 65504)."""

    def __init__(self, keras_input_layer, depth, num_of_channels, kernel_size=3, strides=1, activation='relu',
                 weights_normalization=True, exponential_norm=True, precision='fp32', seed=0, device=None,
                 params=None):
        shape = getattr(keras_input_layer, 'shape', keras_input_layer)
        shape = tuple(int(s) for s in shape if s is not None)[-2:]
        if len(shape) != 2:
            raise ValueError('ConvNetAutoregressive2D needs a 2-D input shape')
        if kernel_size != 3 or strides != 1:
            raise NotImplementedError('the HIP kernels implement kernel_size=3, strides=1')
        if activation not in _ACT:
            raise NotImplementedError(f'activation {activation!r} (supported: relu, elu)')
        self.input_size = shape
        self.H, self.W = shape
        self.depth = int(depth)
        self.num_of_channels = int(num_of_channels)
        self.kernel_size = kernel_size
        self.activation = activation
        self.weights_normalization = bool(weights_normalization)
        self.exponential_norm = bool(exponential_norm)
        self.precision = precision
        self.input_shape = (None,) + shape
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else 'cpu'
        self.device = torch.device(device)
        self.shapes = layer_shapes(self.depth, self.num_of_channels, kernel_size)
        self.n_params = sum(int(np.prod(k)) + int(np.prod(b)) for k, b in self.shapes)
        if params is None:
            params = self.glorot_init(seed)
        self.set_params(params)
        self._packs = {}
        self._maps = {}
        self._fws = {}
        self.conditional_log_probs_model = _ConditionalLogProbsModel(self)

    # ------------------------------------------------------------------ params
    def glorot_init(self, seed):
        """Glorot-uniform kernels and zero biases (Keras Conv2D defaults) drawn
        from a seeded torch generator; with weight normalisation the g of each
        layer starts at log|v| per output channel (wrappers.py:26-39,104-111)."""
        gen = torch.Generator().manual_seed(int(seed))
        out = []
        for kshape, bshape in self.shapes:
            fan_in = kshape[0] * kshape[1] * kshape[2]
            fan_out = kshape[0] * kshape[1] * kshape[3]
            lim = float(np.sqrt(6.0 / (fan_in + fan_out)))
            out.append((torch.rand(kshape, generator=gen, dtype=torch.float64) * 2 - 1) * lim)
            out.append(torch.zeros(bshape, dtype=torch.float64))
        flat = torch.cat([t.reshape(-1) for t in out]).to(torch.float32)
        return flat

    def _wn_layer_views(self, flat):
        views, off = [], 0
        for idx, (kshape, bshape) in enumerate(self.shapes):
            nk = int(np.prod(kshape))
            views.append((idx, off, kshape))
            off += nk + int(np.prod(bshape))
        return views

    def _wn_initial_g(self, flat):
        gs = []
        for idx, off, kshape in self._wn_layer_views(flat):
            if idx == len(self.shapes) - 1:
                continue  # the head is never weight-normalised (:73)
            v = flat[off:off + int(np.prod(kshape))].reshape(kshape).to(torch.float64)
            norm = torch.sqrt((v * v).sum(dim=(0, 1, 2)))
            gs.append(torch.log(norm + 1e-10) if self.exponential_norm else norm)
        return torch.cat(gs).to(torch.float32)

    def _n_wn_g(self):
        return sum(k[3] for k, _ in self.shapes[:-1]) if self.weights_normalization else 0

    def set_params(self, params):
        """Copies ``params`` (the flat kernels and biases; with weight
        normalisation the kernels are the v of wrappers.py:123-135) into the
        machine.  The trainable values live in one device tensor
        ``trainable`` = [params | wn_g]; ``params`` and ``wn_g`` are views of
        it, updated in place so that device pointers (and captured graphs)
        stay valid.  On the first call with weight normalisation the g start
        at log(|v| + 1e-10) per output channel (CopyNormaInitializer,
        wrappers.py:26-39)."""
        p = torch.as_tensor(params, dtype=torch.float32).reshape(-1)
        if p.numel() != self.n_params:
            raise ValueError(f'expected {self.n_params} parameters, got {p.numel()}')
        if getattr(self, 'trainable', None) is None:
            self.trainable = torch.zeros(self.n_params + self._n_wn_g(), dtype=torch.float32, device=self.device)
            self.params = self.trainable[:self.n_params]
            if self.weights_normalization:
                self.wn_g = self.trainable[self.n_params:]
                self.wn_g.copy_(self._wn_initial_g(p.cpu()).to(self.device))
        self.params.copy_(p.to(self.device))
        self._version = getattr(self, '_version', 0) + 1

    @property
    def n_trainable(self):
        return int(self.trainable.numel())

    def set_wn_g(self, g):
        """Sets the weight-normalisation g of every wrapped layer (in place)."""
        if not self.weights_normalization:
            raise ValueError('the machine has no weight normalisation')
        g = torch.as_tensor(g, dtype=torch.float32).reshape(-1)
        if g.numel() != self.wn_g.numel():
            raise ValueError(f'expected {self.wn_g.numel()} g values, got {g.numel()}')
        self.wn_g.copy_(g.to(self.device))
        self._version += 1

    def mark_params_updated(self):
        self._version += 1

    def save_weights(self, filepath):
        """Parameters (and weight-normalisation g) to a safetensors file."""
        from .checkpoint import save_weights
        save_weights(self, filepath)

    def load_weights(self, filepath):
        """Keras ``model.load_weights``: an HDF5 file that the reference's
        ConvNetAutoregressive2D of the same configuration saved
        (experiments/run_evaluation.py:13-14) -- kernels, biases and, with weight
        normalisation, the g of every wrapped layer -- or a file written by
        ``save_weights``."""
        from .checkpoint import is_hdf5, load_weights
        if not is_hdf5(filepath):
            return load_weights(self, filepath)
        from .keras_weights import convnet_params_from_keras, load_keras_weights
        flat, g = convnet_params_from_keras(load_keras_weights(filepath), self.shapes, self.weights_normalization)
        self.set_params(torch.from_numpy(flat))
        if self.weights_normalization:
            self.set_wn_g(torch.from_numpy(g))
        return self

    def effective_params(self):
        """The flat kernel/bias vector the network actually uses: the raw
        parameters, or with weight normalisation kernel = l2_normalize(v)
        * exp(g) per output channel (wrappers.py:123-135)."""
        if not self.weights_normalization:
            return self.params
        parts, gi = [], 0
        p = self.params
        for idx, off, kshape in self._wn_layer_views(p):
            nk = int(np.prod(kshape))
            nb = self.shapes[idx][1][0]
            v = p[off:off + nk].reshape(kshape)
            if idx < len(self.shapes) - 1:
                cout = kshape[3]
                g = self.wn_g[gi:gi + cout]
                gi += cout
                scale = torch.exp(g) if self.exponential_norm else g
                # tf.nn.l2_normalize: v / sqrt(max(sum v^2, 1e-12))
                norm = torch.sqrt(torch.clamp((v * v).sum(dim=(0, 1, 2)), min=1e-12))
                v = v / norm * scale
            parts += [v.reshape(-1), p[off + nk:off + nk + nb]]
        return torch.cat(parts)

    def _pack(self, precision):
        """Packed (fragment-order) weight and side buffers of ``precision``.
        They are allocated once per precision and repacked in place (on the
        current stream) when the parameters changed, so a captured graph that
        reads them sees every update."""
        prec = _PREC[precision]
        if self._packs.get(prec, (None, -1))[1] == self._version:
            return self._packs[prec][0]
        _lib.require_gpu()
        if prec not in self._maps:
            nw, na = _lib.ctypes.c_int64(), _lib.ctypes.c_int64()
            _lib.call('fk_convar_packed_sizes', self.depth, self.num_of_channels, prec, _lib.ctypes.byref(nw),
                      _lib.ctypes.byref(na))
            wmap = np.empty(nw.value, np.int32)
            amap = np.empty(na.value, np.int32)
            _lib.call('fk_convar_pack_maps', self.depth, self.num_of_channels, prec,
                      wmap.ctypes.data_as(_lib.ctypes.c_void_p), amap.ctypes.data_as(_lib.ctypes.c_void_p))
            self._maps[prec] = (torch.from_numpy(wmap).to(self.device), torch.from_numpy(amap).to(self.device))
            from . import autotune
            autotune.apply_cached(self, prec)
        wmap, amap = self._maps[prec]
        eff = self.effective_params().contiguous()
        if prec in self._packs:
            wbuf, abuf = self._packs[prec][0]
        else:
            wdtype = {_lib.FK_PREC_BF16: torch.bfloat16, _lib.FK_PREC_FP16: torch.float16}.get(prec, torch.float32)
            wbuf = torch.empty(wmap.numel(), dtype=wdtype, device=self.device)
            abuf = torch.empty(amap.numel(), dtype=torch.float32, device=self.device)
        s = _lib.stream_ptr()
        _lib.call('fk_pack_gather', _lib.ptr(eff), _lib.ptr(wmap), wmap.numel(), prec, _lib.ptr(wbuf), s)
        _lib.call('fk_pack_gather', _lib.ptr(eff), _lib.ptr(amap), amap.numel(), _lib.FK_PREC_FP32,
                  _lib.ptr(abuf), s)
        self._packs[prec] = ((wbuf, abuf), self._version)
        return wbuf, abuf

    # ------------------------------------------------------------------ eval
    def _spins(self, x):
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        x = x.to(self.device)
        if x.dtype != torch.int8:
            x = x.to(torch.int8)
        return x.reshape(x.shape[0], self.H * self.W).contiguous()

    def forward(self, spins, pairs=None, m=None, m_dev=None, operator=None, want_cond=False, precision=None,
                stream=None):
        """Raw fused launch.  spins (B, H*W) int8; with ``pairs`` (m, 2)
        evaluates connected configurations of ``operator``.  Returns
        (logpsi (m, 2) float32, cond (m, H, W, 2) float32 or None)."""
        precision = precision or self.precision
        prec = _PREC[precision]
        wbuf, abuf = self._pack(precision)
        if pairs is None:
            m = spins.shape[0]
        elif m is None:
            m = pairs.shape[0]
        out = torch.empty((m, 2), dtype=torch.float32, device=self.device)
        cond = torch.empty((m, self.H, self.W, 2), dtype=torch.float32, device=self.device) if want_cond else None
        kind, ca, cb = 0, None, None
        if pairs is not None:
            tb = operator.tables(self.device)
            kind, ca, cb = operator.kind, tb['conn_a'], tb['conn_b']
        ws = self._forward_workspace(prec)
        _lib.call('fk_convar_forward_ws', _lib.ptr(wbuf), _lib.ptr(abuf), prec, self.H, self.W, self.depth,
                  self.num_of_channels, _ACT[self.activation], _lib.ptr(spins), _lib.ptr(pairs), _lib.ptr(m_dev),
                  int(m), kind, _lib.ptr(ca), _lib.ptr(cb), _lib.ptr(out), _lib.ptr(cond), _lib.ptr(ws),
                  0 if ws is None else ws.numel(), _lib.stream_ptr(stream))
        return out, cond

    def _forward_workspace(self, prec):
        """Global scratch of the forward for lattices whose activations exceed
        the LDS (allocated once per precision; None when not needed)."""
        act = _ACT[self.activation]
        key = (prec, _lib.load().fk_convar_get_schedule(self.H, self.W, self.depth, self.num_of_channels, act, prec))
        if key not in self._fws:
            n = int(_lib.load().fk_convar_forward_workspace_bytes(self.H, self.W, self.depth, self.num_of_channels,
                                                                   act, prec))
            if n < 0:
                raise _lib.FlowketHipError(f'ConvNetAutoregressive2D kernels: {self.H}x{self.W} lattice with '
                                           f'{self.num_of_channels} channels is not supported')
            self._fws[key] = torch.empty(n, dtype=torch.uint8, device=self.device) if n > 0 else None
        return self._fws[key]

    def log_psi(self, x, precision=None):
        """(B,) complex64 log psi."""
        out, _ = self.forward(self._spins(x), precision=precision)
        return torch.complex(out[:, 0], out[:, 1])

    def predict(self, x, batch_size=None):
        """Keras ``Model(inputs, machine.predictions).predict``: (B, 1) complex64."""
        return self.log_psi(x)[:, None]

    __call__ = predict

    def conditional_log_probs(self, x, precision=None):
        """(B, H, W, 2) float32; index 0 is spin +1 (deepar/layers/one_hot.py:7-9)."""
        _, cond = self.forward(self._spins(x), want_cond=True, precision=precision)
        return cond

    def log_psi_connected(self, spins, pairs, m, operator, m_dev=None, precision=None, stream=None):
        """log psi of the connected configurations (sample, connection) in
        ``pairs`` (first m rows, or *m_dev rows when given)."""
        out, _ = self.forward(spins, pairs=pairs, m=m, m_dev=m_dev, operator=operator, precision=precision,
                              stream=stream)
        return out

    def count_params(self):
        return self.n_params

    # ------------------------------------------------------------- row reuse
    row_reuse = True
    # connections first changing a row below this one run a restricted pass
    # (rows from there on); the others, the whole lattice
    row_reuse_first = 2

    def uses_row_reuse(self, operator):
        """Whether local energies of this network run the row-reuse forward
        (fk_convar_forward_rows): lattices of 3..64 rows and an operator with
        a connection table."""
        return bool(self.row_reuse) and 3 <= self.H <= 64 and hasattr(operator, 'row_buckets')

    def row_reuse_policies(self):
        """The row-reuse settings the autotuner compares: 0 (off), or the
        first row that takes restricted passes."""
        if not 3 <= self.H <= 64:
            return [0]
        firsts = sorted({max(2, int(round(f * self.H))) for f in (0.2, 0.3, 0.4, 0.5, 0.6)})
        return [0] + [f for f in firsts if f <= self.H - 1]

    def set_row_reuse_policy(self, policy):
        """0: no row reuse; k >= 2: row reuse with restricted passes for
        connections first changing row k or below."""
        policy = int(policy)
        self.row_reuse = policy > 0
        if policy > 0:
            self.row_reuse_first = max(2, policy)

    def row_reuse_policy(self):
        return int(self.row_reuse_first) if self.row_reuse else 0

    def log_psi_connected_rows(self, spins, r, operator, precision=None, stream=None):
        """log psi of every used connected configuration of the ragged map
        ``r`` (operator.compact: use, offsets, pairs) in its sample-major
        order, with row reuse: the samples' forward stores every block's
        vertical input and concatenation; each connection, bucketed by the
        first lattice row it changes (fk_bucket_connections), is then
        evaluated on its rows from there on only (fk_convar_forward_rows).
        No host round trip: bucket sizes stay on the device."""
        precision = precision or self.precision
        prec = _PREC[precision]
        wbuf, abuf = self._pack(precision)
        B = spins.shape[0]
        H, W = self.H, self.W
        dev = self.device
        nc = operator.num_connections
        bucket_of, nbk, sizes = operator.row_buckets(W, dev, first=self.row_reuse_first)
        tb = operator.tables(dev)
        ws = self._rows_workspace(prec, B, nc, nbk)
        _lib.call('fk_bucket_connections', _lib.ptr(r['use']), B, nc, _lib.ptr(bucket_of), nbk,
                  _lib.ptr(r['offsets']), _lib.ptr(ws['pairs']), _lib.ptr(ws['out_index']), _lib.ptr(ws['boff']),
                  _lib.ptr(ws['scratch']), _lib.stream_ptr(stream))
        logpsi = torch.empty((nc * B, 2), dtype=torch.float32, device=dev)
        fws = self._forward_workspace(prec)
        fwn = 0 if fws is None else fws.numel()
        act = _ACT[self.activation]
        common = (_lib.ptr(wbuf), _lib.ptr(abuf), prec, H, W, self.depth, self.num_of_channels, act)
        # the samples themselves (slot 0): write the cache and the row prefixes
        _lib.call('fk_convar_forward_rows', *common, _lib.ptr(spins), None, None, B, operator.kind,
                  _lib.ptr(tb['conn_a']), _lib.ptr(tb['conn_b']), 0, _lib.ptr(ws['rcache']), 1,
                  _lib.ptr(ws['prefix']), _lib.ptr(r['offsets']), _lib.ptr(logpsi), _lib.ptr(fws), fwn,
                  _lib.stream_ptr(stream))
        for k in range(1, nbk):
            upper = int(sizes[k]) * B
            if upper == 0:
                continue
            row0 = k - 1
            if row0 == 1:
                row0 = 0  # two halo rows are needed above the first computed row
            rng = ws['boff'][k:k + 2]
            _lib.call('fk_convar_forward_rows', *common, _lib.ptr(spins), _lib.ptr(ws['pairs']), _lib.ptr(rng), upper,
                      operator.kind, _lib.ptr(tb['conn_a']), _lib.ptr(tb['conn_b']), row0, _lib.ptr(ws['rcache']),
                      0, _lib.ptr(ws['prefix']), _lib.ptr(ws['out_index']), _lib.ptr(logpsi), _lib.ptr(fws), fwn,
                      _lib.stream_ptr(stream))
        return logpsi

    def _rows_workspace(self, prec, B, nc, nbk):
        key = (prec, B, nc, nbk)
        cache = self.__dict__.setdefault('_rows_ws', {})
        if key not in cache:
            dev = self.device
            per = int(_lib.load().fk_convar_rcache_bytes(self.H, self.W, self.depth, self.num_of_channels, prec))
            cache.clear()
            cache[key] = dict(
                pairs=torch.empty((nc * B, 2), dtype=torch.int32, device=dev),
                out_index=torch.empty(nc * B, dtype=torch.int64, device=dev),
                boff=torch.empty(nbk + 1, dtype=torch.int64, device=dev),
                scratch=torch.empty(int(_lib.load().fk_bucket_scratch_bytes(B, nbk)), dtype=torch.uint8, device=dev),
                rcache=torch.empty(per * B, dtype=torch.uint8, device=dev),
                prefix=torch.empty((B, self.H + 1, 2), dtype=torch.float32, device=dev))
        return cache[key]

    def supports_implicit_products(self):
        """vjp / jvp (batched products with O that never form it) run on the
        MFMA kernels, for channel counts that are a multiple of 32."""
        return self.num_of_channels % 32 == 0

    def _implicit_workspace(self, kind, B):
        fn = {'vjp': 'fk_convar_vjp_workspace_bytes', 'jvp': 'fk_convar_jvp_workspace_bytes'}[kind]
        nbytes = int(getattr(_lib.load(), fn)(self.H, self.W, self.depth, self.num_of_channels, B))
        if nbytes < 0:
            raise _lib.FlowketHipError(f'{kind}: {self.num_of_channels} channels are not supported (the implicit '
                                       'products need a multiple of 32)')
        cache = self.__dict__.setdefault('_implicit_ws', {})
        ws = cache.get(kind)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
            cache[kind] = ws
        return ws

    def vjp_effective(self, x, cot, scale=1.0, stream=None):
        """scale * sum_b (cot[b,0] dRe log psi_b + cot[b,1] dIm log psi_b) with
        respect to the effective kernels and biases (P floats): scale *
        Re[O^H u] for u = cot[:,0] + i cot[:,1], through fk_convar_vjp (one
        batched backward pass, O never formed)."""
        _lib.require_gpu()
        spins = self._spins(x)
        B = spins.shape[0]
        cot = torch.as_tensor(cot).to(self.device, torch.float32).reshape(B, 2).contiguous()
        grad = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
        ws = self._implicit_workspace('vjp', B)
        _lib.call('fk_convar_vjp', _lib.ptr(self.effective_params().contiguous()), self.H, self.W, self.depth,
                  self.num_of_channels, _ACT[self.activation], _lib.ptr(spins), B, _lib.ptr(cot), float(scale),
                  _lib.ptr(ws), _lib.ptr(grad), None, _lib.stream_ptr(stream))
        return grad

    def jvp_effective(self, x, tangent, stream=None):
        """(B, 2) float32: (dRe log psi_b . t, dIm log psi_b . t) for a
        direction t over the effective kernels and biases (O t), by
        forward-mode differentiation (fk_convar_jvp)."""
        _lib.require_gpu()
        spins = self._spins(x)
        B = spins.shape[0]
        t = torch.as_tensor(tangent).to(self.device, torch.float32).reshape(-1).contiguous()
        if t.numel() != self.n_params:
            raise ValueError(f'tangent has {t.numel()} values, the network {self.n_params}')
        out = torch.empty((B, 2), dtype=torch.float32, device=self.device)
        ws = self._implicit_workspace('jvp', B)
        _lib.call('fk_convar_jvp', _lib.ptr(self.effective_params().contiguous()), _lib.ptr(t), self.H, self.W,
                  self.depth, self.num_of_channels, _ACT[self.activation], _lib.ptr(spins), B, _lib.ptr(ws),
                  _lib.ptr(out), None, _lib.stream_ptr(stream))
        return out

    def per_sample_gradients(self, x, workspace=None, stream=None):
        """O = d log psi / d theta per sample (machines/abstract_machine.py:24-28,
        utils/jacobian.py:114-123) through fk_convar_grad: returns (o_re, o_im,
        logpsi) with o_* (B, P) float32 in the flat parameter order and logpsi
        (B, 2) float32 (FP32 forward of the gradient kernel).  The columns are
        the trainable values: the flat parameters, and with weight
        normalisation the g of every wrapped layer after them (the kernel
        columns then hold d/dv), n_trainable columns in all."""
        _lib.require_gpu()
        spins = self._spins(x)
        B = spins.shape[0]
        P, PT = self.n_params, self.n_trainable
        o_re = torch.empty((B, PT), dtype=torch.float32, device=self.device)
        o_im = torch.empty((B, PT), dtype=torch.float32, device=self.device)
        lp = torch.empty((B, 2), dtype=torch.float32, device=self.device)
        nbytes = _lib.load().fk_convar_grad_workspace_bytes(self.H, self.W, self.depth, self.num_of_channels, B)
        if workspace is None or workspace.numel() < nbytes:
            workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        _lib.call('fk_convar_grad', _lib.ptr(self.effective_params().contiguous()), self.H, self.W, self.depth,
                  self.num_of_channels, _ACT[self.activation], _lib.ptr(spins), B, _lib.ptr(workspace),
                  _lib.ptr(o_re), _lib.ptr(o_im), PT, _lib.ptr(lp), _lib.stream_ptr(stream))
        if self.weights_normalization:
            for o in (o_re, o_im):
                self._wn_rows(o, B)
        return o_re, o_im, lp

    # ---------------------------------------------------- weight normalisation
    def _wn_table(self):
        """(device int64 table (layers, 4) = kernel offset, kh*kw*cin, cout,
        g offset; number of layers; largest cout) of the wrapped layers."""
        if getattr(self, '_wn_tab', None) is None:
            rows, off, goff = [], 0, 0
            for idx, (kshape, bshape) in enumerate(self.shapes):
                nk = int(np.prod(kshape))
                if idx < len(self.shapes) - 1:
                    rows.append([off, nk // kshape[3], kshape[3], goff])
                    goff += kshape[3]
                off += nk + int(np.prod(bshape))
            tab = torch.as_tensor(np.asarray(rows, dtype=np.int64)).to(self.device)
            self._wn_tab = (tab, len(rows), max(r[2] for r in rows))
        return self._wn_tab

    def _wn_rows(self, rows, nrows):
        """d/dk rows -> d/d(v, g) rows in place (fk_wn_grad_rows)."""
        tab, nl, mc = self._wn_table()
        _lib.call('fk_wn_grad_rows', _lib.ptr(rows), nrows, rows.shape[-1], _lib.ptr(self.params),
                  _lib.ptr(self.wn_g), _lib.ptr(tab), nl, mc, int(self.exponential_norm), self.n_params,
                  _lib.stream_ptr())

    def vjp(self, x, cot, scale=1.0, stream=None):
        """scale * Re[O^H u] (u = cot[:,0] + i cot[:,1]) over the trainable
        values (n_trainable floats), O never formed: fk_convar_vjp, then the
        weight-normalisation chain rule."""
        g = self.vjp_effective(x, cot, scale, stream)
        if not self.weights_normalization:
            return g
        out = torch.zeros(self.n_trainable, dtype=torch.float32, device=self.device)
        out[:self.n_params] = g
        self._wn_rows(out, 1)
        return out

    def jvp(self, x, tangent, stream=None):
        """O t (B, 2) for a direction t over the trainable values."""
        t = torch.as_tensor(tangent).to(self.device, torch.float32).reshape(-1).contiguous()
        if t.numel() != self.n_trainable:
            raise ValueError(f'tangent has {t.numel()} values, the machine {self.n_trainable} trainable values')
        if self.weights_normalization:
            tab, nl, mc = self._wn_table()
            t_eff = torch.empty(self.n_params, dtype=torch.float32, device=self.device)
            _lib.call('fk_wn_tangent', _lib.ptr(self.params), _lib.ptr(self.wn_g), _lib.ptr(tab), nl, mc,
                      int(self.exponential_norm), _lib.ptr(t), _lib.ptr(t_eff), self.n_params, _lib.stream_ptr())
            t = t_eff
        return self.jvp_effective(x, t, stream)


class RBM:
    """Complex RBM (machines/rbm.py:12-48): log psi = sum_j a_j s_j +
    sum_h lncosh(b_h + sum_j W_jh s_j), hidden units alpha * N
    (multiply_units_by_input_dim, layers/complex/dense.py:27-28).  Each complex
    parameter is stored as (real, imag) and used as real - 1j*imag
    (layers/complex/base_layer.py:30-35); the flat parameter order is the
    Keras weight order W_re, W_im, b_re, b_im, a_re, a_im.  Initial values are
    N(0, stddev^2) for every real and imaginary part (rbm.py:23,
    layers/complex/initializers.py:137-145) from a seeded generator."""

    def __init__(self, keras_input_layer, alpha=1, stddev=1.0, use_float64_ops=False, seed=0, device=None,
                 params=None):
        if use_float64_ops:
            raise NotImplementedError('the RBM kernels compute in FP32 (use_float64_ops=False)')
        shape = getattr(keras_input_layer, 'shape', keras_input_layer)
        shape = tuple(int(s) for s in shape if s is not None)
        self.input_size = shape
        self.input_shape = (None,) + shape
        self.n_visible = int(np.prod(shape))
        self.alpha = int(alpha)
        self.n_hidden = self.alpha * self.n_visible
        if device is None:
            device = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() else 'cpu'
        self.device = torch.device(device)
        N, M = self.n_visible, self.n_hidden
        self.sizes = [N * M, N * M, M, M, N, N]
        self.n_params = sum(self.sizes)
        if params is None:
            gen = torch.Generator().manual_seed(int(seed))
            params = torch.randn(self.n_params, generator=gen, dtype=torch.float64) * stddev
        self.set_params(params)

    def set_params(self, params):
        p = torch.as_tensor(params, dtype=torch.float32).reshape(-1)
        if p.numel() != self.n_params:
            raise ValueError(f'expected {self.n_params} parameters, got {p.numel()}')
        if getattr(self, 'params', None) is not None:
            # in place: device pointers stay valid for captured graphs
            self.params.copy_(p.to(self.device))
        else:
            self.params = p.to(self.device).clone().contiguous()

    def mark_params_updated(self):
        pass

    @property
    def trainable(self):
        """The trainable values (the stored real and imaginary parts)."""
        return self.params

    @property
    def n_trainable(self):
        return self.n_params

    def supports_implicit_products(self):
        return False

    def per_sample_gradients(self, x, stream=None):
        """O = d log psi / d(parameter) per sample over the stored real and
        imaginary parts (fk_rbm_grad): (o_re, o_im, logpsi), o_* (B, P)."""
        _lib.require_gpu()
        spins = self._spins(x)
        B = spins.shape[0]
        o_re = torch.empty((B, self.n_params), dtype=torch.float32, device=self.device)
        o_im = torch.empty((B, self.n_params), dtype=torch.float32, device=self.device)
        w = [_lib.ptr(t) for t in self.parts()]
        _lib.call('fk_rbm_grad', *w, self.n_visible, self.n_hidden, _lib.ptr(spins), B, _lib.ptr(o_re),
                  _lib.ptr(o_im), self.n_params, _lib.stream_ptr(stream))
        lp, _ = self.forward(spins, stream=stream)
        return o_re, o_im, lp

    def parts(self):
        """(w_re, w_im, b_re, b_im, a_re, a_im) views of the flat parameters."""
        out, off = [], 0
        for n in self.sizes:
            out.append(self.params[off:off + n])
            off += n
        return out

    def _spins(self, x):
        if not isinstance(x, torch.Tensor):
            x = torch.as_tensor(np.asarray(x))
        x = x.to(self.device)
        if x.dtype != torch.int8:
            x = x.to(torch.int8)
        return x.reshape(x.shape[0], self.n_visible).contiguous()

    def forward(self, spins, pairs=None, m=None, m_dev=None, operator=None, stream=None):
        if pairs is None:
            m = spins.shape[0]
        elif m is None:
            m = pairs.shape[0]
        out = torch.empty((m, 2), dtype=torch.float32, device=self.device)
        kind, ca, cb = 0, None, None
        if pairs is not None:
            tb = operator.tables(self.device)
            kind, ca, cb = operator.kind, tb['conn_a'], tb['conn_b']
        w = [_lib.ptr(t) for t in self.parts()]
        _lib.call('fk_rbm_log_psi', *w, self.n_visible, self.n_hidden, _lib.ptr(spins), _lib.ptr(pairs),
                  _lib.ptr(m_dev), int(m), kind, _lib.ptr(ca), _lib.ptr(cb), _lib.ptr(out), _lib.stream_ptr(stream))
        return out, None

    def log_psi(self, x):
        out, _ = self.forward(self._spins(x))
        return torch.complex(out[:, 0], out[:, 1])

    def predict(self, x, batch_size=None):
        return self.log_psi(x)[:, None]

    __call__ = predict

    def log_psi_connected(self, spins, pairs, m, operator, m_dev=None, precision=None, stream=None):
        out, _ = self.forward(spins, pairs=pairs, m=m, m_dev=m_dev, operator=operator, stream=stream)
        return out

    def save_weights(self, filepath):
        from .checkpoint import save_weights
        save_weights(self, filepath)

    def load_weights(self, filepath):
        from .checkpoint import load_weights
        return load_weights(self, filepath)

    def count_params(self):
        return self.n_params
