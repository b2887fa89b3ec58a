"""Reading Keras HDF5 weight files (``model.save_weights``) without h5py.

A read-only parser of the subset of HDF5 such files use: superblock version 0,
version-1 object headers (with continuation blocks), groups as symbol tables
(version-1 B-trees, symbol-table nodes, local heaps), contiguous or compact
datasets of little-endian IEEE floats, and fixed-length string attributes.
It interprets bytes only; nothing in the file is executed.  Anything outside
that subset raises ``ValueError``.

``load_keras_weights(path)`` returns the weights in the order Keras saved
them: the file's ``layer_names`` attribute, then each layer group's
``weight_names`` attribute (Keras' own load order).
"""

import numpy as np


class _H5:
    def __init__(self, data):
        self.d = data
        if data[:8] != b'\x89HDF\r\n\x1a\n':
            raise ValueError('not an HDF5 file')
        if data[8] != 0:
            raise ValueError(f'HDF5 superblock version {data[8]} is not supported (only 0)')
        self.so, self.sl = data[13], data[14]
        if self.so != 8 or self.sl != 8:
            raise ValueError('only 8-byte offsets and lengths are supported')
        self.base = self._u(24, 8)
        # root group symbol table entry at 56
        self.root = self._entry(56)

    def _u(self, off, n):
        return int.from_bytes(self.d[off:off + n], 'little')

    def _entry(self, off):
        name_off, obj = self._u(off, 8), self._u(off + 8, 8)
        cache = self._u(off + 16, 4)
        scratch = (self._u(off + 24, 8), self._u(off + 32, 8)) if cache == 1 else None
        return name_off, obj, cache, scratch

    # -------------------------------------------------------------- objects
    def messages(self, addr):
        """(type, data bytes) of a version-1 object header, following
        continuation messages."""
        d = self.d
        if d[addr] != 1:
            raise ValueError(f'object header version {d[addr]} is not supported (only 1)')
        nmsg = self._u(addr + 2, 2)
        blocks = [(addr + 16, self._u(addr + 8, 4))]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p, end = start, start + size
            while p + 8 <= end and len(out) < nmsg:
                mtype, msize = self._u(p, 2), self._u(p + 2, 2)
                body = d[p + 8:p + 8 + msize]
                out.append((mtype, body))
                if mtype == 0x10:  # continuation
                    blocks.append((int.from_bytes(body[0:8], 'little'), int.from_bytes(body[8:16], 'little')))
                p += 8 + msize
        return out

    def heap_name(self, heap_addr, off):
        d = self.d
        if d[heap_addr:heap_addr + 4] != b'HEAP':
            raise ValueError('bad local heap')
        seg = self._u(heap_addr + 24, 8)
        end = d.index(b'\x00', seg + off)
        return d[seg + off:end].decode()

    def group_members(self, btree, heap):
        """name -> object header address of a symbol-table group."""
        out = {}
        d = self.d

        def walk(node):
            if d[node:node + 4] != b'TREE':
                raise ValueError('bad B-tree node')
            level, used = d[node + 5], self._u(node + 6, 2)
            p = node + 24
            children = []
            for _ in range(used):
                p += 8  # key
                children.append(self._u(p, 8))
                p += 8
            for c in children:
                if level > 0:
                    walk(c)
                else:
                    if d[c:c + 4] != b'SNOD':
                        raise ValueError('bad symbol table node')
                    nsym = self._u(c + 6, 2)
                    for k in range(nsym):
                        name_off, obj, _, _ = self._entry(c + 8 + 40 * k)
                        out[self.heap_name(heap, name_off)] = obj

        walk(btree)
        return out

    def children(self, obj):
        for mtype, body in self.messages(obj):
            if mtype == 0x11:
                return self.group_members(int.from_bytes(body[0:8], 'little'), int.from_bytes(body[8:16], 'little'))
        raise ValueError('object is not a group')

    def resolve(self, path):
        obj = self.root[1]
        for part in [p for p in path.split('/') if p]:
            obj = self.children(obj)[part]
        return obj

    # -------------------------------------------------------------- data
    @staticmethod
    def _dataspace(body):
        ver, ndim, flags = body[0], body[1], body[2]
        p = 8 if ver == 1 else 4
        return tuple(int.from_bytes(body[p + 8 * i:p + 8 * i + 8], 'little') for i in range(ndim))

    @staticmethod
    def _datatype(body):
        cls, ver = body[0] & 0x0F, body[0] >> 4
        size = int.from_bytes(body[4:8], 'little')
        bits = int.from_bytes(body[1:4], 'little')
        if cls == 1:
            if bits & 1:
                raise ValueError('big-endian floats are not supported')
            return np.dtype('<f%d' % size)
        if cls == 0:
            if bits & 1:
                raise ValueError('big-endian integers are not supported')
            return np.dtype(('<i%d' if bits & 8 else '<u%d') % size)
        if cls == 3:
            return np.dtype('S%d' % size)
        raise ValueError(f'HDF5 datatype class {cls} is not supported')

    def dataset(self, obj):
        shape, dtype, layout = None, None, None
        for mtype, body in self.messages(obj):
            if mtype == 0x01:
                shape = self._dataspace(body)
            elif mtype == 0x03:
                dtype = self._datatype(body)
            elif mtype == 0x08:
                layout = body
            elif mtype == 0x0B:
                raise ValueError('filtered (compressed) datasets are not supported')
        if shape is None or dtype is None or layout is None:
            raise ValueError('not a dataset')
        n = int(np.prod(shape)) if shape else 1
        ver = layout[0]
        if ver == 3:
            cls = layout[1]
            if cls == 1:
                addr = int.from_bytes(layout[2:10], 'little')
                raw = self.d[addr:addr + n * dtype.itemsize]
            elif cls == 0:
                size = int.from_bytes(layout[2:4], 'little')
                raw = layout[4:4 + size]
            else:
                raise ValueError('chunked datasets are not supported')
        elif ver in (1, 2):
            cls = layout[2]
            if cls != 1:
                raise ValueError('only contiguous version-1/2 layouts are supported')
            addr = int.from_bytes(layout[8:16], 'little')
            raw = self.d[addr:addr + n * dtype.itemsize]
        else:
            raise ValueError(f'data layout version {ver} is not supported')
        return np.frombuffer(raw, dtype=dtype, count=n).reshape(shape).copy()

    def attributes(self, obj):
        out = {}
        for mtype, body in self.messages(obj):
            if mtype != 0x0C:
                continue
            ver = body[0]
            nsz, tsz, ssz = (int.from_bytes(body[i:i + 2], 'little') for i in (2, 4, 6))
            p = 8
            if ver == 3:
                p = 9

            def pad(x):
                return (x + 7) & ~7 if ver == 1 else x

            name = body[p:p + nsz].split(b'\x00')[0].decode()
            p += pad(nsz)
            try:
                dtype = self._datatype(body[p:p + tsz])
            except ValueError:
                continue  # e.g. variable-length strings (backend, keras_version): not needed
            p += pad(tsz)
            shape = self._dataspace(body[p:p + ssz])
            p += pad(ssz)
            n = int(np.prod(shape)) if shape else 1
            out[name] = np.frombuffer(body[p:p + n * dtype.itemsize], dtype=dtype, count=n).reshape(shape)
        return out


def load_keras_weights(path):
    """[(weight name, array)] in Keras' save order."""
    with open(path, 'rb') as f:
        h = _H5(f.read())
    root = h.root[1]
    attrs = h.attributes(root)
    if 'layer_names' not in attrs:
        # weights saved inside a full-model file
        root = h.resolve('model_weights')
        attrs = h.attributes(root)
    out = []
    for lname in attrs['layer_names']:
        lname = lname.decode() if isinstance(lname, bytes) else str(lname)
        grp = h.children(root)[lname]
        wnames = h.attributes(grp).get('weight_names', [])
        for wn in wnames:
            wn = wn.decode() if isinstance(wn, bytes) else str(wn)
            obj = grp
            for part in wn.split('/'):
                obj = h.children(obj)[part]
            out.append((wn, h.dataset(obj)))
    return out


def _creation_index(layer_name):
    """Keras names the n-th layer of a class ``<class>_<n>`` (the first has no
    suffix); the suffix orders a model's layers by creation."""
    head, _, tail = layer_name.rpartition('_')
    return int(tail) if head and tail.isdigit() else 0


def convnet_params_from_keras(weights, shapes, weights_normalization=True):
    """Flat parameter vector (and weight-normalisation g vector) of a
    ConvNetAutoregressive2D from ``load_keras_weights`` output.

    Keras saves layers in topological order, which can differ from the order
    the network creates them (machines/conv_net_autoregressive_2D.py:38-73);
    layers are taken here in creation order and every kernel and bias shape is
    checked against ``shapes`` (the machine's (kernel, bias) shapes in network
    order).  With weight normalisation each wrapped layer holds kernel (v),
    bias and g (deepar/layers/wrappers.py:104-111); the head has no g."""
    layers = {}
    for name, arr in weights:
        lname, _, wname = name.partition('/')
        layers.setdefault(lname, {})[wname.split(':')[0]] = arr
    order = sorted(layers, key=_creation_index)
    if len(order) != len(shapes):
        raise ValueError(f'the file holds {len(order)} layers, the network has {len(shapes)}')
    flat, gs = [], []
    for i, (lname, (kshape, bshape)) in enumerate(zip(order, shapes)):
        w = layers[lname]
        if 'kernel' not in w or 'bias' not in w:
            raise ValueError(f'layer {lname} has no kernel/bias')
        if tuple(w['kernel'].shape) != tuple(kshape) or tuple(w['bias'].shape) != tuple(bshape):
            raise ValueError(f'layer {lname}: kernel {w["kernel"].shape} / bias {w["bias"].shape}, '
                             f'expected {kshape} / {bshape}')
        flat += [w['kernel'].astype(np.float32).reshape(-1), w['bias'].astype(np.float32).reshape(-1)]
        last = i == len(shapes) - 1
        if weights_normalization and not last:
            if 'g' not in w or w['g'].shape != (kshape[3],):
                raise ValueError(f'layer {lname} has no weight-normalisation g of {kshape[3]} values')
            gs.append(w['g'].astype(np.float32))
        elif 'g' in w:
            raise ValueError(f'layer {lname} is weight-normalised but the network is not')
    return np.concatenate(flat), (np.concatenate(gs) if gs else None)
