"""A CBOR (RFC 8949) encoder and decoder for checkpoints.

Covers what a serde data model produces: unsigned and negative integers,
byte and text strings, arrays, maps (any key type; array keys decode to
tuples), floats (f16, f32 and f64 decoded; f64 written), booleans and null.
Blocks of conserved tuples -- an array of n arrays of nv floats, the bulk of
a checkpoint -- are written and read with numpy in one pass (`Rows`), and
the decoder takes the same fast path whenever such an array is made of f64
items; any other layout falls back to the general item-by-item decoder.
"""
import struct

import numpy as np


class Rows:
    """An [n][nv] float array written as an array of n arrays of nv floats."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.a.ndim == 2 and self.a.shape[1] < 24


def _head(major, n):
    if n < 24:
        return bytes([(major << 5) | n])
    if n < 1 << 8:
        return bytes([(major << 5) | 24, n])
    if n < 1 << 16:
        return bytes([(major << 5) | 25]) + struct.pack(">H", n)
    if n < 1 << 32:
        return bytes([(major << 5) | 26]) + struct.pack(">I", n)
    return bytes([(major << 5) | 27]) + struct.pack(">Q", n)


def _rows_bytes(a):
    n, nv = a.shape
    fields = [("h", "u1")]
    for k in range(nv):
        fields += [(f"t{k}", "u1"), (f"v{k}", ">f8")]
    rec = np.empty(n, dtype=np.dtype(fields))      # packed: 1 + 9 nv bytes per row
    rec["h"] = 0x80 | nv
    for k in range(nv):
        rec[f"t{k}"] = 0xFB
        rec[f"v{k}"] = a[:, k]
    return _head(4, n) + rec.tobytes()


def _enc(x, out):
    if isinstance(x, Rows):
        out.append(_rows_bytes(x.a))
    elif x is None:
        out.append(b"\xf6")
    elif x is True:
        out.append(b"\xf5")
    elif x is False:
        out.append(b"\xf4")
    elif isinstance(x, (int, np.integer)):
        x = int(x)
        out.append(_head(0, x) if x >= 0 else _head(1, -1 - x))
    elif isinstance(x, (float, np.floating)):
        out.append(b"\xfb" + struct.pack(">d", float(x)))
    elif isinstance(x, str):
        b = x.encode("utf-8")
        out.append(_head(3, len(b)) + b)
    elif isinstance(x, (bytes, bytearray)):
        out.append(_head(2, len(x)) + bytes(x))
    elif isinstance(x, dict):
        out.append(_head(5, len(x)))
        for k, v in x.items():
            _enc(k, out)
            _enc(v, out)
    elif isinstance(x, (list, tuple)):
        out.append(_head(4, len(x)))
        for v in x:
            _enc(v, out)
    elif isinstance(x, np.ndarray):
        _enc(x.tolist(), out)
    else:
        raise TypeError(f"cannot encode {type(x).__name__}")


def dumps(x):
    out = []
    _enc(x, out)
    return b"".join(out)


def _f16(b):
    return float(np.frombuffer(b, dtype=">f2")[0])


class _Reader:
    def __init__(self, buf):
        self.b = memoryview(buf)
        self.p = 0

    def arg(self, info):
        b, p = self.b, self.p
        if info < 24:
            return info
        if info == 24:
            self.p = p + 1
            return b[p]
        if info == 25:
            self.p = p + 2
            return struct.unpack(">H", b[p:p + 2])[0]
        if info == 26:
            self.p = p + 4
            return struct.unpack(">I", b[p:p + 4])[0]
        if info == 27:
            self.p = p + 8
            return struct.unpack(">Q", b[p:p + 8])[0]
        raise ValueError("indefinite-length CBOR items are not supported")

    def rows(self, n):
        """Fast path: n arrays of nv f64 items starting here, else None."""
        b, p = self.b, self.p
        if n == 0 or p >= len(b) or not (0x81 <= b[p] <= 0x97):
            return None
        nv = b[p] & 0x1F
        size = 1 + 9 * nv
        if p + n * size > len(b):
            return None
        raw = np.frombuffer(b[p:p + n * size], dtype=np.uint8).reshape(n, size)
        if not (np.all(raw[:, 0] == 0x80 | nv) and np.all(raw[:, 1::9] == 0xFB)):
            return None
        vals = np.empty((n, nv))
        for k in range(nv):
            vals[:, k] = raw[:, 2 + 9 * k:10 + 9 * k].copy().view(">f8").ravel()
        self.p = p + n * size
        return vals

    def item(self):
        ib = self.b[self.p]
        self.p += 1
        major, info = ib >> 5, ib & 0x1F
        if major == 0:
            return self.arg(info)
        if major == 1:
            return -1 - self.arg(info)
        if major in (2, 3):
            n = self.arg(info)
            s = bytes(self.b[self.p:self.p + n])
            self.p += n
            return s if major == 2 else s.decode("utf-8")
        if major == 4:
            n = self.arg(info)
            fast = self.rows(n)
            if fast is not None:
                return fast
            return [self.item() for _ in range(n)]
        if major == 5:
            n = self.arg(info)
            d = {}
            for _ in range(n):
                k = self.item()
                if isinstance(k, list):
                    k = tuple(k)
                d[k] = self.item()
            return d
        if major == 6:
            self.arg(info)          # tag: ignored, value kept
            return self.item()
        if info == 20:
            return False
        if info == 21:
            return True
        if info in (22, 23):
            return None
        if info == 25:
            v = _f16(bytes(self.b[self.p:self.p + 2]))
            self.p += 2
            return v
        if info == 26:
            v = struct.unpack(">f", self.b[self.p:self.p + 4])[0]
            self.p += 4
            return v
        if info == 27:
            v = struct.unpack(">d", self.b[self.p:self.p + 8])[0]
            self.p += 8
            return v
        raise ValueError(f"unsupported CBOR simple value {info}")


def loads(buf):
    r = _Reader(buf)
    x = r.item()
    if r.p != len(r.b):
        raise ValueError("trailing bytes after the CBOR item")
    return x
