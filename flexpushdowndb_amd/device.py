"""Host-side handles over the C ABI: a device context, device arrays and
Arrow-layout device columns.  All device memory is owned by the library's
context (with its HBM budget); numpy arrays are the host representation."""
import ctypes

import numpy as np

from . import _capi as C


class Context:
    """One fpdb_context (device, stream, HBM budget)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self._lib = C.lib()
        h = ctypes.c_void_p()
        C.check(self._lib.fpdb_context_create(device, hbm_budget_bytes, None, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            self._lib.fpdb_context_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # --- memory ---
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        C.check(self._lib.fpdb_device_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return p.value

    def free(self, ptr):
        if ptr:
            C.check(self._lib.fpdb_device_free(self.handle, ptr))

    def memory(self):
        used, budget = ctypes.c_uint64(), ctypes.c_uint64()
        C.check(self._lib.fpdb_context_memory(self.handle, ctypes.byref(used), ctypes.byref(budget)))
        return used.value, budget.value

    def device_info(self):
        cus, xcds, hbm = ctypes.c_int(), ctypes.c_int(), ctypes.c_uint64()
        C.check(self._lib.fpdb_device_info(self.handle, ctypes.byref(cus), ctypes.byref(xcds), ctypes.byref(hbm)))
        return {"num_cus": cus.value, "num_xcds": xcds.value, "hbm_bytes": hbm.value}

    def synchronize(self):
        C.check(self._lib.fpdb_context_synchronize(self.handle))

    def upload(self, host):
        host = np.ascontiguousarray(host)
        d = DeviceArray(self, max(1, host.nbytes), host.dtype, len(host) if host.ndim else 1)
        if host.nbytes:
            C.check(self._lib.fpdb_copy_to_device(self.handle, d.ptr, host.ctypes.data, host.nbytes))
        return d

    def empty(self, n, dtype):
        dtype = np.dtype(dtype)
        return DeviceArray(self, max(1, int(n) * dtype.itemsize), dtype, int(n))

    def zeros(self, n, dtype):
        d = self.empty(n, dtype)
        self.fill(d.ptr, 0, d.nbytes)
        return d

    def fill(self, ptr, value, nbytes):
        C.check(self._lib.fpdb_fill(self.handle, ptr, value, int(nbytes)))

    def download(self, ptr, n, dtype):
        dtype = np.dtype(dtype)
        out = np.empty(int(n), dtype)
        if out.nbytes:
            C.check(self._lib.fpdb_copy_to_host(self.handle, out.ctypes.data, ptr, out.nbytes))
        return out

    # --- timing on the context stream ---
    def timer_start(self):
        C.check(self._lib.fpdb_timer_start(self.handle))

    def timer_stop(self):
        ms = ctypes.c_float()
        C.check(self._lib.fpdb_timer_stop(self.handle, ctypes.byref(ms)))
        return ms.value


class DeviceArray:
    """A device buffer owned by a Context."""

    def __init__(self, ctx, nbytes, dtype, length):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.dtype = np.dtype(dtype)
        self.length = int(length)
        self.ptr = ctx.alloc(self.nbytes)

    def numpy(self, n=None):
        return self.ctx.download(self.ptr, self.length if n is None else n, self.dtype)

    def free(self):
        if self.ptr and self.ctx.handle:
            self.ctx.free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceColumn:
    """An Arrow-layout column on the device (values + optional validity)."""

    def __init__(self, values, validity=None, type_id=None, length=None, offset=0):
        self.values = values
        self.validity = validity
        self.type_id = type_id if type_id is not None else C.NP_TO_TYPE[values.dtype]
        self.length = values.length if length is None else int(length)
        self.offset = int(offset)

    @classmethod
    def from_numpy(cls, ctx, arr, valid=None, type_id=None):
        """valid: optional boolean mask (True = valid), packed LSB-first."""
        vals = ctx.upload(arr)
        vb = None
        if valid is not None:
            bits = np.packbits(np.asarray(valid, dtype=bool), bitorder="little")
            pad = np.zeros(((len(arr) + 31) // 32) * 4, np.uint8)
            pad[:len(bits)] = bits
            vb = ctx.upload(pad)
        return cls(vals, vb, type_id, len(arr))

    @classmethod
    def empty(cls, ctx, n, dtype, nullable=False, type_id=None, validity_rows=None):
        """n rows of values; a validity bitmap (when nullable) of
        max(n, validity_rows) rows."""
        vals = ctx.empty(max(1, n), dtype)
        vrows = max(1, n, validity_rows or 0)
        vb = ctx.zeros(((vrows + 31) // 32) * 4, np.uint8) if nullable else None
        return cls(vals, vb, type_id, n)

    @property
    def dtype(self):
        return C.TYPE_TO_NP[self.type_id]

    def c(self, length=None):
        return C.Column(self.type_id, 0, self.length if length is None else int(length), self.offset,
                        self.values.ptr, self.validity.ptr if self.validity is not None else None)

    def numpy(self, n=None):
        n = self.length if n is None else int(n)
        return self.values.ctx.download(self.values.ptr + self.offset * self.dtype.itemsize, n, self.dtype)

    def valid_mask(self, n=None):
        """Validity of rows [0, n) of the column: bits offset .. offset + n of
        the validity bitmap (LSB-first)."""
        n = self.length if n is None else int(n)
        if self.validity is None:
            return np.ones(n, bool)
        nbytes = (self.offset + n + 7) // 8
        bits = self.values.ctx.download(self.validity.ptr, nbytes, np.uint8)
        return unpack_validity(bits, self.offset, n)


def unpack_validity(bits, offset, n):
    """Rows [offset, offset + n) of an LSB-first validity bitmap as booleans."""
    return np.unpackbits(np.asarray(bits, np.uint8), bitorder="little")[offset:offset + n].astype(bool)


class Utf8Column:
    """An Arrow utf8 column on the device: int32 offsets, character bytes and
    an optional validity bitmap."""

    type_id = C.TYPE_UTF8

    def __init__(self, offsets, data, data_bytes, validity=None, length=None, offset=0):
        self.offsets = offsets
        self.data = data
        self.data_bytes = int(data_bytes)
        self.validity = validity
        self.length = offsets.length - 1 if length is None else int(length)
        self.offset = int(offset)

    @classmethod
    def from_strings(cls, ctx, strings):
        """strings: a sequence of str / bytes, None for a null row."""
        enc = [None if s is None else (s.encode() if isinstance(s, str) else bytes(s)) for s in strings]
        lens = np.array([0 if b is None else len(b) for b in enc], np.int64)
        offs = np.zeros(len(enc) + 1, np.int64)
        np.cumsum(lens, out=offs[1:])
        if offs[-1] > np.iinfo(np.int32).max:
            raise ValueError("more than 2^31 - 1 bytes in one utf8 column")
        data = np.frombuffer(b"".join(b for b in enc if b is not None) or b"\0", np.uint8)
        vb = None
        if any(b is None for b in enc):
            bits = np.packbits(np.array([b is not None for b in enc], bool), bitorder="little")
            pad = np.zeros(((len(enc) + 31) // 32) * 4, np.uint8)
            pad[:len(bits)] = bits
            vb = ctx.upload(pad)
        return cls(ctx.upload(offs.astype(np.int32)), ctx.upload(data), int(offs[-1]), vb, len(enc))

    @classmethod
    def empty(cls, ctx, n, data_bytes, nullable=False):
        offs = ctx.zeros(n + 1, np.int32)
        data = ctx.empty(max(1, int(data_bytes)), np.uint8)
        vb = ctx.zeros(((max(1, n) + 31) // 32) * 4, np.uint8) if nullable else None
        return cls(offs, data, int(data_bytes), vb, n)

    def c(self, length=None):
        return C.Column(C.TYPE_UTF8, 0, self.length if length is None else int(length), self.offset,
                        self.offsets.ptr, self.validity.ptr if self.validity is not None else None,
                        self.data.ptr, self.data_bytes)

    def valid_mask(self, n=None):
        n = self.length if n is None else int(n)
        if self.validity is None:
            return np.ones(n, bool)
        nbytes = (self.offset + n + 7) // 8
        bits = self.offsets.ctx.download(self.validity.ptr, nbytes, np.uint8)
        return unpack_validity(bits, self.offset, n)

    def to_list(self, n=None):
        """Rows [0, n) as str (None for null)."""
        n = self.length if n is None else int(n)
        ctx = self.offsets.ctx
        offs = ctx.download(self.offsets.ptr + 4 * self.offset, n + 1, np.int32).astype(np.int64)
        lo, hi = (int(offs[0]), int(offs[-1])) if n else (0, 0)
        raw = ctx.download(self.data.ptr + lo, hi - lo, np.uint8).tobytes() if hi > lo else b""
        valid = self.valid_mask(n)
        return [raw[offs[i] - lo:offs[i + 1] - lo].decode() if valid[i] else None for i in range(n)]
