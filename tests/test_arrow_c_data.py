"""The Arrow C Data Interface at the TupleSet boundary, against pyarrow's own
implementation of it: pyarrow exports a record batch (as the reference's
arrow::Table-based TupleSet would), the operator layer imports it and exports
it again, pyarrow imports the result, and the two batches must be equal --
types, names, nulls, offsets (sliced inputs) and utf8 included.  The host leg
runs on the CPU; the device leg goes through a device TupleSet on the GPU."""
import ctypes
import os

import numpy as np
import pytest

pa = pytest.importorskip("pyarrow")

HOOK = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libcdata_roundtrip.so")


class ArrowSchema(ctypes.Structure):
    pass


ArrowSchema._fields_ = [("format", ctypes.c_char_p), ("name", ctypes.c_char_p), ("metadata", ctypes.c_char_p),
                        ("flags", ctypes.c_int64), ("n_children", ctypes.c_int64),
                        ("children", ctypes.POINTER(ctypes.POINTER(ArrowSchema))),
                        ("dictionary", ctypes.POINTER(ArrowSchema)), ("release", ctypes.c_void_p),
                        ("private_data", ctypes.c_void_p)]


class ArrowArray(ctypes.Structure):
    pass


ArrowArray._fields_ = [("length", ctypes.c_int64), ("null_count", ctypes.c_int64), ("offset", ctypes.c_int64),
                       ("n_buffers", ctypes.c_int64), ("n_children", ctypes.c_int64),
                       ("buffers", ctypes.POINTER(ctypes.c_void_p)),
                       ("children", ctypes.POINTER(ctypes.POINTER(ArrowArray))),
                       ("dictionary", ctypes.POINTER(ArrowArray)), ("release", ctypes.c_void_p),
                       ("private_data", ctypes.c_void_p)]


def _hook():
    if not os.path.exists(HOOK):
        pytest.skip("tests/cpp/libcdata_roundtrip.so not built (python -m flexpushdowndb_amd.build)")
    return ctypes.CDLL(HOOK)


def _roundtrip(batch, fn_name):
    lib = _hook()
    fn = getattr(lib, fn_name)
    fn.restype = ctypes.c_int
    in_s, in_a, out_s, out_a = ArrowSchema(), ArrowArray(), ArrowSchema(), ArrowArray()
    batch._export_to_c(ctypes.addressof(in_a), ctypes.addressof(in_s))
    err = ctypes.create_string_buffer(512)
    rc = fn(ctypes.byref(in_s), ctypes.byref(in_a), ctypes.byref(out_s), ctypes.byref(out_a), err, 512)
    # the importer released what it consumed, as the C data interface requires
    assert not in_s.release and not in_a.release
    if rc != 0:
        raise RuntimeError(err.value.decode())
    return pa.RecordBatch._import_from_c(ctypes.addressof(out_a), ctypes.addressof(out_s))


def _batch(n=1000, seed=0):
    rng = np.random.default_rng(seed)
    null = rng.random(n) < 0.1
    strs = ["" if i % 17 == 0 else "vé" * int(i % 5) + str(i) for i in range(n)]
    return pa.record_batch({
        "i8": pa.array(rng.integers(-100, 100, n), pa.int8()),
        "u16": pa.array(rng.integers(0, 60000, n), pa.uint16()),
        "i32": pa.array(np.where(null, 0, rng.integers(-10**9, 10**9, n)), pa.int32(), mask=null),
        "i64": pa.array(rng.integers(-2**62, 2**62, n), pa.int64()),
        "u64": pa.array(rng.integers(0, 2**63, n, dtype=np.uint64), pa.uint64()),
        "f32": pa.array(rng.standard_normal(n).astype(np.float32), pa.float32()),
        "f64": pa.array(rng.standard_normal(n), pa.float64(), mask=rng.random(n) < 0.2),
        "d32": pa.array(rng.integers(0, 20000, n).astype(np.int32), pa.date32()),
        "d64": pa.array(rng.integers(0, 2 * 10**12, n), pa.date64()),
        "s": pa.array(strs, pa.utf8(), mask=rng.random(n) < 0.15),
    })


def test_host_roundtrip_every_type():
    b = _batch()
    out = _roundtrip(b, "cdata_host_roundtrip")
    assert out.schema.names == b.schema.names
    assert [f.type for f in out.schema] == [f.type for f in b.schema]
    assert out.equals(b)


def test_host_roundtrip_of_a_slice_and_empty_batch():
    b = _batch(777, seed=1)
    sl = b.slice(123, 400)          # array offsets inside the children
    assert _roundtrip(sl, "cdata_host_roundtrip").equals(sl)
    empty = b.slice(0, 0)
    assert _roundtrip(empty, "cdata_host_roundtrip").num_rows == 0


def test_large_utf8_imports_as_utf8():
    b = pa.record_batch({"s": pa.array(["a", None, "bcd", ""], pa.large_utf8())})
    out = _roundtrip(b, "cdata_host_roundtrip")
    assert out.column(0).type == pa.utf8()
    assert out.column(0).to_pylist() == ["a", None, "bcd", ""]


def test_unsupported_type_is_an_error():
    b = pa.record_batch({"t": pa.array([1.5, 2.5], pa.float16()) if hasattr(pa, "float16") else
                         pa.array([b"x"], pa.binary())})
    with pytest.raises(RuntimeError, match="not a supported column type"):
        _roundtrip(b, "cdata_host_roundtrip")


@pytest.mark.gpu
def test_device_roundtrip_every_type():
    """Through a device TupleSet (HBM columns) and back."""
    b = _batch(5000, seed=3)
    assert _roundtrip(b, "cdata_device_roundtrip").equals(b)
    sl = b.slice(1001, 2500)
    assert _roundtrip(sl, "cdata_device_roundtrip").equals(sl)
