"""CPU-side checks of the C ABI: the HIP library loads, exports every entry
point include/fpdb_amd.h declares, and host-only entry points behave."""
import ctypes
import os
import random
import re

import numpy as np

from flexpushdowndb_amd import _capi as C
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(fpdb_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_every_declared_symbol():
    lib = C.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    bound = {n for n, _, _ in C.SIGNATURES}
    assert names == bound, (names ^ bound)


def test_host_entry_points():
    lib = C.lib()
    masks = np.zeros(136, np.uint8)
    assert lib.fpdb_bloom_masks(masks.ctypes.data) == 0
    assert np.array_equal(masks, O.bloom_masks())
    for n in [0, 1, 65, 400_000, 20_000_000]:
        assert lib.fpdb_bloom_blocked_log_num_blocks(n) == O.blocked_log_num_blocks(n)
    k, m = ctypes.c_int64(), ctypes.c_int64()
    assert lib.fpdb_bloom_legacy_params(400_000, 0.01, ctypes.byref(k), ctypes.byref(m)) == 0
    assert (k.value, m.value) == (O.legacy_k(0.01), O.legacy_m(400_000, 0.01))
    hf = (C.UniversalHash * 7)()
    assert lib.fpdb_bloom_legacy_make_hash_functions(99, m.value, 7, hf) == 0
    ref = O.legacy_params(99, m.value, 7)
    assert [[h.a, h.b, h.m, h.p] for h in hf] == ref.tolist()
    # errors come back as status codes with a message, never exceptions
    assert lib.fpdb_bloom_legacy_params(10, 0.0, ctypes.byref(k), ctypes.byref(m)) == C.FPDB_ERR_INVALID
    assert b"false positive" in lib.fpdb_last_error()


def test_invariant_division_formula():
    """The legacy filter's modulo uses division by invariant integers
    (Granlund & Montgomery 1994, Fig. 4.1); check the formula on big ints."""
    rnd = random.Random(11)
    M64 = 2 ** 64
    divisors = [1, 2, 3, 7, 64, 3_834_024, 3_834_031, 7_668_047, 2 ** 32 + 15, 2 ** 63 - 25, 2 ** 63 + 1,
                M64 - 59] + [rnd.randrange(1, M64) for _ in range(200)]
    for d in divisors:
        l = (d - 1).bit_length()            # ceil(log2 d)
        mp = (M64 * (2 ** l - d)) // d + 1
        assert mp < M64
        sh1, sh2 = min(l, 1), max(l - 1, 0)
        for n in [0, 1, d - 1, d, d + 1, M64 - 1] + [rnd.randrange(M64) for _ in range(50)]:
            n %= M64
            t1 = (mp * n) >> 64
            q = (t1 + ((n - t1) >> sh1)) >> sh2
            assert q == n // d
