"""Host-side helpers of the device layer that need no GPU."""
import numpy as np

from flexpushdowndb_amd.device import unpack_validity


def test_unpack_validity_honours_offset():
    rng = np.random.default_rng(5)
    valid = rng.random(203) < 0.5
    bits = np.packbits(valid, bitorder="little")
    for off in (0, 1, 7, 8, 13, 64, 100):
        n = len(valid) - off
        assert np.array_equal(unpack_validity(bits, off, n), valid[off:])
    # a slice that ends before the bitmap does
    assert np.array_equal(unpack_validity(bits, 13, 20), valid[13:33])
