"""The synthetic-input generator is bit-reproducible (CPU only)."""
import hashlib

import numpy as np

from jvsip_amd import datagen


def test_splitmix64_known_values():
    # published splitmix64 outputs for seed 0
    assert list(datagen.splitmix64(0, 0, 3)) == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_stream_hashes_pinned():
    assert hashlib.sha256(datagen.splitmix64(1, 0, 4096).tobytes()).hexdigest() == \
        "87e1cb757476e8485b44378c7678a3c25e79e55a9c62f5b5f61a4de892577697"
    assert hashlib.sha256(datagen.normal(1, 4096).tobytes()).hexdigest() == \
        "8a1fbb703f53e336526ad36e855837f49748db1648c1b88a6c59b9160c79b87b"
    assert hashlib.sha256(datagen.uniform(2, 4096).tobytes()).hexdigest() == \
        "336cadaceb19bc1b483c84c8a7952549c6e5ee0f876019e5c95f0e64d7b05be3"


def test_offsets_and_moments():
    a = datagen.normal(7, 1000)
    b = datagen.normal(7, 500, start=500)
    np.testing.assert_array_equal(a[500:], b)
    x = datagen.normal(3, 200000).astype(np.float64)
    assert abs(x.mean()) < 0.01 and abs(x.std() - 1) < 0.01
    u = datagen.uniform(3, 200000).astype(np.float64)
    assert u.min() >= -1 and u.max() < 1 and abs(u.mean()) < 0.01
    c = datagen.complex_data(1, (4, 8))
    r = datagen.normal(1, 64)
    np.testing.assert_array_equal(c.ravel().real, r[0::2])
    np.testing.assert_array_equal(c.ravel().imag, r[1::2])
