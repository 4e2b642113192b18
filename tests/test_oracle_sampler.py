"""The oracle's sampler against what pins it: rocRAND's JOEKUO6 table, the hash's structure,
Sobol stratification (Tools/verify_sampling.cu:16-25) and the reference's golden renders."""
import ctypes as c

import numpy as np
import pytest

import oracle

MASK = 0xFFFFFFFF


def test_direction_table_matches_rocrand_joekuo6():
    lib = c.CDLL("/opt/rocm/lib/librocrand.so")
    ptr = c.POINTER(c.c_uint)()
    assert lib.rocrand_get_direction_vectors32(c.byref(ptr), 101) == 0   # ROCRAND_DIRECTION_VECTORS_32_JOEKUO6
    rocrand = np.ctypeslib.as_array(ptr, shape=(260 * 32,)).reshape(260, 32)
    assert np.array_equal(oracle.sobol_table(260), rocrand)


def test_direction_table_van_der_corput_and_first_dims():
    t = oracle.sobol_table(3)
    assert [int(x) for x in t[0, :4]] == [1 << 31, 1 << 30, 1 << 29, 1 << 28]
    # Dimension 1: primitive polynomial x + 1 with m1 = 1 gives v_k = v_{k-1} ^ (v_{k-1} >> 1).
    v = [1 << 31]
    for _ in range(31):
        v.append(v[-1] ^ (v[-1] >> 1))
    assert [int(x) for x in t[1]] == v


def _unhash(y):
    """Inverse of the reference hash, built from modular inverses of its multipliers."""
    def unxorshift(x, s):
        r = x
        for _ in range(32 // s + 1):
            r = x ^ (r >> s)
        return r & MASK
    inv1, inv2 = pow(0x7FEB352D, -1, 1 << 32), pow(0x846CA68B, -1, 1 << 32)
    x = unxorshift(y, 16)
    x = (x * inv2) & MASK
    x = unxorshift(x, 15)
    x = (x * inv1) & MASK
    return unxorshift(x, 16)


def test_hash_structure():
    assert oracle.hash32(0) == 0
    rng = np.random.default_rng(3)
    for x in [1, 2, 0x9E3779B9, MASK] + [int(v) for v in rng.integers(0, 1 << 32, 200)]:
        assert _unhash(oracle.hash32(x)) == x


@pytest.mark.parametrize("pixel", [0, 177])
@pytest.mark.parametrize("seed", [42, 1234])
def test_stratification(pixel, seed):
    for count in (16, 64, 256):
        vals = oracle.sobol_batch(np.full(count, pixel), np.arange(count), seed, 260)
        assert (vals >= 0).all() and (vals < 1).all()
        strata = np.floor(vals * count).astype(np.int64)
        for d in range(260):
            assert len(np.unique(strata[:, d])) == count, (d, count)


def test_sample_value_closed_form():
    # sample 0 has no Gray-code term: value = hash(pixel ^ hash(seed + dim*golden)) >> 8 * 2^-24.
    for pixel, seed, dim in [(0, 42, 0), (5, 42, 7), (177, 1234, 259)]:
        shift = oracle.hash32((seed + dim * 0x9E3779B9) & MASK)
        expect = np.float32(oracle.hash32(pixel ^ shift) >> 8) * np.float32(2.0 ** -24)
        assert oracle.sobol_batch([pixel], [0], seed, dim + 1)[0, dim] == expect
