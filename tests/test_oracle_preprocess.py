"""preprocessor::downsample_depth and the readers' depth conversion in the oracle (CPU).

Pinned by known answers that follow from the reference's code (src/map/preprocessor.cpp:19-84):
only depths >= 1e-5 count (NaN, 0 and negatives do not), an odd count takes the middle, an even
count takes region[n / 2] of the ascending order (the upper of the two middles, not their mean),
an all-invalid region gives 0 and maps to its first pixel, ratio 1 keeps valid depths and zeroes
the rest; and by an independent formulation (numpy sort of each region) on random images with
many equal depths, for ratios 1 to 5.  The uint16 conversions are checked for all 65536 values
against numpy's IEEE float32 division and multiplication.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _ds(d):
    d = np.asarray(d, np.float32)
    return O.downsample_depth(d, 1, 1)


def test_known_answers_2x2():
    assert _ds([[3, 1], [2, np.nan]]) == (2.0, 2)  # odd count: the middle, at pixel (0, 1)
    out, mp = _ds([[4, 1], [3, 2]])
    assert out[0, 0] == 3.0 and mp[0, 0] == 2  # even count: region[2] of [1, 2, 3, 4]
    out, mp = _ds([[0, -1], [np.nan, 9.99e-6]])
    assert out[0, 0] == 0.0 and mp[0, 0] == 0  # nothing valid
    out, mp = _ds([[1e-5, 0], [0, 0]])
    assert out[0, 0] == np.float32(1e-5) and mp[0, 0] == 0  # the threshold itself is valid
    out, mp = _ds([[2, 2], [1, 1]])
    assert out[0, 0] == 2.0 and mp[0, 0] == 0  # ties: the first of the equal depths in row-major order


def test_ratio_one_and_index_map():
    d = np.array([[1.5, np.nan, -2.0], [0.0, 7.0, 1e-6]], np.float32)
    out, mp = O.downsample_depth(d, 3, 2)
    np.testing.assert_array_equal(out, [[1.5, 0, 0], [0, 7.0, 0]])
    np.testing.assert_array_equal(mp, [[0, 1, 2], [3, 4, 5]])


def test_bad_sizes_are_rejected():
    d = np.zeros((6, 8), np.float32)
    for w, h in ((3, 3), (4, 2), (5, 3)):
        with pytest.raises(ValueError):
            O.downsample_depth(d, w, h)


def _numpy_downsample(d, r):
    h, w = d.shape[0] // r, d.shape[1] // r
    out = np.zeros((h, w), np.float32)
    mp = np.zeros((h, w), np.uint64)
    for y in range(h):
        for x in range(w):
            reg = d[y * r:(y + 1) * r, x * r:(x + 1) * r].reshape(-1)
            idx = np.nonzero(reg >= np.float32(1e-5))[0]
            if len(idx) == 0:
                mp[y, x] = x * r + y * r * d.shape[1]
                continue
            order = idx[np.argsort(reg[idx], kind="stable")]
            k = order[len(order) // 2]
            out[y, x] = reg[k]
            mp[y, x] = (x * r + k % r) + (y * r + k // r) * d.shape[1]
    return out, mp


@pytest.mark.parametrize("r", [1, 2, 3, 4, 5])
def test_random_images_against_numpy(r):
    rng = np.random.default_rng(r)
    h, w = 12 * r, 20 * r
    d = rng.integers(0, 6, (h, w)).astype(np.float32) * np.float32(0.5)  # many ties and zeros
    d[rng.random((h, w)) < 0.1] = np.nan
    d[rng.random((h, w)) < 0.05] = -1.0
    out, mp = O.downsample_depth(d, w // r, h // r)
    eo, em = _numpy_downsample(d, r)
    np.testing.assert_array_equal(out, eo)
    np.testing.assert_array_equal(mp, em)


def test_u16_conversions_all_values():
    raw = np.arange(65536, dtype=np.uint16)
    np.testing.assert_array_equal(O.depth_from_u16(raw, 1000.0, 1), raw.astype(np.float32) / np.float32(1000.0))
    inv = np.float32(1.0) / np.float32(5000.0)
    np.testing.assert_array_equal(O.depth_from_u16(raw, inv, 0), raw.astype(np.float32) * inv)
