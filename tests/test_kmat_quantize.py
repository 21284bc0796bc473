"""Element matrices of the reduced-precision contractions (host only, no GPU).

The FP32 apply uses each element type's K with every row on its own 24-bit
fixed-point grid and the rounding residual of every 27x27 component block
moved onto the row's entry at its own control point, so each block row sums to
exactly zero and the rigid translations stay in the null space.  Checked on
the oracle's element matrices (all types of two grids): the zero sums in exact
integer arithmetic, the distance to K, exactness in FP32 (and, for the 22-bit
grid the function also serves, as an FP16 hi + lo pair at the split apply's
scale), near-symmetry, and that entry-wise rounding alone does NOT keep the
sums (so the test would catch a quantizer that only rounds)."""
import numpy as np
import pytest

import oracle_py as O

icc = pytest.importorskip("immersed_catmull_clark_amd")
from immersed_catmull_clark_amd._abi import lib  # noqa: E402


def _types(nel):
    sp = O.Space([[0, 2.0], [0, 1.0], [0, 1.0]], list(nel), 2)
    types, _ = sp.elem_mats(*O.lame(1.0, 0.3))
    return types


def _quantize(k, bits):
    out = np.zeros((81, 81))
    dev = lib().icc_debug_quantize_kmat(np.ascontiguousarray(k).ctypes.data, bits, out.ctypes.data)
    return out, dev


@pytest.mark.parametrize("nel", [(8, 4, 4), (5, 3, 2)])
@pytest.mark.parametrize("bits", [24, 22])
def test_quantized_kmats_keep_translations_exact(nel, bits):
    for k in _types(nel):
        out, dev = _quantize(k, bits)
        q = np.ldexp(1.0, np.frexp(np.abs(k).max(axis=1))[1] - bits)[:, None]  # per-row grid
        X = out / q
        assert np.array_equal(X, np.round(X))
        assert np.abs(X).max() < 2.0 ** bits
        Xi = X.astype(np.int64)
        assert not Xi.reshape(81, 3, 27).sum(axis=2).any()  # block row sums exactly zero
        err = np.abs(X - k / q)
        assert dev == pytest.approx(err.max())
        assert dev <= 14.0  # at most the sum of a block's 27 half-quantum roundings
        moved = np.zeros((81, 81), dtype=bool)
        for i in range(81):
            moved[i, [27 * c + i % 27 for c in range(3)]] = True
        assert err[~moved].max() <= 0.5  # every other entry is the nearest grid point
        assert np.abs(out - out.T).max() <= 16 * q.max()  # symmetric up to a few quanta
        if bits == 24:
            assert np.all(out.astype(np.float32).astype(np.float64) == out)  # exact in FP32
        else:
            e = np.frexp(np.abs(k).max())[1] - 15  # the split apply's scale: max in [2^14, 2^15)
            v = np.ldexp(out, -e)
            hi = v.astype(np.float16).astype(np.float64)
            lo = v - hi
            assert np.all(lo.astype(np.float16).astype(np.float64) == lo)  # hi + lo exact
        naive = np.round(k / q).astype(np.int64).reshape(81, 3, 27)
        assert naive.sum(axis=2).any()


def test_quantize_rejects_bad_arguments():
    k = np.zeros((81, 81))
    out = np.ones((81, 81))
    assert lib().icc_debug_quantize_kmat(k.ctypes.data, 0, out.ctypes.data) == -1.0
    assert lib().icc_debug_quantize_kmat(k.ctypes.data, 22, out.ctypes.data) == 0.0
    assert not out.any()
