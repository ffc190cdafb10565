"""tensordot through the drop-in names (diaq_gpu.tensordot on diaQMatrix, numpy arrays and DeviceMatrix operands)
against np.tensordot: values within 1e-12 (max-abs error over max-abs value), offsets equal to the oracle's
(src/tensordotKernels.cpp:21-74: permute, reshape, product with the zero-diagonal prune), the result's tensor shape,
and the reference's argument errors."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from oracle import oracle as O  # noqa: E402

TOL = 1e-12


@pytest.fixture(scope="module")
def d():
    import diaq_amd  # noqa: F401
    import diaq_gpu
    return diaq_gpu


def rel(a, b):
    b = np.asarray(b)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(np.asarray(a) - b)) / (den if den > 0 else 1.0))


CASES = [((40, 60), (60, 50), ([1], [0])), ((60, 40), (60, 50), ([0], [0])), ((6, 7, 8), (8, 7, 5), ([1, 2], [1, 0])),
         ((12, 16), (12, 16), ([0, 1], [0, 1])), ((4, 4, 4, 4), (4, 8), ([3], [0])), ((9, 11), (5, 7), ([], [])),
         ((2, 3, 4, 5), (5, 4, 3, 2), ([0, 1, 2, 3], [3, 2, 1, 0])), ((64, 64), (64, 2, 32), ([0], [0]))]


def tensor(shape, rng, fill):
    T = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    T[rng.random(shape) > fill] = 0
    return T


@pytest.mark.parametrize("fill", [1.0, 0.3, 0.02])
@pytest.mark.parametrize("sa,sb,axes", CASES)
def test_tensordot_vs_numpy(d, sa, sb, axes, fill):
    rng = np.random.default_rng(sum(sa) + 7 * sum(sb))
    A, B = tensor(sa, rng, fill), tensor(sb, rng, fill)
    want = np.tensordot(A, B, axes=axes)
    ref = O.tensordot(A, B, axes)
    free_shape = list(want.shape) if want.ndim else [1, 1]
    dense = want.reshape(ref.rows, ref.cols)
    for got in (d.tensordot(d.from_numpy(A), d.from_numpy(B), [list(axes[0]), list(axes[1])]),
                d.tensordot(A, B, [list(axes[0]), list(axes[1])]),
                d.tensordot(d.from_numpy(A), B, [list(axes[0]), list(axes[1])]),
                d.tensordot(A, d.from_numpy(B), [list(axes[0]), list(axes[1])])):
        assert list(got.shape) == free_shape
        assert list(got.dIndices) == ref.offsets
        assert rel(d.to_numpy(got), dense) < TOL
    # device-resident operands: the result stays on the device with its tensor shape
    D = d.tensordot(d.to_device(d.from_numpy(A)), d.to_device(d.from_numpy(B)), [list(axes[0]), list(axes[1])])
    assert list(D.diagonals) == ref.offsets
    H = D.to_host()
    assert list(H.shape) == free_shape
    assert rel(d.to_numpy(H), dense) < TOL


def test_tensordot_errors(d):
    A = d.from_numpy(np.ones((4, 6), complex))
    B = d.from_numpy(np.ones((6, 5), complex))
    with pytest.raises(ValueError, match="same length"):
        d.tensordot(A, B, [[1], [0, 1]])
    with pytest.raises(ValueError, match="out of range for A"):
        d.tensordot(A, B, [[2], [0]])
    with pytest.raises(ValueError, match="different sizes"):
        d.tensordot(A, B, [[0], [0]])
