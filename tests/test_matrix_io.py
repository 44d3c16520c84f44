"""The reference's dense matrix file formats (src/matrixIO.cpp) through the C
ABI, on the host: the binary format (int32 rows, int32 cols, row-major
doubles) checked byte for byte against files laid out independently with
struct/numpy, the text format ("rows cols", then "i j value" lines), round
trips, and the error statuses that replace the reference's exit()."""
import struct

import numpy as np
import pytest

from dense_hodlr_amd import HodlrError, readBinaryIntoMatrixXd, readTxtIntoMatrix, saveMatrixXdToBinary
from dense_hodlr_amd import inputs as I


def _layout(A):
    """The binary file of A as the reference writes it (matrixIO.cpp:165-182)."""
    return struct.pack("<ii", *A.shape) + np.ascontiguousarray(A, dtype="<f8").tobytes(order="C")


@pytest.mark.parametrize("shape", [(5, 3), (1, 7), (64, 64), (0, 4)])
def test_binary_read_matches_independent_layout(tmp_path, shape):
    A = I.SplitMix64(sum(shape) + 1).normal(shape[0] * shape[1]).reshape(shape)
    f = tmp_path / "m.bin"
    f.write_bytes(_layout(A))
    B = readBinaryIntoMatrixXd(f)
    assert B.shape == A.shape and np.array_equal(B, A)


def test_binary_write_is_the_reference_layout(tmp_path):
    A = np.asfortranarray(I.SplitMix64(9).normal(6 * 4).reshape(6, 4))
    A[2, 1] = -0.0
    A[0, 3] = np.inf
    f = tmp_path / "w.bin"
    saveMatrixXdToBinary(A, f)
    assert f.read_bytes() == _layout(A)
    # and back, bit for bit (the negative zero too)
    B = readBinaryIntoMatrixXd(f)
    assert np.array_equal(np.asfortranarray(B).view(np.int64), A.view(np.int64))


def test_text_format(tmp_path):
    A = np.zeros((4, 3))
    A[0, 0], A[3, 2], A[1, 2] = 1.5, -2.25, 1e-300
    f = tmp_path / "m.txt"
    lines = ["4 3"] + [f"{i} {j} {float(A[i, j])!r}" for i, j in ((0, 0), (3, 2), (1, 2))]
    f.write_text("\n".join(lines) + "\n")
    B = readTxtIntoMatrix(f)
    assert np.array_equal(B, A)


def test_errors(tmp_path):
    with pytest.raises(HodlrError) as e:
        readBinaryIntoMatrixXd(tmp_path / "missing.bin")
    assert e.value.code == -1 and "could not be opened" in str(e.value)
    f = tmp_path / "short.bin"
    f.write_bytes(_layout(np.ones((3, 3)))[:-8])
    with pytest.raises(HodlrError) as e:
        readBinaryIntoMatrixXd(f)
    assert e.value.code == -1 and "ends before" in str(e.value)
    g = tmp_path / "bad.txt"
    g.write_text("2 2\n0 0 1.0\n5 0 2.0\n")
    with pytest.raises(HodlrError) as e:
        readTxtIntoMatrix(g)
    assert "out of range" in str(e.value)
