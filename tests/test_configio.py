"""tetragono's configuration files (tnsp_amd/configio.py), restating
write_configurations / read_configurations (tetragono/tetragono/utility.py:
390-418): the header and body layout are checked byte by byte against that
code's layout, and a file is read back into the same chains."""
import os
import tempfile

import numpy as np
import pytest


def test_layout_and_round_trip():
    from tnsp_amd import configio

    rng = np.random.default_rng(0)
    cfg = rng.integers(0, 2, size=(5, 3, 4))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        configio.write_configurations(cfg, path)
        raw = np.fromfile(path, dtype="<i8")
        # head = [mpi_size, len(shape), *shape] with shape = [L1, L2, orbits]
        assert list(raw[:5]) == [5, 3, 3, 4, 1]
        # rank r's array at head.nbytes + r * config.nbytes
        for r in range(5):
            assert np.array_equal(raw[5 + 12 * r:5 + 12 * (r + 1)], cfg[r].ravel())
        back = configio.read_configurations(path)
        assert back.shape == (5, 3, 4, 1)
        assert np.array_equal(configio.occupations(back), cfg)


def test_bad_files_rejected():
    from tnsp_amd import configio

    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        np.array([3, 3, 2, 2, 1, 0, 1], dtype="<i8").tofile(path)
        with pytest.raises(ValueError):
            configio.read_configurations(path)
    with pytest.raises(ValueError):
        configio.occupations(np.array([[[[-2]]]]))


def test_trailing_data_of_a_longer_earlier_file_ignored():
    """tetragono's writer does not truncate: a file first written with 5
    configurations and then rewritten with 3 keeps the old tail; the reader
    takes only the 3 the header announces."""
    from tnsp_amd import configio

    rng = np.random.default_rng(1)
    first, second = rng.integers(0, 2, size=(5, 3, 4)), rng.integers(0, 2, size=(3, 3, 4))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "c.bin")
        configio.write_configurations(first, path)
        size = os.path.getsize(path)
        head = np.array([3, 3, 3, 4, 1], dtype="<i8").tobytes() + second.astype("<i8").tobytes()
        with open(path, "r+b") as f:  # overwrite in place, as MPI.File.Open(WRONLY | CREATE) does
            f.write(head)
        assert os.path.getsize(path) == size
        back = configio.read_configurations(path)
        assert back.shape == (3, 3, 4, 1)
        assert np.array_equal(configio.occupations(back), second)
