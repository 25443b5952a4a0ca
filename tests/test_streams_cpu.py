"""FileStream (src/streams/file/core.rs) host side: the serialized format,
mapping and validation; runs without a GPU."""
import numpy as np
import pytest

import oracle as O
from effsc import _native as N
from effsc.streams import FileStream


def test_write_open_read_roundtrip(tmp_path):
    v = 10
    data = O.uniform(1 << v, 3)
    path = tmp_path / "evals.bin"
    FileStream.write_to_file(path, data)
    # ark-serialize uncompressed: 8 little-endian bytes per base element
    raw = path.read_bytes()
    assert len(raw) == 8 << v
    assert int.from_bytes(raw[8:16], "little") == int(data[1])
    with FileStream(path) as s:
        assert s.num_variables() == v
        assert np.array_equal(s.read(0, 1 << v), data)
        assert s.evaluation(777) == int(data[777])
        with pytest.raises(N.NativeError):
            s.read(1 << v, 1)


def test_ext3_elements(tmp_path):
    data = O.uniform(3 * 64, 4)
    path = tmp_path / "ext.bin"
    FileStream.write_to_file(path, data, ext=3)
    with FileStream(path, ext=3) as s:
        assert s.num_variables() == 6
        assert s.evaluation(5) == tuple(int(x) for x in data[15:18])


def test_rejects_non_power_of_two_and_non_canonical(tmp_path):
    path = tmp_path / "bad.bin"
    FileStream.write_to_file(path, O.uniform(12, 1))
    with pytest.raises(N.NativeError):
        FileStream(path)  # file/core.rs:34-36 asserts a power-of-two length
    d = O.uniform(16, 2)
    d[3] = O.P  # not canonical: the reference's deserialize rejects it
    FileStream.write_to_file(path, d)
    with FileStream(path) as s:
        assert s.evaluation(2) == int(d[2])
        with pytest.raises(N.NativeError):
            s.evaluation(3)
