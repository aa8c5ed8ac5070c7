"""MAS checkpoint format: psum_amd/mas.py (the pypsum.MasFile API) and
include/psum/mas.hpp (tests/cpp/mas_io) against the byte-level restatement in
oracle/mas_format.py, plus the behaviours the reference's own tests check
(test/pypsum/test_mas.py:38-258).
"""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import mas_format as F  # noqa: E402

from psum_amd.mas import MasError, MasFile, load_node_field, load_particles, save_node_field, save_particles  # noqa: E402

MAS_IO = os.path.join(ROOT, "tests", "cpp", "mas_io")


def test_known_block_bytes(tmp_path):
    """one double block, spelled out from the documented layout"""
    m = MasFile(tmp_path / "a.mas")
    m.write("a", np.array([[1.5]]))
    raw = open(tmp_path / "a.mas", "rb").read()
    want = (b"#MAS!#" + struct.pack("<3i", 1, 6, 2) + b"a" + b"double" + struct.pack("<2Q", 1, 1)
            + struct.pack("<d", 1.5))
    assert raw == want
    assert raw == F.block("a", "double", [1, 1], [1.5])


CASES = [
    ("bool", np.array([True, False], dtype=np.bool_)),
    ("int8", np.array([1, -128, 127], dtype=np.int8)),
    ("uint8", np.array([0, 128, 255], dtype=np.uint8)),
    ("int16", np.array([1, -32768, 32767], dtype=np.int16)),
    ("uint16", np.array([0, 32768, 65535], dtype=np.uint16)),
    ("int32", np.array([1, -2147483648, 2147483647], dtype=np.int32)),
    ("uint32", np.array([0, 2147483648, 4294967295], dtype=np.uint32)),
    ("int64", np.array([1, -2 ** 63, 2 ** 63 - 1], dtype=np.int64)),
    ("uint64", np.array([0, 2 ** 63, 2 ** 64 - 1], dtype=np.uint64)),
    ("float32", np.array([1.0, -3.4e38, 3.4e38], dtype=np.float32)),
    ("float64", np.array([1.0, -1.7e308, 1.7e308], dtype=np.float64)),
    ("str", "test string"),
    ("empty", ""),
]


def test_type_roundtrip_and_oracle_parse(tmp_path):
    path = tmp_path / "types.mas"
    m = MasFile(path)
    for name, data in CASES:
        m.write(name, data)
    for name, data in CASES:
        got = m.get_data(name)
        if isinstance(data, str):
            assert got == data
        else:
            assert got.dtype == data.dtype and got.shape == data.shape
            np.testing.assert_array_equal(got, data)
    assert sorted(m.get_names()) == sorted(n for n, _ in CASES)
    # the restatement reads the same blocks, types and values
    blocks = F.parse(open(path, "rb").read())
    assert [b[0] for b in blocks] == [n for n, _ in CASES]
    tmap = {"bool": "logical", "float32": "single", "float64": "double", "str": "char", "empty": "char"}
    for (key, typ, shape, vals), (name, data) in zip(blocks, CASES):
        assert typ == tmap.get(name, name)
        if isinstance(data, str):
            assert vals == data.encode() and shape == [len(data.encode()), 1]
        else:
            assert shape == list(data.shape)
            np.testing.assert_array_equal(np.array(vals, dtype=data.dtype), data)


def test_metadata_access_delete_replace_clear(tmp_path):
    m = MasFile(tmp_path / "m.mas")
    m.write("arr", np.array([1, 2, 3], dtype=np.float64))
    m.write("scalar", 42)
    m.write("string", "test")
    assert set(m.get_names()) == {"arr", "scalar", "string"}
    info = {i["name"]: i for i in m.get_info()}
    assert info["arr"]["type"] == "double" and info["arr"]["shape"] == [3]
    assert info["scalar"]["shape"] == [1, 1]
    np.testing.assert_array_equal(m.get_data(0), [1, 2, 3])
    assert m.get_data(999) is None and m.get_data("nonexistent") is None
    with pytest.raises(MasError):
        m.write("arr", np.zeros(2))                      # duplicate name
    with pytest.raises(MasError):
        m.write("!x", np.zeros(2))
    with pytest.raises(MasError):
        m.write("", np.zeros(2))
    m.replace("arr", np.array([4.0, 5.0]))
    np.testing.assert_array_equal(m.get_data("arr"), [4.0, 5.0])
    # the replaced block is still in the file, masked with '!'
    keys = [b[0] for b in F.parse(open(tmp_path / "m.mas", "rb").read())]
    assert keys.count("!rr") == 1 and keys[-1] == "arr"
    for i in range(5):
        m.write(f"item{i}", i)
    assert m.delete("item2") and not m.delete("nonexistent")
    n = len(m.get_names())
    assert m.delete(0) and len(m.get_names()) == n - 1
    m.clear()
    assert m.get_names() == []
    m.write("new", np.array([1.0]))
    assert m.get_names() == ["new"]


def test_shapes_large_and_special_strings(tmp_path):
    m = MasFile(tmp_path / "s.mas")
    for i, shape in enumerate([(5,), (2, 3), (2, 2, 2), (1, 2, 3, 4)]):
        a = np.arange(np.prod(shape), dtype=np.float64).reshape(shape)
        m.write(f"arr{i}", a)
        np.testing.assert_array_equal(m.get_data(f"arr{i}"), a)
    big = np.arange(100 * 100, dtype=np.int32).reshape(100, 100)
    m.write("large_2d", big)
    np.testing.assert_array_equal(m.get_data("large_2d"), big)
    for i, s in enumerate(["hello world", "中文测试", "emoji 😊", "a" * 1000, ""]):
        m.write(f"str{i}", s)
        assert m.get_data(f"str{i}") == s


def test_corrupt_files_are_rejected(tmp_path):
    p = tmp_path / "bad.mas"
    open(p, "wb").write(b"#MAX!#" + b"\0" * 20)
    with pytest.raises(MasError):
        MasFile(p).get_names()
    good = F.block("a", "double", [4, 1], [1.0, 2.0, 3.0, 4.0])
    open(p, "wb").write(good[:-8])                      # truncated data
    with pytest.raises(MasError):
        MasFile(p).get_names()


def test_checkpoint_layout_helpers(tmp_path):
    m = MasFile(tmp_path / "c.mas")
    rng = np.random.default_rng(0)
    pos, vel = rng.random((17, 2)), rng.standard_normal((17, 2)).astype(np.float32)
    save_particles(m, "electrons", pos, vel)
    save_node_field(m, "phi", [0.5, 0.5], [7.5, 3.5], [7, 3], np.arange(32.0))
    info = {i["name"]: i for i in m.get_info()}
    assert info["electrons.data.content.position"]["shape"] == [17, 2, 1]
    assert info["electrons.data.content.velocity"]["type"] == "single"
    assert info["phi.grid.cells"]["type"] == "int32"
    p2, v2 = load_particles(m, "electrons")
    np.testing.assert_array_equal(p2, pos)
    np.testing.assert_array_equal(v2, vel)
    lo, up, cells, vals = load_node_field(m, "phi")
    np.testing.assert_array_equal(vals, np.arange(32.0))


@pytest.mark.skipif(not os.path.exists(MAS_IO), reason="tests/cpp/mas_io is built by build()")
def test_cpp_writer_matches_restatement_and_python(tmp_path):
    path = str(tmp_path / "cpp.mas")
    subprocess.run([MAS_IO, "write", path], check=True)
    blocks = F.parse(open(path, "rb").read())
    live = [b for b in blocks if not b[0].startswith("!")]
    assert [b[0] for b in live] == ["grid.lower", "single", "cells", "big", "flags", "label", "step",
                                   "replaced"]
    assert [b[0] for b in blocks if b[0].startswith("!")] == ["!asked", "!eplaced"]
    byname = {b[0]: b for b in live}
    assert byname["grid.lower"][2] == [2, 3] and byname["grid.lower"][3] == [1.5, -2.0, 3.25, 4.0, 5.5, -6.75]
    assert byname["big"][3] == [2 ** 64 - 1] and byname["step"][3] == [-42]
    assert byname["label"][3] == b"[#0]MI355X" and byname["replaced"][3] == [10.0, 11.0]
    m = MasFile(path)
    np.testing.assert_array_equal(m.get_data("cells"), np.array([[-7], [0], [2147483647]], np.int32))
    assert m.get_data("label") == "[#0]MI355X"


@pytest.mark.skipif(not os.path.exists(MAS_IO), reason="tests/cpp/mas_io is built by build()")
def test_cpp_reader_on_python_file(tmp_path):
    path = tmp_path / "py.mas"
    m = MasFile(path)
    m.write("x", np.array([[1.0, 2.0], [3.0, 4.5]]))
    m.write("gone", np.array([1], np.int64))
    m.write("n", np.array([[3]], np.int32))
    m.delete("gone")
    m.write("s", "ab")
    out = subprocess.run([MAS_IO, "read", str(path)], check=True, capture_output=True, text=True).stdout
    assert out.splitlines() == ["x double 2 2 | 10.5", "n int32 1 1 | 3", "s char 2 1 | 195"]
