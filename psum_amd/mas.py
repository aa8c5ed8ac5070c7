"""MAS (multi-array set) files: the reference's checkpoint and fixture format
(src/serialization/mas.hpp:49-453; pypsum.MasFile, pypsum/serialization/
mas_file_py.cpp:15-275), read and written with numpy.

A file is a sequence of blocks, each

    b"#MAS!#"  int32 key_len  int32 type_len  int32 shape_len
    key  type  uint64[shape_len] shape  data (prod(shape) elements, C order)

in native (little-endian) byte order.  Type names are MATLAB's (double,
single, int8..uint64, logical, char).  Deleting a block overwrites the first
byte of its key with '!' in place; scans skip such blocks.  Names may not be
empty or start with '!'; writing a name that already exists is an error
(replace deletes first).
"""
import os
import struct

import numpy as np

MAGIC = b"#MAS!#"
_HEAD = struct.Struct("<iii")

# MATLAB type name <-> numpy dtype
_TYPES = {
    "double": np.dtype(np.float64), "single": np.dtype(np.float32),
    "int64": np.dtype(np.int64), "uint64": np.dtype(np.uint64),
    "int32": np.dtype(np.int32), "uint32": np.dtype(np.uint32),
    "int16": np.dtype(np.int16), "uint16": np.dtype(np.uint16),
    "int8": np.dtype(np.int8), "uint8": np.dtype(np.uint8),
    "logical": np.dtype(np.bool_), "char": np.dtype(np.uint8),
}
_NAMES = {v: k for k, v in _TYPES.items() if k != "char"}


class MasError(RuntimeError):
    pass


class MasFile:
    """pypsum.MasFile: get_names, get_info, get_data(name | index), write,
    replace, delete(name | index), clear.  The file is created when absent."""

    def __init__(self, path):
        self.path = os.fspath(path)
        parent = os.path.dirname(self.path)
        if parent and not os.path.isdir(parent):
            os.makedirs(parent, exist_ok=True)
        if not os.path.exists(self.path):
            open(self.path, "wb").close()

    # -- scanning
    def _heads(self):
        """[(name, type, shape, offset of the block, offset of the data)],
        masked blocks excluded"""
        out = []
        with open(self.path, "rb") as f:
            size = os.fstat(f.fileno()).st_size
            pos = 0
            while pos < size:
                f.seek(pos)
                if f.read(len(MAGIC)) != MAGIC:
                    raise MasError(f"{self.path}: magic string mismatch at byte {pos}")
                raw = f.read(_HEAD.size)
                if len(raw) != _HEAD.size:
                    raise MasError(f"{self.path}: truncated block header at byte {pos}")
                kl, tl, sl = _HEAD.unpack(raw)
                key = f.read(kl).decode("utf-8", errors="replace")
                typ = f.read(tl).decode("ascii", errors="replace")
                shape = list(struct.unpack(f"<{sl}Q", f.read(8 * sl))) if sl else []
                if typ not in _TYPES:
                    raise MasError(f"{self.path}: unknown block type {typ!r}")
                n = int(np.prod(shape, dtype=np.int64)) if shape else 1
                data_at = pos + len(MAGIC) + _HEAD.size + kl + tl + 8 * sl
                end = data_at + n * _TYPES[typ].itemsize
                if end > size:
                    raise MasError(f"{self.path}: incomplete block {key!r}")
                if not key.startswith("!"):
                    out.append((key, typ, shape, pos, data_at))
                pos = end
        return out

    def get_names(self):
        return [h[0] for h in self._heads()]

    def get_info(self):
        return [{"name": h[0], "type": h[1], "shape": list(h[2])} for h in self._heads()]

    def _find(self, key):
        if isinstance(key, (int, np.integer)) and not isinstance(key, bool):
            heads = self._heads()
            return heads[key] if 0 <= key < len(heads) else None
        for h in self._heads():
            if h[0] == key:
                return h
        return None

    def get_data(self, key):
        """numpy array with the stored shape; str for char blocks; None when
        the name or index does not exist"""
        h = self._find(key)
        if h is None:
            return None
        name, typ, shape, _, data_at = h
        dt = _TYPES[typ]
        n = int(np.prod(shape, dtype=np.int64)) if shape else 1
        with open(self.path, "rb") as f:
            f.seek(data_at)
            buf = f.read(n * dt.itemsize)
        if typ == "char":
            return buf.decode("utf-8", errors="surrogateescape")
        return np.frombuffer(buf, dtype=dt).reshape(shape).copy()

    # -- writing
    @staticmethod
    def _encode(data):
        if isinstance(data, str):
            b = data.encode("utf-8", errors="surrogateescape")
            return "char", [len(b), 1], b
        if isinstance(data, (bool, np.bool_)):
            return "logical", [1, 1], np.array([data], np.bool_).tobytes()
        if isinstance(data, (int, np.integer)) and not isinstance(data, np.ndarray):
            a = np.array([data], np.int64 if isinstance(data, int) else np.asarray(data).dtype)
            return _NAMES[a.dtype], [1, 1], a.tobytes()
        if isinstance(data, (float, np.floating)) and not isinstance(data, np.ndarray):
            a = np.array([data], np.float64 if isinstance(data, float) else np.asarray(data).dtype)
            return _NAMES[a.dtype], [1, 1], a.tobytes()
        a = np.asarray(data)
        if a.ndim == 0:
            raise MasError("cannot write a 0-dimensional array; write the scalar instead")
        if a.dtype not in _NAMES:
            raise MasError(f"unsupported dtype {a.dtype}")
        return _NAMES[a.dtype], list(a.shape), np.ascontiguousarray(a).tobytes()

    def write(self, name, data):
        if not name:
            raise MasError("key can't be empty")
        if name.startswith("!"):
            raise MasError("key can't start with '!'")
        if self._find(name) is not None:
            raise MasError(f"block already exists: {name}")
        typ, shape, payload = self._encode(data)
        kb, tb = name.encode("utf-8"), typ.encode("ascii")
        with open(self.path, "ab") as f:
            f.write(MAGIC)
            f.write(_HEAD.pack(len(kb), len(tb), len(shape)))
            f.write(kb)
            f.write(tb)
            f.write(struct.pack(f"<{len(shape)}Q", *shape))
            f.write(payload)

    def delete(self, key):
        """mask every block of that name (or the block at that index)"""
        if isinstance(key, (int, np.integer)) and not isinstance(key, bool):
            h = self._find(key)
            if h is None:
                return False
            key = h[0]
        found = False
        with open(self.path, "r+b") as f:
            for name, _, _, pos, _ in self._heads():
                if name == key:
                    f.seek(pos + len(MAGIC) + _HEAD.size)
                    f.write(b"!")
                    found = True
        return found

    def replace(self, name, data):
        self.delete(name)
        self.write(name, data)

    def clear(self):
        open(self.path, "wb").close()


# ---- checkpoint layout of a particle group and a node field, as the
# reference's save() overloads name them (particle_group.hpp:272-275 ->
# device_vector.hpp:294-309 -> tag_sl.hpp:62-98 -> container_sl.hpp:96-109;
# device_field.hpp:327-331, simple_grid.hpp:429-433, device_array.hpp:189-203)

DEVICE_TAG = "[#0]AMD Instinct MI355X"


def save_particles(mas, name, position, velocity):
    """position, velocity: (n, D) arrays -> name.data.device and the
    tag-wise blocks name.data.content.{position,velocity} of shape (n, D, 1)"""
    mas.write(name + ".data.device", DEVICE_TAG)
    for tag, a in (("position", position), ("velocity", velocity)):
        a = np.asarray(a)
        mas.write(f"{name}.data.content.{tag}", a.reshape(a.shape[0], a.shape[1], 1))


def load_particles(mas, name):
    pos = mas.get_data(name + ".data.content.position")
    vel = mas.get_data(name + ".data.content.velocity")
    if pos is None or vel is None:
        raise MasError(f"no particle group {name!r}")
    return pos[:, :, 0], vel[:, :, 0]


def save_node_field(mas, name, lower, upper, cells, values):
    mas.write(name + ".location", np.int32(0))              # var_loc::nodeCentered
    mas.write(name + ".grid.lower", np.asarray(lower, np.float64).reshape(-1, 1))
    mas.write(name + ".grid.upper", np.asarray(upper, np.float64).reshape(-1, 1))
    mas.write(name + ".grid.cells", np.asarray(cells, np.int32).reshape(-1, 1))
    mas.write(name + ".content.device", DEVICE_TAG)
    mas.write(name + ".content.content", np.asarray(values, np.float64).reshape(-1, 1))


def load_node_field(mas, name):
    loc = mas.get_data(name + ".location")
    if loc is None or int(np.asarray(loc).ravel()[0]) != 0:
        raise MasError(f"{name!r} is not a node field")
    cells = mas.get_data(name + ".grid.cells").ravel()
    vals = mas.get_data(name + ".content.content").ravel()
    if vals.size != int(np.prod(cells + 1)):
        raise MasError(f"{name!r}: content size does not match its grid")
    return (mas.get_data(name + ".grid.lower").ravel(), mas.get_data(name + ".grid.upper").ravel(),
            cells, vals)
