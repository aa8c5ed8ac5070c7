"""Byte-level restatement of the MAS block layout, used only by the tests as
the checker of psum_amd/mas.py and include/psum/mas.hpp.  TEST INFRASTRUCTURE
ONLY: nothing in the product imports it.

Layout (src/serialization/mas.hpp:130-160 writeData, :166-228 readHead,
:230-256 smashData; docs/design_en.md:521-542):
    "#MAS!#" | int32 key_len | int32 type_len | int32 shape_len | key | type
    | size_t shape[shape_len] | prod(shape) elements of the type
MATLAB type names (foundation.hpp:91-113); a masked block's key starts '!'.
"""
import struct

ITEM = {"double": ("d", 8), "single": ("f", 4), "int64": ("q", 8), "uint64": ("Q", 8),
        "int32": ("i", 4), "uint32": ("I", 4), "int16": ("h", 2), "uint16": ("H", 2),
        "int8": ("b", 1), "uint8": ("B", 1), "logical": ("?", 1), "char": ("c", 1)}


def block(key, typ, shape, values):
    """bytes of one block; values: flat list of Python numbers (or a bytes
    object for char)"""
    k, t = key.encode(), typ.encode()
    out = b"#MAS!#" + struct.pack("<3i", len(k), len(t), len(shape)) + k + t
    out += b"".join(struct.pack("<Q", s) for s in shape)
    if typ == "char":
        out += bytes(values)
    else:
        out += struct.pack("<%d%s" % (len(values), ITEM[typ][0]), *values)
    return out


def parse(buf):
    """[(key, type, shape, flat values)] of every block, masked ones
    included (their key starts with '!')"""
    out, p = [], 0
    while p < len(buf):
        assert buf[p:p + 6] == b"#MAS!#", f"magic at {p}"
        kl, tl, sl = struct.unpack_from("<3i", buf, p + 6)
        q = p + 18
        key = buf[q:q + kl].decode()
        q += kl
        typ = buf[q:q + tl].decode()
        q += tl
        shape = list(struct.unpack_from("<%dQ" % sl, buf, q))
        q += 8 * sl
        n = 1
        for s in shape:
            n *= s
        code, size = ITEM[typ]
        raw = buf[q:q + n * size]
        vals = raw if typ == "char" else list(struct.unpack("<%d%s" % (n, code), raw))
        out.append((key, typ, shape, vals))
        p = q + n * size
    return out
