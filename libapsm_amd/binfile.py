"""Reader/writer of the reference's binary container format.

Format (cpp/external/binary/include/binary/binary_file.hpp:1-28,159-228): i32 key count;
per key an i32 length and the characters; then per key, in the same order, an i32
number of dimensions, the i32 dimensions and the float32 data in row-major order.
The writer emits keys in sorted order (it iterates a std::map).

Complex signal matrices are stored as [sample][antenna][2]
(cpp/cli/common/binary_load.cpp:55-113); load_complex / add_complex convert to and
from the library's [antenna][sample] complex layout.
"""
import struct

import numpy as np


class BinaryFileError(RuntimeError):
    pass


def read(path):
    try:
        with open(path, "rb") as f:
            data = f.read()
    except FileNotFoundError:
        raise BinaryFileError(f"Error loading the file {path}: File not found.")
    off = 0

    def i32():
        nonlocal off
        if off + 4 > len(data):
            raise BinaryFileError(f"Error loading the file {path}: File truncated.")
        (v,) = struct.unpack_from("<i", data, off)
        off += 4
        return v

    n = i32()
    if n < 0:
        raise BinaryFileError(f"Error loading the file {path}: bad key count")
    keys = []
    for _ in range(n):
        ln = i32()
        if ln < 0 or off + ln > len(data):
            raise BinaryFileError(f"Error loading the file {path}: File truncated.")
        keys.append(data[off:off + ln].decode())
        off += ln
    out = {}
    for k in keys:
        nd = i32()
        dims = [i32() for _ in range(nd)]
        cnt = int(np.prod(dims)) if dims else 1
        if off + 4 * cnt > len(data):
            raise BinaryFileError(f"Error loading the file {path}: File truncated.")
        out[k] = np.frombuffer(data, dtype="<f4", count=cnt, offset=off).reshape(dims).copy()
        off += 4 * cnt
    return out


def write(path, arrays):
    keys = sorted(arrays)
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(keys)))
        for k in keys:
            b = k.encode()
            f.write(struct.pack("<i", len(b)))
            f.write(b)
        for k in keys:
            a = np.ascontiguousarray(arrays[k], dtype="<f4")
            f.write(struct.pack("<i", a.ndim))
            for d in a.shape:
                f.write(struct.pack("<i", d))
            f.write(a.tobytes())


def to_complex(arr, scale=1.0):
    """[sample][antenna][2] float -> complex64 [antenna][sample] (times scale)."""
    c = (arr[..., 0] + 1j * arr[..., 1]).T * scale
    return np.ascontiguousarray(c.astype(np.complex64))


def from_complex(mat):
    """complex [antenna][sample] (or [sample]) -> [sample][antenna][2] float32."""
    m = np.asarray(mat)
    if m.ndim == 1:
        m = m[None, :]
    out = np.stack([m.real.T, m.imag.T], axis=-1)
    return np.ascontiguousarray(out.astype(np.float32))


def load_complex(path, key, scale=1.0):
    return to_complex(read(path)[key], scale)
