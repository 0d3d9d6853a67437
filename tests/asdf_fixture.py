"""Minimal ASDF 1.x writer for test fixtures: a YAML tree followed by
binary blocks laid out byte by byte as the ASDF standard specifies (magic
b'\\xd3BLK', big-endian uint16 header size, then flags uint32, 4-byte
compression, allocated / used / data sizes uint64 and an MD5 checksum)."""
import bz2
import hashlib
import struct
import zlib

import numpy as np


def block(data: bytes, compression=b"\0\0\0\0", pad=0):
    if compression == b"zlib":
        stored = zlib.compress(data)
    elif compression == b"bzp2":
        stored = bz2.compress(data)
    else:
        stored = data
    header = struct.pack(">I4sQQQ", 0, compression, len(stored) + pad, len(stored), len(data))
    header += hashlib.md5(data).digest()
    return b"\xd3BLK" + struct.pack(">H", len(header)) + header + stored + b"\0" * pad


def ndarray_yaml(source, a, indent):
    sp = " " * indent
    dt = {np.dtype("float32"): "float32", np.dtype("float64"): "float64"}[a.dtype]
    return (f"!core/ndarray-1.0.0\n{sp}source: {source}\n{sp}datatype: {dt}\n{sp}byteorder: little\n"
            f"{sp}shape: {list(a.shape)}\n")


def epsf_file(path, data, pixel_x, pixel_y, spectral_type=("A0V", "G2V"), defocus=(0,), oversample=2,
              compression=b"\0\0\0\0", roman_key="roman", subtype="psf"):
    """A Roman-CRDS-like ePSF file: data is row-major
    (defocus, spectral_type, grid_index, y, x) float32."""
    a = np.ascontiguousarray(data, dtype=np.float32)
    text = ("#ASDF 1.0.0\n#ASDF_STANDARD 1.5.0\n%YAML 1.1\n%TAG ! tag:stsci.edu:asdf/\n"
            "--- !core/asdf-1.1.0\n"
            "asdf_library: !core/software-1.0.0 {name: fixture, version: 0.0.1}\n"
            f"{roman_key}: !<asdf://stsci.edu/datamodels/roman/tags/reference_files/epsf-1.0.0>\n"
            "  meta:\n"
            f"    pixel_x: {[float(v) for v in pixel_x]}\n"
            f"    pixel_y: {[float(v) for v in pixel_y]}\n"
            f"    oversample: {oversample}\n"
            f"    spectral_type: [{', '.join(spectral_type)}]\n"
            f"    defocus: {list(defocus)}\n"
            f"  {subtype}: " + ndarray_yaml(0, a, 4) + "...\n")
    with open(path, "wb") as f:
        f.write(text.encode("utf-8"))
        f.write(block(a.tobytes(), compression, pad=13))
    return path
