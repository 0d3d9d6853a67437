"""Roman CRDS ePSF reference files (ASDF) -> GriddedPSFModel.

  roman_crds_gridded_epsf   src/psf/roman_crds.jl:1-27, ext/CrowdPhotASDFExt.jl:7-73
  read_asdf                 the subset of the ASDF 1.x file format those files use

An ASDF file is a YAML tree followed by binary blocks.  Each block starts
with the magic b"\\xd3BLK", a big-endian uint16 header size and a header of
flags (uint32), a 4-byte compression code (b"\\0\\0\\0\\0", b"zlib", b"bzp2"),
allocated, used and data sizes (uint64) and a 16-byte checksum; the stored
(used) bytes follow, padded to the allocated size.  Arrays in the tree are
`ndarray` nodes naming their block (`source`), `datatype`, `byteorder`,
row-major `shape` and optional `offset` / `strides`.  The YAML is read with
PyYAML's safe loader, extended only to turn the files' custom tags into plain
mappings, sequences and scalars, so nothing in a file is executed.
"""
import bz2
import struct
import zlib

import numpy as np
import yaml

from .photometry import ArgumentError
from .psf import GriddedPSFModel, ImagePSF, pixel_response_kernel

BLOCK_MAGIC = b"\xd3BLK"
_DTYPES = {"int8": "i1", "uint8": "u1", "int16": "i2", "uint16": "u2", "int32": "i4", "uint32": "u4",
           "int64": "i8", "uint64": "u8", "float32": "f4", "float64": "f8", "bool8": "?"}


class _NDArray(dict):
    """An ndarray node of the tree, resolved against the blocks after parsing."""


class _Loader(yaml.SafeLoader):
    pass


def _construct_tagged(loader, tag_suffix, node):
    if isinstance(node, yaml.MappingNode):
        value = loader.construct_mapping(node, deep=True)
        if "ndarray" in tag_suffix:
            return _NDArray(value)
        return value
    if isinstance(node, yaml.SequenceNode):
        return loader.construct_sequence(node, deep=True)
    return loader.construct_scalar(node)


for _prefix in ("!", "tag:", "asdf://", "http://", "https://"):
    _Loader.add_multi_constructor(_prefix, _construct_tagged)


def _split(raw):
    """YAML text and the byte offset where the block section starts."""
    if not raw.startswith(b"#ASDF"):
        raise ArgumentError("not an ASDF file (missing the #ASDF header)")
    end = raw.find(b"\n...\n")
    if end < 0:
        end = raw.find(b"\n...\r\n")
    if end < 0:   # a tree without blocks may end at EOF
        return raw.decode("utf-8"), len(raw)
    return raw[: end + 5].decode("utf-8"), end + 5


def _blocks(raw, pos):
    out = []
    while True:
        k = raw.find(BLOCK_MAGIC, pos)
        if k < 0:
            return out
        hsize = struct.unpack(">H", raw[k + 4:k + 6])[0]
        h = raw[k + 6:k + 6 + hsize]
        if len(h) < 48:
            raise ArgumentError("truncated ASDF block header")
        flags, comp, alloc, used, dsize = struct.unpack(">I4sQQQ", h[:32])
        start = k + 6 + hsize
        data = raw[start:start + used]
        if comp == b"\0\0\0\0":
            pass
        elif comp == b"zlib":
            data = zlib.decompress(data)
        elif comp == b"bzp2":
            data = bz2.decompress(data)
        else:
            raise ArgumentError(f"unsupported ASDF block compression {comp!r}")
        if len(data) != dsize:
            raise ArgumentError("ASDF block size does not match its header")
        out.append(data)
        pos = start + alloc


def _resolve(node, blocks):
    if isinstance(node, _NDArray):
        if "data" in node and "source" not in node:
            return np.array(node["data"])
        src = node.get("source")
        if not isinstance(src, int) or not (0 <= src < len(blocks)):
            raise ArgumentError(f"unsupported or missing ndarray source {src!r}")
        dt = node["datatype"]
        if not isinstance(dt, str) or dt not in _DTYPES:
            raise ArgumentError(f"unsupported ndarray datatype {dt!r}")
        order = "<" if node.get("byteorder", "little") == "little" else ">"
        dtype = np.dtype(order + _DTYPES[dt])
        shape = [int(v) for v in node["shape"]]
        buf = np.frombuffer(blocks[src], dtype=np.uint8)
        off = int(node.get("offset", 0))
        strides = node.get("strides")
        if strides is None:
            n = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(blocks[src], dtype=dtype, count=n, offset=off).reshape(shape)
        else:
            a = np.lib.stride_tricks.as_strided(buf[off:].view(dtype), shape=shape,
                                                strides=[int(s) for s in strides])
        return a.astype(dtype.newbyteorder("="), copy=True)
    if isinstance(node, dict):
        return {k: _resolve(v, blocks) for k, v in node.items()}
    if isinstance(node, list):
        return [_resolve(v, blocks) for v in node]
    return node


def read_asdf(path):
    """The tree of an ASDF file with its ndarray nodes as numpy arrays
    (row-major, the file's shape)."""
    with open(path, "rb") as f:
        raw = f.read()
    text, pos = _split(raw)
    body = text[text.find("%YAML"):] if "%YAML" in text else text
    tree = yaml.load(body, Loader=_Loader)
    return _resolve(tree if tree is not None else {}, _blocks(raw, pos))


def roman_crds_gridded_epsf(path, defocus=0, spectral_type="G2V", psf_subtype="psf", origin=None,
                            normalize=False, ctx=None):
    """A Roman CRDS ePSF reference file as a GriddedPSFModel of ImagePSF
    nodes (ext/CrowdPhotASDFExt.jl:7-73): the (defocus, spectral_type) slice
    of the psf (or psf_noipc) stamps, one node per meta.pixel_x / pixel_y,
    oversampling from meta.oversample.  Stamps in the optical-PSF-only
    convention (median sum below oversample^2 / 2) are correlated with the
    pixel-response kernel (zero border, on the GPU) and scaled by
    oversample^2, as romancal does."""
    if psf_subtype not in ("psf", "psf_noipc"):
        raise ArgumentError(
            f'psf_subtype must be "psf" or "psf_noipc" (got {psf_subtype!r}); "extended_psf"/"extended_psf_noipc" '
            "are single non-gridded stamps, not supported by GriddedPSFModel -- build an ImagePSF directly "
            "instead.")
    tree = read_asdf(path)
    if not isinstance(tree, dict) or "roman" not in tree:
        raise ArgumentError(f'{path} does not contain a top-level "roman" key; is this a Roman CRDS reference file?')
    afr = tree["roman"]
    if "meta" not in afr:
        raise ArgumentError(f'{path} does not contain a "roman.meta" key')
    meta = afr["meta"]
    if psf_subtype not in afr:
        raise ArgumentError(f'{path} does not contain "{psf_subtype}"; available: {list(afr.keys())}')
    data = np.asarray(afr[psf_subtype])
    if data.ndim != 5:
        raise ArgumentError(f'{path}\'s "{psf_subtype}" array has {data.ndim} dimensions, expected 5 (defocus, '
                            "spectral_type, grid_index, y, x); is this a gridded (not extended) PSF array?")
    pixel_x, pixel_y = list(meta["pixel_x"]), list(meta["pixel_y"])
    if len(pixel_x) != len(pixel_y):
        raise ArgumentError(f"pixel_x and pixel_y have different lengths in reference file {path}")
    spectral_types, defocus_values = list(meta["spectral_type"]), list(meta["defocus"])
    os_ = int(meta["oversample"])
    if os_ <= 0:
        raise ArgumentError(f"{path} has non-positive meta.oversample={os_}")
    if defocus not in defocus_values:
        raise ArgumentError(f"defocus={defocus} not found in {path}; available: {defocus_values}")
    if spectral_type not in spectral_types:
        raise ArgumentError(f"spectral_type={spectral_type!r} not found in {path}; available: {spectral_types}")
    di, si = defocus_values.index(defocus), spectral_types.index(spectral_type)
    n_grid = len(pixel_x)
    if data.shape[2] != n_grid:
        raise ArgumentError(f'{path}\'s "{psf_subtype}" array has {data.shape[2]} grid nodes but '
                            f"meta.pixel_x/pixel_y list {n_grid}")
    stamps = [np.array(data[di, si, i]) for i in range(n_grid)]   # rows y, columns x
    if np.median([s.sum(dtype=np.float64) for s in stamps]) < os_ ** 2 / 2:
        from .detection import correlate
        kern = pixel_response_kernel(os_).astype(stamps[0].dtype)
        stamps = [(np.float32(os_ ** 2) * correlate(s, kern, "zero", ctx=ctx)).astype(s.dtype) for s in stamps]
    psfs = [ImagePSF(s, origin=origin, oversampling=os_, normalize=normalize) for s in stamps]
    return GriddedPSFModel(psfs, [float(v) for v in pixel_y], [float(v) for v in pixel_x])
