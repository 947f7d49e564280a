"""MRC / MRCS image stacks, the RELION particle format the reference opens with ``mrcfile``
(``getDistanceCTF_local_Conj9combinedS2.py:259-262``: ``mrcfile.mmap(...).data[indiS]``),
read here with numpy.

Layout (MRC2014): a 1024-byte main header of 32-bit words -- nx, ny, nz (words 0-2), mode
(word 3), nsymbt = bytes of extended header (word 23), 'MAP ' (word 52), machine stamp (word
53: 0x44 0x4? little-endian, 0x11 0x11 big-endian) -- then the extended header, then nz images of
ny rows of nx values, x fastest.  ``read_images`` returns images in mrcfile's ``data[z]``
orientation, (ny, nx) with row y.  Only mode 2 (float32), the mode RELION writes particles in,
is read: the reference feeds the native dtype to ``ndimage.shift``, which for integer modes
rounds its output to that dtype, a path this reader does not reproduce (it raises).
"""
from __future__ import annotations

import struct

import numpy as np

HEADER_BYTES = 1024
MODE_DTYPES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}


def header(path):
    """Main-header fields of an MRC file: nx, ny, nz, mode, nsymbt, data offset, byte order."""
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES:
        raise ValueError(f"{path}: shorter than an MRC header")
    stamp = raw[212:214]
    order = ">" if stamp[0] == 0x11 else "<"
    nx, ny, nz, mode = struct.unpack(order + "4i", raw[0:16])
    nsymbt = struct.unpack(order + "i", raw[92:96])[0]
    if nx <= 0 or ny <= 0 or nz <= 0 or nsymbt < 0:
        raise ValueError(f"{path}: bad MRC dimensions ({nx}, {ny}, {nz}) / nsymbt {nsymbt}")
    return dict(nx=nx, ny=ny, nz=nz, mode=mode, nsymbt=nsymbt,
                offset=HEADER_BYTES + nsymbt, order=order)


def read_images(path, idx=None):
    """Images ``idx`` (all when None) of an MRC / MRCS stack as float32 (len(idx), ny, nx), in
    mrcfile's orientation (``data[z]``)."""
    h = header(path)
    if h["mode"] != 2:
        raise NotImplementedError(f"{path}: MRC mode {h['mode']} "
                                  f"({np.dtype(MODE_DTYPES.get(h['mode'], np.void)).name}); "
                                  "only mode 2 (float32) particle stacks are read")
    mm = np.memmap(path, dtype=np.dtype(np.float32).newbyteorder(h["order"]), mode="r",
                   offset=h["offset"], shape=(h["nz"], h["ny"], h["nx"]))
    if idx is None:
        idx = np.arange(h["nz"])
    idx = np.asarray(idx, dtype=np.int64)
    if idx.size and (idx.min() < 0 or idx.max() >= h["nz"]):
        raise IndexError(f"{path}: image index outside the stack of {h['nz']}")
    return np.ascontiguousarray(mm[idx], dtype=np.float32)


def write_stack(path, data, voxel_size=1.0):
    """A mode-2 (float32) MRCS stack of images data (nz, ny, nx), little-endian."""
    data = np.ascontiguousarray(data, dtype="<f4")
    nz, ny, nx = data.shape
    hdr = bytearray(HEADER_BYTES)
    struct.pack_into("<4i", hdr, 0, nx, ny, nz, 2)
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)                  # mx, my, mz
    struct.pack_into("<3f", hdr, 40, nx * voxel_size, ny * voxel_size, nz * voxel_size)
    struct.pack_into("<3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)                      # mapc, mapr, maps
    struct.pack_into("<3f", hdr, 76, float(data.min()), float(data.max()), float(data.mean()))
    struct.pack_into("<i", hdr, 92, 0)                             # nsymbt
    hdr[208:212] = b"MAP "
    hdr[212:216] = bytes([0x44, 0x44, 0x00, 0x00])
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(data.tobytes())


def read_volume(path):
    """A 3-D map (nz, ny, nx) as float32, in mrcfile's ``data`` orientation (the mask volume of
    ``getDistanceCTF_local_Conj9combinedS2.py:305-307``).  Integer and float16 modes are
    converted to float32 here; the reference interpolates them in their own dtype."""
    h = header(path)
    dt = MODE_DTYPES.get(h["mode"])
    if dt is None:
        raise NotImplementedError(f"{path}: MRC mode {h['mode']} is not read")
    mm = np.memmap(path, dtype=np.dtype(dt).newbyteorder(h["order"]), mode="r",
                   offset=h["offset"], shape=(h["nz"], h["ny"], h["nx"]))
    return np.ascontiguousarray(mm, dtype=np.float32)
