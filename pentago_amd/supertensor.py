"""Supertensor (.pentago) files: read into device slabs, write from device slabs.

Format (pentago/data/supertensor.h:1-50, supertensor.cc): an 85-byte little-endian header
(magic "pentago supertensor\\n", version, valid, stones, section, shape, block_size, blocks,
filter, index blob), a compressed index of one (uncompressed size, compressed size, offset)
blob per block, and each block compressed on its own (zlib, or xz/lzma: recognised by its
magic, pentago/data/compress.cc).  A multiple-section file starts with
"pentago sections   \\n", version 3, the section count and the header size, then one header
per section (supertensor.cc:401-436, 512-539).  Block data: per node (black wins, white wins),
interleaved by filter 1 (pentago/data/filter.h).

Host side (this module): headers, indices, compression, file I/O -- Python's zlib and lzma
modules are the compression libraries available here.  Device side (HIP, through the C ABI):
pe_supertensor_decode / pe_supertensor_encode convert between file words and slab nodes
(we-win, we-win-or-tie) in the block's own slab slot; blocks of 8^4 nodes map one to one onto
slab blocks.
"""
import lzma
import struct
import zlib

import numpy as np

from . import PentagoError, _check, hip, lib, section_shape

SINGLE_MAGIC = b"pentago supertensor\n"
MULTIPLE_MAGIC = b"pentago sections   \n"
HEADER = struct.Struct("<20sI?I8s4HI4HI3Q")
assert HEADER.size == 85
XZ_MAGIC = b"\xfd7zXZ\x00"
BLOCK = 8   # slab blocks are 8^4 nodes; files written here use the same block size


def decompress(data, size):
    """Decompress one blob (zlib or xz by magic, pentago/data/compress.cc:71-95).  Output is
    capped one byte past the expected size, so a corrupt blob cannot inflate without bound."""
    if data[:6] == XZ_MAGIC:
        out = lzma.LZMADecompressor(format=lzma.FORMAT_XZ).decompress(data, max_length=size + 1)
    else:
        out = zlib.decompressobj().decompress(data, size + 1)
    if len(out) != size:
        raise PentagoError("supertensor blob decompressed to %d bytes, expected %d" % (len(out), size))
    return out


def compress(data, level):
    """zlib for levels below 20, else xz at preset level - 20 with a CRC64 check
    (pentago/data/compress.cc:42-61)."""
    if level < 20:
        return zlib.compress(data, level)
    return lzma.compress(data, format=lzma.FORMAT_XZ, check=lzma.CHECK_CRC64, preset=level - 20)


class Header:
    """supertensor_header_t (pentago/data/supertensor.h:73-98)."""

    def __init__(self, section, block_size=BLOCK, filter_=1, version=3, valid=False, index=(0, 0, 0)):
        self.section = tuple(tuple(int(x) for x in q) for q in section)
        self.version, self.valid, self.filter = version, valid, filter_
        self.stones = sum(map(sum, self.section))
        self.shape = tuple(section_shape(self.section))
        self.block_size = block_size
        self.blocks = tuple((s + block_size - 1) // block_size for s in self.shape)
        self.index = tuple(index)   # (uncompressed size, compressed size, offset)

    def pack(self):
        sec = bytes(x for q in self.section for x in q)
        return HEADER.pack(SINGLE_MAGIC, self.version, self.valid, self.stones, sec, *self.shape,
                           self.block_size, *self.blocks, self.filter, *self.index)

    @classmethod
    def unpack(cls, buf, where="supertensor"):
        f = HEADER.unpack(buf)
        if f[0] != SINGLE_MAGIC:
            raise PentagoError("%s: incorrect magic string" % where)
        sec = tuple((f[4][2 * q], f[4][2 * q + 1]) for q in range(4))
        # the reference's consistency checks (supertensor.cc:156-171)
        if f[1] not in (2, 3):
            raise PentagoError("%s: unknown version %d" % (where, f[1]))
        if not f[2]:
            raise PentagoError("%s: file is marked invalid" % where)
        if not (1 <= f[9] <= 27) or f[9] & 1:
            raise PentagoError("%s: bad block size %d" % (where, f[9]))
        if any(sum(q) > 9 for q in sec):
            raise PentagoError("%s: invalid section" % where)
        h = cls(sec, f[9], f[14], f[1], f[2], f[15:18])
        if f[3] != h.stones or tuple(f[5:9]) != h.shape or tuple(f[10:14]) != h.blocks:
            raise PentagoError("%s: header fields disagree with the section" % where)
        return h

    def block_shape(self, block):
        return tuple(min(self.block_size, s - self.block_size * b) for s, b in zip(self.shape, block))


class Reader:
    """One section of a supertensor file (supertensor_reader_t)."""

    def __init__(self, f, header_offset, where):
        f.seek(header_offset)
        self.f = f
        self.header = h = Header.unpack(f.read(HEADER.size), where)
        f.seek(h.index[2])
        raw = decompress(f.read(h.index[1]), h.index[0])
        n = int(np.prod(h.blocks))
        self.index = np.frombuffer(raw, dtype="<u8").reshape(n, 3)
        for i in range(n):
            shape = h.block_shape(np.unravel_index(i, h.blocks))
            if int(self.index[i, 0]) != 64 * int(np.prod(shape)):
                raise PentagoError("%s: block %d has the wrong uncompressed size" % (where, i))

    def read_raw(self, block):
        """Decompressed, still filtered words of one block: (nodes, 8) uint64."""
        i = int(np.ravel_multi_index(block, self.header.blocks))
        size, csize, off = (int(x) for x in self.index[i])
        self.f.seek(off)
        return np.frombuffer(decompress(self.f.read(csize), size), dtype="<u8").reshape(-1, 8)

    def slab_fits(self, slab):
        return self.header.block_size == BLOCK and all(
            slab.lo[a] == 0 and slab.hi[a] == self.header.blocks[a] for a in range(4))

    def load(self, slab, stream=None):
        """Every block into a whole-section slab of this section (device): decompress on the
        host, copy into the block's slot, decode in place on the device."""
        if slab.section.counts != self.header.section or not self.slab_fits(slab):
            raise PentagoError("slab does not cover this supertensor section with 8^4 blocks")
        turn = self.header.stones & 1
        for i in range(int(np.prod(self.header.blocks))):
            block = np.unravel_index(i, self.header.blocks)
            words = np.ascontiguousarray(self.read_raw(block))
            dst = slab.data + 64 * 4096 * i
            hip.copy_to_device(dst, words)
            _check(lib().pe_supertensor_decode(self.header.filter, turn, dst, len(words), dst, stream),
                   "pe_supertensor_decode")


def open_supertensors(path):
    """Readers for every section of a single- or multiple-section file (supertensor.cc:401-436)."""
    f = open(path, "rb")
    magic = f.read(20)
    if magic == SINGLE_MAGIC:
        return [Reader(f, 0, path)]
    if magic != MULTIPLE_MAGIC:
        raise PentagoError("%s: bad magic string" % path)
    version, count, size = struct.unpack("<3I", f.read(12))
    if version != 3 or size != HEADER.size or count >= 8239:
        raise PentagoError("%s: unsupported multiple-section header" % path)
    return [Reader(f, 32 + size * s, path) for s in range(count)]


def write_supertensors(path, slabs, level=26, filter_=1, stream=None):
    """One multiple-section file from whole-section slabs (supertensor_writers,
    supertensor.cc:439-465, and finalize, 344-367): each block encoded on the device, copied
    back, compressed; blocks appended in order, then each section's index."""
    slabs = list(slabs)
    headers = [Header(s.section.counts, BLOCK, filter_, valid=True) for s in slabs]
    staging = None
    with open(path, "wb") as f:
        f.write(MULTIPLE_MAGIC + struct.pack("<3I", 3, len(slabs), HEADER.size))
        f.write(b"\0" * (HEADER.size * len(slabs)))
        offset = 32 + HEADER.size * len(slabs)
        for s, h in zip(slabs, headers):
            if not all(s.lo[a] == 0 and s.hi[a] == h.blocks[a] for a in range(4)):
                raise PentagoError("write_supertensors needs whole-section slabs")
            n = int(np.prod(h.blocks))
            index = np.zeros((n, 3), dtype="<u8")
            turn = h.stones & 1
            for i in range(n):
                shape = h.block_shape(np.unravel_index(i, h.blocks))
                nodes = int(np.prod(shape))
                if staging is None:
                    staging = hip.Buffer(64 * 4096)
                _check(lib().pe_supertensor_encode(filter_, turn, s.data + 64 * 4096 * i, nodes, staging.ptr, stream),
                       "pe_supertensor_encode")
                hip.synchronize()
                words = staging.download(np.zeros((nodes, 8), dtype="<u8"))
                blob = compress(words.tobytes(), level)
                f.write(blob)
                index[i] = (64 * nodes, len(blob), offset)
                offset += len(blob)
            blob = compress(index.tobytes(), level)
            f.write(blob)
            h.index = (index.nbytes, len(blob), offset)
            offset += len(blob)
        f.seek(32)
        for h in headers:
            f.write(h.pack())
    return headers
