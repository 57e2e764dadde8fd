"""Readers of the reducer's binary output files (.fmx matrices, _SAM.fsm).

Every file starts with the 40-byte file tag that writetagdb
(src/vpmUtilities/binaryDB.c:643-674) writes through FFa_writeTag: the tag
text blank-padded to 30 characters, the int32 endian marker 0x1234 followed
by two zero bytes, and the uint32 checksum; then the version field up to
a newline (";1.0;\\n" for the reducer's files) and the data.
"""
import struct

import numpy as np

TAG_CHARS = 30
TAG_BYTES = 40
ENDIAN_MARK = 0x1234


def read_tag(raw):
    """(tag text, checksum, version line, data offset) of a tagged file's
    bytes; ValueError when raw does not start with a file tag."""
    if len(raw) < TAG_BYTES:
        raise ValueError("short file: %d bytes" % len(raw))
    for order in "<>":
        mark, = struct.unpack(order + "I", raw[TAG_CHARS:TAG_CHARS + 4])
        if mark == ENDIAN_MARK:
            break
    else:
        raise ValueError("no endian marker: not a tagged binary file")
    if raw[TAG_CHARS + 4:TAG_CHARS + 6] != b"\0\0":
        raise ValueError("malformed endian marker")
    cs, = struct.unpack(order + "I", raw[TAG_CHARS + 6:TAG_BYTES])
    tag = raw[:TAG_CHARS].rstrip(b" \0").decode("ascii")
    end = raw.find(b"\n", TAG_BYTES, TAG_BYTES + 64)
    end = TAG_BYTES + 64 if end < 0 else end + 1
    return tag, cs, raw[TAG_BYTES:end].decode("ascii"), end


def read_fmx(path, shape=None):
    """(column-major FP64 matrix, tag text, checksum) of a .fmx file."""
    raw = open(path, "rb").read()
    tag, cs, ver, off = read_tag(raw)
    if not tag.startswith("#FEDEM ") or ver != ";1.0;\n":
        raise ValueError("%s: not a FEDEM matrix file (tag %r, version %r)" % (path, tag, ver))
    a = np.frombuffer(raw[off:], "<f8")
    return (a.reshape(shape, order="F") if shape is not None else a), tag, cs


def read_sam(path):
    """(checksum, int32 records) of a _SAM.fsm file (saveSAM,
    samReducerModule.f90:586-660)."""
    raw = open(path, "rb").read()
    tag, cs, ver, off = read_tag(raw)
    if tag != "#SAM data" or ver != ";1.0;\n":
        raise ValueError("%s: not a SAM file (tag %r)" % (path, tag))
    return cs, np.frombuffer(raw[off:], "<i4")
