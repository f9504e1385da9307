"""Reader for the RREATBL v1 column container used by the schema-6 bundle.

Format (native_amrex_rrea/include/rrea/RreaTableContainer.H:4-20):

    0   8  magic "RREATBL\\0"
    8   4  version 1                       u32 le
   12   4  endian marker 0x01020304        u32 le
   16   4  header_bytes                    u32 le
   20   4  reserved 0
   24   8  payload_bytes                   u64 le
   32  32  reserved 0
   64   N  JSON header (every value a string)
       ..  one deflate stream holding the byte-shuffled columns in header order

A numeric column of width w is stored as w byte planes (byte k of every item
contiguous), the inverse of which is a reshape + transpose.  Text columns are
newline-separated UTF-8.  The decoder refuses anything the reference reader
refuses (RreaTableContainer.H:92-144 Read, :191-237 Column): bad magic,
version, endianness, size mismatch and non-finite f8 values.
"""

from __future__ import annotations

import json
import struct
import zlib

import numpy as np

MAGIC = b"RREATBL\x00"
_DTYPES = {"f8": "<f8", "i4": "<i4", "u4": "<u4", "u2": "<u2", "u1": "<u1"}


class RreaTableError(RuntimeError):
    """A container that the reference reader would also refuse."""


class Container:
    def __init__(self, header: dict, payload: bytes, path: str):
        self.header = header
        self.payload = payload
        self.path = path
        names = _split(header.get("column_names", ""))
        dtypes = _split(header.get("column_dtypes", ""))
        sizes = _split(header.get("column_bytes", ""))
        if not (len(names) == len(dtypes) == len(sizes)):
            raise RreaTableError(f"RREATBL: column lists disagree in {path}")
        self.columns = {}
        offset = 0
        for name, dtype, size in zip(names, dtypes, sizes):
            size = int(size)
            if offset + size > len(payload):
                raise RreaTableError(f"RREATBL: column runs past the payload: {name}")
            self.columns[name] = (offset, size, dtype)
            offset += size

    def rows(self) -> int:
        return int(self.header["rows"])

    def column(self, name: str) -> np.ndarray:
        """One numeric column decoded to float64 in row order."""
        if name not in self.columns:
            raise RreaTableError(f"RREATBL: {self.path} has no column named {name}")
        offset, size, dtype = self.columns[name]
        if dtype == "str":
            raise RreaTableError(f"RREATBL: column {name} is text, not numeric")
        if dtype not in _DTYPES:
            raise RreaTableError(f"RREATBL: unsupported column dtype: {dtype}")
        width = np.dtype(_DTYPES[dtype]).itemsize
        count = size // width
        planes = np.frombuffer(self.payload, dtype=np.uint8, count=size, offset=offset)
        plain = planes.reshape(width, count).T.copy().reshape(-1)
        values = plain.view(_DTYPES[dtype]).astype(np.float64)
        if dtype == "f8" and not np.all(np.isfinite(values)):
            raise RreaTableError(f"RREATBL: non-finite value in column {name} of {self.path}")
        return values

    def text_column(self, name: str) -> list[str]:
        offset, size, dtype = self.columns[name]
        if dtype != "str":
            raise RreaTableError(f"RREATBL: {self.path} has no text column named {name}")
        return self.payload[offset:offset + size].decode("utf-8").split("\n")

    def opaque(self) -> bytes:
        if self.header.get("kind") != "opaque":
            raise RreaTableError(f"RREATBL: not an opaque container: {self.path}")
        return self.payload


def _split(text: str) -> list[str]:
    return [item.strip() for item in text.split(",")] if text else []


def read(path: str) -> Container:
    with open(path, "rb") as handle:
        blob = handle.read()
    if len(blob) < 64 or blob[:8] != MAGIC:
        raise RreaTableError(f"RREATBL: bad magic in {path}")
    version, endian, header_bytes, reserved, payload_bytes = struct.unpack_from("<IIIIQ", blob, 8)
    if version != 1 or endian != 0x01020304 or reserved != 0:
        raise RreaTableError(f"RREATBL: incompatible version/endianness in {path}")
    if len(blob) != 64 + header_bytes + payload_bytes:
        raise RreaTableError(f"RREATBL: declared header/payload size does not match {path}")
    header = json.loads(blob[64:64 + header_bytes].decode("utf-8"))
    payload = zlib.decompress(blob[64 + header_bytes:])
    return Container(header, payload, path)
