"""Byte-compatible restatement of the reference's DataContainer wire format.

The reference (common/data_container.cc:133-240) stores a parameter list as a
sequence of records, each ``[u8 length][length bytes]`` with the raw object
bytes of the inserted value.  Plugins read the records back in order in their
``Setup*`` methods.  The same bytes configure the HIP path (through the C ABI)
and the CPU oracle, so a driver that builds a container the reference's way
configures either.
"""
from __future__ import annotations

import struct
from typing import Iterable, Sequence


class DataContainer:
    """Sequential parameter blob; ``insert_*`` mirror ``DataContainer::Insert<T>``."""

    def __init__(self) -> None:
        self._buf = bytearray()
        self.n_records = 0

    # -- record primitives ------------------------------------------------------------------
    def _record(self, payload: bytes) -> "DataContainer":
        if len(payload) > 255:
            raise ValueError(f"record of {len(payload)} bytes exceeds the 255-byte limit of the format")
        self._buf.append(len(payload))
        self._buf += payload
        self.n_records += 1
        return self

    def insert_double(self, x: float) -> "DataContainer":
        return self._record(struct.pack("<d", float(x)))

    def insert_int(self, x: int) -> "DataContainer":
        return self._record(struct.pack("<i", int(x)))

    def insert_bool(self, x: bool) -> "DataContainer":
        return self._record(struct.pack("<B", 1 if x else 0))

    def insert_vec(self, v: Sequence[float]) -> "DataContainer":
        """GeoVector: three doubles (common/vectors.hh:28)."""
        if len(v) != 3:
            raise ValueError("GeoVector needs 3 components")
        return self._record(struct.pack("<3d", *map(float, v)))

    def insert_multiindex(self, v: Sequence[int]) -> "DataContainer":
        """MultiIndex: three ints (common/multi_index.hh)."""
        if len(v) != 3:
            raise ValueError("MultiIndex needs 3 components")
        return self._record(struct.pack("<3i", *map(int, v)))

    def insert_matrix(self, m: Sequence[Sequence[float]]) -> "DataContainer":
        flat = [float(x) for row in m for x in row]
        if len(flat) != 9:
            raise ValueError("GeoMatrix needs 9 components")
        return self._record(struct.pack("<9d", *flat))

    def insert_int_vector(self, v: Iterable[int]) -> "DataContainer":
        """std::vector<int>: length-prefixed raw int array (common/data_container.cc:178-195)."""
        vals = [int(x) for x in v]
        return self._record(struct.pack(f"<{len(vals)}i", *vals))

    def insert_double_vector(self, v: Iterable[float]) -> "DataContainer":
        vals = [float(x) for x in v]
        return self._record(struct.pack(f"<{len(vals)}d", *vals))

    def insert_string(self, s: str) -> "DataContainer":
        """std::string: raw characters, no terminator (common/data_container.cc:159-175)."""
        return self._record(s.encode())

    # -- access ----------------------------------------------------------------------------
    def bytes(self) -> bytes:
        return bytes(self._buf)

    def __len__(self) -> int:
        return len(self._buf)

    def clear(self) -> None:
        self._buf = bytearray()
        self.n_records = 0

    @staticmethod
    def records(blob: bytes) -> list[bytes]:
        """Split a blob back into its records (the reader side of the format)."""
        out, i = [], 0
        while i < len(blob):
            n = blob[i]
            if i + 1 + n > len(blob):
                raise ValueError("truncated DataContainer record")
            out.append(blob[i + 1:i + 1 + n])
            i += 1 + n
        return out


class DataReader:
    """Sequential reader over a blob; ``read_*`` mirror ``DataContainer::Read<T>``
    (common/data_container.cc:203-240): each call consumes one record."""

    def __init__(self, blob: bytes) -> None:
        self._recs = DataContainer.records(blob)
        self._i = 0

    def _next(self) -> bytes:
        if self._i >= len(self._recs):
            raise ValueError("read past the last record")
        r = self._recs[self._i]
        self._i += 1
        return r

    def reset(self) -> None:
        self._i = 0

    def read_int(self) -> int:
        return struct.unpack("<i", self._next())[0]

    def read_double(self) -> float:
        return struct.unpack("<d", self._next())[0]

    def read_bool(self) -> bool:
        return self._next()[0] != 0

    def read_vec(self) -> tuple:
        return struct.unpack("<3d", self._next())

    def read_multiindex(self) -> tuple:
        return struct.unpack("<3i", self._next())

    def read_string(self) -> str:
        return self._next().decode()

    def read_int_vector(self) -> list:
        r = self._next()
        return list(struct.unpack(f"<{len(r) // 4}i", r))

    def read_double_vector(self) -> list:
        r = self._next()
        return list(struct.unpack(f"<{len(r) // 8}d", r))
