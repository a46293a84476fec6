"""Minimal GSD file reader and writer for the snapshot harness.

The reference benchmark reads its frames from HOOMD-blue GSD files through HOOMD's GSDReader
(benchmark/lbvh_benchmark.cc:73-100, 123-131).  Neither HOOMD nor the gsd package is part of
this library, so this module implements the GSD file layer itself, as the GSD format
documentation lays it out, plus the few chunks of the HOOMD schema the harness needs:

File layer (little endian):
  header, 256 bytes at offset 0:
    u64 magic (0x65DF65DF65DF65DF), u64 index_location, u64 index_allocated_entries,
    u64 namelist_location, u64 namelist_allocated_entries, u32 schema_version,
    u32 gsd_version ((major << 16) | minor), char application[64], char schema[64],
    char reserved[80]
  index: index_allocated_entries entries of 32 bytes at index_location:
    u64 frame, u64 N, i64 location, u32 M, u16 id, u8 type, u8 flags
    (an entry with location 0 is unused); chunk data: N x M values of `type` at `location`
  names: chunk id k is the k-th name.  Version 1 files keep names in 64-byte slots
    (namelist_allocated_entries slots); version 2 files keep null-terminated strings one after
    another in a buffer of namelist_allocated_entries bytes.
  types: 1 u8, 2 u16, 3 u32, 4 u64, 5 i8, 6 i16, 7 i32, 8 i64, 9 f32, 10 f64.

HOOMD schema chunks used: configuration/box (6 floats: Lx, Ly, Lz, xy, xz, yz), particles/N
(u32) and particles/position (N x 3 floats).  A chunk missing from frame f > 0 takes its value
from frame 0, and from the schema's default when frame 0 lacks it too (box 1, 1, 1, 0, 0, 0;
N 0), as the HOOMD schema specifies.

There is no GSD file in the reference and no GSD library in this environment, so the layout
is checked by round trips through the writer and by a file assembled byte by byte in the
tests, not against files written by the gsd library.
"""
from __future__ import annotations

import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

MAGIC = 0x65DF65DF65DF65DF
HEADER = struct.Struct("<QQQQQII64s64s80s")
INDEX = np.dtype([("frame", "<u8"), ("N", "<u8"), ("location", "<i8"), ("M", "<u4"), ("id", "<u2"), ("type", "u1"),
                  ("flags", "u1")])
TYPES = {1: np.dtype("<u1"), 2: np.dtype("<u2"), 3: np.dtype("<u4"), 4: np.dtype("<u8"), 5: np.dtype("<i1"),
         6: np.dtype("<i2"), 7: np.dtype("<i4"), 8: np.dtype("<i8"), 9: np.dtype("<f4"), 10: np.dtype("<f8")}
TYPE_IDS = {v: k for k, v in TYPES.items()}
NAME_SIZE = 64

assert HEADER.size == 256 and INDEX.itemsize == 32


class GSDFile:
    """Read-only view of a GSD file (format versions 1 and 2)."""

    def __init__(self, path: str):
        with open(path, "rb") as f:
            self._data = f.read()
        if len(self._data) < HEADER.size:
            raise ValueError(f"{path}: too short for a GSD header")
        (magic, index_loc, index_n, names_loc, names_n, self.schema_version, self.gsd_version, app, schema,
         _) = HEADER.unpack_from(self._data, 0)
        if magic != MAGIC:
            raise ValueError(f"{path}: not a GSD file (magic {magic:#x})")
        self.major = self.gsd_version >> 16
        if self.major not in (1, 2):
            raise ValueError(f"{path}: GSD version {self.major}.{self.gsd_version & 0xffff} is not supported")
        self.application = app.split(b"\0", 1)[0].decode()
        self.schema = schema.split(b"\0", 1)[0].decode()
        self.names = self._read_names(names_loc, names_n)
        idx = np.frombuffer(self._data, dtype=INDEX, count=index_n, offset=index_loc) if index_n else \
            np.zeros(0, INDEX)
        idx = idx[idx["location"] != 0]
        self._chunks: Dict[Tuple[int, str], np.ndarray] = {}
        self.nframes = int(idx["frame"].max()) + 1 if idx.size else 0
        for e in idx:
            if int(e["id"]) >= len(self.names):
                raise ValueError(f"{path}: chunk id {int(e['id'])} has no name")
            if int(e["type"]) not in TYPES:
                raise ValueError(f"{path}: unknown chunk type {int(e['type'])}")
            self._chunks[(int(e["frame"]), self.names[int(e["id"])])] = e

    def _read_names(self, loc: int, n: int) -> List[str]:
        if n == 0:
            return []
        if self.major == 1:
            raw = self._data[loc:loc + NAME_SIZE * n]
            names = [raw[k * NAME_SIZE:(k + 1) * NAME_SIZE].split(b"\0", 1)[0].decode() for k in range(n)]
            while names and names[-1] == "":
                names.pop()
            return names
        raw = self._data[loc:loc + n]
        names = raw.split(b"\0")
        out = []
        for s in names:
            if s == b"":
                break
            out.append(s.decode())
        return out

    def chunk(self, frame: int, name: str) -> Optional[np.ndarray]:
        """The chunk's values (N x M array, or N values when M == 1), or None if absent."""
        e = self._chunks.get((frame, name))
        if e is None:
            return None
        dt = TYPES[int(e["type"])]
        n, m = int(e["N"]), int(e["M"])
        a = np.frombuffer(self._data, dtype=dt, count=n * m, offset=int(e["location"]))
        return a.reshape(n, m) if m > 1 else a.copy()

    def hoomd_frame(self, frame: int):
        """(positions N x 3 float32, box 6 float32) of a frame of the HOOMD schema."""
        if not 0 <= frame < self.nframes:
            raise IndexError(f"frame {frame} of {self.nframes}")

        def value(name, default):
            v = self.chunk(frame, name)
            if v is None and frame > 0:
                v = self.chunk(0, name)
            return default if v is None else v

        box = np.asarray(value("configuration/box", np.array([1, 1, 1, 0, 0, 0], np.float32)), np.float32).reshape(-1)
        N = int(np.asarray(value("particles/N", np.zeros(1, np.uint32))).reshape(-1)[0])
        pos = np.asarray(value("particles/position", np.zeros((N, 3), np.float32)), np.float32).reshape(-1, 3)
        if pos.shape[0] != N:
            raise ValueError(f"frame {frame}: particles/position holds {pos.shape[0]} rows for N = {N}")
        if box.shape[0] != 6:
            raise ValueError(f"frame {frame}: configuration/box must hold 6 values")
        return np.ascontiguousarray(pos), box


def write_gsd(path: str, frames: List[Dict[str, np.ndarray]], application: str = "neighbor_amd",
              schema: str = "hoomd", schema_version: Tuple[int, int] = (1, 4)) -> None:
    """Write a version 2.0 GSD file; frames[f] maps chunk names to arrays (1-D: M = 1, 2-D: N x M)."""
    names: List[str] = []
    for fr in frames:
        for k in fr:
            if k not in names:
                names.append(k)
    body = bytearray()
    entries = []
    loc = HEADER.size
    for f, fr in enumerate(frames):
        for name in names:
            if name not in fr:
                continue
            a = np.ascontiguousarray(fr[name])
            dt = a.dtype.newbyteorder("<") if a.dtype.byteorder == ">" else a.dtype
            if np.dtype(dt) not in TYPE_IDS:
                raise ValueError(f"chunk {name}: unsupported dtype {a.dtype}")
            n, m = (a.shape[0], 1) if a.ndim == 1 else (a.shape[0], int(np.prod(a.shape[1:])))
            data = a.astype(np.dtype(dt).newbyteorder("<"), copy=False).tobytes()
            entries.append((f, n, loc + len(body), m, names.index(name), TYPE_IDS[np.dtype(dt)], 0))
            body += data
    names_buf = b"".join(n.encode() + b"\0" for n in names)
    names_buf += b"\0" * (-len(names_buf) % 8)
    names_loc = HEADER.size + len(body)
    index_loc = names_loc + len(names_buf)
    index = np.zeros(len(entries), INDEX)
    for k, e in enumerate(entries):
        index[k] = e
    header = HEADER.pack(MAGIC, index_loc, len(entries), names_loc, len(names_buf),
                         (schema_version[0] << 16) | schema_version[1], 2 << 16, application.encode()[:63],
                         schema.encode()[:63], b"\0" * 80)
    with open(path, "wb") as fh:
        fh.write(header)
        fh.write(bytes(body))
        fh.write(names_buf)
        fh.write(index.tobytes())


def read_hoomd_frames(path: str):
    """(positions list, box-length list) of every frame of a HOOMD-schema GSD file with an
    orthorhombic box (the harness's periodic cube or cuboid; tilt factors must be zero)."""
    g = GSDFile(path)
    positions, boxes = [], []
    for f in range(g.nframes):
        pos, box = g.hoomd_frame(f)
        if np.any(box[3:] != 0):
            raise ValueError(f"{path} frame {f}: triclinic boxes are not supported")
        if not np.all(box[:3] > 0):
            raise ValueError(f"{path} frame {f}: box lengths must be positive")
        positions.append(pos)
        boxes.append(box[:3].astype(np.float64))
    return positions, boxes


def write_hoomd_frames(path: str, positions, boxes) -> None:
    """HOOMD-schema GSD file of frames of positions (N x 3) in orthorhombic boxes (3 lengths)."""
    frames = []
    for pos, L in zip(positions, boxes):
        pos = np.asarray(pos, np.float32)
        box = np.zeros(6, np.float32)
        box[:3] = np.asarray(L, np.float32)
        frames.append({"configuration/box": box, "particles/N": np.array([pos.shape[0]], np.uint32),
                       "particles/position": pos})
    write_gsd(path, frames)
