"""Snapshot input of the benchmark harness (CPU): frames round-trip through the .npz layout and
through HOOMD-schema GSD files (the reference harness's input, benchmark/lbvh_benchmark.cc:123-131),
GSD files assembled field by field from the documented layout (name lists of format versions 1 and
2, unused index slots, chunks missing from a frame taken from frame 0) read back exactly,
single frames and one box for all frames are accepted, malformed inputs are refused; the 27
image offsets are the reference harness's, center first."""
import numpy as np
import pytest

from neighbor_amd.harness import image_offsets, read_frames, write_frames


def test_round_trip_and_shapes(tmp_path):
    rng = np.random.default_rng(0)
    pos = rng.uniform(-5, 5, size=(3, 100, 3)).astype(np.float32)
    box = np.array([[10.0, 10.0, 10.0], [10.5, 10.0, 9.5], [11.0, 11.0, 11.0]])
    write_frames(str(tmp_path / "f.npz"), pos, box)
    p, b = read_frames(str(tmp_path / "f.npz"))
    assert len(p) == 3 and all(np.array_equal(p[i], pos[i]) for i in range(3))
    assert all(np.array_equal(b[i], box[i]) for i in range(3))
    write_frames(str(tmp_path / "g.npz"), pos[0], [8.0, 9.0, 10.0])
    p, b = read_frames(str(tmp_path / "g.npz"))
    assert len(p) == 1 and p[0].dtype == np.float32 and list(b[0]) == [8.0, 9.0, 10.0]


def test_malformed_inputs_are_refused(tmp_path):
    np.savez(str(tmp_path / "a.npz"), positions=np.zeros((10, 2), np.float32), box=np.ones(3))
    with pytest.raises(ValueError):
        read_frames(str(tmp_path / "a.npz"))
    np.savez(str(tmp_path / "b.npz"), positions=np.zeros((2, 10, 3), np.float32), box=np.array([1.0, 0.0, 1.0]))
    with pytest.raises(ValueError):
        read_frames(str(tmp_path / "b.npz"))
    np.savez(str(tmp_path / "c.npz"), positions=np.zeros((2, 10, 3), np.float32), box=np.ones((3, 3)))
    with pytest.raises(ValueError):
        read_frames(str(tmp_path / "c.npz"))


def test_image_offsets_center_first():
    t = image_offsets()
    assert t[0] == (0, 0, 0) and len(t) == 27 and len(set(t)) == 27
    assert t[1] == (-1, -1, -1) and t[-1] == (1, 1, 1)   # loop order of lbvh_benchmark.cc:194-206


# ----------------------------------------------------------------------------- GSD input
def _gsd_bytes(version, names, chunks, index_slots=None, name_slots=None):
    """A GSD file assembled field by field from the documented layout (independent of
    neighbor_amd.gsd.write_gsd).  chunks: (frame, name, dtype id, N, M, raw bytes)."""
    import struct
    body = b""
    entries = []
    for frame, name, tid, n, m, raw in chunks:
        entries.append((frame, n, 256 + len(body), m, names.index(name), tid))
        body += raw
    if version == 1:
        slots = name_slots or len(names) + 3
        nl = b"".join(nm.encode().ljust(64, b"\0") for nm in names) + b"\0" * 64 * (slots - len(names))
        n_names = slots
    else:
        nl = b"".join(nm.encode() + b"\0" for nm in names) + b"\0" * 5
        n_names = len(nl)
    names_loc = 256 + len(body)
    index_loc = names_loc + len(nl)
    slots = index_slots or len(entries)
    idx = b"".join(struct.pack("<QQqIHBB", f, n, loc, m, i, t, 0) for f, n, loc, m, i, t in entries)
    idx += b"\0" * 32 * (slots - len(entries))           # unused slots: location 0
    header = struct.pack("<QQQQQII", 0x65DF65DF65DF65DF, index_loc, slots, names_loc, n_names, (1 << 16) | 4,
                         version << 16)
    header += b"test".ljust(64, b"\0") + b"hoomd".ljust(64, b"\0") + b"\0" * 80
    assert len(header) == 256
    return header + body + nl + idx


@pytest.mark.parametrize("version", [1, 2])
def test_gsd_file_assembled_by_hand(tmp_path, version):
    """Frames of a hand-assembled GSD file (both name-list layouts, unused index slots, a frame
    without a box that takes frame 0's) read back exactly."""
    p0 = np.arange(12, dtype=np.float32).reshape(4, 3) - 5.5
    p1 = (p0 * 0.5).astype(np.float32)
    box = np.array([12, 13, 14, 0, 0, 0], np.float32)
    names = ["configuration/box", "particles/N", "particles/position"]
    chunks = [(0, "configuration/box", 9, 6, 1, box.tobytes()), (0, "particles/N", 3, 1, 1, np.uint32(4).tobytes()),
              (0, "particles/position", 9, 4, 3, p0.tobytes()),
              (1, "particles/position", 9, 4, 3, p1.tobytes())]      # frame 1: N and box from frame 0
    path = tmp_path / "hand.gsd"
    path.write_bytes(_gsd_bytes(version, names, chunks, index_slots=7))
    pos, boxes = read_frames(str(path))
    assert len(pos) == 2
    assert np.array_equal(pos[0], p0) and np.array_equal(pos[1], p1)
    assert list(boxes[0]) == [12.0, 13.0, 14.0] and list(boxes[1]) == [12.0, 13.0, 14.0]


def test_gsd_writer_round_trip(tmp_path):
    from neighbor_amd.gsd import GSDFile, write_gsd, write_hoomd_frames
    rng = np.random.default_rng(3)
    pos = [rng.uniform(-4, 4, (50 + 10 * f, 3)).astype(np.float32) for f in range(3)]
    boxes = [[8.0, 8.5, 9.0], [8.0, 8.0, 8.0], [10.0, 10.0, 10.0]]
    write_hoomd_frames(str(tmp_path / "w.gsd"), pos, boxes)
    p, b = read_frames(str(tmp_path / "w.gsd"))
    assert all(np.array_equal(p[f], pos[f]) for f in range(3))
    assert all(np.array_equal(b[f], np.asarray(boxes[f], np.float32).astype(np.float64)) for f in range(3))
    g = GSDFile(str(tmp_path / "w.gsd"))
    assert g.major == 2 and g.schema == "hoomd" and g.nframes == 3
    # other chunk types round-trip too
    write_gsd(str(tmp_path / "t.gsd"), [{"a": np.arange(5, dtype=np.int64), "b": np.ones((2, 4), np.float64),
                                         "c": np.array([7], np.uint8)}])
    g = GSDFile(str(tmp_path / "t.gsd"))
    assert np.array_equal(g.chunk(0, "a"), np.arange(5)) and g.chunk(0, "b").shape == (2, 4)
    assert g.chunk(0, "c")[0] == 7 and g.chunk(0, "missing") is None


def test_gsd_malformed_inputs_are_refused(tmp_path):
    from neighbor_amd.gsd import write_hoomd_frames, write_gsd
    bad = tmp_path / "bad.gsd"
    bad.write_bytes(b"\0" * 300)
    with pytest.raises(ValueError):
        read_frames(str(bad))
    box = np.array([5, 5, 5, 0.1, 0, 0], np.float32)        # triclinic
    write_gsd(str(tmp_path / "tri.gsd"), [{"configuration/box": box, "particles/N": np.array([1], np.uint32),
                                           "particles/position": np.zeros((1, 3), np.float32)}])
    with pytest.raises(ValueError):
        read_frames(str(tmp_path / "tri.gsd"))
    write_gsd(str(tmp_path / "n.gsd"), [{"configuration/box": np.array([5, 5, 5, 0, 0, 0], np.float32),
                                         "particles/N": np.array([3], np.uint32),
                                         "particles/position": np.zeros((2, 3), np.float32)}])
    with pytest.raises(ValueError):
        read_frames(str(tmp_path / "n.gsd"))
    write_hoomd_frames(str(tmp_path / "ok.gsd"), [np.zeros((2, 3), np.float32)], [[1.0, 2.0, 3.0]])
    assert len(read_frames(str(tmp_path / "ok.gsd"))[0]) == 1
