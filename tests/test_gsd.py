"""The GSD file layer (hmc_gsd_*, host code) against the GSD 2.1 layout read
here independently with struct (the layout of hoomd-gsd/src/file_layer.rs:
a 256-byte header, 32-byte index entries sorted by (frame, id), 64-byte
names), and the reference's own file-layer tests (file_layer.rs:2460-3060):
a new file's header, duplicate chunk names, names too long, index and
name-list expansion, appending after reopening."""
import struct

import numpy as np
import pytest

import _pkg

MAGIC = 0x65DF65DF65DF65DF
TYPES = {1: np.uint8, 2: np.uint16, 3: np.uint32, 4: np.uint64, 5: np.int8, 6: np.int16, 7: np.int32,
         8: np.int64, 9: np.float32, 10: np.float64, 11: np.uint8}


@pytest.fixture(scope="module")
def hmc():
    p = _pkg.pkg()
    p.lib()
    return p


def parse(path):
    """The whole file by the GSD 2.x layout: header fields, names by id,
    index entries, and every chunk as {(frame, name): array}."""
    b = open(path, "rb").read()
    magic, iloc, ialloc, nloc, nalloc, sver, fver = struct.unpack_from("<QQQQQII", b, 0)
    app = b[48:112].split(b"\0")[0].decode()
    sch = b[112:176].split(b"\0")[0].decode()
    names = []
    for k in range(nalloc):
        s = b[nloc + 64 * k: nloc + 64 * (k + 1)]
        if s[0] == 0:
            break
        names.append(s.split(b"\0")[0].decode())
    entries = []
    for k in range(ialloc):
        frame, rows, loc, cols, nid, typ, flags = struct.unpack_from("<QQQIHBB", b, iloc + 32 * k)
        if loc == 0:
            break
        entries.append((frame, rows, loc, cols, nid, typ, flags))
    chunks = {}
    for frame, rows, loc, cols, nid, typ, _ in entries:
        dt = np.dtype(TYPES[typ])
        chunks[(frame, names[nid])] = np.frombuffer(b, dt, rows * cols, loc).reshape(rows, cols)
    return dict(magic=magic, iloc=iloc, ialloc=ialloc, nloc=nloc, nalloc=nalloc, sver=sver, fver=fver, app=app,
                sch=sch, names=names, entries=entries, chunks=chunks, size=len(b))


def test_new_file_header(hmc, tmp_path):
    # file_layer.rs:2460-2478 (create_new) and the initial layout
    p = tmp_path / "new.gsd"
    f = hmc.GsdFile.create(p, "application", "schema", (12, 42))
    assert f.n_frames == 0 and f.header() == ("application", "schema", (12, 42))
    f.close()
    h = parse(p)
    assert h["magic"] == MAGIC and h["fver"] == (2 << 16) | 1 and h["sver"] == (12 << 16) | 42
    assert (h["iloc"], h["ialloc"]) == (256, 128) and (h["nloc"], h["nalloc"]) == (256 + 128 * 32, 16)
    assert h["size"] == 256 + 128 * 32 + 1024
    assert h["app"] == "application" and h["sch"] == "schema" and h["names"] == [] and h["entries"] == []
    g = hmc.GsdFile.open(p)
    assert g.n_frames == 0 and g.header() == ("application", "schema", (12, 42))
    g.close()


def test_names_too_long(hmc, tmp_path):
    # file_layer.rs:2495-2510: application or schema names of 64 bytes or more
    with pytest.raises(hmc.HmcError):
        hmc.GsdFile.create(tmp_path / "a.gsd", "x" * 64, "schema", (1, 0))
    with pytest.raises(hmc.HmcError):
        hmc.GsdFile.create(tmp_path / "b.gsd", "app", "y" * 64, (1, 0))
    f = hmc.GsdFile.create(tmp_path / "c.gsd", "x" * 63, "y" * 63, (1, 0))
    with pytest.raises(hmc.HmcError):
        f.write_chunk("z" * 64, np.zeros(1, np.uint8))
    f.close()


def test_all_types_round_trip(hmc, tmp_path):
    # file_layer.rs:2623-2800 (all_types): every data type, scalars and arrays,
    # over three frames; the index sorted by (frame, id) with ids by first use
    p = tmp_path / "types.gsd"
    rng = np.random.default_rng(3)
    f = hmc.GsdFile.create(p, "app", "sch", (1, 4))
    written = {}
    order = []
    for frame in range(3):
        for t, dt in TYPES.items():
            if t == 11:
                continue
            name = f"chunk/{np.dtype(dt).name}"
            cols = 1 + (t % 3)
            a = (rng.integers(0, 100, size=(5 + frame, cols)) if np.dtype(dt).kind in "iu"
                 else rng.random((5 + frame, cols))).astype(dt)
            f.write_chunk(name, a)
            written[(frame, name)] = a
            if name not in order:
                order.append(name)
        f.end_frame()
    assert f.n_frames == 3
    f.close()
    h = parse(p)
    assert h["names"] == order
    keys = [(e[0], e[4]) for e in h["entries"]]
    assert keys == sorted(keys) and len(keys) == 30
    g = hmc.GsdFile.open(p)
    assert g.n_frames == 3
    for key, a in written.items():
        assert np.array_equal(h["chunks"][key], a)
        assert np.array_equal(g.read(*key), a) and g.read(*key).dtype == a.dtype
    with pytest.raises(KeyError):
        g.read(3, "chunk/float32")
    with pytest.raises(KeyError):
        g.read(0, "missing")
    g.close()


def test_duplicate_chunk_name(hmc, tmp_path):
    # file_layer.rs:2904-2920: a name written twice in one frame
    f = hmc.GsdFile.create(tmp_path / "d.gsd", "app", "sch", (1, 0))
    f.write_chunk("a", np.zeros(2, np.uint32))
    with pytest.raises(hmc.HmcError):
        f.write_chunk("a", np.zeros(2, np.uint32))
    f.end_frame()
    f.write_chunk("a", np.ones(2, np.uint32))  # allowed again in the next frame
    f.end_frame()
    f.close()
    g = hmc.GsdFile.open(tmp_path / "d.gsd")
    assert g.n_frames == 2 and g.read(1, "a").ravel().tolist() == [1, 1]
    g.close()


def test_index_and_name_list_expansion(hmc, tmp_path):
    # file_layer.rs:3000-3025 (expand_index_multi): more entries than the
    # initial 128 and more names than the initial 16 move both blocks to the
    # end of the file; every chunk stays readable
    p = tmp_path / "grow.gsd"
    f = hmc.GsdFile.create(p, "app", "sch", (1, 0))
    for frame in range(70):
        f.write_chunk("value", np.array([frame], np.int64))
        f.write_chunk(f"name/{frame % 40}", np.full(3, frame, np.float64))
        f.end_frame()
    f.close()
    h = parse(p)
    assert h["ialloc"] >= 140 and h["iloc"] > 256 and h["nalloc"] >= 41 and h["nloc"] > 256 + 4096
    assert len(h["entries"]) == 140 and len(h["names"]) == 41
    g = hmc.GsdFile.open(p)
    assert g.n_frames == 70
    for frame in range(70):
        assert g.read(frame, "value")[0, 0] == frame == h["chunks"][(frame, "value")][0, 0]
        assert np.all(g.read(frame, f"name/{frame % 40}") == frame)
    g.close()


def test_append_after_reopen(hmc, tmp_path):
    # file_layer.rs:2481-2492 (open_write): frames appended to an existing file
    p = tmp_path / "app.gsd"
    f = hmc.GsdFile.create(p, "app", "sch", (1, 4))
    for k in range(2):
        f.write_chunk("configuration/step", np.array([k], np.uint64))
        f.end_frame()
    f.close()
    g = hmc.GsdFile.open(p, writable=True)
    assert g.n_frames == 2
    g.write_chunk("configuration/step", np.array([7], np.uint64))
    g.write_chunk("particles/N", np.array([5], np.uint32))
    g.end_frame()
    g.close()
    h = parse(p)
    assert h["names"] == ["configuration/step", "particles/N"]
    r = hmc.GsdFile.open(p)
    assert r.n_frames == 3 and [int(r.read(k, "configuration/step")[0, 0]) for k in range(3)] == [0, 1, 7]
    r.close()


def test_read_only_file_refuses_writes(hmc, tmp_path):
    p = tmp_path / "ro.gsd"
    hmc.GsdFile.create(p, "app", "sch", (1, 0)).close()
    g = hmc.GsdFile.open(p)
    with pytest.raises(hmc.HmcError):
        g.write_chunk("a", np.zeros(1, np.uint8))
    g.close()
    with pytest.raises(hmc.HmcError):
        hmc.GsdFile.open(tmp_path / "missing.gsd")
    (tmp_path / "bad.gsd").write_bytes(b"\0" * 300)
    with pytest.raises(hmc.HmcError):
        hmc.GsdFile.open(tmp_path / "bad.gsd")
