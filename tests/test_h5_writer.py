"""The HDF5 writer (dfn_h5_write_f64) that produces the solver's output files.

libhdf5 is not in the image, so the writer is pinned two ways: every file it
writes reads back bit for bit through the library's reader, and its superblock
and dataset header messages are byte-equal to the encoding libhdf5 itself used
for FP64 datasets in the reference's input file (tests/golden/h5_messages.json,
extracted from problems/BP6QD_A_4000/config by make_h5_messages.py)."""
import json
import os
import struct

import numpy as np
import pytest

from demystification_dfn_amd.device import LIB_PATH

pytestmark = pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="device library not built")
GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "h5_messages.json")))


def _write(tmp_path, ds):
    from demystification_dfn_amd import config_io
    path = str(tmp_path / "f.h5")
    config_io.write_h5(path, ds)
    return path, open(path, "rb").read()


def _members(d):
    """{name: [(type, flags, body)]} of the root group's datasets, by the format."""
    root = struct.unpack_from("<Q", d, 64)[0]
    tree, heap = struct.unpack_from("<QQ", d, 80)
    assert d[root] == 1 and struct.unpack_from("<H", d, root + 16)[0] == 0x11
    assert struct.unpack_from("<QQ", d, root + 24) == (tree, heap)
    assert d[tree:tree + 4] == b"TREE" and d[tree + 4] == 0 and d[tree + 5] == 0
    assert d[heap:heap + 4] == b"HEAP"
    hsize, free, hdata = struct.unpack_from("<QQQ", d, heap + 8)
    assert free < hsize and struct.unpack_from("<Q", d, hdata + free)[0] == 1  # one free block, list end
    snod = struct.unpack_from("<Q", d, tree + 32)[0]
    assert d[snod:snod + 4] == b"SNOD" and d[snod + 4] == 1
    out, names = {}, []
    for s in range(struct.unpack_from("<H", d, snod + 6)[0]):
        noff, oh = struct.unpack_from("<QQ", d, snod + 8 + 40 * s)
        name = d[hdata + noff:d.index(b"\0", hdata + noff)].decode()
        names.append(name)
        nmsg = struct.unpack_from("<H", d, oh + 2)[0]
        p, msgs = oh + 16, []
        for _ in range(nmsg):
            t, sz, fl = struct.unpack_from("<HHB", d, p)
            msgs.append((t, fl, d[p + 8:p + 8 + sz]))
            p += 8 + sz
        out[name] = msgs
    # the B-tree's second key is the heap offset of the last name; entries are in name order
    assert names == sorted(names)
    last = struct.unpack_from("<Q", d, tree + 40)[0]
    assert d[hdata + last:d.index(b"\0", hdata + last)].decode() == names[-1]
    return out


def test_round_trip(tmp_path):
    from demystification_dfn_amd import config_io
    rng = np.random.default_rng(3)
    ds = {"V": rng.standard_normal((11, 4)), "theta": np.exp(rng.standard_normal((11, 4)) * 50),
          "time": rng.random(4) * 1e9, "P": np.array([[np.inf, -0.0, 5e-324, np.nan]]), "e": np.zeros((0, 3))}
    path, _ = _write(tmp_path, ds)
    with config_io.H5File(path) as f:
        for k, v in ds.items():
            assert f.info(k) == (1, v.shape)
            if v.size:
                got = f.f64(k)
                assert got.tobytes() == v.tobytes(), k


def test_encoding_matches_libhdf5(tmp_path):
    ds = {"V": np.arange(12.0).reshape(6, 2), "time": np.arange(2.0)}
    path, d = _write(tmp_path, ds)
    assert d[:8] == b"\x89HDF\r\n\x1a\n"
    # versions, sizes of offsets and lengths, B-tree ranks
    assert d[8:20].hex() == GOLD["superblock_fixed"][:24]
    assert struct.unpack_from("<Q", d, 40)[0] == len(d)  # end-of-file address
    m = _members(d)
    assert sorted(m) == ["V", "time"]
    for name, arr in ds.items():
        g = GOLD["datasets"][f"rank{arr.ndim}"]
        msgs = {t: (fl, body) for t, fl, body in m[name]}
        assert sorted(msgs) == [0x01, 0x03, 0x05, 0x08]
        assert msgs[0x03][0] == g["datatype"][0] and msgs[0x03][1].hex() == g["datatype"][1]
        assert msgs[0x05][0] == g["fill"][0] and msgs[0x05][1].hex() == g["fill"][1]
        # dataspace: version, rank, flags (maximum dimensions present) as libhdf5 writes them
        sp = msgs[0x01][1]
        assert sp[:8].hex() == g["dataspace"][1][:16] and msgs[0x01][0] == g["dataspace"][0]
        dims = struct.unpack_from(f"<{2 * arr.ndim}Q", sp, 8)
        assert dims == arr.shape + arr.shape
        lay = msgs[0x08][1]
        assert msgs[0x08][0] == g["layout_head"][0] and lay[:2].hex() == g["layout_head"][1]
        addr, size = struct.unpack_from("<QQ", lay, 2)
        assert size == arr.nbytes and np.frombuffer(d[addr:addr + size], "<f8").tobytes() == arr.tobytes()


def test_rejects_bad_arguments(tmp_path):
    from demystification_dfn_amd import _abi, config_io
    with pytest.raises(_abi.DfnError):
        config_io.write_h5(str(tmp_path / "a.h5"), {f"d{i}": np.zeros(2) for i in range(9)})  # > 8 datasets
    with pytest.raises(_abi.DfnError):
        config_io.write_h5(str(tmp_path / "b.h5"), {"x": np.zeros((2, 2, 2))})  # rank 3
