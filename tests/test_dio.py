"""UG's data files (SaveData / LoadData, np/udm/data_io.cc:408-1019;
header np/udm/dio.cc:269-370; encodings low/bio.cc) written and read by
libugmg (include/ugmg_dio.h), without a GPU.

No file written by UG itself exists in this repository, so the layout is
pinned byte for byte against the encoding rules of bio.cc spelled out here
(length-prefixed strings with their mode's separators, "%d\\n" / "%g\\n"
text numbers, big-endian XDR words, native binary), plus round trips in
every mode and the DATA_IO_1.6 variant (no ident line)."""
import struct

import numpy as np
import pytest

from ug_amd import dio

TITLE = "####.sparse.data.storage.format.####"


def _ascii_str(s):
    return f"{len(s)}\n{s} ".encode()


def _bin_str(s):
    return f"{len(s)} {s} ".encode()


def _xdr_str(s):
    return f"{len(s)} {s}\n".encode()


def _expected(mode, data, version="DATA_IO_1.7", ident="---", mgfile="saved_without_mg",
              vd=(("sol", 1, 0, "u"),), times=(-1.0, -1.0, -1.0), ints=(1, 0, 0)):
    out = _ascii_str(TITLE) + f"{mode}\n".encode()
    if mode == dio.ASCII:
        S, I, D = _ascii_str, (lambda v: f"{v}\n".encode()), (lambda v: f"{v:g}\n".encode())
    elif mode == dio.BIN:
        S, I, D = _bin_str, (lambda v: struct.pack("=i", v)), (lambda v: struct.pack("=d", v))
    else:
        S, I, D = _xdr_str, (lambda v: struct.pack(">I", v & 0xffffffff)), (lambda v: struct.pack(">d", v))
    out += S(version)
    if version != "DATA_IO_1.6":
        out += S(ident)
    out += S(mgfile) + b"".join(D(t) for t in times)
    out += b"".join(I(v) for v in (*ints, len(vd)))
    for name, ncomp, typ, comps in vd:
        out += S(name) + I(ncomp) + I(typ) + S(comps)
    out += I(len(data)) + b"".join(D(v) for v in data)
    return out


@pytest.mark.parametrize("mode", [dio.XDR, dio.ASCII, dio.BIN])
def test_bytes_follow_bio_encodings(tmp_path, mode):
    data = [0.5, -1.25, 3.0, 1e-300, -0.0, 2.5e10]
    p = str(tmp_path / f"x.ug.data.{['xdr', 'asc', 'bin'][mode]}")
    dio.write(p, data, mode=mode, vd=[("sol", "u")])
    assert open(p, "rb").read() == _expected(mode, data)


@pytest.mark.parametrize("mode", [dio.XDR, dio.BIN])
def test_round_trip_is_exact_in_binary_modes(tmp_path, mode):
    rng = np.random.default_rng(mode)
    data = rng.standard_normal(3 * 1000)
    p = str(tmp_path / "d.ug.data")
    dio.write(p, data, mode=mode, vd=[("x", "xyz")], ident="elasticity run 7", time=0.25, dt=0.01,
              ndt=0.02, magic_cookie=123456)
    h, back = dio.read(p)
    assert np.array_equal(back, data)
    assert h["mode"] == mode and h["version"] == "DATA_IO_1.7" and h["ident"] == "elasticity run 7"
    assert (h["time"], h["dt"], h["ndt"], h["magic_cookie"]) == (0.25, 0.01, 0.02, 123456)
    assert h["vd"] == [("x", 3, dio.MULTIPLE_SCALAR, "xyz")] and h["ndata"] == 3000


def test_ascii_keeps_six_digits_as_ug_writes_them(tmp_path):
    data = np.array([np.pi, -np.e, 1.0 / 3.0, 12345678.9])
    p = str(tmp_path / "a.ug.data.asc")
    dio.write(p, data, mode=dio.ASCII)
    _, back = dio.read(p)
    assert np.array_equal(back, [float(f"{v:g}") for v in data])


def test_version_1_6_has_no_ident(tmp_path):
    p = tmp_path / "old.ug.data.asc"
    p.write_bytes(_expected(dio.ASCII, [1.0, 2.0], version="DATA_IO_1.6"))
    h, back = dio.read(str(p))
    assert h["version"] == "DATA_IO_1.7" and h["ident"] == "" and list(back) == [1.0, 2.0]


def test_rejects_other_files(tmp_path):
    p = tmp_path / "junk"
    p.write_bytes(b"12\nnot.a.title. 1\n")
    with pytest.raises(ValueError):
        dio.read(str(p))
    good = _expected(dio.XDR, [1.0, 2.0, 3.0])
    p.write_bytes(good[:-8])                      # last double cut off
    with pytest.raises(ValueError):
        dio.read(str(p))


def test_per_level_split():
    data = np.arange(2 * (5 + 9), dtype=float)     # 2 components, levels of 5 and 9 nodes
    lv = dio.per_level(data, 2, [5, 9])
    assert np.array_equal(lv[0], np.arange(10)) and np.array_equal(lv[1], np.arange(10, 28))
    lv = dio.per_level(data, 2, [5, 9], ids=[[4, 3, 2, 1, 0], list(range(5, 14))])
    assert np.array_equal(lv[0][:2], [8, 9])
