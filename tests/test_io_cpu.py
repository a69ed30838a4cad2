"""Host I/O rows of the path through the C ABI (no GPU needed): the .cor
writer (cor.rs:10-296), sector normalisation (main.rs:2645-2683), the
packed-sample reader (main.rs:445-516) and the word-aligned seek
(main.rs:1585-1616).  Expected values come from the byte offsets and
constants the reference fixes and from hand-computed cases."""
import math
import os
import struct

import numpy as np
import pytest

import fxcorr as F


def _hdr(**kw):
    d = dict(fs_hz=1024000000, obs_hz=6.7e9, fft_point=1024, nsector_hint=8,
             clock_ref_unix=1700000000, source="3C273", ra=3.2, dec=0.035)
    d.update(kw)
    return F.cor_header(d["fs_hz"], d["obs_hz"], d["fft_point"], d["nsector_hint"],
                        d["clock_ref_unix"], d["source"], d["ra"], d["dec"])


ST1 = ("YAMAGU32", ord("G"), (-3502544.0, 3950966.0, 3566381.0))
ST2 = ("HITACHI_LONGNAME_X", 0xE9, (-3961788.0, 3243597.0, 3790597.0))


def _stations():
    return F.cor_station(*ST1), F.cor_station(*ST2)


# ------------------------------------------------------------ file header
def test_cor_file_header_offsets():
    s1, s2 = _stations()
    b = F.cor_build_header(_hdr(), s1, s2)
    assert len(b) == 256
    assert b[0:4] == bytes([0x83, 0xF9, 0xA2, 0x3E])                  # cor.rs:166
    assert struct.unpack_from("<iiid", b, 4) == (0x01030000, 1, 1024000000, 6.7e9)
    assert struct.unpack_from("<ii", b, 24) == (1024, 8)
    assert b[32:48] == b"YAMAGU32" + bytes(8)
    assert struct.unpack_from("<3d", b, 48) == ST1[2]
    assert b[72] == ord("G") and b[73:80] == bytes(7)
    assert b[80:96] == b"HITACHI_LONGNAME"                            # 16-byte truncation
    assert struct.unpack_from("<3d", b, 96) == ST2[2]
    assert b[120] == 0xC3                                             # UTF-8 lead byte of U+00E9
    assert b[128:144] == b"3C273" + bytes(11)
    assert struct.unpack_from("<dd", b, 144) == (3.2, 0.035)
    assert struct.unpack_from("<II", b, 160) == (1700000000, 0)
    assert struct.unpack_from("<II", b, 208) == (1700000000, 0)
    assert b[168:208] == bytes(40) and b[216:256] == bytes(40)        # clock polynomials zero


@pytest.mark.parametrize("kw", [dict(fft_point=1023), dict(fft_point=0),
                                dict(nsector_hint=-1), dict(clock_ref_unix=-1),
                                dict(clock_ref_unix=2 ** 32)])
def test_cor_file_header_rejects(kw):
    s1, s2 = _stations()
    with pytest.raises(F.FxcError):
        F.cor_build_header(_hdr(**kw), s1, s2)


# ---------------------------------------------------------- sector header
def test_cor_sector_header_times():
    b = F.cor_build_sector_header(1700000000, 1.0)
    assert len(b) == 128
    assert struct.unpack_from("<iIII", b, 0) == (1700000000, 0, 1700000001, 0)
    assert struct.unpack_from("<f", b, 112)[0] == 1.0
    assert b[16:112] == bytes(96) and b[116:128] == bytes(12)         # no model
    b = F.cor_build_sector_header(1700000000, 0.25)
    assert struct.unpack_from("<III", b, 4) == (0, 1700000000, 250000000)
    # cor.rs adds the offset to the Unix seconds in f64 before taking the
    # fraction: 1700000000 + 0.1f32 = 1700000000.0999999046..., nsec 99999905
    b = F.cor_build_sector_header(1700000000, 0.1)
    assert struct.unpack_from("<III", b, 4) == (0, 1700000000, 99999905)
    # non-finite or non-positive integration: end = start + 1 s, stored value kept
    for v in (float("nan"), -2.0, 0.0):
        b = F.cor_build_sector_header(1700000000, v)
        assert struct.unpack_from("<II", b, 8) == (1700000001, 0)
        got = struct.unpack_from("<f", b, 112)[0]
        assert (math.isnan(got) and math.isnan(v)) or got == np.float32(v)


def test_cor_sector_header_model_and_range():
    m = F.CorSectorModel()
    m.station1 = F.CorClockModel(11, 12, 1.5, 2.5, 3.5, 4.5, 5.5)
    m.station2 = F.CorClockModel(21, 22, -1.5, -2.5, -3.5, -4.5, -5.5)
    m.amp[:] = [0.5, 0.25]
    m.phs[:] = [-7, 300]
    b = F.cor_build_sector_header(1000, 2.0, m)
    assert struct.unpack_from("<II5d", b, 16) == (11, 12, 1.5, 2.5, 3.5, 4.5, 5.5)
    assert struct.unpack_from("<II5d", b, 64) == (21, 22, -1.5, -2.5, -3.5, -4.5, -5.5)
    assert struct.unpack_from("<fff", b, 112) == (2.0, 0.5, 0.25)
    assert struct.unpack_from("<hh", b, 124) == (-7, 300)
    with pytest.raises(F.FxcError):
        F.cor_build_sector_header(2 ** 31, 1.0)                       # i32 start time
    with pytest.raises(F.FxcError):
        F.cor_build_sector_header(-5, 1.0)                            # end sec below u32


def test_cor_writer_file(tmp_path):
    s1, s2 = _stations()
    n = 64
    p = str(tmp_path / "x.cor")
    w = F.CorWriter(p, _hdr(fft_point=n, nsector_hint=0), s1, s2)
    rng = np.random.default_rng(0)
    specs = [(rng.standard_normal(n // 2) + 1j * rng.standard_normal(n // 2)).astype(np.complex64)
             for _ in range(3)]
    for i, s in enumerate(specs):
        w.write_sector(1700000000 + i, 0.5, s)
    with pytest.raises(F.FxcError):
        w.write_sector(1700000003, 0.5, specs[0][:-1])                # length mismatch
    w.finalize()
    data = open(p, "rb").read()
    rec = 128 + (n // 2) * 8
    assert len(data) == 256 + 3 * rec
    assert struct.unpack_from("<i", data, 28)[0] == 3                 # sector count at finalize
    for i, s in enumerate(specs):
        off = 256 + i * rec
        assert struct.unpack_from("<i", data, off)[0] == 1700000000 + i
        got = np.frombuffer(data[off + 128:off + rec], dtype="<f4").view(np.complex64)
        assert np.array_equal(got, s)


# ----------------------------------------------------------- normalisation
def test_normalize_sector_closed_form():
    n, nf, half = 16, 1000, 9
    acc = np.zeros((5, half))
    acc[0] = np.arange(half) + 1.0
    acc[1] = 2.0
    acc[2] = 3.0
    acc[3] = 4.0
    acc[4] = -5.0
    p1, p2, w1, w2 = 5.0, 1.25, 0.5, 2.0
    o = F.normalize_sector(acc, n, nf, p1, p2, w1, w2)
    base = nf * n * n
    assert o["11"].shape == (n // 2,)                                  # bins 0..N/2-1
    np.testing.assert_array_equal(o["11"], np.complex64(2.0 / (0.5 * p1 * base)))
    np.testing.assert_array_equal(o["22"], np.complex64(3.0 / (0.5 * p2 * base)))
    c = 0.5 * math.sqrt(p1 * p2) * base
    np.testing.assert_array_equal(o["12"], np.complex64(complex(4.0 / c, -5.0 / c)))
    ph = (np.arange(n // 2) + 1.0) / (0.5 * (w1 * w1 * p1 + w2 * w2 * p2) * base)
    np.testing.assert_array_equal(o["ph"], ph.astype(np.float32).astype(np.complex64))
    o = F.normalize_sector(acc, n, nf, p1, p2, 0.0, 0.0)              # floor 1e-12
    ph = (np.arange(n // 2) + 1.0) / (0.5 * 1e-12 * base)
    np.testing.assert_array_equal(o["ph"].real, ph.astype(np.float32))


# ------------------------------------------------------------------ reader
def test_reader_padding_and_offset(tmp_path):
    data = bytes((i * 37 + 11) & 255 for i in range(1000))
    p = str(tmp_path / "a.raw")
    open(p, "wb").write(data)
    r = F.Reader(p, start_byte=100)
    out, got = r.read(300)
    assert got == 300 and out.tobytes() == data[100:400]
    out, got = r.read(700)                                            # 600 left, 100 zero pad
    assert got == 600 and out.tobytes() == data[400:] + bytes(100)
    out, got = r.read(10)
    assert got == 0 and not out.any()
    r.close()
    for sh in (1, 3, 7):
        r = F.Reader(p, start_byte=5, bit_offset=sh)
        out, _ = r.read(40)
        out2, _ = r.read(20)
        src = data[5:5 + 61]
        exp = [((src[i] >> sh) | (src[i + 1] << (8 - sh))) & 255 for i in range(60)]
        assert list(out) + list(out2) == exp
        r.close()
    with pytest.raises(F.FxcError):
        F.Reader(p, bit_offset=8)
    with pytest.raises(F.FxcError):
        F.Reader(str(tmp_path / "missing.raw"))


@pytest.mark.parametrize("start,bits,exp", [
    (37, 2, (32, 8, -5)),        # B seeks 37 samples -> word-aligned 32, int shift 5
    (37, 4, (32, 16, -5)),
    (37, 1, (32, 4, -5)),
    (37, 8, (36, 36, -1)),
    (0, 2, (0, 0, 0)),
    (1_000_000_007, 2, (1_000_000_000, 250_000_000, -7)),
])
def test_word_aligned_start(start, bits, exp):
    assert F.word_aligned_start(start, bits) == exp
