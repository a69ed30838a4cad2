"""Pins the oracle's decode / shift / shuffle / quantise (utils.rs:126-410,
args.rs:194-210) against known answers, brute force and independent numpy
formulations."""
import numpy as np
import pytest

import fxref as R
from _synth import DEFAULT_EXT, IDENT, pack_codes, shuffle_out_words, uniform_levels

VSREC_EXT = "24,25,26,27,28,29,30,31,16,17,18,19,20,21,22,23,8,9,10,11,12,13,14,15,0,1,2,3,4,5,6,7"


def test_default_shuffle_is_identity():
    # args.rs:9-12 DEFAULT_SHUFFLE_IN (31..0, MSB first) maps to internal identity
    assert np.array_equal(R.parse_shuffle(DEFAULT_EXT), IDENT)


def test_vsrec_shuffle_reverses_bits_within_bytes():
    # xml.rs:376 example map: internal k -> 8*floor(k/8) + 7 - k%8
    sh = R.parse_shuffle(VSREC_EXT)
    k = np.arange(32)
    assert np.array_equal(sh, 8 * (k // 8) + 7 - k % 8)


@pytest.mark.parametrize("bad", [
    ",".join(str(i) for i in range(31)),              # 31 entries
    ",".join(str(i) for i in [0] * 32),               # not a permutation
    ",".join(str(i) for i in range(1, 33)),           # 32 out of range
    "a," + ",".join(str(i) for i in range(31)),       # not a number
    "-1," + ",".join(str(i) for i in range(1, 32)),   # negative
])
def test_parse_shuffle_rejects(bad):
    with pytest.raises(ValueError):
        R.parse_shuffle(bad)


def test_parse_levels_and_power():
    lv = R.parse_levels(2, "-1.5, -0.5, 0.5, 1.5")
    assert list(lv) == [-1.5, -0.5, 0.5, 1.5]
    assert R.mean_power(lv) == 1.25  # README.md:290-291
    assert R.mean_power([-3, -1, 1, 3]) == 5.0
    with pytest.raises(ValueError):
        R.parse_levels(2, "1,2,3")        # wrong count
    with pytest.raises(ValueError):
        R.parse_levels(0, "1")            # bit depth 0
    with pytest.raises(ValueError):
        R.mean_power([0.0, 0.0])          # P must be > 0


def test_kat_e4():
    # byte 0xE4 = 11 10 01 00: LSB-first codes 0,1,2,3 -> levels -3,-1,1,3
    p = R.DecodePlan(2, IDENT, [-3, -1, 1, 3])
    ok, v = p.decode(bytes([0xE4]), 4, False)
    assert ok and list(v) == [-3, -1, 1, 3]
    ok, v = p.decode(bytes([0xE4]), 4, True)      # LSB->USB negates odd samples
    assert ok and list(v) == [-3, 1, 1, -3]


def test_2bit_lut_brute_force():
    # every u16 word through the 16-bit LUT path, USB and LSB, vs bit arithmetic
    lv = np.array([-3.0, -1.0, 1.0, 3.0])
    p = R.DecodePlan(2, IDENT, lv)
    words = np.arange(65536, dtype=np.uint16)
    raw = words.astype("<u2").view(np.uint8)
    n = 65536 * 8
    for lsb in (False, True):
        ok, v = p.decode(raw, n, lsb)
        assert ok
        codes = (words[:, None].astype(np.uint32) >> (2 * np.arange(8))) & 3
        exp = lv[codes].reshape(-1).astype(np.float32)
        if lsb:
            exp[1::2] *= -1
        assert np.array_equal(v, exp)


def _independent_decode(raw, n, bits, sh, lv, lsb):
    """Decode by a different route: bit matrix of shuffled words, then codes."""
    words = np.frombuffer(raw[: len(raw) // 4 * 4].tobytes(), dtype="<u4").astype(np.uint64)
    bitm = ((words[:, None] >> np.arange(32, dtype=np.uint64)) & 1).astype(np.uint8)
    shuffled = bitm[:, sh]                     # out bit i <- in bit sh[i]
    stream = shuffled.reshape(-1)
    ncodes = min(n, len(stream) // bits)
    idx = np.arange(ncodes)[:, None] * bits + np.arange(bits)[None, :]
    codes = (stream[idx].astype(np.int64) << np.arange(bits)).sum(axis=1)
    out = np.asarray(lv, dtype=np.float32)[codes]
    if lsb:
        out[1::2] *= -1
    return ncodes == n, out


@pytest.mark.parametrize("bits", [1, 2, 3, 4, 5, 8, 12, 16])
@pytest.mark.parametrize("kind", ["ident", "pair", "generic"])
@pytest.mark.parametrize("lsb", [False, True])
def test_decode_matches_independent(bits, kind, lsb):
    rng = np.random.default_rng(bits * 7 + len(kind) + lsb)
    sh = {"ident": IDENT, "pair": IDENT ^ 1, "generic": rng.permutation(32).astype(np.int32)}[kind]
    lv = rng.standard_normal(1 << bits)
    n = 256
    bpf = (n * bits + 7) // 8
    raw = rng.integers(0, 256, size=bpf, dtype=np.uint8)
    p = R.DecodePlan(bits, sh, lv)
    expect_kind = {"ident": 0, "pair": 1, "generic": 2}[kind]
    assert p.shuffle_kind == expect_kind
    ok, v = p.decode(raw, n, lsb)
    ok2, v2 = _independent_decode(raw, n, bits, sh, lv, lsb)
    assert ok == ok2
    if ok:
        assert np.array_equal(v, v2)


def test_decode_length_rule():
    # utils.rs:372-374: only whole 32-bit words decode except 2-bit identity
    lv = [-3, -1, 1, 3]
    p_id = R.DecodePlan(2, IDENT, lv)
    p_sh = R.DecodePlan(2, IDENT ^ 1, lv)
    n = 1030                      # bpf = 258: 64 whole words = 1024 samples
    raw = np.zeros((n * 2 + 7) // 8, dtype=np.uint8)
    assert p_id.decode(raw, n, False)[0]
    assert not p_sh.decode(raw, n, False)[0]


def test_pack_roundtrip_generic_shuffle():
    rng = np.random.default_rng(3)
    sh = rng.permutation(32).astype(np.int32)
    codes = rng.integers(0, 4, size=1024)
    raw = pack_codes(codes, 2, 1024, sh)
    ok, v = R.DecodePlan(2, sh, [0, 1, 2, 3]).decode(raw, 1024, False)
    assert ok and np.array_equal(v.astype(int), codes)


def test_pairswap_equals_generic_pairswap_map():
    # the PairSwap fast path must equal the generic bit loop with map i^1
    rng = np.random.default_rng(5)
    w = rng.integers(0, 2**32, size=1000, dtype=np.uint64).astype(np.uint32)
    p = R.DecodePlan(2, IDENT ^ 1, [0, 1, 2, 3])
    got = np.array([p.shuffle_word(int(x)) for x in w], dtype=np.uint32)
    exp = shuffle_out_words(w, IDENT ^ 1)  # pair swap is its own inverse
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("shift", [0, 1, 5, -3, 1023, -1023, 1024, -5000, 2**62])
def test_shift_closed_form(shift):
    x = np.arange(1, 1025, dtype=np.float32)
    y = R.shift_zerofill(x, shift)
    n = np.arange(1024)
    src = n - shift
    exp = np.where((src >= 0) & (src < 1024), x[np.clip(src, 0, 1023)], 0.0)
    assert np.array_equal(y, exp.astype(np.float32))


@pytest.mark.parametrize("bits", [1, 2, 4, 8])
def test_quantise_roundtrip_nearest_level(bits):
    rng = np.random.default_rng(bits)
    n = 512
    lv = uniform_levels(bits) * 0.7
    sh = rng.permutation(32).astype(np.int32)
    x = (rng.standard_normal(n) * lv.max()).astype(np.float32)
    enc = R.quantise_frame(x, bits, lv, sh)
    assert len(enc) == (n * bits + 31) // 32 * 4
    ok, dec = R.DecodePlan(bits, sh, lv).decode(enc, n, False)
    assert ok
    lvf = lv.astype(np.float32)
    err = np.abs(x[:, None] - lvf[None, :])
    nearest = lvf[np.argmin(err, axis=1)]   # argmin = first minimum, as the scan
    assert np.array_equal(dec, nearest)


def test_quantise_tail_pads_word():
    # utils.rs:400-408: partial tail word is zero padded to 32 bits
    enc = R.quantise_frame(np.zeros(3, dtype=np.float32), 2, [-1.5, -0.5, 0.5, 1.5], IDENT)
    assert len(enc) == 4
