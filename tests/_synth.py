"""Small seeded synthetic inputs for the tests (numpy; sizes the oracle runs
in seconds).  The packing follows the decode contract of utils.rs:215-376:
codes LSB-first in little-endian 32-bit words, then the station's bit map
inverted (so that decode(pack(codes)) == codes)."""
from __future__ import annotations

import numpy as np

IDENT = np.arange(32, dtype=np.int32)
DEFAULT_EXT = ",".join(str(i) for i in range(31, -1, -1))  # args.rs:9-12
LEVELS_2BIT = np.array([-3.0, -1.0, 1.0, 3.0])


def uniform_levels(bits: int) -> np.ndarray:
    n = 1 << bits
    return np.arange(n, dtype=np.float64) - (n - 1) / 2.0


def shuffle_out_words(words: np.ndarray, sh: np.ndarray) -> np.ndarray:
    """out bit sh[pos] <- in bit pos (the inverse of the decode map)."""
    w = words.astype(np.uint64)
    out = np.zeros_like(w)
    for pos in range(32):
        out |= ((w >> np.uint64(pos)) & np.uint64(1)) << np.uint64(int(sh[pos]))
    return out.astype(np.uint32)


def pack_codes(codes: np.ndarray, bits: int, n_per_frame: int, sh=IDENT) -> np.ndarray:
    """Pack codes (nframes*n) frame by frame into nframes*bpf bytes."""
    if bits in (1, 2, 4, 8, 16) and (n_per_frame * bits) % 32 == 0:
        c = np.asarray(codes).reshape(-1, n_per_frame)
        nfr = c.shape[0]
        if bits == 16:
            by = c.astype("<u2").view(np.uint8).reshape(nfr, -1)
        else:
            per = 8 // bits
            cc = c.astype(np.uint8).reshape(nfr, -1, per)
            by = np.zeros(cc.shape[:2], dtype=np.uint8)
            for j in range(per):
                by |= (cc[:, :, j] << np.uint8(bits * j)).astype(np.uint8)
        if not np.array_equal(np.asarray(sh), IDENT):
            words = np.ascontiguousarray(by).view("<u4").reshape(-1)
            by = shuffle_out_words(words, np.asarray(sh)).astype("<u4").view(np.uint8)
        return np.ascontiguousarray(by).reshape(-1)
    codes = np.asarray(codes, dtype=np.uint64).reshape(-1, n_per_frame)
    nfr = codes.shape[0]
    bpf = (n_per_frame * bits + 7) // 8
    nbits = n_per_frame * bits
    nwords = (nbits + 31) // 32
    out = np.zeros((nfr, nwords * 4), dtype=np.uint8)
    # bit stream per frame
    bitpos = np.arange(n_per_frame, dtype=np.int64) * bits
    stream = np.zeros((nfr, nwords * 32), dtype=np.uint8)
    for b in range(bits):
        stream[:, bitpos + b] = ((codes >> np.uint64(b)) & np.uint64(1)).astype(np.uint8)
    weights = (np.uint64(1) << np.arange(32, dtype=np.uint64))
    words = (stream.reshape(nfr, nwords, 32).astype(np.uint64) * weights).sum(axis=2).astype(
        np.uint32)
    if not np.array_equal(np.asarray(sh), IDENT):
        words = shuffle_out_words(words.reshape(-1), np.asarray(sh)).reshape(nfr, nwords)
    out[:] = words.astype("<u4").view(np.uint8).reshape(nfr, nwords * 4)
    return np.ascontiguousarray(out[:, :bpf]).reshape(-1)


def quantize_2bit(x: np.ndarray, thr: float = 0.98) -> np.ndarray:
    c = np.full(x.shape, 3, dtype=np.uint8)
    c[x < thr] = 2
    c[x < 0] = 1
    c[x < -thr] = 0
    return c


def quantize_uniform(x: np.ndarray, bits: int, scale: float) -> np.ndarray:
    n = 1 << bits
    c = np.floor(x / scale + n / 2.0)
    return np.clip(c, 0, n - 1).astype(np.uint16)


def frac_delay_fir(x: np.ndarray, delay: float, taps: int = 63) -> np.ndarray:
    """y[n] = x[n - delay] with a Hann-windowed sinc (integer part + fraction)."""
    d_int = int(np.floor(delay))
    frac = delay - d_int
    k = np.arange(taps) - (taps - 1) // 2
    h = np.sinc(k - frac) * np.hanning(taps + 2)[1:-1]
    h /= h.sum()
    y = np.convolve(x, h, mode="same")
    return np.roll(y, d_int)


def correlated_pair(n: int, rho: float, delay: float, seed: int):
    """Two unit-variance Gaussian streams correlated by rho, B delayed."""
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n)
    a = np.sqrt(rho) * c + np.sqrt(1 - rho) * rng.standard_normal(n)
    cd = frac_delay_fir(c, delay)
    b = np.sqrt(rho) * cd + np.sqrt(1 - rho) * rng.standard_normal(n)
    return a, b


def make_pair_raw(nframes: int, n: int, bits=(2, 2), sh=(IDENT, IDENT), lsb=(False, False),
                  rho=0.05, delay=5.3, seed=0):
    """Packed raw bytes for two stations plus the levels used."""
    a, b = correlated_pair(nframes * n, rho, delay, seed)
    out = []
    levels = []
    for i, x in enumerate((a, b)):
        if lsb[i]:
            x = x.copy()
            x[1::2] *= -1.0
        if bits[i] == 2:
            codes = quantize_2bit(x)
            lv = LEVELS_2BIT
        else:
            codes = quantize_uniform(x, bits[i], 4.0 / (1 << bits[i]) * 2 if bits[i] > 1 else 1.0)
            lv = uniform_levels(bits[i])
        out.append(pack_codes(codes, bits[i], n, sh[i]))
        levels.append(lv)
    return out[0], out[1], levels
