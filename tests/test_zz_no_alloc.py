"""fxc_sector_run_device_noalloc against fxc_sector_run_device, through the C ABI.

The allocation-free entry runs the same kernels with its scratch in the
library's static device array.  Where its grid equals the plan's (every case
here) the fixed-order reduction makes the two results identical bit for bit.
Each case runs two calls in a row, the second on a smaller input, so stale
partial slots left in the shared array by the first call would show.  The
refusals (phased plans, chain-table plans, an over-long geometric table)
return before any launch."""
import numpy as np
import pytest

import fxcorr as F
from _synth import LEVELS_2BIT

pytestmark = pytest.mark.gpu

UNSUPPORTED, INVALID_ARG = 2, 1


def _synth(n, nf, bits, seed, shuffle=None, lsb2=False):
    s = F.Stream()
    bpf = n * bits // 8
    r1, r2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
    F.synth_station_device(r1.ptr, nf * n, bits, None, False, seed, 0, 0.05, 3, s.ptr)
    F.synth_station_device(r2.ptr, nf * n, bits, shuffle, lsb2, seed, 1, 0.05, 3, s.ptr)
    s.sync()
    return r1, r2, bpf


def _cfg(n, fs, table):
    return F.make_delay_cfg(table, frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                            extra_delay_rate_sps=2e-9, clock1_delay_s=1.2e-6,
                            clock1_rate_sps=1e-12, lo1_hz=6.7e9, lo2_hz=6.7e9)


def _table(n_sec):
    t = np.arange(n_sec, dtype=np.float64)
    return 2.1e-7 + 3.3e-11 * t - 4.0e-15 * t * t


def _acc(buf, half):
    out = np.empty(F.N_PRODUCTS * half)
    buf.download(out)
    return out


CASES = [
    # fft, precision, bits, lsb2 + shuffle, frames of the two calls
    (1024, F.PRECISION_CHOSEN, 2, True, (120000, 37001)),     # fused k_sector_mf3
    (1024, F.PRECISION_PLAIN_FP32, 2, True, (60000, 20003)),  # plain fused k_sector
    (4096, F.PRECISION_CHOSEN, 2, False, (9000, 4001)),       # Stockham, MFMA first pass
    (512, F.PRECISION_PLAIN_FP32, 4, False, (30000, 777)),    # plain, 4-bit
]


@pytest.mark.parametrize("n,prec,bits,lsb2,nfs", CASES)
def test_noalloc_matches_device_entry(n, prec, bits, lsb2, nfs):
    print(f"[noalloc] fft={n} precision={prec} bits={bits} frames={nfs}", flush=True)
    fs = 1024e6
    shuffle = np.random.default_rng(7).permutation(32).astype(np.int32) if lsb2 else None
    levels = LEVELS_2BIT if bits == 2 else np.arange(-7.5, 8.0, 1.0)
    nf_max = max(nfs)
    r1, r2, bpf = _synth(n, nf_max, bits, 11, shuffle, lsb2)
    p = F.Plan(n, bits=(bits, bits), levels=[levels, levels],
               shuffles=[None, shuffle], lsb=(False, lsb2), df_hz=fs / n, precision=prec)
    s = F.Stream()
    a_ref = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    a_na = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    cfg = _cfg(n, fs, _table(16))
    first = 0
    for nf in nfs:
        p.sector_run_device(r1.ptr, r2.ptr, first, nf, cfg, a_ref.ptr, stream=s.ptr)
        p.sector_run_device_noalloc(r1.ptr, r2.ptr, first, nf, cfg, a_na.ptr, stream=s.ptr)
        s.sync()
        F.plan_check_device_error(p, s)
        ref, got = _acc(a_ref, p.half), _acc(a_na, p.half)
        assert np.all(np.isfinite(got))
        assert np.any(got != 0.0)
        assert np.array_equal(ref.view(np.uint64), got.view(np.uint64)), \
            (nf, np.max(np.abs(ref - got)))
        first += 1000   # the second call is another, shorter stretch of the scan
    p.close()


def test_noalloc_refusals_before_launch():
    fs = 256e6
    cfg = _cfg(1024, fs, _table(16))
    acc = F.DeviceBuffer(F.N_PRODUCTS * 1025 * 8)
    ph = F.Plan(1024, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / 1024,
                products=F.PRODUCT_PHASED, bit_out=2)
    with pytest.raises(F.FxcError) as e:
        ph.sector_run_device_noalloc(0, 0, 0, 0, cfg, acc.ptr)
    assert e.value.status == UNSUPPORTED
    ch = F.Plan(2048, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / 2048,
                precision=F.PRECISION_PLAIN_FP32)
    with pytest.raises(F.FxcError) as e:
        ch.sector_run_device_noalloc(0, 0, 0, 0, cfg, acc.ptr)
    assert e.value.status == UNSUPPORTED
    ok = F.Plan(1024, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / 1024,
                precision=F.PRECISION_CHOSEN)
    with pytest.raises(F.FxcError) as e:
        ok.sector_run_device_noalloc(0, 0, 0, 0, _cfg(1024, fs, _table(8193)), acc.ptr)
    assert e.value.status == INVALID_ARG
    # an empty call still writes zeros, with no input and no allocation
    s = F.Stream()
    acc.upload(np.full(F.N_PRODUCTS * 513, 1.0), stream=s.ptr)
    ok.sector_run_device_noalloc(0, 0, 0, 0, cfg, acc.ptr, stream=s.ptr)
    z = np.full(F.N_PRODUCTS * 513, 1.0)
    acc.download(z, stream=s.ptr)
    s.sync()
    assert np.all(z == 0.0)
    for p in (ph, ch, ok):
        p.close()
