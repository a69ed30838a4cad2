"""Streaming host input (fxc_pipeline): sectors submitted from page-locked
host memory give the same spectra, bit for bit, as the device entry on the
same bytes, for pipeline depths 1 and 2 and both arithmetic forms, with
slots reused and submissions running ahead of the waits."""
import numpy as np
import pytest

import fxcorr as F
from _synth import LEVELS_2BIT

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("depth", [1, 2])
@pytest.mark.parametrize("prec", [F.PRECISION_PLAIN_FP32, F.PRECISION_CHOSEN])
def test_pipeline_matches_device_entry(depth, prec):
    n, fs, nf_sec, n_sec = 1024, 512e6, 4000, 5
    bpf = n * 2 // 8
    s = F.Stream()
    d1, d2 = F.DeviceBuffer(n_sec * nf_sec * bpf), F.DeviceBuffer(n_sec * nf_sec * bpf)
    F.synth_station_device(d1.ptr, n_sec * nf_sec * n, 2, None, False, 5, 0, 0.3, 2, s.ptr)
    F.synth_station_device(d2.ptr, n_sec * nf_sec * n, 2, None, True, 5, 1, 0.3, 2, s.ptr)
    s.sync()
    p = F.Plan(n, levels=[LEVELS_2BIT, LEVELS_2BIT], lsb=(False, True), df_hz=fs / n,
               precision=prec)
    cfg = F.make_delay_cfg(np.zeros(4), frame_dt=n / fs, fs=fs, delay_user_samples=2.3,
                           extra_delay_rate_sps=1e-9, lo1_hz=5e9, lo2_hz=5e9)
    h = p.half
    words = F.N_PRODUCTS * h
    ref = F.DeviceBuffer(n_sec * words * 8)
    for i in range(n_sec):
        p.sector_run_device(d1.ptr + i * nf_sec * bpf, d2.ptr + i * nf_sec * bpf, i * nf_sec,
                            nf_sec, cfg, ref.ptr + i * words * 8, stream=s.ptr)
    s.sync()
    expect = ref.download(np.zeros(n_sec * words)).reshape(n_sec, words)

    # host copies of the sectors in page-locked memory; spectra into page-locked memory
    hb1 = [F.HostBuffer(nf_sec * bpf) for _ in range(n_sec)]
    hb2 = [F.HostBuffer(nf_sec * bpf) for _ in range(n_sec)]
    for i in range(n_sec):
        F.check(F.lib().fxc_memcpy_d2h(F.C.c_void_p(hb1[i].ptr), F.C.c_void_p(d1.ptr + i * nf_sec * bpf),
                                       nf_sec * bpf, F.C.c_void_p(0)))
        F.check(F.lib().fxc_memcpy_d2h(F.C.c_void_p(hb2[i].ptr), F.C.c_void_p(d2.ptr + i * nf_sec * bpf),
                                       nf_sec * bpf, F.C.c_void_p(0)))
    out = F.HostBuffer(n_sec * words * 8)
    acc = out.view(np.float64).reshape(n_sec, words)
    acc[:] = np.nan
    pl = F.Pipeline(p, nf_sec, depth)
    tickets = [pl.submit(hb1[i].ptr, hb2[i].ptr, i * nf_sec, nf_sec, cfg, out.ptr + i * words * 8)
               for i in range(n_sec)]
    for t in reversed(tickets):                   # a late ticket covers the earlier ones
        pl.wait(t)
    assert np.array_equal(acc, expect)
    with pytest.raises(F.FxcError):
        pl.submit(hb1[0].ptr, hb2[0].ptr, 0, nf_sec + 1, cfg, out.ptr)
    with pytest.raises(F.FxcError):
        pl.wait(len(tickets))
    pl.close()
