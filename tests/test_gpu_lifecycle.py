"""Device-memory lifetime of plans and graphs, through the C ABI.

- Destroying a plan returns every byte it allocated, including the phase
  chain and whole-scan scratch that only some paths grow.
- A live graph pins the scratch of the plans it captured: calls that would
  move that memory (growth) or change what the graph computes (another
  geometric table) fail before any launch, and a plan destroyed while a
  graph holds it is released with the graph's destruction.
- A fused MFMA kernel whose hand-over waits all complete leaves the plan's
  device error word clear."""
import numpy as np
import pytest

import fxcorr as F
from _synth import LEVELS_2BIT

pytestmark = pytest.mark.gpu

MiB = 1 << 20
INVALID_ARG = 1   # FXC_ERR_INVALID_ARG


def _free():
    return F.device_mem_info()[0]


def _synth(n, nf, bits, seed, rho=0.3, delay=3):
    s = F.Stream()
    bpf = n * bits // 8
    r1, r2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
    F.synth_station_device(r1.ptr, nf * n, bits, None, False, seed, 0, rho, delay, s.ptr)
    F.synth_station_device(r2.ptr, nf * n, bits, None, False, seed, 1, rho, delay, s.ptr)
    s.sync()
    return r1, r2, bpf


def _cfg(n, fs, table=None):
    return F.make_delay_cfg(np.zeros(8) if table is None else table, frame_dt=n / fs, fs=fs,
                            delay_user_samples=3.3, extra_delay_rate_sps=2e-9, lo1_hz=5e9,
                            lo2_hz=5e9)


def _plans_that_grow_scratch(data):
    """A phased plan with .raw output (rotation chains and combined spectra
    of 16k-frame chunks), a plain FFT-2048 XCF plan (a 256 MiB f32 phase-chain
    slab) that also runs the whole-scan cross spectrum (its own buffer)."""
    (a1, a2, nfa), (b1, b2, nfb) = data
    s = F.Stream()
    p = F.Plan(1024, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=256e6 / 1024,
               products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT, w=(1.0, 0.7), bit_out=2)
    acc = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    raw = F.DeviceBuffer(nfa * p.info.bpf_out)
    p.sector_run_device(a1.ptr, a2.ptr, 0, nfa, _cfg(1024, 256e6), acc.ptr, raw.ptr, stream=s.ptr)
    q = F.Plan(2048, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=256e6 / 2048,
               precision=F.PRECISION_PLAIN_FP32)
    acc2 = F.DeviceBuffer(F.N_PRODUCTS * q.half * 8)
    ac = F.DeviceBuffer(2 * q.half * 8)
    q.sector_run_device(b1.ptr, b2.ptr, 0, nfb, _cfg(2048, 256e6), acc2.ptr, stream=s.ptr)
    F.fringe_accumulate_device(q, b1.ptr, b2.ptr, 0, nfb, _cfg(2048, 256e6), ac.ptr, stream=s.ptr)
    s.sync()
    for b in (acc, raw, acc2, ac):
        b.free()
    return [p, q]


def test_plan_destroy_returns_all_device_memory():
    data = [_synth(1024, 20000, 2, 3)[:2] + (20000,), _synth(2048, 40000, 2, 4)[:2] + (40000,)]
    for p in _plans_that_grow_scratch(data):   # first use: code objects and runtime pools
        p.close()
    before = _free()
    plans = _plans_that_grow_scratch(data)
    held = sum(F.plan_device_bytes(p) for p in plans)
    assert held > 400 * MiB, held          # the chain slabs and spectra chunks
    during = _free()
    assert before - during >= held - 8 * MiB, (before, during, held)
    for p in plans:
        p.close()
    after = _free()
    print(f"[lifecycle] plans held {held / MiB:.1f} MiB; free before {before / MiB:.1f}, "
          f"after {after / MiB:.1f} MiB", flush=True)
    assert before - after <= 8 * MiB, (before, after)


def test_live_graph_pins_plan_scratch():
    n, nf = 2048, 64
    fs = 256e6
    r1, r2, bpf = _synth(n, 4 * nf, 2, 21)
    p = F.Plan(n, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n,
               precision=F.PRECISION_PLAIN_FP32)
    table = np.zeros(8)
    cfg = F.make_delay_cfg(table, frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                           extra_delay_rate_sps=2e-9, lo1_hz=5e9, lo2_hz=5e9)
    h = p.half
    s = F.Stream()
    out = F.DeviceBuffer(F.N_PRODUCTS * h * 8)
    other = F.DeviceBuffer(F.N_PRODUCTS * h * 8)
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg, out.ptr, stream=s.ptr)
    s.sync()
    eager = out.download(np.zeros(F.N_PRODUCTS * h))
    g = F.Graph(s)
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg, out.ptr, stream=s.ptr)
    g.end()
    assert F.plan_live_graphs(p) == 1
    # growth of the chain slab (more frames) and another table: refused before any launch
    with pytest.raises(F.FxcError) as e:
        p.sector_run_device(r1.ptr, r2.ptr, 0, 4 * nf, cfg, other.ptr, stream=s.ptr)
    assert e.value.status == INVALID_ARG and "live graph" in str(e.value)
    cfg2 = F.make_delay_cfg(table + 1e-9, frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                            extra_delay_rate_sps=2e-9, lo1_hz=5e9, lo2_hz=5e9)
    with pytest.raises(F.FxcError) as e:
        p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg2, other.ptr, stream=s.ptr)
    assert e.value.status == INVALID_ARG and "live graph" in str(e.value)
    # calls that fit the captured scratch still run
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf // 2, cfg, other.ptr, stream=s.ptr)
    s.sync()
    # the plan's handle goes away; the graph keeps its memory and replays
    p.close()
    for _ in range(2):
        F.check(F.lib().fxc_memset_d(F.C.c_void_p(out.ptr), 0, F.N_PRODUCTS * h * 8,
                                     F.C.c_void_p(s.ptr)))
        s.sync()
        g.launch()
        s.sync()
        assert np.array_equal(out.download(np.zeros(F.N_PRODUCTS * h)), eager)
    g.close()
    # a fresh plan is not pinned: it grows as before
    q = F.Plan(n, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n,
               precision=F.PRECISION_PLAIN_FP32)
    q.sector_run_device(r1.ptr, r2.ptr, 0, 4 * nf, cfg, other.ptr, stream=s.ptr)
    s.sync()
    assert F.plan_live_graphs(q) == 0
    q.close()


def test_growth_inside_capture_is_refused():
    n, nf = 2048, 64
    fs = 256e6
    r1, r2, bpf = _synth(n, 4 * nf, 2, 22)
    p = F.Plan(n, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n,
               precision=F.PRECISION_PLAIN_FP32)
    cfg = F.make_delay_cfg(np.zeros(8), frame_dt=n / fs, fs=fs, delay_user_samples=1.5,
                           lo1_hz=5e9, lo2_hz=5e9)
    s = F.Stream()
    out = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg, out.ptr, stream=s.ptr)
    s.sync()
    g = F.Graph(s)
    with pytest.raises(F.FxcError) as e:
        p.sector_run_device(r1.ptr, r2.ptr, 0, 4 * nf, cfg, out.ptr, stream=s.ptr)
    assert "during a graph capture" in str(e.value)
    p.sector_run_device(r1.ptr, r2.ptr, 0, nf, cfg, out.ptr, stream=s.ptr)
    g.end()
    g.launch()
    s.sync()
    g.close()
    p.close()


def test_device_error_word_clear_after_mfma_sectors():
    n = 1024
    fs = 1024e6
    r1, r2, bpf = _synth(n, 20000, 2, 23)
    for prec in (F.PRECISION_CHOSEN,):
        p = F.Plan(n, bits=(2, 2), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=fs / n,
                   precision=prec)
        assert p.info.fft_form == 1
        cfg = F.make_delay_cfg(np.zeros(8), frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                               lo1_hz=5e9, lo2_hz=5e9)
        s = F.Stream()
        out = F.DeviceBuffer(F.N_PRODUCTS * p.half * 8)
        for _ in range(3):
            p.sector_run_device(r1.ptr, r2.ptr, 0, 20000, cfg, out.ptr, stream=s.ptr)
        s.sync()
        F.plan_check_device_error(p, s)      # raises when a hand-over wait ran out
        acc = out.download(np.zeros(F.N_PRODUCTS * p.half))
        assert np.all(np.isfinite(acc))
        p.close()
