"""Serving: continuous batching of correlation jobs through the C ABI.

A job's spectra must not depend on what else is admitted with it: the same
job run alone and in a mix (other FFT sizes, forms, bit depths, host and
device input, interleaved chunk arrival) gives bit-identical FP64 sums; and
the chunked sum equals a single sector call over the same frames to the
rounding of the kernel's partial sums."""
import numpy as np
import pytest

import fxcorr as F
from _synth import LEVELS_2BIT, uniform_levels

pytestmark = pytest.mark.gpu


def _synth(n, nf, bits, seed, rho=0.3, delay=3):
    s = F.Stream()
    bpf = n * bits // 8
    r1, r2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
    F.synth_station_device(r1.ptr, nf * n, bits, None, False, seed, 0, rho, delay, s.ptr)
    F.synth_station_device(r2.ptr, nf * n, bits, None, False, seed, 1, rho, delay, s.ptr)
    s.sync()
    return r1, r2, bpf


def _job(n, bits, precision, fs=256e6):
    lv = LEVELS_2BIT if bits == 2 else uniform_levels(bits)
    plan_kw = dict(fft_len=n, bits=(bits, bits), levels=[lv, lv], df_hz=fs / n,
                   precision=precision)
    cfg = F.make_delay_cfg(np.zeros(8), frame_dt=n / fs, fs=fs, delay_user_samples=3.3,
                           extra_delay_rate_sps=2e-9, lo1_hz=5e9, lo2_hz=5e9)
    return plan_kw, cfg


def _run(server, plan_kw, cfg, r1, r2, bpf, chunks, host=False):
    jid = server.submit(plan_kw, cfg, input_on_device=not host)
    off = 0
    for c in chunks:
        if host:
            server.push(jid, r1[off * bpf:(off + c) * bpf], r2[off * bpf:(off + c) * bpf], c)
        else:
            server.push(jid, r1.ptr + off * bpf, r2.ptr + off * bpf, c)
        off += c
    server.close_job(jid)
    return jid


def test_job_alone_equals_job_in_mix_and_single_call():
    nA = 1024
    kwA, cfgA = _job(nA, 2, F.PRECISION_CHOSEN)
    r1, r2, bpf = _synth(nA, 30000, 2, 11)
    chunksA = [7000, 9000, 14000]
    half = nA // 2 + 1

    sv = F.Server(max_active_jobs=4, n_streams=4)
    ja = _run(sv, kwA, cfgA, r1, r2, bpf, chunksA)
    assert sv.wait(ja, 120.0)
    alone, st_alone = sv.result(ja, half)
    sv.release(ja)
    sv.close()

    # the mix: other sizes / forms / bits, host input, A's chunks arrive
    # interleaved with the others'
    kwB, cfgB = _job(4096, 4, F.PRECISION_PLAIN_FP32)
    b1, b2, bpfB = _synth(4096, 3000, 4, 12)
    kwC, cfgC = _job(256, 1, F.PRECISION_CHOSEN)
    c1, c2, bpfC = _synth(256, 50000, 1, 13)
    kwD, cfgD = _job(32768, 2, F.PRECISION_CHOSEN)
    d1, d2, bpfD = _synth(32768, 300, 2, 14)
    hB1 = b1.download(np.zeros(3000 * bpfB, dtype=np.uint8))
    hB2 = b2.download(np.zeros(3000 * bpfB, dtype=np.uint8))

    sv = F.Server(max_active_jobs=4, n_streams=4)
    jb = sv.submit(kwB, cfgB, input_on_device=False)
    ja = sv.submit(kwA, cfgA)
    jc = _run(sv, kwC, cfgC, c1, c2, bpfC, [20000, 30000])
    jd = _run(sv, kwD, cfgD, d1, d2, bpfD, [100, 200])
    off = 0
    for i, c in enumerate(chunksA):
        sv.push(ja, r1.ptr + off * bpf, r2.ptr + off * bpf, c)
        off += c
        if i < 2:
            sv.push(jb, hB1[i * 1500 * bpfB:(i + 1) * 1500 * bpfB],
                    hB2[i * 1500 * bpfB:(i + 1) * 1500 * bpfB], 1500)
    sv.close_job(ja)
    sv.close_job(jb)
    for j in (ja, jb, jc, jd):
        assert sv.wait(j, 300.0)
    mixed, st_mixed = sv.result(ja, half)
    resB, _ = sv.result(jb, 4096 // 2 + 1)
    st = sv.stats()
    sv.close()
    assert np.array_equal(alone, mixed)
    assert st_mixed.frames == sum(chunksA) and st.jobs_done == 4 and st.jobs_failed == 0
    assert st.max_concurrent_jobs >= 2
    assert st.p99_latency_ms >= st.p50_latency_ms > 0

    # chunked sums vs one sector call over the same frames
    p = F.Plan(nA, **{k: v for k, v in kwA.items() if k != "fft_len"})
    acc = F.DeviceBuffer(F.N_PRODUCTS * half * 8)
    p.sector_run_device(r1.ptr, r2.ptr, 0, sum(chunksA), cfgA, acc.ptr)
    one = acc.download(np.zeros(F.N_PRODUCTS * half)).reshape(F.N_PRODUCTS, half)
    got = alone.reshape(F.N_PRODUCTS, half)
    scale = np.abs(one).max(axis=1, keepdims=True) + 1e-300
    # The fused N = 1024 chosen kernel sums each consumer's frames in FP32
    # batches before its FP64 totals; chunking the job regroups frames into
    # other batches, which moves the FP32 rounding (about 1e-7 of the
    # spectrum's scale), far inside the chosen form's 1e-4 bar.
    assert np.all(np.abs(got - one) <= 1e-6 * scale)
    # job B (host input) vs a direct call
    pB = F.Plan(4096, **{k: v for k, v in kwB.items() if k != "fft_len"})
    directB = pB.sector_run(hB1, hB2, 0, 3000, cfgB)
    hB = 4096 // 2 + 1
    rb = resB.reshape(F.N_PRODUCTS, hB)
    for k, row in (("acc11", 1), ("acc22", 2)):
        np.testing.assert_allclose(rb[row], directB[k], rtol=1e-11, atol=0)


def test_budget_limits_concurrency_and_failed_plan_is_isolated():
    n = 1024
    kw, cfg = _job(n, 2, F.PRECISION_CHOSEN)
    r1, r2, bpf = _synth(n, 8000, 2, 21)
    probe = F.Plan(n, **{k: v for k, v in kw.items() if k != "fft_len"})
    plan_bytes = F.plan_device_bytes(probe)
    probe.close()
    # room for one job (plan + spectra), not two
    sv = F.Server(hbm_budget_bytes=int(plan_bytes * 1.5), max_active_jobs=4, n_streams=2)
    jobs = [_run(sv, kw, cfg, r1, r2, bpf, [4000, 4000]) for _ in range(3)]
    bad_kw = dict(kw, fft_len=1000)                      # not a power of two: plan fails
    jbad = sv.submit(bad_kw, cfg)
    sv.close_job(jbad)
    for j in jobs + [jbad]:
        assert sv.wait(j, 120.0)
    res = [sv.result(j, n // 2 + 1)[0] for j in jobs]
    with pytest.raises(F.FxcError):
        sv.result(jbad, n // 2 + 1)
    st = sv.stats()
    sv.close()
    assert st.max_concurrent_jobs == 1 and st.jobs_done == 3 and st.jobs_failed == 1
    assert st.peak_hbm_bytes <= int(plan_bytes * 1.5)
    assert all(np.array_equal(res[0], r) for r in res[1:])


# ------------------------------------------------------------------ graphs
@pytest.mark.parametrize("n,precision", [(1024, F.PRECISION_CHOSEN), (2048, F.PRECISION_PLAIN_FP32),
                                         (32768, F.PRECISION_CHOSEN)])
def test_graph_replays_match_eager(n, precision):
    """Three sectors captured once and replayed three times, with eager work
    and host synchronisation between the replays; every replay equals the
    eager result bit for bit (no memset nodes in the captured work)."""
    nf_sec = 3000 if n < 32768 else 200
    kw, cfg = _job(n, 2, precision)
    r1, r2, bpf = _synth(n, 3 * nf_sec, 2, 31)
    p = F.Plan(n, **{k: v for k, v in kw.items() if k != "fft_len"})
    h = p.half
    s = F.Stream()
    out = F.DeviceBuffer(3 * F.N_PRODUCTS * h * 8)
    other = F.DeviceBuffer(F.N_PRODUCTS * h * 8)

    def sectors():
        for i in range(3):
            p.sector_run_device(r1.ptr + i * nf_sec * bpf, r2.ptr + i * nf_sec * bpf, i * nf_sec,
                                nf_sec, cfg, out.ptr + i * F.N_PRODUCTS * h * 8, stream=s.ptr)

    sectors()                                   # eager (also uploads the delay table)
    s.sync()
    eager = out.download(np.zeros(3 * F.N_PRODUCTS * h))
    g = F.Graph(s)
    sectors()
    g.end()
    for rep in range(3):
        F.check(F.lib().fxc_memset_d(F.C.c_void_p(out.ptr), 0, 3 * F.N_PRODUCTS * h * 8,
                                     F.C.c_void_p(s.ptr)))
        s.sync()
        g.launch()
        s.sync()
        got = out.download(np.zeros(3 * F.N_PRODUCTS * h))
        assert np.array_equal(got, eager), rep
        # eager work on the same plan between replays
        p.sector_run_device(r1.ptr, r2.ptr, 0, nf_sec // 2, cfg, other.ptr, stream=s.ptr)
        s.sync()
    g.close()
    assert np.any(eager != 0)
