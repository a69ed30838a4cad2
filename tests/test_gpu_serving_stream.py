"""A short seeded served stream shaped like the e-VLBI serving workload,
through the C ABI, with every served job checked against the oracle.

14 jobs arrive as a Poisson process (seeded gaps); 80 % are 1024 Msps 2-bit
FFT-1024 scans (station B LSB with a seeded shuffle), 20 % come from two
mixed-scan types (FFT 4096 8/1-bit at 1024 Msps, FFT 16384 1/2-bit at
256 Msps).  Durations are log-normal (median 0.2 s of data, at most 0.6 s).
Each station's bytes are pushed from page-locked host memory in chunks of at
most 64 MB (input_on_device = 0: the server copies each chunk into the job's
device window, which counts against the HBM budget).  All jobs run the
chosen form, and each job's FP64 sums are held to the FP64 oracle on that
job's own bytes (1e-4, per-channel amplitude and phase printed)."""
import math
import os
import time

import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import fmt_amp_phase, spectra_amp_phase, spectra_err, zero_pattern_equal
from _synth import LEVELS_2BIT, uniform_levels

pytestmark = pytest.mark.gpu

CHUNK_BYTES = 64 << 20
N_JOBS = 14


def _pool(stream, n, bits, fs, nf, shuffle_b, lsb_b, seed):
    """Device-generated correlated pair copied into page-locked host memory."""
    bpf = (n * bits[0] // 8, n * bits[1] // 8)
    host = []
    for st in range(2):
        d = F.DeviceBuffer(nf * bpf[st])
        F.synth_station_device(d.ptr, nf * n, bits[st], shuffle_b if st else None,
                               lsb_b if st else False, seed, st, 0.02, 3, stream.ptr)
        stream.sync()
        h = F.HostBuffer(nf * bpf[st])
        F.check(F.lib().fxc_memcpy_d2h(F.C.c_void_p(h.ptr), F.C.c_void_p(d.ptr), nf * bpf[st],
                                       F.C.c_void_p(0)))
        d.free()
        host.append(h)
    lv = [LEVELS_2BIT if b == 2 else uniform_levels(b) for b in bits]
    shuffles = (None, shuffle_b)
    plan_kw = dict(fft_len=n, bits=bits, levels=lv, shuffles=shuffles, lsb=(False, lsb_b),
                   df_hz=fs / n, products=F.PRODUCT_ACF | F.PRODUCT_XCF,
                   precision=F.PRECISION_CHOSEN)
    dcfg = dict(frame_dt=n / fs, fs=fs, delay_user_samples=3.2, clock2_delay_s=1.2e-6,
                clock2_rate_sps=1e-12, extra_delay_rate_sps=0.05 / 6.7e9, lo1_hz=6.7e9,
                lo2_hz=6.7e9)
    return dict(n=n, bits=bits, fs=fs, bpf=bpf, nf=nf, host=host, lv=lv, shuffles=shuffles,
                lsb=(False, lsb_b), plan_kw=plan_kw, dcfg=dcfg, table=np.zeros(4))


def test_served_stream_from_host_matches_oracle():
    rng = np.random.default_rng(77)
    s = F.Stream()
    sh = rng.permutation(32).astype(np.int32)
    main = _pool(s, 1024, (2, 2), 1024e6, 700_000, sh, True, 5)
    mix = [_pool(s, 4096, (8, 1), 1024e6, 16_000, None, False, 6),
           _pool(s, 16384, (1, 2), 256e6, 16_000, None, False, 7)]
    sigma = math.log(15.0) / 3.0
    jobs = []
    for i in range(N_JOBS):
        t = main if (i == 0 or rng.random() < 0.8) else mix[int(rng.integers(len(mix)))]
        dur = 0.6 if i == 0 else min(0.6, 0.2 * math.exp(sigma * rng.standard_normal()))
        nf = max(1, min(int(dur * t["fs"] / t["n"]), t["nf"]))
        off = int(rng.integers(0, t["nf"] - nf + 1))
        jobs.append((t, nf, off))
    gaps = rng.exponential(0.01, size=N_JOBS)
    print(f"[serve] {N_JOBS} jobs: " + ", ".join(f"FFT{t['n']}x{nf}" for t, nf, _ in jobs),
          flush=True)
    sv = F.Server(hbm_budget_bytes=200 << 30, max_active_jobs=8, n_streams=4)
    ids = []
    t_next = time.perf_counter()
    for (t, nf, off), gap in zip(jobs, gaps):
        t_next += gap
        while time.perf_counter() < t_next:
            time.sleep(0.0005)
        jid = sv.submit(t["plan_kw"], F.make_delay_cfg(t["table"], **t["dcfg"]),
                        input_on_device=False)
        chunk = CHUNK_BYTES // max(t["bpf"])
        done = 0
        while done < nf:
            k = min(chunk, nf - done)
            sv.push(jid, t["host"][0].ptr + (off + done) * t["bpf"][0],
                    t["host"][1].ptr + (off + done) * t["bpf"][1], k)
            done += k
        sv.close_job(jid)
        ids.append(jid)
    for jid in ids:
        assert sv.wait(jid, 300.0), jid
    results = [sv.result(jid, t["n"] // 2 + 1) for jid, (t, _, _) in zip(ids, jobs)]
    st = sv.stats()
    sv.close()
    print(f"[serve] done: p50 {st.p50_latency_ms:.1f} ms, p99 {st.p99_latency_ms:.1f} ms, "
          f"up to {st.max_concurrent_jobs} in flight, peak HBM {st.peak_hbm_bytes / 1e9:.2f} GB",
          flush=True)
    assert st.jobs_done == N_JOBS and st.jobs_failed == 0
    assert st.max_concurrent_jobs >= 2
    threads = max(1, min(os.cpu_count() or 1, 16))
    for i, ((t, nf, off), (acc, js)) in enumerate(zip(jobs, results)):
        assert js.frames == nf
        n, half = t["n"], t["n"] // 2 + 1
        raws = [np.ctypeslib.as_array((F.C.c_uint8 * (nf * t["bpf"][k])).from_address(
            t["host"][k].ptr + off * t["bpf"][k])) for k in range(2)]
        desc = dict(fft_len=n, bits1=t["bits"][0], bits2=t["bits"][1], levels1=t["lv"][0],
                    levels2=t["lv"][1], shuffle1=np.arange(32, dtype=np.int32),
                    shuffle2=(np.arange(32, dtype=np.int32) if t["shuffles"][1] is None
                              else t["shuffles"][1]),
                    lsb1=0, lsb2=int(t["lsb"][1]), a1s=0, a1e=n // 2, a2s=0, df_hz=t["fs"] / n,
                    need_acf=1, need_xcf=1, cpu_threads=threads)
        o64 = R.sector_run(desc, raws[0], raws[1], 0, nf,
                           R.make_delay_cfg(t["table"], **t["dcfg"]), mode=R.MODE_F64,
                           threads=threads)
        a = acc.reshape(F.N_PRODUCTS, half)
        got = dict(acc11=a[1], acc22=a[2], acc12=a[3] + 1j * a[4])
        e = spectra_err(got, o64)
        ap = spectra_amp_phase(got, o64)
        print(f"[serve] job {i} FFT {n} {t['bits'][0]}/{t['bits'][1]}-bit {nf} frames, "
              f"latency {js.latency_ms:.1f} ms: {e}; {fmt_amp_phase(ap)}", flush=True)
        for k in ("acc11", "acc22", "acc12"):
            assert zero_pattern_equal(got[k], o64[k]), (i, k)
            assert e[k] < 1e-4, (i, k, e[k])
            assert ap[k]["amp_max_strong"] < 1e-4 and ap[k]["phase_max_strong"] < 1e-4, (i, k)
    for t in [main] + mix:
        for h in t["host"]:
            h.free()
