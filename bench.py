#!/usr/bin/env python3
"""Throughput of the FX hot path on one MI355X (Msamples/s/station).

Workload: two stations, 1024 Msps, 2-bit, FFT 1024, 8 s each (2.048 GB packed
per station, 8,000,000 frames, 8 one-second sectors), station B LSB with a
seeded 32-entry bit shuffle, geometric delay from two ECEF positions ~108 m
apart toward a seeded J2000 RA/Dec (the library's 1-s geometric table), clock
1.2 us / 1e-12 s/s, residual 0.2 samples / 0.05 Hz, correlated Gaussian
signal (rho = 0.005) generated on the device.  One step = the whole 8-sector
job (ACF + XCF) with the input already resident in HBM.  Both arithmetic
forms are timed; `value` is the chosen form.  `--sizes` adds both forms'
throughput at FFT sizes 256..65536, `--phased` the phased-array path's.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--sizes]
       [--phased]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "fx-corr_amd"), os.path.join(ROOT, "oracle")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

import fxcorr as F  # noqa: E402

FS = 1024e6
N = 1024
BITS = 2
SECONDS = 8.0
OBS_HZ = 6.7e9
LEVELS = np.array([-3.0, -1.0, 1.0, 3.0])
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
ECEF1 = np.array([-3.50e6, 3.66e6, 3.50e6])
ECEF2 = ECEF1 + np.array([-22.989, -80.501, 67.923])
MJD0 = 60000.25


def workload_setup(seed=2025):
    rng = np.random.default_rng(seed)
    ra, dec = rng.uniform(0, 2 * math.pi), math.radians(rng.uniform(-30, 80))
    # J2000 source precessed to the date; 1-s table over the 8 s scan (main.rs:1628-1642)
    ra, dec = F.precess_j2000_to_date(ra, dec, MJD0)
    n_points = F.geometric_table_points(int(SECONDS * FS / N), N / FS)
    table = F.geometric_delay_table(ECEF1, ECEF2, ra, dec, MJD0, 0.0, n_points)
    clock2, clock2_rate, user, rate_user = 1.2e-6, 1e-12, 0.2, 0.05
    net_d0 = table[0] + user / FS + clock2                       # main.rs:1183-1184
    init = int(round(net_d0 * FS))                               # main.rs:1279
    start2 = (max(init, 0) // 16) * 16                           # word-aligned seek, 2-bit
    start1 = (max(-init, 0) // 16) * 16
    d_seek = (start2 - start1) / FS                              # main.rs:1615-1616
    cfg = dict(frame_dt=N / FS, fs=FS, time_offset_s=0.0, coarse_delay_s=0.0,
               delay_user_samples=user, extra_delay_rate_sps=rate_user / OBS_HZ,
               clock1_delay_s=0.0, clock1_rate_sps=0.0, clock2_delay_s=clock2,
               clock2_rate_sps=clock2_rate, d_seek=d_seek, net_d1_base=0.0,
               total_rate1_base=0.0, total_accel1_base=0.0, net_d2_base=0.0,
               total_rate2_base=0.0, total_accel2_base=0.0, lo1_hz=OBS_HZ, lo2_hz=OBS_HZ)
    residual = int(round((net_d0 - d_seek) * FS))
    shuffle_b = rng.permutation(32).astype(np.int32)
    return table, cfg, shuffle_b, residual


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_1024msps_2bit_fft1024.json")
# Issue counts per frame pair of the fused sector kernels, from the counter
# pass in profiles/sector_kernels_1024msps_2bit_fft1024_kernel_stats.md:
# (VALU instructions, matrix-busy cycles).  A wave64 VALU instruction takes
# 4 cycles of its SIMD; MI355X has 1,024 SIMDs at up to 2.4 GHz.
ISSUE_PER_FRAME = {"k_sector_mf3": (664.8, 512.0), "k_sector_mf3<FP32 batches>": (618.0, 512.0),
                   "k_sector<8, 64, 0>": (2973.0, 0.0)}
SIMDS, CLOCK_HZ = 1024, 2.4e9
COUNTS_FILE = "profiles/sector_kernels_1024msps_2bit_fft1024_kernel_stats.md"
# fused kernel of each form (the MFMA form by kernel organization variant)
KERNEL_KEYS = {("chosen", 0): "k_sector_mf2", ("chosen", 1): "k_sector_mf3",
               ("chosen", 2): "k_sector_mf3<FP32 batches>", ("plain", 0): "k_sector<8, 64, 0>"}
KERNEL_LABELS = {"k_sector_mf2": "fused decode+shift+MFMA four-step FFT+MAC, whole-frame waves",
                 "k_sector_mf3": "fused decode+shift+MFMA four-step FFT+MAC, producer/consumer waves",
                 "k_sector_mf3<FP32 batches>": "fused decode+shift+MFMA four-step FFT+MAC, "
                                               "producer/consumer waves, FP32 batch sums",
                 "k_sector<8, 64, 0>": "fused decode+shift+FFT+MAC (plain FP32)"}


def _pmc_rows(d):
    """(kernel name, counter value) per dispatch from a rocprofv3 --pmc CSV run."""
    import csv
    import glob

    files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no *counter_collection.csv under {d}")
    rows = []
    with open(files[0], newline="") as fh:
        for r in csv.DictReader(fh):
            rows.append((r["Kernel_Name"], float(r["Counter_Value"])))
    return rows


def pmc_summary(fetch_dir, write_dir, out_json, copy_mib):
    """HBM bytes per launch of each kernel: FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) scaled by what they read for the device-to-device copy of known
    size that 'bench.py --copy-mib M' issues last."""
    out = {"kernels": {}}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = _pmc_rows(d)
        copies = [v for k, v in rows if "copyBuffer" in k]
        if not copies:
            raise SystemExit(f"{counter}: no copy kernel dispatch found")
        copy_kib = max(copies)   # the calibration copy is the largest one
        out.setdefault("copy", {})[counter] = copy_kib
        per = {}
        for k, v in rows:
            if "copyBuffer" in k:
                continue
            per.setdefault(k, []).append(v)
        for k, vals in per.items():
            e = out["kernels"].setdefault(k, {"dispatches": len(vals)})
            e[counter + "_kib_avg"] = sum(vals) / len(vals)
    copy_bytes = float(copy_mib) * 2 ** 20
    fs = copy_bytes / 1024.0 / out["copy"]["FETCH_SIZE"]
    ws = copy_bytes / 1024.0 / out["copy"]["WRITE_SIZE"]
    out["copy"].update(bytes=copy_bytes, fetch_scale=fs, write_scale=ws)
    for k, e in out["kernels"].items():
        e["traffic_bytes"] = 1024.0 * (e.get("FETCH_SIZE_kib_avg", 0.0) * fs +
                                       e.get("WRITE_SIZE_kib_avg", 0.0) * ws)
    with open(out_json, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
    print(json.dumps(out, indent=1, sort_keys=True))


def _mix_types(stream, seed=5, n_types=4, pool_bytes=256 << 20):
    """Seeded job types of the mixed-scan set (FFT 256/1024/4096/16384, 1/2/4/8
    bits per station, 128/256/512/1024 Msps), each with a device pool of
    synthetic correlated data (pool_bytes per station at most)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_types):
        n = int(rng.choice([256, 1024, 4096, 16384]))
        bits = (int(rng.choice([1, 2, 4, 8])), int(rng.choice([1, 2, 4, 8])))
        fs = float(rng.choice([128e6, 256e6, 512e6, 1024e6]))
        bpf = (n * bits[0] // 8, n * bits[1] // 8)
        nf = pool_bytes // max(bpf)
        d = (F.DeviceBuffer(nf * bpf[0]), F.DeviceBuffer(nf * bpf[1]))
        for st in range(2):
            F.synth_station_device(d[st].ptr, nf * n, bits[st], None, False, 11, st, 0.01, 3,
                                   stream.ptr)
        lv = [np.arange(1 << b, dtype=np.float64) - ((1 << b) - 1) / 2.0 for b in bits]
        plan_kw = dict(fft_len=n, bits=bits, levels=lv, df_hz=fs / n,
                       products=F.PRODUCT_ACF | F.PRODUCT_XCF)
        cfg = F.make_delay_cfg(np.zeros(4), frame_dt=n / fs, fs=fs, delay_user_samples=3.2,
                               extra_delay_rate_sps=1e-9, lo1_hz=6.7e9, lo2_hz=6.7e9)
        out.append(dict(n=n, bits=bits, fs=fs, bpf=bpf, nf=nf, dev=d, plan_kw=plan_kw, cfg=cfg))
    stream.sync()
    return out


def _to_host(dev, nbytes):
    """A page-locked host copy of the first nbytes of a device buffer."""
    h = F.HostBuffer(nbytes)
    F.check(F.lib().fxc_memcpy_d2h(F.C.c_void_p(h.ptr), F.C.c_void_p(dev.ptr), nbytes,
                                   F.C.c_void_p(0)))
    return h


def serve_bench(raw1, raw2, bpf, cfg, shuffle_b, n_jobs, load, seed=7,
                host_pool_frames=2_000_000):
    """Served correlation stream.  Jobs arrive as a Poisson process; 80 % are
    1024 Msps 2-bit FFT-1024 scans, 20 % are drawn from four seeded types of
    the mixed-scan set.  Durations are log-normal (median 4 s of data, at
    most 60 s).  Each station's data is pushed in 64 MB chunks from
    page-locked host memory (the server copies each chunk into the job's
    device input window, which counts against the 200 GB HBM budget), and
    every job runs the chosen form; the latencies therefore include the PCIe
    copies.  The same job list is then run with the plain FP32 form one job
    at a time from the same host memory, and once more in the chosen form
    with the data already resident in HBM (device pointers pushed)."""
    stream = F.Stream()
    mix = _mix_types(stream)
    rng = np.random.default_rng(seed)
    main_pool = int(SECONDS * FS / N)
    main = dict(n=N, fs=FS, bpf=(bpf, bpf), nf=main_pool, dev=(raw1, raw2), cfg=cfg,
                plan_kw=dict(fft_len=N, bits=(BITS, BITS), levels=[LEVELS, LEVELS],
                             shuffles=(None, shuffle_b), lsb=(False, True),
                             band=(0, 0, N // 2, 0), df_hz=FS / N,
                             products=F.PRODUCT_ACF | F.PRODUCT_XCF))
    # host copies of the data pools (jobs cycle through them)
    main["host_nf"] = min(main_pool, host_pool_frames)
    main["host"] = tuple(_to_host(d, main["host_nf"] * bpf) for d in (raw1, raw2))
    for t in mix:
        t["host_nf"] = t["nf"]
        t["host"] = tuple(_to_host(t["dev"][k], t["nf"] * t["bpf"][k]) for k in range(2))
    # job list: type, duration (s of data) -> frames
    sigma = math.log(15.0) / 3.0                          # 60 s at about +3 sigma of 4 s
    jobs = []
    for _ in range(n_jobs):
        t = main if rng.random() < 0.8 else mix[int(rng.integers(len(mix)))]
        dur = min(60.0, 4.0 * math.exp(sigma * rng.standard_normal()))
        jobs.append((t, max(1, int(dur * t["fs"] / t["n"]))))

    def run(precision, max_active, arrivals, host):
        sv = F.Server(hbm_budget_bytes=200 << 30, max_active_jobs=max_active, n_streams=4)
        for t in [main] + mix:                            # warm the plan cache
            kw = dict(t["plan_kw"], precision=precision)
            j = sv.submit(kw, t["cfg"])
            sv.push(j, t["dev"][0].ptr, t["dev"][1].ptr, min(1024, t["nf"]))
            sv.close_job(j)
            if not sv.wait(j, 120.0):
                raise RuntimeError("serving warm-up job did not finish")
            sv.release(j)
        ids, samples = [], 0
        t0 = time.perf_counter()
        t_next = t0
        for (t, nfj), gap in zip(jobs, arrivals):
            t_next += gap
            while time.perf_counter() < t_next:
                pass
            kw = dict(t["plan_kw"], precision=precision)
            jid = sv.submit(kw, t["cfg"], input_on_device=not host)
            chunk = (64 << 20) // max(t["bpf"])
            src, pool = (t["host"], t["host_nf"]) if host else (t["dev"], t["nf"])
            done = 0
            while done < nfj:
                k = min(chunk, nfj - done)
                off = done % max(1, pool - k)             # cycle through the data pool
                sv.push(jid, src[0].ptr + off * t["bpf"][0], src[1].ptr + off * t["bpf"][1], k)
                done += k
            sv.close_job(jid)
            ids.append(jid)
            samples += nfj * t["n"]
        for jid in ids:
            if not sv.wait(jid, 600.0):
                raise RuntimeError(f"serving job {jid} did not finish")
        wall = time.perf_counter() - t0
        st = sv.stats()
        sv.close()
        return samples, wall, st

    def served(host):
        # service rate of the chosen form from one closed-loop pass (no
        # arrivals), then Poisson arrivals at `load` times that rate
        samples, wall, _ = run(F.PRECISION_CHOSEN, 8, [0.0] * n_jobs, host)
        cap = samples / wall / 1e6
        gaps = list(rng.exponential(wall / (load * n_jobs), size=n_jobs))
        samples, wall, st = run(F.PRECISION_CHOSEN, 8, gaps, host)
        where = "host" if host else "device"
        print(f"[bench] serving ({where} input): closed loop {cap:,.0f}, Poisson "
              f"{samples / wall / 1e6:,.0f} Msamples/s/station, p50 {st.p50_latency_ms:.1f} ms "
              f"p99 {st.p99_latency_ms:.1f} ms", file=sys.stderr, flush=True)
        return {"value": round(samples / wall / 1e6, 1), "unit": "Msamples/s/station",
                "closed_loop_capacity": round(cap, 1),
                "p50_latency_ms": round(st.p50_latency_ms, 3),
                "p99_latency_ms": round(st.p99_latency_ms, 3),
                "mean_latency_ms": round(st.mean_latency_ms, 3),
                "max_concurrent_jobs": st.max_concurrent_jobs,
                "peak_hbm_gb": round(st.peak_hbm_bytes / 1e9, 3), "failed": st.jobs_failed}

    out = served(True)
    p_samples, p_wall, p_st = run(F.PRECISION_PLAIN_FP32, 1, [0.0] * n_jobs, True)
    dev = served(False)
    for t in [main] + mix:
        for h in t["host"]:
            h.free()
    out.update({"jobs": n_jobs, "offered_load": load,
                "job": "80% 1024 Msps 2-bit FFT-1024, 20% from 4 mixed-scan types; log-normal "
                       "durations (median 4 s, max 60 s of data); 64 MB chunks per station "
                       "pushed from page-locked host memory (PCIe copies inside the latencies)",
                "plain_one_at_a_time": {"value": round(p_samples / p_wall / 1e6, 1),
                                        "p50_latency_ms": round(p_st.p50_latency_ms, 3),
                                        "p99_latency_ms": round(p_st.p99_latency_ms, 3),
                                        "input": "the same host memory"},
                "device_resident": dev,
                "mix_types": [f"FFT {t['n']}, {t['bits'][0]}/{t['bits'][1]}-bit, "
                              f"{t['fs'] / 1e6:.0f} Msps" for t in mix]})
    return out


def issue_fractions(key, frames, launch_ms):
    """The compute bound of a fused kernel made explicit: the kernels issue
    VALU and MFMA work from the same SIMDs (compute-bound, the FFT flops
    priced against the FP32 vector peak), so report the share of the SIMDs'
    cycles that its VALU instructions and its matrix-busy cycles take at the
    measured launch time."""
    if key not in ISSUE_PER_FRAME:
        return {}
    valu, mfma = ISSUE_PER_FRAME[key]
    cycles = SIMDS * CLOCK_HZ * launch_ms * 1e-3
    return {"issue": "VALU + MFMA issue on the same SIMDs (peak: FP32 vector)",
            "valu_issue_frac": round(frames * valu * 4 / cycles, 4),
            "mfma_busy_frac": round(frames * mfma / cycles, 4),
            "issue_counts_source": COUNTS_FILE}


def load_traffic(key):
    """Bytes per launch of a fused kernel from the committed summary."""
    try:
        with open(TRAFFIC_FILE) as fh:
            kern = json.load(fh)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    for name, e in kern.items():
        if key in name:
            return e.get("traffic_bytes")
    return None


FORM_NAMES = {0: "FP32 Stockham", 1: "FP16-MFMA four-step", 2: "FP32 four-step through HBM"}


def form_name(info):
    """The FFT a plan runs (fxc_plan_info.fft_form / fft_mfma)."""
    if info.fft_mfma and info.variant == 1 and (info.fft_form == 2 or info.fft_len == 16384):
        return "fused four-step: decode + FP16-MFMA DFT-128 columns, row DFTs + products"
    if info.fft_form == 2 and info.fft_mfma:
        return "four-step through HBM, FP16-MFMA DFT-32 first pass"
    if info.fft_form == 0 and info.fft_mfma:
        return "Stockham, FP16-MFMA DFT-32 first pass"
    return FORM_NAMES[info.fft_form]


def size_sweep(stream, samples_per_station=1 << 28, reps=5):
    """Throughput of both forms at FFT sizes 256..65536: two 2-bit stations at
    1024 Msps, device-resident synthetic input (station B LSB), the plans'
    default configuration, `reps` timed sector runs after one warm-up."""
    rows = {}
    fs = 1024e6
    for n in [256, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536]:
        nf = samples_per_station // n
        bpf = n * 2 // 8
        d1, d2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
        F.synth_station_device(d1.ptr, nf * n, 2, None, False, 11, 0, 0.01, 3, stream.ptr)
        F.synth_station_device(d2.ptr, nf * n, 2, None, True, 11, 1, 0.01, 3, stream.ptr)
        stream.sync()
        cfg = F.make_delay_cfg(np.zeros(2 + int(nf * n / fs)), frame_dt=n / fs, fs=fs,
                               delay_user_samples=2.3, lo1_hz=OBS_HZ, lo2_hz=OBS_HZ)
        row = {"frames": nf}
        for name, prec in (("plain", F.PRECISION_PLAIN_FP32), ("chosen", F.PRECISION_CHOSEN)):
            p = F.Plan(n, levels=[LEVELS, LEVELS], lsb=(False, True), df_hz=fs / n, precision=prec)
            acc = F.DeviceBuffer(p.half * 5 * 8)
            p.sector_run_device(d1.ptr, d2.ptr, 0, nf, cfg, acc.ptr, stream=stream.ptr)
            stream.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                p.sector_run_device(d1.ptr, d2.ptr, 0, nf, cfg, acc.ptr, stream=stream.ptr)
            stream.sync()
            el = (time.perf_counter() - t0) / reps
            row[name] = round(nf * n / el / 1e6, 1)
            row[name + "_fft"] = form_name(p.info)
            acc.free()
            p.close()
        d1.free()
        d2.free()
        rows[str(n)] = row
        print(f"[bench] FFT {n:6d}: plain {row['plain']:,.0f}, chosen {row['chosen']:,.0f} "
              f"Msamples/s/station ({row['chosen_fft']})", file=sys.stderr, flush=True)
    return {"workload": f"two 2-bit stations at 1024 Msps, {samples_per_station} samples per "
                        f"station per size, input resident in HBM, default configuration, "
                        f"mean of {reps} sector runs", "unit": "Msamples/s/station", "sizes": rows}


def phased_sweep(stream, samples_per_station=1 << 28, reps=3):
    """Phased-array path (yi-phasedarray): two 2-bit stations at 1024 Msps
    summed into the phased power spectrum, without and with the 2-bit .raw
    output, at FFT sizes 1024, 4096, 16384 (fused LDS kernel) and 65536 (HBM
    four-step), in both forms (plain: the reference's f32 rotation chains;
    chosen: FP64 rotation phases and the FP16-MFMA first FFT pass);
    device-resident input, mean of `reps` sector runs after one warm-up."""
    rows = {}
    fs = 1024e6
    for n in [1024, 4096, 16384, 65536]:
        nf = samples_per_station // n
        bpf = n * 2 // 8
        d1, d2 = F.DeviceBuffer(nf * bpf), F.DeviceBuffer(nf * bpf)
        F.synth_station_device(d1.ptr, nf * n, 2, None, False, 13, 0, 0.3, 2, stream.ptr)
        F.synth_station_device(d2.ptr, nf * n, 2, None, False, 13, 1, 0.3, 2, stream.ptr)
        stream.sync()
        cfg = F.make_delay_cfg(np.zeros(2 + int(nf * n / fs)), frame_dt=n / fs, fs=fs,
                               delay_user_samples=2.3, lo1_hz=OBS_HZ, lo2_hz=OBS_HZ)
        row = {"frames": nf}
        for name, raw, prec in (("power", False, F.PRECISION_PLAIN_FP32),
                                ("power_and_raw", True, F.PRECISION_PLAIN_FP32),
                                ("chosen_power", False, F.PRECISION_CHOSEN),
                                ("chosen_power_and_raw", True, F.PRECISION_CHOSEN)):
            prods = F.PRODUCT_PHASED | (F.PRODUCT_RAW_OUT if raw else 0)
            p = F.Plan(n, levels=[LEVELS, LEVELS], df_hz=fs / n, products=prods, bit_out=2,
                       precision=prec)
            acc = F.DeviceBuffer(p.half * 5 * 8)
            out = F.DeviceBuffer(nf * p.info.bpf_out) if raw else None
            optr = out.ptr if raw else None
            p.sector_run_device(d1.ptr, d2.ptr, 0, nf, cfg, acc.ptr, optr, stream=stream.ptr)
            stream.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                p.sector_run_device(d1.ptr, d2.ptr, 0, nf, cfg, acc.ptr, optr, stream=stream.ptr)
            stream.sync()
            el = (time.perf_counter() - t0) / reps
            row[name] = round(nf * n / el / 1e6, 1)
            acc.free()
            if out is not None:
                out.free()
            p.close()
        d1.free()
        d2.free()
        rows[str(n)] = row
        print(f"[bench] phased FFT {n:6d}: power {row['power']:,.0f}, power + .raw "
              f"{row['power_and_raw']:,.0f}; chosen {row['chosen_power']:,.0f} / "
              f"{row['chosen_power_and_raw']:,.0f} Msamples/s/station", file=sys.stderr, flush=True)
    return {"workload": f"two 2-bit stations at 1024 Msps, weights 1/1, 2-bit .raw output, "
                        f"{samples_per_station} samples per station per size, input resident "
                        f"in HBM, mean of {reps} sector runs", "unit": "Msamples/s/station",
            "sizes": rows}


def _uniform_levels(bits):
    return np.arange(1 << bits, dtype=np.float64) - ((1 << bits) - 1) / 2.0


def batch_bench(stream, scale=0.125, reps=3, seed=48, branches=4):
    """Batched uneven scans in one graph: 48 seeded scans, FFT sizes from
    {256, 1024, 4096, 16384}, per-station bits from {1, 2, 4, 8}, sampling
    rates from {128, 256, 512, 1024} Msps, lengths uniform in 0.5..20 s
    scaled by `scale` (device-generated input), each with its own delay
    model and clock terms; ACF+XCF of every scan, and 12 scans also phased
    into an 8-bit series.  Per form, every ACF/XCF plan is autotuned on its
    own scan, all sector calls are captured into one graph (as parallel
    branches on `branches` forked streams), and the graph is
    replayed `reps` times after one warm-up replay; the phased plans run the
    form's own arithmetic (chosen: rotation phases in FP64, no f32 chains)."""
    rng = np.random.default_rng(seed)
    scans, total = [], 0
    for i in range(48):
        n = int(rng.choice([256, 1024, 4096, 16384]))
        bits = [int(rng.choice([1, 2, 4, 8])), int(rng.choice([1, 2, 4, 8]))]
        phased = i < 12
        if phased and max(bits) < 8:
            bits[int(rng.integers(2))] = 8
        fs = float(rng.choice([128e6, 256e6, 512e6, 1024e6]))
        nf = max(16, int(float(rng.uniform(0.5, 20.0)) * scale * fs / n))
        delay = float(rng.uniform(-12.0, 12.0))
        lo = float(rng.uniform(2e9, 9e9))
        sh1 = rng.permutation(32).astype(np.int32) if rng.integers(2) else None
        lsb2 = bool(rng.integers(2))
        w = (float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.5, 1.5)))
        d1 = F.DeviceBuffer(nf * n * bits[0] // 8)
        d2 = F.DeviceBuffer(nf * n * bits[1] // 8)
        F.synth_station_device(d1.ptr, nf * n, bits[0], sh1, False, 500 + i, 0, 0.3, 3, stream.ptr)
        F.synth_station_device(d2.ptr, nf * n, bits[1], None, lsb2, 500 + i, 1, 0.3, 3, stream.ptr)
        cfg = F.make_delay_cfg(np.zeros(2 + int(nf * n / fs)), frame_dt=n / fs, fs=fs,
                               delay_user_samples=delay,
                               extra_delay_rate_sps=float(rng.uniform(-5e-9, 5e-9)),
                               clock1_delay_s=float(rng.uniform(-2, 2)) / fs,
                               clock1_rate_sps=float(rng.uniform(-1e-12, 1e-12)),
                               clock2_delay_s=float(rng.uniform(-2, 2)) / fs,
                               clock2_rate_sps=float(rng.uniform(-1e-12, 1e-12)),
                               lo1_hz=lo, lo2_hz=lo)
        scans.append(dict(n=n, bits=tuple(bits), fs=fs, nf=nf, d1=d1, d2=d2, cfg=cfg, sh1=sh1,
                          lsb=(False, lsb2), phased=phased, w=w,
                          output_lsb=bool(rng.integers(2))))
        total += nf * n
    stream.sync()
    gbytes = sum(sc["d1"].nbytes + sc["d2"].nbytes for sc in scans) / 1e9
    out = {"workload": f"48 scans (12 also phased, 8-bit series), FFT 256/1024/4096/16384, "
                       f"1/2/4/8-bit, 128..1024 Msps, lengths 0.5..20 s x {scale}, "
                       f"device-generated, one graph per form",
           "samples_per_station": total, "input_gb": round(gbytes, 2), "unit": "Msamples/s/station"}
    for form, prec in (("plain", F.PRECISION_PLAIN_FP32), ("chosen", F.PRECISION_CHOSEN)):
        jobs = []
        for sc in scans:
            n, lv = sc["n"], [_uniform_levels(b) for b in sc["bits"]]
            p = F.Plan(n, bits=sc["bits"], levels=lv, shuffles=(sc["sh1"], None), lsb=sc["lsb"],
                       df_hz=sc["fs"] / n, products=F.PRODUCT_ACF | F.PRODUCT_XCF, precision=prec)
            F.plan_autotune(p, sc["d1"].ptr, sc["d2"].ptr, sc["nf"], sc["cfg"], stream.ptr)
            job = dict(sc=sc, plan=p, acc=F.DeviceBuffer(F.N_PRODUCTS * (n // 2 + 1) * 8))
            if sc["phased"]:
                pp = F.Plan(n, bits=sc["bits"], levels=lv, shuffles=(sc["sh1"], None),
                            lsb=sc["lsb"], df_hz=sc["fs"] / n,
                            products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT, w=sc["w"],
                            bit_out=max(sc["bits"]), output_lsb=sc["output_lsb"], precision=prec)
                job.update(pplan=pp, pacc=F.DeviceBuffer(F.N_PRODUCTS * (n // 2 + 1) * 8),
                           praw=F.DeviceBuffer(sc["nf"] * pp.info.bpf_out))
            jobs.append(job)

        def enqueue_all(streams=(stream,)):
            for k, j in enumerate(jobs):
                sc, st = j["sc"], streams[k % len(streams)]
                j["plan"].sector_run_device(sc["d1"].ptr, sc["d2"].ptr, 0, sc["nf"], sc["cfg"],
                                            j["acc"].ptr, stream=st.ptr)
                if "pplan" in j:
                    j["pplan"].sector_run_device(sc["d1"].ptr, sc["d2"].ptr, 0, sc["nf"], sc["cfg"],
                                                 j["pacc"].ptr, j["praw"].ptr, stream=st.ptr)

        enqueue_all()                       # eager: delay tables, scratch sizes
        stream.sync()
        # the graph forks onto `branches` streams: the scans run as parallel
        # branches, so one scan's launch tail overlaps the next scan's work
        g = F.Graph(stream)
        side = [F.Stream() for _ in range(branches)]
        fork = F.Event()
        fork.record(stream)
        for st in side:
            st.wait(fork)
        enqueue_all(side)
        for st in side:
            ev = F.Event()
            ev.record(st)
            stream.wait(ev)
        g.end()
        g.launch()
        stream.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            g.launch()
        stream.sync()
        el = (time.perf_counter() - t0) / reps
        out[form] = {"value": round(total / el / 1e6, 1), "ms_per_batch": round(el * 1e3, 3),
                     "graph_branches": branches}
        print(f"[bench] batch {form}: {total / el / 1e6:,.0f} Msamples/s/station "
              f"({el * 1e3:.1f} ms per 48-scan graph)", file=sys.stderr, flush=True)
        g.close()
        for j in jobs:
            for k in ("acc", "pacc", "praw"):
                if k in j:
                    j[k].free()
            j["plan"].close()
            if "pplan" in j:
                j["pplan"].close()
    for sc in scans:
        sc["d1"].free()
        sc["d2"].free()
    out["chosen_over_plain"] = round(out["chosen"]["value"] / out["plain"]["value"], 3)
    return out


def host_identity(threads_used):
    """CPU model and logical core count of the machine the baseline ran on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = None
    return {"cpu_model": model, "logical_cores": os.cpu_count(), "cores_available": affinity,
            "threads_used": threads_used}


def amp_phase(got, ref):
    """Per-channel amplitude error ||got|-|ref||/|ref| and phase error
    |arg(got conj(ref))| (radians) over channels with ref != 0: max and RMS,
    and the max over channels >= 0.1x the reference's RMS magnitude."""
    got = np.asarray(got).astype(np.complex128)
    ref = np.asarray(ref).astype(np.complex128)
    m = np.abs(ref) > 0
    g, r = got[m], ref[m]
    amp = np.abs(np.abs(g) - np.abs(r)) / np.abs(r)
    ph = np.abs(np.angle(g * np.conj(r)))
    strong = np.abs(r) >= 0.1 * np.sqrt(np.mean(np.abs(r) ** 2))
    return {"amp_max": float(amp.max()), "amp_rms": float(np.sqrt(np.mean(amp ** 2))),
            "phase_max_rad": float(ph.max()), "phase_rms_rad": float(np.sqrt(np.mean(ph ** 2))),
            "amp_max_strong": float(amp[strong].max()),
            "phase_max_strong_rad": float(ph[strong].max())}


def sector_parity(acc, ref, half, tol):
    """One sector's device sums ([ph|11|22|12re|12im] x half) against an
    oracle result dict: per-channel amplitude/phase errors of ACF1, ACF2, XCF."""
    acc = np.asarray(acc).reshape(F.N_PRODUCTS, half)
    got = {"acf1": acc[1], "acf2": acc[2], "xcf": acc[3] + 1j * acc[4]}
    want = {"acf1": ref["acc11"], "acf2": ref["acc22"], "xcf": ref["acc12"]}
    out = {k: {kk: float(f"{vv:.3e}") for kk, vv in amp_phase(got[k], want[k]).items()}
           for k in got}
    out["tol"] = tol
    out["within_tol"] = all(v["amp_max_strong"] < tol and v["phase_max_strong_rad"] < tol
                            for k, v in out.items() if isinstance(v, dict))
    return out


def c1_cpu_legs(stream, threads, seed=0):
    """The 64 Msps 2-bit FFT-1024 one-second two-station case on the host
    cores: compute only (input in RAM) and file-inclusive (both stations'
    .raw files read from disk at their word-aligned starts, then the same
    sector).  Input: device-generated correlated 2-bit pair (rho 0.02, B
    delayed 37 samples), written to a temporary directory first."""
    import tempfile

    import fxref as R

    fs, n, bits = 64e6, 1024, 2
    samples = int(fs)
    bpf = n * bits // 8
    d1, d2 = F.DeviceBuffer(samples * bits // 8), F.DeviceBuffer(samples * bits // 8)
    F.synth_station_device(d1.ptr, samples, bits, None, False, seed, 0, 0.02, 37, stream.ptr)
    F.synth_station_device(d2.ptr, samples, bits, None, False, seed, 1, 0.02, 37, stream.ptr)
    stream.sync()
    h1 = d1.download(np.zeros(samples * bits // 8, dtype=np.uint8))
    h2 = d2.download(np.zeros(samples * bits // 8, dtype=np.uint8))
    d1.free()
    d2.free()
    start2 = (37 // 16) * 16                                       # word-aligned seek: 32
    nf = (samples - start2) // n                                   # 62,499 frames
    desc = dict(fft_len=n, bits1=bits, bits2=bits, levels1=LEVELS, levels2=LEVELS,
                shuffle1=np.arange(32, dtype=np.int32), shuffle2=np.arange(32, dtype=np.int32),
                lsb1=0, lsb2=0, a1s=0, a1e=n // 2, a2s=0, df_hz=fs / n, need_acf=1, need_xcf=1,
                cpu_threads=max(1, threads - 2))
    rcfg = R.make_delay_cfg(np.zeros(3), frame_dt=n / fs, fs=fs, delay_user_samples=37.3,
                            d_seek=start2 / fs, lo1_hz=6.7e9, lo2_hz=6.7e9)
    o1 = h1[:nf * bpf]
    o2 = h2[start2 * bits // 8:start2 * bits // 8 + nf * bpf]
    t0 = time.perf_counter()
    R.sector_run(desc, o1, o2, 0, nf, rcfg, mode=R.MODE_F32, threads=threads)
    el_c = time.perf_counter() - t0
    with tempfile.TemporaryDirectory() as tmp:
        p1, p2 = os.path.join(tmp, "A.raw"), os.path.join(tmp, "B.raw")
        h1.tofile(p1)
        h2.tofile(p2)
        t0 = time.perf_counter()
        r1 = np.fromfile(p1, dtype=np.uint8, count=nf * bpf, offset=0)
        r2 = np.fromfile(p2, dtype=np.uint8, count=nf * bpf, offset=start2 * bits // 8)
        R.sector_run(desc, r1, r2, 0, nf, rcfg, mode=R.MODE_F32, threads=threads)
        el_f = time.perf_counter() - t0
    return {"compute_only": round(nf * n / el_c / 1e6, 1),
            "file_inclusive": round(nf * n / el_f / 1e6, 1), "unit": "Msamples/s/station",
            "sample": f"64 Msps 2-bit FFT 1024, {nf} frames (1 s, seek 32), oracle FP32 sector "
                      f"runner, {threads} threads; file-inclusive reads both .raw files (just "
                      f"written, so from the page cache) with the sector; "
                      f"{el_c:.2f} s / {el_f:.2f} s wall"}


def c3_bench(stream, reps=2, seed=2303, forms=("plain", "chosen"), n_sectors=30):
    """Precision-stress workload: 512 Msps, A 4-bit / B 1-bit, FFT 65536,
    16 MHz centre offset (band shift 2048, overlap 30,720), continuum
    coherence 1e-3 plus a 3-channel Gaussian line of peak coherence 0.05 at a
    seeded channel, 30 s = 234,375 frames in 30 sectors, generated on the
    device.  Times the 30-sector job (ACF + XCF) in both forms with the input
    resident in HBM, and checks each form's integrated spectrum: the line is
    the strongest channel with the expected coherence, the continuum has the
    expected mean coherence, and the forms agree.  A profile may run the
    first n_sectors sectors of one form only."""
    n, fs, bits = 65536, 512e6, (4, 1)
    total = int(30 * fs // n)
    sectors = F.sector_partition(total, n / fs)[:n_sectors]
    total = int(sum(sectors))
    band = F.band_alignment(n, fs, 8400.0, 8384.0, 256.0, 256.0)
    line_bin = int(np.random.default_rng(seed).integers(2000, 28000))
    lv = (_uniform_levels(4), _uniform_levels(1))
    bpf = (n * bits[0] // 8, n * bits[1] // 8)
    d1, d2 = F.DeviceBuffer(total * bpf[0]), F.DeviceBuffer(total * bpf[1])
    t0 = time.perf_counter()
    F.synth_spectral_pair_device(d1.ptr, d2.ptr, n, 0, total, bits=bits, shift_bins=band[0],
                                 rho_continuum=1e-3, rho_line=0.05, line_bin=line_bin,
                                 line_fwhm_bins=3.0, clip_sigma=4.0, seed=seed, stream=stream.ptr)
    gen_s = time.perf_counter() - t0
    cfg = F.make_delay_cfg(None, frame_dt=n / fs, fs=fs)
    half = n // 2 + 1
    acc = F.DeviceBuffer(len(sectors) * F.N_PRODUCTS * half * 8)
    out, tot = {}, {}
    for form, prec in (("plain", F.PRECISION_PLAIN_FP32), ("chosen", F.PRECISION_CHOSEN)):
        if form not in forms:
            continue
        plan = F.Plan(n, bits=bits, levels=list(lv), band=band, df_hz=fs / n,
                      products=F.PRODUCT_ACF | F.PRODUCT_XCF, precision=prec)

        def job():
            first = 0
            for i, nf in enumerate(sectors):
                plan.sector_run_device(d1.ptr + first * bpf[0], d2.ptr + first * bpf[1], first, nf,
                                       cfg, acc.ptr + i * F.N_PRODUCTS * half * 8, stream=stream.ptr)
                first += nf

        job()
        stream.sync()
        print(f"[bench] C3 {form}: warm-up job done", file=sys.stderr, flush=True)
        plan.set_timing(True)
        e0, e1 = F.Event(), F.Event()
        e0.record(stream)
        for _ in range(reps):
            job()
        e1.record(stream)
        stream.sync()
        ms = e0.elapsed_ms(e1) / reps
        kms, launches = plan.get_timing()
        plan.set_timing(False)
        F.plan_check_device_error(plan, stream)
        a = acc.download(np.zeros(len(sectors) * F.N_PRODUCTS * half)).reshape(
            len(sectors), F.N_PRODUCTS, half).sum(axis=0)
        tot[form] = a
        out[form] = {"value": round(total * n / (ms * 1e-3) / 1e6, 1), "ms_per_job": round(ms, 3),
                     "ms_per_sector": round(ms / len(sectors), 3),
                     "kernel": "fused four-step" if plan.info.fft_form == 2 and plan.info.variant == 1
                               else "four-step through HBM"}
        plan.close()
        print(f"[bench] C3 {form}: {out[form]['value']:,.0f} Msamples/s/station, "
              f"{ms:.1f} ms per {len(sectors)}-sector job", file=sys.stderr, flush=True)
    acc.free()
    d1.free()
    d2.free()
    lo, hi = band[1], band[2]
    far = np.abs(np.arange(lo, hi) - line_bin) > 20
    checks = {}
    for form, a in tot.items():
        coh = (a[3, lo:hi] + 1j * a[4, lo:hi]) / np.sqrt(a[1, lo:hi] * a[2, lo:hi])
        checks[form] = {"peak_channel": int(np.argmax(np.abs(coh))) + lo,
                        "line_coherence": round(float(coh[line_bin - lo].real), 5),
                        "continuum_coherence": float(f"{np.mean(coh[far].real):.4e}")}
    res = {"workload": "512 Msps, A 4-bit / B 1-bit, FFT 65536, 16 MHz offset (shift 2048, "
                       "overlap 30720), continuum coherence 1e-3 + 3-channel line of 0.05, "
                       f"{total * n / fs:.0f} s = {total} frames in {len(sectors)} sectors, "
                       "input resident in HBM",
           "forms": out, "line_channel": line_bin, "checks": checks,
           "generation_s": round(gen_s, 2), "reps": reps}
    if "chosen" in tot and "plain" in tot:
        x = [tot[f][3, lo:hi] + 1j * tot[f][4, lo:hi] for f in ("chosen", "plain")]
        rms = np.sqrt(np.mean(np.abs(x[1]) ** 2))
        agree = amp_phase(x[0], x[1])
        res["chosen_over_plain"] = round(out["chosen"]["value"] / out["plain"]["value"], 3)
        res["chosen_vs_plain_xcf"] = {
            "max_abs_over_rms": float(f"{np.max(np.abs(x[0] - x[1])) / rms:.3e}"),
            "amp_max_strong": float(f"{agree['amp_max_strong']:.3e}"),
            "phase_max_strong_rad": float(f"{agree['phase_max_strong_rad']:.3e}")}
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=1_000_000)
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the 64 Msps (C1) compute-only and file-inclusive CPU legs")
    ap.add_argument("--forms", default="plain,chosen")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after timing, one device-to-device copy of this many MiB "
                         "(reference kernel for rocprofv3 FETCH_SIZE/WRITE_SIZE scaling)")
    ap.add_argument("--no-pcie", action="store_true",
                    help="skip the PCIe-inclusive (host-buffer) sector timing")
    ap.add_argument("--serve-jobs", type=int, default=200,
                    help="jobs of the continuous-batching run (0: skip)")
    ap.add_argument("--serve-load", type=float, default=0.9,
                    help="offered load of the serving run (arrival rate x service time)")
    ap.add_argument("--tune", choices=("auto", "default"), default="auto",
                    help="auto: run the plan autotuner on 250,000 frames of the workload "
                         "before timing (untimed); default: the plans' default configuration")
    ap.add_argument("--mfma-variant", type=int, choices=(0, 1, 2), default=None,
                    help="pin the MFMA kernel organization (0 whole-frame waves, "
                         "1 producer/consumer waves) instead of tuning it")
    ap.add_argument("--sizes", action="store_true",
                    help="also time both forms at FFT sizes 256..65536 (adds 'fft_sizes')")
    ap.add_argument("--phased", action="store_true",
                    help="also time the phased-array path at FFT 1024/4096/16384/65536 (adds 'phased')")
    ap.add_argument("--c3", action="store_true",
                    help="also time the 512 Msps 4/1-bit FFT-65536 30-sector precision-stress "
                         "job with its line + continuum signal (adds 'c3')")
    ap.add_argument("--c3-forms", default="plain,chosen",
                    help="forms the --c3 leg runs (a profile may run one)")
    ap.add_argument("--c3-sectors", type=int, default=30,
                    help="first sectors of the 30-sector --c3 job to run (a profile may run a few)")
    ap.add_argument("--c3-reps", type=int, default=2, help="timed --c3 jobs per form")
    ap.add_argument("--batch", action="store_true",
                    help="also time the 48-scan uneven batch captured in one graph (adds 'batch')")
    ap.add_argument("--batch-scale", type=float, default=0.125,
                    help="scan lengths (0.5..20 s) are scaled by this factor")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of "
                         "'bench.py --copy-mib M' into OUT_JSON and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        if args.copy_mib <= 0:
            raise SystemExit("--pmc-summary needs --copy-mib M (the calibration copy size)")
        pmc_summary(*args.pmc_summary, args.copy_mib)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    total_frames = int(math.floor(SECONDS * FS / N))             # 8,000,000
    sectors = F.sector_partition(total_frames, N / FS)           # 8 x 1,000,000
    bpf = N * BITS // 8
    samples = total_frames * N
    table, dcfg, shuffle_b, residual = workload_setup()

    stream = F.Stream()
    raw1 = F.DeviceBuffer(total_frames * bpf)
    raw2 = F.DeviceBuffer(total_frames * bpf)
    F.synth_station_device(raw1.ptr, samples, BITS, None, False, 7, 0, 0.005, residual,
                           stream.ptr)
    F.synth_station_device(raw2.ptr, samples, BITS, shuffle_b, True, 7, 1, 0.005, residual,
                           stream.ptr)
    stream.sync()
    half = N // 2 + 1
    acc = F.DeviceBuffer(len(sectors) * F.N_PRODUCTS * half * 8)
    cfg = F.make_delay_cfg(table, **dcfg)
    print(f"[bench] workload: {len(sectors)} sectors, {total_frames} frames, "
          f"{samples / 1e9:.3f} Gsamples/station, residual delay {residual} samples",
          file=sys.stderr, flush=True)

    results = {}
    for form in [f.strip() for f in args.forms.split(",") if f.strip()]:
        prec = F.PRECISION_CHOSEN if form == "chosen" else F.PRECISION_PLAIN_FP32
        plan = F.Plan(N, bits=(BITS, BITS), levels=[LEVELS, LEVELS],
                      shuffles=(None, shuffle_b), lsb=(False, True), band=(0, 0, N // 2, 0),
                      df_hz=FS / N, products=F.PRODUCT_ACF | F.PRODUCT_XCF, precision=prec)
        tuning = {"variant": plan.info.variant, "grid_divisor": plan.info.grid_divisor,
                  "source": "default"}
        if form == "chosen" and args.mfma_variant is not None:
            F.plan_set_tuning(plan, variant=args.mfma_variant)
            tuning = {"variant": args.mfma_variant, "grid_divisor": 1, "source": "pinned"}
        elif args.tune == "auto":
            rep = F.plan_autotune(plan, raw1.ptr, raw2.ptr, min(250_000, sectors[0]), cfg,
                                  stream.ptr)
            tuning = {"variant": rep["variant"], "grid_divisor": rep["grid_divisor"],
                      "source": f"autotuned ({rep['candidates']} candidates, 250000 frames)",
                      "best_ms": round(rep["best_ms"], 4), "default_ms": round(rep["default_ms"], 4)}
        kernel_key = KERNEL_KEYS[(form, tuning["variant"])]

        def step():
            first = 0
            for si, nf in enumerate(sectors):
                plan.sector_run_device(raw1.ptr + first * bpf, raw2.ptr + first * bpf, first,
                                       nf, cfg, acc.ptr + si * F.N_PRODUCTS * half * 8,
                                       stream=stream.ptr)
                first += nf

        for _ in range(args.warmup):
            step()
        stream.sync()
        plan.set_timing(True)
        e0, e1 = F.Event(), F.Event()
        e0.record(stream)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        e1.record(stream)
        stream.sync()
        wall = time.perf_counter() - t0
        ms_total = e0.elapsed_ms(e1)
        kern_ms, launches = plan.get_timing()
        plan.set_timing(False)
        out = acc.download(np.zeros(len(sectors) * F.N_PRODUCTS * half))
        assert np.all(np.isfinite(out)), "non-finite accumulators"
        ms_step = ms_total / args.steps
        value = samples / (ms_step * 1e-3) / 1e6
        avg_launch_ms = kern_ms / max(launches, 1)
        nf_avg = total_frames / len(sectors)
        flops = nf_avg * 2 * 2.5 * N * math.log2(N)
        bytes_in = nf_avg * 2 * bpf
        results[form] = dict(value=value, ms_per_step=ms_step, wall_ms_per_step=wall * 1e3 / args.steps,
                             avg_launch_ms=avg_launch_ms, launches=launches,
                             tflops=flops / (avg_launch_ms * 1e-3) / 1e12,
                             gbs=bytes_in / (avg_launch_ms * 1e-3) / 1e9,
                             acc11_sum=float(out[half:2 * half].sum()),
                             sector1=out[:F.N_PRODUCTS * half].copy(),
                             tuning=tuning, kernel=kernel_key)
        print(f"[bench] {form}: {value:,.0f} Msamples/s/station, {ms_step:.3f} ms/step, "
              f"fused kernel {avg_launch_ms:.3f} ms/launch ({kernel_key}, {tuning['source']})",
              file=sys.stderr, flush=True)
        plan.close()

    if args.copy_mib > 0:
        nb = args.copy_mib * 2 ** 20
        src, dst = F.DeviceBuffer(nb), F.DeviceBuffer(nb)
        dst.copy_from(src, nb, stream=stream.ptr)
        stream.sync()
        print(f"[bench] calibration copy: {args.copy_mib} MiB device to device", file=sys.stderr,
              flush=True)
        src.free()
        dst.free()

    main_form = "chosen" if "chosen" in results else next(iter(results))
    serving = None
    if args.serve_jobs > 0:
        serving = serve_bench(raw1, raw2, bpf, cfg, shuffle_b, args.serve_jobs, args.serve_load)
        print(f"[bench] serving (host input): {serving['value']:,.0f} Msamples/s/station, job "
              f"latency p50 {serving['p50_latency_ms']:.2f} ms p99 {serving['p99_latency_ms']:.2f} "
              f"ms, up to {serving['max_concurrent_jobs']} jobs in flight; plain one job at a "
              f"time {serving['plain_one_at_a_time']['value']:,.0f}; device-resident "
              f"{serving['device_resident']['value']:,.0f}", file=sys.stderr, flush=True)
    h1 = h2 = None
    if not (args.no_cpu_baseline and args.no_pcie):
        nfc = min(args.cpu_frames, sectors[0])
        h1 = raw1.download(np.zeros(nfc * bpf, dtype=np.uint8))
        h2 = raw2.download(np.zeros(nfc * bpf, dtype=np.uint8))

    pcie = None
    if not args.no_pcie:
        # host buffers in, spectra out: H2D of both stations' bytes + kernels + D2H
        prec = F.PRECISION_CHOSEN if main_form == "chosen" else F.PRECISION_PLAIN_FP32
        plan = F.Plan(N, bits=(BITS, BITS), levels=[LEVELS, LEVELS],
                      shuffles=(None, shuffle_b), lsb=(False, True), band=(0, 0, N // 2, 0),
                      df_hz=FS / N, products=F.PRODUCT_ACF | F.PRODUCT_XCF, precision=prec)
        nfp = len(h1) // bpf
        # (1) the streaming pipeline: both stations' bytes of two sectors in
        # page-locked memory, the 8-sector job submitted back to back (sectors
        # alternate between the two host copies), copies overlapping kernels
        nsec_h = 2
        hb = [[F.HostBuffer(nfp * bpf) for _ in range(2)] for _ in range(nsec_h)]
        for i in range(nsec_h):
            for st, d in enumerate((raw1, raw2)):
                F.check(F.lib().fxc_memcpy_d2h(F.C.c_void_p(hb[i][st].ptr),
                                               F.C.c_void_p(d.ptr + i * nfp * bpf), nfp * bpf,
                                               F.C.c_void_p(0)))
        outp = F.HostBuffer(len(sectors) * F.N_PRODUCTS * half * 8)
        pl = F.Pipeline(plan, nfp, 2)

        def job():
            tk = [pl.submit(hb[i % nsec_h][0].ptr, hb[i % nsec_h][1].ptr, (i % nsec_h) * nfp, nfp,
                            cfg, outp.ptr + i * F.N_PRODUCTS * half * 8)
                  for i in range(len(sectors))]
            pl.wait(tk[-1])

        job()
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            job()
        el = (time.perf_counter() - t0) / reps
        pipe_value = len(sectors) * nfp * N / el / 1e6
        pl.close()
        # (2) one sector through the synchronous host entry from pageable memory
        plan.sector_run(h1, h2, 0, nfp, cfg)
        t0 = time.perf_counter()
        for _ in range(reps):
            plan.sector_run(h1, h2, 0, nfp, cfg)
        el_p = (time.perf_counter() - t0) / reps
        pcie = {"value": round(pipe_value, 1), "unit": "Msamples/s/station",
                "sample": f"{len(sectors)} x {nfp}-frame sectors from page-locked host memory "
                          f"through fxc_pipeline (depth 2, {2 * nfp * bpf / 2 ** 20:.0f} MiB H2D "
                          f"per sector, copies overlapping kernels), mean of {reps} jobs",
                "pageable_sync": {"value": round(nfp * N / el_p / 1e6, 1),
                                  "sample": f"one {nfp}-frame sector from pageable host memory "
                                            f"through fxc_sector_run, mean of {reps}"}}
        print(f"[bench] PCIe-inclusive ({main_form}): {pcie['value']:,.0f} Msamples/s/station "
              f"(pipeline, page-locked); {pcie['pageable_sync']['value']:,.0f} "
              f"(synchronous, pageable)", file=sys.stderr, flush=True)
        for row in hb:
            for b in row:
                b.free()
        outp.free()
        plan.close()

    cpu = None
    parity = None
    if not args.no_cpu_baseline:
        import fxref as R

        nfc = len(h1) // bpf
        threads = max(1, min(os.cpu_count() or 1, 16))
        desc = dict(fft_len=N, bits1=BITS, bits2=BITS, levels1=LEVELS, levels2=LEVELS,
                    shuffle1=np.arange(32, dtype=np.int32), shuffle2=shuffle_b, lsb1=0, lsb2=1,
                    a1s=0, a1e=N // 2, a2s=0, df_hz=FS / N, need_acf=1, need_xcf=1,
                    cpu_threads=max(1, threads - 2))
        rcfg = R.make_delay_cfg(table, **dcfg)
        t0 = time.perf_counter()
        o32 = R.sector_run(desc, h1, h2, 0, nfc, rcfg, mode=R.MODE_F32, threads=threads)
        el = time.perf_counter() - t0
        cpu = {"value": nfc * N / el / 1e6, "unit": "Msamples/s/station", "cores": threads,
               "kind": "port",
               "sample": f"first {nfc} frames (sector 1, {nfc * N / FS:.3f} s) of the same "
                         f"workload, oracle FP32 sector runner, {threads} threads, "
                         f"{el:.1f} s wall, input in RAM (compute only)",
               "host": host_identity(threads)}
        print(f"[bench] cpu baseline: {cpu['value']:.1f} Msamples/s/station "
              f"({threads} threads, {el:.1f} s)", file=sys.stderr, flush=True)
        if nfc == sectors[0]:
            # The same oracle runs check the timed runs' own sector-1 sums:
            # plain against the reference-faithful FP32 mode, chosen against FP64.
            o64 = R.sector_run(desc, h1, h2, 0, nfc, rcfg, mode=R.MODE_F64, threads=threads)
            parity = {}
            for form, ref, tol in (("plain", o32, 1e-5), ("chosen", o64, 1e-4)):
                if form in results:
                    parity[form] = sector_parity(results[form]["sector1"], ref, half, tol)
                    parity[form]["oracle"] = "FP32 (reference-faithful)" if form == "plain" else "FP64"
            parity["sample"] = (f"sector 1 ({nfc} frames) of each form's timed run against the "
                                f"oracle on the same bytes")
            print(f"[bench] parity: {json.dumps(parity)}", file=sys.stderr, flush=True)
        if not args.no_c1:
            cpu["c1"] = c1_cpu_legs(stream, threads)
            print(f"[bench] cpu baseline C1: {json.dumps(cpu['c1'])}", file=sys.stderr, flush=True)

    sizes = size_sweep(stream) if args.sizes else None
    phased = phased_sweep(stream) if args.phased else None
    batch = batch_bench(stream, args.batch_scale) if args.batch else None
    c3 = None
    if args.c3:
        c3 = c3_bench(stream, reps=args.c3_reps, n_sectors=args.c3_sectors,
                      forms=[f.strip() for f in args.c3_forms.split(",") if f.strip()])

    r = results[main_form]
    line = {
        "metric": "Msamples/s/station @2-bit FFT1024",
        "value": round(r["value"], 1),
        "unit": "Msamples/s/station",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated correlated 2-bit Gaussian pair, seeded)",
        "config": {"workload": "two stations, 1024 Msps 2-bit, FFT 1024, 8 s per station "
                               "(8 sectors x 1,000,000 frames), B LSB + seeded shuffle, "
                               "ACF+XCF, input resident in HBM",
                   "form": main_form, "fft_len": N, "frames": total_frames,
                   "sectors": len(sectors)},
        "forms": {k: {"value": round(v["value"], 1), "ms_per_step": round(v["ms_per_step"], 4),
                      "fused_kernel_ms": round(v["avg_launch_ms"], 4),
                      "fft_tflops": round(v["tflops"], 2), "input_gbs": round(v["gbs"], 1),
                      "kernel": v["kernel"], "tuning": v["tuning"]}
                  for k, v in results.items()},
        "roofline": {"bound": "mfma", "achieved": round(r["tflops"], 3),
                     "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(r["tflops"] / FP32_PEAK_TFLOPS, 4),
                     "traffic": load_traffic(r["kernel"]),
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT),
                     "kernel": f"{KERNEL_LABELS[r['kernel']]} ({r['kernel']})",
                     "hbm_frac": round(r["gbs"] / HBM_PEAK_GBS, 4),
                     **issue_fractions(r["kernel"], total_frames / len(sectors), r["avg_launch_ms"])},
        "cpu_baseline": cpu,
        "parity": parity,
        "pcie_inclusive": pcie,
        "serving": serving,
    }
    if sizes is not None:
        line["fft_sizes"] = sizes
    if phased is not None:
        line["phased"] = phased
    if batch is not None:
        line["batch"] = batch
    if c3 is not None:
        line["c3"] = c3
    if "plain" in results and "chosen" in results:
        line["chosen_over_plain"] = round(results["chosen"]["value"] / results["plain"]["value"], 3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
