"""Batched uneven scans replayed as one graph.

48 seeded scans, each with its own delay model and clock terms; FFT sizes
from {256, 1024, 4096, 16384}, per-station bits from {1, 2, 4, 8}, sampling
rates from {128, 256, 512, 1024} Msps, lengths uniform in 0.5..20 s (kept in
proportion, scaled to at most 2^20 samples per station so the oracle checks
each scan in seconds).  Each scan produces ACF + XCF in the chosen form; 12
of them also run the phased-array branch: the delay-aligned weighted sum
into an 8-bit series plus its power spectrum (main.rs:2264-2302; the output
width is max(bits), main.rs:1067, so these scans carry one 8-bit station),
alternately in the chosen form (FP64 rotation phases) and the plain,
reference-faithful one.

Every ACF/XCF plan is tuned by fxc_plan_autotune on its own scan, all the
sector calls are captured into one graph (forked onto three streams, so the
scans are parallel branches), and every replay must equal the eager
results bit for bit.  Each eager result is held to the
oracle: ACF/XCF to the FP64 oracle within 1e-4, the phased power to the
FP64 oracle within 1e-4 (chosen) or the reference-faithful FP32 oracle
within 1e-5 (plain), and the 8-bit series to the same oracle's levels
except at quantiser decision boundaries (at most 1e-4 of the samples, each
to a neighbouring level)."""
import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import (gpu_plan, make_case, oracle_run, out_decode_plan, spectra_err,
                    zero_pattern_equal)

pytestmark = pytest.mark.gpu

N_SCANS, N_PHASED = 48, 12
MAX_SAMPLES = 1 << 20
CHOSEN_TOL = 1e-4
PLAIN_TOL = 1e-5


def _scans():
    rng = np.random.default_rng(48)
    scans = []
    for i in range(N_SCANS):
        n = int(rng.choice([256, 1024, 4096, 16384]))
        bits = [int(rng.choice([1, 2, 4, 8])), int(rng.choice([1, 2, 4, 8]))]
        phased = i < N_PHASED
        if phased and max(bits) < 8:
            bits[int(rng.integers(2))] = 8
        fs = float(rng.choice([128e6, 256e6, 512e6, 1024e6]))
        length_s = float(rng.uniform(0.5, 20.0))
        nf = max(16, int(length_s / 20.0 * MAX_SAMPLES / n))
        delay = float(rng.uniform(-12.0, 12.0))
        lo = float(rng.uniform(2e9, 9e9))
        sh1 = rng.permutation(32).astype(np.int32) if rng.integers(2) else None
        lsb = (False, bool(rng.integers(2)))
        c = make_case(n=n, fs=fs, nf=nf, bits=tuple(bits), shuffles=(sh1, None), lsb=lsb,
                      rho=0.4, delay_samples=delay, seed=1000 + i,
                      delay_user_samples=delay, extra_delay_rate_sps=float(rng.uniform(-5e-9, 5e-9)),
                      clock1_delay_s=float(rng.uniform(-2, 2)) / fs,
                      clock1_rate_sps=float(rng.uniform(-1e-12, 1e-12)),
                      clock2_delay_s=float(rng.uniform(-2, 2)) / fs,
                      clock2_rate_sps=float(rng.uniform(-1e-12, 1e-12)),
                      lo1_hz=lo, lo2_hz=lo)
        w = (float(rng.uniform(0.5, 1.5)), float(rng.uniform(0.5, 1.5)))
        scans.append(dict(case=c, phased=phased, w=w, output_lsb=bool(rng.integers(2)),
                          phased_prec=F.PRECISION_CHOSEN if i % 2 == 0 else F.PRECISION_PLAIN_FP32))
    return scans


def _phased_plan(c, w, output_lsb, precision):
    n = c["n"]
    return F.Plan(n, bits=c["bits"], levels=c["levels"], shuffles=c["shuffles"], lsb=c["lsb"],
                  band=c["band"], df_hz=c["fs"] / n, products=F.PRODUCT_PHASED | F.PRODUCT_RAW_OUT,
                  out_grid=c["grid"], w=w, bit_out=max(c["bits"]), output_lsb=output_lsb,
                  precision=precision)


def _phased_oracle(c, w, output_lsb, mode):
    n = c["n"]
    d = dict(fft_len=n, bits1=c["bits"][0], bits2=c["bits"][1], levels1=c["levels"][0],
             levels2=c["levels"][1], shuffle1=c["shuffles"][0], shuffle2=c["shuffles"][1],
             lsb1=int(c["lsb"][0]), lsb2=int(c["lsb"][1]), a1s=c["band"][1], a1e=c["band"][2],
             a2s=c["band"][3], out_grid=c["grid"], df_hz=c["fs"] / n, need_acf=0, need_xcf=0,
             need_phased=1, write_raw=1, w1=w[0], w2=w[1], bit_out=max(c["bits"]),
             output_lsb=int(output_lsb), cpu_threads=8)
    return R.sector_run(d, c["raw1"], c["raw2"], 0, c["nf"],
                        R.make_delay_cfg(c["table"], **c["delay"]), mode=mode, threads=8)


def test_48_uneven_scans_one_graph():
    scans = _scans()
    s = F.Stream()
    jobs = []
    for sc in scans:
        c = sc["case"]
        h = c["n"] // 2 + 1
        d1, d2 = F.DeviceBuffer(c["raw1"].nbytes), F.DeviceBuffer(c["raw2"].nbytes)
        d1.upload(c["raw1"])
        d2.upload(c["raw2"])
        cfg = F.make_delay_cfg(c["table"], **c["delay"])
        p = gpu_plan(c, F.PRECISION_CHOSEN)
        rep = F.plan_autotune(p, d1.ptr, d2.ptr, c["nf"], cfg, s.ptr)
        assert rep["candidates"] >= 1
        job = dict(sc=sc, d1=d1, d2=d2, cfg=cfg, plan=p, h=h,
                   acc=F.DeviceBuffer(F.N_PRODUCTS * h * 8))
        if sc["phased"]:
            pp = _phased_plan(c, sc["w"], sc["output_lsb"], sc["phased_prec"])
            job.update(pplan=pp, pacc=F.DeviceBuffer(F.N_PRODUCTS * h * 8),
                       praw=F.DeviceBuffer(c["nf"] * pp.info.bpf_out))
        jobs.append(job)

    def enqueue_all(streams=(s,)):
        for k, j in enumerate(jobs):
            c = j["sc"]["case"]
            st = streams[k % len(streams)]
            j["plan"].sector_run_device(j["d1"].ptr, j["d2"].ptr, 0, c["nf"], j["cfg"],
                                        j["acc"].ptr, stream=st.ptr)
            if "pplan" in j:
                j["pplan"].sector_run_device(j["d1"].ptr, j["d2"].ptr, 0, c["nf"], j["cfg"],
                                             j["pacc"].ptr, j["praw"].ptr, stream=st.ptr)

    def collect():
        out = []
        for j in jobs:
            r = [j["acc"].download(np.zeros(F.N_PRODUCTS * j["h"]))]
            if "pplan" in j:
                r.append(j["pacc"].download(np.zeros(F.N_PRODUCTS * j["h"])))
                r.append(j["praw"].download(np.zeros(j["praw"].nbytes, dtype=np.uint8)))
            out.append(r)
        return out

    enqueue_all()                                   # eager (also uploads the delay tables)
    s.sync()
    eager = collect()
    # the capture forks onto three streams (the scans as parallel branches)
    # and joins back before it ends
    side = [F.Stream() for _ in range(3)]
    g = F.Graph(s)
    fork = F.Event()
    fork.record(s)
    for st in side:
        st.wait(fork)
    enqueue_all(side)
    joins = [F.Event() for _ in side]
    for ev, st in zip(joins, side):
        ev.record(st)
        s.wait(ev)
    g.end()
    for j in jobs:                                  # every plan is held, whichever branch ran it
        for key in ("plan", "pplan"):
            if key in j:
                assert F.plan_live_graphs(j[key]) == 1, key
    for rep in range(2):
        for j in jobs:
            for key in ("acc", "pacc", "praw"):
                if key in j:
                    F.check(F.lib().fxc_memset_d(F.C.c_void_p(j[key].ptr), 0, j[key].nbytes,
                                                 F.C.c_void_p(s.ptr)))
        s.sync()
        g.launch()
        s.sync()
        got = collect()
        for i, (a, b) in enumerate(zip(eager, got)):
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (rep, i)
    g.close()

    worst = {"chosen": 0.0, "phased chosen": 0.0, "phased plain": 0.0}
    for j, res in zip(jobs, eager):
        sc, c, h = j["sc"], j["sc"]["case"], j["h"]
        acc = res[0].reshape(F.N_PRODUCTS, h)
        got = dict(acc11=acc[1], acc22=acc[2], acc12=acc[3] + 1j * acc[4])
        o64 = oracle_run(c, R.MODE_F64)
        e = spectra_err(got, o64)
        for k in ("acc11", "acc22", "acc12"):
            assert zero_pattern_equal(got[k], o64[k]), (k, c["n"], c["bits"])
            assert e[k] < CHOSEN_TOL, (k, e[k], c["n"], c["bits"])
        worst["chosen"] = max(worst["chosen"], *e.values())
        if not sc["phased"]:
            continue
        ph = res[1].reshape(F.N_PRODUCTS, h)[0]
        chosen = sc["phased_prec"] == F.PRECISION_CHOSEN
        ref = _phased_oracle(c, sc["w"], sc["output_lsb"], R.MODE_F64 if chosen else R.MODE_F32)
        assert ref["n_failed"] == 0
        m = ref["acc_ph"] != 0
        eph = float(np.max(np.abs(ph[m] - ref["acc_ph"][m]) / np.abs(ref["acc_ph"][m])))
        assert eph < (CHOSEN_TOL if chosen else PLAIN_TOL), (eph, c["n"], c["bits"])
        key = "phased chosen" if chosen else "phased plain"
        worst[key] = max(worst[key], eph)
        bo = max(c["bits"])
        bpf_o = c["n"] * bo // 8
        dp = out_decode_plan(bo, c["shuffles"][0], c["levels"][0])
        lv = np.sort(np.asarray(c["levels"][0]))
        step = float(np.min(np.diff(lv)))
        raw = res[2]
        diff, total = 0, 0
        for f in range(c["nf"]):
            seg = slice(f * bpf_o, (f + 1) * bpf_o)
            _, vg = dp.decode(raw[seg], c["n"], False)
            _, vr = dp.decode(ref["raw_out"][seg], c["n"], False)
            dd = np.abs(vg - vr)
            assert np.all((dd == 0) | (dd <= step * 1.0001)), (f, c["n"])   # neighbours only
            diff += int(np.count_nonzero(dd))
            total += c["n"]
        assert diff <= 1e-4 * total + 2, (diff, total, c["n"])
    print(f"[batch] worst rel err: " + ", ".join(f"{k} {v:.3e}" for k, v in worst.items()))
