"""The 64 Msps 2-bit FFT-1024 one-second case end to end, .raw files to .cor
files, through the C ABI (main.rs:2082-2096, 2645-2683; cor.rs:205-297):

two .raw files on disk (station B's stream delayed 37.3 samples) ->
word-aligned seek of B (37 -> 32, residual 5; main.rs:1585-1616) ->
fxc_reader_* (main.rs:458-516) -> fxc_sector_run (plain FP32, the whole
62,499-frame sector) -> fxc_normalize_sector -> fxc_cor_create /
write_sector / finalize for ACF A, ACF B and the cross spectrum.

Checked: three 4,480-byte files (256-byte header, one 128-byte sector
header, 512 complex f32 channels); header fields at the offsets cor.rs
fixes; and every channel of each file within 1e-5 of the oracle's sums for
the same bytes (read here with numpy, independently of the library's
reader) normalised by 1/(0.5 P nf N^2)."""
import struct

import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import amp_phase_err, rel_err, xcf_err
from _synth import LEVELS_2BIT, correlated_pair, pack_codes, quantize_2bit

pytestmark = pytest.mark.gpu

FS, N, BITS = 64e6, 1024, 2
NSAMP = 64_000_000
UNIX0 = 1_700_000_000


def test_64msps_raw_to_cor_end_to_end(tmp_path):
    print("[cor-e2e] generating 2 x 64 Msamples", flush=True)
    a, b = correlated_pair(NSAMP, 0.02, 37.3, seed=0)
    paths = [str(tmp_path / "A.raw"), str(tmp_path / "B.raw")]
    for x, p in zip((a, b), paths):
        pack_codes(quantize_2bit(x), BITS, N).tofile(p)
    del a, b
    bpf = N * BITS // 8
    # seek: B starts 37 samples late -> word-aligned 32 samples (byte 8), residual 5
    actual, start_byte, residual = F.word_aligned_start(37, BITS)
    assert (actual, start_byte, residual) == (32, 8, -5)
    nf = int(min(NSAMP, NSAMP - actual) // N)                     # main.rs:1277-1304
    sectors = F.sector_partition(nf, N / FS)
    assert nf == 62499 and list(sectors) == [62499]
    ra, rb = F.Reader(paths[0], 0, 0), F.Reader(paths[1], start_byte, 0)
    h1, g1 = ra.read(nf * bpf)
    h2, g2 = rb.read(nf * bpf)
    ra.close()
    rb.close()
    assert g1 == nf * bpf and g2 == nf * bpf
    dcfg = dict(frame_dt=N / FS, fs=FS, delay_user_samples=37.3, d_seek=actual / FS,
                extra_delay_rate_sps=0.5 / 6.7e9, lo1_hz=6.7e9, lo2_hz=6.7e9)
    table = np.zeros(3)
    plan = F.Plan(N, bits=(BITS, BITS), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=FS / N,
                  precision=F.PRECISION_PLAIN_FP32)
    res = plan.sector_run(h1, h2, 0, nf, F.make_delay_cfg(table, **dcfg))
    plan.close()
    assert res["n_failed"] == 0
    acc = np.zeros((F.N_PRODUCTS, N // 2 + 1))
    acc[0], acc[1], acc[2] = res["acc_ph"], res["acc11"], res["acc22"]
    acc[3], acc[4] = res["acc12"].real, res["acc12"].imag
    p = F.mean_power(LEVELS_2BIT)
    assert p == 5.0
    spec = F.normalize_sector(acc, N, nf, p, p)
    hdr = F.cor_header(FS, 6.7e9, N, 1, UNIX0, "SYNTH", 1.0, 0.5)
    st = (F.cor_station("YAMAGU32", ord("G"), (-3502544.587, 3950966.235, 3566381.192)),
          F.cor_station("YAMAGU34", ord("H"), (-3502567.576, 3950885.734, 3566449.115)))
    outs = {"A_A.cor": spec["11"], "B_B.cor": spec["22"], "A_B.cor": spec["12"]}
    for name, s in outs.items():
        w = F.CorWriter(str(tmp_path / name), hdr, *st)
        w.write_sector(UNIX0, 1.0, s)
        w.finalize()
    print("[cor-e2e] .cor files written; oracle on the same bytes", flush=True)
    # the oracle reads the files itself (numpy, at the same offsets)
    r1 = np.fromfile(paths[0], dtype=np.uint8, count=nf * bpf, offset=0)
    r2 = np.fromfile(paths[1], dtype=np.uint8, count=nf * bpf, offset=start_byte)
    assert np.array_equal(r1, h1) and np.array_equal(r2, h2)
    desc = dict(fft_len=N, bits1=BITS, bits2=BITS, levels1=LEVELS_2BIT, levels2=LEVELS_2BIT,
                a1s=0, a1e=N // 2, a2s=0, df_hz=FS / N, need_acf=1, need_xcf=1, cpu_threads=8)
    o = R.sector_run(desc, r1, r2, 0, nf, R.make_delay_cfg(table, **dcfg), mode=R.MODE_F32,
                     threads=8)
    base = 0.5 * nf * float(N) * float(N)
    want = {"A_A.cor": (o["acc11"][:N // 2] / (p * base)).astype(np.float32),
            "B_B.cor": (o["acc22"][:N // 2] / (p * base)).astype(np.float32),
            "A_B.cor": (o["acc12"][:N // 2] / (np.sqrt(p * p) * base)).astype(np.complex64)}
    for name in outs:
        data = open(str(tmp_path / name), "rb").read()
        assert len(data) == 4480, (name, len(data))
        assert data[0:4] == bytes([0x83, 0xF9, 0xA2, 0x3E])                   # cor.rs:166
        assert struct.unpack_from("<iiid", data, 4) == (0x01030000, 1, 64000000, 6.7e9)
        assert struct.unpack_from("<ii", data, 24) == (N, 1)                  # fft, sectors
        assert data[32:40] == b"YAMAGU32" and data[80:88] == b"YAMAGU34"
        assert data[128:133] == b"SYNTH"
        assert struct.unpack_from("<iIII", data, 256) == (UNIX0, 0, UNIX0 + 1, 0)
        assert struct.unpack_from("<f", data, 256 + 112)[0] == 1.0
        got = np.frombuffer(data[384:], dtype="<f4").view(np.complex64)
        assert got.size == N // 2
        w = want[name]
        if name == "A_B.cor":
            e = xcf_err(got, w)
        else:
            assert np.all(got.imag == 0)
            e = rel_err(got.real, w.real if np.iscomplexobj(w) else w)
        ap = amp_phase_err(got, w)
        print(f"[cor-e2e] {name}: {e:.3e}; amp {ap['amp_max']:.2e} phase {ap['phase_max']:.2e} rad",
              flush=True)
        assert e < 1e-5, (name, e)
    # the 37.3-sample delay was removed: the cross spectrum's channels add
    # coherently (an uncorrected 5.3-sample residual would wind the phase
    # 2.65 turns across the band and cancel the sum)
    xc = np.frombuffer(open(str(tmp_path / "A_B.cor"), "rb").read()[384:],
                       dtype="<f4").view(np.complex64)[1:]
    coherent = np.abs(np.sum(xc)) / np.sum(np.abs(xc))
    print(f"[cor-e2e] cross-spectrum coherence across channels {coherent:.3f}", flush=True)
    assert coherent > 0.8
