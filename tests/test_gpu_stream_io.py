"""Streaming host I/O end to end, through the C ABI: the reader thread and
per-sector loop of the reference (main.rs:458-516, 2098-2213) replaced by
fxc_reader_* filling page-locked buffers (fxc_host_alloc) that
fxc_pipeline_submit copies to the device while earlier sectors run, with
the buffers recycled after fxc_pipeline_wait.

Two 2-bit .raw files (16 Msps, 3 s, station B's stream 37 samples late);
B is read from its word-aligned start (37 -> 32 samples, main.rs:1585-1616)
and the scan is cut into three sectors (main.rs:2058-2072).  Every
sector's spectra from the pipeline are held to the oracle run on the same
bytes, read here with numpy independently of the library's reader: plain
against the reference-faithful FP32 mode (1e-5), chosen against FP64 (1e-4)."""
import ctypes as C

import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import fmt_amp_phase, spectra_amp_phase, spectra_err, zero_pattern_equal
from _synth import LEVELS_2BIT, pack_codes, quantize_2bit

pytestmark = pytest.mark.gpu

FS, N, BITS, SECONDS = 16e6, 1024, 2, 3.0


def _files(tmp_path, rho=0.05, shift=37, seed=9):
    n = int(FS * SECONDS)
    rng = np.random.default_rng(seed)
    c = rng.standard_normal(n + shift, dtype=np.float32)
    a = np.sqrt(rho) * c[shift:] + np.sqrt(1 - rho) * rng.standard_normal(n, dtype=np.float32)
    b = np.sqrt(rho) * c[:n] + np.sqrt(1 - rho) * rng.standard_normal(n, dtype=np.float32)
    paths = [str(tmp_path / "A.raw"), str(tmp_path / "B.raw")]
    for x, p in zip((a, b), paths):
        pack_codes(quantize_2bit(x), BITS, N).tofile(p)
    return paths, n


@pytest.mark.parametrize("prec", [F.PRECISION_PLAIN_FP32, F.PRECISION_CHOSEN])
def test_reader_pipeline_oracle(tmp_path, prec):
    paths, nsamp = _files(tmp_path)
    bpf = N * BITS // 8
    actual, start_byte, _ = F.word_aligned_start(37, BITS)
    nf_total = (nsamp - actual) // N
    sectors = [int(x) for x in F.sector_partition(nf_total, N / FS)]
    assert len(sectors) == 3 and sum(sectors) == nf_total
    dcfg = dict(frame_dt=N / FS, fs=FS, delay_user_samples=37.0, d_seek=actual / FS,
                lo1_hz=6.7e9, lo2_hz=6.7e9)
    table = np.zeros(5)
    plan = F.Plan(N, bits=(BITS, BITS), levels=[LEVELS_2BIT, LEVELS_2BIT], df_hz=FS / N,
                  precision=prec)
    half = plan.half
    words = F.N_PRODUCTS * half
    max_nf = max(sectors)
    depth = 2
    pl = F.Pipeline(plan, max_nf, depth)
    bufs = [(F.HostBuffer(max_nf * bpf), F.HostBuffer(max_nf * bpf)) for _ in range(depth)]
    out = F.HostBuffer(len(sectors) * words * 8)
    readers = [F.Reader(paths[0], 0, 0), F.Reader(paths[1], start_byte, 0)]
    read = F.lib().fxc_reader_read
    tickets, first = [None] * depth, 0
    for i, nf in enumerate(sectors):
        slot = i % depth
        if tickets[slot] is not None:
            pl.wait(tickets[slot])                # the slot's last sector is done: recycle it
        for st in range(2):
            got = C.c_size_t()
            F.check(read(readers[st].h, C.c_void_p(bufs[slot][st].ptr), nf * bpf, C.byref(got)))
            assert got.value == nf * bpf
        tickets[slot] = pl.submit(bufs[slot][0].ptr, bufs[slot][1].ptr, first, nf,
                                  F.make_delay_cfg(table, **dcfg), out.ptr + i * words * 8)
        first += nf
    pl.wait(max(t for t in tickets if t is not None))
    acc = out.view(np.float64).reshape(len(sectors), F.N_PRODUCTS, half).copy()
    for r in readers:
        r.close()
    pl.close()
    plan.close()
    # the oracle reads the files itself
    mode, tol = (R.MODE_F32, 1e-5) if prec == F.PRECISION_PLAIN_FP32 else (R.MODE_F64, 1e-4)
    desc = dict(fft_len=N, bits1=BITS, bits2=BITS, levels1=LEVELS_2BIT, levels2=LEVELS_2BIT,
                a1s=0, a1e=N // 2, a2s=0, df_hz=FS / N, need_acf=1, need_xcf=1, cpu_threads=8)
    first = 0
    for i, nf in enumerate(sectors):
        r1 = np.fromfile(paths[0], dtype=np.uint8, count=nf * bpf, offset=first * bpf)
        r2 = np.fromfile(paths[1], dtype=np.uint8, count=nf * bpf, offset=start_byte + first * bpf)
        o = R.sector_run(desc, r1, r2, first, nf, R.make_delay_cfg(table, **dcfg), mode=mode,
                         threads=8)
        got = dict(acc11=acc[i, 1], acc22=acc[i, 2], acc12=acc[i, 3] + 1j * acc[i, 4])
        e = spectra_err(got, o)
        print(f"[stream-io] sector {i} ({nf} frames): {e}; "
              f"{fmt_amp_phase(spectra_amp_phase(got, o))}", flush=True)
        for k in ("acc11", "acc22", "acc12"):
            assert zero_pattern_equal(got[k], o[k]), (i, k)
            assert e[k] < tol, (i, k, e[k])
        first += nf
    # the 37-sample delay is removed: the cross spectrum adds coherently
    xc = acc[:, 3, 1:-1].sum(axis=0) + 1j * acc[:, 4, 1:-1].sum(axis=0)
    assert np.abs(xc.sum()) / np.abs(xc).sum() > 0.8
