"""Whole-scan cross spectrum (the reference's --fringe XCF pipeline,
main.rs:1692-1947) through the C ABI: per-chunk XCF sums scaled by 1/N^2
into one FP64 spectrum, then finalize_cross_spectrum (xcf.rs:13-104).

The oracle restates the same accumulation from its own sector runs (XCF
only).  The plain form is held to 1e-5 of the reference-faithful oracle, the
chosen form to 1e-4 of the FP64 oracle.  The host and device entries must
agree bit for bit, and with no delay model the fitted delay recovers the
delay planted in the synthetic signal."""
import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _cases import make_case, rel_err, xcf_err

pytestmark = pytest.mark.gpu


def _oracle_scan(c, chunks, mode):
    n = c["n"]
    ac = np.zeros(n // 2 + 1, dtype=np.complex128)
    d = dict(fft_len=n, bits1=c["bits"][0], bits2=c["bits"][1], levels1=c["levels"][0],
             levels2=c["levels"][1], shuffle1=c["shuffles"][0], shuffle2=c["shuffles"][1],
             lsb1=int(c["lsb"][0]), lsb2=int(c["lsb"][1]), a1s=c["band"][1], a1e=c["band"][2],
             a2s=c["band"][3], out_grid=c["grid"], df_hz=c["fs"] / n, need_acf=0, need_xcf=1,
             cpu_threads=8)
    cfg = R.make_delay_cfg(c["table"], **c["delay"])
    bpf1, bpf2 = n * c["bits"][0] // 8, n * c["bits"][1] // 8
    off = 0
    for k in chunks:
        o = R.sector_run(d, c["raw1"][off * bpf1:(off + k) * bpf1],
                         c["raw2"][off * bpf2:(off + k) * bpf2], off, k, cfg, mode=mode, threads=8)
        ac += o["acc12"] * (1.0 / float(n * n))
        off += k
    return ac


@pytest.mark.parametrize("prec", [F.PRECISION_PLAIN_FP32, F.PRECISION_CHOSEN])
def test_fringe_scan_vs_oracle_and_delay_fit(prec):
    n, fs, delay = 1024, 64e6, 2.4
    chunks = [1700, 2300, 1000]
    c = make_case(n=n, fs=fs, nf=sum(chunks), rho=0.3, delay_samples=delay, seed=77)
    p = F.Plan(n, bits=c["bits"], levels=c["levels"], shuffles=c["shuffles"], lsb=c["lsb"],
               band=c["band"], df_hz=fs / n, products=F.PRODUCT_XCF, precision=prec)
    cfg = F.make_delay_cfg(c["table"], **c["delay"])
    bpf = p.info.bpf1
    ac = np.zeros(p.half, dtype=np.complex128)
    off = 0
    for k in chunks:
        assert F.fringe_accumulate(p, c["raw1"][off * bpf:(off + k) * bpf],
                                   c["raw2"][off * bpf:(off + k) * bpf], off, k, cfg, ac) == 0
        off += k
    # device entry: the same sums added on the device
    d1, d2 = F.DeviceBuffer(c["raw1"].nbytes), F.DeviceBuffer(c["raw2"].nbytes)
    d1.upload(c["raw1"])
    d2.upload(c["raw2"])
    dac = F.DeviceBuffer(p.half * 16)
    F.check(F.lib().fxc_memset_d(F.C.c_void_p(dac.ptr), 0, p.half * 16, F.C.c_void_p(0)))
    s = F.Stream()
    off = 0
    for k in chunks:
        F.fringe_accumulate_device(p, d1.ptr + off * bpf, d2.ptr + off * bpf, off, k, cfg, dac.ptr,
                                   stream=s.ptr)
        off += k
    s.sync()
    dev = dac.download(np.zeros(p.half, dtype=np.complex128))
    assert np.array_equal(dev, ac)

    mode = R.MODE_F32 if prec == F.PRECISION_PLAIN_FP32 else R.MODE_F64
    ref = _oracle_scan(c, chunks, mode)
    tol = 1e-5 if prec == F.PRECISION_PLAIN_FP32 else 1e-4
    assert xcf_err(ac, ref) < tol
    # finalize both: the same delay; and it is the planted one (no delay model)
    _, g = F.finalize_cross_spectrum(ac, n, fs, 6700.0)
    _, o = R.finalize_cross_spectrum(ref, n, fs, 6700.0)
    assert g["delay"] is not None and o["delay"] is not None
    assert g["delay"] == pytest.approx(o["delay"], rel=1e-4)
    assert abs(g["delay"]) == pytest.approx(delay / fs, rel=0.05)
    assert rel_err(g["mag"][1:512], o["mag"][1:512]) < 10 * tol


def test_fringe_needs_xcf_plan():
    lv = [np.array([-3.0, -1.0, 1.0, 3.0])] * 2
    p = F.Plan(1024, levels=lv, df_hz=1.0, products=F.PRODUCT_ACF)
    with pytest.raises(F.FxcError):
        F.fringe_accumulate(p, np.zeros(256, np.uint8), np.zeros(256, np.uint8), 0, 1,
                            F.make_delay_cfg(None, frame_dt=1.6e-5, fs=64e6),
                            np.zeros(p.half, dtype=np.complex128))
