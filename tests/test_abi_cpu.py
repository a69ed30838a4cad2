"""CPU-only checks of the C ABI: the library loads, exports every function
include/fxcorr.h declares, and its host-side helpers (args.rs / main.rs
boundary logic) agree with the oracle.  No GPU compute here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fxcorr as F
import fxref as R
from _synth import DEFAULT_EXT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "fxcorr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fxc_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(F.LIB_PATH)
    names = _declared()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(F.EXPORTED) == names


def test_version_and_strerror():
    assert F.lib().fxc_version().startswith(b"fxcorr")
    assert F.lib().fxc_strerror(F.OK) == b"ok"
    assert F.lib().fxc_strerror(3) == b"HIP runtime error"


def test_parse_shuffle_matches_oracle():
    rng = np.random.default_rng(0)
    for _ in range(20):
        ext = ",".join(str(int(v)) for v in rng.permutation(32))
        assert np.array_equal(F.parse_shuffle(ext), R.parse_shuffle(ext))
    assert np.array_equal(F.parse_shuffle(DEFAULT_EXT), np.arange(32))
    for bad in ["1,2,3", ",".join(["0"] * 32), ""]:
        with pytest.raises(F.FxcError):
            F.parse_shuffle(bad)


def test_parse_levels_and_power():
    assert list(F.parse_levels(2, "-1.5,+0.5,-0.5,+1.5")) == [-1.5, 0.5, -0.5, 1.5]
    assert F.mean_power([-1.5, -0.5, 0.5, 1.5]) == 1.25
    with pytest.raises(F.FxcError):
        F.parse_levels(2, "1,2")
    with pytest.raises(F.FxcError):
        F.mean_power([0.0])


@pytest.mark.parametrize("args", [
    (1024, 64e6, 132.0, 132.0, 32.0, 32.0),
    (65536, 512e6, 228.0, 244.0, 256.0, 256.0),
    (65536, 512e6, 244.0, 228.0, 256.0, 256.0),
    (4096, 256e6, 300.0, 300.5, 128.0, 128.0),
    (1000, 64e6, 100.0, 99.0, 32.0, 32.0),
])
def test_band_alignment_matches_oracle(args):
    assert F.band_alignment(*args) == R.band_alignment(*args)


def test_band_alignment_no_overlap():
    with pytest.raises(F.FxcError) as e:
        F.band_alignment(1024, 64e6, 100.0, 200.0, 32.0, 32.0)
    assert e.value.status == 4


def test_sector_partition():
    # main.rs:2058-2072: round(total seconds) sectors, remainder to the first ones
    assert F.sector_partition(62499, 1024 / 64e6) == [62499]
    c2 = F.sector_partition(7999999, 1024 / 1024e6)
    assert len(c2) == 8 and sum(c2) == 7999999 and max(c2) - min(c2) <= 1
    assert c2[0] == 1000000 and c2[-1] == 999999
    c3 = F.sector_partition(234375, 65536 / 512e6)
    assert len(c3) == 30 and sum(c3) == 234375 and set(c3) == {7812, 7813}
    assert F.sector_partition(0, 1.0) == []
    assert F.sector_partition(3, 10.0) == [1, 1, 1]   # sectors capped at total frames


def test_split_delay_matches_oracle():
    for d, fs in [(12.983, 1.0), (12.345, 1.0), (37.3 / 64e6, 64e6), (-2.5, 1.0),
                  (float("nan"), 1.0), (1e300, 1.0)]:
        assert F.split_delay(d, fs) == R.split_delay(d, fs) or (
            np.isnan(F.split_delay(d, fs)[1]) and np.isnan(R.split_delay(d, fs)[1]))


def test_plan_rejects_bad_descriptions():
    lv = [[-3.0, -1.0, 1.0, 3.0]] * 2
    # these fail validation before any device is touched
    with pytest.raises(F.FxcError) as e:
        F.Plan(1000, levels=lv, df_hz=1.0)
    assert e.value.status == 2
    with pytest.raises(F.FxcError):
        F.Plan(1024, levels=[[1.0, 2.0]] * 2, df_hz=1.0)            # wrong level count
    with pytest.raises(F.FxcError):
        F.Plan(1024, levels=lv, shuffles=(np.zeros(32, np.int32), None), df_hz=1.0)
    with pytest.raises(F.FxcError):
        F.Plan(1024, levels=lv, band=(0, 10, 10, 0), df_hz=1.0)     # empty overlap


def test_no_cpu_fallback_without_gpu():
    # without a GPU the product must fail loudly, never compute on the CPU
    if F.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(F.FxcError) as e:
        F.Plan(1024, levels=[[-3.0, -1.0, 1.0, 3.0]] * 2, df_hz=62500.0)
    assert e.value.status == 3
