"""Pin the CPU oracle of the time-series conditioning (oracle/ts_oracle.py) and the file
readers of pyloki_amd.io.sigproc (no GPU)."""

from __future__ import annotations

import struct

import numpy as np
import pytest

from oracle import ts_oracle as to
from pyloki_amd.io import sigproc
from pyloki_amd.io.timeseries import TimeSeries


def test_quantiles_match_numpy():
    rng = np.random.default_rng(1)
    for m in (1, 2, 3, 10, 101, 1000):
        x = rng.standard_normal(m).astype(np.float32)
        for q in (0.0, 0.25, 0.5, 0.75, 1.0):
            got = float(to.quantile(x, q))
            want = float(np.percentile(x.astype(np.float64), 100 * q))
            assert abs(got - want) <= 2 * np.spacing(np.float32(abs(want)) + np.float32(1e-30))
        assert abs(float(to.quantile(x, 0.5)) - float(np.median(x))) <= 2 * np.spacing(
            np.float32(abs(np.median(x))))


def test_quantiles_on_integers_are_exact_order_statistics():
    rng = np.random.default_rng(2)
    x = rng.integers(-1000, 1000, 999).astype(np.float32)  # ties included
    s = np.sort(x)
    assert to.quantile(x, 0.5) == s[499]
    assert to.quantile(x, 0.25) == s[int(0.25 * 998)] + (s[250] - s[249]) * (0.25 * 998 - 249)


def test_ramp_baseline_is_the_ramp():
    """The median of a block of a linear ramp is the ramp at the block centre, so the
    interpolated baseline equals the ramp between the first and the last centre."""
    w, n = 100, 1050
    ramp = (0.5 + 0.01 * np.arange(n)).astype(np.float64)
    x = ramp.astype(np.float32)
    out, med = to.deredden_median(x, w)
    assert len(med) == 11
    inner = slice(50, 1000)
    assert np.max(np.abs(out[inner])) < 1e-5
    # a constant offset is removed everywhere
    out2, _ = to.deredden_median(np.full(n, 7.25, dtype=np.float32), w)
    assert np.all(out2 == 0.0)


def test_normalise_closed_forms():
    x = np.arange(1, 10001, dtype=np.float32)  # 1..10000
    out, loc, scale = to.normalise(x, "std")
    assert loc == 5000.5 and abs(scale - np.sqrt((10000**2 - 1) / 12)) < 1e-9
    _, _, s_iqr = to.normalise(x, "iqr")
    assert abs(s_iqr - (7500.25 - 2500.75) / to.IQR_NORMAL) < 1e-9
    _, _, s_mad = to.normalise(x, "mad")
    assert abs(s_mad - 2500.0 * to.MAD_NORMAL) < 1e-9  # |x - 5000.5| has median 2500
    assert abs(float(np.mean(out.astype(np.float64)))) < 1e-6


def test_sigproc_header_parse_from_spec_bytes(tmp_path):
    """A header assembled byte by byte from the SIGPROC layout (int32 length + keyword,
    then the typed value) parses to its values and its samples follow."""
    def s(x):
        return struct.pack("<i", len(x)) + x.encode()

    data = np.arange(8, dtype="<f4")
    raw = (s("HEADER_START") + s("source_name") + s("J0000+00") + s("tsamp")
           + struct.pack("<d", 6.4e-5) + s("nbits") + struct.pack("<i", 32) + s("nchans")
           + struct.pack("<i", 1) + s("tstart") + struct.pack("<d", 60000.5) + s("signed")
           + bytes([0]) + s("HEADER_END") + data.tobytes())
    path = tmp_path / "a.tim"
    path.write_bytes(raw)
    got, hdr = sigproc.read_sigproc_tim(path)
    assert hdr["tsamp"] == 6.4e-5 and hdr["source_name"] == "J0000+00"
    assert hdr["tstart"] == 60000.5 and hdr["nbits"] == 32
    assert np.array_equal(got, data)


def test_readers_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(1000).astype(np.float32)
    sigproc.write_sigproc_tim(tmp_path / "b.tim", x, 6.4e-5, source_name="fake")
    got, hdr = sigproc.read_sigproc_tim(tmp_path / "b.tim")
    assert np.array_equal(got, x) and hdr["tsamp"] == 6.4e-5
    sigproc.write_presto_dat(tmp_path / "c.dat", x, 2.56e-4)
    got, inf = sigproc.read_presto_dat(tmp_path / "c.dat")
    assert np.array_equal(got, x) and inf["tsamp"] == 2.56e-4
    with pytest.raises(ValueError, match="HEADER_START"):
        (tmp_path / "bad.tim").write_bytes(struct.pack("<i", 3) + b"abc")
        sigproc.read_sigproc_tim(tmp_path / "bad.tim")
    with pytest.raises(ValueError, match="Invalid tim type"):
        TimeSeries.from_tim(str(tmp_path / "c.dat"), tim_type="fits")


def test_downsample_is_the_mean():
    ts = TimeSeries(np.arange(12, dtype=np.float32), np.ones(12, dtype=np.float32), 1e-3)
    d = ts.downsample(4)
    assert np.array_equal(d.ts_e, [1.5, 5.5, 9.5]) and np.array_equal(d.ts_v, [1, 1, 1])
    with pytest.raises(ValueError):
        ts.downsample(5)
