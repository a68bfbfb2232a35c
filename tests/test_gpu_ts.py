"""GPU parity of the time-series conditioning (pyl_ts_quantiles, pyl_ts_deredden_median,
pyl_ts_normalise) against the CPU oracle (oracle/ts_oracle.py), and TimeSeries.from_tim
end to end on written .tim and .dat files. Order statistics, block medians and the
dereddened series are bit-exact; the normalisation's FP64 sums run in another order, so
its location and scale agree within 1e-12 (relative) and the series within 2 ulps."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import ts_oracle as to
from pyloki_amd import _lib
from pyloki_amd.device import ptr, stream_ptr
from pyloki_amd.io import sigproc
from pyloki_amd.io.timeseries import TimeSeries, deredden_median, normalise

pytestmark = pytest.mark.gpu


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _quantiles(x, seg_len, qs, center=None):
    xd = torch.from_numpy(x).cuda()
    nseg = -(-len(x) // seg_len)
    out = torch.empty((nseg, len(qs)), dtype=torch.float32, device="cuda")
    q = (ctypes.c_double * len(qs))(*qs)
    cd = torch.from_numpy(np.asarray(center, dtype=np.float32)).cuda() if center is not None \
        else None
    _lib.check(_lib.load().pyl_ts_quantiles(ptr(xd), len(x), seg_len, q, len(qs), ptr(cd),
                                            ptr(out), stream_ptr()))
    return out.cpu().numpy()


@pytest.mark.parametrize("n,seg_len,qs", [
    (100_003, 1000, [0.5]), (4096, 4096, [0.25, 0.75]), (77, 10, [0.0, 1.0]),
    (2**20, 2**20, [0.5]), (70_001, 1, [0.5])])
def test_quantiles_bit_exact(gpu_device, n, seg_len, qs):
    _case(f"quantiles n={n} seg_len={seg_len} q={qs}")
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n).astype(np.float32)
    x[::7] = np.round(x[::7])  # ties, signed zeros
    x[5] = -0.0
    got = _quantiles(x, seg_len, qs)
    want = to.segment_quantiles(x, seg_len, qs)
    assert np.array_equal(got, want)


def test_quantiles_of_absolute_deviations(gpu_device):
    _case("quantiles of |x - c| per segment")
    rng = np.random.default_rng(4)
    x = rng.standard_normal(30_000).astype(np.float32)
    c = rng.standard_normal(3).astype(np.float32)
    assert np.array_equal(_quantiles(x, 10_000, [0.5], c),
                          to.segment_quantiles(x, 10_000, [0.5], c))


@pytest.mark.parametrize("n,w", [(2**20, 15625), (100_000, 1000), (999, 1000), (5000, 1)])
def test_deredden_bit_exact(gpu_device, n, w):
    _case(f"deredden n={n} w={w}")
    rng = np.random.default_rng(w)
    t = np.arange(n)
    x = (rng.standard_normal(n) + 3 * np.sin(t / 20000.0) + 1e-4 * t).astype(np.float32)
    out, med = deredden_median(torch.from_numpy(x).cuda(), w)
    want, wmed = to.deredden_median(x, w)
    assert np.array_equal(med.cpu().numpy(), wmed)
    assert np.array_equal(out.cpu().numpy(), want)


@pytest.mark.parametrize("method", ["std", "iqr", "mad"])
def test_normalise(gpu_device, method):
    _case(f"normalise {method}")
    rng = np.random.default_rng(8)
    x = (rng.standard_normal(2**20) * 3 + 5).astype(np.float32)
    out, loc, scale = normalise(torch.from_numpy(x).cuda(), method)
    want, wloc, wscale = to.normalise(x, method)
    assert abs(loc - wloc) <= 1e-12 * abs(wloc) and abs(scale / wscale - 1) <= 1e-12
    got = out.cpu().numpy()
    assert np.max(np.abs(got - want) / np.spacing(np.abs(want) + 1e-30)) <= 2
    with pytest.raises(ValueError, match="zero scale"):
        normalise(torch.zeros(100, device="cuda"), method)


@pytest.mark.parametrize("tim_type", ["tim", "dat"])
def test_from_tim_end_to_end(gpu_device, tmp_path, tim_type):
    """A red-noise series with an offset written to disk, loaded by from_tim: the result
    is the oracle's dereddened and normalised series; its IQR scale is ~1."""
    _case(f"from_tim {tim_type}")
    n, tsamp = 2**20, 6.4e-5
    rng = np.random.default_rng(11)
    t = np.arange(n) * tsamp
    x = (100 + 5 * np.sin(2 * np.pi * t / 30.0) + rng.standard_normal(n)).astype(np.float32)
    path = tmp_path / f"s.{tim_type}"
    if tim_type == "tim":
        sigproc.write_sigproc_tim(path, x, tsamp)
    else:
        sigproc.write_presto_dat(path, x, tsamp)
    ts = TimeSeries.from_tim(str(path), tim_type=tim_type, filter_window=1.0)
    red, _ = to.deredden_median(x, int(1.0 / tsamp))
    want, _, _ = to.normalise(red, "iqr")
    assert ts.dt == tsamp and np.all(ts.ts_v == 1)
    assert np.max(np.abs(ts.ts_e - want) / np.spacing(np.abs(want) + 1e-30)) <= 2
    q25, q75 = np.percentile(ts.ts_e, [25, 75])
    assert abs((q75 - q25) / to.IQR_NORMAL - 1) < 1e-3
