"""GPU parity of the complex (use_fourier=True) FFA against the oracle, within FP32
tolerance: the leaf fold (a GEMM whose summation order is the library's), every level
(complex64 phasor recurrences), the irfft scorer, and the whole search's S/N grid
(<= 1e-5 of the peak S/N)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import complex_oracle as co
from oracle import oracle as orc
from pyloki_amd import kernels
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming, FFAPlan
from pyloki_amd.search import ffa_search
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _cfg(P: int, nsamps: int, nbins: int) -> PulsarSearchConfig:
    if P == 1:
        lim = np.array([[1.8, 2.2]])
        tsamp = 256e-6
    else:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
        tsamp = 64e-6
    return PulsarSearchConfig(nsamps=nsamps, tsamp=tsamp, nbins=nbins, eta=1.0,
                              param_limits=lim, ducy_max=0.3, wtsp=1.5, use_fourier=True)


def _series(cfg: PulsarSearchConfig, seed: int = 2):
    period = 0.5 if cfg.nparams == 1 else 1 / 200.37
    mod = {} if cfg.nparams == 1 else {"acc": 12.0}
    return PulseSignalConfig(period=period, dt=cfg.tsamp, nsamps=cfg.nsamps, snr=25.0,
                             ducy=0.08, mod_kwargs=mod).generate(seed=seed)


@pytest.mark.parametrize("nbins,seg_len,nfreq", [(64, 4096, 37), (32, 2048, 9), (128, 1024, 5)])
def test_complex_leaf_fold(gpu_device, nbins, seg_len, nfreq):
    _case(f"complex leaf fold nbins={nbins} seg_len={seg_len} nfreq={nfreq}")
    rng = np.random.default_rng(nbins)
    nseg = 6
    ts_e = rng.normal(size=nseg * seg_len).astype(np.float32)
    ts_v = rng.uniform(0.5, 1.5, nseg * seg_len).astype(np.float32)
    freqs = np.linspace(150.0, 210.0, nfreq)
    t_ref = seg_len * 64e-6 / 2
    plan = kernels.ComplexFoldPlan(freqs, seg_len, nbins, 64e-6, t_ref)
    got = plan.fold(_dev(ts_e), _dev(ts_v)).cpu().numpy()
    exp = co.brutefold_complex(ts_e, ts_v, freqs, seg_len, nbins, 64e-6, t_ref)
    scale_e = np.abs(ts_e.reshape(nseg, seg_len)).sum(axis=1)[:, None, None]
    scale_v = np.abs(ts_v.reshape(nseg, seg_len)).sum(axis=1)[:, None, None]
    err_e = np.abs(got[:, :, 0] - exp[:, :, 0]) / scale_e
    err_v = np.abs(got[:, :, 1] - exp[:, :, 1]) / scale_v
    print(f"[result] leaf fold max error / sum|x|: e {err_e.max():.2e}, v {err_v.max():.2e}",
          flush=True)
    assert err_e.max() < 2e-6 and err_v.max() < 2e-6


@pytest.mark.parametrize("P,nsamps,nbins", [(1, 2**18, 64), (2, 2**18, 64), (2, 2**17, 32)])
def test_complex_levels(gpu_device, P, nsamps, nbins):
    _case(f"complex levels P={P} nsamps={nsamps} nbins={nbins}")
    cfg = _cfg(P, nsamps, nbins)
    ts = _series(cfg)
    plan = FFAPlan(cfg)
    got = plan.run(_dev(ts.ts_e), _dev(ts.ts_v)).cpu().numpy()
    exp, _ = co.ffa_fold_complex(ts.ts_e, ts.ts_v, cfg)
    assert got.shape == exp.shape
    scale = np.abs(exp).max(axis=-1, keepdims=True)
    rel = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"[result] complex levels max error / max |X|: {rel:.2e}", flush=True)
    assert np.all(np.abs(got - exp) <= 1e-5 * np.maximum(scale, 1e-3 * np.abs(exp).max()))


def test_complex_scorer(gpu_device):
    _case("complex scorer")
    rng = np.random.default_rng(4)
    nbins, n = 64, 500
    prof = rng.normal(size=(n, 2, nbins)).astype(np.float32)
    prof[:, 1] = rng.uniform(50.0, 60.0, (n, nbins))
    folds = np.fft.rfft(prof).astype(np.complex64)
    widths = np.array([1, 2, 3, 5, 8, 13])
    got = kernels.score_boxcar_complex(_dev(folds), widths, nbins).cpu().numpy()
    exp = orc.boxcar_snr(co.fold_norm_complex(folds, nbins), widths)
    err = np.abs(got - exp).max() / np.abs(exp).max()
    print(f"[result] complex scorer max error / max S/N: {err:.2e}", flush=True)
    assert err <= 1e-5


@pytest.mark.parametrize("P", [1, 2])
def test_complex_search_snr(gpu_device, P):
    _case(f"complex ffa_search P={P}")
    cfg = _cfg(P, 2**18, 64)
    ts = _series(cfg)
    dyp, pgram = ffa_search(ts, cfg)
    exp = co.ffa_search_snr_complex(ts.ts_e, ts.ts_v, cfg)
    got = pgram.snrs.reshape(exp.shape)
    peak = float(np.abs(exp).max())
    err = float(np.abs(got - exp).max()) / peak
    print(f"[result] complex search S/N max error / peak: {err:.2e} (peak {peak:.2f})",
          flush=True)
    assert err <= 1e-5
    assert np.unravel_index(np.argmax(got), got.shape) == np.unravel_index(np.argmax(exp),
                                                                           exp.shape)
    # DynamicProgramming keeps the reference's complex layout and normalisation
    assert dyp.fold.dtype == np.complex64 and dyp.fold.shape[-1] == cfg.nbins // 2 + 1
    norm = dyp.get_fold_norm()
    assert norm.shape[-1] == cfg.nbins
