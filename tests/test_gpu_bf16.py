"""BF16-storage / FP32-accumulate FFA: bit-exact against the oracle's BF16 emulation, and
the S/N error of FP32 and BF16 storage against the FP64 oracle, stated and bounded."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from pyloki_amd import kernels
from pyloki_amd.config import ParamLimits, PulsarSearchConfig
from pyloki_amd.ffa import FFAPlan, ffa_batch
from pyloki_amd.simulation.pulse import PulseSignalConfig

pytestmark = pytest.mark.gpu


def _case(msg: str) -> None:
    print(f"[case] {msg}", flush=True)


def _jerk_cfg(nsamps: int, nbins: int = 128, nseg_ffa: int = 1) -> PulsarSearchConfig:
    lim = ParamLimits.from_taylor((330.0, 336.7), accel=(-30.0, 30.0), jerk=(-2.0, 2.0)).limits
    return PulsarSearchConfig(nsamps=nsamps, tsamp=64e-6, nbins=nbins, eta=1.0, param_limits=lim,
                              bseg_brute=nsamps // 512, bseg_ffa=nsamps // nseg_ffa,
                              ducy_max=0.2, wtsp=1.5, use_fourier=False)


def _series(cfg: PulsarSearchConfig, seed: int = 3):
    sim = PulseSignalConfig(period=3e-3, dt=cfg.tsamp, nsamps=cfg.nsamps, snr=20.0, ducy=0.08,
                            mod_kwargs={"acc": 10.0, "jerk": 0.5})
    return sim.generate_analytic(seed=seed)


@pytest.mark.parametrize("nbins,P", [(64, 2), (128, 3)])
def test_bf16_levels_bit_exact(gpu_device, nbins, P):
    _case(f"bf16 levels nbins={nbins} P={P}")
    if P == 2:
        lim = ParamLimits.from_taylor((195.0, 205.0), accel=(-50.0, 50.0)).limits
        cfg = PulsarSearchConfig(nsamps=2**18, tsamp=64e-6, nbins=nbins, eta=1.0,
                                 param_limits=lim, bseg_brute=4096, use_fourier=False)
        ts = PulseSignalConfig(period=5e-3, dt=64e-6, nsamps=cfg.nsamps, snr=20.0,
                               ducy=0.08).generate_analytic(seed=1)
    else:
        cfg = _jerk_cfg(2**19, nbins)
        ts = _series(cfg)
    plan = FFAPlan(cfg)
    e = torch.from_numpy(ts.ts_e).cuda()
    v = torch.from_numpy(ts.ts_v).cuda()
    got = plan.run(e, v, dtype=torch.bfloat16)
    exp, _ = orc.ffa_fold_bf16(ts.ts_e, ts.ts_v, cfg)
    np.testing.assert_array_equal(got.view(torch.int16).cpu().numpy().astype(np.uint16), exp)
    # fused last level + score on BF16 storage == score of the stored BF16 fold
    folds = plan.run(e, v, dtype=torch.bfloat16, keep=True)
    lp = plan.levels[-1]
    widths = cfg.score_widths
    fused = kernels.ffa_shift_add_score(folds[-2], lp.table, lp.ncoords, widths).cpu().numpy()
    last = orc.bf16_to_f32(folds[-1].view(torch.int16).cpu().numpy().astype(np.uint16))
    np.testing.assert_array_equal(fused, orc.boxcar_snr(orc.fold_norm(last), widths))
    sc = kernels.score_boxcar(folds[-1], widths).cpu().numpy()
    np.testing.assert_array_equal(sc, fused)


def test_snr_error_fp32_and_bf16_vs_fp64(gpu_device):
    """Jerk search (Taylor order 3, nbins 128, accel +-30, jerk +-2, P = 3 ms) on a 2^20
    sample series: S/N error of FP32 folds (target <= 1e-5 relative) and of BF16 fold
    storage, both against the FP64 oracle, relative to the grid's peak S/N."""
    _case("snr error fp32/bf16 vs fp64")
    cfg = _jerk_cfg(2**20)
    ts = _series(cfg)
    plan = FFAPlan(cfg)
    series = [(torch.from_numpy(ts.ts_e).cuda(), torch.from_numpy(ts.ts_v).cuda())]
    snr32 = ffa_batch(plan, series, cfg.score_widths)[0].cpu().numpy()
    snr16 = ffa_batch(plan, series, cfg.score_widths, dtype=torch.bfloat16)[0].cpu().numpy()
    snr64 = orc.ffa_search_snr_f64(ts.ts_e, ts.ts_v, cfg).reshape(snr32.shape)
    peak = float(np.max(np.abs(snr64)))
    err32 = float(np.max(np.abs(snr32 - snr64))) / peak
    err16 = float(np.max(np.abs(snr16 - snr64))) / peak
    med16 = float(np.median(np.abs(snr16 - snr64)))
    print(f"[result] peak S/N {peak:.3f}: FP32 max err/peak {err32:.3e}; BF16 max err/peak "
          f"{err16:.3e}, median abs err {med16:.3e}", flush=True)
    assert err32 <= 1e-5
    assert err16 <= 5e-2
    best64 = np.unravel_index(np.argmax(snr64), snr64.shape)
    assert abs(snr16[best64] - snr64[best64]) / snr64[best64] <= 2e-2
