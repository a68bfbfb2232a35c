"""``ffa_search``: FFA, normalise, score, periodogram (``src/pyloki/search.py:17-53``)."""

from __future__ import annotations

from pyloki_amd import kernels
from pyloki_amd.config import PulsarSearchConfig
from pyloki_amd.ffa import DynamicProgramming
from pyloki_amd.io.timeseries import TimeSeries
from pyloki_amd.periodogram import Periodogram


def ffa_search(
    tseries: TimeSeries,
    search_cfg: PulsarSearchConfig,
    *,
    quiet: bool = False,  # noqa: ARG001
    show_progress: bool = False,
) -> tuple[DynamicProgramming, Periodogram]:
    """Fast Folding Algorithm search of a time series on the GPU.

    Scores segment 0 of the last level (``ffa.py:407-411``, ``search.py:46-47``) with the
    per-width boxcar S/N; the normalisation ``e/sqrt(v)`` is fused into the scorer.
    """
    dyp = DynamicProgramming(tseries, search_cfg)
    dyp.initialize()
    dyp.execute(show_progress=show_progress)
    widths = search_cfg.score_widths
    if search_cfg.use_fourier:
        snr_dev = kernels.score_boxcar_complex(dyp.fold_device[0], widths, search_cfg.nbins, 1.0)
    else:
        snr_dev = kernels.score_boxcar(dyp.fold_device[0], widths, 1.0)
    counts = [len(p) for p in dyp.param_arr]
    snrs = snr_dev.cpu().numpy().reshape(*counts, len(widths))
    pgram = Periodogram(params={"width": widths, **dyp.param_arr_dict}, snrs=snrs,
                        tobs=tseries.tobs)
    return dyp, pgram
