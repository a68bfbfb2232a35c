"""Search read-outs the notebooks use (``src/pyloki/periodogram.py``).

- ``Periodogram``: the FFA S/N grid over (accel..., freq, width) with best-parameter
  summaries, slices, 1-D / 2-D plots and save / load (``periodogram.py:24-200``).
- ``ScatteredPeriodogram``: the candidates of EP runs as a table, read from a results file
  (``:203-672``).
- ``PruningStatsPlotter``: the per-level statistics of EP runs (``:675-765``).

The reference builds on xarray and seaborn, which this package does not depend on: the
grid is a small labelled array (``LabeledArray``: dims, coords, ``isel``, ``max``) and
the plots are plain matplotlib. ``Periodogram.save`` writes ``.npz`` (the reference writes
netCDF); the EP results file is the ``.npz`` archive of ``pyloki_amd.io.cands``.
"""

from __future__ import annotations

from pathlib import Path
from typing import ClassVar

import numpy as np

from pyloki_amd.io.cands import read_results


def get_precision(d: float) -> int:
    """Decimal places to show for a value of resolution ``d`` (``periodogram.py:17-21``)."""
    if d == 0:
        return 1
    return max(0, -int(np.floor(np.log10(abs(d))))) + 1


def find_nearest_sorted_idx(array: np.ndarray, value: float, rtol: float = 1e-5,
                            atol: float = 1e-8) -> int:
    """Index of the closest value in a sorted array, ties to the smaller value
    (``np_utils.py:13-55``)."""
    n = len(array)
    if n == 0:
        msg = "Array must not be empty"
        raise ValueError(msg)
    idx = int(np.searchsorted(array, value, side="left"))
    if idx > 0:
        diff_prev = abs(value - array[idx - 1])
        diff_curr = abs(value - array[idx]) if idx < n else np.inf
        if diff_prev <= diff_curr * (1 + rtol) + atol:
            return idx - 1
    return min(idx, n - 1)


class _Coord:
    def __init__(self, values: np.ndarray) -> None:
        self.values = np.asarray(values)

    def to_numpy(self) -> np.ndarray:
        return self.values


class LabeledArray:
    """An n-d array with named dimensions and a coordinate array per dimension: the part
    of ``xarray.DataArray`` the periodogram uses."""

    def __init__(self, values: np.ndarray, dims: list[str], coords: dict[str, np.ndarray],
                 attrs: dict | None = None) -> None:
        self.values = np.asarray(values)
        self.dims = tuple(dims)
        self.coords = {d: _Coord(coords[d]) for d in self.dims}
        self.attrs = dict(attrs or {})
        if self.values.shape != tuple(len(coords[d]) for d in self.dims):
            msg = "values do not match the coordinates"
            raise ValueError(msg)

    @property
    def shape(self) -> tuple[int, ...]:
        return self.values.shape

    def __getitem__(self, name: str) -> _Coord:
        return self.coords[name]

    def to_numpy(self) -> np.ndarray:
        return self.values

    def isel(self, indexers: dict[str, int] | None = None) -> LabeledArray:
        """Fix the named dimensions at integer positions (they are dropped)."""
        indexers = indexers or {}
        for d in indexers:
            if d not in self.dims:
                msg = f"Invalid dimension: {d}"
                raise ValueError(msg)
        sl = tuple(int(indexers[d]) if d in indexers else slice(None) for d in self.dims)
        keep = [d for d in self.dims if d not in indexers]
        return LabeledArray(self.values[sl], keep,
                            {d: self.coords[d].values for d in keep}, self.attrs)

    def max(self, dim: list[str] | str | None = None) -> LabeledArray | _Scalar:
        if dim is None:
            return _Scalar(float(np.max(self.values)))
        dims = [dim] if isinstance(dim, str) else list(dim)
        if not dims:
            return self
        axes = tuple(self.dims.index(d) for d in dims)
        keep = [d for d in self.dims if d not in dims]
        return LabeledArray(np.max(self.values, axis=axes), keep,
                            {d: self.coords[d].values for d in keep}, self.attrs)


class _Scalar:
    def __init__(self, v: float) -> None:
        self.v = v

    def item(self) -> float:
        return self.v


def _plt():
    from matplotlib import pyplot as plt

    return plt


class Periodogram:
    """FFA S/N grid (``periodogram.py:24-200``); dims are the search parameters in
    ``crackle..accel, freq`` order, then ``width``."""

    REQUIRED_PARAMS: ClassVar[tuple[str, str]] = ("freq", "width")
    OPTIONAL_PARAMS: ClassVar[tuple[str, ...]] = ("crackle", "snap", "jerk", "accel")

    def __init__(self, params: dict[str, np.ndarray], snrs: np.ndarray, tobs: float) -> None:
        self.params = params
        self.snrs = np.asarray(snrs)
        self.tobs = tobs
        if not all(p in params for p in self.REQUIRED_PARAMS):
            msg = f"Missing required parameters: {self.REQUIRED_PARAMS}"
            raise ValueError(msg)
        for p in params:
            if p not in self.valid_params:
                msg = f"Invalid parameter: {p}"
                raise ValueError(msg)
        shape = tuple(len(params[n]) for n in self.param_names)
        if self.snrs.shape != shape:
            msg = f"SNR shape {self.snrs.shape} does not match expected shape {shape}"
            raise ValueError(msg)
        self.data = LabeledArray(self.snrs, self.param_names,
                                 {n: params[n] for n in self.param_names}, {"tobs": tobs})

    @property
    def valid_params(self) -> tuple[str, ...]:
        return self.OPTIONAL_PARAMS + self.REQUIRED_PARAMS

    @property
    def param_names(self) -> list[str]:
        return [p for p in self.valid_params if p in self.params]

    @property
    def dims(self) -> list[str]:
        return self.param_names

    @property
    def ndim(self) -> int:
        return len(self.param_names)

    def get_slice(self, fixed_param_idx: dict[str, int]) -> LabeledArray:
        return self.data.isel(fixed_param_idx)

    def find_best_indices(self) -> tuple[int, ...]:
        flat_index = np.argmax(self.snrs)
        return tuple(int(i) for i in np.unravel_index(int(flat_index), self.snrs.shape))

    def find_best_params(self) -> dict[str, float]:
        best = self.find_best_indices()
        out = {"snr": float(self.snrs[best])}
        for name, idx in zip(self.param_names, best):
            out[name] = self.params[name][idx]
        return out

    def get_indices_summary(self, true_values: dict[str, float] | None = None) -> str:
        best = self.find_best_indices()
        summary = []
        if true_values:
            true_idx = [find_nearest_sorted_idx(self.params[n], true_values[n])
                        for n in self.param_names if n in true_values]
            summary.append(f"True param indices: {tuple(true_idx)}")
        summary.append(f"Best param indices: {best}")
        return "\n".join(summary)

    def get_summary(self) -> str:
        best = self.find_best_params()
        summary = [f"Best S/N: {best['snr']:.2f}", f"Best Period: {1 / best['freq']}"]
        summary += [f"Best {n}: {best[n]}" for n in self.param_names if n != "snr"]
        return "\n".join(summary)

    def plot_1d(self, param: str, fixed_param_idx: dict[str, int] | None = None,
                x_lim: tuple[float, float] | None = None, figsize: tuple[float, float] = (10, 5),
                dpi: int = 100):
        """Best S/N over the other dimensions against ``param`` (``periodogram.py:89-126``)."""
        if param not in self.param_names:
            msg = f"Invalid parameter: {param}"
            raise ValueError(msg)
        fixed_param_idx = fixed_param_idx or {}
        fixed = ", ".join(f"{k}={v}" for k, v in fixed_param_idx.items())
        title = f"Best S/N vs {param}" + (f" for {fixed}" if fixed else "")
        x = self.data[param].to_numpy()
        sl = self.get_slice(fixed_param_idx)
        y = sl.max(dim=[d for d in sl.dims if d != param]).to_numpy()
        plt = _plt()
        fig, ax = plt.subplots(figsize=figsize, dpi=dpi)
        ax.plot(x, y, marker="o", markersize=2, alpha=0.5)
        ax.set_xlim(x_lim if x_lim else (x.min(), x.max()))
        ax.set_xlabel(f"Trial {param}", fontsize=16)
        ax.set_ylabel("S/N", fontsize=16)
        ax.set_title(title, fontsize=18)
        ax.grid(linestyle=":")
        return fig

    def plot_2d(self, param_x: str, param_y: str, fixed_param_idx: dict[str, int] | None = None,
                x_lim: tuple[float, float] | None = None, y_lim: tuple[float, float] | None = None,
                figsize: tuple[float, float] = (10, 8), dpi: int = 100):
        """Best S/N over the other dimensions on the (``param_x``, ``param_y``) plane
        (``periodogram.py:128-163``)."""
        if param_x not in self.param_names or param_y not in self.param_names:
            msg = f"Invalid parameter: {param_x} or {param_y}"
            raise ValueError(msg)
        fixed_param_idx = fixed_param_idx or {}
        x = self.data[param_x].to_numpy()
        y = self.data[param_y].to_numpy()
        sl = self.get_slice(fixed_param_idx)
        z = sl.max(dim=[d for d in sl.dims if d not in (param_x, param_y)])
        zz = z.to_numpy()
        if z.dims.index(param_y) > z.dims.index(param_x):
            zz = zz.T  # rows along param_y
        plt = _plt()
        fig, ax = plt.subplots(figsize=figsize, dpi=dpi)
        im = ax.imshow(zz, aspect="auto", origin="lower",
                       extent=(x.min(), x.max(), y.min(), y.max()), cmap="viridis")
        ax.set_xlabel(f"Trial {param_x}", fontsize=16)
        ax.set_ylabel(f"Trial {param_y}", fontsize=16)
        ax.set_title(f"Best S/N: {param_x} vs {param_y}", fontsize=18)
        if x_lim:
            ax.set_xlim(x_lim)
        if y_lim:
            ax.set_ylim(y_lim)
        fig.colorbar(im, ax=ax, label="S/N")
        return fig

    def save(self, filename: str | Path) -> None:
        """The grid, its coordinates and ``tobs`` as ``.npz``."""
        np.savez(filename, snrs=self.snrs, tobs=np.array(self.tobs),
                 dims=np.array(self.param_names),
                 **{f"coord_{n}": self.params[n] for n in self.param_names})

    @classmethod
    def load(cls, filename: str | Path) -> Periodogram:
        with np.load(filename, allow_pickle=False) as z:
            dims = [str(d) for d in z["dims"]]
            params = {d: z[f"coord_{d}"] for d in dims}
            return cls(params=params, snrs=z["snrs"], tobs=float(z["tobs"]))

    def __str__(self) -> str:
        info = ", ".join(f"{k}: {len(v)}" for k, v in self.params.items())
        return f"Periodogram({info}, tobs: {self.tobs})"

    __repr__ = __str__


class ScatteredPeriodogram:
    """Candidates of EP runs (``periodogram.py:203-672``): one row per reported parameter
    set with its values and widths (``<p>``, ``d<p>``), ``score_ep``, ``score`` and
    ``run_id``."""

    def __init__(self, param_names: list[str]) -> None:
        import pandas as pd

        self.param_names = list(param_names)
        self.data = pd.DataFrame()

    @property
    def n_runs(self) -> int:
        if self.data.empty or "run_id" not in self.data.columns:
            return 0
        return int(self.data["run_id"].nunique())

    @staticmethod
    def _score_filtered(data, score_col: str, snr_min: float | None):
        return data if snr_min is None else data.loc[data[score_col] >= snr_min]

    def add_run(self, param_sets: np.ndarray, scores_ep: np.ndarray, scores: np.ndarray,
                run_id: str) -> None:
        import pandas as pd

        self._validate_inputs(param_sets, scores_ep, scores)
        run = {}
        for i, name in enumerate(self.param_names):
            run[name] = param_sets[:, i, 0]
            run[f"d{name}"] = param_sets[:, i, 1]
        run["score_ep"] = scores_ep
        run["score"] = scores
        df = pd.DataFrame.from_dict(run)
        df["run_id"] = run_id
        self.data = pd.concat([self.data, df], ignore_index=True)

    def _formatters(self, data) -> tuple[dict, str]:
        first = data.iloc[0]
        fmt = {p: f"{{:.{get_precision(float(first[f'd{p}']))}f}}" for p in self.param_names}
        dparams = ", ".join(f"d{p}: {first[f'd{p}']:.10g}" for p in self.param_names)
        return fmt, dparams

    def _row_str(self, row, fmt) -> str:
        params = ", ".join(f"{p}: {fmt[p].format(row[p])}" for p in self.param_names)
        return (f"Run: {row['run_id']}, S/N (EP): {row['score_ep']:.2f}, "
                f"S/N: {row['score']:.2f}, {params}")

    def get_summary_cands(self, n: int = 10, score_type: str = "score",
                          run_id: str | None = None, snr_min: float | None = None) -> str:
        """The top ``n`` candidates (``periodogram.py:250-314``)."""
        if score_type not in ("score_ep", "score"):
            msg = f"Invalid score type: {score_type}"
            raise ValueError(msg)
        data = self.data if run_id is None else self.data[self.data["run_id"] == run_id]
        scope = " (no candidates in this selection)" if run_id is not None else ""
        if data.empty:
            return f"Top candidates:{scope}\nNo candidates."
        data = self._score_filtered(data, score_type, snr_min)
        if data.empty:
            thresh = f" (S/N >= {snr_min:g})" if snr_min is not None else ""
            return f"Top candidates:{scope}{thresh}\nNo candidates."
        top = data.nlargest(n, score_type)
        fmt, dparams = self._formatters(data)
        lines = ["Top candidates:", f"dparams: {dparams}"]
        lines += [self._row_str(row, fmt) for _, row in top.iterrows()]
        return "\n".join(lines)

    def get_best_in_each_run(self, score_type: str = "score",
                             snr_min: float | None = None) -> str:
        """The best candidate of every run (``periodogram.py:316-370``)."""
        if self.data.empty:
            return "Best candidate in each run:\nNo candidates."
        if score_type not in ("score_ep", "score"):
            msg = f"Invalid score type: {score_type}"
            raise ValueError(msg)
        data = self._score_filtered(self.data, score_type, snr_min)
        if data.empty:
            thresh = f" (S/N >= {snr_min:g})" if snr_min is not None else ""
            return f"Best candidate in each run:{thresh}\nNo candidates."
        best = data.loc[data.groupby("run_id")[score_type].idxmax()]
        best = best.sort_values("run_id", ascending=True)
        fmt, dparams = self._formatters(data)
        lines = ["Best candidate in each run:", f"dparams: {dparams}"]
        lines += [self._row_str(row, fmt) for _, row in best.iterrows()]
        return "\n".join(lines)

    def plot_correlation(self, param_x: str, param_y: str, score_type: str = "score",
                         run_id: str | None = None, true_values: dict[str, float] | None = None,
                         x_lim: tuple[float, float] | None = None,
                         y_lim: tuple[float, float] | None = None,
                         figsize: tuple[float, float] = (8, 6), dpi: int = 100,
                         cmap: str = "magma_r", snr_min: float | None = None):
        """Scatter of two parameters coloured and sized by score (``periodogram.py:372-450``)."""
        if score_type not in ("score_ep", "score"):
            msg = f"Invalid score type: {score_type}"
            raise ValueError(msg)
        data = self.data if run_id is None else self.data[self.data["run_id"] == run_id]
        data = self._score_filtered(data, score_type, snr_min)
        plt = _plt()
        fig, ax = plt.subplots(figsize=figsize, dpi=dpi)
        if data.empty:
            ax.text(0.5, 0.5, "No candidates to plot", ha="center", va="center",
                    transform=ax.transAxes)
            ax.set_xlabel(f"Trial {param_x}")
            ax.set_ylabel(f"Trial {param_y}")
            ax.set_title(f"{param_x} vs {param_y}")
            return fig
        sc = data[score_type].to_numpy()
        lo, hi = float(sc.min()), float(sc.max())
        sizes = 20 + 180 * ((sc - lo) / (hi - lo) if hi > lo else np.zeros_like(sc))
        pts = ax.scatter(data[param_x], data[param_y], c=sc, s=sizes, cmap=cmap, alpha=0.7,
                         edgecolors="black", linewidths=0.2, vmin=lo, vmax=hi)
        label = "Score (EP)" if score_type == "score_ep" else "Score (Integrated)"
        fig.colorbar(pts, ax=ax, pad=0.02, aspect=30, label=label)
        if true_values:
            if param_x in true_values:
                ax.axvline(true_values[param_x], c="r", ls="--", label=f"True {param_x}")
            if param_y in true_values:
                ax.axhline(true_values[param_y], c="r", ls="--", label=f"True {param_y}")
        if x_lim:
            ax.set_xlim(x_lim)
        if y_lim:
            ax.set_ylim(y_lim)
        ax.set_xlabel(f"Trial {param_x}")
        ax.set_ylabel(f"Trial {param_y}")
        ax.set_title(f"Best S/N: {param_x} vs {param_y}")
        return fig

    def plot_scores(self, kind: str = "scatter", score_type: str = "score",
                    run_id: str | None = None, figsize: tuple[float, float] = (7, 3.5),
                    dpi: int = 100, snr_min: float | None = None):
        """Score distribution, scatter or histogram (``periodogram.py:452-563``)."""
        if score_type not in ("score_ep", "score"):
            msg = f"Invalid score type: {score_type}"
            raise ValueError(msg)
        label = "Score (EP)" if score_type == "score_ep" else "Score (Integrated)"
        data = self.data if run_id is None else self.data[self.data["run_id"] == run_id]
        data = self._score_filtered(data, score_type, snr_min)
        plt = _plt()
        fig, ax = plt.subplots(figsize=figsize, dpi=dpi)
        if data.empty:
            ax.text(0.5, 0.5, "No candidates to plot", ha="center", va="center",
                    transform=ax.transAxes)
            ax.set_title(f"{label} Distribution")
            return fig
        if kind == "scatter":
            ax.scatter(range(len(data)), data[score_type], s=12, edgecolor="tab:blue",
                       facecolor="white", linewidth=0.6)
            if run_id is None:
                prev = 0
                for i, rid in enumerate(data["run_id"].unique()):
                    idx = prev + int((data["run_id"] == rid).sum())
                    ax.axvspan(prev, idx, color=f"C{i % 5}", alpha=0.08)
                    if idx < len(data):
                        ax.axvline(idx, color="gray", linestyle=":", alpha=0.5)
                    prev = idx
                ax.set_title(f"{label} Distribution Across Runs")
            else:
                ax.set_title(f"{label} Distribution for Run {run_id}")
            ax.set_xlabel("Candidate Index")
            ax.set_ylabel("Score (S/N)")
        else:
            if run_id is not None:
                ax.hist(data[score_type], bins=50, histtype="step", color="black")
                ax.set_title(f"{label} Distribution for Run {run_id}")
            else:
                for rid in data["run_id"].unique():
                    ax.hist(data.loc[data["run_id"] == rid, score_type], bins=50,
                            histtype="step", label=str(rid))
                ax.set_title(f"{label} Distribution Across Runs")
                ax.legend(loc="upper right", ncol=3, fontsize="xx-small", frameon=False,
                          title="run", title_fontsize="xx-small")
            ax.set_yscale("log")
            ax.set_xlabel("Score (S/N)")
            ax.set_ylabel("Count")
        return fig

    def get_stats(self):
        """Mean, spread and median of every parameter and its mean width
        (``periodogram.py:565-592``)."""
        import pandas as pd

        cols = ["parameter", "mean", "std", "median", "mean_uncertainty"]
        if self.data.empty:
            return pd.DataFrame(columns=cols)
        rows = []
        for p in self.param_names:
            s = self.data[p].describe()
            u = self.data[f"d{p}"].describe()
            rows.append({"parameter": p, "mean": s["mean"], "std": s["std"],
                         "median": s["50%"], "mean_uncertainty": u["mean"]})
        return pd.DataFrame(rows)

    @classmethod
    def load(cls, filename: str | Path) -> ScatteredPeriodogram:
        """Every run of an EP results file (``periodogram.py:594-609``)."""
        res = read_results(filename)
        pgram = cls(param_names=list(res["attrs"]["param_names"]))
        for run_name, run in res["runs"].items():
            pgram.add_run(run["param_sets"], run["scores_ep"], run["scores"],
                          run_name.split("seg_")[-1])
        return pgram

    def _validate_inputs(self, param_sets: np.ndarray, scores_ep: np.ndarray,
                         scores: np.ndarray) -> None:
        if param_sets.ndim != 3 or param_sets.shape[-1] != 2:
            msg = f"param_sets should be (n_sugg, n_params, 2), got {param_sets.shape}"
            raise ValueError(msg)
        shape = (param_sets.shape[0],)
        if scores.shape != shape:
            msg = f"scores shape {scores.shape} does not match expected shape {shape}"
            raise ValueError(msg)
        if scores_ep.shape != shape:
            msg = f"scores_ep shape {scores_ep.shape} does not match expected shape {shape}"
            raise ValueError(msg)
        nparams = param_sets.shape[1] - 2
        if len(self.param_names) != nparams:
            msg = f"Got {len(self.param_names)} names for {nparams} parameters"
            raise ValueError(msg)

    def __str__(self) -> str:
        lines = ["ScatteredPeriodogram Summary:", f"Parameters: {', '.join(self.param_names)}"]
        if self.data.empty:
            return "\n".join([*lines, "Total runs: 0", "Total candidates: 0",
                              "\nRun Statistics:", "(no data)"])
        stats = self.data.groupby("run_id").agg({"score": ["count", "min", "max"]})
        lines += [f"Total runs: {len(stats)}", f"Total candidates: {len(self.data)}",
                  "\nRun Statistics:"]
        for rid, st in stats.iterrows():
            lines.append(f"Run {rid}: {int(st['score']['count'])} candidates, "
                         f"max S/N: {st['score']['max']:.2f}, "
                         f"min S/N: {st['score']['min']:.2f}")
        return "\n".join(lines)

    __repr__ = __str__


class PruningStatsPlotter:
    """Per-level statistics of EP runs (``periodogram.py:675-765``)."""

    def __init__(self) -> None:
        import pandas as pd

        self.data = pd.DataFrame()

    @property
    def n_runs(self) -> int:
        if self.data.empty or "run_id" not in self.data.columns:
            return 0
        return int(self.data["run_id"].nunique())

    def add_run(self, level_stats: np.ndarray, run_id: str) -> None:
        import pandas as pd

        if not isinstance(level_stats, np.ndarray):
            msg = "level_stats should be a numpy array"
            raise TypeError(msg)
        if level_stats.dtype.names is None or len(level_stats.dtype.names) != 9:
            msg = "level_stats should have 9 fields"
            raise ValueError(msg)
        df = pd.DataFrame.from_records(level_stats)
        df["run_id"] = run_id
        self.data = pd.concat([self.data, df], ignore_index=True)

    def get_level_stats(self, run_id: str | None = None):
        return self.data if run_id is None else self.data[self.data["run_id"] == run_id]

    def plot_level_stats(self, run_id: str | None = None, figsize: tuple[float, float] = (10, 5),
                         dpi: int = 100):
        """Surviving leaves per level, one line per run, log scale."""
        data = self.get_level_stats(run_id)
        plt = _plt()
        fig, ax = plt.subplots(figsize=figsize, dpi=dpi)
        for rid in data["run_id"].unique() if not data.empty else []:
            d = data[data["run_id"] == rid]
            ax.plot(d["level"], d["n_leaves_surv"], label=str(rid))
        if not data.empty:
            ax.legend(loc="lower center", ncol=3, fontsize="xx-small", frameon=False,
                      title="run", title_fontsize="xx-small")
        ax.set_yscale("log")
        ax.set_title("Pruning Complexity")
        ax.set_xlabel("Level")
        ax.set_ylabel("Number of Leaves Surviving")
        return fig

    @classmethod
    def load(cls, filename: str | Path) -> PruningStatsPlotter:
        res = read_results(filename)
        out = cls()
        for run_name, run in res["runs"].items():
            out.add_run(run["level_stats"], run_name)
        return out
