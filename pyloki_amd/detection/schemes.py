"""Threshold schedules for the EP search and their state records.

Mirrors ``pyloki.detection.schemes`` (``src/pyloki/detection/schemes.py``): the closed-form
schedules ``bound_scheme`` and ``trials_scheme``, the per-stage record ``StateInfo`` and
its sequence ``StatesInfo``, and ``DynamicThresholdSchemeAnalyser`` (backtracking of the
dynamic program's table and its plots). The reference's Tk front end
(``ThresholdAnalyzerApp``) is a desktop UI and is not part of this package.
"""

from __future__ import annotations

import json
import logging
from dataclasses import asdict, dataclass, field, fields
from pathlib import Path

import numpy as np
from scipy import stats

logger = logging.getLogger(__name__)


def bound_scheme(nstages: int, snr_bound: float) -> np.ndarray:
    """``schemes.py:29-50``: thresholds whose squares grow linearly with the stage, up to
    ``snr_bound**2`` at the last segment; FP32, one per stage (level 1 .. nstages)."""
    nsegments = nstages + 1
    thresh_sn2 = (np.arange(1, nsegments + 1) * snr_bound**2 / nsegments).astype(np.float32)
    return np.sqrt(thresh_sn2[1:])


def trials_scheme(branching_pattern: np.ndarray, trials_start: int = 1) -> np.ndarray:
    """``schemes.py:53-76``: the Gaussian tail level of one false alarm in the cumulative
    number of trials, ``isf(1 / (trials_start * prod(bp[:k+1])))``."""
    trials = np.cumprod(branching_pattern) * trials_start
    return stats.norm.isf(1 / trials)


@dataclass(frozen=True, kw_only=True)
class StateInfo:
    """One stage of a threshold scheme (``schemes.py:79-122``)."""

    success_h0: float
    success_h1: float
    complexity: float
    complexity_cumul: float
    success_h1_cumul: float
    nbranches: float
    threshold: float

    @property
    def cost(self) -> float:
        return self.complexity_cumul / self.success_h1_cumul

    @classmethod
    def from_record(cls, state) -> StateInfo:
        names = {f.name for f in fields(cls)}
        dtype_names = state.dtype.names
        if dtype_names is None:
            msg = "State dtype has no field names"
            raise ValueError(msg)
        missing = names - set(dtype_names)
        if missing:
            msg = f"State dtype is missing required fields: {missing}"
            raise ValueError(msg)
        return cls(**{n: float(np.asarray(state[n]).item()) for n in names})


@dataclass(frozen=True)
class StatesInfo:
    """The stages of one scheme (``schemes.py:125-204``)."""

    entries: list[StateInfo] = field(default_factory=list)

    @property
    def thresholds(self) -> np.ndarray:
        return np.array([e.threshold for e in self.entries], dtype=np.float32)

    def get_info(self, key: str) -> np.ndarray:
        return np.array([getattr(e, key) for e in self.entries], dtype=np.float32)

    def summary_metrics(self) -> dict[str, float]:
        """The figures ``print_summary`` tabulates (``schemes.py:142-191``)."""
        bp = self.get_info("nbranches")
        survive_prob = self.get_info("success_h1_cumul")[-1]
        n_independent = len(self.thresholds) / self.thresholds[-1]
        return {
            "Branching mean": float(2 ** (np.mean(np.log2(bp)))),
            "Branching max": float(np.max(bp)),
            "Total enumerated options": float(np.sum(np.log2(bp))),
            "Pruning complexity": float(np.log2(self.get_info("complexity_cumul")[-1])),
            "Crude survival probability": float(survive_prob),
            "Total cost": float(np.log2(self.get_info("cost")[-1])),
            "Number of independent trials": float(n_independent),
            "Total survival probability": float(1 - (1 - survive_prob) ** n_independent),
        }

    def print_summary(self) -> None:
        """The branching pattern and thresholds (copy-pasteable) and the metrics table."""
        from rich.console import Console
        from rich.table import Table

        console = Console()
        kw = {"precision": 1, "floatmode": "fixed", "separator": ",",
              "threshold": np.iinfo(np.int32).max}
        console.print(f"[bold cyan]Branching Pattern:[/bold cyan] "
                      f"{np.array2string(self.get_info('nbranches'), **kw)}")
        console.print(f"[bold cyan]Thresholds:[/bold cyan] "
                      f"{np.array2string(self.thresholds, **kw)}")
        console.print()
        table = Table(title="Scheme Metrics Summary", show_header=True,
                      header_style="bold magenta")
        table.add_column("Metric", style="cyan")
        table.add_column("Value", justify="right", style="green")
        for name, val in self.summary_metrics().items():
            table.add_row(name, f"{val:.2f}")
        console.print(table)

    def save(self, filename: str) -> None:
        with Path(filename).open("w") as fp:
            json.dump({"entries": [asdict(e) for e in self.entries]}, fp)

    @classmethod
    def load(cls, filename: str) -> StatesInfo:
        with Path(filename).open("r") as fp:
            data = json.load(fp)
        return cls(entries=[StateInfo(**e) for e in data["entries"]])


def _backtrack_failed(istage: int, thr: float, prob: float) -> ValueError:
    return ValueError(f"Backtracking failed at stage {istage} for threshold {thr} and "
                      f"success probability {prob}")


class DynamicThresholdSchemeAnalyser:
    """Paths through the dynamic program's table (``schemes.py:207-654``): ``states`` is
    the (stage, threshold, success bin) table of state records."""

    def __init__(self, states: np.ndarray, thresholds: np.ndarray, probs: np.ndarray,
                 branching_pattern: np.ndarray, guess_path: np.ndarray,
                 beam_width: float) -> None:
        self.states = states
        self.thresholds = thresholds
        self.probs = probs
        self.branching_pattern = branching_pattern
        self.guess_path = guess_path
        self.beam_width = beam_width
        self.nstages, self.nthresholds, self.nprobs = states.shape

    def _backtrack(self, final_state) -> StatesInfo:
        out = [StateInfo.from_record(final_state)]
        thr = final_state["threshold_prev"]
        prob = final_state["success_h1_cumul_prev"]
        for istage in range(self.nstages - 2, -1, -1):
            ithres = np.argmin(np.abs(self.thresholds - thr))
            iprob = np.digitize(prob, self.probs) - 1
            if iprob < 0:
                raise _backtrack_failed(istage, thr, prob)
            prev = self.states[istage, ithres, iprob]
            if prev["is_empty"]:
                raise _backtrack_failed(istage, thr, prob)
            out.insert(0, StateInfo.from_record(prev))
            thr = prev["threshold_prev"]
            prob = prev["success_h1_cumul_prev"]
        return StatesInfo(out)

    def backtrack_best(self, min_probs: list[float] | None) -> list[StatesInfo]:
        """Per minimum final H1 success, the least-cost complete scheme
        (``schemes.py:225-277``)."""
        if min_probs is None:
            min_probs = [self.probs[0]]
        final = self._final_cloud()
        if len(final) == 0:
            msg = "No non-empty final states to plot"
            raise ValueError(msg)
        out = []
        for min_prob in min_probs:
            sel = final[final["success_h1_cumul"] >= min_prob]
            if len(sel) == 0:
                logger.warning("No states found for min_prob %s", min_prob)
                out.append(StatesInfo([]))
                continue
            out.append(self._backtrack(sel[np.argmin(sel["cost"])]))
        return out

    def backtrack_all(self, min_prob: float) -> list[StatesInfo]:
        """Every complete scheme with final H1 success >= ``min_prob``
        (``schemes.py:279-326``)."""
        final = self._final_cloud()
        if len(final) == 0:
            msg = "No non-empty final states to plot"
            raise ValueError(msg)
        return [self._backtrack(s) for s in final[final["success_h1_cumul"] >= min_prob]]

    def _final_cloud(self) -> np.ndarray:
        final = self.states[-1]
        return final[~final["is_empty"]]

    def _frontier_by_bin(self, metric: str) -> np.ndarray:
        """One least-``metric`` final state per success bin (``schemes.py:530-560``)."""
        names = self.states.dtype.names
        if names is None or metric not in names:
            msg = f"metric must be one of {names}, got {metric!r}"
            raise ValueError(msg)
        final = self.states[-1]
        winners = []
        for r in range(self.nprobs):
            col = final[:, r]
            valid = col[~col["is_empty"]]
            if len(valid) == 0:
                continue
            score = valid[metric].astype(float)
            prob = valid["success_h1_cumul"].astype(float)
            good = np.isfinite(score) & np.isfinite(prob) & (score > 0) & (prob > 0)
            valid, score = valid[good], score[good]
            if len(valid):
                winners.append(valid[np.argmin(score)])
        out = np.empty(len(winners), dtype=final.dtype)
        for i, w in enumerate(winners):
            out[i] = w
        return out

    def plot_paths(self, best_prob: float, min_prob: float, fig=None):
        """Thresholds, H0 options, cumulative complexity and H1 success of every complete
        path (and the best one) (``schemes.py:328-424``)."""
        from matplotlib import pyplot as plt
        from matplotlib.ticker import FormatStrFormatter

        paths = self.backtrack_all(min_prob)
        best_path = self.backtrack_best(min_probs=[best_prob])[0]
        if len(best_path.entries) == 0:
            msg = f"No best path found for probability {best_prob}"
            raise ValueError(msg)
        if fig is None:
            fig, axes = plt.subplots(2, 2, figsize=(12, 8), dpi=120, constrained_layout=True)
            ax1, ax2, ax3, ax4 = axes.flatten()
        else:
            ax1, ax2, ax3, ax4 = fig.get_axes()
            for ax in (ax1, ax2, ax3, ax4):
                ax.clear()
        x = np.arange(1, self.nstages + 1)
        for path in paths:
            ax1.plot(x, path.thresholds**2, "b-", alpha=0.2)
        ax1.plot(x, best_path.thresholds**2, "r-", label=f"Best: P(H1) = {best_prob:.2f}")
        ax1.plot(x, self.guess_path**2, color="navy", ls="--", label="Guess path")
        upper = np.minimum(self.guess_path + self.beam_width, self.thresholds[-1])
        lower = np.maximum(self.guess_path - self.beam_width, 0)
        ax1.fill_between(x, lower**2, upper**2, color="cornflowerblue", alpha=0.2,
                         label="Beam width")
        ax1.set_xlabel("Pruning stage")
        ax1.set_ylabel("S/N squared")
        ax1.set_title("Threshold scheme")
        ax1.set_ylim(-0.5, self.thresholds[-1] ** 2 + 0.5)
        for ax, info, title, ylabel in (
            (ax2, "complexity", "False Alarm", r"Number of $H_{0}$ options"),
            (ax3, "complexity_cumul", r"Cumulative complexity $H_{0}$",
             r"Cumulative $H_{0}$ complexity"),
            (ax4, "success_h1_cumul", r"Cumulative Success $H_{1}$", "Detection Probability"),
        ):
            for path in paths:
                ax.plot(x, path.get_info(info), "b-", alpha=0.2)
            ax.plot(x, best_path.get_info(info), "r-", label="Best path")
            ax.set_xlabel("Pruning stage")
            ax.set_ylabel(ylabel)
            ax.set_title(title)
            ax.set_yscale("log")
        ylim = ax2.get_ylim()
        ax2.plot(x, np.cumprod(self.branching_pattern), color="k", ls="--",
                 label="Total options")
        ax2.set_ylim(bottom=0.05, top=ylim[1])
        ax1.legend(loc="upper left")
        ax2.legend(loc="upper right")
        ax3.legend(loc="lower right")
        ax4.legend(loc="upper right")
        for pval in self.probs[1:]:
            ax4.axhline(y=pval, color="gray", alpha=0.15, linestyle="-", zorder=0)
        ax4.yaxis.set_major_formatter(FormatStrFormatter("%.2f"))
        return fig, best_path

    def update_best_path(self, best_prob: float, fig):
        """Redraw only the best path of a ``plot_paths`` figure (``schemes.py:426-466``)."""
        best_path = self.backtrack_best(min_probs=[best_prob])[0]
        if len(best_path.entries) == 0:
            msg = f"No best path found for probability {best_prob}"
            raise ValueError(msg)
        x = np.arange(1, self.nstages + 1)
        for ax, info in zip(fig.get_axes(), [None, "complexity", "complexity_cumul",
                                             "success_h1_cumul"]):
            for line in list(ax.get_lines()):
                label = line.get_label()
                if label == "Best path" or (isinstance(label, str)
                                            and label.startswith("Best: P(H1)")):
                    line.remove()
            if info is None:
                ax.plot(x, best_path.thresholds**2, "r-", label=f"Best: P(H1) = {best_prob:.2f}")
            else:
                ax.plot(x, best_path.get_info(info), "r-", label="Best path")
            ax.legend(loc="upper right")
        fig.canvas.draw_idle()
        return fig, best_path

    def plot_slice(self, stage: int, attribute: str = "success_h1_cumul", fmt: str = ".3f",
                   cmap: str = "viridis", figsize: tuple[float, float] = (12, 8),
                   annot_size: float = 8):
        """Heat map of one stage's table (``schemes.py:468-517``; drawn with matplotlib,
        empty cells blank)."""
        from matplotlib import pyplot as plt

        if not 0 <= stage < self.nstages:
            msg = f"Stage must be between 0 and {self.nstages - 1}, got {stage}"
            raise ValueError(msg)
        if self.states.dtype.names is None or attribute not in self.states.dtype.names:
            msg = f"Attribute must be one of {self.states.dtype.names}, got {attribute}"
            raise ValueError(msg)
        vals = self.states[stage][attribute].astype(float)
        vals[self.states[stage]["is_empty"]] = np.nan
        fig, ax = plt.subplots(figsize=figsize)
        im = ax.imshow(vals, origin="lower", aspect="auto", cmap=cmap)
        for (i, j), v in np.ndenumerate(vals):
            if np.isfinite(v):
                ax.text(j, i, format(v, fmt), ha="center", va="center", fontsize=annot_size)
        ax.set_yticks(range(self.nthresholds), [f"{t:.2f}" for t in self.thresholds])
        ax.set_xlabel("Success Prob bins")
        ax.set_ylabel("Thresholds")
        title = attribute.replace("_", " ").title()
        ax.set_title(f"Stage {stage}: {title}")
        cbar = fig.colorbar(im, ax=ax, shrink=0.8, pad=0.02)
        cbar.ax.set_ylabel(title)
        fig.tight_layout()
        return fig

    def plot_pareto(self, nparam_vol: float = 1):
        """Cost-efficiency frontier ``L = C / P(H1)`` against ``P(H1)``
        (``schemes.py:562-625``)."""
        from matplotlib import pyplot as plt

        frontier = self._frontier_by_bin("cost")
        if len(frontier) == 0:
            msg = "No non-empty final states to plot"
            raise ValueError(msg)
        prob = frontier["success_h1_cumul"].astype(float)
        cost = frontier["cost"].astype(float)
        order = np.argsort(prob)
        prob, cost = prob[order], cost[order]
        fig, ax = plt.subplots(figsize=(7, 6), constrained_layout=True)
        cloud = self._final_cloud()
        cprob = cloud["success_h1_cumul"].astype(float)
        ccost = cloud["cost"].astype(float) * nparam_vol
        keep = np.isfinite(cprob) & np.isfinite(ccost) & (cprob > 0) & (ccost > 0)
        ax.scatter(cprob[keep], ccost[keep], color="0.82", s=14, alpha=0.5, edgecolor="none",
                   zorder=1, rasterized=True, label="All schemes")
        ax.plot(prob, cost, c="#0072B2", lw=1.5, zorder=2, label="DP frontier")
        ax.scatter(prob, cost, facecolors="white", edgecolors="#0072B2", s=45, linewidth=0.4,
                   zorder=3)
        ax.set_xlabel(r"P(H$_{1}$)")
        ax.set_ylabel(r"Ensemble Cumulative cost $L = C_{\mathrm{total}}\,/\,P_d$")
        ax.set_xscale("log")
        ax.set_yscale("log")
        ax.set_title("Cost-efficiency frontier")
        ax.legend(loc="upper left", frameon=False)
        return fig

    def save(self, filename: str) -> None:
        np.savez(filename, thresholds=self.thresholds, probs=self.probs, states=self.states)

    @classmethod
    def from_file(cls, filename: str) -> DynamicThresholdSchemeAnalyser:
        """Load a scheme written by ``DynamicThresholdScheme.save``."""
        with np.load(filename, allow_pickle=False) as f:
            return cls(f["states"], f["thresholds"], f["probs"], f["branching_pattern"],
                       f["guess_path"], float(f["attrs/beam_width"]))
