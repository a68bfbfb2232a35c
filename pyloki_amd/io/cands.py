"""EP search results on disk: the pruning statistics and the results file.

Mirrors ``pyloki.io.cands`` (``src/pyloki/io/cands.py``): ``PruneStats`` per level,
``PruneStatsCollection`` with the same array layouts (``to_array``, ``:203-239``), and
``PruneResultWriter`` writing the same datasets (``:277-354``): file attributes
``pruning_version``, ``param_names``, ``nsegments``, ``max_sugg``; the ``threshold_scheme``;
and per run ``runs/<ref>_<task>/{scheme, param_sets, scores_ep, scores, level_stats,
timer_stats}``.

The reference stores them in HDF5. h5py is not part of this package's dependencies, so
the file is an ``.npz`` archive (a zip of ``.npy`` members, readable with
``numpy.load(..., allow_pickle=False)``) whose member names are the HDF5 paths:
``attrs/pruning_version``, ``threshold_scheme``, ``runs/003_00/param_sets``, ... Structured
arrays (``level_stats``, ``timer_stats``) keep their field names.
"""

from __future__ import annotations

import io
import zipfile
from dataclasses import dataclass, field
from pathlib import Path
from typing import ClassVar

import numpy as np

PRUNING_VERSION = "0.8.0+mi355x"

LEVEL_DTYPE = np.dtype([
    ("level", np.int32), ("seg_idx", np.int32), ("threshold", np.float32),
    ("score_min", np.float32), ("score_max", np.float32), ("n_branches", np.int32),
    ("n_leaves", np.int32), ("n_leaves_phy", np.int32), ("n_leaves_surv", np.int32)])


@dataclass
class PruneStats:
    """Pruning statistics of one level (``io/cands.py:16-101``)."""

    level: int
    seg_idx: int
    threshold: float
    score_min: float = 0.0
    score_max: float = 0.0
    n_branches: int = 1
    n_leaves: int = 1
    n_leaves_phy: int = 1
    n_leaves_surv: int = 1

    @property
    def lb_leaves(self) -> float:
        return float(np.round(np.log2(max(self.n_leaves, 1)), 2))

    @property
    def lb_leaves_phys(self) -> float:
        return float(np.round(np.log2(max(self.n_leaves_phy, 1)), 2))

    @property
    def branch_frac(self) -> float:
        return 0.0 if self.n_branches == 0 else float(np.round(self.n_leaves_phy / self.n_branches, 2))

    @property
    def phys_frac(self) -> float:
        return 0.0 if self.n_leaves == 0 else float(np.round(self.n_leaves_phy / self.n_leaves, 2))

    @property
    def surv_frac(self) -> float:
        return 0.0 if self.n_leaves_phy == 0 else float(np.round(self.n_leaves_surv / self.n_leaves_phy, 2))

    def update(self, stats_dict: dict[str, float]) -> None:
        for key, value in stats_dict.items():
            setattr(self, key, value)

    def get_summary(self) -> str:
        return (f"Prune level: {self.level:3d}, seg_idx: {self.seg_idx:3d}, "
                f"leaves: {self.lb_leaves:5.2f}, leaves_phys: {self.lb_leaves_phys:5.2f}, "
                f"branch_frac: {self.branch_frac:5.2f},"
                f"score thresh: {self.threshold:5.2f}, max: {self.score_max:5.2f}, "
                f"min: {self.score_min:5.2f}, P(surv): {self.surv_frac:4.2f}\n")


@dataclass(eq=False)
class PruneStatsCollection:
    """Statistics of every level of a run and its phase timers (``io/cands.py:104-274``).

    Also a sequence of its levels' ``PruneStats`` (indexing, slicing, ``len``, iteration,
    ``append``); two collections compare equal when their levels do."""

    stats_list: list[PruneStats] = field(default_factory=list)
    timers: dict[str, float] = field(default_factory=dict)

    def __getitem__(self, i):
        return self.stats_list[i]

    def __len__(self) -> int:
        return len(self.stats_list)

    def __iter__(self):
        return iter(self.stats_list)

    def __eq__(self, other: object) -> bool:
        if isinstance(other, PruneStatsCollection):
            return self.stats_list == other.stats_list
        if isinstance(other, list):
            return self.stats_list == other
        return NotImplemented

    def append(self, stats: PruneStats) -> None:
        self.update_stats(stats)

    TIMER_NAMES: ClassVar[list[str]] = ["branch", "validate", "resolve", "shift_add", "score",
                                        "threshold", "transform", "batch_add"]

    def __post_init__(self) -> None:
        for name in self.TIMER_NAMES:
            self.timers.setdefault(name, 0.0)

    @property
    def nstages(self) -> int:
        return len(self.stats_list)

    def update_stats(self, stats: PruneStats, timer_arr: np.ndarray | None = None) -> None:
        self.stats_list.append(stats)
        if timer_arr is not None:
            if len(timer_arr) != len(self.TIMER_NAMES):
                msg = f"Invalid timer array length: {len(timer_arr)}"
                raise ValueError(msg)
            for i, name in enumerate(self.TIMER_NAMES):
                self.timers[name] += float(timer_arr[i])

    def get_stats(self, level: int) -> PruneStats | None:
        for st in self.stats_list:
            if st.level == level:
                return st
        return None

    def get_all_summaries(self) -> str:
        return "".join(st.get_summary() for st in sorted(self.stats_list, key=lambda x: x.level))

    def get_stats_summary(self) -> str:
        last = self.stats_list[-1]
        return f"Score: {last.score_max:.2f}, Leaves: {last.lb_leaves:.2f}"

    def get_timer_summary(self) -> str:
        total = sum(self.timers.values())
        lines = [f"Timing breakdown: {total:.2f}s"]
        for name, value in sorted(self.timers.items(), key=lambda x: x[1], reverse=True):
            pct = 100.0 * value / total if total > 0 else 0.0
            lines.append(f"  {name:10s}: {pct:6.1f}%")
        return "\n".join(lines)

    def to_array(self) -> tuple[np.ndarray, np.ndarray]:
        """``level_stats`` (one row per level, ``LEVEL_DTYPE``) and ``timer_stats`` (one row,
        one float32 field per timer), as the reference stores them."""
        level_stats = np.zeros(len(self.stats_list), dtype=LEVEL_DTYPE)
        for i, st in enumerate(self.stats_list):
            level_stats[i] = (st.level, st.seg_idx, st.threshold, st.score_min, st.score_max,
                              st.n_branches, st.n_leaves, st.n_leaves_phy, st.n_leaves_surv)
        timer_stats = np.zeros(1, dtype=np.dtype([(n, np.float32) for n in self.TIMER_NAMES]))
        for name in self.TIMER_NAMES:
            timer_stats[0][name] = self.timers[name]
        return level_stats, timer_stats

    @classmethod
    def from_arrays(cls, level_stats: np.ndarray, timer_stats: np.ndarray
                    ) -> PruneStatsCollection:
        out = cls()
        for r in level_stats:
            out.stats_list.append(PruneStats(
                level=int(r["level"]), seg_idx=int(r["seg_idx"]), threshold=float(r["threshold"]),
                score_min=float(r["score_min"]), score_max=float(r["score_max"]),
                n_branches=int(r["n_branches"]), n_leaves=int(r["n_leaves"]),
                n_leaves_phy=int(r["n_leaves_phy"]), n_leaves_surv=int(r["n_leaves_surv"])))
        for name in cls.TIMER_NAMES:
            out.timers[name] = float(timer_stats[0][name])
        return out


def _npy_bytes(arr: np.ndarray) -> bytes:
    buf = io.BytesIO()
    np.save(buf, np.asarray(arr), allow_pickle=False)
    return buf.getvalue()


class PruneResultWriter:
    """Writes an EP results file (``io/cands.py:277-354``), ``mode`` "w" (new file) or "a"
    (append runs to an existing one)."""

    def __init__(self, filename: str | Path, mode: str = "w") -> None:
        if mode not in ("w", "a"):
            msg = f"mode must be 'w' or 'a', got {mode!r}"
            raise ValueError(msg)
        self.filename = Path(filename)
        self.mode = mode
        self.file: zipfile.ZipFile | None = None

    def __enter__(self) -> PruneResultWriter:
        mode = "a" if self.mode == "a" and self.filename.exists() else "w"
        self.file = zipfile.ZipFile(self.filename, mode, compression=zipfile.ZIP_DEFLATED)
        return self

    def __exit__(self, exc_type, exc_val, exc_tb) -> None:  # noqa: ANN001
        if self.file is not None:
            self.file.close()
            self.file = None

    def _put(self, name: str, arr: np.ndarray) -> None:
        if self.file is None:
            msg = "File not open."
            raise ValueError(msg)
        self.file.writestr(f"{name}.npy", _npy_bytes(arr))

    def write_metadata(self, param_names: list[str], nsegments: int, max_sugg: int,
                       threshold_scheme: np.ndarray) -> None:
        self._put("attrs/pruning_version", np.array(PRUNING_VERSION))
        self._put("attrs/param_names", np.array(list(param_names)))
        self._put("attrs/nsegments", np.array(int(nsegments)))
        self._put("attrs/max_sugg", np.array(int(max_sugg)))
        self._put("threshold_scheme", np.asarray(threshold_scheme))

    def write_run_results(self, run_name: str, scheme: np.ndarray, leaves_report: np.ndarray,
                          scores: np.ndarray, scores_ep: np.ndarray,
                          pstats: PruneStatsCollection) -> None:
        if self.file is None:
            msg = "File not open."
            raise ValueError(msg)
        if any(n.startswith(f"runs/{run_name}/") for n in self.file.namelist()):
            msg = f"Run name {run_name} already exists in the file."
            raise ValueError(msg)
        level_stats, timer_stats = pstats.to_array()
        for name, data in (("scheme", scheme), ("param_sets", leaves_report),
                           ("scores_ep", scores_ep), ("scores", scores),
                           ("level_stats", level_stats), ("timer_stats", timer_stats)):
            self._put(f"runs/{run_name}/{name}", data)


def read_results(filename: str | Path) -> dict:
    """A results file as ``{"attrs": {...}, "threshold_scheme": ..., "runs": {name: {...}}}``
    (runs in name order, as h5py iterates a group)."""
    path = Path(filename)
    with np.load(path, allow_pickle=False) as z:
        keys = list(z.keys())
        if "attrs/pruning_version" not in keys:
            msg = "Not a valid pruning results file"
            raise ValueError(msg)
        attrs = {k.split("/", 1)[1]: z[k] for k in keys if k.startswith("attrs/")}
        attrs = {k: (v.tolist() if v.ndim else v.item()) for k, v in attrs.items()}
        runs: dict[str, dict] = {}
        for k in keys:
            if k.startswith("runs/"):
                _, run, name = k.split("/", 2)
                runs.setdefault(run, {})[name] = z[k]
        thr = z["threshold_scheme"] if "threshold_scheme" in keys else None
    return {"attrs": attrs, "threshold_scheme": thr,
            "runs": {k: runs[k] for k in sorted(runs)}}


def mkdir_p(dir_path: str | Path, must_not_exist: list[str] | None = None) -> Path:
    """Create a directory (and parents); refuse if any of ``must_not_exist`` is already
    there (``utils/misc.py:42-67``)."""
    path = Path(dir_path)
    try:
        path.mkdir(parents=True, exist_ok=True)
    except OSError as e:
        msg = f"Failed to create directory {path}: {e}"
        raise OSError(msg) from e
    if must_not_exist:
        existing = [f for f in must_not_exist if (path / f).exists()]
        if existing:
            msg = (f"Files already exist in directory {path}: {existing}\n"
                   f"Provide a different directory or remove the files.")
            raise FileExistsError(msg)
    return path
