"""``pyloki`` import names for the MI355X search path.

The reference's notebooks import ``pyloki.config``, ``pyloki.search``, ``pyloki.ffa``,
``pyloki.prune``, ``pyloki.periodogram``, ``pyloki.simulation.pulse``,
``pyloki.detection`` ...; this package makes those lines run unchanged by binding each
``pyloki.<module>`` to the same module of ``pyloki_amd`` (one module object under two
names, so classes and functions are the same objects).
"""

from __future__ import annotations

import importlib
import pkgutil
import sys

import pyloki_amd

__all__ = ["pyloki_amd"]


def _alias(pkg, prefix: str) -> None:
    for info in pkgutil.walk_packages(pkg.__path__, pkg.__name__ + "."):
        if info.name.startswith("pyloki_amd.csrc") or info.name.endswith("libpylokihip"):
            continue  # the HIP library is loaded through ctypes, not imported
        mod = importlib.import_module(info.name)
        sys.modules.setdefault(prefix + info.name[len("pyloki_amd"):], mod)


_alias(pyloki_amd, "pyloki")


def __getattr__(name: str):
    mod = sys.modules.get(f"pyloki.{name}")
    if mod is None:
        raise AttributeError(name)
    return mod
