"""Readers (and writers, for tests and tools) of the two time-series file formats that
``TimeSeries.from_tim`` takes (``src/pyloki/io/timeseries.py:197-235``, through sigpyproc's
``TimeSeries.from_tim`` / ``from_dat``):

- SIGPROC ``.tim``: a header of keyword records, each string stored as a little-endian
  int32 length and its characters, from ``HEADER_START`` to ``HEADER_END``; after a
  keyword comes its value (int32, float64, one byte, or a string, by keyword), then the
  samples (``nbits`` 32: float32; 8: uint8, or int8 when ``signed``; 16: uint16);
- PRESTO ``.dat`` + ``.inf``: raw float32 samples and a text description whose
  ``Width of each time series bin (sec)`` line gives the sampling time.

These are file formats, read on the host; the conditioning that follows runs on the GPU.
"""

from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

_INT_KEYS = {"telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric",
             "nbits", "nsamples", "nchans", "nifs", "nbeams", "ibeam"}
_DBL_KEYS = {"tstart", "tsamp", "fch1", "foff", "refdm", "az_start", "za_start", "src_raj",
             "src_dej", "period", "refrf"}
_STR_KEYS = {"rawdatafile", "source_name"}
_BYTE_KEYS = {"signed"}


def _read_string(buf: bytes, pos: int) -> tuple[str, int]:
    (n,) = struct.unpack_from("<i", buf, pos)
    if n < 0 or n > 4096 or pos + 4 + n > len(buf):
        msg = "not a SIGPROC header (bad keyword length)"
        raise ValueError(msg)
    return buf[pos + 4:pos + 4 + n].decode("ascii", errors="replace"), pos + 4 + n


def parse_sigproc_header(buf: bytes) -> tuple[dict, int]:
    """The header's keywords and the byte offset of the first sample."""
    key, pos = _read_string(buf, 0)
    if key != "HEADER_START":
        msg = "not a SIGPROC file (no HEADER_START)"
        raise ValueError(msg)
    hdr: dict = {}
    while True:
        key, pos = _read_string(buf, pos)
        if key == "HEADER_END":
            return hdr, pos
        if key in _INT_KEYS:
            (hdr[key],) = struct.unpack_from("<i", buf, pos)
            pos += 4
        elif key in _DBL_KEYS:
            (hdr[key],) = struct.unpack_from("<d", buf, pos)
            pos += 8
        elif key in _BYTE_KEYS:
            hdr[key] = buf[pos]
            pos += 1
        elif key in _STR_KEYS:
            hdr[key], pos = _read_string(buf, pos)
        else:
            msg = f"unknown SIGPROC header keyword {key!r}"
            raise ValueError(msg)


def read_sigproc_tim(path) -> tuple[np.ndarray, dict]:
    """Samples (float32) and header of a SIGPROC time series (one channel, one IF)."""
    buf = Path(path).read_bytes()
    hdr, off = parse_sigproc_header(buf[:65536])
    if hdr.get("nchans", 1) != 1 or hdr.get("nifs", 1) != 1:
        msg = "a .tim file holds one channel and one IF"
        raise ValueError(msg)
    nbits = hdr.get("nbits", 32)
    if nbits == 32:
        dt = np.dtype("<f4")
    elif nbits == 16:
        dt = np.dtype("<u2")
    elif nbits == 8:
        dt = np.dtype(np.int8 if hdr.get("signed", 0) else np.uint8)
    else:
        msg = f"unsupported nbits {nbits}"
        raise ValueError(msg)
    data = np.frombuffer(buf, dtype=dt, offset=off)
    if "tsamp" not in hdr or hdr["tsamp"] <= 0:
        msg = "the header has no positive tsamp"
        raise ValueError(msg)
    return data.astype(np.float32), hdr


def write_sigproc_tim(path, data: np.ndarray, tsamp: float, **header) -> None:
    """A float32 SIGPROC time series with the given header keywords."""

    def s(x: str) -> bytes:
        b = x.encode("ascii")
        return struct.pack("<i", len(b)) + b

    out = [s("HEADER_START")]
    hdr = {"nbits": 32, "nchans": 1, "nifs": 1, "data_type": 2, "tsamp": float(tsamp)}
    hdr.update(header)
    for k, v in hdr.items():
        out.append(s(k))
        if k in _INT_KEYS:
            out.append(struct.pack("<i", int(v)))
        elif k in _DBL_KEYS:
            out.append(struct.pack("<d", float(v)))
        elif k in _BYTE_KEYS:
            out.append(bytes([int(v)]))
        elif k in _STR_KEYS:
            out.append(s(str(v)))
        else:
            msg = f"unknown SIGPROC header keyword {k!r}"
            raise ValueError(msg)
    out.append(s("HEADER_END"))
    out.append(np.ascontiguousarray(data, dtype="<f4").tobytes())
    Path(path).write_bytes(b"".join(out))


def read_presto_inf(path) -> dict:
    """The ``key = value`` lines of a PRESTO .inf file; ``tsamp`` from the bin width."""
    info: dict = {}
    for line in Path(path).read_text().splitlines():
        if "=" not in line:
            continue
        k, v = line.split("=", 1)
        info[k.strip()] = v.strip()
    width = next((v for k, v in info.items() if k.startswith("Width of each time series bin")),
                 None)
    if width is None:
        msg = f"{path}: no 'Width of each time series bin' line"
        raise ValueError(msg)
    info["tsamp"] = float(width)
    return info


def read_presto_dat(path) -> tuple[np.ndarray, dict]:
    """Samples (float32) of a PRESTO .dat file and its .inf description."""
    path = Path(path)
    inf = read_presto_inf(path.with_suffix(".inf"))
    return np.fromfile(path, dtype="<f4").astype(np.float32), inf


def write_presto_dat(path, data: np.ndarray, tsamp: float, source: str = "fake") -> None:
    path = Path(path)
    np.ascontiguousarray(data, dtype="<f4").tofile(path)
    path.with_suffix(".inf").write_text(
        f" Data file name without suffix          =  {path.stem}\n"
        f" Object being observed                  =  {source}\n"
        f" Number of bins in the time series      =  {len(data)}\n"
        f" Width of each time series bin (sec)    =  {tsamp!r}\n")
