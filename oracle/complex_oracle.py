"""CPU oracle of the complex (Fourier-domain) FFA -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.
It restates, in NumPy, the reference's ``use_fourier=True`` FFA path:

- ``cfold_basis`` / ``brutefold_complex``: ``core/fold.py:395-481``
  (``brutefold_start_complex_opt``) via ``fold.py:572-614`` (``ffa_taylor_init_complex``):
  for harmonics m = 0..B/2 of every trial frequency f, the FP32 basis
  ``cos/sin(f32(f*m) * (t * f32(-2 pi tsamp)))`` (angle in FP64, stored in FP32), the
  segment-by-basis product, then the ``e^{+2 pi i f32(f*m) f32(2 pi t_ref)}`` correction in
  complex64. The reference's product is an FP32 BLAS GEMM whose summation order is the
  library's; here it is summed in FP64 and rounded once to FP32.
- ``shift_add_complex``: ``core/common.py:344-372``: the complex64 recurrence
  ``phase *= delta`` with ``delta = (cos, sin)(f32(-2 pi f32(shift) / B))`` in FP32.
- ``unify_fold_complex``: ``dyn_pffa.py:435-489`` with the same resolve table as the real
  path (``oracle.unify_fold`` computes it).
- ``fold_norm_complex``: ``ffa.py:400-411``: ``irfft`` (NumPy, FP32 input), then e / sqrt(v).

Pinning (tests/test_complex_oracle.py): the brute fold against a direct FP64 DFT of the
samples (and its DC term against the plain sums); the complex shift-add at integer shifts
against the real shift-add through the circular-shift theorem, and its phasors against
the closed form e^{-2 pi i k s / B}; the last level's profile against the real-domain FFA
of the same data (same peak trial, S/N within the interpolation difference).
Parity unpinned: the reference's BLAS summation order and numba fastmath contractions;
the GPU path is compared with this oracle within a stated FP32 tolerance.
"""

from __future__ import annotations

import numpy as np

from oracle import oracle as orc


def cfold_basis(freqs: np.ndarray, seg_len: int, nbins: int, tsamp: float):
    """FP32 basis [n_comb, seg_len] (real, imag) and the FP32 comb frequencies."""
    nb_f = nbins // 2 + 1
    target = (np.asarray(freqs, dtype=np.float64)[:, None] * np.arange(nb_f)[None, :])
    target = target.astype(np.float32).reshape(-1)
    time_factor = np.float32(-2.0 * np.pi * tsamp)
    time_val = np.arange(seg_len, dtype=np.float64) * np.float64(time_factor)
    ang = target.astype(np.float64)[:, None] * time_val[None, :]
    return np.cos(ang).astype(np.float32), np.sin(ang).astype(np.float32), target


def brutefold_complex(ts_e, ts_v, freqs, seg_len: int, nbins: int, tsamp: float,
                      t_ref: float = 0.0) -> np.ndarray:
    """complex64 [nseg, nfreq, 2, nbins // 2 + 1]."""
    nfreq = len(freqs)
    nb_f = nbins // 2 + 1
    n = len(ts_e)
    nseg = -(-n // seg_len)
    x = np.asarray(ts_e, dtype=np.float32).reshape(nseg, seg_len).astype(np.float64)
    y = np.asarray(ts_v, dtype=np.float32).reshape(nseg, seg_len).astype(np.float64)
    wr, wi, target = cfold_basis(freqs, seg_len, nbins, tsamp)
    wr64, wi64 = wr.astype(np.float64), wi.astype(np.float64)
    out = np.empty((nseg, nfreq, 2, nb_f), dtype=np.complex64)
    for comp, data in ((0, x), (1, y)):
        re = (data @ wr64.T).astype(np.float32).reshape(nseg, nfreq, nb_f)
        im = (data @ wi64.T).astype(np.float32).reshape(nseg, nfreq, nb_f)
        out[:, :, comp, :] = re + 1j * im
    corr = np.float32(2.0 * np.pi * t_ref)
    phasor = np.exp(1j * np.float64(corr) * target.astype(np.float64)).astype(np.complex64)
    re, im = _cmul(out, phasor.reshape(1, nfreq, 1, nb_f))
    return (re + 1j * im).astype(np.complex64)


def _phasors(shift: np.ndarray, nb_f: int) -> np.ndarray:
    """[n, nb_f] complex64 powers of delta by the reference's sequential recurrence."""
    nbins = (nb_f - 1) * 2
    s32 = np.asarray(shift, dtype=np.float64).astype(np.float32)
    step = (-2.0 * np.pi * s32.astype(np.float64) / nbins).astype(np.float32)
    dre, dim = np.cos(step), np.sin(step)          # FP32 cos / sin
    out = np.empty((len(s32), nb_f), dtype=np.complex64)
    pre = np.ones(len(s32), dtype=np.float32)
    pim = np.zeros(len(s32), dtype=np.float32)
    for k in range(nb_f):
        out[:, k] = pre + 1j * pim
        nre = pre * dre - pim * dim
        nim = pre * dim + pim * dre
        pre, pim = nre.astype(np.float32), nim.astype(np.float32)
    return out


def _cmul(a: np.ndarray, b: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """complex64 product in FP32 (ac - bd, ad + bc), no contraction."""
    ar, ai = a.real.astype(np.float32), a.imag.astype(np.float32)
    br, bi = b.real.astype(np.float32), b.imag.astype(np.float32)
    return (ar * br - ai * bi).astype(np.float32), (ar * bi + ai * br).astype(np.float32)


def shift_add_complex(tail: np.ndarray, head: np.ndarray, sh_t: np.ndarray,
                      sh_h: np.ndarray) -> np.ndarray:
    """Rows [n, 2, nb_f] complex64: tail * phi_t^k + head * phi_h^k."""
    nb_f = tail.shape[-1]
    pt = _phasors(sh_t, nb_f)[:, None, :]
    ph = _phasors(sh_h, nb_f)[:, None, :]
    tr, ti = _cmul(tail, pt)
    hr, hi = _cmul(head, ph)
    return ((tr + hr).astype(np.float32) + 1j * (ti + hi).astype(np.float32)).astype(
        np.complex64)


def unify_fold_complex(fold_in: np.ndarray, counts_prev, param_arr_cur, limits, level: int,
                       tseg_brute: float, nbins: int) -> np.ndarray:
    """One complex level: the real path's resolve table, complex shift-add."""
    nseg = fold_in.shape[0]
    nprev = int(np.prod(counts_prev))
    nb_f = fold_in.shape[-1]
    _, rows, phases = orc.unify_fold(np.zeros((2, nprev, 2, nbins), dtype=np.float32),
                                     counts_prev, param_arr_cur, limits, level, tseg_brute,
                                     nbins, return_table=True)
    fin = fold_in.reshape(nseg, nprev, 2, nb_f)
    out = np.empty((nseg // 2, len(rows), 2, nb_f), dtype=np.complex64)
    for q in range(nseg // 2):
        out[q] = shift_add_complex(fin[2 * q][rows[:, 0]], fin[2 * q + 1][rows[:, 1]],
                                   phases[:, 0], phases[:, 1])
    return out


def ffa_fold_complex(ts_e, ts_v, cfg, levels: int | None = None):
    """All levels of the complex FFA; returns (fold [nseg_last, ncoords, 2, nb_f],
    param_arr of the last level)."""
    dparams0 = cfg.get_dparams(0)
    pa0 = cfg.get_param_arr(dparams0)
    t_ref = 0.0 if cfg.nparams == 1 else cfg.bseg_brute * cfg.tsamp / 2
    fold = brutefold_complex(ts_e, ts_v, pa0[-1], cfg.bseg_brute, cfg.nbins, cfg.tsamp, t_ref)
    nseg = fold.shape[0]
    fold = fold.reshape(nseg, -1, 2, fold.shape[-1])
    counts_prev = cfg.get_param_grid_count(0)
    pa = pa0
    nl = cfg.niters_ffa if levels is None else levels
    for level in range(1, nl + 1):
        pa = cfg.get_param_arr(cfg.get_dparams(level))
        fold = unify_fold_complex(fold, counts_prev, pa, cfg.param_limits, level,
                                  cfg.tseg_brute, cfg.nbins)
        counts_prev = np.array([len(a) for a in pa])
    return fold, pa


def fold_norm_complex(fold: np.ndarray, nbins: int) -> np.ndarray:
    """irfft to nbins real bins (FP32), then e / sqrt(v)."""
    t = np.fft.irfft(fold, n=nbins).astype(np.float32)
    return (t[..., 0, :] / np.sqrt(t[..., 1, :])).astype(np.float32)


def ffa_search_snr_complex(ts_e, ts_v, cfg) -> np.ndarray:
    fold, _ = ffa_fold_complex(ts_e, ts_v, cfg)
    norm = fold_norm_complex(fold[0], cfg.nbins)
    return orc.boxcar_snr(norm, cfg.score_widths)


# ---------------------------------------------------------------- EP with complex folds
# common.py:403-429 shift_add_complex_batch (the segment row rotated, the parent row not),
# scoring.py:285-292 snr_score_batch_func_complex, common.py:463-490
# shift_add_ascend_complex_batch.

def ep_phasors(shift: np.ndarray, nb_f: int) -> np.ndarray:
    """[n, nb_f] complex64 phasors of the EP complex shift-add: the angle is
    f32(-2 * f32(pi) * f32(shift) / B) (common.py:419-421), then the recurrence."""
    nbins = (nb_f - 1) * 2
    s32 = np.asarray(shift, dtype=np.float64).astype(np.float32)
    ang = ((-2.0 * float(np.float32(np.pi))) * s32.astype(np.float64) / nbins).astype(
        np.float32)
    dre, dim = np.cos(ang), np.sin(ang)
    out = np.empty((len(s32), nb_f), dtype=np.complex64)
    pre = np.ones(len(s32), dtype=np.float32)
    pim = np.zeros(len(s32), dtype=np.float32)
    for k in range(nb_f):
        out[:, k] = pre + 1j * pim
        nre = pre * dre - pim * dim
        nim = pre * dim + pim * dre
        pre, pim = nre.astype(np.float32), nim.astype(np.float32)
    return out


def cshift_add_batch(seg: np.ndarray, shifts: np.ndarray, folds: np.ndarray,
                     parents: np.ndarray) -> np.ndarray:
    """res = seg * phi^k + parent fold (complex64, [n, 2, nb_f])."""
    ph = ep_phasors(shifts, seg.shape[-1])[:, None, :]
    sr, si = _cmul(seg, ph)
    par = folds[parents]
    re = (sr + par.real.astype(np.float32)).astype(np.float32)
    im = (si + par.imag.astype(np.float32)).astype(np.float32)
    return (re + 1j * im).astype(np.complex64)


def cscore_batch(folds: np.ndarray, widths: np.ndarray) -> np.ndarray:
    """irfft (FP32) then the real scorer's max over widths."""
    return orc.snr_score_batch(np.fft.irfft(folds).astype(np.float32), widths)


def cascend_accumulate(rows: list[np.ndarray], shifts: list[np.ndarray]) -> np.ndarray:
    """sum over segments (in order, from zero) of row * phi^k, complex64."""
    n, _, nb_f = rows[0].shape
    re = np.zeros((n, 2, nb_f), dtype=np.float32)
    im = np.zeros((n, 2, nb_f), dtype=np.float32)
    for r, s in zip(rows, shifts):
        ph = ep_phasors(s, nb_f)[:, None, :]
        pr, pi = _cmul(r, ph)
        re = (re + pr).astype(np.float32)
        im = (im + pi).astype(np.float32)
    return (re + 1j * im).astype(np.complex64)
