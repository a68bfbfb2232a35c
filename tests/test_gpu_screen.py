"""The boxcar S/N threshold screen on the matrix cores (pyl_score_boxcar_screen) against
the exact scorer (pyl_score_boxcar, mode PYL_SCORE_MAX, itself bit-exact to the oracle):
the profiles at or above the threshold, and their scores, are the exact scorer's bit for
bit; every other output is an upper bound of the exact score below the threshold."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pyloki_amd import _lib, kernels
from pyloki_amd.detection.scoring import generate_box_width_trials
from pyloki_amd.simulation.pulse import generate_folded_profile

pytestmark = pytest.mark.gpu


def _folds(n, B, seed):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal((n, B)) * 4.0
    v = rng.uniform(12.0, 20.0, (n, B))
    k = rng.random(n) < 0.03  # pulsed profiles over a range of S/N
    prof = generate_folded_profile(B, 0.1)
    amp = rng.uniform(0.0, 12.0, n) * 4.0
    sh = rng.integers(0, B, n)
    for i in np.nonzero(k)[0]:
        e[i] += amp[i] * np.roll(prof, sh[i])
    e[:5] = 0.0  # all-zero profiles
    e[5] = 1e-3
    return np.stack([e, v], axis=1).astype(np.float32)


@pytest.mark.parametrize("B", [32, 64])
@pytest.mark.parametrize("thr", [3.5, 6.0, 9.0])
def test_screen_selects_the_exact_survivors(gpu_device, B, thr):
    print(f"[case] screen B={B} thr={thr}", flush=True)
    widths = generate_box_width_trials(B, ducy_max=0.3, wtsp=1.5)
    assert len(widths) <= 16
    f = torch.from_numpy(_folds(100_003, B, B)).cuda()
    exact = kernels.score_boxcar(f, widths, mode=_lib.PYL_SCORE_MAX).cpu().numpy()
    out, nres = kernels.score_boxcar_screen(f, widths, thr)
    out = out.cpu().numpy()
    sel = exact >= thr
    assert np.array_equal(out >= thr, sel)
    assert np.array_equal(out[sel], exact[sel])
    assert np.all(out[~sel] >= exact[~sel]) and np.all(out[~sel] < thr)
    assert nres >= int(sel.sum())
    print(f"[result] survivors {int(sel.sum())}, rescored {nres} of {len(out)}", flush=True)
