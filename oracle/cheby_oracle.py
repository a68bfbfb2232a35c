"""CPU oracle of the EP search in the Chebyshev basis -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this.
It restates, in NumPy with element-wise IEEE operations (no BLAS, no contraction), the
reference's Chebyshev pruning plugin with a moving grid:

- connection coefficients and matrices: ``utils/maths.py:399-557``
  (``compute_connection_coefficient_s/_r``, ``compute_transformation_coefficient_c``,
  ``compute_connection_matrix_s/_r``, ``poly_chebyshev_transform_matrix``), with ``x**k``
  for an integer k by binary powering (``powi``), as numba lowers a float to an integer
  power.
- transforms: ``utils/transforms.py:435-737`` (``taylor_to_cheby[_full/_errors]``,
  ``cheby_to_taylor[_full]``, ``shift_cheby_full/_errors``, ``cheby_to_taylor_param_shift``).
  A row vector times a matrix is summed in ascending row order.
- ``cheby_seed``: ``core/chebyshev.py:14-69``; ``cheby_branch_batch``: ``:72-177`` (the
  parent moved to coord_cur first, then the padded split of every coefficient but alpha_0
  with the step ``eta/nbins * C / f0``, ``psr_utils.py:258-278``); ``cheby_resolve_batch``:
  ``:180-258``; ``cheby_transform_batch``: ``:352-368``; ``cheby_report_batch``: ``:371-398``.
- ``cheby_prune_level``: ``prune.py:165-332`` with the plugin ``dyn_poly_cheby.py:587-730``;
  ``cheby_ascend``: ``dyn_poly_cheby.py:733-776`` with ``chebyshev.py:261-349``.
- ``generate_bp_cheby``: ``chebyshev.py:458-546``. The reference leaves the alpha_0 column
  of its next-level steps uninitialised (``np.empty``); that column is never read by the
  count and no other column depends on it (C maps T_0 only to T_0), so it is 0 here.

Leaves hold ``[alpha_P, ..., alpha_1, alpha_0, (f0, flag)]`` with (value, cell width) per
coefficient, referenced to the current coordinate (t0, scale).

Pinning (tests/test_cheby_oracle.py): the reference's known answers
(``tests/test_maths.py:55-149``: S, R entries, the closed forms at t_s = 4.2, round trips,
polynomial evaluation); the domain-transfer identity ``T_n(p x + q) = sum_k C_nk T_k(x)``
and the matrix ``M1 inv(M2)`` of ``tests/mat_inv_tests.py``; the parameter shift against
numpy's Chebyshev-series derivatives; the branch against the exact-tiling invariant; a
level against a child-by-child brute force; the branching pattern against the counts of a
search; and the search's candidate against the Taylor-basis search of the same data.
Parity unpinned: the reference runs under numba ``fastmath`` and sums its matrix products
in BLAS order, so its last-ulp choices are not reproduced.
"""

from __future__ import annotations

import math

import numpy as np

from oracle import prune_oracle as po

C_VAL = po.C_VAL
FLOAT_EPSILON = po.FLOAT_EPSILON
powi = po.powi


# ------------------------------------------------------------------ coefficient tables
def nbinom(n: int, k: int) -> int:
    return math.comb(n, k) if 0 <= k <= n else 0


def conn_s(k: int, m: int) -> float:
    """S_{k,m}: x^k = sum_m S_{k,m} T_m(x)."""
    if k < 0 or m < 0 or m > k or (k - m) % 2 != 0:
        return 0.0
    n = (k - m) // 2
    d = 1 if m == 0 else 0
    return 2.0 ** (1 - k - d) * nbinom(k, n)


def conn_r(k: int, m: int) -> float:
    """R_{k,m}: T_k(x) = sum_m R_{k,m} x^m."""
    if k < 0 or m < 0 or m > k or (k - m) % 2 != 0:
        return 0.0
    if m == 0 and k % 2 == 0:
        return float((-1) ** (k // 2))
    n = (k - m) // 2
    r = (k + m) // 2
    return ((-1) ** n) * powi(2.0, m - 1) * (2.0 * k / (k + m)) * nbinom(r, n)


def trans_c(n: int, k: int, p: float, q: float) -> float:
    """C_{n,k}(p, q): T_n(p x + q) = sum_k C_{n,k} T_k(x)."""
    if k > n:
        return 0.0
    res = 0.0
    for m in range(k, n + 1):
        r_nm = conn_r(n, m)
        inner = 0.0
        for i in range(k, m + 1):
            inner += nbinom(m, i) * powi(p, i) * powi(q, m - i) * conn_s(i, k)
        res += r_nm * inner
    return res


def matrix_s(kmax: int) -> np.ndarray:
    return np.array([[conn_s(k, m) for m in range(kmax + 1)] for k in range(kmax + 1)])


def matrix_r(kmax: int) -> np.ndarray:
    return np.array([[conn_r(k, m) for m in range(kmax + 1)] for k in range(kmax + 1)])


def transform_matrix(poly_order: int, tc1: float, ts1: float, tc2: float, ts2: float,
                     descending: bool = True) -> np.ndarray:
    """C[i, j] = C_{n,k}(ts2/ts1, (tc2 - tc1)/ts1), indices in the storage order."""
    if ts1 <= 0 or ts2 <= 0:
        msg = "ts1 and ts2 must be positive."
        raise ValueError(msg)
    p = ts2 / ts1
    q = (tc2 - tc1) / ts1
    K = poly_order
    c = np.zeros((K + 1, K + 1))
    for i in range(K + 1):
        for j in range(K + 1):
            n, k = (K - i, K - j) if descending else (i, j)
            c[i, j] = trans_c(n, k, p, q)
    return c


def vecmat(v: np.ndarray, M: np.ndarray) -> np.ndarray:
    """Rows of v times M, out[:, j] = sum_i v[:, i] M[i, j] (ascending i)."""
    out = np.zeros((v.shape[0], M.shape[1]))
    for j in range(M.shape[1]):
        acc = np.zeros(v.shape[0])
        for i in range(M.shape[0]):
            acc = acc + v[:, i] * M[i, j]
        out[:, j] = acc
    return out


# ------------------------------------------------------------------------- transforms
def taylor_to_cheby_full(vec: np.ndarray, ts: float) -> np.ndarray:
    """[n, K+1, 2] Taylor (d_K..d_0, widths) -> Chebyshev (alpha_K..alpha_0, widths)."""
    if ts <= 0:
        msg = "t_s must be a positive."
        raise ValueError(msg)
    K = vec.shape[-2] - 1
    S = matrix_s(K)
    scale = np.array([powi(ts, k) / po.fact(k) for k in range(K + 1)])
    dv = vec[:, ::-1, 0] * scale
    de = vec[:, ::-1, 1] * scale
    av = vecmat(dv, S)
    ae = np.sqrt(vecmat(de**2, S**2))
    out = np.empty_like(vec)
    out[:, :, 0] = av[:, ::-1]
    out[:, :, 1] = ae[:, ::-1]
    return out


def taylor_to_cheby(vec: np.ndarray, ts: float) -> np.ndarray:
    """[n, K+1] Taylor values -> Chebyshev values (``transforms.py:435-463``)."""
    K = vec.shape[-1] - 1
    S = matrix_s(K)
    scale = np.array([powi(ts, k) / po.fact(k) for k in range(K + 1)])
    return vecmat(vec[:, ::-1] * scale, S)[:, ::-1]


def taylor_to_cheby_errors(err: np.ndarray, ts: float) -> np.ndarray:
    K = err.shape[-1] - 1
    S = matrix_s(K)
    scale = np.array([powi(ts, k) / po.fact(k) for k in range(K + 1)])
    return np.sqrt(vecmat((err[:, ::-1] * scale) ** 2, S**2))[:, ::-1]


def cheby_to_taylor(alpha: np.ndarray, ts: float) -> np.ndarray:
    """[n, K+1] Chebyshev values -> Taylor values: (a @ R) * k! / ts^k."""
    if ts <= 0:
        msg = "t_s must be a positive."
        raise ValueError(msg)
    K = alpha.shape[-1] - 1
    R = matrix_r(K)
    d = vecmat(alpha[:, ::-1], R)
    for m in range(K + 1):
        d[:, m] = (d[:, m] * po.fact(m)) / powi(ts, m)
    return d[:, ::-1]


def cheby_to_taylor_full(alpha: np.ndarray, ts: float) -> np.ndarray:
    K = alpha.shape[-2] - 1
    R = matrix_r(K)
    dv = vecmat(alpha[:, ::-1, 0], R)
    var = vecmat(alpha[:, ::-1, 1] ** 2, R**2)
    scale = np.array([po.fact(k) / powi(ts, k) for k in range(K + 1)])
    out = np.empty_like(alpha)
    out[:, :, 0] = (dv * scale)[:, ::-1]
    out[:, :, 1] = np.sqrt(var * scale**2)[:, ::-1]
    return out


def _tile_errors(err: np.ndarray, C: np.ndarray, tiling: str) -> np.ndarray:
    if tiling == "aggressive":
        return err * np.abs(np.diag(C))[None, :]
    if tiling == "conservative":
        return vecmat(err, np.abs(C))
    if tiling == "quadrature":
        return np.sqrt(vecmat(err**2, C**2))
    msg = f"Invalid tiling strategy: {tiling}"
    raise ValueError(msg)


def shift_cheby_full(alpha: np.ndarray, coord_next, coord_cur, tiling: str) -> np.ndarray:
    """[n, K+1, 2] from the domain coord_cur to coord_next (values and widths)."""
    K = alpha.shape[-2] - 1
    C = transform_matrix(K, coord_cur[0], coord_cur[1], coord_next[0], coord_next[1])
    out = np.empty_like(alpha)
    out[:, :, 0] = vecmat(alpha[:, :, 0], C)
    out[:, :, 1] = _tile_errors(alpha[:, :, 1], C, tiling)
    return out


def shift_cheby_errors(err: np.ndarray, coord_next, coord_cur, tiling: str) -> np.ndarray:
    K = err.shape[-1] - 1
    C = transform_matrix(K, coord_cur[0], coord_cur[1], coord_next[0], coord_next[1])
    return _tile_errors(err, C, tiling)


def cheby_param_shift(alpha: np.ndarray, t0: float, ts: float, t_eval: float) -> np.ndarray:
    """Taylor coefficients at t_eval of the series on (t0, ts)."""
    K = alpha.shape[-1] - 1
    C = transform_matrix(K, t0, ts, t_eval, ts)
    return cheby_to_taylor(vecmat(alpha, C), ts)


# ------------------------------------------------------------------------- the plugin
def cheby_seed(param_arr, dparams, poly_order: int, coord_init) -> np.ndarray:
    scale_init = coord_init[1]
    mesh = np.meshgrid(*param_arr, indexing="ij")
    cart = np.stack([m.reshape(-1) for m in mesh], axis=1)
    n = cart.shape[0]
    P = poly_order
    d = np.zeros((n, P + 1, 2))
    for k in range(P - 1):
        d[:, k, 0] = cart[:, k]
        d[:, k, 1] = dparams[k]
    f0 = cart[:, -1]
    d[:, -2, 1] = np.full(n, float(dparams[-1])) * (C_VAL / f0)
    leaves = np.zeros((n, P + 2, 2))
    leaves[:, :-1] = taylor_to_cheby_full(d, scale_init)
    leaves[:, -1, 0] = f0
    return leaves


def cheby_branch_batch(leaves, coord_cur, coord_prev, nbins: int, eta: float,
                       poly_order: int, branch_max: int, tiling: str):
    trans = shift_cheby_full(leaves[:, :-1], coord_cur, coord_prev, tiling)
    pcur = trans[:, :-1, 0]
    dcur = trans[:, :-1, 1]
    f0 = leaves[:, -1, 0]
    dphi = eta / nbins
    dnew = np.empty_like(dcur)
    for j in range(poly_order):
        dnew[:, j] = dphi * C_VAL / f0
    shift = np.abs(dcur - dnew) * (nbins * (f0 / C_VAL))[:, None]
    return po.odometer_children(pcur, dcur, dnew, shift, eta, branch_max, trans[:, -1, 0], f0,
                                leaves[:, -1, 1])


def cheby_resolve_batch(leaves, coord_add, coord_cur, coord_init, counts_init, limits,
                        nbins: int):
    t0_cur, scale_cur = coord_cur
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    d_add = cheby_param_shift(vec, t0_cur, scale_cur, coord_add[0])
    d_init = cheby_param_shift(vec, t0_cur, scale_cur, coord_init[0])
    accel = d_add[:, -3]
    vel = d_add[:, -2] - d_init[:, -2]
    freq = f0 * (1 - vel / C_VAL)
    delay = (d_add[:, -1] - d_init[:, -1]) / C_VAL
    phase = po.phase_idx_vec(coord_add[0] - coord_init[0], f0, nbins, delay)
    idx = po.nearest_2d(accel, freq, counts_init, limits)
    return idx, phase


def cheby_fixed_resolve_batch(leaves, coord_add, coord_cur_fixed, coord_init, counts_init,
                              limits, nbins: int):
    """poly_chebyshev_fixed_resolve_batch (chebyshev.py:261-300): the leaf's domain is
    (t_init, scale of the fixed coord); velocity and delay at t_add used whole."""
    scale = coord_cur_fixed[1]
    t0_init = coord_init[0]
    vec = leaves[:, :-1, 0]
    f0 = leaves[:, -1, 0]
    d_add = cheby_param_shift(vec, t0_init, scale, coord_add[0])
    freq = f0 * (1 - d_add[:, -2] / C_VAL)
    delay = d_add[:, -1] / C_VAL
    phase = po.phase_idx_vec(coord_add[0] - t0_init, f0, nbins, delay)
    return po.nearest_2d(d_add[:, -3], freq, counts_init, limits), phase


def cheby_transform_batch(leaves, coord_next, coord_cur, tiling: str):
    out = np.zeros_like(leaves)
    out[:, :-1] = shift_cheby_full(leaves[:, :-1], coord_next, coord_cur, tiling)
    out[:, -1] = leaves[:, -1]
    return out


def cheby_prune_level(tree: po.Tree, fold_segment, coords: dict, threshold: float, cfg: dict,
                      max_sugg: int, batch_size: int = 1024, moving: bool = True):
    P = cfg["poly_order"]
    tiling = cfg["tiling"]
    if moving:
        def resolve(ch):
            return cheby_resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                                       cfg["counts_init"], cfg["param_limits"], cfg["nbins"])

        def transform(ch):
            return cheby_transform_batch(ch, coords["next"], coords["cur"], tiling)
    else:
        def resolve(ch):
            return cheby_fixed_resolve_batch(ch, coords["add"], coords["cur"], coords["init"],
                                             cfg["counts_init"], cfg["param_limits"],
                                             cfg["nbins"])

        def transform(ch):
            return ch
    return po.level_loop(
        tree, fold_segment, threshold, cfg, max_sugg, batch_size,
        lambda lv: cheby_branch_batch(lv, coords["cur"], coords["prev"], cfg["nbins"],
                                      cfg["eta"], P, cfg["branch_max"], tiling),
        resolve, transform)


def cheby_ascend(tree: po.Tree, dyp_fold, idx_segments, coord_segments, coord_cur,
                 cfg: dict) -> po.Tree:
    if tree.size == 0:
        return tree
    t0_cur, scale_cur = coord_cur
    vec = tree.leaves[:, :-1, 0]
    f0 = tree.leaves[:, -1, 0]
    parts = []
    for iseg, (t0_seg, _) in zip(idx_segments, coord_segments):
        d = cheby_param_shift(vec, t0_cur, scale_cur, t0_seg)
        freq = f0 * (1 - d[:, -2] / C_VAL)
        delay = d[:, -1] / C_VAL
        phase = po.phase_idx_vec(t0_seg - t0_cur, f0, cfg["nbins"], delay)
        idx = po.nearest_2d(d[:, -3], freq, cfg["counts_init"], cfg["param_limits"])
        parts.append((po.seg_rows(dyp_fold[iseg], idx), phase))
    res = po._ascend_sum(parts, tree.folds)
    return po.Tree(tree.leaves.copy(), res, po._score_folds(res, cfg["widths"]),
                   tree.backtracks.copy(), scores_ep=tree.scores.copy())


def cheby_report_batch(leaves: np.ndarray, coord_report) -> np.ndarray:
    out = leaves.copy()
    d = cheby_to_taylor_full(leaves[:, :-1, :], coord_report[1])
    vals = d[:, :-2, 0]
    sigs = d[:, :-2, 1]
    out[:, -2, 0] = d[:, -1, 0]
    out[:, -2, 1] = d[:, -1, 1]
    v = d[:, -2, 0]
    dv = d[:, -2, 1]
    f0 = leaves[:, -1, 0]
    s = 1 - v / C_VAL
    out[:, :-3, 0] = vals / s[:, None]
    out[:, :-3, 1] = np.sqrt((sigs / s[:, None]) ** 2
                             + ((vals / (C_VAL * s[:, None] ** 2)) ** 2) * (dv[:, None] ** 2))
    out[:, -3, 0] = f0 * s
    out[:, -3, 1] = f0 * dv / C_VAL
    return out


def level_coords(scheme: po.MiddleOut, level: int, moving: bool = True) -> dict:
    return {"init": scheme.get_coord(0), "cur": scheme.get_current_coord(level, moving),
            "prev": scheme.get_previous_coord(level, moving), "next": scheme.get_coord(level),
            "add": scheme.get_segment_coord(level)}


def cheby_prune_run(dyp_fold, param_arr, dparams, tseg, thresholds, ref_seg, cfg, max_sugg,
                    batch_size=1024, nlevels=None, moving=True):
    """``Pruning.execute`` with the Chebyshev plugin (moving or fixed grid; the fixed
    report converts with the fixed report coord, dyn_poly_cheby.py:824-839)."""
    nseg = dyp_fold.shape[0]
    scheme = po.MiddleOut(nseg, ref_seg, tseg)
    P = cfg["poly_order"]
    leaves = cheby_seed(param_arr, dparams, P, scheme.get_coord(0))
    B = dyp_fold.shape[-1]
    folds = dyp_fold[scheme.ref_idx].reshape(-1, 2, B)
    folds = folds.astype(np.complex64 if np.iscomplexobj(folds) else np.float32)
    tree = po.Tree(leaves, folds.copy(), po._score_folds(folds, cfg["widths"]),
                   np.zeros((len(leaves), P + 2), dtype=np.int32))
    seed = tree
    stats, survivors, trees = [], [], []
    nl = nseg - 1 if nlevels is None else nlevels
    level = 0
    for level in range(1, nl + 1):
        tree, st, surv = cheby_prune_level(tree, dyp_fold[scheme.get_segment_idx(level)],
                                           level_coords(scheme, level, moving),
                                           float(thresholds[level - 1]), cfg, max_sugg,
                                           batch_size, moving)
        stats.append(st)
        survivors.append(surv)
        trees.append(tree)
        if tree.size == 0:
            break
    if tree.size > 0:
        idx_seg, coord_seg = scheme.get_segment_coords_so_far(level)
        tree = cheby_ascend(tree, dyp_fold, idx_seg, coord_seg,
                            scheme.get_previous_coord(level + 1, moving), cfg)
    report = (cheby_report_batch(tree.leaves, scheme.get_report_coord(moving)) if tree.size
              else tree.leaves)
    return {"tree": tree, "report": report, "stats": stats, "survivors": survivors,
            "trees": trees, "scheme": scheme, "seed": seed}


def generate_bp_cheby(param_arr, dparams_act, tseg_ffa: float, nsegments: int, nbins: int,
                      eta: float, ref_seg: int, tiling: str) -> np.ndarray:
    """Expected children per parent at every level, moving grid (``chebyshev.py:458-546``)."""
    P = len(dparams_act)
    f0 = np.asarray(param_arr[-1], dtype=np.float64)
    nf = len(f0)
    scheme = po.MiddleOut(nsegments, ref_seg, tseg_ffa)
    d = np.zeros((nf, P + 1))
    d[:, :P] = np.asarray(dparams_act, dtype=np.float64)[None, :]
    d[:, P - 1] *= C_VAL / f0
    dcur = taylor_to_cheby_errors(d, scheme.get_coord(0)[1])
    weights = np.ones(nf)
    bp = np.empty(nsegments - 1)
    dphi = eta / nbins
    for level in range(1, nsegments):
        c = level_coords(scheme, level)
        dcur = shift_cheby_errors(dcur, c["cur"], c["prev"], tiling)
        dnew = np.empty_like(dcur)
        for j in range(P + 1):
            dnew[:, j] = dphi * C_VAL / f0
        shift = np.abs(dcur - dnew) * (nbins * (f0 / C_VAL))[:, None]
        nbr = np.ones(nf, dtype=np.int64)
        dnext = np.zeros_like(dcur)
        for i in range(nf):
            for j in range(P):
                if shift[i, j] < (eta - FLOAT_EPSILON):
                    dnext[i, j] = dcur[i, j]
                    continue
                num = max(1, math.ceil(dcur[i, j] / dnew[i, j] - FLOAT_EPSILON))
                nbr[i] *= num
                dnext[i, j] = dcur[i, j] / num
        bp[level - 1] = np.sum(weights * nbr) / np.sum(weights)
        weights = weights * nbr
        dcur = shift_cheby_errors(dnext, c["next"], c["cur"], tiling)
    return bp
