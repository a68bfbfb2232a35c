/*
 * pyloki_hip.h -- C ABI of the MI355X (gfx950) search hot path of pyloki.
 *
 * The reference (pravirkr/pyloki v0.8.0) is Python + numba; its "FFI" for this path is
 * the set of numba kernels and plugin methods listed beside each entry point below
 * (paths relative to src/pyloki/). A Python caller binds these symbols with ctypes
 * (pyloki_amd/_lib.py); the ctypes stub a pyloki maintainer would add is shown in
 * INTEGRATION.md.
 *
 * Conventions
 *  - Every entry point returns an int status: PYL_OK (0) or a negative code.
 *    PYL_ERR_INVALID maps 1:1 to the reference's ValueError cases; the message is
 *    available from pyl_last_error() (thread-local) and is re-raised as ValueError.
 *  - Pointers named *_dev are device pointers (HBM); all other pointers are host
 *    pointers read synchronously during the call. Arrays are C-contiguous.
 *  - `stream` is a hipStream_t (NULL = the default stream). Kernels are enqueued on it;
 *    the call returns without synchronising unless stated otherwise.
 *  - A plan/handle is not re-entrant: one caller thread per handle at a time.
 */
#ifndef PYLOKI_HIP_H
#define PYLOKI_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PYL_OK 0
#define PYL_ERR_INVALID (-1) /* invalid argument: ValueError in the reference */
#define PYL_ERR_HIP (-2)     /* HIP runtime error */
#define PYL_ERR_ALLOC (-3)   /* device allocation failed */

/* Last error message of the calling thread ("" if none). */
const char* pyl_last_error(void);

/* ABI version (major*100 + minor). */
int pyl_abi_version(void);

/* ------------------------------------------------------------------------------------
 * Leaf brute fold.
 * Replaces core/fold.py:108-171 brutefold_start (via ffa_taylor_init fold.py:527-569,
 * FFATaylorDPFuncts.init dyn_pffa.py:48-55,190-204).
 * out fold_dev: f32 [nseg][nfreq][2][nbins], nseg = ceil(nsamps/seg_len); per-bin FP32
 * sums in sample order (bit-exact with the sequential reference loop). The plan holds the
 * FP64-resolved phase-bin lists of one (freqs, seg_len, nbins, tsamp, t_ref) and can be
 * reused for any number of series.
 * ------------------------------------------------------------------------------------ */
typedef struct pyl_bf_plan pyl_bf_plan;

int pyl_brutefold_plan_create(const double* freqs, int64_t nfreq, int64_t seg_len,
                              int32_t nbins, double tsamp, double t_ref, void* stream,
                              pyl_bf_plan** plan_out);
int pyl_brutefold_plan_destroy(pyl_bf_plan* plan);
int pyl_brutefold_execute(const pyl_bf_plan* plan, const float* ts_e_dev, const float* ts_v_dev,
                          int64_t nsamps, float* fold_dev, void* stream);
/* BF16-storage variant: FP32 accumulation, each bin rounded to BF16 (RNE) when stored. */
int pyl_brutefold_execute_bf16(const pyl_bf_plan* plan, const float* ts_e_dev,
                               const float* ts_v_dev, int64_t nsamps, uint16_t* fold_dev,
                               void* stream);
/* A batch of series sharing the plan in one launch (per 32 series): ts_e / ts_v are host
 * arrays of nseries device pointers, nsamps a host array of lengths. Series i's rows are
 * segments [sum_{j<i} ceil(nsamps[j]/seg_len), ...) of fold_dev, in the given order; each
 * series equals its own pyl_brutefold_execute. (FFATaylorDPFuncts.init per series.) */
int pyl_brutefold_execute_batch(const pyl_bf_plan* plan, const float* const* ts_e_dev,
                                const float* const* ts_v_dev, const int64_t* nsamps,
                                int32_t nseries, float* fold_dev, void* stream);
int pyl_brutefold_execute_batch_bf16(const pyl_bf_plan* plan, const float* const* ts_e_dev,
                                     const float* const* ts_v_dev, const int64_t* nsamps,
                                     int32_t nseries, uint16_t* fold_dev, void* stream);
/* Phase bins of the plan (nfreq x seg_len FP64 evaluations of psr_utils.py:49-56) that a
 * +-1 ulp change of the FP64 phase would move to another bin: the bound on disagreement
 * with an evaluation that rounds differently in the last bit. */
int pyl_brutefold_plan_boundary_count(const pyl_bf_plan* plan, int64_t* count_out);
/* Launch shape: segments staged per block (0 = default, 1, 2, 4); never changes a result. */
int pyl_brutefold_set_segments_per_block(int32_t segs);

/* ------------------------------------------------------------------------------------
 * Complex (Fourier-domain) FFA, use_fourier=True.
 * pyl_cfold_*: core/fold.py:395-481 brutefold_start_complex_opt (via
 *   ffa_taylor_init_complex fold.py:572-614, FFATaylorComplexDPFuncts.init
 *   dyn_pffa.py:82-140): fold_dev complex64 (float pairs) [nseg][nfreq][2][nbins/2 + 1],
 *   nsamps a multiple of seg_len. The plan holds the FP32 basis of the comb of
 *   harmonics; an execute is one launch of the library's FP32 matrix-core product of
 *   the segments (both components) and the basis, with the t_ref correction and the fold
 *   layout in its epilogue. Summation order is the matrix unit's (the reference's is its
 *   BLAS library's). No scratch memory.
 * pyl_ffa_shift_add_complex: the segment-pair loop of unify_fold with
 *   common.py:344-372 shift_add_complex: out[k] = tail[k] phi_t^k + head[k] phi_h^k, the
 *   phasors by the reference's complex64 recurrence from the FP64 phases of
 *   pyl_ffa_resolve (phases_dev f64 [ncoords_cur][2]); nbins in {32, 64, 128}.
 * pyl_score_boxcar_complex: ffa.py:400-411 / scoring.py:285-292: irfft of each complex
 *   row pair to nbins real bins (FP32), e / sqrt(v), boxcar S/N (modes as for
 *   pyl_score_boxcar); folds_dev complex64 [n][2][nbins/2 + 1].
 * ------------------------------------------------------------------------------------ */
typedef struct pyl_cfold_plan pyl_cfold_plan;

int pyl_cfold_plan_create(const double* freqs, int64_t nfreq, int64_t seg_len, int32_t nbins,
                          double tsamp, double t_ref, void* stream, pyl_cfold_plan** plan_out);
int pyl_cfold_plan_destroy(pyl_cfold_plan* plan);
int pyl_cfold_execute(pyl_cfold_plan* plan, const float* ts_e_dev, const float* ts_v_dev,
                      int64_t nsamps, float* fold_dev, void* stream);
int pyl_ffa_shift_add_complex(const float* fold_in_dev, int64_t nseg_in, int64_t ncoords_prev,
                              const int32_t* table_dev, const double* phases_dev,
                              int64_t ncoords_cur, int32_t nbins, float* fold_out_dev,
                              void* stream);
int pyl_score_boxcar_complex(const float* folds_dev, int64_t n, int32_t nbins,
                             const int64_t* widths, int32_t nw, double stdnoise, int32_t mode,
                             float* out_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * FFA level: resolve + shift-add.
 * pyl_ffa_resolve replaces core/fold.py:617-671 ffa_taylor_resolve (with
 * transforms.py:140-170 and psr_utils.py:12-56,466-514), evaluated in FP64 on the GPU
 * for every output parameter set of the level (dyn_pffa.py:465-480).
 *  params_dev: the current-level grids, concatenated (lengths counts_cur[0..P-1]);
 *              output parameter set j is their cartesian product, last param fastest.
 *  table_dev : int32 [ncoords_cur][4] = {row_tail, shift_tail, row_head, shift_head},
 *              row = flattened index into the previous level's coordinates and
 *              shift = int(f32(phase)+0.5f) with nbins -> 0 (common.py:313-318).
 *  phases_dev: optional f64 [ncoords_cur][2] FP64 phases (tail, head); may be NULL.
 * pyl_ffa_shift_add replaces the segment-pair loop of unify_fold (dyn_pffa.py:481-489)
 * with core/common.py:304-324 shift_add:
 *  out[q][j][c][b] = in[2q][row_t][c][(b-s_t) mod B] + in[2q+1][row_h][c][(b-s_h) mod B].
 * pyl_ffa_level = resolve + shift_add in one call (table_dev is the caller's scratch).
 * ------------------------------------------------------------------------------------ */
int pyl_ffa_resolve(int32_t nparams, const double* params_dev, const int64_t* counts_cur,
                    const int64_t* counts_prev, const double* param_limits, int32_t ffa_level,
                    double tseg_brute, int32_t nbins, int32_t* table_dev, double* phases_dev,
                    void* stream);
/* pyl_ffa_resolve_boundary_count: the resolve of one level without writing a table: adds
 * to counts_dev[0] the grid indices (int(raw + 1e-6), psr_utils.py:508) and to
 * counts_dev[1] the bin shifts (int(f32(x) + 0.5f), common.py:313-318) that a +-1 ulp
 * change of their FP64 input would flip -- the bound on disagreement with an evaluation
 * that differs in the last bit (the reference runs under numba fastmath). */
int pyl_ffa_resolve_boundary_count(int32_t nparams, const double* params_dev,
                                   const int64_t* counts_cur, const int64_t* counts_prev,
                                   const double* param_limits, int32_t ffa_level,
                                   double tseg_brute, int32_t nbins,
                                   unsigned long long* counts_dev, void* stream);
int pyl_ffa_shift_add(const float* fold_in_dev, int64_t nseg_in, int64_t ncoords_prev,
                      const int32_t* table_dev, int64_t ncoords_cur, int32_t nbins,
                      float* fold_out_dev, void* stream);
/* pyl_ffa_shift_add_score: the shift-add of one level fused with the per-width boxcar S/N
 * of every output row (unify_fold dyn_pffa.py:481-489 + get_fold_norm ffa.py:411 +
 * boxcar_snr_nd scoring.py:156-165), without storing the level's fold: snr_out f32
 * [nseg_in/2][ncoords_cur][nw]. Used for a series' last level (search.py:44-47).
 * n_inner = grid count of the last (frequency) axis; it must divide ncoords_cur and sets
 * the output tiling (rows of neighbouring outer indices share source rows).
 * nbins in {32, 64, 128}; widths as for pyl_score_boxcar. */
int pyl_ffa_shift_add_score(const float* fold_in_dev, int64_t nseg_in, int64_t ncoords_prev,
                            const int32_t* table_dev, int64_t ncoords_cur, int64_t n_inner,
                            int32_t nbins, const int64_t* widths, int32_t nw, double stdnoise,
                            float* snr_out_dev, void* stream);
/* BF16-storage / FP32-accumulate variants of the level kernels: rows are loaded as BF16,
 * added in FP32 and stored rounded to BF16 (RNE); the fused variant normalises and scores
 * the BF16-rounded row. nbins in {32, 64, 128, 256} ({32, 64, 128} fused). */
int pyl_ffa_shift_add_bf16(const uint16_t* fold_in_dev, int64_t nseg_in, int64_t ncoords_prev,
                           const int32_t* table_dev, int64_t ncoords_cur, int32_t nbins,
                           uint16_t* fold_out_dev, void* stream);
int pyl_ffa_shift_add_score_bf16(const uint16_t* fold_in_dev, int64_t nseg_in,
                                 int64_t ncoords_prev, const int32_t* table_dev,
                                 int64_t ncoords_cur, int64_t n_inner, int32_t nbins,
                                 const int64_t* widths, int32_t nw, double stdnoise,
                                 float* snr_out_dev, void* stream);
/* Fold row layouts of the FP32 level kernels (no reference counterpart: an internal
 * storage choice between levels). PYL_FOLD_DENSE: [..][2][nbins], the reference's layout.
 * PYL_FOLD_WRAP: [..][2][nbins + 4], each component followed by a copy of its first 4 bins,
 * so that every 4 consecutive bins of a circular rotation are one contiguous 16-byte load;
 * written by the level before a series' last level and read by the fused last level.
 * The wrap layout needs nbins in {32, 64, 128, 256} (and FP32 storage). */
#define PYL_FOLD_DENSE 0
#define PYL_FOLD_WRAP 1
/* pyl_ffa_shift_add / pyl_ffa_shift_add_score with explicit input and output row layouts
 * (the plain entry points are the dense-to-dense case). */
int pyl_ffa_shift_add_layout(const float* fold_in_dev, int32_t in_layout, int64_t nseg_in,
                             int64_t ncoords_prev, const int32_t* table_dev,
                             int64_t ncoords_cur, int32_t nbins, float* fold_out_dev,
                             int32_t out_layout, void* stream);
int pyl_ffa_shift_add_score_layout(const float* fold_in_dev, int32_t in_layout,
                                   int64_t nseg_in, int64_t ncoords_prev,
                                   const int32_t* table_dev, int64_t ncoords_cur,
                                   int64_t n_inner, int32_t nbins, const int64_t* widths,
                                   int32_t nw, double stdnoise, float* snr_out_dev,
                                   void* stream);
/* A series' last two levels fused with scoring: fold_in_dev holds the fold two levels
 * before the last ([nseg_in][ncoords_prev2][2][nbins (+4 in the wrap layout)], nseg_in a
 * multiple of 4), table_prev the resolve table of the level before the last (ncoords_prev
 * rows, indices into ncoords_prev2) and table_dev that of the last level. The result equals
 * pyl_ffa_shift_add followed by pyl_ffa_shift_add_score_layout bit for bit (the sums are
 * formed in the same order and rounded to FP32 at the same points); the intermediate fold
 * is never written. nbins 32 or 64, at most 16 widths. No reference counterpart as an
 * interface: two unify_fold calls (dyn_pffa.py:435-489) and the final scoring. */
/* Launch shape of the round-based fused scoring kernels (pyl_ffa_shift_add_score_layout,
 * pyl_ffa_shift_add2_score_layout; nbins 32/64): a block owns 256 output rows, `rows`
 * along the outer grid axis by 256 / rows along the inner one. 0 restores the default
 * (16). Process-wide; set by the autotuner (pyloki_amd/autotune.py). No reference
 * counterpart: it changes no result, only which rows share a block's L2 reuse. */
int pyl_ffa_set_tile_rows(int32_t rows);
int pyl_ffa_get_tile_rows(void);
/* Gather descriptors of the two-level kernel (pyl_ffa_shift_add2_score_layout): 1 (the
 * default) packs each output row's four source rows and shifts into 16 B of LDS wherever
 * an input segment holds fewer than 2^32 bytes, so four blocks fit a CU; 0 keeps the
 * 48 B table entries. Process-wide. No reference counterpart: the results are identical,
 * the switch exists so that tests can check both variants. */
int pyl_ffa_set_packed_tables(int32_t enable);

int pyl_ffa_shift_add2_score_layout(const float* fold_in_dev, int32_t in_layout,
                                    int64_t nseg_in, int64_t ncoords_prev2,
                                    const int32_t* table_prev_dev, int64_t ncoords_prev,
                                    const int32_t* table_dev, int64_t ncoords_cur,
                                    int64_t n_inner, int32_t nbins, const int64_t* widths,
                                    int32_t nw, double stdnoise, float* snr_out_dev,
                                    void* stream);
/* pyl_ffa_shift_add2_layout: two consecutive levels in one pass (FP32, input
 * [nseg_in][ncoords_prev2][2][nbins] or its PYL_FOLD_WRAP form, nseg_in a multiple of 4):
 * equal bit for bit to pyl_ffa_shift_add with table_prev (level L+1) followed by
 * pyl_ffa_shift_add with table_dev (level L+2), but the intermediate fold is never stored.
 * Output [nseg_in/4][ncoords_cur][2][nbins] (dense) or [..][2][nbins + 4] (wrap).
 * No reference counterpart: the reference runs unify_fold once per level
 * (ffa.py:332-380). */
int pyl_ffa_shift_add2_layout(const float* fold_in_dev, int32_t in_layout, int64_t nseg_in,
                              int64_t ncoords_prev2, const int32_t* table_prev_dev, int64_t ncoords_prev,
                              const int32_t* table_dev, int64_t ncoords_cur, int32_t nbins,
                              float* fold_out_dev, int32_t out_layout, void* stream);
int pyl_ffa_level(int32_t nparams, const double* params_dev, const int64_t* counts_cur,
                  const int64_t* counts_prev, const double* param_limits, int32_t ffa_level,
                  double tseg_brute, int32_t nbins, const float* fold_in_dev, int64_t nseg_in,
                  int32_t* table_dev, float* fold_out_dev, void* stream);
/* pyl_ffa_level_noalloc: no counterpart in the reference. pyl_ffa_level for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops): it
 * takes no workspace and calls no allocator; the resolve table lives in a fixed device
 * array of the library, at most pyl_ffa_level_noalloc_max_coords() parameter sets
 * (2^21, 32 MiB per GPU, held until the process ends). Larger levels are rejected with
 * PYL_ERR_INVALID. Calls on one GPU share that array: calls that overlap on one GPU (two
 * streams, two threads) race and must be serialised by the caller. */
int pyl_ffa_level_noalloc(int32_t nparams, const double* params_dev, const int64_t* counts_cur,
                          const int64_t* counts_prev, const double* param_limits,
                          int32_t ffa_level, double tseg_brute, int32_t nbins,
                          const float* fold_in_dev, int64_t nseg_in, float* fold_out_dev,
                          void* stream);
int64_t pyl_ffa_level_noalloc_max_coords(void);

/* ------------------------------------------------------------------------------------
 * Boxcar matched-filter S/N.
 * Replaces detection/scoring.py:99-165 boxcar_snr_1d/2d/nd (mode PYL_SCORE_PER_WIDTH,
 * out f32 [n][nw]) and scoring.py:249-282 snr_score_batch_func (mode PYL_SCORE_MAX,
 * out f32 [n]). layout PYL_LAYOUT_EV: folds_dev is [n][2][nbins] (e, v) and the profile
 * is e/sqrt(v) in FP32 (ffa.py:411, scoring.py:265-266); PYL_LAYOUT_NORM: folds_dev is
 * [n][nbins] already normalised. widths: host int64 [nw], 1 <= w < nbins, nw <= 64.
 * ------------------------------------------------------------------------------------ */
#define PYL_LAYOUT_EV 0
#define PYL_LAYOUT_NORM 1
#define PYL_LAYOUT_EV_BF16 2 /* [n][2][nbins] BF16 (e, v); folds_dev points at uint16 */
#define PYL_SCORE_PER_WIDTH 0
#define PYL_SCORE_MAX 1

int pyl_score_boxcar(const float* folds_dev, int64_t n, int32_t nbins, int32_t layout,
                     const int64_t* widths, int32_t nw, double stdnoise, int32_t mode,
                     float* out_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * EP (enumerate-and-prune) Taylor search, real folds, moving or fixed grid.
 * One pruning level of prune.py:165-332 pruning_iteration_batched is three calls:
 *   pyl_prune_branch      core/taylor.py:68-169 poly_taylor_branch_batch: per parent and
 *                         parameter the child count, first child centre and step
 *                         (psr_utils.py:347-379), child counts per parent;
 *   pyl_prune_children    for every child (parent-major, odometer order of
 *                         np_utils.py:174-238): resolve (taylor.py:172-241, FP64),
 *                         load + shift-add (common.py:375-400) and score
 *                         (scoring.py:249-282) -> per-child score, segment row, shift,
 *                         phase;
 *   pyl_prune_materialize for the surviving children: transform to coord_next
 *                         (taylor.py:323-339), fold, score and backtrack
 *                         (prune.py:302-314) of the new tree.
 * pyl_prune_ascend replaces dyn_poly_taylor.py:587-628 ascend_func (re-integration of
 * every survivor over all segments so far, taylor.py:280-320, common.py:432-459).
 * Which children survive (threshold and the WorldTree overflow rule of
 * world_tree.py:482-551) is decided between calls 2 and 3 by the caller.
 *
 * Layouts: leaves f64 [n][P+2][2]; folds f32 [n][2][nbins]; seg_fold f32 rows of the FFA
 * fold of one segment, row = accel_idx * n_freq + freq_idx (leading axes at index 0,
 * common.py:178-235); backtracks int32 [n][P+2].
 * ------------------------------------------------------------------------------------ */
#define PYL_TILING_AGGRESSIVE 0
#define PYL_TILING_QUADRATURE 1
#define PYL_TILING_CONSERVATIVE 2
#define PYL_MAX_WIDTHS 64

typedef struct pyl_prune_params {
  int32_t poly_order;     /* P: Taylor parameters d_P..d_1 branched; leaf rows P + 2 */
  int32_t nbins;
  int32_t branch_max;
  int32_t tiling;         /* PYL_TILING_* */
  double eta;
  int32_t nparams_ffa;    /* FFA grid parameters (backtrack width = P + 2) */
  int32_t n_accel;        /* FFA grid count of the acceleration axis */
  int32_t n_freq;         /* FFA grid count of the frequency axis */
  int32_t nw;             /* number of boxcar widths */
  double accel_lim[2];
  double freq_lim[2];
  int32_t widths[PYL_MAX_WIDTHS];
  int32_t fixed_grid;     /* 0: moving grid; 1: fixed grid (use_moving_grid=False): the
                             coords are the fixed ones (snail.py:286-345), the resolve is the
                             fixed variant (taylor.py:244-277, circular.py:531-590,
                             chebyshev.py:261-300) and the transform is the identity */
} pyl_prune_params;

typedef struct pyl_prune_coords {
  double t0_init;         /* coord_init[0] */
  double t0_cur;          /* coord_cur[0] */
  double scale_cur;       /* coord_cur[1] */
  double t0_next;         /* coord_next[0] */
  double t0_add;          /* coord_add[0] */
  double t0_prev;         /* coord_prev[0]  (Chebyshev basis only) */
  double scale_prev;      /* coord_prev[1]  (Chebyshev basis only) */
  double scale_next;      /* coord_next[1]  (Chebyshev basis only) */
} pyl_prune_coords;

/* branch_dev: f64 [n_parents][P][2] = {first child centre, child step}; counts_dev int32
 * [n_parents][P] = children per parameter, or -1 when the parameter is not split (the
 * child keeps the parent value); nchild_dev int64 [n_parents]. err_dev (int64, one word,
 * set to INT64_MAX by the caller) receives the smallest (parent * P + param) * 4 + kind
 * over failing parameters: kind 1 = "increase branch_max", kind 2 = non-positive step
 * (psr_utils.py:358-365). */
int pyl_prune_branch(const pyl_prune_params* p, const pyl_prune_coords* c,
                     const double* leaves_dev, int64_t n_parents, double* branch_dev,
                     int32_t* counts_dev, int64_t* nchild_dev, int64_t* err_dev, void* stream);

/* child_off_dev: int64 [n_parents + 1] exclusive prefix of nchild. Outputs per child:
 * parent (int32), seg row (int32), shift (int32), phase (f64), score (f32).
 * minmax_dev (f32 [2], caller-initialised to {+inf, -inf}) receives the score min/max. */
int pyl_prune_children(const pyl_prune_params* p, const pyl_prune_coords* c,
                       const double* leaves_dev, const float* folds_dev, int64_t n_parents,
                       const double* branch_dev, const int32_t* counts_dev,
                       const int64_t* child_off_dev, int64_t n_children,
                       const float* seg_fold_dev, int32_t* child_parent_dev,
                       int32_t* child_row_dev, int32_t* child_shift_dev, double* child_phase_dev,
                       float* child_score_dev, float* minmax_dev, void* stream);

/* surv_dev: int64 [n_surv] child ids in increasing order. Writes the new tree. */
int pyl_prune_materialize(const pyl_prune_params* p, const pyl_prune_coords* c,
                          const double* leaves_dev, const float* folds_dev,
                          const double* branch_dev, const int32_t* counts_dev,
                          const int64_t* child_off_dev, const float* seg_fold_dev,
                          const int32_t* child_parent_dev, const int32_t* child_row_dev,
                          const int32_t* child_shift_dev, const double* child_phase_dev,
                          const float* child_score_dev, const int64_t* surv_dev, int64_t n_surv,
                          double* out_leaves_dev, float* out_folds_dev, float* out_scores_dev,
                          int32_t* out_backtracks_dev, void* stream);

/* Complex-fold variants (use_fourier=True; PrunePolyTaylorComplexDPFuncts,
 * dyn_poly_taylor.py): folds, segment rows and outputs are complex64 [..][2][nbins/2 + 1]
 * (float pairs). A child's fold is common.py:403-429 shift_add_complex_batch (segment row
 * times the phasor powers of its FP64 phase, plus the parent row) and its score
 * scoring.py:285-292 (irfft, then the boxcar scorer); the ascend sums the rotated complex
 * rows (common.py:463-490). Arguments as for the real entry points. */
int pyl_prune_children_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                               const double* leaves_dev, const float* folds_dev,
                               int64_t n_parents, const double* branch_dev,
                               const int32_t* counts_dev, const int64_t* child_off_dev,
                               int64_t n_children, const float* seg_fold_dev,
                               int32_t* child_parent_dev, int32_t* child_row_dev,
                               int32_t* child_shift_dev, double* child_phase_dev,
                               float* child_score_dev, float* minmax_dev, void* stream);
int pyl_prune_materialize_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                                  const double* leaves_dev, const float* folds_dev,
                                  const double* branch_dev, const int32_t* counts_dev,
                                  const int64_t* child_off_dev, const float* seg_fold_dev,
                                  const int32_t* child_parent_dev, const int32_t* child_row_dev,
                                  const int32_t* child_shift_dev, const double* child_phase_dev,
                                  const float* child_score_dev, const int64_t* surv_dev,
                                  int64_t n_surv, double* out_leaves_dev, float* out_folds_dev,
                                  float* out_scores_dev, int32_t* out_backtracks_dev,
                                  void* stream);
int pyl_prune_ascend_complex(const pyl_prune_params* p, const double* leaves_dev, int64_t n,
                             const float* dyp_fold_dev, int64_t ncoords_ffa,
                             const int64_t* seg_idx, const double* seg_t0, int32_t nseg_used,
                             double t0_cur, float* out_folds_dev, float* out_scores_dev,
                             void* stream);

/* dyp_fold_dev: f32 [nseg][ncoords][2][nbins] (the FFA fold handed to the EP).
 * seg_idx / seg_t0: host [nseg_used] segment indices and reference times (segments so
 * far, in scheme order). Writes the re-integrated folds and their scores. */
int pyl_prune_ascend(const pyl_prune_params* p, const double* leaves_dev, int64_t n,
                     const float* dyp_fold_dev, int64_t ncoords_ffa, const int64_t* seg_idx,
                     const double* seg_t0, int32_t nseg_used, double t0_cur,
                     float* out_folds_dev, float* out_scores_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * EP, Chebyshev basis (PrunePolyChebyshevDPFuncts, dyn_poly_cheby.py:587-842; moving grid).
 * Leaves hold Chebyshev coefficients [alpha_P .. alpha_1, alpha_0, (f0, flag)] with their
 * cell widths, on the domain (t0, scale) of the current coordinate. One level is the same
 * three calls as the Taylor basis:
 *  pyl_cheb_branch      chebyshev.py:72-177 poly_chebyshev_branch_batch: every parent moved
 *                       from coord_prev to coord_cur (transforms.py:650-700 shift_cheby_full,
 *                       values and widths by the tiling strategy) into leaves_cur_dev
 *                       (f64 [n][P+2][2]), then the Taylor split rule with the step
 *                       eta/nbins * C / f0 for every coefficient (psr_utils.py:258-278);
 *  pyl_cheb_children    as pyl_prune_children on leaves_cur_dev, resolved by
 *                       chebyshev.py:180-258 (the series moved to t_add and t_init, then
 *                       converted to Taylor coefficients, transforms.py:532-560,703-737);
 *  pyl_cheb_materialize as pyl_prune_materialize on leaves_cur_dev, the survivors moved to
 *                       coord_next by chebyshev.py:352-368.
 * pyl_cheb_ascend replaces dyn_poly_cheby.py:733-776 (chebyshev.py:261-349) and needs the
 * scale of the survivors' domain. The domain-transfer matrices (maths.py:399-557) are built
 * per call on the host. Arguments otherwise as for the Taylor entry points.
 * ------------------------------------------------------------------------------------ */
int pyl_cheb_branch(const pyl_prune_params* p, const pyl_prune_coords* c,
                    const double* leaves_dev, int64_t n_parents, double* leaves_cur_dev,
                    double* branch_dev, int32_t* counts_dev, int64_t* nchild_dev,
                    int64_t* err_dev, void* stream);
int pyl_cheb_children(const pyl_prune_params* p, const pyl_prune_coords* c,
                      const double* leaves_cur_dev, const float* folds_dev, int64_t n_parents,
                      const double* branch_dev, const int32_t* counts_dev,
                      const int64_t* child_off_dev, int64_t n_children,
                      const float* seg_fold_dev, int32_t* child_parent_dev,
                      int32_t* child_row_dev, int32_t* child_shift_dev, double* child_phase_dev,
                      float* child_score_dev, float* minmax_dev, void* stream);
int pyl_cheb_materialize(const pyl_prune_params* p, const pyl_prune_coords* c,
                         const double* leaves_cur_dev, const float* folds_dev,
                         const double* branch_dev, const int32_t* counts_dev,
                         const int64_t* child_off_dev, const float* seg_fold_dev,
                         const int32_t* child_parent_dev, const int32_t* child_row_dev,
                         const int32_t* child_shift_dev, const double* child_phase_dev,
                         const float* child_score_dev, const int64_t* surv_dev, int64_t n_surv,
                         double* out_leaves_dev, float* out_folds_dev, float* out_scores_dev,
                         int32_t* out_backtracks_dev, void* stream);
int pyl_cheb_ascend(const pyl_prune_params* p, const double* leaves_dev, int64_t n,
                    const float* dyp_fold_dev, int64_t ncoords_ffa, const int64_t* seg_idx,
                    const double* seg_t0, int32_t nseg_used, double t0_cur, double scale_cur,
                    float* out_folds_dev, float* out_scores_dev, void* stream);
/* Complex-fold variants (PrunePolyChebyshevComplexDPFuncts, dyn_poly_cheby.py:304-469). */
int pyl_cheb_children_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                              const double* leaves_cur_dev, const float* folds_dev,
                              int64_t n_parents, const double* branch_dev,
                              const int32_t* counts_dev, const int64_t* child_off_dev,
                              int64_t n_children, const float* seg_fold_dev,
                              int32_t* child_parent_dev, int32_t* child_row_dev,
                              int32_t* child_shift_dev, double* child_phase_dev,
                              float* child_score_dev, float* minmax_dev, void* stream);
int pyl_cheb_materialize_complex(const pyl_prune_params* p, const pyl_prune_coords* c,
                                 const double* leaves_cur_dev, const float* folds_dev,
                                 const double* branch_dev, const int32_t* counts_dev,
                                 const int64_t* child_off_dev, const float* seg_fold_dev,
                                 const int32_t* child_parent_dev, const int32_t* child_row_dev,
                                 const int32_t* child_shift_dev, const double* child_phase_dev,
                                 const float* child_score_dev, const int64_t* surv_dev,
                                 int64_t n_surv, double* out_leaves_dev, float* out_folds_dev,
                                 float* out_scores_dev, int32_t* out_backtracks_dev,
                                 void* stream);
int pyl_cheb_ascend_complex(const pyl_prune_params* p, const double* leaves_dev, int64_t n,
                            const float* dyp_fold_dev, int64_t ncoords_ffa,
                            const int64_t* seg_idx, const double* seg_t0, int32_t nseg_used,
                            double t0_cur, double scale_cur, float* out_folds_dev,
                            float* out_scores_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * EP level loop without host synchronisation (Taylor and Chebyshev bases, real or complex
 * folds). Replaces the reference's per-level driver loop prune.py:615-669 execute_iter
 * (branch, resolve, shift-add, score, threshold, add to the WorldTree) for `nlev`
 * consecutive levels in one call: every count (parents, children, survivors) stays in
 * device memory, kernels stride over capacity-sized grids and read their counts there,
 * and the call only enqueues work. Results equal those of the per-level entry points above
 * bit for bit.
 *
 * A level stops the run, and leaves its input tree untouched, when it cannot be finished
 * without the host: the branch step failed (the reference's ValueError), the children
 * exceed cap_child, the passing children exceed max_sugg (the batch-sequential overflow
 * rule, world_tree.py:482-551, is then replayed on the host) or the backtrack history is
 * full. Its record then holds a PYL_EP_ST_* status and every later level has no parents;
 * the caller runs that level through the per-level entry points and calls again.
 *
 * Level l reads the tree of level l - 1 (the input tree for l = 0) and writes tree buffer
 * (first_out + l) % 2 of L. Its survivors' backtracks go to rows rec[l].HIST.. of L->hist
 * (hist_row0 for l = 0).
 * coords / thresholds / seg_idx: host arrays [nlev] (the level's coordinates, S/N threshold
 * and segment index); the segment rows of level l start at dyp_fold_dev + seg_idx[l] *
 * seg_stride floats.
 * flags: PYL_EP_COMPLEX (complex64 folds) | PYL_EP_CHEBYSHEV.
 * ------------------------------------------------------------------------------------ */
#define PYL_EP_COMPLEX 1
#define PYL_EP_CHEBYSHEV 2
#define PYL_EP_SCAN_BLOCKS 256

#define PYL_EP_REC_NPAR 0   /* parents of the level */
#define PYL_EP_REC_NCHILD 1 /* children (circular: before validation) */
#define PYL_EP_REC_NSURV 2  /* survivors (the next level's parents) */
#define PYL_EP_REC_STATUS 3 /* PYL_EP_ST_* */
#define PYL_EP_REC_ERR 4    /* branch error word, as err_dev of pyl_prune_branch */
#define PYL_EP_REC_HIST 5   /* first history row of the level's survivors */
#define PYL_EP_REC_NVALID 6 /* children scored (circular: after validation) */
#define PYL_EP_REC_NHELD 7  /* circular: held children (before the crackle expansion) */
#define PYL_EP_REC_NBATCH 8 /* circular: batches of parents */
#define PYL_EP_REC_NEXPAND 9 /* circular: held children to expand (0 if the level stops) */
#define PYL_EP_REC_NKEPT 10 /* survivors kept by the budget (= the next level's parents) */
#define PYL_EP_REC_NRANK 11 /* survivors the budget ranks: NSURV above the budget, else 0 */
#define PYL_EP_REC_WORDS 12

#define PYL_EP_ST_OK 0
#define PYL_EP_ST_BRANCH 1    /* branch error: rerun the level to raise it */
#define PYL_EP_ST_CHILD_CAP 2 /* children exceed cap_child */
#define PYL_EP_ST_OVERFLOW 3  /* passing children exceed max_sugg */
#define PYL_EP_ST_HIST_FULL 4 /* backtrack history full */
#define PYL_EP_ST_BUDGET 5    /* survivors exceed a budget the level was enqueued without */
#define PYL_EP_ST_TREE_OVERFLOW 6 /* forest: a tree's passing children exceed its max_sugg,
                                     or its held children times branch_max reach 2^31 */

/* Optional per-phase timing of pyl_ep_levels: HIP events recorded on the stream around
 * each level's phases (PYL_EP_PH_*), read back by pyl_ep_timer_read (which waits for them
 * and clears the timer). Records beyond the capacity are dropped. */
typedef struct pyl_ep_timer pyl_ep_timer;
#define PYL_EP_PH_BRANCH 0      /* branch + child offsets */
#define PYL_EP_PH_CHILDREN 1    /* resolve + shift-add + score of every child */
#define PYL_EP_PH_SELECT 2      /* survivor indices */
#define PYL_EP_PH_MATERIALIZE 3 /* survivors' leaves, folds, backtracks */
#define PYL_EP_PH_BUDGET 4      /* the survivor budget (pyl_ep_budget) */
int pyl_ep_timer_create(int32_t capacity, pyl_ep_timer** out);
int pyl_ep_timer_destroy(pyl_ep_timer* t);
int pyl_ep_timer_read(pyl_ep_timer* t, int32_t* phase_out, float* ms_out, int32_t cap,
                      int32_t* n_out);

/* Survivor budget of a level (no counterpart in the reference: a learned, best-first
 * search, for callers that trade recall for speed). After a level's survivors are
 * materialised, the `budget` of highest key are kept, in their order: the key is the S/N
 * (PYL_BUDGET_SCORE) or the ranking network's logit (PYL_BUDGET_CNN, pyl_cnn_forward on
 * the survivors' folds and pyl_cnn_features). Ties at the cut keep the lowest rows. The
 * exhaustive search (no budget) stays the reference: every kept survivor is still scored
 * exactly. Scratch arrays hold `cap` rows (the tree buffers' rows). */
#define PYL_BUDGET_SCORE 1
#define PYL_BUDGET_CNN 2
#define PYL_BUDGET_HYBRID 3 /* the larger of the rank fractions by logit and by S/N */
#define PYL_BUDGET_FP32 0
#define PYL_BUDGET_BF16 1
#define PYL_BUDGET_SEL_WORDS 272
#define PYL_BUDGET_HIST_BINS 1024
#define PYL_BUDGET_STATS 1032 /* f64 words of `stats` */
typedef struct pyl_ep_budget {
  int32_t kind;          /* PYL_BUDGET_* */
  int32_t nbins;         /* fold bins (the network's input) */
  int32_t complex_folds; /* folds are complex64 [2][nbins/2 + 1] */
  int32_t nlevels;       /* levels of the search (the feature level / nlevels) */
  int32_t level0;        /* level loop: the search level of the call's first level */
  int32_t precision;     /* PYL_BUDGET_FP32, or PYL_BUDGET_BF16: pyl_cnn_forward_bf16 */
  int64_t budget;        /* survivors kept per level (>= 1) */
  const float* params;   /* the network's parameters (PYL_BUDGET_CNN) */
  float* keys;           /* f32 [cap]: ranking key of every survivor */
  float* feats;          /* f32 [cap][8]: network features (PYL_BUDGET_CNN) */
  int64_t* kept;         /* i64 [cap]: kept rows */
  int32_t* bt_tmp;       /* i32 [cap][P + 2]: the level loop's backtracks before the cut */
  uint32_t* sel;         /* u32 [PYL_BUDGET_SEL_WORDS]: radix-select state and histogram */
  double* stats;         /* f64 [PYL_BUDGET_STATS]: per-level score statistics */
  uint32_t* shist;       /* u32 [2 * PYL_BUDGET_HIST_BINS]: score histogram and its prefix */
  int64_t* part;         /* i64 [PYL_EP_SCAN_BLOCKS] */
  /* Level loop: host flags indexed by search level (0..nlevels), or null. A level
   * whose flag is 0 is enqueued without the ranking and the cut: its survivors are copied
   * as they are when they fit the budget, else the level stops with PYL_EP_ST_BUDGET (and
   * runs per level, where the budget always applies). Results do not depend on the flags;
   * the launches of a level without the cut are 1 instead of 15. */
  const uint8_t* level_mask;
  /* Host budgets indexed by search level (0..nlevels), each >= 1, or null: every level
   * keeps `budget`. A schedule keeps more survivors on the levels where the ranking is
   * least reliable. */
  const int64_t* level_budget;
  /* Host keys indexed by search level (0..nlevels), each PYL_BUDGET_SCORE or `kind`, or
   * null: every level ranks by `kind`. A network budget can rank by S/N where S/N ranking
   * suffices and skip the network's inference there. */
  const int32_t* level_kind;
} pyl_ep_budget;

/* The budget on one tree of n rows (the per-level path): writes the kept rows to out_*
 * (out_bt: int32 [budget][P + 2]) and their count to rec_dev[PYL_EP_REC_NKEPT]; rec_dev is
 * an int64 [2][PYL_EP_REC_WORDS] workspace. `level` feeds the network's features. */
int pyl_ep_budget_apply(const pyl_ep_budget* b, int32_t P, int32_t level,
                        const double* leaves_dev, const float* folds_dev,
                        const float* scores_dev, const int32_t* bt_dev, int64_t n,
                        int64_t* rec_dev, double* out_leaves_dev, float* out_folds_dev,
                        float* out_scores_dev, int32_t* out_bt_dev, void* stream);

/* The ranking network's 8 features of n survivors of one level (f32 [n][8]): S/N / 10,
 * level / nlevels, the S/N's z-score over the level, its rank fraction over the level
 * (from a 1024-bin histogram between the level's smallest and largest S/N: rows below its
 * bin plus half of its bin's other rows, over n - 1), then log10(1 + cell width) / 10 of
 * the (up to) four lowest-order branched parameters. stats / shist as in pyl_ep_budget. */
int pyl_cnn_features(const double* leaves_dev, const float* scores_dev, int64_t n, int32_t P,
                     int32_t level, int32_t nlevels, double* stats_dev, uint32_t* shist_dev,
                     float* feats_dev, void* stream);

typedef struct pyl_ep_loop {
  int64_t cap_par;      /* rows of each tree buffer (>= max_sugg and the input tree) */
  int64_t cap_child;    /* children of one level */
  int64_t cap_hist;     /* rows of the backtrack history */
  int64_t max_sugg;     /* survivor buffer of the reference's WorldTree */
  double* leaves[2];    /* tree buffers: f64 [cap_par][P + 2][2] */
  float* folds[2];      /*   f32 [cap_par][2][nbins] or complex64 [cap_par][2][nbins/2 + 1] */
  float* scores[2];     /*   f32 [cap_par] */
  int32_t* hist;        /* int32 [cap_hist][P + 2]: every level's backtracks, level by level */
  double* leaves_cur;   /* Chebyshev: f64 [cap_par][P + 2][2] (parents on coord_cur) */
  double* branch;       /* f64 [cap_par][P][2] */
  int32_t* counts;      /* int32 [cap_par][P] */
  int64_t* nchild;      /* int64 [cap_par] */
  int64_t* off;         /* int64 [cap_par + 1] */
  int32_t* c_parent;    /* int32 [cap_child] */
  int32_t* c_row;       /* int32 [cap_child] */
  int32_t* c_shift;     /* int32 [cap_child] */
  double* c_phase;      /* f64 [cap_child] */
  float* c_score;       /* f32 [cap_child] */
  int64_t* surv;        /* int64 [max_sugg] */
  int64_t* part;        /* int64 [PYL_EP_SCAN_BLOCKS] */
  int64_t* rec;         /* int64 [nlev + 1][PYL_EP_REC_WORDS]: level records (written) */
  float* minmax;        /* f32 [nlev][2]: children's score min/max per level (written) */
  pyl_ep_timer* timer;  /* NULL, or the phases are timed */
  const pyl_ep_budget* budget; /* NULL, or every level keeps at most budget->budget
                                  survivors: they are materialised into tree buffer 0 (and
                                  budget->bt_tmp) and the kept rows gathered into tree
                                  buffer 1, which the next level reads (first_out unused) */
} pyl_ep_loop;

int pyl_ep_levels(const pyl_prune_params* p, int32_t flags, const pyl_prune_coords* coords,
                  const double* thresholds, const int64_t* seg_idx, int32_t nlev,
                  const float* dyp_fold_dev, int64_t seg_stride, const double* in_leaves_dev,
                  const float* in_folds_dev, int64_t n_in, int64_t hist_row0,
                  int32_t first_out, const pyl_ep_loop* L, void* stream);

/* ------------------------------------------------------------------------------------
 * EP, circular basis (prune_poly_order = 5: crackle, snap, jerk, accel, vel; leaf rows 7).
 * The entry points replace the circular plugin (dyn_circular_taylor.py:484-678) driven
 * by pruning_iteration_batched (prune.py:165-332):
 *  pyl_circ_branch     circ_taylor_branch_batch (circular.py:204-394) per parent: every
 *                      parameter but crackle as in the Taylor branch; crackle is held at the
 *                      parent value (its step already the split step when it needs a split)
 *                      and its split (first centre, count) kept for the expansion;
 *  pyl_circ_stage1     per held child: parent and whether it expands (snap-accel "hole",
 *                      circular.py:169-201, with a crackle that needs a split);
 *  (caller)            final child order per batch of batch_size parents: the kept
 *                      children in held order, then every expanding child's crackle children
 *                      (circular.py:359-394) -> s1_pos;
 *  pyl_circ_expand     per final child: its held child and crackle index, validation
 *                      (circ_taylor_validate_batch, circular.py:82-124,397-440) and region
 *                      (get_circ_taylor_mask, circular.py:13-79,127-166);
 *  pyl_circ_children   per validated child: resolve with the region's propagator
 *                      (circ_taylor_resolve_batch, circular.py:443-528; transforms.py:248-313),
 *                      shift-add and score;
 *  pyl_circ_materialize per survivor: basis flag (snap 1, crackle 2, or 0 when its batch
 *                      holds a Taylor cell, circular.py:478-511), transform
 *                      (circ_taylor_transform_batch, circular.py:669-705), backtrack, fold;
 *  pyl_circ_ascend     ascend_func (dyn_circular_taylor.py:636-678, circular.py:605-666).
 * ------------------------------------------------------------------------------------ */
typedef struct pyl_circ_params {
  double prop_sig;        /* propagator_significance (region classification) */
  double val_sig;         /* validation_significance */
  double omega_max_sq;    /* (2 pi / p_orb_min)^2 */
  double x_mass_const;    /* PulsarSearchConfig.x_mass_const */
  int64_t batch_size;     /* parents per batch (the reference's batch_size) */
} pyl_circ_params;

/* Device buffers of one circular level (allocated by the caller). */
typedef struct pyl_circ_level {
  double* branch;         /* f64 [n_par][5][2] {first centre or held value, step} */
  int32_t* counts;        /* i32 [n_par][5]: children per parameter, -1 = held (crackle) */
  double* cr_first;       /* f64 [n_par] first crackle centre of an expansion */
  int32_t* cr_num;        /* i32 [n_par] crackle children of an expansion (0: no split) */
  int64_t* off1;          /* i64 [n_par + 1] exclusive prefix of held children per parent */
  int32_t* s1_parent;     /* i32 [n_s1] */
  uint8_t* s1_expand;     /* u8  [n_s1] */
  int64_t* s1_pos;        /* i64 [n_s1] final position of the held child's first output */
  int64_t* c_s1;          /* i64 [n_child] held child of every final child */
  int32_t* c_sub;         /* i32 [n_child] crackle index, -1 = held crackle */
  uint8_t* c_keep;        /* u8  [n_child] passes validation */
  int8_t* c_region;       /* i8  [n_child] 0 snap-circular, 1 crackle-circular, 2 Taylor */
  int64_t n_par, n_s1, n_child;
} pyl_circ_level;

#define PYL_REGION_SNAP 0
#define PYL_REGION_CRACKLE 1
#define PYL_REGION_TAYLOR 2

/* Writes branch, counts, cr_first, cr_num and nchild1_dev (i64 [n_par], held children per
 * parent). err_dev as for pyl_prune_branch; the key orders failures the way the reference
 * raises them: per batch, the non-crackle parameters of every parent, then crackle. */
int pyl_circ_branch(const pyl_prune_params* p, const pyl_circ_params* q,
                    const pyl_prune_coords* c, const double* leaves_dev,
                    const pyl_circ_level* L, int64_t* nchild1_dev, int64_t* err_dev,
                    void* stream);
/* Reads branch, counts, cr_num, off1; writes s1_parent, s1_expand. */
int pyl_circ_stage1(const pyl_prune_params* p, const pyl_circ_params* q,
                    const pyl_prune_coords* c, const double* leaves_dev,
                    const pyl_circ_level* L, void* stream);
/* Reads the stage-1 buffers and s1_pos; writes c_s1, c_sub, c_keep, c_region. */
int pyl_circ_expand(const pyl_prune_params* p, const pyl_circ_params* q,
                    const pyl_prune_coords* c, const double* leaves_dev,
                    const pyl_circ_level* L, void* stream);
/* valid_dev: i64 [n_valid] final-child ids that pass validation, increasing. Outputs per
 * validated child as for pyl_prune_children. */
int pyl_circ_children(const pyl_prune_params* p, const pyl_circ_params* q,
                      const pyl_prune_coords* c, const double* leaves_dev,
                      const float* folds_dev, const pyl_circ_level* L, const int64_t* valid_dev,
                      int64_t n_valid, const float* seg_fold_dev, int32_t* child_parent_dev,
                      int32_t* child_row_dev, int32_t* child_shift_dev, double* child_phase_dev,
                      float* child_score_dev, float* minmax_dev, void* stream);
/* surv_dev: i64 [n_surv] indices into valid_dev, increasing; has_taylor_dev: u8 [n_batches]
 * whether the batch's validated children include a Taylor cell. Writes the new tree. */
int pyl_circ_materialize(const pyl_prune_params* p, const pyl_circ_params* q,
                         const pyl_prune_coords* c, const double* leaves_dev,
                         const float* folds_dev, const pyl_circ_level* L,
                         const int64_t* valid_dev, const float* seg_fold_dev,
                         const int32_t* child_parent_dev, const int32_t* child_row_dev,
                         const int32_t* child_shift_dev, const double* child_phase_dev,
                         const float* child_score_dev, const uint8_t* has_taylor_dev,
                         const int64_t* surv_dev, int64_t n_surv, double* out_leaves_dev,
                         float* out_folds_dev, float* out_scores_dev,
                         int32_t* out_backtracks_dev, void* stream);
/* Complex-fold variants of the circular entry points (as for pyl_prune_*_complex). */
int pyl_circ_children_complex(const pyl_prune_params* p, const pyl_circ_params* q,
                              const pyl_prune_coords* c, const double* leaves_dev,
                              const float* folds_dev, const pyl_circ_level* L,
                              const int64_t* valid_dev, int64_t n_valid,
                              const float* seg_fold_dev, int32_t* child_parent_dev,
                              int32_t* child_row_dev, int32_t* child_shift_dev,
                              double* child_phase_dev, float* child_score_dev,
                              float* minmax_dev, void* stream);
int pyl_circ_materialize_complex(const pyl_prune_params* p, const pyl_circ_params* q,
                                 const pyl_prune_coords* c, const double* leaves_dev,
                                 const float* folds_dev, const pyl_circ_level* L,
                                 const int64_t* valid_dev, const float* seg_fold_dev,
                                 const int32_t* child_parent_dev, const int32_t* child_row_dev,
                                 const int32_t* child_shift_dev, const double* child_phase_dev,
                                 const float* child_score_dev, const uint8_t* has_taylor_dev,
                                 const int64_t* surv_dev, int64_t n_surv,
                                 double* out_leaves_dev, float* out_folds_dev,
                                 float* out_scores_dev, int32_t* out_backtracks_dev,
                                 void* stream);
int pyl_circ_ascend_complex(const pyl_prune_params* p, const pyl_circ_params* q,
                            const double* leaves_dev, int64_t n, const float* dyp_fold_dev,
                            int64_t ncoords_ffa, const int64_t* seg_idx, const double* seg_t0,
                            int32_t nseg_used, double t0_cur, float* out_folds_dev,
                            float* out_scores_dev, void* stream);
/* As pyl_prune_ascend, each survivor propagated with its region's propagator. */
int pyl_circ_ascend(const pyl_prune_params* p, const pyl_circ_params* q,
                    const double* leaves_dev, int64_t n, const float* dyp_fold_dev,
                    int64_t ncoords_ffa, const int64_t* seg_idx, const double* seg_t0,
                    int32_t nseg_used, double t0_cur, float* out_folds_dev,
                    float* out_scores_dev, void* stream);

/* Sync-free level loop in the circular basis: pyl_ep_levels for prune_poly_order 5 (same
 * records, statuses, tree buffers and history; flags: PYL_EP_COMPLEX). The per-batch
 * child positions, the validation compaction and the per-batch Taylor flags that the
 * per-level path computes between its calls are computed on the device; C holds their
 * buffers (n_child below = L->cap_child, n_par = L->cap_par). */
typedef struct pyl_circ_loop {
  double* cr_first;    /* f64 [n_par] */
  int32_t* cr_num;     /* i32 [n_par] */
  int64_t* off1;       /* i64 [n_par + 1] */
  int32_t* s1_parent;  /* i32 [n_child] */
  uint8_t* s1_expand;  /* u8  [n_child] */
  int64_t* s1_pos;     /* i64 [n_child] */
  int64_t* held_off;   /* i64 [n_child + 1]: packed prefix sums over held children */
  int64_t* c_s1;       /* i64 [n_child] */
  int32_t* c_sub;      /* i32 [n_child] */
  uint8_t* c_keep;     /* u8  [n_child] */
  int8_t* c_region;    /* i8  [n_child] */
  int64_t* valid;      /* i64 [n_child] */
  int64_t* batch_tot;  /* i64 [n_par]: children per batch of parents */
  int64_t* batch_nk;   /* i64 [n_par]: held children per batch that do not expand */
  int64_t* batch_off;  /* i64 [n_par]: first child of every batch */
  uint8_t* has_taylor; /* u8  [n_par] */
} pyl_circ_loop;

int pyl_circ_levels(const pyl_prune_params* p, const pyl_circ_params* q, int32_t flags,
                    const pyl_prune_coords* coords, const double* thresholds,
                    const int64_t* seg_idx, int32_t nlev, const float* dyp_fold_dev,
                    int64_t seg_stride, const double* in_leaves_dev, const float* in_folds_dev,
                    int64_t n_in, int64_t hist_row0, int32_t first_out, const pyl_ep_loop* L,
                    const pyl_circ_loop* C, void* stream);

/* A forest: ntree searches of one configuration (same grids, reference segment, coords
 * and thresholds; one series each) through one run of the circular level loop. The
 * reference fans the searches of a survey out over processes (prune.py:892-922, one
 * Pruning per task); here every launch of a level serves all of them.
 * A level's parents are the trees' parents one tree after the other, and so are its held
 * children, its batches (each tree's own batches of batch_size parents), its children,
 * survivors and kept rows. Per level l, tab + l * PYL_EP_FOREST_TAB * (ntree + 1) holds,
 * for tree t, the offsets (each [ntree + 1], non-decreasing) of its parents (PO), batches
 * (BO), children (CB), scored children (VB), survivors (SO) and kept rows (KO: the budget's,
 * else SO) in those sequences; row l + 1's PO is level l's KO. Level records hold the
 * forest's totals; minmax holds every tree's children's score min/max per level.
 * Each tree's results equal those of pyl_circ_levels on that tree alone, bit for bit: its
 * backtracks hold parent indices within its own parents. Besides the single loop's stops
 * (PYL_EP_ST_OVERFLOW: the forest's survivor buffer L->max_sugg), a level stops with
 * PYL_EP_ST_TREE_OVERFLOW when any tree's survivors exceed F->max_sugg or its held children
 * times branch_max reach 2^31 (the caller then runs the trees one by one from that level)
 * and with PYL_EP_ST_BUDGET when a level enqueued without the cut finds any tree's
 * survivors above the budget.
 * The FFA folds of the trees are stacked: tree t's segment rows of level l start at
 * dyp_fold_dev + t * tree_stride + seg_idx[l] * seg_stride floats. L->max_sugg and
 * L->cap_* are the forest's capacities (all trees together); n_in: host [ntree], the
 * input trees' sizes (their rows one after the other in in_*). With L->budget, every tree
 * keeps at most budget->budget survivors per level, ranked and cut per tree (the network's
 * level statistics are per tree); budget scratch is sized for L->max_sugg rows. A
 * PYL_BUDGET_SCORE budget is cut on the survivors' scores (copied to budget->keys) before
 * anything is materialised: only the kept survivors are, into tree buffer
 * (first_out + l) % 2 as without a budget (budget->bt_tmp unused); the network's budgets
 * use the buffers as pyl_ep_loop describes. */
#define PYL_EP_FOREST_MAX 64
#define PYL_EP_FOREST_TAB 6
#define PYL_EP_TAB_PO 0
#define PYL_EP_TAB_BO 1
#define PYL_EP_TAB_CB 2
#define PYL_EP_TAB_VB 3
#define PYL_EP_TAB_SO 4
#define PYL_EP_TAB_KO 5
typedef struct pyl_ep_forest {
  int32_t ntree;       /* trees, 1..PYL_EP_FOREST_MAX */
  int32_t reserved;
  int64_t tree_stride; /* floats between consecutive trees' FFA folds */
  int64_t max_sugg;    /* survivor buffer of each tree (the reference's max_sugg) */
  int64_t* tab;        /* i64 [nlev + 1][PYL_EP_FOREST_TAB][ntree + 1] (written) */
  float* minmax;       /* f32 [nlev][ntree][2] (written) */
  /* per-tree budget state (with L->budget) */
  uint32_t* sel;       /* u32 [ntree][PYL_BUDGET_SEL_WORDS] */
  double* stats;       /* f64 [ntree][PYL_BUDGET_STATS] */
  uint32_t* shist;     /* u32 [ntree][2 * PYL_BUDGET_HIST_BINS] */
  int64_t* part;       /* i64 [ntree][PYL_EP_SCAN_BLOCKS] */
  int64_t* kcount;     /* i64 [ntree] */
} pyl_ep_forest;

int pyl_circ_forest_levels(const pyl_prune_params* p, const pyl_circ_params* q, int32_t flags,
                           const pyl_prune_coords* coords, const double* thresholds,
                           const int64_t* seg_idx, int32_t nlev, const float* dyp_fold_dev,
                           int64_t seg_stride, const double* in_leaves_dev,
                           const float* in_folds_dev, const int64_t* n_in, int64_t hist_row0,
                           int32_t first_out, const pyl_ep_loop* L, const pyl_circ_loop* C,
                           const pyl_ep_forest* F, void* stream);

/* pyl_circ_ascend(_complex) of a forest's final trees in one launch: tree t's leaves are
 * rows tree_off[t]..tree_off[t + 1] (host [ntree + 1], tree_off[0] = 0) of leaves_dev and
 * its FFA fold starts dyp_fold_dev + t * tree_stride floats; out_* as pyl_circ_ascend for
 * the tree_off[ntree] rows. flags: PYL_EP_COMPLEX. Replaces one Pruning's ascend
 * (prune.py:671-693) per task. */
int pyl_circ_forest_ascend(const pyl_prune_params* p, const pyl_circ_params* q, int32_t flags,
                           const double* leaves_dev, const int64_t* tree_off, int32_t ntree,
                           const float* dyp_fold_dev, int64_t tree_stride, int64_t ncoords,
                           const int64_t* seg_idx, const double* seg_t0, int32_t nseg,
                           double t0_cur, float* out_folds_dev, float* out_scores_dev,
                           void* stream);

/* ------------------------------------------------------------------------------------
 * Learned branch ranking (no counterpart in the reference). A small 1-D CNN scores EP
 * survivors from their folds and 8 parameter features so that a budgeted, best-first
 * search keeps the most promising ones; the exhaustive EP above stays the judge.
 * Network: profile e/sqrt(v) (irfft first for complex folds) standardised over the bins;
 * three circular Conv1d layers (32 channels, kernel 5, ReLU); max over bins; with the
 * features, a 32-unit ReLU layer and one logit. Parameters are one flat FP32 vector of
 * pyl_cnn_num_params() values: W1[32][1][5], b1[32], W2[32][32][5], b2[32], W3[32][32][5],
 * b3[32], Wf1[32][40], bf1[32], Wf2[32], bf2[1]. nbins must be 32, 64 or 128.
 * Loss: sum_n w_pos y softplus(-logit) + w_neg (1 - y) softplus(logit).
 * ------------------------------------------------------------------------------------ */
int pyl_cnn_num_params(void);
/* Floats of training cache per sample (-1 for an unsupported nbins). */
int64_t pyl_cnn_cache_floats(int nbins);
/* folds_dev: f32 [n][2][nbins], or complex64 [n][2][nbins/2 + 1] with complex_folds = 1;
 * feats_dev f32 [n][8]; logits_dev f32 [n]. cache_dev (nullable): f32
 * [n][pyl_cnn_cache_floats(nbins)], written for pyl_cnn_backward. */
int pyl_cnn_forward(const float* params_dev, const float* folds_dev, int64_t n, int nbins,
                    int complex_folds, const float* feats_dev, float* logits_dev,
                    float* cache_dev, void* stream);
/* Gradient of the loss (summed over the n samples) into grad_dev [num_params], from the
 * cache and logits of pyl_cnn_forward; labels_dev f32 [n] in {0, 1}. work_dev holds the
 * per-block partial sums and, past 32 blocks, one group sum per 32 blocks (at least
 * num_params floats; more blocks with more room). */
int pyl_cnn_backward(const float* params_dev, const float* cache_dev, const float* logits_dev,
                     const float* labels_dev, int64_t n, int nbins, float w_pos, float w_neg,
                     float* work_dev, int64_t work_floats, float* grad_dev, void* stream);
/* One Adam step (bias-corrected; step counts from 1). */
/* Inference only: the forward pass with the two 32 -> 32 convolutions on MFMA (BF16
 * operands, FP32 accumulation; the profile, the first convolution and the head in FP32).
 * Arguments as pyl_cnn_forward without the training cache. */
int pyl_cnn_forward_bf16(const float* params, const float* folds, int64_t n, int nbins,
                         int complex_folds, const float* feats, float* logits, void* stream);
int pyl_cnn_adam(float* params_dev, const float* grad_dev, float* m_dev, float* v_dev,
                 int64_t nparams, float lr, float beta1, float beta2, float eps, int step,
                 void* stream);

/* ------------------------------------------------------------------------------------
 * Monte-Carlo threshold design (detection/thresholding.py).
 *
 * pyl_thresh_simulate replaces, per task, one simulate_folds (:45-95) of a fold set
 * followed by its scoring (prune_folds :98-134 / fold_max_scores :483-499: folds / sqrt(var)
 * in FP64 rounded to FP32, boxcar S/N of every width with stdnoise 1, maximum over widths)
 * and either the survivor counts of a list of thresholds (PYL_SIM_COUNT) or the survivors
 * themselves (PYL_SIM_COMPACT: prune_folds's `score > thr` selection, stable order).
 *
 * Row i of a task's simulated folds (i < n_sim) is
 *   f32( f64( src[src_row + i % n_prev][j] + f32(sqrt(var_add) * z) ) + prof_scaled[hyp][j] )
 * with src rows all zero when src_row < 0 and z = normal deviate (seed, stream, row i,
 * bin j) of pyl_rng_normal. A task simulated twice with the same stream gives the same
 * folds, which is how survivors are materialised after the caller chose which to keep.
 * Tasks run concurrently (one wavefront each); dst rows of different tasks must not
 * overlap. nbins in {16, 32, 64}; widths as for pyl_score_boxcar; thr_count <= 256 (the
 * kernel counts at most the first 256 thresholds of a task).
 * ------------------------------------------------------------------------------------ */
#define PYL_SIM_COUNT 0
#define PYL_SIM_COMPACT 1
#define PYL_SIM_NO_NOISE 1 /* PylSimTask.flags */

typedef struct {
  int64_t src_row;   /* first source row in src_dev; < 0: the n_prev source rows are zero */
  int64_t dst_row;   /* PYL_SIM_COMPACT: first output row in dst_dev */
  int64_t score_off; /* first score written to scores_dev (when scores_dev is not NULL) */
  int64_t count_off; /* first count written to counts_dev */
  uint64_t stream;   /* RNG stream of this simulation */
  double var_norm;   /* variance the simulated folds are normalised by (var_cur + var_add) */
  double thr;        /* PYL_SIM_COMPACT: rows with score > thr are kept */
  int32_t n_prev;    /* source rows (> 0) */
  int32_t n_sim;     /* rows simulated (> 0; simulate_folds: n_prev * ceil(ntrials / n_prev)) */
  int32_t hyp;       /* row of prof_scaled_dev added: 0 (H0) or 1 (H1) */
  int32_t thr_first; /* PYL_SIM_COUNT: first of this task's thresholds in thr_dev */
  int32_t thr_count; /* PYL_SIM_COUNT: number of thresholds */
  int32_t flags;     /* PYL_SIM_NO_NOISE: row i is src row i % n_prev itself (scoring only) */
} PylSimTask;

/* tasks_dev: PylSimTask [ntasks]; prof_scaled_dev: f64 [2][nbins] (profile * bias of H0 and
 * H1); thr_dev: f64 thresholds. PYL_SIM_COUNT: counts_dev[count_off + t] = #(score >
 * thr_dev[thr_first + t]) and, if scores_dev, scores_dev[score_off + i] = score of row i.
 * PYL_SIM_COMPACT: survivors to dst_dev rows dst_row.., their scores to scores_dev
 * [score_off..] (if not NULL), their number to counts_dev[count_off]. */
int pyl_thresh_simulate(const float* src_dev, const PylSimTask* tasks_dev, int32_t ntasks,
                        int32_t nbins, const double* prof_scaled_dev, double var_add,
                        uint64_t seed, const int64_t* widths, int32_t nw, const double* thr_dev,
                        int32_t mode, float* dst_dev, float* scores_dev, int32_t* counts_dev,
                        void* stream);

/* Normal deviates (no reference counterpart: the reference draws numpy Generator.normal,
 * whose bits are not reproducible on a GPU): out_dev[i][j] = f32(std * z(seed, stream, i,
 * j)) for i < n_rows, j < row_len (row_len a multiple of 4). z: Philox4x32-10 of counter
 * {j / 4, i, stream} and key seed, Box-Muller with an IEEE-basic-operation ln and sincos. */
int pyl_rng_normal(uint64_t seed, uint64_t stream, int64_t n_rows, int32_t row_len, double std,
                   float* out_dev, void* stream_);

/* Rows gather: dst_dev[r] = src_dev[idx[r]] (rows of row_len floats); idx_dev int64 [n]. */
int pyl_rows_gather(const float* src_dev, const int64_t* idx_dev, int64_t n, int32_t row_len,
                    float* dst_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * Seeded signal simulation.
 * Replaces simulation/pulse.py:310-406 PulseSignalConfig.generate (its sample-level
 * parts: generate_pulse_template :39-66, the noise draw, and the two brutefold calls of
 * the calibration, core/fold.py:57-105) with the proper time of simulation/modulate.py
 * (DerivativeModulating :34-54, DerivativeSeriesModulating :92-97, CircularModulating
 * :147-150, CircularT0Modulating :222-226, KeplerianModulating :240-250 with
 * kepler.py:64-140).
 *
 * pyl_modulation.p by kind:
 *   DERIVATIVE         {shift, vel, acc, jerk, snap}
 *   DERIVATIVE_SERIES  coeffs[k] = d_k / k!, k < n_coeffs
 *   CIRCULAR           {p_orb, psi, x_orb}
 *   CIRCULAR_T0        {a, p_orb, t0}
 *   KEPLERIAN          {p_orb, ecc, phi, a, om, inc}
 * ------------------------------------------------------------------------------------ */
#define PYL_MOD_DERIVATIVE 0
#define PYL_MOD_DERIVATIVE_SERIES 1
#define PYL_MOD_CIRCULAR 2
#define PYL_MOD_CIRCULAR_T0 3
#define PYL_MOD_KEPLERIAN 4
#define PYL_MOD_MAX_COEFFS 24
#define PYL_SIM_MAX_SAMPS (((int64_t)1) << 27)

typedef struct {
  int32_t kind;     /* PYL_MOD_* */
  int32_t n_coeffs; /* DERIVATIVE_SERIES: order + 1 */
  double t_ref;     /* reference epoch of the modulation (mod_tref) */
  double p[6];
  double coeffs[PYL_MOD_MAX_COEFFS];
} pyl_modulation;

/* For every sample i < nsamps (t = i dt, t' = proper time): tmpl_dev[i] = the pulse
 * template (the CDF difference between the phases of t' and t' + dt, interpolated in
 * cdf_lut_dev, f32 [ngrid + 1] with the wrap element last); noise_dev[i] = f32(z) with z the
 * normal deviate of pyl_rng_normal(seed, stream_id) row 0, column i; and the calibration
 * folds at f = 1/period into folds_dev int64 [3][nbins]: fixed-point sums (value * 2^32,
 * rounded to nearest even) of the noise and of the template, and the sample count, per bin
 * int(f32(phase * nbins) + 0.5f) (nbins -> 0). nsamps <= PYL_SIM_MAX_SAMPS keeps the sums
 * inside int64. Deterministic: the same arguments give the same bits. */
int pyl_sim_series(const pyl_modulation* mod, int64_t nsamps, double dt, double period,
                   const float* cdf_lut_dev, int32_t ngrid, uint64_t seed, uint64_t stream_id,
                   int32_t nbins, float* noise_dev, float* tmpl_dev, int64_t* folds_dev,
                   void* stream);
/* ts_e_dev[i] = f32(noise_scale * noise[i]) + f32(signal_scale * tmpl[i]) and
 * ts_v_dev[i] = variance, in FP32 (pulse.py:402-405). */
int pyl_sim_combine(const float* noise_dev, const float* tmpl_dev, int64_t nsamps,
                    float noise_scale, float signal_scale, float variance, float* ts_e_dev,
                    float* ts_v_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * Boxcar S/N threshold screen on the matrix cores.
 * The boxcar scores of detection/scoring.py:99-165 (snr_score_batch_func :249-282, max
 * over widths, stdnoise 1) as a product of profiles and 0/1 box templates (the templates
 * of scoring.py:383-471), BF16 operands, FP32 accumulation, with exact rescoring:
 * out[i] = the exact score of pyl_score_boxcar (PYL_SCORE_MAX) for every profile whose
 * screened upper bound reaches `threshold` (so out >= threshold selects the same profiles
 * with the same scores, bit for bit), an upper bound below `threshold` for the others.
 * folds_dev: f32 [n][2][nbins] (e, v), nbins 32 or 64, 1..16 widths; nrescored_dev
 * (nullable, zeroed by the caller) is increased by the number of profiles scored exactly.
 * ------------------------------------------------------------------------------------ */
int pyl_score_boxcar_screen(const float* folds_dev, int64_t n, int32_t nbins,
                            const int64_t* widths, int32_t nw, double threshold, float* out_dev,
                            unsigned long long* nrescored_dev, void* stream);

/* ------------------------------------------------------------------------------------
 * Branching patterns on the device.
 * Replaces the per-frequency loops of core/taylor.py:414-499 generate_bp_poly_taylor
 * (kind 0) and core/circular.py:905-1010 generate_bp_circ_taylor (kind 1, n_params 5):
 * nbr_out (host, f64 [nlev][nf]) = each final-grid frequency's branch count at every
 * pruning level, from its cell widths [d_k .. d_1] (dinit, the last one scaled by C/f0).
 * Per level l (host, FP64): dpf [nlev][n_params] the target step before the C/f0 scaling,
 * fac [nlev][n_params] the bin shift per unit width (span powers / factorials, /2^k with
 * Chebyshev coarsening), tmat [nlev][n_params + 1][n_params + 1] the Taylor shift matrix to
 * the next level's epoch (moving grid; tiling 0 aggressive, 1 conservative, 2
 * quadrature). The caller forms the weighted averages. Synchronises the stream.
 * ------------------------------------------------------------------------------------ */
int pyl_bp_counts(int32_t kind, int32_t n_params, int32_t nlev, int64_t nf, const double* f0,
                  const double* dinit, const double* dpf, const double* fac, const double* tmat,
                  int32_t moving, int32_t tiling, double eta, double* nbr_out, void* stream);

/* ------------------------------------------------------------------------------------
 * Time-series conditioning.
 * Replaces the sigpyproc calls of io/timeseries.py:197-235 TimeSeries.from_tim:
 * deredden(method="median", window, fast=True) and normalise(loc_method="mean",
 * scale_method) (sigpyproc is an unpinned git dependency of the reference, absent here).
 *
 * pyl_ts_quantiles: for each segment s of seg_len consecutive samples (the last may be
 *   shorter), out_dev[s][t] = numpy's "linear" quantile q[t] (nq = 1 or 2) of the
 *   segment's values, or of |x - center_dev[s]| when center_dev is not NULL: the order
 *   statistics at floor and ceil of q (m - 1), found exactly by radix select, interpolated
 *   in FP64. Deterministic.
 * pyl_ts_deredden_median: medians_dev [ceil(n / window)] = the medians of consecutive
 *   blocks of `window` samples; out_dev = x minus those medians interpolated linearly
 *   between the block centres (held constant beyond the first and last centre).
 * pyl_ts_normalise: out_dev = (x - mean) / scale in FP64 stored as FP32, scale =
 *   PYL_TS_SCALE_STD the standard deviation, PYL_TS_SCALE_IQR the interquartile range /
 *   1.3489795003921634, PYL_TS_SCALE_MAD the median absolute deviation x
 *   1.482602218505602; loc_scale (host, 2 doubles) receives (mean, scale). Synchronises
 *   the stream (the statistics come back to the host). A zero scale is a ValueError.
 * ------------------------------------------------------------------------------------ */
#define PYL_TS_SCALE_STD 0
#define PYL_TS_SCALE_IQR 1
#define PYL_TS_SCALE_MAD 2
int pyl_ts_quantiles(const float* x_dev, int64_t n, int64_t seg_len, const double* q,
                     int32_t nq, const float* center_dev, float* out_dev, void* stream);
int pyl_ts_deredden_median(const float* x_dev, int64_t n, int64_t window, float* medians_dev,
                           float* out_dev, void* stream);
int pyl_ts_normalise(const float* x_dev, int64_t n, int32_t scale_method, float* out_dev,
                     double* loc_scale, void* stream);

/* ------------------------------------------------------------------------------------
 * Bandwidth probes of known size (no reference counterpart; measurement only):
 * pyl_hbm_read loads n_float4 16-byte elements once each (one float per block written to
 * partial_dev [nblocks]); pyl_hbm_copy copies n_float4 elements. Used to measure the
 * attainable HBM bandwidth and to scale the FETCH_SIZE / WRITE_SIZE counters to bytes.
 * ------------------------------------------------------------------------------------ */
int pyl_hbm_read(const float* src_dev, int64_t n_float4, float* partial_dev, int32_t nblocks,
                 void* stream);
int pyl_hbm_copy(const float* src_dev, float* dst_dev, int64_t n_float4, int32_t nblocks,
                 void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PYLOKI_HIP_H */
