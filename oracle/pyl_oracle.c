/*
 * CPU oracle for the pyloki search hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the thing measured or shipped: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it. The product
 * path (pyloki_amd) calls the HIP library and fails loudly without it.
 *
 * It is a strict restatement of the reference's numba kernels in their operation
 * order: no FMA contraction (-ffp-contract=off), IEEE division and sqrt, sequential
 * sums in the reference's loop order. Each function names the reference lines it
 * follows (paths relative to pyloki's src/pyloki/).
 *
 * Pinning (what checks each function beyond its own restatement) is listed in the
 * header comment of oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_C_VAL 299792458.0
#define ORC_FLOAT_EPSILON 1e-6

/* Threads of the parallel loops below. Every output element is computed by one thread
 * in the reference's order, so results do not depend on the count. The parallel loops
 * are the reference's prange loops: segments in the brute fold (fold.py:161), parameter
 * sets in unify_fold (dyn_pffa.py:465) and profiles in the scorer (scoring.py:151). */
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* utils/psr_utils.py:12-46 get_phase_idx */
double orc_phase_idx(double proper_time, double freq, int64_t nbins, double delay) {
    double phase = (proper_time - delay) * freq;
    phase -= floor(phase);
    double iphase = phase * (double)nbins;
    if (iphase >= (double)nbins) iphase = 0.0;
    return iphase;
}

/* utils/psr_utils.py:49-56 get_phase_idx_int: int(f32(x) + f32(0.5)), nbins -> 0 */
int64_t orc_shift_int(double iphase, int64_t nbins) {
    float xf = (float)iphase;
    float r = xf + 0.5f;
    int64_t ibin = (int64_t)r;
    if (ibin == nbins) ibin = 0;
    return ibin;
}

int64_t orc_phase_idx_int(double proper_time, double freq, int64_t nbins, double delay) {
    return orc_shift_int(orc_phase_idx(proper_time, freq, nbins, delay), nbins);
}

/* x**n by binary powering: what numba's fastmath lowering of a float raised to an
 * integer array evaluates (LLVM rewrites pow(x, sitofp(n)) to llvm.powi). */
double orc_powi(double a, int b) {
    int recip = b < 0;
    if (recip) b = -b;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

static double orc_fact(int n) {
    double r = 1.0;
    for (int i = 2; i <= n; ++i) r *= (double)i;
    return r;
}

/* transforms.py:10-47 shift_taylor_params for one vector [d_k..d_0] (n entries):
 * out[i] = sum_{j=0..n-1} vec[j] * T[i][j], T[i][j] = dt**(i-j)/(i-j)! for j<=i else 0
 * (orc_powi), summed in ascending j. */
void orc_shift_taylor_params(const double* vec, int n, double dt, double* out) {
    for (int i = 0; i < n; ++i) {
        double acc = 0.0;
        for (int j = 0; j < n; ++j) {
            double t = 0.0;
            if (j <= i) t = orc_powi(dt, i - j) / orc_fact(i - j);
            acc += vec[j] * t;
        }
        out[i] = acc;
    }
}

/* utils/psr_utils.py:466-514 get_nearest_indices_analytical (one parameter) */
int64_t orc_nearest_idx(double val, int64_t n_grid, double vmin, double vmax) {
    if (n_grid <= 1) return 0;
    double raw_idx = (double)n_grid * (val - vmin) / (vmax - vmin);
    int64_t idx = (int64_t)(raw_idx + ORC_FLOAT_EPSILON);
    if (idx < 0) idx = 0;
    else if (idx >= n_grid) idx = n_grid - 1;
    return idx;
}

/* core/fold.py:617-671 ffa_taylor_resolve (+ transforms.py:140-170 shift_taylor_params_d_f).
 * pset: P values [.., accel, freq]; limits: P x 2. Writes P indices and the phase. */
void orc_ffa_resolve(int P, const double* pset, const int64_t* counts_prev,
                     const double* limits, int level, int latter, double tseg_brute,
                     int64_t nbins, int64_t* idx_out, double* phase_out) {
    double pnew[8];
    double delta_t, delay;
    if (P == 1) {
        delta_t = (double)((int64_t)latter * ((int64_t)1 << (level - 1))) * tseg_brute;
        pnew[0] = pset[0];
        delay = 0.0;
    } else {
        delta_t = ((latter - 0.5) * (double)((int64_t)1 << (level - 1))) * tseg_brute;
        double tv[8], tn[8];
        for (int k = 0; k < P + 1; ++k) tv[k] = 0.0;
        for (int k = 0; k < P - 1; ++k) tv[k] = pset[k];
        orc_shift_taylor_params(tv, P + 1, delta_t, tn);
        for (int k = 0; k < P - 1; ++k) pnew[k] = tn[k];
        pnew[P - 1] = pset[P - 1] * (1.0 - tn[P - 1] / ORC_C_VAL);
        delay = tn[P] / ORC_C_VAL;
    }
    *phase_out = orc_phase_idx(delta_t, pset[P - 1], nbins, delay);
    for (int k = 0; k < P; ++k)
        idx_out[k] = orc_nearest_idx(pnew[k], counts_prev[k], limits[2 * k], limits[2 * k + 1]);
}

/* core/fold.py:108-171 brutefold_start: out[nseg][nfreq][2][nbins], sequential
 * per-bin FP32 sums in sample order. proper_time = i*tsamp - t_ref. */
void orc_brutefold(const float* ts_e, const float* ts_v, int64_t nsamps,
                   const double* freqs, int64_t nfreq, int64_t seg_len, int64_t nbins,
                   double tsamp, double t_ref, float* out) {
    int64_t nseg = (nsamps + seg_len - 1) / seg_len;
    int32_t* pmap = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nfreq * seg_len));
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nfreq; ++f)
        for (int64_t i = 0; i < seg_len; ++i) {
            double pt = (double)i * tsamp - t_ref;
            pmap[f * seg_len + i] = (int32_t)orc_phase_idx_int(pt, freqs[f], nbins, 0.0);
        }
    memset(out, 0, sizeof(float) * (size_t)(nseg * nfreq * 2 * nbins));
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < nseg; ++s) {
        int64_t start = s * seg_len;
        int64_t len = nsamps - start < seg_len ? nsamps - start : seg_len;
        for (int64_t f = 0; f < nfreq; ++f) {
            float* o = out + (s * nfreq + f) * 2 * nbins;
            for (int64_t i = 0; i < len; ++i) {
                int32_t b = pmap[f * seg_len + i];
                o[b] += ts_e[start + i];
                o[nbins + b] += ts_v[start + i];
            }
        }
    }
    free(pmap);
}

/* core/common.py:304-324 shift_add for one (tail, head) pair of [2][nbins] rows */
void orc_shift_add(const float* tail, const float* head, double sh_t, double sh_h,
                   int64_t nbins, float* res) {
    int64_t st = orc_shift_int(sh_t, nbins);
    int64_t shh = orc_shift_int(sh_h, nbins);
    for (int64_t j = 0; j < nbins; ++j) {
        int64_t i1 = ((j - st) % nbins + nbins) % nbins;
        int64_t i2 = ((j - shh) % nbins + nbins) % nbins;
        res[j] = tail[i1] + head[i2];
        res[nbins + j] = tail[nbins + i1] + head[nbins + i2];
    }
}

/* dynamic/dyn_pffa.py:435-489 unify_fold. fold_in: [nseg_in][ncoords_prev][2][nbins];
 * params: P concatenated current-level grids (lengths counts_cur); cartesian product
 * with the last parameter fastest (np_utils.cartesian_prod_st). Optionally exports the
 * resolved table: rows[ncoords][2] (flattened prev index of tail, head) and
 * phases[ncoords][2]. */
void orc_unify_fold(const float* fold_in, int64_t nseg_in, int P, const int64_t* counts_prev,
                    const double* params, const int64_t* counts_cur, const double* limits,
                    int level, double tseg_brute, int64_t nbins, float* fold_out,
                    int64_t* rows_out, double* phases_out) {
    int64_t ncur = 1, nprev = 1;
    for (int k = 0; k < P; ++k) { ncur *= counts_cur[k]; nprev *= counts_prev[k]; }
    int64_t offs[8];
    offs[0] = 0;
    for (int k = 1; k < P; ++k) offs[k] = offs[k - 1] + counts_cur[k - 1];
    int64_t nseg_out = nseg_in / 2;
    int64_t rowlen = 2 * nbins;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ncur; ++j) {
        double pset[8];
        int64_t r = j;
        for (int k = P - 1; k >= 0; --k) {
            int64_t ik = r % counts_cur[k];
            r /= counts_cur[k];
            pset[k] = params[offs[k] + ik];
        }
        int64_t row[2];
        double ph[2];
        for (int latter = 0; latter < 2; ++latter) {
            int64_t idx[8];
            orc_ffa_resolve(P, pset, counts_prev, limits, level, latter, tseg_brute, nbins,
                            idx, &ph[latter]);
            int64_t flat = 0;
            for (int k = 0; k < P; ++k) flat = flat * counts_prev[k] + idx[k];
            row[latter] = flat;
        }
        if (rows_out) { rows_out[2 * j] = row[0]; rows_out[2 * j + 1] = row[1]; }
        if (phases_out) { phases_out[2 * j] = ph[0]; phases_out[2 * j + 1] = ph[1]; }
        for (int64_t q = 0; q < nseg_out; ++q) {
            const float* tail = fold_in + ((2 * q) * nprev + row[0]) * rowlen;
            const float* head = fold_in + ((2 * q + 1) * nprev + row[1]) * rowlen;
            orc_shift_add(tail, head, ph[0], ph[1], nbins, fold_out + (q * ncur + j) * rowlen);
        }
    }
}

/* detection/scoring.py:42-71 circular_prefix_sum for nsum <= 2*nbins (one wrap) and
 * the general multi-wrap form. */
static void orc_circular_prefix_sum(const float* x, int64_t nbins, float* out, int64_t nsum) {
    if (nbins == 0 || nsum == 0) return;
    out[0] = x[0];
    int64_t lim = nbins < nsum ? nbins : nsum;
    for (int64_t i = 1; i < lim; ++i) out[i] = out[i - 1] + x[i];
    if (nsum <= nbins) return;
    float last_sum = out[nbins - 1];
    int64_t n_wraps = nsum / nbins;
    int64_t extra = nsum % nbins;
    for (int64_t i = 1; i < n_wraps; ++i)
        for (int64_t j = 0; j < nbins; ++j)
            out[i * nbins + j] = (float)((double)out[j] + (double)i * (double)last_sum);
    for (int64_t j = 0; j < extra; ++j)
        out[n_wraps * nbins + j] = (float)((double)out[j] + (double)n_wraps * (double)last_sum);
}

/* detection/scoring.py:74-84 diff_max */
static float orc_diff_max(const float* x, const float* y, int64_t n) {
    float m = -FLT_MAX;
    for (int64_t i = 0; i < n; ++i) {
        float d = x[i] - y[i];
        if (d > m) m = d;
    }
    return m;
}

/* detection/scoring.py:99-140 boxcar_snr_1d over n profiles (boxcar_snr_2d, :143-153).
 * norm: [n][nbins]; out: [n][nw]. */
void orc_boxcar_snr_2d(const float* norm, int64_t n, int64_t nbins, const int64_t* widths,
                       int64_t nw, double stdnoise, float* out) {
    int64_t wmax = 0;
    for (int64_t k = 0; k < nw; ++k) if (widths[k] > wmax) wmax = widths[k];
    double inv_std = 1.0 / stdnoise;
#pragma omp parallel
    {
    float* ps = (float*)malloc(sizeof(float) * (size_t)(nbins + wmax));
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const float* x = norm + p * nbins;
        orc_circular_prefix_sum(x, nbins, ps, nbins + wmax);
        float total = ps[nbins - 1];
        for (int64_t k = 0; k < nw; ++k) {
            int64_t w = widths[k];
            float size_w = (float)(nbins - w);
            double height = sqrt((double)size_w / (double)(nbins * w));
            double b = (double)w * height / (double)size_w;
            float dmax = orc_diff_max(ps + w, ps, nbins);
            double t1 = (height + b) * (double)dmax;
            double t2 = b * (double)total;
            out[p * nw + k] = (float)((t1 - t2) * inv_std);
        }
    }
    free(ps);
    }
}

/* detection/scoring.py:249-282 snr_score_batch_func: folds [n][2][nbins] -> max S/N */
void orc_snr_score_batch(const float* folds, int64_t n, int64_t nbins, const int64_t* widths,
                         int64_t nw, float* out) {
    int64_t wmax = 0;
    for (int64_t k = 0; k < nw; ++k) if (widths[k] > wmax) wmax = widths[k];
#pragma omp parallel
    {
    float* ps = (float*)malloc(sizeof(float) * (size_t)(nbins + wmax));
    float* fn = (float*)malloc(sizeof(float) * (size_t)nbins);
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const float* e = folds + p * 2 * nbins;
        const float* v = e + nbins;
        for (int64_t j = 0; j < nbins; ++j) fn[j] = e[j] / sqrtf(v[j]);
        orc_circular_prefix_sum(fn, nbins, ps, nbins + wmax);
        float total = ps[nbins - 1];
        double max_snr = -(double)FLT_MAX;
        for (int64_t k = 0; k < nw; ++k) {
            int64_t w = widths[k];
            int64_t size_w = nbins - w;
            double height = sqrt((double)size_w / (double)(nbins * w));
            double b = (double)w * height / (double)size_w;
            float dmax = orc_diff_max(ps + w, ps, nbins);
            double snr = (height + b) * (double)dmax - b * (double)total;
            if (snr > max_snr) max_snr = snr;
        }
        out[p] = (float)max_snr;
    }
    free(fn);
    free(ps);
    }
}

/* ffa.py:382-411 get_fold_norm (real path): e / sqrt(v) in FP32 */
void orc_fold_norm(const float* folds, int64_t n, int64_t nbins, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p)
        for (int64_t j = 0; j < nbins; ++j)
            out[p * nbins + j] = folds[p * 2 * nbins + j] / sqrtf(folds[p * 2 * nbins + nbins + j]);
}

/* ---- FP64 reference of the FFA (for the error of reduced-precision fold storage) ----
 * Same algorithm as orc_brutefold / orc_unify_fold / boxcar S/N, with every fold
 * accumulation, normalisation and prefix sum in FP64. Grid indices and shifts are the
 * same FP64 resolution as the FP32 path, so only the data arithmetic differs. */
void orc_brutefold_f64(const float* ts_e, const float* ts_v, int64_t nsamps,
                       const double* freqs, int64_t nfreq, int64_t seg_len, int64_t nbins,
                       double tsamp, double t_ref, double* out) {
    int64_t nseg = (nsamps + seg_len - 1) / seg_len;
    memset(out, 0, sizeof(double) * (size_t)(nseg * nfreq * 2 * nbins));
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < nfreq; ++f)
        for (int64_t i = 0; i < seg_len; ++i) {
            double pt = (double)i * tsamp - t_ref;
            int64_t b = orc_phase_idx_int(pt, freqs[f], nbins, 0.0);
            for (int64_t s = 0; s < nseg; ++s) {
                int64_t k = s * seg_len + i;
                if (k >= nsamps) continue;
                double* o = out + (s * nfreq + f) * 2 * nbins;
                o[b] += (double)ts_e[k];
                o[nbins + b] += (double)ts_v[k];
            }
        }
}

void orc_unify_fold_f64(const double* fold_in, int64_t nseg_in, int P, const int64_t* counts_prev,
                        const double* params, const int64_t* counts_cur, const double* limits,
                        int level, double tseg_brute, int64_t nbins, double* fold_out) {
    int64_t ncur = 1, nprev = 1;
    for (int k = 0; k < P; ++k) { ncur *= counts_cur[k]; nprev *= counts_prev[k]; }
    int64_t offs[8];
    offs[0] = 0;
    for (int k = 1; k < P; ++k) offs[k] = offs[k - 1] + counts_cur[k - 1];
    int64_t nseg_out = nseg_in / 2, rowlen = 2 * nbins;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ncur; ++j) {
        double pset[8];
        int64_t r = j;
        for (int k = P - 1; k >= 0; --k) {
            int64_t ik = r % counts_cur[k];
            r /= counts_cur[k];
            pset[k] = params[offs[k] + ik];
        }
        int64_t row[2], sh[2];
        for (int latter = 0; latter < 2; ++latter) {
            int64_t idx[8];
            double ph;
            orc_ffa_resolve(P, pset, counts_prev, limits, level, latter, tseg_brute, nbins, idx, &ph);
            int64_t flat = 0;
            for (int k = 0; k < P; ++k) flat = flat * counts_prev[k] + idx[k];
            row[latter] = flat;
            sh[latter] = orc_shift_int(ph, nbins);
        }
        for (int64_t q = 0; q < nseg_out; ++q) {
            const double* t = fold_in + ((2 * q) * nprev + row[0]) * rowlen;
            const double* h = fold_in + ((2 * q + 1) * nprev + row[1]) * rowlen;
            double* o = fold_out + (q * ncur + j) * rowlen;
            for (int64_t b = 0; b < nbins; ++b) {
                int64_t i1 = ((b - sh[0]) % nbins + nbins) % nbins;
                int64_t i2 = ((b - sh[1]) % nbins + nbins) % nbins;
                o[b] = t[i1] + h[i2];
                o[nbins + b] = t[nbins + i1] + h[nbins + i2];
            }
        }
    }
}

/* Boxcar S/N in FP64 of FP64 (e, v) rows: per width. */
void orc_boxcar_snr_f64(const double* folds, int64_t n, int64_t nbins, const int64_t* widths,
                        int64_t nw, double* out) {
    int64_t wmax = 0;
    for (int64_t k = 0; k < nw; ++k) if (widths[k] > wmax) wmax = widths[k];
#pragma omp parallel
    {
    double* ps = (double*)malloc(sizeof(double) * (size_t)(nbins + wmax));
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        const double* e = folds + p * 2 * nbins;
        const double* v = e + nbins;
        double run = 0.0;
        for (int64_t i = 0; i < nbins; ++i) { run += e[i] / sqrt(v[i]); ps[i] = run; }
        for (int64_t j = 0; j < wmax; ++j) ps[nbins + j] = ps[j] + run;
        for (int64_t k = 0; k < nw; ++k) {
            int64_t w = widths[k];
            double h = sqrt((double)(nbins - w) / (double)(nbins * w));
            double b = (double)w * h / (double)(nbins - w);
            double dmax = -DBL_MAX;
            for (int64_t i = 0; i < nbins; ++i) {
                double d = ps[i + w] - ps[i];
                if (d > dmax) dmax = d;
            }
            out[p * nw + k] = (h + b) * dmax - b * run;
        }
    }
    free(ps);
    }
}

/* ---- BF16 storage emulation (FP32 arithmetic, every stored fold element rounded to
 * BF16 with round-to-nearest-even) ---- */
static uint16_t orc_f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

static float orc_bf16_to_f32(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

void orc_round_bf16(const float* in, int64_t n, uint16_t* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_f32_to_bf16(in[i]);
}

void orc_bf16_to_f32_array(const uint16_t* in, int64_t n, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = orc_bf16_to_f32(in[i]);
}

/* unify_fold on BF16 storage: out = bf16(f32(tail) + f32(head)). */
void orc_unify_fold_bf16(const uint16_t* fold_in, int64_t nseg_in, int P, const int64_t* counts_prev,
                         const double* params, const int64_t* counts_cur, const double* limits,
                         int level, double tseg_brute, int64_t nbins, uint16_t* fold_out) {
    int64_t ncur = 1, nprev = 1;
    for (int k = 0; k < P; ++k) { ncur *= counts_cur[k]; nprev *= counts_prev[k]; }
    int64_t offs[8];
    offs[0] = 0;
    for (int k = 1; k < P; ++k) offs[k] = offs[k - 1] + counts_cur[k - 1];
    int64_t nseg_out = nseg_in / 2, rowlen = 2 * nbins;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < ncur; ++j) {
        double pset[8];
        int64_t r = j;
        for (int k = P - 1; k >= 0; --k) {
            int64_t ik = r % counts_cur[k];
            r /= counts_cur[k];
            pset[k] = params[offs[k] + ik];
        }
        int64_t row[2], sh[2];
        for (int latter = 0; latter < 2; ++latter) {
            int64_t idx[8];
            double ph;
            orc_ffa_resolve(P, pset, counts_prev, limits, level, latter, tseg_brute, nbins, idx, &ph);
            int64_t flat = 0;
            for (int k = 0; k < P; ++k) flat = flat * counts_prev[k] + idx[k];
            row[latter] = flat;
            sh[latter] = orc_shift_int(ph, nbins);
        }
        for (int64_t q = 0; q < nseg_out; ++q) {
            const uint16_t* t = fold_in + ((2 * q) * nprev + row[0]) * rowlen;
            const uint16_t* h = fold_in + ((2 * q + 1) * nprev + row[1]) * rowlen;
            uint16_t* o = fold_out + (q * ncur + j) * rowlen;
            for (int64_t b = 0; b < 2 * nbins; ++b) {
                int64_t c = b / nbins, bb = b % nbins;
                int64_t i1 = ((bb - sh[0]) % nbins + nbins) % nbins;
                int64_t i2 = ((bb - sh[1]) % nbins + nbins) % nbins;
                float v = orc_bf16_to_f32(t[c * nbins + i1]) + orc_bf16_to_f32(h[c * nbins + i2]);
                o[b] = orc_f32_to_bf16(v);
            }
        }
    }
}

/* ---------------------------------------------------------------------------------------
 * Monte-Carlo threshold design (detection/thresholding.py:45-95 simulate_folds).
 *
 * The reference draws numpy Generator.normal (bits not reproducible off the CPU). The
 * GPU path draws its deviates from Philox4x32-10 (Salmon et al., SC'11; counter {quad,
 * row, stream_lo, stream_hi}, key = 64-bit seed) through Box-Muller with a logarithm and
 * a sincos(2 pi u) built from IEEE basic operations; this is the same published
 * generator and the same series, written out independently, so the simulated folds can be
 * compared bit for bit. Pinned by the Random123 known-answer vectors (philox) and against
 * libm (ln, sin, cos) in tests/test_thresh_oracle.py.
 * ------------------------------------------------------------------------------------- */
static uint32_t orc_mulhi32(uint32_t a, uint32_t b) {
  return (uint32_t)(((uint64_t)a * (uint64_t)b) >> 32);
}

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint32_t h0 = orc_mulhi32(0xD2511F53u, c0), l0 = 0xD2511F53u * c0;
    const uint32_t h1 = orc_mulhi32(0xCD9E8D57u, c2), l1 = 0xCD9E8D57u * c2;
    const uint32_t n0 = h1 ^ c1 ^ k0, n2 = h0 ^ c3 ^ k1;
    c0 = n0; c1 = l1; c2 = n2; c3 = l0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ln x, x normal in (0, 1): x = m 2^e, m in [sqrt(1/2), sqrt(2)); atanh series in
 * s = (m-1)/(m+1) to s^21, Horner from the highest term. */
double orc_ln_det(double x) {
  static const double c[11] = {
      1.0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
      0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4,
      0x1.1111111111111p-4, 0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5,
      0x1.8618618618618p-5};
  uint64_t b;
  memcpy(&b, &x, sizeof b);
  int e = (int)((b >> 52) & 0x7ff) - 1022;
  const uint64_t mb = (b & 0x000fffffffffffffull) | 0x3fe0000000000000ull;
  double m;
  memcpy(&m, &mb, sizeof m);
  if (m < 0x1.6a09e667f3bcdp-1) { m = m * 2.0; e -= 1; }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double p = c[10];
  for (int k = 9; k >= 0; --k) p = p * s2 + c[k];
  return (double)e * 0x1.62e42fefa39efp-1 + 2.0 * s * p;
}

/* sin and cos of 2 pi u, u in [0, 1): exact quadrant split of 4u, first-quadrant angle
 * <= pi/4, Taylor series to a^17 (sin) and a^18 (cos). */
void orc_sincos_2pi(double u, double* sn, double* cs) {
  static const double cs_k[9] = {
      1.0, -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
      0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33,
      -0x1.ae7f3e733b81fp-41, 0x1.952c77030ad4ap-49};
  static const double cc_k[10] = {
      1.0, -0.5, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-16,
      -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37,
      0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53};
  const double t = 4.0 * u;
  const int q = (int)t;
  const double f = t - (double)q;
  const int swap = f > 0.5;
  const double g = swap ? 1.0 - f : f;
  const double a = g * 0x1.921fb54442d18p+0;
  const double a2 = a * a;
  double ps = cs_k[8];
  for (int k = 7; k >= 0; --k) ps = ps * a2 + cs_k[k];
  double pc = cc_k[9];
  for (int k = 8; k >= 0; --k) pc = pc * a2 + cc_k[k];
  double s0 = a * ps, c0 = pc;
  if (swap) { const double tmp = s0; s0 = c0; c0 = tmp; }
  switch (q & 3) {
    case 0: *sn = s0; *cs = c0; break;
    case 1: *sn = c0; *cs = -s0; break;
    case 2: *sn = -s0; *cs = -c0; break;
    default: *sn = -c0; *cs = s0; break;
  }
}

void orc_normal4(uint64_t seed, uint64_t stream, uint32_t row, uint32_t quad, double z[4]) {
  const uint32_t ctr[4] = {quad, row, (uint32_t)stream, (uint32_t)(stream >> 32)};
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  orc_philox4x32_10(ctr, key, r);
  double sn, cs;
  const double r0 = sqrt(-2.0 * orc_ln_det(((double)r[0] + 0.5) * 0x1p-32));
  orc_sincos_2pi(((double)r[1] + 0.5) * 0x1p-32, &sn, &cs);
  z[0] = r0 * cs;
  z[1] = r0 * sn;
  const double r1 = sqrt(-2.0 * orc_ln_det(((double)r[2] + 0.5) * 0x1p-32));
  orc_sincos_2pi(((double)r[3] + 0.5) * 0x1p-32, &sn, &cs);
  z[2] = r1 * cs;
  z[3] = r1 * sn;
}

/* out[i][j] = f32(std * z(seed, stream, i, j)); row_len a multiple of 4. */
void orc_rng_normal(uint64_t seed, uint64_t stream, int64_t n_rows, int64_t row_len, double sd,
                    float* out) {
  for (int64_t i = 0; i < n_rows; ++i)
    for (int64_t q = 0; q < row_len / 4; ++q) {
      double z[4];
      orc_normal4(seed, stream, (uint32_t)i, (uint32_t)q, z);
      for (int r = 0; r < 4; ++r) out[i * row_len + 4 * q + r] = (float)(sd * z[r]);
    }
}

/* thresholding.py:45-95: folds_sim[i][j] = folds[i % n_prev][j] + noise[i][j] +
 * profile_scaled[j] -- the first addition in FP32 (both operands FP32), the second in FP64
 * (profile_scaled is FP64), stored as FP32. src == NULL: zero source rows. */
void orc_simulate_folds(const float* src, int64_t n_prev, int64_t n_sim, int64_t nbins,
                        const double* prof_scaled, double var_add, uint64_t seed,
                        uint64_t stream, float* out) {
  const double sd = sqrt(var_add);
  for (int64_t i = 0; i < n_sim; ++i)
    for (int64_t q = 0; q < nbins / 4; ++q) {
      double z[4];
      orc_normal4(seed, stream, (uint32_t)i, (uint32_t)q, z);
      for (int r = 0; r < 4; ++r) {
        const int64_t j = 4 * q + r;
        const float noise = (float)(sd * z[r]);
        const float base = src ? src[(i % n_prev) * nbins + j] : 0.0f;
        const float a = base + noise;
        out[i * nbins + j] = (float)((double)a + prof_scaled[j]);
      }
    }
}
