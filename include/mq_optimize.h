/*
 * mq_optimize.h -- the Optimize search over drive settings, run on the engine.
 *
 * Replaces the research driver of the Tauri app (src-tauri/src/research.rs:
 * fit_quality :67-75, summarize :76-95, eligible/best :96-107, experiment
 * :133-163), which runs frequency sweeps on the live mqviz model. Here the
 * sweeps are batches of the HIP engine (mq_engine.h), and a second search
 * climbs the differentiable objective with mq_forward / mq_backward.
 */
#ifndef MQ_OPTIMIZE_H
#define MQ_OPTIMIZE_H

#include <stddef.h>
#include <stdint.h>

#include "mq_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Gaussian peak on a baseline fitted to a transmission curve,
 * T(f) = base + amp exp(-(f - f0)^2 / (2 s^2)), by weighted least squares with
 * binomial errors; fwhm = 2 sqrt(2 ln 2) s; r_f = f0 / fwhm (the frequency
 * resolving power the reference's fit reports as r_frequency,
 * research.rs:83-87). Errors are one standard error from the covariance. */
typedef struct {
    double f0_khz, f0_err;
    double fwhm_khz, fwhm_err;
    double amp, base;
    double r_f, r_f_err;
    double chi2_red;
    int32_t converged, npts;
} mq_peak_fit;

/* Fits npts points (frequencies [kHz], transmitted fractions, ions per point).
 * Host only. Returns 0, or < 0 on bad input; a fit that does not converge
 * returns 0 with out->converged = 0. */
int mq_fit_peak(const double* f_khz, const double* t, const uint64_t* n, size_t npts, mq_peak_fit* out);

/* Options of the grid search (research.rs:133-163). */
typedef struct {
    double f_lo_khz, f_hi_khz, f_step_khz; /* coarse AC-frequency scan (the reference: 70-100 kHz, 0.5 kHz) */
    uint64_t n;                            /* ions per coarse point; a refined scan uses 2n, the confirmation 3n */
    double min_r;                          /* required resolving power minus two standard errors */
    double dlnf_dlnmz;                     /* |d ln f / d ln(m/z)| at the operating point, to turn the
                                              frequency resolving power into the mass resolving power
                                              (the reference takes it from the Mathieu terrain); 0 means 1 */
    uint64_t seed;                         /* seed of the first drive; drive i uses seed + 2 i */
    uint64_t leak_n;                       /* ions of the leakage species per point (0: none) */
} mq_optimize_opts;

/* Issues of a row (fit_quality, research.rs:67-75). */
enum {
    MQ_OPT_OK = 0,
    MQ_OPT_NO_FIT = 1,        /* the fit did not converge */
    MQ_OPT_TOO_NARROW = 2,    /* FWHM below 4 scan steps */
    MQ_OPT_WINGS_OUTSIDE = 3, /* f0 -/+ FWHM beyond the scan */
    MQ_OPT_R_UNCERTAIN = 4,   /* resolving-power error above 30 % */
    MQ_OPT_BAD_SHAPE = 5      /* reduced chi^2 above 10 */
};

/* One drive amplitude of the search: its scan, fit and the scan point nearest
 * the fitted centre. */
typedef struct {
    double v_ac;
    double lo_khz, hi_khz, step_khz;
    uint64_t ions_per_point;
    mq_peak_fit fit;
    double r_mz, r_mz_err;
    double f_best_khz, t_best, t_best_err; /* t_best_err: Wilson-score error of t_k / t_n */
    uint64_t t_k, t_n;                     /* transmitted and ions at f_best */
    uint64_t leak_k, leak_n;               /* the leakage species at f_best (0, 0 without one) */
    double leakage, leakage_err;           /* leak_k/leak_n over t_k/t_n (mq_ratio; NaN when undefined) */
    double leakage_ul95;                   /* its 95 % upper limit (mq_ratio) */
    int32_t refined; /* 1: a narrow peak (FWHM < 2 kHz) was rescanned around its centre with 2n ions */
    int32_t issue;   /* MQ_OPT_* */
} mq_optimize_row;

/* Statistics of the Optimize rows (mqviz/src/stats.rs). mq_frac_sigma: the
 * Wilson-score 1-sigma error of k/n, sqrt(p(1-p)/n + 1/(4 n^2)) / (1 + 1/n)
 * (Frac::sigma, stats.rs:17-22; NaN for n = 0). mq_ratio: the ratio of two
 * fractions (ka/na) / (kb/nb) with its log-delta error on Jeffreys-shifted
 * counts (ratio, stats.rs:51-62) and the 95 % upper limit from the
 * Wilson-Hilferty Poisson limit of ka (ratio_ul95 and poisson_ul95,
 * stats.rs:29-33, 65-69); NaN where the reference returns NaN. Host only. */
int mq_frac_sigma(uint64_t k, uint64_t n, double* sigma);
int mq_ratio(uint64_t ka, uint64_t na, uint64_t kb, uint64_t nb, double* v, double* e, double* ul95);

/* experiment (research.rs:133-163) on the engine, with research_start's
 * input check (research.rs:166-167: min_r 5-150 and 500-5000 ions per coarse
 * point). For each of the n_ac AC amplitudes: a scan of the AC frequency (one
 * batch: one job of o->n ions per frequency, drive rf with v_ac and f_ac
 * replaced, plus one job of o->leak_n ions of the leakage species leak per
 * frequency when leak_n > 0; the reference draws both species from one live
 * beam by abundance), a peak fit, and for a narrow peak a rescan of 60 steps
 * over centre -/+ clamp(3 FWHM, 2, 6) kHz with 2n ions; each row reports the
 * scan point nearest the fitted centre with its leakage (annotate_leakage,
 * research.rs:88-95). Then the eligible row (no issue, r_mz - 2 r_mz_err >=
 * min_r) with the largest t_best - 2 t_best_err (research.rs:96-107), and an
 * independent confirmation of it with 3n ions and fresh seeds. rows[n_ac] is
 * filled; *best = its index or -1; *confirmation and *confirmed (1 when the
 * confirmation is eligible as well) are written when *best >= 0. */
int mq_optimize_grid(mq_engine* eng, const mq_species* sp, const mq_species* leak, const mq_rfac* rf,
                     const double dc[4], const mq_physics* ph, int32_t path, const double* v_ac, size_t n_ac,
                     const mq_optimize_opts* o, mq_optimize_row* rows, int32_t* best,
                     mq_optimize_row* confirmation, int32_t* confirmed);

/* Options of the gradient search on the differentiable objective. */
typedef struct {
    uint32_t mask;     /* bit k set: theta_k is searched (V_rf, f_rf, V_exc, f_exc, V_dc = dc_scale) */
    uint32_t max_iter; /* forward evaluations at most */
    double lo[5], hi[5];
    double step0[5];   /* initial trust-region radius per parameter, in its own units */
    double tol;        /* stop when the radius falls below tol (in units of step0) */
} mq_grad_opts;

/* Projected gradient ascent of J (mq_objective) over the masked parameters of
 * *job, starting from its values and staying inside [lo, hi]: each step moves
 * the parameters along the gradient scaled by step0, by the trust radius in the
 * largest coordinate; a step that raises J is kept and widens the radius
 * (x 1.5, at most 4), one that does not is undone and halves it. No reference
 * counterpart (its Optimize search is the grid above). On return *job holds the
 * best parameters, *J_best their objective, *n_eval the forward evaluations;
 * history (optional, 6 doubles per evaluation: theta[5], J) records every
 * evaluated point in order. */
int mq_optimize_grad(mq_engine* eng, mq_job* job, const double dc[4], const mq_physics* ph,
                     const mq_objective* obj, const mq_grad_opts* o, double* J_best, uint32_t* n_eval,
                     double* history);

#ifdef __cplusplus
}
#endif

#endif /* MQ_OPTIMIZE_H */
