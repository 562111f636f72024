// optimize.cpp -- the Optimize search over drive settings on the engine
// (include/mq_optimize.h).
//
// mq_optimize_grid follows the research driver of the Tauri app
// (src-tauri/src/research.rs:133-163, with fit_quality :67-75, the row summary
// :76-95 and eligible/best :96-107): AC-frequency scans per AC amplitude, a
// peak fit per scan, a rescan of narrow peaks, the best eligible row and an
// independent confirmation. Its sweeps run as batches of the HIP engine
// instead of the live mqviz model, one job per frequency point. The peak fit
// is a weighted Gaussian-plus-baseline least-squares fit (the reference's
// analysis::fit_sweep is not in the snapshot's hot path).
//
// mq_optimize_grad climbs the differentiable objective (mq_forward /
// mq_backward) with a trust-region projected gradient ascent.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mq_engine.h"
#include "../../include/mq_optimize.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace {

struct OptErr : std::runtime_error {
    int code;
    OptErr(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Solves the 4 x 4 system A x = b in place by Gaussian elimination with
// partial pivoting; false if singular.
bool solve4(double A[4][4], double b[4], double x[4])
{
    int p[4] = {0, 1, 2, 3};
    double M[4][5];
    for (int i = 0; i < 4; ++i) {
        for (int j = 0; j < 4; ++j) M[i][j] = A[i][j];
        M[i][4] = b[i];
    }
    for (int c = 0; c < 4; ++c) {
        int piv = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[piv][c])) piv = r;
        if (!(std::fabs(M[piv][c]) > 0.0)) return false;
        if (piv != c) {
            for (int j = 0; j < 5; ++j) std::swap(M[c][j], M[piv][j]);
            std::swap(p[c], p[piv]);
        }
        for (int r = c + 1; r < 4; ++r) {
            const double f = M[r][c] / M[c][c];
            for (int j = c; j < 5; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 3; i >= 0; --i) {
        double s = M[i][4];
        for (int j = i + 1; j < 4; ++j) s -= M[i][j] * x[j];
        x[i] = s / M[i][i];
    }
    return true;
}

bool invert4(double A[4][4], double inv[4][4])
{
    for (int c = 0; c < 4; ++c) {
        double B[4][4], e[4] = {0, 0, 0, 0}, x[4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) B[i][j] = A[i][j];
        e[c] = 1.0;
        if (!solve4(B, e, x)) return false;
        for (int i = 0; i < 4; ++i) inv[i][c] = x[i];
    }
    return true;
}

constexpr double FWHM_PER_SIGMA = 2.3548200450309493;  // 2 sqrt(2 ln 2)

// Weighted least-squares fit of base + amp exp(-(f - f0)^2 / (2 s^2)) by
// Levenberg-Marquardt; weights 1/var with var = t(1 - t)/n, t kept at least
// half an ion away from 0 and 1.
mq_peak_fit fit_peak(const double* f, const double* t, const uint64_t* n, size_t np)
{
    mq_peak_fit out{};
    out.npts = (int32_t)np;
    if (np < 5) return out;
    std::vector<double> w(np);
    for (size_t i = 0; i < np; ++i) {
        const double nn = std::max<double>(1.0, (double)n[i]);
        const double tt = std::min(std::max(t[i], 0.5 / nn), 1.0 - 0.5 / nn);
        w[i] = nn / (tt * (1.0 - tt));
    }
    // start: the highest point, baseline the lowest point, width from the half-maximum crossings
    size_t im = 0;
    double tmin = t[0];
    for (size_t i = 1; i < np; ++i) {
        if (t[i] > t[im]) im = i;
        tmin = std::min(tmin, t[i]);
    }
    double p[4] = {t[im] - tmin, f[im], 0.0, tmin};  // amp, f0, s, base
    if (!(p[0] > 0.0)) return out;
    const double half = tmin + 0.5 * p[0];
    size_t a = im, b = im;
    while (a > 0 && t[a - 1] >= half) --a;
    while (b + 1 < np && t[b + 1] >= half) ++b;
    double step = std::fabs(f[np - 1] - f[0]) / (double)(np - 1);
    p[2] = std::max(0.5 * step, (std::fabs(f[b] - f[a]) + step) / FWHM_PER_SIGMA);
    auto chi2_of = [&](const double q[4]) {
        double c = 0.0;
        for (size_t i = 0; i < np; ++i) {
            const double u = (f[i] - q[1]) / q[2];
            const double r = t[i] - (q[3] + q[0] * std::exp(-0.5 * u * u));
            c += w[i] * r * r;
        }
        return c;
    };
    auto normal = [&](const double q[4], double JtJ[4][4], double Jtr[4]) {
        for (int i = 0; i < 4; ++i) {
            Jtr[i] = 0.0;
            for (int j = 0; j < 4; ++j) JtJ[i][j] = 0.0;
        }
        for (size_t i = 0; i < np; ++i) {
            const double d = f[i] - q[1], u = d / q[2];
            const double e = std::exp(-0.5 * u * u);
            const double g[4] = {e, q[0] * e * d / (q[2] * q[2]), q[0] * e * d * d / (q[2] * q[2] * q[2]), 1.0};
            const double r = t[i] - (q[3] + q[0] * e);
            for (int k = 0; k < 4; ++k) {
                Jtr[k] += w[i] * g[k] * r;
                for (int l = 0; l < 4; ++l) JtJ[k][l] += w[i] * g[k] * g[l];
            }
        }
    };
    double chi2 = chi2_of(p), lambda = 1e-3;
    bool conv = false;
    for (int it = 0; it < 500; ++it) {
        double JtJ[4][4], Jtr[4], A[4][4], d[4];
        normal(p, JtJ, Jtr);
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) A[i][j] = JtJ[i][j] + (i == j ? lambda * std::max(JtJ[i][i], 1e-300) : 0.0);
        if (!solve4(A, Jtr, d)) break;
        double q[4] = {p[0] + d[0], p[1] + d[1], p[2] + d[2], p[3] + d[3]};
        q[2] = std::fabs(q[2]);
        const double c2 = q[2] > 0.0 ? chi2_of(q) : INFINITY;
        if (c2 <= chi2) {
            const double rel = (chi2 - c2) / std::max(chi2, 1e-300);
            bool small = true;
            for (int k = 0; k < 4; ++k) small = small && std::fabs(d[k]) <= 1e-9 * (std::fabs(p[k]) + 1e-12);
            std::copy(q, q + 4, p);
            chi2 = c2;
            lambda = std::max(lambda * 0.3, 1e-12);
            if (rel < 1e-12 || small) { conv = true; break; }
        } else {
            lambda *= 10.0;
            if (lambda > 1e12) { conv = true; break; }  // no step lowers chi^2: at the minimum
        }
    }
    double JtJ[4][4], Jtr[4], C[4][4];
    normal(p, JtJ, Jtr);
    const double dof = (double)np - 4.0;
    const double red = chi2 / std::max(dof, 1.0);
    if (!conv || !invert4(JtJ, C) || !(p[0] > 0.0) || !(p[2] > 0.0)) return out;
    const double scale = std::max(1.0, red);
    double e[4];
    for (int k = 0; k < 4; ++k) e[k] = std::sqrt(std::max(C[k][k], 0.0) * scale);
    out.converged = 1;
    out.amp = p[0];
    out.f0_khz = p[1];
    out.f0_err = e[1];
    out.fwhm_khz = FWHM_PER_SIGMA * p[2];
    out.fwhm_err = FWHM_PER_SIGMA * e[2];
    out.base = p[3];
    out.chi2_red = red;
    out.r_f = p[1] / out.fwhm_khz;
    out.r_f_err = out.r_f * std::sqrt((e[1] / p[1]) * (e[1] / p[1]) + (e[2] / p[2]) * (e[2] / p[2]));
    return out;
}

// Frac::sigma (mqviz/src/stats.rs:17-22): Wilson-score 1-sigma error of k/n.
double frac_sigma(uint64_t k, uint64_t n)
{
    if (n == 0) return NAN;
    const double nn = (double)n, p = (double)k / nn;
    return std::sqrt(p * (1.0 - p) / nn + 0.25 / (nn * nn)) / (1.0 + 1.0 / nn);
}

// poisson_ul95 (stats.rs:29-33): Wilson-Hilferty one-sided 95 % upper limit of a Poisson mean.
double poisson_ul95(uint64_t k)
{
    const double a = (double)(k + 1);
    const double c = 1.0 - 1.0 / (9.0 * a) + 1.644854 / (3.0 * std::sqrt(a));
    return a * c * c * c;
}

// ratio and ratio_ul95 (stats.rs:51-69): (ka/na) / (kb/nb), log-delta error on Jeffreys-shifted counts.
void ratio(uint64_t ka, uint64_t na, uint64_t kb, uint64_t nb, double& v, double& e, double& ul)
{
    if (na == 0 || nb == 0 || kb == 0) {
        v = e = ul = NAN;
        return;
    }
    const double pa = (double)ka / (double)na, pb = (double)kb / (double)nb;
    const double vb = (1.0 - pb) / ((double)kb + 0.5);
    if (ka == 0) {
        v = 0.0;
        e = (1.0 / (double)na) / pb * std::sqrt(1.0 + vb);
    } else {
        const double va = (1.0 - pa) / ((double)ka + 0.5);
        v = pa / pb;
        e = v * std::sqrt(va + vb);
    }
    ul = (poisson_ul95(ka) / (double)na) / pb * (1.0 + 1.645 * std::sqrt(vb));
}

uint64_t point_seed(uint64_t base, size_t k) { return base * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull * (k + 1); }

// One scan of the AC frequency: a batch of one job per frequency.
struct Scan {
    std::vector<double> f, t;
    std::vector<uint64_t> k, n, lk, ln;  // transmitted / ions of the main and the leakage species per point
};

// One scan of the AC frequency: a batch of one job per frequency (and one of the
// leakage species per frequency when leak_n > 0).
void scan(mq_engine* eng, const mq_species& sp, const mq_species* leak, uint64_t leak_n, const mq_rfac& rf,
          double v_ac, const double dc[4], const mq_physics& ph, int32_t path, double lo, double hi, double step,
          uint64_t n, uint64_t seed, Scan& S)
{
    std::vector<double>& fs = S.f;
    fs.clear();
    for (int i = 0;; ++i) {
        const double f = lo + step * (double)i;
        if (!(f <= hi + 1e-9)) break;
        fs.push_back(f);
    }
    const size_t m = fs.size();
    const bool with_leak = leak && leak_n > 0;
    std::vector<mq_job> jobs(with_leak ? 2 * m : m);
    for (size_t q = 0; q < jobs.size(); ++q) {
        const size_t k = q % m;
        mq_job& j = jobs[q];
        j = mq_job{};
        j.sp = q < m ? sp : *leak;
        j.drive = rf;
        j.drive.v_ac = v_ac;
        j.drive.f_ac = fs[k] * 1e3;
        j.n = q < m ? n : leak_n;
        j.seed = q < m ? point_seed(seed, k) : point_seed(seed ^ 0x1EA4ull, k);
        j.dc_scale = 1.0;
    }
    std::vector<mq_outcome> out(jobs.size());
    const int rc = mq_run_batch(eng, jobs.data(), jobs.size(), dc, &ph, path, out.data());
    if (rc != 0) throw OptErr(rc, mq_last_error());
    S.t.assign(m, 0.0);
    S.k.assign(m, 0);
    S.n.assign(m, 0);
    S.lk.assign(m, 0);
    S.ln.assign(m, 0);
    for (size_t k = 0; k < m; ++k) {
        S.n[k] = out[k].n;
        S.k[k] = out[k].transmitted;
        S.t[k] = out[k].n ? (double)out[k].transmitted / (double)out[k].n : 0.0;
        if (with_leak) {
            S.ln[k] = out[m + k].n;
            S.lk[k] = out[m + k].transmitted;
        }
    }
}

// The row summary of one scan (research.rs:76-95 and fit_quality :67-75).
mq_optimize_row summarize_row(double v_ac, double lo, double hi, double step, uint64_t n, const Scan& S, double dlnf)
{
    const std::vector<double>& fs = S.f;
    const std::vector<double>& ts = S.t;
    const std::vector<uint64_t>& ns = S.n;
    mq_optimize_row r{};
    r.leakage = r.leakage_err = r.leakage_ul95 = NAN;
    r.v_ac = v_ac;
    r.lo_khz = lo;
    r.hi_khz = fs.empty() ? hi : fs.back();
    r.step_khz = step;
    r.ions_per_point = n;
    r.fit = fit_peak(fs.data(), ts.data(), ns.data(), fs.size());
    if (!r.fit.converged) {
        r.issue = MQ_OPT_NO_FIT;
        return r;
    }
    size_t i = 0;
    for (size_t k = 1; k < fs.size(); ++k)
        if (std::fabs(fs[k] - r.fit.f0_khz) < std::fabs(fs[i] - r.fit.f0_khz)) i = k;
    r.f_best_khz = fs[i];
    r.t_best = ts[i];
    r.t_best_err = frac_sigma(S.k[i], ns[i]);
    r.t_k = S.k[i];
    r.t_n = ns[i];
    r.leak_k = S.lk[i];
    r.leak_n = S.ln[i];
    if (r.leak_n) ratio(r.leak_k, r.leak_n, r.t_k, r.t_n, r.leakage, r.leakage_err, r.leakage_ul95);
    r.r_mz = r.fit.r_f * dlnf;
    r.r_mz_err = r.fit.r_f_err * dlnf;
    if (r.fit.fwhm_khz < 4.0 * step) r.issue = MQ_OPT_TOO_NARROW;
    else if (r.fit.f0_khz - r.fit.fwhm_khz < fs.front() || r.fit.f0_khz + r.fit.fwhm_khz > fs.back())
        r.issue = MQ_OPT_WINGS_OUTSIDE;
    else if (r.r_mz_err > 0.3 * r.r_mz) r.issue = MQ_OPT_R_UNCERTAIN;
    else if (r.fit.chi2_red > 10.0) r.issue = MQ_OPT_BAD_SHAPE;
    return r;
}

bool eligible(const mq_optimize_row& r, double min_r) { return r.issue == MQ_OPT_OK && r.r_mz - 2.0 * r.r_mz_err >= min_r; }

double theta_get(const mq_job& j, int k)
{
    switch (k) {
    case 0: return j.drive.v_rf;
    case 1: return j.drive.f_rf;
    case 2: return j.drive.v_ac;
    case 3: return j.drive.f_ac;
    default: return j.dc_scale;
    }
}

void theta_set(mq_job& j, int k, double v)
{
    switch (k) {
    case 0: j.drive.v_rf = v; break;
    case 1: j.drive.f_rf = v; break;
    case 2: j.drive.v_ac = v; break;
    case 3: j.drive.f_ac = v; break;
    default: j.dc_scale = v; break;
    }
}

}  // namespace

extern "C" {

int mq_fit_peak(const double* f_khz, const double* t, const uint64_t* n, size_t npts, mq_peak_fit* out)
{
    if (!f_khz || !t || !n || !out) { mq_set_last_error("NULL argument"); return -1; }
    *out = fit_peak(f_khz, t, n, npts);
    return 0;
}

int mq_frac_sigma(uint64_t k, uint64_t n, double* sigma)
{
    if (!sigma) { mq_set_last_error("NULL argument"); return -1; }
    *sigma = frac_sigma(k, n);
    return 0;
}

int mq_ratio(uint64_t ka, uint64_t na, uint64_t kb, uint64_t nb, double* v, double* e, double* ul95)
{
    if (!v || !e || !ul95) { mq_set_last_error("NULL argument"); return -1; }
    ratio(ka, na, kb, nb, *v, *e, *ul95);
    return 0;
}

int mq_optimize_grid(mq_engine* eng, const mq_species* sp, const mq_species* leak, const mq_rfac* rf,
                     const double dc[4], const mq_physics* ph, int32_t path, const double* v_ac, size_t n_ac,
                     const mq_optimize_opts* o, mq_optimize_row* rows, int32_t* best,
                     mq_optimize_row* confirmation, int32_t* confirmed)
{
    try {
        if (!eng || !sp || !rf || !dc || !ph || !v_ac || !o || !rows || !best || !confirmation || !confirmed)
            throw OptErr(-1, "NULL argument");
        if (o->leak_n > 0 && !leak) throw OptErr(-1, "leak_n > 0 needs the leakage species");
        // research_start's check (research.rs:167)
        if (!std::isfinite(o->min_r) || o->min_r < 5.0 || o->min_r > 150.0 || o->n < 500 || o->n > 5000)
            throw OptErr(-1, "Choose resolving power 5-150 and 500-5000 ions per coarse point");
        if (n_ac == 0) throw OptErr(-1, "no AC amplitudes to search");
        if (!(o->f_step_khz > 0.0) || !(o->f_hi_khz > o->f_lo_khz))
            throw OptErr(-1, "bad scan: need f_hi > f_lo and f_step > 0");
        const double dlnf = o->dlnf_dlnmz > 0.0 ? o->dlnf_dlnmz : 1.0;
        Scan S;
        for (size_t i = 0; i < n_ac; ++i) {
            uint64_t seed = o->seed + 2 * (uint64_t)i;
            scan(eng, *sp, leak, o->leak_n, *rf, v_ac[i], dc, *ph, path, o->f_lo_khz, o->f_hi_khz, o->f_step_khz, o->n,
                 seed, S);
            mq_optimize_row row = summarize_row(v_ac[i], o->f_lo_khz, o->f_hi_khz, o->f_step_khz, o->n, S, dlnf);
            if (row.fit.converged && row.fit.fwhm_khz < 2.0) {
                // a narrow peak: resolve it around its centre (research.rs:143-150)
                const double halfw = std::min(std::max(3.0 * row.fit.fwhm_khz, 2.0), 6.0);
                const double lo = std::max(row.fit.f0_khz - halfw, o->f_lo_khz);
                const double hi = std::min(row.fit.f0_khz + halfw, o->f_hi_khz);
                const double step = (hi - lo) / 60.0;
                seed += 100;
                scan(eng, *sp, leak, 2 * o->leak_n, *rf, v_ac[i], dc, *ph, path, lo, hi, step, 2 * o->n, seed, S);
                row = summarize_row(v_ac[i], lo, hi, step, 2 * o->n, S, dlnf);
                row.refined = 1;
            }
            rows[i] = row;
        }
        // best eligible row by transmission minus two standard errors (research.rs:101-107)
        int32_t b = -1;
        double score = -INFINITY;
        for (size_t i = 0; i < n_ac; ++i) {
            if (!eligible(rows[i], o->min_r)) continue;
            const double s = rows[i].t_best - 2.0 * rows[i].t_best_err;
            if (s > score) { score = s; b = (int32_t)i; }
        }
        *best = b;
        *confirmed = 0;
        if (b >= 0) {
            // independent confirmation with three times the ions and fresh seeds (research.rs:154-160)
            const mq_optimize_row& r = rows[b];
            scan(eng, *sp, leak, 3 * o->leak_n, *rf, r.v_ac, dc, *ph, path, r.lo_khz, r.hi_khz, r.step_khz, 3 * o->n,
                 o->seed + 0x5EED0000ull, S);
            *confirmation = summarize_row(r.v_ac, r.lo_khz, r.hi_khz, r.step_khz, 3 * o->n, S, dlnf);
            *confirmed = eligible(*confirmation, o->min_r) ? 1 : 0;
        }
        return 0;
    } catch (const OptErr& e) {
        if (e.code != 0 && std::string(e.what()) != mq_last_error()) mq_set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -9;
    }
}

int mq_optimize_grad(mq_engine* eng, mq_job* job, const double dc[4], const mq_physics* ph,
                     const mq_objective* obj, const mq_grad_opts* o, double* J_best, uint32_t* n_eval,
                     double* history)
{
    try {
        if (!eng || !job || !dc || !ph || !obj || !o || !J_best || !n_eval) throw OptErr(-1, "NULL argument");
        if ((o->mask & 31u) == 0) throw OptErr(-1, "mask selects no parameter");
        if (o->max_iter == 0) throw OptErr(-1, "max_iter must be positive");
        for (int k = 0; k < 5; ++k) {
            if (!((o->mask >> k) & 1u)) continue;
            if (!(o->step0[k] > 0.0) || !(o->lo[k] <= o->hi[k])) throw OptErr(-1, "need step0 > 0 and lo <= hi");
            const double v = theta_get(*job, k);
            if (v < o->lo[k] || v > o->hi[k]) throw OptErr(-1, "the starting point lies outside [lo, hi]");
        }
        mq_job cur = *job;
        uint32_t evals = 0;
        auto record = [&](const mq_job& j, double J) {
            if (history) {
                for (int k = 0; k < 5; ++k) history[6 * (size_t)evals + k] = theta_get(j, k);
                history[6 * (size_t)evals + 5] = J;
            }
            ++evals;
        };
        auto forward = [&](const mq_job& j) {
            double J = 0.0;
            const int rc = mq_forward(eng, &j, dc, ph, obj, &J);
            if (rc != 0) throw OptErr(rc, mq_last_error());
            return J;
        };
        auto backward = [&](double g[5]) {
            const int rc = mq_backward(eng, 1.0, g, nullptr);
            if (rc != 0) throw OptErr(rc, mq_last_error());
        };
        double J = forward(cur);
        record(cur, J);
        double g[5];
        backward(g);
        double radius = 1.0;
        while (evals < o->max_iter && radius >= o->tol) {
            double m = 0.0;
            for (int k = 0; k < 5; ++k)
                if ((o->mask >> k) & 1u) m = std::max(m, std::fabs(g[k] * o->step0[k]));
            if (!(m > 0.0)) break;
            mq_job next = cur;
            bool moved = false;
            for (int k = 0; k < 5; ++k) {
                if (!((o->mask >> k) & 1u)) continue;
                const double v = theta_get(cur, k);
                const double nv = std::min(std::max(v + radius * o->step0[k] * (g[k] * o->step0[k] / m), o->lo[k]),
                                           o->hi[k]);
                moved = moved || nv != v;
                theta_set(next, k, nv);
            }
            if (!moved) break;
            const double Jn = forward(next);
            record(next, Jn);
            if (Jn > J) {
                cur = next;
                J = Jn;
                backward(g);
                radius = std::min(radius * 1.5, 4.0);
            } else {
                radius *= 0.5;
            }
        }
        *job = cur;
        *J_best = J;
        *n_eval = evals;
        return 0;
    } catch (const OptErr& e) {
        if (std::string(e.what()) != mq_last_error()) mq_set_last_error(e.what());
        return e.code;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -9;
    }
}

}  // extern "C"
