// scan.cpp -- chsim::scan over the C ABI: the persistent result cache and the
// adaptive two-stage frequency scan (simulation/chsim/src/scan.rs:1-78).
//
// The cache file is the reference's TSV: one line per job,
//   <key>\t<n>,<crashed>,<returned>,<transmitted>,<left>,<timeout>,<mean_coll>,<mean_time>
// with key = "{geom:?}|{species:?}|{drive:?}|{n}|{seed}|{dc:?}|{physics:?}" in Rust
// Debug notation (scan.rs:9-11) and the two means in Rust Display notation
// (scan.rs:41), so files written by either side, and the spectrum model's
// reader (spectrum_trans.rs:28-48), keep working.
#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <system_error>
#include <vector>

#include "../../include/mq_engine.h"

void mq_set_last_error(const std::string& msg);  // engine.cpp

namespace {

// Shortest round-trip digits and decimal exponent of |x|: x = 0.d1d2... * 10^e10.
void shortest_digits(double x, std::string& digits, int& e10)
{
    char buf[64];
    auto r = std::to_chars(buf, buf + sizeof buf, std::fabs(x), std::chars_format::scientific);
    std::string s(buf, r.ptr);
    const size_t epos = s.find('e');
    std::string mant = s.substr(0, epos);
    const int exp = std::stoi(s.substr(epos + 1));
    digits.clear();
    for (char c : mant)
        if (c != '.') digits.push_back(c);
    while (digits.size() > 1 && digits.back() == '0') digits.pop_back();
    e10 = exp + 1;  // position of the decimal point after the first `e10` digits
}

std::string plain_decimal(const std::string& d, int e10, int min_frac)
{
    std::string out;
    if (e10 <= 0) {
        out = "0.";
        out.append((size_t)(-e10), '0');
        out += d;
    } else if ((size_t)e10 >= d.size()) {
        out = d;
        out.append((size_t)e10 - d.size(), '0');
        if (min_frac > 0) out += ".0";
    } else {
        out = d.substr(0, (size_t)e10) + "." + d.substr((size_t)e10);
    }
    return out;
}

}  // namespace

// Rust `{:?}` of an f64: shortest round-trip digits; decimal notation with at
// least one fractional digit for 1e-4 <= |x| < 1e16 (and 0), else d.ddde<exp>.
std::string mq_rust_debug_f64(double x)
{
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    const std::string sign = std::signbit(x) ? "-" : "";
    if (x == 0.0) return sign + "0.0";
    std::string d;
    int e10;
    shortest_digits(x, d, e10);
    const double a = std::fabs(x);
    if (a < 1e-4 || a >= 1e16) {
        std::string m = d.substr(0, 1);
        if (d.size() > 1) m += "." + d.substr(1);
        return sign + m + "e" + std::to_string(e10 - 1);
    }
    return sign + plain_decimal(d, e10, 1);
}

// Rust `{}` of an f64: shortest round-trip digits in plain decimal notation,
// no fractional part for integral values.
std::string mq_rust_display_f64(double x)
{
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return x > 0 ? "inf" : "-inf";
    const std::string sign = std::signbit(x) ? "-" : "";
    if (x == 0.0) return sign + "0";
    std::string d;
    int e10;
    shortest_digits(x, d, e10);
    return sign + plain_decimal(d, e10, 0);
}

namespace {

// The reference's key holds no random-stream or path field (its Engine has one
// stream and its paths are statistically equivalent). This build's extension
// fields are appended inside the Physics braces only when they differ from the
// default, so default requests keep the reference's key byte for byte while a
// request for the reference's own streams (rng) or for the FP64 plain path
// never reads a line the default Philox / FP32 batched run wrote, or the reverse.
std::string key_of(const char* geom, const mq_job& j, const double dc[4], const mq_physics& ph, int32_t path)
{
    auto D = mq_rust_debug_f64;
    auto B = [](int32_t v) { return std::string(v ? "true" : "false"); };
    std::ostringstream k;
    k << geom << "|Species { mass: " << D(j.sp.mass) << ", z: " << D(j.sp.z) << ", ccs_n2: " << D(j.sp.ccs_n2) << " }";
    k << "|RfAc { v_rf: " << D(j.drive.v_rf) << ", f_rf: " << D(j.drive.f_rf) << ", v_ac: " << D(j.drive.v_ac)
      << ", f_ac: " << D(j.drive.f_ac);
    if (j.dc_scale != 1.0) k << ", dc_scale: " << D(j.dc_scale);
    if (j.pressure_pa > 0.0) k << ", pressure_pa: " << D(j.pressure_pa);
    k << " }|" << j.n << "|" << j.seed << "|[" << D(dc[0]) << ", " << D(dc[1]) << ", " << D(dc[2]) << ", "
      << D(dc[3]) << "]";
    k << "|Physics { legacy_force: " << B(ph.legacy_force) << ", legacy_collisions: " << B(ph.legacy_collisions)
      << ", legacy_fields: " << B(ph.legacy_fields) << ", legacy_injection: " << B(ph.legacy_injection)
      << ", legacy_criteria: " << B(ph.legacy_criteria) << ", pressure_pa: " << D(ph.pressure_pa)
      << ", temp_k: " << D(ph.temp_k) << ", dt: " << D(ph.dt) << ", t_max: " << D(ph.t_max)
      << ", r_inject_mm: " << D(ph.r_inject_mm);
    if (ph.injection != MQ_INJ_REFERENCE) k << ", injection: " << ph.injection << ", inj_sigma_mm: " << D(ph.inj_sigma_mm);
    if (ph.max_steps) k << ", max_steps: " << ph.max_steps;
    if (ph.rng != MQ_RNG_PHILOX) k << ", rng: " << ph.rng;
    if (path != MQ_PATH_BATCHED) k << ", path: " << path;
    k << " }";
    return k.str();
}

std::map<std::string, mq_outcome> load_cache(const std::string& path)
{
    std::map<std::string, mq_outcome> m;
    std::ifstream f(path);
    std::string line;
    while (std::getline(f, line)) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos) continue;
        std::string v = line.substr(tab + 1);
        const size_t tab2 = v.find('\t');
        if (tab2 != std::string::npos) v = v.substr(0, tab2);
        std::vector<double> x;
        std::stringstream ss(v);
        std::string tok;
        bool ok = true;
        while (std::getline(ss, tok, ',')) {
            char* end = nullptr;
            const double d = std::strtod(tok.c_str(), &end);
            if (end == tok.c_str()) { ok = false; break; }
            x.push_back(d);
        }
        if (!ok || x.size() != 8) continue;
        mq_outcome o{};
        o.n = (uint64_t)x[0]; o.crashed = (uint64_t)x[1]; o.returned = (uint64_t)x[2];
        o.transmitted = (uint64_t)x[3]; o.left = (uint64_t)x[4]; o.timeout = (uint64_t)x[5];
        o.mean_coll = x[6]; o.mean_time = x[7];
        m[line.substr(0, tab)] = o;
    }
    return m;
}

}  // namespace

extern "C" {

int mq_scan_key(const char* geom_key, const mq_job* job, const double dc[4], const mq_physics* ph, int32_t path,
                char* buf, size_t len)
{
    if (!geom_key || !job || !dc || !ph || !buf) { mq_set_last_error("NULL argument"); return -1; }
    const std::string k = key_of(geom_key, *job, dc, *ph, path);
    if (k.size() + 1 > len) return (int)(k.size() + 1);
    std::memcpy(buf, k.c_str(), k.size() + 1);
    return 0;
}

int mq_format_f64(double x, int32_t display, char* buf, size_t len)
{
    if (!buf) { mq_set_last_error("NULL argument"); return -1; }
    const std::string s = display ? mq_rust_display_f64(x) : mq_rust_debug_f64(x);
    if (s.size() + 1 > len) return (int)(s.size() + 1);
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return 0;
}

// scan::cached_batch (scan.rs:31-47).
int mq_cached_batch(mq_engine* eng, const char* geom_key, const mq_job* jobs, size_t njobs, const double dc[4],
                    const mq_physics* ph, int32_t path, const char* cache_path, mq_outcome* out, size_t* n_computed)
{
    try {
        if (!eng || !geom_key || !jobs || !dc || !ph || !cache_path || !out) throw std::runtime_error("NULL argument");
        auto cache = load_cache(cache_path);
        std::vector<std::string> keys(njobs);
        std::vector<size_t> miss;
        for (size_t i = 0; i < njobs; ++i) {
            keys[i] = key_of(geom_key, jobs[i], dc, *ph, path);
            if (!cache.count(keys[i])) miss.push_back(i);
        }
        // duplicated keys inside one call are computed once
        std::vector<size_t> uniq;
        std::map<std::string, size_t> seen;
        for (size_t i : miss)
            if (!seen.count(keys[i])) { seen[keys[i]] = i; uniq.push_back(i); }
        if (!uniq.empty()) {
            std::vector<mq_job> mj;
            for (size_t i : uniq) mj.push_back(jobs[i]);
            std::vector<mq_outcome> o(mj.size());
            const int rc = mq_run_batch(eng, mj.data(), mj.size(), dc, ph, path, o.data());
            if (rc != 0) return rc;  // mq_last_error() holds the engine's message
            std::ofstream fh(cache_path, std::ios::app);
            if (!fh) throw std::runtime_error(std::string("cannot append to ") + cache_path);
            for (size_t k = 0; k < uniq.size(); ++k) {
                const mq_outcome& r = o[k];
                fh << keys[uniq[k]] << '\t' << r.n << ',' << r.crashed << ',' << r.returned << ',' << r.transmitted
                   << ',' << r.left << ',' << r.timeout << ',' << mq_rust_display_f64(r.mean_coll) << ','
                   << mq_rust_display_f64(r.mean_time) << '\n';
                cache[keys[uniq[k]]] = r;
            }
        }
        std::fprintf(stderr, "  [cache: %zu hits, %zu computed]\n", njobs - miss.size(), uniq.size());
        for (size_t i = 0; i < njobs; ++i) out[i] = cache[keys[i]];
        if (n_computed) *n_computed = uniq.size();
        return 0;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -1;
    }
}

// Refinement rule of scan::adaptive_scan (scan.rs:59-67): around every coarse point
// with T >= 0.35 T_max and T - base > 4 sigma, the frequencies f - step .. f + step
// in steps of step / 4 (rounded to 1e-4 kHz), sorted and deduplicated.
int mq_adaptive_fine_points(const double* f_khz, const double* t, size_t nf, uint64_t n_lo, double base,
                            double f_step, double* out, size_t max_out, size_t* n_out)
{
    if (!f_khz || !t || !n_out || n_lo == 0) { mq_set_last_error("bad argument"); return -1; }
    double tmax = 0.0;
    for (size_t i = 0; i < nf; ++i) tmax = std::max(tmax, t[i]);
    std::vector<double> fine;
    for (size_t i = 0; i < nf; ++i) {
        const double sig = std::sqrt(std::max(t[i], 1.0 / (double)n_lo) / (double)n_lo);
        if (t[i] >= 0.35 * tmax && t[i] - base > 4.0 * sig) {
            double g = f_khz[i] - f_step;
            while (g <= f_khz[i] + f_step + 1e-9) {
                fine.push_back(std::round(g * 1e4) / 1e4);
                g += f_step / 4.0;
            }
        }
    }
    std::sort(fine.begin(), fine.end());
    fine.erase(std::unique(fine.begin(), fine.end()), fine.end());
    *n_out = fine.size();
    if (out) {
        if (fine.size() > max_out) return (int)fine.size();
        std::copy(fine.begin(), fine.end(), out);
    }
    return 0;
}

// scan::adaptive_scan (scan.rs:52-78): coarse pass at f_step with n_lo ions
// (seeds seed0 + i), fine pass at f_step / 4 with n_hi ions (seeds seed0 +
// 100000 + i) near the resonance; returns (f_kHz, T, C, n) sorted by f.
int mq_adaptive_scan(mq_engine* eng, const char* geom_key, const mq_species* sp, const mq_rfac* rf,
                     const double dc[4], const mq_physics* ph, double f_lo, double f_hi, double f_step, uint64_t n_lo,
                     uint64_t n_hi, double base, uint64_t seed0, int32_t path, const char* cache_path,
                     mq_scan_point* out, size_t max_out, size_t* n_out)
{
    try {
        if (!eng || !sp || !rf || !n_out || !(f_step > 0.0)) throw std::runtime_error("bad argument");
        std::vector<double> fs;
        for (int i = 0;; ++i) {
            const double f = f_lo + f_step * (double)i;
            if (!(f <= f_hi + 1e-9)) break;
            fs.push_back(f);
        }
        std::vector<mq_job> j1;
        for (size_t i = 0; i < fs.size(); ++i) {
            mq_job j{};
            j.sp = *sp; j.drive = *rf; j.drive.f_ac = fs[i] * 1e3; j.n = n_lo; j.seed = seed0 + i;
            j.dc_scale = 1.0;
            j1.push_back(j);
        }
        std::vector<mq_outcome> o1(j1.size());
        if (!j1.empty()) {
            const int rc = mq_cached_batch(eng, geom_key, j1.data(), j1.size(), dc, ph, path, cache_path, o1.data(), nullptr);
            if (rc != 0) return rc;
        }
        std::vector<double> t1(fs.size());
        for (size_t i = 0; i < fs.size(); ++i) t1[i] = o1[i].n ? (double)o1[i].transmitted / (double)o1[i].n : 0.0;
        size_t nfine = 0;
        mq_adaptive_fine_points(fs.data(), t1.data(), fs.size(), n_lo, base, f_step, nullptr, 0, &nfine);
        std::vector<double> fine(nfine);
        mq_adaptive_fine_points(fs.data(), t1.data(), fs.size(), n_lo, base, f_step, fine.data(), nfine, &nfine);
        std::vector<mq_job> j2;
        for (size_t i = 0; i < fine.size(); ++i) {
            mq_job j{};
            j.sp = *sp; j.drive = *rf; j.drive.f_ac = fine[i] * 1e3; j.n = n_hi; j.seed = seed0 + 100000 + i;
            j.dc_scale = 1.0;
            j2.push_back(j);
        }
        std::vector<mq_outcome> o2(j2.size());
        if (!j2.empty()) {
            const int rc = mq_cached_batch(eng, geom_key, j2.data(), j2.size(), dc, ph, path, cache_path, o2.data(), nullptr);
            if (rc != 0) return rc;
        }
        std::vector<mq_scan_point> pts;
        for (size_t i = 0; i < fs.size(); ++i) {
            bool in_fine = false;
            for (double g : fine) in_fine = in_fine || std::fabs(g - fs[i]) < 1e-6;
            if (in_fine) continue;
            const mq_outcome& o = o1[i];
            pts.push_back({fs[i], o.n ? (double)o.transmitted / o.n : 0.0, o.n ? (double)o.crashed / o.n : 0.0, n_lo});
        }
        for (size_t i = 0; i < fine.size(); ++i) {
            const mq_outcome& o = o2[i];
            pts.push_back({fine[i], o.n ? (double)o.transmitted / o.n : 0.0, o.n ? (double)o.crashed / o.n : 0.0, n_hi});
        }
        std::stable_sort(pts.begin(), pts.end(), [](const mq_scan_point& a, const mq_scan_point& b) { return a.f_khz < b.f_khz; });
        const double cost_full = (double)fs.size() * 4.0 * (double)n_hi;
        const double cost = (double)fs.size() * (double)n_lo + (double)fine.size() * (double)n_hi;
        std::fprintf(stderr, "  [adaptive scan: %zu coarse x %llu + %zu fine x %llu ions = %.0f%% of a uniform fine scan]\n",
                     fs.size(), (unsigned long long)n_lo, fine.size(), (unsigned long long)n_hi,
                     cost_full > 0 ? 100.0 * cost / cost_full : 0.0);
        *n_out = pts.size();
        if (out) {
            if (pts.size() > max_out) return (int)pts.size();
            std::copy(pts.begin(), pts.end(), out);
        }
        return 0;
    } catch (const std::exception& e) {
        mq_set_last_error(e.what());
        return -1;
    }
}

}  // extern "C"
