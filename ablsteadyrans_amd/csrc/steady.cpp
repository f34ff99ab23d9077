// steady.cpp -- the steady criterion of the pseudo-time loop (host only):
// residual and monitor histories recorded one iteration at a time and the
// verdict "steady", "not steady" or "diverged" (Source/SteadyCriterion.cpp),
// plus the two pieces of the driver's step that read it: the relative
// residual r_n and the next CFL number (Source/Driver.cpp:325-334, 421-437).
//
// Every comparison tests finiteness first, as the reference does, so no
// ordered comparison ever sees a NaN. Messages keep the reference's text,
// numbers printed with three significant digits.
#include "errors.hpp"
#include "gust_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace gust_amd {

namespace {

std::string fmt(double x)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.3g", x);
    return b;
}

std::vector<std::string> validate(const gust_steady_params& p)
{
    std::vector<std::string> err;
    if (!std::isfinite(p.residual_tol) || !(p.residual_tol > 0)) {
        err.emplace_back("steady.residual_tol must be finite and > 0");
    }
    // 10^-orders underflows to 0 above max_exponent10 orders.
    const int emax = std::numeric_limits<double>::max_exponent10;
    if (!std::isfinite(p.residual_drop_orders) || !(p.residual_drop_orders >= 0) ||
        !(p.residual_drop_orders <= emax)) {
        err.emplace_back("steady.residual_drop_orders must be >= 0 and <= " + std::to_string(emax));
    }
    if (p.residual_ref_iterations < 1) { err.emplace_back("steady.residual_ref_iterations must be >= 1"); }
    if (!std::isfinite(p.monitor_rel_tol) || !(p.monitor_rel_tol > 0)) {
        err.emplace_back("steady.monitor_rel_tol must be finite and > 0");
    }
    if (p.monitor_window < 2) { err.emplace_back("steady.monitor_window must be >= 2"); }
    return err;
}

void put_text(const std::string& s, char* buf, int64_t cap)
{
    if (!buf || cap <= 0) { return; }
    const size_t n = std::min<size_t>(s.size(), size_t(cap - 1));
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
}

struct Monitor {
    std::string name;
    double scale;
};

struct History {
    int count = 0;
    bool nonfinite = false;
    std::string nonfinite_what;
    std::vector<double> res_ref, res_last;
    std::vector<std::vector<double>> mon_hist;   // the last window + 1 values, oldest first
};

struct Verdict {
    int status = GUST_STEADY_NOT_STEADY;
    int iteration = -1;
    std::vector<std::string> reasons;
    bool residuals_met = false, monitors_met = false;
};

}   // namespace

struct Steady {
    gust_steady_params p;
    std::vector<std::string> res_names;
    std::vector<Monitor> mon;
    History h;

    void add(const double* r, int nr, const double* m, int nm)
    {
        if (nr != int(res_names.size()) || nm != int(mon.size())) {
            GUST_FAIL("gust: SteadyCriterion::add got %d residuals and %d monitors, expected %d and %d", nr, nm,
                      int(res_names.size()), int(mon.size()));
        }
        for (int q = 0; q < nr; ++q) {
            if (!std::isfinite(r[q])) {
                if (!h.nonfinite) { h.nonfinite_what = "residual '" + res_names[q] + "'"; }
                h.nonfinite = true;
                continue;
            }
            if (h.count < p.residual_ref_iterations) { h.res_ref[q] = std::max(h.res_ref[q], std::fabs(r[q])); }
            h.res_last[q] = std::fabs(r[q]);
        }
        const size_t keep = size_t(p.monitor_window) + 1;
        for (int k = 0; k < nm; ++k) {
            if (!std::isfinite(m[k])) {
                if (!h.nonfinite) { h.nonfinite_what = "monitor '" + mon[k].name + "'"; }
                h.nonfinite = true;
                continue;
            }
            auto& hist = h.mon_hist[k];
            hist.push_back(m[k]);
            if (hist.size() > keep) { hist.erase(hist.begin()); }
        }
        ++h.count;
    }

    Verdict verdict() const
    {
        Verdict v;
        v.iteration = h.count - 1;
        if (h.count == 0) {
            v.reasons.emplace_back("no iterations recorded");
            return v;
        }
        if (h.nonfinite) {
            v.status = GUST_STEADY_DIVERGED;
            v.reasons.emplace_back(h.nonfinite_what + " became NaN or infinite");
            return v;
        }
        // Residuals: below the tolerance and dropped the required orders below
        // the reference (floored at the tolerance).
        const double drop = std::pow(10.0, -p.residual_drop_orders);
        for (size_t q = 0; q < res_names.size(); ++q) {
            const double r = h.res_last[q];
            const double ref = std::max(h.res_ref[q], p.residual_tol);
            if (!(r < p.residual_tol)) {
                v.reasons.emplace_back("residual '" + res_names[q] + "' = " + fmt(r) +
                                       " is not below steady.residual_tol = " + fmt(p.residual_tol));
            }
            if (!(r <= ref * drop)) {
                const double orders = (r > 0) ? std::log10(ref / r) : 0.0;
                v.reasons.emplace_back("residual '" + res_names[q] + "' dropped " + fmt(orders) +
                                       " orders, less than steady.residual_drop_orders = " +
                                       fmt(p.residual_drop_orders));
            }
        }
        v.residuals_met = v.reasons.empty();
        const size_t residual_reasons = v.reasons.size();

        // Monitors: a full window whose values all lie within the tolerance of
        // the latest, and an Aitken estimate of the remaining distance to the
        // limit, from the changes over the window's two halves, within it too.
        const size_t need = size_t(p.monitor_window) + 1;
        for (size_t k = 0; k < mon.size(); ++k) {
            const auto& hist = h.mon_hist[k];
            if (hist.size() < need) {
                v.reasons.emplace_back("monitor '" + mon[k].name + "': " + std::to_string(hist.size()) +
                                       " of the " + std::to_string(need) +
                                       " values steady.monitor_window needs");
                continue;
            }
            const double last = hist.back();
            double change = 0;
            for (const double x : hist) { change = std::max(change, std::fabs(x - last)); }
            const double ref = std::max(std::fabs(last), mon[k].scale);
            const double tol = p.monitor_rel_tol * ref;
            if (!(change <= tol)) {
                v.reasons.emplace_back("monitor '" + mon[k].name + "' changed by " + fmt(change / ref) +
                                       " of max(|value|, scale) over the last " +
                                       std::to_string(p.monitor_window) +
                                       " iterations; steady.monitor_rel_tol = " + fmt(p.monitor_rel_tol));
                continue;
            }
            const size_t W = hist.size() - 1;
            const size_t half = W / 2;
            const double d1 = hist[W - half] - hist[W - 2 * half];
            const double d2 = hist[W] - hist[W - half];
            const double noise = 1.0e-3 * tol;   // round-off says nothing about the rate
            if (std::max(std::fabs(d1), std::fabs(d2)) <= noise) { continue; }
            const double rho = (d1 != 0) ? d2 / d1 : std::numeric_limits<double>::infinity();
            if (!(rho > -1 && rho < 1)) {
                v.reasons.emplace_back("monitor '" + mon[k].name +
                                       "' is not converging over the window (half-window change ratio " +
                                       fmt(rho) + ")");
                continue;
            }
            const double remaining = std::fabs(d2 * rho / (1 - rho));
            if (!(remaining <= tol)) {
                v.reasons.emplace_back("monitor '" + mon[k].name + "' is an extrapolated " +
                                       fmt(remaining / ref) +
                                       " of max(|value|, scale) from its limit; steady.monitor_rel_tol = " +
                                       fmt(p.monitor_rel_tol));
            }
        }
        v.monitors_met = v.reasons.size() == residual_reasons;
        v.status = v.reasons.empty() ? GUST_STEADY_STEADY : GUST_STEADY_NOT_STEADY;
        return v;
    }

    std::vector<double> thresholds() const
    {
        const double drop = std::pow(10.0, -p.residual_drop_orders);
        std::vector<double> t(res_names.size(), p.residual_tol);
        for (size_t q = 0; q < t.size(); ++q) {
            t[q] = std::min(p.residual_tol, std::max(h.res_ref[q], p.residual_tol) * drop);
        }
        return t;
    }
};

namespace {

// Little-endian byte image of a History.
struct Writer {
    std::vector<unsigned char> b;
    void i32(int32_t x) { raw(&x, sizeof x); }
    void f64(double x) { raw(&x, sizeof x); }
    void raw(const void* p, size_t n)
    {
        const auto* c = static_cast<const unsigned char*>(p);
        b.insert(b.end(), c, c + n);
    }
};

struct Reader {
    const unsigned char* p;
    size_t n, at = 0;
    bool ok = true;
    void raw(void* out, size_t k)
    {
        if (!ok || at + k > n) {
            ok = false;
            std::memset(out, 0, k);
            return;
        }
        std::memcpy(out, p + at, k);
        at += k;
    }
    int32_t i32()
    {
        int32_t x;
        raw(&x, sizeof x);
        return x;
    }
    double f64()
    {
        double x;
        raw(&x, sizeof x);
        return x;
    }
};

std::vector<unsigned char> save(const History& h)
{
    Writer w;
    w.i32(h.count);
    w.i32(h.nonfinite ? 1 : 0);
    w.i32(int32_t(h.nonfinite_what.size()));
    w.raw(h.nonfinite_what.data(), h.nonfinite_what.size());
    w.i32(int32_t(h.res_ref.size()));
    for (const double x : h.res_ref) { w.f64(x); }
    for (const double x : h.res_last) { w.f64(x); }
    w.i32(int32_t(h.mon_hist.size()));
    for (const auto& m : h.mon_hist) {
        w.i32(int32_t(m.size()));
        for (const double x : m) { w.f64(x); }
    }
    return w.b;
}

bool load(const void* buf, int64_t size, History& h)
{
    if (!buf || size < 0) { return false; }
    Reader r{static_cast<const unsigned char*>(buf), size_t(size)};
    h.count = r.i32();
    h.nonfinite = r.i32() != 0;
    const int32_t wl = r.i32();
    if (!r.ok || wl < 0 || size_t(wl) > r.n - r.at) { return false; }
    h.nonfinite_what.assign(reinterpret_cast<const char*>(r.p + r.at), size_t(wl));
    r.at += size_t(wl);
    const int32_t nres = r.i32();
    if (!r.ok || nres < 0 || size_t(nres) * 16 > r.n - r.at) { return false; }
    h.res_ref.resize(size_t(nres));
    h.res_last.resize(size_t(nres));
    for (auto& x : h.res_ref) { x = r.f64(); }
    for (auto& x : h.res_last) { x = r.f64(); }
    const int32_t nmon = r.i32();
    if (!r.ok || nmon < 0 || size_t(nmon) * 4 > r.n - r.at) { return false; }
    h.mon_hist.assign(size_t(nmon), {});
    for (auto& m : h.mon_hist) {
        const int32_t k = r.i32();
        if (!r.ok || k < 0 || size_t(k) * 8 > r.n - r.at) { return false; }
        m.resize(size_t(k));
        for (auto& x : m) { x = r.f64(); }
    }
    return r.ok && r.at == r.n;
}

// Grid sequencing's loose stop (Source/Sequencing.cpp:196-252): after the
// last recorded iteration, (a) every residual drop_orders below its floored
// reference or below residual_tol, once the level's hold is over (a
// released seeded level: min(3, residual_ref_iterations) iterations); or
// (b) with monitor_tol > 0 and after residual_ref_iterations iterations,
// every monitor's full window within monitor_tol of max(|latest|, scale).
// Never on an empty or non-finite history.
std::string fmt_e(double x)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.3e", x);
    return b;
}

constexpr int kReleasedMinIterations = 3;

bool loose_stop(const Steady& c, const gust_sequencing_options& o, int hold, std::string& reason)
{
    const History& h = c.h;
    const gust_steady_params& p = c.p;
    if (h.count == 0 || h.nonfinite) { return false; }
    const int first = hold > 0 ? hold : std::min(p.residual_ref_iterations, kReleasedMinIterations);
    if (h.count >= first && !h.res_last.empty()) {
        const double drop = std::pow(10.0, -o.drop_orders);
        bool all = true;
        for (size_t q = 0; q < h.res_last.size(); ++q) {
            const double ref = std::max(h.res_ref[q], p.residual_tol);
            const double r = h.res_last[q];
            all = all && (r <= ref * drop || r < p.residual_tol);
        }
        if (all) {
            reason = "every residual " + fmt_e(o.drop_orders) + " orders below its reference";
            return true;
        }
    }
    const size_t need = size_t(p.monitor_window) + 1;
    if (o.monitor_tol > 0 && h.count >= p.residual_ref_iterations && !c.mon.empty() &&
        h.mon_hist.size() == c.mon.size()) {
        bool all = true;
        for (size_t m = 0; m < c.mon.size() && all; ++m) {
            const auto& w = h.mon_hist[m];
            if (w.size() < need) {
                all = false;
                break;
            }
            const double last = w.back();
            double change = 0;
            for (const double x : w) { change = std::max(change, std::fabs(x - last)); }
            all = change <= o.monitor_tol * std::max(std::fabs(last), c.mon[m].scale);
        }
        if (all) {
            reason = "every monitor within " + fmt_e(o.monitor_tol) + " over " + std::to_string(p.monitor_window) +
                     " iterations";
            return true;
        }
    }
    return false;
}

// Source/Sequencing.cpp:118-176 (check_sequencing_options): each problem
// names its key; the ranges are checked only with sequencing on.
std::vector<std::string> check_sequencing(const gust_sequencing_options& o, const gust_steady_params& strict,
                                          int levels, int nx, int ny)
{
    std::vector<std::string> err;
    if (o.enabled) {
        if (!std::isfinite(o.drop_orders) || !(o.drop_orders > 0) ||
            !(std::isfinite(strict.residual_drop_orders) && o.drop_orders <= strict.residual_drop_orders)) {
            err.emplace_back("steady.sequencing_drop_orders = " + fmt_e(o.drop_orders) +
                             ": must be > 0 and <= steady.residual_drop_orders = " +
                             fmt_e(strict.residual_drop_orders) +
                             " (the loose stop of the coarse levels is not stricter than the strict one)");
        }
        if (!std::isfinite(o.monitor_tol) || !(o.monitor_tol >= 0) || !(o.monitor_tol < 1)) {
            err.emplace_back("steady.sequencing_monitor_tol = " + fmt_e(o.monitor_tol) +
                             ": must be 0 (off) or > 0 and < 1");
        }
    }
    if (o.enabled && levels < 2) {
        err.emplace_back("steady.sequencing = 1 needs mesh.hierarchy_levels >= 2 (it is " + std::to_string(levels) +
                         ")");
    }
    if (o.enabled && levels >= 2 && levels <= GUST_MAX_HIERARCHY_LEVELS) {
        const int r = 1 << (levels - 1);
        const int cx = nx / r, cy = ny / r;
        if (cx < 2 || cy < 2) {
            err.emplace_back("steady.sequencing = 1 needs at least 2 cells in x and in y on the coarsest level: "
                             "mesh.n_cell x and y / 2^(mesh.hierarchy_levels - 1) = " +
                             std::to_string(cx) + " " + std::to_string(cy) + "; use fewer levels");
        }
    }
    if (o.enabled && o.cfl_previous && o.hold == 1) {
        err.emplace_back("steady.sequencing_cfl_start = previous needs steady.sequencing_hold = 0 "
                         "(the hold runs the first steady.residual_ref_iterations iterations at steady.cfl0)");
    }
    if (!o.enabled) {
        const char* keys[4] = {"sequencing_drop_orders", "sequencing_monitor_tol", "sequencing_hold",
                               "sequencing_cfl_start"};
        for (int b = 0; b < 4; ++b) {
            if (o.given & (1u << b)) {
                err.emplace_back(std::string("steady.") + keys[b] + " has no effect unless steady.sequencing = 1");
            }
        }
    }
    return err;
}

Steady* as(gust_steady* c) { return reinterpret_cast<Steady*>(c); }
const Steady* as(const gust_steady* c) { return reinterpret_cast<const Steady*>(c); }

const Steady& need(const gust_steady* c, const char* op)
{
    if (!c) { GUST_FAIL("gust: %s: null criterion", op); }
    return *as(c);
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_steady_default_params(gust_steady_params* p)
{
    return guarded([&] {
        if (!p) { GUST_FAIL("gust: SteadyParams: null output"); }
        p->residual_tol = 5.0e-5;
        p->residual_drop_orders = 4;
        p->residual_ref_iterations = 5;
        p->monitor_rel_tol = 1.0e-3;
        p->monitor_window = 50;
    });
}

int gust_steady_validate(const gust_steady_params* p, int32_t* count, char* buf, int64_t cap)
{
    return guarded([&] {
        if (!p || !count) { GUST_FAIL("gust: validate: null argument"); }
        const auto err = validate(*p);
        std::string all;
        for (size_t e = 0; e < err.size(); ++e) { all += (e ? "\n" : "") + err[e]; }
        *count = int32_t(err.size());
        put_text(all, buf, cap);
    });
}

int gust_steady_create(const gust_steady_params* p, int32_t nres, const char* const* residual_names, int32_t nmon,
                       const char* const* monitor_names, const double* monitor_scales, gust_steady** out)
{
    return guarded([&] {
        if (!p || !out || nres < 0 || nmon < 0 || (nres > 0 && !residual_names) ||
            (nmon > 0 && (!monitor_names || !monitor_scales))) {
            GUST_FAIL("gust: SteadyCriterion: null argument");
        }
        *out = nullptr;
        auto err = validate(*p);
        if (nres == 0) { err.emplace_back("SteadyCriterion needs at least one residual"); }
        for (int m = 0; m < nmon; ++m) {
            const double s = monitor_scales[m];
            if (!std::isfinite(s) || !(s > 0)) {
                err.emplace_back(std::string("monitor '") + (monitor_names[m] ? monitor_names[m] : "") +
                                 "': scale must be finite and > 0");
            }
        }
        if (!err.empty()) {
            std::string all = "gust: invalid steady criterion:";
            for (const auto& e : err) { all += "\n  " + e; }
            GUST_FAIL("%s", all.c_str());
        }
        auto* c = new Steady;
        c->p = *p;
        for (int q = 0; q < nres; ++q) { c->res_names.emplace_back(residual_names[q] ? residual_names[q] : ""); }
        for (int m = 0; m < nmon; ++m) {
            c->mon.push_back({monitor_names[m] ? monitor_names[m] : "", monitor_scales[m]});
        }
        c->h.res_ref.assign(size_t(nres), 0.0);
        c->h.res_last.assign(size_t(nres), 0.0);
        c->h.mon_hist.assign(size_t(nmon), {});
        *out = reinterpret_cast<gust_steady*>(c);
    });
}

int gust_steady_destroy(gust_steady* c)
{
    delete as(c);
    return 0;
}

int gust_steady_add(gust_steady* c, const double* residuals, int32_t nres, const double* monitors, int32_t nmon)
{
    return guarded([&] {
        need(c, "SteadyCriterion::add");
        if ((nres > 0 && !residuals) || (nmon > 0 && !monitors)) {
            GUST_FAIL("gust: SteadyCriterion::add: null values");
        }
        as(c)->add(residuals, nres, monitors, nmon);
    });
}

int gust_steady_verdict(const gust_steady* c, gust_steady_result* v, char* reasons, int64_t cap)
{
    return guarded([&] {
        const Steady& s = need(c, "SteadyCriterion::verdict");
        if (!v) { GUST_FAIL("gust: SteadyCriterion::verdict: null output"); }
        const Verdict r = s.verdict();
        std::string all;
        for (size_t e = 0; e < r.reasons.size(); ++e) { all += (e ? "\n" : "") + r.reasons[e]; }
        v->status = r.status;
        v->iteration = r.iteration;
        v->residuals_met = r.residuals_met ? 1 : 0;
        v->monitors_met = r.monitors_met ? 1 : 0;
        v->nreasons = int32_t(r.reasons.size());
        v->reasons_len = int64_t(all.size());
        put_text(all, reasons, cap);
    });
}

int gust_steady_iterations(const gust_steady* c, int32_t* n)
{
    return guarded([&] {
        const Steady& s = need(c, "SteadyCriterion::iterations");
        if (!n) { GUST_FAIL("gust: SteadyCriterion::iterations: null output"); }
        *n = s.h.count;
    });
}

int gust_steady_residual_thresholds(const gust_steady* c, double* out, int32_t n)
{
    return guarded([&] {
        const Steady& s = need(c, "SteadyCriterion::residual_thresholds");
        if (!out || n != int32_t(s.res_names.size())) {
            GUST_FAIL("gust: SteadyCriterion::residual_thresholds: needs room for %d values",
                      int(s.res_names.size()));
        }
        const auto t = s.thresholds();
        std::copy(t.begin(), t.end(), out);
    });
}

int gust_steady_references(const gust_steady* c, double* out, int32_t n)
{
    return guarded([&] {
        const Steady& s = need(c, "SteadyCriterion::history");
        if (!out || n != int32_t(s.h.res_ref.size())) {
            GUST_FAIL("gust: SteadyCriterion::history: %d references asked for, the criterion has %d", int(n),
                      int(s.h.res_ref.size()));
        }
        for (int q = 0; q < n; ++q) { out[q] = s.h.res_ref[size_t(q)]; }
    });
}

int gust_steady_seed_reference(gust_steady* c, const double* ref, int32_t n)
{
    return guarded([&] {
        need(c, "SteadyCriterion::seed_reference");
        Steady& s = *as(c);
        bool ok = s.h.count == 0 && n == int32_t(s.res_names.size()) && (n == 0 || ref);
        for (int q = 0; ok && q < n; ++q) { ok = std::isfinite(ref[q]) && ref[q] >= 0; }
        if (!ok) {
            GUST_FAIL("gust: SteadyCriterion::seed_reference: needs one finite reference >= 0 per residual, "
                      "before the first iteration");
        }
        s.h.res_ref.assign(ref, ref + n);
    });
}

int gust_steady_history_save(const gust_steady* c, void* buf, int64_t cap, int64_t* size)
{
    return guarded([&] {
        const Steady& s = need(c, "SteadyCriterion::history");
        if (!size) { GUST_FAIL("gust: SteadyCriterion::history: null size"); }
        const auto img = save(s.h);
        *size = int64_t(img.size());
        if (buf && cap >= int64_t(img.size())) { std::memcpy(buf, img.data(), img.size()); }
    });
}

int gust_steady_history_restore(gust_steady* c, const void* buf, int64_t size)
{
    return guarded([&] {
        need(c, "SteadyCriterion::restore");
        Steady& s = *as(c);
        History h;
        const size_t keep = size_t(s.p.monitor_window) + 1;
        bool ok = load(buf, size, h) && h.count >= 0 && h.res_ref.size() == s.res_names.size() &&
                  h.res_last.size() == s.res_names.size() && h.mon_hist.size() == s.mon.size();
        for (const auto& m : h.mon_hist) { ok = ok && m.size() <= keep; }
        if (!ok) {
            GUST_FAIL("gust: SteadyCriterion::restore: the saved history does not match this criterion's "
                      "residuals, monitors or steady.monitor_window");
        }
        s.h = h;
    });
}

int gust_steady_relative_residual(const gust_steady* c, const double* residuals, int32_t nres, int32_t hold,
                                  double previous, double* r)
{
    return guarded([&] {
        const Steady& s = need(c, "Driver::step");
        if (!r || nres != int32_t(s.res_names.size()) || (nres > 0 && !residuals)) {
            GUST_FAIL("gust: Driver::step: needs one residual per criterion residual");
        }
        double out = 1;
        if (s.h.nonfinite) {
            out = previous;   // no arithmetic on a NaN or infinity
        } else if (s.h.count > hold) {
            out = 0;
            for (int q = 0; q < nres; ++q) {
                const double ref = std::max(s.h.res_ref[q], s.p.residual_tol);
                out = std::max(out, std::fabs(residuals[q]) / ref);
            }
        }
        *r = out;
    });
}

int gust_steady_next_cfl(const gust_steady* c, int32_t hold, double cfl, double r, double cfl0, double cfl_max,
                         double ser_exponent, double max_growth, double* out)
{
    return guarded([&] {
        const Steady& s = need(c, "Driver::next_cfl");
        if (!out) { GUST_FAIL("gust: Driver::next_cfl: null output"); }
        if (s.h.count <= hold) {
            *out = cfl0;
            return;
        }
        const int rc = gust_ser_cfl(cfl, r, cfl0, cfl_max, ser_exponent, max_growth, out);
        if (rc != 0) { throw Fail{rc}; }
    });
}

int gust_sequencing_default_options(gust_sequencing_options* o)
{
    return guarded([&] {
        if (!o) { GUST_FAIL("gust: SequencingOptions: null output"); }
        std::memset(o, 0, sizeof *o);
        o->drop_orders = 2;
    });
}

int gust_check_sequencing_options(const gust_sequencing_options* o, const gust_steady_params* strict,
                                  int32_t levels, int32_t nx, int32_t ny, int32_t* count, char* buf, int64_t cap)
{
    return guarded([&] {
        if (!o || !strict || !count) { GUST_FAIL("gust: check_sequencing_options: null argument"); }
        const auto err = check_sequencing(*o, *strict, levels, nx, ny);
        std::string all;
        for (size_t n = 0; n < err.size(); ++n) { all += (n ? "\n" : "") + err[n]; }
        *count = int32_t(err.size());
        put_text(all, buf, cap);
    });
}

int gust_loose_stop(const gust_steady* c, const gust_sequencing_options* o, int32_t hold_iterations, int32_t* stop,
                    char* reason, int64_t cap)
{
    return guarded([&] {
        const Steady& s = need(c, "loose_stop");
        if (!o || !stop) { GUST_FAIL("gust: loose_stop: null argument"); }
        std::string why;
        *stop = loose_stop(s, *o, hold_iterations, why) ? 1 : 0;
        put_text(why, reason, cap);
    });
}

}   // extern "C"
