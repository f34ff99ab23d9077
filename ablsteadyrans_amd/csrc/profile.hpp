// profile.hpp -- the inflow profile by height above the ground, shared by
// the initializer and the boundary fills: the device evaluation
// (Source/InflowProfile.H:137-200), the option checks (Source/
// InflowProfile.cpp:14-47, the table's rules 49-118, the sounding's
// Source/Forcing.cpp:30-49) and the view's constants (Source/
// InflowProfile.cpp:178-205) with the tables in device memory. The upload is
// cached per (device, stream): a steady loop that passes the same profile
// every iteration copies its tables once.
#pragma once

#include "common.hpp"

#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

namespace gust_amd::profile {

struct ProfileView {
    int kind;   // 0 log_law, 1 power_law, 2 table
    double ustar, z0, speed_ref, z_ref, alpha, cosd, sind, Cmu0, kappa, inv_lmax;
    const double *tz, *tu, *tv, *tt, *tk;
    int n, has_k;
    const double *rz, *rt;
    int rn;
};

__device__ __forceinline__ double dmaxd(double a, double b) { return (a > b) ? a : b; }

// Linear in a table of n increasing abscissae, constant beyond the ends
// (InflowProfile.H:137-147).
__device__ inline double table_interp(const double* x, const double* y, int n, double xq)
{
    if (n == 1 || xq <= x[0]) { return y[0]; }
    if (xq >= x[n - 1]) { return y[n - 1]; }
    int m = 1;
    while (m < n - 1 && !(xq < x[m])) { ++m; }
    const double w = (xq - x[m - 1]) / (x[m] - x[m - 1]);
    return y[m - 1] + w * (y[m] - y[m - 1]);
}

// The profile at height z above the ground (InflowProfile.H:149-200): the
// wind, theta, the equilibrium k of the k-L closure and nu_t = l u*_l.
__device__ inline void profile_eval(const ProfileView& p, double z, double& u, double& v, double& theta, double& k,
                             double& nut)
{
    z = dmaxd(z, 0.0);
    // l / l_raw = 1 / (1 + l_raw / l_max), exactly 1 when unbounded
    const double c = 1.0 / (1.0 + p.kappa * (z + p.z0) * p.inv_lmax);
    double ustar_l;
    if (p.kind == 2) {
        u = table_interp(p.tz, p.tu, p.n, z);
        v = table_interp(p.tz, p.tv, p.n, z);
        theta = table_interp(p.tz, p.tt, p.n, z);
        int m = 1;
        while (m < p.n - 1 && !(z < p.tz[m])) { ++m; }
        const double s0 = sqrt(p.tu[m - 1] * p.tu[m - 1] + p.tv[m - 1] * p.tv[m - 1]);
        const double s1 = sqrt(p.tu[m] * p.tu[m] + p.tv[m] * p.tv[m]);
        const double dsdz = (p.n > 1) ? (s1 - s0) / (p.tz[m] - p.tz[m - 1]) : 0.0;
        ustar_l = p.kappa * (z + p.z0) * dmaxd(dsdz, 0.0) * c;
        if (p.has_k) {
            k = table_interp(p.tz, p.tk, p.n, z);
            ustar_l = p.Cmu0 * sqrt(dmaxd(k, 0.0));
        } else {
            k = ustar_l * ustar_l / (p.Cmu0 * p.Cmu0);
        }
        nut = p.kappa * ustar_l * (z + p.z0) * c;
        return;
    }
    double S;
    if (p.kind == 0) {
        S = p.ustar / p.kappa * log((z + p.z0) / p.z0);
        ustar_l = p.ustar * c;
    } else {
        S = p.speed_ref * pow(z / p.z_ref, p.alpha);
        // the slope at z0 below z0: max(z, z0) written in arithmetic
        const double dsdz = p.alpha * S / (0.5 * (z + p.z0 + fabs(z - p.z0)));
        ustar_l = p.kappa * (z + p.z0) * dsdz * c;
    }
    u = S * p.cosd;
    v = S * p.sind;
    theta = table_interp(p.rz, p.rt, p.rn, z);
    k = ustar_l * ustar_l / (p.Cmu0 * p.Cmu0);
    nut = p.kappa * ustar_l * (z + p.z0) * c;
}


inline bool finite_pos(double x) { return !std::isnan(x) && x > 0 && std::isfinite(x); }

// The problems with a profile, each naming its key.
inline std::vector<std::string> problems(const gust_inflow_profile& p)
{
    std::vector<std::string> err;
    auto num = [](double x) { return std::to_string(x); };
    if (!finite_pos(p.z0)) { err.push_back("bc.inflow_z0 = " + num(p.z0) + ": must be finite and > 0"); }
    if (p.kind == GUST_INFLOW_LOG_LAW) {
        if (!finite_pos(p.ustar)) { err.push_back("bc.inflow_ustar = " + num(p.ustar) + ": must be finite and > 0"); }
    } else if (p.kind == GUST_INFLOW_POWER_LAW) {
        if (std::isnan(p.speed_ref) || !(p.speed_ref >= 0 && std::isfinite(p.speed_ref))) {
            err.push_back("bc.inflow_speed = " + num(p.speed_ref) + ": must be finite and >= 0");
        }
        if (!finite_pos(p.z_ref)) { err.push_back("bc.inflow_zref = " + num(p.z_ref) + ": must be finite and > 0"); }
        if (!finite_pos(p.alpha)) { err.push_back("bc.inflow_alpha = " + num(p.alpha) + ": must be finite and > 0"); }
    } else if (p.kind == GUST_INFLOW_TABLE) {
        if (p.n < 2 || !p.z || !p.u || !p.v || !p.theta) {
            err.push_back("bc.inflow_file: the table needs at least two rows of z u v theta [k]");
        } else {
            for (int r = 0; r < p.n; ++r) {
                const double vals[5] = {p.z[r], p.u[r], p.v[r], p.theta[r], p.k ? p.k[r] : 0.0};
                bool fin = true;
                for (double x : vals) { fin = fin && std::isfinite(x); }
                const std::string row = std::to_string(r + 1);
                if (!fin) {
                    err.push_back("bc.inflow_file: row " + row + " holds a value that is not finite");
                } else if (r == 0 && p.z[0] != 0) {
                    err.push_back("bc.inflow_file: the first height must be 0 (the ground)");
                } else if (r > 0 && !(p.z[r] > p.z[r - 1])) {
                    err.push_back("bc.inflow_file: heights must increase strictly (row " + row + ")");
                } else if (!(p.theta[r] > 0)) {
                    err.push_back("bc.inflow_file: theta must be > 0 (row " + row + ")");
                } else if (p.k && p.k[r] < 0) {
                    err.push_back("bc.inflow_file: k must be >= 0 (row " + row + ")");
                }
                if (!err.empty()) { break; }
            }
        }
    } else {
        err.push_back("bc.inflow_profile: must be log_law, power_law or a table");
    }
    if (!std::isfinite(p.direction)) {
        err.push_back("bc.inflow_direction = " + num(p.direction) + ": must be finite (degrees)");
    }
    if (!finite_pos(p.Cmu0)) { err.push_back("bc.inflow_Cmu0 = " + num(p.Cmu0) + ": must be finite and > 0"); }
    if (!finite_pos(p.kappa)) { err.push_back("bc.inflow_kappa = " + num(p.kappa) + ": must be finite and > 0"); }
    if (!p.unbounded && !finite_pos(p.l_max)) {
        err.push_back("turb.max_geom_lscale = " + num(p.l_max) +
                      ": must be finite and > 0 (the length cap of bc.inflow_tke_length = closure)");
    }
    if (p.n_theta_ref < 1 || p.n_theta_ref > 64) {
        err.push_back("reference profile: z and theta need the same number of points, at least one (at most 64)");
    } else {
        for (int m = 0; m < p.n_theta_ref; ++m) {
            const double z = p.theta_ref_z[m], t = p.theta_ref[m];
            if (!std::isfinite(z) || std::isnan(t) || !(t > 0 && std::isfinite(t))) {
                err.push_back("reference profile point " + std::to_string(m) +
                              ": z must be finite and theta finite and > 0");
                break;
            }
            if (m > 0 && !(z > p.theta_ref_z[m - 1])) {
                err.push_back("reference profile: heights must increase strictly (point " + std::to_string(m) + ")");
                break;
            }
        }
    }
    return err;
}

// The profile's view with its tables in device memory, for kernels on st.
// The tables are uploaded when the profile differs from the last one seen
// on this (device, stream), with a synchronisation; otherwise the previous
// upload is reused.
// The device copies of the profile tables, one per (device, stream),
// re-uploaded when a stream sees other tables.
struct TableEntry {
    std::vector<double> key;
    double* dev = nullptr;
    size_t cap = 0;
};
inline std::mutex& table_mutex()
{
    static std::mutex mu;
    return mu;
}
inline std::map<std::pair<int, hipStream_t>, TableEntry>& table_map()
{
    static std::map<std::pair<int, hipStream_t>, TableEntry> m;
    return m;
}
inline size_t table_bytes(int device, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(table_mutex());
    auto it = table_map().find({device, st});
    return it == table_map().end() ? 0 : it->second.cap * sizeof(double);
}
inline void table_drop(int device, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(table_mutex());
    auto it = table_map().find({device, st});
    if (it == table_map().end()) { return; }
    if (it->second.dev) { (void)hipFree(it->second.dev); }
    table_map().erase(it);
}

inline ProfileView device_view(hipStream_t st, const gust_inflow_profile& p)
{
    const int n = p.kind == GUST_INFLOW_TABLE ? p.n : 0;
    const int rn = p.n_theta_ref;
    std::vector<double> key;
    key.reserve(16 + 5 * size_t(n) + 2 * size_t(rn));
    for (double x : {double(p.kind), p.ustar, p.z0, p.speed_ref, p.z_ref, p.alpha, p.direction, p.Cmu0, p.kappa,
                     double(p.unbounded), p.l_max, double(n), double(rn), double(p.k != nullptr)}) {
        key.push_back(x);
    }
    const size_t head = key.size();
    key.resize(head + 5 * size_t(n) + 2 * size_t(rn), 0.0);
    double* h = key.data() + head;
    for (int r = 0; r < n; ++r) {
        h[r] = p.z[r];
        h[n + r] = p.u[r];
        h[2 * n + r] = p.v[r];
        h[3 * n + r] = p.theta[r];
        h[4 * n + r] = p.k ? p.k[r] : 0.0;
    }
    for (int m = 0; m < rn; ++m) {
        h[5 * n + m] = p.theta_ref_z[m];
        h[5 * n + rn + m] = p.theta_ref[m];
    }
    static const bool hooked = [] {
        register_stream_cache(StreamCacheHook{&table_bytes, &table_drop});
        return true;
    }();
    (void)hooked;
    std::mutex& mu = table_mutex();
    auto& cache = table_map();
    int device = 0;
    GUST_HIP(hipGetDevice(&device));
    std::lock_guard<std::mutex> lock(mu);
    TableEntry& e = cache[{device, st}];
    const size_t total = key.size() - head;
    if (e.key != key) {
        if (total > e.cap) {
            if (e.dev) {
                GUST_HIP(hipStreamSynchronize(st));
                GUST_HIP(hipFree(e.dev));
                e.dev = nullptr;
                e.cap = 0;
            }
            GUST_HIP(hipMalloc(&e.dev, std::max<size_t>(total, 16) * sizeof(double)));
            e.cap = std::max<size_t>(total, 16);
        }
        GUST_HIP(hipStreamSynchronize(st));   // kernels still reading the previous tables
        GUST_HIP(hipMemcpyAsync(e.dev, h, total * sizeof(double), hipMemcpyHostToDevice, st));
        GUST_HIP(hipStreamSynchronize(st));
        e.key = std::move(key);
    }
    ProfileView v{};
    v.kind = p.kind;
    v.ustar = p.ustar;
    v.z0 = p.z0;
    v.speed_ref = p.speed_ref;
    v.z_ref = p.z_ref;
    v.alpha = p.alpha;
    const double a = p.direction * 3.14159265358979323846 / 180;
    v.cosd = std::cos(a);
    v.sind = std::sin(a);
    v.Cmu0 = p.Cmu0;
    v.kappa = p.kappa;
    v.inv_lmax = p.unbounded ? 0.0 : 1.0 / p.l_max;
    double* d = e.dev;
    v.tz = d;
    v.tu = d + n;
    v.tv = d + 2 * n;
    v.tt = d + 3 * n;
    v.tk = d + 4 * n;
    v.n = n;
    v.has_k = (n > 0 && p.k) ? 1 : 0;
    v.rz = d + 5 * n;
    v.rt = d + 5 * n + rn;
    v.rn = rn;
    return v;
}

}   // namespace gust_amd::profile
