// line_sweeps.hip -- batched z-line Jacobi and line LU-SGS sweeps for gfx950.
//
// Reference behaviour (GUST):
//   Source/LineSolve_K.H:22-45     the exact column solve (Thomas, x/y
//                                  neighbours of the iterate on the rhs)
//   Source/Predictor.cpp:87-127    zline_jacobi: sweeps of independent columns
//   Source/LineSweeps.cpp:181-305  line LU-SGS, lagged / zero ghosts
//   Source/LineSweeps.cpp:361-413  ownership, sweep region, checks
//
// Design (MI355X):
//  * Jacobi: one 64-lane wavefront per 64 consecutive columns (flattened
//    (i, j) with i fastest), lane = column, loop over k. In the Array4
//    layout (i fastest) every k-step is one coalesced 512 B (FP64) load per
//    operand across the wave. The Thomas factor c' lives in LDS (64 x nz
//    reals per wave); e' is staged in the output array and overwritten by
//    the back substitution (the re-read hits L2). Many systems (the five
//    variables of an iteration, many tiles) share one launch through a
//    device table; blocks are remapped so each XCD sweeps a contiguous range
//    of column rows (the j +- 1 neighbour reads then hit its L2).
//  * LU-SGS: one workgroup per system executes the diagonal wavefronts
//    i + j = s (or i - j for the j-reversed pair) with a workgroup barrier
//    between planes; the sequential order performs the same operations on
//    the same values, so both orders return the same bits.
//  * Arithmetic follows the reference expression by expression and the
//    library is compiled without floating-point contraction, so FP64 results
//    equal the CPU build bit for bit.
#include "common.hpp"
#include "deferred.hpp"
#include "simplec.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <map>
#include <memory>

namespace gust_amd {

namespace {

constexpr int kWave = 64;
constexpr int kStagedMaxTiles = 1024;

// The z-line Jacobi schedule: 0 autotuned, 1 streaming, 2 LDS-staged (where
// it fits), 3 the fixed rule (staged under kStagedMaxTiles tiles), 4 the
// register kernel (where the columns fit a level bucket, else streaming).
std::atomic<int> g_jacobi_schedule{0};
int jacobi_schedule() { return g_jacobi_schedule.load(std::memory_order_relaxed); }

// The autotuner's candidates: the LDS-staged kernel, the streaming kernel
// with c' in LDS at prefetch depths 2, 4 and 8, the streaming kernel with c'
// in global memory (no LDS limit on occupancy), the register kernel (e' in
// registers, c' in LDS), the register kernel with c' in registers too
// (FP32, or FP64 columns of at most 65 levels) and, for FP64 columns of up
// to 97 levels, with c' split between LDS and registers. Every one gives the same
// bits; the first run of a plan whose shapes are new times each (all its
// sweeps, HIP events on its stream) and keeps the fastest for those shapes.
enum TuneChoice : int { kChoiceStaged = 0, kChoiceP4 = 1, kChoiceP2 = 2, kChoiceP8 = 3, kChoiceGlobal = 4,
                        kChoiceRegs = 5, kChoiceRegsC = 6, kChoiceRegsSplit = 7, kNumChoices = 8 };
const char* kChoiceNames[kNumChoices] = {"staged", "streaming-p4", "streaming-p2", "streaming-p8",
                                         "streaming-global-cprime", "register-e", "register-ce",
                                         "register-split"};

struct TuneKey {
    int dtype = 0, sweeps = 0;
    std::vector<int> shape;   // (nx, ny, nz) of every system
    bool operator<(const TuneKey& o) const
    {
        if (dtype != o.dtype) { return dtype < o.dtype; }
        if (sweeps != o.sweeps) { return sweeps < o.sweeps; }
        return shape < o.shape;
    }
};

struct TuneRecord {
    int choice = kChoiceP4;
    float ms[kNumChoices] = {-1, -1, -1, -1, -1, -1, -1, -1};
};

std::mutex g_tune_mu;
std::map<TuneKey, TuneRecord> g_tuned;

// One system of a Jacobi launch. Pointers are at the valid box's low corner.
template <typename T>
struct JacobiSys {
    const T* A;
    int64_t a_j, a_k, a_n;
    const T* ap;
    int64_t p_j, p_k;
    const T* r;
    int64_t r_j, r_k;
    T* d;            // the caller's delta (valid lo corner)
    int64_t d_j, d_k;
    T* s;            // scratch iterate (compact)
    int64_t s_j, s_k;
    int nx, ny, nz;  // valid extents at the field's location
    int Lx, Ly;      // periodic shift in cells along x / y, 0 if not periodic
    int tile0;       // first column tile of this system
    int pad;
};

using dev::gld;
using dev::gst;
using dev::ldg;

// Which system holds column tile t (tile0 ascending).
template <typename S>
__device__ __forceinline__ int find_sys(const S* __restrict__ sys, int nsys, int t)
{
    int lo = 0, hi = nsys - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sys[mid].tile0 <= t) { lo = mid; } else { hi = mid - 1; }
    }
    return lo;
}

// One z-line Jacobi sweep over every system of the table.
//   kFirst: the input iterate is zero (the first sweep after the zeroing of
//           check_line_system); its neighbour values are 0, not read.
//   kLast:  record max |new - old| per system into changes (as ordered bits).
//   kP:     prefetch depth: the operands of level k + kP are loaded while
//           level k is eliminated (a register ring indexed at compile time),
//           so every wave keeps ~kP levels of loads in flight instead of
//           waiting one memory latency per level. kP x 12 loads stay under
//           the 63 outstanding vector memory operations a wave can count.
//   dst_scratch: 1 = write the scratch iterate, read the caller's delta;
//                0 = the reverse.
//   cglob: c' in global memory ([tile][k][lane]) when the LDS cannot hold
//          64 x nz reals (null: c' in dynamic LDS).
template <typename T>
struct LevelOps {
    T axm, axp, aym, ayp, azm, azp, dg, rr, vm, vp, vjm, vjp;
};

template <typename T, bool kFirst, int kP>
__device__ __forceinline__ void load_level(LevelOps<T>& o, const T* __restrict__ a, int64_t ak, int64_t an,
                                           const T* __restrict__ pap, int64_t pk, const T* __restrict__ pr,
                                           int64_t rk, const T* sm, const T* sp, const T* sjm, const T* sjp,
                                           int64_t sk, bool okm, bool okp, bool okjm, bool okjp, int k)
{
    const T* ak0 = a + k * ak;
    o.axm = ldg(ak0 + 1 * an);
    o.axp = ldg(ak0 + 2 * an);
    o.aym = ldg(ak0 + 3 * an);
    o.ayp = ldg(ak0 + 4 * an);
    o.azm = ldg(ak0 + 5 * an);
    o.azp = ldg(ak0 + 6 * an);
    o.dg = ldg(pap + k * pk);
    o.rr = ldg(pr + k * rk);
    if (!kFirst) {
        const T tm = gld(sm + k * sk), tp = gld(sp + k * sk), tjm = gld(sjm + k * sk), tjp = gld(sjp + k * sk);
        o.vm = okm ? tm : T(0);
        o.vp = okp ? tp : T(0);
        o.vjm = okjm ? tjm : T(0);
        o.vjp = okjp ? tjp : T(0);
    } else {
        o.vm = o.vp = o.vjm = o.vjp = T(0);
    }
}

template <typename T, bool kFirst, bool kLast, bool kLds, int kP = 4>
__global__ __launch_bounds__(kWave) void zjacobi_sweep_kernel(const JacobiSys<T>* __restrict__ table, int nsys,
                                                              int total_tiles, int dst_scratch,
                                                              unsigned long long* __restrict__ changes,
                                                              T* __restrict__ cglob, int nzmax)
{
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x;
    const int tile = dev::xcd_remap(blockIdx.x, total_tiles);
    T* cgl = cglob + (kLds ? 0 : size_t(tile) * kWave * size_t(nzmax));
    T* cls = reinterpret_cast<T*>(lds_raw);
    const int si = find_sys(table, nsys, tile);
    const JacobiSys<T> S = table[si];
    const int c = (tile - S.tile0) * kWave + lane;
    const int ncols = S.nx * S.ny;
    unsigned long long chg = 0;
    if (c < ncols) {
        const int i = c % S.nx, j = c / S.nx;
        const T* src = dst_scratch ? S.d : S.s;
        T* dst = dst_scratch ? S.s : S.d;
        const int64_t src_j = dst_scratch ? S.d_j : S.s_j, src_k = dst_scratch ? S.d_k : S.s_k;
        const int64_t dst_j = dst_scratch ? S.s_j : S.d_j, dst_k = dst_scratch ? S.s_k : S.d_k;
        // Neighbour columns of the input iterate: in the domain, across a
        // periodic seam (the image; nodal or cell), or 0 outside.
        int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
        bool okm = true, okp = true, okjm = true, okjp = true;
        if (im < 0) { if (S.Lx) { im += S.Lx; } else { okm = false; im = i; } }
        if (ip > S.nx - 1) { if (S.Lx) { ip -= S.Lx; } else { okp = false; ip = i; } }
        if (jm < 0) { if (S.Ly) { jm += S.Ly; } else { okjm = false; jm = j; } }
        if (jp > S.ny - 1) { if (S.Ly) { jp -= S.Ly; } else { okjp = false; jp = j; } }
        const T* __restrict__ a = S.A + i + int64_t(j) * S.a_j;
        const T* __restrict__ pap = S.ap + i + int64_t(j) * S.p_j;
        const T* __restrict__ pr = S.r + i + int64_t(j) * S.r_j;
        const T* sm = src + im + int64_t(j) * src_j;
        const T* sp = src + ip + int64_t(j) * src_j;
        const T* sjm = src + i + int64_t(jm) * src_j;
        const T* sjp = src + i + int64_t(jp) * src_j;
        const T* sc = src + i + int64_t(j) * src_j;
        T* out = dst + i + int64_t(j) * dst_j;
        const int64_t ak = S.a_k, an = S.a_n;
        const int nz = S.nz;
        // Forward elimination (Source/LineSolve_K.H:29-42) over a ring of kP
        // prefetched levels. Full chunks of kP levels run branch-free (the
        // prefetch level is clamped to the column, a harmless re-read near
        // its top); the tail of nz % kP levels issues no loads.
        LevelOps<T> ring[kP];
#pragma unroll
        for (int u = 0; u < kP; ++u) {
            load_level<T, kFirst, kP>(ring[u], a, ak, an, pap, S.p_k, pr, S.r_k, sm, sp, sjm, sjp, src_k, okm, okp,
                                      okjm, okjp, u < nz ? u : nz - 1);
        }
        T cprev = T(0), eprev = T(0);
        auto eliminate = [&](const LevelOps<T>& o, int k) {
            const T rk = o.rr + o.axm * o.vm + o.axp * o.vp + o.aym * o.vjm + o.ayp * o.vjp;
            const T lower = -o.azm, upper = -o.azp;
            T ck, ek;
            if (k == 0) {
                const T m = o.dg;
                ck = upper / m;
                ek = rk / m;
            } else {
                const T m = o.dg - lower * cprev;
                ck = upper / m;
                ek = (rk - lower * eprev) / m;
            }
            if (kLds) { cls[k * kWave + lane] = ck; } else { gst(cgl + k * kWave + lane, ck); }
            gst(out + k * dst_k, ek);
            cprev = ck;
            eprev = ek;
        };
        const int nfull = (nz / kP) * kP;
        for (int k0 = 0; k0 < nfull; k0 += kP) {
#pragma unroll
            for (int u = 0; u < kP; ++u) {
                const int k = k0 + u;
                const LevelOps<T> o = ring[u];
                const int kn = k + kP < nz ? k + kP : nz - 1;
                load_level<T, kFirst, kP>(ring[u], a, ak, an, pap, S.p_k, pr, S.r_k, sm, sp, sjm, sjp, src_k, okm, okp,
                                          okjm, okjp, kn);
                eliminate(o, k);
            }
        }
#pragma unroll
        for (int u = 0; u < kP; ++u) {
            if (nfull + u < nz) { eliminate(ring[u], nfull + u); }
        }
        // Back substitution (Source/LineSolve_K.H:43-44); e' is re-read from
        // the output column kB levels ahead (clamped the same way), c' from
        // LDS. Level m of the ring walk is k = nz - 2 - m.
        constexpr int kB = 8;
        T o = eprev;
        if (kLast) {
            const T old = kFirst ? T(0) : gld(sc + (nz - 1) * src_k);
            const unsigned long long b = dev::abs_bits(double(o - old));
            chg = b > chg ? b : chg;
        }
        const int nb = nz - 1;   // levels to substitute
        T eb[kB], ob[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            const int k = nz - 2 - u >= 0 ? nz - 2 - u : 0;
            eb[u] = gld(out + k * dst_k);
            ob[u] = (kLast && !kFirst) ? gld(sc + k * src_k) : T(0);
        }
        auto substitute = [&](T ek, T oldk, int k) {
            const T ck = kLds ? cls[k * kWave + lane] : gld(cgl + k * kWave + lane);
            o = ek - ck * o;
            gst(out + k * dst_k, o);
            if (kLast) {
                const unsigned long long b = dev::abs_bits(double(o - oldk));
                chg = b > chg ? b : chg;
            }
        };
        const int nbfull = nb > 0 ? (nb / kB) * kB : 0;
        for (int m0 = 0; m0 < nbfull; m0 += kB) {
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int k = nz - 2 - (m0 + u);
                const T ek = eb[u], oldk = ob[u];
                const int kn = k - kB >= 0 ? k - kB : 0;
                eb[u] = gld(out + kn * dst_k);
                ob[u] = (kLast && !kFirst) ? gld(sc + kn * src_k) : T(0);
                substitute(ek, oldk, k);
            }
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (nbfull + u < nb) { substitute(eb[u], ob[u], nz - 2 - (nbfull + u)); }
        }
    }
    if (kLast) {
        chg = dev::wave_max_u64(chg);
        if (lane == 0 && chg) { atomicMax(changes + si, chg); }
    }
}

// The streaming sweep with e' held in registers for the whole column instead
// of staged in the output column: the output is then written once, by the
// back substitution, and e' never travels to L2 or HBM (in the kernel above
// it is written, re-read and overwritten, and a wave's 64 x nz e' values
// outlive the L2 once a few hundred waves are in flight). The level loops are
// unrolled to kNZ levels (every register index is a constant). Loads are
// issued for every unrolled level (a column with nz < kNZ re-reads its top
// level) and only the arithmetic of levels past the column's top is skipped,
// by a branch that is uniform over the wave (all lanes of a wave sweep the
// same system): no loaded value is merged at a branch, so the prefetched
// loads stay in flight across levels. Every array is read and written
// through a buffer resource: the level is a scalar offset shared by the wave,
// the column a 32-bit per-lane offset, so a load is one instruction with no
// per-lane address arithmetic (the unrolled body stays small enough for the
// instruction cache). c' of levels below kCL lives in LDS, of the levels
// above in registers (kCL = kNZ: all in LDS; 0: all in registers, no LDS;
// in between: less LDS per wave, so more waves per CU). Same operations in
// the same order as zjacobi_sweep_kernel: the same bits.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;   // raw (untyped) buffer access on gfx9
constexpr int kAuxNontemporal = 2;       // the nt cache policy: streamed once

__device__ __forceinline__ Rsrc rsrc_of(const void* base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, kRsrcWord3);
}

template <typename T, int kAux>
__device__ __forceinline__ T bload(Rsrc r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, kAux));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kAux));
    }
}

template <typename T>
__device__ __forceinline__ void bstore(T v, Rsrc r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 8) {
        using U2 = unsigned int __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, voff, soff, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
    }
}

template <typename T, bool kFirst, bool kLast, int kNZ, int kCL>
__global__ __launch_bounds__(kWave, 1) void zjacobi_regs_kernel(const JacobiSys<T>* __restrict__ table, int nsys,
                                                                int total_tiles, int dst_scratch,
                                                                unsigned long long* __restrict__ changes)
{
    constexpr int kP = 4;   // prefetch depth of the forward elimination
    constexpr int kB = 8;   // prefetch depth of the old iterate (last sweep)
    constexpr bool kOld = kLast && !kFirst;
    constexpr uint32_t kS = sizeof(T);
    extern __shared__ unsigned char lds_raw[];
    const int lane = threadIdx.x;
    const int tile = dev::xcd_remap(blockIdx.x, total_tiles);
    T* cls = reinterpret_cast<T*>(lds_raw);
    const int si = find_sys(table, nsys, tile);
    const JacobiSys<T> S = table[si];
    const int c = (tile - S.tile0) * kWave + lane;
    const int ncols = S.nx * S.ny;
    unsigned long long chg = 0;
    if (c < ncols) {
        const int i = c % S.nx, j = c / S.nx;
        const T* src = dst_scratch ? S.d : S.s;
        T* dst = dst_scratch ? S.s : S.d;
        const int64_t src_j = dst_scratch ? S.d_j : S.s_j, src_k = dst_scratch ? S.d_k : S.s_k;
        const int64_t dst_j = dst_scratch ? S.s_j : S.d_j, dst_k = dst_scratch ? S.s_k : S.d_k;
        int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
        bool okm = true, okp = true, okjm = true, okjp = true;
        if (im < 0) { if (S.Lx) { im += S.Lx; } else { okm = false; im = i; } }
        if (ip > S.nx - 1) { if (S.Lx) { ip -= S.Lx; } else { okp = false; ip = i; } }
        if (jm < 0) { if (S.Ly) { jm += S.Ly; } else { okjm = false; jm = j; } }
        if (jp > S.ny - 1) { if (S.Ly) { jp -= S.Ly; } else { okjp = false; jp = j; } }
        // buffer resources (the host checks every offset fits 31 bits)
        const Rsrc rA1 = rsrc_of(S.A + 1 * S.a_n), rA2 = rsrc_of(S.A + 2 * S.a_n), rA3 = rsrc_of(S.A + 3 * S.a_n);
        const Rsrc rA4 = rsrc_of(S.A + 4 * S.a_n), rA5 = rsrc_of(S.A + 5 * S.a_n), rA6 = rsrc_of(S.A + 6 * S.a_n);
        const Rsrc rP = rsrc_of(S.ap), rR = rsrc_of(S.r), rS = rsrc_of(src), rD = rsrc_of(dst);
        // per-lane byte offsets within a level, per-level byte strides
        const uint32_t oa = uint32_t(i + int64_t(j) * S.a_j) * kS;
        const uint32_t op = uint32_t(i + int64_t(j) * S.p_j) * kS;
        const uint32_t orr = uint32_t(i + int64_t(j) * S.r_j) * kS;
        const uint32_t om = uint32_t(im + int64_t(j) * src_j) * kS;
        const uint32_t opp = uint32_t(ip + int64_t(j) * src_j) * kS;
        const uint32_t ojm = uint32_t(i + int64_t(jm) * src_j) * kS;
        const uint32_t ojp = uint32_t(i + int64_t(jp) * src_j) * kS;
        const uint32_t oc = uint32_t(i + int64_t(j) * src_j) * kS;
        const uint32_t od = uint32_t(i + int64_t(j) * dst_j) * kS;
        const uint32_t sa = uint32_t(S.a_k) * kS, sp = uint32_t(S.p_k) * kS, sr = uint32_t(S.r_k) * kS;
        const uint32_t ss = uint32_t(src_k) * kS, sd = uint32_t(dst_k) * kS;
        const int nz = S.nz;
        auto load = [&](LevelOps<T>& o, int k) {
            const uint32_t ka = uint32_t(k) * sa;
            o.axm = bload<T, kAuxNontemporal>(rA1, oa, ka);
            o.axp = bload<T, kAuxNontemporal>(rA2, oa, ka);
            o.aym = bload<T, kAuxNontemporal>(rA3, oa, ka);
            o.ayp = bload<T, kAuxNontemporal>(rA4, oa, ka);
            o.azm = bload<T, kAuxNontemporal>(rA5, oa, ka);
            o.azp = bload<T, kAuxNontemporal>(rA6, oa, ka);
            o.dg = bload<T, kAuxNontemporal>(rP, op, uint32_t(k) * sp);
            o.rr = bload<T, kAuxNontemporal>(rR, orr, uint32_t(k) * sr);
            if (!kFirst) {
                const uint32_t ks = uint32_t(k) * ss;
                const T tm = bload<T, 0>(rS, om, ks), tp = bload<T, 0>(rS, opp, ks);
                const T tjm = bload<T, 0>(rS, ojm, ks), tjp = bload<T, 0>(rS, ojp, ks);
                o.vm = okm ? tm : T(0);
                o.vp = okp ? tp : T(0);
                o.vjm = okjm ? tjm : T(0);
                o.vjp = okjp ? tjp : T(0);
            } else {
                o.vm = o.vp = o.vjm = o.vjp = T(0);
            }
        };
        T ev[kNZ];
        T cv[kCL < kNZ ? kNZ - kCL : 1];
        LevelOps<T> ring[kP];
#pragma unroll
        for (int u = 0; u < kP; ++u) { load(ring[u], u < nz ? u : nz - 1); }
        // Forward elimination (Source/LineSolve_K.H:29-42).
        T cprev = T(0), eprev = T(0);
        auto level = [&](int k) {
            const LevelOps<T> o = ring[k % kP];
            if (k + kP < kNZ) { load(ring[k % kP], k + kP < nz ? k + kP : nz - 1); }
            if (k < nz) {
                const T rk = o.rr + o.axm * o.vm + o.axp * o.vp + o.aym * o.vjm + o.ayp * o.vjp;
                const T lower = -o.azm, upper = -o.azp;
                T ck, ek;
                if (k == 0) {
                    const T m = o.dg;
                    ck = upper / m;
                    ek = rk / m;
                } else {
                    const T m = o.dg - lower * cprev;
                    ck = upper / m;
                    ek = (rk - lower * eprev) / m;
                }
                if (k >= kCL) { cv[k >= kCL ? k - kCL : 0] = ck; } else { cls[k * kWave + lane] = ck; }
                ev[k] = ek;
                cprev = ck;
                eprev = ek;
            }
        };
        // two loops of at most 49 levels: each stays under the compiler's
        // full-unroll limit (a partly unrolled loop would index ev at run
        // time and move it to scratch memory)
        constexpr int kHalf = (kNZ + 1) / 2;
#pragma unroll
        for (int k = 0; k < kHalf; ++k) { level(k); }
#pragma unroll
        for (int k = kHalf; k < kNZ; ++k) { level(k); }
        // Back substitution (Source/LineSolve_K.H:43-44) from the registers.
        T o = eprev;
        bstore<T>(o, rD, od, uint32_t(nz - 1) * sd);
        if (kLast) {
            const T old = kFirst ? T(0) : bload<T, 0>(rS, oc, uint32_t(nz - 1) * ss);
            const unsigned long long b = dev::abs_bits(double(o - old));
            chg = b > chg ? b : chg;
        }
        // The old iterate of level k, read kB unrolled levels ahead (clamped
        // to the column like the forward loads).
        T ob[kB];
        if (kOld) {
#pragma unroll
            for (int u = 0; u < kB; ++u) {
                const int k = kNZ - 2 - u;
                if (k >= 0) { ob[k % kB] = bload<T, 0>(rS, oc, uint32_t(k <= nz - 2 ? k : nz - 1) * ss); }
            }
        }
#pragma unroll
        for (int k = kNZ - 2; k >= 0; --k) {
            T oldk = T(0);
            if (kOld) {
                oldk = ob[k % kB];
                if (k - kB >= 0) { ob[k % kB] = bload<T, 0>(rS, oc, uint32_t(k - kB <= nz - 2 ? k - kB : nz - 1) * ss); }
            }
            if (k <= nz - 2) {
                const T ck = k >= kCL ? cv[k >= kCL ? k - kCL : 0] : cls[k * kWave + lane];
                o = ev[k] - ck * o;
                bstore<T>(o, rD, od, uint32_t(k) * sd);
                if (kLast) {
                    const unsigned long long b = dev::abs_bits(double(o - oldk));
                    chg = b > chg ? b : chg;
                }
            }
        }
    }
    if (kLast) {
        chg = dev::wave_max_u64(chg);
        if (lane == 0 && chg) { atomicMax(changes + si, chg); }
    }
}

// The register kernel's level buckets: the columns of the configurations
// (48, 64 and 96 cells, one more level at z-faces) fit exactly.
constexpr int kRegBuckets[3] = {49, 65, 97};
// FP64 columns of the largest bucket with c' split: this many levels in LDS
// (40 KB per wave: four waves per CU), the rest in registers.
constexpr int kSplitLevels = 80;

// The same sweep for launches with too few tiles to hide memory latency
// (small tiles: one wave per CU or fewer). A block of kStagedThreads stages
// one tile's operands in LDS first -- every level of every column loaded at
// once, the neighbour terms summed into r_k in the order of the streaming
// kernel -- and its first wave then runs the columns' Thomas recurrences
// (Source/LineSolve_K.H:29-44) out of LDS: the serial chain waits on LDS
// instead of one memory latency per level. Results are bitwise those of
// zjacobi_sweep_kernel. LDS: [array][k][W] with W >= the tile's active
// columns; arrays d_P, lower, upper, r (c' and e' overwrite upper and r)
// and, for the last sweep, the old iterate.
constexpr int kStagedThreads = 256;

template <typename T, bool kFirst, bool kLast>
__global__ __launch_bounds__(kStagedThreads) void zjacobi_staged_kernel(const JacobiSys<T>* __restrict__ table,
                                                                        int nsys, int total_tiles, int dst_scratch,
                                                                        unsigned long long* __restrict__ changes,
                                                                        int W)
{
    extern __shared__ unsigned char lds_raw[];
    const int tile = dev::xcd_remap(blockIdx.x, total_tiles);
    const int si = find_sys(table, nsys, tile);
    const JacobiSys<T> S = table[si];
    const int ncols = S.nx * S.ny;
    const int c0 = (tile - S.tile0) * kWave;
    const int nact = ncols - c0 < kWave ? ncols - c0 : kWave;
    const int nz = S.nz;
    const int64_t plane = int64_t(nz) * W;
    T* Ldg = reinterpret_cast<T*>(lds_raw);
    T* Llo = Ldg + plane;
    T* Lup = Llo + plane;
    T* Lr = Lup + plane;
    T* Lold = Lr + plane;
    constexpr bool kOld = kLast && !kFirst;
    const T* src = dst_scratch ? S.d : S.s;
    T* dst = dst_scratch ? S.s : S.d;
    const int64_t src_j = dst_scratch ? S.d_j : S.s_j, src_k = dst_scratch ? S.d_k : S.s_k;
    const int64_t dst_j = dst_scratch ? S.s_j : S.d_j, dst_k = dst_scratch ? S.s_k : S.d_k;
    // 1. stage: thread t takes column t % nact of level t / nact (consecutive
    // threads, consecutive i: coalesced rows)
    for (int t = threadIdx.x; t < nz * nact; t += kStagedThreads) {
        const int lane = t % nact, k = t / nact;
        const int c = c0 + lane;
        const int i = c % S.nx, j = c / S.nx;
        const T* ak0 = S.A + i + int64_t(j) * S.a_j + k * S.a_k;
        const int64_t an = S.a_n;
        const T axm = ldg(ak0 + 1 * an), axp = ldg(ak0 + 2 * an), aym = ldg(ak0 + 3 * an), ayp = ldg(ak0 + 4 * an);
        const T azm = ldg(ak0 + 5 * an), azp = ldg(ak0 + 6 * an);
        const T dg = ldg(S.ap + i + int64_t(j) * S.p_j + k * S.p_k);
        const T rr = ldg(S.r + i + int64_t(j) * S.r_j + k * S.r_k);
        T vm = T(0), vp = T(0), vjm = T(0), vjp = T(0);
        if (!kFirst) {
            int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
            bool okm = true, okp = true, okjm = true, okjp = true;
            if (im < 0) { if (S.Lx) { im += S.Lx; } else { okm = false; im = i; } }
            if (ip > S.nx - 1) { if (S.Lx) { ip -= S.Lx; } else { okp = false; ip = i; } }
            if (jm < 0) { if (S.Ly) { jm += S.Ly; } else { okjm = false; jm = j; } }
            if (jp > S.ny - 1) { if (S.Ly) { jp -= S.Ly; } else { okjp = false; jp = j; } }
            const int64_t kk = k * src_k;
            const T tm = gld(src + im + int64_t(j) * src_j + kk), tp = gld(src + ip + int64_t(j) * src_j + kk);
            const T tjm = gld(src + i + int64_t(jm) * src_j + kk), tjp = gld(src + i + int64_t(jp) * src_j + kk);
            vm = okm ? tm : T(0);
            vp = okp ? tp : T(0);
            vjm = okjm ? tjm : T(0);
            vjp = okjp ? tjp : T(0);
        }
        const int64_t at = int64_t(k) * W + lane;
        Ldg[at] = dg;
        Llo[at] = -azm;
        Lup[at] = -azp;
        Lr[at] = rr + axm * vm + axp * vp + aym * vjm + ayp * vjp;
        if (kOld) { Lold[at] = gld(src + i + int64_t(j) * src_j + k * src_k); }
    }
    __syncthreads();
    // 2. the recurrences, one lane per column of the first wave
    if (threadIdx.x >= kWave) { return; }
    const int lane = threadIdx.x;
    unsigned long long chg = 0;
    if (lane < nact) {
        const int c = c0 + lane;
        const int i = c % S.nx, j = c / S.nx;
        T* out = dst + i + int64_t(j) * dst_j;
        T cprev = T(0), eprev = T(0);
        for (int k = 0; k < nz; ++k) {
            const int64_t at = int64_t(k) * W + lane;
            const T lower = Llo[at], upper = Lup[at], dg = Ldg[at], rk = Lr[at];
            T ck, ek;
            if (k == 0) {
                const T m = dg;
                ck = upper / m;
                ek = rk / m;
            } else {
                const T m = dg - lower * cprev;
                ck = upper / m;
                ek = (rk - lower * eprev) / m;
            }
            Lup[at] = ck;
            Lr[at] = ek;
            cprev = ck;
            eprev = ek;
        }
        T o = eprev;
        gst(out + (nz - 1) * dst_k, o);
        if (kLast) {
            const T old = kFirst ? T(0) : Lold[int64_t(nz - 1) * W + lane];
            const unsigned long long b = dev::abs_bits(double(o - old));
            chg = b > chg ? b : chg;
        }
        for (int k = nz - 2; k >= 0; --k) {
            const int64_t at = int64_t(k) * W + lane;
            o = Lr[at] - Lup[at] * o;
            gst(out + k * dst_k, o);
            if (kLast) {
                const T old = kFirst ? T(0) : Lold[at];
                const unsigned long long b = dev::abs_bits(double(o - old));
                chg = b > chg ? b : chg;
            }
        }
    }
    if (kLast) {
        chg = dev::wave_max_u64(chg);
        if (lane == 0 && chg) { atomicMax(changes + si, chg); }
    }
}

// The copy of a periodic seam face OverrideSync keeps (Source/Predictor.cpp:125
// with AMReX's owner mask: the low copy owns): delta(hi seam) = delta(lo seam).
template <typename T>
__global__ void seam_sync_kernel(T* d, int64_t d_j, int64_t d_k, int nx, int ny, int nz, int axis, int low_owner)
{
    const int64_t n = axis == 0 ? int64_t(ny) * nz : int64_t(nx) * nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        int64_t lo, hi;
        if (axis == 0) {
            const int j = int(t % ny), k = int(t / ny);
            lo = int64_t(j) * d_j + k * d_k;
            hi = lo + (nx - 1);
        } else {
            const int i = int(t % nx), k = int(t / nx);
            lo = i + int64_t(k) * d_k;
            hi = lo + int64_t(ny - 1) * d_j;
        }
        if (low_owner) { d[hi] = d[lo]; } else { d[lo] = d[hi]; }
    }
}

// Zero the points of a view's box outside its valid box [vlo, vhi] (local
// indices): the ghosts of delta. The valid points are all overwritten by the
// last sweep. A fill kernel, not a memset, so the call can live inside a
// captured graph.
template <typename T>
__global__ void zero_ghosts_kernel(T* p, int64_t jstride, int64_t kstride, int nx, int ny, int nz, int vlo0,
                                   int vhi0, int vlo1, int vhi1, int vlo2, int vhi2)
{
    const int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        const int i = int(t % nx);
        const int64_t r = t / nx;
        const int j = int(r % ny), k = int(r / ny);
        if (i >= vlo0 && i <= vhi0 && j >= vlo1 && j <= vhi1 && k >= vlo2 && k <= vhi2) { continue; }
        p[i + j * jstride + k * kstride] = T(0);
    }
}

// The same for up to kGhostSets views in one launch (blockIdx.y: the view).
constexpr int kGhostSets = 8;
template <typename T>
struct GhostSet {
    T* p[kGhostSets];
    int64_t js[kGhostSets], ks[kGhostSets];
    int n[kGhostSets][3], vlo[kGhostSets][3], vhi[kGhostSets][3];
};

template <typename T>
__global__ void zero_ghosts_multi_kernel(GhostSet<T> g)
{
    const int s = blockIdx.y;
    const int nx = g.n[s][0], ny = g.n[s][1], nz = g.n[s][2];
    const int64_t n = int64_t(nx) * ny * nz;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
        const int i = int(t % nx);
        const int64_t r = t / nx;
        const int j = int(r % ny), k = int(r / ny);
        if (i >= g.vlo[s][0] && i <= g.vhi[s][0] && j >= g.vlo[s][1] && j <= g.vhi[s][1] && k >= g.vlo[s][2] &&
            k <= g.vhi[s][2]) {
            continue;
        }
        g.p[s][i + j * g.js[s] + k * g.ks[s]] = T(0);
    }
}

__global__ void copy_u64_kernel(const unsigned long long* src, unsigned long long* dst, int n)
{
    for (int t = threadIdx.x; t < n; t += blockDim.x) { dst[t] = src[t]; }
}

__global__ void zero_u64_kernel(unsigned long long* p, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) { p[t] = 0ull; }
}

__global__ void bits_to_double_kernel(const unsigned long long* bits, double* out, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) { out[t] = __longlong_as_double(static_cast<long long>(bits[t])); }
}

// ---------------------------------------------------------------------------
// Host side.

// check_line_system (Source/LineSweeps.cpp:391-413) for a one-box domain.
void check_line_system(const char* op, int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs,
                       int sweeps, const gust_geometry* g, const gust_array4* delta, IBox& V)
{
    if (!g) { GUST_FAIL("gust: %s: geometry is null", op); }
    if (sweeps < 1) { GUST_FAIL("gust: %s: sweeps must be >= 1 (got %d)", op, sweeps); }
    if (!valid_loc(loc)) { GUST_FAIL("gust: %s: unknown location %d", op, loc); }
    for (int d = 0; d < 3; ++d) {
        if (g->hi[d] < g->lo[d]) { GUST_FAIL("gust: %s: empty domain", op); }
    }
    if (g->periodic[2]) { GUST_FAIL("gust: %s: the domain cannot be periodic in z", op); }
    V = loc_box(*g, loc);
    if (!A || A->ncomp != GUST_N_COEFS) { GUST_FAIL("gust: %s: A needs the implicit operator's components", op); }
    if (!covers(A, V, GUST_N_COEFS)) { GUST_FAIL("gust: %s: A is not on the valid box at the %s", op, loc_words(loc)); }
    if (!covers(ap, V, 1)) { GUST_FAIL("gust: %s: ap is not on the valid box at the %s", op, loc_words(loc)); }
    if (!covers(rhs, V, 1)) { GUST_FAIL("gust: %s: rhs is not on the valid box at the %s", op, loc_words(loc)); }
    if (!covers(delta, grow(V, 1, 1, 1), 1)) {
        GUST_FAIL("gust: %s: delta needs at least 1 ghost cells at the %s", op, loc_words(loc));
    }
}

// delta = 0 outside its valid box V (check_line_system's setVal; the valid
// points are overwritten by the sweeps).
template <typename T>
void zero_ghosts(const gust_array4& v, const IBox& V, hipStream_t st)
{
    const int nx = v.end[0] - v.begin[0], ny = v.end[1] - v.begin[1], nz = v.end[2] - v.begin[2];
    const int64_t n = int64_t(nx) * ny * nz;
    if (n <= 0) { return; }
    const int blocks = int(std::min<int64_t>((n + 255) / 256, 8192));
    zero_ghosts_kernel<T><<<blocks, 256, 0, st>>>(static_cast<T*>(v.p), v.jstride, v.kstride, nx, ny, nz,
                                                  V.lo[0] - v.begin[0], V.hi[0] - v.begin[0], V.lo[1] - v.begin[1],
                                                  V.hi[1] - v.begin[1], V.lo[2] - v.begin[2], V.hi[2] - v.begin[2]);
    GUST_HIP(hipGetLastError());
}

// zero_ghosts of many views, kGhostSets per launch.
template <typename T>
void zero_ghosts_all(const std::vector<gust_line_system>& sys, const std::vector<IBox>& valid, hipStream_t st)
{
    for (size_t s0 = 0; s0 < sys.size(); s0 += kGhostSets) {
        GhostSet<T> g{};
        int64_t most = 1;
        const int cnt = int(std::min<size_t>(kGhostSets, sys.size() - s0));
        for (int q = 0; q < cnt; ++q) {
            const gust_array4& v = sys[s0 + size_t(q)].delta;
            const IBox& V = valid[s0 + size_t(q)];
            g.p[q] = static_cast<T*>(v.p);
            g.js[q] = v.jstride;
            g.ks[q] = v.kstride;
            for (int d = 0; d < 3; ++d) {
                g.n[q][d] = v.end[d] - v.begin[d];
                g.vlo[q][d] = V.lo[d] - v.begin[d];
                g.vhi[q][d] = V.hi[d] - v.begin[d];
            }
            most = std::max<int64_t>(most, int64_t(g.n[q][0]) * g.n[q][1] * g.n[q][2]);
        }
        const int blocks = int(std::min<int64_t>((most + 255) / 256, 8192));
        zero_ghosts_multi_kernel<T><<<dim3(blocks, cnt), 256, 0, st>>>(g);
        GUST_HIP(hipGetLastError());
    }
}

template <typename T>
T* at_lo(const gust_array4& v, const IBox& V, int n = 0)
{
    return static_cast<T*>(v.p) + offset_of(v, V.lo[0], V.lo[1], V.lo[2], n);
}

// A batched Jacobi plan: the device table of systems, the scratch iterate
// and the change accumulators, built once. Running it enqueues kernels only
// (no host synchronisation, no copies, no memsets), so a caller can capture
// it in a HIP graph and replay it every iteration.
struct JacobiPlanBase {
    virtual ~JacobiPlanBase() = default;
    virtual void run(int sweeps, double* changes_dev, hipStream_t st) = 0;
    virtual void sweep(bool first, bool last, int dst_scratch, hipStream_t st) = 0;
    virtual void change_host(double* out, int n, hipStream_t st) = 0;
    int nsys = 0, tiles = 0, nzmax = 1;
    int64_t unknowns = 0;
};

// A device copy of a system table, uploaded the first time its bytes are
// seen on the stream: an iteration that sweeps the same systems every time
// copies nothing and waits for nothing. Entries are dropped by
// Workspace::release of their stream.
struct TableEntry {
    std::vector<char> key;
    void* dev;
};
std::mutex g_table_mu;
std::map<std::pair<int, hipStream_t>, std::vector<TableEntry>> g_tables;

size_t table_cache_bytes(int device, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(g_table_mu);
    auto it = g_tables.find({device, st});
    size_t n = 0;
    if (it != g_tables.end()) {
        for (const TableEntry& e : it->second) { n += e.key.size(); }
    }
    return n;
}

void table_cache_drop(int device, hipStream_t st)
{
    std::lock_guard<std::mutex> lock(g_table_mu);
    auto it = g_tables.find({device, st});
    if (it == g_tables.end()) { return; }
    for (TableEntry& e : it->second) { (void)hipFree(e.dev); }
    g_tables.erase(it);
}

void* table_cache_get(hipStream_t st, const void* tab, size_t bytes)
{
    static const bool hooked = [] {
        register_stream_cache(StreamCacheHook{&table_cache_bytes, &table_cache_drop});
        return true;
    }();
    (void)hooked;
    std::vector<char> key(bytes);
    std::memcpy(key.data(), tab, bytes);
    int device = 0;
    GUST_HIP(hipGetDevice(&device));
    std::lock_guard<std::mutex> lock(g_table_mu);
    std::vector<TableEntry>& list = g_tables[{device, st}];
    for (TableEntry& e : list) {
        if (e.key == key) { return e.dev; }
    }
    if (list.size() >= 64) {
        GUST_HIP(hipStreamSynchronize(st));
        for (TableEntry& e : list) { (void)hipFree(e.dev); }
        list.clear();
    }
    void* dev = nullptr;
    GUST_HIP(hipMalloc(&dev, bytes));
    // on the sweep's stream (a null-stream copy need not finish before
    // that stream's kernels start), complete before the host returns
    GUST_HIP(hipMemcpyAsync(dev, tab, bytes, hipMemcpyHostToDevice, st));
    GUST_HIP(hipStreamSynchronize(st));
    list.push_back(TableEntry{std::move(key), dev});
    return dev;
}

template <typename T>
struct JacobiPlan final : JacobiPlanBase {
    std::vector<gust_line_system> sys;
    std::vector<JacobiSys<T>> tab;
    std::vector<IBox> valid;
    JacobiSys<T>* dtab = nullptr;
    T* scratch = nullptr;
    T* cglob = nullptr;
    unsigned long long* chg = nullptr;
    bool owned = false;   // buffers from hipMalloc (a plan handle), not the stream workspace
    bool chg_zeroed = false;   // chg already holds zeros (a fresh deferred slot)
    int staged_w = 0;          // > 0: the LDS-staged schedule, W columns per LDS row
    int wmax_ = 1;             // the widest tile (columns)
    bool staged_fits = false;  // a tile's operands fit the LDS
    int kp = 4;                // the streaming kernel's prefetch depth
    bool tune = false;         // the first run times the candidates (schedule 0)
    bool no_alloc = false;     // scratch carved from the caller's bytes
    int regs_nz = 0;           // > 0: the register kernel, unrolled to this many levels
    int regs_cl = 0;           // the register kernel's levels of c' in LDS

    ~JacobiPlan() override
    {
        if (owned) {
            (void)hipFree(dtab);
            (void)hipFree(scratch);
            (void)hipFree(cglob);
            (void)hipFree(chg);
        }
    }

    // own: buffers from hipMalloc (a plan handle); ext != nullptr: carved
    // from the caller's ext_cap bytes (no allocation of any kind); neither:
    // the stream workspace.
    void init(int n, const gust_line_system* in, bool own, hipStream_t st, char* ext = nullptr,
              size_t ext_cap = 0)
    {
        const char* op = ext ? "zline_jacobi_noalloc" : "zline_jacobi";
        if (n < 1 || !in) { GUST_FAIL("gust: %s: no systems", op); }
        nsys = n;
        owned = own;
        sys.assign(in, in + n);
        tab.resize(size_t(n));
        valid.resize(size_t(n));
        size_t scratch_elems = 0;
        for (int s = 0; s < n; ++s) {
            const gust_line_system& q = sys[size_t(s)];
            IBox V;
            check_line_system(op, q.loc, &q.A, &q.ap, &q.rhs, 1, &q.geom, &q.delta, V);
            valid[size_t(s)] = V;
            JacobiSys<T>& t = tab[size_t(s)];
            t.A = at_lo<T>(q.A, V);
            t.a_j = q.A.jstride; t.a_k = q.A.kstride; t.a_n = q.A.nstride;
            t.ap = at_lo<T>(q.ap, V);
            t.p_j = q.ap.jstride; t.p_k = q.ap.kstride;
            t.r = at_lo<T>(q.rhs, V);
            t.r_j = q.rhs.jstride; t.r_k = q.rhs.kstride;
            t.d = at_lo<T>(q.delta, V);
            t.d_j = q.delta.jstride; t.d_k = q.delta.kstride;
            t.nx = V.len(0); t.ny = V.len(1); t.nz = V.len(2);
            t.Lx = q.geom.periodic[0] ? (q.geom.hi[0] - q.geom.lo[0] + 1) : 0;
            t.Ly = q.geom.periodic[1] ? (q.geom.hi[1] - q.geom.lo[1] + 1) : 0;
            t.s_j = t.nx;
            t.s_k = int64_t(t.nx) * t.ny;
            t.s = reinterpret_cast<T*>(scratch_elems);   // offset, fixed up below
            scratch_elems += size_t(V.npts());
            unknowns += V.npts();
            t.tile0 = tiles;
            t.pad = 0;
            const int64_t cols = int64_t(t.nx) * t.ny;
            const int64_t nt = (cols + kWave - 1) / kWave;
            if (int64_t(tiles) + nt > (int64_t(1) << 30)) { GUST_FAIL("gust: %s: too many columns in one batch", op); }
            tiles += int(nt);
            nzmax = std::max(nzmax, t.nz);
        }
        const size_t lds = size_t(kWave) * size_t(nzmax) * sizeof(T);
        // Schedule: few tiles (under kStagedMaxTiles, about four waves per
        // CU or fewer) take the LDS-staged kernel when a tile's operands fit
        // the LDS; many tiles the register kernel or the streaming kernel
        // (below).
        {
            int wmax = 1;
            for (const auto& t : tab) {
                const int64_t cols = int64_t(t.nx) * t.ny;
                wmax = int(std::max<int64_t>(wmax, std::min<int64_t>(cols, kWave)));
            }
            const size_t staged = 5 * size_t(nzmax) * size_t(wmax) * sizeof(T);
            const int sched = jacobi_schedule();
            const bool fits = staged <= device_info().lds_per_block;
            wmax_ = wmax;
            staged_fits = fits;
            if (fits && (sched == 2 || ((sched == 0 || sched == 3) && tiles < kStagedMaxTiles))) { staged_w = wmax; }
            if (sched == 4 && offsets_fit32()) { regs_nz = reg_bucket(nzmax); regs_cl = regs_nz; }
            tune = sched == 0 && !ext;
        }
        no_alloc = ext != nullptr;
        // c' of the streaming kernel: in LDS (default) or in global memory
        // (GUST_JACOBI_CPRIME = global: no LDS limit on occupancy, c'
        // round-trips through L2).
        const char* cp = std::getenv("GUST_JACOBI_CPRIME");
        const bool force_global = cp && std::strcmp(cp, "global") == 0;
        if (force_global) { tune = false; }
        // Many tiles under the fixed rule (and the autotuner's starting
        // point): the register kernel where the columns fit a level bucket
        // (the fastest candidate on the configurations' tiles), else the
        // streaming kernel.
        if (staged_w == 0 && regs_nz == 0 && !force_global && (jacobi_schedule() == 0 || jacobi_schedule() == 3)) {
            if (!apply_choice(sizeof(T) == 8 ? kChoiceRegsSplit : kChoiceRegsC, st)) { apply_choice(kChoiceRegs, st); }
        }
        const bool use_lds = !force_global && lds <= device_info().lds_per_block;
        const size_t tab_bytes = tab.size() * sizeof(JacobiSys<T>);
        const size_t cbytes = use_lds ? 0 : size_t(tiles) * lds;
        if (ext) {
            // [changes | table | iterate], each 256-byte aligned
            if (cbytes) {
                GUST_FAIL("gust: %s: %d levels need c' outside the LDS, which this entry point does not hold",
                          op, nzmax);
            }
            auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
            const size_t need = up(size_t(n) * 8) + up(tab_bytes) + up(scratch_elems * sizeof(T));
            if (need > ext_cap) {
                GUST_FAIL("gust: %s: the systems need %zu bytes of scratch, more than the fixed %zu bytes "
                          "this entry point holds per device", op, need, ext_cap);
            }
            chg = reinterpret_cast<unsigned long long*>(ext);
            dtab = reinterpret_cast<JacobiSys<T>*>(ext + up(size_t(n) * 8));
            scratch = reinterpret_cast<T*>(ext + up(size_t(n) * 8) + up(tab_bytes));
            for (auto& t : tab) { t.s = scratch + reinterpret_cast<size_t>(t.s); }
            // The table's host copy lives in this plan; the wait ends the
            // copy before the plan can go away.
            GUST_HIP(hipMemcpyAsync(dtab, tab.data(), tab_bytes, hipMemcpyHostToDevice, st));
            GUST_HIP(hipStreamSynchronize(st));
            return;
        }
        if (own) {
            GUST_HIP(hipMalloc(&scratch, scratch_elems * sizeof(T) + 256));
            GUST_HIP(hipMalloc(&chg, size_t(n) * 8 + 8));
            GUST_HIP(hipMalloc(&dtab, tab_bytes));
            if (cbytes) { GUST_HIP(hipMalloc(&cglob, cbytes)); }
        } else {
            scratch = static_cast<T*>(Workspace::get(st, 0, scratch_elems * sizeof(T) + 256));
            chg = static_cast<unsigned long long*>(Workspace::get(st, 1, size_t(n) * 8 + 8));
            if (cbytes) { cglob = static_cast<T*>(Workspace::get(st, 5, cbytes)); }
        }
        for (auto& t : tab) { t.s = scratch + reinterpret_cast<size_t>(t.s); }
        if (own) {
            // The table travels once; the synchronisation keeps the host
            // copy's lifetime out of the picture.
            GUST_HIP(hipMemcpyAsync(dtab, tab.data(), tab_bytes, hipMemcpyHostToDevice, st));
            GUST_HIP(hipStreamSynchronize(st));
        } else {
            dtab = cached_table(st, tab);
        }
    }

    static JacobiSys<T>* cached_table(hipStream_t st, const std::vector<JacobiSys<T>>& tab)
    {
        return static_cast<JacobiSys<T>*>(
            table_cache_get(st, tab.data(), tab.size() * sizeof(JacobiSys<T>)));
    }

    // The register kernel addresses every array through a buffer resource
    // with a 32-bit lane offset plus a 32-bit level offset: every byte it
    // reaches (one component of A, ap, rhs, delta, the scratch iterate) must
    // lie within 2^31 bytes of the array's base.
    bool offsets_fit32() const
    {
        const int64_t lim = int64_t(1) << 31;
        for (const auto& t : tab) {
            const int64_t last = int64_t(t.nx - 1), top = int64_t(t.nz - 1), ny1 = int64_t(t.ny - 1);
            const int64_t js[5] = {t.a_j, t.p_j, t.r_j, t.d_j, t.s_j};
            const int64_t ks[5] = {t.a_k, t.p_k, t.r_k, t.d_k, t.s_k};
            for (int q = 0; q < 5; ++q) {
                if (js[q] < 0 || ks[q] < 0) { return false; }
                if ((last + ny1 * js[q]) * int64_t(sizeof(T)) >= lim) { return false; }
                if (top * ks[q] * int64_t(sizeof(T)) >= lim) { return false; }
            }
        }
        return true;
    }

    static int reg_bucket(int nz)
    {
        for (int b : kRegBuckets) {
            if (nz <= b) { return b; }
        }
        return 0;
    }

    template <bool F, bool L, int NZ, int CL>
    void launch_regs_at(int dst_scratch, hipStream_t st)
    {
        const size_t lds = size_t(kWave) * size_t(std::min(nzmax, CL)) * sizeof(T);
        if (lds > 65536) {
            GUST_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&zjacobi_regs_kernel<T, F, L, NZ, CL>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        }
        zjacobi_regs_kernel<T, F, L, NZ, CL><<<tiles, kWave, lds, st>>>(dtab, nsys, tiles, dst_scratch, chg);
    }

    // regs_cl: levels of c' in LDS (regs_nz: all of them; 0: none).
    template <bool F, bool L>
    void launch_regs(int dst_scratch, hipStream_t st)
    {
        constexpr int b0 = kRegBuckets[0], b1 = kRegBuckets[1], b2 = kRegBuckets[2];
        if (regs_nz == b0) {
            if (regs_cl == 0) { launch_regs_at<F, L, b0, 0>(dst_scratch, st); }
            else { launch_regs_at<F, L, b0, b0>(dst_scratch, st); }
        } else if (regs_nz == b1) {
            if (regs_cl == 0) { launch_regs_at<F, L, b1, 0>(dst_scratch, st); }
            else { launch_regs_at<F, L, b1, b1>(dst_scratch, st); }
        } else if (regs_cl == b2) {
            launch_regs_at<F, L, b2, b2>(dst_scratch, st);
        } else if constexpr (sizeof(T) == 4) {
            launch_regs_at<F, L, b2, 0>(dst_scratch, st);
        } else {
            launch_regs_at<F, L, b2, kSplitLevels>(dst_scratch, st);
        }
    }

    template <bool F, bool L>
    void launch(int dst_scratch, hipStream_t st)
    {
        if (regs_nz > 0) {
            launch_regs<F, L>(dst_scratch, st);
            GUST_HIP(hipGetLastError());
            return;
        }
        if (staged_w > 0) {
            const size_t lds = size_t(L && !F ? 5 : 4) * size_t(nzmax) * size_t(staged_w) * sizeof(T);
            if (lds > 65536) {
                GUST_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&zjacobi_staged_kernel<T, F, L>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
            }
            zjacobi_staged_kernel<T, F, L><<<tiles, kStagedThreads, lds, st>>>(dtab, nsys, tiles, dst_scratch, chg,
                                                                              staged_w);
            GUST_HIP(hipGetLastError());
            return;
        }
        if (cglob) {
            zjacobi_sweep_kernel<T, F, L, false><<<tiles, kWave, 0, st>>>(dtab, nsys, tiles, dst_scratch, chg, cglob,
                                                                       nzmax);
        } else if (kp == 2) {
            launch_lds<F, L, 2>(dst_scratch, st);
        } else if (kp == 8) {
            launch_lds<F, L, 8>(dst_scratch, st);
        } else {
            launch_lds<F, L, 4>(dst_scratch, st);
        }
        GUST_HIP(hipGetLastError());
    }

    template <bool F, bool L, int P>
    void launch_lds(int dst_scratch, hipStream_t st)
    {
        const size_t lds = size_t(kWave) * size_t(nzmax) * sizeof(T);
        if (lds > 65536) {
            GUST_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&zjacobi_sweep_kernel<T, F, L, true, P>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
        }
        zjacobi_sweep_kernel<T, F, L, true, P><<<tiles, kWave, lds, st>>>(dtab, nsys, tiles, dst_scratch, chg,
                                                                          cglob, nzmax);
    }

    // Makes candidate c this plan's schedule (false: it does not apply).
    bool apply_choice(int c, hipStream_t st)
    {
        const size_t lds = size_t(kWave) * size_t(nzmax) * sizeof(T);
        regs_nz = 0;
        regs_cl = 0;
        if (c == kChoiceStaged) {
            if (!staged_fits) { return false; }
            staged_w = wmax_;
            return true;
        }
        staged_w = 0;
        if (c == kChoiceRegs || c == kChoiceRegsC || c == kChoiceRegsSplit) {
            const int b = offsets_fit32() ? reg_bucket(nzmax) : 0;
            if (b == 0) { return false; }
            int cl = b;
            if (c == kChoiceRegsC) {
                // c' in registers too: FP32 columns, or FP64 columns of the
                // two smaller buckets (twice the registers per level)
                if (sizeof(T) == 8 && b == kRegBuckets[2]) { return false; }
                cl = 0;
            } else if (c == kChoiceRegsSplit) {
                // FP64 columns of the largest bucket: the top levels' c' in
                // registers, four waves per CU by LDS instead of three
                if (sizeof(T) != 8 || b != kRegBuckets[2] || nzmax <= kSplitLevels) { return false; }
                cl = kSplitLevels;
            }
            if (size_t(kWave) * size_t(std::min(nzmax, cl)) * sizeof(T) > device_info().lds_per_block) { return false; }
            regs_nz = b;
            regs_cl = cl;
            return true;
        }
        if (c == kChoiceGlobal) {
            if (no_alloc) { return false; }
            if (!cglob) {
                if (owned) { return false; }   // a plan handle holds c' in global memory only when LDS cannot
                cglob = static_cast<T*>(Workspace::get(st, 5, size_t(tiles) * lds));
            }
            return true;
        }
        if (lds > device_info().lds_per_block) { return false; }
        if (!owned) { cglob = nullptr; }
        kp = c == kChoiceP2 ? 2 : c == kChoiceP8 ? 8 : 4;
        return true;
    }

    TuneKey key(int sweeps) const
    {
        TuneKey k;
        k.dtype = sizeof(T) == 8 ? 0 : 1;
        k.sweeps = sweeps;
        for (const auto& t : tab) {
            k.shape.push_back(t.nx);
            k.shape.push_back(t.ny);
            k.shape.push_back(t.nz);
        }
        return k;
    }

    void sweep(bool first, bool last, int dst_scratch, hipStream_t st) override
    {
        if (first && last) { launch<true, true>(dst_scratch, st); }
        else if (first) { launch<true, false>(dst_scratch, st); }
        else if (last) { launch<false, true>(dst_scratch, st); }
        else { launch<false, false>(dst_scratch, st); }
    }

    void run(int sweeps, double* changes_dev, hipStream_t st) override
    {
        if (sweeps < 1) { GUST_FAIL("gust: zline_jacobi: sweeps must be >= 1 (got %d)", sweeps); }
        if (tune) {
            tune = false;
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            GUST_HIP(hipStreamIsCapturing(st, &cs));
            autotune(sweeps, st, cs == hipStreamCaptureStatusNone);
        }
        run_once(sweeps, changes_dev, st);
    }

    // The first plan of a set of shapes times every candidate that applies
    // (each run gives the same bits: delta starts at 0, the change maxima are
    // idempotent); later plans of those shapes take the recorded fastest.
    void autotune(int sweeps, hipStream_t st, bool may_time)
    {
        const TuneKey k = key(sweeps);
        {
            std::lock_guard<std::mutex> lock(g_tune_mu);
            auto it = g_tuned.find(k);
            if (it != g_tuned.end()) {
                apply_choice(it->second.choice, st);
                return;
            }
        }
        if (!may_time) { return; }   // under graph capture: the fixed rule
        TuneRecord rec;
        hipEvent_t e0, e1;
        GUST_HIP(hipEventCreate(&e0));
        GUST_HIP(hipEventCreate(&e1));
        float best = -1;
        for (int c = 0; c < kNumChoices; ++c) {
            if (!apply_choice(c, st)) { continue; }
            run_once(sweeps, nullptr, st);   // the warm-up of this candidate
            GUST_HIP(hipEventRecord(e0, st));
            run_once(sweeps, nullptr, st);
            GUST_HIP(hipEventRecord(e1, st));
            GUST_HIP(hipEventSynchronize(e1));
            float ms = 0;
            GUST_HIP(hipEventElapsedTime(&ms, e0, e1));
            rec.ms[c] = ms;
            if (best < 0 || ms < best) {
                best = ms;
                rec.choice = c;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        apply_choice(rec.choice, st);
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned[k] = rec;
    }

    void run_once(int sweeps, double* changes_dev, hipStream_t st)
    {
        if (!chg_zeroed) {
            zero_u64_kernel<<<(nsys + 255) / 256, 256, 0, st>>>(chg, nsys);
            GUST_HIP(hipGetLastError());
        }
        // delta = 0 outside the domain (check_line_system's setVal).
        zero_ghosts_all<T>(sys, valid, st);
        // The sweeps ping-pong between delta and the scratch iterate, arranged
        // so the last sweep writes delta; sweep 0 reads the zero iterate
        // implicitly.
        for (int s = 0; s < sweeps; ++s) {
            const int dst_scratch = ((sweeps - 1 - s) % 2 == 1) ? 1 : 0;
            sweep(s == 0, s == sweeps - 1, dst_scratch, st);
        }
        // OverrideSync of the periodic seam faces (the low copy owns).
        for (int s = 0; s < nsys; ++s) {
            const gust_line_system& q = sys[size_t(s)];
            const JacobiSys<T>& t = tab[size_t(s)];
            for (int d = 0; d < 2; ++d) {
                if (!((q.loc >> d) & 1) || !q.geom.periodic[d]) { continue; }
                const int64_t n = d == 0 ? int64_t(t.ny) * t.nz : int64_t(t.nx) * t.nz;
                const int blocks = int(std::min<int64_t>((n + 255) / 256, 2048));
                seam_sync_kernel<T><<<blocks, 256, 0, st>>>(t.d, t.d_j, t.d_k, t.nx, t.ny, t.nz, d, 1);
                GUST_HIP(hipGetLastError());
            }
        }
        if (changes_dev) {
            bits_to_double_kernel<<<(nsys + 255) / 256, 256, 0, st>>>(chg, changes_dev, nsys);
            GUST_HIP(hipGetLastError());
        }
    }

    void change_host(double* out, int n, hipStream_t st) override
    {
        auto* h = static_cast<unsigned long long*>(Workspace::get_host_pinned(0, size_t(n) * 8));
        GUST_HIP(hipMemcpyAsync(h, chg, size_t(n) * 8, hipMemcpyDeviceToHost, st));
        GUST_HIP(hipStreamSynchronize(st));
        std::memcpy(out, h, size_t(n) * 8);
    }
};

JacobiPlanBase* make_plan(int dtype)
{
    if (dtype == 0) { return new JacobiPlan<double>(); }
    if (dtype == 1) { return new JacobiPlan<float>(); }
    GUST_FAIL("gust: zline_jacobi: unknown dtype %d (0 = f64, 1 = f32)", dtype);
}

template <typename T>
void zjacobi_run(int nsys, const gust_line_system* sys, int sweeps, double* changes_dev, double* change_host,
                 hipStream_t st)
{
    if (sweeps < 1) { GUST_FAIL("gust: zline_jacobi: sweeps must be >= 1 (got %d)", sweeps); }
    JacobiPlan<T> plan;
    plan.init(nsys, sys, false, st);
    Deferred* dq = change_host && Deferred::current() && Deferred::current()->stream() == st ? Deferred::current()
                                                                                               : nullptr;
    if (dq && nsys == 1) {
        // the sweep records the change straight into a slot of the batch,
        // read with its other values
        plan.chg = dq->zero_slot([change_host](unsigned long long b) { *change_host = bits_to_double(b); });
        plan.chg_zeroed = true;
    }
    plan.run(sweeps, changes_dev, st);
    if (!change_host || plan.chg_zeroed) { return; }
    if (dq) {
        unsigned long long* slot = dq->zero_slot([change_host](unsigned long long b) { *change_host = bits_to_double(b); });
        copy_u64_kernel<<<1, 64, 0, st>>>(plan.chg, slot, 1);
        GUST_HIP(hipGetLastError());
        return;
    }
    plan.change_host(change_host, 1, st);
}

// ---------------------------------------------------------------------------
// Line LU-SGS. The Thomas denominators m_k and ratios c'_k of a column
// depend on its coefficients only, not on the right-hand side, so one
// parallel pre-pass over every column of the sweep region computes them for
// all the sweeps of the call. A sweep then walks the (i, j) diagonals in
// order (a diagonal's columns are independent): all threads of the
// workgroup form the diagonal's right-hand sides (every column, every level)
// from the neighbours the previous diagonal left, then one thread per column
// runs the forward recurrence e_k = (r_k - l_k e_{k-1}) / m_k and the back
// substitution with their operands loaded ahead in chunks. Each operation
// is the column Thomas solve's own, in its order (the bits of the sequential
// column-by-column pass). One workgroup per system; a batch of systems (u,
// v, w of the momentum predictor) shares the launch.

constexpr int kLusgsMaxBatch = 3;
constexpr int kLusgsChunk = 8;
constexpr int kLusgsRhsBatch = 6;
constexpr size_t kLusgsMaxLds = 144 * 1024;

template <typename T>
struct LusgsSys {
    const T* A;          // at the valid lo corner
    int64_t a_j, a_k, a_n;
    const T* ap;
    int64_t p_j, p_k;
    const T* r;
    int64_t r_j, r_k;
    T* d;                // the caller's delta at the valid lo corner (1 ghost)
    int64_t d_j, d_k;
    T* w;                // zero ghosts: work array at valid lo (ng + 1 ghosts)
    int64_t w_j, w_k;
    int w_gx, w_gy;      // work ghost widths
    T* c;                // c'_k over the sweep region
    T* e;                // right-hand sides, then e'_k
    T* m;                // the denominators m_k
    int64_t c_j, c_k;
    int nx, ny, nz;      // valid extents at loc (local index 0 = valid lo)
    int Lx, Ly;          // periodic shifts (cells), 0 if not periodic
    int s_lo[2], s_hi[2];// sweep region S (local indices)
    int o_lo[2], o_hi[2];// owned region (local indices)
    int fold[2];         // one-cell periodic axis folded into the diagonal
    int zero_mode;       // 1: zero ghosts (work array), 0: lagged (in delta)
    int loc;
    T* sk[10];           // the skewed copies (lagged, no alternation): kSk*
    int sP, sJ;          // their extents: sP = nx + ny + 3 diagonals, sJ = ny + 2
};

template <typename T>
struct LusgsBatch {
    LusgsSys<T> s[kLusgsMaxBatch];
};

// Coefficient/ap/rhs index (local) of a point of the grown region: the
// periodic image inside the valid box.
__device__ __forceinline__ int wrap_idx(int i, int n, int L)
{
    if (L) {
        if (i < 0) { i += L; }
        else if (i > n - 1) { i -= L; }
    }
    return i;
}

// m_k and c'_k of every column of the sweep region (blockIdx.y: the
// system). The fold of a periodic one-cell axis (Source/LineSweeps.cpp:
// 242-256) moves its two neighbour coefficients into the diagonal.
template <typename T>
__global__ __launch_bounds__(256) void lusgs_factor_kernel(const LusgsBatch<T> B)
{
    const LusgsSys<T>& S = B.s[blockIdx.y];
    const int sx = S.s_hi[0] - S.s_lo[0] + 1, sy = S.s_hi[1] - S.s_lo[1] + 1;
    const int64_t ncol = int64_t(sx) * sy;
    for (int64_t t = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; t < ncol;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int i = S.s_lo[0] + int(t % sx), j = S.s_lo[1] + int(t / sx);
        const int ii = wrap_idx(i, S.nx, S.Lx), jj = wrap_idx(j, S.ny, S.Ly);
        const T* a = S.A + ii + int64_t(jj) * S.a_j;
        const T* pap = S.ap + ii + int64_t(jj) * S.p_j;
        T* cc = S.c + t;
        T* mm = S.m + t;
        T cprev = T(0);
        for (int k = 0; k < S.nz; ++k) {
            const T* ak = a + k * S.a_k;
            T dg = pap[k * S.p_k];
            if (S.fold[0]) { dg -= ak[1 * S.a_n] + ak[2 * S.a_n]; }
            if (S.fold[1]) { dg -= ak[3 * S.a_n] + ak[4 * S.a_n]; }
            const T lower = -ak[5 * S.a_n], upper = -ak[6 * S.a_n];
            const T mk = k == 0 ? dg : dg - lower * cprev;
            const T ck = upper / mk;
            cc[k * S.c_k] = ck;
            mm[k * S.c_k] = mk;
            cprev = ck;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void lusgs_kernel(const LusgsBatch<T> B, int sweeps, int alternate,
                                                    unsigned long long* __restrict__ changes)
{
    const LusgsSys<T>& S = B.s[blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    T* W;
    int64_t wj, wk;
    if (S.zero_mode) {
        W = S.w; wj = S.w_j; wk = S.w_k;
        // The box-local iterate: 0 everywhere, ghosts included.
        const int gx = S.w_gx, gy = S.w_gy;
        const int tx = S.nx + 2 * gx, ty = S.ny + 2 * gy;
        const int64_t n = int64_t(tx) * ty * S.nz;
        for (int64_t t = tid; t < n; t += nth) {
            const int i = int(t % tx) - gx;
            const int64_t rr = t / tx;
            const int j = int(rr % ty) - gy, k = int(rr / ty);
            W[i + j * wj + k * wk] = T(0);
        }
    } else {
        W = S.d; wj = S.d_j; wk = S.d_k;
        // delta starts at 0 (check_line_system's setVal): the passes read
        // the columns not yet updated, so nothing a previous call left there
        // may remain (the ghosts are zeroed before the launch).
        const int64_t n = int64_t(S.nx) * S.ny * S.nz;
        for (int64_t t = tid; t < n; t += nth) {
            const int i = int(t % S.nx);
            const int64_t rr = t / S.nx;
            const int j = int(rr % S.ny), k = int(rr / S.ny);
            W[i + j * wj + k * wk] = T(0);
        }
    }
    __syncthreads();
    const int nz = S.nz;
    const int sx = S.s_hi[0] - S.s_lo[0] + 1, sy = S.s_hi[1] - S.s_lo[1] + 1;
    const int last = sx + sy - 2;
    unsigned long long chg = 0;
    for (int s = 0; s < sweeps; ++s) {
        const bool jrev = alternate && (s % 2 == 1);
        for (int f = 1; f >= 0; --f) {
            if (!S.zero_mode) {
                // FillBoundary: the periodic images into delta's ghosts.
                if (S.Lx) {
                    const int64_t n = int64_t(S.ny) * S.nz;
                    for (int64_t t = tid; t < n; t += nth) {
                        const int j = int(t % S.ny), k = int(t / S.ny);
                        T* row = W + j * wj + k * wk;
                        row[-1] = row[S.Lx - 1];
                        row[S.nx] = row[S.nx - S.Lx];
                    }
                }
                if (S.Ly) {
                    const int64_t n = int64_t(S.nx) * S.nz;
                    for (int64_t t = tid; t < n; t += nth) {
                        const int i = int(t % S.nx), k = int(t / S.nx);
                        T* colp = W + i + k * wk;
                        colp[-wj] = colp[(S.Ly - 1) * wj];
                        colp[S.ny * wj] = colp[(S.ny - S.Ly) * wj];
                    }
                }
                __syncthreads();
            }
            const bool track = (s == sweeps - 1) && f == 0;
            for (int tt = 0; tt <= last; ++tt) {
                const int p = f ? tt : last - tt;
                const int amin = p - (sy - 1) > 0 ? p - (sy - 1) : 0;
                const int amax = p < sx - 1 ? p : sx - 1;
                const int ncd = amax - amin + 1;
                // 1. the right-hand sides of the diagonal's columns
                const int nq = ncd * nz;
                for (int q = tid; q < nq; q += nth) {
                    const int ia = amin + q % ncd, k = q / ncd;
                    const int i = S.s_lo[0] + ia;
                    const int j = jrev ? S.s_hi[1] - (p - ia) : S.s_lo[1] + p - ia;
                    const int ii = wrap_idx(i, S.nx, S.Lx), jj = wrap_idx(j, S.ny, S.Ly);
                    const T* ak = S.A + ii + int64_t(jj) * S.a_j + k * S.a_k;
                    T axm = ak[1 * S.a_n], axp = ak[2 * S.a_n], aym = ak[3 * S.a_n], ayp = ak[4 * S.a_n];
                    if (S.fold[0]) { axm = T(0); axp = T(0); }
                    if (S.fold[1]) { aym = T(0); ayp = T(0); }
                    const T* wk0 = W + i + int64_t(j) * wj + k * wk;
                    const T rk = S.r[ii + int64_t(jj) * S.r_j + k * S.r_k] + axm * wk0[-1] + axp * wk0[1] +
                                 aym * wk0[-wj] + ayp * wk0[wj];
                    S.e[int64_t(i - S.s_lo[0]) + int64_t(j - S.s_lo[1]) * S.c_j + k * S.c_k] = rk;
                }
                __syncthreads();
                // 2. the recurrences, one thread per column
                for (int mi = tid; mi < ncd; mi += nth) {
                    const int ia = amin + mi;
                    const int i = S.s_lo[0] + ia;
                    const int j = jrev ? S.s_hi[1] - (p - ia) : S.s_lo[1] + p - ia;
                    const int ii = wrap_idx(i, S.nx, S.Lx), jj = wrap_idx(j, S.ny, S.Ly);
                    const T* az = S.A + ii + int64_t(jj) * S.a_j + 5 * S.a_n;
                    const int64_t cb = int64_t(i - S.s_lo[0]) + int64_t(j - S.s_lo[1]) * S.c_j;
                    T* ee = S.e + cb;
                    const T* cc = S.c + cb;
                    const T* mm = S.m + cb;
                    T* col = W + i + int64_t(j) * wj;
                    const bool tr = track && i >= S.o_lo[0] && i <= S.o_hi[0] && j >= S.o_lo[1] && j <= S.o_hi[1];
                    T eprev = T(0);
                    for (int k0 = 0; k0 < nz; k0 += kLusgsChunk) {
                        T rr[kLusgsChunk], ll[kLusgsChunk], dd[kLusgsChunk];
#pragma unroll
                        for (int u = 0; u < kLusgsChunk; ++u) {
                            const int k = k0 + u < nz ? k0 + u : nz - 1;
                            rr[u] = ee[k * S.c_k];
                            ll[u] = -az[k * S.a_k];
                            dd[u] = mm[k * S.c_k];
                        }
#pragma unroll
                        for (int u = 0; u < kLusgsChunk; ++u) {
                            const int k = k0 + u;
                            if (k < nz) {
                                const T ek = k == 0 ? rr[u] / dd[u] : (rr[u] - ll[u] * eprev) / dd[u];
                                ee[k * S.c_k] = ek;
                                eprev = ek;
                            }
                        }
                    }
                    T o = eprev;
                    {
                        T* pz = col + (nz - 1) * wk;
                        if (tr) {
                            const unsigned long long b = dev::abs_bits(double(*pz - o));
                            chg = b > chg ? b : chg;
                        }
                        *pz = o;
                    }
                    for (int k0 = nz - 2; k0 >= 0; k0 -= kLusgsChunk) {
                        T ev[kLusgsChunk], cv[kLusgsChunk], ov[kLusgsChunk];
#pragma unroll
                        for (int u = 0; u < kLusgsChunk; ++u) {
                            const int k = k0 - u >= 0 ? k0 - u : 0;
                            ev[u] = ee[k * S.c_k];
                            cv[u] = cc[k * S.c_k];
                            ov[u] = tr ? col[k * wk] : T(0);
                        }
#pragma unroll
                        for (int u = 0; u < kLusgsChunk; ++u) {
                            const int k = k0 - u;
                            if (k >= 0) {
                                o = ev[u] - cv[u] * o;
                                if (tr) {
                                    const unsigned long long b = dev::abs_bits(double(ov[u] - o));
                                    chg = b > chg ? b : chg;
                                }
                                col[k * wk] = o;
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    // max |change| over the block.
    __shared__ unsigned long long red[256 / 64];
    chg = dev::wave_max_u64(chg);
    if ((tid & 63) == 0) { red[tid >> 6] = chg; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long mx = 0;
        for (int w = 0; w < (nth + 63) / 64; ++w) { mx = red[w] > mx ? red[w] : mx; }
        changes[blockIdx.x] = mx;
    }
    // The result into delta (zero ghosts) and the seam copies.
    if (S.zero_mode) {
        const int64_t n = int64_t(S.nx) * S.ny * S.nz;
        for (int64_t t = tid; t < n; t += nth) {
            const int i = int(t % S.nx);
            const int64_t rr = t / S.nx;
            const int j = int(rr % S.ny), k = int(rr / S.ny);
            S.d[i + j * S.d_j + k * S.d_k] = W[i + j * wj + k * wk];
        }
        __syncthreads();
    } else {
        // lagged leaves stale periodic images in delta's ghosts: zero them
        // (no caller reads them; zero outside the domain, as documented).
        if (S.Lx) {
            const int64_t n = int64_t(S.ny) * S.nz;
            for (int64_t t = tid; t < n; t += nth) {
                const int j = int(t % S.ny), k = int(t / S.ny);
                T* row = S.d + j * S.d_j + k * S.d_k;
                row[-1] = T(0);
                row[S.nx] = T(0);
            }
        }
        if (S.Ly) {
            const int64_t n = int64_t(S.nx) * S.nz;
            for (int64_t t = tid; t < n; t += nth) {
                const int i = int(t % S.nx), k = int(t / S.nx);
                T* colp = S.d + i + k * S.d_k;
                colp[-S.d_j] = T(0);
                colp[S.ny * S.d_j] = T(0);
            }
        }
        __syncthreads();
    }
    for (int d = 0; d < 2; ++d) {
        if (!((S.loc >> d) & 1) || !(d == 0 ? S.Lx : S.Ly)) { continue; }
        if (S.zero_mode && !(S.loc == GUST_LOC_XFACE || S.loc == GUST_LOC_YFACE)) { continue; }
        const int64_t n = d == 0 ? int64_t(S.ny) * S.nz : int64_t(S.nx) * S.nz;
        for (int64_t t = tid; t < n; t += nth) {
            int64_t lo, hi;
            if (d == 0) {
                const int j = int(t % S.ny), k = int(t / S.ny);
                lo = j * S.d_j + k * S.d_k;
                hi = lo + (S.nx - 1);
            } else {
                const int i = int(t % S.nx), k = int(t / S.nx);
                lo = i + k * S.d_k;
                hi = lo + (S.ny - 1) * S.d_j;
            }
            // lagged: AMReX's default owner (low copy); zero: line_owned_box
            // (the high copy, Source/LineSweeps.cpp:361-374).
            if (S.zero_mode) { S.d[lo] = S.d[hi]; } else { S.d[hi] = S.d[lo]; }
        }
    }
}


// The lagged, non-alternating sweeps (the default) on skewed copies. Point
// (i, j, k), i in [-1, nx], j in [-1, ny], sits at
// (((i + j + 2) sJ + j + 1) nz + k): one (i + j) diagonal is one contiguous
// run (its columns consecutive in j, each column's levels consecutive), so
// the diagonal's loads coalesce and touch few pages (in the caller's layout
// each column of a diagonal is its own cache line and each level its own
// page range); its four neighbours sit at -sJ nz, +sJ nz, -(sJ + 1) nz and
// +(sJ + 1) nz. The arithmetic is the column Thomas solve's, in its order
// (the same bits).
enum { kSkW = 0, kSkAxm, kSkAxp, kSkAym, kSkAyp, kSkR, kSkLz, kSkM, kSkC, kSkN };

template <typename T>
__device__ __forceinline__ int64_t skew_at(const LusgsSys<T>& S, int i, int j, int k)
{
    return (int64_t(i + j + 2) * S.sJ + (j + 1)) * S.nz + k;
}

// The skewed copies of each system (blockIdx.y): the iterate zeroed, ghosts
// included; per valid column its neighbour coefficients, rhs, lower
// diagonal -a_zm and the Thomas factors m_k, c'_k.
template <typename T>
__global__ __launch_bounds__(256) void lusgs_skew_layout_kernel(const LusgsBatch<T> B)
{
    const LusgsSys<T>& S = B.s[blockIdx.y];
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    const int64_t t0 = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
    const int64_t nw = int64_t(S.sP) * S.sJ * S.nz;
    for (int64_t t = t0; t < nw; t += stride) { S.sk[kSkW][t] = T(0); }
    const int64_t ncol = int64_t(S.nx) * S.ny;
    for (int64_t t = t0; t < ncol; t += stride) {
        const int i = int(t % S.nx), j = int(t / S.nx);
        const T* a = S.A + i + int64_t(j) * S.a_j;
        const T* pap = S.ap + i + int64_t(j) * S.p_j;
        const T* pr = S.r + i + int64_t(j) * S.r_j;
        int64_t x = skew_at(S, i, j, 0);
        T cprev = T(0);
        for (int k = 0; k < S.nz; ++k, ++x) {
            const T* ak = a + k * S.a_k;
            const T dg = pap[k * S.p_k];
            const T lower = -ak[5 * S.a_n], upper = -ak[6 * S.a_n];
            const T mk = k == 0 ? dg : dg - lower * cprev;
            const T ck = upper / mk;
            S.sk[kSkAxm][x] = ak[1 * S.a_n];
            S.sk[kSkAxp][x] = ak[2 * S.a_n];
            S.sk[kSkAym][x] = ak[3 * S.a_n];
            S.sk[kSkAyp][x] = ak[4 * S.a_n];
            S.sk[kSkR][x] = pr[k * S.r_k];
            S.sk[kSkLz][x] = lower;
            S.sk[kSkM][x] = mk;
            S.sk[kSkC][x] = ck;
            cprev = ck;
        }
    }
}

// NZB > 0: columns of at most NZB levels hold their operands in registers
// (every diagonal fits the workgroup: at most 256 columns); else chunks of
// kLusgsChunk levels. Dynamic LDS: min(nx, ny) * nz reals.
template <typename T, int NZB>
__global__ __launch_bounds__(256) void lusgs_skew_kernel(const LusgsBatch<T> B, int sweeps,
                                                         unsigned long long* __restrict__ changes)
{
    // register path: the diagonal's right-hand sides, -a_zm, m_k, c'_k and
    // (last pass) the old iterate, each [column][level] with an odd column
    // stride cs (coalesced copies in, few bank conflicts out); chunked path:
    // the right-hand sides, then e'_k, [level][column]
    extern __shared__ unsigned char lds_raw[];
    T* Es = reinterpret_cast<T*>(lds_raw);
    const LusgsSys<T>& S = B.s[blockIdx.x];
    const int tid = threadIdx.x, nth = blockDim.x;
    const int nx = S.nx, ny = S.ny, nz = S.nz, J = S.sJ;
    const int ldj = nx < ny ? nx : ny;   // the longest diagonal
    const int cs = nz | 1;
    const int64_t lda = int64_t(ldj) * cs;
    T* Ls = Es + lda;
    T* Ms = Ls + lda;
    T* Cs = Ms + lda;
    T* Os = Cs + lda;
    const int64_t dst = int64_t(J) * nz;   // one diagonal
    T* __restrict__ W = S.sk[kSkW];
    const T* __restrict__ Axm = S.sk[kSkAxm];
    const T* __restrict__ Axp = S.sk[kSkAxp];
    const T* __restrict__ Aym = S.sk[kSkAym];
    const T* __restrict__ Ayp = S.sk[kSkAyp];
    const T* __restrict__ Rr = S.sk[kSkR];
    const T* __restrict__ Lz = S.sk[kSkLz];
    const T* __restrict__ Mm = S.sk[kSkM];
    const T* __restrict__ Cc = S.sk[kSkC];
    const Rsrc rW = rsrc_of(W);
    const int last = nx + ny - 2;
    unsigned long long chg = 0;
    for (int s = 0; s < sweeps; ++s) {
        for (int f = 1; f >= 0; --f) {
            // FillBoundary: the periodic images into the iterate's ghosts
            if (S.Lx) {
                const int64_t n = int64_t(ny) * nz;
                for (int64_t t = tid; t < n; t += nth) {
                    const int j = int(t % ny), k = int(t / ny);
                    W[skew_at(S, -1, j, k)] = W[skew_at(S, S.Lx - 1, j, k)];
                    W[skew_at(S, nx, j, k)] = W[skew_at(S, nx - S.Lx, j, k)];
                }
            }
            if (S.Ly) {
                const int64_t n = int64_t(nx) * nz;
                for (int64_t t = tid; t < n; t += nth) {
                    const int i = int(t % nx), k = int(t / nx);
                    W[skew_at(S, i, -1, k)] = W[skew_at(S, i, S.Ly - 1, k)];
                    W[skew_at(S, i, ny, k)] = W[skew_at(S, i, ny - S.Ly, k)];
                }
            }
            __syncthreads();
            const bool track = (s == sweeps - 1) && f == 0;
            for (int tt = 0; tt <= last; ++tt) {
                const int p = f ? tt : last - tt;
                const int jmin = p - (nx - 1) > 0 ? p - (nx - 1) : 0;
                const int jmax = p < ny - 1 ? p : ny - 1;
                const int ncd = jmax - jmin + 1;
                const int64_t x0 = (int64_t(p + 2) * J + (jmin + 1)) * nz;
                // b. the diagonal's right-hand sides into LDS (and, register
                // path, its recurrence operands), each thread's share loaded
                // before any of it is stored
                const int nq = ncd * nz;
                for (int q0 = 0; q0 < nq; q0 += nth * kLusgsRhsBatch) {
                    T rv[kLusgsRhsBatch], lv[kLusgsRhsBatch], mv[kLusgsRhsBatch], cv2[kLusgsRhsBatch],
                        wv[kLusgsRhsBatch];
#pragma unroll
                    for (int u = 0; u < kLusgsRhsBatch; ++u) {
                        const int qu = q0 + tid + u * nth;
                        const int q = qu < nq ? qu : nq - 1;
                        const int64_t x = x0 + q;   // column q / nz, level q % nz
                        rv[u] = Rr[x] + Axm[x] * W[x - dst] + Axp[x] * W[x + dst] + Aym[x] * W[x - dst - nz] +
                                Ayp[x] * W[x + dst + nz];
                        if (NZB > 0) {
                            lv[u] = Lz[x];
                            mv[u] = Mm[x];
                            cv2[u] = Cc[x];
                            wv[u] = track ? W[x] : T(0);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < kLusgsRhsBatch; ++u) {
                        const int q = q0 + tid + u * nth;
                        if (q < nq) {
                            if (NZB > 0) {
                                const int at = (q / nz) * cs + q % nz;
                                Es[at] = rv[u];
                                Ls[at] = lv[u];
                                Ms[at] = mv[u];
                                Cs[at] = cv2[u];
                                Os[at] = wv[u];
                            } else {
                                Es[(q % nz) * ldj + q / nz] = rv[u];
                            }
                        }
                    }
                }
                __syncthreads();
                // c. the recurrences, one thread per column
                if (NZB > 0) {
                    if (tid < ncd) {
                        constexpr int NB = NZB > 0 ? NZB : 1;
                        const uint32_t vo = uint32_t(x0 + int64_t(tid) * nz) * uint32_t(sizeof(T));
                        const int cb = tid * cs;
                        T rr[NB], ll[NB], dd[NB], cv[NB], ov[NB];
#pragma unroll
                        for (int u = 0; u < NB; ++u) {
                            if (u < nz) {
                                rr[u] = Es[cb + u];
                                ll[u] = Ls[cb + u];
                                dd[u] = Ms[cb + u];
                                cv[u] = Cs[cb + u];
                                ov[u] = Os[cb + u];
                            }
                        }
                        T ep = T(0);
#pragma unroll
                        for (int u = 0; u < NB; ++u) {
                            if (u < nz) {
                                const T ek = u == 0 ? rr[u] / dd[u] : (rr[u] - ll[u] * ep) / dd[u];
                                rr[u] = ek;
                                ep = ek;
                            }
                        }
                        T o = ep;
#pragma unroll
                        for (int u = NB - 1; u >= 0; --u) {
                            if (u < nz) {
                                if (u < nz - 1) { o = rr[u] - cv[u] * o; }
                                if (track) {
                                    const unsigned long long b = dev::abs_bits(double(ov[u] - o));
                                    chg = b > chg ? b : chg;
                                }
                                bstore<T>(o, rW, vo, uint32_t(u) * uint32_t(sizeof(T)));
                            }
                        }
                    }
                } else {
                    for (int m = tid; m < ncd; m += nth) {
                        const int64_t xc = x0 + int64_t(m) * nz;
                        T eprev = T(0);
                        for (int k0 = 0; k0 < nz; k0 += kLusgsChunk) {
                            T rr[kLusgsChunk], l2[kLusgsChunk], d2[kLusgsChunk];
#pragma unroll
                            for (int u = 0; u < kLusgsChunk; ++u) {
                                if (k0 + u < nz) {
                                    const int64_t x = xc + (k0 + u);
                                    rr[u] = Es[(k0 + u) * ldj + m];
                                    l2[u] = Lz[x];
                                    d2[u] = Mm[x];
                                }
                            }
#pragma unroll
                            for (int u = 0; u < kLusgsChunk; ++u) {
                                const int k = k0 + u;
                                if (k < nz) {
                                    const T ek = k == 0 ? rr[u] / d2[u] : (rr[u] - l2[u] * eprev) / d2[u];
                                    Es[k * ldj + m] = ek;
                                    eprev = ek;
                                }
                            }
                        }
                        T o = eprev;
                        {
                            const int64_t x = xc + (nz - 1);
                            if (track) {
                                const unsigned long long b = dev::abs_bits(double(W[x] - o));
                                chg = b > chg ? b : chg;
                            }
                            W[x] = o;
                        }
                        for (int k0 = nz - 2; k0 >= 0; k0 -= kLusgsChunk) {
                            T ev[kLusgsChunk], c2[kLusgsChunk], o2[kLusgsChunk];
#pragma unroll
                            for (int u = 0; u < kLusgsChunk; ++u) {
                                if (k0 - u >= 0) {
                                    const int64_t x = xc + (k0 - u);
                                    ev[u] = Es[(k0 - u) * ldj + m];
                                    c2[u] = Cc[x];
                                    o2[u] = track ? W[x] : T(0);
                                }
                            }
#pragma unroll
                            for (int u = 0; u < kLusgsChunk; ++u) {
                                const int k = k0 - u;
                                if (k >= 0) {
                                    o = ev[u] - c2[u] * o;
                                    if (track) {
                                        const unsigned long long b = dev::abs_bits(double(o2[u] - o));
                                        chg = b > chg ? b : chg;
                                    }
                                    W[xc + k] = o;
                                }
                            }
                        }
                    }
                }
                __syncthreads();
            }
        }
    }
    // max |change| over the block.
    __shared__ unsigned long long red[256 / 64];
    chg = dev::wave_max_u64(chg);
    if ((tid & 63) == 0) { red[tid >> 6] = chg; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long mx = 0;
        for (int w = 0; w < (nth + 63) / 64; ++w) { mx = red[w] > mx ? red[w] : mx; }
        changes[blockIdx.x] = mx;
    }
    // The iterate into delta (whose ghosts were zeroed before the launch and
    // stay zero), then the seam copies (the low copy owns).
    {
        const int64_t n = int64_t(nx) * ny * nz;
        for (int64_t t = tid; t < n; t += nth) {
            const int i = int(t % nx);
            const int64_t rr = t / nx;
            const int j = int(rr % ny), k = int(rr / ny);
            S.d[i + j * S.d_j + k * S.d_k] = W[skew_at(S, i, j, k)];
        }
    }
    __syncthreads();
    for (int d = 0; d < 2; ++d) {
        if (!((S.loc >> d) & 1) || !(d == 0 ? S.Lx : S.Ly)) { continue; }
        const int64_t n = d == 0 ? int64_t(ny) * nz : int64_t(nx) * nz;
        for (int64_t t = tid; t < n; t += nth) {
            int64_t lo, hi;
            if (d == 0) {
                const int j = int(t % ny), k = int(t / ny);
                lo = j * S.d_j + k * S.d_k;
                hi = lo + (nx - 1);
            } else {
                const int i = int(t % nx), k = int(t / nx);
                lo = i + k * S.d_k;
                hi = lo + (ny - 1) * S.d_j;
            }
            S.d[hi] = S.d[lo];
        }
    }
}

int check_line_options_impl(const gust_line_options* opt)
{
    if (!opt) { GUST_FAIL("gust: zline_lusgs: options are null"); }
    if (opt->overlap < 0 || opt->overlap > GUST_MAX_LINE_OVERLAP) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: must be an integer >= 0 and <= %d",
                  opt->overlap, GUST_MAX_LINE_OVERLAP);
    }
    if (opt->overlap > 0 && opt->ghosts != GUST_GHOSTS_ZERO) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: needs steady.line_ghosts = zero "
                  "(the lagged ghosts sweep the valid box alone)", opt->overlap);
    }
    if (opt->alternate != 0 && opt->alternate != 1) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_alternate = %d: must be 0 or 1", opt->alternate);
    }
    if (opt->ghosts != GUST_GHOSTS_LAGGED && opt->ghosts != GUST_GHOSTS_ZERO) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_ghosts = %d: must be lagged or zero", opt->ghosts);
    }
    if (opt->order != GUST_ORDER_SEQUENTIAL && opt->order != GUST_ORDER_DIAGONAL) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_order = %d: must be sequential or diagonal",
                  opt->order);
    }
    return 0;
}

// One system of a batch: checked, its views at the valid lo corner, its
// sweep and owned regions; *scratch_bytes: its share of the workspace (c',
// e', m over the sweep region and, zero ghosts, the work array).
template <typename T>
LusgsSys<T> lusgs_system(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                         const gust_geometry* g, gust_array4* delta, const gust_line_options* opt,
                         size_t* scratch_bytes, IBox& V)
{
    const char* op = "zline_lusgs";
    check_line_system(op, loc, A, ap, rhs, sweeps, g, delta, V);
    check_line_options_impl(opt);
    // check_line_overlap_layout (Source/LineSweeps.cpp:343-359), one box.
    const int ngx = (g->hi[0] - g->lo[0] + 1) > 1 ? opt->overlap : 0;
    const int ngy = (g->hi[1] - g->lo[1] + 1) > 1 ? opt->overlap : 0;
    int narrowest = 0x7fffffff;
    if (ngx > 0) { narrowest = std::min(narrowest, g->hi[0] - g->lo[0] + 1); }
    if (ngy > 0) { narrowest = std::min(narrowest, g->hi[1] - g->lo[1] + 1); }
    if (opt->overlap > 0 && narrowest != 0x7fffffff && opt->overlap > narrowest) {
        GUST_FAIL("gust: zline_lusgs: invalid options:\n  steady.line_overlap = %d: must be <= the narrowest box's "
                  "width in x and y (%d cells; mesh.max_grid_size)", opt->overlap, narrowest);
    }
    LusgsSys<T> t{};
    t.A = at_lo<T>(*A, V);
    t.a_j = A->jstride; t.a_k = A->kstride; t.a_n = A->nstride;
    t.ap = at_lo<T>(*ap, V);
    t.p_j = ap->jstride; t.p_k = ap->kstride;
    t.r = at_lo<T>(*rhs, V);
    t.r_j = rhs->jstride; t.r_k = rhs->kstride;
    t.d = at_lo<T>(*delta, V);
    t.d_j = delta->jstride; t.d_k = delta->kstride;
    t.nx = V.len(0); t.ny = V.len(1); t.nz = V.len(2);
    t.Lx = g->periodic[0] ? (g->hi[0] - g->lo[0] + 1) : 0;
    t.Ly = g->periodic[1] ? (g->hi[1] - g->lo[1] + 1) : 0;
    t.loc = loc;
    const bool zero = opt->ghosts == GUST_GHOSTS_ZERO;
    t.zero_mode = zero ? 1 : 0;
    // owned (line_owned_box) and the sweep region (line_sweep_box), local.
    int olo[2] = {0, 0}, ohi[2] = {t.nx - 1, t.ny - 1};
    if (zero) {
        for (int d = 0; d < 2; ++d) {
            if (((loc >> d) & 1) && g->periodic[d]) { olo[d] += 1; }
        }
    }
    int slo[2] = {olo[0], olo[1]}, shi[2] = {ohi[0], ohi[1]};
    if (zero) {
        const int ng[2] = {ngx, ngy};
        const int n[2] = {t.nx, t.ny};
        for (int d = 0; d < 2; ++d) {
            slo[d] -= ng[d];
            shi[d] += ng[d];
            if (!g->periodic[d]) {
                slo[d] = std::max(slo[d], 0);
                shi[d] = std::min(shi[d], n[d] - 1);
            }
        }
        t.fold[0] = g->periodic[0] && (g->hi[0] - g->lo[0] + 1) == 1;
        t.fold[1] = g->periodic[1] && (g->hi[1] - g->lo[1] + 1) == 1;
    }
    for (int d = 0; d < 2; ++d) {
        t.s_lo[d] = slo[d]; t.s_hi[d] = shi[d];
        t.o_lo[d] = olo[d]; t.o_hi[d] = ohi[d];
    }
    const int sx = shi[0] - slo[0] + 1, sy = shi[1] - slo[1] + 1;
    t.c_j = sx;
    t.c_k = int64_t(sx) * sy;
    const size_t cbytes = (size_t(sx) * sy * t.nz * sizeof(T) + 255) / 256 * 256;
    size_t wbytes = 0;
    if (zero) {
        t.w_gx = ngx + 1;
        t.w_gy = ngy + 1;
        const int64_t tx = t.nx + 2 * t.w_gx, ty = t.ny + 2 * t.w_gy;
        t.w_j = tx;
        t.w_k = tx * ty;
        wbytes = (size_t(tx * ty * t.nz) * sizeof(T) + 255) / 256 * 256;
    }
    *scratch_bytes = 3 * cbytes + wbytes;
    return t;
}

// Points a system's scratch pointers into ws (its share, as lusgs_system
// sized it).
template <typename T>
void lusgs_place(LusgsSys<T>& t, char* ws)
{
    const int sx = int(t.c_j), sy = int(t.c_k / t.c_j);
    const size_t cbytes = (size_t(sx) * sy * t.nz * sizeof(T) + 255) / 256 * 256;
    t.c = reinterpret_cast<T*>(ws);
    t.e = reinterpret_cast<T*>(ws + cbytes);
    t.m = reinterpret_cast<T*>(ws + 2 * cbytes);
    if (t.zero_mode) { t.w = reinterpret_cast<T*>(ws + 3 * cbytes) + t.w_gx + t.w_gy * t.w_j; }
}

// LU-SGS sweeps of a batch of systems (one workgroup each, one launch);
// each system's last-sweep change as the ordered bits of |change| into
// changes[s]. The systems go to the kernels by value: nothing waits.
template <typename T>
void lusgs_batch(int nsys, const gust_line_system* sys, const gust_line_options* opt, int sweeps,
                 unsigned long long* changes, hipStream_t st)
{
    if (nsys < 1 || nsys > kLusgsMaxBatch) {
        GUST_FAIL("gust: zline_lusgs: a batch holds 1 to %d systems (got %d)", kLusgsMaxBatch, nsys);
    }
    LusgsBatch<T> B{};
    size_t bytes[kLusgsMaxBatch] = {}, total = 0;
    IBox V[kLusgsMaxBatch];
    gust_array4 dv[kLusgsMaxBatch];
    int64_t most = 1;
    for (int s = 0; s < nsys; ++s) {
        dv[s] = sys[s].delta;
        B.s[s] = lusgs_system<T>(sys[s].loc, &sys[s].A, &sys[s].ap, &sys[s].rhs, sweeps, &sys[s].geom, &dv[s], opt,
                                 &bytes[s], V[s]);
        total += bytes[s];
        most = std::max<int64_t>(most, B.s[s].c_k);
    }
    // lagged ghosts without alternation: the skewed copies
    bool skew = opt->ghosts == GUST_GHOSTS_LAGGED && opt->alternate == 0;
    for (int s = 0; s < nsys && skew; ++s) {
        const LusgsSys<T>& t = B.s[s];
        // the buffer offsets are 32-bit; the diagonal's right-hand sides sit
        // in LDS
        skew = double(t.nx + t.ny + 3) * double(t.ny + 2) * double(t.nz) * sizeof(T) < 2.0e9 &&
               size_t(std::min(t.nx, t.ny)) * t.nz * sizeof(T) <= kLusgsMaxLds;
    }
    if (skew) {
        total = 0;
        for (int s = 0; s < nsys; ++s) {
            LusgsSys<T>& t = B.s[s];
            t.sP = t.nx + t.ny + 3;
            t.sJ = t.ny + 2;
            bytes[s] = kSkN * ((size_t(t.sP) * t.sJ * t.nz * sizeof(T) + 255) / 256 * 256);
            total += bytes[s];
            most = std::max<int64_t>(most, int64_t(t.sP) * t.sJ * t.nz);
        }
    }
    char* ws = static_cast<char*>(Workspace::get(st, 3, total + 1024));
    for (int s = 0; s < nsys; ++s) {
        if (skew) {
            LusgsSys<T>& t = B.s[s];
            const size_t one = bytes[s] / kSkN;
            for (int a = 0; a < kSkN; ++a) { t.sk[a] = reinterpret_cast<T*>(ws + a * one); }
        } else {
            lusgs_place(B.s[s], ws);
        }
        ws += bytes[s];
        zero_ghosts<T>(dv[s], V[s], st);
    }
    if (skew) {
        int nzmax = 0, diag = 0;
        size_t lds = 0;
        size_t lds_regs = 0;
        for (int s = 0; s < nsys; ++s) {
            const LusgsSys<T>& t = B.s[s];
            nzmax = std::max(nzmax, t.nz);
            diag = std::max(diag, std::min(t.nx, t.ny));
            lds = std::max(lds, size_t(std::min(t.nx, t.ny)) * t.nz * sizeof(T));
            lds_regs = std::max(lds_regs, 5 * size_t(std::min(t.nx, t.ny)) * (t.nz | 1) * sizeof(T));
        }
        const bool regs = diag <= 256 && nzmax <= 32 && lds_regs <= kLusgsMaxLds;
        if (regs) { lds = lds_regs; }
        if (lds > 64 * 1024) {
            // above the default dynamic LDS limit (gfx950 has 160 KiB per CU)
            for (const void* fn : {reinterpret_cast<const void*>(&lusgs_skew_kernel<T, 16>),
                                   reinterpret_cast<const void*>(&lusgs_skew_kernel<T, 32>),
                                   reinterpret_cast<const void*>(&lusgs_skew_kernel<T, 0>)}) {
                GUST_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(kLusgsMaxLds)));
            }
        }
        const unsigned lx = unsigned(std::min<int64_t>((most + 255) / 256, 2048));
        lusgs_skew_layout_kernel<T><<<dim3(lx, unsigned(nsys)), 256, 0, st>>>(B);
        GUST_HIP(hipGetLastError());
        if (regs && nzmax <= 16) {
            lusgs_skew_kernel<T, 16><<<unsigned(nsys), 256, lds, st>>>(B, sweeps, changes);
        } else if (regs) {
            lusgs_skew_kernel<T, 32><<<unsigned(nsys), 256, lds, st>>>(B, sweeps, changes);
        } else {
            lusgs_skew_kernel<T, 0><<<unsigned(nsys), 256, lds, st>>>(B, sweeps, changes);
        }
        GUST_HIP(hipGetLastError());
        return;
    }
    const unsigned gx = unsigned(std::min<int64_t>((most + 255) / 256, 1024));
    lusgs_factor_kernel<T><<<dim3(gx, unsigned(nsys)), 256, 0, st>>>(B);
    GUST_HIP(hipGetLastError());
    lusgs_kernel<T><<<unsigned(nsys), 256, 0, st>>>(B, sweeps, opt->alternate, changes);
    GUST_HIP(hipGetLastError());
}

// One system: its change recorded in the stream's deferred batch when one is
// open (read with its other values), else read back here.
template <typename T>
void lusgs_run(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
               const gust_geometry* g, gust_array4* delta, const gust_line_options* opt, double* change,
               hipStream_t st)
{
    {
        // every refusal before the first device call
        size_t bytes = 0;
        IBox V;
        gust_array4 dv = *delta;
        (void)lusgs_system<T>(loc, A, ap, rhs, sweeps, g, &dv, opt, &bytes, V);
    }
    gust_line_system sy1{loc, *g, *A, *ap, *rhs, *delta};
    Deferred* dq = change && Deferred::current() && Deferred::current()->stream() == st ? Deferred::current()
                                                                                          : nullptr;
    if (dq) {
        unsigned long long* slot = dq->zero_slot([change](unsigned long long b) { *change = bits_to_double(b); });
        lusgs_batch<T>(1, &sy1, opt, sweeps, slot, st);
        return;
    }
    auto* chg = static_cast<unsigned long long*>(Workspace::get(st, 1, 16));
    lusgs_batch<T>(1, &sy1, opt, sweeps, chg, st);
    if (change) {
        auto* h = static_cast<unsigned long long*>(Workspace::get_host_pinned(0, 8));
        GUST_HIP(hipMemcpyAsync(h, chg, 8, hipMemcpyDeviceToHost, st));
        GUST_HIP(hipStreamSynchronize(st));
        *change = bits_to_double(h[0]);
    }
}

}   // namespace

// A batch of FP64 LU-SGS systems (up to three) in one launch, their
// last-sweep changes into `changes` as ordered bits of |change|.
void zline_lusgs_into(int nsys, const gust_line_system* sys, const gust_line_options* opt, int sweeps,
                      unsigned long long* changes, hipStream_t st)
{
    lusgs_batch<double>(nsys, sys, opt, sweeps, changes, st);
}

// The same on FP32 systems (the views' pointers are float data).
void zline_lusgs_into_f32(int nsys, const gust_line_system* sys, const gust_line_options* opt, int sweeps,
                          unsigned long long* changes, hipStream_t st)
{
    lusgs_batch<float>(nsys, sys, opt, sweeps, changes, st);
}

// A batch of z-line Jacobi systems whose last-sweep changes go straight
// into `changes` (nsys slots that hold zeros, as ordered bits of |change|).
void zline_jacobi_into(int nsys, const gust_line_system* sys, int sweeps, unsigned long long* changes, hipStream_t st)
{
    JacobiPlan<double> plan;
    plan.init(nsys, sys, false, st);
    plan.chg = changes;
    plan.chg_zeroed = true;
    plan.run(sweeps, nullptr, st);
}

// ---------------------------------------------------------------------------
// The entry point that allocates nothing: its scratch (the last-sweep
// change, the system table and the iterate) lives in this fixed array, one
// copy per device, held until the process ends. Calls that overlap on one
// device share it and race; each call clears what must start at zero.

constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once
// per device with that device current (in host code the array's name is not
// a device address).
char* noalloc_scratch()
{
    int dev = 0;
    GUST_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) { GUST_FAIL("gust: zline_jacobi_noalloc: device %d out of range", dev); }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            GUST_FAIL("gust: zline_jacobi_noalloc: hipGetSymbolAddress of the scratch array failed (%s)",
                      hipGetErrorString(e));
        }
        g_scratch_ptr[dev] = static_cast<char*>(p);
    }
    return g_scratch_ptr[dev];
}

void zjacobi_noalloc(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                     const gust_geometry* geom, gust_array4* delta, double* change, hipStream_t st)
{
    if (!A || !ap || !rhs || !delta || !geom) { GUST_FAIL("gust: zline_jacobi_noalloc: null argument"); }
    if (sweeps < 1) { GUST_FAIL("gust: zline_jacobi: sweeps must be >= 1 (got %d)", sweeps); }
    gust_line_system q{};
    q.loc = loc;
    q.geom = *geom;
    q.A = *A;
    q.ap = *ap;
    q.rhs = *rhs;
    q.delta = *delta;
    char* base = noalloc_scratch();
    JacobiPlan<double> plan;
    plan.init(1, &q, false, st, base, kNoAllocScratchBytes);
    plan.run(sweeps, nullptr, st);
    if (change) {
        unsigned long long bits = 0;
        GUST_HIP(hipMemcpyAsync(&bits, plan.chg, 8, hipMemcpyDeviceToHost, st));
        GUST_HIP(hipStreamSynchronize(st));
        *change = bits_to_double(bits);
    } else {
        GUST_HIP(hipStreamSynchronize(st));
    }
}

}   // namespace

void zline_jacobi_into_f32(int nsys, const gust_line_system* sys, int sweeps, unsigned long long* changes,
                           hipStream_t st)
{
    JacobiPlan<float> plan;
    plan.init(nsys, sys, false, st);
    plan.chg = changes;
    plan.chg_zeroed = true;
    plan.run(sweeps, nullptr, st);
}

namespace {

template <typename T>
void zjacobi_one(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                 const gust_geometry* geom, gust_array4* delta, double* change, hipStream_t st)
{
    if (!A || !ap || !rhs || !delta || !geom) { GUST_FAIL("gust: zline_jacobi: null argument"); }
    gust_line_system q{};
    q.loc = loc;
    q.geom = *geom;
    q.A = *A;
    q.ap = *ap;
    q.rhs = *rhs;
    q.delta = *delta;
    zjacobi_run<T>(1, &q, sweeps, nullptr, change, st);
}

}   // namespace

}   // namespace gust_amd

using namespace gust_amd;

extern "C" {

int gust_zline_jacobi(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                      const gust_geometry* geom, gust_array4* delta, double* change, void* stream)
{
    return guarded([&] { zjacobi_one<double>(loc, A, ap, rhs, sweeps, geom, delta, change, (hipStream_t)stream); });
}

int gust_zline_jacobi_noalloc(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs,
                              int sweeps, const gust_geometry* geom, gust_array4* delta, double* change, void* stream)
{
    return guarded([&] { zjacobi_noalloc(loc, A, ap, rhs, sweeps, geom, delta, change, (hipStream_t)stream); });
}

int gust_zline_jacobi_f32(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                          const gust_geometry* geom, gust_array4* delta, double* change, void* stream)
{
    return guarded([&] { zjacobi_one<float>(loc, A, ap, rhs, sweeps, geom, delta, change, (hipStream_t)stream); });
}

int gust_zline_jacobi_batched_async(int nsys, const gust_line_system* sys, int sweeps, double* changes_dev,
                                    void* stream)
{
    return guarded([&] { zjacobi_run<double>(nsys, sys, sweeps, changes_dev, nullptr, (hipStream_t)stream); });
}

int gust_zline_jacobi_batched_async_f32(int nsys, const gust_line_system* sys, int sweeps, double* changes_dev,
                                        void* stream)
{
    return guarded([&] { zjacobi_run<float>(nsys, sys, sweeps, changes_dev, nullptr, (hipStream_t)stream); });
}

struct gust_jacobi_plan {
    gust_amd::JacobiPlanBase* impl;
};

int gust_jacobi_plan_create(int nsys, const gust_line_system* sys, int dtype, gust_jacobi_plan** plan, void* stream)
{
    return guarded([&] {
        if (!plan) { GUST_FAIL("gust: jacobi_plan_create: plan is null"); }
        *plan = nullptr;
        std::unique_ptr<JacobiPlanBase> p(make_plan(dtype));
        if (dtype == 0) { static_cast<JacobiPlan<double>*>(p.get())->init(nsys, sys, true, (hipStream_t)stream); }
        else { static_cast<JacobiPlan<float>*>(p.get())->init(nsys, sys, true, (hipStream_t)stream); }
        *plan = new gust_jacobi_plan{p.release()};
    });
}

int gust_jacobi_plan_run(gust_jacobi_plan* plan, int sweeps, double* changes_dev, void* stream)
{
    return guarded([&] {
        if (!plan || !plan->impl) { GUST_FAIL("gust: jacobi_plan_run: plan is null"); }
        plan->impl->run(sweeps, changes_dev, (hipStream_t)stream);
    });
}

int gust_jacobi_plan_sweep(gust_jacobi_plan* plan, int first, int last, int into_scratch, void* stream)
{
    return guarded([&] {
        if (!plan || !plan->impl) { GUST_FAIL("gust: jacobi_plan_sweep: plan is null"); }
        plan->impl->sweep(first != 0, last != 0, into_scratch ? 1 : 0, (hipStream_t)stream);
    });
}

int gust_jacobi_set_schedule(int32_t schedule)
{
    return guarded([&] {
        if (schedule < 0 || schedule > 4) {
            GUST_FAIL("gust: jacobi_set_schedule: schedule must be 0 (autotuned), 1 (streaming), 2 (staged), 3 "
                      "(the fixed rule) or 4 (register)");
        }
        g_jacobi_schedule.store(schedule, std::memory_order_relaxed);
    });
}

int gust_jacobi_tuned(int32_t index, int32_t* count, int32_t* dtype, int32_t* sweeps, int32_t* nsys,
                      int32_t* choice, float* ms, char* name, int64_t cap)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        if (count) { *count = int32_t(g_tuned.size()); }
        if (index < 0 || index >= int32_t(g_tuned.size())) { return; }
        auto it = g_tuned.begin();
        std::advance(it, index);
        if (dtype) { *dtype = it->first.dtype; }
        if (sweeps) { *sweeps = it->first.sweeps; }
        if (nsys) { *nsys = int32_t(it->first.shape.size() / 3); }
        if (choice) { *choice = it->second.choice; }
        if (ms) {
            for (int c = 0; c < kNumChoices; ++c) { ms[c] = it->second.ms[c]; }
        }
        if (name && cap > 0) { std::snprintf(name, size_t(cap), "%s", kChoiceNames[it->second.choice]); }
    });
}

int gust_jacobi_tune_reset(void)
{
    return guarded([&] {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        g_tuned.clear();
    });
}

int gust_jacobi_plan_info(const gust_jacobi_plan* plan, int64_t* unknowns, int32_t* tiles)
{
    return guarded([&] {
        if (!plan || !plan->impl) { GUST_FAIL("gust: jacobi_plan_info: plan is null"); }
        if (unknowns) { *unknowns = plan->impl->unknowns; }
        if (tiles) { *tiles = plan->impl->tiles; }
    });
}

void gust_jacobi_plan_destroy(gust_jacobi_plan* plan)
{
    if (plan) {
        delete plan->impl;
        delete plan;
    }
}

int gust_zline_lusgs(int loc, const gust_array4* A, const gust_array4* ap, const gust_array4* rhs, int sweeps,
                     const gust_geometry* geom, gust_array4* delta, const gust_line_options* opt, double* change,
                     void* stream)
{
    return guarded([&] {
        if (!A || !ap || !rhs || !delta || !geom || !opt) { GUST_FAIL("gust: zline_lusgs: null argument"); }
        lusgs_run<double>(loc, A, ap, rhs, sweeps, geom, delta, opt, change, (hipStream_t)stream);
    });
}

int gust_line_sweeps(const gust_line_options* opt, int loc, const gust_array4* A, const gust_array4* ap,
                     const gust_array4* rhs, int sweeps, const gust_geometry* geom, gust_array4* delta,
                     double* change, void* stream)
{
    return guarded([&] {
        if (!opt) { GUST_FAIL("gust: line_sweeps: options are null"); }
        if (opt->kind == GUST_SWEEP_LUSGS) {
            if (!A || !ap || !rhs || !delta || !geom) { GUST_FAIL("gust: zline_lusgs: null argument"); }
            lusgs_run<double>(loc, A, ap, rhs, sweeps, geom, delta, opt, change, (hipStream_t)stream);
        } else {
            zjacobi_one<double>(loc, A, ap, rhs, sweeps, geom, delta, change, (hipStream_t)stream);
        }
    });
}

int gust_check_line_options(const gust_line_options* opt)
{
    return guarded([&] { check_line_options_impl(opt); });
}

}
