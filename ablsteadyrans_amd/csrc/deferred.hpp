// deferred.hpp -- device-side checks and reported values of a sequence of
// operations, read back with one copy and one synchronisation.
//
// Each request takes a 64-bit slot in device memory that kernels update
// atomically (a maximum or a count starts at 0, a minimum at all ones) and a
// host callback. resolve() copies the slots to pinned host memory, waits for
// the stream once and runs the callbacks in request order; the first failing
// check raises its error there. An entry point that finds a batch current on
// its thread adds its requests to it (a SIMPLEC iteration resolves once);
// otherwise it opens its own and resolves before it returns, so its results
// and errors are the same as when every check was read at once.
#pragma once

#include "common.hpp"

#include <cstring>
#include <functional>
#include <utility>
#include <vector>

namespace gust_amd {

class Deferred {
public:
    static constexpr int kZero = 512, kOnes = 512;
    using Callback = std::function<void(unsigned long long)>;

    explicit Deferred(hipStream_t st) : st_(st) {}
    Deferred(const Deferred&) = delete;
    Deferred& operator=(const Deferred&) = delete;

    hipStream_t stream() const { return st_; }

    // A slot that starts at 0 (atomicMax of ordered bits, atomicAdd, atomicOr).
    unsigned long long* zero_slot(Callback f)
    {
        init();
        if (nz_ >= kZero) { overflow(); }
        const int s = nz_++;
        cbs_.emplace_back(s, std::move(f));
        return dev_ + s;
    }
    // n consecutive slots that start at 0, callback i for slot i (for a
    // kernel that writes an array of values).
    unsigned long long* zero_run(std::vector<Callback> fs)
    {
        const int n = int(fs.size());
        init();
        if (nz_ + n > kZero) { overflow(); }
        const int s0 = nz_;
        nz_ += n;
        for (int i = 0; i < n; ++i) { cbs_.emplace_back(s0 + i, std::move(fs[size_t(i)])); }
        return dev_ + s0;
    }
    // A slot that starts at all ones (atomicMin).
    unsigned long long* ones_slot(Callback f)
    {
        init();
        if (no_ >= kOnes) { overflow(); }
        const int s = kZero + no_++;
        cbs_.emplace_back(s, std::move(f));
        return dev_ + s;
    }
    // A host callback with no slot, run in order at resolve (a check of
    // values other callbacks have stored).
    void then(std::function<void()> f)
    {
        cbs_.emplace_back(-1, [g = std::move(f)](unsigned long long) { g(); });
    }
    bool pending() const { return !cbs_.empty(); }

    // One copy, one synchronisation, the callbacks in order.
    void resolve()
    {
        if (cbs_.empty()) { return; }
        if (recording_) {
            GUST_FAIL("gust: internal: a host read inside a captured iteration (use the device-resident forms)");
        }
        flush();
    }

    // A batch recorded into a captured graph: the graph holds the slots'
    // initialisation, the kernels that write them and the copy to pinned
    // host memory; after each launch of the graph (and a wait) replay()
    // runs the callbacks on the copied values. Every callback must refer
    // only to objects that outlive the graph.
    struct Recorded {
        std::vector<std::pair<int, Callback>> cbs;
        const unsigned long long* host = nullptr;
        void replay() const
        {
            for (const auto& c : cbs) { c.second(c.first >= 0 && host ? host[c.first] : 0ull); }
        }
    };
    void begin_recording() { recording_ = true; }
    // Enqueues (records) the slots' copy to the host and hands the callbacks
    // over; the batch is empty afterwards.
    Recorded end_recording()
    {
        Recorded r;
        if (inited_) {
            GUST_HIP(hipMemcpyAsync(host_, dev_, sizeof(unsigned long long) * (kZero + kOnes),
                                    hipMemcpyDeviceToHost, st_));
            r.host = host_;
        }
        r.cbs.swap(cbs_);
        inited_ = false;
        nz_ = no_ = 0;
        recording_ = false;
        return r;
    }

    static Deferred*& current()
    {
        static thread_local Deferred* cur = nullptr;
        return cur;
    }

    // Makes a batch the thread's current one for its lifetime.
    class Scope {
    public:
        explicit Scope(Deferred& d) : prev_(current()) { current() = &d; }
        ~Scope() { current() = prev_; }
        Scope(const Scope&) = delete;
        Scope& operator=(const Scope&) = delete;

    private:
        Deferred* prev_;
    };

private:
    void init();
    // A full batch is an error, not an early flush: queued device work (a
    // Reducer's pending jobs) may still target slots of this batch, and a
    // flush would hand their callbacks the initial values.
    [[noreturn]] void overflow() const
    {
        GUST_FAIL("gust: internal: more than %d deferred device checks in one batch (%d at 0, %d at all ones)",
                  kZero, nz_, no_);
    }
    void flush()
    {
        if (!inited_) {
            std::vector<std::pair<int, Callback>> cbs;
            cbs.swap(cbs_);
            for (auto& c : cbs) { c.second(0); }
            return;
        }
        GUST_HIP(hipMemcpyAsync(host_, dev_, sizeof(unsigned long long) * (kZero + kOnes), hipMemcpyDeviceToHost, st_));
        GUST_HIP(hipStreamSynchronize(st_));
        std::vector<std::pair<int, Callback>> cbs;
        cbs.swap(cbs_);
        std::vector<unsigned long long> vals(host_, host_ + kZero + kOnes);
        inited_ = false;
        nz_ = no_ = 0;
        for (auto& c : cbs) { c.second(c.first >= 0 ? vals[size_t(c.first)] : 0ull); }
    }

    hipStream_t st_;
    unsigned long long* dev_ = nullptr;
    unsigned long long* host_ = nullptr;
    int nz_ = 0, no_ = 0;
    bool inited_ = false, recording_ = false;
    std::vector<std::pair<int, Callback>> cbs_;
};

// The batch an entry point reports into: the thread's current one when it
// is on the same stream, or its own (resolved by finish()).
class DeferredUse {
public:
    explicit DeferredUse(hipStream_t st)
        : own_(st), cur_(Deferred::current() && Deferred::current()->stream() == st ? Deferred::current() : &own_),
          scope_(*cur_) {}
    Deferred& operator*() { return *cur_; }
    Deferred* operator->() { return cur_; }
    // Resolves the batch if this entry point opened it.
    void finish()
    {
        if (cur_ == &own_) { own_.resolve(); }
    }

private:
    Deferred own_;
    Deferred* cur_;
    Deferred::Scope scope_;
};

// Ordered bits of |x| <-> x (non-negative doubles order as their bits).
inline double bits_to_double(unsigned long long b)
{
    double d;
    std::memcpy(&d, &b, sizeof d);
    return d;
}

}   // namespace gust_amd
