// errors.hpp -- the C-ABI's error contract, shared by device and host-only
// sources: a failing entry point stores the reference's "gust: <op>: ..."
// message per thread and returns nonzero; nothing throws across the ABI.
#pragma once

#include <exception>

namespace gust_amd {

// ---------------------------------------------------------------------------
// Errors: the C-ABI never throws. A failing entry point stores the message
// (the reference's "gust: <op>: ..." text) per thread and returns nonzero.

void set_error(const char* fmt, ...);
void clear_error();

struct Fail {
    int code;
};

#define GUST_FAIL(...)                        \
    do {                                      \
        ::gust_amd::set_error(__VA_ARGS__);   \
        throw ::gust_amd::Fail{1};            \
    } while (0)

// Runs f; maps a Fail (or any C++ exception) to a nonzero status.
template <class F>
int guarded(F&& f) noexcept
{
    try {
        clear_error();
        f();
        return 0;
    } catch (const Fail& e) {
        return e.code;
    } catch (const std::exception& e) {
        set_error("gust: internal error: %s", e.what());
        return 3;
    } catch (...) {
        set_error("gust: internal error");
        return 3;
    }
}

}   // namespace gust_amd
