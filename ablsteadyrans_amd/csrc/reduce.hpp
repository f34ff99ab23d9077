// reduce.hpp -- the fixed-tree reductions of reductions.hip for the other
// sources of the library.
#pragma once

#include "common.hpp"

namespace gust_amd {

// The sum of squares of a over the box b (a fixed two-level tree, the same
// bits on every call); synchronises the stream. MultiFab::norm2 is its root.
double sum_squares(hipStream_t st, const gust_array4* a, const IBox& b);

}   // namespace gust_amd
