// simplec.hpp -- steps of simplec.hip used by the SIMPLEC stage itself.
#pragma once

#include "common.hpp"

namespace gust_amd {

// theta's and k's scalar sweeps of one iteration with their z-line Jacobi
// solves sharing one launch per sweep (the bits of two scalar_sweep calls:
// theta without a clip, then k with the floor clip). Throws on failure.
void scalar_sweeps_theta_tke(const gust_geometry& g, gust_array4* theta, const gust_array4* R_theta,
                             const gust_array4* A_theta, gust_array4* tke, const gust_array4* R_tke,
                             const gust_array4* A_tke, const gust_array4* dtau, int sweeps, double floor,
                             const gust_line_options& sweep, gust_scalar_sweep_result* res_theta,
                             gust_scalar_sweep_result* res_tke, hipStream_t st, int sweep_precision = 0);

// FP64 z-line Jacobi sweeps of a batch of systems (one launch per sweep),
// each system's last-sweep change written as the ordered bits of |change|
// into changes[s] (nsys slots holding zeros); defined in line_sweeps.hip.
void zline_jacobi_into(int nsys, const gust_line_system* sys, int sweeps, unsigned long long* changes,
                       hipStream_t st);
// The same on FP32 systems (the views' pointers are float data).
void zline_jacobi_into_f32(int nsys, const gust_line_system* sys, int sweeps, unsigned long long* changes,
                           hipStream_t st);
// FP64 line LU-SGS of up to three systems in one launch (one workgroup
// each), with the options' ghosts, overlap and alternation; the changes as
// zline_jacobi_into's.
void zline_lusgs_into(int nsys, const gust_line_system* sys, const gust_line_options* opt, int sweeps,
                      unsigned long long* changes, hipStream_t st);
// The same on FP32 systems.
void zline_lusgs_into_f32(int nsys, const gust_line_system* sys, const gust_line_options* opt, int sweeps,
                          unsigned long long* changes, hipStream_t st);

}   // namespace gust_amd
