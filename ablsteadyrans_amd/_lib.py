"""ctypes binding of ``libgust_amd.so`` (the C-ABI in ``include/gust_amd.h``).

The library is required: importing this module raises if it is missing, and
there is no fallback path of any kind.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgust_amd.so")


class Array4(C.Structure):
    """``gust_array4``: amrex::Array4 (i fastest, begin inclusive, end exclusive)."""

    _fields_ = [
        ("p", C.c_void_p),
        ("jstride", C.c_int64),
        ("kstride", C.c_int64),
        ("nstride", C.c_int64),
        ("begin", C.c_int32 * 3),
        ("end", C.c_int32 * 3),
        ("ncomp", C.c_int32),
    ]


class GeometryC(C.Structure):
    """``gust_geometry``: cell-centred domain box, periodicity, spacing."""

    _fields_ = [
        ("lo", C.c_int32 * 3),
        ("hi", C.c_int32 * 3),
        ("periodic", C.c_int32 * 3),
        ("dx", C.c_double * 3),
    ]


class LineOptionsC(C.Structure):
    """``gust_line_options``."""

    _fields_ = [
        ("kind", C.c_int32),
        ("ghosts", C.c_int32),
        ("overlap", C.c_int32),
        ("order", C.c_int32),
        ("alternate", C.c_int32),
    ]


class LineSystemC(C.Structure):
    """``gust_line_system``."""

    _fields_ = [
        ("loc", C.c_int32),
        ("geom", GeometryC),
        ("A", Array4),
        ("ap", Array4),
        ("rhs", Array4),
        ("delta", Array4),
    ]


class PoissonOptionsC(C.Structure):
    """``gust_poisson_options``."""

    _fields_ = [
        ("reltol", C.c_double),
        ("abstol", C.c_double),
        ("max_iter", C.c_int32),
        ("restart", C.c_int32),
        ("verbose", C.c_int32),
        ("precision", C.c_int32),
        ("pre_sweeps", C.c_int32),
        ("post_sweeps", C.c_int32),
        ("coarse_sweeps", C.c_int32),
        ("max_levels", C.c_int32),
        ("fixed_steps", C.c_int32),
        ("seam_copy", C.c_int32),
    ]


class FaceCoeffsC(C.Structure):
    """``gust_face_coeffs``."""

    _fields_ = [
        ("x", C.POINTER(Array4)),
        ("y", C.POINTER(Array4)),
        ("z", C.POINTER(Array4)),
        ("z_cross", C.POINTER(Array4)),
        ("precond_ratio", C.c_double),
    ]


class PoissonResultC(C.Structure):
    """``gust_poisson_result``."""

    _fields_ = [
        ("iterations", C.c_int32),
        ("initial_residual", C.c_double),
        ("final_residual", C.c_double),
        ("singular", C.c_int32),
        ("removed_mean", C.c_double),
    ]


class PressureCorrectionResultC(C.Structure):
    """``gust_pressure_correction_result``."""

    _fields_ = [
        ("poisson", PoissonResultC),
        ("divergence_before", C.c_double),
        ("divergence_after", C.c_double),
        ("max_dp", C.c_double),
        ("max_du", C.c_double),
        ("max_dv", C.c_double),
        ("max_dw", C.c_double),
    ]


class FacesC(C.Structure):
    """``gust_faces``: one view per velocity component."""

    _fields_ = [("c", C.POINTER(Array4) * 3)]


class PredictorOptionsC(C.Structure):
    """``gust_predictor_options``."""

    _fields_ = [
        ("sweeps", C.c_int32),
        ("rotation_relax", C.c_double),
        ("w_vnn_cap", C.c_double),
        ("w_vnn_slope_margin", C.c_double),
        ("coriolis_f", C.c_double),
        ("gravity", C.c_double),
        ("theta0", C.c_double),
        ("sweep", LineOptionsC),
        ("sweep_precision", C.c_int32),
    ]


class MomentumPredictorResultC(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("max_du", "max_dv", "max_dw", "last_sweep_du", "last_sweep_dv",
                                          "last_sweep_dw")]


class ScalarSweepResultC(C.Structure):
    _fields_ = [("max_dq", C.c_double), ("last_sweep", C.c_double), ("clipped", C.c_int64)]


class JvpScalesC(C.Structure):
    """``gust_jvp_scales`` (Source/Residual.H:72-83)."""

    _fields_ = [("velocity", C.c_double), ("theta", C.c_double), ("tke", C.c_double), ("pressure", C.c_double),
                ("step", C.c_double)]


class ResidualsC(C.Structure):
    """``gust_residuals``: u, v, w, p, theta, k residual views."""

    _fields_ = [("R", C.POINTER(Array4) * 6)]


class SteadyParamsC(C.Structure):
    """``gust_steady_params`` (Source/SteadyCriterion.H:13-33)."""

    _fields_ = [("residual_tol", C.c_double), ("residual_drop_orders", C.c_double),
                ("residual_ref_iterations", C.c_int32), ("monitor_rel_tol", C.c_double),
                ("monitor_window", C.c_int32)]


class SteadyResultC(C.Structure):
    """``gust_steady_result`` (Source/SteadyCriterion.H:57-64)."""

    _fields_ = [("status", C.c_int32), ("iteration", C.c_int32), ("residuals_met", C.c_int32),
                ("monitors_met", C.c_int32), ("nreasons", C.c_int32), ("reasons_len", C.c_int64)]


class ClosureOptionsC(C.Structure):
    """``gust_closure_options`` (Source/Closure.H:78-105)."""

    _fields_ = [(n, C.c_double) for n in ("Cmu0", "Cb", "Rt_crit", "Rt_min", "max_geom_lscale", "tke_floor",
                                          "sigma_k", "nu_molecular", "alpha_molecular", "kappa")]


class SurfaceOptionsC(C.Structure):
    """``gust_surface_options`` (Source/SurfaceLayer.H:100-146)."""

    _fields_ = [("z0", C.c_double), ("Cmu0", C.c_double), ("kappa", C.c_double), ("wall_k", C.c_int32),
                ("denominator_floor", C.c_int32)]


class ForcingC(C.Structure):
    """``gust_forcing`` (Source/Forcing.H:40-96)."""

    _fields_ = [("gravity", C.c_double), ("theta0", C.c_double), ("coriolis", C.c_int32), ("coriolis_3d", C.c_int32),
                ("geostrophic", C.c_int32), ("latitude", C.c_double), ("rotation_period", C.c_double),
                ("ug", C.c_double), ("vg", C.c_double), ("pressure_gradient", C.c_double * 3)]


class InflowProfileC(C.Structure):
    """``gust_inflow_profile`` (Source/InflowProfile.H:42-66, the table in memory)."""

    _fields_ = [("kind", C.c_int32), ("ustar", C.c_double), ("z0", C.c_double), ("speed_ref", C.c_double),
                ("z_ref", C.c_double), ("alpha", C.c_double), ("direction", C.c_double), ("Cmu0", C.c_double),
                ("kappa", C.c_double), ("unbounded", C.c_int32), ("l_max", C.c_double), ("n", C.c_int32),
                ("z", C.POINTER(C.c_double)), ("u", C.POINTER(C.c_double)), ("v", C.POINTER(C.c_double)),
                ("theta", C.POINTER(C.c_double)), ("k", C.POINTER(C.c_double)), ("n_theta_ref", C.c_int32),
                ("theta_ref_z", C.c_double * 64), ("theta_ref", C.c_double * 64)]


class DampingC(C.Structure):
    """``gust_damping`` (Source/BoundaryConditions.H:53-57, 107-121)."""

    _fields_ = [("rayleigh_zdamp", C.c_double), ("rayleigh_dampcoef", C.c_double),
                ("rayleigh_fields", C.c_int32 * 5), ("rayleigh_above_ground", C.c_int32),
                ("sponge_width", C.c_double), ("sponge_strength", C.c_double), ("ztop", C.c_double),
                ("prob_lo", C.c_double * 2), ("prob_hi", C.c_double * 2)]


class PhysicsC(C.Structure):
    """``gust_physics`` (Source/Residual.H:90-128)."""

    _fields_ = [("use_closure", C.c_int32), ("closure", ClosureOptionsC), ("nu", C.POINTER(Array4)),
                ("gamma_theta", C.POINTER(Array4)), ("gamma_tke", C.POINTER(Array4)), ("surface", SurfaceOptionsC),
                ("forcing", ForcingC), ("use_bc", C.c_int32), ("face", C.c_int32 * 6),
                ("zhi_theta_grad", C.c_double), ("wall_velocity", (C.c_double * 3) * 6),
                ("n_theta_ref", C.c_int32), ("theta_ref_z", C.c_double * 64), ("theta_ref", C.c_double * 64),
                ("momentum_limiter", C.c_int32), ("theta_limiter", C.c_int32), ("momentum_eps", C.c_double),
                ("theta_eps", C.c_double), ("face_mean", C.c_int32), ("inflow", InflowProfileC),
                ("damping", DampingC)]


class StateC(C.Structure):
    """``gust_state``: u, v, w, p, theta, k views."""

    _fields_ = [("f", C.POINTER(Array4) * 6)]


class MeshC(C.Structure):
    """``gust_mesh``."""

    _fields_ = [("z_nd", C.POINTER(Array4)), ("detJ", C.POINTER(Array4)), ("ax", C.POINTER(Array4)),
                ("ay", C.POINTER(Array4)), ("az", C.POINTER(Array4))]


class ClosureFieldsC(C.Structure):
    """``gust_closure_fields``."""

    _fields_ = [(n, C.POINTER(Array4)) for n in ("nu", "gamma_theta", "gamma_tke", "Ps", "Pb", "eps", "dSdk")]


class InitOptionsC(C.Structure):
    """``gust_init_options`` (Source/Initializer.H:63-98)."""

    _fields_ = [("first_guess", C.c_int32), ("projection", C.c_int32), ("z_agl_min", C.c_double),
                ("alpha_h", C.c_double), ("alpha_v", C.c_double), ("k_decay_height", C.c_double),
                ("obrien", C.c_int32), ("column_theta", C.c_int32)]


class InitResultC(C.Structure):
    """``gust_init_result`` (Source/Initializer.H:110-120)."""

    _fields_ = [("poisson", PoissonResultC), ("divergence_before", C.c_double), ("divergence_after", C.c_double),
                ("obrien_top_residual", C.c_double)]


class SimplecOptionsC(C.Structure):
    """``gust_simplec_options`` (Source/Simplec.H:71-83 and the stage's settings)."""

    _fields_ = [("pressure_reltol", C.c_double), ("alpha_p", C.c_double), ("rotation_relax", C.c_double),
                ("w_vnn_cap", C.c_double), ("w_vnn_cap_slope_margin", C.c_double), ("sweeps", C.c_int32),
                ("sweep", LineOptionsC), ("tke_floor", C.c_double), ("scales", JvpScalesC), ("height", C.c_double),
                ("poisson", PoissonOptionsC), ("sweep_precision", C.c_int32)]


class SimplecInfoC(C.Structure):
    """``gust_simplec_info`` (Source/Simplec.H:178-184)."""

    _fields_ = [("predictor", MomentumPredictorResultC), ("pressure", PressureCorrectionResultC),
                ("theta", ScalarSweepResultC), ("tke", ScalarSweepResultC)]


class DriverOptionsC(C.Structure):
    """``gust_driver_options`` (Source/Driver.H:58-66)."""

    _fields_ = [("max_iterations", C.c_int32), ("cfl0", C.c_double), ("cfl_max", C.c_double),
                ("ser_exponent", C.c_double), ("ser_max_growth", C.c_double), ("criterion", SteadyParamsC)]


class DriverResultC(C.Structure):
    """``gust_driver_result``."""

    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32), ("cfl", C.c_double),
                ("relative_residual", C.c_double), ("res", C.c_double * 6), ("mon", C.c_double * 3),
                ("nmon", C.c_int32)]


class SequencingOptionsC(C.Structure):
    """``gust_sequencing_options``."""

    _fields_ = [("enabled", C.c_int32), ("drop_orders", C.c_double), ("monitor_tol", C.c_double),
                ("hold", C.c_int32), ("cfl_previous", C.c_int32), ("given", C.c_uint32)]


class LevelControlC(C.Structure):
    """``gust_level_control``."""

    _fields_ = [("loose", C.POINTER(SequencingOptionsC)), ("seeded", C.c_int32), ("seed_ref", C.c_double * 6),
                ("release_hold", C.c_int32), ("start_cfl", C.c_double)]


class LevelResultC(C.Structure):
    """``gust_level_result``."""

    _fields_ = [("driver", DriverResultC), ("stop", C.c_int32), ("hold", C.c_int32), ("res_ref", C.c_double * 6),
                ("reason", C.c_char * 256)]


class ProlongResultC(C.Structure):
    """``gust_prolong_result``."""

    _fields_ = [("projection", PressureCorrectionResultC), ("tke_clipped", C.c_int64)]


class NewtonOptionsC(C.Structure):
    """``gust_newton_options`` (Source/Newton.H:78-99, Continuation.H:63-79)."""

    _fields_ = [("gmres_restart", C.c_int32), ("gmres_max_iters", C.c_int32), ("fd_eps_scale", C.c_double),
                ("krylov_rtol", C.c_double), ("preconditioner", C.c_int32), ("cfl_max", C.c_double),
                ("start_cfl", C.c_double), ("ser_exponent", C.c_double), ("ser_max_growth", C.c_double),
                ("armijo_c", C.c_double), ("backtrack", C.c_double), ("lambda_min", C.c_double),
                ("cfl_cut", C.c_double), ("max_retries", C.c_int32), ("dtau_inf_at_cap", C.c_int32),
                ("ptc_growth", C.c_double), ("pc_denominator", C.c_int32)]


class NewtonStepInfoC(C.Structure):
    """``gust_newton_step_info`` (Source/Newton.H:236-252)."""

    _fields_ = [("gmres_iterations", C.c_int32), ("gmres_status", C.c_int32), ("residual_norm", C.c_double),
                ("krylov_residual", C.c_double), ("krylov_relative", C.c_double), ("eps_min", C.c_double),
                ("eps_max", C.c_double), ("jvp_calls", C.c_int32), ("applied", C.c_int32)]


class ContinuationStateC(C.Structure):
    """``gust_continuation_state`` (Source/Continuation.H:118-126)."""

    _fields_ = [("steps", C.c_int32), ("cfl", C.c_double), ("infinite", C.c_int32), ("r0", C.c_double),
                ("r_last", C.c_double)]


class ContinuationRecordC(C.Structure):
    """``gust_continuation_record`` (Source/Continuation.H:133-151)."""

    _fields_ = [("step", C.c_int32), ("attempt", C.c_int32), ("residual", C.c_double),
                ("residual_new", C.c_double), ("cfl", C.c_double), ("infinite", C.c_int32), ("lambda_", C.c_double),
                ("backtracks", C.c_int32), ("gmres_iterations", C.c_int32), ("gmres_status", C.c_int32),
                ("krylov_relative", C.c_double), ("outcome", C.c_int32), ("by_growth", C.c_int32)]


class SwitchOptionsC(C.Structure):
    """``gust_switch_options`` (Source/Driver.H:67-97)."""

    _fields_ = [("switch_window", C.c_int32), ("switch_residual", C.c_double),
                ("switch_back_growth", C.c_double), ("switch_monitor_tol", C.c_double),
                ("newton_fallback_failures", C.c_int32)]


class NewtonRunResultC(C.Structure):
    """``gust_newton_run_result``."""

    _fields_ = [("driver", DriverResultC), ("simplec_iterations", C.c_int32), ("newton_iterations", C.c_int32),
                ("switches", C.c_int32), ("first_switch", C.c_int32), ("newton_steps", C.c_int32),
                ("stage", C.c_int32), ("handed_back", C.c_int32)]


def _load() -> C.CDLL:
    # One HIP runtime per process: torch ships its own libamdhip64 and the
    # device tensors, streams and events handed to the library come from it.
    # Loading torch first makes the library bind to that runtime instead of
    # bringing a second one in (a second runtime sees no usable device).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python ablsteadyrans_amd/build.py` "
            "(there is no non-HIP fallback)")
    return C.CDLL(LIB_PATH)


lib = _load()

_P = C.POINTER
_sig = {
    "gust_last_error": (C.c_char_p, []),
    "gust_version": (C.c_char_p, []),
    "gust_stream_bytes_held": (C.c_int, [C.c_void_p, _P(C.c_uint64)]),
    "gust_release_stream": (C.c_int, [C.c_void_p]),
    "gust_zline_jacobi": (C.c_int, [C.c_int, _P(Array4), _P(Array4), _P(Array4), C.c_int, _P(GeometryC),
                                    _P(Array4), _P(C.c_double), C.c_void_p]),
    "gust_zline_jacobi_noalloc": (C.c_int, [C.c_int, _P(Array4), _P(Array4), _P(Array4), C.c_int, _P(GeometryC),
                                    _P(Array4), _P(C.c_double), C.c_void_p]),
    "gust_zline_jacobi_f32": (C.c_int, [C.c_int, _P(Array4), _P(Array4), _P(Array4), C.c_int, _P(GeometryC),
                                        _P(Array4), _P(C.c_double), C.c_void_p]),
    "gust_zline_jacobi_batched_async": (C.c_int, [C.c_int, _P(LineSystemC), C.c_int, C.c_void_p, C.c_void_p]),
    "gust_zline_jacobi_batched_async_f32": (C.c_int, [C.c_int, _P(LineSystemC), C.c_int, C.c_void_p,
                                                      C.c_void_p]),
    "gust_zline_lusgs": (C.c_int, [C.c_int, _P(Array4), _P(Array4), _P(Array4), C.c_int, _P(GeometryC),
                                   _P(Array4), _P(LineOptionsC), _P(C.c_double), C.c_void_p]),
    "gust_line_sweeps": (C.c_int, [_P(LineOptionsC), C.c_int, _P(Array4), _P(Array4), _P(Array4), C.c_int,
                                   _P(GeometryC), _P(Array4), _P(C.c_double), C.c_void_p]),
    "gust_check_line_options": (C.c_int, [_P(LineOptionsC)]),
    "gust_jacobi_plan_create": (C.c_int, [C.c_int, _P(LineSystemC), C.c_int, _P(C.c_void_p), C.c_void_p]),
    "gust_jacobi_plan_run": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "gust_jacobi_plan_sweep": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "gust_jacobi_plan_info": (C.c_int, [C.c_void_p, _P(C.c_int64), _P(C.c_int32)]),
    "gust_jacobi_plan_destroy": (None, [C.c_void_p]),
    "gust_jacobi_set_schedule": (C.c_int, [C.c_int32]),
    "gust_jacobi_tuned": (C.c_int, [C.c_int32, _P(C.c_int32), _P(C.c_int32), _P(C.c_int32), _P(C.c_int32),
                                    _P(C.c_int32), _P(C.c_float), C.c_char_p, C.c_int64]),
    "gust_jacobi_tune_reset": (C.c_int, []),
    "gust_build_mesh": (C.c_int, [_P(GeometryC), _P(C.c_double), _P(C.c_double), _P(Array4), _P(Array4),
                                  _P(Array4), _P(Array4), _P(Array4), _P(Array4), C.c_void_p]),
    "gust_poisson_default_options": (C.c_int, [_P(PoissonOptionsC)]),
    "gust_poisson_create": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), _P(Array4), _P(Array4), _P(Array4),
                                      _P(PoissonOptionsC), _P(C.c_void_p), C.c_void_p]),
    "gust_poisson_destroy": (None, [C.c_void_p]),
    "gust_poisson_levels": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "gust_poisson_solve": (C.c_int, [C.c_void_p, _P(Array4), _P(Array4), _P(FaceCoeffsC), _P(C.c_int32),
                                     _P(Array4), _P(Array4), _P(Array4), _P(PoissonResultC), C.c_void_p]),
    "gust_poisson_apply": (C.c_int, [C.c_void_p, _P(Array4), _P(Array4), _P(FaceCoeffsC), _P(C.c_int32),
                                     C.c_void_p]),
    "gust_poisson_fill_ghosts": (C.c_int, [C.c_void_p, _P(Array4), _P(C.c_int32), C.c_void_p]),
    "gust_poisson_fluxes": (C.c_int, [C.c_void_p, _P(Array4), _P(FaceCoeffsC), _P(C.c_int32), _P(Array4),
                                      _P(Array4), _P(Array4), C.c_void_p]),
    "gust_poisson_vcycle": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, _P(FaceCoeffsC), _P(C.c_int32), C.c_int,
                                      C.c_void_p]),
    "gust_face_divergence": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), _P(Array4), _P(Array4), _P(Array4),
                                       _P(Array4), _P(Array4), _P(Array4), C.c_void_p]),
    "gust_omega_from_w": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), _P(Array4), _P(Array4), _P(Array4),
                                    C.c_int32, C.c_void_p]),
    "gust_w_from_omega": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), _P(Array4), _P(Array4), _P(Array4),
                                    C.c_void_p]),
    "gust_transport_fluxes": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 9 + [C.c_void_p]),
    "gust_velocity_divergence": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 9 + [C.c_void_p]),
    "gust_pressure_gradient": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 8 + [C.c_void_p]),
    "gust_pressure_correction": (C.c_int, [C.c_void_p, _P(GeometryC)] + [_P(Array4)] * 12 +
                                 [_P(C.c_int32), C.c_double, _P(Array4), C.c_int32, _P(PressureCorrectionResultC),
                                  C.c_void_p]),
    "gust_implicit_coefficients": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 6 +
                                   [C.c_int32, _P(Array4), C.c_void_p]),
    "gust_apply_implicit": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 3 + [C.c_void_p]),
    "gust_convective_fluxes": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 4 +
                               [C.c_int32, C.c_double] + [_P(Array4)] * 3 + [C.c_void_p]),
    "gust_flux_divergence": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 5 + [C.c_void_p]),
    "gust_convection": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 5 +
                        [C.c_int32, C.c_double, C.c_int32, _P(Array4), C.c_void_p]),
    "gust_scalar_diffusive_fluxes": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 3 + [C.c_int32] + [_P(Array4)] * 3 +
                                     [C.c_void_p]),
    "gust_scalar_diffusion": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 4 + [C.c_int32, _P(Array4), C.c_void_p]),
    "gust_momentum_diffusive_fluxes": (C.c_int, [C.c_int32, _P(GeometryC)] + [_P(Array4)] * 5 + [C.c_int32] +
                                       [_P(Array4)] * 3 + [C.c_void_p]),
    "gust_momentum_diffusion": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 6 + [C.c_int32] + [_P(Array4)] * 3 +
                                [C.c_void_p]),
    "gust_momentum_predictor": (C.c_int, [_P(GeometryC), _P(FacesC), _P(Array4), _P(Array4), _P(FacesC), _P(Array4),
                                          _P(FacesC), _P(Array4), _P(FacesC), _P(FacesC), _P(PredictorOptionsC),
                                          _P(FacesC), _P(FacesC), _P(MomentumPredictorResultC), C.c_void_p]),
    "gust_scalar_sweep": (C.c_int, [_P(GeometryC), C.c_int32, _P(Array4), _P(Array4), _P(Array4), _P(Array4), C.c_int32,
                                    C.c_double, C.c_int32, _P(LineOptionsC), _P(ScalarSweepResultC), C.c_void_p]),
    "gust_local_dtau": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 6 + [C.c_double, _P(Array4), C.c_void_p]),
    "gust_ser_cfl": (C.c_int, [C.c_double] * 6 + [_P(C.c_double)]),
    "gust_face_dtau": (C.c_int, [_P(GeometryC)] + [_P(Array4)] * 4 + [C.c_void_p]),
    "gust_deterministic_sum": (C.c_int, [_P(GeometryC), _P(Array4), C.c_int32, _P(C.c_double), C.c_void_p]),
    "gust_global_max": (C.c_int, [_P(GeometryC), C.c_int32, _P(Array4), C.c_int32, _P(C.c_double), C.c_void_p]),
    "gust_scaled_sums": (C.c_int, [_P(GeometryC), C.c_int32, _P(Array4), _P(Array4), _P(Array4), C.c_double,
                                   _P(C.c_double), C.c_void_p]),
    "gust_residual_scales": (C.c_int, [_P(JvpScalesC), C.c_double, _P(C.c_double)]),
    "gust_residual_scale_height": (C.c_int, [_P(C.c_double), C.c_int32, C.c_double, C.c_double, _P(C.c_double)]),
    "gust_stage_norms": (C.c_int, [_P(GeometryC), _P(ResidualsC)] + [_P(Array4)] * 5 +
                         [_P(JvpScalesC), C.c_double, _P(C.c_double), _P(C.c_double), _P(C.c_int32), C.c_void_p]),
    "gust_confidence": (C.c_int, [_P(GeometryC), _P(ResidualsC), _P(Array4), _P(JvpScalesC), C.c_double,
                                  _P(C.c_double), _P(C.c_double), _P(C.c_double), C.c_void_p]),
    "gust_height_above_ground": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), C.c_void_p]),
    "gust_closure_evaluate": (C.c_int, [_P(GeometryC), _P(StateC), _P(Array4), _P(Array4), C.c_double,
                                        _P(ClosureOptionsC), C.c_double, C.c_double, _P(ClosureFieldsC), C.c_void_p]),
    "gust_surface_state": (C.c_int, [_P(GeometryC), _P(Array4), _P(Array4), _P(Array4), _P(SurfaceOptionsC),
                                     _P(Array4), C.c_void_p]),
    "gust_fill_ghosts": (C.c_int, [_P(GeometryC), _P(StateC), _P(MeshC), _P(PhysicsC), _P(Array4), C.c_void_p]),
    "gust_residual": (C.c_int, [_P(GeometryC), _P(StateC), _P(MeshC), _P(PhysicsC), _P(ResidualsC),
                                _P(ClosureFieldsC), _P(Array4), C.c_void_p]),
    "gust_implicit_rows": (C.c_int, [C.c_int32, _P(GeometryC), _P(FacesC), _P(MeshC), _P(PhysicsC),
                                     _P(ClosureFieldsC), _P(Array4), _P(Array4), C.c_void_p]),
    "gust_simplec_default_options": (C.c_int, [_P(SimplecOptionsC)]),
    "gust_simplec_create": (C.c_int, [_P(GeometryC), _P(MeshC), _P(PhysicsC), _P(SimplecOptionsC), _P(C.c_void_p),
                                      C.c_void_p]),
    "gust_simplec_destroy": (C.c_int, [C.c_void_p]),
    "gust_simplec_init": (C.c_int, [C.c_void_p, _P(StateC), C.c_void_p]),
    "gust_simplec_reset": (C.c_int, [C.c_void_p, C.c_void_p]),
    "gust_simplec_iterate": (C.c_int, [C.c_void_p, _P(StateC), _P(Array4), _P(SimplecInfoC), C.c_void_p]),
    "gust_simplec_norms": (C.c_int, [C.c_void_p, _P(C.c_double), _P(C.c_double), _P(C.c_int32)]),
    "gust_simplec_nu": (C.c_int, [C.c_void_p, _P(Array4)]),
    "gust_steady_solve": (C.c_int, [C.c_void_p, _P(StateC), _P(DriverOptionsC), _P(DriverResultC), C.c_void_p]),
    "gust_steady_solve_level": (C.c_int, [C.c_void_p, _P(StateC), _P(DriverOptionsC), _P(LevelControlC),
                                          _P(LevelResultC), C.c_void_p]),
    "gust_sequencing_default_options": (C.c_int, [_P(SequencingOptionsC)]),
    "gust_check_sequencing_options": (C.c_int, [_P(SequencingOptionsC), _P(SteadyParamsC), C.c_int32, C.c_int32,
                                                C.c_int32, _P(C.c_int32), C.c_char_p, C.c_int64]),
    "gust_loose_stop": (C.c_int, [C.c_void_p, _P(SequencingOptionsC), C.c_int32, _P(C.c_int32), C.c_char_p,
                                  C.c_int64]),
    "gust_steady_references": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_int32]),
    "gust_prolong_fields": (C.c_int, [_P(GeometryC), _P(GeometryC), C.c_int32, _P(StateC), _P(StateC), C.c_void_p]),
    "gust_restrict_state": (C.c_int, [_P(GeometryC), _P(MeshC), _P(StateC), _P(GeometryC), _P(MeshC), C.c_int32,
                                      _P(StateC), _P(Array4), C.c_void_p]),
    "gust_prolong_state": (C.c_int, [C.c_void_p, C.c_void_p, _P(GeometryC), C.c_int32, _P(StateC), _P(StateC),
                                     C.c_double, _P(ProlongResultC), C.c_void_p]),
    "gust_write_plotfile": (C.c_int, [C.c_char_p, C.c_int32, _P(GeometryC), _P(C.c_double), _P(C.c_double),
                                      _P(StateC), _P(Array4), _P(Array4), C.c_int32, C.c_char_p, C.c_int64,
                                      C.c_void_p]),
    "gust_write_checkpoint": (C.c_int, [C.c_char_p, C.c_int32, _P(GeometryC), _P(C.c_double), _P(C.c_double),
                                        _P(StateC), _P(Array4), C.c_int32, C.c_char_p, C.c_int64, C.c_void_p]),
    "gust_read_checkpoint": (C.c_int, [C.c_char_p, _P(GeometryC), _P(C.c_double), _P(C.c_double), _P(StateC),
                                       _P(Array4), _P(C.c_int32), C.c_void_p]),
    "gust_poisson_set_fixed_steps": (C.c_int, [C.c_void_p, C.c_int32]),
    "gust_simplec_set_pressure_steps": (C.c_int, [C.c_void_p, C.c_int32]),
    "gust_poisson_fixed_steps_misses": (C.c_int, [C.c_void_p, C.c_int32, _P(C.c_int64)]),
    "gust_simplec_capture": (C.c_int, [C.c_void_p, _P(StateC), _P(Array4), C.c_void_p]),
    "gust_simplec_replay": (C.c_int, [C.c_void_p, _P(SimplecInfoC), C.c_void_p]),
    "gust_steady_solve_graphed": (C.c_int, [C.c_void_p, _P(StateC), _P(DriverOptionsC), C.c_int32,
                                            _P(DriverResultC), C.c_void_p]),
    "gust_steady_default_params": (C.c_int, [_P(SteadyParamsC)]),
    "gust_steady_validate": (C.c_int, [_P(SteadyParamsC), _P(C.c_int32), C.c_char_p, C.c_int64]),
    "gust_steady_create": (C.c_int, [_P(SteadyParamsC), C.c_int32, _P(C.c_char_p), C.c_int32, _P(C.c_char_p),
                                     _P(C.c_double), _P(C.c_void_p)]),
    "gust_steady_destroy": (C.c_int, [C.c_void_p]),
    "gust_steady_add": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_int32, _P(C.c_double), C.c_int32]),
    "gust_steady_verdict": (C.c_int, [C.c_void_p, _P(SteadyResultC), C.c_char_p, C.c_int64]),
    "gust_steady_iterations": (C.c_int, [C.c_void_p, _P(C.c_int32)]),
    "gust_steady_residual_thresholds": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_int32]),
    "gust_steady_seed_reference": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_int32]),
    "gust_steady_history_save": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, _P(C.c_int64)]),
    "gust_steady_history_restore": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64]),
    "gust_steady_relative_residual": (C.c_int, [C.c_void_p, _P(C.c_double), C.c_int32, C.c_int32, C.c_double,
                                                _P(C.c_double)]),
    "gust_steady_next_cfl": (C.c_int, [C.c_void_p, C.c_int32] + [C.c_double] * 6 + [_P(C.c_double)]),
    "gust_inflow_default_profile": (C.c_int, [_P(InflowProfileC)]),
    "gust_check_inflow_profile": (C.c_int, [_P(InflowProfileC)]),
    "gust_inflow_eval": (C.c_int, [_P(InflowProfileC), C.c_int32] + [C.c_void_p] * 6 + [C.c_void_p]),
    "gust_init_default_options": (C.c_int, [_P(InitOptionsC)]),
    "gust_check_init_options": (C.c_int, [_P(InitOptionsC)]),
    "gust_initializer_poisson_bc": (C.c_int, [_P(GeometryC), _P(C.c_int32), _P(C.c_int32)]),
    "gust_obrien_omega": (C.c_int, [_P(GeometryC), _P(MeshC), _P(Array4), _P(Array4), _P(Array4), _P(C.c_double),
                                    C.c_void_p]),
    "gust_first_guess_state": (C.c_int, [_P(GeometryC), _P(MeshC), _P(InflowProfileC), _P(InitOptionsC), _P(StateC),
                                         _P(C.c_double), C.c_void_p]),
    "gust_initialize_state": (C.c_int, [_P(GeometryC), _P(MeshC), _P(InflowProfileC), _P(InitOptionsC),
                                        _P(C.c_int32), C.c_void_p, _P(StateC), _P(Array4), _P(InitResultC),
                                        C.c_void_p]),
    "gust_fill_ghosts_homogeneous": (C.c_int, [_P(GeometryC), _P(StateC), _P(MeshC), _P(PhysicsC), C.c_void_p]),
    "gust_newton_default_options": (C.c_int, [_P(NewtonOptionsC)]),
    "gust_check_newton_options": (C.c_int, [_P(NewtonOptionsC), C.c_char_p, C.c_int64, _P(C.c_int32)]),
    "gust_jvp": (C.c_int, [C.c_void_p, _P(StateC), _P(StateC), _P(ResidualsC), _P(ResidualsC), _P(C.c_double),
                           C.c_void_p]),
    "gust_newton_create": (C.c_int, [C.c_void_p, _P(NewtonOptionsC), _P(C.c_void_p), C.c_void_p]),
    "gust_newton_destroy": (None, [C.c_void_p]),
    "gust_newton_pc_setup": (C.c_int, [C.c_void_p, _P(StateC), _P(Array4), C.c_void_p]),
    "gust_newton_pc_apply": (C.c_int, [C.c_void_p, _P(ResidualsC), _P(ResidualsC), C.c_void_p]),
    "gust_newton_step": (C.c_int, [C.c_void_p, _P(StateC), _P(Array4), _P(NewtonStepInfoC), C.c_void_p]),
    "gust_newton_residual_norm": (C.c_int, [C.c_void_p, _P(StateC), _P(C.c_double), C.c_void_p]),
    "gust_continuation_step": (C.c_int, [C.c_void_p, _P(StateC), _P(C.c_int32), _P(ContinuationRecordC),
                                         C.c_void_p]),
    "gust_continuation_get": (C.c_int, [C.c_void_p, _P(ContinuationStateC)]),
    "gust_continuation_restore": (C.c_int, [C.c_void_p, _P(ContinuationStateC)]),
    "gust_switch_default_options": (C.c_int, [_P(SwitchOptionsC)]),
    "gust_steady_solve_newton": (C.c_int, [C.c_void_p, C.c_void_p, _P(StateC), _P(DriverOptionsC),
                                           _P(SwitchOptionsC), _P(NewtonRunResultC), C.c_void_p]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class GustError(RuntimeError):
    """A refused call; the message is the reference's "gust: <op>: ..." text."""


def check(status: int) -> None:
    if status != 0:
        raise GustError(lib.gust_last_error().decode())


def exported_symbols() -> list[str]:
    return list(_sig)


def stream_bytes_held(stream) -> int:
    """Device bytes the library holds for a torch stream (its scratch and caches)."""
    out = C.c_uint64(0)
    check(lib.gust_stream_bytes_held(C.c_void_p(stream.cuda_stream), C.byref(out)))
    return int(out.value)


def release_stream(stream) -> None:
    """Frees what the library holds for a torch stream (synchronises it first)."""
    check(lib.gust_release_stream(C.c_void_p(stream.cuda_stream)))
