"""The initializer's host-side checks in the HIP library (no device work):
the refusals of check_inflow_profile_options (Source/InflowProfile.cpp:
14-47), the table's rules (49-118), the sounding's (Source/Forcing.cpp:
30-49) and check_init_options (Source/Initializer.cpp:109-148), each naming
its key; initializer_poisson_bc (Source/Initializer.cpp:181-195) against
the oracle's for every face type and periodicity."""
import itertools

import numpy as np
import pytest

from ablsteadyrans_amd import _lib
from ablsteadyrans_amd import initializer as GI
from ablsteadyrans_amd.fab import Geometry
from oracle import initializer as OI
from oracle import physics as P


def _refused(fn, *args, match):
    with pytest.raises(_lib.GustError, match=match):
        fn(*args)


def test_valid_defaults_pass():
    GI.check_inflow_profile(GI.InflowProfile())
    GI.check_init_options(GI.InitOptions(first_guess=1))


@pytest.mark.parametrize("kw,key", [
    (dict(z0=0.0), "bc.inflow_z0"),
    (dict(ustar=-1.0), "bc.inflow_ustar"),
    (dict(ustar=float("nan")), "bc.inflow_ustar"),
    (dict(kind=GI.POWER_LAW, speed_ref=-1.0), "bc.inflow_speed"),
    (dict(kind=GI.POWER_LAW, z_ref=0.0), "bc.inflow_zref"),
    (dict(kind=GI.POWER_LAW, alpha=float("inf")), "bc.inflow_alpha"),
    (dict(direction=float("inf")), "bc.inflow_direction"),
    (dict(Cmu0=0.0), "bc.inflow_Cmu0"),
    (dict(kappa=-0.41), "bc.inflow_kappa"),
    (dict(l_max=0.0), "turb.max_geom_lscale"),
    (dict(theta_ref_z=(0.0, 0.0), theta_ref=(300.0, 301.0)), "heights must increase"),
    (dict(theta_ref_z=(0.0,), theta_ref=(-1.0,)), "theta finite and > 0"),
    (dict(kind=GI.TABLE), "bc.inflow_file"),
    (dict(kind=GI.TABLE, table=([1.0, 2.0], [1, 2], [0, 0], [300, 300], None)), "first height must be 0"),
    (dict(kind=GI.TABLE, table=([0.0, 2.0, 2.0], [1, 2, 3], [0, 0, 0], [300, 300, 300], None)), "increase strictly"),
    (dict(kind=GI.TABLE, table=([0.0, 2.0], [1, 2], [0, 0], [300, 0], None)), "theta must be > 0"),
    (dict(kind=GI.TABLE, table=([0.0, 2.0], [1, 2], [0, 0], [300, 300], [0.1, -0.1])), "k must be >= 0"),
    (dict(kind=GI.TABLE, table=([0.0, 2.0], [1, float("nan")], [0, 0], [300, 300], None)), "not finite"),
])
def test_profile_refusals_name_the_key(kw, key):
    _refused(GI.check_inflow_profile, GI.InflowProfile(**kw), match=key)


def test_unbounded_length_needs_no_cap():
    GI.check_inflow_profile(GI.InflowProfile(unbounded=True, l_max=0.0))


@pytest.mark.parametrize("kw,key", [
    (dict(z_agl_min=-1.0), "init.z_agl_min"),
    (dict(alpha_h=0.0), "init.alpha_h"),
    (dict(alpha_v=float("nan")), "init.alpha_v"),
    (dict(k_decay_height=-5.0), "init.k_decay_height"),
    (dict(first_guess=2), "init.first_guess"),
])
def test_init_refusals_name_the_key(kw, key):
    _refused(GI.check_init_options, GI.InitOptions(**kw), match=key)


def test_poisson_bc_matches_the_oracle():
    types = (P.PERIODIC, P.INFLOW, P.OUTFLOW, P.SLIP_WALL, P.NO_SLIP_WALL, P.SURFACE_LAYER, P.GEOSTROPHIC)
    rng = np.random.default_rng(3)
    for per in itertools.product((0, 1), repeat=2):
        g = Geometry.of_size(4, 4, 4, periodic=per)
        for _ in range(20):
            faces = tuple(int(x) for x in rng.choice(types, 6))
            assert GI.initializer_poisson_bc(g, faces) == OI.poisson_bc(faces, per), (per, faces)
