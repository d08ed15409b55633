"""The benchmark workloads of this library, shared by bench.py and the tests.

hills_c2:  the real-data-shaped 119x89x44 domain at 15 km (seeded hills up to 2.5 km, 30-150 km wide, US standard
           atmosphere, jets of 5-40 m/s with +-1 m/s noise, map factors in [0.95, 1.05], Coriolis, hybrid
           coordinate with a 50 hPa top, six moist species with the positive-definite limiter), dt 90 s, 4 acoustic
           steps, damp_opt 3 in the top 5 km, specified lateral boundaries (a 5-point zone nudged toward slowly
           drifting boundary data).
hills_c3:  the same recipe at 425x300x50, 3 km, hills 10-60 km wide, dt 18 s, 6 acoustic steps, sixth-order
           diffusion (factor 0.12).
ensemble:  32 members of uneven sizes (nx 60-220, ny 60-180, nz 35/44/50, dx 3-15 km), each the 3-D warm bubble
           (N = 0.01, +2 K cos^2 bubble of radius 4 km at 2 km, +-0.05 K noise) with its own seed plus +-0.5 K theta
           and +-1 m/s wind perturbations, dt 6 s per km of dx, 6 acoustic steps, damp_opt 3.
serving:   Poisson arrivals (cppwrf_amd.serving.poisson_trace).
"""
from __future__ import annotations

import numpy as np

from . import _abi as A
from .boundary import drifting_boundary
from .dyn import Case, RK3

# the six moist species besides qv: seeded fractions of qv (cloud, rain, ice, snow, graupel), zero below 0
_SPECIES_FRACTIONS = (0.05, 0.02, 0.01, 0.01, 0.005)


def moist_species(case, n_moist, seed):
    """initial values of moist species 1 .. n_moist-1: seeded patchy fractions of qv (species 0 is the case's
    qv), so the positive-definite limiter sees sharp edges"""
    qv = case.field(A.MI_Q)
    rng = np.random.default_rng(seed)
    out = {}
    for n in range(1, n_moist):
        frac = _SPECIES_FRACTIONS[(n - 1) % len(_SPECIES_FRACTIONS)]
        mask = rng.random(qv.shape) > 0.7
        out[A.MI_Q + n] = frac * qv * mask
    return out


def hills_c2():
    """(case, params keywords, description, boundary data, extra initial fields)"""
    case = Case.hills(119, 89, 44, dx=15000.0, dt=90.0, seed=1, bc=A.MI_BC_SPECIFIED)
    kw = dict(time_step_sound=4, damp_opt=3, zdamp=5000.0, dampcoef=0.2, n_moist=6, n_scalars=6, moist_adv_opt=1)
    desc = {"workload": "real-data-shaped 119x89x44 at 15 km: 12 seeded hills, US standard atmosphere, jets, map "
                        "factors, Coriolis, hybrid coordinate (50 hPa top), 6 moist species with the PD limiter, dt "
                        "90 s, 4 acoustic steps; specified lateral boundaries (5-point relaxation zone, seeded slowly "
                        "varying boundary data)",
            "nx": 119, "ny": 89, "nz": 44, "dx_m": 15000.0, "dt_s": 90.0, "time_step_sound": 4, "moist_species": 6,
            "lateral_bc": "specified"}
    return case, kw, desc, drifting_boundary(case, seed=17), moist_species(case, 6, seed=23)


def hills_c3():
    case = Case.hills(425, 300, 50, dx=3000.0, dt=18.0, seed=3, hill_wmin=10000.0, hill_wmax=60000.0,
                      bc=A.MI_BC_SPECIFIED)
    kw = dict(time_step_sound=6, damp_opt=3, zdamp=5000.0, dampcoef=0.2, n_moist=6, n_scalars=6, moist_adv_opt=1,
              diff_6th_opt=1, diff_6th_factor=0.12)
    desc = {"workload": "real-data-shaped 425x300x50 at 3 km: 12 seeded hills 10-60 km wide, US standard "
                        "atmosphere, jets, map factors, Coriolis, hybrid coordinate (50 hPa top), 6 moist species "
                        "with the PD limiter, sixth-order diffusion, dt 18 s, 6 acoustic steps; specified lateral "
                        "boundaries with seeded slowly varying boundary data",
            "nx": 425, "ny": 300, "nz": 50, "dx_m": 3000.0, "dt_s": 18.0, "time_step_sound": 6, "moist_species": 6,
            "lateral_bc": "specified", "diff_6th_opt": 1}
    return case, kw, desc, drifting_boundary(case, seed=19), moist_species(case, 6, seed=29)


def make_domain(ctx, case, kw, bdy, extra, precision):
    rk = RK3(ctx, case.grid, case.params(precision=precision, **kw), case)
    if bdy:
        rk.load_boundary(bdy)
    for fid, a in extra.items():
        rk.set_field(fid, a)
    return rk


def ensemble_specs(members=32, seed=2024):
    """(nx, ny, nz, dx, member seed) of each member"""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(members):
        nx, ny = int(rng.integers(60, 221)), int(rng.integers(60, 181))
        nz, dx = int(rng.choice([35, 44, 50])), float(rng.uniform(3000.0, 15000.0))
        specs.append((nx, ny, nz, dx, int(rng.integers(1 << 30))))
    return specs


def ensemble_member(ctx, spec, precision):
    """one member of the ensemble as a domain: the warm-bubble recipe with its seed's perturbations"""
    nx, ny, nz, dx, sd = spec
    case = Case.bubble3d(nx, ny, nz, dx=dx, ztop=20000.0, dt=6.0 * dx / 1000.0, seed=sd, noise_amp=0.05)
    rk = RK3(ctx, case.grid, case.params(time_step_sound=6, damp_opt=3, zdamp=5000.0, precision=precision), case)
    prng = np.random.default_rng(sd)
    for fid, amp in ((A.MI_T, 0.5), (A.MI_U, 1.0), (A.MI_V, 1.0)):
        f = case.field(fid)
        rk.set_field(fid, f + prng.uniform(-amp, amp, size=f.shape))
    return rk


def rel_l2(a, b):
    nb = float(np.linalg.norm(b))
    return float(np.linalg.norm(a - b)) / nb if nb > 0 else float(np.linalg.norm(a - b))


ERROR_FIELDS = {"U": A.MI_U, "V": A.MI_V, "W": A.MI_W, "T": A.MI_T, "PH": A.MI_PH, "MU": A.MI_MU}


def error_report(d32, d64, moist):
    """relative L2 and max |difference| of the fields as WRF writes them, FP32-bulk against FP64"""
    ids = dict(ERROR_FIELDS)
    if moist:
        ids["QVAPOR"] = A.MI_Q
    out = {}
    for name, fid in ids.items():
        a, b = d32.field(fid), d64.field(fid)
        out[name] = {"rel_l2": rel_l2(a, b), "max_abs": float(np.abs(a - b).max())}
    return out
