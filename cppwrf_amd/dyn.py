"""Python view of the library, named after the reference's src/dyn surface.

    ctx = Context(device=0)
    case = Case.bubble3d(64, 64, 40, dx=1000.0, ztop=20000.0, dt=6.0, seed=1)
    rk = RK3(ctx, case.grid, case.params(time_step_sound=6, damp_opt=3), case)   # dyn::RK3 + State
    rk.step(10)                                                                  # RK3::step x 10
    u = rk.field(MI_U)

The arrays are numpy float64 in the C ABI's dense layout: shape (nzs, nys, nxs), i fastest.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
from dataclasses import dataclass

import numpy as np

from . import _abi as A
from ._abi import check, field_shape, lib


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class Grid:
    """reference Grid (src/core/grid.hpp:21-31)"""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dt: float
    bc_x: int = A.MI_BC_PERIODIC
    bc_y: int = A.MI_BC_PERIODIC

    def c(self) -> A.MiGrid:
        return A.MiGrid(nx=self.nx, ny=self.ny, nz=self.nz, bc_x=self.bc_x, bc_y=self.bc_y, dx=self.dx, dy=self.dy,
                        dt=self.dt)

    @property
    def points(self) -> int:
        return self.nx * self.ny * self.nz


class Case:
    """A synthetic initial state built by the library's host generator (mi_case_*)."""

    STATE = (A.MI_U, A.MI_V, A.MI_W, A.MI_T, A.MI_PH, A.MI_MU, A.MI_P, A.MI_AL)

    # the fields a real-data-shaped case carries beyond the state and the base profiles
    HILLS_EXTRA = (A.MI_HT, A.MI_MSFTX, A.MI_MSFTY, A.MI_MSFUX, A.MI_MSFUY, A.MI_MSFVX, A.MI_MSFVY, A.MI_MSFVX_INV,
                   A.MI_F, A.MI_E)

    def __init__(self, kind: int, grid: Grid, ztop: float, seed: int = 0, noise_amp: float = 0.0,
                 bubble_dtheta: float = 2.0, bubble_radius: float = 4000.0, bubble_zc: float = 2000.0, u0: float = 0.0,
                 **hills):
        self.grid = grid
        self.kind = kind
        self.spec = A.MiCaseSpec(kind=kind, grid=grid.c(), ztop=ztop, seed=seed, noise_amp=noise_amp,
                                 bubble_dtheta=bubble_dtheta, bubble_radius=bubble_radius, bubble_zc=bubble_zc, u0=u0,
                                 **hills)
        h = C.c_void_p()
        check(lib().mi_case_create(C.byref(self.spec), C.byref(h)))
        self._h = h
        self.levels = A.MiLevels()
        check(lib().mi_case_levels(self._h, C.byref(self.levels)))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().mi_case_destroy(self._h)
            self._h = None

    @property
    def handle(self):
        return self._h

    def field(self, fid: int) -> np.ndarray:
        a = np.zeros(field_shape(fid, self.grid.nx, self.grid.ny, self.grid.nz))
        check(lib().mi_case_get_field(self._h, fid, _ptr(a), a.size))
        return a

    def profile(self, fid: int) -> np.ndarray:
        """a base-state field's vertical profile (the generator's base state is horizontally uniform)"""
        return np.ascontiguousarray(self.field(fid)[:, 0, 0])

    @property
    def full_base(self) -> bool:
        """the base state varies horizontally (terrain): base fields are full arrays"""
        return self.kind == A.MI_CASE_HILLS

    @property
    def moist(self) -> bool:
        return ((self.kind == A.MI_CASE_HILLS and bool(self.spec.moist)) or self.spec.qv0 > 0 or
                self.kind == A.MI_CASE_SQUALL2D)

    def params(self, **kw) -> A.MiParams:
        if self.kind == A.MI_CASE_HILLS:  # the flags a real-data-shaped domain needs
            kw.setdefault("terrain", 1)
            kw.setdefault("mapped", 1)
            kw.setdefault("coriolis", 1)
            if self.spec.moist:
                kw.setdefault("n_moist", 1)
                kw.setdefault("n_scalars", max(1, kw.get("n_scalars", 1)))
        elif self.moist:
            kw.setdefault("n_moist", 1)
            kw.setdefault("n_scalars", max(1, kw.get("n_scalars", 1)))
        return A.default_params(**kw)

    # the workloads of this library and the reference's idealized cases
    @classmethod
    def bubble3d(cls, nx, ny, nz, dx, ztop, dt, seed=1, noise_amp=0.05, dy=None, bc=A.MI_BC_PERIODIC, bc_y=None, **kw):
        g = Grid(nx, ny, nz, dx, dy if dy is not None else dx, dt, bc, bc if bc_y is None else bc_y)
        return cls(A.MI_CASE_BUBBLE3D, g, ztop, seed=seed, noise_amp=noise_amp, **kw)

    @classmethod
    def igw(cls):
        return cls(A.MI_CASE_IGW, Grid(300, 1, 10, 1000.0, 1000.0, 6.0), 10000.0)

    @classmethod
    def warm_bubble(cls):
        return cls(A.MI_CASE_WARM_BUBBLE, Grid(200, 1, 100, 100.0, 100.0, 1.0, A.MI_BC_SYMMETRIC), 10000.0)

    @classmethod
    def straka(cls, nx=512, nz=64, dx=100.0, dt=1.0):
        """the reference's straka case (configs/straka.ini: x-z density current, periodic, K = 75 m2/s with
        pr_inv = 1; run it with params(diff_opt=1, khdif=75, kvdif=75, pr_inv=1))"""
        return cls(A.MI_CASE_STRAKA, Grid(nx, 1, nz, dx, dx, dt, A.MI_BC_PERIODIC), 6400.0)

    @classmethod
    def squall2d(cls, nx=120, nz=40, dx=1000.0, dt=6.0):
        """the reference's squall2d case (configs/squall2d.ini: x-z moist bubble in the Weisman-Klemp sounding,
        periodic; run it with params(n_moist=3, n_scalars=3, mp_physics=1): qv, qc, qr and Kessler)"""
        return cls(A.MI_CASE_SQUALL2D, Grid(nx, 1, nz, dx, dx, dt, A.MI_BC_PERIODIC), 20000.0)

    @classmethod
    def rest_state(cls, nx=24, ny=24, nz=40, u0=0.0, bc=A.MI_BC_PERIODIC, bc_y=None, qv0=0.0, qv_hscale=0.0):
        return cls(A.MI_CASE_REST_STATE, Grid(nx, ny, nz, 2000.0, 2000.0, 10.0, bc, bc if bc_y is None else bc_y), 20000.0,
                   u0=u0, qv0=qv0, qv_hscale=qv_hscale)

    @classmethod
    def hills(cls, nx, ny, nz, dx, dt, seed=1, n_hills=12, hill_height=2500.0, hill_wmin=30000.0, hill_wmax=150000.0,
              jet_min=5.0, jet_max=40.0, map_amp=0.05, etac=0.2, p_top=5000.0, moist=True, bc=A.MI_BC_PERIODIC,
              bc_y=None):
        """real-data-shaped domain: seeded hills, US standard atmosphere, jets, map factors, Coriolis, hybrid
        coordinate with a p_top top, qv from a relative humidity of 80 % falling to 10 % at 10 km"""
        g = Grid(nx, ny, nz, dx, dx, dt, bc, bc if bc_y is None else bc_y)
        return cls(A.MI_CASE_HILLS, g, 0.0, seed=seed, n_hills=n_hills, moist=1 if moist else 0, hill_height=hill_height,
                   hill_wmin=hill_wmin, hill_wmax=hill_wmax, jet_min=jet_min, jet_max=jet_max, map_amp=map_amp,
                   etac=etac, p_top=p_top)


class Context:
    """A GPU and the stream its domains launch on (mi_ctx)."""

    def __init__(self, device: int = 0, stream: int | None = None):
        h = C.c_void_p()
        check(lib().mi_ctx_create(device, C.c_void_p(stream) if stream else None, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().mi_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class Batch:
    """Domains of one context stepped together, one HIP graph per step (mi_batch): a branch per member, or each
    acoustic substep's kernels as one launch over all members (the fused schedule)."""

    def __init__(self, ctx: Context, members, schedule: int = A.MI_BATCH_BRANCHES):
        self.ctx = ctx
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.handle.value for m in self.members])
        h = C.c_void_p()
        check(lib().mi_batch_create(ctx.handle, arr, len(self.members), C.byref(h)))
        self._h = h
        if schedule != A.MI_BATCH_BRANCHES:
            self.set_schedule(schedule)

    def set_schedule(self, schedule: int) -> None:
        """MI_BATCH_BRANCHES (a branch per member) or MI_BATCH_FUSED (each acoustic substep one launch over all
        members)"""
        check(lib().mi_batch_set_schedule(self._h, schedule))

    def step(self, n: int = 1) -> None:
        check(lib().mi_batch_step(self._h, n))

    def sync(self) -> None:
        for m in self.members:
            m.sync()

    @property
    def graph_count(self) -> int:
        n = lib().mi_batch_graph_count(self._h)
        check(min(n, 0))
        return n

    def close(self):
        if getattr(self, "_h", None):
            lib().mi_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


class RK3:
    """reference dyn::RK3 together with the State / BaseState it steps (src/dyn/rk3.hpp:34-78)."""

    def __init__(self, ctx: Context, grid: Grid, params: A.MiParams, case: Case | None = None,
                 levels: A.MiLevels | None = None):
        if levels is None:
            if case is None:
                raise ValueError("RK3 needs the level coefficients: pass a Case or levels")
            levels = case.levels
        self.ctx = ctx
        self.grid = grid
        self.params = params
        h = C.c_void_p()
        check(lib().mi_domain_create(ctx.handle, C.byref(grid.c()), C.byref(params), C.byref(levels), C.byref(h)))
        self._h = h
        if case is not None:
            check(lib().mi_case_load(case.handle, self._h))

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            lib().mi_domain_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def n_sound(self) -> int:
        return lib().mi_domain_n_sound(self._h)

    @property
    def device_bytes(self) -> int:
        return lib().mi_domain_device_bytes(self._h)

    def step(self, n: int = 1) -> None:
        check(lib().mi_step(self._h, n))
        self._raise_hook_error()

    def field_view(self, fid: int) -> "A.MiFieldView":
        """the device address, element size and strides of a field (reference Field3::view)"""
        v = A.MiFieldView()
        check(lib().mi_domain_field_view(self._h, fid, C.byref(v)))
        return v

    def field_layout(self, fid: int) -> "A.MiFieldLayout":
        """the layout a caller-owned buffer of the field must have (mi_domain_field_layout)"""
        v = A.MiFieldLayout()
        check(lib().mi_domain_field_layout(self._h, fid, C.byref(v)))
        return v

    def bind(self, fid: int, dev_ptr, nbytes: int) -> None:
        """make a caller-owned device allocation the field's storage (mi_field_bind); None gives the field storage
        of the domain's own again"""
        check(lib().mi_field_bind(self._h, fid, dev_ptr, int(nbytes)))

    def set_column_physics(self, cp) -> None:
        """RK3::set_column_physics (reference src/dyn/column_physics.hpp:21-34): `cp` has the methods part1(rk3),
        part2(rk3), after_rk(rk3), after_all_rk_steps(rk3) (any may be missing) and the flags th_forcing,
        qv_forcing; None detaches.  The methods read and write fields with field / set_field, among them the held
        tendencies MI_RU_TENDF, MI_RV_TENDF, MI_T_TENDF, MI_Q_TENDF + n in part2.  An exception raised in a method
        is raised again when the step returns."""
        self._phys = cp
        if cp is None:
            self._phys_struct = None
            check(lib().mi_domain_set_column_physics(self._h, None))
            return
        self._hook_error = None

        def thunk(name):
            f = getattr(cp, name, None)
            if f is None:
                return A.PHYS_FN()

            def call(_u, _d):
                if self._hook_error is None:
                    try:
                        f(self)
                    except BaseException as e:  # noqa: BLE001 - re-raised after the step
                        self._hook_error = e
            return A.PHYS_FN(call)
        s = A.MiColumnPhysics(None, thunk("part1"), thunk("part2"), thunk("after_rk"), thunk("after_all_rk_steps"),
                              1 if getattr(cp, "th_forcing", False) else 0, 1 if getattr(cp, "qv_forcing", False) else 0)
        self._phys_struct = s
        check(lib().mi_domain_set_column_physics(self._h, C.byref(s)))

    def _raise_hook_error(self) -> None:
        e = getattr(self, "_hook_error", None)
        if e is not None:
            self._hook_error = None
            raise e

    def sync(self) -> None:
        check(lib().mi_domain_sync(self._h))

    def field(self, fid: int) -> np.ndarray:
        a = np.zeros(field_shape(fid, self.grid.nx, self.grid.ny, self.grid.nz))
        check(lib().mi_domain_get_field(self._h, fid, _ptr(a), a.size))
        return a

    def set_field(self, fid: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        check(lib().mi_domain_set_field(self._h, fid, _ptr(a), a.size))

    def autotune(self, trials: int = 1) -> list[int]:
        """time every block shape of the column kernels on this domain, keep the fastest (state restored)"""
        check(lib().mi_domain_autotune(self._h, trials))
        return self.tuning

    @property
    def tuning(self) -> list[int]:
        a = (C.c_int * 4)()
        check(lib().mi_domain_get_tuning(self._h, a, 4))
        return list(a)

    @tuning.setter
    def tuning(self, shapes) -> None:
        a = (C.c_int * 4)(*shapes)
        check(lib().mi_domain_set_tuning(self._h, a, 4))

    def save_state(self) -> np.ndarray:
        """the carried state (halos included) and the boundary clock as an opaque byte buffer (checkpoint)"""
        n = lib().mi_domain_state_bytes(self._h)
        if n < 0:
            check(int(n))
        buf = np.zeros(int(n), dtype=np.uint8)
        check(lib().mi_domain_save_state(self._h, buf.ctypes.data_as(C.c_void_p), buf.size))
        return buf

    def load_state(self, buf: np.ndarray) -> None:
        """resume from a buffer of save_state (a domain of the same grid, parameters and inputs)"""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        check(lib().mi_domain_load_state(self._h, buf.ctypes.data_as(C.c_void_p), buf.size))

    @staticmethod
    def adaptive_sound_steps(dt: float, dx: float, dy: float, max_msft: float) -> int:
        """RK3::adaptive_sound_steps: the acoustic step count of an adaptive time step (rk3.cpp:74-77)"""
        return lib().mi_adaptive_sound_steps(dt, dx, dy, max_msft)

    def set_dt(self, dt: float, n_sound: int) -> None:
        """RK3::set_dt with the boundary relaxation weights and the diffusion's dt (the adaptive time step)"""
        check(lib().mi_domain_set_dt(self._h, dt, n_sound))
        self.grid = dataclasses.replace(self.grid, dt=dt)

    @property
    def on_cfl(self):
        """RK3::on_cfl: f(vert, horiz) after each stage's tendencies (steps then run launch by launch)"""
        return getattr(self, "_on_cfl", None)

    @on_cfl.setter
    def on_cfl(self, f) -> None:
        self._on_cfl = f
        self._cfl_thunk = A.CFL_FN(lambda _u, v, h: f(v, h)) if f is not None else A.CFL_FN()
        check(lib().mi_domain_set_cfl_callback(self._h, self._cfl_thunk, None))

    @property
    def on_first_rk_prep(self):
        """RK3::on_first_rk_prep: f() after stage 1's rk_step_prep"""
        return getattr(self, "_on_prep", None)

    @on_first_rk_prep.setter
    def on_first_rk_prep(self, f) -> None:
        self._on_prep = f
        self._prep_thunk = A.STAGE_FN(lambda _u: f()) if f is not None else A.STAGE_FN()
        check(lib().mi_domain_set_first_rk_prep_callback(self._h, self._prep_thunk, None))

    def cfl_max(self) -> tuple[float, float]:
        """w_damp's vertical and horizontal Courant-number maxima of the current fields (LargeStep::cfl_max)"""
        v, h = C.c_double(), C.c_double()
        check(lib().mi_cfl_max(self._h, C.byref(v), C.byref(h)))
        return v.value, h.value

    def check_finite(self) -> int:
        bad = C.c_int(-1)
        rc = lib().mi_check_finite(self._h, C.byref(bad))
        if rc not in (A.MI_OK, A.MI_ENONFINITE):
            check(rc)
        return bad.value

    # the phases of a step, for tests
    def step_begin(self):
        check(lib().mi_step_begin(self._h))

    def rk_prep(self, rk):
        check(lib().mi_rk_prep(self._h, rk))
        self._raise_hook_error()

    def rk_tendency(self, rk):
        check(lib().mi_rk_tendency(self._h, rk))
        self._raise_hook_error()

    def small_steps(self, rk):
        check(lib().mi_small_steps(self._h, rk))

    def scalar_transport(self, rk):
        check(lib().mi_scalar_transport(self._h, rk))

    def calc_p_rho_phi(self, rk):
        check(lib().mi_calc_p_rho_phi(self._h, rk))

    def set_diffusion(self, diff_6th_opt: int, diff_6th_factor: float = 0.12) -> None:
        """RK3::set_diffusion with the sixth-order option: attach (1 linear, 2 monotonic), change or detach (0)"""
        check(lib().mi_domain_set_diffusion(self._h, diff_6th_opt, diff_6th_factor))

    def calc_p_rho_phi_interior(self, spec_zone: int) -> None:
        """SmallStep::calc_p_rho_phi_interior: p, al again away from the spec zone"""
        check(lib().mi_calc_p_rho_phi_interior(self._h, spec_zone))

    def step_end(self):
        check(lib().mi_step_end(self._h))
        self._raise_hook_error()

    # specified lateral boundaries (reference LateralBC::set / set_scalar / dtbc)
    def bdy_set(self, var: int, side: int, values: np.ndarray, tends: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        t = np.ascontiguousarray(tends, dtype=np.float64)
        if v.size != t.size:
            raise ValueError("values and tendencies differ in size")
        check(lib().mi_bdy_set(self._h, var, side, _ptr(v), _ptr(t), v.size))

    def load_boundary(self, bdy: dict) -> None:
        """every slab of a {(var, side): (values, tendencies)} dict (cppwrf_amd.boundary)"""
        for (var, side), (v, t) in bdy.items():
            self.bdy_set(var, side, v, t)

    @property
    def bdy_time(self) -> float:
        d = C.c_double()
        check(lib().mi_bdy_get_time(self._h, C.byref(d)))
        return d.value

    @bdy_time.setter
    def bdy_time(self, v: float) -> None:
        check(lib().mi_bdy_set_time(self._h, float(v)))
