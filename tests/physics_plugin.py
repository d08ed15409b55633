"""A column physics plugin for the tests (the reference's ColumnPhysics interface, src/dyn/column_physics.hpp:21-34):
part2 sets tendencies that depend on the state (a Rayleigh friction of u and v, a Newtonian relaxation of theta, a
moistening), the other hooks record the order of the calls.  It works on any domain with field / set_field, the
GPU's RK3 and the oracle alike."""
import numpy as np

from cppwrf_amd import _abi as A


class Plugin:
    def __init__(self, th_forcing=False, qv_forcing=False, n_q=0):
        self.th_forcing = th_forcing
        self.qv_forcing = qv_forcing
        self.n_q = n_q
        self.calls = []
        self.tendf_at_part1 = []

    def part1(self, d):
        self.calls.append("part1")
        # init_zero_tendency ran before part1: every held tendency is zero here
        self.tendf_at_part1.append(max(float(np.abs(d.field(f)).max()) for f in (A.MI_RU_TENDF, A.MI_T_TENDF)))

    def part2(self, d):
        self.calls.append("part2")
        d.set_field(A.MI_RU_TENDF, -1e-3 * d.field(A.MI_U) * 1e4)
        if d.grid.ny > 1:
            d.set_field(A.MI_RV_TENDF, -2e-3 * d.field(A.MI_V) * 1e4)
        t = d.field(A.MI_T)
        d.set_field(A.MI_T_TENDF, -1e-2 * (t - t.mean(axis=(1, 2), keepdims=True)) * 1e4)
        for n in range(self.n_q):
            q = d.field(A.MI_Q + n)
            d.set_field(A.MI_Q_TENDF + n, 1e-7 * (1.0 + n) * (1.0 + 0.1 * np.tanh(q * 1e3)) * 1e4)

    def after_rk(self, d):
        self.calls.append("after_rk")

    def after_all_rk_steps(self, d):
        self.calls.append("after_all")
