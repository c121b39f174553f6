"""Host logic of the learned preconditioner: the guarded solve's fallback and
state restore, driven by a scripted stand-in for the solver (no GPU)."""
import numpy as np

from hpmp import _lib
from hpmp.learned import GuardedSolver


class ScriptedSolver:
    """Residual follows rate_learned per cycle in learned mode, rate_analytic otherwise;
    the state is a single number equal to the residual."""

    def __init__(self, rate_learned, rate_analytic=0.9, blow_up_at=None):
        self.mode = _lib.HPG_PRCN_ANALYTIC
        self.x = 1.0
        self.rl, self.ra = rate_learned, rate_analytic
        self.blow_up_at = blow_up_at
        self.ncycle = 0
        self.modes = []

    def set_preconditioner_mode(self, m):
        self.mode = m
        self.modes.append(m)

    def _check(self, rc, what):
        return rc

    def cycle(self, **kw):
        self.ncycle += 1
        if self.mode == _lib.HPG_PRCN_LEARNED and self.blow_up_at and self.ncycle >= self.blow_up_at:
            raise FloatingPointError("maxres is not positive")
        self.x *= self.rl if self.mode == _lib.HPG_PRCN_LEARNED else self.ra

    def rsdl(self):
        pass

    def to_finest(self):
        pass

    def resnorm(self):
        return self.x

    def get_state(self):
        return (np.array([self.x]),)

    def set_state(self, x):
        self.x = float(x[0])


def _guard(s):
    g = GuardedSolver.__new__(GuardedSolver)
    g.s, g.mlp, g.check_every, g.grow_tol, g.window, g.learned, g.fell_back_at = s, object(), 10, 10.0, 4, True, None
    g.grids = None
    g.cfl_scale, g.cfl0 = None, [2.5, 1.5, 1.0, 0.5]
    s.set_preconditioner_mode(_lib.HPG_PRCN_LEARNED)
    return g


def test_contracting_learned_run_never_falls_back():
    s = ScriptedSolver(0.5)
    g = _guard(s)
    it, hist = g.solve(tol_rel=1e-10, max_cycles=1000)
    assert g.fell_back_at is None and hist[-1][1] <= 1e-10 and it == 40


def test_growth_restores_best_state_and_continues_analytically():
    s = ScriptedSolver(1.5)
    g = _guard(s)
    it, hist = g.solve(tol_rel=1e-6, max_cycles=1000)
    # 1.5^10 = 57.7 > 10 x best (1.0) at the first check
    assert g.fell_back_at == 10
    assert s.modes[-1] == _lib.HPG_PRCN_ANALYTIC
    # restored to the initial state (1.0), then 0.9^k
    assert abs(hist[-1][1] - 0.9 ** (it - 10)) < 1e-12 and hist[-1][1] <= 1e-6


def test_stagnation_over_the_window_falls_back():
    s = ScriptedSolver(1.0)
    g = _guard(s)
    it, hist = g.solve(tol_rel=1e-6, max_cycles=1000)
    # no contraction: the 4th check (cycle 40) is not below the value 4 checks earlier
    assert g.fell_back_at == 40 and hist[-1][1] <= 1e-6


class BumpySolver(ScriptedSolver):
    """Contracts by 0.5 per cycle but jumps up x8 once at cycle 13."""

    def cycle(self, **kw):
        super().cycle(**kw)
        if self.ncycle == 13:
            self.x *= 2.0 ** 13


def test_transient_rise_within_tolerance_does_not_fall_back():
    s = BumpySolver(0.5)
    g = _guard(s)
    it, hist = g.solve(tol_rel=1e-10, max_cycles=1000)
    assert g.fell_back_at is None and hist[-1][1] <= 1e-10


def test_exception_in_learned_cycle_falls_back():
    s = ScriptedSolver(0.8, blow_up_at=15)
    g = _guard(s)
    it, hist = g.solve(tol_rel=1e-8, max_cycles=2000)
    assert g.fell_back_at == 15 and hist[-1][1] <= 1e-8
    # the best check was at cycle 10: state 0.8^10
    assert abs(hist[-1][1] - 0.8 ** 10 * 0.9 ** (it - 15)) < 1e-12


class ScriptedCflSolver(ScriptedSolver):
    """Residual follows rate_learned per cycle while the CFL differs from the reference
    one, rate_analytic otherwise."""

    def __init__(self, rate_tuned, rate_analytic=0.9):
        super().__init__(rate_tuned, rate_analytic)
        self.cfl = [2.5, 1.5, 1.0, 0.5]
        self.cfls = []

    def set_cfl(self, cfl):
        self.cfl = list(cfl)
        self.cfls.append(self.cfl)

    def cycle(self, **kw):
        self.ncycle += 1
        self.x *= self.rl if self.cfl[0] != 2.5 else self.ra


def test_tuned_cfl_guard_restores_the_reference_cfl():
    from hpmp.tri_hp import CFL_DEFAULT
    s = ScriptedCflSolver(1.5)                      # the tuned CFL diverges
    g = GuardedSolver(s, None, check_every=10, cfl_scale=2.5)
    assert s.cfl == [c * 2.5 for c in CFL_DEFAULT]
    it, hist = g.solve(tol_rel=1e-6, max_cycles=2000)
    assert g.fell_back_at is not None
    assert s.cfl == list(CFL_DEFAULT)
    assert hist[-1][1] <= 1e-6 * hist[0][1]


def test_tuned_cfl_guard_keeps_a_contracting_cfl():
    s = ScriptedCflSolver(0.5)
    g = GuardedSolver(s, None, check_every=10, cfl_scale=2.0)
    it, hist = g.solve(tol_rel=1e-6, max_cycles=2000)
    assert g.fell_back_at is None and it <= 30
