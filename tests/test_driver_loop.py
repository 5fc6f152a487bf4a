"""Big-step loop semantics (runMPMYDW.py:638-877) on the host: a scripted stand-in
for the simulation supplies the residual and damage sequences, so every
termination rule, the gravity pulse and the XPBD-only phase are checked without
a GPU."""
from types import SimpleNamespace

from mpmwarpydw_amd import driver


class Scripted:
    """Substeps that only record what the loop asked for; residual(k) and
    damage(k) give the scalars read back after substep k (0-based)."""

    def __init__(self, residual=lambda k: 1.0, damage=lambda k: 0.0, n_steps=50, threshold=1e-8, stall_thr=1e-9,
                 stall_dur=0.1, dt=1e-3, pulse=(1.0, 0.0), xpbd_only=0, sleep=(10, 0), sleep_ratio=0.95,
                 big_steps=1):
        self.sim = SimpleNamespace(dt=dt, t=0.0, nSteps=n_steps, residualThreshold=threshold,
                                   damage_stall_threshold=stall_thr, damage_stall_duration=stall_dur,
                                   gravity=(0.0, 0.0, -9.81), xpbdOnlySteps=xpbd_only, dtxpbd=10 * dt,
                                   xpbd_sleep_termination_ratio=sleep_ratio, bigSteps=big_steps)
        self.args = SimpleNamespace(gravity_pulse_factor=pulse[0], gravity_pulse_duration=pulse[1])
        self.res, self.dmg, self.sleep = residual, damage, sleep
        self.calls = []
        self.k = 0

    def substep(self, counter):
        self.calls.append(("mpm", counter, self.sim.gravity))

    def loop_scalars(self):
        k = self.k
        self.k += 1
        return self.res(k), self.dmg(k)

    def mean_damage(self):
        return self.dmg(-1)

    def xpbd_only_step(self):
        self.calls.append(("xpbd_only",))

    def sleep_counts(self):
        return self.sleep

    def deactivate_below_datum(self):
        self.calls.append(("deactivate",))


def test_runs_n_steps_below_min_steps():
    S = Scripted(n_steps=50)
    assert driver.run_phase1(S) == (50, "nSteps")
    assert [c[1] for c in S.calls] == list(range(50))
    assert abs(S.sim.t - 50 * 1e-3) < 1e-12


def test_residual_stops_at_the_first_step_past_min_steps():
    # residual is below the threshold from the start; minSteps = 1000 still holds it
    S = Scripted(n_steps=5000, residual=lambda k: 1e-12)
    assert driver.run_phase1(S) == (1000, "residual")
    # above the threshold until substep 1499 (counter 1500 after it); damage
    # keeps growing so that the stall rule stays out of the way
    S = Scripted(n_steps=5000, residual=lambda k: 1e-12 if k >= 1499 else 1.0, damage=lambda k: 1e-6 * (k + 1))
    assert driver.run_phase1(S) == (1500, "residual")


def test_damage_stall():
    # damage never changes: the stall counter starts past minSteps and needs
    # int(duration/dt) = 5 steps
    S = Scripted(n_steps=5000, stall_dur=5e-3, threshold=0.0)
    assert driver.run_phase1(S) == (1005, "damage_stall")
    # damage that grows by more than the threshold every step never stalls
    S = Scripted(n_steps=1200, stall_dur=5e-3, threshold=0.0, damage=lambda k: 1e-6 * (k + 1))
    assert driver.run_phase1(S) == (1200, "nSteps")
    # one jump resets the count
    S = Scripted(n_steps=5000, stall_dur=5e-3, threshold=0.0, damage=lambda k: 1.0 if k >= 1002 else 0.0)
    assert driver.run_phase1(S) == (1008, "damage_stall")


def test_gravity_pulse():
    S = Scripted(n_steps=8, pulse=(2.0, 3.5e-3))  # int(3.5e-3 / 1e-3) = 3 pulsed substeps
    driver.run_phase1(S)
    g = [c[2][2] for c in S.calls]
    assert g == [-19.62] * 3 + [-9.81] * 5
    # a phase shorter than the pulse keeps the pulsed gravity
    S = Scripted(n_steps=2, pulse=(2.0, 3.5e-3))
    driver.run_phase1(S)
    assert S.sim.gravity[2] == -19.62


def test_xpbd_only_phase_and_big_steps():
    S = Scripted(n_steps=3, xpbd_only=7, sleep=(10, 5), sleep_ratio=0.95, big_steps=2)
    saved = []
    rec = driver.run(S, checkpoint=lambda s, b: saved.append(b))
    assert [(r["phase1_steps"], r["phase1_end"], r["phase2_steps"], r["phase2_end"]) for r in rec] == \
        [(3, "nSteps", 7, "steps")] * 2
    assert saved == [0, 1]
    kinds = [c[0] for c in S.calls]
    assert kinds == (["mpm"] * 3 + ["xpbd_only"] * 7 + ["deactivate"]) * 2
    assert abs(S.sim.t - 2 * (3e-3 + 7e-2)) < 1e-12
    # everything asleep: the phase ends after its first step
    S = Scripted(n_steps=3, xpbd_only=7, sleep=(10, 10))
    assert driver.run_phase2(S) == (1, "asleep")
    # no XPBD particles: the sleep test never fires
    S = Scripted(n_steps=3, xpbd_only=4, sleep=(0, 0))
    assert driver.run_phase2(S) == (4, "steps")
    S = Scripted(xpbd_only=0)
    assert driver.run_phase2(S) == (0, "off")


def test_pulse_rescales_the_base_gravity_each_big_step():
    S = Scripted(n_steps=2, pulse=(3.0, 5e-3), big_steps=2)
    driver.run(S)
    # both big steps start from 3x the initial gravity (not 9x)
    g = [c[2][2] for c in S.calls if c[0] == "mpm"]
    assert g == [-9.81 * 3] * 4
