"""Integration control (host, CPU): the known answers of the reference's
schedule and scheduler tests (tests/cli/test_schedules.py:71-390) and the
legacy interval countdown (run/engine.py:206-240) restated as a plain loop,
which the scheduler must reproduce bit for bit."""
import math

import pytest

from aeolus_amd.run.schedule import (SECONDS_PER_DAY, IntegrationScheduler,
                                     advective_cfl_timestep, count_snapshot_times, integrate,
                                     integration_plan, interval_snapshot_times,
                                     scheduler_tolerance, validate_snapshot_schedule)

MISALIGNED = 600.0000003


def _stores(plan):
    return [t for k, _, t in plan if k == "store"]


def _final(plan):
    s = [t for k, _, t in plan if k == "step"]
    return s[-1] if s else 0.0


def _countdown(t_end, dts, cfl):
    """The legacy interval bookkeeping as one loop (stores, steps, final t)."""
    t, tol, left = 0.0, 1e-6 * dts, 0.0
    stores, steps = [], []
    while t <= t_end + tol:
        if left <= tol:
            stores.append(t)
            left = dts
        rem = t_end - t
        if rem <= tol:
            break
        d = min(cfl, left, rem)
        if d <= 0:
            break
        t += d
        left = max(0.0, left - d)
        steps.append(d)
    return stores, steps, t


def test_count_schedules():
    assert count_snapshot_times(0, 86400.0) == []
    assert count_snapshot_times(1, 86400.0) == [86400.0]
    s = count_snapshot_times(4, MISALIGNED)
    assert s[0] == 0.0 and s[-1] == MISALIGNED and len(s) == 4
    assert interval_snapshot_times(21600.0, 86400.0) == [0.0, 21600.0, 43200.0, 64800.0, 86400.0]


def test_count_n1_and_n0_end_exactly():
    for n, stores in ((1, [MISALIGNED]), (0, [])):
        plan = integration_plan(MISALIGNED, 600.0, mode="count",
                                snapshot_times=count_snapshot_times(n, MISALIGNED))
        assert _stores(plan) == stores and _final(plan) == MISALIGNED


def test_count_lands_on_every_target():
    sched = count_snapshot_times(4, MISALIGNED)
    plan = integration_plan(MISALIGNED, 71.0, mode="count", snapshot_times=sched)
    assert _stores(plan) == sched and _final(plan) == MISALIGNED
    assert all(d > 0 for k, d, _ in plan if k == "step")
    big = 8.64e8
    plan = integration_plan(big, big, mode="count", snapshot_times=count_snapshot_times(1, big))
    assert _stores(plan) == [big] and _final(plan) == big


def test_count_positive_residual_gets_a_step():
    target = math.nextafter(1.0, math.inf)
    plan = integration_plan(2.0, 1.0, mode="count", snapshot_times=[target, 2.0])
    assert plan[:3] == [("step", 1.0, 1.0), ("step", target - 1.0, target),
                        ("store", 0.0, target)]
    assert _final(plan) == 2.0


@pytest.mark.parametrize("n", [0, 1])
def test_count_tiny_duration_is_integrated(n):
    t_end = 1e-16
    plan = integration_plan(t_end, 1.0, mode="count",
                            snapshot_times=count_snapshot_times(n, t_end))
    assert ("step", t_end, t_end) in plan and _final(plan) == t_end
    assert _stores(plan) == ([] if n == 0 else [t_end])


@pytest.mark.parametrize("t_end,dts,cfl", [
    (86400.0, 21600.0, 600.0), (1.1 * SECONDS_PER_DAY, 21600.0, 600.0),
    (0.02 * SECONDS_PER_DAY, 864.0, 137.0), (MISALIGNED, 200.0, 250.0),
    (5 * SECONDS_PER_DAY, 21600.0, 611.7)])
def test_interval_plan_matches_the_countdown(t_end, dts, cfl):
    plan = integration_plan(t_end, cfl, mode="interval", dt_snapshots=dts)
    stores, steps, final = _countdown(t_end, dts, cfl)
    assert _stores(plan) == stores
    assert [d for k, d, _ in plan if k == "step"] == steps
    assert _final(plan) == final


def test_interval_misaligned_omits_final_state():
    t_end = 1.1 * SECONDS_PER_DAY
    plan = integration_plan(t_end, 600.0, mode="interval", dt_snapshots=21600.0)
    assert all(abs(s - t_end) > 1e-6 for s in _stores(plan))
    cp = integration_plan(t_end, 600.0, mode="count", snapshot_times=count_snapshot_times(5, t_end))
    assert _final(cp) == t_end


def test_cfl_timestep():
    assert advective_cfl_timestep(1000.0, 20.0) == 25.0
    assert advective_cfl_timestep(800.0, 20.0) == 0.5 * 800.0 / 20.0
    assert advective_cfl_timestep(1000.0, 40.0) < advective_cfl_timestep(1000.0, 20.0)
    assert advective_cfl_timestep(1000.0, 0.0) == 600.0
    assert advective_cfl_timestep(1000.0, 0.0, fallback=42.0) == 42.0
    assert advective_cfl_timestep(None, 20.0) == 600.0
    assert advective_cfl_timestep(0.0, 20.0) == 600.0
    for bad in (float("nan"), float("inf"), -1.0, "x"):
        with pytest.raises(ValueError):
            advective_cfl_timestep(1000.0, bad)
    for bad in (float("nan"), -5.0):
        with pytest.raises(ValueError):
            advective_cfl_timestep(bad, 20.0)
    with pytest.raises(ValueError):
        advective_cfl_timestep(1e308, 1e-308)   # overflows to inf


def test_scheduler_uses_the_ceiling_of_each_call():
    s = IntegrationScheduler(1000.0, mode="count", snapshot_times=[0.0, 1000.0])
    assert s.next_event(100.0) == ("store", 0.0, 0.0)
    assert s.next_event(300.0) == ("step", 300.0, 300.0)
    assert s.next_event(400.0) == ("step", 400.0, 700.0)
    assert s.next_event(1000.0) == ("step", 300.0, 1000.0)
    assert s.next_event(500.0) == ("store", 0.0, 1000.0)
    assert s.next_event(500.0) is None


def test_scheduler_exact_stores_under_varying_ceilings():
    sched = count_snapshot_times(4, MISALIGNED)
    s = IntegrationScheduler(MISALIGNED, mode="count", snapshot_times=sched)
    ceil = iter([50.0, 71.0, 123.0, 40.0, 200.0, 99.0])
    stores, final = [], 0.0
    while (ev := s.next_event(next(ceil, 250.0))) is not None:
        if ev[0] == "store":
            stores.append(ev[2])
        else:
            final = ev[2]
    assert stores == sched and final == MISALIGNED


def test_scheduler_interval_steps_follow_ceiling():
    s = IntegrationScheduler(10000.0, mode="interval", dt_snapshots=4000.0)
    ceil = [500.0, 1000.0, 1500.0, 2000.0]
    stores, steps, i = [], [], 0
    while (ev := s.next_event(ceil[i] if i < len(ceil) else 2000.0)) is not None:
        i += 1
        (stores if ev[0] == "store" else steps).append(ev[2] if ev[0] == "store" else ev[1])
    assert {0.0, 4000.0, 8000.0} <= set(stores)
    assert all(abs(x - 10000.0) > 1e-6 for x in stores)
    assert steps[0] == 1000.0 and steps[1] == 1500.0


def test_scheduler_errors():
    s = IntegrationScheduler(100.0, mode="count", snapshot_times=[100.0])
    for bad in (0.0, float("inf")):
        with pytest.raises(ValueError, match="dt_cfl"):
            s.next_event(bad)
    s = IntegrationScheduler(2.0, mode="count", snapshot_times=[2.0])
    assert s.next_event(1.0) == ("step", 1.0, 1.0)
    with pytest.raises(FloatingPointError, match="stagnated"):
        s.next_event(math.ulp(1.0) / 2.0)
    with pytest.raises(ValueError):
        integration_plan(10.0, 1.0, mode="bogus")
    with pytest.raises(ValueError):
        integration_plan(10.0, 1.0, mode="interval")


def test_schedule_validation_and_tolerance():
    assert validate_snapshot_schedule([0.0, 5.0, 10.0 + 1e-12], 10.0) == [0.0, 5.0, 10.0]
    for bad in ([0.0, 0.0], [5.0, 4.0], [float("nan")], [11.0]):
        with pytest.raises(ValueError):
            validate_snapshot_schedule(bad, 10.0)
    with pytest.raises(ValueError):
        validate_snapshot_schedule([], 0.0)
    assert scheduler_tolerance(1e4, [0.0, 1e-6]) == 0.25e-6
    assert scheduler_tolerance(100.0, [0.0, 1e-6]) == 1e-9 * 100.0
    assert scheduler_tolerance(0.0, []) == 1e-12


def test_integrate_recomputes_the_ceiling_after_each_step():
    s = IntegrationScheduler(100.0, mode="count", snapshot_times=[0.0, 100.0])
    speeds = iter([1.0, 2.0])
    seen, stores = [], []

    def on_step(t0, t1, dt, k):
        seen.append((t0, t1, dt, k))
        return next(speeds, 0.1)

    t, n = integrate(s, 40.0, 10.0, on_step=on_step, on_store=stores.append)
    # ceilings: 40 (given), then 0.5*10/speed = 5, 2.5, then 50 (clipped to land on 100)
    assert seen[0] == (0.0, 40.0, 40.0, 1) and seen[1][2] == 5.0 and seen[2][2] == 2.5
    assert seen[3] == (47.5, 97.5, 50.0, 4) and seen[4] == (97.5, 100.0, 2.5, 5)
    assert stores == [0.0, 100.0] and t == 100.0 and n == len(seen)
