"""Graph replay of steady steps after the rings moved: a graph is captured
per ring phase, so steps taken outside it (an eager step here) do not make a
later replay use another phase's slots."""

import numpy as np
import pytest

from jaxfun_amd import PlaneCouetteFluctuationHIP
from jaxfun_amd.ic import seeded_state

pytestmark = pytest.mark.gpu


def test_graph_replay_after_phase_shift_matches_eager():
    a = PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                   nonlinear_form="rotational")
    b = PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3, time_integrator="SBDF3",
                                   nonlinear_form="rotational")
    s = seeded_state(a, seed=6, energy=5e-2, K=4)
    for x in (a, b):
        x.load(s)
        x.advance(2)  # start-up
    a.advance_steady(12, graph_block=12)
    b.advance_steady(12)
    for x in (a, b):
        x.advance(1)  # one step outside the graph: the rings stand one phase on
    a.advance_steady(24, graph_block=12)
    b.advance_steady(24)
    sa, sb = a.fetch(), b.fetch()
    for x, y in zip(list(sa.u) + [sa.g] + list(sa.nonlinear_old) + list(sa.solution_old),
                    list(sb.u) + [sb.g] + list(sb.nonlinear_old) + list(sb.solution_old)):
        err = float(np.max(np.abs(x - y))) / max(float(np.max(np.abs(y))), 1e-300)
        assert err < 1e-13, err


def _leaves(s):
    out = list(s.u) + [s.g]
    for grp in (s.nonlinear_old, s.nonlinear_older, s.solution_old, s.solution_older):
        out.extend(grp)
    return out


def test_graphs_of_any_length_and_history_scalars_match_eager():
    """Graphs of 5, 7 and 12 steps (cached per ring phase and length, some
    replayed from the cache after the rings came back to their phase) give the
    eager steps bit for bit, and every replay leaves the history scalars a
    steady step leaves: a state loaded with have_old = 0 and previous_dt != dt
    fetches back with have_old = 1 and previous_dt = dt."""
    from dataclasses import replace

    mk = lambda: PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3,  # noqa: E731
                                            time_integrator="SBDF3", nonlinear_form="rotational")
    a, b = mk(), mk()
    s = seeded_state(a, seed=9, energy=5e-2, K=4)
    a.load(s)
    a.advance(3)
    s3 = a.fetch()
    odd = replace(s3, have_old=0.0, previous_dt=7e-3)  # steady steps read neither
    for x in (a, b):
        x.load(odd)
    plan = (5, 7, 12, 5, 7, 12, 24)  # the second 5 and 7 start at the phases of the first
    for n in plan:
        a.advance_steady(n, graph_block=n)
        b.advance_steady(n)
        ha, hb = a.fetch(), b.fetch()
        assert (ha.history_steps, ha.have_old, ha.previous_dt) == (2.0, 1.0, 2e-3)
        assert (hb.history_steps, hb.have_old, hb.previous_dt) == (2.0, 1.0, 2e-3)
        for x, y in zip(_leaves(ha), _leaves(hb)):
            assert np.array_equal(x, y), n
        a.load(ha)
        b.load(hb)


def test_prepared_graph_runs_nothing_until_replayed():
    """prepare_steady captures without running: the state and the history
    scalars stay as they were, and the replay that follows equals eager steps."""
    mk = lambda: PlaneCouetteFluctuationHIP(N=(17, 16, 16), Re=400.0, dt=2e-3,  # noqa: E731
                                            time_integrator="SBDF3", nonlinear_form="rotational")
    a, b = mk(), mk()
    s = seeded_state(a, seed=10, energy=5e-2, K=4)
    for x in (a, b):
        x.load(s)
        x.advance(4)
    before = a.fetch()
    a.load(before)
    a.prepare_steady(31, 12)
    after = a.fetch()
    assert all(np.array_equal(x, y) for x, y in zip(_leaves(before), _leaves(after)))
    assert (after.history_steps, after.have_old, after.previous_dt) == (before.history_steps, before.have_old,
                                                                        before.previous_dt)
    a.load(after)
    a.prepare_steady(31, 12)
    a.advance_steady(31, graph_block=12)
    b.advance_steady(31)
    assert all(np.array_equal(x, y) for x, y in zip(_leaves(a.fetch()), _leaves(b.fetch())))
