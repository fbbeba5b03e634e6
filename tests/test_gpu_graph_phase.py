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
