"""hipGraph replay of the projected-gradient rounds (SCOPI_TUNE_GRAPH_ROUNDS): the captured
rounds are the eager loop's kernels in the eager loop's order, replayed several times per solve
with eager work (search, assembly, integration) between the solves, so the trajectory is
bitwise the eager one."""
import numpy as np
import pytest

from scopi_amd import hip, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [hip.FP64, hip.FP32_REFINE], ids=["fp64", "fp32_refine"])
@pytest.mark.parametrize("rounds", [1, 15])
def test_graph_rounds_bitwise_eager(precision, rounds):
    sc = scenes.spheres_3d_count(3000, seed=7, exact=True)
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6, precision=precision)
    print(f"[graph] precision={precision} rounds={rounds}", flush=True)
    a = hip.HipScene(sc)
    ia = a.run(prm, 1e-3, 4)
    b = hip.HipScene(sc)
    b.set_tuning(hip.TUNE_GRAPH_ROUNDS, rounds)
    ib = b.run(prm, 1e-3, 4)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    for k in ("iterations", "backtracks", "n_contacts", "converged"):
        assert ia[k] == ib[k], k
    ua, ub = a.solution(), b.solution()
    for k in ua:
        assert np.array_equal(ua[k], ub[k]), k


def test_graph_rounds_must_be_odd():
    g = hip.HipScene(scenes.spheres_3d_count(64, seed=1, exact=True))
    with pytest.raises(hip.ScopiHipError):
        g.set_tuning(hip.TUNE_GRAPH_ROUNDS, 16)


def test_graph_kept_across_solves_and_shapes():
    """The executable graph is kept across solves and updated in place; a change of the round count (another
    topology) instantiates a new one.  Steps with 15-round graphs, then 3-round graphs, then eager, against
    the eager loop throughout."""
    sc = scenes.spheres_3d_count(2000, seed=9, exact=True)
    prm = hip.params(vap=hip.VAP_FPD, dmax=0.02, kd_tree_radius=0.042 ** 2, tolerance=1e-6,
                     precision=hip.FP32_REFINE)
    a = hip.HipScene(sc)
    a.run(prm, 1e-3, 7)
    b = hip.HipScene(sc)
    for rounds, steps in ((15, 3), (3, 2), (0, 2)):
        b.set_tuning(hip.TUNE_GRAPH_ROUNDS, rounds)
        b.run(prm, 1e-3, steps)
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
