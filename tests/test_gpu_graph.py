"""Graph mode (cfg.graph = 1): each stage's fixed head (stage operands, explicit part, preparation, shock budget) and each
Krylov chunk run as captured hipGraphs, re-captured per launch and applied to the instantiated graph of the same shape.
The kernels are the eager path's, so whole runs are bit-identical to it (and so to the oracle), in the plain FP64 form
and in the FP32-Krylov form (where a chunk also carries the defect, the sweep's close and the trial residual); the graphs
are instantiated once per shape and updated in place afterwards."""
import pytest

from tests.helpers import bits_diff, field_l1

pytestmark = pytest.mark.gpu


def _cases():
    from diffusionimex_amd import cases
    return [("bubble64x32", lambda **kw: cases.shock_bubble(64, 32, seed=3, cfl=0.5, **kw), 25),
            ("tube200", lambda **kw: cases.water_air_tube(200, cfl=0.5, newton_tol=1e-10, **kw), 30)]


@pytest.mark.parametrize("form", ["plain", "mixed"])
@pytest.mark.parametrize("idx", [0, 1])
def test_graph_mode_bitwise_equal_to_eager(idx, form):
    from diffusionimex_amd import spb6
    name, mk, steps = _cases()[idx]
    kw = dict(krylov_fp32=1, analytic_jacobian=1) if form == "mixed" else {}
    out = []
    for graph in (0, 1):
        cfg, U = mk(graph=graph, **kw)
        st = spb6.Stepper(cfg)
        st.set_state(U)
        st.enable_timing(True)
        assert st.run(1.0, steps) == 0
        out.append((st.state(), st.pressure()[0], st.stats(), st.kernel_times()))
    (U0, P0, s0, t0), (U1, P1, s1, t1) = out
    for k in ("steps", "rk1", "rejects", "stages", "newton_it", "krylov_it", "floor_acc", "linsolves"):
        assert s0[k] == s1[k], f"{name}/{form}: {k} {s0[k]} vs {s1[k]}"
    assert bits_diff(U1, U0) == 0, f"{name}/{form}: state differs in {bits_diff(U1, U0)} values; {field_l1(U1, U0)}"
    assert bits_diff(P1, P0) == 0
    assert t0["graph_launches"] == 0
    if cfg.ny > 1 or form == "plain":   # the 1D tube's FP32 form solves with Thomas: heads only
        assert t1["graph_launches"] >= s1["stages"]
    assert t1["graph_rebuilds"] <= 40, t1["graph_rebuilds"]
