"""Whole steps captured in hipGraphs (mom_gpu_graph_*).

A graph of two steps is replayed several times; after every replay the
state equals, bit for bit, the same number of eagerly launched steps from
the same initial state.  Eager work and host synchronisations (content sums,
copies) run between the replays.
"""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

REPLAYS = 3


def _same(a, b):
    return np.array_equal(a.to_host(), b.to_host())


@pytest.mark.parametrize("streams", [1, 2])
def test_batched_mixed_step_graph_replays_equal_eager(streams):
    """streams = 2: the FP32 tracers of each captured step run on the
    context's second stream (graph branches joined at the end)."""
    from mom5_amd.step import BatchedStep
    from mom5_amd.synth import make_case_device
    case = make_case_device(150, 90, 14, seed=9, nts=10, n_passive=3, stf_scale=1e-3)
    print(f"[graph: batched mixed 150x90x14, 2 FP64 + 3 FP32, barotropic nts=10, "
          f"streams={streams}]", flush=True)
    g = case.info["gpu"]
    g.set_tuning("step_streams", streams)
    dt = [np.float64, np.float64] + [np.float32] * 3

    def fresh():
        return dataclasses.replace(case, tracers=[
            (nm, case.info["tracer"](i, dt[i])) for i, (nm, _) in enumerate(case.tracers)])

    eager, graphed = BatchedStep(fresh(), gpu=g), BatchedStep(fresh(), gpu=g)
    eager.run(1)
    graphed.run(1)                  # first FP32 step allocates the solver scratch
    graph = graphed.capture(2)
    assert graph.nodes > 0
    for r in range(REPLAYS):
        graph.launch()
        eager.run(2)
        g.sync()
        rho = eager.dev["rho_dzt"]
        for n in range(eager.ntr):
            assert g.tracer_content(eager.tracer(n), rho) == g.tracer_content(graphed.tracer(n), rho)
            assert _same(eager.tracer(n), graphed.tracer(n)), (r, n)
        assert _same(eager.eta_bar[eager.cur], graphed.eta_bar[graphed.cur]), r
        assert _same(eager.udrho[eager.cur], graphed.udrho[graphed.cur]), r
    graph.close()


def test_plain_step_graph_replays_equal_eager():
    from mom5_amd.step import PlainStep
    from mom5_amd.synth import make_case
    case = make_case(64, 40, 10, land_frac=0.3, nts=12, seed=17, n_passive=1, stf_scale=1e-3)
    print("[graph: plain FP64 64x40x10, MDFL, MDFL, Quicker, barotropic nts=12]", flush=True)
    eager = PlainStep(case)
    graphed = PlainStep(case, gpu=eager.gpu)
    g = eager.gpu
    graph = graphed.capture(2)
    for r in range(REPLAYS):
        graph.launch()
        eager.run(2)
        g.sync()
        for n in range(eager.ntr):
            assert _same(eager.tracer(n), graphed.tracer(n)), (r, n)
        assert _same(eager.eta_bar[eager.cur], graphed.eta_bar[graphed.cur]), r
    graph.close()


def test_capture_rejects_host_synchronisation():
    from mom5_amd import MomGpuError
    from mom5_amd.synth import make_case_device
    case = make_case_device(40, 30, 6, seed=3, nts=4)
    g = case.info["gpu"]
    g.graph_begin()
    with pytest.raises(MomGpuError):
        g.tracer_content(case.tracers[0][1], case.fields["rho_dzt"])
        g.graph_end()
    # the context stays usable after a failed capture
    try:
        g.graph_end()
    except MomGpuError:
        pass
    assert g.tracer_content(case.tracers[0][1], case.fields["rho_dzt"]) > 0.0
