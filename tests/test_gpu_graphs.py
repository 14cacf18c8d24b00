"""HIP-graph replay of the evaluation (fts_solver_set_graphs): device-input
compute_propagators and compute_concentrations are captured once and replayed. Every
replay must equal an eager evaluation of the same inputs bit for bit, with eager work
and host synchronisation between replays, after the fields' contents change in place,
after their addresses change (capture again) and after the box changes (capture
again)."""
import numpy as np
import pytest

from tests.conftest import golden

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def capi():
    from langevin_fts_amd import capi as c
    return c


@pytest.fixture(scope="module")
def ctx(capi):
    return capi.Context(0)


def outputs(s, names):
    return [s.get_total_partition(0)] + [s.get_total_concentration(m) for m in names]


def same(a, b):
    return a[0] == b[0] and all(np.array_equal(x, y) for x, y in zip(a[1:], b[1:]))


@pytest.mark.parametrize("case", [("diblock_f0.5_ds1_90|discrete|0", 32, "fp64"),
                                  ("diblock_f0.5_ds1_90|discrete|0", 64, "fp32"),
                                  ("grafted|continuous|0", 24, "fp64"),
                                  ("branched|continuous|0", 32, "fp32")])
def test_graph_replays_equal_eager(capi, ctx, case):
    name, n, prec = case
    say(f"graphs {name} {n}^3 {prec}")
    precision = capi.FTS_FP64 if prec == "fp64" else capi.FTS_FP32
    plan = golden("plans")[name]
    inp = plan["input"]
    M = n ** 3
    g = capi.Grid(ctx, [n] * 3, [4.0] * 3)
    P = capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"])
    mons = list(P.monomers)
    eager = capi.Solver(g, P, precision)
    graph = capi.Solver(g, P, precision)
    graph.set_graphs(True)
    rng = np.random.default_rng(n)
    fields = {m: capi.DeviceArray(8 * M) for m in mons}
    qinit = [capi.DeviceArray(8 * M) for _ in P.qinit_names]

    def run(s, dev):
        s.compute_propagators_device([dev[m].ptr for m in mons], [q.ptr for q in qinit])
        s.compute_concentrations()
        return outputs(s, mons)

    for rep in range(4):
        for m in mons:
            fields[m].upload(0.5 * rng.standard_normal(M))
        for q in qinit:
            q.upload(0.5 + rng.random(M))
        ref = run(eager, fields)
        got = run(graph, fields)
        assert same(got, ref), f"replay {rep}"
        eager.compute_statistics({m: rng.standard_normal(M) for m in mons},
                                 {k: 0.5 + rng.random(M) for k in P.qinit_names} if qinit else None)
    # new addresses: captured again
    moved = {m: capi.DeviceArray.from_host(0.3 * rng.standard_normal(M)) for m in mons}
    assert same(run(graph, moved), run(eager, moved))
    # a new box: the volume is a kernel argument, captured again
    g.set_lx([4.4, 4.0, 3.6])
    eager.update_laplacian()
    graph.update_laplacian()
    assert same(run(graph, moved), run(eager, moved))
    assert same(run(graph, fields), run(eager, fields))


def test_graphs_with_kernel_timing_run_eagerly(capi, ctx):
    plan = golden("plans")["diblock_f0.5_ds1_90|discrete|0"]
    inp = plan["input"]
    M = 32 ** 3
    g = capi.Grid(ctx, [32] * 3, [4.0] * 3)
    s = capi.Solver(g, capi.Plan(plan, inp["ds"], inp["bonds"], plan["schedules"]["4"]), capi.FTS_FP64)
    s.set_graphs(True)
    a = capi.DeviceArray.from_host(np.linspace(-1, 1, M))
    b = capi.DeviceArray.from_host(np.linspace(1, -1, M))
    s.compute_propagators_device([a.ptr, b.ptr])
    s.compute_concentrations()
    ref = outputs(s, ["A", "B"])
    ctx.set_kernel_timing(4)
    try:
        s.compute_propagators_device([a.ptr, b.ptr])
        s.compute_concentrations()
        assert same(outputs(s, ["A", "B"]), ref)
    finally:
        ctx.set_kernel_timing(0)
