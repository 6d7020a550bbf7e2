"""Batches: independent Riemann shock tubes on Cartesian meshes of uneven sizes, advanced
together by one handle (teal_gpu_create_batch), each with its own time step.

Every instance must match the oracle stepping that mesh alone: bit for bit in FP64 (state,
time and residual to 1e-12), within a stated bound in the mixed path.  Graph replay and the
autotuner's launch shapes must reproduce eager launches bit for bit, replay after replay."""
import ctypes as C

import numpy as np
import pytest

import teal_amd
from oracle_lib import Case, dp, oracle, oracle_eqns, oracle_mesh
from teal_amd import _abi

pytestmark = pytest.mark.gpu

BCS = [("left", "supersonic outflow", None), ("right", "supersonic outflow", None)] + [
    (side, "symmetry", None) for side in ("bottom", "top", "back", "front")]
# sizes chosen so no instance fills whole 256-cell blocks (padding rows in every instance)
SIZES = [(5, 3, 3), (7, 4, 4), (9, 3, 5), (4, 4, 4), (11, 6, 5)]


def tube_case(n):
    return Case(n, periodic=(0, 0, 0), flux="hllc", limiter="minmod", bcs=BCS)


def tube_state(center, ni, nc, seed):
    """Seeded left/right states (rho, p in [0.1, 1], u in [-1, 1]) split at x = 0.5."""
    rng = np.random.default_rng(seed)
    left = np.array([rng.uniform(0.1, 1), rng.uniform(-1, 1), 0, 0, rng.uniform(0.1, 1)])
    right = np.array([rng.uniform(0.1, 1), rng.uniform(-1, 1), 0, 0, rng.uniform(0.1, 1)])
    P = np.where((center[:nc, 0] < 0.5)[:, None], left, right)
    P[ni:] = 0
    return P


def build(sizes, precision=teal_amd.FP64, seed=11):
    lib = teal_amd.lib()
    meshes, states, oracles = [], [], []
    for i, n in enumerate(sizes):
        case = tube_case(n)
        meshes.append(teal_amd.mesh_box(n, periodic=case.periodic))
        om = oracle_mesh(case)
        oe = oracle_eqns(case, om)
        a = _abi.mesh_arrays(om)
        P = tube_state(a["center"], a["num_inner"], a["num_cells"], seed + i)
        states.append(P)
        oracles.append((om, oe, a["num_cells"], a["num_inner"]))
    batch = teal_amd.Batch(meshes, precision)
    batch.configure(space_order=2, flux="hllc", limiter="minmod",
                    bcs={ent: (name, ctx) for ent, name, ctx in BCS})
    return lib, meshes, states, oracles, batch


def oracle_run(om, oe, nc, P, steps, courant=0.5, t0=0.0):
    po = np.ctypeslib.as_array(oracle().oracle_eqns_primitive(oe), shape=(nc * 5,))
    po[:] = P.ravel()
    t = C.c_double(t0)
    res = (C.c_double * 5)()
    for _ in range(steps):
        oracle().oracle_lserk(oe, C.byref(t), res, 1e300, courant, 2, 3)
    return po.reshape(-1, 5).copy(), t.value, np.array(res[:])


def test_batch_layout():
    print("batch layout", flush=True)
    _, meshes, _, _, batch = build(SIZES)
    assert batch.lib.teal_gpu_batch_size(batch.h) == len(SIZES)
    info = batch.info()
    inner = sum(np.prod(n) for n in SIZES)
    assert batch.arrays["num_inner"] == inner
    padded = sum(-(-int(np.prod(n)) // 256) * 256 for n in SIZES)
    assert info["num_inner"] == padded  # device rows: each instance padded to whole blocks
    assert info["fused"] == 1
    batch.close()


def test_batch_fp64_matches_each_instance_alone():
    """Each instance starts at its own time; times advance by the instance's own steps."""
    steps = 12
    print(f"batch fp64 {len(SIZES)} tubes {steps} steps", flush=True)
    _, _, states, oracles, batch = build(SIZES)
    batch.upload_batch(states)
    t0 = 0.25 * np.arange(len(SIZES))
    times, res = batch.run_lserk_batch(steps, t0, 0.5)
    got = batch.download_batch()
    for i, ((om, oe, nc, ni), P) in enumerate(zip(oracles, states)):
        Po, t_o, r_o = oracle_run(om, oe, nc, P, steps, t0=t0[i])
        assert times[i] == t_o, (i, times[i], t_o)
        np.testing.assert_allclose(res[i], r_o, rtol=1e-12, atol=1e-300)
        np.testing.assert_array_equal(got[i][:ni], Po[:ni])
    batch.close()


def test_batch_of_tets_and_hexes():
    """Instances of different cell types in one batch (jittered tetrahedra next to hexahedra):
    each matches the oracle stepping it alone, bit for bit, over 8 steps."""
    steps = 8
    print(f"batch tets + hexes {steps} steps", flush=True)
    lib = teal_amd.lib()
    o = oracle()
    o.oracle_mesh_copy_created.restype = C.c_void_p
    o.oracle_mesh_copy_created.argtypes = [C.c_void_p]
    meshes, states, oracles = [], [], []
    for i, (kind, n) in enumerate([("tets", 4), ("hex", (5, 3, 3)), ("tets", 3), ("hex", (4, 4, 6))]):
        case = tube_case(n if kind == "hex" else (n, n, n))
        if kind == "tets":
            pm = lib.mesh_create_tets(n, 0.2, 100 + i)
            om = o.oracle_mesh_copy_created(pm)
            o.oracle_mesh_generate(om)
            lib.mesh_generate(pm)
        else:
            pm = teal_amd.mesh_box(n, periodic=case.periodic)
            om = oracle_mesh(case)
        oe = oracle_eqns(case, om)
        a = _abi.mesh_arrays(om)
        meshes.append(pm)
        states.append(tube_state(a["center"], a["num_inner"], a["num_cells"], 50 + i))
        oracles.append((om, oe, a["num_cells"], a["num_inner"]))
    batch = teal_amd.Batch(meshes, teal_amd.FP64)
    batch.configure(space_order=2, flux="hllc", limiter="minmod",
                    bcs={ent: (name, ctx) for ent, name, ctx in BCS})
    batch.upload_batch(states)
    times, res = batch.run_lserk_batch(steps, np.zeros(len(meshes)), 0.5)
    got = batch.download_batch()
    for i, ((om, oe, nc, ni), P) in enumerate(zip(oracles, states)):
        Po, t_o, r_o = oracle_run(om, oe, nc, P, steps)
        assert times[i] == t_o, (i, times[i], t_o)
        np.testing.assert_allclose(res[i], r_o, rtol=1e-12, atol=1e-300)
        np.testing.assert_array_equal(got[i][:ni], Po[:ni])
    batch.close()


def test_batch_of_one_equals_single_handle():
    """A batch of one mesh has no padding and steps exactly like a single-mesh handle."""
    print("batch of one", flush=True)
    _, meshes, states, _, batch = build(SIZES[4:5])
    single = teal_amd.Gpu(meshes[0])
    single.configure(space_order=2, flux="hllc", limiter="minmod",
                     bcs={ent: (name, ctx) for ent, name, ctx in BCS})
    assert batch.info()["num_inner"] == single.info()["num_inner"]
    batch.upload_batch(states)
    single.upload(states[0])
    times, _ = batch.run_lserk_batch(7, np.zeros(1), 0.5)
    t = single.run_lserk(7, 0.0, 0.5)
    assert times[0] == t
    np.testing.assert_array_equal(batch.download_batch()[0], single.download())
    batch.close()
    single.close()


def test_batch_derivative_and_residuals():
    """The per-call routines on a batch handle: the derivative of the union (host layout of
    teal_mesh_union) equals each instance's oracle derivative bit for bit, and
    teal_gpu_residual gives one residual norm per instance (rtol 1e-12)."""
    print("batch derivative + residuals", flush=True)
    _, _, states, oracles, batch = build(SIZES)
    batch.upload_batch(states)
    union = batch.download()  # the union's host layout
    R, _ = batch.derivative(union)
    res = batch.residual(R)
    assert res.shape == (len(SIZES), 5)
    row = 0
    for i, ((om, oe, nc, ni), P) in enumerate(zip(oracles, states)):
        Ro = np.zeros((ni, 5))
        o = oracle()
        o.oracle_derivative(oe, dp(P.copy()), dp(Ro))
        np.testing.assert_array_equal(R[row:row + ni], Ro)
        ro = np.zeros(5)
        o.oracle_residual(oe, dp(Ro), dp(ro))
        np.testing.assert_allclose(res[i], ro, rtol=1e-12, atol=1e-300)
        row += ni
    batch.close()


def test_batch_union_upload_layout():
    """teal_gpu_upload of the union layout and teal_gpu_upload_batch agree."""
    print("batch union upload", flush=True)
    _, _, states, _, batch = build(SIZES[:3])
    batch.upload_batch(states)
    union = batch.download()
    batch.upload(union)
    back = batch.download_batch()
    for P, Q, a in zip(states, back, batch.inst_arrays):
        np.testing.assert_array_equal(Q[:a["num_inner"]], P[:a["num_inner"]])
    batch.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_graph_replays_match_eager(precision):
    """Graph replay (two steps per launch) against eager launches, replay after replay, with
    the program's own downloads and eager work between replays: bit for bit in both paths."""
    print(f"batch graph replays {precision}", flush=True)
    prec = teal_amd.FP64 if precision == "fp64" else teal_amd.MIXED
    _, _, states, _, eager = build(SIZES, prec)
    _, _, _, _, graph = build(SIZES, prec)
    graph.set_graphs(True)
    eager.upload_batch(states)
    graph.upload_batch(states)
    t_e = np.zeros(len(SIZES))
    t_g = np.zeros(len(SIZES))
    for k, steps in enumerate((5, 4, 6, 1, 4)):
        t_e, r_e = eager.run_lserk_batch(steps, t_e, 0.5)
        t_g, r_g = graph.run_lserk_batch(steps, t_g, 0.5)
        np.testing.assert_array_equal(t_g, t_e)
        np.testing.assert_array_equal(r_g, r_e)
        for a, b in zip(graph.download_batch(), eager.download_batch()):
            np.testing.assert_array_equal(a, b)
        print(f"  replay round {k}: {steps} steps equal", flush=True)
    eager.close()
    graph.close()


def test_single_mesh_graph_matches_eager():
    """teal_gpu_run_lserk with graphs on one mesh (TGV 8^3, Navier-Stokes) against eager."""
    print("single graph", flush=True)
    from cases import initial_state
    case = Case((8, 8, 8), lo=(0, 0, 0), hi=(2 * np.pi,) * 3, kind="navierstokes", mu=1e-3)
    gs = []
    for graphs in (False, True):
        g = teal_amd.Gpu(teal_amd.mesh_box(case.n, case.lo, case.hi, case.periodic))
        g.configure(space_order=2, flux="hllc", viscous=1, limiter="minmod", navierstokes=1,
                    mu=case.mu)
        g.set_graphs(graphs)
        gs.append(g)
    a = gs[0].arrays
    P = initial_state(case, a, "tgv")
    t = [0.0, 0.0]
    for k, g in enumerate(gs):
        g.upload(P)
    for steps in (7, 2, 3):
        for k, g in enumerate(gs):
            t[k] = g.run_lserk(steps, t[k], 0.5)
        assert t[0] == t[1]
        np.testing.assert_array_equal(gs[1].download(), gs[0].download())
    for g in gs:
        g.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_autotune_keeps_state_and_results(precision):
    """The autotuner restores the state it started from; its pick gives the same steps as the
    default shape: bit for bit in FP64; in the mixed path the fused kernel rebuilds face
    tangents and corrections the split kernels read from the stored geometry, so the two
    agree to FP32 rounding (state within 1e-6 of max |P|, times within 1e-6 relative)."""
    print(f"batch autotune {precision}", flush=True)
    prec = teal_amd.FP64 if precision == "fp64" else teal_amd.MIXED
    _, _, states, _, ref = build(SIZES, prec)
    _, _, _, _, tuned = build(SIZES, prec)
    ref.upload_batch(states)
    tuned.upload_batch(states)
    ms = tuned.autotune(4, 0.5)
    print(f"  autotune ms/step {dict(zip(teal_amd.AUTOTUNE_SHAPES, ms.round(4)))}", flush=True)
    assert np.isfinite(ms).all() and (ms > 0).all()
    print(f"  picked: {tuned.info()}", flush=True)
    for a, b in zip(tuned.download_batch(), ref.download_batch()):
        np.testing.assert_array_equal(a, b)
    t_r, r_r = ref.run_lserk_batch(6, np.zeros(len(SIZES)), 0.5)
    t_t, r_t = tuned.run_lserk_batch(6, np.zeros(len(SIZES)), 0.5)
    got, want = tuned.download_batch(), ref.download_batch()
    if precision == "fp64":
        np.testing.assert_array_equal(t_t, t_r)
        np.testing.assert_array_equal(r_t, r_r)
        for a, b in zip(got, want):
            np.testing.assert_array_equal(a, b)
    else:
        np.testing.assert_allclose(t_t, t_r, rtol=1e-6)
        for a, b, ia in zip(got, want, tuned.inst_arrays):
            ni = ia["num_inner"]
            assert np.abs(a[:ni] - b[:ni]).max() <= 1e-6 * np.abs(b[:ni]).max()
    ref.close()
    tuned.close()


def test_batch_mixed_error_bound():
    """Mixed batch after 10 steps, each instance with the time step of its own FP32 state:
    per instance, time within 1e-6 relative and state error / max |P| <= 1e-5 against the
    oracle stepping that instance alone."""
    steps = 10
    print(f"batch mixed {steps} steps", flush=True)
    _, _, states, oracles, batch = build(SIZES, teal_amd.MIXED)
    batch.upload_batch(states)
    times, _ = batch.run_lserk_batch(steps, np.zeros(len(SIZES)), 0.5)
    got = batch.download_batch()
    for i, ((om, oe, nc, ni), P) in enumerate(zip(oracles, states)):
        Po, t_o, _ = oracle_run(om, oe, nc, P, steps)
        assert abs(times[i] - t_o) <= 1e-6 * t_o, (i, times[i], t_o)
        err = np.abs(got[i][:ni] - Po[:ni]).max() / np.abs(Po[:ni]).max()
        print(f"  instance {i}: time rel diff {abs(times[i] - t_o) / t_o:.2e}, "
              f"state error / max |P| {err:.2e}", flush=True)
        assert err <= 1e-5, (i, err)
    batch.close()


def test_batch_refuses_single_mesh_steps():
    print("batch refusals", flush=True)
    _, _, states, _, batch = build(SIZES[:2])
    batch.upload_batch(states)
    with pytest.raises(RuntimeError, match="run_lserk_batch"):
        batch.lserk(0.0, 0.5)
    with pytest.raises(RuntimeError, match="batch"):
        batch.implicit_euler(0.0, 1.0)
    batch.close()


@pytest.mark.parametrize("shape", teal_amd.AUTOTUNE_STAGE_SHAPES)
def test_every_shape_steps_alike(shape):
    """Each stage shape (fused 256- or 128-row block layouts, split; eager or graph replay) and
    each gradient workgroup size, with direct or LDS-staged gathers, gives the default shape's
    steps bit for bit in FP64: state, time and residual (the residual partial sums are kept
    per 64 rows, whatever the block), on the batch and on one Navier-Stokes Taylor-Green
    mesh."""
    print(f"shape {shape}", flush=True)
    _, _, states, _, ref = build(SIZES)
    ref.upload_batch(states)
    t_r, r_r = ref.run_lserk_batch(5, np.zeros(len(SIZES)), 0.5)
    want = ref.download_batch()
    ref.close()
    for gb, st in ((256, 0), (128, 1), (64, 0), (256, 1)):
        _, _, _, _, b = build(SIZES)
        b.set_shape(shape, gb, st)
        b.upload_batch(states)
        t_g, r_g = b.run_lserk_batch(5, np.zeros(len(SIZES)), 0.5)
        np.testing.assert_array_equal(t_g, t_r)
        np.testing.assert_array_equal(r_g, r_r)
        for got, w in zip(b.download_batch(), want):
            np.testing.assert_array_equal(got, w)
        b.close()
    from oracle_lib import tgv_state
    n = (12, 10, 9)
    P = None
    out = []
    for sh, gb, st in (("fused256", 256, 0), (shape, 128, 1)):
        m = teal_amd.mesh_box(n, (0, 0, 0), (2 * np.pi,) * 3, (1, 1, 1))
        g = teal_amd.Gpu(m).configure(flux="hllc", viscous=1, limiter="minmod", navierstokes=1,
                                      mu=1 / 1600)
        if P is None:
            a = g.arrays
            P = tgv_state(a["center"][:a["num_cells"]])
            P[a["num_inner"]:] = 0
        g.set_shape(sh, gb, st)
        g.upload(P)
        t = g.run_lserk(4, 0.0, 0.5)
        out.append((t, g.download()))
        g.close()
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_instances_of_several_chunks_match_single_handles():
    """The step reductions cut each instance's groups of 64 rows into chunks (2048 groups for
    the residual, 4096 for the time step) and combine them in chunk order.  Instances of 1, 2
    and 3 residual chunks (8,000 to 287,496 cells), the largest not starting on a chunk
    boundary, step exactly as single-mesh handles: time, residual and state bit for bit."""
    print("batch of multi-chunk instances", flush=True)
    sizes = [(20, 20, 20), (66, 66, 66), (54, 53, 52)]
    meshes = [teal_amd.mesh_box(n, periodic=(0, 0, 0)) for n in sizes]
    bcs = {ent: (name, ctx) for ent, name, ctx in BCS}
    states = []
    for i, m in enumerate(meshes):
        a = _abi.mesh_arrays(m)
        states.append(tube_state(a["center"], a["num_inner"], a["num_cells"], 40 + i))
    batch = teal_amd.Batch(meshes, teal_amd.FP64)
    batch.configure(space_order=2, flux="hllc", limiter="minmod", bcs=bcs)
    batch.upload_batch(states)
    times, res = batch.run_lserk_batch(4, np.zeros(len(sizes)), 0.5)
    got = batch.download_batch()
    batch.close()
    for i, m in enumerate(meshes):
        single = teal_amd.Gpu(m)
        single.configure(space_order=2, flux="hllc", limiter="minmod", bcs=bcs)
        single.upload(states[i])
        t = 0.0
        for _ in range(4):
            _, t, r = single.lserk(t, 0.5)
        assert times[i] == t
        np.testing.assert_array_equal(res[i], r)
        np.testing.assert_array_equal(got[i], single.download())
        single.close()
