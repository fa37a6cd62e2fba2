"""HIP path (libtomo.so through the C ABI) against the oracle.

Tolerances (written here):
  * material maps, coefficient tables, contour indexing: bit-exact;
  * fields after a few hundred steps: relative L2 <= 1e-5 against the FP32
    oracle's restatement of the reference, <= 1e-4 against the FP64 oracle;
  * running-DFT accumulators, far field and fitness: <= 1e-4 relative to FP64.
"""
import numpy as np
import pytest

import tomography_amd as t
from cases import (CYL, CYL_PROBES, CYL_REF, DISC_PROBES, FAT, PHYS_PARAMS, cylinder_field, fat_medium,
                   fat_with_cylinder, line_source_ratio, phys_small, ref_cpu_params, ref_cpu_small, ref_debye_disc,
                   ref_params, ref_small, rel_l2, skin_cylinder, spectrum)

pytestmark = pytest.mark.gpu

FIELD_TOL_F32 = 1e-5
FIELD_TOL_F64 = 1e-4
FF_TOL = 1e-4


def oracle_fields(oracle, cfg, p, exc, steps, fp64):
    s = oracle.sim(cfg, p, exc=exc, fp64=fp64)
    s.step(steps)
    return s


def test_ref_materials_and_coefficients_bit_exact(ctx, oracle):
    cfg = ref_small()
    p = ref_params(11)
    g = ctx.sim(cfg, p, exc=0)
    mat, coef = g.materials()
    om, oc = oracle.sim(cfg, p, exc=0).materials()
    assert np.array_equal(mat, om)
    for m in np.unique(mat):
        sel = om == m
        # oracle planes: Ceze, Cezhy, Cezhx, Cezeic, Cezeip, Cezjp, Beta_p, Beta_p/dt, Kp
        want = np.array([oc[0][sel][0], oc[1][sel][0], oc[3][sel][0], oc[4][sel][0], oc[5][sel][0],
                         oc[8][sel][0], oc[7][sel][0]], np.float32)
        assert np.array_equal(coef[m, :7], want), m


def test_phys_materials_bit_exact(ctx, oracle):
    cfg = phys_small()
    for pv in (PHYS_PARAMS, np.array([-6.0, 5.0, 5.5, 25.0, 40.0, 10.0, 0.5, 3.0, 4.0, 1.0, 2.5])):
        g = ctx.sim(cfg, pv, exc=1)
        mat, coef = g.materials()
        om, oc = oracle.sim(cfg, pv, exc=1).materials()
        assert np.array_equal(mat, om)
        assert len(np.unique(mat)) >= 4
        for m in np.unique(mat):
            sel = om == m
            want = np.array([oc[0][sel][0], oc[1][sel][0], oc[5][sel][0], oc[8][sel][0], oc[7][sel][0]],
                            np.float32)
            got = coef[m, [0, 1, 4, 5, 6]]
            assert np.array_equal(got, want), (m, got, want)


@pytest.mark.parametrize("exc", [0, 1])
def test_ref_fields_vs_oracle(ctx, oracle, exc):
    cfg = ref_small(steps=300)
    p = ref_params(5)
    g = ctx.sim(cfg, p, exc=exc)
    g.step(300)
    ez, hx, hy, jp = g.fields()
    o32 = oracle_fields(oracle, cfg, p, exc, 300, False).fields()
    o64 = oracle_fields(oracle, cfg, p, exc, 300, True).fields()
    assert np.abs(o64[0]).max() > 1e-3  # the pulse has reached the phantom
    for name, a, b32, b64 in zip(("ez", "hx", "hy", "jp"), (ez, hx, hy, jp), o32, o64):
        assert rel_l2(a, b32) < FIELD_TOL_F32, (name, rel_l2(a, b32))
        assert rel_l2(a, b64) < FIELD_TOL_F64, (name, rel_l2(a, b64))


def test_phys_fields_vs_oracle(ctx, oracle):
    cfg = phys_small(steps=400)
    g = ctx.sim(cfg, PHYS_PARAMS, exc=1)
    g.step(400)
    got = g.fields()
    o32 = oracle_fields(oracle, cfg, PHYS_PARAMS, 1, 400, False).fields()
    o64 = oracle_fields(oracle, cfg, PHYS_PARAMS, 1, 400, True).fields()
    assert np.abs(o64[0]).max() > 1e-3
    for name, a, b32, b64 in zip(("ez", "hx", "hy", "jp"), got, o32, o64):
        assert rel_l2(a, b32) < FIELD_TOL_F32, (name, rel_l2(a, b32))
        assert rel_l2(a, b64) < FIELD_TOL_F64, (name, rel_l2(a, b64))


@pytest.mark.parametrize("n, steps, sched", [(256, 300, (4, 4, 1, 60, 1, 16, t.FORM_FUSED, 1, 10)),
                                             (512, 700, (6, 48, 1, 60, 4, 16, t.FORM_STREAMS, 24)),
                                             (512, 700, (8, 12, 1, 64, 2, 16, t.FORM_STREAMS, 4, 40))])
def test_placed_interior_bands_vs_oracle(oracle, n, steps, sched):
    """The full interior bands need not start at the first clean column: the
    tile builder moves them right so that the column-edge regions on both
    sides fit in one band each (256^2 at T = 4: 33 and 56 columns instead of
    21 and 68; 512^2 at T = 6: 45 and 52 instead of 23 and 74; at T = 8: 33
    and 48 instead of 25 and 56).  The fields must still match the oracle and
    equal those of the streams form at T = 4 with 32-row tiles (whose bands
    stay in place on 512^2: 21 and 44 columns).  By the last step the pulse
    has reached both column-edge regions (checked on the oracle's Ez)."""
    cfg = phys_small(steps=steps, n=n, n_exc=1)
    with t.Context() as ctx:
        ctx.set_schedule(*sched)
        g = ctx.sim(cfg, PHYS_PARAMS, exc=0)
        g.step(steps)
        got = g.fields()
    with t.Context() as base:
        base.set_schedule(4, 32, 4, 32, form=t.FORM_STREAMS)
        b = base.sim(cfg, PHYS_PARAMS, exc=0)
        b.step(steps)
        ref = b.fields()
    o32 = oracle_fields(oracle, cfg, PHYS_PARAMS, 0, steps, False).fields()
    assert np.abs(o32[0][:, :60]).max() > 1e-7 and np.abs(o32[0][:, -60:]).max() > 1e-4
    for name, a, r, o in zip(("ez", "hx", "hy", "jp"), got, ref, o32):
        assert np.array_equal(a, r), name
        assert rel_l2(a, o) < FIELD_TOL_F32, (name, rel_l2(a, o))


@pytest.mark.parametrize("sched", [(1, 32, 4, 32), (2, 16, 4, 32), (4, 7, 1, 32), (6, 50, 3, 48), (8, 64, 2, 64),
                                   (2, 8, 2, 32, 4, 2, t.FORM_FUSED), (8, 5, 4, 64, 1, 1, t.FORM_FUSED),
                                   (4, 16, 1, 32, 3, 8, t.FORM_FUSED), (8, 8, 4, 64, 4, 16, t.FORM_FUSED, 2),
                                   (4, 32, 1, 64, 4, 16, t.FORM_STREAMS, 3), (6, 48, 1, 60, 4, 16, t.FORM_STREAMS, 24),
                                   (4, 4, 1, 64, 1, 16, t.FORM_FUSED, 1, 8), (2, 5, 2, 32, 4, 4, t.FORM_STREAMS, 2, 12),
                                   (4, 4, 1, 64, 1, 16, t.FORM_FUSED, 1, 14), (4, 4, 1, 60, 1, 16, t.FORM_FUSED, 1, 10),
                                   (4, 8, 1, 32, 1, 16, t.FORM_RESIDENT), (8, 12, 1, 64, 1, 16, t.FORM_RESIDENT),
                                   (2, 12, 1, 32, 1, 16, t.FORM_RESIDENT), (4, 12, 1, 64, 1, 16, t.FORM_RESIDENT),
                                   (4, 4, 1, 64, 1, 16, t.FORM_TILE), (2, 5, 1, 32, 1, 16, t.FORM_TILE),
                                   (6, 4, 1, 60, 1, 16, t.FORM_TILE), (8, 6, 1, 64, 1, 16, t.FORM_TILE),
                                   (4, 6, 1, 32, 1, 16, t.FORM_TILE)])
def test_schedule_does_not_change_results(oracle, sched):
    """Fused step count, tile heights (boundary and interior), waves per
    workgroup, DFT chunk and the launch form (streams, fused, resident, tile)
    only reorder work: every schedule gives the same fields and accumulators."""
    cfg = phys_small(steps=200)
    base = t.Context()
    base.set_schedule(4, 32, 4, 32, form=t.FORM_STREAMS)
    a = base.sim(cfg, PHYS_PARAMS, exc=0)
    a.step(200)
    other = t.Context()
    other.set_schedule(*sched)
    b = other.sim(cfg, PHYS_PARAMS, exc=0)
    b.step(90)
    b.step(110)  # an odd split of the steps exercises partial launches and chunks
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    Ja, Ma = a.dft()
    Jb, Mb = b.dft()
    assert rel_l2(Jb, Ja) < 1e-6 and rel_l2(Mb, Ma) < 1e-6
    a.close(); b.close(); base.close(); other.close()


def test_ref_dft_and_farfield_vs_oracle(ctx, oracle):
    cfg = ref_small(steps=300)
    p = ref_params(8)
    g = ctx.sim(cfg, p, exc=0)
    g.step(300)
    J, M = g.dft()
    o = oracle_fields(oracle, cfg, p, 0, 300, True)
    Jo, Mo = o.dft()
    assert rel_l2(J, Jo) < FF_TOL and rel_l2(M, Mo) < FF_TOL
    D, P = g.farfield()
    Do, Po = o.farfield()
    assert rel_l2(D, Do) < FF_TOL
    # P through the complex far field sqrt(D) e^{jP}: phase differences count
    # in proportion to the field's magnitude (the phase of a null is arbitrary)
    Do, Po = np.asarray(Do).reshape(D.shape), np.asarray(Po).reshape(P.shape)
    assert rel_l2(np.sqrt(D) * np.exp(1j * P), np.sqrt(Do) * np.exp(1j * Po)) < FF_TOL


def test_phys_dft_and_farfield_vs_oracle(ctx, oracle):
    cfg = phys_small(steps=500)
    g = ctx.sim(cfg, PHYS_PARAMS, exc=0)
    g.step(500)
    o = oracle_fields(oracle, cfg, PHYS_PARAMS, 0, 500, True)
    J, M = g.dft()
    Jo, Mo = o.dft()
    assert rel_l2(J, Jo) < FF_TOL and rel_l2(M, Mo) < FF_TOL
    D, P = g.farfield()
    Do, Po = o.farfield()
    assert rel_l2(D, Do) < FF_TOL
    assert rel_l2(np.sqrt(D) * np.exp(1j * P), np.sqrt(Do) * np.exp(1j * Po)) < FF_TOL


def test_ref_eval_fitness_vs_oracle(ctx, oracle):
    cfg = ref_small(steps=300)
    p = ref_params(21)
    rng = np.random.default_rng(0)
    target = rng.uniform(0.0, 0.05, cfg.farfield_len()).astype(np.float32)
    fit, D, P = ctx.eval(cfg, p, target=target, want_farfield=True)
    fo, Do, Po = oracle.eval(cfg, p, target=target, fp64=True)
    assert rel_l2(D, Do) < FF_TOL
    assert abs(fit - fo) <= FF_TOL * abs(fo)


def test_phys_eval_fitness_vs_oracle(ctx, oracle):
    """Physical mode, two conditionings.  Far from the target (a large
    tumour elsewhere) the fitness itself is within FF_TOL of FP64.  Near the
    target (D - T is 4e-4 of D: the small tumour barely moves the total far
    field) the fitness -||D - T||^2 / n amplifies the far-field error by
    ~2||T|| / ||D - T||, so the bar is the first-order propagation of the
    far-field error of both D and T: |dfit| <= 2 ||D - T|| (||dD|| + ||dT||) / n,
    with ||dT|| bounded like ||dD||."""
    cfg = phys_small(steps=400)
    n = cfg.n_exc * cfg.n_freq
    far = np.array([-8.0, 9.0, 6.0, 30.0, 45.0, 15.0, 0.5, 0.0])
    fit, D, _ = ctx.eval(cfg, far, want_farfield=True)
    fo, Do, _ = oracle.eval(cfg, far, fp64=True)
    assert rel_l2(D, Do) < FF_TOL
    assert abs(fit - fo) <= FF_TOL * abs(fo), (fit, fo)
    fit, D, _ = ctx.eval(cfg, PHYS_PARAMS, want_farfield=True)
    fo, Do, _ = oracle.eval(cfg, PHYS_PARAMS, fp64=True)
    assert rel_l2(D, Do) < FF_TOL
    dD = np.linalg.norm(D.astype(np.float64) - Do)
    bound = 2 * np.sqrt(abs(fo) * n) * 2 * dD / n
    assert abs(fit - fo) <= FF_TOL * abs(fo) + bound, (fit, fo, bound)


def test_batch_equals_plain(ctx):
    cfg = phys_small(steps=300)
    vecs = np.array([PHYS_PARAMS, PHYS_PARAMS + [1, 1, 0.5, 1, 2, 1, 0.1, 0], PHYS_PARAMS - [2, 0, 1, 2, 3, 2, 0.1, 0]])
    fb, Db = ctx.eval_batch(cfg, vecs, want_farfield=True)
    for i, v in enumerate(vecs):
        fp, Dp, _ = ctx.eval(cfg, v, want_farfield=True)
        assert np.array_equal(Db[i], Dp)
        assert fb[i] == fp


def test_requests_alone_equal_requests_in_mix(ctx):
    """Continuous batching: a request's result does not depend on what else is
    resident (different grids, antenna counts and admission times)."""
    cfgs = [phys_small(steps=256, n=96, n_exc=1), phys_small(steps=320, n=128, n_exc=2),
            phys_small(steps=200, n=160, n_exc=3)]
    cfgs[0] = cfgs[0].copy(breast_radius=0.012, skin_thickness=0.001, src_radius=0.015)
    vecs = [PHYS_PARAMS, PHYS_PARAMS + [1, -1, 0.5, 0, 0, 0, 0, 0], PHYS_PARAMS + [-2, 2, 1, 2, 2, 1, 0.1, 0]]
    alone = [ctx.eval(c, v) for c, v in zip(cfgs, vecs)]
    tickets = [ctx.submit(cfgs[0], vecs[0]), ctx.submit(cfgs[1], vecs[1])]
    ctx.progress()
    ctx.progress()
    tickets.append(ctx.submit(cfgs[2], vecs[2]))  # admitted while the others are resident
    got = [ctx.wait(tk) for tk in tickets]
    for (f, lat), a in zip(got, alone):
        assert f == a and lat > 0


def test_error_paths(ctx):
    cfg = ref_small()
    with pytest.raises(t.TomoError) as e:
        ctx.eval(cfg, np.ones(20), target=np.zeros(cfg.farfield_len(), np.float32))
    assert e.value.code == t.TOMO_E_SHORT
    bad = ref_params(1)
    bad[0] = -1.0
    with pytest.raises(t.TomoError) as e:
        ctx.eval(cfg, bad, target=np.zeros(cfg.farfield_len(), np.float32))
    assert e.value.code == t.TOMO_E_INVALID
    pc = phys_small()
    with pytest.raises(t.TomoError):
        ctx.submit(pc, np.array([0, 0, 4, 20, 34, np.nan, 0, 0.0]))
    small = t.Context(hbm_budget=1 << 20)
    tk = small.submit(pc, PHYS_PARAMS)
    with pytest.raises(t.TomoError) as e:
        small.wait(tk)
    assert e.value.code == t.TOMO_E_NOMEM
    tk2 = ctx.submit(pc, PHYS_PARAMS)
    ctx.cancel(tk2)
    with pytest.raises(t.TomoError) as e:
        ctx.poll(tk2)
    assert e.value.code == t.TOMO_E_CANCELLED
    small.close()


def test_reused_slot_gets_its_own_tiles(oracle):
    """A simulation created after another one was destroyed reuses its device
    table slot; its tile list must still be built for its own grid."""
    c = t.Context()
    a = c.sim(ref_small(steps=60), ref_params(2), exc=0)
    a.step(60)
    a.close()
    cfg = phys_small(steps=200)
    b = c.sim(cfg, PHYS_PARAMS, exc=1)
    b.step(200)
    o = oracle.sim(cfg, PHYS_PARAMS, exc=1)
    o.step(200)
    assert rel_l2(b.fields()[0], o.fields()[0]) < FIELD_TOL_F32
    b.close()
    c.close()


def test_autotune_installs_a_schedule_that_keeps_results(oracle):
    cfg = phys_small(steps=200)
    c = t.Context()
    sched, rate = c.autotune(cfg, 2)
    assert rate > 0 and c.schedule() == sched
    g = c.sim(cfg, PHYS_PARAMS, exc=0)
    g.step(200)
    o = oracle.sim(cfg, PHYS_PARAMS, exc=0)
    o.step(200)
    assert rel_l2(g.fields()[0], o.fields()[0]) < FIELD_TOL_F32
    g.close()
    c.close()


@pytest.mark.parametrize("form", [1, 2, 3, 4])
def test_kernel_stats_account_for_every_cell_step(form):
    """Per-kernel timing: the update kernels' cell-steps add up to the
    simulations' cell-steps, every tile class of a split grid runs (in its
    own kernel, or inside the fused one), and the algorithmic bytes per
    cell-step stay within the 25..49 bytes the formula allows."""
    cfg = phys_small(steps=96, n=256, n_exc=2)
    c = t.Context()
    c.set_schedule(4, {3: 8, 4: 4}.get(form, 16), 2, 32, 2, 2, form)
    c.target_farfield(cfg)  # the target's own simulations stay out of the counters
    c.profile(True)
    c.eval(cfg, PHYS_PARAMS)
    st = c.stats()
    upd = [st[k] for k in ("k_tblock", "k_tedge", "k_trow", "k_tcol", "k_tinner", "k_tfused", "k_tres", "k_twave")]
    want = float(cfg.nx) * cfg.ny * cfg.steps * cfg.n_exc
    assert sum(k["cell_steps"] for k in upd) == want
    assert st["update"]["cell_steps"] == want
    ran = {1: ("k_tedge", "k_trow", "k_tcol", "k_tinner"), 2: ("k_tfused",), 3: ("k_tres",), 4: ("k_twave",)}[form]
    for k in ran + ("k_fold", "k_project_mfma", "k_fitness"):
        assert st[k]["launches"] > 0 and st[k]["ms"] > 0, k
    for k in set(("k_tblock", "k_tedge", "k_trow", "k_tcol", "k_tinner", "k_tfused", "k_tres", "k_twave")) - set(ran):
        assert st[k]["launches"] == 0, k
    b = sum(k["alg_bytes"] for k in upd) / want
    assert 25.0 < b < 49.0
    c.close()


@pytest.mark.parametrize("method", [t.SEARCH_DE, t.SEARCH_PSO])
def test_search_driver(ctx, method):
    """The asynchronous DE / PSO drivers run every evaluation through the
    request loop; the best vector they report stays inside the bounds and
    evaluates (plain path) to exactly the best fitness they report."""
    cfg = phys_small(steps=160, n=128, n_exc=1)
    lo = np.array([-10.0, -10.0, 2.0, 10.0, 10.0, 8.0, 0.0, 0.0])
    hi = np.array([10.0, 10.0, 6.0, 30.0, 50.0, 20.0, 1.0, 0.0])
    best, res = ctx.search(cfg, lo, hi, method, pop=8, max_evals=32, inflight=8, seed=3)
    assert res["evals"] == 32 and res["failed"] == 0
    assert np.all(best >= lo) and np.all(best <= hi)
    assert ctx.eval(cfg, best) == res["best_fitness"]
    assert res["p99_latency_s"] >= res["p50_latency_s"] > 0


def test_ref_cpu_semantics_fields_vs_oracle(ctx, oracle):
    """The FDTD_CPU() semantics (9 x 9 patches = 82 materials, absolute
    incident origin, no Debye, no NaN clamp) on the GPU against the FP32
    oracle, which reproduces the compiled FDTD_CPU.cxx bit for bit."""
    cfg = ref_cpu_small(steps=300)
    p = ref_cpu_params(4)
    g = ctx.sim(cfg, p, exc=1)
    g.step(300)
    mat, coef = g.materials()
    om, _ = oracle.sim(cfg, p, exc=1).materials()
    assert np.array_equal(mat, om) and coef.shape[0] == 82
    o32 = oracle_fields(oracle, cfg, p, 1, 300, False).fields()
    for name, a, b in zip(("ez", "hx", "hy"), g.fields()[:3], o32[:3]):
        assert rel_l2(a, b) < FIELD_TOL_F32, (name, rel_l2(a, b))


def test_objective_entry_matches_eval(ctx):
    """tomo_fitness (preset + vector -> fitness, packaged measurement) equals
    tomo_eval of the same preset against the same measurement; the generic
    FDTD_common normalisation is n_freq times the relative FDTD_CPU one."""
    p = ref_params(6)
    m = t.measurement("gpu")
    cfg = t.preset(t.PRESET_REF_GPU)
    f_obj = ctx.fitness(t.PRESET_REF_GPU, p)
    f_eval = ctx.eval(cfg, p, target=m[:cfg.farfield_len()])
    assert f_obj == f_eval and f_obj < 0
    f_rel = ctx.fitness(t.PRESET_REF_GPU, p, flags=t.FLAG_REL_FITNESS)
    f_com = ctx.fitness(t.PRESET_REF_GPU, p, flags=t.FLAG_REL_FITNESS | t.FLAG_FITNESS_PER_EXC)
    assert abs(f_com - f_rel * cfg.n_freq) <= 1e-12 * abs(f_com)


@pytest.mark.parametrize("form", [t.FORM_STREAMS, t.FORM_FUSED])
def test_graph_replay_equals_direct_launches(form):
    """Epochs replayed from the captured hipGraph (several replays, with host
    reads of the fields between them, and a re-capture when a simulation
    joins) give the same fields, accumulators and fitness as launching every
    kernel directly."""
    cfg = phys_small(steps=256)
    out = []
    for graphs in (False, True):
        c = t.Context()
        c.set_schedule(4, 16, 2, 32, 4, 8, form, 4)
        c.use_graphs(graphs)
        a = c.sim(cfg, PHYS_PARAMS, exc=0)
        snaps = []
        for _ in range(4):
            a.step(64)  # two full chunks: two replays of one graph
            snaps.append([f.copy() for f in a.fields()])
        b = c.sim(cfg, PHYS_PARAMS + [1, 1, 0.5, 1, 2, 1, 0.1, 0], exc=1)
        b.step(96)
        snaps.append([f.copy() for f in b.fields()])
        snaps.append(list(a.dft()) + list(b.dft()))
        vecs = np.array([PHYS_PARAMS, PHYS_PARAMS - [2, 0, 1, 2, 3, 2, 0.1, 0]])
        fb, Db = c.eval_batch(cfg, vecs, want_farfield=True)
        snaps.append([np.asarray(fb), np.asarray(Db)])
        out.append(snaps)
        a.close(); b.close(); c.close()
    for sa, sb in zip(*out):
        for x, y in zip(sa, sb):
            assert np.array_equal(x, y)


def test_large_batch_layouts_equal_single_simulation_layouts(ctx):
    """A batch large enough for the wide fold (8 frequencies per block) and
    the 32-simulation projection layout gives the same far fields and fitness
    as plain evaluations, which use the frequency-split fold and the
    angle-split projection; many staged epochs (a 700-step run, K = 64) agree
    with the same run stepped one epoch at a time by the request loop."""
    cfg = phys_small(steps=700, n=96, n_exc=4).copy(n_freq=5, breast_radius=0.012, skin_thickness=0.001,
                                                     src_radius=0.015)
    rng = np.random.default_rng(7)
    vecs = PHYS_PARAMS + rng.uniform(-1, 1, (40, 8)) * [2, 2, 0.5, 2, 3, 1, 0.1, 0]
    vecs[:, 7] = 0
    fb, Db = ctx.eval_batch(cfg, vecs, want_farfield=True)  # 160 simulations
    for i in (0, 17, 39):
        fp, Dp, _ = ctx.eval(cfg, vecs[i], want_farfield=True)
        assert np.array_equal(Db[i], Dp)
        assert fb[i] == fp
    tk = ctx.submit(cfg, vecs[17])
    f_served, _ = ctx.wait(tk)
    assert f_served == fb[17]



@pytest.mark.parametrize("form", [t.FORM_STREAMS, t.FORM_FUSED])
def test_debye_medium_propagation_matches_hankel_closed_form(form):
    """The HIP update itself against the closed form (no oracle): in a
    homogeneous Debye medium the ratio of the field's spectra 40 and 20 cells
    from the soft source is H0^(2)(40 k dx) / H0^(2)(20 k dx) with the Debye
    eps(w), to 1e-3 in magnitude and 2e-3 rad, along both axes; the probes lie
    in interior tiles and the source in an interior tile's row band, in both
    launch forms."""
    cfg, p = fat_medium()
    c = t.Context()
    c.set_schedule(tsteps=4, seg_rows=6, edge_waves=4, inner_waves=4, chunk=64, grid_mult=16, form=form, edge_rows=2)
    g = c.sim(cfg, p, exc=0)
    mat, _ = g.materials()
    assert np.all(mat == 2)
    sx = sy = cfg.nx // 2
    d1, d2, every = 20, 40, 4
    tr = []
    for _ in range(cfg.steps // every):
        g.step(every)
        ez = g.fields()[0]
        tr.append((ez[sy, sx + d1], ez[sy, sx + d2], ez[sy + d1, sx], ez[sy + d2, sx]))
    tr = np.array(tr, np.float64)
    for f in (1.5e9, 2.5e9, 3.5e9):
        want = line_source_ratio(f, d1, d2, cfg.dx, *FAT)
        for a, b in ((0, 1), (2, 3)):
            got = spectrum(tr[:, b], f, every * cfg.dt) / spectrum(tr[:, a], f, every * cfg.dt)
            assert abs(abs(got) / abs(want) - 1) < 1e-3, (f, got, want)
            assert abs(np.angle(got / want)) < 2e-3, (f, got, want)
    g.close(); c.close()



def test_debye_cylinder_scattering_matches_mie_series(ctx):
    """The HIP update with material interfaces against the cylinder's series
    solution (no oracle): a Debye cylinder in the fat medium, six probes around
    it over a reference probe, within 1% at 1.5-3.5 GHz."""
    cfg, p = fat_with_cylinder()
    g = ctx.sim(cfg, p, exc=0)
    mat, _ = g.materials()
    assert set(np.unique(mat)) == {2, 4}
    a = np.sqrt((mat == 4).sum() / np.pi) * cfg.dx
    sx = sy = cfg.nx // 2
    every = 4
    pts = CYL_PROBES + [CYL_REF]
    tr = []
    for _ in range(cfg.steps // every):
        g.step(every)
        ez = g.fields()[0]
        tr.append([ez[sy + v, sx + u] for u, v in pts])
    tr = np.array(tr, np.float64)
    centre = (10e-3, 0.0)
    for f in (1.5e9, 2.5e9, 3.5e9):
        ref = spectrum(tr[:, -1], f, every * cfg.dt)
        for i, (u, v) in enumerate(CYL_PROBES):
            got = spectrum(tr[:, i], f, every * cfg.dt) / ref
            want = cylinder_field(f, (u * cfg.dx, v * cfg.dx), centre, a, FAT, CYL) / \
                cylinder_field(f, (CYL_REF[0] * cfg.dx, CYL_REF[1] * cfg.dx), centre, a, FAT, CYL)
            assert abs(got / want - 1) < 1e-2, (f, (u, v), got, want)
    g.close()



def test_far_field_of_lossy_cylinder_matches_series(ctx, oracle):
    """The HIP evaluation's far field (fold, matrix-core projection) against
    the lossy cylinder's series far field, normalised over the angles: within
    1% relative L2 and 2% at any angle, at every frequency (the antenna cell
    comes from the oracle's geometry)."""
    from test_oracle import check_cylinder_far_field
    cfg, p = skin_cylinder()
    g = ctx.sim(cfg, p, exc=0)
    mat, _ = g.materials()
    g.close()
    _, D, _ = ctx.eval(cfg, p, want_farfield=True)
    sx, sy = oracle.source_cell(cfg, 0)
    check_cylinder_far_field(np.asarray(D, np.float64).reshape(cfg.n_freq, cfg.n_angles), cfg, mat,
                             ((sx - cfg.nx // 2) * cfg.dx, (sy - cfg.ny // 2) * cfg.dy))


def test_ref_gpu_debye_disc_scattering_matches_plane_wave_series(ctx):
    """The HIP form of the FDTD_GPU.cu update (scattered-field formulation,
    Debye current, centred incident-wave origin) against the Debye disc's
    plane-wave series (no oracle): ten probes within 1.5% at 2-3 GHz, as
    ratios between probes and absolutely against the incident pulse."""
    from test_oracle import check_disc_scattering
    cfg, p = ref_debye_disc()
    g = ctx.sim(cfg, p, exc=0)
    mat, _ = g.materials()
    assert set(np.unique(mat)) == set(range(10))
    c, every = cfg.nx // 2, 2
    tr = []
    for _ in range(cfg.steps // every):
        g.step(every)
        ez = g.fields()[0]
        tr.append([ez[c + v, c + u] for u, v in DISC_PROBES])
    g.close()
    check_disc_scattering(np.array(tr, np.float64), cfg, mat, every)


def test_ref_gpu_far_field_phase_of_debye_disc_matches_series(ctx):
    """The HIP far field of the FDTD_GPU.cu update, D and P together, against
    the Debye disc's series far field (no oracle): within 1.5% complex
    relative L2 at 2-5 GHz."""
    from test_oracle import check_disc_far_field
    cfg, p = ref_debye_disc(steps=2500)
    g = ctx.sim(cfg, p, exc=0)
    mat, _ = g.materials()
    g.step(cfg.steps)
    D, P = g.farfield()
    g.close()
    check_disc_far_field(D, P, cfg, mat)


def test_resident_form_batches_requests_and_fallback_equal_other_forms():
    """The resident form with several simulations in one cooperative launch
    (a batch of vectors x antennas, and a request-loop mix of grid sizes)
    gives exactly the fitness and far fields of the streams form; a grid whose
    tiles cannot all be co-resident (2049 x 2049 nodes) falls back to the
    fused launches and keeps the same fields."""
    cfg = phys_small(steps=300, n=256, n_exc=2)
    vecs = np.array([PHYS_PARAMS, PHYS_PARAMS + [1, 1, 0.5, 1, 2, 1, 0.1, 0], PHYS_PARAMS - [2, 0, 1, 2, 3, 2, 0.1, 0]])
    mix = [(phys_small(steps=200, n=96, n_exc=1).copy(breast_radius=0.012, skin_thickness=0.001, src_radius=0.015),
            PHYS_PARAMS), (phys_small(steps=260, n=160, n_exc=3), PHYS_PARAMS + [-2, 2, 1, 2, 2, 1, 0.1, 0])]
    out = []
    for form, rows in ((t.FORM_STREAMS, 16), (t.FORM_RESIDENT, 8)):
        with t.Context() as c:
            c.set_schedule(tsteps=4, seg_rows=rows, edge_waves=1, chunk=64, inner_waves=4, grid_mult=16, form=form)
            fb, Db = c.eval_batch(cfg, vecs, want_farfield=True)
            tks = [c.submit(cf, v) for cf, v in mix]
            served = [c.wait(tk)[0] for tk in tks]
            big = t.preset(t.PRESET_CFG4, 2048).copy(steps=64)
            g = c.sim(big, PHYS_PARAMS, exc=0)
            g.step(64)
            ez = g.fields()[0]
            g.close()
            out.append((fb, Db, served, ez))
    (fa, Da, sa, ea), (fr, Dr, sr, er) = out
    assert np.array_equal(fa, fr) and np.array_equal(Da, Dr) and sa == sr
    assert np.array_equal(ea, er) and np.abs(er).max() > 0
