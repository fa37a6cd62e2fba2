"""The HIP path at the benchmark workloads' own sizes and at the drop-in's own
configuration, against the FP64 oracle (and the FP32 oracle, which restates
the reference's FP32 arithmetic).

Tolerances (written here):
  * far field D, relative L2 per simulation: <= 1e-4 against FP64;
  * fitness: <= 1e-4 relative against FP64 when the fitness is well
    conditioned (a candidate far from the target); near the target the bar
    adds the first-order propagation of the far-field error;
  * fields, relative L2: <= 1e-4 against FP64 while the pulse crosses the
    grid; at the end of a run, when the pulse has left through the CPML, the
    error is stated against the norm of the mid-run field (the residual is
    orders of magnitude below it, and its own relative error measures
    rounding, not parity);
  * batched vs plain, alone vs in the serving mix: identical.

The oracle runs on the host's cores in threads (its C functions release the
GIL), one simulation per thread.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import tomography_amd as t
from cases import rel_l2
from tomography_amd.workloads import (BATCH_SCHEDULE, PLAIN_SCHEDULE, REF_SCHEDULE, batch_vectors, seeded_vectors,
                                      serve_requests, uneven_requests)

pytestmark = pytest.mark.gpu

FF_TOL = 1e-4
FIELD_TOL_F64 = 1e-4
# at most 16 host worker threads beside the one GPU
POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
# a candidate far from every preset's seeded target (well-conditioned fitness)
FAR = np.array([-8.0, 9.0, 5.0, 28.0, 45.0, 15.0, 0.5, 1.0])


def oracle_farfields(oracle, jobs):
    """[(cfg, params, exc)] -> [D (freq x angle)], FP64, in parallel threads."""
    futs = [POOL.submit(oracle.farfield, c, p, e, True) for c, p, e in jobs]
    return [f.result()[0] for f in futs]


def oracle_eval(oracle, cfg, p, target=None):
    """FP64 fitness and D of one vector: every excitation (and, without an
    explicit target, every excitation of the seeded target phantom) in its own
    thread, the fitness by the oracle's own arithmetic."""
    jobs = [(cfg, p, e) for e in range(cfg.n_exc)]
    if target is None:
        tv = oracle.target_vector(cfg)
        jobs += [(cfg, tv, e) for e in range(cfg.n_exc)]
    D = oracle_farfields(oracle, jobs)
    Dc = np.concatenate([d.ravel() for d in D[:cfg.n_exc]])
    T = np.concatenate([d.ravel() for d in D[cfg.n_exc:]]) if target is None else np.asarray(target)
    return oracle.fitness(cfg, Dc, T), Dc, T


def check_fitness(fit, fo, D, Do, T, n):
    """|fit - fo| <= FF_TOL |fo| + the first-order propagation of the far-field
    error: |d(-||D - T||^2 / n)| <= 2 ||D - T|| ||dD|| / n."""
    dD = np.linalg.norm(np.asarray(D, np.float64) - Do)
    bound = 2 * np.linalg.norm(np.asarray(Do) - T) * dD / n
    assert abs(fit - fo) <= FF_TOL * abs(fo) + bound, (fit, fo, bound)


def fields_mid_and_end(ctx, oracle, cfg, p, mid):
    """GPU and FP64-oracle fields of excitation 0 at step `mid` and at the end,
    and both far fields; the oracle runs in a thread while the GPU steps."""
    def run_oracle():
        s = oracle.sim(cfg, p, exc=0, fp64=True)
        s.step(mid)
        a = [f.copy() for f in s.fields()]
        s.step(cfg.steps - mid)
        b = [f.copy() for f in s.fields()]
        return a, b, s.farfield()[0]
    fut = POOL.submit(run_oracle)
    g = ctx.sim(cfg, p, exc=0)
    g.step(mid)
    ga = g.fields()
    g.step(cfg.steps - mid)
    gb = g.fields()
    gD = g.farfield()[0]
    g.close()
    oa, ob, oD = fut.result()
    return ga, gb, gD, oa, ob, oD


def check_fields(ga, gb, oa, ob, label):
    for name, a, b in zip(("ez", "hx", "hy", "jp"), ga, oa):
        e = rel_l2(a, b)
        print(f"{label} mid {name}: rel L2 {e:.2e}")
        assert e < FIELD_TOL_F64, (label, "mid", name, e)
    for name, a, b, ref in zip(("ez", "hx", "hy", "jp"), gb, ob, oa):
        e = np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(ref)
        print(f"{label} end {name}: error / mid-run norm {e:.2e} (end norm / mid norm "
              f"{np.linalg.norm(b) / np.linalg.norm(ref):.2e})")
        assert e < FIELD_TOL_F64, (label, "end", name, e)


def test_cfg1_full_size_vs_fp64_oracle(oracle):
    """256^2, dx 0.5 mm, 2048 steps, 8 frequencies x 180 angles, the plain
    schedule: fields mid-run and at the end, D and fitness against FP64."""
    cfg = t.preset(t.PRESET_CFG1)
    with t.Context() as ctx:
        ctx.set_schedule(**PLAIN_SCHEDULE)
        ga, gb, gD, oa, ob, oD = fields_mid_and_end(ctx, oracle, cfg, FAR, 512)
        check_fields(ga, gb, oa, ob, "cfg1")
        assert rel_l2(gD, oD) < FF_TOL
        fut = POOL.submit(oracle_eval, oracle, cfg, FAR)
        fit, D, _ = ctx.eval(cfg, FAR, want_farfield=True)
        fo, Do, T = fut.result()
        e = rel_l2(D, Do)
        print(f"cfg1 D rel L2 {e:.2e}, fitness {fit:.9e} vs {fo:.9e}")
        assert e < FF_TOL
        assert abs(fit - fo) <= FF_TOL * abs(fo)


def test_cfg2_full_size_plain_schedule_vs_fp64_oracle(oracle):
    """1024^2, dx 0.25 mm, 8192 steps under the bench's plain schedule (fused
    T = 4, 1-row boundary, 4-row column-edge and 10-row interior tiles): fields
    at step 2048 and at the end, D and fitness against the FP64 oracle."""
    cfg = t.preset(t.PRESET_CFG2)
    v = seeded_vectors(1, 8, 20261202)[0]
    with t.Context() as ctx:
        ctx.set_schedule(**PLAIN_SCHEDULE)
        fut = POOL.submit(oracle_eval, oracle, cfg, v)
        ga, gb, gD, oa, ob, oD = fields_mid_and_end(ctx, oracle, cfg, v, 2048)
        check_fields(ga, gb, oa, ob, "cfg2")
        assert rel_l2(gD, oD) < FF_TOL
        fit, D, _ = ctx.eval(cfg, v, want_farfield=True)
        fo, Do, T = fut.result()
        e = rel_l2(D, Do)
        print(f"cfg2 D rel L2 {e:.2e}, fitness {fit:.9e} vs {fo:.9e}")
        assert e < FF_TOL
        check_fitness(fit, fo, D, Do, T, cfg.n_exc * cfg.n_freq)


def test_cfg3_batch_sample_vs_fp64_oracle(oracle):
    """The whole 1024-simulation batch (64 vectors x 16 antennas, 512^2 x
    4096 steps, 16 frequencies x 360 angles) under the batched schedule; two
    of its vectors (32 simulations) against the FP64 oracle, per simulation,
    and against the plain path bit for bit."""
    cfg = t.preset(t.PRESET_CFG3)
    vecs = batch_vectors()
    pick = (0, 37)
    futs = [POOL.submit(oracle_eval, oracle, cfg, vecs[i]) for i in pick]
    with t.Context() as ctx:
        ctx.set_schedule(**BATCH_SCHEDULE)
        fb, Db = ctx.eval_batch(cfg, vecs, want_farfield=True)
        assert np.all(np.isfinite(fb))
        ctx.set_schedule(**PLAIN_SCHEDULE)
        for i in pick:
            fp, Dp, _ = ctx.eval(cfg, vecs[i], want_farfield=True)
            assert fp == fb[i] and np.array_equal(Dp, Db[i])
    rec = cfg.n_freq * cfg.n_angles
    for i, fut in zip(pick, futs):
        fo, Do, T = fut.result()
        for e in range(cfg.n_exc):
            err = rel_l2(Db[i][e * rec:(e + 1) * rec], Do[e * rec:(e + 1) * rec])
            assert err < FF_TOL, (i, e, err)
        print(f"cfg3 vector {i}: D rel L2 {rel_l2(Db[i], Do):.2e}, fitness {fb[i]:.9e} vs {fo:.9e}")
        check_fitness(fb[i], fo, Db[i], Do, T, cfg.n_exc * cfg.n_freq)


def test_cfg4_grids_vs_fp64_oracle_and_dft_precision(oracle):
    """The uneven batch's smallest grid and a mid-size one (about 724^2,
    5800 steps) evaluated inside an uneven mix through the request loop,
    against the FP64 oracle; and the compensated-FP32 running DFT against the
    FP64 one (TOMO_FLAG_DFT_FP64) on 24 results of the mix: far fields within
    1e-6 of each other."""
    reqs = uneven_requests(512)
    grids = np.array([c.nx for c, _ in reqs])
    small = int(np.argmin(grids))
    mid = int(np.argmin(np.abs(grids - 724)))
    futs = {i: POOL.submit(oracle_eval, oracle, reqs[i][0], reqs[i][1]) for i in (small, mid)}
    # the mix: the two checked requests among eight others of the smaller half
    others = [i for i in np.argsort(grids)[1:200:20] if i not in (small, mid)][:8]
    order = [others[0], small, others[1], others[2], mid] + others[3:]
    with t.Context() as ctx:
        ctx.set_schedule(**BATCH_SCHEDULE)
        tks = {}
        for k, i in enumerate(order):
            tks[i] = ctx.submit(*reqs[i])
            if k % 3 == 2:
                ctx.progress()  # later requests join while earlier ones are resident
        got = {i: ctx.wait(tk)[0] for i, tk in tks.items()}
        for i in (small, mid):
            cfg, v = reqs[i]
            fo, Do, T = futs[i].result()
            fp, Dp, _ = ctx.eval(cfg, v, want_farfield=True)
            assert got[i] == fp  # the mix and the plain path agree exactly
            e = rel_l2(Dp, Do)
            print(f"cfg4 grid {cfg.nx}^2 x {cfg.steps}: D rel L2 {e:.2e}, fitness {fp:.9e} vs {fo:.9e}")
            assert e < FF_TOL
            check_fitness(fp, fo, Dp, Do, T, cfg.n_exc * cfg.n_freq)
        # precision study: 24 results of the mix's smallest grids, both DFT precisions
        worst = 0.0
        for i in np.argsort(grids)[:24]:
            cfg, v = reqs[int(i)]
            c64 = cfg.copy(flags=cfg.flags | t.FLAG_DFT_FP64)
            _, D32, _ = ctx.eval(cfg, v, want_farfield=True)
            _, D64, _ = ctx.eval(c64, v, want_farfield=True)
            worst = max(worst, rel_l2(D32, D64))
        print(f"cfg4 Kahan-FP32 vs FP64 running DFT, 24 results: max rel L2 {worst:.2e}")
        assert worst < 1e-6


def test_cfg5_requests_alone_equal_requests_in_mix():
    """Continuous batching at the serving mix's own sizes (grids 256^2 to
    2048^2, 1-16 antennas, 8-32-dim vectors): every request's fitness is the
    same alone (plain path) and served among the others, admitted at
    different times while others are resident, under an HBM budget that holds
    only part of the mix at once."""
    pool = serve_requests(96)
    work = [c.nx * c.ny * c.steps * c.n_exc for c, _ in pool]
    # a bounded sample that still spans the size range: the largest grid with
    # the fewest antennas, and requests of every other grid size
    chosen, seen = [], set()
    for i in np.argsort(work):
        c = pool[i][0]
        if c.nx not in seen and work[i] < 1.5e11:
            chosen.append(int(i))
            seen.add(c.nx)
    assert {pool[i][0].nx for i in chosen} >= {256, 512, 1024, 2048}
    reqs = [pool[i] for i in chosen]
    need = max(t.request_bytes(c) for c, _ in reqs)
    with t.Context() as plain:
        plain.set_schedule(**PLAIN_SCHEDULE)
        alone = [plain.eval(c, v) for c, v in reqs]
    # a budget of twice the largest request: admission has to wait for slots to free up
    with t.Context(hbm_budget=2 * need + (64 << 20)) as ctx:
        ctx.set_schedule(**BATCH_SCHEDULE)
        for c, _ in reqs:
            ctx.target_farfield(c)  # targets outside the budgeted mix
        tks = []
        for k, (c, v) in enumerate(reqs):
            tks.append(ctx.submit(c, v))
            ctx.progress()
        got = [ctx.wait(tk) for tk in tks]
        assert ctx.stats()["requests"] == 0
    for (f, lat), a, (c, _) in zip(got, alone, reqs):
        print(f"cfg5 {c.nx}^2 x {c.n_exc} antennas: alone {a:.9e} in mix {f:.9e}, latency {lat:.2f} s")
        assert f == a and lat > 0


def test_ref_gpu_dropin_full_size_vs_oracle(oracle):
    """The FDTD_GPU() configuration itself (401^2, 4 excitations x 2000 steps,
    10 frequencies x 100 angles, 27 parameters) through the objective entry
    the shims call, against the packaged measurement: D and fitness against
    the FP64 oracle; and the FDTD_CPU() configuration (162 parameters, CPU
    semantics) likewise."""
    from cases import ref_cpu_params, ref_params
    with t.Context() as ctx:
        ctx.set_schedule(**REF_SCHEDULE)
        for pid, p, kind in ((t.PRESET_REF_GPU, ref_params(6), "gpu"), (t.PRESET_REF_CPU, ref_cpu_params(4), "cpu")):
            cfg = t.preset(pid)
            m = t.measurement(kind)[:cfg.farfield_len()]
            fut = POOL.submit(oracle_eval, oracle, cfg, p, m)
            f_obj = ctx.fitness(pid, p)
            fit, D, _ = ctx.eval(cfg, p, target=m, want_farfield=True)
            assert f_obj == fit
            fo, Do, T = fut.result()
            e = rel_l2(D, Do)
            print(f"{kind} drop-in: D rel L2 {e:.2e}, fitness {fit:.9e} vs {fo:.9e}")
            assert e < FF_TOL
            check_fitness(fit, fo, D, Do, T, cfg.n_exc * cfg.n_freq)
