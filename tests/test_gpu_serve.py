"""Serving pool on the GPU: jobs admitted while others run, advanced together in multi-job
launches and hipGraph bundles, must give every job exactly the responses of the same job run
alone on its own context (which the parity suite ties to the oracle), bit for bit.
"""
import numpy as np
import pytest

import oracle
from parallelfdtd_amd import native as N
from parallelfdtd_amd import serve as S

pytestmark = pytest.mark.gpu


def small_jobs(n=6, first=0):
    """Rooms 8..64 voxels per side, 100..400 steps, FP32/FP16 as the stream draws them."""
    return [S.make_job(j, scale=8, step_scale=20) for j in range(first, first + n)]


def drive(pool, jobs, schedule):
    """Submit jobs[k] before advance number schedule[k]; run until all finish."""
    ids, results, step = {}, {}, 0
    order = sorted(range(len(jobs)), key=lambda k: schedule[k])
    nxt = 0
    while len(results) < len(jobs):
        while nxt < len(order) and schedule[order[nxt]] <= step:
            k = order[nxt]
            jid = pool.submit(jobs[k])
            if jid is None:
                break
            ids[jid] = k
            nxt += 1
        pool.advance()
        step += 1
        for jid in pool.poll():
            results[ids[jid]] = pool.responses(jid)
            pool.release(jid)
        assert step < 100000
    return results


@pytest.fixture(scope="module")
def jobs_and_alone():
    jobs = small_jobs(6)
    bf = S.make_job(40, scale=8, step_scale=20)
    bf.precision, bf.scale_exp = N.BF16, 12
    jobs.append(bf)
    precs = {j.precision for j in jobs}
    assert N.F32 in precs and N.F16 in precs
    return jobs, [S.run_alone(j) for j in jobs]


@pytest.mark.parametrize("use_graph", [True, False])
def test_pool_alone_vs_mix_bit_exact(gpu, jobs_and_alone, use_graph):
    jobs, alone = jobs_and_alone
    with S.Pool(0, arena_bytes=1 << 30, bundle=16, use_graph=use_graph) as pool:
        got = drive(pool, jobs, schedule=[0, 0, 0, 2, 2, 5, 9])
        st = pool.stats()
    for k, j in enumerate(jobs):
        assert got[k].shape == alone[k].shape
        assert np.array_equal(got[k], alone[k]), f"job {k} ({j.labels.shape}, prec {j.precision})"
    if use_graph:
        assert st["graph_builds"] >= 3
    assert st["resident"] == 0


def test_pool_oracle(gpu):
    """One FP32 job through the pool against the CPU oracle directly."""
    spec = S.make_job(3, scale=8, step_scale=20)
    spec.precision, spec.scale_exp = N.F32, 0
    pos, mat, _ = oracle.classify(spec.labels)
    ref = oracle.run(pos, mat, np.repeat(spec.betas[:, None], 20, axis=1), spec.src_xyz,
                     spec.src_kind, spec.signals, spec.rec_xyz, spec.n_steps, threads=8)
    with S.Pool(0, arena_bytes=256 << 20, bundle=7) as pool:
        got = drive(pool, [spec], [0])[0]
    assert np.array_equal(got, ref)


def test_pool_admission_waits_for_memory(gpu, jobs_and_alone):
    """An arena that holds one job at a time: the second waits until the first is released."""
    jobs, alone = jobs_and_alone
    a, b = jobs[0], jobs[1]
    need = max(S.Pool.job_bytes(a), S.Pool.job_bytes(b))
    with S.Pool(0, arena_bytes=need + 4096, bundle=16) as pool:
        ida = pool.submit(a)
        assert ida is not None
        assert pool.submit(b) is None  # does not fit while a is resident
        res = {}
        while ida not in res:
            pool.advance()
            for jid in pool.poll():
                res[jid] = pool.responses(jid)
                pool.release(jid)
        idb = pool.submit(b)
        assert idb is not None
        while idb not in res:
            pool.advance()
            for jid in pool.poll():
                res[jid] = pool.responses(jid)
                pool.release(jid)
        assert pool.mem()[1] == pool.mem()[0]  # everything returned to the arena
    assert np.array_equal(res[ida], alone[0]) and np.array_equal(res[idb], alone[1])


def test_pool_rejects_bad_jobs(gpu):
    spec = S.make_job(1, scale=8, step_scale=20)
    with S.Pool(0, arena_bytes=256 << 20) as pool:
        bad = S.JobSpec(**{**spec.__dict__, "labels": spec.labels.copy()})
        bad.labels[0, 3, 3] = 0  # hole in the floor
        with pytest.raises(N.PfdtdError) as ei:
            pool.submit(bad)
        assert ei.value.code == N.E_DOMAIN
        bad = S.JobSpec(**{**spec.__dict__, "labels": spec.labels.copy()})
        bad.labels[0, 3, 3] = len(spec.betas) + 1  # names a material outside the table
        with pytest.raises(N.PfdtdError) as ei:
            pool.submit(bad)
        assert ei.value.code == N.E_DOMAIN
        bad = S.JobSpec(**{**spec.__dict__, "precision": N.F64})
        with pytest.raises(N.PfdtdError) as ei:
            pool.submit(bad)
        assert ei.value.code == N.E_ARG
        assert pool.stats()["resident"] == 0
        assert pool.mem()[1] == pool.mem()[0]


def test_serve_stream_matches_alone(gpu):
    """The request loop (Poisson arrivals, admission, bundles, completion) on a short stream."""
    rep = S.serve(10, rate=200.0, scale=8, step_scale=40, arena_bytes=1 << 30, bundle=16,
                  keep_results=True, log=lambda m: None)
    assert rep["jobs"] == 10 and len(rep["results"]) == 10
    assert rep["latency_s"]["p50"] > 0 and rep["latency_s"]["p99"] >= rep["latency_s"]["p50"]
    for j in (0, 4, 9):
        assert np.array_equal(rep["results"][j], S.run_alone(S.make_job(j, 8, 40))), f"job {j}"


@pytest.mark.parametrize("bundle", [1, 5, 16])
def test_pool_two_level_and_single_step_jobs_mixed(gpu, bundle):
    """Jobs advanced by the two-level sweep (four buffers, two steps per launch, one for the
    last odd step) and jobs advanced by single steps (a source on the never-updated first z
    slice) in one pool, with step counts 1, 2, 3 and odd lengths: each equals the job alone."""
    base = small_jobs(5, first=20)
    specs = []
    for k, n_steps in enumerate([1, 2, 3, 37, 64]):
        j = base[k]
        specs.append(S.JobSpec(**{**j.__dict__, "n_steps": n_steps}))
    single = S.JobSpec(**{**base[0].__dict__, "n_steps": 45})
    src = single.src_xyz.copy()
    src[0, 2] = 0  # on the first z slice: this job takes single steps
    single.src_xyz = src
    specs.append(single)
    alone = [S.run_alone(j) for j in specs]
    with S.Pool(0, arena_bytes=1 << 30, bundle=bundle) as pool:
        got = drive(pool, specs, schedule=[0, 0, 1, 1, 2, 0])
    for k, j in enumerate(specs):
        assert np.array_equal(got[k], alone[k]), f"job {k} ({j.n_steps} steps, prec {j.precision})"
