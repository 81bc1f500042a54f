"""Batched execution: many uneven jobs in one persistent fused launch, eager and
hipGraph-replayed, each job checked against the oracle.  Calls go through the
C ABI (pph_batch_*)."""
import numpy as np
import pytest

import oracle as O
import parpar_amd
from parpar_amd import batch as B
from golden_io import expected_valid

pytestmark = pytest.mark.gpu
SEED = 9

# (n_in, slice_size, recovery exponents, length of the last input)
SPECS = [
    (7, 4100, [0, 1, 2, 3, 65534], 1001),
    (40, 65536, list(range(33)), 65536),
    (1, 1024, [5], 1024),
    (130, 20000, list(range(100, 117)), 7),
    (3, 8192, list(range(0, 64000, 2000)), 8192),
    (19, 2, [0, 9], 1),
    # 400 rows: in the bilinear form's batch, row tiles of 256, 128 and 64 rows
    (21, 3000, list(range(400)), 2999),
]


def say(*a):
    print(*a, flush=True)


def make_jobs():
    jobs = []
    for j, (n, S, exps, last) in enumerate(SPECS):
        lens = [S] * n
        lens[-1] = last
        nums = [(7 * k + j) % 32768 for k in range(n)]
        jobs.append(B.Job(n, S, exps, input_nums=nums, lens=lens))
    return jobs


def fill_all(jobs):
    for j, job in enumerate(jobs):
        job.fill(SEED, j)
    parpar_amd.check(parpar_amd.lib().pph_synchronize())


def expected(j, job):
    n, S = job.n_in, job.slice_size
    inp = np.zeros((n, S), np.uint8)
    for k in range(n):
        ln = int(job.lens[k])
        if ln:
            inp[k, :ln] = O.prng_bytes(SEED, j, k, ln)
    ref = O.recover(inp, job.lens, S, job.exps, input_nums=job.input_nums)
    flags = expected_valid(inp, job.lens, S, job.exps, job.input_nums)
    return ref, flags


@pytest.fixture(scope="module")
def refs():
    jobs = make_jobs()
    return [expected(j, job) for j, job in enumerate(jobs)]


@pytest.mark.parametrize("mode", ["plain", "dense", "bilinear", "mixed", "auto"])
def test_batch_eager_and_graph(mode, refs):
    say("batch", mode)
    jobs = make_jobs()
    tuning = {"plain": B.fixed_tuning(plain=1.0, dense=0.001), "dense": B.fixed_tuning(plain=0.001, dense=1.0),
              "bilinear": B.fixed_tuning(plain=0.001, dense=0.002, bilinear=1.0), "auto": None}.get(mode)
    if mode == "mixed":
        # every form in one batch: plain for R <= 4, dense for 5..16, bilinear
        # above where the padding of its GEMM (inputs to 256, outputs to 128)
        # leaves it faster than plain
        tuning = B.fixed_tuning(plain=1.0, dense=0.0, bilinear=0.0)
        for b in range(8):
            tuning.dense[b] = 2e12 if b == 2 else 1e9
            tuning.bilinear[b] = tuning.bilinear_wide[b] = 3e13 if b >= 3 else 1e9
    batch = B.Batch(jobs, tuning=tuning)
    forms = [batch.plan(j)["form"] for j in range(len(jobs))]
    say(" forms", forms, "group", batch.plan(0)["group"])
    if mode in ("plain", "dense", "bilinear"):
        assert all(f == mode for f in forms)
    if mode == "mixed":
        assert set(forms) == {"plain", "dense", "bilinear"}, forms

    fill_all(jobs)
    batch.run(use_graph=False)
    parpar_amd.check(parpar_amd.lib().pph_synchronize())
    eager = [job.outputs().copy() for job in jobs]
    for j, job in enumerate(jobs):
        ref, flags = refs[j]
        assert (eager[j] == ref).all(), (mode, j)
        assert ((job.valid() == 1) == flags).all(), (mode, j)

    # graph replays, with eager device work and synchronisation between them
    scratch = parpar_amd.proc.DeviceBuffer(1 << 20)
    for rep in range(3):
        fill_all(jobs)
        for job in jobs:
            job.d_out.zero()
        scratch.zero()
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        batch.run(use_graph=True)
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        for j, job in enumerate(jobs):
            assert (job.outputs() == eager[j]).all(), (mode, rep, j)
            assert ((job.valid() == 1) == refs[j][1]).all(), (mode, rep, j)
    batch.close()


def test_autotune_rates_are_sane():
    t = B.autotune()
    d = t.as_dict()
    say("tuning", d)
    for k in ("plain_g8", "plain_g16", "dense", "bilinear", "bilinear_wide"):
        assert all(v > 1e9 for v in d[k]), k
    # the plain form's rate grows with the recovery count until it is compute bound
    assert d["plain_g16"][4] > d["plain_g16"][0]


@pytest.mark.parametrize("overlap", ["1", "0"])
def test_batch_bilinear_in_waves(refs, monkeypatch, overlap):
    """The bilinear jobs' cross-job plan under a small workspace: 1024-word
    pieces in waves of a few, each wave one word-plane launch, one GEMM launch
    per row-tile height over the pieces' items and one recombination launch,
    pipelined on two streams over two halves of the workspace (overlap "1") or
    in sequence on one (PPH_BATCH_OVERLAP=0); eager and graph replays equal the
    oracle."""
    monkeypatch.setenv("PPH_BATCH_BILINEAR_WORKSPACE", str(40 << 20))
    monkeypatch.setenv("PPH_BATCH_PIECE_WORDS", "1024")
    monkeypatch.setenv("PPH_BATCH_OVERLAP", overlap)
    jobs = make_jobs()
    batch = B.Batch(jobs, tuning=B.fixed_tuning(plain=0.001, dense=0.002, bilinear=1.0))
    monkeypatch.delenv("PPH_BATCH_BILINEAR_WORKSPACE")
    monkeypatch.delenv("PPH_BATCH_PIECE_WORDS")
    monkeypatch.delenv("PPH_BATCH_OVERLAP")
    assert all(batch.plan(j)["form"] == "bilinear" for j in range(len(jobs)))
    for use_graph in (False, True, True):
        fill_all(jobs)
        for job in jobs:
            job.d_out.zero()
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        batch.run(use_graph=use_graph)
        parpar_amd.check(parpar_amd.lib().pph_synchronize())
        for j, job in enumerate(jobs):
            ref, flags = refs[j]
            assert (job.outputs() == ref).all(), (use_graph, j)
            assert ((job.valid() == 1) == flags).all(), (use_graph, j)
    batch.close()
