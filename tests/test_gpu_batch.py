"""GPU parity for the ragged multi-job gap-block batch (persistent launch, hipGraph replay,
autotuned tiles): every job must equal its own single-job score matrix, which the gap-block
suite pins to the oracle; a few jobs are also checked against the oracle directly."""
from __future__ import annotations

import random

import numpy as np
import pytest

from conftest import cdr3_like, rand_seq

pytestmark = pytest.mark.gpu


def _jobs(st, rng):
    from seqtree_amd.batch import GapblockJob

    b62, structural = st.SubstitutionMatrix.blosum62(), st.SubstitutionMatrix.structural()
    jobs = []
    for t in range(9):
        nq, nr = rng.randint(1, 140), rng.randint(1, 3300)
        m = b62 if t % 2 == 0 else structural
        lam = int(1.5 * m.scale())
        jobs.append(GapblockJob(cdr3_like(rng, nq, 6, 30), cdr3_like(rng, nr, 6, 30), m,
                                2 * m.scale(), 1, st.central_prior(lam)))
    # unit cost on nucleotides, no prior; a job with an empty side; long sequences (table path)
    jobs.append(GapblockJob([rand_seq(rng, 0, 20, "ACGT") for _ in range(30)],
                            [rand_seq(rng, 0, 20, "ACGT") for _ in range(700)], None, 2, 1, None,
                            "nt"))
    jobs.append(GapblockJob([], cdr3_like(rng, 50, 6, 30), b62, 28, 1, None))
    jobs.append(GapblockJob([rand_seq(rng, 500, 650) for _ in range(3)],
                            [rand_seq(rng, 480, 700) for _ in range(40)], b62, 28, 1, None))
    # enough references (20 chunks) for two chunk groups per query block
    jobs.append(GapblockJob(cdr3_like(rng, 5, 6, 30), cdr3_like(rng, 20_000, 6, 30), b62, 28, 1,
                            st.central_prior(21)))
    # a wrapping prior forces the exact evaluation order inside the batch
    jobs.append(GapblockJob(cdr3_like(rng, 20, 6, 16), cdr3_like(rng, 90, 6, 16), b62, 28, 1,
                            lambda i, d, n: (2**31 - 1) - 5 * i))
    return jobs


def _dense(st, job):
    return np.asarray(st.score_matrix(job.queries, job.refs, job.matrix, job.gap_open,
                                      job.gap_extend, job.gap_prior, alphabet=job.alphabet))


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_equals_single_jobs_across_replays(gpu, autotune):
    st = gpu
    from seqtree_amd._device import DeviceArray
    from seqtree_amd.batch import GapblockBatch

    jobs = _jobs(st, random.Random(1 + autotune))
    want = [_dense(st, j) for j in jobs]
    batch = GapblockBatch(jobs, autotune=autotune)
    tiles = batch.tiles()
    assert tiles["queries_per_unit"] >= 1 and tiles["threads"] in (256, 512)
    for replay in range(3):
        for o in batch.outputs:  # poison the outputs between replays
            o.fill_from(np.full(o.shape, -7, np.int32))
        batch.run()
        got = batch.results()
        for t, (g, w) in enumerate(zip(got, want)):
            assert g.shape == w.shape, t
            assert np.array_equal(g, w), f"replay {replay}, job {t}"
        # eager library work between replays
        st.score_matrix(jobs[0].queries[:3], jobs[0].refs[:50], jobs[0].matrix, 28)
        scratch = DeviceArray.from_numpy(np.arange(1 << 16, dtype=np.int32))
        assert scratch.numpy()[-1] == (1 << 16) - 1
    batch.close()


def test_batch_against_the_oracle(gpu, oracle_lib):
    st = gpu
    from seqtree_amd.batch import GapblockBatch, GapblockJob
    from seqtree_amd.gapblock import _prior_cube

    rng = random.Random(5)
    m = st.SubstitutionMatrix.blosum62()
    syms = st.amino_acids()
    pen = np.array([[m.penalty(a, b) for b in syms] for a in syms], np.int32)
    jobs = [GapblockJob(cdr3_like(rng, 50, 6, 30), cdr3_like(rng, 2000, 6, 30), m, 28, 1,
                        st.central_prior(21)) for _ in range(3)]
    batch = GapblockBatch(jobs)
    batch.run()
    for job, got in zip(jobs, batch.results()):
        w = max(len(s) for s in list(job.queries) + list(job.refs))
        want = oracle_lib.gapblock_matrix(job.queries, job.refs, pen, 28, 1,
                                          _prior_cube(job.gap_prior, w), w, threads=8)
        assert np.array_equal(got, want)
    batch.close()


def test_batch_encoded_sets_equal_strings(gpu):
    st = gpu
    from seqtree_amd.batch import EncodedSet, GapblockBatch, GapblockJob

    rng = random.Random(11)
    syms = st.amino_acids()
    m = st.SubstitutionMatrix.structural()

    def enc(seqs):
        off = np.zeros(len(seqs) + 1, np.uint64)
        off[1:] = np.cumsum([len(s) for s in seqs])
        codes = np.array([syms.index(c) for s in seqs for c in s], np.uint8)
        return EncodedSet(codes, off)

    q, r = cdr3_like(rng, 33, 6, 30), cdr3_like(rng, 1500, 6, 30)
    jobs = [GapblockJob(q, r, m, 14, 1, st.central_prior(10)),
            GapblockJob(enc(q), enc(r), m, 14, 1, st.central_prior(10))]
    batch = GapblockBatch(jobs)
    batch.run()
    a, b = batch.results()
    assert np.array_equal(a, b) and np.array_equal(a, _dense(st, jobs[0]))
    batch.close()


def test_batch_replay_after_query_and_reference_refill(gpu):
    """The caller may rewrite its sequence buffers between runs (same sizes, same maximum
    lengths): each replay re-packs the references and re-pairs the queries by length (the
    twin rows of short jobs), so the results follow the new contents."""
    st = gpu
    from seqtree_amd.batch import GapblockBatch, GapblockJob, _encode

    rng = random.Random(21)
    m = st.SubstitutionMatrix.blosum62()
    jobs = [GapblockJob(cdr3_like(rng, 61, 6, 30), cdr3_like(rng, 2500, 6, 30), m, 28, 1,
                        st.central_prior(21)) for _ in range(2)]
    batch = GapblockBatch(jobs)
    batch.run()
    first = batch.results()
    assert all(np.array_equal(g, _dense(st, j)) for g, j in zip(first, jobs))
    syms = st.alphabet_symbols("aa")
    for t, job in enumerate(jobs):  # a permutation keeps every buffer's size and the maxima
        q = list(job.queries)
        r = list(job.refs)
        rng.shuffle(q)
        rng.shuffle(r)
        qb, qo = _encode(q, syms)
        rb, ro = _encode(r, syms)
        tq_b, tq_o, tr_b, tr_o = batch._keep[6 * t + 2: 6 * t + 6]
        tq_b.fill_from(qb)
        tq_o.fill_from(qo)
        tr_b.fill_from(rb)
        tr_o.fill_from(ro)
        jobs[t] = GapblockJob(q, r, m, 28, 1, st.central_prior(21))
    batch.run()
    for t, (g, j) in enumerate(zip(batch.results(), jobs)):
        assert np.array_equal(g, _dense(st, j)), f"job {t}"
    batch.close()
