// Ragged multi-job gap-block scoring in one persistent launch.
//
// A batch is many independent gap-block jobs (each its own query set, reference set,
// penalty table, gap costs and prior) of uneven sizes. Every job's query rows are cut into
// blocks of `queries_per_unit` queries and its reference chunks into groups small enough to
// stay in L2; a work unit is one query block against one chunk group. Units are numbered job
// by job (heaviest jobs first) and, inside a job, group-major, so the units in flight at one
// time stream the same reference chunks. They are handed out through one device atomic
// counter to a fixed grid of workgroups that loop until the counter runs past the last unit
// -- so the launch shape does not depend on the batch, and the whole batch (reference
// packing, counter reset, scoring) can be captured once in a hipGraph and replayed.
//
// Inside a unit the workgroup builds each query's penalty profile in LDS and sweeps the
// job's chunk-sorted references exactly as the single-job kernel does (same per-pair
// routine, gapblock_core.hpp), staging each chunk's results in LDS for one coalesced row
// write. Short jobs whose sums fit int16 run twin rows: their query blocks are blocks of
// twins (two queries of one length in the 16-bit halves of one table), paired by length on
// every run by a small kernel ahead of the persistent one.

#include "gapblock_core.hpp"

namespace st {
namespace {

using gb::Look;
using gb::PackedWords;
using gb::pair_score;
using gb::pair_score_fast;

template <bool PROFILE, bool EXACT, bool SHORT, uint32_t THREADS>
__device__ __forceinline__ void run_unit(const BatchJob& J, uint64_t q_begin, uint32_t qt, uint32_t c_begin,
                         uint32_t c_end, int32_t* lds, int32_t* stage) {
    constexpr uint32_t kSlots = kChunk / THREADS;
    const uint32_t tid = threadIdx.x;
    const uint32_t A = J.A;
    const uint64_t q_end = q_begin + qt < J.n_queries ? q_begin + qt : J.n_queries;
    int32_t* slice = lds + (SHORT ? J.prof_rows : 0u) * gb::kRowInts;
    if constexpr (!PROFILE) {
        gb::build_rows(lds, A, A, J.pen, A, [](uint32_t r) { return r; }, tid, THREADS);
        __syncthreads();
    }
    for (uint64_t i = q_begin; i < q_end; ++i) {
        const uint64_t q0 = gb::global_ptr(J.q_off)[i];
        const uint32_t m = uint32_t(gb::global_ptr(J.q_off)[i + 1] - q0);
        const uint8_t* q = J.q_codes + q0;
        if constexpr (PROFILE) {
            __syncthreads();  // the previous query's profile is no longer read
            const uint32_t rows = SHORT ? J.prof_rows : m;
            gb::build_rows(lds, rows, m, J.pen, A, [&](uint32_t t) { return uint32_t(q[t]); },
                           tid, THREADS);
            if constexpr (SHORT)
                gb::build_slice(slice, J.slice_rows - 1, m, J.prior, J.w1, J.big, tid, THREADS);
            __syncthreads();
        }
        const Look<PROFILE> look{lds, q};
        for (uint32_t c = c_begin; c < c_end; ++c) {
#pragma unroll
            for (uint32_t it = 0; it < kSlots; ++it) {
                const uint32_t slot = c * kChunk + it * THREADS + tid;
                const uint32_t k = gb::global_ptr(J.refs.orig)[slot];
                const uint32_t n = gb::global_ptr(J.refs.len)[slot];
                const PackedWords words{gb::global_ptr(J.refs.words) + slot, J.refs.npad};
                int32_t score = 0;
                if constexpr (SHORT) {
                    score = gb::pair_score_short(
                        lds, slice,
                        gb::buffer_words(J.refs.words, J.refs.npad, J.refs.nwords, slot), k, m,
                        n, J.prior, J.w1, J.gap_open, J.gap_extend);
                } else {
                    const uint32_t n0 = __builtin_amdgcn_readfirstlane(n);
                    if (PROFILE && !EXACT && __all(k == kNoSlot || n == n0) &&
                        (m < n0 ? m : n0) <= gb::kFastMaxL) {
                        score = pair_score_fast(lds, words, m, n0, J.prior, J.w1, J.gap_open,
                                                J.gap_extend);
                    } else if (k != kNoSlot) {
                        score = pair_score<PROFILE, EXACT>(look, words, m, n, J.prior, J.w1,
                                                           J.gap_open, J.gap_extend);
                    }
                }
                if (k != kNoSlot) stage[k - c * kChunk] = score;
            }
            __syncthreads();
            const uint64_t col0 = uint64_t(c) * kChunk;
            const uint64_t lim64 = J.n_refs - col0;
            const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
            int32_t* row = J.out + i * J.n_refs + col0;
            for (uint32_t t = tid; t < lim; t += THREADS) row[t] = stage[t];
            __syncthreads();
        }
    }
}

// A unit of a job with twin rows: twins [t_begin, t_begin + qt) (each two queries of one
// length sharing the 16-bit halves of the tables, gapblock_core.hpp) against chunks
// [c_begin, c_end); the stage holds both rows of a chunk.
template <uint32_t THREADS>
__device__ __forceinline__ void run_unit_twin(const BatchJob& J, uint64_t t_begin, uint32_t qt, uint32_t c_begin,
                              uint32_t c_end, int32_t* lds, int32_t* stage) {
    constexpr uint32_t kSlots = kChunk / THREADS;
    const uint32_t tid = threadIdx.x;
    const uint64_t cap = twin_capacity(J.n_queries);
    const uint64_t t_end = t_begin + qt < cap ? t_begin + qt : cap;
    uint32_t* prof = reinterpret_cast<uint32_t*>(lds);
    uint32_t* slice = prof + J.prof_rows * gb::kTwinRowInts;
    for (uint64_t p = t_begin; p < t_end; ++p) {
        // the twin (a, b) as one 8-byte global load: a in the low word
        const uint64_t tw = gb::global_ptr(reinterpret_cast<const uint64_t*>(J.twins))[p];
        if (uint32_t(tw) == kNoSlot) break;  // past the job's last twin (the same for every thread)
        const uint64_t ia = uint32_t(tw), ib = uint32_t(tw >> 32);
        const uint64_t qa0 = gb::global_ptr(J.q_off)[ia], qb0 = gb::global_ptr(J.q_off)[ib];
        const uint32_t m = uint32_t(gb::global_ptr(J.q_off)[ia + 1] - qa0);
        __syncthreads();  // the previous twin's tables are no longer read
        gb::build_rows2(prof, J.prof_rows, m, J.pen, J.A, J.q_codes + qa0, J.q_codes + qb0, tid,
                        THREADS);
        gb::build_slice2(slice, J.slice_rows - 1, m, J.prior, J.w1, J.big, tid, THREADS);
        __syncthreads();
        // the next slot's (index, length) load while the current slot is scored
        uint32_t next_k = gb::global_ptr(J.refs.orig)[c_begin * kChunk + tid];
        uint32_t next_n = gb::global_ptr(J.refs.len)[c_begin * kChunk + tid];
        for (uint32_t c = c_begin; c < c_end; ++c) {
#pragma unroll 1  // one slot's twin scorer at a time: bounds the registers
            for (uint32_t it = 0; it < kSlots; ++it) {
                const uint32_t slot = c * kChunk + it * THREADS + tid;
                const uint32_t k = next_k;
                const uint32_t n = next_n;
                {
                    const uint32_t it2 = it + 1 < kSlots ? it + 1 : 0;
                    const uint32_t c2 = it + 1 < kSlots ? c : c + 1;
                    if (c2 < c_end) {
                        next_k = gb::global_ptr(J.refs.orig)[c2 * kChunk + it2 * THREADS + tid];
                        next_n = gb::global_ptr(J.refs.len)[c2 * kChunk + it2 * THREADS + tid];
                    }
                }
                int32_t sa = 0, sb = 0;
                gb::twin_score_short(
                    prof, slice, gb::buffer_words(J.refs.words, J.refs.npad, J.refs.nwords, slot),
                    k, m, n, J.gap_open, J.gap_extend, sa, sb);
                if (k != kNoSlot) {
                    stage[k - c * kChunk] = sa;
                    stage[kChunk + k - c * kChunk] = sb;
                }
            }
            __syncthreads();
            const uint64_t col0 = uint64_t(c) * kChunk;
            const uint64_t lim64 = J.n_refs - col0;
            const uint32_t lim = lim64 < kChunk ? uint32_t(lim64) : kChunk;
            int32_t* row_a = J.out + ia * J.n_refs + col0;
            for (uint32_t t = tid; t < lim; t += THREADS) row_a[t] = stage[t];
            if (ib != ia) {
                int32_t* row_b = J.out + ib * J.n_refs + col0;
                for (uint32_t t = tid; t < lim; t += THREADS) row_b[t] = stage[kChunk + t];
            }
            __syncthreads();
        }
    }
}

// Dynamic LDS: the chunk stage (stage_ints int32), then the unit's tables. TWIN_ONLY: every
// job has twin rows, so the kernel holds only the twin path (and needs far fewer registers
// than the general one, whose other paths set its register count).
template <uint32_t THREADS, bool TWIN_ONLY>
__global__ __launch_bounds__(THREADS) void gapblock_batch_kernel(
    const BatchJob* __restrict__ jobs, uint32_t n_jobs, uint64_t total_units, uint32_t qt,
    uint32_t stage_ints, unsigned long long* __restrict__ next_unit) {
    extern __shared__ int32_t dyn[];
    int32_t* stage = dyn;
    int32_t* lds = dyn + stage_ints;
    __shared__ unsigned long long unit_sh;
    __shared__ uint32_t job_sh;

    for (;;) {
        if (threadIdx.x == 0) {
            const unsigned long long u = atomicAdd(next_unit, 1ull);
            unit_sh = u;
            uint32_t lo = 0, hi = n_jobs;  // last job whose first unit is <= u
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (jobs[mid].unit_begin <= u) lo = mid;
                else hi = mid;
            }
            job_sh = lo;
        }
        __syncthreads();
        const uint64_t u = unit_sh;
        const uint32_t j = job_sh;
        __syncthreads();  // every thread has read the unit before thread 0 takes the next
        if (u >= total_units) break;
        const BatchJob& J = jobs[j];
        const uint64_t local = u - J.unit_begin;
        const uint32_t g = uint32_t(local / J.qblocks);
        const uint64_t q_begin = (local % J.qblocks) * qt;
        const uint32_t c0 = g * J.group;
        const uint32_t c1 = c0 + J.group < J.refs.nchunks ? c0 + J.group : J.refs.nchunks;
        if (TWIN_ONLY || J.twin) {
            run_unit_twin<THREADS>(J, q_begin, qt, c0, c1, lds, stage);
        } else if (J.short_path) {
            run_unit<true, false, true, THREADS>(J, q_begin, qt, c0, c1, lds, stage);
        } else if (J.profile) {
            if (J.exact) run_unit<true, true, false, THREADS>(J, q_begin, qt, c0, c1, lds, stage);
            else run_unit<true, false, false, THREADS>(J, q_begin, qt, c0, c1, lds, stage);
        } else {
            if (J.exact) run_unit<false, true, false, THREADS>(J, q_begin, qt, c0, c1, lds, stage);
            else run_unit<false, false, false, THREADS>(J, q_begin, qt, c0, c1, lds, stage);
        }
        __syncthreads();
    }
}

}  // namespace

void launch_gapblock_batch(const BatchJob* jobs, uint32_t n_jobs, uint64_t total_units,
                           const BatchTiles& tiles, uint32_t stage_ints, size_t lds_bytes,
                           bool twin_only, unsigned long long* next_unit, hipStream_t s) {
    if (total_units == 0) return;
    const dim3 grid(tiles.workgroups);
    const auto go = [&](auto kernel, uint32_t threads) {
        hipLaunchKernelGGL(kernel, grid, dim3(threads), lds_bytes, s, jobs, n_jobs, total_units,
                           tiles.queries_per_unit, stage_ints, next_unit);
    };
    if (tiles.threads == 512) {
        if (twin_only) go(gapblock_batch_kernel<512, true>, 512);
        else go(gapblock_batch_kernel<512, false>, 512);
    } else {
        if (twin_only) go(gapblock_batch_kernel<256, true>, 256);
        else go(gapblock_batch_kernel<256, false>, 256);
    }
}

}  // namespace st
