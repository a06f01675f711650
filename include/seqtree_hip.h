/* seqtree_hip.h -- C-ABI of the MI355X (gfx950) all-vs-all scoring library.
 *
 * These entry points replace the compute of the reference's five dense-matrix functions
 * (include/seqtree/seqtree.hpp:198-242 in seqtree 0.6.0). The C++ API layer above them keeps
 * the reference's signatures, validation order, exceptions and messages; it encodes the
 * strings, packs them as CSR and calls in here. Nothing but plain pointers and sizes crosses
 * this boundary, and no exception escapes it.
 *
 * Sequences are passed as CSR: sequence i is bytes[offsets[i] .. offsets[i+1]), offsets has
 * count+1 entries and offsets[0] == 0. For the gap-block and Gotoh functions the bytes are
 * symbol codes already produced by the alphabet codec (each < alphabet_size); for Hamming and
 * Levenshtein they are the raw, case-sensitive characters.
 *
 * Output is the dense row-major matrix out[i * n_refs + k] (query i against reference k),
 * exactly as the reference's std::vector<int32_t> result.
 *
 * Every function returns the status code it also stores in *status (status may be NULL):
 *   ST_OK                 success
 *   ST_INVALID_ARGUMENT   the reference throws std::invalid_argument here (message kept)
 *   ST_OVERFLOW           the reference throws std::overflow_error here (message kept)
 *   ST_DEVICE_ERROR       a HIP runtime error, or no usable gfx950 device
 *   ST_OUT_OF_MEMORY      a device or pinned host allocation failed
 *
 * The functions are reentrant: each call runs on its own stream (host-memory variants) or
 * on the caller's stream (device-memory variants), with its own stream-ordered scratch.
 */
#ifndef SEQTREE_HIP_H
#define SEQTREE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ST_OK = 0,
    ST_INVALID_ARGUMENT = 1,
    ST_OVERFLOW = 2,
    ST_DEVICE_ERROR = 3,
    ST_OUT_OF_MEMORY = 4
};

typedef struct st_status {
    int32_t code;
    char message[256];
} st_status;

/* ---- host-memory entry points (the drop-in path) ------------------------------------ */

/* Single-gap-block penalty matrix.
 * Replaces the compute of seqtree::gapblock_matrix (src/gapblock.cpp:66-133; per-pair
 * `cell` at src/gapblock.cpp:39-62). pen is the alphabet_size x alphabet_size penalty table
 * indexed [query code][ref code] (src/gapblock.cpp:25-31); prior is NULL or the
 * (prior_width+1)^3 cube indexed [longer length][block length][block start]. The caller has
 * already validated gap costs, symbols, the prior's size and that every sequence is at most
 * prior_width long (src/gapblock.cpp:71-102). */
int32_t st_gapblock_matrix(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                           const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                           const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                           int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                           int32_t* out, st_status* status);

/* Gotoh affine Needleman-Wunsch (local == 0) or Smith-Waterman (local != 0) similarity
 * matrix, or with dist != 0 the distance d = s(a,a) + s(b,b) - 2 s(a,b).
 * Replaces the compute of seqtree::align_score_matrix / align_dist_matrix
 * (src/pairwise.cpp:273-352; per-pair DP score_encoded at src/pairwise.cpp:76-125). sim is
 * the signed similarity table indexed [query code][ref code]. Returns ST_OVERFLOW with the
 * reference's narrow() message (src/pairwise.cpp:42-49) if a result does not fit int32. */
int32_t st_align_matrix(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                        const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                        const int32_t* sim, uint32_t alphabet_size, int32_t local, int32_t dist,
                        int32_t gap_open, int32_t gap_extend, int32_t* out, st_status* status);

/* Unit-cost Hamming distance over raw bytes. Replaces the compute of seqtree::hamming_matrix
 * (src/distance.cpp:29-76,97-102). Returns ST_INVALID_ARGUMENT with the reference's message
 * for the first length-mismatched pair in row-major order. */
int32_t st_hamming_matrix(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                          const uint8_t* b, const uint64_t* b_offsets, uint64_t n_b, int32_t* out,
                          st_status* status);

/* Unit-cost Levenshtein distance over raw bytes. Replaces the compute of
 * seqtree::levenshtein_matrix (src/distance.cpp:29-64,78-95,104-109). */
int32_t st_levenshtein_matrix(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                              const uint8_t* b, const uint64_t* b_offsets, uint64_t n_b,
                              int32_t* out, st_status* status);

/* Thresholded (sparse) gap-block output: every (query, reference) pair whose score is
 * <= cutoff, as parallel arrays hit_query / hit_ref / hit_score in unspecified order.
 * *hit_count receives the total number of such pairs; only the first `capacity` are stored,
 * so a count above capacity tells the caller to retry with more room. This is the exhaustive
 * counterpart of the reference's GapBlockIndex.search (python/seqtree/gapblock.py:567-673)
 * without its d_max limit; it has no direct C++ counterpart in the reference. */
int32_t st_gapblock_threshold(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, const int32_t* pen,
                              uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                              const int32_t* prior, uint32_t prior_width, int32_t cutoff,
                              uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                              int32_t* hit_score, uint64_t* hit_count, st_status* status);

/* Gap-block scores of an explicit list of pairs: out[t] = score(query pair_query[t],
 * reference pair_ref[t]) with the same semantics as st_gapblock_matrix (the per-pair `cell`
 * of src/gapblock.cpp:39-62). Indices past either set return ST_INVALID_ARGUMENT. */
int32_t st_gapblock_pairs(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                          const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                          const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                          int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                          const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                          int32_t* out, st_status* status);

/* ---- per-query score histograms (control-calibrated counts; no counterpart in the
 * reference, whose E-values count Index search hits, python/seqtree/evalue.py:45-178) ------
 * histogram[i][bins + 2] (uint64) of the exact gap-block scores of query i against every
 * reference, as st_gapblock_matrix computes them: bin 0 counts scores below zero, bin 1 + s
 * the score s in [0, bins), bin bins + 1 scores >= bins. 1 <= bins <= 16382. The counts of
 * references within any cutoff below `bins`, and every order statistic of the scores up to
 * it, follow from one histogram. Up to 1022 bins the scoring kernels add each score to an LDS
 * histogram of their query rows and no score matrix is formed; wider histograms go through
 * dense row blocks (at most 512 MB of scratch) and a reduction pass. */
int32_t st_gapblock_histogram(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, const int32_t* pen,
                              uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                              const int32_t* prior, uint32_t prior_width, uint32_t bins,
                              uint64_t* histogram, st_status* status);
int32_t st_gapblock_histogram_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                     uint64_t n_queries, uint32_t q_max_len,
                                     const uint8_t* r_codes, const uint64_t* r_offsets,
                                     uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                     uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                                     const int32_t* prior, uint32_t prior_width, uint32_t bins,
                                     uint64_t* histogram, void* stream, st_status* status);

/* ---- Gotoh alignment with traceback (src/pairwise.cpp:141-267, `align_pair`) ----------
 * Scores and traces back every (pair_query[p], pair_ref[p]) with the reference's recurrences,
 * boundary values, argmax order and traceback rules. Outputs per pair: score[p] (int64; the
 * reference narrows it to int32 and raises OverflowError past that), the ops in forward order
 * at ops + p * ops_stride ('M' equal codes, 'S' substitution, 'D' query residue against a
 * gap, 'I' reference residue against a gap; ops_stride >= q_max_len + r_max_len), their count
 * n_ops[p], and the first aligned residue of each side (q_begin[p], r_begin[p]: 0 in global
 * mode; in local mode where the matched sub-sequences start). The aligned strings are those
 * residues, decoded, in op order. align_pair is the one-pair case; the batch has no
 * counterpart in the reference (it is the traceback step for candidate pairs). */
int32_t st_align_pairs(const uint8_t* q_codes, const uint64_t* q_offsets, uint64_t n_queries,
                       const uint8_t* r_codes, const uint64_t* r_offsets, uint64_t n_refs,
                       const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                       const int32_t* sim, uint32_t alphabet_size, int32_t local,
                       int32_t gap_open, int32_t gap_extend, uint8_t* ops, uint64_t ops_stride,
                       int64_t* score, uint32_t* n_ops, uint32_t* q_begin, uint32_t* r_begin,
                       st_status* status);
/* Device-resident sets, pairs and outputs on `stream`. Pair indices must be in range, and
 * q_max_len / r_max_len must bound the lengths of the queries and references the pairs name:
 * the direction scratch is sized by them. A pair longer than either bound is skipped and the
 * call returns ST_INVALID_ARGUMENT; the check reads back one device flag, so the call
 * synchronises `stream` once. */
int32_t st_align_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                              uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                              const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                              const uint32_t* pair_query, const uint32_t* pair_ref,
                              uint64_t n_pairs, const int32_t* sim, uint32_t alphabet_size,
                              int32_t local, int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                              uint64_t ops_stride, int64_t* score, uint32_t* n_ops,
                              uint32_t* q_begin, uint32_t* r_begin, void* stream,
                              st_status* status);

/* ---- penalty-space global alignment with traceback (src/index.cpp:105-179) -----------
 * Replaces the compute of seqtree::Index::align (src/index.cpp:105-179, declared at
 * include/seqtree/seqtree.hpp:118-120) for a list of pairs: the global alignment of query
 * pair_query[p] against reference pair_ref[p] that minimises substitution penalty plus affine
 * gap cost (a run of L gap columns costs gap_open + (L - 1) gap_extend; an X<->Y switch opens a
 * new run), with the reference's tie order (the minimum over the M, X, Y states keeps the first
 * of equal values) and its int32 score. pen is the alphabet_size x alphabet_size penalty table
 * indexed [query code][reference code] (SubstitutionMatrix::penalty), or NULL for unit cost (0
 * for equal codes, 1 otherwise). Outputs per pair: score[p], the ops in forward order at
 * ops + p * ops_stride ('M' equal codes, 'S' substitution, 'D' query residue against a gap,
 * 'I' reference residue against a gap; ops_stride >= q_max_len + r_max_len) and their count
 * n_ops[p]; the alignment is global, so it starts at residue 0 of both sides. Negative gap
 * costs are ST_INVALID_ARGUMENT with the reference's message. The results equal the
 * reference's wherever its int32 arithmetic stays finite (every cost below INT32_MAX / 4),
 * beyond which the reference's own walk reads outside its strings. */
int32_t st_index_align_pairs(const uint8_t* q_codes, const uint64_t* q_offsets,
                             uint64_t n_queries, const uint8_t* r_codes,
                             const uint64_t* r_offsets, uint64_t n_refs,
                             const uint32_t* pair_query, const uint32_t* pair_ref,
                             uint64_t n_pairs, const int32_t* pen, uint32_t alphabet_size,
                             int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                             uint64_t ops_stride, int32_t* score, uint32_t* n_ops,
                             st_status* status);
/* Device-resident sets, pairs and outputs on `stream`, with the length-bound contract and the
 * one synchronisation of st_align_pairs_device; pen stays a host pointer (or NULL). */
int32_t st_index_align_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                    uint64_t n_queries, uint32_t q_max_len,
                                    const uint8_t* r_codes, const uint64_t* r_offsets,
                                    uint64_t n_refs, uint32_t r_max_len,
                                    const uint32_t* pair_query, const uint32_t* pair_ref,
                                    uint64_t n_pairs, const int32_t* pen, uint32_t alphabet_size,
                                    int32_t gap_open, int32_t gap_extend, uint8_t* ops,
                                    uint64_t ops_stride, int32_t* score, uint32_t* n_ops,
                                    void* stream, st_status* status);

/* ---- device-memory entry points ------------------------------------------------------
 * Same computations with the sequences (CSR bytes and offsets) and the output already in
 * device memory; the small tables (pen, sim, prior) stay host pointers. The work is enqueued
 * on `stream` (a hipStream_t, NULL for the null stream). q_max_len / r_max_len must bound
 * the longest sequence of each side. A call may synchronise `stream` once to read back a few
 * counters; its scratch is stream-ordered and released on `stream`. These are the same
 * kernels the host-memory entry points run, without the host<->device copies. */
int32_t st_gapblock_matrix_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                  uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                                  const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                                  const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                                  int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                                  int32_t* out, void* stream, st_status* status);

/* The allocation-free form of st_gapblock_matrix_device: same arguments, same results. It has
 * no counterpart in the reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Its scratch
 * (the device penalty table and prior cube, the packed references and the twin list) lives
 * in a static 32 MiB device array that each GPU the library runs on holds until the process
 * ends; no call allocates or frees device memory, and the tables go in with stream-ordered
 * copies. Sets whose scratch would exceed 32 MiB (roughly 700,000 references of up to 30
 * residues) return ST_INVALID_ARGUMENT before any work is enqueued. Every call on one GPU
 * shares that GPU's array and leaves its contents behind: calls that overlap in time on one
 * GPU (on different streams, or from different threads) race and must be serialised by the
 * caller; consecutive calls on one stream are ordered and safe. */
int32_t st_gapblock_matrix_device_noalloc(const uint8_t* q_codes, const uint64_t* q_offsets,
                                          uint64_t n_queries, uint32_t q_max_len,
                                          const uint8_t* r_codes, const uint64_t* r_offsets,
                                          uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                          uint32_t alphabet_size, int32_t gap_open,
                                          int32_t gap_extend, const int32_t* prior,
                                          uint32_t prior_width, int32_t* out, void* stream,
                                          st_status* status);

/* The device form of st_gapblock_threshold: hit_query / hit_ref / hit_score AND hit_count are
 * device memory. *hit_count is a device counter (zeroed by the call, then counted into by the
 * kernels); read it back once the stream's work is done. */
int32_t st_gapblock_threshold_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                     uint64_t n_queries, uint32_t q_max_len,
                                     const uint8_t* r_codes, const uint64_t* r_offsets,
                                     uint64_t n_refs, uint32_t r_max_len, const int32_t* pen,
                                     uint32_t alphabet_size, int32_t gap_open, int32_t gap_extend,
                                     const int32_t* prior, uint32_t prior_width, int32_t cutoff,
                                     uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                                     int32_t* hit_score, uint64_t* hit_count, void* stream,
                                     st_status* status);

int32_t st_gapblock_pairs_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                                 uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                                 const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                                 const int32_t* pen, uint32_t alphabet_size, int32_t gap_open,
                                 int32_t gap_extend, const int32_t* prior, uint32_t prior_width,
                                 const uint32_t* pair_query, const uint32_t* pair_ref,
                                 uint64_t n_pairs, int32_t* out, void* stream,
                                 st_status* status);

int32_t st_align_matrix_device(const uint8_t* q_codes, const uint64_t* q_offsets,
                               uint64_t n_queries, uint32_t q_max_len, const uint8_t* r_codes,
                               const uint64_t* r_offsets, uint64_t n_refs, uint32_t r_max_len,
                               const int32_t* sim, uint32_t alphabet_size, int32_t local,
                               int32_t dist, int32_t gap_open, int32_t gap_extend, int32_t* out,
                               void* stream, st_status* status);

int32_t st_hamming_matrix_device(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                                 uint32_t a_max_len, const uint8_t* b, const uint64_t* b_offsets,
                                 uint64_t n_b, uint32_t b_max_len, int32_t* out, void* stream,
                                 st_status* status);

int32_t st_levenshtein_matrix_device(const uint8_t* a, const uint64_t* a_offsets, uint64_t n_a,
                                     uint32_t a_max_len, const uint8_t* b,
                                     const uint64_t* b_offsets, uint64_t n_b, uint32_t b_max_len,
                                     int32_t* out, void* stream, st_status* status);

/* ---- ragged batch of gap-block jobs ---------------------------------------------------
 * Many independent gap-block matrices of uneven sizes in one persistent launch. No
 * counterpart in the reference (whose gapblock_matrix scores one job per call): each job is
 * what one st_gapblock_matrix_device call would compute, with its sequences and output in
 * device memory and its tables (pen, prior) on the host.
 *
 * st_gapblock_batch_plan_create uploads the tables, allocates the packed-reference buffers,
 * chooses the tile shape (autotune != 0 times the candidate shapes on this batch once per
 * process and workload signature; 0 takes the default or a memoised choice) and captures the
 * whole batch -- reference packing, work-counter reset and the persistent scoring kernel --
 * in a hipGraph. st_gapblock_batch_plan_run replays that graph on `stream`; the caller's
 * sequence buffers must stay valid and unchanged in size for the plan's lifetime. Their
 * contents may change between runs within each job's declared maximum lengths: every run
 * re-packs the references and re-pairs the queries by length (short jobs score two queries
 * of one length per table), so the results follow the current contents. */
typedef struct st_gapblock_job {
    const uint8_t* q_codes;
    const uint64_t* q_offsets;
    uint64_t n_queries;
    uint32_t q_max_len;
    const uint8_t* r_codes;
    const uint64_t* r_offsets;
    uint64_t n_refs;
    uint32_t r_max_len;
    const int32_t* pen;
    uint32_t alphabet_size;
    int32_t gap_open;
    int32_t gap_extend;
    const int32_t* prior;
    uint32_t prior_width;
    int32_t* out;
} st_gapblock_job;

typedef struct st_batch_plan st_batch_plan;

int32_t st_gapblock_batch_plan_create(const st_gapblock_job* jobs, uint32_t n_jobs,
                                      int32_t autotune, void* stream, st_batch_plan** plan,
                                      st_status* status);
int32_t st_gapblock_batch_plan_run(st_batch_plan* plan, void* stream, st_status* status);
/* The chosen tile shape and the number of work units of the plan. */
int32_t st_gapblock_batch_plan_tiles(const st_batch_plan* plan, uint32_t* queries_per_unit,
                                     uint32_t* threads, uint32_t* workgroups,
                                     uint64_t* work_units);
/* How many of the plan's jobs run twin rows (two queries of one length per 16-bit packed
 * table, chosen when the job's sums provably fit 16 bits). */
int32_t st_gapblock_batch_plan_twin_jobs(const st_batch_plan* plan, uint32_t* twin_jobs);
void st_gapblock_batch_plan_destroy(st_batch_plan* plan);

/* ---- library information ------------------------------------------------------------- */

/* Version string of this library. */
const char* st_version(void);

/* Number of visible HIP devices (0 when none); never faults on a machine without a GPU. */
int32_t st_device_count(void);

/* ---- pinned host buffers for matrix outputs (no counterpart in the reference) -----------
 * Page-locked host memory owned by the library. A host-memory entry point whose output lies
 * inside such a buffer transfers the matrix straight into it by DMA, with no staging copy;
 * any other output goes through the library's pinned staging slots. The reference returns a
 * std::vector<int32_t> that the Python ScoreMatrix owns (src/_bindings.cpp:203-209,
 * 470-479); this library's ScoreMatrix takes its storage from here. st_host_alloc fails
 * with ST_OUT_OF_MEMORY when the memory cannot be pinned (the caller then uses ordinary
 * memory), st_host_free accepts only pointers st_host_alloc returned (and NULL). */
int32_t st_host_alloc(uint64_t bytes, void** ptr, st_status* status);
void st_host_free(void* ptr);

/* ---- learned prefilter (no counterpart in the reference) -------------------------------
 * A 1D-conv encoder (embedding 32, three width-5 convolutions of 128 channels with ReLU,
 * mean pool, 128-d output; BF16 operands, FP32 accumulation) whose squared embedding
 * distance |e(a) - e(b)|^2 = s(a,a) + s(b,b) - 2 s(a,b) screens (query, reference) pairs
 * before exact gap-block rescoring. Sequences are codec codes as CSR in device memory, at most
 * 64 residues each. All pointers below are device pointers unless noted; the calls are
 * enqueued on `stream` (NULL: the null stream) and the ones returning host values synchronise
 * it. A model is used by one thread at a time. */
typedef struct st_prefilter st_prefilter;
uint64_t st_prefilter_param_count(void);
int32_t st_prefilter_create(uint64_t seed, st_prefilter** model, st_status* status);
void st_prefilter_destroy(st_prefilter* model);
/* FP32 parameters to / from host memory (count = st_prefilter_param_count()). */
int32_t st_prefilter_get_params(const st_prefilter* model, float* params, uint64_t count,
                                st_status* status);
int32_t st_prefilter_set_params(st_prefilter* model, const float* params, uint64_t count,
                                st_status* status);
/* The gradient of the last train step (host copy), for checking the backward pass. */
int32_t st_prefilter_get_grads(const st_prefilter* model, float* grads, uint64_t count,
                               st_status* status);
/* One Adam step on the pair loss sum_t weight[t] * (|e(a_t) - e(b_t)|^2 - target[t])^2 over
 * the n_seqs sequences given (pair_a / pair_b index them). *loss_out (host) receives the mean
 * loss before the step. */
int32_t st_prefilter_train_step(st_prefilter* model, const uint8_t* codes,
                                const uint64_t* offsets, uint32_t n_seqs, uint32_t max_len,
                                const uint32_t* pair_a, const uint32_t* pair_b,
                                const float* target, const float* weight, uint32_t n_pairs,
                                float lr, float* loss_out, void* stream, st_status* status);
/* Embeddings (n x 128 BF16, row-major) and their squared norms (n floats). */
int32_t st_prefilter_encode(const st_prefilter* model, const uint8_t* codes,
                            const uint64_t* offsets, uint32_t n, uint32_t max_len, void* emb,
                            float* norm, void* stream, st_status* status);
/* Every (query, reference) with |q|^2 + |r|^2 - 2 q.r <= tau (BF16 MFMA, FP32 accumulation),
 * in no particular order; the first `capacity` go to hit_query / hit_ref and *hit_count (host)
 * receives the total, which may exceed capacity (call again with more room). */
int32_t st_prefilter_screen(const void* q_emb, const float* q_norm, uint32_t n_queries,
                            const void* r_emb, const float* r_norm, uint32_t n_refs, float tau,
                            uint64_t capacity, uint32_t* hit_query, uint32_t* hit_ref,
                            uint64_t* hit_count, void* stream, st_status* status);

/* The model distance |q|^2 + |r|^2 - 2 q.r (FP32 sum of the BF16 products) of each listed
 * (query, reference) pair of two encoded sets, into distance[t] (device): what the screen
 * estimates for that pair, used to calibrate its threshold on the exact hits of a sample. */
int32_t st_prefilter_pair_distance(const void* q_emb, const float* q_norm, const void* r_emb,
                                   const float* r_norm, const uint32_t* pair_query,
                                   const uint32_t* pair_ref, uint64_t n_pairs, float* distance,
                                   void* stream, st_status* status);

/* Indices t < n with values[t] <= cutoff (device int32 values), in no particular order; the
 * first `capacity` go to index (device) and *count (host) receives the total. Used to keep the
 * exactly rescored prefilter candidates that are real hits. */
int32_t st_select_le(const int32_t* values, uint64_t n, int32_t cutoff, uint64_t capacity,
                     uint32_t* index, uint64_t* count, void* stream, st_status* status);
/* The same selection gathering (pair_a[t], pair_b[t], values[t]) into out_a / out_b /
 * out_value (device) instead of writing indices. */
int32_t st_select_pairs_le(const int32_t* values, const uint32_t* pair_a, const uint32_t* pair_b,
                           uint64_t n, int32_t cutoff, uint64_t capacity, uint32_t* out_a,
                           uint32_t* out_b, int32_t* out_value, uint64_t* count, void* stream,
                           st_status* status);

/* ---- device memory (no counterpart in the reference) ------------------------------------
 * Minimal allocation and synchronous copies so a caller without its own GPU allocator can
 * use the device-memory entry points. kind: 1 host->device, 2 device->host, 3 device->device.
 * st_synchronize waits for all work on the device. */
int32_t st_device_alloc(uint64_t bytes, void** ptr, st_status* status);
int32_t st_device_free(void* ptr);
int32_t st_memcpy(void* dst, const void* src, uint64_t bytes, int32_t kind, st_status* status);
int32_t st_synchronize(st_status* status);

/* ---- instrumentation (no counterpart in the reference; used by the benchmark) ---------
 * st_timing_enable(1) makes every kernel launch record a pair of hipEvents on the stream it
 * runs on. st_timing_collect waits for the recorded events and writes one line per kernel
 * label, "label launches total_ms", into report (NUL-terminated, truncated to capacity),
 * then forgets them. It returns the number of labels, or minus the number of unresolved
 * event pairs. */
void st_timing_enable(int32_t on);
int32_t st_timing_collect(char* report, uint64_t capacity);

/* Measured VALU issue rate of one instruction kind (0 v_add_u32, 1 v_max_i16, 2 v_max_i32,
 * 3 v_pk_max_i16, 4 v_max3_i32, 5 v_bfi_b32) over the whole device, in lane-instructions per
 * second: the measured peaks the VALU-bound kernels' rooflines are checked against. Writes the
 * instruction's name into name (NUL-terminated, truncated to name_capacity); an unknown kind
 * is ST_INVALID_ARGUMENT. */
int32_t st_probe_valu(int32_t kind, char* name, uint64_t name_capacity, double* lane_ops_per_s,
                      st_status* status);
/* Measured rate of v_mfma_f32_32x32x16_bf16 (four independent chains per wave, registers
 * only) over the whole device, in FLOP/s: the attainable BF16 MFMA peak. */
int32_t st_probe_mfma(double* flops_per_s, st_status* status);
/* Measured HBM write rate of a plain 16-byte-store fill of `bytes` device bytes (at least
 * 1 MiB, a multiple of 16), in bytes per second: the attainable write bandwidth the
 * write-bound kernels' rooflines are checked against. */
int32_t st_probe_hbm_write(uint64_t bytes, double* bytes_per_s, st_status* status);
/* Measured LDS read rate over the whole device, in bytes per second: 16 workgroups of 256
 * threads per CU, each thread reading `dwords` (1, 2 or 4) consecutive dwords per load in a
 * bank-conflict-free pattern. The attainable LDS bandwidth the table-lookup kernels'
 * rooflines are checked against; other dword counts are ST_INVALID_ARGUMENT. */
int32_t st_probe_lds(int32_t dwords, double* bytes_per_s, st_status* status);

#ifdef __cplusplus
}
#endif

#endif /* SEQTREE_HIP_H */
