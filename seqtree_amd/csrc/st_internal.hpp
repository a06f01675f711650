// Internal declarations shared by the HIP translation units of libseqtree_hip.so.
//
// Data layout in HBM (one call):
//   * query side: the caller's CSR (codes + uint64 offsets), read directly; a workgroup owns
//     one query, so its residues are read as uniform (scalar / broadcast) values.
//   * reference side: packed once per call by pack_refs into
//       words[(nwords + 1) * npad]   word w (residues 4w..4w+3, little-endian bytes) of slot s
//                                    at words[w * npad + s]: a wave reading word w of 64
//                                    consecutive slots issues one 256-byte coalesced load;
//                                    row nwords is an all-zero pad row;
//       len[npad], orig[npad]        slot length and original column index (kNoSlot = pad).
//     Slots are grouped in chunks of kChunk consecutive references; inside a chunk they are
//     sorted by length so that the 64 lanes of a wave see (almost) equal lengths and the DP
//     loops stay convergent. A scoring workgroup owns (query, chunk), stages its kChunk
//     results in LDS by original column and writes them back as one coalesced row segment.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace st {

constexpr uint32_t kChunk = 1024;        // reference slots per scoring workgroup
constexpr uint32_t kThreads = 256;       // threads per scoring workgroup (4 wave64s)
constexpr uint32_t kSlotsPerThread = kChunk / kThreads;
constexpr uint32_t kHistFusedBins = 1024;  // LDS histogram bins per query row of a workgroup
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;
constexpr uint32_t kMaxAlphabet = 32;    // codec sizes are <= 32 (seqtree.hpp:15-16)

struct PackedRefs {
    const uint32_t* words;
    const uint32_t* len;
    const uint32_t* orig;
    uint32_t npad;
    uint32_t nwords;
    uint32_t nchunks;
};

// ---- prep.hip --------------------------------------------------------------------------
// scale multiplies every byte (1: raw bytes / codes, zero past the end; 4: codes as LDS byte
// offsets, the gap-block kernels' form -- codes are < 32, so code * 4 still fits a byte --
// with code 31 (byte 124) past the end).
// max_len bounds the reference lengths (below 256: a counting sort by length). When
// tw.twins is set the same launch also builds that query set's twin list (below).
struct TwinJob {
    const uint64_t* off;
    uint64_t n;
    uint32_t* sorted;
    uint2* twins;
};
void launch_pack_refs(const uint8_t* bytes, const uint64_t* off, uint64_t n, uint32_t nwords,
                      uint32_t npad, uint32_t* words, uint32_t* len, uint32_t* orig,
                      uint32_t scale, uint32_t max_len, const TwinJob& tw, hipStream_t s);

// Buckets the n queries by length: class c receives every query with
// length <= bounds[c] (and > bounds[c-1]); the last bound must be 0xFFFFFFFF.
constexpr int kMaxClasses = 24;
struct ClassBounds {
    uint32_t bound[kMaxClasses];
    int count;
};
void launch_classify(const uint64_t* off, uint64_t n, const ClassBounds& b, uint32_t* counts,
                     uint32_t* lists, hipStream_t s);

void launch_fill_u32(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s);

// Adds each row of the rows x cols int32 matrix m to its histogram out[row][H + 2]: bin 0
// counts values below zero, bin 1 + v the value v in [0, H), bin H + 1 values >= H.
// H + 2 <= kMaxHistBins. rows < 65536 per launch.
constexpr uint32_t kMaxHistBins = 16384;
void launch_score_histogram(const int32_t* m, uint64_t rows, uint64_t cols, uint32_t H,
                            unsigned long long* out, hipStream_t s);

// Twin list of a query set whose lengths are at most kTwinMaxLen (built by a packing launch
// given a TwinJob): the queries grouped by length and paired inside each length, a length
// with an odd count ending in a query paired with itself; entries past the twin count are
// (kNoSlot, kNoSlot). `sorted` is n entries of scratch; `twins` holds twin_capacity(n)
// entries.
constexpr uint32_t kTwinMaxLen = 32;
__host__ __device__ inline uint64_t twin_capacity(uint64_t n) { return (n + kTwinMaxLen + 1) / 2; }

// ---- prefilter.hip ---------------------------------------------------------------------
namespace pf {
constexpr uint32_t kEmbDim = 32;         // embedding width
constexpr uint32_t kChannels = 128;      // conv channels and output width
constexpr uint32_t kEmbRows = kMaxAlphabet;
constexpr uint32_t kMaxLen = 64;         // longest sequence the encoder takes
// FP32 parameter vector: emb [32][32], then per layer W [5 * cin][128] (row kk = k * cin + ci)
// and b [128].
constexpr uint32_t kOffEmb = 0;
constexpr uint32_t kOffW1 = kOffEmb + kEmbRows * kEmbDim;
constexpr uint32_t kOffB1 = kOffW1 + 5 * kEmbDim * kChannels;
constexpr uint32_t kOffW2 = kOffB1 + kChannels;
constexpr uint32_t kOffB2 = kOffW2 + 5 * kChannels * kChannels;
constexpr uint32_t kOffW3 = kOffB2 + kChannels;
constexpr uint32_t kOffB3 = kOffW3 + 5 * kChannels * kChannels;
constexpr uint32_t kParams = kOffB3 + kChannels;
// BF16 operand copies (element offsets; every table starts on a 16-byte boundary):
// emb, Wt_l [128][5 cin] for the forward GEMMs, Wf_l [cin][5 * 128] for the data gradients.
constexpr uint32_t kOffEmbB = 0;
constexpr uint32_t kOffWt1 = kOffEmbB + kEmbRows * kEmbDim;
constexpr uint32_t kOffWt2 = kOffWt1 + 5 * kEmbDim * kChannels;
constexpr uint32_t kOffWt3 = kOffWt2 + 5 * kChannels * kChannels;
constexpr uint32_t kOffWf1 = kOffWt3 + 5 * kChannels * kChannels;
constexpr uint32_t kOffWf2 = kOffWf1 + 5 * kEmbDim * kChannels;
constexpr uint32_t kOffWf3 = kOffWf2 + 5 * kChannels * kChannels;
constexpr uint32_t kBf16Params = kOffWf3 + 5 * kChannels * kChannels;

// Sequence s of a set occupies rows (off[s] - off[0]) + 2 s onward (see prefilter.hip).
// tile_first (or nullptr): per 128-row tile of the fused encoder, the sequence whose span holds
// the tile's first row (tiles past the used rows keep their prior contents).
void launch_embed(const uint8_t* codes, const uint64_t* off, uint32_t n, const __bf16* emb,
                  __bf16* X0, uint8_t* valid, uint32_t* tile_first, hipStream_t s);
// layer 1..3; epi 0 forward (bias, relu, valid rows), 1 backward data masked by mask > 0,
// 2 backward data of the embedding layer (valid rows). R: rows, a multiple of 128.
void launch_conv(int layer, int epi, const __bf16* X, const __bf16* Bt, const float* bias,
                 const __bf16* mask, const uint8_t* valid, __bf16* Y, uint64_t R, hipStream_t s);
// Fused inference forward and mean pool over R rows (a multiple of 128): the BF16 embeddings
// eb and their squared norms for the n sequences of offsets off, from X0 and `valid`, with the
// BF16 operand copies B and the FP32 parameters P (biases). tile_first comes from
// launch_embed (tiles past the used rows hold n); part (R / 64 x 128 floats) and head_seq
// (R / 128) are scratch.
void launch_encode(const __bf16* X0, const uint8_t* valid, const __bf16* B, const float* P,
                   uint64_t R, const uint64_t* off, uint32_t n, const uint32_t* tile_first,
                   __bf16* eb, float* norm, float* part, uint32_t* head_seq, hipStream_t s);
void launch_pool(const __bf16* H, const uint64_t* off, uint32_t n, float* ef, __bf16* eb,
                 float* norm, hipStream_t s);
void launch_gram(const __bf16* Q, const float* q_norm, uint32_t nq, const __bf16* Rm,
                 const float* r_norm, uint32_t nr, float tau, uint32_t* hit_q, uint32_t* hit_r,
                 uint64_t cap, unsigned long long* count, uint32_t cus, hipStream_t s);
// Model distances |q|^2 + |r|^2 - 2 q.r of the pairs (Q[pq[t]], Rm[pr[t]]), FP32; n < 2^32 * 4.
void launch_pair_distance(const __bf16* Q, const float* q_norm, const __bf16* Rm,
                          const float* r_norm, const uint32_t* pq, const uint32_t* pr, uint64_t n,
                          float* out, hipStream_t s);
void launch_pair_loss(const float* ef, const uint32_t* pa, const uint32_t* pb, const float* y,
                      const float* w, uint32_t n_pairs, float* dE, float* loss, hipStream_t s);
void launch_pool_bwd(const __bf16* H, const float* dE, const uint64_t* off, uint32_t n,
                     __bf16* dY, hipStream_t s);
void launch_dw(int layer, const __bf16* X, const __bf16* dY, uint64_t R, float* dW, float* db,
               hipStream_t s);
void launch_embed_bwd(const __bf16* dX0, const uint8_t* codes, const uint64_t* off, uint32_t n,
                      float* dEmb, hipStream_t s);
void launch_adam(float* p, const float* g, float* m, float* v, uint32_t n, float lr, float b1,
                 float b2, float c1, float c2, float eps, hipStream_t s);
void launch_pack_weights(const float* P, __bf16* B, hipStream_t s);
void launch_gather_offsets(const uint64_t* src, uint32_t stride, uint32_t last, uint32_t count,
                           uint64_t* dst, hipStream_t s);
// Entries with v[t] <= cutoff: indices into out_a, or (pa, pb given) (pa[t], pb[t], v[t]).
void launch_select_le(const int32_t* v, uint64_t n, int32_t cutoff, const uint32_t* pa,
                      const uint32_t* pb, uint32_t* out_a, uint32_t* out_b, int32_t* out_v,
                      uint64_t cap, unsigned long long* count, hipStream_t s);
}  // namespace pf

// ---- probe.hip -------------------------------------------------------------------------
// VALU issue-rate probe: `blocks` workgroups of 256 threads, each thread 16 chains of
// `iters` instructions of the named kind (probe_name returns null past the last kind).
const char* probe_name(int kind);
void launch_probe(int kind, uint32_t* out, uint32_t blocks, uint32_t iters, hipStream_t s);
// v_mfma_f32_32x32x16_bf16 chains (256-thread workgroups); returns the FLOPs launched.
double launch_probe_mfma(float* out, uint32_t blocks, uint32_t iters, hipStream_t s);
void launch_probe_fill(void* p, uint64_t bytes, uint32_t blocks, uint32_t v, hipStream_t s);
// LDS read-rate probe with `dwords` (1, 2 or 4) per read; returns the bytes it reads.
double launch_probe_lds(int dwords, uint32_t* out, uint32_t blocks, uint32_t iters,
                        hipStream_t s);

// ---- gapblock.hip ----------------------------------------------------------------------
// Thresholded output: every pair with score <= cutoff is appended as (query, ref, score);
// count is a device counter that keeps counting past capacity (entries past it are dropped).
struct GapblockHits {
    int32_t cutoff;
    uint64_t capacity;
    uint32_t* query;
    uint32_t* ref;
    int32_t* score;
    unsigned long long* count;  // nullptr: dense output
};
struct GapblockArgs {
    const uint8_t* q_codes;
    const uint64_t* q_off;
    uint64_t n_queries;
    uint64_t n_refs;
    PackedRefs refs;
    const int32_t* pen;    // device, A*A
    uint32_t A;
    int32_t gap_open, gap_extend;
    const int32_t* prior;  // device cube or nullptr
    uint32_t w1;           // prior_width + 1
    int32_t* out;          // dense N x K output (or one score per pair in pair-list mode)
    GapblockHits hits;     // hits.count != nullptr selects the thresholded output
    // hist != nullptr selects the histogram output (before hits and the dense matrix):
    // hist[i][hist_bins + 2] += bin 0 scores < 0, bin 1 + s scores s < hist_bins, bin
    // hist_bins + 1 the rest; hist_bins + 2 <= kHistFusedBins
    unsigned long long* hist;
    uint32_t hist_bins;
    // short launches (queries <= 32, references <= 64, A <= 31): profile rows per workgroup
    // (query length + 32 zero rows), prior-slice rows (reference lengths 0..r_max) and the
    // slice's past-the-end value
    uint32_t prof_rows;
    uint32_t slice_rows;
    int32_t big;
    uint32_t group;        // reference chunks per workgroup (>= 1)
};
// use_profile: stage a per-query profile (one table row per query position) in LDS.
// exact_wrap: evaluate the candidates in the reference's order with wrapping int32
// arithmetic (needed only when an int32 sum may overflow).
// short_path: the short-launch kernels (a.prof_rows / a.slice_rows / a.big set).
void launch_gapblock(const GapblockArgs& a, bool use_profile, uint32_t q_max_len, bool exact_wrap,
                     bool short_path, hipStream_t s);
// Short launch with two queries of one length per workgroup (the twin rows of
// gapblock_core.hpp; a.big holds the 16-bit sentinel): twins[0 .. n_twins), a twin list
// (see TwinJob).
void launch_gapblock_twin(const GapblockArgs& a, const uint2* twins, uint64_t n_twins,
                          hipStream_t s);
// Scores an explicit list of (query, ref) pairs into a.out[t]; references come as CSR.
void launch_gapblock_pairs(const GapblockArgs& a, const uint8_t* r_codes, const uint64_t* r_off,
                           const uint32_t* pair_query, const uint32_t* pair_ref, uint64_t n_pairs,
                           bool exact_wrap, hipStream_t s);

// ---- batch.hip -------------------------------------------------------------------------
// One job of a ragged gap-block batch, as the persistent kernel reads it.
struct BatchJob {
    const uint8_t* q_codes;
    const uint64_t* q_off;
    uint64_t n_queries;
    uint64_t n_refs;
    PackedRefs refs;
    const int32_t* pen;    // device, A*A
    uint32_t A;
    int32_t gap_open, gap_extend;
    const int32_t* prior;  // device cube or nullptr
    uint32_t w1;
    int32_t* out;          // device, n_queries * n_refs
    uint32_t profile;      // LDS profile (+ negation) rather than the A x A table
    uint32_t exact;        // wrapping-order evaluation
    uint32_t short_path;   // short-launch evaluation (as GapblockArgs)
    uint32_t prof_rows;
    uint32_t slice_rows;
    int32_t big;
    uint32_t group;        // reference chunks per work unit
    uint32_t qblocks;      // query blocks (of queries_per_unit) per chunk group
    uint64_t unit_begin;   // index of the job's first work unit
    // twin rows (short path, two equal-length queries per table, big = the 16-bit sentinel):
    // the job's twin list (twin_capacity(n_queries) entries, rebuilt on every run by the
    // job's packing launch) and its scratch; a query block is then a block of twins
    uint32_t twin;
    const uint2* twins;
    uint32_t* twin_sorted;
};
// Tile shape of the persistent launch: queries per work unit, threads per workgroup
// (256 or 512) and the number of resident workgroups.
struct BatchTiles {
    uint32_t queries_per_unit;
    uint32_t threads;
    uint32_t workgroups;
};
// lds_bytes: the chunk stage (stage_ints int32: kChunk, or 2 kChunk when a job has twin rows)
// followed by the largest job's tables. twin_only: every job has twin rows (a kernel with
// only that path).
void launch_gapblock_batch(const BatchJob* jobs, uint32_t n_jobs, uint64_t total_units,
                           const BatchTiles& tiles, uint32_t stage_ints, size_t lds_bytes,
                           bool twin_only, unsigned long long* next_unit, hipStream_t s);


// ---- gotoh.hip -------------------------------------------------------------------------
struct GotohArgs {
    const uint8_t* q_codes;
    const uint64_t* q_off;
    const uint8_t* r_codes;  // CSR, for the generic kernels
    const uint64_t* r_off;
    uint64_t n_queries;
    uint64_t n_refs;
    PackedRefs refs;
    const int32_t* sim;   // device, A*A   [query code][ref code]
    const int32_t* simT;  // device, A*A   transposed
    uint32_t A;
    int32_t local;
    int32_t dist;
    int32_t gap_open, gap_extend;
    const int64_t* q_self;  // dist only
    const int64_t* r_self;
    int32_t* out;
    int32_t* overflow;      // device flag, set to 1 when a result does not fit int32
};
// Register-stripe kernel for queries of length <= rows (rows a multiple of 4, <= 64); the
// query indices come from `list` (count entries). packed16: two queries per workgroup in
// int16 halves (needs the int16 range proof), else one query in int32 (int32 proof).
void launch_gotoh_fast(const GotohArgs& a, int rows, bool packed16, const uint32_t* list,
                       uint32_t count, hipStream_t s);
// Register kernel over stripes of kStripeRows query rows, for queries longer than one stripe
// (same range proofs with the query's padded length in the bound). Persistent grid of `grid`
// workgroups over gotoh_stripes_units(...) units; bnd holds gotoh_stripes_words(grid, W,
// gap_open >= gap_extend || packed16) words, W >= the longest reference.
// q_max: the longest query of the list (queries padded to at most kStripeFullRows rows keep
// their whole profile in LDS).
constexpr int kStripeRows = 32;
constexpr uint32_t kStripeFullRows = 256;
uint64_t gotoh_stripes_units(uint32_t count, bool packed16, uint32_t nchunks);
uint64_t gotoh_stripes_words(uint32_t grid, uint32_t W, bool hform);
void launch_gotoh_stripes(const GotohArgs& a, bool packed16, const uint32_t* list,
                          uint32_t count, uint32_t grid, uint32_t* bnd, uint32_t W,
                          uint32_t q_max, hipStream_t s);
// int64 per-thread DP for arbitrary lengths and gap costs: every (list[t], k) pair, k < n_refs.
// scratch holds `slots` regions of 6 * (stride) int64.
void launch_gotoh_generic(const GotohArgs& a, const uint32_t* list, uint32_t count,
                          int64_t* scratch, uint32_t slots, uint64_t stride, hipStream_t s);
// self scores s(x, x) for every sequence of a CSR set, into int64 res.
void launch_gotoh_self(const GotohArgs& a, const uint8_t* codes, const uint64_t* off, uint64_t n,
                       int64_t* res, int64_t* scratch, uint32_t slots, uint64_t stride,
                       hipStream_t s);
// Self-scores with the DP rows in LDS (strings of at most 160 residues and (2 max_len + 2) *
// max(|s|, gap costs) < 2^28: gotoh_self_fits_lds); the same results as launch_gotoh_self.
// Two CSR sets in one launch (either may be empty); max_len: the longest string of both
// (sizes the LDS rows); nb0 is set by the launcher.
struct SelfSets {
    const uint8_t* codes[2];
    const uint64_t* off[2];
    uint64_t n[2];
    int64_t* res[2];
    uint32_t nb0;
};
bool gotoh_self_fits_lds(uint32_t max_len, int64_t mx);
// Intra-pair wavefront kernel (one wave per pair, the query's rows spread over the lanes): for
// long queries when the pairs are too few to give every lane a pair of its own. Matrix mode
// (list != nullptr): every list[x] (x < count) against every reference, pairs count * n_refs,
// written to a.out like the register kernels (dist and overflow included). Self mode
// (list == nullptr): s(x, x) for the strings of two sets in one launch, the first n_first of
// the n_pairs from (codes[0], off[0]) into res[0], the rest from set 1 into res[1]. Needs
// gap_open >= gap_extend and (q_max + r_max + 2) * max(|s|, go, ge) < 2^28 (int32 cells);
// the matrix mode also needs the reference side as CSR (a.r_codes, a.r_off). bnd holds
// gotoh_wave_bnd_words(grid, q_max, r_max) int32 (none when one stripe covers q_max).
struct GotohWaveJob {
    const uint32_t* list;
    uint64_t n_pairs;
    uint32_t q_max, r_max;  // length bounds (pairs beyond them are skipped, not scored)
    const uint8_t* codes[2];
    const uint64_t* off[2];
    int64_t* res[2];
    uint64_t n_first;
};
uint32_t gotoh_wave_rows(uint32_t q_max);
uint64_t gotoh_wave_bnd_words(uint32_t grid, uint32_t q_max, uint32_t r_max);
void launch_gotoh_wave(const GotohArgs& a, const GotohWaveJob& job, uint32_t q_max,
                       uint32_t r_max, uint32_t grid, int32_t* bnd, hipStream_t s);
void launch_gotoh_self_lds(const GotohArgs& a, const SelfSets& sets, uint32_t max_len,
                           hipStream_t s);

// ---- traceback.hip ---------------------------------------------------------------------
// Gotoh alignment with traceback of (pair_q[p], pair_r[p]) for p < n_pairs: score[p] (int64),
// the ops in forward order at ops + p * ops_stride (ops_stride >= q_max + r_max), their count
// and the first aligned residue of each side.
struct TracebackArgs {
    const uint8_t* q_codes;
    const uint64_t* q_off;
    const uint8_t* r_codes;
    const uint64_t* r_off;
    const uint32_t* pair_q;
    const uint32_t* pair_r;
    uint64_t n_pairs;
    const int32_t* sim;  // device, A*A [query code][ref code]
    uint32_t A;
    int32_t local;
    int32_t gap_open, gap_extend;
    uint32_t q_max, r_max;  // bounds of the sequence lengths
    uint8_t* ops;
    uint64_t ops_stride;
    int64_t* score;
    uint32_t* n_ops;
    uint32_t* q_begin;
    uint32_t* r_begin;
    const uint32_t* order;  // processing order of the pairs (a permutation), or nullptr
    // minus infinity of the int64 cells: INT64_MIN / 4 for align_pair's similarity DP,
    // -(INT32_MAX / 4) for the penalty form (the negation of Index::align's int32 kInf)
    int64_t neg_inf;
    // penalty form (Index::align, global only; sim holds the negated penalties): the pair's
    // penalty -score goes to score32[p] as int32 and score is not written
    int32_t* score32;
    // device flag, set to 1 by a pair longer than q_max or r_max (the pair is skipped: no ops,
    // score 0)
    uint32_t* too_long;
};
// Longest reference whose DP rows fit the LDS (int32 cells).
constexpr uint32_t kTracebackLdsMaxLen = 64;
// wide: int64 cells (else int32, valid under the host's range bound); lds_rows: int32 rows in
// LDS (r_max <= kTracebackLdsMaxLen), else in `rows` (3 (r_max + 1) 64 cells per workgroup);
// dirs: (q_max + 1) (r_max + 1) 64 bytes per workgroup.
void launch_traceback(const TracebackArgs& a, bool wide, bool lds_rows, uint32_t workgroups,
                      void* rows, uint8_t* dirs, hipStream_t s);
// Register variant (int32 cells): references of at most kTbRegCols residues and queries of at
// most kTbRegMaxQ; dirs holds (q_max + 1) * kTbRegGroups * 64 dwords per workgroup.
constexpr uint32_t kTbRegCols = 32;
constexpr uint32_t kTbRegGroups = (kTbRegCols + 1 + 3) / 4;
constexpr uint32_t kTbRegMaxQ = 64;
void launch_traceback_reg(const TracebackArgs& a, uint32_t workgroups, uint32_t* dirs,
                          hipStream_t s);
// Nibble variant of the register kernel (gap_open >= gap_extend): dirs holds
// q_max * kTbNibDwords * 64 dwords per workgroup.
constexpr uint32_t kTbNibDwords = (kTbRegCols + 7) / 8;
void launch_traceback_nib(const TracebackArgs& a, uint32_t workgroups, uint32_t* dirs,
                          hipStream_t s);
// Shape order for the traceback (n_pairs < 2^32): counting sort of the pairs by their
// clamped (query, reference) lengths into order[n_pairs], over traceback_order_slices(n)
// contiguous slices; table holds (slices + 1) * kTracebackShapeBins counters, keys n_pairs.
constexpr uint32_t kTracebackShapeBins = 1024;
constexpr uint32_t kTracebackShapeSlices = 256;
uint32_t traceback_order_slices(uint64_t n_pairs);
void launch_traceback_order(const TracebackArgs& a, uint32_t* table, uint32_t* keys,
                            uint32_t* order, hipStream_t s);

// ---- distance.hip ----------------------------------------------------------------------
struct DistanceArgs {
    const uint8_t* a_bytes;
    const uint64_t* a_off;
    const uint8_t* b_bytes;  // CSR, for the generic kernel
    const uint64_t* b_off;
    uint64_t n_a;
    uint64_t n_b;
    PackedRefs refs;
    int32_t* out;
};
// Queries packed row-major as a_words[i * nwords + w] (equal lengths, zero padded).
void launch_pack_rows(const uint8_t* bytes, const uint64_t* off, uint64_t n, uint32_t nwords,
                      uint32_t* words, hipStream_t s);
void launch_hamming(const DistanceArgs& a, const uint32_t* a_words, uint32_t nwords,
                    hipStream_t s);
// Locates the first length-mismatched pair in row-major order; res = {i, k} or {~0, ~0}.
void launch_hamming_check(const uint64_t* a_off, uint64_t n_a, const uint64_t* b_off,
                          uint64_t n_b, uint64_t* res, hipStream_t s);
// Myers bit-parallel Levenshtein for queries of length <= 64 * words (words in 1..4).
void launch_levenshtein_myers(const DistanceArgs& a, int words, const uint32_t* list,
                              uint32_t count, hipStream_t s);
// Two-row DP, one pair per thread, for arbitrary lengths.
void launch_levenshtein_generic(const DistanceArgs& a, const uint32_t* list, uint32_t count,
                                int32_t* scratch, uint32_t slots, uint64_t stride,
                                hipStream_t s);

// ---- single pairs (single.hip): the scalar entry points ---------------------------------
// One wave per call, the pair in the kernel arguments, the results in pinned host memory.
// Gotoh and Levenshtein rows are the query (one per lane, 1..kSingleMaxQ residues), columns
// the reference (1..kSingleMaxR); Hamming takes up to kSingleMaxR positions of either.
constexpr uint32_t kSingleMaxQ = 64;
constexpr uint32_t kSingleMaxR = 256;
struct SingleOut {
    int64_t score;  // Gotoh similarity (or minus the penalty), distances
    uint32_t n_ops, q_begin, r_begin, unused;
    uint8_t ops[kSingleMaxQ + kSingleMaxR];  // traceback ops in forward order
};
struct SingleArgs {
    uint8_t q[kSingleMaxR];
    uint8_t r[kSingleMaxR];
    uint32_t m, n;
    const int32_t* sim;  // device A x A table [query code][reference code] (Gotoh)
    uint32_t A;
    int32_t go, ge;
    SingleOut* out;  // pinned host memory
};
enum class SingleKind {
    GotohGlobal, GotohLocal, TracebackGlobal, TracebackLocal, Levenshtein, Hamming
};
void launch_single(SingleKind kind, const SingleArgs& a, hipStream_t s);

}  // namespace st
