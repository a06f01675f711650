// The reference's scoring functions, with their validation order, exception types and
// messages, in front of the GPU C-ABI (seqtree_hip.h). Host work is limited to validation,
// encoding and CSR packing; all scoring happens on the device.
//
// Reference behaviour followed here:
//   gapblock_matrix     src/gapblock.cpp:12-22 (encode), 66-103 (checks, order, early return)
//   align_*_matrix      src/pairwise.cpp:51-68 (encode, check), 273-286, 338-352
//   align_score         src/pairwise.cpp:131-139
//   align_pair          src/pairwise.cpp:141-267 (traceback on the device; the empty-input
//                       case and the narrowing to int32 on the host, as the reference)
//   hamming/levenshtein src/distance.cpp:68-109
//   Index::build        src/index.cpp:93-97, src/trie.cpp:43-70 (stored strings, symbol check)
//   Index::align        src/index.cpp:105-179 (checks and their order; the traceback on the
//                       device in penalty form, the aligned strings built here from the raw
//                       characters, as the reference does)
#include "seqtree.hpp"
#include "seqtree_hip.h"

#include <algorithm>
#include <array>
#include <climits>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace seqtree {

namespace {

// Strings packed back to back with N+1 offsets, the form the C-ABI takes.
struct Csr {
    std::vector<uint8_t> bytes;
    std::vector<uint64_t> off;
    uint32_t max_len = 0;
};

// Codes of every character of `set` through `codec`, back to back, with the set's own offsets.
// One table lookup per character; only when some character is outside the alphabet does a
// second pass find the first such character in set order and call `bad(ch)`, which throws.
// (Codes are below 32, so the top bit of the OR of all codes is set by kInvalid alone.)
template <class Bad>
Csr pack(const StringSet& set, const Codec& codec, Bad&& bad) {
    Csr c;
    const size_t n = set.total();
    c.bytes.resize(std::max<size_t>(n, 1));
    const char* ch = set.chars();
    uint8_t* out = c.bytes.data();
    uint8_t seen = 0;
    for (size_t t = 0; t < n; ++t) {
        const uint8_t code = codec.encode(ch[t]);
        out[t] = code;
        seen |= code;
    }
    if (seen & 0x80)
        for (size_t t = 0; t < n; ++t)
            if (out[t] == Codec::kInvalid) bad(ch[t]);
    c.off.assign(set.offsets(), set.offsets() + set.size() + 1);
    c.max_len = set.max_len();
    return c;
}

// Encoding for gap-block: "symbol 'x' is not in the alphabet" (gapblock.cpp:12-22).
Csr pack_gapblock(const StringSet& v, const Codec& codec) {
    return pack(v, codec, [](char ch) {
        throw std::invalid_argument(std::string("symbol '") + ch + "' is not in the alphabet");
    });
}

// Encoding for the Gotoh paths: "invalid symbol 'x' in sequence" (pairwise.cpp:51-60).
Csr pack_pairwise(const StringSet& v, const Codec& codec) {
    return pack(v, codec, [](char ch) {
        throw std::invalid_argument(std::string("invalid symbol '") + ch + "' in sequence");
    });
}

// A one-string set.
StringSet one(std::string_view s) {
    StringSet set;
    set.push_back(s);
    return set;
}

// The raw bytes of a set (the distance kernels compare bytes, not codes), and the offsets of a
// single string.
const uint8_t* raw(const StringSet& s) { return reinterpret_cast<const uint8_t*>(s.chars()); }
const uint8_t* raw(std::string_view s) {
    static const uint8_t kEmpty = 0;
    return s.empty() ? &kEmpty : reinterpret_cast<const uint8_t*>(s.data());
}
std::array<uint64_t, 2> offsets(std::string_view s) { return {0, s.size()}; }

void raise(int code, const st_status& st) {
    switch (code) {
        case ST_OK: return;
        case ST_INVALID_ARGUMENT: throw std::invalid_argument(st.message);
        case ST_OVERFLOW: throw std::overflow_error(st.message);
        case ST_OUT_OF_MEMORY: throw std::bad_alloc();
        default: throw std::runtime_error(st.message);
    }
}

void check_pairwise(const SubstitutionMatrix& mat, const Codec& codec, int32_t open,
                    int32_t extend) {
    if (open < 0 || extend < 0)
        throw std::invalid_argument("seqtree: gap_open and gap_extend must be >= 0 (they are "
                                    "magnitudes, and are subtracted)");
    if (mat.size() != codec.size())
        throw std::invalid_argument("matrix size does not match the alphabet");
}

void pairwise_matrix(const StringSet& queries, const StringSet& refs, const SubstitutionMatrix& mat,
                     Alphabet alphabet, AlignMode mode, int32_t go, int32_t ge, bool dist,
                     int32_t* out) {
    const Codec codec(alphabet);
    check_pairwise(mat, codec, go, ge);
    const size_t N = queries.size(), K = refs.size();
    if (N == 0 || K == 0) return;
    const Csr q = pack_pairwise(queries, codec);
    const Csr r = pack_pairwise(refs, codec);
    st_status st{};
    const int rc = st_align_matrix(q.bytes.data(), q.off.data(), N, r.bytes.data(), r.off.data(),
                                   K, mat.similarities().data(), mat.size(),
                                   mode == AlignMode::Local, dist, go, ge, out, &st);
    raise(rc, st);
}

}  // namespace

int32_t align_score(std::string_view query, std::string_view ref, const SubstitutionMatrix& mat,
                    Alphabet alphabet, AlignMode mode, int32_t gap_open, int32_t gap_extend) {
    const Codec codec(alphabet);
    check_pairwise(mat, codec, gap_open, gap_extend);
    const Csr q = pack_pairwise(one(query), codec);
    const Csr r = pack_pairwise(one(ref), codec);
    int32_t out = 0;
    st_status st{};
    const int rc = st_align_matrix(q.bytes.data(), q.off.data(), 1, r.bytes.data(), r.off.data(), 1,
                                   mat.similarities().data(), mat.size(),
                                   mode == AlignMode::Local, 0, gap_open, gap_extend, &out, &st);
    raise(rc, st);
    return out;
}

namespace {

// The reference's narrow() for Alignment::score (pairwise.cpp:40-50).
int32_t narrow_score(int64_t v) {
    if (v < INT32_MIN || v > INT32_MAX)
        throw std::overflow_error("seqtree: the alignment score does not fit in 32 bits -- the "
                                  "gap costs are absurdly large for these sequences");
    return int32_t(v);
}

// The Alignment the reference builds from a traceback: decoded residues in op order, or --
// when a side is empty -- the caller's own characters (pairwise.cpp:150-165).
Alignment build_alignment(std::string_view query, std::string_view ref, const Codec& codec,
                          const uint8_t* q, const uint8_t* r, int64_t score, const uint8_t* ops,
                          uint32_t n_ops, uint32_t qb, uint32_t rb) {
    Alignment al;
    al.score = narrow_score(score);
    const bool empty = query.empty() || ref.empty();
    al.ops.assign(reinterpret_cast<const char*>(ops), n_ops);
    al.aligned_query.reserve(n_ops);
    al.aligned_ref.reserve(n_ops);
    uint32_t i = qb, j = rb;
    for (uint32_t t = 0; t < n_ops; ++t) {
        const char o = char(ops[t]);
        if (o == 'M' || o == 'S') {
            al.aligned_query += empty ? query[i] : codec.decode(q[i]);
            al.aligned_ref += empty ? ref[j] : codec.decode(r[j]);
            ++i;
            ++j;
        } else if (o == 'D') {
            al.aligned_query += empty ? query[i] : codec.decode(q[i]);
            al.aligned_ref += '-';
            ++i;
        } else {
            al.aligned_query += '-';
            al.aligned_ref += empty ? ref[j] : codec.decode(r[j]);
            ++j;
        }
    }
    return al;
}

}  // namespace

std::vector<Alignment> align_pairs(const StringSet& queries, const StringSet& refs,
                                   const std::vector<uint32_t>& pair_query,
                                   const std::vector<uint32_t>& pair_ref,
                                   const SubstitutionMatrix& mat, Alphabet alphabet,
                                   AlignMode mode, int32_t gap_open, int32_t gap_extend) {
    const Codec codec(alphabet);
    check_pairwise(mat, codec, gap_open, gap_extend);
    if (pair_query.size() != pair_ref.size())
        throw std::invalid_argument("seqtree: pair_query and pair_ref differ in length");
    const size_t P = pair_query.size();
    std::vector<Alignment> out(P);
    if (P == 0) return out;
    const Csr q = pack_pairwise(queries, codec);
    const Csr r = pack_pairwise(refs, codec);
    const uint64_t stride = uint64_t(q.max_len) + r.max_len + 1;
    std::vector<uint8_t> ops(P * stride);
    std::vector<int64_t> score(P);
    std::vector<uint32_t> n_ops(P), qb(P), rb(P);
    st_status st{};
    raise(st_align_pairs(q.bytes.data(), q.off.data(), queries.size(), r.bytes.data(),
                         r.off.data(), refs.size(), pair_query.data(), pair_ref.data(), P,
                         mat.similarities().data(), mat.size(), mode == AlignMode::Local,
                         gap_open, gap_extend, ops.data(), stride, score.data(), n_ops.data(),
                         qb.data(), rb.data(), &st),
          st);
    for (size_t p = 0; p < P; ++p) {
        const uint32_t a = pair_query[p], b = pair_ref[p];
        out[p] = build_alignment(queries[a], refs[b], codec, q.bytes.data() + q.off[a],
                                 r.bytes.data() + r.off[b], score[p], ops.data() + p * stride,
                                 n_ops[p], qb[p], rb[p]);
    }
    return out;
}

Alignment align_pair(std::string_view query, std::string_view ref, const SubstitutionMatrix& mat,
                     Alphabet alphabet, AlignMode mode, int32_t gap_open, int32_t gap_extend) {
    return align_pairs(one(query), one(ref), {0u}, {0u}, mat, alphabet, mode,
                       gap_open, gap_extend)[0];
}

void align_score_matrix_into(const StringSet& queries, const StringSet& refs,
                             const SubstitutionMatrix& mat,
                             Alphabet alphabet, AlignMode mode, int32_t gap_open,
                             int32_t gap_extend, int32_t* out) {
    pairwise_matrix(queries, refs, mat, alphabet, mode, gap_open, gap_extend, false, out);
}

void align_dist_matrix_into(const StringSet& queries, const StringSet& refs,
                            const SubstitutionMatrix& mat,
                            Alphabet alphabet, AlignMode mode, int32_t gap_open,
                            int32_t gap_extend, int32_t* out) {
    pairwise_matrix(queries, refs, mat, alphabet, mode, gap_open, gap_extend, true, out);
}

std::vector<int32_t> align_score_matrix(const std::vector<std::string>& queries,
                                        const std::vector<std::string>& refs,
                                        const SubstitutionMatrix& mat, Alphabet alphabet,
                                        AlignMode mode, int32_t gap_open, int32_t gap_extend,
                                        int /*threads*/) {
    std::vector<int32_t> out(queries.size() * refs.size());
    align_score_matrix_into(queries, refs, mat, alphabet, mode, gap_open, gap_extend, out.data());
    return out;
}

std::vector<int32_t> align_dist_matrix(const std::vector<std::string>& queries,
                                       const std::vector<std::string>& refs,
                                       const SubstitutionMatrix& mat, Alphabet alphabet,
                                       AlignMode mode, int32_t gap_open, int32_t gap_extend,
                                       int /*threads*/) {
    std::vector<int32_t> out(queries.size() * refs.size());
    align_dist_matrix_into(queries, refs, mat, alphabet, mode, gap_open, gap_extend, out.data());
    return out;
}

namespace {

// Validation and encoding shared by the gap-block entry points, in the reference's order
// (gapblock.cpp:71-102): gap costs, matrix size, [empty -> nothing to do], query symbols,
// ref symbols, prior size, prior width.
struct GapblockInput {
    Csr q, r;
    std::vector<int32_t> pen;
    uint8_t A = 0;
    bool empty = false;
};

GapblockInput gapblock_input(const StringSet& queries, const StringSet& refs, Alphabet alphabet,
                             const SubstitutionMatrix* matrix, int32_t gap_open,
                             int32_t gap_extend, const std::vector<int32_t>& prior,
                             uint32_t prior_width) {
    if (gap_open < 0 || gap_extend < 0)
        throw std::invalid_argument("gap_open and gap_extend must be >= 0");
    const Codec codec(alphabet);
    GapblockInput in;
    in.A = codec.size();
    const uint8_t A = in.A;
    if (matrix && matrix->size() != A)
        throw std::invalid_argument("matrix size does not match the alphabet");
    if (queries.empty() || refs.empty()) {
        in.empty = true;
        return in;
    }
    in.pen.resize(size_t(A) * A);
    for (uint32_t a = 0; a < A; ++a)
        for (uint32_t b = 0; b < A; ++b)
            in.pen[a * A + b] = matrix ? matrix->penalty(uint8_t(a), uint8_t(b)) : int32_t(a != b);
    in.q = pack_gapblock(queries, codec);
    in.r = pack_gapblock(refs, codec);
    const uint32_t longest = std::max(in.q.max_len, in.r.max_len);
    const size_t W1 = size_t(prior_width) + 1;
    if (!prior.empty()) {
        if (prior.size() != W1 * W1 * W1)
            throw std::invalid_argument("prior table must have (prior_width + 1)^3 entries");
        if (longest > prior_width)
            throw std::invalid_argument("a sequence is longer than the prior table's width");
    }
    return in;
}

}  // namespace

GapblockHits gapblock_threshold(const StringSet& queries, const StringSet& refs, Alphabet alphabet,
                                const SubstitutionMatrix* matrix, int32_t gap_open,
                                int32_t gap_extend, const std::vector<int32_t>& prior,
                                uint32_t prior_width, int32_t cutoff) {
    const GapblockInput in = gapblock_input(queries, refs, alphabet, matrix, gap_open, gap_extend,
                                            prior, prior_width);
    GapblockHits hits;
    if (in.empty) return hits;
    const size_t N = queries.size(), K = refs.size();
    uint64_t capacity = std::min<uint64_t>(uint64_t(N) * K, uint64_t(1) << 22);
    for (;;) {
        hits.query.resize(capacity);
        hits.ref.resize(capacity);
        hits.score.resize(capacity);
        uint64_t count = 0;
        st_status st{};
        raise(st_gapblock_threshold(in.q.bytes.data(), in.q.off.data(), N, in.r.bytes.data(),
                                    in.r.off.data(), K, in.pen.data(), in.A, gap_open,
                                    gap_extend, prior.empty() ? nullptr : prior.data(),
                                    prior_width, cutoff, capacity, hits.query.data(),
                                    hits.ref.data(), hits.score.data(), &count, &st),
              st);
        if (count <= capacity) {
            hits.query.resize(count);
            hits.ref.resize(count);
            hits.score.resize(count);
            break;
        }
        capacity = count;  // the first pass counted every hit; the second stores them all
    }
    // row-major order, independent of the device's append order
    std::vector<uint64_t> order(hits.query.size());
    for (size_t t = 0; t < order.size(); ++t) order[t] = t;
    std::sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
        return hits.query[x] != hits.query[y] ? hits.query[x] < hits.query[y]
                                              : hits.ref[x] < hits.ref[y];
    });
    GapblockHits sorted;
    sorted.query.reserve(order.size());
    sorted.ref.reserve(order.size());
    sorted.score.reserve(order.size());
    for (uint64_t t : order) {
        sorted.query.push_back(hits.query[t]);
        sorted.ref.push_back(hits.ref[t]);
        sorted.score.push_back(hits.score[t]);
    }
    return sorted;
}

std::vector<int32_t> gapblock_pairs(const StringSet& queries, const StringSet& refs,
                                    Alphabet alphabet,
                                    const SubstitutionMatrix* matrix, int32_t gap_open,
                                    int32_t gap_extend, const std::vector<int32_t>& prior,
                                    uint32_t prior_width, const std::vector<uint32_t>& pair_query,
                                    const std::vector<uint32_t>& pair_ref) {
    if (pair_query.size() != pair_ref.size())
        throw std::invalid_argument("pair_query and pair_ref must have the same length");
    const GapblockInput in = gapblock_input(queries, refs, alphabet, matrix, gap_open, gap_extend,
                                            prior, prior_width);
    std::vector<int32_t> out(pair_query.size());
    if (out.empty()) return out;
    if (in.empty) throw std::invalid_argument("pairs index an empty query or reference set");
    st_status st{};
    raise(st_gapblock_pairs(in.q.bytes.data(), in.q.off.data(), queries.size(), in.r.bytes.data(),
                            in.r.off.data(), refs.size(), in.pen.data(), in.A, gap_open,
                            gap_extend, prior.empty() ? nullptr : prior.data(), prior_width,
                            pair_query.data(), pair_ref.data(), out.size(), out.data(), &st),
          st);
    return out;
}

std::vector<uint64_t> gapblock_histogram(const StringSet& queries, const StringSet& refs,
                                         Alphabet alphabet,
                                         const SubstitutionMatrix* matrix, int32_t gap_open,
                                         int32_t gap_extend, const std::vector<int32_t>& prior,
                                         uint32_t prior_width, uint32_t bins) {
    if (bins < 1 || bins > 16382) throw std::invalid_argument("bins must be in [1, 16382]");
    const GapblockInput in = gapblock_input(queries, refs, alphabet, matrix, gap_open, gap_extend,
                                            prior, prior_width);
    std::vector<uint64_t> out(queries.size() * (size_t(bins) + 2), 0);
    if (in.empty) return out;  // no queries or no references
    st_status st{};
    raise(st_gapblock_histogram(in.q.bytes.data(), in.q.off.data(), queries.size(),
                                in.r.bytes.data(), in.r.off.data(), refs.size(), in.pen.data(),
                                in.A, gap_open, gap_extend,
                                prior.empty() ? nullptr : prior.data(), prior_width, bins,
                                out.data(), &st),
          st);
    return out;
}

std::vector<int32_t> gapblock_matrix(const std::vector<std::string>& queries,
                                     const std::vector<std::string>& refs, Alphabet alphabet,
                                     const SubstitutionMatrix* matrix, int32_t gap_open,
                                     int32_t gap_extend, const std::vector<int32_t>& prior,
                                     uint32_t prior_width, int /*threads*/) {
    std::vector<int32_t> out(queries.size() * refs.size());
    gapblock_matrix_into(queries, refs, alphabet, matrix, gap_open, gap_extend, prior,
                         prior_width, out.data());
    return out;
}

void gapblock_matrix_into(const StringSet& queries, const StringSet& refs, Alphabet alphabet,
                          const SubstitutionMatrix* matrix, int32_t gap_open, int32_t gap_extend,
                          const std::vector<int32_t>& prior, uint32_t prior_width,
                          int32_t* out) {
    const GapblockInput in = gapblock_input(queries, refs, alphabet, matrix, gap_open, gap_extend,
                                            prior, prior_width);
    const size_t N = queries.size(), K = refs.size();
    if (in.empty) return;
    const Csr& q = in.q;
    const Csr& r = in.r;
    const std::vector<int32_t>& pen = in.pen;
    const uint8_t A = in.A;

    st_status st{};
    const int rc = st_gapblock_matrix(q.bytes.data(), q.off.data(), N, r.bytes.data(),
                                      r.off.data(), K, pen.data(), A, gap_open, gap_extend,
                                      prior.empty() ? nullptr : prior.data(), prior_width,
                                      out, &st);
    raise(rc, st);
}

int hamming(std::string_view a, std::string_view b) {
    if (a.size() != b.size())
        throw std::invalid_argument("seqtree: hamming distance is only defined for equal-length "
                                    "sequences (got " + std::to_string(a.size()) + " and " +
                                    std::to_string(b.size()) + ")");
    int32_t out = 0;
    st_status st{};
    raise(st_hamming_matrix(raw(a), offsets(a).data(), 1, raw(b), offsets(b).data(), 1, &out, &st),
          st);
    return out;
}

int levenshtein(std::string_view a, std::string_view b) {
    int32_t out = 0;
    st_status st{};
    raise(st_levenshtein_matrix(raw(a), offsets(a).data(), 1, raw(b), offsets(b).data(), 1, &out,
                                &st),
          st);
    return out;
}

void hamming_matrix_into(const StringSet& a, const StringSet& b, int32_t* out) {
    const size_t N = a.size(), K = b.size();
    if (N == 0 || K == 0) return;
    st_status st{};
    raise(st_hamming_matrix(raw(a), a.offsets(), N, raw(b), b.offsets(), K, out, &st), st);
}

std::vector<int32_t> hamming_matrix(const std::vector<std::string>& a,
                                   const std::vector<std::string>& b, int /*threads*/) {
    std::vector<int32_t> out(a.size() * b.size());
    hamming_matrix_into(a, b, out.data());
    return out;
}

void levenshtein_matrix_into(const StringSet& a, const StringSet& b, int32_t* out) {
    const size_t N = a.size(), K = b.size();
    if (N == 0 || K == 0) return;
    st_status st{};
    raise(st_levenshtein_matrix(raw(a), a.offsets(), N, raw(b), b.offsets(), K, out, &st), st);
}

std::vector<int32_t> levenshtein_matrix(const std::vector<std::string>& a,
                                   const std::vector<std::string>& b, int /*threads*/) {
    std::vector<int32_t> out(a.size() * b.size());
    levenshtein_matrix_into(a, b, out.data());
    return out;
}

std::unique_ptr<Index> Index::build(std::vector<std::string> refs, Alphabet a) {
    std::unique_ptr<Index> idx(new Index(a));
    const Codec& codec = idx->codec_;
    size_t total = 0;
    for (const auto& s : refs) total += s.size();
    idx->data_.reserve(total);
    idx->codes_.resize(total);
    idx->off_.assign(1, 0);
    idx->off_.reserve(refs.size() + 1);
    for (uint32_t id = 0; id < refs.size(); ++id) {
        const std::string& s = refs[id];
        for (size_t t = 0; t < s.size(); ++t) {
            const uint8_t code = codec.encode(s[t]);
            if (code == Codec::kInvalid)
                throw std::invalid_argument(std::string("invalid symbol '") + s[t] +
                                            "' in reference #" + std::to_string(id));
            idx->codes_[idx->data_.size() + t] = code;
        }
        idx->data_ += s;
        idx->off_.push_back(idx->data_.size());
    }
    return idx;
}

std::string_view Index::ref_seq(uint32_t ref_id) const {
    if (ref_id >= size()) throw std::out_of_range("seqtree: ref_id out of range");
    return std::string_view(data_.data() + off_[ref_id], off_[ref_id + 1] - off_[ref_id]);
}

Alignment Index::align(std::string_view query, uint32_t ref_id, const SearchParams& p) const {
    return align_batch(one(query), {ref_id}, p)[0];
}

std::vector<Alignment> Index::align_batch(const StringSet& queries,
                                          const std::vector<uint32_t>& ref_ids,
                                          const SearchParams& p) const {
    if (queries.size() != ref_ids.size())
        throw std::invalid_argument("seqtree: queries and ref_ids differ in length");
    // the reference's order: query symbols, then gap costs (src/index.cpp:109-113)
    const Csr q = pack(queries, codec_, [](char ch) {
        throw std::invalid_argument(std::string("invalid symbol '") + ch + "' in query");
    });
    if (p.gap_open < 0 || p.gap_extend < 0)
        throw std::invalid_argument("seqtree: gap_open and gap_extend must be >= 0");
    if (p.matrix && p.matrix->size() != codec_.size())
        throw std::invalid_argument("matrix size does not match the alphabet");
    for (uint32_t id : ref_ids)
        if (id >= size()) throw std::out_of_range("seqtree: ref_id out of range");
    const size_t P = queries.size();
    std::vector<Alignment> out(P);
    if (P == 0) return out;
    // only the references the pairs name travel to the device, each once
    std::unordered_map<uint32_t, uint32_t> slot;
    std::vector<uint32_t> pair_ref(P), pair_query(P);
    Csr r;
    r.off.push_back(0);
    for (size_t t = 0; t < P; ++t) {
        pair_query[t] = uint32_t(t);
        const uint32_t id = ref_ids[t];
        auto it = slot.find(id);
        if (it == slot.end()) {
            it = slot.emplace(id, uint32_t(r.off.size() - 1)).first;
            r.bytes.insert(r.bytes.end(), codes_.begin() + off_[id], codes_.begin() + off_[id + 1]);
            r.off.push_back(r.bytes.size());
            r.max_len = std::max<uint32_t>(r.max_len, uint32_t(off_[id + 1] - off_[id]));
        }
        pair_ref[t] = it->second;
    }
    if (r.bytes.empty()) r.bytes.push_back(0);  // a valid pointer for all-empty references
    std::vector<uint8_t> qbytes = q.bytes;
    if (qbytes.empty()) qbytes.push_back(0);
    const uint64_t stride = uint64_t(q.max_len) + r.max_len + 1;
    std::vector<uint8_t> ops(P * stride);
    std::vector<int32_t> score(P);
    std::vector<uint32_t> n_ops(P);
    st_status st{};
    raise(st_index_align_pairs(qbytes.data(), q.off.data(), P, r.bytes.data(), r.off.data(),
                               r.off.size() - 1, pair_query.data(), pair_ref.data(), P,
                               p.matrix ? p.matrix->penalties().data() : nullptr, codec_.size(),
                               p.gap_open, p.gap_extend, ops.data(), stride, score.data(),
                               n_ops.data(), &st),
          st);
    for (size_t t = 0; t < P; ++t) {
        const std::string_view query = queries[t];
        const std::string_view ref = ref_seq(ref_ids[t]);
        Alignment& al = out[t];
        al.score = score[t];
        al.ops.assign(reinterpret_cast<const char*>(ops.data() + t * stride), n_ops[t]);
        al.aligned_query.reserve(n_ops[t]);
        al.aligned_ref.reserve(n_ops[t]);
        size_t i = 0, j = 0;
        for (const char o : al.ops) {
            const bool takes_q = o != 'I', takes_r = o != 'D';
            al.aligned_query += takes_q ? query[i++] : '-';
            al.aligned_ref += takes_r ? ref[j++] : '-';
        }
    }
    return out;
}

}  // namespace seqtree
