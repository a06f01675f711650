// C++ API of the MI355X scoring path: the same declarations, argument meaning, validation
// order and exceptions as the reference's include/seqtree/seqtree.hpp (Codec :15-31,
// SubstitutionMatrix :43-75, AlignMode :186, the scalar and matrix functions :188-242) and
// include/seqtree/types.hpp:9 (Alphabet), restricted to the all-vs-all scoring path, plus
// Index::align (seqtree.hpp:108-120, types.hpp:31-54 SearchParams) without the trie search.
//
// The dense-matrix functions run on the GPU through the C-ABI in seqtree_hip.h. Their
// `threads` argument is accepted for signature compatibility and ignored: the host does only
// validation and encoding, and the device does the scoring.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <string_view>
#include <vector>

namespace seqtree {

enum class Alphabet : uint8_t { AminoAcid, Nucleotide, NucleotideIUPAC };

// Maps characters to small integer codes (< 32) and back, case-insensitively on input.
// Characters outside the alphabet encode to kInvalid.
class Codec {
public:
    static constexpr uint8_t kInvalid = 0xFF;
    explicit Codec(Alphabet);
    uint8_t size() const { return size_; }
    Alphabet alphabet() const { return alphabet_; }
    uint8_t encode(char c) const { return enc_[static_cast<uint8_t>(c)]; }
    char decode(uint8_t code) const { return dec_[code]; }

private:
    Alphabet alphabet_;
    uint8_t size_ = 0;
    uint8_t enc_[256];
    char dec_[32];
};

// A substitution matrix in two views: the signed similarity (maximised by Gotoh) and the
// non-negative Gram penalty pen(a,b) = max(0, s(a,a) + s(b,b) - 2 s(a,b)) (minimised by the
// gap-block score). unit(A) is the identity: penalty 0/1, similarity 1/0.
class SubstitutionMatrix {
public:
    static SubstitutionMatrix unit(uint8_t size);
    static SubstitutionMatrix blosum62();
    static SubstitutionMatrix blosum45();
    static SubstitutionMatrix blosum80();
    static SubstitutionMatrix pam250();
    static SubstitutionMatrix pam100();
    static SubstitutionMatrix structural();
    static SubstitutionMatrix from_similarity(uint8_t size, const int32_t* sim);

    uint8_t size() const { return size_; }
    int32_t penalty(uint8_t a, uint8_t b) const { return pen_[a * size_ + b]; }
    int32_t similarity(uint8_t a, uint8_t b) const { return sim_[a * size_ + b]; }
    // Median penalty over the mismatched symbol pairs.
    int32_t scale() const;

    const std::vector<int32_t>& penalties() const { return pen_; }
    const std::vector<int32_t>& similarities() const { return sim_; }

private:
    uint8_t size_ = 0;
    std::vector<int32_t> pen_;
    std::vector<int32_t> sim_;
};

// A set of strings stored back to back: characters plus n + 1 offsets (no counterpart in the
// reference, whose functions take std::vector<std::string>). The matrix functions encode from
// this form in one pass; the Python binding fills it straight from a list of str, and a
// std::vector<std::string> converts implicitly.
class StringSet {
public:
    StringSet() = default;
    StringSet(const std::vector<std::string>& v) {  // NOLINT: implicit on purpose
        size_t total = 0;
        for (const auto& s : v) total += s.size();
        reserve(v.size(), total);
        for (const auto& s : v) push_back(s);
    }
    void reserve(size_t n, size_t chars) {
        chars_.reserve(chars);
        off_.reserve(n + 1);
    }
    void push_back(std::string_view s) {
        chars_.append(s.data(), s.size());
        off_.push_back(chars_.size());
        if (s.size() > max_len_) max_len_ = uint32_t(s.size());
    }
    size_t size() const { return off_.size() - 1; }
    bool empty() const { return off_.size() == 1; }
    std::string_view operator[](size_t i) const {
        return std::string_view(chars_.data() + off_[i], off_[i + 1] - off_[i]);
    }
    const char* chars() const { return chars_.data(); }
    const uint64_t* offsets() const { return off_.data(); }
    size_t total() const { return chars_.size(); }
    uint32_t max_len() const { return max_len_; }

private:
    std::string chars_;
    std::vector<uint64_t> off_{0};
    uint32_t max_len_ = 0;
};

enum class AlignMode { Global, Local };

// Optimal Gotoh score of one pair (int64 DP semantics, narrowed to int32 or
// std::overflow_error).
int32_t align_score(std::string_view query, std::string_view ref, const SubstitutionMatrix&,
                    Alphabet, AlignMode, int32_t gap_open, int32_t gap_extend);

// Optimal score plus the aligned strings and ops (reference types.hpp:67-72).
struct Alignment {
    std::string aligned_query;
    std::string aligned_ref;
    std::string ops;  // per column: 'M' match, 'S' substitution, 'I' insertion, 'D' deletion
    int32_t score = 0;
};

// Gotoh alignment with traceback of one pair (reference pairwise.cpp:141-267).
Alignment align_pair(std::string_view query, std::string_view ref, const SubstitutionMatrix&,
                     Alphabet, AlignMode, int32_t gap_open, int32_t gap_extend);

// align_pair for every (queries[pair_query[p]], refs[pair_ref[p]]), in one device batch (no
// reference counterpart: the traceback step for candidate pairs).
std::vector<Alignment> align_pairs(const StringSet& queries,
                                   const StringSet& refs,
                                   const std::vector<uint32_t>& pair_query,
                                   const std::vector<uint32_t>& pair_ref,
                                   const SubstitutionMatrix&, Alphabet, AlignMode,
                                   int32_t gap_open, int32_t gap_extend);

// Dense N*K similarity matrix, row-major.
std::vector<int32_t> align_score_matrix(const std::vector<std::string>& queries,
                                        const std::vector<std::string>& refs,
                                        const SubstitutionMatrix&, Alphabet, AlignMode,
                                        int32_t gap_open, int32_t gap_extend, int threads = 0);

// Dense N*K distance matrix d(a,b) = s(a,a) + s(b,b) - 2*s(a,b).
std::vector<int32_t> align_dist_matrix(const std::vector<std::string>& queries,
                                       const std::vector<std::string>& refs,
                                       const SubstitutionMatrix&, Alphabet, AlignMode,
                                       int32_t gap_open, int32_t gap_extend, int threads = 0);

// Dense N*K single-gap-block penalty matrix, row-major. `prior` is the gap prior flattened to
// [m][d][i] with stride prior_width + 1, or empty; `matrix` may be null for unit cost.
std::vector<int32_t> gapblock_matrix(const std::vector<std::string>& queries,
                                     const std::vector<std::string>& refs, Alphabet,
                                     const SubstitutionMatrix* matrix, int32_t gap_open,
                                     int32_t gap_extend, const std::vector<int32_t>& prior,
                                     uint32_t prior_width, int threads = 0);

// Every (query, ref) pair whose gap-block score is <= cutoff, sorted by (query, ref). The
// exhaustive counterpart of the reference's GapBlockIndex.search without a d_max bound.
struct GapblockHits {
    std::vector<uint32_t> query;
    std::vector<uint32_t> ref;
    std::vector<int32_t> score;
};
GapblockHits gapblock_threshold(const StringSet& queries,
                                const StringSet& refs, Alphabet,
                                const SubstitutionMatrix* matrix, int32_t gap_open,
                                int32_t gap_extend, const std::vector<int32_t>& prior,
                                uint32_t prior_width, int32_t cutoff);

// Per-query histograms of the exact gap-block scores against every reference, row-major
// [queries.size()][bins + 2]: bin 0 counts scores below zero, bin 1 + s the score s in
// [0, bins), bin bins + 1 scores >= bins (the counts behind control-calibrated E-values).
std::vector<uint64_t> gapblock_histogram(const StringSet& queries,
                                         const StringSet& refs, Alphabet,
                                         const SubstitutionMatrix* matrix, int32_t gap_open,
                                         int32_t gap_extend, const std::vector<int32_t>& prior,
                                         uint32_t prior_width, uint32_t bins);

// Gap-block scores of explicit (query index, ref index) pairs.
std::vector<int32_t> gapblock_pairs(const StringSet& queries,
                                    const StringSet& refs, Alphabet,
                                    const SubstitutionMatrix* matrix, int32_t gap_open,
                                    int32_t gap_extend, const std::vector<int32_t>& prior,
                                    uint32_t prior_width, const std::vector<uint32_t>& pair_query,
                                    const std::vector<uint32_t>& pair_ref);

// Unit-cost distances on raw, case-sensitive bytes.
int hamming(std::string_view a, std::string_view b);
int levenshtein(std::string_view a, std::string_view b);

std::vector<int32_t> hamming_matrix(const std::vector<std::string>& a,
                                    const std::vector<std::string>& b, int threads = 0);
std::vector<int32_t> levenshtein_matrix(const std::vector<std::string>& a,
                                        const std::vector<std::string>& b, int threads = 0);

// The dense-matrix functions above, writing the row-major matrix into caller memory of
// queries.size() * refs.size() int32 instead of returning a vector (no counterpart in the
// reference): the same checks, order and exceptions. The pybind11 ScoreMatrix allocates its
// buffer and calls these, which spares zero-filling a vector that the call overwrites.
void gapblock_matrix_into(const StringSet& queries,
                          const StringSet& refs, Alphabet,
                          const SubstitutionMatrix* matrix, int32_t gap_open, int32_t gap_extend,
                          const std::vector<int32_t>& prior, uint32_t prior_width, int32_t* out);
void align_score_matrix_into(const StringSet& queries,
                             const StringSet& refs, const SubstitutionMatrix&,
                             Alphabet, AlignMode, int32_t gap_open, int32_t gap_extend,
                             int32_t* out);
void align_dist_matrix_into(const StringSet& queries,
                            const StringSet& refs, const SubstitutionMatrix&,
                            Alphabet, AlignMode, int32_t gap_open, int32_t gap_extend,
                            int32_t* out);
void hamming_matrix_into(const StringSet& a, const StringSet& b,
                         int32_t* out);
void levenshtein_matrix_into(const StringSet& a,
                             const StringSet& b, int32_t* out);

// Search scope and budget (reference include/seqtree/types.hpp:23-54), kept field for field
// so code that fills one compiles unchanged. Index::align reads the scoring fields (matrix,
// gap_open, gap_extend); the search scope and budget fields belong to the reference's trie
// search, which this library does not provide.
enum class Engine : uint8_t { Auto, SeqTrie, SeqTm };
enum class Mode : uint8_t { AllHits, TopHit };
struct SearchParams {
    Engine engine = Engine::Auto;
    Mode mode = Mode::AllHits;
    uint16_t max_substitutions = 0;
    uint16_t max_insertions = 0;
    uint16_t max_deletions = 0;
    uint16_t max_total_edits = 0;
    int32_t max_score_penalty = 0;
    const SubstitutionMatrix* matrix = nullptr;  // null => unit cost
    int32_t gap_open = 1;
    int32_t gap_extend = 1;
    uint32_t max_hits = 0;
};

// A reference set addressed by position (reference seqtree.hpp:108-142): build() validates and
// stores the references, align() is the reference's on-demand global alignment of a query to
// one of them (src/index.cpp:105-179), on the GPU. align_batch() traces back many
// (query, reference id) pairs in one device batch (no reference counterpart). The trie, its
// search and its serialisation are not part of this library.
class Index {
public:
    // Throws std::invalid_argument("invalid symbol 'c' in reference #id") (src/trie.cpp:60-66).
    static std::unique_ptr<Index> build(std::vector<std::string> refs, Alphabet);

    uint32_t size() const { return uint32_t(off_.size() - 1); }
    Alphabet alphabet() const { return codec_.alphabet(); }
    const Codec& codec() const { return codec_; }
    std::string_view ref_seq(uint32_t ref_id) const;

    // Minimum-penalty global alignment with affine gaps: substitution penalties from
    // p.matrix (unit cost when null), gap runs costing gap_open + (L - 1) * gap_extend.
    // Alignment::score is that penalty. Throws std::invalid_argument for a query symbol
    // outside the alphabet ("invalid symbol 'c' in query"), then for negative gap costs, and
    // std::out_of_range for a ref_id past the set.
    Alignment align(std::string_view query, uint32_t ref_id, const SearchParams& p) const;
    std::vector<Alignment> align_batch(const StringSet& queries,
                                       const std::vector<uint32_t>& ref_ids,
                                       const SearchParams& p) const;

private:
    explicit Index(Alphabet a) : codec_(a) {}
    Codec codec_;
    std::string data_;             // the references as given, back to back
    std::vector<uint64_t> off_;    // size() + 1 offsets into data_ (and codes_)
    std::vector<uint8_t> codes_;   // their codec codes
};

}  // namespace seqtree
