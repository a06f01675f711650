// Test infrastructure only -- never part of the shipped library.
//
// A C-ABI shim over the reference's own C++ API (include/seqtree/seqtree.hpp:186-242 in the
// reference tree), compiled together with the reference's hot-path sources by oracle/Makefile
// into oracle/_ref/libseqtree_ref.so. tests/ load it through ctypes to pin the plain-C
// restatement in oracle/seqtree_oracle.c, and bench.py times it as the CPU baseline
// ("kind": "reference"). Every entry point returns 0 on success, 1 for
// std::invalid_argument, 2 for std::overflow_error and 3 for anything else, copying the
// exception text into `err`.

#include "seqtree/seqtree.hpp"

#include <atomic>
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace seqtree;

namespace {

std::vector<std::string> unpack(const char* bytes, const uint64_t* off, uint64_t n) {
    std::vector<std::string> v(n);
    for (uint64_t i = 0; i < n; ++i) v[i].assign(bytes + off[i], bytes + off[i + 1]);
    return v;
}

// kind: -1 none (null matrix), 0 blosum62, 1 blosum45, 2 blosum80, 3 pam250, 4 pam100,
// 5 structural, 6 unit(custom_size), 7 from_similarity(custom_size, custom_sim).
bool make_matrix(int kind, const int32_t* custom_sim, int32_t custom_size, SubstitutionMatrix& m) {
    switch (kind) {
        case 0: m = SubstitutionMatrix::blosum62(); return true;
        case 1: m = SubstitutionMatrix::blosum45(); return true;
        case 2: m = SubstitutionMatrix::blosum80(); return true;
        case 3: m = SubstitutionMatrix::pam250(); return true;
        case 4: m = SubstitutionMatrix::pam100(); return true;
        case 5: m = SubstitutionMatrix::structural(); return true;
        case 6: m = SubstitutionMatrix::unit(uint8_t(custom_size)); return true;
        case 7: m = SubstitutionMatrix::from_similarity(uint8_t(custom_size), custom_sim); return true;
        default: return false;
    }
}

template <class F>
int guarded(char* err, int errlen, F&& f) {
    auto put = [&](const char* s) {
        if (err && errlen > 0) {
            std::strncpy(err, s, size_t(errlen) - 1);
            err[errlen - 1] = 0;
        }
    };
    try {
        f();
        return 0;
    } catch (const std::invalid_argument& e) {
        put(e.what());
        return 1;
    } catch (const std::overflow_error& e) {
        put(e.what());
        return 2;
    } catch (const std::exception& e) {
        put(e.what());
        return 3;
    }
}

}  // namespace

extern "C" {

int ref_matrix_tables(int kind, int32_t custom_size, const int32_t* custom_sim, int32_t* sim,
                      int32_t* pen, int32_t* size, int32_t* scale) {
    SubstitutionMatrix m;
    if (!make_matrix(kind, custom_sim, custom_size, m)) return 3;
    const int A = m.size();
    *size = A;
    *scale = m.scale();
    for (int a = 0; a < A; ++a)
        for (int b = 0; b < A; ++b) {
            sim[a * A + b] = m.similarity(uint8_t(a), uint8_t(b));
            pen[a * A + b] = m.penalty(uint8_t(a), uint8_t(b));
        }
    return 0;
}

int ref_gapblock_matrix(const char* qb, const uint64_t* qo, uint64_t N, const char* rb,
                        const uint64_t* ro, uint64_t K, int alphabet, int kind,
                        const int32_t* custom_sim, int32_t custom_size, int32_t go, int32_t ge,
                        const int32_t* prior, uint64_t prior_len, uint32_t prior_width,
                        int threads, int32_t* out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        auto q = unpack(qb, qo, N);
        auto r = unpack(rb, ro, K);
        std::vector<int32_t> pr(prior, prior + prior_len);
        SubstitutionMatrix m;
        const bool has = make_matrix(kind, custom_sim, custom_size, m);
        auto v = gapblock_matrix(q, r, Alphabet(alphabet), has ? &m : nullptr, go, ge, pr,
                                 prior_width, threads);
        std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    });
}

int ref_align_matrix(const char* qb, const uint64_t* qo, uint64_t N, const char* rb,
                     const uint64_t* ro, uint64_t K, int alphabet, int kind,
                     const int32_t* custom_sim, int32_t custom_size, int local, int dist,
                     int32_t go, int32_t ge, int threads, int32_t* out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        auto q = unpack(qb, qo, N);
        auto r = unpack(rb, ro, K);
        SubstitutionMatrix m;
        if (!make_matrix(kind, custom_sim, custom_size, m))
            throw std::runtime_error("pairwise needs a matrix");
        const AlignMode md = local ? AlignMode::Local : AlignMode::Global;
        auto v = dist ? align_dist_matrix(q, r, m, Alphabet(alphabet), md, go, ge, threads)
                      : align_score_matrix(q, r, m, Alphabet(alphabet), md, go, ge, threads);
        std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    });
}

int ref_align_score(const char* q, uint64_t m, const char* r, uint64_t n, int alphabet, int kind,
                    const int32_t* custom_sim, int32_t custom_size, int local, int32_t go,
                    int32_t ge, int32_t* out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        SubstitutionMatrix mat;
        if (!make_matrix(kind, custom_sim, custom_size, mat))
            throw std::runtime_error("pairwise needs a matrix");
        *out = align_score(std::string_view(q, m), std::string_view(r, n), mat, Alphabet(alphabet),
                           local ? AlignMode::Local : AlignMode::Global, go, ge);
    });
}

// align_pair (src/pairwise.cpp:141-267): the score and the three strings, each written to
// a buffer of at least m + n bytes, their common length in *len.
int ref_align_pair(const char* q, uint64_t m, const char* r, uint64_t n, int alphabet, int kind,
                   const int32_t* custom_sim, int32_t custom_size, int local, int32_t go,
                   int32_t ge, int32_t* score, char* aq, char* ar, char* ops, uint64_t* len,
                   char* err, int errlen) {
    return guarded(err, errlen, [&] {
        SubstitutionMatrix mat;
        if (!make_matrix(kind, custom_sim, custom_size, mat))
            throw std::runtime_error("pairwise needs a matrix");
        const Alignment a =
            align_pair(std::string_view(q, m), std::string_view(r, n), mat, Alphabet(alphabet),
                       local ? AlignMode::Local : AlignMode::Global, go, ge);
        *score = a.score;
        *len = a.ops.size();
        std::memcpy(aq, a.aligned_query.data(), a.aligned_query.size());
        std::memcpy(ar, a.aligned_ref.data(), a.aligned_ref.size());
        std::memcpy(ops, a.ops.data(), a.ops.size());
    });
}

// align_pair over a list of (query, reference) index pairs on `threads` workers (pairs taken
// from an atomic counter): the CPU baseline of batched traceback. Returns the total op count.
int ref_align_pairs(const char* qb, const uint64_t* qo, uint64_t N, const char* rb,
                    const uint64_t* ro, uint64_t K, const uint32_t* pq, const uint32_t* pr,
                    uint64_t n_pairs, int alphabet, int kind, int local, int32_t go, int32_t ge,
                    int threads, uint64_t* total_ops, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        SubstitutionMatrix mat;
        if (!make_matrix(kind, nullptr, 0, mat))
            throw std::runtime_error("pairwise needs a matrix");
        (void)N;
        (void)K;
        std::atomic<uint64_t> next{0}, ops{0};
        auto work = [&] {
            uint64_t mine = 0;
            for (uint64_t t; (t = next.fetch_add(1)) < n_pairs;) {
                const uint32_t a = pq[t], b = pr[t];
                const Alignment al = align_pair(
                    std::string_view(qb + qo[a], qo[a + 1] - qo[a]),
                    std::string_view(rb + ro[b], ro[b + 1] - ro[b]), mat, Alphabet(alphabet),
                    local ? AlignMode::Local : AlignMode::Global, go, ge);
                mine += al.ops.size();
            }
            ops += mine;
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < threads; ++w) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        *total_ops = ops.load();
    });
}

// Index::align (src/index.cpp:105-179) for every (queries[pq[t]], refs[pr[t]]) against one
// Index built over refs, on `threads` workers taking pairs from an atomic counter. kind -1 is
// unit cost (a null matrix). Per pair: the penalty score, and the ops and aligned strings
// (at most m + n bytes each, at t * stride) with their common length.
int ref_index_align_pairs(const char* qb, const uint64_t* qo, uint64_t N, const char* rb,
                          const uint64_t* ro, uint64_t K, const uint32_t* pq, const uint32_t* pr,
                          uint64_t n_pairs, int alphabet, int kind, const int32_t* custom_sim,
                          int32_t custom_size, int32_t go, int32_t ge, int threads,
                          int32_t* score, char* ops, char* aq, char* ar, uint64_t stride,
                          uint64_t* len, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        (void)N;
        auto idx = Index::build(unpack(rb, ro, K), Alphabet(alphabet));
        SubstitutionMatrix mat;
        const bool has = make_matrix(kind, custom_sim, custom_size, mat);
        SearchParams p;
        p.matrix = has ? &mat : nullptr;
        p.gap_open = go;
        p.gap_extend = ge;
        std::atomic<uint64_t> next{0};
        std::exception_ptr first;
        std::atomic<bool> failed{false};
        auto work = [&] {
            try {
                for (uint64_t t; !failed && (t = next.fetch_add(1)) < n_pairs;) {
                    const uint32_t a = pq[t];
                    const Alignment al =
                        idx->align(std::string_view(qb + qo[a], qo[a + 1] - qo[a]), pr[t], p);
                    score[t] = al.score;
                    len[t] = al.ops.size();
                    if (ops) {
                        std::memcpy(ops + t * stride, al.ops.data(), al.ops.size());
                        std::memcpy(aq + t * stride, al.aligned_query.data(), al.ops.size());
                        std::memcpy(ar + t * stride, al.aligned_ref.data(), al.ops.size());
                    }
                }
            } catch (...) {
                if (!failed.exchange(true)) first = std::current_exception();
            }
        };
        std::vector<std::thread> pool;
        for (int w = 1; w < threads; ++w) pool.emplace_back(work);
        work();
        for (auto& th : pool) th.join();
        if (first) std::rethrow_exception(first);
    });
}

int ref_distance_matrix(int which, const char* ab, const uint64_t* ao, uint64_t N, const char* bb,
                        const uint64_t* bo, uint64_t K, int threads, int32_t* out, char* err,
                        int errlen) {
    return guarded(err, errlen, [&] {
        auto a = unpack(ab, ao, N);
        auto b = unpack(bb, bo, K);
        auto v = which == 0 ? hamming_matrix(a, b, threads) : levenshtein_matrix(a, b, threads);
        std::memcpy(out, v.data(), v.size() * sizeof(int32_t));
    });
}

int ref_distance(int which, const char* a, uint64_t m, const char* b, uint64_t n, int32_t* out,
                 char* err, int errlen) {
    return guarded(err, errlen, [&] {
        std::string_view x(a, m), y(b, n);
        *out = which == 0 ? hamming(x, y) : levenshtein(x, y);
    });
}

}  // extern "C"
