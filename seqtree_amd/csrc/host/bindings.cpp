// pybind11 module `_core`: the reference's hot-path bindings (src/_bindings.cpp:34-40
// parse_alphabet, 203-233 ScoreMatrix and gapblock_matrix, 240-315 SubstitutionMatrix,
// 346-350 alphabet helpers, 470-615 ScoreMatrix buffer protocol and the matrix functions;
// 42-55, 64-121, 197-201, 352-397, 420-455 SearchParams and Index.align, without the trie
// search)
// with the same names, argument names, defaults and error types, over the GPU C++ API.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <sys/mman.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <optional>
#include <string>
#include <vector>

#include "seqtree.hpp"
#include "seqtree_hip.h"

namespace py = pybind11;
using namespace seqtree;

// The gap prior cube as the matrix functions take it: a list of int as in the reference, or a
// C-contiguous numpy int32 array (copied in one block; the list form converts per element).
struct PriorCube {
    std::vector<int32_t> v;
};

namespace pybind11::detail {

// A list or tuple of str whose items are all compact ASCII is read straight from the str
// objects into one StringSet (one copy of the characters, no std::string per item). Anything
// else -- other sequences, bytes items, non-ASCII text -- takes pybind11's own conversion to
// std::vector<std::string> first, so every input the reference's binding accepts (and every
// TypeError it raises) is unchanged.
template <>
struct type_caster<StringSet> {
    PYBIND11_TYPE_CASTER(StringSet, const_name("list[str]"));

    bool load(handle src, bool convert) {
        PyObject* o = src.ptr();
        if (!o || PyUnicode_Check(o) || PyBytes_Check(o) || !PySequence_Check(o)) return false;
        if (PyList_Check(o) || PyTuple_Check(o)) {
            const Py_ssize_t n = PySequence_Fast_GET_SIZE(o);
            PyObject** items = PySequence_Fast_ITEMS(o);
            size_t total = 0;
            bool ascii = true;
            for (Py_ssize_t i = 0; i < n && ascii; ++i) {
                PyObject* it = items[i];
                ascii = PyUnicode_Check(it) && PyUnicode_IS_COMPACT_ASCII(it);
                if (ascii) total += size_t(PyUnicode_GET_LENGTH(it));
            }
            if (ascii) {
                value = StringSet();
                value.reserve(size_t(n), total);
                for (Py_ssize_t i = 0; i < n; ++i)
                    value.push_back(std::string_view(
                        static_cast<const char*>(PyUnicode_DATA(items[i])),
                        size_t(PyUnicode_GET_LENGTH(items[i]))));
                return true;
            }
        }
        make_caster<std::vector<std::string>> generic;
        if (!generic.load(src, convert)) return false;
        value = StringSet(cast_op<std::vector<std::string>&>(generic));
        return true;
    }
};

template <>
struct type_caster<PriorCube> {
    PYBIND11_TYPE_CASTER(PriorCube, const_name("list[int]"));

    bool load(handle src, bool convert) {
        if (py::isinstance<py::array>(src)) {
            auto a = py::reinterpret_borrow<py::array>(src);
            if (a.dtype().is(py::dtype::of<int32_t>()) && a.ndim() == 1 &&
                (a.flags() & py::array::c_style)) {
                const auto* p = static_cast<const int32_t*>(a.data());
                value.v.assign(p, p + a.size());
                return true;
            }
        }
        make_caster<std::vector<int32_t>> generic;
        if (!generic.load(src, convert)) return false;
        value.v = std::move(cast_op<std::vector<int32_t>&>(generic));
        return true;
    }
    static handle cast(const PriorCube& c, return_value_policy p, handle parent) {
        return make_caster<std::vector<int32_t>>::cast(c.v, p, parent);
    }
};

}  // namespace pybind11::detail

namespace {

std::string lower(std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::tolower(c); });
    return s;
}

Alphabet parse_alphabet(const std::string& a) {
    const std::string l = lower(a);
    if (l == "aa" || l == "amino" || l == "protein") return Alphabet::AminoAcid;
    if (l == "nt" || l == "dna" || l == "nucleotide") return Alphabet::Nucleotide;
    if (l == "nt_iupac" || l == "iupac") return Alphabet::NucleotideIUPAC;
    throw py::value_error("unknown alphabet '" + a + "' (use 'aa', 'nt', or 'iupac')");
}

AlignMode parse_align_mode(const std::string& m) {
    const std::string l = lower(m);
    if (l == "global" || l == "nw" || l == "needleman-wunsch") return AlignMode::Global;
    if (l == "local" || l == "sw" || l == "smith-waterman") return AlignMode::Local;
    throw py::value_error("unknown mode '" + m + "' (use 'global' or 'local')");
}

std::string alphabet_symbols(Alphabet a) {
    const Codec c(a);
    std::string s;
    for (uint8_t i = 0; i < c.size(); ++i) s += c.decode(i);
    return s;
}

// Uninitialised int32 storage for a ScoreMatrix, never zero-filled (the matrix functions write
// every element). Up to kPinnedMax bytes it is page-locked memory from st_host_alloc, which
// the library fills by DMA straight from the device; past that, or where pinning fails, it is
// 2 MiB-aligned ordinary memory with transparent huge pages requested, filled through the
// library's pinned staging. Buffers of released matrices are kept for reuse (at most
// kCacheBuffers, kCacheBytes in all, the oldest released first): pinning or first-touching
// fresh pages costs the operating system about as much as the copy-out itself. A request
// takes a cached buffer of at least its size and at most twice it.
class HostBuffer {
public:
    HostBuffer() = default;
    explicit HostBuffer(size_t count) {
        const size_t align = size_t(2) << 20;
        bytes_ = (std::max<size_t>(count, 16) * sizeof(int32_t) + align - 1) / align * align;
        if (take(bytes_, &p_, &pinned_)) return;
        if (bytes_ <= kPinnedMax) {
            void* v = nullptr;
            st_status st{};
            if (st_host_alloc(bytes_, &v, &st) == ST_OK) {
                p_ = static_cast<int32_t*>(v);
                pinned_ = true;
                return;
            }
        }
        void* v = nullptr;
        if (posix_memalign(&v, align, bytes_) != 0) throw std::bad_alloc();
        (void)madvise(v, bytes_, MADV_HUGEPAGE);  // advice only; fine where unsupported
        p_ = static_cast<int32_t*>(v);
    }
    HostBuffer(HostBuffer&& o) noexcept : p_(o.p_), bytes_(o.bytes_), pinned_(o.pinned_) {
        o.p_ = nullptr;
    }
    HostBuffer& operator=(HostBuffer&& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(bytes_, o.bytes_);
        std::swap(pinned_, o.pinned_);
        return *this;
    }
    HostBuffer(const HostBuffer&) = delete;
    HostBuffer& operator=(const HostBuffer&) = delete;
    ~HostBuffer() {
        if (p_) give({p_, bytes_, pinned_});
    }
    int32_t* data() const { return p_; }

private:
    static constexpr size_t kPinnedMax = size_t(2) << 30;
    static constexpr size_t kCacheBuffers = 4;
    static constexpr size_t kCacheBytes = size_t(8) << 30;
    struct Cached {
        int32_t* p;
        size_t bytes;
        bool pinned;
    };
    static std::mutex& mu() {
        static std::mutex m;
        return m;
    }
    static std::vector<Cached>& cache() {
        static auto* c = new std::vector<Cached>();  // outlives every ScoreMatrix
        return *c;
    }
    static void release(const Cached& e) {
        if (e.pinned) st_host_free(e.p);
        else std::free(e.p);
    }
    static bool take(size_t bytes, int32_t** p, bool* pinned) {
        std::lock_guard<std::mutex> lk(mu());
        auto& c = cache();
        size_t best = c.size();
        for (size_t i = 0; i < c.size(); ++i)
            if (c[i].bytes >= bytes && c[i].bytes <= 2 * bytes &&
                (best == c.size() || c[i].bytes < c[best].bytes))
                best = i;
        if (best == c.size()) return false;
        *p = c[best].p;
        *pinned = c[best].pinned;
        c.erase(c.begin() + best);
        return true;
    }
    static void give(const Cached& e) {
        std::lock_guard<std::mutex> lk(mu());
        auto& c = cache();
        size_t total = e.bytes;
        for (const Cached& x : c) total += x.bytes;
        while (!c.empty() && (c.size() >= kCacheBuffers || total > kCacheBytes)) {
            total -= c.front().bytes;  // the oldest goes first
            release(c.front());
            c.erase(c.begin());
        }
        if (total <= kCacheBytes) c.push_back(e);
        else release(e);
    }

    int32_t* p_ = nullptr;
    size_t bytes_ = 0;
    bool pinned_ = false;
};

// Dense (rows, cols) int32 block lent to Python through the buffer protocol.
struct ScoreMatrix {
    HostBuffer data;
    size_t rows = 0, cols = 0;
};

// Index of a single amino-acid character in codec order, checked against the matrix size.
uint8_t residue(const SubstitutionMatrix& m, const std::string& a, const char* what) {
    if (a.size() != 1)
        throw py::value_error(std::string(what) + "() takes two single amino-acid characters");
    static const std::string aa = alphabet_symbols(Alphabet::AminoAcid);
    const auto i = aa.find(a[0]);
    if (i == std::string::npos) throw py::value_error("unknown amino acid; expected one of " + aa);
    if (i >= m.size()) throw py::value_error("residue out of range for this matrix's alphabet");
    return uint8_t(i);
}

Engine parse_engine(const std::string& e) {
    const std::string l = lower(e);
    if (l == "auto") return Engine::Auto;
    if (l == "seqtrie") return Engine::SeqTrie;
    if (l == "seqtm") return Engine::SeqTm;
    throw py::value_error("unknown engine '" + e + "' (use 'auto', 'seqtrie', or 'seqtm')");
}

Mode parse_mode(const std::string& m) {
    const std::string l = lower(m);
    if (l == "all") return Mode::AllHits;
    if (l == "top") return Mode::TopHit;
    throw py::value_error("unknown mode '" + m + "' (use 'all' or 'top')");
}

constexpr const char* kMatrixNames =
    "'identity', 'BLOSUM62', 'BLOSUM45', 'BLOSUM80', 'PAM250', 'PAM100', 'structural'";

bool is_matrix_name(const std::string& l) {
    static const char* names[] = {"",       "identity", "blosum62", "blosum45",
                                  "blosum80", "pam250",  "pam100",   "structural"};
    return std::any_of(std::begin(names), std::end(names), [&](const char* n) { return l == n; });
}

// The Python-side search parameters: a matrix given by name or as an object.
struct PyParams {
    int max_subs = 0, max_ins = 0, max_dels = 0, max_total_edits = 0;
    long max_penalty = 0;
    std::string matrix;                             // a built-in name ("" = unit cost)
    std::optional<SubstitutionMatrix> matrix_obj;   // or an explicit matrix, which wins
    int gap_open = 1, gap_extend = 1;
    std::string engine = "auto", mode = "all";
};

void set_matrix(PyParams& p, const py::object& m) {
    p.matrix.clear();
    p.matrix_obj.reset();
    if (m.is_none()) return;
    if (py::isinstance<py::str>(m)) {
        std::string name = m.cast<std::string>();
        if (!is_matrix_name(lower(name)))
            throw py::value_error("unknown matrix '" + name + "' (use '', " + kMatrixNames +
                                  ", or a SubstitutionMatrix)");
        p.matrix = std::move(name);
    } else if (py::isinstance<SubstitutionMatrix>(m)) {
        p.matrix_obj = m.cast<SubstitutionMatrix>();
    } else {
        throw py::type_error("matrix must be a name string or a SubstitutionMatrix");
    }
}

// The matrix a parameter set names for an alphabet; nullopt is unit cost.
std::optional<SubstitutionMatrix> params_matrix(const PyParams& p, Alphabet a) {
    const uint8_t size = Codec(a).size();
    if (p.matrix_obj) {
        if (p.matrix_obj->size() != size)
            throw py::value_error("matrix size does not match the alphabet");
        return p.matrix_obj;
    }
    if (p.matrix.empty()) return std::nullopt;
    const std::string l = lower(p.matrix);
    if (l == "identity") return SubstitutionMatrix::unit(size);
    if (a != Alphabet::AminoAcid) throw py::value_error(l + " requires the amino-acid alphabet");
    if (l == "blosum62") return SubstitutionMatrix::blosum62();
    if (l == "blosum45") return SubstitutionMatrix::blosum45();
    if (l == "blosum80") return SubstitutionMatrix::blosum80();
    if (l == "pam250") return SubstitutionMatrix::pam250();
    if (l == "pam100") return SubstitutionMatrix::pam100();
    return SubstitutionMatrix::structural();
}

SearchParams to_cpp(const PyParams& pp, const SubstitutionMatrix* mat) {
    SearchParams p;
    p.engine = parse_engine(pp.engine);
    p.mode = parse_mode(pp.mode);
    p.max_substitutions = uint16_t(pp.max_subs);
    p.max_insertions = uint16_t(pp.max_ins);
    p.max_deletions = uint16_t(pp.max_dels);
    p.max_total_edits = uint16_t(pp.max_total_edits);
    p.max_score_penalty = int32_t(pp.max_penalty);
    p.matrix = mat;
    p.gap_open = pp.gap_open;
    p.gap_extend = pp.gap_extend;
    return p;
}

// A rows x cols ScoreMatrix filled by compute(out) with the GIL released.
template <class F>
ScoreMatrix make_matrix(size_t rows, size_t cols, F&& compute) {
    ScoreMatrix out;
    out.rows = rows;
    out.cols = cols;
    {
        py::gil_scoped_release release;
        out.data = HostBuffer(rows * cols);
        compute(out.data.data());
    }
    return out;
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() = "seqtree all-vs-all scoring on AMD Instinct MI355X (gfx950 HIP kernels)";

    py::class_<SubstitutionMatrix>(m, "SubstitutionMatrix",
                                   "Substitution matrix with a signed similarity view and a "
                                   "non-negative Gram penalty view (penalty(a,a) == 0).")
        .def_static("blosum62", &SubstitutionMatrix::blosum62)
        .def_static("blosum45", &SubstitutionMatrix::blosum45)
        .def_static("blosum80", &SubstitutionMatrix::blosum80)
        .def_static("pam250", &SubstitutionMatrix::pam250)
        .def_static("pam100", &SubstitutionMatrix::pam100)
        .def_static("structural", &SubstitutionMatrix::structural)
        .def_static("unit", &SubstitutionMatrix::unit, py::arg("size"))
        .def_static(
            "from_similarity",
            [](const std::vector<std::vector<int32_t>>& grid) {
                const size_t n = grid.size();
                if (n == 0 || n > 32) throw py::value_error("matrix size must be 1..32");
                std::vector<int32_t> flat;
                flat.reserve(n * n);
                for (const auto& row : grid) {
                    if (row.size() != n) throw py::value_error("similarity matrix must be square");
                    flat.insert(flat.end(), row.begin(), row.end());
                }
                return SubstitutionMatrix::from_similarity(uint8_t(n), flat.data());
            },
            py::arg("grid"),
            "Build from a square similarity grid in the alphabet's symbol order; penalties are "
            "s[a,a] + s[b,b] - 2*s[a,b] clamped at 0.")
        .def("size", &SubstitutionMatrix::size)
        .def(
            "similarity",
            [](const SubstitutionMatrix& self, const std::string& a, const std::string& b) {
                return self.similarity(residue(self, a, "similarity"),
                                       residue(self, b, "similarity"));
            },
            py::arg("a"), py::arg("b"), "Raw signed similarity of two amino acids.")
        .def("scale", &SubstitutionMatrix::scale,
             "Median penalty over all mismatched symbol pairs (the matrix's natural unit).")
        .def(
            "penalty",
            [](const SubstitutionMatrix& self, const std::string& a, const std::string& b) {
                return self.penalty(residue(self, a, "penalty"), residue(self, b, "penalty"));
            },
            py::arg("a"), py::arg("b"), "Gram-distance penalty of two amino acids.")
        .def("__repr__", [](const SubstitutionMatrix& s) {
            return "SubstitutionMatrix(size=" + std::to_string(s.size()) + ")";
        });

    m.def("alphabet_symbols",
          [](const std::string& a) { return alphabet_symbols(parse_alphabet(a)); },
          py::arg("alphabet") = "aa", "Symbols of an alphabet in codec (matrix) order.");
    m.def("amino_acids", [] { return alphabet_symbols(Alphabet::AminoAcid); },
          "The 24 amino-acid symbols in codec (matrix) order.");

    py::class_<ScoreMatrix>(m, "ScoreMatrix", py::buffer_protocol(),
                            "A read-only (n_queries, n_refs) int32 matrix, row-major. "
                            "numpy.asarray(sm) and memoryview(sm) wrap it without copying.")
        .def_buffer([](ScoreMatrix& s) {
            return py::buffer_info(s.data.data(), sizeof(int32_t),
                                   py::format_descriptor<int32_t>::format(), 2, {s.rows, s.cols},
                                   {sizeof(int32_t) * s.cols, sizeof(int32_t)});
        })
        .def_property_readonly("shape",
                               [](const ScoreMatrix& s) { return py::make_tuple(s.rows, s.cols); })
        .def("__len__", [](const ScoreMatrix& s) { return s.rows; })
        .def(
            "row",
            [](const ScoreMatrix& s, size_t i) {
                if (i >= s.rows) throw py::index_error("row out of range");
                return std::vector<int32_t>(s.data.data() + i * s.cols,
                                            s.data.data() + (i + 1) * s.cols);
            },
            py::arg("i"), "Row i as a list, one value per reference.")
        .def("__getitem__",
             [](const ScoreMatrix& s, std::pair<size_t, size_t> ik) {
                 if (ik.first >= s.rows || ik.second >= s.cols)
                     throw py::index_error("index out of range");
                 return s.data.data()[ik.first * s.cols + ik.second];
             })
        .def("__repr__", [](const ScoreMatrix& s) {
            return "ScoreMatrix(" + std::to_string(s.rows) + ", " + std::to_string(s.cols) + ")";
        });

    m.def(
        "align_score",
        [](const std::string& q, const std::string& r, const SubstitutionMatrix& mat,
           const std::string& mode, int32_t gap_open, int32_t gap_extend,
           const std::string& alphabet) {
            const Alphabet a = parse_alphabet(alphabet);
            const AlignMode md = parse_align_mode(mode);
            py::gil_scoped_release release;
            return align_score(q, r, mat, a, md, gap_open, gap_extend);
        },
        py::arg("query"), py::arg("ref"), py::arg("matrix"), py::arg("mode") = "global",
        py::arg("gap_open") = 11, py::arg("gap_extend") = 1, py::arg("alphabet") = "aa",
        "Optimal Gotoh similarity score ('global' Needleman-Wunsch, 'local' Smith-Waterman).");

    py::class_<Alignment>(m, "Alignment",
                          "An alignment: aligned strings, ops and the score (the similarity for "
                          "align_pair / align_pairs, the penalty for Index.align).")
        .def_readonly("aligned_query", &Alignment::aligned_query)
        .def_readonly("aligned_ref", &Alignment::aligned_ref)
        .def_readonly("ops", &Alignment::ops)
        .def_readonly("score", &Alignment::score)
        .def("__repr__", [](const Alignment& a) {
            return "Alignment(score=" + std::to_string(a.score) + ", ops='" + a.ops + "')";
        });

    m.def(
        "align_pair",
        [](const std::string& q, const std::string& r, const SubstitutionMatrix& mat,
           const std::string& mode, int32_t gap_open, int32_t gap_extend,
           const std::string& alphabet) {
            const Alphabet a = parse_alphabet(alphabet);
            const AlignMode md = parse_align_mode(mode);
            py::gil_scoped_release release;
            return align_pair(q, r, mat, a, md, gap_open, gap_extend);
        },
        py::arg("query"), py::arg("ref"), py::arg("matrix"), py::arg("mode") = "global",
        py::arg("gap_open") = 11, py::arg("gap_extend") = 1, py::arg("alphabet") = "aa",
        "As align_score, but also returns the aligned strings and ops (Alignment.score is the "
        "similarity, not a penalty).");

    m.def(
        "align_pairs",
        [](const StringSet& q, const StringSet& r,
           const std::vector<uint32_t>& pq, const std::vector<uint32_t>& pr,
           const SubstitutionMatrix& mat, const std::string& mode, int32_t gap_open,
           int32_t gap_extend, const std::string& alphabet) {
            const Alphabet a = parse_alphabet(alphabet);
            const AlignMode md = parse_align_mode(mode);
            for (size_t t = 0; t < pq.size() && t < pr.size(); ++t)
                if (pq[t] >= q.size() || pr[t] >= r.size())
                    throw py::index_error("pair index out of range");
            py::gil_scoped_release release;
            return align_pairs(q, r, pq, pr, mat, a, md, gap_open, gap_extend);
        },
        py::arg("queries"), py::arg("refs"), py::arg("pair_query"), py::arg("pair_ref"),
        py::arg("matrix"), py::arg("mode") = "global", py::arg("gap_open") = 11,
        py::arg("gap_extend") = 1, py::arg("alphabet") = "aa",
        "align_pair for every (queries[pair_query[p]], refs[pair_ref[p]]) in one device batch.");

    m.def(
        "align_score_matrix",
        [](const StringSet& q, const StringSet& r,
           const SubstitutionMatrix& mat, const std::string& mode, int32_t gap_open,
           int32_t gap_extend, const std::string& alphabet, int threads) {
            const Alphabet a = parse_alphabet(alphabet);
            const AlignMode md = parse_align_mode(mode);
            (void)threads;
            return make_matrix(q.size(), r.size(), [&](int32_t* out) {
                align_score_matrix_into(q, r, mat, a, md, gap_open, gap_extend, out);
            });
        },
        py::arg("queries"), py::arg("refs"), py::arg("matrix"), py::arg("mode") = "global",
        py::arg("gap_open") = 11, py::arg("gap_extend") = 1, py::arg("alphabet") = "aa",
        py::arg("threads") = 0, "Dense (n_queries, n_refs) similarity matrix, GIL released.");

    m.def(
        "align_dist_matrix",
        [](const StringSet& q, const StringSet& r,
           const SubstitutionMatrix& mat, const std::string& mode, int32_t gap_open,
           int32_t gap_extend, const std::string& alphabet, int threads) {
            const Alphabet a = parse_alphabet(alphabet);
            const AlignMode md = parse_align_mode(mode);
            (void)threads;
            return make_matrix(q.size(), r.size(), [&](int32_t* out) {
                align_dist_matrix_into(q, r, mat, a, md, gap_open, gap_extend, out);
            });
        },
        py::arg("queries"), py::arg("refs"), py::arg("matrix"), py::arg("mode") = "global",
        py::arg("gap_open") = 11, py::arg("gap_extend") = 1, py::arg("alphabet") = "aa",
        py::arg("threads") = 0,
        "Dense distance matrix d = s(a,a) + s(b,b) - 2*s(a,b), GIL released.");

    m.def(
        "gapblock_matrix",
        [](const StringSet& queries, const StringSet& refs,
           const std::string& alphabet, const std::optional<SubstitutionMatrix>& matrix,
           int32_t gap_open, int32_t gap_extend, const PriorCube& prior,
           uint32_t prior_width, int threads) {
            const Alphabet a = parse_alphabet(alphabet);
            (void)threads;
            return make_matrix(queries.size(), refs.size(), [&](int32_t* out) {
                gapblock_matrix_into(queries, refs, a, matrix ? &*matrix : nullptr, gap_open,
                                     gap_extend, prior.v, prior_width, out);
            });
        },
        py::arg("queries"), py::arg("refs"), py::arg("alphabet") = "aa",
        py::arg("matrix") = std::nullopt, py::arg("gap_open") = 1, py::arg("gap_extend") = 1,
        py::arg("prior") = PriorCube{}, py::arg("prior_width") = 0,
        py::arg("threads") = 0,
        "Exhaustive single-gap-block penalties for every (query, ref) pair, GIL released.");

    m.def(
        "gapblock_threshold",
        [](const StringSet& queries, const StringSet& refs,
           int32_t cutoff, const std::string& alphabet,
           const std::optional<SubstitutionMatrix>& matrix, int32_t gap_open, int32_t gap_extend,
           const PriorCube& prior, uint32_t prior_width) {
            const Alphabet a = parse_alphabet(alphabet);
            GapblockHits h;
            {
                py::gil_scoped_release release;
                h = gapblock_threshold(queries, refs, a, matrix ? &*matrix : nullptr, gap_open,
                                       gap_extend, prior.v, prior_width, cutoff);
            }
            auto arr = [](const auto& v) {
                using T = typename std::decay_t<decltype(v)>::value_type;
                py::array_t<T> out(v.size());
                std::copy(v.begin(), v.end(), out.mutable_data());
                return out;
            };
            return py::make_tuple(arr(h.query), arr(h.ref), arr(h.score));
        },
        py::arg("queries"), py::arg("refs"), py::arg("cutoff"), py::arg("alphabet") = "aa",
        py::arg("matrix") = std::nullopt, py::arg("gap_open") = 1, py::arg("gap_extend") = 1,
        py::arg("prior") = PriorCube{}, py::arg("prior_width") = 0,
        "Every (query, ref) pair with gap-block score <= cutoff as three arrays (query index, "
        "ref index, score), sorted by (query, ref). GIL released.");

    m.def(
        "gapblock_histogram",
        [](const StringSet& queries, const StringSet& refs,
           uint32_t bins, const std::string& alphabet,
           const std::optional<SubstitutionMatrix>& matrix, int32_t gap_open, int32_t gap_extend,
           const PriorCube& prior, uint32_t prior_width) {
            const Alphabet a = parse_alphabet(alphabet);
            std::vector<uint64_t> v;
            {
                py::gil_scoped_release release;
                v = gapblock_histogram(queries, refs, a, matrix ? &*matrix : nullptr, gap_open,
                                       gap_extend, prior.v, prior_width, bins);
            }
            py::array_t<uint64_t> out({py::ssize_t(queries.size()), py::ssize_t(bins) + 2});
            std::copy(v.begin(), v.end(), out.mutable_data());
            return out;
        },
        py::arg("queries"), py::arg("refs"), py::arg("bins"), py::arg("alphabet") = "aa",
        py::arg("matrix") = std::nullopt, py::arg("gap_open") = 1, py::arg("gap_extend") = 1,
        py::arg("prior") = PriorCube{}, py::arg("prior_width") = 0,
        "Per-query histograms of the exact gap-block scores, uint64 [queries][bins + 2]: bin 0 "
        "scores < 0, bin 1 + s the score s < bins, the last bin scores >= bins. GIL released.");

    m.def(
        "gapblock_pairs",
        [](const StringSet& queries, const StringSet& refs,
           const std::vector<uint32_t>& pair_query, const std::vector<uint32_t>& pair_ref,
           const std::string& alphabet, const std::optional<SubstitutionMatrix>& matrix,
           int32_t gap_open, int32_t gap_extend, const PriorCube& prior,
           uint32_t prior_width) {
            const Alphabet a = parse_alphabet(alphabet);
            std::vector<int32_t> v;
            {
                py::gil_scoped_release release;
                v = gapblock_pairs(queries, refs, a, matrix ? &*matrix : nullptr, gap_open,
                                   gap_extend, prior.v, prior_width, pair_query, pair_ref);
            }
            py::array_t<int32_t> out(v.size());
            std::copy(v.begin(), v.end(), out.mutable_data());
            return out;
        },
        py::arg("queries"), py::arg("refs"), py::arg("pair_query"), py::arg("pair_ref"),
        py::arg("alphabet") = "aa", py::arg("matrix") = std::nullopt, py::arg("gap_open") = 1,
        py::arg("gap_extend") = 1, py::arg("prior") = PriorCube{},
        py::arg("prior_width") = 0,
        "Gap-block scores of explicit (query index, ref index) pairs, as an int32 array.");

    m.def(
        "hamming",
        [](const std::string& a, const std::string& b) {
            py::gil_scoped_release release;
            return hamming(a, b);
        },
        py::arg("a"), py::arg("b"),
        "Hamming distance of two equal-length sequences (ValueError otherwise), byte-wise.");
    m.def(
        "levenshtein",
        [](const std::string& a, const std::string& b) {
            py::gil_scoped_release release;
            return levenshtein(a, b);
        },
        py::arg("a"), py::arg("b"), "Unit-cost edit distance, byte-wise and case-sensitive.");

    m.def(
        "hamming_matrix",
        [](const StringSet& a, const StringSet& b, int threads) {
            (void)threads;
            return make_matrix(a.size(), b.size(),
                               [&](int32_t* out) { hamming_matrix_into(a, b, out); });
        },
        py::arg("a"), py::arg("b"), py::arg("threads") = 0,
        "Dense (len(a), len(b)) int32 Hamming matrix, GIL released.");
    m.def(
        "levenshtein_matrix",
        [](const StringSet& a, const StringSet& b, int threads) {
            (void)threads;
            return make_matrix(a.size(), b.size(),
                               [&](int32_t* out) { levenshtein_matrix_into(a, b, out); });
        },
        py::arg("a"), py::arg("b"), py::arg("threads") = 0,
        "Dense (len(a), len(b)) int32 Levenshtein matrix, GIL released.");

    py::class_<PyParams>(m, "SearchParams",
                         "Scope and budget of the reference's search, and the scoring of "
                         "Index.align: matrix (a built-in name, '' for unit cost, or a "
                         "SubstitutionMatrix), gap_open and gap_extend. engine is "
                         "'auto'|'seqtrie'|'seqtm', mode is 'all'|'top'.")
        .def(py::init([](int max_subs, int max_ins, int max_dels, int max_total_edits,
                         long max_penalty, py::object matrix, int gap_open, int gap_extend,
                         std::string engine, std::string mode) {
                 PyParams p;
                 p.max_subs = max_subs;
                 p.max_ins = max_ins;
                 p.max_dels = max_dels;
                 p.max_total_edits = max_total_edits;
                 p.max_penalty = max_penalty;
                 set_matrix(p, matrix);
                 p.gap_open = gap_open;
                 p.gap_extend = gap_extend;
                 parse_engine(engine);  // rejected here, as the reference does
                 parse_mode(mode);
                 p.engine = std::move(engine);
                 p.mode = std::move(mode);
                 return p;
             }),
             py::arg("max_subs") = 0, py::arg("max_ins") = 0, py::arg("max_dels") = 0,
             py::arg("max_total_edits") = 0, py::arg("max_penalty") = 0,
             py::arg("matrix") = "", py::arg("gap_open") = 1, py::arg("gap_extend") = 1,
             py::arg("engine") = "auto", py::arg("mode") = "all")
        .def_readwrite("max_subs", &PyParams::max_subs)
        .def_readwrite("max_ins", &PyParams::max_ins)
        .def_readwrite("max_dels", &PyParams::max_dels)
        .def_readwrite("max_total_edits", &PyParams::max_total_edits)
        .def_readwrite("max_penalty", &PyParams::max_penalty)
        .def_property(
            "matrix",
            [](const PyParams& p) -> py::object {
                if (p.matrix_obj) return py::cast(*p.matrix_obj);
                return py::cast(p.matrix);
            },
            [](PyParams& p, const py::object& v) { set_matrix(p, v); })
        .def_readwrite("gap_open", &PyParams::gap_open)
        .def_readwrite("gap_extend", &PyParams::gap_extend)
        .def_property(
            "engine", [](const PyParams& p) { return p.engine; },
            [](PyParams& p, std::string v) {
                parse_engine(v);
                p.engine = std::move(v);
            })
        .def_property(
            "mode", [](const PyParams& p) { return p.mode; },
            [](PyParams& p, std::string v) {
                parse_mode(v);
                p.mode = std::move(v);
            });

    py::class_<Index>(m, "Index",
                      "A reference set addressed by position (ref_id = index in ``refs``), "
                      "with on-demand global alignment of a query to any reference on the GPU.")
        .def_static(
            "build",
            [](std::vector<std::string> refs, const std::string& alphabet) {
                return Index::build(std::move(refs), parse_alphabet(alphabet));
            },
            py::arg("refs"), py::arg("alphabet") = "aa",
            "Store the references. ``alphabet`` is 'aa', 'nt', or 'iupac'. Raises ValueError "
            "on a symbol outside the alphabet.")
        .def("__len__", &Index::size)
        .def(
            "ref_seq", [](const Index& i, uint32_t id) { return std::string(i.ref_seq(id)); },
            py::arg("ref_id"), "The reference sequence string of a reference id.")
        .def(
            "align",
            [](const Index& idx, uint32_t ref_id, const std::string& q, const PyParams& pp) {
                const auto mat = params_matrix(pp, idx.alphabet());
                const SearchParams cp = to_cpp(pp, mat ? &*mat : nullptr);
                py::gil_scoped_release release;
                return idx.align(q, ref_id, cp);
            },
            py::arg("ref_id"), py::arg("query"), py::arg("params"),
            "Minimum-penalty global alignment of ``query`` to a reference (score is the "
            "penalty: substitution penalties plus gap_open + (L-1)*gap_extend per gap run).")
        .def(
            "align_batch",
            [](const Index& idx, const std::vector<uint32_t>& ref_ids,
               const StringSet& queries, const PyParams& pp) {
                const auto mat = params_matrix(pp, idx.alphabet());
                const SearchParams cp = to_cpp(pp, mat ? &*mat : nullptr);
                py::gil_scoped_release release;
                return idx.align_batch(queries, ref_ids, cp);
            },
            py::arg("ref_ids"), py::arg("queries"), py::arg("params"),
            "align() of every (ref_ids[t], queries[t]) in one device batch.");

    m.def(
        "_longest",
        [](const py::args& sets) {
            Py_ssize_t best = 0;
            auto take = [&](PyObject* it) {
                const Py_ssize_t n = PyUnicode_Check(it) ? PyUnicode_GET_LENGTH(it)
                                                         : PyObject_Length(it);
                if (n < 0) throw py::error_already_set();
                best = std::max(best, n);
            };
            for (const py::handle set : sets) {
                PyObject* o = set.ptr();
                if (PyList_Check(o) || PyTuple_Check(o)) {
                    PyObject** items = PySequence_Fast_ITEMS(o);
                    for (Py_ssize_t i = 0, n = PySequence_Fast_GET_SIZE(o); i < n; ++i)
                        take(items[i]);
                } else {
                    for (const py::handle it : set) take(it.ptr());
                }
            }
            return best;
        },
        "Longest len() over the items of the given sequences (0 when all are empty).");

    m.def("version", [] { return std::string(st_version()); }, "Native library version.");
    m.def("device_count", [] { return st_device_count(); }, "Visible HIP devices.");
}
