// Substitution matrices (reference: src/substitution_matrix.cpp:15-86).
#include "matrix_data.hpp"
#include "seqtree.hpp"

#include <algorithm>

namespace seqtree {

namespace {

SubstitutionMatrix from_lower(const int8_t* lower) {
    int32_t grid[tables::kAminoSize * tables::kAminoSize];
    int p = 0;
    for (int a = 0; a < tables::kAminoSize; ++a)
        for (int b = 0; b <= a; ++b, ++p) {
            grid[a * tables::kAminoSize + b] = lower[p];
            grid[b * tables::kAminoSize + a] = lower[p];
        }
    return SubstitutionMatrix::from_similarity(uint8_t(tables::kAminoSize), grid);
}

}  // namespace

SubstitutionMatrix SubstitutionMatrix::unit(uint8_t size) {
    SubstitutionMatrix m;
    m.size_ = size;
    m.pen_.assign(size_t(size) * size, 1);
    m.sim_.assign(size_t(size) * size, 0);
    for (uint8_t i = 0; i < size; ++i) {
        m.pen_[size_t(i) * size + i] = 0;
        m.sim_[size_t(i) * size + i] = 1;
    }
    return m;
}

SubstitutionMatrix SubstitutionMatrix::from_similarity(uint8_t size, const int32_t* sim) {
    SubstitutionMatrix m;
    m.size_ = size;
    m.sim_.assign(sim, sim + size_t(size) * size);
    m.pen_.resize(size_t(size) * size);
    for (uint32_t a = 0; a < size; ++a)
        for (uint32_t b = 0; b < size; ++b) {
            // Gram transform in wrapping int32 arithmetic, clamped at zero
            const uint32_t saa = uint32_t(sim[a * size + a]);
            const uint32_t sbb = uint32_t(sim[b * size + b]);
            const uint32_t sab = uint32_t(sim[a * size + b]);
            const int32_t p = int32_t(saa + sbb - 2u * sab);
            m.pen_[a * size + b] = p < 0 ? 0 : p;
        }
    return m;
}

SubstitutionMatrix SubstitutionMatrix::blosum62() { return from_lower(tables::kBlosum62Lower); }
SubstitutionMatrix SubstitutionMatrix::blosum45() { return from_lower(tables::kBlosum45Lower); }
SubstitutionMatrix SubstitutionMatrix::blosum80() { return from_lower(tables::kBlosum80Lower); }
SubstitutionMatrix SubstitutionMatrix::pam250() { return from_lower(tables::kPam250Lower); }
SubstitutionMatrix SubstitutionMatrix::pam100() { return from_lower(tables::kPam100Lower); }
SubstitutionMatrix SubstitutionMatrix::structural() {
    return from_lower(tables::kStructuralLower);
}

int32_t SubstitutionMatrix::scale() const {
    if (size_ < 2) return 0;
    std::vector<int32_t> off;
    off.reserve(size_t(size_) * (size_ - 1));
    for (uint32_t a = 0; a < size_; ++a)
        for (uint32_t b = 0; b < size_; ++b)
            if (a != b) off.push_back(pen_[a * size_ + b]);
    // the element of rank |off| / 2 in ascending order
    std::sort(off.begin(), off.end());
    return off[off.size() / 2];
}

}  // namespace seqtree
