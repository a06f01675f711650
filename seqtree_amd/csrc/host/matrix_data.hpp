// Built-in similarity grids as lower triangles (see matrix_data.cpp for provenance).
#pragma once

#include <cstdint>

namespace seqtree {
namespace tables {

constexpr int kAminoSize = 24;
constexpr int kTriangle = kAminoSize * (kAminoSize + 1) / 2;

extern const int8_t kBlosum62Lower[kTriangle];
extern const int8_t kBlosum45Lower[kTriangle];
extern const int8_t kBlosum80Lower[kTriangle];
extern const int8_t kPam250Lower[kTriangle];
extern const int8_t kPam100Lower[kTriangle];
extern const int8_t kStructuralLower[kTriangle];

}  // namespace tables
}  // namespace seqtree
