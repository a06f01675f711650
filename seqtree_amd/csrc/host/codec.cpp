// Alphabet codec (reference: src/codec.cpp:9-35). The symbol orders are fixed by the
// substitution matrices: amino acids follow the BLOSUM row order, so a code indexes a matrix
// row directly.
#include "seqtree.hpp"

#include <cstring>

namespace seqtree {

namespace {

const char* symbols(Alphabet a) {
    switch (a) {
        case Alphabet::Nucleotide: return "ACGT";
        case Alphabet::NucleotideIUPAC: return "ACGTRYSWKMBDHVN";
        case Alphabet::AminoAcid: break;
    }
    return "ARNDCQEGHILKMFPSTWYVBZX*";
}

}  // namespace

Codec::Codec(Alphabet a) : alphabet_(a) {
    std::memset(enc_, kInvalid, sizeof(enc_));
    std::memset(dec_, 0, sizeof(dec_));
    uint8_t code = 0;
    for (const char* p = symbols(a); *p; ++p, ++code) {
        const unsigned char up = static_cast<unsigned char>(*p);
        enc_[up] = code;
        if (up >= 'A' && up <= 'Z') enc_[up - 'A' + 'a'] = code;  // case-insensitive input
        dec_[code] = *p;
    }
    size_ = code;
}

}  // namespace seqtree
