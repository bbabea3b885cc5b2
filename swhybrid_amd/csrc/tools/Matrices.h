// Matrices.h -- named substitution matrices as 32x32 int32 tables in residue
// code order (row = query code, column = database code; codes 23..31 score 0).
#pragma once
#include <cstdint>

namespace swhy {
// "blosum62" or "blosum50"; false for other names.
bool named_matrix(const char* name, int32_t* out32x32);
}  // namespace swhy
