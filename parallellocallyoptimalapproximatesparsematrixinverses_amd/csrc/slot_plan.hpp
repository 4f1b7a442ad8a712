// Slot plan of the fixed-pattern products (A Y)(:, j) on S2(:, j): for every
// column j and every product term -- k over S(:, j) ascending, i over A(:, k)
// ascending, the order in which the products visit them -- the position of
// row i in S2(:, j).  With it a product adds each term into its accumulator
// slot without searching.  Built once per pattern (fixed-pattern builders and
// the differentiation form).
#pragma once

#include <cstdint>

#include "ops.hpp"

namespace spai {

struct SlotPlan {
  DBuf<int64_t> ptr;      // n + 1 term offsets (ptr[j] = first term of column j)
  DBuf<uint8_t> plan8;    // slots, when every S2 column has at most 256 entries
  DBuf<uint16_t> plan16;  // slots, when some S2 column is longer (at most 65535)
  bool slot16 = false;
  bool built = false;     // false: an S2 column exceeds 65535 entries (no plan)
  int64_t terms = 0;
  int max_a_col = 0, max_s2_col = 0, max_s_col = 0;
};

// Builds the plan of A (sp/si: S, ap/ai: A, s2p/s2i: S2, all n columns).  A
// term whose row is missing from S2(:, j) raises a device consistency error
// at the context's next synchronisation.  One host read-back (the sizes).
SlotPlan BuildSlotPlan(Ctx& c, int64_t n, const int* sp, const int* si, const int* ap, const int* ai,
                       const int* s2p, const int* s2i);

// out[0..n] = exclusive scan of in[0..n) (out[n] = the sum), int64; no read-back.
void ScanInt64(Ctx& c, const int64_t* in, int64_t n, int64_t* out);

}  // namespace spai
