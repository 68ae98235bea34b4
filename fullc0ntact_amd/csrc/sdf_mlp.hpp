// Learned signed-distance surrogate: one small MLP per rigid body.
//
//   x = (q - center) / halfExtent                       (model frame, [-1,1]^3)
//   f = [sin(2 pi B x), cos(2 pi B x)]                   (B: 16 x 3, fixed)
//   h1 = SiLU(f W1 + b1), h2 = SiLU(h1 W2 + b2), h3 = SiLU(h2 W3 + b3)
//   y = h3 w4 + b4,   distance = y * halfExtent
//
// Widths 32 -> 64 -> 64 -> 64 -> 1.  Parameters per body, FP32, in the
// order W1[32][64] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] w4[64] b4
// (W[i][o]: input i, output o).
#pragma once

#include <cstdint>

namespace fc {

constexpr int kSdfFreqs = 16;
constexpr int kSdfIn = 2 * kSdfFreqs;  // 32 features
constexpr int kSdfHidden = 64;
constexpr int kSdfOffW1 = 0;
constexpr int kSdfOffB1 = kSdfOffW1 + kSdfIn * kSdfHidden;        // 2048
constexpr int kSdfOffW2 = kSdfOffB1 + kSdfHidden;                 // 2112
constexpr int kSdfOffB2 = kSdfOffW2 + kSdfHidden * kSdfHidden;    // 6208
constexpr int kSdfOffW3 = kSdfOffB2 + kSdfHidden;                 // 6272
constexpr int kSdfOffB3 = kSdfOffW3 + kSdfHidden * kSdfHidden;    // 10368
constexpr int kSdfOffW4 = kSdfOffB3 + kSdfHidden;                 // 10432
constexpr int kSdfOffB4 = kSdfOffW4 + kSdfHidden;                 // 10496
constexpr int kSdfParams = kSdfOffB4 + 1;                         // 10497
// Per-body stride of the parameter arrays (keeps every body's block and its
// bias vectors 16-byte aligned for vector loads); entries beyond kSdfParams
// are padding.
constexpr int kSdfStride = 10500;

// Packed BF16 A-operand fragments of the inference kernel per body:
// layer 1: 4 output blocks x 1 k-block, layers 2 and 3: 4 x 2; each
// fragment is 64 lanes x 8 bf16.
constexpr int kSdfFrags = 4 + 8 + 8;
constexpr int kSdfFragElems = 64 * 8;
constexpr int kSdfPackedBf16 = kSdfFrags * kSdfFragElems;  // 10240

// Row of the K dimension held by element j (0..7) of lane group g (lane/16)
// in one 32-row k-block of the 16x16x32 BF16 MFMA operands, as used here:
// the first four elements take rows 4g..4g+3, the last four 16+4g..16+4g+3.
// These are exactly the rows a 16x16 accumulator block leaves in lane group
// g (rows 4g+t) for the two 16-row halves of the k-block, so one layer's
// outputs feed the next layer without a transpose.  Any permutation of K is
// a valid contraction as long as both operands use the same one.
inline constexpr int sdf_krow(int g, int j) { return j < 4 ? 4 * g + j : 16 + 4 * g + (j - 4); }

}  // namespace fc
