// mlp_pack.hpp — the classifier's weights in the matrix-core operand layout
// (host side).  The inference runs the layers transposed on 16x16x16 BF16
// MFMAs; the A operand of lane l (column col = l % 16, k block kb = l / 16)
// holds W^T[i = col][k = 4 kb + r], r = 0..3.  Each lane needs eight such
// 4-vectors:
//   0, 1: W1 columns  0..15 / 16..31 (k = input index)
//   2, 3: W2 rows 0..15 / 16..31 for output columns 0..15
//   4, 5: the same for output columns 16..31
//   6, 7: W3 (only the lanes with col == 0; zero elsewhere), k halves 0 / 1
// Packed entry (v, lane) holds 4 BF16 values (round to nearest even) at
// uint16 index ((v * 64) + lane) * 4 + r, so a wave reads each vector with
// one coalesced 8-byte load per lane.
#pragma once
#include <cstdint>
#include <cstring>

namespace mlp {

constexpr int kPackVecs = 8;
constexpr int kPackHalves = kPackVecs * 64 * 4;  // uint16 entries

inline uint16_t bf16_rne(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// P: the packed float parameter vector of gks_learn.h
inline void pack_bf16(const float* P, uint16_t* out) {
  constexpr int W1 = 0, W2 = 544, W3 = 1600;
  for (int lane = 0; lane < 64; lane++) {
    const int col = lane & 15, kb = lane >> 4;
    for (int r = 0; r < 4; r++) {
      const int k = 4 * kb + r;
      float v[kPackVecs];
      for (int m = 0; m < 2; m++) {
        v[m] = P[W1 + k * 32 + m * 16 + col];
        for (int kk = 0; kk < 2; kk++) v[2 + 2 * m + kk] = P[W2 + (kk * 16 + k) * 32 + m * 16 + col];
        v[6 + m] = col == 0 ? P[W3 + m * 16 + k] : 0.0f;
      }
      for (int q = 0; q < kPackVecs; q++) out[((q * 64) + lane) * 4 + r] = bf16_rne(v[q]);
    }
  }
}

}  // namespace mlp
