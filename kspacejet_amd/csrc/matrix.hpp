// Helpers of the matrix-core gather kernels (2-D and 3-D): the v_mfma_f32_16x16x4_f32
// accumulator type and vector loads / stores of a lane's NF consecutive floats.
#pragma once

#include "common.hpp"

namespace ksj::gpu {

typedef float MfmaAcc __attribute__((ext_vector_type(4)));

template <int NF>
struct RowPiece {
  float v[NF];
};

// Which floats of a 16 NF-float row lane j of a 16-lane group holds: NF consecutive
// floats from NF j, except for NF = 8 (64 coils), where the lane holds the float4s at 4 j
// and 64 + 4 j, so each load or store instruction of the 16 lanes covers 256 contiguous
// bytes (whole cache lines) instead of every other 16 bytes of the 512-byte row.  Float
// v[n] of the lane is the column-j operand of product n; any lane-to-float mapping is
// valid as long as loads and stores share it.
template <int NF>
__device__ inline int piece_offset(int j) {
  return NF == 8 ? 4 * j : NF * j;
}
template <int NF>
constexpr int piece_chunk_stride() {  // floats between the lane's float4 chunks
  return NF == 8 ? 64 : 4;
}

template <int NF>
__device__ inline RowPiece<NF> load_piece(const float* p) {
  RowPiece<NF> r;
  if constexpr (NF == 1) {
    r.v[0] = *p;
  } else if constexpr (NF == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    r.v[0] = t.x;
    r.v[1] = t.y;
  } else {
#pragma unroll
    for (int h = 0; h < NF / 4; ++h) {
      const float4 t = *reinterpret_cast<const float4*>(p + h * piece_chunk_stride<NF>());
      r.v[4 * h] = t.x;
      r.v[4 * h + 1] = t.y;
      r.v[4 * h + 2] = t.z;
      r.v[4 * h + 3] = t.w;
    }
  }
  return r;
}

template <int NF>
__device__ inline void store_piece(float* p, const float (&v)[NF]) {
  if constexpr (NF == 1) {
    *p = v[0];
  } else if constexpr (NF == 2) {
    *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
  } else {
#pragma unroll
    for (int h = 0; h < NF / 4; ++h)
      *reinterpret_cast<float4*>(p + h * piece_chunk_stride<NF>()) =
        make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
  }
}

template <int NF>
__device__ inline void store_piece_nt(float* p, const float (&v)[NF]) {
  if constexpr (NF == 1) {
    __builtin_nontemporal_store(v[0], p);
  } else if constexpr (NF == 2) {
    __builtin_nontemporal_store(v[0], p);
    __builtin_nontemporal_store(v[1], p + 1);
  } else {
#pragma unroll
    for (int h = 0; h < NF / 4; ++h) {
      typedef float F4 __attribute__((ext_vector_type(4)));
      const F4 t = {v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]};
      __builtin_nontemporal_store(t, reinterpret_cast<F4*>(p + h * piece_chunk_stride<NF>()));
    }
  }
}

}  // namespace ksj::gpu
