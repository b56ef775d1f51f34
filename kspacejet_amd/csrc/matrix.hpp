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
      const float4 t = reinterpret_cast<const float4*>(p)[h];
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
      reinterpret_cast<float4*>(p)[h] = make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
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
      __builtin_nontemporal_store(t, reinterpret_cast<F4*>(p) + h);
    }
  }
}

}  // namespace ksj::gpu
