// 16-bit MFMA helpers shared by the mixed-precision scan kernels: operand
// types, the v_mfma_f32_16x16x32_{f16,bf16} wrapper and f32 <-> 16-bit
// conversions (round to nearest even).
#pragma once
#include "common.hpp"

namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bfv8 __attribute__((ext_vector_type(8)));
typedef float float4v __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <int TY>
__device__ __forceinline__ float4v mfma(u16x8 a, u16x8 b, float4v c) {
    if constexpr (TY == RR_PREC_F16) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a),
                                                      __builtin_bit_cast(half8, b), c, 0, 0, 0);
    } else {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bfv8, a),
                                                       __builtin_bit_cast(bfv8, b), c, 0, 0, 0);
    }
}

template <int TY>
__device__ __forceinline__ unsigned short to16(float f) {
    if constexpr (TY == RR_PREC_F16) return rr::f32_to_f16_bits(f);
    else return rr::f32_to_bf16_bits(f);
}

typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef __bf16 bfv4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));

// two MFMA outputs -> one 8-wide 16-bit operand, round to nearest even
// (v_cvt_pk_{f16,bf16}_f32: one instruction per pair)
template <int TY>
__device__ __forceinline__ u16x8 to16x8(float4v a, float4v b) {
    u16x4 lo, hi;
    if constexpr (TY == RR_PREC_F16) {
        lo = __builtin_bit_cast(u16x4, __builtin_convertvector(a, half4));
        hi = __builtin_bit_cast(u16x4, __builtin_convertvector(b, half4));
    } else {
        lo = __builtin_bit_cast(u16x4, __builtin_convertvector(a, bfv4));
        hi = __builtin_bit_cast(u16x4, __builtin_convertvector(b, bfv4));
    }
    return u16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
}

template <int TY>
__device__ __forceinline__ float from16(unsigned short h) {
    if constexpr (TY == RR_PREC_F16) return __half2float(__ushort_as_half(h));
    else return __uint_as_float(((unsigned)h) << 16);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace
