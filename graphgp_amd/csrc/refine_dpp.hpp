// refine_dpp.hpp -- FP32 trailing updates of the register Cholesky with the column
// broadcast folded into the FMA: v_fmac_f32_dpp acc, -src(quad_perm S), m gives
// acc += -(lane S's src) * m, the same fused operation as a separate DPP move followed by
// v_fma_f32 (bit-identical), without the move.  One asm statement updates NR rows from one
// broadcast source.  A DPP read of a VGPR that a VALU instruction wrote in the two
// preceding wait states is a hazard the compiler does not see through inline asm, so each
// statement starts with s_nop 1; within a statement the source is not written.
#pragma once

namespace ggp {

template <int S, int NR>
struct FmacDpp;

template <int S>
struct FmacDpp<S, 1> {
  __device__ static __forceinline__ void run(float src, float& a0, float m0) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0)
        : [src] "v"(src), [m0] "v"(m0), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 2> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float m0, float m1) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 3> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float m0, float m1, float m2) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 4> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float m0, float m1, float m2, float m3) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a3], -%[src], %[m3] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 5> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float m0, float m1, float m2, float m3, float m4) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a3], -%[src], %[m3] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a4], -%[src], %[m4] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 6> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float m0, float m1, float m2, float m3, float m4, float m5) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a3], -%[src], %[m3] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a4], -%[src], %[m4] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a5], -%[src], %[m5] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 7> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float& a6, float m0, float m1, float m2, float m3, float m4, float m5, float m6) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a3], -%[src], %[m3] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a4], -%[src], %[m4] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a5], -%[src], %[m5] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a6], -%[src], %[m6] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [s] "i"(S));
  }
};

template <int S>
struct FmacDpp<S, 8> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float& a6, float& a7, float m0, float m1, float m2, float m3, float m4, float m5, float m6, float m7) {
    asm("s_nop 1\n"
        "v_fmac_f32_dpp %[a0], -%[src], %[m0] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a1], -%[src], %[m1] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a2], -%[src], %[m2] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a3], -%[src], %[m3] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a4], -%[src], %[m4] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a5], -%[src], %[m5] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a6], -%[src], %[m6] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        "v_fmac_f32_dpp %[a7], -%[src], %[m7] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"
        : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6), [a7] "+v"(a7)
        : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [m7] "v"(m7), [s] "i"(S));
  }
};

}  // namespace ggp
