// refine_dpp.hpp -- FP32 trailing updates of the register Cholesky with the column
// broadcast folded into the FMA: v_fmac_f32_dpp acc, -src(quad_perm S), m gives
// acc += -(lane S's src) * m, the same fused operation as a separate DPP move followed by
// v_fma_f32 (bit-identical), without the move.  One asm statement updates NR rows from one
// broadcast source.  A DPP read of a VGPR that a VALU instruction wrote in the two
// preceding wait states is a hazard the compiler does not see through inline asm, so a
// statement whose source may have been written just before it starts with s_nop 1 (NOP);
// within a statement the source is not written.  (The compiler does see the statements'
// writes, and pads its own DPP reads after them.)
#pragma once

namespace ggp {

#define GGP_FMAC(i) \
  "v_fmac_f32_dpp %[a" #i "], -%[src], %[m" #i "] quad_perm:[%[s],%[s],%[s],%[s]] row_mask:0xf bank_mask:0xf bound_ctrl:1\n"

template <int S, int NR, bool NOP>
struct FmacDpp;

template <int S, bool NOP>
struct FmacDpp<S, 1, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float m0) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0)
          : [a0] "+v"(a0)
          : [src] "v"(src), [m0] "v"(m0), [s] "i"(S));
    else
      asm(GGP_FMAC(0)
          : [a0] "+v"(a0)
          : [src] "v"(src), [m0] "v"(m0), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 2, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float m0, float m1) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1)
          : [a0] "+v"(a0), [a1] "+v"(a1)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1)
          : [a0] "+v"(a0), [a1] "+v"(a1)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 3, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float m0, float m1, float m2) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 4, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float m0, float m1, float m2, float m3) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 5, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float m0, float m1, float m2, float m3, float m4) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 6, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float m0, float m1, float m2, float m3, float m4, float m5) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 7, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float& a6, float m0, float m1, float m2, float m3, float m4, float m5, float m6) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5) GGP_FMAC(6)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5) GGP_FMAC(6)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [s] "i"(S));
  }
};

template <int S, bool NOP>
struct FmacDpp<S, 8, NOP> {
  __device__ static __forceinline__ void run(float src, float& a0, float& a1, float& a2, float& a3, float& a4, float& a5, float& a6, float& a7, float m0, float m1, float m2, float m3, float m4, float m5, float m6, float m7) {
    if constexpr (NOP)
      asm("s_nop 1\n" GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5) GGP_FMAC(6) GGP_FMAC(7)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6), [a7] "+v"(a7)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [m7] "v"(m7), [s] "i"(S));
    else
      asm(GGP_FMAC(0) GGP_FMAC(1) GGP_FMAC(2) GGP_FMAC(3) GGP_FMAC(4) GGP_FMAC(5) GGP_FMAC(6) GGP_FMAC(7)
          : [a0] "+v"(a0), [a1] "+v"(a1), [a2] "+v"(a2), [a3] "+v"(a3), [a4] "+v"(a4), [a5] "+v"(a5), [a6] "+v"(a6), [a7] "+v"(a7)
          : [src] "v"(src), [m0] "v"(m0), [m1] "v"(m1), [m2] "v"(m2), [m3] "v"(m3), [m4] "v"(m4), [m5] "v"(m5), [m6] "v"(m6), [m7] "v"(m7), [s] "i"(S));
  }
};

#undef GGP_FMAC

}  // namespace ggp
