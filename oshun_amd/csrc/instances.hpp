// Instantiation lists of the heavy kernel templates.
//
// Each family below is compiled in a translation unit of its own
// (k_*.hip: explicit instantiation definitions) so that the library builds in
// parallel; oshun_hip.hip, which launches them, sees explicit instantiation
// declarations only (OSHUN_EXTERN_KERNELS) and compiles no device code for
// them.  A launch of an instance missing here fails at link time
// (-Wl,--no-undefined), never at run time.
#pragma once

#include "kernels.hpp"
#include "stage_lm.hpp"
#include "stage_spec.hpp"

// X(T, L0): 1D stage kernel, l0 = 1..12
#define OSHUN_L0_1_12(X, T) X(T, 1) X(T, 2) X(T, 3) X(T, 4) X(T, 5) X(T, 6) X(T, 7) X(T, 8) X(T, 9) X(T, 10) X(T, 11) X(T, 12)
#define OSHUN_L0_1_8(X, T, ...) X(T, 1, __VA_ARGS__) X(T, 2, __VA_ARGS__) X(T, 3, __VA_ARGS__) X(T, 4, __VA_ARGS__) \
  X(T, 5, __VA_ARGS__) X(T, 6, __VA_ARGS__) X(T, 7, __VA_ARGS__) X(T, 8, __VA_ARGS__)
#define OSHUN_L0_9_12(X, T, ...) X(T, 9, __VA_ARGS__) X(T, 10, __VA_ARGS__) X(T, 11, __VA_ARGS__) X(T, 12, __VA_ARGS__)
#define OSHUN_L0_3_12(X, T, ...) X(T, 3, __VA_ARGS__) X(T, 4, __VA_ARGS__) X(T, 5, __VA_ARGS__) X(T, 6, __VA_ARGS__) \
  X(T, 7, __VA_ARGS__) X(T, 8, __VA_ARGS__) OSHUN_L0_9_12(X, T, __VA_ARGS__)

#define OSHUN_I_STAGE1D(EXT, T, L0) EXT template __global__ void oshun::vlasov_stage_1d<T, L0>(oshun::Stage1DArgs<T>);
#define OSHUN_I_PAIR(EXT, T, L0, R, P) \
  EXT template __global__ void oshun::vlasov_pair_1d<T, L0, R, P>(oshun::Stage1DArgs<T>);
#define OSHUN_I_MARCH(EXT, T, L0, P) \
  EXT template __global__ void oshun::vlasov_march_1d<T, L0, P>(oshun::Stage1DArgs<T>);
#define OSHUN_I_MARCHB(EXT, T, L0, P) \
  EXT template __global__ void oshun::vlasov_march_1d_batch<T, L0, P>(const oshun::Stage1DArgs<T>*, const int*, int);
#define OSHUN_I_PAIRB(EXT, T, L0, R, P) \
  EXT template __global__ void oshun::vlasov_pair_1d_batch<T, L0, R, P>(const oshun::Stage1DArgs<T>*, const int*, int);
#define OSHUN_I_COLL1(EXT, T, NPL, COLS) EXT template __global__ void oshun::collide<T, NPL, COLS>(oshun::CollArgs<T>);
#define OSHUN_I_COLL(EXT, T, NPL) OSHUN_I_COLL1(EXT, T, NPL, 1) OSHUN_I_COLL1(EXT, T, NPL, 2) \
  OSHUN_I_COLL1(EXT, T, NPL, 4) OSHUN_I_COLL1(EXT, T, NPL, 8)
#define OSHUN_I_ROWS(EXT, T, NPL, SEGW, MODE) \
  EXT template __global__ void oshun::collide_rows<T, NPL, SEGW, MODE>(oshun::CollArgs<T>);
#define OSHUN_I_ROWSB(EXT, T, NPL, SEGW, MODE) \
  EXT template __global__ void oshun::collide_rows_batch<T, NPL, SEGW, MODE>(const oshun::CollArgs<T>*, const int*, int);
#define OSHUN_I_GEN(EXT, T) EXT template __global__ void oshun::vlasov_stage<T>(oshun::StageArgs<T>); \
  EXT template __global__ void oshun::vlasov_current_tiles<T>(oshun::StageArgs<T>);
#define OSHUN_I_SPEC(EXT, T, ID, L, M, TD) \
  EXT template __global__ void oshun::vlasov_stage_spec<T, L, M, TD, oshun::kSpecDepth>(oshun::StageArgs<T>);

// ---- families, each with the storage type T ---------------------------------
// 1D stage kernels and single-run stage pairs (R = 3 all l0; R = 2 for l0 <= 8;
// FP64 implicit-E modes 2, 3 for l0 = 3..12 and the probe mode 4 at l0 = 3)
#define OSHUN_P_ONE(T, L0, EXT) OSHUN_I_STAGE1D(EXT, T, L0) OSHUN_I_PAIR(EXT, T, L0, 3, 0) OSHUN_I_PAIR(EXT, T, L0, 3, 1)
#define OSHUN_P_R2(T, L0, EXT) OSHUN_I_PAIR(EXT, T, L0, 2, 0) OSHUN_I_PAIR(EXT, T, L0, 2, 1)
#define OSHUN_P_IMP(T, L0, EXT) OSHUN_I_PAIR(EXT, T, L0, 3, 2) OSHUN_I_PAIR(EXT, T, L0, 3, 3)
#define OSHUN_FAMILY_PAIR(EXT, T)                 \
  OSHUN_L0_1_8(OSHUN_P_ONE, T, EXT)               \
  OSHUN_L0_9_12(OSHUN_P_ONE, T, EXT)              \
  OSHUN_L0_1_8(OSHUN_P_R2, T, EXT)
// single-run stage pairs marched per wavefront: every l0, both pairs
#define OSHUN_P_MARCH(T, L0, EXT) OSHUN_I_MARCH(EXT, T, L0, 0) OSHUN_I_MARCH(EXT, T, L0, 1)
#define OSHUN_FAMILY_MARCH(EXT, T) OSHUN_L0_1_8(OSHUN_P_MARCH, T, EXT) OSHUN_L0_9_12(OSHUN_P_MARCH, T, EXT)
#define OSHUN_FAMILY_PAIR_IMP(EXT) OSHUN_L0_3_12(OSHUN_P_IMP, double, EXT) OSHUN_I_PAIR(EXT, double, 3, 3, 4)

// batched stage pairs: l0 <= 8, R = 2, 3 (and 4 in FP32)
#define OSHUN_PB(T, L0, EXT, R) OSHUN_I_PAIRB(EXT, T, L0, R, 0) OSHUN_I_PAIRB(EXT, T, L0, R, 1)
#define OSHUN_FAMILY_PAIRB(EXT, T) OSHUN_L0_1_8(OSHUN_PB, T, EXT, 2) OSHUN_L0_1_8(OSHUN_PB, T, EXT, 3)
#define OSHUN_FAMILY_PAIRB_F32_R4(EXT) OSHUN_L0_1_8(OSHUN_PB, float, EXT, 4)
// batched marched stage pairs: l0 <= 8
#define OSHUN_MB(T, L0, EXT) OSHUN_I_MARCHB(EXT, T, L0, 0) OSHUN_I_MARCHB(EXT, T, L0, 1)
#define OSHUN_FAMILY_MARCHB(EXT, T) OSHUN_L0_1_8(OSHUN_MB, T, EXT)

// collisions: one wavefront per column (NPL 1, 2, 4, 8) and the lane layouts
// (NPL, SEGW) = (4, 16), (8, 16), (8, 32) in modes 1 (e-e), 2 (l >= 1), 3 (both)
#define OSHUN_ROWS_LAYOUTS(X, EXT, T, MODE) X(EXT, T, 4, 16, MODE) X(EXT, T, 8, 16, MODE) X(EXT, T, 8, 32, MODE)
#define OSHUN_FAMILY_COLL(EXT, T)                                                                          \
  OSHUN_I_COLL(EXT, T, 1) OSHUN_I_COLL(EXT, T, 2) OSHUN_I_COLL(EXT, T, 4) OSHUN_I_COLL(EXT, T, 8)          \
  OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWS, EXT, T, 1) OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWS, EXT, T, 2)                 \
  OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWS, EXT, T, 3)
#define OSHUN_FAMILY_COLLB(EXT, T)                                                                          \
  OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWSB, EXT, T, 1) OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWSB, EXT, T, 2)               \
  OSHUN_ROWS_LAYOUTS(OSHUN_I_ROWSB, EXT, T, 3)

// general system: the term-list interpreter and the specialised shapes
// the l-streaming kernel, m0 = 1..4: every program (kind) in FP64, the full
// program in FP32; with and without y advection
#define OSHUN_I_LM(EXT, T, M0, KIND, TD) \
  EXT template __global__ void oshun::vlasov_stage_lm<T, M0, KIND, TD>(oshun::StageArgs<T>);
#define OSHUN_LM_TD(EXT, T, M0, KIND) OSHUN_I_LM(EXT, T, M0, KIND, false) OSHUN_I_LM(EXT, T, M0, KIND, true)
#define OSHUN_LM_KINDS(EXT, T, M0) OSHUN_LM_TD(EXT, T, M0, 0) OSHUN_LM_TD(EXT, T, M0, 1) OSHUN_LM_TD(EXT, T, M0, 2) \
  OSHUN_LM_TD(EXT, T, M0, 3) OSHUN_LM_TD(EXT, T, M0, 4)
#define OSHUN_FAMILY_LM_F64(EXT) OSHUN_LM_KINDS(EXT, double, 1) OSHUN_LM_KINDS(EXT, double, 2) \
  OSHUN_LM_KINDS(EXT, double, 3) OSHUN_LM_KINDS(EXT, double, 4)
#define OSHUN_FAMILY_LM_F32(EXT) OSHUN_LM_TD(EXT, float, 1, 0) OSHUN_LM_TD(EXT, float, 2, 0) \
  OSHUN_LM_TD(EXT, float, 3, 0) OSHUN_LM_TD(EXT, float, 4, 0)
#define OSHUN_SPEC_F64(ID, L, M, TD) OSHUN_I_SPEC(OSHUN_SPEC_EXT, double, ID, L, M, TD)
#define OSHUN_SPEC_F32(ID, L, M, TD) OSHUN_I_SPEC(OSHUN_SPEC_EXT, float, ID, L, M, TD)

#define OSHUN_EXTERN_KERNELS                                                                  \
  OSHUN_FAMILY_PAIR(extern, double) OSHUN_FAMILY_PAIR(extern, float) OSHUN_FAMILY_PAIR_IMP(extern) \
  OSHUN_FAMILY_MARCH(extern, double) OSHUN_FAMILY_MARCH(extern, float)                        \
  OSHUN_FAMILY_MARCHB(extern, double) OSHUN_FAMILY_MARCHB(extern, float)                      \
  OSHUN_FAMILY_PAIRB(extern, double) OSHUN_FAMILY_PAIRB(extern, float) OSHUN_FAMILY_PAIRB_F32_R4(extern) \
  OSHUN_FAMILY_COLL(extern, double) OSHUN_FAMILY_COLL(extern, float)                          \
  OSHUN_FAMILY_COLLB(extern, double) OSHUN_FAMILY_COLLB(extern, float)                        \
  OSHUN_I_GEN(extern, double) OSHUN_I_GEN(extern, float)                                      \
  OSHUN_FAMILY_LM_F64(extern) OSHUN_FAMILY_LM_F32(extern)
