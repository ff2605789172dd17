// The general system's stage kernel for any l0 at m0 = 1..4 (the production
// decks' l0 = 16..48 among them): one RK stage of a coupling program for all
// planes of one (x, p, y) point.
//
// Every term of Actions::Adv (and of the implicit step's Adv1 / Adv2) that
// feeds a target harmonic (l, m) comes from a source at l-1, l or l+1
// (def-VlasovMax.cpp:346-644, 757-951), so the thread walks l upward and
// keeps three layers of source values in registers: for layer l-1 the G of
// MakeGH / G00 and the x, y differences, for layer l the value itself (the B
// rotation), for layer l+1 the H and the differences.  At each l the stencil
// of layer l+1 is loaded and formed together with the RK operands of layer l;
// the other resident workgroups cover that wait (no look-ahead buffer, which
// would cost the registers that keep several waves per SIMD resident).
//
// The targets' sums are formed in the coupling program's order (tables.hpp
// build_tables: E field by source, m outer and l inner; then x advection, y
// advection, B rotation) one add_complex call at a time: each call adds
// OP(a + i b) (Mr + i Mi) with at most one field term in Mr (c times a field
// or v(p)) and one in Mi (d times a field), i.e. re += OP(a) Mr - OP(b) Mi,
// im += OP(a) Mi + OP(b) Mr.  The products Mr = field * c and Mi = field * d
// are formed once per call; a term -OP(b) Mi is a subtraction, which IEEE
// arithmetic makes identical to the program's addition of OP(b) (field *
// (-d)).  So every slope is bitwise the term-list interpreter's
// (vlasov_stage).  The host replays this call order against the program's
// term list at set-up (lm_build) and keeps the interpreter for a program it
// does not match.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <type_traits>
#include <vector>

#include "kernels.hpp"

namespace oshun {

enum LmPhase { LM_E = 1, LM_X = 2, LM_Y = 4, LM_B = 8 };
// The coupling programs the kernel is built for: Actions::Adv, and
// Implicit_Actions::Adv1 with and without the B rotation, Adv2, and the
// field-solve probe (Implicit_Ex / Implicit_Eyz: the E-field couplings of
// sources l <= 3 into targets l <= 2, Ex only from sources m <= 1;
// def-Actions.cpp:73-94, 118-160, def-VlasovMax.cpp:157-196, 231-291)
enum LmKind { kLmFull = 0, kLmAdv1 = 1, kLmAdv1Es = 2, kLmAdv2 = 3, kLmProbe = 4, kLmKinds = 5 };
__host__ __device__ constexpr int lm_mask(int kind) {
  return kind == kLmFull ? (LM_E | LM_X | LM_Y | LM_B)
                         : (kind == kLmAdv1 ? (LM_X | LM_Y | LM_B) : (kind == kLmAdv1Es ? (LM_X | LM_Y) : LM_E));
}
// the probe's top source layer
constexpr int kLmProbeTop = 3;
// CLEAN_LOW_PC rows of a plane of harmonic l (def-Actions.cpp:240-253) in the
// programs that have them (Actions::Adv and Adv2; tables.hpp build_tables)
__host__ __device__ constexpr int lm_clean_rows(int kind, int l) {
  return (kind == kLmFull || kind == kLmAdv2) ? (l >= 1 ? (l - 1 < 80 ? l - 1 : 80) : 0) : 0;
}

constexpr int kLmMaxM0 = 4;
constexpr int kLmNone = -1;  // no field term in Mr (Mi)

// slot of (m, re|im) within a layer: m = 0, then re and im of each m > 0
__host__ __device__ constexpr int lm_slot(int m, int im) { return m == 0 ? 0 : 2 * m - 1 + im; }
// first plane of harmonic layer l (planes in ind(l, m) order, re then im)
__host__ __device__ constexpr int lm_layer_start(int l, int m0) {
  return l <= m0 + 1 ? l * l : (m0 + 1) * (m0 + 1) + (l - m0 - 1) * (2 * m0 + 1);
}

// The add_complex calls into target (l, TM) in program order.  S provides
//   template <int DL, int SM, int OP, int MR, int MI, bool TB> void g()
// -- target += OP(source (l + DL, SM)) (Mr + i Mi), MR / MI the field (Mul)
// of Mr / Mi or kLmNone, TB: the target has an imaginary plane (TM > 0).
// IN: l is interior, M0 + 1 <= l <= L - 1, so that every source l-1, l, l+1
// with m' <= M0 exists and l >= 2: the same calls, the existence tests
// resolved at compile time.
template <int M0, int TM, int KIND, bool TD, bool IN, class S>
__host__ __device__ __forceinline__ void lm_groups(S& s, int l, int L) {
  constexpr int MASK = lm_mask(KIND);
  constexpr bool PROBE = KIND == kLmProbe;
  constexpr bool TB = TM > 0;
  if (PROBE && l > kLmProbeTop - 1) return;
  // source (l + dl, sm) exists
  auto ok = [&](int dl, int sm) -> bool {
    if constexpr (IN) return true;
    else return sm >= 0 && sm <= M0 && l + dl >= 0 && l + dl <= L && sm <= l + dl;
  };
  // ---- E field (ElectricField::Exyz): sources m' = TM-1, TM, TM+1; G of
  // l' = l-1 then H of l' = l+1.  Ex (m' = TM): Mr = Ex c; Em (m' = TM-1) and
  // Ep (m' = TM+1): Mr = Ey c, Mi = Ez d
  if constexpr ((MASK & LM_E) != 0) {
    auto from = [&](auto smc) {
      constexpr int SM = decltype(smc)::value;
      if constexpr (SM >= 0 && SM <= M0) {
        constexpr int MR = SM == TM ? MUL_EX : MUL_EY;
        constexpr int MI = SM == TM ? kLmNone : MUL_EZ;
        if constexpr (PROBE && MR == MUL_EX && SM > 1) return;  // the probe's Ex: sources m <= 1
        if (ok(-1, SM)) {
          if (!IN && l == 1) {  // G00 of f00 into (1,0) (Ex) and (1,1) (Em)
            if constexpr (SM == 0 && TM <= 1) s.template g<-1, 0, OP_G00, MR, MI, TB>();
          } else {
            s.template g<-1, SM, OP_G, MR, MI, TB>();
          }
        }
        if (ok(1, SM)) s.template g<1, SM, OP_H, MR, MI, TB>();
      }
    };
    from(std::integral_constant<int, TM - 1>{});
    from(std::integral_constant<int, TM>{});
    from(std::integral_constant<int, TM + 1>{});
  }
  // ---- x advection (Spatial_Advection::Axy): (l-1, TM) then (l+1, TM), Mr = v c
  if constexpr ((MASK & LM_X) != 0) {
    if (ok(-1, TM)) s.template g<-1, TM, OP_DX, MUL_V, kLmNone, TB>();
    if (ok(1, TM)) s.template g<1, TM, OP_DX, MUL_V, kLmNone, TB>();
  }
  // ---- y advection: sources m' = TM-1 then TM+1, each from l-1 then l+1
  if constexpr ((MASK & LM_Y) != 0 && TD) {
    auto yfrom = [&](auto smc) {
      constexpr int SM = decltype(smc)::value;
      if constexpr (SM >= 0 && SM <= M0) {
        if (ok(-1, SM)) s.template g<-1, SM, OP_DY, MUL_V, kLmNone, TB>();
        if (ok(1, SM)) s.template g<1, SM, OP_DY, MUL_V, kLmNone, TB>();
      }
    };
    yfrom(std::integral_constant<int, TM - 1>{});
    yfrom(std::integral_constant<int, TM + 1>{});
  }
  // ---- B rotation (MagneticField::Bxyz): (l, TM-1) Bp/2 (Mr = Bz c, Mi = By d),
  // (l, TM) -i TM Bx (Mi only), (l, TM+1) Bm (Mr = Bz c, Mi = By d)
  if constexpr ((MASK & LM_B) != 0) {
    if (IN || l >= 1) {
      if constexpr (TM >= 1) {
        s.template g<0, TM - 1, OP_ID, MUL_BZ, MUL_BY, TB>();
        s.template g<0, TM, OP_ID, kLmNone, MUL_BX, TB>();
      }
      if constexpr (TM + 1 <= M0) {
        if (ok(0, TM + 1)) s.template g<0, TM + 1, OP_ID, MUL_BZ, MUL_BY, TB>();
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Host: the call order against a program's term list

// One expected term: source plane, operator, source l, multiplier, and the
// coefficient it carries: +c, +d or -d of call `grp` (role 0: c, 1: d)
struct LmTerm {
  int src, op, l, mul, grp, role, sign;
};

struct LmRecord {
  const HostTables* T;
  int l;
  std::vector<LmTerm>* out[2];              // the target's real and imaginary term lists
  std::vector<std::pair<int, int>>* slots;  // coefficients the kernel reads: (call, role) in order
  int ngrp = 0;
  template <int DL, int SM, int OP, int MR, int MI, bool TB>
  void g() {
    const int sl = l + DL;
    const int h = hind(sl, SM, T->m0);
    const int sa = T->comp_re[h], sb = SM > 0 ? T->comp_im[h] : -1;
    const int id = ngrp++;
    if (MR != kLmNone) out[0]->push_back({sa, OP, sl, MR, id, 0, 1});
    if (MI != kLmNone && sb >= 0) out[0]->push_back({sb, OP, sl, MI, id, 1, -1});
    if (TB) {
      if (MI != kLmNone) out[1]->push_back({sa, OP, sl, MI, id, 1, 1});
      if (MR != kLmNone && sb >= 0) out[1]->push_back({sb, OP, sl, MR, id, 0, 1});
    }
    if (MR != kLmNone) slots->push_back({id, 0});
    if (MI != kLmNone && (TB || sb >= 0)) slots->push_back({id, 1});
  }
};

template <int M0, int KIND, bool TD>
void lm_record_target(LmRecord& r, int m, int L) {
  switch (m) {
    case 0: lm_groups<M0, 0, KIND, TD, false>(r, r.l, L); break;
    case 1: if constexpr (M0 >= 1) lm_groups<M0, 1, KIND, TD, false>(r, r.l, L); break;
    case 2: if constexpr (M0 >= 2) lm_groups<M0, 2, KIND, TD, false>(r, r.l, L); break;
    case 3: if constexpr (M0 >= 3) lm_groups<M0, 3, KIND, TD, false>(r, r.l, L); break;
    case 4: if constexpr (M0 >= 4) lm_groups<M0, 4, KIND, TD, false>(r, r.l, L); break;
    default: break;
  }
}

// The kernel's coefficients of a program: gstart[plane] (real plane of each
// target harmonic; nc + 1 entries) indexes coef, which holds for each target
// the c and d of its add_complex calls in the order the kernel reads them.
struct LmPlan {
  int kind = -1;  // LmKind, or -1: the interpreter runs the program
  std::vector<int> gstart;
  std::vector<double> coef;
  // interior layers l = m0+1 .. l0-1: the coefficients of target (l, m)
  // start at istart + (l - m0 - 1) istride + ioff[m]
  int istart = 0, istride = 0, ioff[kLmMaxM0 + 1] = {0, 0, 0, 0, 0};
};

// True (and *P filled) when the program (terms, tstart over all planes)
// holds, target by target, exactly the terms of the kernel's add_complex
// calls for KIND, in the same order (source plane, operator, multiplier, and
// source l), with each call's c and d consistent across its terms.
template <int M0, int KIND, bool TD>
bool lm_build_kind(const HostTables& T, const std::vector<Term>& terms, const std::vector<int>& tstart,
                   const std::vector<int>& clean, LmPlan* P) {
  if (T.m0 != M0 || T.two_d != TD || (int)tstart.size() != T.nc + 1 || (int)clean.size() != T.nc) return false;
  for (int c = 0; c < T.nc; ++c)
    if (clean[c] != lm_clean_rows(KIND, T.comp_l[c])) return false;
  P->gstart.assign(T.nc + 1, 0);
  P->coef.clear();
  for (int l = 0; l <= T.l0; ++l)
    for (int m = 0; m <= (l < M0 ? l : M0); ++m) {
      std::vector<LmTerm> e[2];
      std::vector<std::pair<int, int>> slots;
      LmRecord r{&T, l, {&e[0], &e[1]}, &slots};
      lm_record_target<M0, KIND, TD>(r, m, T.l0);
      const int h = hind(l, m, M0);
      const int ca = T.comp_re[h];
      if (ca != lm_layer_start(l, M0) + lm_slot(m, 0)) return false;
      if (m == 0 && !e[1].empty()) return false;
      std::vector<double> val(2 * r.ngrp, 0.0);
      std::vector<char> have(2 * r.ngrp, 0);
      for (int im = 0; im < (m > 0 ? 2 : 1); ++im) {
        const int tc = im ? T.comp_im[h] : ca;
        if ((int)e[im].size() != tstart[tc + 1] - tstart[tc]) return false;
        for (size_t k = 0; k < e[im].size(); ++k) {
          const Term& a = terms[tstart[tc] + k];
          const LmTerm& b = e[im][k];
          if (a.src != b.src || a.op != b.op || a.mul != b.mul || a.l != b.l) return false;
          const double v = b.sign > 0 ? a.c : -a.c;
          const int q = 2 * b.grp + b.role;
          if (have[q] && std::memcmp(&val[q], &v, sizeof v) != 0) return false;
          val[q] = v;
          have[q] = 1;
        }
      }
      P->gstart[ca] = (int)P->coef.size();
      for (const auto& sl : slots) {
        const int q = 2 * sl.first + sl.second;
        if (!have[q]) return false;
        P->coef.push_back(val[q]);
      }
    }
  P->gstart[T.nc] = (int)P->coef.size();
  // the interior layout (the same calls at every interior l of the kernel's
  // layers: 0..l0, or 0..3 for the probe)
  const int l1 = M0 + 1;
  const int lk = KIND == kLmProbe ? (T.l0 < kLmProbeTop ? T.l0 : kLmProbeTop) : T.l0;
  if (l1 <= lk - 1) {
    P->istart = P->gstart[lm_layer_start(l1, M0)];
    P->istride = l1 + 1 <= lk - 1 ? P->gstart[lm_layer_start(l1 + 1, M0)] - P->istart : 0;
    for (int m = 0; m <= M0; ++m) P->ioff[m] = P->gstart[lm_layer_start(l1, M0) + lm_slot(m, 0)] - P->istart;
    for (int l = l1; l <= lk - 1; ++l)
      for (int m = 0; m <= M0; ++m)
        if (P->gstart[lm_layer_start(l, M0) + lm_slot(m, 0)] != P->istart + (l - l1) * P->istride + P->ioff[m])
          return false;
  }
  P->kind = KIND;
  return true;
}

template <int M0, bool TD>
bool lm_build_m(const HostTables& T, const std::vector<Term>& terms, const std::vector<int>& tstart,
                const std::vector<int>& clean, LmPlan* P) {
  return lm_build_kind<M0, kLmFull, TD>(T, terms, tstart, clean, P) ||
         lm_build_kind<M0, kLmAdv1, TD>(T, terms, tstart, clean, P) ||
         lm_build_kind<M0, kLmAdv1Es, TD>(T, terms, tstart, clean, P) ||
         lm_build_kind<M0, kLmAdv2, TD>(T, terms, tstart, clean, P) ||
         lm_build_kind<M0, kLmProbe, TD>(T, terms, tstart, clean, P);
}

// The kernel's plan of a program (terms, tstart, CLEAN_LOW_PC rows over all
// planes); P->kind = -1 when no kernel matches it
inline void lm_build(const HostTables& T, const std::vector<Term>& terms, const std::vector<int>& tstart,
                     const std::vector<int>& clean, LmPlan* P) {
  bool ok = false;
  switch (T.m0 * 2 + (T.two_d ? 1 : 0)) {
    case 2: ok = lm_build_m<1, false>(T, terms, tstart, clean, P); break;
    case 3: ok = lm_build_m<1, true>(T, terms, tstart, clean, P); break;
    case 4: ok = lm_build_m<2, false>(T, terms, tstart, clean, P); break;
    case 5: ok = lm_build_m<2, true>(T, terms, tstart, clean, P); break;
    case 6: ok = lm_build_m<3, false>(T, terms, tstart, clean, P); break;
    case 7: ok = lm_build_m<3, true>(T, terms, tstart, clean, P); break;
    case 8: ok = lm_build_m<4, false>(T, terms, tstart, clean, P); break;
    case 9: ok = lm_build_m<4, true>(T, terms, tstart, clean, P); break;
    default: break;
  }
  if (!ok) {
    P->kind = -1;
    P->gstart.clear();
    P->coef.clear();
  }
}

// ---------------------------------------------------------------------------
// Device

// Read-only tables read at wave-uniform addresses: the constant address
// space turns those reads into scalar loads
typedef const __attribute__((address_space(4))) double* LmCoefPtr;
typedef const __attribute__((address_space(4))) int* LmIntPtr;

// Three layers (0: l-1, 1: l, 2: l+1) of source values and the sums of the
// target being formed.  The layers live in this one object, indexed by
// compile-time constants only, so that they stay in registers.
template <typename T, int M0>
struct LmSums {
  static constexpr int NS = 2 * M0 + 1;
  T G[3][NS], H[3][NS], Dx[3][NS], Dy[3][NS], f[3][NS];
  LmCoefPtr cf;  // the target's coefficients
  int kc;
  double F[6];
  T v;
  T sa, sb;  // sums of the target's real and imaginary planes
  template <int J, int K, int OP>
  __device__ __forceinline__ T val() const {
    if constexpr (OP == OP_G || OP == OP_G00) return G[J][K];
    else if constexpr (OP == OP_H) return H[J][K];
    else if constexpr (OP == OP_DX) return Dx[J][K];
    else if constexpr (OP == OP_DY) return Dy[J][K];
    else return f[J][K];
  }
  template <int MUL>
  __device__ __forceinline__ T prod(double c) const {
    if constexpr (MUL == MUL_V) return v * static_cast<T>(c);
    else return static_cast<T>(F[MUL] * c);
  }
  template <int DL, int SM, int OP, int MR, int MI, bool TB>
  __device__ __forceinline__ void g() {
    constexpr int J = DL + 1;
    const T va = val<J, lm_slot(SM, 0), OP>();
    T vb = T(0);
    if constexpr (SM > 0) vb = val<J, lm_slot(SM, 1), OP>();
    T P = T(0), Q = T(0);
    if constexpr (MR != kLmNone) P = prod<MR>(cf[kc++]);
    if constexpr (MI != kLmNone && (TB || SM > 0)) Q = prod<MI>(cf[kc++]);
    if constexpr (MR != kLmNone) sa += va * P;
    if constexpr (MI != kLmNone && SM > 0) sa -= vb * Q;
    if constexpr (TB) {
      if constexpr (MI != kLmNone) sb += va * Q;
      if constexpr (MR != kLmNone && SM > 0) sb += vb * P;
    }
  }
};

template <typename T, int M0>
struct LmRaw {
  static constexpr int NS = 2 * M0 + 1;
  T fm[NS], f0[NS], fp[NS], xm[NS], xp[NS], ym[NS], yp[NS];
};

constexpr int kLmBlock = 256;

// KIND: the program (LmKind); TD: it has y advection.  a.coef, a.tstart and
// a.lm_* are an LmPlan's coef, gstart and interior layout; a.l0 is the top
// source layer (l0, or 3 for the probe) and a.lm_ltop the top layer written;
// a.mode is one of the modes the kind runs (launch_stage: 0..4 with y0 the
// stage input in mode 1 for the full program, 0, 5 with y0 the stage input
// and 6 for the others).  One thread per element of a plane in memory order
// (a workgroup writes kLmBlock consecutive elements of each plane: whole
// cache lines whatever the row length); the current partials of the stage
// input come from vlasov_current_tiles.
template <typename T, int M0, int KIND, bool TD>
__global__ __launch_bounds__(kLmBlock) void vlasov_stage_lm(StageArgs<T> a) {
  using Raw = LmRaw<T, M0>;
  constexpr int NS = 2 * M0 + 1;
  constexpr int MASK = lm_mask(KIND);
  constexpr bool RK4 = KIND == kLmFull;
  const int nx = a.nx, np = a.np, ny = a.ny;
  const int nxy = nx * ny;
  const int L = a.l0;
  const long long idx = (long long)blockIdx.x * kLmBlock + threadIdx.x;
  if (idx < a.plane) {
    const long long ystride = (long long)np * nx;
    const int y = static_cast<int>(idx / ystride);
    const int rem = static_cast<int>(idx - (long long)y * ystride);
    const int p = rem / nx;
    const int x = rem - p * nx;
    long long xm_off = (x == 0) ? (nx - 1) : -1;
    long long xp_off = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    long long ym_off = (y == 0) ? (long long)(ny - 1) * ystride : -ystride;
    long long yp_off = (y == ny - 1) ? -(long long)(ny - 1) * ystride : ystride;
    bool xlast = false, ylast = false;
    if (a.ext) ext_offsets(x, y, nx, ny, ystride, &xm_off, &xp_off, &ym_off, &yp_off, &xlast, &ylast);
    const int ixy = y * nx + x;
    LmSums<T, M0> e;
#pragma unroll
    for (int k = 0; k < 6; ++k) e.F[k] = a.fin[(long long)k * nxy + ixy];
    e.v = a.vtab[p];
    const T bgg = a.bgG ? a.bgG[p] : T(0);
    typedef const __attribute__((address_space(4))) T* TabPtr;
    const TabPtr hp0 = (TabPtr)a.hp0;
    const TabPtr cg = (TabPtr)a.cg;
    // in-bounds offsets of every stencil load (a clamped one is not used:
    // p - 1 at p = 0, p + 1 at p = np - 1, the y pair of a 1D program); the
    // "last" points' subtrahend is zeroed after the load
    const long long pm_off = p > 0 ? -(long long)nx : 0;
    const long long pp_off = p < np - 1 ? (long long)nx : 0;
    if (!TD) ym_off = yp_off = 0;
    const bool yzero = ylast || !TD;
    // Addresses: a wave-uniform plane base plus the lane's 32-bit byte offset
    // within the plane (the host keeps a plane below 4 GiB), so that every
    // load and store takes a scalar base and one offset register per stencil
    // point instead of a 64-bit address per plane and point.
    auto boff = [&](long long o) { return static_cast<unsigned>((idx + o) * (long long)sizeof(T)); };
    const unsigned o0 = boff(0), opm = boff(pm_off), opp = boff(pp_off), oxm = boff(xm_off), oxp = boff(xp_off),
                   oym = boff(ym_off), oyp = boff(yp_off);
    const size_t pbytes = (size_t)a.plane * sizeof(T);
    auto at = [](const void* b, unsigned o) -> const T& { return *reinterpret_cast<const T*>((const char*)b + o); };
    auto atw = [](void* b, unsigned o) -> T& { return *reinterpret_cast<T*>((char*)b + o); };
    // the stencil of layer l's planes (m <= min(l, M0)); IN: all 2 M0 + 1
    auto load = [&](auto inc, int l, Raw& r) {
      constexpr bool IN = decltype(inc)::value;
      const char* lb = (const char*)a.in + (size_t)lm_layer_start(l, M0) * pbytes;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        if (!IN && k > 2 * l) continue;
        const char* s = lb + (size_t)k * pbytes;
        r.f0[k] = at(s, o0);
        r.fm[k] = at(s, opm);
        r.fp[k] = at(s, opp);
        r.xm[k] = at(s, oxm);
        r.xp[k] = at(s, oxp);
        if (TD) {
          r.ym[k] = at(s, oym);
          r.yp[k] = at(s, oyp);
        } else {
          r.ym[k] = T(0);
          r.yp[k] = T(0);
        }
      }
    };
    // G, H (MakeGH and G00, formed as op_value forms them) and differences of
    // layer l into layer slot J; IN: l >= 1 and all slots
    auto form = [&](auto jc, auto inc, int l, const Raw& r) {
      constexpr int J = decltype(jc)::value;
      constexpr bool IN = decltype(inc)::value;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        if (!IN && k > 2 * l) continue;
        const T f0 = r.f0[k];
        e.Dx[J][k] = r.xm[k] - (xlast ? T(0) : r.xp[k]);
        e.Dy[J][k] = r.ym[k] - (yzero ? T(0) : r.yp[k]);
        // values first, then one unconditional store of each (stores under
        // branches would keep the layers out of registers)
        T gv, hv = T(0);
        if (p == 0) {
          const T f1 = r.fp[k];
          if (!IN && l == 0) {
            const T fe = (f0 - f1 * a.p0p1_sq) * a.inv_mp0p1_sq;
            gv = (f1 - fe) * a.g_r;
            if (a.bgG) gv += bgg;
          } else {
            gv = T(0);
            hv = f1 * hp0[l];
          }
        } else {
          const T fm = r.fm[k];
          const T dpf = (p == np - 1) ? (fm - f0) : (fm - r.fp[k]);
          if (!IN && l == 0) {
            gv = dpf;
            if (a.bgG) gv += bgg;
          } else {
            hv = f0 * a.invpax[l * np + p];
            hv = hv + dpf;
            gv = dpf * cg[l] + hv;
          }
        }
        e.G[J][k] = gv;
        e.H[J][k] = hv;
      }
    };
    {
      Raw r0;
      load(std::false_type{}, 0, r0);
      form(std::integral_constant<int, 1>{}, std::false_type{}, 0, r0);
    }
    const LmCoefPtr coef = (LmCoefPtr)a.coef;
    const LmIntPtr gstart = (LmIntPtr)a.tstart;
    // one RK mode, all layers
    auto run = [&](auto modec) {
      constexpr int MODE = decltype(modec)::value;
      // read with the next layer's stencil at the top of each l: the
      // accumulator, y0 (modes 2, 3; in modes 1 and 5 it is the stage input)
      // and the layer's own values (the B rotation, and y0 = the input)
      constexpr bool NEED_ACC = MODE == 2 || MODE == 3 || MODE == 4 || MODE == 6;
      constexpr bool NEED_Y0 = MODE == 2 || MODE == 3;
      constexpr bool NEED_F = (MASK & LM_B) != 0 || MODE == 1 || MODE == 5;
      // layer l; IN: l is interior (every neighbour exists, every slot too)
      auto body = [&](auto inc, int l) {
        constexpr bool IN = decltype(inc)::value;
        const int c0 = lm_layer_start(l, M0);
        const size_t lo = (size_t)c0 * pbytes;
        T opa[NS], opy[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          if (!IN && k > 2 * l) continue;
          if constexpr (NEED_ACC) opa[k] = at((const char*)a.acc + lo + k * pbytes, o0);
          if constexpr (NEED_Y0) opy[k] = at((const char*)a.y0 + lo + k * pbytes, o0);
          if constexpr (NEED_F) e.f[1][k] = at((const char*)a.in + lo + k * pbytes, o0);
        }
        if (IN || l + 1 <= L) {  // layer l+1, which the sums of layer l read
          Raw rn;
          load(inc, l + 1, rn);
          form(std::integral_constant<int, 2>{}, inc, l + 1, rn);
        }
        // the sums of layer l's targets
        T sum[NS];
        const int ib = a.lm_istart + (l - M0 - 1) * a.lm_istride;
        auto target = [&](auto tmc) {
          constexpr int TM = decltype(tmc)::value;
          if (!IN && TM > l) return;
          e.cf = IN ? coef + ib + a.lm_ioff[TM] : coef + gstart[c0 + lm_slot(TM, 0)];
          e.kc = 0;
          e.sa = T(0);
          e.sb = T(0);
          lm_groups<M0, TM, KIND, TD, IN>(e, l, L);
          sum[lm_slot(TM, 0)] = e.sa;
          if constexpr (TM > 0) sum[lm_slot(TM, 1)] = e.sb;
        };
        target(std::integral_constant<int, 0>{});
        if constexpr (M0 >= 1) target(std::integral_constant<int, 1>{});
        if constexpr (M0 >= 2) target(std::integral_constant<int, 2>{});
        if constexpr (M0 >= 3) target(std::integral_constant<int, 3>{});
        if constexpr (M0 >= 4) target(std::integral_constant<int, 4>{});
        // RK update of layer l's planes, def-RK.cpp:116-138 and :478-488, as vlasov_stage
        const int rows = lm_clean_rows(KIND, l);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          if (!IN && k > 2 * l) continue;
          T s = sum[k];
          if (p < rows) s = T(0);  // CLEAN_LOW_PC, def-Actions.cpp:240-253
          T& out = atw((char*)a.out + lo + k * pbytes, o0);
          T& acc = atw((char*)a.acc + lo + k * pbytes, o0);
          if constexpr (MODE == 0) {
            out = s;
          } else if constexpr (MODE == 1) {
            const T kh = s * a.c05h;
            out = e.f[1][k] + kh;
            acc = e.f[1][k] + kh * a.third;
          } else if constexpr (MODE == 2) {
            const T kh = s * a.c05h;
            out = opy[k] + kh;
            acc = opa[k] + kh * a.twothird;
          } else if constexpr (MODE == 3) {
            const T kh = s * a.hh;
            out = opy[k] + kh;
            acc = opa[k] + kh * a.third;
          } else if constexpr (MODE == 5) {
            const T kh = s * a.hh;
            out = e.f[1][k] + kh;
            acc = e.f[1][k] + kh * T(0.5);
          } else if constexpr (MODE == 6) {
            const T kh = s * a.c05h;
            out = opa[k] + kh;
          } else {
            const T kh = s * a.h6;
            out = opa[k] + kh;
          }
        }
        // l-1 <- l, l <- l+1
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          e.G[0][k] = e.G[1][k];
          e.Dx[0][k] = e.Dx[1][k];
          e.Dy[0][k] = e.Dy[1][k];
          e.G[1][k] = e.G[2][k];
          e.Dx[1][k] = e.Dx[2][k];
          e.Dy[1][k] = e.Dy[2][k];
        }
      };
      // layers 0..m0 and L at the edges, m0+1..L-1 interior (up to lm_ltop)
      const int lt = a.lm_ltop;
      const int le = lt < M0 ? lt : M0;
      const int li = L - 1 < lt ? L - 1 : lt;
      for (int l = 0; l <= le; ++l) body(std::false_type{}, l);
      for (int l = M0 + 1; l <= li; ++l) body(std::true_type{}, l);
      if (L >= M0 + 1 && L <= lt) body(std::false_type{}, L);
    };
    if constexpr (RK4) {
      switch (a.mode) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 3: run(std::integral_constant<int, 3>{}); break;
        default: run(std::integral_constant<int, 4>{}); break;
      }
    } else {
      switch (a.mode) {
        case 0: run(std::integral_constant<int, 0>{}); break;
        case 5: run(std::integral_constant<int, 5>{}); break;
        default: run(std::integral_constant<int, 6>{}); break;
      }
    }
  }
}

// The per-tile current partials of Current::Jxyz from the stage input, as
// vlasov_stage forms them alongside its slopes (same tiles, same sums)
template <typename T>
__global__ __launch_bounds__(kTileX* kTileP) void vlasov_current_tiles(StageArgs<T> a) {
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = blockIdx.x * kTileX + tx;
  const int p = blockIdx.y * kTileP + ty;
  const int y = blockIdx.z;
  const int nx = a.nx, np = a.np;
  const int nxy = nx * a.ny;
  __shared__ double jsh[3][kTileP][kTileX];
  double j10 = 0.0, j11r = 0.0, j11i = 0.0;
  if (x < nx && p < np) {
    const long long idx = ((long long)y * np + p) * nx + x;
    // trapezoid weights of Current::Jxyz: 1/2 at both ends of the p axis
    const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
    const double g = a.p3og[p];
    if (a.c10 >= 0) j10 = w * (static_cast<double>(a.in[(long long)a.c10 * a.plane + idx]) * g);
    if (a.c11r >= 0) {
      j11r = w * (static_cast<double>(a.in[(long long)a.c11r * a.plane + idx]) * g);
      j11i = w * (static_cast<double>(a.in[(long long)a.c11i * a.plane + idx]) * g);
    }
  }
  jsh[0][ty][tx] = j10;
  jsh[1][ty][tx] = j11r;
  jsh[2][ty][tx] = j11i;
  __syncthreads();
  if (ty < 3 && x < nx) {
    double s = jsh[ty][0][tx];
#pragma unroll
    for (int k = 1; k < kTileP; ++k) s += jsh[ty][k][tx];
    a.jpart[((long long)blockIdx.y * 3 + ty) * nxy + y * nx + x] = s;
  }
}

}  // namespace oshun
