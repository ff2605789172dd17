// Compile-time specialised stage kernel of the general system (m0 > 0 and/or
// 2D): one RK4 stage of Actions::Adv (def-Actions.cpp:73-94) for all planes
// of one (x, p, y) point, with L0 and M0 template constants.
//
// The interpreter kernel vlasov_stage walks the host's per-target term list
// and recomputes the source operator (G, H, Dx, Dy, identity) of every term
// from memory: at l0 = m0 = 4 that is 298 terms and about a thousand loads
// per point.  Here the terms are visited source-major, in the order in which
// build_tables emits them (E field by source (m outer, l inner), then x
// advection, y advection and the B rotation, def-VlasovMax.cpp:346-644 and
// :757-951): each source harmonic's stencil is loaded once, its operator
// values are formed once, and every term adds into a per-target accumulator
// held in registers.  Every target still receives its terms in the host
// program's order, so the sums are those of the interpreter.  The host checks
// the visit sequence against the term list at set-up (spec_coefficients) and
// takes the coefficients from it in visit order; a mismatch falls back to the
// interpreter.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <utility>
#include <vector>

#include "kernels.hpp"

namespace oshun {

enum SpecPhase { PH_E = 0, PH_X = 1, PH_Y = 2, PH_B = 3 };

template <int L0, int M0>
struct SpecShape {
  // plane of the real part of (l, m) in the device layout: harmonics in ind(l,m)
  // order (def-State.cpp:422-426), each m > 0 harmonic followed by its imaginary plane
  __host__ __device__ static constexpr int re(int l, int m) {
    int c = 0;
    for (int j = 0; j < l; ++j) c += 1 + 2 * (j < M0 ? j : M0);
    return c + (m == 0 ? 0 : 2 * m - 1);
  }
  __host__ __device__ static constexpr int im(int l, int m) { return re(l, m) + 1; }
  __host__ __device__ static constexpr bool valid(int l, int m) {
    return l >= 0 && l <= L0 && m >= 0 && m <= M0 && m <= l;
  }
  static constexpr int NC = re(L0 + 1, 0);
};

// target (tl,tm) += OP(source (sl,sm)) * (Mr + i Mi), Mr and Mi one field (or
// v(p)) each, -1 when absent; the same pushes as build_tables' add_complex
template <int L0, int M0, class S>
__host__ __device__ constexpr __forceinline__ void spec_cplx(S& s, int tl, int tm, int sm, int op, int mr, int mi,
                                                   bool re_only) {
  using Sh = SpecShape<L0, M0>;
  const int ta = Sh::re(tl, tm);
  const int tb = (tm > 0 && !re_only) ? Sh::im(tl, tm) : -1;
  const bool sb = sm > 0;
  if (mr >= 0) s.emit(ta, 0, op, mr);
  if (sb && mi >= 0) s.emit(ta, 1, op, mi);
  if (tb >= 0) {
    if (mi >= 0) s.emit(tb, 0, op, mi);
    if (sb && mr >= 0) s.emit(tb, 1, op, mr);
  }
}

// E-field term of source (sl, sm): which 0 = Ex, 1 = Em = Ey - i Ez, 2 = Ep = Ey + i Ez
// (def-VlasovMax.cpp:352-353; the sign of the Ez part is in the coefficient)
template <int L0, int M0, class S>
__host__ __device__ constexpr __forceinline__ void spec_e(S& s, int tl, int tm, int sm, int op, int which, bool re_only) {
  if (!SpecShape<L0, M0>::valid(tl, tm)) return;
  if (which == 0) spec_cplx<L0, M0>(s, tl, tm, sm, op, MUL_EX, -1, re_only);
  else spec_cplx<L0, M0>(s, tl, tm, sm, op, MUL_EY, MUL_EZ, re_only);
}

template <int L0, int M0, class S>
__host__ __device__ constexpr __forceinline__ void spec_a(S& s, int tl, int tm, int sm, int op, bool re_only) {
  if (!SpecShape<L0, M0>::valid(tl, tm)) return;
  spec_cplx<L0, M0>(s, tl, tm, sm, op, MUL_V, -1, re_only);
}

// Source harmonics in visit order, m outer and l inner: index j -> (l, m)
template <int L0, int M0>
__host__ __device__ constexpr int src_m(int j) {
  for (int m = 0; m <= M0; ++m) {
    if (j <= L0 - m) return m;
    j -= L0 - m + 1;
  }
  return -1;
}
template <int L0, int M0>
__host__ __device__ constexpr int src_l(int j) {
  for (int m = 0; m <= M0; ++m) {
    if (j <= L0 - m) return m + j;
    j -= L0 - m + 1;
  }
  return -1;
}
template <int L0, int M0>
__host__ __device__ constexpr int num_sources() {
  int n = 0;
  for (int m = 0; m <= M0; ++m) n += L0 - m + 1;
  return n;
}

// The terms of one source harmonic in one phase (l and m compile-time)
template <int L0, int M0, int PH, int J, class S>
__host__ __device__ constexpr __forceinline__ void spec_source(S& s) {
  constexpr int m = src_m<L0, M0>(J), l = src_l<L0, M0>(J);
  if constexpr (PH == PH_E) {  // ElectricField::Exyz, def-VlasovMax.cpp:346-491
    s.template source<PH_E, J>(l, m);
    if constexpr (l == 0) {  // G00 feeds (1,0) and (1,1)
      if (L0 >= 1) {
        spec_e<L0, M0>(s, 1, 0, 0, OP_G00, 0, false);
        if (M0 >= 1) spec_e<L0, M0>(s, 1, 1, 0, OP_G00, 1, false);
      }
    } else {
      spec_e<L0, M0>(s, l - 1, m, m, OP_H, 0, false);
      if (m + 1 <= M0) spec_e<L0, M0>(s, l - 1, m + 1, m, OP_H, 1, false);
      if (m == 1) spec_e<L0, M0>(s, l - 1, 0, m, OP_H, 2, true);
      else if (m >= 2) spec_e<L0, M0>(s, l - 1, m - 1, m, OP_H, 2, false);
      if (l + 1 <= L0) {
        spec_e<L0, M0>(s, l + 1, m, m, OP_G, 0, false);
        if (m + 1 <= M0) spec_e<L0, M0>(s, l + 1, m + 1, m, OP_G, 1, false);
        if (m == 1) spec_e<L0, M0>(s, l + 1, 0, m, OP_G, 2, true);
        else if (m >= 2) spec_e<L0, M0>(s, l + 1, m - 1, m, OP_G, 2, false);
      }
    }
  } else if constexpr (PH == PH_X) {  // Spatial_Advection::Axy, x part, def-VlasovMax.cpp:842-853
    s.template source<PH_X, J>(l, m);
    spec_a<L0, M0>(s, l - 1, m, m, OP_DX, false);
    if (l + 1 <= L0) spec_a<L0, M0>(s, l + 1, m, m, OP_DX, false);
  } else if constexpr (PH == PH_Y) {  // y part, def-VlasovMax.cpp:857-947
    s.template source<PH_Y, J>(l, m);
    if (m + 1 <= M0) {
      spec_a<L0, M0>(s, l - 1, m + 1, m, OP_DY, false);
      if (l + 1 <= L0) spec_a<L0, M0>(s, l + 1, m + 1, m, OP_DY, false);
    }
    if (m == 1) {
      spec_a<L0, M0>(s, l - 1, 0, m, OP_DY, true);
      if (l + 1 <= L0) spec_a<L0, M0>(s, l + 1, 0, m, OP_DY, true);
    } else if (m >= 2) {
      spec_a<L0, M0>(s, l - 1, m - 1, m, OP_DY, false);
      if (l + 1 <= L0) spec_a<L0, M0>(s, l + 1, m - 1, m, OP_DY, false);
    }
  } else {  // MagneticField::Bxyz, def-VlasovMax.cpp:597-644 (no terms from l = 0)
    s.template source<PH_B, J>(l, m);
    if (l >= 1 && m + 1 <= M0 && m + 1 <= l) spec_cplx<L0, M0>(s, l, m + 1, m, OP_ID, MUL_BZ, MUL_BY, false);
    if (l >= 1 && m >= 1) {
      spec_cplx<L0, M0>(s, l, m, m, OP_ID, -1, MUL_BX, false);
      if (m == 1) spec_cplx<L0, M0>(s, l, 0, m, OP_ID, MUL_BZ, MUL_BY, true);
      else spec_cplx<L0, M0>(s, l, m - 1, m, OP_ID, MUL_BZ, MUL_BY, false);
    }
  }
}

template <int L0, int M0, int PH, class S, int... J>
__host__ __device__ constexpr __forceinline__ void spec_phase(S& s, std::integer_sequence<int, J...>) {
  (spec_source<L0, M0, PH, J>(s), ...);
}

// Every term of Actions::Adv in emission order, as straight-line code (no
// loops, so every target index is a constant).  S provides
//   source<phase, j>(l, m)            before the terms of source harmonic j = (l, m)
//   emit(target, src_im, op, mul)     one term (src_im: the source's imaginary plane)
template <int L0, int M0, bool TWO_D, class S>
__host__ __device__ constexpr __forceinline__ void spec_terms(S& s) {
  using Seq = std::make_integer_sequence<int, num_sources<L0, M0>()>;
  spec_phase<L0, M0, PH_E>(s, Seq{});
  spec_phase<L0, M0, PH_X>(s, Seq{});
  if constexpr (TWO_D && M0 >= 1) spec_phase<L0, M0, PH_Y>(s, Seq{});
  if constexpr (M0 >= 1) spec_phase<L0, M0, PH_B>(s, Seq{});
}

struct SpecCount {
  int n = 0;
  template <int PH, int J>
  __host__ __device__ constexpr void source(int, int) {}
  __host__ __device__ constexpr void emit(int, int, int, int) { ++n; }
};
template <int L0, int M0, bool TWO_D>
__host__ __device__ constexpr int spec_nterms() {
  SpecCount c;
  spec_terms<L0, M0, TWO_D>(c);
  return c.n;
}

// ---------------------------------------------------------------------------
// Host: the visit sequence against the term list, and the coefficients in
// visit order.

template <int L0, int M0>
struct SpecRecord {
  int cl = 0, cm = 0;
  std::vector<std::array<int, 5>> e;  // target, source plane, op, mul, source l
  template <int PH, int J>
  void source(int l, int m) { cl = l; cm = m; }
  void emit(int t, int sim, int op, int mul) { e.push_back({t, SpecShape<L0, M0>::re(cl, cm) + sim, op, mul, cl}); }
};

template <int L0, int M0, bool TWO_D>
bool spec_coefficients(const HostTables& T, std::vector<double>& coef) {
  if (T.l0 != L0 || T.m0 != M0 || T.two_d != TWO_D || T.nc != SpecShape<L0, M0>::NC) return false;
  SpecRecord<L0, M0> r;
  spec_terms<L0, M0, TWO_D>(r);
  std::vector<int> pos(T.nc, 0);
  coef.clear();
  for (const auto& q : r.e) {
    const int t = q[0];
    if (t < 0 || t >= T.nc) return false;
    const int i = T.tstart[t] + pos[t]++;
    if (i >= T.tstart[t + 1]) return false;
    const Term& tm = T.terms[i];
    if (tm.src != q[1] || tm.op != q[2] || tm.mul != q[3]) return false;
    if ((tm.op == OP_G || tm.op == OP_H || tm.op == OP_G00) && tm.l != q[4]) return false;
    coef.push_back(tm.c);
  }
  for (int t = 0; t < T.nc; ++t)
    if (pos[t] != T.tstart[t + 1] - T.tstart[t]) return false;
  return true;
}

// ---------------------------------------------------------------------------
// Device.  A wavefront is 64 x at one (p, y).  The work is a fixed sequence of
// steps: every source harmonic of every phase, then the RK4 combination of
// the planes in chunks of kSpecChunk.  The loads of step g + D - 1 are issued
// at step g (a ring of D register slots), so each wavefront has D steps of
// loads in flight while it adds the terms of the current one.

template <int L0, int M0>
__host__ __device__ constexpr int plane_l(int c) {
  for (int l = 0; l <= L0; ++l)
    for (int m = 0; m <= (l < M0 ? l : M0); ++m) {
      const int n = m == 0 ? 1 : 2;
      if (c < n) return l;
      c -= n;
    }
  return -1;
}

constexpr int kSpecDepth = 4;  // steps of loads in flight per wavefront
constexpr int kSpecChunk = 3;  // planes per RK4 step (y0 and acc of each: 6 values, the size of a slot)

template <int L0, int M0, bool TWO_D>
struct SpecSeq {
  static constexpr int NS = num_sources<L0, M0>();
  static constexpr int NPH = 2 + ((TWO_D && M0 >= 1) ? 1 : 0) + (M0 >= 1 ? 1 : 0);
  static constexpr int NSRC = NPH * NS;
  static constexpr int NC = SpecShape<L0, M0>::NC;
  static constexpr int NCH = (NC + kSpecChunk - 1) / kSpecChunk;
  static constexpr int NG = NSRC + NCH;
  // k-th phase present: E, X, then Y (2D with m0 >= 1), then B (m0 >= 1)
  __host__ __device__ static constexpr int phase(int k) {
    return k == 0 ? PH_E : k == 1 ? PH_X : (k == 2 && TWO_D && M0 >= 1) ? PH_Y : PH_B;
  }
  __host__ __device__ static constexpr int base(int ph) {
    for (int k = 0; k < NPH; ++k)
      if (phase(k) == ph) return k * NS;
    return -1;
  }
};

// the LDS row cache of a workgroup (4 wavefronts x every plane x 64 lanes),
// used where it leaves room for three workgroups per CU
template <typename T, int L0, int M0>
__host__ __device__ constexpr bool spec_lds() {
  return SpecShape<L0, M0>::NC * kTileX * kTileP * sizeof(T) <= 52 * 1024;
}

template <typename T, int L0, int M0, bool TWO_D, int D, bool LDS = spec_lds<T, L0, M0>()>
struct SpecEval {
  using Sq = SpecSeq<L0, M0, TWO_D>;
  using Sh = SpecShape<L0, M0>;
  static constexpr int NC = Sh::NC;
  const StageArgs<T>& a;
  // byte offsets within a plane of (x,p,y) and of its stencil neighbours;
  // with a wave-uniform plane base the loads take the SGPR-base form
  unsigned o0, oxm, oxp, oym, oyp, opm, opn;
  int p;
  double F[6];
  T v, bgg;
  T acc[NC];
  T raw[D][6];  // loads in flight, one slot per step
  // LDS row cache (when it fits, see spec_lds): the E field stores f(p) of
  // every plane in the wavefront's own region, the x advection and the B
  // rotation read it back (no re-fetch of rows the L2 has already evicted);
  // each wavefront reads only what it wrote, so no barrier is needed
  T* lrow;     // this lane's slot of plane 0
  bool xl, xr;  // the x-1 / x+1 neighbour is in the cached row (else from memory)
  T val[2][2];  // [re | im][G, Dx, Dy or f | H] of the current source
  const double* coef;  // the coefficients in LDS
  int k;               // next coefficient

  static __device__ __forceinline__ T at(const T* base, unsigned off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
  }
  static __device__ __forceinline__ void put(T* base, unsigned off, T v) {
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
  }

  __device__ __forceinline__ SpecEval(const StageArgs<T>& args, const double* c) : a(args), coef(c), k(0) {
#pragma unroll
    for (int q = 0; q < NC; ++q) acc[q] = T(0);
  }

  // loads of step G into slot G % D
  template <int G>
  __device__ __forceinline__ void issue() {
    if constexpr (G < Sq::NG) {
      T* r = raw[G % D];
      if constexpr (G < Sq::NSRC) {
        constexpr int ph = Sq::phase(G / Sq::NS), J = G % Sq::NS;
        constexpr int l = src_l<L0, M0>(J), m = src_m<L0, M0>(J);
        if constexpr (!(ph == PH_B && l == 0)) {
#pragma unroll
          for (int s = 0; s < (m > 0 ? 2 : 1); ++s) {
            const T* __restrict__ f = a.in + (long long)(Sh::re(l, m) + s) * a.plane;
            const int c = Sh::re(l, m) + s;
            if constexpr (ph == PH_E) {
              r[3 * s] = at(f, opm);
              r[3 * s + 1] = at(f, o0);
              r[3 * s + 2] = at(f, opn);
            } else if constexpr (ph == PH_X) {
              if constexpr (LDS) {
                r[3 * s] = xl ? lrow[c * kTileX - 1] : at(f, oxm);
                r[3 * s + 2] = xr ? lrow[c * kTileX + 1] : at(f, oxp);
              } else {
                r[3 * s] = at(f, oxm);
                r[3 * s + 2] = at(f, oxp);
              }
            } else if constexpr (ph == PH_Y) {
              r[3 * s] = at(f, oym);
              r[3 * s + 2] = at(f, oyp);
            } else {
              if constexpr (LDS) r[3 * s + 1] = lrow[c * kTileX];
              else r[3 * s + 1] = at(f, o0);
            }
          }
        }
      } else {
        constexpr int c0 = (G - Sq::NSRC) * kSpecChunk;
#pragma unroll
        for (int q = 0; q < kSpecChunk; ++q) {
          if (c0 + q < NC) {
            const long long pl = (long long)(c0 + q) * a.plane;
            if (a.mode >= 1 && a.mode <= 3) r[2 * q] = at(a.y0 + pl, o0);
            if (a.mode >= 2) r[2 * q + 1] = at(a.acc + pl, o0);
          }
        }
      }
    }
  }
  template <int... I>
  __device__ __forceinline__ void prologue(std::integer_sequence<int, I...>) {
    (issue<I>(), ...);
  }

  // step G: fence, the loads of step G + D - 1, then wait for this step's own
  template <int G>
  __device__ __forceinline__ void advance() {
    // the fence keeps the loads in step order (the compiler would otherwise
    // issue all of them up front); the accumulators are pinned as they stand,
    // or every sum is sunk to the stores at the end with all its operands live
    asm volatile("" ::: "memory");
#pragma unroll
    for (int q = 0; q < NC; ++q) asm volatile("" : "+v"(acc[q]));
    if constexpr (G == 0) prologue(std::make_integer_sequence<int, D - 1>{});
    issue<G + D - 1>();
  }

  template <int PH, int J>
  __device__ __forceinline__ void source(int l, int m) {
    constexpr int G = Sq::base(PH) + J;
    advance<G>();
    const T* r = raw[G % D];
#pragma unroll
    for (int s = 0; s < (m > 0 ? 2 : 1); ++s) {
      if (PH == PH_X || PH == PH_Y) {
        val[s][0] = r[3 * s] - r[3 * s + 2];
      } else if (PH == PH_B) {
        val[s][0] = r[3 * s + 1];
      } else {  // G and H of MakeGH / G00, def-VlasovMax.cpp:496-535, as op_value
        const T fm = r[3 * s], f0 = r[3 * s + 1], fn = r[3 * s + 2];
        if constexpr (LDS) lrow[(Sh::re(l, m) + s) * kTileX] = f0;
        if (p == 0) {
          if (l == 0) {
            const T fe = (f0 - fn * a.p0p1_sq) * a.inv_mp0p1_sq;
            val[s][0] = (fn - fe) * a.g_r;
            if (a.bgG) val[s][0] += bgg;
          } else {
            val[s][0] = T(0);
            val[s][1] = fn * ldc(a.hp0, l);
          }
        } else {
          // SHarmonic::Dp, def-State.cpp:359-380: f(N-2) - f(N-1) at pmax
          const T dpf = (p == a.np - 1) ? (fm - f0) : (fm - fn);
          if (l == 0) {
            val[s][0] = dpf;
            if (a.bgG) val[s][0] += bgg;
          } else {
            T hv = f0 * ldc(a.invpax, l * a.np + p);
            hv = hv + dpf;
            val[s][1] = hv;
            val[s][0] = dpf * ldc(a.cg, l) + hv;
          }
        }
      }
    }
  }

  __device__ __forceinline__ void emit(int t, int sim, int op, int mul) {
    const T x = (op == OP_H) ? val[sim][1] : val[sim][0];
    const double c = coef[k];
    ++k;
    T mult;
    if (mul == MUL_V) mult = v * static_cast<T>(c);
    else mult = static_cast<T>(F[mul] * c);
    acc[t] += x * mult;
  }

  // CLEAN_LOW_PC (def-Actions.cpp:240-253) and the RK4 combination
  // (def-RK.cpp:116-138) of the planes of chunk CH
  template <int CH>
  __device__ __forceinline__ void rk_chunk() {
    constexpr int G = Sq::NSRC + CH;
    advance<G>();
    const T* r = raw[G % D];
#pragma unroll
    for (int q = 0; q < kSpecChunk; ++q) {
      const int c = CH * kSpecChunk + q;
      if (c < NC) {
        const int l = plane_l<L0, M0>(c);
        T sl = acc[c < NC ? c : 0];
        if (l >= 1 && p < (l - 1 < 80 ? l - 1 : 80)) sl = T(0);
        const long long pl = (long long)c * a.plane;
        T* out = a.out + pl;
        T* acw = a.acc + pl;
        const T y0 = r[2 * q], ac = r[2 * q + 1];
        switch (a.mode) {
          case 0:
            put(out, o0, sl);
            break;
          case 1: {
            const T kh = sl * a.c05h;
            put(out, o0, y0 + kh);
            put(acw, o0, y0 + kh * a.third);
          } break;
          case 2: {
            const T kh = sl * a.c05h;
            put(out, o0, y0 + kh);
            put(acw, o0, ac + kh * a.twothird);
          } break;
          case 3: {
            const T kh = sl * a.hh;
            put(out, o0, y0 + kh);
            put(acw, o0, ac + kh * a.third);
          } break;
          default: {
            const T kh = sl * a.h6;
            put(out, o0, ac + kh);
          } break;
        }
      }
    }
  }
  template <int... CH>
  __device__ __forceinline__ void rk_all(std::integer_sequence<int, CH...>) {
    (rk_chunk<CH>(), ...);
  }
};

template <typename T, int L0, int M0, bool TWO_D, int D>
__global__ __launch_bounds__(kTileX* kTileP) void vlasov_stage_spec(StageArgs<T> a) {
  using Sh = SpecShape<L0, M0>;
  using Sq = SpecSeq<L0, M0, TWO_D>;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = blockIdx.x * kTileX + tx;
  const int p = __builtin_amdgcn_readfirstlane(blockIdx.y * kTileP + ty);  // one p per wavefront
  const int y = blockIdx.z;
  const int nx = a.nx, np = a.np, ny = a.ny;
  const int nxy = nx * ny;
  // the coefficients, read by every lane at the same address (LDS broadcast):
  // as scalar loads the compiler hoists all of them and spills SGPRs
  constexpr int NT = spec_nterms<L0, M0, TWO_D>();
  __shared__ double csh[NT];
  constexpr bool LDS = spec_lds<T, L0, M0>();
  constexpr int kJsh = 3 * kTileP * kTileX * sizeof(double) / sizeof(T);
  constexpr int kRows = LDS ? Sh::NC * kTileX * kTileP : 1;
  // the row cache, and after the barrier below the current partial sums
  __shared__ __align__(16) T lds_rows[kRows > kJsh ? kRows : kJsh];
  double(*jsh)[kTileP][kTileX] = reinterpret_cast<double(*)[kTileP][kTileX]>(lds_rows);
  for (int i = ty * kTileX + tx; i < NT; i += kTileX * kTileP) csh[i] = a.coef[i];
  __syncthreads();
  double j10 = 0.0, j11r = 0.0, j11i = 0.0;
  if (x < nx && p < np) {
    SpecEval<T, L0, M0, TWO_D, D> e(a, csh);
    e.p = p;
    // element offsets within a plane (the host requires plane * 8 bytes < 4 GiB)
    const unsigned idx = ((unsigned)y * np + p) * nx + x;
    const unsigned ystride = (unsigned)np * nx;
    const unsigned es = sizeof(T);
    e.o0 = idx * es;
    e.oxm = (x == 0 ? idx + (nx - 1) : idx - 1) * es;
    e.oxp = (x == nx - 1 ? idx - (nx - 1) : idx + 1) * es;
    e.oym = (y == 0 ? idx + (ny - 1) * ystride : idx - ystride) * es;
    e.oyp = (y == ny - 1 ? idx - (ny - 1) * ystride : idx + ystride) * es;
    e.opm = (p > 0 ? idx - nx : idx) * es;        // p-1 and p+1 rows, clamped at the ends where
    e.opn = (p < np - 1 ? idx + nx : idx) * es;   // MakeGH / G00 do not use them
    e.lrow = lds_rows + (LDS ? ty * Sh::NC * kTileX + tx : 0);
    e.xl = tx > 0;                                // lane x-1 of this row (x = 0 wraps: memory)
    e.xr = tx < kTileX - 1 && x < nx - 1;         // lane x+1 (tile edge and x = nx-1: memory)
    const int ixy = y * nx + x;
#pragma unroll
    for (int c = 0; c < 6; ++c) e.F[c] = a.fin[(long long)c * nxy + ixy];
    e.v = ldc(a.vtab, p);
    e.bgg = a.bgG ? ldc(a.bgG, p) : T(0);
    spec_terms<L0, M0, TWO_D>(e);
    e.rk_all(std::make_integer_sequence<int, Sq::NCH>{});
    // trapezoid weights of Current::Jxyz: 1/2 at both ends of the p axis
    const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
    const double g = a.p3og[p];
    // f10, f11 at the point: from the row cache where there is one
    auto fpt = [&](int c) {
      if constexpr (LDS) return e.lrow[c * kTileX];
      else return a.in[(long long)c * a.plane + idx];
    };
    j10 = w * (static_cast<double>(fpt(Sh::re(1, 0))) * g);
    if (M0 >= 1) {
      j11r = w * (static_cast<double>(fpt(Sh::re(1, 1))) * g);
      j11i = w * (static_cast<double>(fpt(Sh::im(1, 1))) * g);
    }
  }
  __syncthreads();  // every wavefront is done with its cached rows
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

// (l0, m0, 2D) shapes with a specialised stage kernel; others use vlasov_stage
#define OSHUN_STAGE_SPECS(X) \
  X(0, 4, 4, true)           \
  X(1, 4, 3, true)           \
  X(2, 6, 3, true)           \
  X(3, 7, 3, true)           \
  X(4, 8, 2, false)          \
  X(5, 3, 2, false)

}  // namespace oshun
