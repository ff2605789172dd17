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
    s.source(PH_E, l, m);
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
    s.source(PH_X, l, m);
    spec_a<L0, M0>(s, l - 1, m, m, OP_DX, false);
    if (l + 1 <= L0) spec_a<L0, M0>(s, l + 1, m, m, OP_DX, false);
  } else if constexpr (PH == PH_Y) {  // y part, def-VlasovMax.cpp:857-947
    s.source(PH_Y, l, m);
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
  } else if constexpr (l >= 1) {  // MagneticField::Bxyz, def-VlasovMax.cpp:597-644
    s.source(PH_B, l, m);
    if (m + 1 <= M0 && m + 1 <= l) spec_cplx<L0, M0>(s, l, m + 1, m, OP_ID, MUL_BZ, MUL_BY, false);
    if (m >= 1) {
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
//   source(phase, l, m)               before the terms of source harmonic (l, m)
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
  __host__ __device__ constexpr void source(int, int, int) {}
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
  void source(int, int l, int m) { cl = l; cm = m; }
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
// Device: operator values of the current source harmonic, accumulators of
// every target plane.

template <typename T, int L0, int M0>
struct SpecEval {
  static constexpr int NC = SpecShape<L0, M0>::NC;
  const StageArgs<T>& a;
  long long idx, xm, xp, ym, yp;
  int p;
  double F[6];
  T v, bgg;
  T acc[NC];
  T val[2][2];  // [re | im][G, Dx, Dy or f | H]
  const double* coef;  // the coefficients in LDS
  int k;               // next coefficient

  __device__ __forceinline__ SpecEval(const StageArgs<T>& args, const double* c) : a(args), coef(c), k(0) {
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = T(0);
  }

  __device__ __forceinline__ void source(int ph, int l, int m) {
    // loads of one source harmonic at a time: without this fence the compiler
    // issues every load of the kernel up front (and keeps f(p) of every
    // plane from the E field for the B rotation), which needs more registers
    // than a SIMD has for two wavefronts
    asm volatile("" ::: "memory");
    // and the accumulators as they stand: otherwise every sum is sunk to the
    // stores at the end, keeping all operator values and coefficients live
#pragma unroll
    for (int c = 0; c < NC; ++c) asm volatile("" : "+v"(acc[c]));
    plane_ops(ph, l, SpecShape<L0, M0>::re(l, m), 0);
    if (m > 0) plane_ops(ph, l, SpecShape<L0, M0>::im(l, m), 1);
  }

  __device__ __forceinline__ void plane_ops(int ph, int l, int c, int s) {
    {
      const T* __restrict__ f = a.in + (long long)c * a.plane;
      if (ph == PH_X) {
        val[s][0] = f[idx + xm] - f[idx + xp];
      } else if (ph == PH_Y) {
        val[s][0] = f[idx + ym] - f[idx + yp];
      } else if (ph == PH_B) {
        val[s][0] = f[idx];
      } else {  // G and H of MakeGH / G00, def-VlasovMax.cpp:496-535, as op_value
        const T f0 = f[idx];
        if (p == 0) {
          const T f1 = f[idx + a.nx];
          if (l == 0) {
            const T fe = (f0 - f1 * a.p0p1_sq) * a.inv_mp0p1_sq;
            val[s][0] = (f1 - fe) * a.g_r;
            if (a.bgG) val[s][0] += bgg;
          } else {
            val[s][0] = T(0);
            val[s][1] = f1 * ldc(a.hp0, l);
          }
        } else {
          const T fm = f[idx - a.nx];
          const T dpf = (p == a.np - 1) ? (fm - f0) : (fm - f[idx + a.nx]);
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
};

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

// CLEAN_LOW_PC (def-Actions.cpp:240-253) and the RK4 combination of one
// plane (def-RK.cpp:116-138)
template <typename T, int L0, int M0, int C>
__device__ __forceinline__ void spec_rk_plane(const StageArgs<T>& a, T acc, long long idx, int p) {
  constexpr int l = plane_l<L0, M0>(C);
  if (l >= 1 && p < (l - 1 < 80 ? l - 1 : 80)) acc = T(0);
  const long long o = (long long)C * a.plane + idx;
  switch (a.mode) {
    case 0:
      a.out[o] = acc;
      break;
    case 1: {
      const T y0 = a.y0[o];
      const T kh = acc * a.c05h;
      a.out[o] = y0 + kh;
      a.acc[o] = y0 + kh * a.third;
    } break;
    case 2: {
      const T kh = acc * a.c05h;
      a.out[o] = a.y0[o] + kh;
      a.acc[o] = a.acc[o] + kh * a.twothird;
    } break;
    case 3: {
      const T kh = acc * a.hh;
      a.out[o] = a.y0[o] + kh;
      a.acc[o] = a.acc[o] + kh * a.third;
    } break;
    default: {
      const T kh = acc * a.h6;
      a.out[o] = a.acc[o] + kh;
    } break;
  }
}
template <typename T, int L0, int M0, int NC, int... C>
__device__ __forceinline__ void spec_rk(const StageArgs<T>& a, const T (&acc)[NC], long long idx, int p,
                                        std::integer_sequence<int, C...>) {
  (spec_rk_plane<T, L0, M0, C>(a, acc[C], idx, p), ...);
}

template <typename T, int L0, int M0, bool TWO_D>
__global__ __launch_bounds__(kTileX* kTileP) void vlasov_stage_spec(StageArgs<T> a) {
  using Sh = SpecShape<L0, M0>;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = blockIdx.x * kTileX + tx;
  const int p = __builtin_amdgcn_readfirstlane(blockIdx.y * kTileP + ty);  // one p per wavefront
  const int y = blockIdx.z;
  const int nx = a.nx, np = a.np, ny = a.ny;
  const int nxy = nx * ny;
  __shared__ double jsh[3][kTileP][kTileX];
  // the coefficients, read by every lane at the same address (LDS broadcast):
  // as scalar loads the compiler hoists all of them and spills SGPRs
  constexpr int NT = spec_nterms<L0, M0, TWO_D>();
  __shared__ double csh[NT];
  for (int i = ty * kTileX + tx; i < NT; i += kTileX * kTileP) csh[i] = a.coef[i];
  __syncthreads();
  double j10 = 0.0, j11r = 0.0, j11i = 0.0;
  if (x < nx && p < np) {
    SpecEval<T, L0, M0> e(a, csh);
    e.p = p;
    e.idx = ((long long)y * np + p) * nx + x;
    e.xm = (x == 0) ? (nx - 1) : -1;
    e.xp = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    const long long ystride = (long long)np * nx;
    e.ym = (y == 0) ? (long long)(ny - 1) * ystride : -ystride;
    e.yp = (y == ny - 1) ? -(long long)(ny - 1) * ystride : ystride;
    const int ixy = y * nx + x;
#pragma unroll
    for (int c = 0; c < 6; ++c) e.F[c] = a.fin[(long long)c * nxy + ixy];
    e.v = ldc(a.vtab, p);
    e.bgg = a.bgG ? ldc(a.bgG, p) : T(0);
    spec_terms<L0, M0, TWO_D>(e);
    asm volatile("" ::: "memory");
    const long long idx = e.idx;
    spec_rk<T, L0, M0>(a, e.acc, idx, p, std::make_integer_sequence<int, Sh::NC>{});
    // trapezoid weights of Current::Jxyz: 1/2 at both ends of the p axis
    const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
    const double g = a.p3og[p];
    j10 = w * (static_cast<double>(a.in[(long long)Sh::re(1, 0) * a.plane + idx]) * g);
    if (M0 >= 1) {
      j11r = w * (static_cast<double>(a.in[(long long)Sh::re(1, 1) * a.plane + idx]) * g);
      j11i = w * (static_cast<double>(a.in[(long long)Sh::im(1, 1) * a.plane + idx]) * g);
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

// (l0, m0, 2D) shapes with a specialised stage kernel; others use vlasov_stage
#define OSHUN_STAGE_SPECS(X) \
  X(0, 4, 4, true)           \
  X(1, 4, 3, true)           \
  X(2, 6, 3, true)           \
  X(3, 7, 3, true)           \
  X(4, 8, 2, false)          \
  X(5, 3, 2, false)

}  // namespace oshun
