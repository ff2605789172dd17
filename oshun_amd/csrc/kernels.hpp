// gfx950 kernels of the distribution-function update.
//
// Device layout: every real plane (the real part of each harmonic, plus the
// imaginary part when m > 0) is [y][p][x] with x fastest, planes back to
// back; the six fields are double planes [y][x].  Loads along x are coalesced;
// the p-stencil of the E-field operators is three rows of the same plane.
//
//   vlasov_stage  one RK4 stage of Actions::Adv for every plane, fused with
//                 the RK4 combination of def-RK.cpp:116-138 and with the
//                 partial p-sums of Current::Jxyz (def-VlasovMax.cpp:680-745)
//   field_stage   J + Ampere + Faraday (def-VlasovMax.cpp:680-745,979-1035)
//                 and the same RK4 combination for the six fields
//   collide       Explicit_EC_FP (e-e on f00, RK4 sub-cycled) followed by
//                 Implicit_FLM_step for every l >= 1 plane, one wavefront per
//                 (x,y) column with p across the 64 lanes; the Rosenbluth
//                 integrals are wavefront prefix scans and the tridiagonal
//                 solves are wavefront scans (Moebius scan for the Thomas
//                 pivots, affine scans for the two substitutions).
#pragma once

#include <hip/hip_runtime.h>

#include "tables.hpp"

namespace oshun {

constexpr int kTileX = 64;
constexpr int kTileP = 4;

template <typename T>
struct StageArgs {
  const T* in;    // stage input planes
  const T* y0;    // state at the start of the step
  T* out;         // stage output (slope when mode == 0)
  T* acc;         // RK accumulator
  const double* fin;  // fields of the stage input, 6 planes of nxy
  double* jpart;      // [npb][3][nxy] partial trapezoid sums of p^3/gamma f1m
  const Term* terms;
  const int* tstart;
  const int* clean;
  const T* invpax;
  const T* vtab;
  const T* cg;
  const T* hp0;
  const double* p3og;
  const T* bgG;  // G00 of the f00 background (plane 0 stores f00 minus it), or null
  // term coefficients in visit order (vlasov_stage_spec); add_complex c and d
  // per target (vlasov_stage_lm, which reads tstart as their per-plane start)
  const double* coef;
  T p0p1_sq, inv_mp0p1_sq, g_r;
  T c05h, hh, h6, third, twothird;
  int np, nx, ny, nc;
  // vlasov_stage_lm only: l0 and the interior coefficient layout (LmPlan)
  int l0, lm_ltop, lm_istart, lm_istride, lm_ioff[5];
  long long plane;
  int mode;  // 0: slope, 1..4: RK4 stage, 5..6: RK2 stage (vlasov_stage only)
  int c10, c11r, c11i;
  // 1: the planes hold the reference's whole arrays, guard cells included,
  // and x/y differences follow its flattened-array rule (ext_offsets)
  int ext = 0;
};

// x and y neighbours of a point of the whole (guard-including) arrays as
// Matrix3D::Dd2 / Dd3 (matrices.h:811-829) and Matrix2D::Dd1 / Dd2 (:511-533)
// take them: minus central differences over the flattened array, so that
//   x = 0 of row y > 0 pairs with x = nx-1 of row y-1, x = nx-1 of row y < ny-1
//   with x = 0 of row y+1; the very first point takes f(x=0) - f(x=2) and the
//   very last keeps f itself (no subtraction);
//   y = 0 takes f(y=0) - f(y=2), y = ny-1 keeps f.
// Offsets are in units of the x stride 1 and the row stride ys; *last set
// means "no subtrahend" (the minus-neighbour offset is then 0: f itself).
__host__ __device__ __forceinline__ void ext_offsets(int x, int y, int nx, int ny, long long ys, long long* xm,
                                                     long long* xp, long long* ym, long long* yp, bool* xlast,
                                                     bool* ylast) {
  *xlast = false;
  *ylast = false;
  if (x > 0) *xm = -1;
  else if (y > 0) *xm = -ys + (nx - 1);
  else *xm = 0;
  if (x == 0 && y == 0) *xp = 2;
  else if (x < nx - 1) *xp = 1;
  else if (y < ny - 1) *xp = ys - (nx - 1);
  else { *xm = 0; *xp = 0; *xlast = true; }
  if (y == 0) { *ym = 0; *yp = 2 * ys; }
  else if (y < ny - 1) { *ym = -ys; *yp = ys; }
  else { *ym = 0; *yp = 0; *ylast = true; }
}

template <typename T>
__device__ __forceinline__ T op_value(const Term& tm, const T* __restrict__ s, long long idx, int p, int np, int nx,
                                      long long xm_off, long long xp_off, long long ym_off, long long yp_off,
                                      bool xlast, bool ylast, const StageArgs<T>& a) {
  switch (tm.op) {
    case OP_DX:
      return s[idx + xm_off] - (xlast ? T(0) : s[idx + xp_off]);
    case OP_DY:
      return s[idx + ym_off] - (ylast ? T(0) : s[idx + yp_off]);
    case OP_ID:
      return s[idx];
    default: {
      const T f0 = s[idx];
      if (p == 0) {
        const T f1 = s[idx + nx];
        if (tm.op == OP_G) return T(0);
        if (tm.op == OP_H) return f1 * a.hp0[tm.l];                       // H(p0), def-VlasovMax.cpp:513-514
        const T fe = (f0 - f1 * a.p0p1_sq) * a.inv_mp0p1_sq;              // G00, :531-532
        return (f1 - fe) * a.g_r;
      }
      const T fm = s[idx - nx];
      const T dpf = (p == np - 1) ? (fm - f0) : (fm - s[idx + nx]);       // SHarmonic::Dp, def-State.cpp:359-380
      if (tm.op == OP_G00) return dpf;
      T hv = f0 * a.invpax[tm.l * np + p];                                 // MakeGH, :504-509
      hv = hv + dpf;
      if (tm.op == OP_H) return hv;
      return dpf * a.cg[tm.l] + hv;
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kTileX* kTileP) void vlasov_stage(StageArgs<T> a) {
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int x = blockIdx.x * kTileX + tx;
  const int p = blockIdx.y * kTileP + ty;
  const int y = blockIdx.z;
  const int nx = a.nx, np = a.np, ny = a.ny;
  const int nxy = nx * ny;
  __shared__ double jsh[3][kTileP][kTileX];
  double j10 = 0.0, j11r = 0.0, j11i = 0.0;
  if (x < nx && p < np) {
    const long long idx = ((long long)y * np + p) * nx + x;
    const long long ystride = (long long)np * nx;
    long long xm_off = (x == 0) ? (nx - 1) : -1;
    long long xp_off = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    long long ym_off = (y == 0) ? (long long)(ny - 1) * ystride : -ystride;
    long long yp_off = (y == ny - 1) ? -(long long)(ny - 1) * ystride : ystride;
    bool xlast = false, ylast = false;
    if (a.ext) ext_offsets(x, y, nx, ny, ystride, &xm_off, &xp_off, &ym_off, &yp_off, &xlast, &ylast);
    const int ixy = y * nx + x;
    double F[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) F[k] = a.fin[(long long)k * nxy + ixy];
    const T v = a.vtab[p];
    for (int tc = 0; tc < a.nc; ++tc) {
      T acc = T(0);
      const int t1 = a.tstart[tc + 1];
      for (int t = a.tstart[tc]; t < t1; ++t) {
        const Term tm = a.terms[t];
        const T* s = a.in + (long long)tm.src * a.plane;
        T val = op_value<T>(tm, s, idx, p, np, nx, xm_off, xp_off, ym_off, yp_off, xlast, ylast, a);
        if (tm.op == OP_G00 && a.bgG) val += a.bgG[p];  // G00 is linear: G00(bg + df) = G00(bg) + G00(df)
        T mult;
        switch (tm.mul) {  // wave-uniform; keeps F[] in registers
          case MUL_EX: mult = static_cast<T>(F[0] * tm.c); break;
          case MUL_EY: mult = static_cast<T>(F[1] * tm.c); break;
          case MUL_EZ: mult = static_cast<T>(F[2] * tm.c); break;
          case MUL_BX: mult = static_cast<T>(F[3] * tm.c); break;
          case MUL_BY: mult = static_cast<T>(F[4] * tm.c); break;
          case MUL_BZ: mult = static_cast<T>(F[5] * tm.c); break;
          default: mult = v * static_cast<T>(tm.c); break;
        }
        acc += val * mult;
      }
      if (p < a.clean[tc]) acc = T(0);  // CLEAN_LOW_PC, def-Actions.cpp:240-253
      const long long o = (long long)tc * a.plane + idx;
      switch (a.mode) {  // def-RK.cpp:116-138
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
        case 5: {  // Runge_Kutta_2_Imp stage 1: Y0 = Y + h k, Y += (h k) / 2   (def-RK.cpp:478-482)
          const T kh = acc * a.hh;
          const T y0 = a.y0[o];
          a.out[o] = y0 + kh;
          a.acc[o] = y0 + kh * T(0.5);
        } break;
        case 6: {  // stage 2: Y += (h / 2) k   (:486-488)
          const T kh = acc * a.c05h;
          a.out[o] = a.acc[o] + kh;
        } break;
        default: {
          const T kh = acc * a.h6;
          a.out[o] = a.acc[o] + kh;
        } break;
      }
    }
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

// ---------------------------------------------------------------------------
// 1D, m0 = 0 specialisation (every real plane is one harmonic l, no y, only
// Ex couples).  L0 is a compile-time constant so the l loops unroll and every
// operand stays in registers.  One thread per (x, p): the 5-point stencil of
// all L0+1 planes (p-1, p, p+1, x-1, x+1) and the RK operands are loaded up
// front, so a wavefront keeps ~5 (L0+1) independent 512-byte loads in flight;
// the p+-1 and x+-1 reloads are served by L1/L2.  The sums are those of the
// coupling program in the same order: E-field terms of source l-1 then l+1
// (def-VlasovMax.cpp:358-379), x-advection of source l-1 then l+1 (:842-853).
constexpr int kRows1D = 4;  // p rows per block (block = 64 x 4)

template <typename T>
struct Stage1DArgs {
  const T* in;
  const T* y0;
  T* out;
  T* acc;
  const double* fin;  // Ex plane
  double* jpart;      // [npb][3][nx]
  const T* invpax;    // (l0+1) x np
  const T* vtab;
  const T* cg;
  const T* hp0;
  const double* p3og;
  const double* eA1;  // E-field G coefficient of source l   (l = 0: G00)
  const double* eA2;  // E-field H coefficient of source l
  const double* xA1;  // x-advection coefficient of source l towards l+1
  const double* xA2;  // x-advection coefficient of source l towards l-1
  const T* bgG;       // G00 of the f00 background (plane 0 stores f00 minus it), or null
  const T* xv1;       // v(p) * xA1[l] at [l][p]
  const T* xv2;       // v(p) * xA2[l] at [l][p]
  T p0p1_sq, inv_mp0p1_sq, g_r;
  T c05h, hh, h6, third, twothird;
  int np, nx, mode;
  long long plane;
  int pw;            // marched stage pairs (vlasov_march_1d): p rows per wavefront
  const T* rowc;     // their coefficients by row: [p][invpax l = 1..l0 | xv1 l = 0..l0-1 | xv2 l = 1..l0]
  const T* eAT;      // and eA1 | eA2 in the storage precision (the E multiplier is T(Ex) * eA)
  // stage pairs (vlasov_pair_1d) only
  const double* fin2;  // Ex of the second stage's input
  double* jpA;         // [ptile][3][nx] current partials of the second stage's input
  double* jpB;         // same for the pair's output (first pair only)
  double* f1A;         // f10(p1) of the second stage's input, per x
};

// Read-only table access through the constant address space: with a
// wave-uniform index the load becomes a scalar load even after the kernel's
// own global stores (which a generic pointer could alias).
template <typename T>
__device__ __forceinline__ T ldc(const T* ptr, long long i) {
  return ((const __attribute__((address_space(4))) T*)ptr)[i];
}

// Vlasov RHS of every l at one (x,p) of the 1D m0 = 0 system from the values
// of all planes at (p-1, p, p+1) and (x-1, x+1):
// Actions::Adv (def-Actions.cpp:73-94) = ElectricField::Exyz + Spatial_Advection::Axy
// restricted to m = 0, with CLEAN_LOW_PC (def-Actions.cpp:240-253).
template <typename T, int NL>
struct EMul {  // E-field multipliers of one x: T(Ex * eA1[l]), T(Ex * eA2[l]) per source l
  T e1[NL], e2[NL];
};
template <typename T, int NL>
__device__ __forceinline__ EMul<T, NL> emul(const Stage1DArgs<T>& a, double Ex) {
  EMul<T, NL> m;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    m.e1[l] = static_cast<T>(Ex * ldc(a.eA1, l));
    m.e2[l] = static_cast<T>(Ex * ldc(a.eA2, l));
  }
  return m;
}
template <typename T, int NL, bool CLEAN = true>
__device__ __forceinline__ void rhs_point_1d(const Stage1DArgs<T>& a, int p, const EMul<T, NL>& em, T v, T bgg,
                                             const T (&fm)[NL], const T (&fc)[NL], const T (&fn)[NL],
                                             const T (&fl)[NL], const T (&fr)[NL], T (&k)[NL]) {
  const int np = a.np;
  T G[NL], Hh[NL];
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    if (p == 0) {
      if (l == 0) {
        const T fe = (fc[0] - fn[0] * a.p0p1_sq) * a.inv_mp0p1_sq;   // G00, def-VlasovMax.cpp:531-532
        G[0] = (fn[0] - fe) * a.g_r;
        if (a.bgG) G[0] += bgg;
        Hh[0] = T(0);
      } else {
        G[l] = T(0);
        Hh[l] = fn[l] * ldc(a.hp0, l);                                    // H(p0), :513-514
      }
    } else {
      // SHarmonic::Dp, def-State.cpp:359-380: f(p-1) - f(p+1), and f(N-2) - f(N-1)
      // at pmax, which is the same expression because callers pass fn = fc there
      const T dpf = fm[l] - fn[l];
      if (l == 0) {
        G[0] = dpf;
        if (a.bgG) G[0] += bgg;
        Hh[0] = T(0);
      } else {
        T hv = fc[l] * ldc(a.invpax, l * np + p);                           // MakeGH, :504-509
        hv = hv + dpf;
        Hh[l] = hv;
        G[l] = dpf * ldc(a.cg, l) + hv;
      }
    }
  }
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    T acc = T(0);
    if (l >= 1) acc += G[l - 1] * em.e1[l - 1];
    if (l + 1 < NL) acc += Hh[l + 1] * em.e2[l + 1];
    if (l >= 1) acc += (fl[l - 1] - fr[l - 1]) * ldc(a.xv1, (l - 1) * np + p);  // v * xA1, wave-uniform
    if (l + 1 < NL) acc += (fl[l + 1] - fr[l + 1]) * ldc(a.xv2, (l + 1) * np + p);
    k[l] = acc;
  }
  if (CLEAN && p < NL - 2) {  // CLEAN_LOW_PC, def-Actions.cpp:240-253 (rows p < min(l-1, 80) of l >= 1)
#pragma unroll
    for (int l = 2; l < NL; ++l)
      if (p < min(l - 1, 80)) k[l] = T(0);
  }
}

// The same right-hand side for the marched kernels: the x differences given
// (dx[l] = f_l(x-1) - f_l(x+1)), the E multipliers formed where they are
// used, and the coefficients of row p read from one contiguous block (a few
// wide scalar loads): identical arithmetic in FP64.  GENERAL = false leaves out the
// rows that need a special case (p = 0, CLEAN_LOW_PC rows; callers pass fn =
// fc at pmax themselves).
template <typename T, int NL, bool GENERAL>
__device__ __forceinline__ void rhs_point_1d_dx(const Stage1DArgs<T>& a, int p, double Ex, T bgg, const T (&fm)[NL],
                                                const T (&fc)[NL], const T (&fn)[NL], const T (&dx)[NL], T (&k)[NL]) {
  constexpr int L0 = NL - 1;
  const T* rc = a.rowc + (long long)p * (3 * L0);
  T G[NL], Hh[NL];
  // one branch for the p = 0 row, not one per harmonic
  if (GENERAL && p == 0) {
    const T fe = (fc[0] - fn[0] * a.p0p1_sq) * a.inv_mp0p1_sq;   // G00, def-VlasovMax.cpp:531-532
    G[0] = (fn[0] - fe) * a.g_r;
    if (a.bgG) G[0] += bgg;
    Hh[0] = T(0);
#pragma unroll
    for (int l = 1; l < NL; ++l) {
      G[l] = T(0);
      Hh[l] = fn[l] * ldc(a.hp0, l);                                    // H(p0), :513-514
    }
  } else {
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const T dpf = fm[l] - fn[l];                                      // SHarmonic::Dp, def-State.cpp:359-380
      if (l == 0) {
        G[0] = dpf;
        if (a.bgG) G[0] += bgg;
        Hh[0] = T(0);
      } else {
        T hv = fc[l] * ldc(rc, l - 1);                                  // MakeGH, :504-509 (invpax)
        hv = hv + dpf;
        Hh[l] = hv;
        G[l] = dpf * ldc(a.cg, l) + hv;
      }
    }
  }
  // E multipliers in the storage precision: in FP64 the same products as
  // rhs_point_1d; in FP32 a float product (the slope is float throughout)
  const T ex = static_cast<T>(Ex);
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    T acc = T(0);
    if (l >= 1) acc += G[l - 1] * (ex * ldc(a.eAT, l - 1));
    if (l + 1 < NL) acc += Hh[l + 1] * (ex * ldc(a.eAT, NL + l + 1));
    if (l >= 1) acc += dx[l - 1] * ldc(rc, L0 + l - 1);       // v * xA1 of source l-1
    if (l + 1 < NL) acc += dx[l + 1] * ldc(rc, 2 * L0 + l);   // v * xA2 of source l+1
    k[l] = acc;
  }
  if (GENERAL) {  // CLEAN_LOW_PC, def-Actions.cpp:240-253 (rows p < min(l-1, 80) of l >= 1), by selection
#pragma unroll
    for (int l = 2; l < NL; ++l) k[l] = (p < min(l - 1, 80)) ? T(0) : k[l];
  }
}

template <typename T, int L0>
__global__ __launch_bounds__(64 * kRows1D) void vlasov_stage_1d(Stage1DArgs<T> a) {
  constexpr int NL = L0 + 1;
  const int lane = threadIdx.x;
  const int row = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one row per wavefront
  const int x = blockIdx.x * 64 + lane;
  const int p = blockIdx.y * kRows1D + row;
  const int np = a.np, nx = a.nx;
  __shared__ double jsh[kRows1D][64];
  double jv = 0.0;
  if (x < nx && p < np) {
    const long long idx = (long long)p * nx + x;
    const long long xm = (x == 0) ? (nx - 1) : -1;
    const long long xp = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    const long long up = (p > 0) ? -(long long)nx : 0;          // row p-1 (unused at p = 0)
    const long long dn = (p < np - 1) ? (long long)nx : 0;      // row p+1 (unused at np-1)
    T fm[NL], fc[NL], fn[NL], fl[NL], fr[NL];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const T* s = a.in + (long long)l * a.plane + idx;
      fc[l] = s[0];
      fm[l] = s[up];
      fn[l] = s[dn];
      fl[l] = s[xm];
      fr[l] = s[xp];
    }
    T y0v[NL], accv[NL];
    if (a.mode == 2 || a.mode == 3) {
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        y0v[l] = a.y0[(long long)l * a.plane + idx];
        accv[l] = a.acc[(long long)l * a.plane + idx];
      }
    } else if (a.mode == 4) {
#pragma unroll
      for (int l = 0; l < NL; ++l) accv[l] = a.acc[(long long)l * a.plane + idx];
    }
    const T bgg = a.bgG ? a.bgG[p] : T(0);
    T kv[NL];
    rhs_point_1d<T, NL>(a, p, emul<T, NL>(a, a.fin[x]), a.vtab[p], bgg, fm, fc, fn, fl, fr, kv);
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const T acc = kv[l];
      const long long o = (long long)l * a.plane + idx;
      switch (a.mode) {  // def-RK.cpp:116-138
        case 0: a.out[o] = acc; break;
        case 1: { const T kh = acc * a.c05h; a.out[o] = fc[l] + kh; a.acc[o] = fc[l] + kh * a.third; } break;
        case 2: { const T kh = acc * a.c05h; a.out[o] = y0v[l] + kh; a.acc[o] = accv[l] + kh * a.twothird; } break;
        case 3: { const T kh = acc * a.hh; a.out[o] = y0v[l] + kh; a.acc[o] = accv[l] + kh * a.third; } break;
        default: { const T kh = acc * a.h6; a.out[o] = accv[l] + kh; } break;
      }
    }
    if (NL > 1) {  // Current::Jxyz integrand, trapezoid weights
      const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
      jv = w * (static_cast<double>(fc[1]) * a.p3og[p]);
    }
  }
  jsh[row][lane] = jv;
  __syncthreads();
  if (row == 0 && x < nx) {
    double s2 = jsh[0][lane];
#pragma unroll
    for (int k = 1; k < kRows1D; ++k) s2 += jsh[k][lane];
    a.jpart[((long long)blockIdx.y * 3) * nx + x] = s2;
  }
}

// Workgroups are dealt to the 8 XCDs round-robin; remap so that each XCD
// works on one contiguous run of tiles and neighbouring x columns share its L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks) {
  constexpr int kXcd = 8;
  if (nblocks % kXcd != 0) return bid;
  return (bid % kXcd) * (nblocks / kXcd) + bid / kXcd;
}

// Batched launches: the workgroups of one launch cover the tiles of many
// independent runs.  ts[r] is the first global tile of run r (ts[nruns] =
// total); a workgroup finds its run by binary search and reads that run's
// arguments from device memory.
__device__ __forceinline__ int batch_run(const int* __restrict__ ts, int nruns, int g) {
  int lo = 0, hi = nruns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (ts[mid] <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Two RK4 stages in one pass (1D, m0 = 0).  A workgroup owns kPairX x columns
// and TP = 4R-2 p rows.  Stage one is evaluated on the tile plus one column
// and one row on each side and kept in LDS; stage two reads its stencil from
// there.  Per RK4 step the state crosses HBM 7 times (read Y0; write Y2 and
// the accumulator; read Y2, Y0 and the accumulator; write Y) instead of 16
// for four separate stages.  The stage arithmetic is the same as in
// vlasov_stage_1d (def-RK.cpp:116-138): PAIR 0 = stages 1+2, PAIR 1 = 3+4.
// PAIR 2 and 3 are the two stages of one RK2 of Runge_Kutta_2_Imp
// (def-RK.cpp:470-524: Y0 = Y + h k1, Y += (h k1)/2, Y += (h/2) k2): PAIR 2 on
// the x-advection alone (Adv1), PAIR 3 on the E couplings alone with
// CLEAN_LOW_PC (Adv2, Exyz1D), PAIR 4 (L0 = 3) on the E couplings of planes
// 0..3 into targets l <= 2 without CLEAN_LOW_PC (Implicit_Ex1D, the field
// solve's probe); no current partials.
constexpr int kPairX = 62;  // interior columns; lanes 0 and 63 are the halo columns

// raw buffer access: 32-bit byte offsets, plane offset in an SGPR, and
// hardware range checking against the buffer size
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ void bstore<double>(double v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}
template <>
__device__ __forceinline__ void bstore<float>(float v, __amdgpu_buffer_rsrc_t r, int voff, int soff) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, soff, 0);
}

// value of lane-1 (wave_shr:1) or lane+1 (wave_shl:1) across the whole
// wavefront; the lane without a source (0, resp. 63) keeps `fill`
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;
template <int CTRL>
__device__ __forceinline__ double wave_shift(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float wave_shift(float v, float fill) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, fill), __builtin_bit_cast(int, v),
                                                               CTRL, 0xF, 0xF, false));
}

template <typename T, int L0, int R, int PAIR>
__device__ __forceinline__ void pair_body(const Stage1DArgs<T>& a, const int tile) {
  constexpr int NL = L0 + 1;
  constexpr int TP = 4 * R - 2;  // interior rows
  constexpr int S1 = TP + 2;     // rows of the first stage
  constexpr int SZ = sizeof(T);
  __shared__ T ys[S1][NL][64];
  __shared__ double jsh[2][4][64];
  const int lane = threadIdx.x;
  const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one row group per wavefront
  const int nx = a.nx, np = a.np;
  // p-tiles of one x strip are consecutive and each XCD gets a contiguous run
  // of tiles, so the halo rows shared by neighbouring p-tiles are L2 hits
  const int npt = (np + TP - 1) / TP;
  if (tile >= ((nx + kPairX - 1) / kPairX) * npt) return;  // grid padding (whole workgroup)
  const int x0 = (tile / npt) * kPairX;
  const int pt = tile % npt;
  const int p0 = pt * TP;
  const int xe = ((x0 - 1 + lane) % nx + nx) % nx;
  const int xm = (xe == 0) ? nx - 1 : xe - 1, xp = (xe == nx - 1) ? 0 : xe + 1;
  const bool inner = lane >= 1 && lane <= kPairX && x0 + lane - 1 < nx;
  const unsigned nbytes = static_cast<unsigned>(a.plane * NL * SZ);
  const int pb = static_cast<int>(a.plane) * SZ;  // plane stride in bytes
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(a.in, nbytes);
  const __amdgpu_buffer_rsrc_t rout = buf_rsrc(a.out, nbytes);
  const __amdgpu_buffer_rsrc_t racc = buf_rsrc(a.acc, nbytes);
  const __amdgpu_buffer_rsrc_t ry0 = buf_rsrc(PAIR == 1 ? a.y0 : a.in, nbytes);
  // Y0 of the tile points stays in registers for stage 2 (a re-read would
  // come from HBM: the tile's inputs are out of L2 by then)
  constexpr bool kKeepY0 = true;
  T accr[R][NL], y0r[R][NL];
  // ---- first stage on rows p0-1 .. p0+TP, columns x0-1 .. x0+kPairX.  Wave ty
  // marches over rows s = ty*R .. ty*R+R-1 with rows p-1, p, p+1 in registers;
  // the x neighbours come from the adjacent lanes, except for the outer
  // columns of lanes 0 and 63 (one load), and the next row is in flight while
  // the current one is evaluated.
  constexpr int kOut = 0x7ffffff0;  // offset outside the buffer: the load returns 0
  const int xedge = (lane == 0) ? xm : xp;
  auto ld_row = [&](int q, T (&v)[NL]) {
    const int o = (q >= 0 && q < np) ? (q * nx + xe) * SZ : kOut;
#pragma unroll
    for (int l = 0; l < NL; ++l) v[l] = bload<T>(rin, o, l * pb);
  };
  auto ld_edge = [&](int q, T (&v)[NL]) {
    const int o = (q >= 0 && q < np && (lane == 0 || lane == 63)) ? (q * nx + xedge) * SZ : kOut;
#pragma unroll
    for (int l = 0; l < NL; ++l) v[l] = bload<T>(rin, o, l * pb);
  };
  // stage 3 reads Y0 at every point it evaluates and the accumulator at the tile's own points
  // (0 elsewhere)
  auto ld_y0acc = [&](int r, T (&y0)[NL], T (&ac)[NL]) {
    const int s = ty * R + r, q = p0 - 1 + s;
    const bool ok = q >= 0 && q < np;
    const int o = ok ? (q * nx + xe) * SZ : kOut;
    const int oa = (ok && s >= 1 && s <= TP && inner) ? o : kOut;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      y0[l] = bload<T>(ry0, o, l * pb);
      ac[l] = bload<T>(racc, oa, l * pb);
    }
  };
  // the E multipliers are re-formed per row: holding them for the whole tile
  // costs a wavefront of occupancy (measured slower in FP32 and FP64)
  const int pa = p0 - 1 + ty * R;  // first row of this wave (wave-uniform)
  T fm[NL], fc[NL], fn[NL], fe[NL], cy0[NL], cac[NL];
  ld_row(pa - 1, fm);
  ld_row(pa, fc);
  ld_row(pa + 1, fn);
  ld_edge(pa, fe);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = ty * R + r;
    const int p = pa + r;  // wave-uniform
    // issued ahead of the next row's loads, so that waiting for them leaves those in flight
    if (PAIR == 1) ld_y0acc(r, cy0, cac);
    T nn[NL], ne[NL];
    if (r + 1 < R) {
      ld_row(p + 2, nn);
      ld_edge(p + 1, ne);
    }
    if (p >= 0 && p < np) {
      T fnv[NL], fl[NL], fr[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        fnv[l] = (p < np - 1) ? fn[l] : fc[l];  // f(N-1) itself at pmax; fm is unused at p = 0
        fl[l] = wave_shift<kWaveShr1>(fc[l], fe[l]);
        fr[l] = wave_shift<kWaveShl1>(fc[l], fe[l]);
      }
      if (PAIR >= 3) {  // E couplings only: fl - fr = 0 adds +-0
#pragma unroll
        for (int l = 0; l < NL; ++l) { fl[l] = T(0); fr[l] = T(0); }
      }
      T kv[NL];
      rhs_point_1d<T, NL, PAIR != 2 && PAIR != 4>(a, p, emul<T, NL>(a, PAIR == 2 ? 0.0 : a.fin[xe]), ldc(a.vtab, p),
                                     a.bgG ? ldc(a.bgG, p) : T(0), fm, fc, fnv, fl, fr, kv);
      if (PAIR == 4) kv[NL - 1] = T(0);  // Implicit_Ex1D: targets l <= 2 (NL = 4)
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        T y;
        if (PAIR >= 2) {  // RK2 stage 1: Y0 = Y + h k1, acc = Y + (h k1)/2
          const T kh = kv[l] * a.hh;
          y = fc[l] + kh;
          accr[r][l] = fc[l] + kh * T(0.5);
          y0r[r][l] = T(0);
        } else if (PAIR == 0) {  // stage 1: Y1 = Y0 + h/2 k1, acc = Y0 + h/6 k1
          const T kh = kv[l] * a.c05h;
          y = fc[l] + kh;
          accr[r][l] = fc[l] + kh * a.third;
          y0r[r][l] = fc[l];
        } else {          // stage 3: Y3 = Y0 + h k3, acc += h/3 k3 (0 outside the tile)
          const T kh = kv[l] * a.hh;
          y = cy0[l] + kh;
          accr[r][l] = cac[l] + kh * a.third;
          y0r[r][l] = T(0);
        }
        ys[s][l][lane] = y;
      }
    } else {
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        ys[s][l][lane] = T(0);
        accr[r][l] = T(0);
        y0r[r][l] = T(0);
      }
    }
    if (r + 1 < R) {
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        fm[l] = fc[l]; fc[l] = fn[l]; fn[l] = nn[l]; fe[l] = ne[l];
      }
    }
  }
  __syncthreads();
  // ---- second stage on the tile, stencil from LDS
  double jin = 0.0, jout = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int s = ty * R + r;
    const int p = p0 - 1 + s;  // wave-uniform
    if (s >= 1 && s <= TP && p < np && inner) {
      // rows outside 0..np-1 hold zeros and are not read: p = 0 uses no f(p-1),
      // and at pmax f(p+1) is replaced by f(p) as in the first stage
      T fm[NL], fc[NL], fn[NL], fl[NL], fr[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        fc[l] = ys[s][l][lane];
        fm[l] = ys[s - 1][l][lane];
        fn[l] = (p < np - 1) ? ys[s + 1][l][lane] : fc[l];
        fl[l] = (PAIR >= 3) ? T(0) : ys[s][l][lane - 1];
        fr[l] = (PAIR >= 3) ? T(0) : ys[s][l][lane + 1];
      }
      T kv[NL];
      rhs_point_1d<T, NL, PAIR != 2 && PAIR != 4>(a, p, emul<T, NL>(a, PAIR == 2 ? 0.0 : a.fin2[xe]), ldc(a.vtab, p),
                                     a.bgG ? ldc(a.bgG, p) : T(0), fm, fc, fn, fl, fr, kv);
      if (PAIR == 4) kv[NL - 1] = T(0);
      const int vc = (p * nx + xe) * SZ;
      if (PAIR >= 2) {  // RK2 stage 2: Y = acc + (h/2) k2
#pragma unroll
        for (int l = 0; l < NL; ++l) bstore<T>(accr[r][l] + kv[l] * a.c05h, rout, vc, l * pb);
        continue;
      }
      const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;  // Current::Jxyz trapezoid weights
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        if (PAIR == 0) {  // stage 2: Y2 = Y0 + h/2 k2, acc += h/3 k2
          const T kh = kv[l] * a.c05h;
          const T y2 = (kKeepY0 ? y0r[r][l] : bload<T>(rin, vc, l * pb)) + kh;
          bstore<T>(y2, rout, vc, l * pb);
          bstore<T>(accr[r][l] + kh * a.twothird, racc, vc, l * pb);
          if (l == 1) jout += w * (static_cast<double>(y2) * ldc(a.p3og, p));
        } else {          // stage 4: Y = acc + h/6 k4
          const T kh = kv[l] * a.h6;
          bstore<T>(accr[r][l] + kh, rout, vc, l * pb);
        }
      }
      if (NL > 1) {
        jin += w * (static_cast<double>(fc[1]) * ldc(a.p3og, p));
        if (p == 1) a.f1A[xe] = static_cast<double>(fc[1]);
      }
    }
  }
  if (PAIR >= 2) return;  // (uniform) no current partials for the RK2 stages
  jsh[0][ty][lane] = jin;
  jsh[1][ty][lane] = jout;
  __syncthreads();
  if (ty == 0 && inner) {
    const long long o = ((long long)pt * 3) * nx + xe;
    a.jpA[o] = ((jsh[0][0][lane] + jsh[0][1][lane]) + jsh[0][2][lane]) + jsh[0][3][lane];
    if (PAIR == 0) a.jpB[o] = ((jsh[1][0][lane] + jsh[1][1][lane]) + jsh[1][2][lane]) + jsh[1][3][lane];
  }
}

template <typename T, int L0, int R, int PAIR>
__global__ __launch_bounds__(256) void vlasov_pair_1d(Stage1DArgs<T> a) {
  pair_body<T, L0, R, PAIR>(a, xcd_tile(blockIdx.x, gridDim.x));
}
template <typename T, int L0, int R, int PAIR>
__global__ __launch_bounds__(256) void vlasov_pair_1d_batch(const Stage1DArgs<T>* __restrict__ args,
                                                            const int* __restrict__ ts, int nruns) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  if (g >= ts[nruns]) return;  // grid padding
  const int r = batch_run(ts, nruns, g);
  pair_body<T, L0, R, PAIR>(args[r], g - ts[r]);
}

// ---------------------------------------------------------------------------
// Two RK4 stages in one pass with every wavefront marching along p on its
// own (1D, m0 = 0).  A wavefront owns kMarchX x columns and PW p rows (one
// "chunk"): lanes 2..61 are its columns, lanes 0, 1, 62 and 63 the halo
// (stage one is exact on lanes 1..62, stage two on 2..61).  At step s it holds
// the input rows q-1, q, q+1 with the next rows in flight, evaluates stage
// one on row q = p0-1+s and stage two on row q-1 from the stage-one rows
// q-2, q-1, q it keeps: no stage goes through LDS, no wavefront waits for
// another, and loads are issued on every row rather than in a burst at the
// start of a tile.  x neighbours come from the adjacent lanes (wave_shr:1,
// wave_shl:1).  The point arithmetic is rhs_point_1d and the stage
// epilogues are those of pair_body (def-RK.cpp:116-138): PAIR 0 = stages
// 1+2, PAIR 1 = 3+4.  The current partials of the second stage's input (and,
// for PAIR 0, of Y2) are summed per chunk in row order.
constexpr int kMarchX = 60;

template <typename T, int L0, int PAIR, bool GENERAL, int DIR>
__device__ __forceinline__ void march_body(const Stage1DArgs<T>& a, const int strip, const int chunk) {
  constexpr int NL = L0 + 1;
  constexpr int SZ = sizeof(T);
  constexpr bool kDeep = sizeof(T) == 4;  // FP32: a second input row and the stage-3 rows one step ahead
  constexpr int kOut = 0x7ffffff0;        // offset outside the buffer: loads return 0, stores drop
  const int PW = a.pw;
  const int NS = PW + 2;  // stage-one rows p0-1 .. p0+PW
  const int lane = threadIdx.x & 63;
  const int nx = a.nx, np = a.np;
  const int x0 = strip * kMarchX;
  const int xe = ((x0 - 2 + lane) % nx + nx) % nx;
  const bool inner = lane >= 2 && lane < 2 + kMarchX && x0 + lane - 2 < nx;
  const int p0 = chunk * PW;  // wave-uniform
  // DIR = +1 walks p upwards from p0-1, DIR = -1 downwards from p0+PW
  const int q0 = DIR > 0 ? p0 - 1 : p0 + PW;
  const unsigned nbytes = static_cast<unsigned>(a.plane * NL * SZ);
  const int pb = static_cast<int>(a.plane) * SZ;
  const __amdgpu_buffer_rsrc_t rin = buf_rsrc(a.in, nbytes);
  const __amdgpu_buffer_rsrc_t rout = buf_rsrc(a.out, nbytes);
  const __amdgpu_buffer_rsrc_t racc = buf_rsrc(a.acc, nbytes);
  const __amdgpu_buffer_rsrc_t ry0 = buf_rsrc(PAIR == 1 ? a.y0 : a.in, nbytes);
  // rows outside 0..np-1 read 0 (only chunks at the ends of p have them)
  auto row_ofs = [&](int q) { return (!GENERAL || (q >= 0 && q < np)) ? (q * nx + xe) * SZ : kOut; };
  auto ld_in = [&](int q, T (&v)[NL]) {
    const int o = row_ofs(q);
#pragma unroll
    for (int l = 0; l < NL; ++l) v[l] = bload<T>(rin, o, l * pb);
  };
  // stage 3 reads Y0 at every stage-one row and the accumulator at the chunk's rows
  auto ld_y0acc = [&](int s, T (&y0)[NL], T (&acv)[NL]) {
    const int o = row_ofs(q0 + DIR * s);
    const int oa = (s >= 1 && s <= PW) ? o : kOut;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      y0[l] = bload<T>(ry0, o, l * pb);
      acv[l] = bload<T>(racc, oa, l * pb);
    }
  };
  // input rows behind (q - DIR), at and ahead of q and the rows in flight;
  // stage-one rows q - 2 DIR, q - DIR; the accumulator of row q - DIR (the
  // windows move by one row per step, in the direction of the march)
  T im[NL], ic[NL], inx[NL], in1[NL], in2[NL];
  ld_in(q0 - DIR, im);
  ld_in(q0, ic);
  ld_in(q0 + DIR, inx);
  if (kDeep) ld_in(q0 + 2 * DIR, in1);
  T cy0[NL], cac[NL], ny0[NL], nac[NL];
  if (PAIR == 1 && kDeep) ld_y0acc(0, cy0, cac);
  T ym[NL], yc[NL], acp[NL];
  double jin = 0.0, jout = 0.0;
#pragma unroll 1
  for (int s = 0; s < NS; ++s) {
    const int q = q0 + DIR * s;  // wave-uniform
    if (kDeep) {
      if (s + 2 < NS) ld_in(q + 3 * DIR, in2);
    } else if (s + 1 < NS) {
      ld_in(q + 2 * DIR, in1);
    }
    if (PAIR == 1) {
      if (!kDeep) ld_y0acc(s, cy0, cac);
      else if (s + 1 < NS) ld_y0acc(s + 1, ny0, nac);
    }
    // ---- stage one on row q (stage 1, or stage 3)
    T yn[NL], acn[NL];
    if (!GENERAL || (q >= 0 && q < np)) {
      const T (&fm)[NL] = DIR > 0 ? im : inx;   // row q - 1
      const T (&fp)[NL] = DIR > 0 ? inx : im;   // row q + 1
      T fnv[NL], dx[NL], kv[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        fnv[l] = (!GENERAL || q < np - 1) ? fp[l] : ic[l];  // f(N-1) itself at pmax
        dx[l] = wave_shift<kWaveShr1>(ic[l], ic[l]) - wave_shift<kWaveShl1>(ic[l], ic[l]);
      }
      rhs_point_1d_dx<T, NL, GENERAL>(a, q, a.fin[xe], a.bgG ? ldc(a.bgG, q) : T(0), fm, ic, fnv, dx, kv);
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        if (PAIR == 0) {  // stage 1: Y1 = Y0 + h/2 k1, acc = Y0 + h/6 k1
          const T kh = kv[l] * a.c05h;
          yn[l] = ic[l] + kh;
          acn[l] = ic[l] + kh * a.third;
        } else {          // stage 3: Y3 = Y0 + h k3, acc += h/3 k3
          const T kh = kv[l] * a.hh;
          yn[l] = cy0[l] + kh;
          acn[l] = cac[l] + kh * a.third;
        }
      }
    } else {
#pragma unroll
      for (int l = 0; l < NL; ++l) { yn[l] = T(0); acn[l] = T(0); }
    }
    // ---- stage two on row r = q - DIR (stage 2, or stage 4)
    const int r = q - DIR;
    if (s >= 2 && (!GENERAL || r < np)) {
      const T (&gm)[NL] = DIR > 0 ? ym : yn;    // stage-one row r - 1
      const T (&gp)[NL] = DIR > 0 ? yn : ym;    // stage-one row r + 1
      T fnv[NL], dx[NL], kv[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        fnv[l] = (!GENERAL || r < np - 1) ? gp[l] : yc[l];
        dx[l] = wave_shift<kWaveShr1>(yc[l], yc[l]) - wave_shift<kWaveShl1>(yc[l], yc[l]);
      }
      rhs_point_1d_dx<T, NL, GENERAL>(a, r, a.fin2[xe], a.bgG ? ldc(a.bgG, r) : T(0), gm, yc, fnv, dx, kv);
      const int vc = inner ? (r * nx + xe) * SZ : kOut;
      const double w = (GENERAL && (r == 0 || r == np - 1)) ? 0.5 : 1.0;  // Current::Jxyz trapezoid weights
#pragma unroll
      for (int l = 0; l < NL; ++l) {
        if (PAIR == 0) {  // stage 2: Y2 = Y0 + h/2 k2, acc += h/3 k2
          const T kh = kv[l] * a.c05h;
          const T y2 = im[l] + kh;
          bstore<T>(y2, rout, vc, l * pb);
          bstore<T>(acp[l] + kh * a.twothird, racc, vc, l * pb);
          if (l == 1) jout += w * (static_cast<double>(y2) * ldc(a.p3og, r));
        } else {          // stage 4: Y = acc + h/6 k4
          bstore<T>(acp[l] + kv[l] * a.h6, rout, vc, l * pb);
        }
      }
      if (NL > 1) {
        jin += w * (static_cast<double>(yc[1]) * ldc(a.p3og, r));
        if (GENERAL && r == 1 && inner) a.f1A[xe] = static_cast<double>(yc[1]);
      }
    }
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      im[l] = ic[l]; ic[l] = inx[l]; inx[l] = in1[l];
      if (kDeep) in1[l] = in2[l];
      ym[l] = yc[l]; yc[l] = yn[l]; acp[l] = acn[l];
      if (PAIR == 1 && kDeep) { cy0[l] = ny0[l]; cac[l] = nac[l]; }
    }
  }
  if (inner) {
    const long long o = ((long long)chunk * 3) * nx + xe;
    a.jpA[o] = jin;
    if (PAIR == 0) a.jpB[o] = jout;
  }
}

// wavefront g of a run: strip g / nchunk, chunk g % nchunk
template <typename T, int L0, int PAIR>
__device__ __forceinline__ void march_wave(const Stage1DArgs<T>& a, const int g) {
  const int nchunk = (a.np + a.pw - 1) / a.pw;
  const int chunk = g % nchunk, p0 = chunk * a.pw;
  // interior chunks: every stage row has both p neighbours inside 1..np-2 and
  // lies above the CLEAN_LOW_PC rows (no special rows, no range checks).
  // Even chunks march down and odd ones up, so that the two chunks on either
  // side of a boundary read its rows at the same time (the start of both
  // marches or the end of both) and the second read is an L2 hit (measured:
  // FP64 pair launches 3-5% faster, FP32 pairs 66.3 -> 64.2 us per step).
  const bool interior = p0 - 1 >= max(1, L0 - 1) && p0 + a.pw <= a.np - 2;
  const int strip = g / nchunk;
  if (chunk & 1) {
    if (interior) march_body<T, L0, PAIR, false, 1>(a, strip, chunk);
    else march_body<T, L0, PAIR, true, 1>(a, strip, chunk);
  } else {
    if (interior) march_body<T, L0, PAIR, false, -1>(a, strip, chunk);
    else march_body<T, L0, PAIR, true, -1>(a, strip, chunk);
  }
}

// one wavefront per (strip, chunk); the four wavefronts of a workgroup take
// consecutive chunks of one strip, and each XCD a contiguous run of workgroups
template <typename T, int L0, int PAIR>
__global__ __launch_bounds__(256) void vlasov_march_1d(Stage1DArgs<T> a) {
  const int nchunk = (a.np + a.pw - 1) / a.pw;
  const int nstrip = (a.nx + kMarchX - 1) / kMarchX;
  const int g = xcd_tile(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (g >= nstrip * nchunk) return;  // grid padding (whole wavefront)
  march_wave<T, L0, PAIR>(a, g);
}
// Batched runs: ws[r] is the first global wavefront of run r (ws[nruns] =
// total); a wavefront finds its run by binary search.
template <typename T, int L0, int PAIR>
__global__ __launch_bounds__(256) void vlasov_march_1d_batch(const Stage1DArgs<T>* __restrict__ args,
                                                             const int* __restrict__ ws, int nruns) {
  const int g = xcd_tile(blockIdx.x, gridDim.x) * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (g >= ws[nruns]) return;  // grid padding (whole wavefront)
  const int r = batch_run(ws, nruns, g);
  march_wave<T, L0, PAIR>(args[r], g - ws[r]);
}

template <typename T>
struct FieldArgs {
  const double* fin;
  const double* f0;
  double* fout;
  double* facc;
  const double* jpart;
  const T* fpl;  // distribution planes of the stage input
  const double* f1p1;  // if set: f10 at p1 per x of the stage input (instead of fpl)
  const double* p3og;  // if set: the current of f10 is summed here over p (no partials)
  int npb, np, nx, ny, mode;
  int c10, c11r, c11i;
  long long plane;
  double p3og1, small, delta_p, midx, midy;
  double c05h, hh, h6, third, twothird;
  int ext = 0;  // whole arrays with guards: flattened-array differences (ext_offsets)
};

__device__ __forceinline__ void rk_combine(int mode, double k, double y0, double* out, double* acc, double c05h,
                                           double hh, double h6, double third, double twothird) {
  switch (mode) {
    case 0: *out = k; break;
    case 1: { const double kh = k * c05h; *out = y0 + kh; *acc = y0 + kh * third; } break;
    case 2: { const double kh = k * c05h; *out = y0 + kh; *acc = *acc + kh * twothird; } break;
    case 3: { const double kh = k * hh; *out = y0 + kh; *acc = *acc + kh * third; } break;
    default: { const double kh = k * h6; *out = *acc + kh; } break;
  }
}

// partial sums of the current: kFieldRows rows of kFieldX lanes split the p
// blocks, then the row sums are added in LDS in a fixed tree (deterministic)
constexpr int kFieldX = 16, kFieldRows = 16;
__device__ __forceinline__ void sum_parts(const double* jpart, int npb, int nxy, int i, int comp, int row,
                                          double (&sh)[kFieldRows][kFieldX], double* out) {
  double s = 0.0;
  if (i < nxy) {
    for (int b0 = row; b0 < npb; b0 += 4 * kFieldRows) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int b = b0 + kFieldRows * k;
        v[k] = (b < npb) ? jpart[((long long)b * 3 + comp) * nxy + i] : 0.0;
      }
      s += (v[0] + v[1]) + (v[2] + v[3]);
    }
  }
  const int lx = threadIdx.x;
  sh[row][lx] = s;
  __syncthreads();
#pragma unroll
  for (int w = kFieldRows / 2; w >= 1; w /= 2) {
    if (row < w) sh[row][lx] += sh[row + w][lx];
    __syncthreads();
  }
  *out = sh[0][lx];
  __syncthreads();
}

// two sum_parts in one pass: each sum keeps its own order, the tree levels
// share their barriers
__device__ __forceinline__ void sum_parts2(const double* ja, const double* jb, int npb, int nxy, int i, int row,
                                           double (&sh)[2][kFieldRows][kFieldX], double* oa, double* ob) {
  double sa = 0.0, sb = 0.0;
  if (i < nxy) {
    for (int b0 = row; b0 < npb; b0 += 4 * kFieldRows) {
      double va[4], vb[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int b = b0 + kFieldRows * k;
        va[k] = (b < npb) ? ja[((long long)b * 3) * nxy + i] : 0.0;
        vb[k] = (b < npb) ? jb[((long long)b * 3) * nxy + i] : 0.0;
      }
      sa += (va[0] + va[1]) + (va[2] + va[3]);
      sb += (vb[0] + vb[1]) + (vb[2] + vb[3]);
    }
  }
  const int lx = threadIdx.x;
  sh[0][row][lx] = sa;
  sh[1][row][lx] = sb;
  __syncthreads();
#pragma unroll
  for (int w = kFieldRows / 2; w >= 1; w /= 2) {
    if (row < w) {
      sh[0][row][lx] += sh[0][row + w][lx];
      sh[1][row][lx] += sh[1][row + w][lx];
    }
    __syncthreads();
  }
  *oa = sh[0][0][lx];
  *ob = sh[1][0][lx];
}

// the current integrand of one plane summed over p by the rows of the block
// (row r takes p = r, r + kFieldRows, ...), then the same fixed tree
// row `row` of a column's current integrand (p = row, row + kFieldRows, ...,
// Current::Jxyz trapezoid weights), added in p order; eight rows' loads are
// issued before their adds
template <typename T>
__device__ __forceinline__ double direct_rows(const T* col, const double* p3og, int np, int nx, int row) {
  double s = 0.0;
  for (int p0 = row; p0 < np; p0 += 8 * kFieldRows) {
    double v[8], q[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int p = p0 + k * kFieldRows;
      v[k] = p < np ? static_cast<double>(col[(long long)p * nx]) : 0.0;
      q[k] = p < np ? p3og[p] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int p = p0 + k * kFieldRows;
      const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
      if (p < np) s += w * (v[k] * q[k]);
    }
  }
  return s;
}
template <typename T>
__device__ __forceinline__ void sum_direct(const T* plane, const double* p3og, int np, int nx, int nxy, int i,
                                           int row, double (&sh)[kFieldRows][kFieldX], double* out) {
  double s = 0.0;
  if (i < nxy) {
    const int x = i % nx, y = i / nx;
    s = direct_rows<T>(plane + (long long)y * np * nx + x, p3og, np, nx, row);
  }
  const int lx = threadIdx.x;
  sh[row][lx] = s;
  __syncthreads();
#pragma unroll
  for (int w = kFieldRows / 2; w >= 1; w /= 2) {
    if (row < w) sh[row][lx] += sh[row + w][lx];
    __syncthreads();
  }
  *out = sh[0][lx];
  __syncthreads();
}

// MaxwellEq::Ampere (def-VlasovMax.cpp:1009-1035) and Faraday (:979-1005)
// added to k; F(c, at) is field c (Ex..Bz) at x-1, x+1, y-1, y+1 (at = 0..3)
template <typename Get>
__device__ __forceinline__ void maxwell_terms(Get F, double mx, double my, double (&k)[6]) {
  k[0] += F(5, 2) * my - F(5, 3) * my;
  k[1] += F(5, 0) * (-mx) - F(5, 1) * (-mx);
  k[2] += F(3, 2) * (-my) - F(3, 3) * (-my);
  k[2] += F(4, 0) * mx - F(4, 1) * mx;
  k[3] += F(2, 2) * (-my) - F(2, 3) * (-my);
  k[4] += F(2, 0) * mx - F(2, 1) * mx;
  k[5] += F(0, 2) * my - F(0, 3) * my;
  k[5] += F(1, 0) * (-mx) - F(1, 1) * (-mx);
}

template <typename T>
__device__ __forceinline__ void field_body(const FieldArgs<T>& a, const int block) {
  const int i = block * kFieldX + threadIdx.x;
  const int row = threadIdx.y;
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  __shared__ double sh[kFieldRows][kFieldX];
  double jx = 0.0, jr = 0.0, ji = 0.0;
  if (a.c10 >= 0) {
    if (a.p3og) sum_direct<T>(a.fpl + (long long)a.c10 * a.plane, a.p3og, a.np, nx, nxy, i, row, sh, &jx);
    else sum_parts(a.jpart, a.npb, nxy, i, 0, row, sh, &jx);
  }
  if (a.c11r >= 0) {
    sum_parts(a.jpart, a.npb, nxy, i, 1, row, sh, &jr);
    sum_parts(a.jpart, a.npb, nxy, i, 2, row, sh, &ji);
  }
  if (i >= nxy || row != 0) return;
  const int x = i % nx, y = i / nx;
  int xm = y * nx + (x == 0 ? nx - 1 : x - 1), xp = y * nx + (x == nx - 1 ? 0 : x + 1);
  int ym = (y == 0 ? ny - 1 : y - 1) * nx + x, yp = (y == ny - 1 ? 0 : y + 1) * nx + x;
  bool xlast = false, ylast = false;
  if (a.ext) {
    long long oxm, oxp, oym, oyp;
    ext_offsets(x, y, nx, ny, nx, &oxm, &oxp, &oym, &oyp, &xlast, &ylast);
    xm = i + static_cast<int>(oxm); xp = i + static_cast<int>(oxp);
    ym = i + static_cast<int>(oym); yp = i + static_cast<int>(oyp);
  }
  const double* Ex = a.fin;
  const double* Ey = a.fin + nxy;
  const double* Ez = a.fin + 2 * nxy;
  const double* Bx = a.fin + 3 * nxy;
  const double* By = a.fin + 4 * nxy;
  const double* Bz = a.fin + 5 * nxy;
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // Current::Jxyz (def-VlasovMax.cpp:680-745)
  if (a.c10 >= 0) {
    double j = jx;
    j *= a.delta_p;
    const long long o1 = (long long)a.c10 * a.plane + ((long long)y * a.np + 1) * nx + x;
    j += a.small * ((a.f1p1 ? a.f1p1[i] : static_cast<double>(a.fpl[o1])) * a.p3og1);
    j *= 4.0 * M_PI / 3.0;
    k[0] += j;
  }
  if (a.c11r >= 0) {
    jr *= a.delta_p;
    ji *= a.delta_p;
    const long long o1 = ((long long)y * a.np + 1) * nx + x;
    jr += a.small * (static_cast<double>(a.fpl[(long long)a.c11r * a.plane + o1]) * a.p3og1);
    ji += a.small * (static_cast<double>(a.fpl[(long long)a.c11i * a.plane + o1]) * a.p3og1);
    jr *= 8.0 * M_PI / 3.0;
    ji *= 8.0 * M_PI / 3.0;
    k[1] += jr;
    k[2] += -ji;
  }
  maxwell_terms(
      [&](int c, int at) {
        const double* F = (c == 0) ? Ex : (c == 1) ? Ey : (c == 2) ? Ez : (c == 3) ? Bx : (c == 4) ? By : Bz;
        if ((at == 1 && xlast) || (at == 3 && ylast)) return 0.0;
        return F[at == 0 ? xm : at == 1 ? xp : at == 2 ? ym : yp];
      },
      a.midx, a.midy, k);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const long long o = (long long)c * nxy + i;
    double* acc = a.facc + o;
    rk_combine(a.mode, k[c], a.f0 ? a.f0[o] : 0.0, a.fout + o, acc, a.c05h, a.hh, a.h6, a.third, a.twothird);
  }
}

template <typename T>
__global__ __launch_bounds__(kFieldX * kFieldRows) void field_stage(FieldArgs<T> a) {
  field_body<T>(a, blockIdx.x);
}

// RK4 field stages 2 and 3 of the 1D m0 = 0 system in one pass (a2: mode 2,
// a3: mode 3, same arithmetic as two field_stage launches).  Only Ex couples
// to f and its update is local; the stage-3 update of the other five fields
// at x needs their stage-2 values at x +- 1, which are recomputed here from
// the stage-1 fields at x +- 2.  a3.fin is not read (the stage-2 fields are
// in registers) and a3.fout must not alias a2.fin.
template <typename T>
struct Field23Args {
  FieldArgs<T> a2, a3;
};
template <typename T>
__device__ __forceinline__ void field23_body(const Field23Args<T>& A, const int block) {
  const FieldArgs<T>& a = A.a2;
  const FieldArgs<T>& b = A.a3;
  const int i = block * kFieldX + threadIdx.x;
  const int row = threadIdx.y;
  const int nx = a.nx, nxy = nx;  // ny = 1
  __shared__ double sh[2][kFieldRows][kFieldX];
  auto wrap = [nx](int x) {  // x - 2 .. x + 2 into 0 .. nx-1 for any nx >= 1
    x %= nx;
    return x < 0 ? x + nx : x;
  };
  // the field inputs of row 0 are loaded before the current sums, so that
  // their latency overlaps the partial-sum loads and the reduction barriers
  double fin[6][5], f0v[6][3], accv[6], f1a = 0.0, f2b = 0.0;
  if (i < nxy && row == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int d = 0; d < 5; ++d) fin[c][d] = a.fin[(long long)c * nxy + wrap(i + d - 2)];
#pragma unroll
      for (int d = 0; d < 3; ++d) f0v[c][d] = a.f0[(long long)c * nxy + wrap(i + d - 1)];
      accv[c] = a.facc[(long long)c * nxy + i];
    }
    f1a = a.f1p1 ? a.f1p1[i] : static_cast<double>(a.fpl[(long long)a.c10 * a.plane + nx + i]);
    f2b = b.f1p1 ? b.f1p1[i] : static_cast<double>(b.fpl[(long long)b.c10 * b.plane + nx + i]);
  }
  double j1 = 0.0, j2 = 0.0;
  sum_parts2(a.jpart, b.jpart, a.npb, nxy, i, row, sh, &j1, &j2);  // a.npb == b.npb: the same pair tiles
  if (i >= nxy || row != 0) return;
  // Current::Jxyz (def-VlasovMax.cpp:680-745) of the stage inputs Y1 and Y2
  auto current = [&](const FieldArgs<T>& f, double j, double f1) {
    j *= f.delta_p;
    j += f.small * (f1 * f.p3og1);
    j *= 4.0 * M_PI / 3.0;
    return j;
  };
  // stage 2 at x-1, x, x+1 (Ex only where it is used: at x)
  double F2[3][6], acc2[6];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (d == 1) k[0] += current(a, j1, f1a);
    maxwell_terms([&](int c, int at) { return fin[c][at == 0 ? d : at == 1 ? d + 2 : d + 1]; }, a.midx, a.midy, k);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double kh = k[c] * a.c05h;
      F2[d][c] = f0v[c][d] + kh;  // rk_combine mode 2: out = y0 + kh
      if (d == 1) {
        acc2[c] = accv[c] + kh * a.twothird;
        a.fout[(long long)c * nxy + i] = F2[1][c];
      }
    }
  }
  // stage 3 at x from the stage-2 fields (rk_combine mode 3, with the
  // accumulator of stage 2 kept in registers)
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  k[0] += current(b, j2, f2b);
  maxwell_terms([&](int c, int at) { return at == 0 ? F2[0][c] : at == 1 ? F2[2][c] : F2[1][c]; }, b.midx, b.midy,
                k);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const long long o = (long long)c * nxy + i;
    const double kh = k[c] * b.hh;
    b.fout[o] = f0v[c][1] + kh;
    b.facc[o] = acc2[c] + kh * b.third;
  }
}
// The last RK4 field stage of a step (a4, mode 4) and the first of the next
// (a1, mode 1) of the 1D m0 = 0 system in one pass, for steps that end
// without a collision launch to carry them (collisions off).  The sums and
// operations are those of the two field_stage launches: the stage-4 current
// from the pair partials, the stage-1 current summed directly over p of the
// new state; the stage-4 fields at x +- 1 are recomputed here (Ex there is not
// needed: the stage-1 update at x reads Ex only at x), and the new
// accumulator goes to af2 because neighbours still read a4.facc.
template <typename T>
struct Field41Args {
  FieldArgs<T> a4, a1;
  double* af2;
};
template <typename T>
__global__ __launch_bounds__(kFieldX * kFieldRows) void field_stage41_1d(Field41Args<T> A) {
  const FieldArgs<T>& a = A.a4;
  const FieldArgs<T>& b = A.a1;
  const int i = blockIdx.x * kFieldX + threadIdx.x;
  const int row = threadIdx.y;
  const int nx = a.nx, nxy = nx;  // ny = 1
  __shared__ double sh[2][kFieldRows][kFieldX];
  auto wrap = [nx](int x) {
    x %= nx;
    return x < 0 ? x + nx : x;
  };
  // the field inputs of row 0 are loaded before the current sums, so that
  // their latency overlaps the sums' loads and barriers
  double fin[6][5], accv[6][3], f1a = 0.0, f1b = 0.0;
  if (i < nxy && row == 0) {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
#pragma unroll
      for (int d = 0; d < 5; ++d) fin[c][d] = a.fin[(long long)c * nxy + wrap(i + d - 2)];
#pragma unroll
      for (int d = 0; d < 3; ++d) accv[c][d] = a.facc[(long long)c * nxy + wrap(i + d - 1)];
    }
    f1a = a.f1p1[i];
    f1b = static_cast<double>(b.fpl[(long long)b.c10 * b.plane + nx + i]);
  }
  // the stage-4 current from the pair partials (sum_parts) and the stage-1
  // current summed directly over p (sum_direct): every load of both is issued
  // before the trees, each sum keeps its own order, the tree levels share
  // their barriers
  double s4 = 0.0, s1 = 0.0;
  if (i < nxy) {
    for (int b0 = row; b0 < a.npb; b0 += 4 * kFieldRows) {
      double v[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int pb = b0 + kFieldRows * k;
        v[k] = (pb < a.npb) ? a.jpart[((long long)pb * 3) * nxy + i] : 0.0;
      }
      s4 += (v[0] + v[1]) + (v[2] + v[3]);
    }
    s1 = direct_rows<T>(b.fpl + (long long)b.c10 * b.plane + i, b.p3og, b.np, nx, row);
  }
  const int lx = threadIdx.x;
  sh[0][row][lx] = s4;
  sh[1][row][lx] = s1;
  __syncthreads();
#pragma unroll
  for (int w = kFieldRows / 2; w >= 1; w /= 2) {
    if (row < w) {
      sh[0][row][lx] += sh[0][row + w][lx];
      sh[1][row][lx] += sh[1][row + w][lx];
    }
    __syncthreads();
  }
  if (i >= nxy || row != 0) return;
  const double j4 = sh[0][0][lx], j1 = sh[1][0][lx];
  // stage 4 at x-1, x, x+1: out = acc + (h/6) k
  double F4[3][6];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (d == 1) {  // Current::Jxyz (def-VlasovMax.cpp:680-745), as field_body
      double j = j4;
      j *= a.delta_p;
      j += a.small * (f1a * a.p3og1);
      j *= 4.0 * M_PI / 3.0;
      k[0] += j;
    }
    maxwell_terms([&](int c, int at) { return fin[c][at == 0 ? d : at == 1 ? d + 2 : d + 1]; }, a.midx, a.midy, k);
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      const double kh = k[c] * a.h6;
      F4[d][c] = accv[c][d] + kh;
    }
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) a.fout[(long long)c * nxy + i] = F4[1][c];
  // stage 1 of the next step at x: out = y0 + (h/2) k, acc = y0 + (h/2) k / 3
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  {
    double j = j1;
    j *= b.delta_p;
    j += b.small * (f1b * b.p3og1);
    j *= 4.0 * M_PI / 3.0;
    k[0] += j;
  }
  maxwell_terms([&](int c, int at) { return at == 0 ? F4[0][c] : at == 1 ? F4[2][c] : F4[1][c]; }, b.midx, b.midy, k);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const long long o = (long long)c * nxy + i;
    const double kh = k[c] * b.c05h;
    b.fout[o] = F4[1][c] + kh;
    A.af2[o] = F4[1][c] + kh * b.third;
  }
}

template <typename T>
__global__ __launch_bounds__(kFieldX * kFieldRows) void field_stage23_1d(Field23Args<T> a) {
  field23_body<T>(a, blockIdx.x);
}
template <typename T>
__global__ __launch_bounds__(kFieldX * kFieldRows) void field_stage23_1d_batch(const Field23Args<T>* __restrict__ args,
                                                                              const int* __restrict__ ts, int nruns) {
  const int g = blockIdx.x;
  const int r = batch_run(ts, nruns, g);
  field23_body<T>(args[r], g - ts[r]);
}
template <typename T>
__global__ __launch_bounds__(kFieldX * kFieldRows) void field_stage_batch(const FieldArgs<T>* __restrict__ args,
                                                                         const int* __restrict__ ts, int nruns) {
  const int g = blockIdx.x;
  const int r = batch_run(ts, nruns, g);
  field_body<T>(args[r], g - ts[r]);
}

// ------------------------------------------------------------------ collide
struct CollTables {
  const double *U4, *U4m1, *U2, *U2m1, *U1s, *U1m1s, *U3, *Qn, *Pn, *vc, *ivdm, *ivdp, *ivs2, *pw3, *pw5, *pw7;
};

template <typename T>
struct CollArgs {
  T* f;  // distribution planes
  CollTables t;
  const int* comp_l;   // l of each plane
  int nc, np, nx, ny, nb, l0;
  // the columns collided: x in [cx0, cx0 + cnx), y in [cy0, cy0 + cny) (the
  // interior when the planes keep guard cells); nx is the row stride
  int cx0 = 0, cy0 = 0, cnx = 0, cny = 0;
  long long plane;
  int do_ee, do_f1, do_flm, numh;
  double h_ee, dt_flm;
  double c_kpre, kpre, density_np, zeta, v0sq, v1sq, g_den;
  const double* rt;             // collide_rows tables, interleaved for its NPL
  const double* bg;             // f00 background per p (plane 0 stores f00 minus it), or null
  const double* tclk;           // batched runs: the run's clock before the step (else null)
  double hstep, small_dt;       // batched runs: RK4 step and e-e sub-step target
  double r01, inv_gden, c2v0;   // v0^2/v1^2, 1/g_den, 2/v0^2
  // single 1D m0 = 0 runs: the step's last RK4 field stage and, with next1,
  // the next step's first field stage done by this launch (fuse = 0: off)
  int fuse;
  struct {
    const double* cf;   // stage-3 fields (input of the last field stage)
    const double* af;   // RK4 field accumulator
    double* sf;         // out: fields after the step
    double* e1;         // out (next1): stage-1 fields of the next step
    double* af2;        // out (next1): the next step's accumulator
    const double* jp;   // current partials of the stage-4 input, [tile][3][nx]
    const double* f1y;  // f10(p1) of the stage-4 input per x
    int npt, next1, c10;
    double p3og1, small, delta_p, midx, midy, c05h, h6, third;
  } fz;
};

__device__ __forceinline__ double log_ee(double ne, double Te, double density_np) {
  // LOGee, def-FokkerPlanck.cpp:41-67
  if (ne > 0.000000001) {
    Te /= (3.0 * ne);
    Te *= 511000;
    ne *= density_np;
    Te = log(Te);
    ne = log(ne);
    const double lnee = 23.5 - 0.5 * ne + 1.25 * Te - sqrt(0.00001 + 0.0625 * (Te - 2.0) * (Te - 2.0));
    if (lnee > 2.0) return lnee;
  }
  return 2.0;
}
__device__ __forceinline__ double zlog_ei(double ne, double Te, double density_np, double zeta) {
  // ZLOGei, def-FokkerPlanck.cpp:71-95
  if (ne > 0.0000001) {
    Te /= (3.0 * ne);
    Te *= 511000;
    ne *= density_np;
    Te = log(Te);
    ne = log(ne);
    const double lnei = 24.0 - 0.5 * ne + Te;
    if (lnei > 2.0) return lnei * zeta;
  }
  return 2.0 * zeta;
}

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// ---- wavefront data movement without LDS -----------------------------------
// DPP row shifts move data inside each 16-lane row; readlane carries values
// across the four rows.  Doubles move as two 32-bit halves.
template <int N> constexpr int row_shl() { return 0x100 + N; }
template <int N> constexpr int row_shr() { return 0x110 + N; }

// lane i takes v of lane i -+ N inside its row; lanes without a source keep
// `fill` (the DPP move's old value: with bound_ctrl off, a lane whose source is
// outside its row is not written)
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_read(double v, int l) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
  return __hiloint2double(hi, lo);
}
// value held by lane-1 (lane 0 gets fill)
__device__ __forceinline__ double lane_prev(double v, double fill) {
  double e = dpp_mov<row_shr<1>()>(v, fill);
  const double b15 = lane_read(v, 15), b31 = lane_read(v, 31), b47 = lane_read(v, 47);
  const int lane = lane_id();
  e = (lane == 16) ? b15 : (lane == 32) ? b31 : (lane == 48) ? b47 : e;
  return e;
}
// value held by lane+1 (lane 63 gets fill)
__device__ __forceinline__ double lane_next(double v, double fill) {
  double e = dpp_mov<row_shl<1>()>(v, fill);
  const double b16 = lane_read(v, 16), b32 = lane_read(v, 32), b48 = lane_read(v, 48);
  const int lane = lane_id();
  e = (lane == 15) ? b16 : (lane == 31) ? b32 : (lane == 47) ? b48 : e;
  return e;
}

// inclusive prefix sum over the 64 lanes
__device__ __forceinline__ double wave_scan_incl(double v) {
  v += dpp_mov<row_shr<1>()>(v, 0.0);
  v += dpp_mov<row_shr<2>()>(v, 0.0);
  v += dpp_mov<row_shr<4>()>(v, 0.0);
  v += dpp_mov<row_shr<8>()>(v, 0.0);
  const double r0 = lane_read(v, 15), r1 = lane_read(v, 31), r2 = lane_read(v, 47);
  const int row = lane_id() >> 4;
  if (row == 1) v += r0;
  else if (row == 2) v += (r0 + r1);
  else if (row == 3) v += ((r0 + r1) + r2);
  return v;
}
// inclusive suffix sum over lanes
__device__ __forceinline__ double wave_scan_incl_rev(double v) {
  v += dpp_mov<row_shl<1>()>(v, 0.0);
  v += dpp_mov<row_shl<2>()>(v, 0.0);
  v += dpp_mov<row_shl<4>()>(v, 0.0);
  v += dpp_mov<row_shl<8>()>(v, 0.0);
  const double r1 = lane_read(v, 16), r2 = lane_read(v, 32), r3 = lane_read(v, 48);
  const int row = lane_id() >> 4;
  if (row == 2) v += r3;
  else if (row == 1) v += (r2 + r3);
  else if (row == 0) v += (r1 + (r2 + r3));
  return v;
}

// Rows held by one lane: p = lane*NPL + i.  Prefix of t over p (inclusive).
template <int NPL>
__device__ __forceinline__ void prefix_p(const double (&t)[NPL], double (&out)[NPL], double& total) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) { s += t[i]; out[i] = s; }
  const double inc = wave_scan_incl(s);
  const double ex = lane_prev(inc, 0.0);
#pragma unroll
  for (int i = 0; i < NPL; ++i) out[i] += ex;
  total = lane_read(inc, 63);
}
// Suffix of t over p (inclusive).
template <int NPL>
__device__ __forceinline__ void suffix_p(const double (&t)[NPL], double (&out)[NPL]) {
  double s = 0.0;
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { s += t[i]; out[i] = s; }
  const double inc = wave_scan_incl_rev(s);
  const double ex = lane_next(inc, 0.0);
#pragma unroll
  for (int i = 0; i < NPL; ++i) out[i] += ex;
}
// value of row p+1 (next row) for each held row
template <int NPL>
__device__ __forceinline__ void next_row(const double (&v)[NPL], double (&nx)[NPL]) {
  const double first_of_next = lane_next(v[0], 0.0);
#pragma unroll
  for (int i = 0; i < NPL - 1; ++i) nx[i] = v[i + 1];
  nx[NPL - 1] = first_of_next;
}
template <int NPL>
__device__ __forceinline__ void prev_row(const double (&v)[NPL], double (&pv)[NPL]) {
  const double last_of_prev = lane_prev(v[NPL - 1], 0.0);
#pragma unroll
  for (int i = NPL - 1; i > 0; --i) pv[i] = v[i - 1];
  pv[0] = last_of_prev;
}

// The p tables of the e-e slope in LDS, interleaved (row i of every lane
// contiguous: conflict-free reads), shared by the workgroup's columns; zero
// past np except U3 (one)
enum { ET_U4 = 0, ET_U4M1, ET_U2, ET_U2M1, ET_U1S, ET_U1M1S, ET_U3, ET_QN, ET_PN, ET_COUNT };
template <int NPL>
struct EETab {
  const double* base;  // ET_COUNT x NPL x 64
  int lane;
  __device__ __forceinline__ double get(int t, int i) const { return base[(t * NPL + i) * 64 + lane]; }
  __device__ __forceinline__ double U4(int i) const { return get(ET_U4, i); }
  __device__ __forceinline__ double U4m1(int i) const { return get(ET_U4M1, i); }
  __device__ __forceinline__ double U2(int i) const { return get(ET_U2, i); }
  __device__ __forceinline__ double U2m1(int i) const { return get(ET_U2M1, i); }
  __device__ __forceinline__ double U1s(int i) const { return get(ET_U1S, i); }
  __device__ __forceinline__ double U1m1s(int i) const { return get(ET_U1M1S, i); }
  __device__ __forceinline__ double U3(int i) const { return get(ET_U3, i); }
  __device__ __forceinline__ double Qn(int i) const { return get(ET_QN, i); }
  __device__ __forceinline__ double Pn(int i) const { return get(ET_PN, i); }
};

// EE_isotropic_conserva::Getslope, def-FokkerPlanck.cpp:192-267
template <typename T, int NPL>
__device__ __forceinline__ void ee_slope(const CollArgs<T>& a, const EETab<NPL>& tb, const double (&f)[NPL],
                                         double (&fh)[NPL]) {
  const int lane = lane_id();
  const int np = a.np;
  double fp[NPL], fn[NPL];
  prev_row<NPL>(f, fp);
  next_row<NPL>(f, fn);
  double t4[NPL], t2[NPL], t1[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    const bool in1 = p >= 1 && p < np;
    t4[i] = in1 ? tb.U4(i) * f[i] + tb.U4m1(i) * fp[i] : 0.0;
    t2[i] = in1 ? tb.U2(i) * f[i] + tb.U2m1(i) * fp[i] : 0.0;
    t1[i] = (p < np - 1) ? tb.U1s(i) * fn[i] + tb.U1m1s(i) * f[i] : 0.0;
  }
  double I4[NPL], I2[NPL], J1[NPL], tot4, tot2;
  prefix_p<NPL>(t4, I4, tot4);
  prefix_p<NPL>(t2, I2, tot2);
  suffix_p<NPL>(t1, J1);
  const double lnee = log_ee(4.0 * M_PI * tot2, 4.0 * M_PI * tot4, a.density_np);
  // f(p0), f(p1) for the parabolic G at low v
  const double f0 = lane_read(f[0], 0);
  const double f1 = (NPL >= 2) ? lane_read(f[NPL >= 2 ? 1 : 0], 0) : lane_read(f[0], 1);
  const double f00 = (f0 - f1 * a.v0sq / a.v1sq) / a.g_den;
  double G[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    double i2 = I2[i] * J1[i];
    i2 *= -3.0;
    double g = I4[i] + tb.U3(i) * J1[i];
    g *= f[i];
    g += i2;
    if (p < a.nb) {  // EE_isotropic_conserva::G, :178-188
      const double p3 = a.t.pw3[p], p5 = a.t.pw5[p], p7 = a.t.pw7[p];
      const double i2s = f00 * p3 / 3.0 + (f1 - f00) * p5 / a.v1sq * 0.2;
      const double i4s = f00 * p5 * 0.2 + (f1 - f00) * p7 / (a.v1sq * 7.0);
      g = f[i] * i4s + (p3 * f[i] - 3.0 * i2s) * J1[i];
    }
    G[i] = (p < np) ? g : 0.0;
  }
  double Gn[NPL], D1[NPL];
  next_row<NPL>(G, Gn);
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    D1[i] = (p < np - 1) ? G[i] - Gn[i] : G[i];
  }
  const double G0 = lane_read(G[0], 0);
  double fh0 = (-1.0) * G0;
  fh0 *= 2.0 / a.v0sq;
  double D2[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    D2[i] = (p < np - 1) ? D1[i] * tb.Pn(i) : D1[i];
  }
  fh0 -= lane_read(D2[0], 0);
  double D2n[NPL], D3[NPL];
  next_row<NPL>(D2, D2n);
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    D3[i] = (p < np - 1) ? D2[i] - D2n[i] : D2[i];
  }
  fh0 *= a.t.Qn[0];
  double D3p[NPL];
  prev_row<NPL>(D3, D3p);
  const double s = a.c_kpre * lnee;
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    double v = (p == 0) ? fh0 : tb.Qn(i) * D3p[i];
    fh[i] = (p < np) ? v * s : 0.0;
  }
}

// 2x2 projective scan element [[m00 m01][m10 m11]] with power-of-two rescaling
struct Mob {
  double a, b, c, d;
};
__device__ __forceinline__ Mob mob_mul(const Mob& x, const Mob& y) {  // x @ y
  return {x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
__device__ __forceinline__ Mob mob_norm(Mob m) {
  const double mx = fmax(fmax(fabs(m.a), fabs(m.b)), fmax(fabs(m.c), fabs(m.d)));
  if (mx > 0.0 && isfinite(mx)) {
    int e;
    frexp(mx, &e);
    m.a = ldexp(m.a, -e); m.b = ldexp(m.b, -e); m.c = ldexp(m.c, -e); m.d = ldexp(m.d, -e);
  }
  return m;
}

// Solve the tridiagonal system a x_{i-1} + b x_i + c x_{i+1} = d over the
// wavefront (Thomas algorithm of def-Nmethods.cpp:142-167 as three scans).
// Pivot part (depends on the matrix only): cp[i] = c'_i, w[i] = 1/(b_i - a_i c'_{i-1}).
template <int NPL>
__device__ __forceinline__ void tri_pivots(const double (&A)[NPL], const double (&B)[NPL], const double (&C)[NPL],
                                           double (&cp)[NPL], double (&w)[NPL]) {
  const int lane = lane_id();
  // per-lane product M_last ... M_first, M_i = [[0, c_i], [-a_i, b_i]]
  Mob P = {1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const Mob Mi = {0.0, C[i], -A[i], B[i]};
    P = mob_norm(mob_mul(Mi, P));
  }
  // inclusive scan S_j = P_j S_{j-1}: four DPP steps inside each row ...
#define OSHUN_MOB_STEP(N)                                                                    \
  {                                                                                          \
    Mob q;                                                                                   \
    q.a = dpp_mov<row_shr<N>()>(P.a, 1.0); q.b = dpp_mov<row_shr<N>()>(P.b, 0.0);            \
    q.c = dpp_mov<row_shr<N>()>(P.c, 0.0); q.d = dpp_mov<row_shr<N>()>(P.d, 1.0);            \
    P = mob_norm(mob_mul(P, q));                                                             \
  }
  OSHUN_MOB_STEP(1)
  OSHUN_MOB_STEP(2)
  OSHUN_MOB_STEP(4)
  OSHUN_MOB_STEP(8)
#undef OSHUN_MOB_STEP
  // ... then the products of the earlier rows
  {
    auto rd = [](const Mob& m, int l) {
      return Mob{lane_read(m.a, l), lane_read(m.b, l), lane_read(m.c, l), lane_read(m.d, l)};
    };
    const Mob R0 = rd(P, 15), R1 = rd(P, 31), R2 = rd(P, 47);
    const Mob C2 = mob_norm(mob_mul(R1, R0));
    const Mob C3 = mob_norm(mob_mul(R2, C2));
    const int row = lane >> 4;
    if (row == 1) P = mob_norm(mob_mul(P, R0));
    else if (row == 2) P = mob_norm(mob_mul(P, C2));
    else if (row == 3) P = mob_norm(mob_mul(P, C3));
  }
  Mob E;
  E.a = lane_prev(P.a, 1.0); E.b = lane_prev(P.b, 0.0);
  E.c = lane_prev(P.c, 0.0); E.d = lane_prev(P.d, 1.0);
  // state vector before this lane's first row: E @ [0, 1]
  double Pv = E.b, Qv = E.d;
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const double Pn = C[i] * Qv;
    const double Qn = -A[i] * Pv + B[i] * Qv;
    const double r = 1.0 / Qn;
    cp[i] = Pn * r;
    w[i] = Qv * r;
    // keep the vector O(1): rescale by a power of two
    int e;
    frexp(fmax(fabs(Pn), fabs(Qn)), &e);
    Pv = ldexp(Pn, -e);
    Qv = ldexp(Qn, -e);
  }
}

// d'_i = w_i (d_i - a_i d'_{i-1}); x_i = d'_i - c'_i x_{i+1}
template <int NPL>
__device__ __forceinline__ void tri_solve(const double (&A)[NPL], const double (&cp)[NPL], const double (&w)[NPL],
                                          double (&d)[NPL]) {
  const int lane = lane_id();
  // forward affine scan y -> al*y + be
  double al = 1.0, be = 0.0;
  double ali[NPL], bei[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    ali[i] = -A[i] * w[i];
    bei[i] = w[i] * d[i];
    be = ali[i] * be + bei[i];
    al = ali[i] * al;
  }
#define OSHUN_AFF_FWD(N)                                                                         \
  {                                                                                              \
    const double qa = dpp_mov<row_shr<N>()>(al, 1.0), qb = dpp_mov<row_shr<N>()>(be, 0.0);       \
    be = al * qb + be;                                                                           \
    al = al * qa;                                                                                \
  }
  OSHUN_AFF_FWD(1)
  OSHUN_AFF_FWD(2)
  OSHUN_AFF_FWD(4)
  OSHUN_AFF_FWD(8)
#undef OSHUN_AFF_FWD
  {
    const double a15 = lane_read(al, 15), b15 = lane_read(be, 15);
    const double a31 = lane_read(al, 31), b31 = lane_read(be, 31);
    const double a47 = lane_read(al, 47), b47 = lane_read(be, 47);
    const double c2a = a31 * a15, c2b = a31 * b15 + b31;  // rows 1 o 0
    const double c3a = a47 * c2a, c3b = a47 * c2b + b47;  // rows 2 o 1 o 0
    const int row = lane >> 4;
    double ca = 1.0, cb = 0.0;
    if (row == 1) { ca = a15; cb = b15; }
    else if (row == 2) { ca = c2a; cb = c2b; }
    else if (row == 3) { ca = c3a; cb = c3b; }
    if (row > 0) { be = al * cb + be; al = al * ca; }
  }
  double y = lane_prev(be, 0.0);
#pragma unroll
  for (int i = 0; i < NPL; ++i) { y = ali[i] * y + bei[i]; d[i] = y; }
  // backward affine scan x -> ga*x + de
  double ga = 1.0, de = 0.0;
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { de = -cp[i] * de + d[i]; ga = -cp[i] * ga; }
#define OSHUN_AFF_BWD(N)                                                                         \
  {                                                                                              \
    const double qa = dpp_mov<row_shl<N>()>(ga, 1.0), qb = dpp_mov<row_shl<N>()>(de, 0.0);       \
    de = ga * qb + de;                                                                           \
    ga = ga * qa;                                                                                \
  }
  OSHUN_AFF_BWD(1)
  OSHUN_AFF_BWD(2)
  OSHUN_AFF_BWD(4)
  OSHUN_AFF_BWD(8)
#undef OSHUN_AFF_BWD
  {
    const double a16 = lane_read(ga, 16), b16 = lane_read(de, 16);
    const double a32 = lane_read(ga, 32), b32 = lane_read(de, 32);
    const double a48 = lane_read(ga, 48), b48 = lane_read(de, 48);
    const double c1a = a32 * a48, c1b = a32 * b48 + b32;  // rows 2 o 3
    const double c0a = a16 * c1a, c0b = a16 * c1b + b16;  // rows 1 o 2 o 3
    const int row = lane >> 4;
    double ca = 1.0, cb = 0.0;
    if (row == 2) { ca = a48; cb = b48; }
    else if (row == 1) { ca = c1a; cb = c1b; }
    else if (row == 0) { ca = c0a; cb = c0b; }
    if (row < 3) { de = ga * cb + de; ga = ga * ca; }
  }
  double x = lane_next(de, 0.0);
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { x = d[i] - cp[i] * x; d[i] = x; }
}

// One wavefront per (x,y) column (lane holds rows lane*NPL .. lane*NPL+NPL-1),
// COLS adjacent columns per workgroup (a column past the window computes on
// zeros and stores nothing; the host takes the widest COLS that still gives
// a workgroup per CU).  Planes cross HBM in row order (COLS adjacent x of
// one p row per COLS lanes) and are turned to the lane layout through LDS
// (one pad per NPL rows keeps the lane-layout reads free of bank
// conflicts); the next plane's row-order loads are in flight while the
// current one is solved.
template <typename T, int NPL, int COLS>
__global__ __launch_bounds__(64 * COLS) void collide(CollArgs<T> a) {
  constexpr int kCollCols = COLS;
  constexpr int kRows = 64 * NPL;
  // LDS column stride: kRows + kRows / NPL is a multiple of 32 doubles; the
  // extra 8 put the kCollCols columns of a row-order pass on distinct banks
  constexpr int kS = kRows + kRows / NPL + 8;
  __shared__ double tr[kCollCols * kS];
  const int nx = a.nx, np = a.np;
  const int ntx = (a.cnx + kCollCols - 1) / kCollCols;
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int y = tile / ntx + a.cy0;
  const int xw0 = (tile % ntx) * kCollCols;
  const int wv = threadIdx.x >> 6;
  const int lane = lane_id();
  const int lrow = wv * kS + lane * (NPL + 1);  // LDS index of lane-layout row i: lrow + i
  // row-order: thread t moves column cc = t % kCollCols, rows r0 + 64 k
  const int cc = threadIdx.x % kCollCols, r0 = threadIdx.x / kCollCols;
  const bool ccv = xw0 + cc < a.cnx;
  const long long cb = (long long)y * np * nx + (a.cx0 + (ccv ? xw0 + cc : 0));
  auto rok = [&](int k) { return ccv && r0 + 64 * k < np; };
  auto gidx = [&](int k) { return cb + (long long)(r0 + 64 * k) * nx; };
  auto lidx = [&](int k) { const int r = r0 + 64 * k; return cc * kS + r + r / NPL; };
  auto load_rows = [&](const T* pl, T (&v)[NPL]) {
#pragma unroll
    for (int k = 0; k < NPL; ++k) v[k] = rok(k) ? pl[gidx(k)] : T(0);
  };
  // LDS round trips (callers keep the barriers)
  auto rows_to_lds = [&](const T (&v)[NPL]) {
#pragma unroll
    for (int k = 0; k < NPL; ++k) tr[lidx(k)] = static_cast<double>(v[k]);
  };
  auto lds_to_rows_store = [&](T* pl) {
#pragma unroll
    for (int k = 0; k < NPL; ++k)
      if (rok(k)) pl[gidx(k)] = static_cast<T>(tr[lidx(k)]);
  };
  __shared__ double etab[ET_COUNT * kRows];
  for (int q = threadIdx.x; q < kRows; q += 64 * kCollCols) {  // row p = q -> (i, lane) = (q % NPL, q / NPL)
    const bool in = q < np;
    const int o = (q % NPL) * 64 + q / NPL;
    etab[ET_U4 * kRows + o] = in ? a.t.U4[q] : 0.0;
    etab[ET_U4M1 * kRows + o] = in ? a.t.U4m1[q] : 0.0;
    etab[ET_U2 * kRows + o] = in ? a.t.U2[q] : 0.0;
    etab[ET_U2M1 * kRows + o] = in ? a.t.U2m1[q] : 0.0;
    etab[ET_U1S * kRows + o] = in ? a.t.U1s[q] : 0.0;
    etab[ET_U1M1S * kRows + o] = in ? a.t.U1m1s[q] : 0.0;
    etab[ET_U3 * kRows + o] = in ? a.t.U3[q] : 1.0;
    etab[ET_QN * kRows + o] = in ? a.t.Qn[q] : 0.0;
    etab[ET_PN * kRows + o] = in ? a.t.Pn[q] : 0.0;
  }
  const EETab<NPL> tb{etab, lane};
  double f[NPL];
  {
    T v[NPL];
    load_rows(a.f, v);
    rows_to_lds(v);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int p = lane * NPL + i;
      f[i] = (p < np) ? tr[lrow + i] + (a.bg ? a.bg[p] : 0.0) : 0.0;
    }
  }
  if (a.do_ee) {
    // Explicit_EC_FP::loop + RungeKutta4_ee::advance, def-FokkerPlanck.cpp:518-553,590-625
    const double h = a.h_ee;
    for (int s = 0; s < a.numh; ++s) {
      double f0[NPL], f1[NPL], fh[NPL];
#pragma unroll
      for (int i = 0; i < NPL; ++i) { f0[i] = f[i]; f1[i] = f[i]; }
      ee_slope<T, NPL>(a, tb, f1, fh);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { fh[i] *= (0.5 * h); f1[i] += fh[i]; fh[i] *= (1.0 / 3.0); f[i] += fh[i]; }
      ee_slope<T, NPL>(a, tb, f1, fh);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { f1[i] = f0[i]; fh[i] *= (0.5 * h); f1[i] += fh[i]; fh[i] *= (2.0 / 3.0); f[i] += fh[i]; }
      ee_slope<T, NPL>(a, tb, f1, fh);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { fh[i] *= h; f0[i] += fh[i]; fh[i] *= (1.0 / 3.0); f[i] += fh[i]; }
      ee_slope<T, NPL>(a, tb, f0, fh);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { fh[i] *= (h / 6.0); f[i] += fh[i]; }
    }
    __syncthreads();  // every lane has read its rows of f00
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int p = lane * NPL + i;
      tr[lrow + i] = static_cast<double>(static_cast<T>(f[i] - ((a.bg && p < np) ? a.bg[p] : 0.0)));
    }
    __syncthreads();
    lds_to_rows_store(a.f);
    // the implicit step reads f00 as stored
    if (sizeof(T) < sizeof(double)) {
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int p = lane * NPL + i;
        const double b = (a.bg && p < np) ? a.bg[p] : 0.0;
        f[i] = static_cast<double>(static_cast<T>(f[i] - b)) + b;
      }
    }
  }
  if (!(a.do_f1 || a.do_flm)) return;
  // Implicit_FLM_step::reset_coeff, def-FokkerPlanck.cpp:808-939
  double fp[NPL], fn[NPL];
  prev_row<NPL>(f, fp);
  next_row<NPL>(f, fn);
  double t4[NPL], t2[NPL], t1[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    const bool in1 = p >= 1 && p < np;
    t4[i] = in1 ? tb.U4(i) * f[i] + tb.U4m1(i) * fp[i] : 0.0;
    t2[i] = in1 ? tb.U2(i) * f[i] + tb.U2m1(i) * fp[i] : 0.0;
    t1[i] = (p < np - 1) ? tb.U1s(i) * fn[i] + tb.U1m1s(i) * f[i] : 0.0;
  }
  double I2[NPL], I0[NPL], J1m[NPL], tot2, tot0;
  prefix_p<NPL>(t4, I2, tot2);
  prefix_p<NPL>(t2, I0, tot0);
  suffix_p<NPL>(t1, J1m);
  const double I2_temperature = tot2 * (4.0 * M_PI);
  const double I0_density = tot0 * (4.0 * M_PI);
  const double lnee = log_ee(I0_density, I2_temperature, a.density_np);
  const double zlnei = zlog_ei(I0_density, I2_temperature, a.density_np, a.zeta);
  const double zi = zlnei * I0_density;
  const double kd = a.kpre * a.dt_flm;
  const double sc = (-1.0) * lnee * a.kpre * a.dt_flm;
  double Asub[NPL], Bdia[NPL], Csup[NPL], Sct[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    const double v = (p < np) ? a.t.vc[p] : 1.0;
    double i2 = I2[i] * (4.0 * M_PI);
    i2 /= v * v;
    const double i0 = I0[i] * (4.0 * M_PI);
    const double j1 = J1m[i] * (4.0 * M_PI * v);
    const double T1 = i0 + (2.0 * j1 - i2) / 3.0;
    const double T2 = (i2 + j1) / 3.0;
    double s = (p == 0) ? 0.0 : T1 * lnee;
    s += zi;
    s /= tb.U3(i);
    s *= kd;
    Sct[i] = (p < np) ? s : 0.0;
    double dg = 0.0, lo = 0.0, up = 0.0;
    if (p == 0) {
      dg = 8.0 * M_PI * f[i];
    } else if (p < np - 1) {
      const double dm = a.t.ivdm[p], dp = a.t.ivdp[p], s2 = a.t.ivs2[p];
      dg = 8.0 * M_PI * f[i] - T2 * (dm + dp);
      lo = T2 * dm - T1 * s2;
      up = T2 * dp + T1 * s2;
    }
    Asub[i] = lo * sc;
    Bdia[i] = dg * sc;
    Csup[i] = up * sc;
  }
  // Implicit_FLM_step::advance for every l >= 1 plane, def-FokkerPlanck.cpp:943-1035.
  // Planes come in ind(l,m) order (l non-decreasing): the pivots are formed
  // when l changes (f1_loop handles m = 0, 1 only, :1130; flm_loop m <=
  // min(m0, l)); every plane of each l in [lmin, lmax] is solved.
  const int lmin = a.do_f1 ? 1 : 2;
  const int lmax = a.do_flm ? a.l0 : 1;
  int c = 1;
  while (c < a.nc && ldc(a.comp_l, c) < lmin) ++c;
  int cend = c;
  while (cend < a.nc && ldc(a.comp_l, cend) <= lmax) ++cend;
  // a grid of gridDim.y > 1 splits the planes of each column group among
  // that many workgroups, in contiguous runs (each forms the coefficients
  // from f00 as stored, so the host launches it so only without the e-e
  // step); a plane's solve is the same in any workgroup
  if (gridDim.y > 1) {
    const int n = cend - c, g = blockIdx.y, ng = gridDim.y;
    cend = c + (n * (g + 1)) / ng;
    c += (n * g) / ng;
  }
  T vn[NPL];  // next plane, row order
  if (c < cend) load_rows(a.f + (long long)c * a.plane, vn);
  double cpv[NPL], w[NPL];
  int lcur = -1;
  for (; c < cend; ++c) {
    __syncthreads();  // the previous plane (or f00) has left LDS
    rows_to_lds(vn);
    __syncthreads();
    double d[NPL];
#pragma unroll
    for (int i = 0; i < NPL; ++i) d[i] = (lane * NPL + i < np) ? tr[lrow + i] : 0.0;
    if (c + 1 < cend) load_rows(a.f + (long long)(c + 1) * a.plane, vn);  // in flight during the solve
    const int l = ldc(a.comp_l, c);
    if (l != lcur) {
      lcur = l;
      double ll1 = static_cast<double>(l);
      ll1 *= (-0.5) * (ll1 + 1.0);
      double Bl[NPL];
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int p = lane * NPL + i;
        double b = (p == 0 && l > 1) ? 0.0 : Bdia[i];
        b += 1.0 - ll1 * Sct[i];
        Bl[i] = (p < np) ? b : 1.0;
      }
      tri_pivots<NPL>(Asub, Bl, Csup, cpv, w);
    }
    tri_solve<NPL>(Asub, cpv, w, d);
    __syncthreads();  // every lane has read its rows of this plane
#pragma unroll
    for (int i = 0; i < NPL; ++i) tr[lrow + i] = d[i];
    __syncthreads();
    lds_to_rows_store(a.f + (long long)c * a.plane);
  }
}

// ---------------------------------------------------------------------------
// collide_rows: the collision step with SEGW = 16 or 32 lanes per (x,y)
// column (one or two 16-lane DPP rows), 64/SEGW columns per wavefront.  Each
// lane owns NPL consecutive p rows and runs the reference's recurrences over
// them; the lanes of a column are joined by 4-step DPP scans inside each row
// plus, for 32 lanes, one exchange between the two rows (ds_swizzle, lane
// xor 16).  The p tables sit in LDS in an interleaved order (row i of every
// lane is one contiguous line), all p < np masks are folded into the tables
// (zero where the reference's loops skip a row), and divisions by tabulated
// or uniform quantities are multiplications by reciprocals.
constexpr int kRowP = 256;    // LDS table length: SEGW lanes x NPL rows

// DPP move with bound_ctrl: lanes without a source lane read 0 (full row and
// bank masks, so every lane is written and no old value is needed)
template <int CTRL>
__device__ __forceinline__ double dpp_mov0(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

template <int N>
__device__ __forceinline__ double row_bcast(double v) {  // lane N of the 16-lane row, to the whole row
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x150 + N, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x150 + N, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double row_scan(double v) {  // inclusive prefix over the row
  v += dpp_mov0<row_shr<1>()>(v);
  v += dpp_mov0<row_shr<2>()>(v);
  v += dpp_mov0<row_shr<4>()>(v);
  v += dpp_mov0<row_shr<8>()>(v);
  return v;
}
__device__ __forceinline__ double row_scan_rev(double v) {  // inclusive suffix over the row
  v += dpp_mov0<row_shl<1>()>(v);
  v += dpp_mov0<row_shl<2>()>(v);
  v += dpp_mov0<row_shl<4>()>(v);
  v += dpp_mov0<row_shl<8>()>(v);
  return v;
}
// Joining the two 16-lane rows of a 32-lane column (rows 0+1 and 2+3 of the
// wavefront) without LDS: v_permlane16_swap_b32 exchanges rows 1 and 3 of x
// with rows 0 and 2 of y (x' = x with rows 1,3 taken from y's rows 0,2;
// y' = y with rows 0,2 taken from x's rows 1,3), and the DPP row_bcast:15 with
// row mask 0xA moves lane 15 (47) to every lane of row 1 (3).
__device__ __forceinline__ void pl16_swap(double& x, double& y) {
  const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(x), __double2loint(y), false, false);
  const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(x), __double2hiint(y), false, false);
  x = __hiloint2double(static_cast<int>(hi[0]), static_cast<int>(lo[0]));
  y = __hiloint2double(static_cast<int>(hi[1]), static_cast<int>(lo[1]));
}
// lower row of each column: `fill`; upper row: lane 15 of the lower row
__device__ __forceinline__ double bcast15_up(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x142, 0xA, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x142, 0xA, 0xF, false);
  return __hiloint2double(hi, lo);
}
// lower row of each column: lane 0 of the upper row; upper row: `fill`
__device__ __forceinline__ double bcast0_down(double v, double fill) {
  double x = row_bcast<0>(v), y = fill;
  pl16_swap(x, y);
  return y;
}
// whole-wavefront shifts by one lane (lane 0, resp. 63, keeps `fill`)
__device__ __forceinline__ double wave_shr1(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x138, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x138, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_shl1(double v, double fill) {
  const int lo = __builtin_amdgcn_update_dpp(__double2loint(fill), __double2loint(v), 0x130, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(__double2hiint(fill), __double2hiint(v), 0x130, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// column-wide (SEGW lanes) versions of the row primitives
template <int SEGW>
__device__ __forceinline__ double seg_scan(double v) {  // inclusive prefix
  v = row_scan(v);
  if (SEGW == 32) v += bcast15_up(v, 0.0);
  return v;
}
template <int SEGW>
__device__ __forceinline__ double seg_scan_rev(double v) {  // inclusive suffix
  v = row_scan_rev(v);
  if (SEGW == 32) v += bcast0_down(v, 0.0);
  return v;
}
template <int SEGW>
__device__ __forceinline__ double seg_last(double v) {  // the column's last lane, to all its lanes
  double t = row_bcast<15>(v);
  if (SEGW == 16) return t;
  double u = t;
  pl16_swap(u, t);  // t: rows 0,2 from rows 1,3; rows 1,3 kept
  return t;
}
template <int SEGW>
__device__ __forceinline__ double seg_first(double v) {  // the column's lane 0, to all its lanes
  double b = row_bcast<0>(v);
  if (SEGW == 16) return b;
  double u = b;
  pl16_swap(b, u);  // b: rows 1,3 from rows 0,2; rows 0,2 kept
  return b;
}
template <int SEGW>
__device__ __forceinline__ double seg_prev(double v, double fill) {  // lane - 1 (fill at lane 0)
  if (SEGW == 16) return dpp_mov<row_shr<1>()>(v, fill);
  const double r = wave_shr1(v, fill);
  return ((threadIdx.x & 31) == 0) ? fill : r;
}
template <int SEGW>
__device__ __forceinline__ double seg_prev0(double v) {  // lane - 1 (0 at lane 0)
  if (SEGW == 16) return dpp_mov0<row_shr<1>()>(v);
  const double r = dpp_mov0<0x138>(v);
  return ((threadIdx.x & 31) == 0) ? 0.0 : r;
}
template <int SEGW>
__device__ __forceinline__ double seg_next0(double v) {  // lane + 1 (0 at the last lane)
  if (SEGW == 16) return dpp_mov0<row_shl<1>()>(v);
  const double r = dpp_mov0<0x130>(v);
  return ((threadIdx.x & 31) == 31) ? 0.0 : r;
}
template <int SEGW>
__device__ __forceinline__ double seg_next(double v, double fill) {  // lane + 1 (fill at the last lane)
  if (SEGW == 16) return dpp_mov<row_shl<1>()>(v, fill);
  const double r = wave_shl1(v, fill);
  return ((threadIdx.x & 31) == 31) ? fill : r;
}
// LOGee and ZLOGei (def-FokkerPlanck.cpp:41-95) of one column: the two
// 16-lane rows of a 32-lane column evaluate one logarithm each (lower row
// ln T_e, upper row ln n_e) and swap the results, so a wavefront issues one
// vector log instead of two; the clamps are applied by selection.
template <int SEGW>
__device__ __forceinline__ void coulomb_logs(double ne, double Te, double density_np, double zeta, double* lnee,
                                             double* zlnei) {
  if (SEGW != 32) {
    *lnee = log_ee(ne, Te, density_np);
    if (zlnei) *zlnei = zlog_ei(ne, Te, density_np, zeta);
    return;
  }
  double te = Te / (3.0 * ne);
  te *= 511000;
  const double nn = ne * density_np;
  double lt = log(((threadIdx.x & 16) != 0) ? nn : te);
  double ln = lt;
  pl16_swap(lt, ln);  // lt: ln T_e on every lane; ln: ln n_e on every lane
  const double ee = 23.5 - 0.5 * ln + 1.25 * lt - sqrt(0.00001 + 0.0625 * (lt - 2.0) * (lt - 2.0));
  *lnee = (ne > 0.000000001 && ee > 2.0) ? ee : 2.0;
  if (zlnei) {
    const double ei = 24.0 - 0.5 * ln + lt;
    *zlnei = ((ne > 0.0000001 && ei > 2.0) ? ei : 2.0) * zeta;
  }
}

// 1/d by v_rcp_f64 and two Newton steps (within an ulp or so of the correctly rounded quotient)
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = fma(-d, r, 1.0);
  r = fma(r, e, r);
  e = fma(-d, r, 1.0);
  return fma(r, e, r);
}

// LDS/host table rows of collide_rows; element p of row k sits at
// k*kRowP + (p % NPL)*SEGW + p / NPL.
enum {
  RT_U4 = 0,  // 0.5 v^4 dv          (0 at p = 0 and p >= np)
  RT_U4M1,    // 0.5 v_{p-1}^4 dv    (same zeros)
  RT_U2,      // 0.5 v^2 dv
  RT_U2M1,
  RT_U1S,     // U1[p+1]            (0 at p >= np-1)
  RT_U1M1S,   // U1m1[p+1]
  RT_U3,      // v^3                 (1 beyond np)
  RT_PN,      // Pn, 1 at np-1, 0 beyond
  RT_QN,      // Qn, 0 at np-1 and beyond
  RT_VC,      // v                   (1 beyond np)
  RT_IV2,     // 1/v^2               (1 beyond np)
  RT_IU3,     // 1/v^3               (0 beyond np)
  RT_IVDM,    // IvDnDm1             (0 at p = 0 and p >= np-1)
  RT_IVDP,    // IvDnDp1
  RT_IVS2,    // Ivsq2Dn
  RT_ROW,     // 1 for p < np-1, else 0 (Alpha_Tri row np-1 is empty)
  RT_P3,      // v^3 for p < NB (low-p G)
  RT_A2,      // v^3/3
  RT_B2,      // 0.2 v^5/v1^2
  RT_A4,      // 0.2 v^5
  RT_B4,      // v^7/(7 v1^2)
  RT_JW,      // w p^3/gamma with the trapezoid weights w of Current::Jxyz (0 beyond np)
  RT_COUNT
};

template <int SEGW>
__device__ __forceinline__ double rt(const double* tb, int k, int i, int seg) {
  return tb[k * kRowP + i * SEGW + seg];
}

// EE_isotropic_conserva::Getslope (def-FokkerPlanck.cpp:192-267) for the NPL
// rows p = seg*NPL + i of one lane; f in, slope times `scale` out (same array)
template <typename T, int NPL, int SEGW>
__device__ __forceinline__ void ee_slope_rows(const CollArgs<T>& a, const double* tb, int seg, double (&f)[NPL],
                                              double scale) {
  // contracting a*b+c into one rounding is allowed in the collision kernels:
  // their scans already reorder the reference's sums (results agree to ~1e-15)
#pragma clang fp contract(fast)
  // re-read the tables from LDS on every call instead of letting the
  // compiler keep all 9*NPL of them in registers across the subcycles
  __asm__ volatile("" ::: "memory");
  const double fprev = seg_prev0<SEGW>(f[NPL - 1]);
  const double fnext = seg_next0<SEGW>(f[0]);
  // trapezoid prefix sums I4, I2 (from p = 1) and suffix J1, lane-local first
  double I4[NPL], I2[NPL], J1[NPL];
  double s4 = 0.0, s2 = 0.0, s1 = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const double fm = (i > 0) ? f[i - 1] : fprev;
    s4 = fma(rt<SEGW>(tb, RT_U4M1, i, seg), fm, fma(rt<SEGW>(tb, RT_U4, i, seg), f[i], s4));
    s2 = fma(rt<SEGW>(tb, RT_U2M1, i, seg), fm, fma(rt<SEGW>(tb, RT_U2, i, seg), f[i], s2));
    I4[i] = s4;
    I2[i] = s2;
  }
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) {
    const double fn = (i < NPL - 1) ? f[i + 1] : fnext;
    s1 = fma(rt<SEGW>(tb, RT_U1M1S, i, seg), f[i], fma(rt<SEGW>(tb, RT_U1S, i, seg), fn, s1));
    J1[i] = s1;
  }
  const double inc4 = seg_scan<SEGW>(s4), inc2 = seg_scan<SEGW>(s2), inc1 = seg_scan_rev<SEGW>(s1);
  // the lanes before (after) this one: inclusive scan minus the lane's own sum
  const double ex4 = inc4 - s4, ex2 = inc2 - s2, ex1 = inc1 - s1;
  double lnee;
  coulomb_logs<SEGW>(4.0 * M_PI * seg_last<SEGW>(inc2), 4.0 * M_PI * seg_last<SEGW>(inc4), a.density_np, 0.0, &lnee,
                     nullptr);
  const double f0 = seg_first<SEGW>(f[0]);
  const double f1 = seg_first<SEGW>(f[1]);
  const double f00 = (f0 - f1 * a.r01) * a.inv_gden;
  // G on every row first, in the basic block of the Coulomb logarithm so
  // that the two chains interleave; then the parabolic rows (one branch each)
  double G[NPL], J1x[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const double j1 = J1[i] + ex1;
    J1x[i] = j1;
    double i2 = (I2[i] + ex2) * j1;
    i2 *= -3.0;
    double g = (I4[i] + ex4) + rt<SEGW>(tb, RT_U3, i, seg) * j1;
    g *= f[i];
    g += i2;
    G[i] = g;
  }
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    if (i < a.nb) {  // parabolic G for p < NB (EE_isotropic_conserva::G, :178-188): rows of lane 0 only
      const double p3 = tb[RT_P3 * kRowP + i * SEGW];
      const double i2s = f00 * tb[RT_A2 * kRowP + i * SEGW] + (f1 - f00) * tb[RT_B2 * kRowP + i * SEGW];
      const double i4s = f00 * tb[RT_A4 * kRowP + i * SEGW] + (f1 - f00) * tb[RT_B4 * kRowP + i * SEGW];
      const double gl = f[i] * i4s + (p3 * f[i] - 3.0 * i2s) * J1x[i];
      G[i] = (seg == 0) ? gl : G[i];
    }
  }
  // -DG, -DG/(v Dv), DDG/(v Dv), DDG/(v^3 DDv)   (:247-260)
  const double Gnext = seg_next0<SEGW>(G[0]);
  double D2[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const double gn = (i < NPL - 1) ? G[i + 1] : Gnext;
    D2[i] = (G[i] - gn) * rt<SEGW>(tb, RT_PN, i, seg);
  }
  double fh0 = (-1.0) * seg_first<SEGW>(G[0]);
  fh0 *= a.c2v0;
  fh0 -= seg_first<SEGW>(D2[0]);
  fh0 *= tb[RT_QN * kRowP];
  const double D2next = seg_next0<SEGW>(D2[0]);
  double D3[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) D3[i] = D2[i] - ((i < NPL - 1) ? D2[i + 1] : D2next);
  const double D3prev = seg_prev0<SEGW>(D3[NPL - 1]);
  const double sc = (a.c_kpre * lnee) * scale;
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    double v = rt<SEGW>(tb, RT_QN, i, seg) * ((i > 0) ? D3[i - 1] : D3prev);
    if (i == 0) v = (seg == 0) ? fh0 : v;
    f[i] = v * sc;
  }
}

// the current of Y0 at x from its p sum j and f10(p1), as field_body forms it
template <typename Z>
__device__ __forceinline__ double fused_current(const Z& z, double j, double f1) {
  j *= z.delta_p;
  j += z.small * (f1 * z.p3og1);
  j *= 4.0 * M_PI / 3.0;
  return j;
}
// Field stage 4 of the step and field stage 1 of the next one for column x of
// a 1D m0 = 0 run (FieldArgs modes 4 and 1, def-RK.cpp:116-138 with
// Current::Jxyz and MaxwellEq of def-VlasovMax.cpp:680-745,979-1035), the
// same arithmetic as field_body.  In 1D only Ex couples to f and its update is
// local; the stage-1 curl at x needs the stage-4 fields at x +- 1, which the
// launch forms once per position of its columns and the two beside them, from
// the prefetched stage-3 fields and accumulator; the current sums are the
// stage-4 input's (its pair kernels' partials) and the collided f10's.
constexpr int kFuseTiles = 64;  // most stage-pair p tiles a fused collision launch takes (LDS slots per column)
static_assert(kFuseTiles <= 4 * kFieldRows, "fused_stage4_1d sums at most 4 partials per field row");
// row r of sum_direct (p = r, r + kFieldRows, ...) over a plane held in the
// row-order LDS layout of collide_rows (row p at p + p / NPL), as stored in T;
// jw = w p3og with w = 1 or 1/2, so f (w p3og) = w (f p3og) exactly
template <typename T, int NPL, int SEGW>
__device__ __forceinline__ double current_row(const double* trc, const double* jw, int r, int np) {
  // all kRowP / kFieldRows terms are read (the LDS rows exist up to kRowP) and
  // those of p < np added in order, so the loads do not wait on the sum
  double js = 0.0;
#pragma unroll
  for (int k = 0; k < kRowP / kFieldRows; ++k) {
    const int p = r + k * kFieldRows;
    const double t = static_cast<double>(static_cast<T>(trc[p + p / NPL])) * jw[(p % NPL) * SEGW + p / NPL];
    if (p < np) js += t;
  }
  return js;
}
// the fixed tree of sum_parts / sum_direct over kFieldRows row sums
__device__ __forceinline__ double field_tree(double (&sh)[kFieldRows]) {
#pragma unroll
  for (int w = kFieldRows / 2; w >= 1; w /= 2)
#pragma unroll
    for (int r = 0; r < w; ++r) sh[r] += sh[r + w];
  return sh[0];
}
// Stage 4 of the step at one position (its six fields into F): it needs only
// the stage inputs, so it runs before the collisions.  cf: the stage-3 fields
// at the position -1 .. +1 (component c at cf[c * SC]), af: the accumulator at
// the position (af[c * SA]), both read from LDS where used.  jp3 (the
// position's current partials) is null at the positions beside a column, whose
// Ex the next stage does not read; at a column's own x the Ex slope takes the
// current first and the fields are stored at x.
template <typename T, int SC, int SA>
__device__ void fused_stage4_1d(const CollArgs<T>& a, int x, const double* cf, const double* af, const double* jp3,
                                double f1y3, double (&F)[6]) {
  const auto& z = a.fz;
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (jp3) {
    double sh[kFieldRows];
#pragma unroll
    for (int r = 0; r < kFieldRows; ++r) {  // sum_parts with npb <= 4 kFieldRows partials
      double v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = (r + kFieldRows * q < z.npt) ? jp3[r + kFieldRows * q] : 0.0;
      double s = 0.0;
      if (r < z.npt) s += (v[0] + v[1]) + (v[2] + v[3]);
      sh[r] = s;
    }
    k[0] += fused_current(z, field_tree(sh), f1y3);
  }
  maxwell_terms([&](int c, int at) { return cf[c * SC + (at == 0 ? 0 : at == 1 ? 2 : 1)]; }, z.midx, z.midy, k);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double kh = k[c] * z.h6;
    F[c] = af[c * SA] + kh;
  }
  if (jp3) {
#pragma unroll
    for (int c = 0; c < 6; ++c) z.sf[(long long)c * a.nx + x] = F[c];
  }
}
// The next step's stage 1 at x from the stage-4 fields F4 at x-1, x, x+1 and
// the current of the collided f10 (its kFieldRows row sums and f10(p1)).
template <typename T>
__device__ void fused_stage1_1d(const CollArgs<T>& a, int x, const double (&F4)[3][6], const double* jrows0,
                                double f10p1) {
  const auto& z = a.fz;
  const int nx = a.nx;
  double sh[kFieldRows];
#pragma unroll
  for (int r = 0; r < kFieldRows; ++r) sh[r] = jrows0[r];
  const double j0 = field_tree(sh);
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  k[0] += fused_current(z, j0, f10p1);
  maxwell_terms([&](int c, int at) { return F4[at == 0 ? 0 : at == 1 ? 2 : 1][c]; }, z.midx, z.midy, k);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double kh = k[c] * z.c05h;
    z.e1[(long long)c * nx + x] = F4[1][c] + kh;
    z.af2[(long long)c * nx + x] = F4[1][c] + kh * z.third;
  }
}

// MODE 1: e-e subcycles on f_00 (Explicit_EC_FP::loop); MODE 2: implicit
// l >= 1 step (Euler_Backward::advance_1/advance_lm); MODE 3: both in one
// launch (the implicit step continues from the collided f00 in registers).
template <typename T, int NPL, int SEGW, int MODE>
__device__ __forceinline__ void collide_body(const CollArgs<T>& a, const int tile) {
#pragma clang fp contract(fast)
  constexpr int kCols = 256 / SEGW;  // columns per workgroup
  // Planes cross HBM in row order (a wavefront moves 64/kCols rows of kCols
  // adjacent x) and are turned to the lane layout (lane seg of column w holds
  // rows seg*NPL .. seg*NPL+NPL-1) through LDS.  One pad per NPL rows keeps
  // the lane-layout reads free of bank conflicts (16 lanes read 16 distinct
  // 8-byte bank pairs); the column stride kS does the same for the row-order
  // accesses, whose 16 lanes cover kCols columns x 16/kCols rows: kS = 16/kCols
  // modulo 16 puts column cc at pair 2cc (kCols = 8) or cc (kCols = 16).
  constexpr int kCols0 = 256 / SEGW;
  constexpr int kSmin = SEGW * NPL + SEGW;
  constexpr int kSres = (16 / kCols0) % 16;
  constexpr int kS = kSmin + ((kSres - kSmin % 16) + 16) % 16;
  static_assert(kCols0 == 8 || kCols0 == 16, "row-order lane mapping assumes 8 or 16 columns per workgroup");
  __shared__ double tb[RT_COUNT * kRowP];
  __shared__ double tr[kCols * kS];
  __shared__ double sbg[kRowP];  // f00 background per p (0 beyond np or without one), read in the lane layout
  // Every global load of the prologue (tables, background, fused-field
  // inputs, f00) is issued before the first LDS write: the workgroup then
  // waits for one memory latency instead of one per fill-loop iteration.
  // (blockDim.x == 256 == kRowP)
  constexpr int kTbPer = RT_COUNT * kRowP / 256;
  static_assert(RT_COUNT * kRowP % 256 == 0 && kRowP == 256, "one table element per thread and pass");
  double tbv[kTbPer];
#pragma unroll
  for (int k = 0; k < kTbPer; ++k) tbv[k] = a.rt[threadIdx.x + 256 * k];
  const double bgv = (a.bg && static_cast<int>(threadIdx.x) < a.np) ? a.bg[threadIdx.x] : 0.0;
  const int nx = a.nx, np = a.np;
  const int ntx = (a.cnx + kCols - 1) / kCols;
  const int y = tile / ntx + a.cy0;
  const int x0 = (tile % ntx) * kCols + a.cx0;
  // fused field stages: their inputs are fetched now, while the collisions run
  // (stage-3 fields at x0-2 .. x0+kCols+1, accumulator at x0-1 .. x0+kCols,
  // f10(p1) and the current partials of the stage-4 input at the columns)
  constexpr int kZcf = kCols + 4, kZaf = kCols + 2, kZpt = kFuseTiles;  // the host fuses only npt <= kFuseTiles
  constexpr int kZoAf = 6 * kZcf, kZoF1 = kZoAf + 6 * kZaf, kZoJp = kZoF1 + kCols, kZn = kZoJp + kCols * kZpt;
  __shared__ double zpre[kZn];
  constexpr int kZper = (kZn + 255) / 256;
  double zv[kZper];
  int zi[kZper];
  if (a.fuse) {
    const int npt = a.fz.npt;
    auto wrap = [nx](int xx) { return ((xx % nx) + nx) % nx; };
#pragma unroll
    for (int k = 0; k < kZper; ++k) {
      const int i = threadIdx.x + 256 * k;
      zi[k] = -1;
      zv[k] = 0.0;
      if (i < kZoJp) {
        zi[k] = i;
        if (i < kZoAf) zv[k] = a.fz.cf[(long long)(i / kZcf) * nx + wrap(x0 - 2 + i % kZcf)];
        else if (i < kZoF1) zv[k] = a.fz.af[(long long)((i - kZoAf) / kZaf) * nx + wrap(x0 - 1 + (i - kZoAf) % kZaf)];
        else zv[k] = a.fz.f1y[wrap(x0 + i - kZoF1)];
      } else if (i < kZoJp + kCols * npt) {
        const int j = i - kZoJp, col = j / npt, b = j % npt;
        zi[k] = kZoJp + col * kZpt + b;
        zv[k] = a.fz.jp[((long long)b * 3) * nx + wrap(x0 + col)];
      }
    }
  }
  __shared__ double zj[kCols][kFieldRows], zf1[kCols];  // the collided f10's current rows and f10(p1)
  __shared__ double zf4[kCols + 2][6];                   // stage-4 fields at x0-1 .. x0+kCols
  const int seg = threadIdx.x & (SEGW - 1);
  const int p0 = seg * NPL;
  const int lrow = (threadIdx.x / SEGW) * kS + p0 + seg;  // LDS index of lane-layout row i: lrow + i
  // row-order layout: thread t moves column cc = t % kCols, rows r0 + k*SEGW (k < NPL)
  const int cc = threadIdx.x % kCols, r0 = threadIdx.x / kCols;
  const bool ccv = x0 + cc < a.cx0 + a.cnx;
  const int crow = cc * kS + r0 + r0 / NPL;  // LDS index of row-order element k: crow + k*(SEGW + SEGW/NPL)
  constexpr int kStepL = SEGW + SEGW / NPL;
  // byte offsets within a plane; rows past np and idle columns read 0 / drop their stores
  constexpr int kOut = 0x7ffffff0;
  const int cofs = static_cast<int>(((long long)y * np * nx + (ccv ? x0 + cc : 0) + (long long)r0 * nx) * sizeof(T));
  const int crowb = SEGW * nx * static_cast<int>(sizeof(T));
  auto ld_ofs = [&](int k) { return (r0 + k * SEGW < np) ? cofs + k * crowb : kOut; };
  auto st_ofs = [&](int k) { return (ccv && r0 + k * SEGW < np) ? cofs + k * crowb : kOut; };
  const unsigned pbytes = static_cast<unsigned>(a.plane * sizeof(T));
  int numh = a.numh;
  double h_ee = a.h_ee, dt_flm = a.dt_flm;
  if (a.tclk) {  // batched runs: the sub-steps of Explicit_EC_FP::loop from the run's clock (:597-599)
    const double t = *a.tclk, tnew = t + a.hstep;
    numh = static_cast<int>((tnew - t) / a.small_dt) + 1;
    h_ee = (tnew - t) / static_cast<double>(numh);
    dt_flm = tnew - t;
  }
  double f[NPL];
  {
    const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.f, pbytes);
    T v[NPL];
#pragma unroll
    for (int k = 0; k < NPL; ++k) v[k] = bload<T>(r, ld_ofs(k), 0);
#pragma unroll
    for (int k = 0; k < kTbPer; ++k) tb[threadIdx.x + 256 * k] = tbv[k];
    sbg[threadIdx.x] = bgv;
    if (a.fuse) {
#pragma unroll
      for (int k = 0; k < kZper; ++k)
        if (zi[k] >= 0) zpre[zi[k]] = zv[k];
      __syncthreads();  // the fused stages' inputs are in LDS
      // stage 4 of the step while f00 is still in flight (it reads no state),
      // once per position x0-1 .. x0+kCols: lane 0 of each column takes its
      // own x, lane 1 of the first and last columns the positions beside them
      const int col = threadIdx.x / SEGW;
      const int j = seg == 0 ? col + 1 : (seg == 1 && col == 0) ? 0 : (seg == 1 && col == kCols - 1) ? kCols + 1 : -1;
      if (j >= 0) {
        const bool own = seg == 0 && x0 + col < nx;
        double F[6];
        fused_stage4_1d<T, kZcf, kZaf>(a, x0 + col, &zpre[j], &zpre[kZoAf + j],
                                       own ? &zpre[kZoJp + col * kZpt] : nullptr, zpre[kZoF1 + col], F);
#pragma unroll
        for (int c6 = 0; c6 < 6; ++c6) zf4[j][c6] = F[c6];
      }
    }
#pragma unroll
    for (int k = 0; k < NPL; ++k) tr[crow + k * kStepL] = static_cast<double>(v[k]);  // rows >= np load 0
    __syncthreads();  // also covers the table fill
#pragma unroll
    for (int i = 0; i < NPL; ++i) f[i] = tr[lrow + i] + sbg[p0 + i];
  }
  // the implicit step's planes (l in [lmin, lmax], ind(l,m) order); the first
  // one's row-order loads are issued now, so that they are in flight during
  // the e-e sub-steps (MODE 3) or the coefficients (MODE 2)
  const int lmin = a.do_f1 ? 1 : 2;
  const int lmax = a.do_flm ? a.l0 : 1;
  int c = 1, cend = 1;
  T vn[NPL];  // next plane, row order
  if (MODE != 1) {
    while (c < a.nc && ldc(a.comp_l, c) < lmin) ++c;
    cend = c;
    while (cend < a.nc && ldc(a.comp_l, cend) <= lmax) ++cend;
    if (c < cend) {
      const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.f + (long long)c * a.plane, pbytes);
#pragma unroll
      for (int k = 0; k < NPL; ++k) vn[k] = bload<T>(r, ld_ofs(k), 0);
    }
  }
  if (MODE == 1 || MODE == 3) {
    // RungeKutta4_ee::advance, def-FokkerPlanck.cpp:518-553, numh times (:590-625)
    const double h = h_ee;
    for (int s = 0; s < numh; ++s) {
      // f: running sum, f0: start of the subcycle, fs: stage input, overwritten by its slope
      double f0[NPL], fs[NPL];
#pragma unroll
      for (int i = 0; i < NPL; ++i) { f0[i] = f[i]; fs[i] = f[i]; }
      // each slope comes back multiplied by its RK4 factor h/2, h/2, h, h/6
      ee_slope_rows<T, NPL, SEGW>(a, tb, seg, fs, 0.5 * h);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { const double hs = fs[i]; fs[i] += f0[i]; f[i] += hs * (1.0 / 3.0); }
      ee_slope_rows<T, NPL, SEGW>(a, tb, seg, fs, 0.5 * h);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { const double hs = fs[i]; fs[i] += f0[i]; f[i] += hs * (2.0 / 3.0); }
      ee_slope_rows<T, NPL, SEGW>(a, tb, seg, fs, h);
#pragma unroll
      for (int i = 0; i < NPL; ++i) { const double hs = fs[i]; fs[i] += f0[i]; f[i] += hs * (1.0 / 3.0); }
      ee_slope_rows<T, NPL, SEGW>(a, tb, seg, fs, h / 6.0);
#pragma unroll
      for (int i = 0; i < NPL; ++i) f[i] += fs[i];
    }
    __syncthreads();  // every lane has read its rows of f00
#pragma unroll
    for (int i = 0; i < NPL; ++i) tr[lrow + i] = f[i] - sbg[p0 + i];
    __syncthreads();
    {
      const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.f, pbytes);
#pragma unroll
      for (int k = 0; k < NPL; ++k) bstore<T>(static_cast<T>(tr[crow + k * kStepL]), r, st_ofs(k), 0);
    }
    if (MODE == 1) return;
    // MODE 3 goes on with the implicit step, which reads f00 as stored
    if (sizeof(T) < sizeof(double)) {
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const double b = sbg[p0 + i];
        f[i] = (p0 + i < np) ? static_cast<double>(static_cast<T>(f[i] - b)) + b : 0.0;
      }
    }
  }
  // Implicit_FLM_step::reset_coeff, def-FokkerPlanck.cpp:808-939
  double Asub[NPL], Bdia[NPL], Csup[NPL], Sct[NPL];
  {
    const double fprev = seg_prev0<SEGW>(f[NPL - 1]);
    const double fnext = seg_next0<SEGW>(f[0]);
    double I2[NPL], I0[NPL], J1[NPL];
    double s4 = 0.0, s2 = 0.0, s1 = 0.0;
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const double fm = (i > 0) ? f[i - 1] : fprev;
      s4 = fma(rt<SEGW>(tb, RT_U4M1, i, seg), fm, fma(rt<SEGW>(tb, RT_U4, i, seg), f[i], s4));
      s2 = fma(rt<SEGW>(tb, RT_U2M1, i, seg), fm, fma(rt<SEGW>(tb, RT_U2, i, seg), f[i], s2));
      I2[i] = s4;
      I0[i] = s2;
    }
#pragma unroll
    for (int i = NPL - 1; i >= 0; --i) {
      const double fn = (i < NPL - 1) ? f[i + 1] : fnext;
      s1 = fma(rt<SEGW>(tb, RT_U1M1S, i, seg), f[i], fma(rt<SEGW>(tb, RT_U1S, i, seg), fn, s1));
      J1[i] = s1;
    }
    const double inc4 = seg_scan<SEGW>(s4), inc2 = seg_scan<SEGW>(s2), inc1 = seg_scan_rev<SEGW>(s1);
    const double ex4 = inc4 - s4, ex2 = inc2 - s2, ex1 = inc1 - s1;
    const double I2_temperature = seg_last<SEGW>(inc4) * (4.0 * M_PI);
    const double I0_density = seg_last<SEGW>(inc2) * (4.0 * M_PI);
    double lnee, zlnei;
    coulomb_logs<SEGW>(I0_density, I2_temperature, a.density_np, a.zeta, &lnee, &zlnei);
    const double zi = zlnei * I0_density;
    const double kd = a.kpre * dt_flm;
    const double sc = (-1.0) * lnee * a.kpre * dt_flm;
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      double i2 = (I2[i] + ex4) * (4.0 * M_PI);
      i2 *= rt<SEGW>(tb, RT_IV2, i, seg);
      const double i0 = (I0[i] + ex2) * (4.0 * M_PI);
      const double j1 = (J1[i] + ex1) * (4.0 * M_PI * rt<SEGW>(tb, RT_VC, i, seg));
      const double T1 = i0 + (2.0 * j1 - i2) * (1.0 / 3.0);
      const double T2 = (i2 + j1) * (1.0 / 3.0);
      double sct = T1 * lnee;
      if (i == 0) sct = (seg == 0) ? 0.0 : sct;  // Scattering_Term[0] has no e-e part (:874)
      sct += zi;
      sct *= rt<SEGW>(tb, RT_IU3, i, seg);
      Sct[i] = sct * kd;
      const double dm = rt<SEGW>(tb, RT_IVDM, i, seg), dpp = rt<SEGW>(tb, RT_IVDP, i, seg);
      const double s2v = rt<SEGW>(tb, RT_IVS2, i, seg);
      const double dg = 8.0 * M_PI * f[i] * rt<SEGW>(tb, RT_ROW, i, seg) - T2 * (dm + dpp);
      Asub[i] = (T2 * dm - T1 * s2v) * sc;
      Bdia[i] = dg * sc;
      Csup[i] = (T2 * dpp + T1 * s2v) * sc;
    }
  }
  // Implicit_FLM_step::advance for every l >= 1 plane, def-FokkerPlanck.cpp:943-1035,
  // TridiagonalSolve def-Nmethods.cpp:142-167 with the lane boundaries joined by
  // scans.  Planes come in ind(l,m) order (l non-decreasing): the pivots are
  // formed when l changes, and the next plane's row-order loads are in flight
  // while the current plane is solved.
  double w[NPL];
  int lcur = -1;
  for (; c < cend; ++c) {
    double d[NPL];
    __syncthreads();  // the previous plane's results have left LDS (and f00 in MODE 3)
#pragma unroll
    for (int k = 0; k < NPL; ++k) tr[crow + k * kStepL] = static_cast<double>(vn[k]);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NPL; ++i) d[i] = tr[lrow + i];
    if (c + 1 < cend) {  // next plane in flight while this one is solved
      const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.f + (long long)(c + 1) * a.plane, pbytes);
#pragma unroll
      for (int k = 0; k < NPL; ++k) vn[k] = bload<T>(r, ld_ofs(k), 0);
    }
    const int l = ldc(a.comp_l, c);  // scalar load: a vector load's wait would also drain the prefetch
    if (l != lcur) {
      lcur = l;
      double ll1 = static_cast<double>(l);
      ll1 *= (-0.5) * (ll1 + 1.0);
      double bl[NPL];  // diagonal of this l
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        double bd = Bdia[i];
        if (i == 0 && l > 1) bd = (seg == 0) ? 0.0 : bd;  // row 0 of Alpha_Tri dropped for l > 1 (:953-955)
        bl[i] = bd + (1.0 - ll1 * Sct[i]);
      }
      // pivots: c'_i = P_i/Q_i with (P,Q)_i = M_i (P,Q)_{i-1}, M_i = [[0, c_i], [-a_i, b_i]]
      Mob P = {1.0, 0.0, 0.0, 1.0};
#pragma unroll
      for (int i = 0; i < NPL; ++i) {  // P <- [[0, c], [-a, b]] P
        const Mob Q = {Csup[i] * P.c, Csup[i] * P.d, fma(bl[i], P.c, -Asub[i] * P.a), fma(bl[i], P.d, -Asub[i] * P.b)};
        P = Q;
        if (i == NPL - 1) P = mob_norm(P);  // exact power-of-two rescaling against overflow (NPL <= 8 rows)
      }
#define OSHUN_RMOB(N)                                                                    \
      {                                                                                    \
        Mob q;                                                                             \
        q.a = dpp_mov<row_shr<N>()>(P.a, 1.0); q.b = dpp_mov0<row_shr<N>()>(P.b);      \
        q.c = dpp_mov0<row_shr<N>()>(P.c); q.d = dpp_mov<row_shr<N>()>(P.d, 1.0);      \
        P = mob_mul(P, q);  /* normalised lane products: no range issue over 5 steps */    \
      }
      OSHUN_RMOB(1)
      OSHUN_RMOB(2)
      OSHUN_RMOB(4)
      OSHUN_RMOB(8)
#undef OSHUN_RMOB
      if (SEGW == 32) {  // the upper row continues from the lower row's product (identity below)
        Mob q;
        q.a = bcast15_up(P.a, 1.0); q.b = bcast15_up(P.b, 0.0);
        q.c = bcast15_up(P.c, 0.0); q.d = bcast15_up(P.d, 1.0);
        P = mob_mul(P, q);
      }
      // The state entering the lane is the exclusive prefix applied to (0, 1);
      // then (P, Q)_i = M_i (P, Q)_{i-1} row by row, and the reference's pivot
      // 1 / (b_i - a_i c'_{i-1}) is Q_{i-1} / Q_i: the recurrence carries two
      // products per row and the divisions hang off it instead of lying on it.
      double Pv = seg_prev0<SEGW>(P.b), Qv = seg_prev<SEGW>(P.d, 1.0);
      double Qin[NPL];
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        Qin[i] = Qv;
        const double Pn = Csup[i] * Qv;
        const double Qn = bl[i] * Qv - Asub[i] * Pv;
        w[i] = Qn;
        Pv = Pn;
        Qv = Qn;
        if (i == NPL / 2 - 1) {  // exact power-of-two rescaling halfway (the ratios are unchanged)
          int e;
          frexp(Qv, &e);
          Pv = ldexp(Pv, -e);
          Qv = ldexp(Qv, -e);
        }
      }
#pragma unroll
      for (int i = 0; i < NPL; ++i) w[i] = Qin[i] * rcp_nr(w[i]);
    }
    // forward substitution d'_i = (d_i - a_i d'_{i-1}) w_i: affine map per lane, scanned
    double al = 1.0, be = 0.0;
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const double aw = -Asub[i] * w[i];
      be = aw * be + w[i] * d[i];
      al = aw * al;
    }
#define OSHUN_RAFF(N)                                                                              \
    {                                                                                            \
      const double qa = dpp_mov<row_shr<N>()>(al, 1.0), qb = dpp_mov0<row_shr<N>()>(be);     \
      be = al * qb + be;                                                                         \
      al = al * qa;                                                                              \
    }
    OSHUN_RAFF(1)
    OSHUN_RAFF(2)
    OSHUN_RAFF(4)
    OSHUN_RAFF(8)
#undef OSHUN_RAFF
    if (SEGW == 32) {  // identity map on the lower row
      const double qa = bcast15_up(al, 1.0), qb = bcast15_up(be, 0.0);
      be = al * qb + be;
      al = al * qa;
    }
    double yv = seg_prev0<SEGW>(be);
#pragma unroll
    for (int i = 0; i < NPL; ++i) { yv = (d[i] - yv * Asub[i]) * w[i]; d[i] = yv; }
    // back substitution x_i = d'_i - c'_i x_{i+1}: affine map per lane, reverse scan
    double ga = 1.0, de = 0.0;
#pragma unroll
    for (int i = NPL - 1; i >= 0; --i) {
      const double cp = Csup[i] * w[i];
      de = -cp * de + d[i];
      ga = -cp * ga;
    }
#define OSHUN_RAFB(N)                                                                              \
    {                                                                                            \
      const double qa = dpp_mov<row_shl<N>()>(ga, 1.0), qb = dpp_mov0<row_shl<N>()>(de);     \
      de = ga * qb + de;                                                                         \
      ga = ga * qa;                                                                              \
    }
    OSHUN_RAFB(1)
    OSHUN_RAFB(2)
    OSHUN_RAFB(4)
    OSHUN_RAFB(8)
#undef OSHUN_RAFB
    if (SEGW == 32) {  // identity map on the upper row
      const double qa = bcast0_down(ga, 1.0), qb = bcast0_down(de, 0.0);
      de = ga * qb + de;
      ga = ga * qa;
    }
    double xv = seg_next0<SEGW>(de);
#pragma unroll
    for (int i = NPL - 1; i >= 0; --i) { xv = d[i] - (Csup[i] * w[i]) * xv; d[i] = xv; }
    __syncthreads();  // every lane has read its rows of this plane
#pragma unroll
    for (int i = 0; i < NPL; ++i) tr[lrow + i] = d[i];
    __syncthreads();
    {
      const __amdgpu_buffer_rsrc_t r = buf_rsrc(a.f + (long long)c * a.plane, pbytes);
#pragma unroll
      for (int k = 0; k < NPL; ++k) bstore<T>(static_cast<T>(tr[crow + k * kStepL]), r, st_ofs(k), 0);
    }
    if (a.fuse && a.fz.next1 && c == a.fz.c10 && seg < kFieldRows) {
      // the next step's Y0 current from f10 as stored, summed like sum_direct:
      // row r of the field block takes p = r, r + kFieldRows, ... (then the tree)
      const int col = threadIdx.x / SEGW;
      zj[col][seg] = current_row<T, NPL, SEGW>(&tr[col * kS], &tb[RT_JW * kRowP], seg, np);
      if (seg == 0) zf1[col] = static_cast<double>(static_cast<T>(tr[col * kS + 1]));
    }
  }
  if (a.fuse && a.fz.next1 && seg == 0 && x0 + static_cast<int>(threadIdx.x) / SEGW < nx) {
    const int col = threadIdx.x / SEGW;
    double F4[3][6];  // stage-4 fields at x-1, x, x+1 (positions col .. col+2 of zf4)
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
      for (int c6 = 0; c6 < 6; ++c6) F4[d][c6] = zf4[col + d][c6];
    fused_stage1_1d<T>(a, x0 + col, F4, zj[col], zf1[col]);
  }
}

template <typename T, int NPL, int SEGW, int MODE>
__global__ __launch_bounds__(64 * 4) void collide_rows(CollArgs<T> a) {
  collide_body<T, NPL, SEGW, MODE>(a, xcd_tile(blockIdx.x, gridDim.x));
}
template <typename T, int NPL, int SEGW, int MODE>
__global__ __launch_bounds__(64 * 4) void collide_rows_batch(const CollArgs<T>* __restrict__ args,
                                                             const int* __restrict__ ts, int nruns) {
  const int g = xcd_tile(blockIdx.x, gridDim.x);
  if (g >= ts[nruns]) return;  // grid padding
  const int r = batch_run(ts, nruns, g);
  collide_body<T, NPL, SEGW, MODE>(args[r], g - ts[r]);
}

#ifndef OSHUN_FAMILY_TU  // non-template kernels: defined once, in the library's main unit
// batched runs: every run's clock advances by its step after the collisions
__global__ __launch_bounds__(256) void batch_clock(double* __restrict__ t, const double* __restrict__ h, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) t[i] += h[i];
}
#endif

// Nonrelativistic moments, Parallel_Output::nonrelativistic_moments
// (def-Parallel.cpp:1677-2253) in double, in three launches:
//   moments_partial  the ten trapezoid sums over p, split into p chunks
//                    (one thread per (x, chunk, y), coalesced along x) into
//                    part[chunk][sum][y][x];
//   moments_final    per (x,y): the chunks summed in order, the p0 terms, and
//                    the 16 pointwise moments;
//   moments_cond     per (x,y): Kx and Ky, which difference <v^2>^2 over x, y.
// out[k][y][x], k = density, Vx, Vy, VxVx, VyVy, VzVz, Qx, VxVy, VxVz, VyVz,
// Vsq, T[eV], P[Mbar], N_D, nu, Qy, Kx, Ky.
constexpr int kMomSums = 10;
constexpr int kMomOut = 18;
constexpr int kMomChunkMin = 16;  // p rows per chunk at least
constexpr int kMomChunkMax = 32;  // chunks at most

template <typename T>
struct MomArgs {
  const T* f;
  double* part;  // [nchunk][kMomSums][ny][nx]
  double* out;   // [kMomOut][ny][nx]
  const double *w2, *w3, *w4, *w5;
  const double* bg;  // f00 background per p added to plane 0, or null
  int np, nx, ny, nchunk, plen;
  // real planes of f00, f10, Re f11, f20, Re f21, Im f21, Im f22 (-1: not in the run)
  int c00, c10, c11r, c20, c21r, c21i, c22i;
  long long plane;
  double dp, r, inv, c5, c7, c00k, c04k, density_np, zeta, dx, dy;
};

template <typename T>
__global__ __launch_bounds__(64) void moments_partial(MomArgs<T> a) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  if (x >= a.nx) return;
  const int ch = blockIdx.y, y = blockIdx.z;
  const int nx = a.nx, np = a.np;
  const int p0 = ch * a.plen, p1 = min(np, p0 + a.plen);
  const long long base = (long long)y * np * nx + x;
  auto row = [&](int c, int p) -> double {
    return c < 0 ? 0.0 : static_cast<double>(a.f[(long long)c * a.plane + base + (long long)p * nx]);
  };
  double s[kMomSums];
#pragma unroll
  for (int k = 0; k < kMomSums; ++k) s[k] = 0.0;
  for (int p = p0; p < p1; ++p) {
    const double we = (p == 0 || p == np - 1) ? 0.5 : 1.0;  // trapezoid end weights
    const double wi = p == 0 ? 0.0 : we;                     // sums that start at p1 (f20(p0) = 0)
    const double f00 = row(a.c00, p) + (a.bg ? a.bg[p] : 0.0), f10 = row(a.c10, p), f11 = row(a.c11r, p);
    const double W2 = a.w2[p] * we, W3 = a.w3[p] * we, W4 = a.w4[p] * we, W4i = a.w4[p] * wi, W5 = a.w5[p] * we;
    s[0] += W2 * f00;
    s[1] += W3 * f10;
    s[2] += W3 * f11;
    s[3] += W4 * f00;
    s[4] += W4i * row(a.c20, p);
    s[5] += W4i * row(a.c21r, p);
    s[6] += W4i * row(a.c21i, p);
    s[7] += W4i * row(a.c22i, p);
    s[8] += W5 * f10;
    s[9] += W5 * f11;
  }
  const long long nxy = (long long)nx * a.ny, o = (long long)y * nx + x;
#pragma unroll
  for (int k = 0; k < kMomSums; ++k) a.part[((long long)ch * kMomSums + k) * nxy + o] = s[k];
}

template <typename T>
__global__ __launch_bounds__(256) void moments_final(MomArgs<T> a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nxy = a.nx * a.ny;
  if (i >= nxy) return;
  const int x = i % a.nx, y = i / a.nx;
  double s[kMomSums];
#pragma unroll
  for (int k = 0; k < kMomSums; ++k) {
    double v = 0.0;
    for (int ch = 0; ch < a.nchunk; ++ch) v += a.part[((long long)ch * kMomSums + k) * nxy + i];
    s[k] = v;
  }
  const long long base = (long long)y * a.np * a.nx + x;
  auto at = [&](int c, int p) -> double {
    return c < 0 ? 0.0 : static_cast<double>(a.f[(long long)c * a.plane + base + (long long)p * a.nx]);
  };
  const double f0_0 = at(a.c00, 0) + (a.bg ? a.bg[0] : 0.0);
  const double f0_1 = at(a.c00, 1) + (a.bg ? a.bg[1] : 0.0);
  // density, with the p0 extrapolation term (:1712-1717) and the FLT_EPSILON floor (:1722-1726)
  double n = s[0];
  double c = f0_1 / 5.0 * a.r + (f0_0 - f0_1 * a.r) * a.inv * (1.0 / 3.0 - 1.0 / 5.0 * a.r);
  c *= a.c00k;
  n += c;
  n *= 4.0 * M_PI * a.dp;
  if (n < 4.0 * 1.1920928955078125e-07) n = 1.0;
  double vx = s[1] + at(a.c10, 1) * a.c5;
  vx *= (4.0 * M_PI / 3.0) * a.dp;
  vx /= n;
  double vy = 0.0;
  if (a.c11r >= 0) {
    vy = s[2] + at(a.c11r, 1) * a.c5;
    vy *= (8.0 * M_PI / 3.0) * a.dp;
    vy /= n;
  }
  double m1 = s[3];
  double c1 = f0_1 / 7.0 * a.r + (f0_0 - f0_1 * a.r) * a.inv * (1.0 / 5.0 - 1.0 / 7.0 * a.r);
  c1 *= a.c04k;
  m1 += c1;
  m1 *= 4.0 * M_PI / 3.0 * a.dp;
  double m2 = s[4] * (8.0 * M_PI / 15.0 * a.dp);
  const double vxvx = (m1 + m2) / n;
  m2 *= 0.5;
  const double m1b = m1 - m2;
  const double t22 = s[4] * (16.0 * M_PI / 5.0 * a.dp);  // f20 as in def-Parallel.cpp:1875-1886
  const double vyvy = (t22 + m1b) / n;
  const double vzvz = (m1b - t22) / n;
  const double vxvy = s[5] * (8.0 * M_PI / 5.0 * a.dp) / n;
  const double vxvz = s[6] * (-(8.0 * M_PI / 5.0) * a.dp) / n;
  const double vyvz = s[7] * (-(16.0 * M_PI / 5.0) * a.dp) / n;
  double vsq = 0.0;
  vsq += vxvx; vsq += vyvy; vsq += vzvz;
  vsq -= vx * vx; vsq -= vy * vy;
  const double teV = vsq * 170880.776;
  const double pre = vsq * (2.73785179e-19 * a.density_np) * n;
  const double nd = 1.72e+9 * sqrt((teV * teV * teV) / (a.density_np * n));
  double lnei = 24.0 - 0.5 * log(a.density_np * n) + log(teV);
  lnei *= a.zeta;
  const double nu = sqrt(2.0 / M_PI) / 9.0 * lnei / nd;
  double q = 0.0;
  q -= vx * (vxvx + vyvy + vzvz);
  q -= 2.0 * (vxvx * vx + vxvy * vy);
  q += 2.0 * (vx * vx + vy * vy) * vx;
  q *= n;
  double m5 = s[8] + at(a.c10, 1) * a.c7;
  m5 *= (4.0 * M_PI / 3.0) * a.dp;
  q += m5;
  q *= 0.5;
  double qy = 0.0;
  qy -= vy * (vxvx + vyvy + vzvz);
  qy -= 2.0 * (vxvy * vx + vyvy * vy);
  qy += 2.0 * (vx * vx + vy * vy) * vy;
  qy *= n;
  double m5y = a.c11r >= 0 ? s[9] + at(a.c11r, 1) * a.c7 : 0.0;
  m5y *= (8.0 * M_PI / 3.0) * a.dp;
  qy += m5y;
  qy *= 0.5;
  const double vals[16] = {n, vx, vy, vxvx, vyvy, vzvz, q, vxvy, vxvz, vyvz, vsq, teV, pre, nd, nu, qy};
#pragma unroll
  for (int k = 0; k < 16; ++k) a.out[(long long)k * nxy + i] = vals[k];
}

// Kx, Ky (def-Parallel.cpp:2135-2165, :2214-2244): d<v^2>^2 as Matrix2D::Dd1/Dd2
// (matrices.h:511-533) over the rank's interior, the two end cells taking
// their neighbour's difference; Ky is 0 below 3 rows of y
template <typename T>
__global__ __launch_bounds__(256) void moments_cond(MomArgs<T> a) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  if (i >= nxy) return;
  const int x = i % nx, y = i / nx;
  const double* V = a.out + 10LL * nxy;
  const double n = a.out[i], nu = a.out[14LL * nxy + i];
  const double fmin4 = 4.0 * 1.17549435082228750797e-38;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    const int len = d ? ny : nx, j = d ? y : x;
    double k = 0.0;
    if (len >= 3) {
      const int jj = j == 0 ? 1 : (j == len - 1 ? len - 2 : j);
      const int im = d ? (jj - 1) * nx + x : y * nx + jj - 1;
      const int ip = d ? (jj + 1) * nx + x : y * nx + jj + 1;
      double cond = V[im] * V[im] - V[ip] * V[ip];
      cond *= 0.5 / (d ? -a.dy : -a.dx);
      cond *= n;
      cond /= nu;
      cond *= (-1.0) / 18.0;
      if (cond < fmin4 && cond > -fmin4) cond = 1.0;
      k = a.out[(d ? 15LL : 6LL) * nxy + i] / cond;
    }
    a.out[(16LL + d) * nxy + i] = k;
  }
}

#ifndef OSHUN_FAMILY_TU
// plain copy, 8 B per lane per iteration like the Vlasov and collision kernels (counter calibration)
__global__ __launch_bounds__(256) void diag_copy(const double* __restrict__ src, double* __restrict__ dst,
                                                 long long n) {
  for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += (long long)gridDim.x * 256) dst[i] = src[i];
}
#endif

}  // namespace oshun
