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
  T p0p1_sq, inv_mp0p1_sq, g_r;
  T c05h, hh, h6, third, twothird;
  int np, nx, ny, nc;
  long long plane;
  int mode;  // 0: slope, 1..4: RK4 stage
  int c10, c11r, c11i;
};

template <typename T>
__device__ __forceinline__ T op_value(const Term& tm, const T* __restrict__ s, long long idx, int p, int np, int nx,
                                      long long xm_off, long long xp_off, long long ym_off, long long yp_off,
                                      const StageArgs<T>& a) {
  switch (tm.op) {
    case OP_DX:
      return s[idx + xm_off] - s[idx + xp_off];
    case OP_DY:
      return s[idx + ym_off] - s[idx + yp_off];
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
    const long long xm_off = (x == 0) ? (nx - 1) : -1;
    const long long xp_off = (x == nx - 1) ? -(long long)(nx - 1) : 1;
    const long long ystride = (long long)np * nx;
    const long long ym_off = (y == 0) ? (long long)(ny - 1) * ystride : -ystride;
    const long long yp_off = (y == ny - 1) ? -(long long)(ny - 1) * ystride : ystride;
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
        const T val = op_value<T>(tm, s, idx, p, np, nx, xm_off, xp_off, ym_off, yp_off, a);
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

template <typename T>
struct FieldArgs {
  const double* fin;
  const double* f0;
  double* fout;
  double* facc;
  const double* jpart;
  const T* fpl;  // distribution planes of the stage input
  int npb, np, nx, ny, mode;
  int c10, c11r, c11i;
  long long plane;
  double p3og1, small, delta_p, midx, midy;
  double c05h, hh, h6, third, twothird;
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

template <typename T>
__global__ __launch_bounds__(256) void field_stage(FieldArgs<T> a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.nx, ny = a.ny, nxy = nx * ny;
  if (i >= nxy) return;
  const int x = i % nx, y = i / nx;
  const int xm = y * nx + (x == 0 ? nx - 1 : x - 1), xp = y * nx + (x == nx - 1 ? 0 : x + 1);
  const int ym = (y == 0 ? ny - 1 : y - 1) * nx + x, yp = (y == ny - 1 ? 0 : y + 1) * nx + x;
  const double* Ex = a.fin;
  const double* Ey = a.fin + nxy;
  const double* Ez = a.fin + 2 * nxy;
  const double* Bx = a.fin + 3 * nxy;
  const double* By = a.fin + 4 * nxy;
  const double* Bz = a.fin + 5 * nxy;
  double k[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  // Current::Jxyz (def-VlasovMax.cpp:680-745)
  if (a.c10 >= 0) {
    double j = 0.0;
    for (int b = 0; b < a.npb; ++b) j += a.jpart[((long long)b * 3 + 0) * nxy + i];
    j *= a.delta_p;
    const long long o1 = (long long)a.c10 * a.plane + ((long long)y * a.np + 1) * nx + x;
    j += a.small * (static_cast<double>(a.fpl[o1]) * a.p3og1);
    j *= 4.0 * M_PI / 3.0;
    k[0] += j;
  }
  if (a.c11r >= 0) {
    double jr = 0.0, ji = 0.0;
    for (int b = 0; b < a.npb; ++b) {
      jr += a.jpart[((long long)b * 3 + 1) * nxy + i];
      ji += a.jpart[((long long)b * 3 + 2) * nxy + i];
    }
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
  const double mx = a.midx, my = a.midy;
  // MaxwellEq::Ampere (def-VlasovMax.cpp:1009-1035)
  k[0] += Bz[ym] * my - Bz[yp] * my;
  k[1] += Bz[xm] * (-mx) - Bz[xp] * (-mx);
  k[2] += Bx[ym] * (-my) - Bx[yp] * (-my);
  k[2] += By[xm] * mx - By[xp] * mx;
  // MaxwellEq::Faraday (def-VlasovMax.cpp:979-1005)
  k[3] += Ez[ym] * (-my) - Ez[yp] * (-my);
  k[4] += Ez[xm] * mx - Ez[xp] * mx;
  k[5] += Ex[ym] * my - Ex[yp] * my;
  k[5] += Ey[xm] * (-mx) - Ey[xp] * (-mx);
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const long long o = (long long)c * nxy + i;
    double* acc = a.facc + o;
    rk_combine(a.mode, k[c], a.f0 ? a.f0[o] : 0.0, a.fout + o, acc, a.c05h, a.hh, a.h6, a.third, a.twothird);
  }
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
  long long plane;
  int do_ee, do_f1, do_flm, numh;
  double h_ee, dt_flm;
  double c_kpre, kpre, density_np, zeta, v0sq, v1sq, g_den;
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

// inclusive prefix over the 64 lanes of per-lane totals
__device__ __forceinline__ double wave_scan_incl(double v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}
// inclusive suffix over lanes
__device__ __forceinline__ double wave_scan_incl_rev(double v) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double t = __shfl_down(v, o, 64);
    if (lane + o < 64) v += t;
  }
  return v;
}

// Rows held by one lane: p = lane*NPL + i.  Prefix of t over p (inclusive).
template <int NPL>
__device__ __forceinline__ void prefix_p(const double (&t)[NPL], double (&out)[NPL], double& total) {
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) { s += t[i]; out[i] = s; }
  const double inc = wave_scan_incl(s);
  double ex = __shfl_up(inc, 1, 64);
  if (lane_id() == 0) ex = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) out[i] += ex;
  total = __shfl(inc, 63, 64);
}
// Suffix of t over p (inclusive).
template <int NPL>
__device__ __forceinline__ void suffix_p(const double (&t)[NPL], double (&out)[NPL]) {
  double s = 0.0;
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { s += t[i]; out[i] = s; }
  const double inc = wave_scan_incl_rev(s);
  double ex = __shfl_down(inc, 1, 64);
  if (lane_id() == 63) ex = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) out[i] += ex;
}
// value of row p+1 (next row) for each held row
template <int NPL>
__device__ __forceinline__ void next_row(const double (&v)[NPL], double (&nx)[NPL]) {
  const double first_of_next = __shfl_down(v[0], 1, 64);
#pragma unroll
  for (int i = 0; i < NPL - 1; ++i) nx[i] = v[i + 1];
  nx[NPL - 1] = first_of_next;
}
template <int NPL>
__device__ __forceinline__ void prev_row(const double (&v)[NPL], double (&pv)[NPL]) {
  const double last_of_prev = __shfl_up(v[NPL - 1], 1, 64);
#pragma unroll
  for (int i = NPL - 1; i > 0; --i) pv[i] = v[i - 1];
  pv[0] = last_of_prev;
}

template <int NPL>
struct EETab {
  double U4[NPL], U4m1[NPL], U2[NPL], U2m1[NPL], U1s[NPL], U1m1s[NPL], U3[NPL], Qn[NPL], Pn[NPL];
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
    t4[i] = in1 ? tb.U4[i] * f[i] + tb.U4m1[i] * fp[i] : 0.0;
    t2[i] = in1 ? tb.U2[i] * f[i] + tb.U2m1[i] * fp[i] : 0.0;
    t1[i] = (p < np - 1) ? tb.U1s[i] * fn[i] + tb.U1m1s[i] * f[i] : 0.0;
  }
  double I4[NPL], I2[NPL], J1[NPL], tot4, tot2;
  prefix_p<NPL>(t4, I4, tot4);
  prefix_p<NPL>(t2, I2, tot2);
  suffix_p<NPL>(t1, J1);
  const double lnee = log_ee(4.0 * M_PI * tot2, 4.0 * M_PI * tot4, a.density_np);
  // f(p0), f(p1) for the parabolic G at low v
  const double f0 = __shfl(f[0], 0, 64);
  const double f1 = (NPL >= 2) ? __shfl(f[NPL >= 2 ? 1 : 0], 0, 64) : __shfl(f[0], 1, 64);
  const double f00 = (f0 - f1 * a.v0sq / a.v1sq) / a.g_den;
  double G[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    double i2 = I2[i] * J1[i];
    i2 *= -3.0;
    double g = I4[i] + tb.U3[i] * J1[i];
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
  const double G0 = __shfl(G[0], 0, 64);
  double fh0 = (-1.0) * G0;
  fh0 *= 2.0 / a.v0sq;
  double D2[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    D2[i] = (p < np - 1) ? D1[i] * tb.Pn[i] : D1[i];
  }
  fh0 -= __shfl(D2[0], 0, 64);
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
    double v = (p == 0) ? fh0 : tb.Qn[i] * D3p[i];
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
  // inclusive scan: S_j = P_j S_{j-1}
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    Mob q;
    q.a = __shfl_up(P.a, o, 64); q.b = __shfl_up(P.b, o, 64);
    q.c = __shfl_up(P.c, o, 64); q.d = __shfl_up(P.d, o, 64);
    if (lane >= o) P = mob_norm(mob_mul(P, q));
  }
  Mob E;
  E.a = __shfl_up(P.a, 1, 64); E.b = __shfl_up(P.b, 1, 64);
  E.c = __shfl_up(P.c, 1, 64); E.d = __shfl_up(P.d, 1, 64);
  if (lane == 0) E = {1.0, 0.0, 0.0, 1.0};
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
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double qa = __shfl_up(al, o, 64), qb = __shfl_up(be, o, 64);
    if (lane >= o) { be = al * qb + be; al = al * qa; }
  }
  double y = __shfl_up(be, 1, 64);
  if (lane == 0) y = 0.0;
#pragma unroll
  for (int i = 0; i < NPL; ++i) { y = ali[i] * y + bei[i]; d[i] = y; }
  // backward affine scan x -> ga*x + de
  double ga = 1.0, de = 0.0;
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { de = -cp[i] * de + d[i]; ga = -cp[i] * ga; }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double qa = __shfl_down(ga, o, 64), qb = __shfl_down(de, o, 64);
    if (lane + o < 64) { de = ga * qb + de; ga = ga * qa; }
  }
  double x = __shfl_down(de, 1, 64);
  if (lane == 63) x = 0.0;
#pragma unroll
  for (int i = NPL - 1; i >= 0; --i) { x = d[i] - cp[i] * x; d[i] = x; }
}

template <typename T, int NPL>
__global__ __launch_bounds__(64) void collide(CollArgs<T> a) {
  const int col = blockIdx.x;
  const int nx = a.nx, np = a.np;
  const int x = col % nx, y = col / nx;
  const int lane = lane_id();
  const long long base = (long long)y * np * nx + x;
  EETab<NPL> tb;
  double f[NPL];
#pragma unroll
  for (int i = 0; i < NPL; ++i) {
    const int p = lane * NPL + i;
    const bool in = p < np;
    tb.U4[i] = in ? a.t.U4[p] : 0.0;
    tb.U4m1[i] = in ? a.t.U4m1[p] : 0.0;
    tb.U2[i] = in ? a.t.U2[p] : 0.0;
    tb.U2m1[i] = in ? a.t.U2m1[p] : 0.0;
    tb.U1s[i] = in ? a.t.U1s[p] : 0.0;
    tb.U1m1s[i] = in ? a.t.U1m1s[p] : 0.0;
    tb.U3[i] = in ? a.t.U3[p] : 1.0;
    tb.Qn[i] = in ? a.t.Qn[p] : 0.0;
    tb.Pn[i] = in ? a.t.Pn[p] : 0.0;
    f[i] = in ? static_cast<double>(a.f[base + (long long)p * nx]) : 0.0;
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
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int p = lane * NPL + i;
      if (p < np) a.f[base + (long long)p * nx] = static_cast<T>(f[i]);
    }
    // the implicit step reads f00 as stored
    if (sizeof(T) < sizeof(double)) {
#pragma unroll
      for (int i = 0; i < NPL; ++i) f[i] = static_cast<double>(static_cast<T>(f[i]));
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
    t4[i] = in1 ? tb.U4[i] * f[i] + tb.U4m1[i] * fp[i] : 0.0;
    t2[i] = in1 ? tb.U2[i] * f[i] + tb.U2m1[i] * fp[i] : 0.0;
    t1[i] = (p < np - 1) ? tb.U1s[i] * fn[i] + tb.U1m1s[i] * f[i] : 0.0;
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
    s /= tb.U3[i];
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
  // Implicit_FLM_step::advance for every l >= 1 plane, def-FokkerPlanck.cpp:943-1035
  const int lmin = a.do_f1 ? 1 : 2;
  const int lmax = a.do_flm ? a.l0 : 1;
  for (int l = lmin; l <= lmax; ++l) {
    double ll1 = static_cast<double>(l);
    ll1 *= (-0.5) * (ll1 + 1.0);
    double Bl[NPL], cp[NPL], w[NPL];
#pragma unroll
    for (int i = 0; i < NPL; ++i) {
      const int p = lane * NPL + i;
      double b = (p == 0 && l > 1) ? 0.0 : Bdia[i];
      b += 1.0 - ll1 * Sct[i];
      Bl[i] = (p < np) ? b : 1.0;
    }
    tri_pivots<NPL>(Asub, Bl, Csup, cp, w);
    for (int c = 1; c < a.nc; ++c) {
      if (a.comp_l[c] != l) continue;
      // f1_loop handles m = 0, 1 only (def-FokkerPlanck.cpp:1130); flm_loop m <= min(m0,l)
      T* fc = a.f + (long long)c * a.plane;
      double d[NPL];
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int p = lane * NPL + i;
        d[i] = (p < np) ? static_cast<double>(fc[base + (long long)p * nx]) : 0.0;
      }
      tri_solve<NPL>(Asub, cp, w, d);
#pragma unroll
      for (int i = 0; i < NPL; ++i) {
        const int p = lane * NPL + i;
        if (p < np) fc[base + (long long)p * nx] = static_cast<T>(d[i]);
      }
    }
  }
}

}  // namespace oshun
