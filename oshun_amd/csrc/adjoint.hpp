// adjoint.hpp -- discrete adjoint (vector-Jacobian product) of one oshun_step
// for 1D runs with m0 = 0 in FP64: RK4 of the Vlasov-Maxwell right-hand side,
// then the e-e sub-cycles on f00 and the implicit l >= 1 solves.  Included at
// the end of oshun_hip.hip.
//
// The step is y_{n+1} = C(R(y_n)) with R the RK4 map (Runge_Kutta_4::advance_p1,
// def-RK.cpp:110-145) and C the collisions (Explicit_EC_FP::loop and
// Euler_Backward::advance_1/advance_lm, def-FokkerPlanck.cpp:590-625,1270-1304).
// Given the cotangent lambda of y_{n+1}, the adjoint returns R'^T C'^T lambda
// at the trajectory point y_n.  Each operator's transpose is written out by
// hand from the forward formulas:
//   * RHS (Actions::Adv for m0 = 0, 1D): x-advection, E-effect through
//     G/H/G00 with the p0 and pmax closures, CLEAN_LOW_PC, the current of f10
//     into Ex, and the vacuum Maxwell terms;
//   * e-e: Getslope (EE_isotropic_conserva, def-FokkerPlanck.cpp:178-267)
//     including the parabolic G below NB and the Coulomb logarithm, inside the
//     RK4 sub-cycles (RungeKutta4_ee::advance, :518-553);
//   * implicit: A_l(f00) x = b per l (reset_coeff + advance, :808-1035): the
//     transposed tridiagonal solve for b and -z x^T for the bands, carried back
//     through the band formulas (integrals, LOGee, ZLOGei) to f00.
// For m0 = 0 in 1D, Ey, Ez and B do not couple to f; their cotangents follow
// the vacuum Maxwell terms only.

namespace oshun {

// ---- right-hand side ----------------------------------------------------------
struct AdjRhsArgs {
  const double* f;      // stage input planes [l][p][x]
  const double* fld;    // stage input fields [6][nx]
  const double* mu;     // cotangent of the slope planes
  const double* muf;    // cotangent of the field slopes [6][nx]
  double* nu;           // out: planes
  double* nuf;          // out: fields
  double* epart;        // scratch [p][x]: Ex cotangent before the sum over p
  const double* invpax; // [l][p]
  const double* cg;     // [l]
  const double* hp0;    // [l]
  const double* xv1;    // [l][p] source l -> l+1
  const double* xv2;    // [l][p] source l -> l-1
  const double* eA1;    // [l] E-field G coefficient of source l
  const double* eA2;    // [l] E-field H coefficient of source l
  const double* jw;     // [p] d(dEx/dt)/d f10(p): (4 pi/3)(delta_p w_p p^3/g + small p^3/g(p1) [p = 1])
  double p0p1_sq, inv_mp0p1_sq, g_r, midx;
  int np, nx;
  long long plane;
};

// nu_j(p,x) = sum over the slope entries that read f_j(p,x) (gather form)
template <int NL>
__global__ __launch_bounds__(256) void adj_rhs_1d(AdjRhsArgs a) {
  const int x = blockIdx.x * 64 + threadIdx.x;
  const int p = blockIdx.y * 4 + threadIdx.y;
  const int np = a.np, nx = a.nx;
  if (x >= nx || p >= np) return;
  const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
  const double Ex = a.fld[x];
  // cotangent of slope l at (q, xx) after CLEAN_LOW_PC (def-Actions.cpp:240-253)
  auto MU = [&](int l, int q, int xx) -> double {
    if (q < 0 || q >= np) return 0.0;
    if (l >= 2 && q < min(l - 1, 80)) return 0.0;
    return a.mu[(long long)l * a.plane + (long long)q * nx + xx];
  };
#pragma unroll
  for (int j = 0; j < NL; ++j) {
    double v = 0.0;
    // x-advection: slope l(x) += c (f_j(x-1) - f_j(x+1))
    if (j + 1 < NL) v += a.xv1[j * np + p] * (MU(j + 1, p, xp) - MU(j + 1, p, xm));
    if (j >= 1) v += a.xv2[j * np + p] * (MU(j - 1, p, xp) - MU(j - 1, p, xm));
    // E-effect: slope j+1 += G_j Ex eA1[j], slope j-1 += H_j Ex eA2[j]
    auto nG = [&](int q) { return (j + 1 < NL) ? MU(j + 1, q, x) * (Ex * a.eA1[j]) : 0.0; };
    auto nH = [&](int q) { return (j >= 1) ? MU(j - 1, q, x) * (Ex * a.eA2[j]) : 0.0; };
    // G_j = (cg+1) Dp f + invpax f, H_j = Dp f + invpax f for p >= 1 (j >= 1); G_0 = Dp f.
    // Dp f(q) = f(q-1) - f(q+1) for 1 <= q <= np-2 and f(np-2) - f(np-1) at np-1.
    auto comb = [&](int q) { return (j == 0) ? nG(q) : (a.cg[j] + 1.0) * nG(q) + nH(q); };
    double td = 0.0;
    if (p + 1 <= np - 1) td += comb(p + 1);
    if (p - 1 >= 1 && p - 1 <= np - 2) td -= comb(p - 1);
    if (p == np - 1) td -= comb(np - 1);
    v += td;
    if (j == 0) {  // G00 at p0 (def-VlasovMax.cpp:531-532)
      const double ng0 = nG(0);
      if (p == 0) v -= a.g_r * a.inv_mp0p1_sq * ng0;
      if (p == 1) v += a.g_r * (1.0 + a.p0p1_sq * a.inv_mp0p1_sq) * ng0;
    } else {
      if (p >= 1) v += a.invpax[j * np + p] * (nG(p) + nH(p));
      if (p == 1) v += a.hp0[j] * nH(0);  // H(p0) = f(p1) Hp0 (:513-514)
    }
    if (j == 1) v += a.muf[x] * a.jw[p];  // Current::Jxyz into Ex
    a.nu[(long long)j * a.plane + (long long)p * nx + x] = v;
  }
  // Ex cotangent of this point: sum_l mu_l (eA1[l-1] G_{l-1} + eA2[l+1] H_{l+1})
  double G[NL], Hh[NL];
  {
    auto F = [&](int l, int q) { return a.f[(long long)l * a.plane + (long long)q * nx + x]; };
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const double fc = F(l, p);
      const double fn = (p < np - 1) ? F(l, p + 1) : fc;
      if (p == 0) {
        if (l == 0) {
          const double fe = (fc - fn * a.p0p1_sq) * a.inv_mp0p1_sq;
          G[0] = (fn - fe) * a.g_r;
          Hh[0] = 0.0;
        } else {
          G[l] = 0.0;
          Hh[l] = fn * a.hp0[l];
        }
      } else {
        const double dpf = F(l, p - 1) - fn;
        if (l == 0) {
          G[0] = dpf;
          Hh[0] = 0.0;
        } else {
          Hh[l] = fc * a.invpax[l * np + p] + dpf;
          G[l] = dpf * a.cg[l] + Hh[l];
        }
      }
    }
  }
  double e = 0.0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    const double m = MU(l, p, x);
    if (l >= 1) e += m * (a.eA1[l - 1] * G[l - 1]);
    if (l + 1 < NL) e += m * (a.eA2[l + 1] * Hh[l + 1]);
  }
  a.epart[(long long)p * nx + x] = e;
}

// field cotangents: Ex from the E-effect partials; vacuum Maxwell terms
// (MaxwellEq::Ampere/Faraday in 1D, def-VlasovMax.cpp:979-1035) transposed
__global__ __launch_bounds__(256) void adj_field_1d(AdjRhsArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.nx;
  if (x >= nx) return;
  const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
  double ex = 0.0;
  for (int p = 0; p < a.np; ++p) ex += a.epart[(long long)p * nx + x];
  const double* m = a.muf;
  const double mx = a.midx;
  a.nuf[x] = ex;
  a.nuf[nx + x] = -mx * (m[5 * nx + xp] - m[5 * nx + xm]);
  a.nuf[2 * nx + x] = mx * (m[4 * nx + xp] - m[4 * nx + xm]);
  a.nuf[3 * nx + x] = 0.0;
  a.nuf[4 * nx + x] = mx * (m[2 * nx + xp] - m[2 * nx + xm]);
  a.nuf[5 * nx + x] = -mx * (m[nx + xp] - m[nx + xm]);
}

// out = a*x + b*y over n values (y may be null)
__global__ __launch_bounds__(256) void adj_axpby(long long n, double a, const double* x, double b,
                                                 const double* y, double* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = a * x[i];
  if (y) v += b * y[i];
  out[i] = v;
}
// out = y0 + k*c (an RK stage input, rounded as in the stage kernels)
__global__ __launch_bounds__(256) void adj_stage_in(long long n, const double* y0, const double* k, double c,
                                                    double* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = y0[i] + k[i] * c;
}
// acc = y0 + (k1 h/2)/3, then += (k2 h/2) 2/3, += (k3 h)/3, then y = acc + k4 h/6 (def-RK.cpp:116-138)
__global__ __launch_bounds__(256) void adj_rk_acc(long long n, int stage, const double* y0, const double* k,
                                                  double h, double* acc) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  switch (stage) {
    case 1: { const double kh = k[i] * (0.5 * h); acc[i] = y0[i] + kh * (1.0 / 3.0); } break;
    case 2: { const double kh = k[i] * (0.5 * h); acc[i] = acc[i] + kh * (2.0 / 3.0); } break;
    case 3: { const double kh = k[i] * h; acc[i] = acc[i] + kh * (1.0 / 3.0); } break;
    default: { const double kh = k[i] * (h / 6.0); acc[i] = acc[i] + kh; } break;
  }
}

// ---- collisions (one thread per x column, sequential in p) ----------------------
// per-p tables, rows of np
enum AdjTab {
  AT_U4, AT_U4M1, AT_U2, AT_U2M1, AT_U1S, AT_U1M1S, AT_U3, AT_QN, AT_PN, AT_VC, AT_IV2, AT_IU3,
  AT_IVDM, AT_IVDP, AT_IVS2, AT_ROW, AT_P3, AT_A2, AT_B2, AT_A4, AT_B4, AT_COUNT
};
// work vectors of one column, [slot][p][x]; slots >= W_SUB hold the e-e sub-cycle inputs
enum AdjSlot {
  W_Y0, W_Y1, W_Y2, W_Y3, W_K, W_I4, W_I2, W_J1, W_G, W_GI4, W_GI2, W_GJ1, W_LAM, W_MU, W_NU, W_ACC,
  W_CI2, W_CI0, W_CJ1, W_T1, W_T2, W_SUB0, W_DIAG0, W_SUP0, W_X, W_Z, W_CP, W_GSUB, W_GDIAG, W_GSUP, W_GSC,
  W_SUB
};

struct AdjCollArgs {
  const double* frk;    // collision input planes (the RK4 result)
  double* lam;          // in: cotangent of the collision output; out: of its input
  double* W;            // work vectors
  const double* tb;     // [AT_COUNT][np]
  const int* comp_l;
  long long plane;
  int np, nx, nc, l0, nb, numh, do_ee, do_f1, do_flm;
  double h_ee, dt_flm, c_kpre, kpre, density_np, zeta, r01, g_den, v0sq;
};

// Coulomb logarithms with their derivatives in (ne, Te) (0 where clamped)
__device__ __forceinline__ void log_ee_d(double ne, double Te, double dens, double* dne, double* dTe) {
  *dne = 0.0;
  *dTe = 0.0;
  if (ne > 0.000000001) {
    const double a = log(Te / (3.0 * ne) * 511000), b = log(ne * dens);
    const double s = sqrt(0.00001 + 0.0625 * (a - 2.0) * (a - 2.0));
    if (23.5 - 0.5 * b + 1.25 * a - s > 2.0) {
      const double dLda = 1.25 - 0.0625 * (a - 2.0) / s;
      *dTe = dLda / Te;
      *dne = (-dLda - 0.5) / ne;
    }
  }
}
__device__ __forceinline__ void zlog_ei_d(double ne, double Te, double dens, double zeta, double* dne, double* dTe) {
  *dne = 0.0;
  *dTe = 0.0;
  if (ne > 0.0000001) {
    const double a = log(Te / (3.0 * ne) * 511000), b = log(ne * dens);
    if (24.0 - 0.5 * b + a > 2.0) {
      *dTe = zeta / Te;
      *dne = zeta * (-1.5) / ne;
    }
  }
}

struct ColView {
  double* W;
  const double* tb;
  int np, nx, x;
  __device__ double& w(int k, int p) const { return W[((long long)k * np + p) * nx + x]; }
  __device__ double t(int k, int p) const { return tb[k * np + p]; }
};

// Getslope forward: slope of slot fin into slot out (I4, I2, J1, G kept in their slots)
__device__ void adj_ee_slope(const AdjCollArgs& a, const ColView& V, int fin, int out, double* Lout) {
  const int N = a.np;
  V.w(W_I4, 0) = 0.0;
  V.w(W_I2, 0) = 0.0;
  for (int k = 1; k < N; ++k) {
    V.w(W_I4, k) = (V.t(AT_U4, k) * V.w(fin, k) + V.t(AT_U4M1, k) * V.w(fin, k - 1)) + V.w(W_I4, k - 1);
    V.w(W_I2, k) = (V.t(AT_U2, k) * V.w(fin, k) + V.t(AT_U2M1, k) * V.w(fin, k - 1)) + V.w(W_I2, k - 1);
  }
  V.w(W_J1, N - 1) = 0.0;
  for (int k = N - 2; k >= 0; --k)
    V.w(W_J1, k) = (V.t(AT_U1S, k) * V.w(fin, k + 1) + V.t(AT_U1M1S, k) * V.w(fin, k)) + V.w(W_J1, k + 1);
  const double L = log_ee(4.0 * M_PI * V.w(W_I2, N - 1), 4.0 * M_PI * V.w(W_I4, N - 1), a.density_np);
  const double f0 = V.w(fin, 0), f1 = V.w(fin, 1);
  const double f00 = (f0 - f1 * a.r01) / a.g_den;
  for (int k = 0; k < N; ++k) {
    const double fk = V.w(fin, k), j1 = V.w(W_J1, k);
    double g;
    if (k < a.nb) {  // EE_isotropic_conserva::G (:178-188)
      const double i2s = f00 * V.t(AT_A2, k) + (f1 - f00) * V.t(AT_B2, k);
      const double i4s = f00 * V.t(AT_A4, k) + (f1 - f00) * V.t(AT_B4, k);
      g = fk * i4s + (V.t(AT_P3, k) * fk - 3.0 * i2s) * j1;
    } else {
      g = (V.w(W_I4, k) + V.t(AT_U3, k) * j1) * fk - 3.0 * (V.w(W_I2, k) * j1);
    }
    V.w(W_G, k) = g;
  }
  const double cl = a.c_kpre * L;
  auto D = [&](int k) { return (V.w(W_G, k) - V.w(W_G, k + 1)) * V.t(AT_PN, k); };  // k <= N-2
  double dprev = D(0);
  V.w(out, 0) = ((-V.w(W_G, 0)) * (2.0 / a.v0sq) - dprev) * V.t(AT_QN, 0) * cl;
  for (int k = 1; k < N - 1; ++k) {
    const double dk = D(k);
    V.w(out, k) = V.t(AT_QN, k) * (dprev - dk) * cl;
    dprev = dk;
  }
  V.w(out, N - 1) = 0.0;
  if (Lout) *Lout = L;
}

// Getslope VJP: slot gout += (d slope/d f)^T slot nu, at f = slot fin
__device__ void adj_ee_slope_vjp(const AdjCollArgs& a, const ColView& V, int fin, int nu, int gout) {
  const int N = a.np;
  double L;
  adj_ee_slope(a, V, fin, W_K, &L);  // I4, I2, J1, G of this input
  const double cl = a.c_kpre * L;
  // gL = sum_k nu_k c stuff_k with slope_k = c L stuff_k
  double gL = 0.0;
  for (int k = 0; k < N; ++k) gL += V.w(nu, k) * V.w(W_K, k);
  gL /= (L != 0.0 ? L : 1.0);
  // gD_k = Qn_{k+1} nu'_{k+1} - Qn_k nu'_k (k <= N-2), nu' = nu c L; gG from D_k = (G_k - G_{k+1}) Pn_k
  auto nup = [&](int k) { return (k <= N - 2) ? V.w(nu, k) * cl * V.t(AT_QN, k) : 0.0; };
  double gdprev = 0.0;  // gD_{k-1}
  for (int k = 0; k < N; ++k) {
    const double gd = (k <= N - 2) ? nup(k + 1) - nup(k) : 0.0;
    double gg = 0.0;
    if (k <= N - 2) gg += V.t(AT_PN, k) * gd;
    if (k >= 1) gg -= V.t(AT_PN, k - 1) * gdprev;
    if (k == 0) gg += -(2.0 / a.v0sq) * nup(0);
    V.w(W_GI4, k) = gg;  // gG kept in GI4 for the moment
    gdprev = gd;
  }
  // through G: gf (direct), gI4, gI2, gJ1
  const double f0 = V.w(fin, 0), f1 = V.w(fin, 1);
  const double f00 = (f0 - f1 * a.r01) / a.g_den;
  double gf00 = 0.0, gf1x = 0.0;
  for (int k = 0; k < N; ++k) {
    const double gg = V.w(W_GI4, k);
    const double fk = V.w(fin, k), j1 = V.w(W_J1, k);
    if (k < a.nb) {
      const double i2s = f00 * V.t(AT_A2, k) + (f1 - f00) * V.t(AT_B2, k);
      const double i4s = f00 * V.t(AT_A4, k) + (f1 - f00) * V.t(AT_B4, k);
      V.w(gout, k) += gg * (i4s + V.t(AT_P3, k) * j1);
      V.w(W_GJ1, k) = gg * (V.t(AT_P3, k) * fk - 3.0 * i2s);
      const double gi4s = gg * fk, gi2s = -3.0 * gg * j1;
      gf00 += gi4s * (V.t(AT_A4, k) - V.t(AT_B4, k)) + gi2s * (V.t(AT_A2, k) - V.t(AT_B2, k));
      gf1x += gi4s * V.t(AT_B4, k) + gi2s * V.t(AT_B2, k);
      V.w(W_GI4, k) = 0.0;
      V.w(W_GI2, k) = 0.0;
    } else {
      V.w(gout, k) += gg * (V.w(W_I4, k) + V.t(AT_U3, k) * j1);
      V.w(W_GJ1, k) = gg * (V.t(AT_U3, k) * fk - 3.0 * V.w(W_I2, k));
      V.w(W_GI4, k) = gg * fk;
      V.w(W_GI2, k) = -3.0 * gg * j1;
    }
  }
  V.w(gout, 0) += gf00 / a.g_den;
  V.w(gout, 1) += gf1x - gf00 * a.r01 / a.g_den;
  // Coulomb logarithm of the totals
  double dne, dTe;
  log_ee_d(4.0 * M_PI * V.w(W_I2, N - 1), 4.0 * M_PI * V.w(W_I4, N - 1), a.density_np, &dne, &dTe);
  V.w(W_GI2, N - 1) += gL * 4.0 * M_PI * dne;
  V.w(W_GI4, N - 1) += gL * 4.0 * M_PI * dTe;
  // prefix sums (from p = 1) transposed: suffix sums of their cotangents
  double s4 = 0.0, s2 = 0.0;
  for (int q = N - 1; q >= 1; --q) {
    s4 += V.w(W_GI4, q);
    s2 += V.w(W_GI2, q);
    V.w(gout, q) += V.t(AT_U4, q) * s4 + V.t(AT_U2, q) * s2;
    V.w(gout, q - 1) += V.t(AT_U4M1, q) * s4 + V.t(AT_U2M1, q) * s2;
  }
  // suffix sum J1_k = sum_{q=k}^{N-2} (U1s_q f_{q+1} + U1m1s_q f_q) transposed: prefix sums
  double s1 = 0.0;
  for (int q = 0; q <= N - 2; ++q) {
    s1 += V.w(W_GJ1, q);
    V.w(gout, q + 1) += V.t(AT_U1S, q) * s1;
    V.w(gout, q) += V.t(AT_U1M1S, q) * s1;
  }
}

// one RK4 e-e sub-cycle: slot f advanced in place (stage inputs left in W_Y0..W_Y3)
__device__ void adj_ee_rk4(const AdjCollArgs& a, const ColView& V, int f, double h) {
  const int N = a.np;
  for (int k = 0; k < N; ++k) V.w(W_Y0, k) = V.w(f, k);
  adj_ee_slope(a, V, W_Y0, W_K, nullptr);
  for (int k = 0; k < N; ++k) {
    const double hs = V.w(W_K, k) * (0.5 * h);
    V.w(W_Y1, k) = V.w(W_Y0, k) + hs;
    V.w(f, k) += hs * (1.0 / 3.0);
  }
  adj_ee_slope(a, V, W_Y1, W_K, nullptr);
  for (int k = 0; k < N; ++k) {
    const double hs = V.w(W_K, k) * (0.5 * h);
    V.w(W_Y2, k) = V.w(W_Y0, k) + hs;
    V.w(f, k) += hs * (2.0 / 3.0);
  }
  adj_ee_slope(a, V, W_Y2, W_K, nullptr);
  for (int k = 0; k < N; ++k) {
    const double hs = V.w(W_K, k) * h;
    V.w(W_Y3, k) = V.w(W_Y0, k) + hs;
    V.w(f, k) += hs * (1.0 / 3.0);
  }
  adj_ee_slope(a, V, W_Y3, W_K, nullptr);
  for (int k = 0; k < N; ++k) V.w(f, k) += V.w(W_K, k) * (h / 6.0);
}

// transposed RK4 sub-cycle: slot lam (cotangent of the output) -> cotangent of
// the input, at the input in slot fin
__device__ void adj_ee_rk4_vjp(const AdjCollArgs& a, const ColView& V, int fin, int lam, double h) {
  const int N = a.np;
  {  // stage inputs
    for (int k = 0; k < N; ++k) V.w(W_ACC, k) = V.w(fin, k);
    adj_ee_rk4(a, V, W_ACC, h);
  }
  // acc = lam + sum of the stage cotangents; mu = cotangent of the stage slope
  for (int k = 0; k < N; ++k) { V.w(W_ACC, k) = V.w(lam, k); V.w(W_MU, k) = V.w(lam, k) * (h / 6.0); V.w(W_NU, k) = 0.0; }
  adj_ee_slope_vjp(a, V, W_Y3, W_MU, W_NU);  // nu = lambda of Y3
  for (int k = 0; k < N; ++k) {
    V.w(W_ACC, k) += V.w(W_NU, k);
    V.w(W_MU, k) = V.w(lam, k) * (h / 3.0) + V.w(W_NU, k) * h;
    V.w(W_NU, k) = 0.0;
  }
  adj_ee_slope_vjp(a, V, W_Y2, W_MU, W_NU);
  for (int k = 0; k < N; ++k) {
    V.w(W_ACC, k) += V.w(W_NU, k);
    V.w(W_MU, k) = V.w(lam, k) * (h / 3.0) + V.w(W_NU, k) * (0.5 * h);
    V.w(W_NU, k) = 0.0;
  }
  adj_ee_slope_vjp(a, V, W_Y1, W_MU, W_NU);
  for (int k = 0; k < N; ++k) {
    V.w(W_ACC, k) += V.w(W_NU, k);
    V.w(W_MU, k) = V.w(lam, k) * (h / 6.0) + V.w(W_NU, k) * (0.5 * h);
    V.w(W_NU, k) = 0.0;
  }
  adj_ee_slope_vjp(a, V, W_Y0, W_MU, W_NU);
  for (int k = 0; k < N; ++k) V.w(lam, k) = V.w(W_ACC, k) + V.w(W_NU, k);
}

// reset_coeff (def-FokkerPlanck.cpp:808-939) from slot f00: bands (before the
// factor s) in W_SUB0/W_DIAG0/W_SUP0, scattering term in W_GSC's partner
// W_T1/W_T2 kept; returns lnee, zl (ZLOGei), the totals and s.
struct AdjCoef {
  double lnee, zl, ne, Te, s, kd;
};
__device__ AdjCoef adj_coeff(const AdjCollArgs& a, const ColView& V, int f) {
  const int N = a.np;
  AdjCoef k;
  V.w(W_CI2, 0) = 0.0;
  V.w(W_CI0, 0) = 0.0;
  for (int q = 1; q < N; ++q) {
    V.w(W_CI2, q) = (V.t(AT_U4, q) * V.w(f, q) + V.t(AT_U4M1, q) * V.w(f, q - 1)) + V.w(W_CI2, q - 1);
    V.w(W_CI0, q) = (V.t(AT_U2, q) * V.w(f, q) + V.t(AT_U2M1, q) * V.w(f, q - 1)) + V.w(W_CI0, q - 1);
  }
  V.w(W_CJ1, N - 1) = 0.0;
  for (int q = N - 2; q >= 0; --q)
    V.w(W_CJ1, q) = (V.t(AT_U1S, q) * V.w(f, q + 1) + V.t(AT_U1M1S, q) * V.w(f, q)) + V.w(W_CJ1, q + 1);
  k.Te = V.w(W_CI2, N - 1) * (4.0 * M_PI);
  k.ne = V.w(W_CI0, N - 1) * (4.0 * M_PI);
  k.lnee = log_ee(k.ne, k.Te, a.density_np);
  k.zl = zlog_ei(k.ne, k.Te, a.density_np, a.zeta);
  k.kd = a.kpre * a.dt_flm;
  k.s = (-1.0) * k.lnee * a.kpre * a.dt_flm;
  for (int i = 0; i < N; ++i) {
    const double i2 = V.w(W_CI2, i) * (4.0 * M_PI) * V.t(AT_IV2, i);
    const double i0 = V.w(W_CI0, i) * (4.0 * M_PI);
    const double j1 = V.w(W_CJ1, i) * (4.0 * M_PI * V.t(AT_VC, i));
    const double T1 = i0 + (2.0 * j1 - i2) * (1.0 / 3.0);
    const double T2 = (i2 + j1) * (1.0 / 3.0);
    V.w(W_T1, i) = T1;
    V.w(W_T2, i) = T2;
    const double dm = V.t(AT_IVDM, i), dp = V.t(AT_IVDP, i), s2 = V.t(AT_IVS2, i);
    V.w(W_DIAG0, i) = 8.0 * M_PI * V.w(f, i) * V.t(AT_ROW, i) - T2 * (dm + dp);
    V.w(W_SUB0, i) = T2 * dm - T1 * s2;
    V.w(W_SUP0, i) = T2 * dp + T1 * s2;
  }
  return k;
}
__device__ __forceinline__ double adj_scatt(const AdjCollArgs& a, const ColView& V, const AdjCoef& k, int i) {
  double sct = (i == 0) ? 0.0 : V.w(W_T1, i) * k.lnee;
  sct += k.zl * k.ne;
  return sct * V.t(AT_IU3, i) * k.kd;
}

__global__ __launch_bounds__(64) void adj_collide_col(AdjCollArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= a.nx) return;
  const int N = a.np;
  ColView V{a.W, a.tb, N, a.nx, x};
  auto F = [&](int c, int p) -> const double& { return a.frk[(long long)c * a.plane + (long long)p * a.nx + x]; };
  auto LAM = [&](int c, int p) -> double& { return a.lam[(long long)c * a.plane + (long long)p * a.nx + x]; };
  // e-e forward: sub-cycle inputs W_SUB + s, post-e-e f00 in W_LAM for now
  for (int p = 0; p < N; ++p) V.w(W_LAM, p) = F(0, p);
  if (a.do_ee) {
    for (int s = 0; s < a.numh; ++s) {
      for (int p = 0; p < N; ++p) V.w(W_SUB + s, p) = V.w(W_LAM, p);
      adj_ee_rk4(a, V, W_LAM, a.h_ee);
    }
  }
  // W_LAM now holds f00 after e-e: the implicit coefficients come from it
  const int lmin = a.do_f1 ? 1 : 2;
  const int lmax = a.do_flm ? a.l0 : 1;
  const bool any = (a.do_f1 || a.do_flm);
  for (int p = 0; p < N; ++p) { V.w(W_GSUB, p) = 0.0; V.w(W_GDIAG, p) = 0.0; V.w(W_GSUP, p) = 0.0; V.w(W_GSC, p) = 0.0; }
  AdjCoef K{};
  if (any) {
    K = adj_coeff(a, V, W_LAM);
    for (int c = 1; c < a.nc; ++c) {
      const int l = a.comp_l[c];
      if (l < lmin || l > lmax) continue;
      double ll1 = static_cast<double>(l);
      ll1 *= (-0.5) * (ll1 + 1.0);
      auto bb = [&](int i) {
        double d = (i == 0 && l > 1) ? 0.0 : V.w(W_DIAG0, i) * K.s;
        return d + (1.0 - ll1 * adj_scatt(a, V, K, i));
      };
      auto aa = [&](int i) { return (i >= 1) ? V.w(W_SUB0, i) * K.s : 0.0; };
      auto cc = [&](int i) { return (i <= N - 2) ? V.w(W_SUP0, i) * K.s : 0.0; };
      // forward solve x = A^{-1} f_rk (Thomas, def-Nmethods.cpp:142-167)
      {
        double b0 = bb(0);
        V.w(W_CP, 0) = cc(0) / b0;
        V.w(W_X, 0) = F(c, 0) / b0;
        for (int i = 1; i < N; ++i) {
          const double id = 1.0 / (bb(i) - V.w(W_CP, i - 1) * aa(i));
          V.w(W_CP, i) = cc(i) * id;
          V.w(W_X, i) = (F(c, i) - V.w(W_X, i - 1) * aa(i)) * id;
        }
        for (int i = N - 2; i >= 0; --i) V.w(W_X, i) = V.w(W_X, i) - V.w(W_CP, i) * V.w(W_X, i + 1);
      }
      // transposed solve A^T z = lambda: sub of A^T is c_{i-1}, super is a_{i+1}
      {
        double b0 = bb(0);
        V.w(W_CP, 0) = aa(1) / b0;
        V.w(W_Z, 0) = LAM(c, 0) / b0;
        for (int i = 1; i < N; ++i) {
          const double sub = cc(i - 1), sup = (i + 1 < N) ? aa(i + 1) : 0.0;
          const double id = 1.0 / (bb(i) - V.w(W_CP, i - 1) * sub);
          V.w(W_CP, i) = sup * id;
          V.w(W_Z, i) = (LAM(c, i) - V.w(W_Z, i - 1) * sub) * id;
        }
        for (int i = N - 2; i >= 0; --i) V.w(W_Z, i) = V.w(W_Z, i) - V.w(W_CP, i) * V.w(W_Z, i + 1);
      }
      // band cotangents -z x^T; the input cotangent is z
      for (int i = 0; i < N; ++i) {
        const double z = V.w(W_Z, i), xi = V.w(W_X, i);
        if (i >= 1) V.w(W_GSUB, i) += -z * V.w(W_X, i - 1);
        if (i <= N - 2) V.w(W_GSUP, i) += -z * V.w(W_X, i + 1);
        if (!(i == 0 && l > 1)) V.w(W_GDIAG, i) += -z * xi;
        V.w(W_GSC, i) += -z * xi * (-ll1);
        LAM(c, i) = z;
      }
    }
  }
  // cotangent of f00 after e-e: lambda_00 plus the bands' dependence on it (into W_NU)
  for (int p = 0; p < N; ++p) V.w(W_NU, p) = LAM(0, p);
  if (any) {
    // bands: sub = s sub0 etc., s = -lnee kpre Dt
    double gs = 0.0;
    for (int i = 0; i < N; ++i)
      gs += V.w(W_GSUB, i) * V.w(W_SUB0, i) + V.w(W_GDIAG, i) * V.w(W_DIAG0, i) + V.w(W_GSUP, i) * V.w(W_SUP0, i);
    double glnee = gs * (-a.kpre * a.dt_flm);
    double gzi = 0.0;
    // T1, T2 cotangents (reuse W_GI4 for gT1, W_GI2 for gT2)
    for (int i = 0; i < N; ++i) {
      const double gsub0 = V.w(W_GSUB, i) * K.s, gdiag0 = V.w(W_GDIAG, i) * K.s, gsup0 = V.w(W_GSUP, i) * K.s;
      const double dm = V.t(AT_IVDM, i), dp = V.t(AT_IVDP, i), s2 = V.t(AT_IVS2, i);
      double gT1 = -gsub0 * s2 + gsup0 * s2;
      double gT2 = gsub0 * dm + gsup0 * dp - gdiag0 * (dm + dp);
      V.w(W_NU, i) += 8.0 * M_PI * V.t(AT_ROW, i) * gdiag0;
      const double u = V.w(W_GSC, i) * V.t(AT_IU3, i) * K.kd;
      if (i != 0) { gT1 += u * K.lnee; glnee += u * V.w(W_T1, i); }
      gzi += u;
      V.w(W_GI4, i) = gT1;
      V.w(W_GI2, i) = gT2;
    }
    // zi = zl ne; zl = ZLOGei(ne, Te); lnee = LOGee(ne, Te)
    double gne = gzi * K.zl, gTe = 0.0;
    const double gzl = gzi * K.ne;
    double dne, dTe;
    zlog_ei_d(K.ne, K.Te, a.density_np, a.zeta, &dne, &dTe);
    gne += gzl * dne; gTe += gzl * dTe;
    log_ee_d(K.ne, K.Te, a.density_np, &dne, &dTe);
    gne += glnee * dne; gTe += glnee * dTe;
    // T1 = i0 + (2 j1 - i2)/3, T2 = (i2 + j1)/3 with i2 = 4pi CI2/v^2, i0 = 4pi CI0, j1 = 4pi v CJ1
    for (int i = 0; i < N; ++i) {
      const double gT1 = V.w(W_GI4, i), gT2 = V.w(W_GI2, i);
      const double gi0 = gT1, gj1 = gT1 * (2.0 / 3.0) + gT2 * (1.0 / 3.0), gi2 = gT1 * (-1.0 / 3.0) + gT2 * (1.0 / 3.0);
      V.w(W_GI4, i) = gi2 * (4.0 * M_PI) * V.t(AT_IV2, i);   // cotangent of CI2 (prefix U4)
      V.w(W_GI2, i) = gi0 * (4.0 * M_PI);                     // of CI0 (prefix U2)
      V.w(W_GJ1, i) = gj1 * (4.0 * M_PI * V.t(AT_VC, i));     // of CJ1 (suffix U1)
    }
    V.w(W_GI4, N - 1) += gTe * (4.0 * M_PI);
    V.w(W_GI2, N - 1) += gne * (4.0 * M_PI);
    double s4 = 0.0, s2 = 0.0;
    for (int q = N - 1; q >= 1; --q) {
      s4 += V.w(W_GI4, q);
      s2 += V.w(W_GI2, q);
      V.w(W_NU, q) += V.t(AT_U4, q) * s4 + V.t(AT_U2, q) * s2;
      V.w(W_NU, q - 1) += V.t(AT_U4M1, q) * s4 + V.t(AT_U2M1, q) * s2;
    }
    double s1 = 0.0;
    for (int q = 0; q <= N - 2; ++q) {
      s1 += V.w(W_GJ1, q);
      V.w(W_NU, q + 1) += V.t(AT_U1S, q) * s1;
      V.w(W_NU, q) += V.t(AT_U1M1S, q) * s1;
    }
  }
  // e-e sub-cycles in reverse; the running cotangent lives in W_LAM
  for (int p = 0; p < N; ++p) V.w(W_LAM, p) = V.w(W_NU, p);
  if (a.do_ee) {
    for (int s = a.numh - 1; s >= 0; --s) adj_ee_rk4_vjp(a, V, W_SUB + s, W_LAM, a.h_ee);
  }
  for (int p = 0; p < N; ++p) LAM(0, p) = V.w(W_LAM, p);
}

}  // namespace oshun

// ---- host side -------------------------------------------------------------------
namespace {

struct AdjState {
  bool ready = false;
  double* Y = nullptr;     // 3 stage inputs Y1..Y3 (planes), then K, Yrk
  double* Yf = nullptr;    // their fields: Y1..Y3, K, Yrk (6 nxy each)
  double* mu = nullptr;    // planes: mu, nu, acc
  double* muf = nullptr;   // fields: mu, nu, acc
  double* epart = nullptr;
  double* W = nullptr;
  int wslots = 0;
  double* tb = nullptr;
  double* jw = nullptr;
  // checkpointed runs
  std::vector<double*> ck;        // states every `every` steps (planes + fields + 3 clocks)
  std::vector<double> ck_t;       // clocks of the checkpoints
  double* seg = nullptr;          // step inputs of one segment
  std::vector<double> seg_t;
  int seg_cap = 0;
  double* yN = nullptr;           // final state of the recorded run
  double yN_t[3] = {0, 0, 0};
  double h = 0.0;
  int nsteps = 0, every = 0;
  long long generation = 0;  // of the recorded run: one more per oshun_adjoint_forward
};

std::map<const oshun_ctx*, AdjState>& adj_map() {
  static std::map<const oshun_ctx*, AdjState> m;
  return m;
}

long long adj_state_elems(const oshun_ctx* c) { return (long long)c->nc * c->plane; }

int adj_check(oshun_ctx* c) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (c->fp32 || c->ny != 1 || c->T.m0 != 0 || c->T.l0 < 1 || c->T.l0 > 8)
    return fail(c, OSHUN_ERR_ARG, "adjoint: 1D runs with m0 = 0, 1 <= l0 <= 8, FP64 only");
  return OSHUN_OK;
}

void adj_free(AdjState& s) {
  for (double* p : {s.Y, s.Yf, s.mu, s.muf, s.epart, s.W, s.tb, s.jw, s.seg, s.yN})
    if (p) (void)hipFree(p);
  for (double* p : s.ck) (void)hipFree(p);
  s = AdjState();
}

int adj_alloc(oshun_ctx* c, AdjState& s, int numh) {
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  const int slots = oshun::W_SUB + std::max(numh, 1);
  if (s.ready && slots <= s.wslots) return OSHUN_OK;
  if (s.W) { (void)hipFree(s.W); s.W = nullptr; }
  if (!s.ready) {
    HIPCHK(c, hipMalloc(&s.Y, sizeof(double) * S * 5));
    HIPCHK(c, hipMalloc(&s.Yf, sizeof(double) * Fn * 5));
    HIPCHK(c, hipMalloc(&s.mu, sizeof(double) * S * 3));
    HIPCHK(c, hipMalloc(&s.muf, sizeof(double) * Fn * 3));
    HIPCHK(c, hipMalloc(&s.epart, sizeof(double) * c->plane));
    // per-p tables of the collision adjoint
    const oshun::HostTables& T = c->T;
    const int np = c->np, nb = std::min(c->ph.nb_algorithms, np);
    std::vector<double> tb((size_t)oshun::AT_COUNT * np, 0.0);
    auto put = [&](int k, int p, double v) { tb[(size_t)k * np + p] = v; };
    for (int p = 0; p < np; ++p) {
      const double v = T.vc[p];
      put(oshun::AT_U4, p, T.U4[p]); put(oshun::AT_U4M1, p, T.U4m1[p]);
      put(oshun::AT_U2, p, T.U2[p]); put(oshun::AT_U2M1, p, T.U2m1[p]);
      put(oshun::AT_U1S, p, T.U1s[p]); put(oshun::AT_U1M1S, p, T.U1m1s[p]);
      put(oshun::AT_U3, p, T.U3[p]); put(oshun::AT_QN, p, T.Qn[p]); put(oshun::AT_PN, p, T.Pn[p]);
      put(oshun::AT_VC, p, v); put(oshun::AT_IV2, p, 1.0 / (v * v)); put(oshun::AT_IU3, p, 1.0 / std::pow(v, 3));
      const bool mid = p >= 1 && p < np - 1;
      put(oshun::AT_IVDM, p, mid ? T.ivdm[p] : 0.0); put(oshun::AT_IVDP, p, mid ? T.ivdp[p] : 0.0);
      put(oshun::AT_IVS2, p, mid ? T.ivs2[p] : 0.0); put(oshun::AT_ROW, p, p < np - 1 ? 1.0 : 0.0);
      if (p < nb) {
        put(oshun::AT_P3, p, T.pw3[p]); put(oshun::AT_A2, p, T.pw3[p] / 3.0);
        put(oshun::AT_B2, p, T.pw5[p] / T.v1sq * 0.2); put(oshun::AT_A4, p, T.pw5[p] * 0.2);
        put(oshun::AT_B4, p, T.pw7[p] / (T.v1sq * 7.0));
      }
    }
    if (upload_vec(c, &s.tb, tb)) return OSHUN_ERR_HIP;
    std::vector<double> jw(np);
    for (int p = 0; p < np; ++p) {
      const double w = (p == 0 || p == np - 1) ? 0.5 : 1.0;
      jw[p] = (4.0 * M_PI / 3.0) * (T.delta_p * w * T.p3og[p] + (p == 1 ? T.small * T.p3og[1] : 0.0));
    }
    if (upload_vec(c, &s.jw, jw)) return OSHUN_ERR_HIP;
    s.ready = true;
  }
  HIPCHK(c, hipMalloc(&s.W, sizeof(double) * (long long)slots * c->np * c->nx));
  s.wslots = slots;
  return OSHUN_OK;
}

dim3 adj_grid(long long n) { return dim3(static_cast<unsigned>((n + 255) / 256)); }

// F = the RK4 right-hand side (mode 0 stage + field stage) of (yin, fin) into (k, kf)
int adj_rhs(oshun_ctx* c, const double* yin, const double* fin, double* k, double* kf) {
  int rc = launch_stage<double>(c, 0, yin, nullptr, k, c->A, fin, 0.0);
  if (rc) return rc;
  return launch_field<double>(c, 0, fin, nullptr, kf, c->Af, yin, 0.0);
}

// nu = F'(y)^T mu (planes and fields)
template <int NL>
void adj_rhs_vjp_launch(const oshun::AdjRhsArgs& a, hipStream_t s) {
  hipLaunchKernelGGL((oshun::adj_rhs_1d<NL>), dim3((a.nx + 63) / 64, (a.np + 3) / 4), dim3(64, 4), 0, s, a);
}
int adj_rhs_vjp(oshun_ctx* c, AdjState& s, const double* y, const double* yf, const double* mu, const double* muf,
                double* nu, double* nuf) {
  oshun::AdjRhsArgs a;
  const oshun::HostTables& T = c->T;
  a.f = y; a.fld = yf; a.mu = mu; a.muf = muf; a.nu = nu; a.nuf = nuf; a.epart = s.epart;
  a.invpax = vt_ptr<double>(c, 0);
  a.cg = vt_ptr<double>(c, 2);
  a.hp0 = vt_ptr<double>(c, 3);
  a.xv1 = vt_ptr<double>(c, 4);
  a.xv2 = vt_ptr<double>(c, 5);
  a.eA1 = c->d_e1d;
  a.eA2 = c->d_e1d + (T.l0 + 1);
  a.jw = s.jw;
  a.p0p1_sq = T.p0p1_sq; a.inv_mp0p1_sq = T.inv_mp0p1_sq; a.g_r = T.g_r; a.midx = T.midx;
  a.np = c->np; a.nx = c->nx; a.plane = c->plane;
  switch (T.l0) {
    case 1: adj_rhs_vjp_launch<2>(a, c->stream); break;
    case 2: adj_rhs_vjp_launch<3>(a, c->stream); break;
    case 3: adj_rhs_vjp_launch<4>(a, c->stream); break;
    case 4: adj_rhs_vjp_launch<5>(a, c->stream); break;
    case 5: adj_rhs_vjp_launch<6>(a, c->stream); break;
    case 6: adj_rhs_vjp_launch<7>(a, c->stream); break;
    case 7: adj_rhs_vjp_launch<8>(a, c->stream); break;
    default: adj_rhs_vjp_launch<9>(a, c->stream); break;
  }
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(oshun::adj_field_1d, dim3((c->nx + 255) / 256), dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

// one step's adjoint at the context's current state (left unchanged)
int adj_step(oshun_ctx* c, AdjState& s, double h, double* lam, double* lamf) {
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  const int coll = c->ph.collisions;
  const int do_ee = (coll & OSHUN_COLL_EE) ? 1 : 0;
  const int do_flm = (coll & OSHUN_COLL_FLM) ? 1 : 0;
  const double tnew = c->t_rk + h;
  int numh = 0;
  double hee = 0.0;
  if (do_ee) ee_substeps(c, tnew, &numh, &hee);
  int rc = adj_alloc(c, s, numh);
  if (rc) return rc;
  double* Y0 = static_cast<double*>(c->S);
  double* Y0f = c->Sf;
  double* Ys[4] = {Y0, s.Y, s.Y + S, s.Y + 2 * S};
  double* Yfs[4] = {Y0f, s.Yf, s.Yf + Fn, s.Yf + 2 * Fn};
  double* K = s.Y + 3 * S;
  double* Kf = s.Yf + 3 * Fn;
  double* Yrk = s.Y + 4 * S;
  double* Yrkf = s.Yf + 4 * Fn;
  const hipStream_t st = c->stream;
  // ---- forward: stage inputs and the RK4 result
  const double cst[3] = {0.5 * h, 0.5 * h, h};
  for (int k = 0; k < 4; ++k) {
    if ((rc = adj_rhs(c, Ys[k], Yfs[k], K, Kf))) return rc;
    hipLaunchKernelGGL(oshun::adj_rk_acc, adj_grid(S), dim3(256), 0, st, S, k + 1, Y0, K, h, Yrk);
    hipLaunchKernelGGL(oshun::adj_rk_acc, adj_grid(Fn), dim3(256), 0, st, Fn, k + 1, Y0f, Kf, h, Yrkf);
    if (k < 3) {
      hipLaunchKernelGGL(oshun::adj_stage_in, adj_grid(S), dim3(256), 0, st, S, Y0, K, cst[k], Ys[k + 1]);
      hipLaunchKernelGGL(oshun::adj_stage_in, adj_grid(Fn), dim3(256), 0, st, Fn, Y0f, Kf, cst[k], Yfs[k + 1]);
    }
  }
  HIPCHK(c, hipGetLastError());
  // ---- collisions: lam (cotangent of the step output) -> cotangent of Yrk
  if (do_ee || do_flm) {
    oshun::AdjCollArgs a;
    a.frk = Yrk; a.lam = lam; a.W = s.W; a.tb = s.tb; a.comp_l = c->d_comp_l; a.plane = c->plane;
    a.np = c->np; a.nx = c->nx; a.nc = c->nc; a.l0 = c->T.l0; a.nb = std::min(c->ph.nb_algorithms, c->np);
    a.numh = numh; a.do_ee = do_ee; a.do_f1 = do_flm; a.do_flm = do_flm;
    a.h_ee = hee; a.dt_flm = tnew - c->t_eb;
    a.c_kpre = c->T.c_kpre; a.kpre = c->T.kpre; a.density_np = c->ph.density_np; a.zeta = c->ph.zeta;
    a.r01 = c->T.v0sq / c->T.v1sq; a.g_den = c->T.g_den; a.v0sq = c->T.v0sq;
    hipLaunchKernelGGL(oshun::adj_collide_col, dim3((c->nx + 63) / 64), dim3(64), 0, st, a);
    HIPCHK(c, hipGetLastError());
  }
  // ---- RK4 transposed: lam_k4 = h/6 lam, lam_Y3 = F'(Y3)^T lam_k4, ...
  double* mu = s.mu;
  double* nu = s.mu + S;
  double* acc = s.mu + 2 * S;
  double* muf = s.muf;
  double* nuf = s.muf + Fn;
  double* accf = s.muf + 2 * Fn;
  HIPCHK(c, hipMemcpyAsync(acc, lam, sizeof(double) * S, hipMemcpyDeviceToDevice, st));
  HIPCHK(c, hipMemcpyAsync(accf, lamf, sizeof(double) * Fn, hipMemcpyDeviceToDevice, st));
  // (coefficient on lam, coefficient on the previous stage cotangent) per stage 3, 2, 1, 0
  const double ca[4] = {h / 6.0, h / 3.0, h / 3.0, h / 6.0};
  const double cb[4] = {0.0, h, 0.5 * h, 0.5 * h};
  for (int k = 3, i = 0; k >= 0; --k, ++i) {
    hipLaunchKernelGGL(oshun::adj_axpby, adj_grid(S), dim3(256), 0, st, S, ca[i], lam, cb[i], i ? nu : nullptr, mu);
    hipLaunchKernelGGL(oshun::adj_axpby, adj_grid(Fn), dim3(256), 0, st, Fn, ca[i], lamf, cb[i], i ? nuf : nullptr,
                       muf);
    if ((rc = adj_rhs_vjp(c, s, Ys[k], Yfs[k], mu, muf, nu, nuf))) return rc;
    hipLaunchKernelGGL(oshun::adj_axpby, adj_grid(S), dim3(256), 0, st, S, 1.0, acc, 1.0, nu, acc);
    hipLaunchKernelGGL(oshun::adj_axpby, adj_grid(Fn), dim3(256), 0, st, Fn, 1.0, accf, 1.0, nuf, accf);
  }
  HIPCHK(c, hipMemcpyAsync(lam, acc, sizeof(double) * S, hipMemcpyDeviceToDevice, st));
  HIPCHK(c, hipMemcpyAsync(lamf, accf, sizeof(double) * Fn, hipMemcpyDeviceToDevice, st));
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

int adj_copy_state(oshun_ctx* c, double* dst, const double* src_f, const double* src_fld) {
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  HIPCHK(c, hipMemcpyAsync(dst, src_f, sizeof(double) * S, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dst + S, src_fld, sizeof(double) * Fn, hipMemcpyDeviceToDevice, c->stream));
  return OSHUN_OK;
}
int adj_restore_state(oshun_ctx* c, const double* src) {
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  HIPCHK(c, hipMemcpyAsync(c->S, src, sizeof(double) * S, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->Sf, src + S, sizeof(double) * Fn, hipMemcpyDeviceToDevice, c->stream));
  return OSHUN_OK;
}

}  // namespace

extern "C" {

int oshun_adjoint_step(oshun_ctx* c, double h, double* lam_f, double* lam_fld) {
  int rc = adj_check(c);
  if (rc) return rc;
  if (!lam_f || !lam_fld) return fail(c, OSHUN_ERR_ARG, "adjoint: null cotangent");
  return adj_step(c, adj_map()[c], h, lam_f, lam_fld);
}

int oshun_adjoint_forward(oshun_ctx* c, double h, int nsteps, int every, long long* generation) {
  int rc = adj_check(c);
  if (rc) return rc;
  if (nsteps < 1 || every < 1) return fail(c, OSHUN_ERR_ARG, "adjoint: nsteps and every must be >= 1");
  AdjState& s = adj_map()[c];
  for (double* p : s.ck) (void)hipFree(p);
  s.ck.clear();
  s.ck_t.clear();
  ++s.generation;  // the previous run's checkpoints are gone
  if (generation) *generation = s.generation;
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  s.h = h; s.nsteps = nsteps; s.every = every;
  for (int i = 0; i < nsteps; ++i) {
    if (i % every == 0) {
      double* p = nullptr;
      HIPCHK(c, hipMalloc(&p, sizeof(double) * (S + Fn)));
      s.ck.push_back(p);
      s.ck_t.insert(s.ck_t.end(), {c->t_rk, c->t_ee, c->t_eb});
      if ((rc = adj_copy_state(c, p, static_cast<double*>(c->S), c->Sf))) return rc;
    }
    if ((rc = full_step(c, h))) return rc;
  }
  if (!s.yN) HIPCHK(c, hipMalloc(&s.yN, sizeof(double) * (S + Fn)));
  if ((rc = adj_copy_state(c, s.yN, static_cast<double*>(c->S), c->Sf))) return rc;
  s.yN_t[0] = c->t_rk; s.yN_t[1] = c->t_ee; s.yN_t[2] = c->t_eb;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_adjoint_backward(oshun_ctx* c, long long generation, double* lam_f, double* lam_fld) {
  int rc = adj_check(c);
  if (rc) return rc;
  AdjState& s = adj_map()[c];
  if (s.ck.empty() || !s.yN) return fail(c, OSHUN_ERR_STATE, "adjoint: no recorded run (oshun_adjoint_forward)");
  if (generation > 0 && generation != s.generation)
    return fail(c, OSHUN_ERR_STATE, "adjoint: run " + std::to_string(generation) +
                                        " was replaced by a later oshun_adjoint_forward on this context (run " +
                                        std::to_string(s.generation) + ")");
  if (!lam_f || !lam_fld) return fail(c, OSHUN_ERR_ARG, "adjoint: null cotangent");
  const long long S = adj_state_elems(c), Fn = 6LL * c->nxy;
  if (s.seg_cap < s.every) {
    if (s.seg) (void)hipFree(s.seg);
    HIPCHK(c, hipMalloc(&s.seg, sizeof(double) * (S + Fn) * s.every));
    s.seg_cap = s.every;
  }
  for (int k = static_cast<int>(s.ck.size()) - 1; k >= 0; --k) {
    const int i0 = k * s.every, i1 = std::min(s.nsteps, i0 + s.every);
    // recompute the segment's step inputs from its checkpoint
    if ((rc = adj_restore_state(c, s.ck[k]))) return rc;
    c->t_rk = s.ck_t[3 * k]; c->t_ee = s.ck_t[3 * k + 1]; c->t_eb = s.ck_t[3 * k + 2];
    s.seg_t.assign(3 * (i1 - i0), 0.0);
    for (int i = i0; i < i1; ++i) {
      const int j = i - i0;
      if ((rc = adj_copy_state(c, s.seg + (S + Fn) * j, static_cast<double*>(c->S), c->Sf))) return rc;
      s.seg_t[3 * j] = c->t_rk; s.seg_t[3 * j + 1] = c->t_ee; s.seg_t[3 * j + 2] = c->t_eb;
      if (i + 1 < i1 && (rc = full_step(c, s.h))) return rc;
    }
    for (int i = i1 - 1; i >= i0; --i) {
      const int j = i - i0;
      if ((rc = adj_restore_state(c, s.seg + (S + Fn) * j))) return rc;
      c->t_rk = s.seg_t[3 * j]; c->t_ee = s.seg_t[3 * j + 1]; c->t_eb = s.seg_t[3 * j + 2];
      if ((rc = adj_step(c, s, s.h, lam_f, lam_fld))) return rc;
    }
  }
  // back to the end of the recorded run
  if ((rc = adj_restore_state(c, s.yN))) return rc;
  c->t_rk = s.yN_t[0]; c->t_ee = s.yN_t[1]; c->t_eb = s.yN_t[2];
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_adjoint_release(oshun_ctx* c) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  auto it = adj_map().find(c);
  if (it != adj_map().end()) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    adj_free(it->second);
    adj_map().erase(it);
  }
  return OSHUN_OK;
}

}  // extern "C"
