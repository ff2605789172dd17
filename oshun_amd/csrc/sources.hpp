// Laser heating, inverse bremsstrahlung and the external laser field of
// OSHUN-1's def-PLSource.cpp, on the device.  Included at the end of
// oshun_hip.hip (uses its context and helpers).
//
//   pls_kernel   PLaserSource::Laser_Source (:138-252): per (x,y) cell the
//                density and temperature of f00, the unit Maxwellian at the
//                heated temperature (Make_Gaussian_f, :1057-1124), and f00
//                relaxed toward it at the rate alpha
//   ib_kernel    InverseBremsstrahlung::loop (:780-861): per cell numh RK4
//                sub-steps of IB_Operator::Getslope (:550-608), then the
//                density restored
//   efld_kernel  ExternalEfield::Laser_Efield (:368-418)
//
// These operators are per-cell sequences of p sums and p-local updates,
// applied once per step.  One wavefront per (x,y) cell: the cell's column
// (and, for the inverse bremsstrahlung sub-steps, the integrated f00, the two
// RK4 stage inputs and the slope) in LDS, each p sum in one lane in the
// reference's order, the p-local work across the lanes.  Per-p and per-cell
// tables are formed on the host with the reference's expressions
// (oshun_set_sources).

namespace oshun {

// table rows in SrcArgs::tab (host layout in oshun_set_sources)
enum {
  ST_P = 0,     // p
  ST_G,         // gamma
  ST_P4,        // pow(p, 4)
  ST_G2,        // pow(gamma, 2)
  ST_W4,        // pow(p, 4) / pow(gamma, 2)
  ST_U4, ST_U4M1, ST_U2, ST_U2M1,  // IB_Operator integral weights on v = p/gamma (:514-519)
  ST_IUAV6,     // Inv_Uav6 (:523-526; the last entry is never read)
  ST_QN, ST_PN, // Qn, Pn (:529-537; 0 at np-1)
  ST_U2P, ST_U2PM1,  // the renormalisation weights on p (:902-905)
  ST_NROWS
};

template <typename T>
struct SrcArgs {
  T* f00;                  // plane 0 of the state
  const double* bg;        // FP32 split: f00 background per p (null in FP64)
  const double* tab;       // ST_NROWS rows of np
  const double *ILx, *ILy, *Ty, *ILxe, *ILye, *xpos;
  double* fld;             // EFLD: the field plane written
  int* flag;               // set when a heating rate alpha leaves [0, 0.99]
  int np, nx, ny, numh;
  int cx0, cy0, cnx, cny;  // InverseBremsstrahlung's cells: the interior (def-PLSource.cpp:838-839)
  double dp, p0p1_sq, inv_mp0p1_sq;
  // PLaserSource
  double ILt, dt, IL, IL_sqrt, IL_min, acoeff, nmin;
  // InverseBremsstrahlung
  double h, vos, vw_coeff_cube, Qn_coeff, Inv_Uav6_nm1, Pnm1, p0overp1_sq, density_np, zeta;
  // ExternalEfield
  double env, tnew, w0overwp;
};

template <typename T>
__device__ __forceinline__ double src_row(const SrcArgs<T>& a, int k, int p) {
  return a.tab[k * a.np + p];
}

// The p sums of the operators below: the lanes form the terms (each as the
// reference forms it), then one lane adds them in the reference's order, so
// every sum is the reference's to the last bit; the p-local work runs across
// the wavefront's lanes.  LDS holds the cell's column and the terms.
constexpr int kSrcLanes = 64;

// the ordered sum t[b] + t[b+1] + ... + t[e-1] onto s (one lane)
__device__ __forceinline__ double src_tsum(double s, const double* t, int b, int e) {
#pragma unroll 8
  for (int n = b; n < e; ++n) s += t[n];
  return s;
}

template <typename T>
__global__ __launch_bounds__(kSrcLanes) void pls_kernel(SrcArgs<T> a) {
  // the column F(p), the unit Maxwellian gauss(p), two term columns
  extern __shared__ double pls_sh[];
  __shared__ double red[4];
  const int N = a.np, lane = threadIdx.x;
  const int i = blockIdx.x;
  const int ix = i % a.nx, iy = i / a.nx;
  double* Fc = pls_sh;
  double* gs = pls_sh + N;
  double* tA = pls_sh + 2 * N;
  double* tB = pls_sh + 3 * N;
  const double* pr = a.tab + ST_P * N;
  const double* P4 = a.tab + ST_P4 * N;
  const double* G2 = a.tab + ST_G2 * N;
  const double* W4 = a.tab + ST_W4 * N;
  T* col = a.f00 + (long long)iy * N * a.nx + ix;
  for (int p = lane; p < N; p += kSrcLanes) {
    const double F = static_cast<double>(col[(long long)p * a.nx]) + (a.bg ? a.bg[p] : 0.0);
    Fc[p] = F;
    tA[p] = pr[p] * pr[p] * F;  // the density sum's terms
    tB[p] = W4[p] * F;          // the temperature sum's
  }
  __syncthreads();
  if (lane == 0) {  // density (:149-170)
    double n0 = Fc[0] * pr[0] * pr[0];
    n0 += Fc[N - 1] * pr[N - 1] * pr[N - 1];
    n0 *= 0.5;
    n0 = src_tsum(n0, tA, 1, N - 1);
    double f00 = Fc[1] / 5.0 * a.p0p1_sq;
    f00 += (Fc[0] - Fc[1] * a.p0p1_sq) * a.inv_mp0p1_sq * (1.0 / 3.0 - 1.0 / 5.0 * a.p0p1_sq);
    f00 *= pr[0] * (pr[0] / a.dp) * pr[0];
    n0 += f00;
    n0 *= 4.0 * M_PI * a.dp;
    red[0] = n0;
  } else if (lane == 1) {  // temperature in eV (:176-206), before the division by n0
    double vsq = Fc[0] * P4[0] / G2[0];
    vsq += Fc[N - 1] * P4[N - 1] / G2[N - 1];
    vsq *= 0.5;
    vsq = src_tsum(vsq, tB, 1, N - 1);
    double f00 = Fc[1] / 7.0 * a.p0p1_sq;
    f00 += (Fc[0] - Fc[1] * a.p0p1_sq) * a.inv_mp0p1_sq * (1.0 / 5.0 - 1.0 / 7.0 * a.p0p1_sq);
    f00 *= P4[0] * (pr[0] / a.dp);
    vsq += f00;
    vsq *= 4.0 * M_PI / 3.0 * a.dp;
    red[1] = vsq;
  }
  __syncthreads();
  const double n0 = red[0];
  double vsq = red[1];
  vsq /= n0;
  vsq *= (3000.0 / 4.19) * (3000.0 / 4.19);
  // Make_Gaussian_f (:1057-1124) at T_hot = Ty + T
  double T_hot = a.Ty[iy] + vsq;
  const double PTH = 4.19 / 3000.0 * sqrt(T_hot);
  const double e_c = -0.5 / (PTH * PTH);
  const double g_c = 1.0 / pow(sqrt(2 * M_PI) * PTH, 3);
  const bool hot = PTH > pr[1];
  if (hot) {
    __syncthreads();  // lanes 0 and 1 have read the term columns
    for (int p = lane; p < N; p += kSrcLanes) {
      const double g = exp((pr[p] * pr[p]) * e_c) * g_c;
      gs[p] = g;
      tA[p] = g * pr[p] * pr[p];  // its density sum's terms
    }
    __syncthreads();
    if (lane == 0) {
      double g0 = gs[0] * pr[0] * pr[0];
      g0 += gs[N - 1] * pr[N - 1] * pr[N - 1];
      g0 *= 0.5;
      g0 = src_tsum(g0, tA, 1, N - 1);
      double gf = gs[1] / 5.0 * a.p0p1_sq;
      gf += (gs[0] - gs[1] * a.p0p1_sq) * a.inv_mp0p1_sq * (1.0 / 3.0 - 1.0 / 5.0 * a.p0p1_sq);
      gf *= pr[0] * (pr[0] / a.dp) * pr[0];
      g0 += gf;
      g0 *= 4.0 * M_PI * a.dp;
      red[2] = 1.0 / g0;  // 1 / (its density before normalisation)
    }
    __syncthreads();
    const double gnorm = red[2];
    auto fg = [&](int p) { return gs[p] * gnorm; };
    for (int p = lane; p < N; p += kSrcLanes) tB[p] = fg(p) * P4[p] / G2[p];  // its temperature sum's terms
    __syncthreads();
    if (lane == 0) {
      double vt = fg(0) * P4[0] / G2[0];
      vt += fg(N - 1) * P4[N - 1] / G2[N - 1];
      vt *= 0.5;
      vt = src_tsum(vt, tB, 1, N - 1);
      double gf = fg(1) / 7.0 * a.p0p1_sq;
      gf += (fg(0) - fg(1) * a.p0p1_sq) * a.inv_mp0p1_sq * (1.0 / 5.0 - 1.0 / 7.0 * a.p0p1_sq);
      gf *= P4[0] * (pr[0] / a.dp);
      vt += gf;
      vt *= 4.0 * M_PI / 3.0 * a.dp;
      red[3] = vt * (3000.0 / 4.19) * (3000.0 / 4.19);
    }
    __syncthreads();
    T_hot = red[3];
  } else {
    T_hot = -1.0;
  }
  if ((n0 > a.nmin) && (a.ILy[iy] * a.ILx[ix] * a.ILt * a.IL > a.IL_min) && (T_hot > vsq)) {
    double alpha = a.acoeff * a.IL_sqrt * a.dt / n0;
    alpha *= a.ILx[ix] * sqrt(a.ILy[iy] * a.ILt);
    if ((alpha < 0) || (alpha > 0.99)) {  // the reference stops the run here (:230-233)
      if (lane == 0) atomicOr(a.flag, 1);
      return;
    }
    const double gnorm = red[2];
    for (int p = lane; p < N; p += kSrcLanes) {
      double v = Fc[p];
      v *= (1.0 - alpha);
      v += alpha * (gs[p] * gnorm) * n0;
      col[(long long)p * a.nx] = static_cast<T>(v - (a.bg ? a.bg[p] : 0.0));
    }
  }
}

// ZLOGei_v2 (:42-67)
__device__ __forceinline__ double zlogei_v2(double ne, double Te, double density_np, double zeta) {
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

// the terms A[n] in[n] + B[n] in[n-1], n = 1..N-1, of the IB_Operator sums into t (all lanes)
__device__ __forceinline__ void ib_terms(const double* A, const double* B, const double* in, double* t, int N) {
  for (int n = 1 + (int)threadIdx.x; n < N; n += kSrcLanes) t[n] = A[n] * in[n] + B[n] * in[n - 1];
}

// IB_Operator::Getslope (:550-608) of the column `in` into fh (rows 0..np-2),
// both in LDS; lanes 0 and 1 add the sums I4 and I2 from the term columns t4, t2
template <typename T>
__device__ void ib_slope(const SrcArgs<T>& a, const double* in, double* fh, double* red, double* t4, double* t2,
                         double vos) {
  const int N = a.np, lane = threadIdx.x;
  ib_terms(a.tab + ST_U4 * N, a.tab + ST_U4M1 * N, in, t4, N);
  ib_terms(a.tab + ST_U2 * N, a.tab + ST_U2M1 * N, in, t2, N);
  __syncthreads();
  if (lane < 2) red[lane] = src_tsum(0.0, lane == 0 ? t4 : t2, 1, N);
  __syncthreads();
  const double I4 = red[0], I2 = red[1];
  const double f00 = (in[0] - in[1] * a.p0overp1_sq) / (1.0 - a.p0overp1_sq);
  const double ZLn_ei = zlogei_v2(4.0 * M_PI * I2, 4.0 * M_PI * I4, a.density_np, a.zeta);
  const double vw_cube = a.vw_coeff_cube * ZLn_ei * 4.0 * M_PI * I2;
  auto gn = [&](int i) {
    double g = src_row(a, ST_IUAV6, i) * vw_cube * vw_cube;
    return 1.0 / (1.0 + g);
  };
  const double qn_c = a.Qn_coeff * vos * vos * ZLn_ei * 4.0 * M_PI * I2;
  for (int i = lane; i < N - 1; i += kSrcLanes) {
    if (i == 0) {
      fh[0] = qn_c * src_row(a, ST_QN, 0) *
              (src_row(a, ST_PN, 0) * gn(0) * (in[1] - in[0]) -
               a.Pnm1 / (1.0 + a.Inv_Uav6_nm1 * vw_cube * vw_cube) * (in[0] - f00));
    } else {
      fh[i] = qn_c * src_row(a, ST_QN, i) *
              (src_row(a, ST_PN, i) * gn(i) * (in[i + 1] - in[i]) -
               src_row(a, ST_PN, i - 1) * gn(i - 1) * (in[i] - in[i - 1]));
    }
  }
  __syncthreads();
}

// InverseBremsstrahlung::loop (:780-861) of one interior cell per workgroup
template <typename T>
__global__ __launch_bounds__(kSrcLanes) void ib_kernel(SrcArgs<T> a) {
  // the integrated column fc, the stage inputs f0, f1, the slope fh and two
  // term columns
  extern __shared__ double ib_sh[];
  __shared__ double red[2];
  const int N = a.np, lane = threadIdx.x;
  const int w = blockIdx.x;
  const int ix = a.cx0 + w % a.cnx, iy = a.cy0 + w / a.cnx;
  double* fc = ib_sh;
  double* f0 = ib_sh + N;
  double* f1 = ib_sh + 2 * N;
  double* fh = ib_sh + 3 * N;
  double* t4 = ib_sh + 4 * N;
  double* t2 = ib_sh + 5 * N;
  T* col = a.f00 + (long long)iy * N * a.nx + ix;
  const double vos = a.vos * sqrt(a.ILxe[ix] * a.ILye[iy]);
  for (int p = lane; p < N; p += kSrcLanes) {
    fc[p] = static_cast<double>(col[(long long)p * a.nx]) + (a.bg ? a.bg[p] : 0.0);
    fh[p] = 0.0;  // row np-1 is never written by Getslope
  }
  __syncthreads();
  ib_terms(a.tab + ST_U2P * N, a.tab + ST_U2PM1 * N, fc, t4, N);
  __syncthreads();
  if (lane == 0) red[0] = src_tsum(0.0, t4, 1, N);
  __syncthreads();
  const double I2_before = red[0];
  const double h = a.h;
  for (int s = 0; s < a.numh; ++s) {  // RungeKutta4_ib::advance (:713-757)
    for (int p = lane; p < N; p += kSrcLanes) { f0[p] = fc[p]; f1[p] = fc[p]; }
    __syncthreads();
    ib_slope(a, f1, fh, red, t4, t2, vos);
    for (int p = lane; p < N; p += kSrcLanes) {
      double q = fh[p];
      q *= (0.5 * h); f1[p] += q;
      q *= (1.0 / 3.0); fc[p] += q;
      fh[p] = q;
    }
    __syncthreads();
    ib_slope(a, f1, fh, red, t4, t2, vos);
    for (int p = lane; p < N; p += kSrcLanes) {
      double q = fh[p];
      f1[p] = f0[p];
      q *= (0.5 * h); f1[p] += q;
      q *= (2.0 / 3.0); fc[p] += q;
      fh[p] = q;
    }
    __syncthreads();
    ib_slope(a, f1, fh, red, t4, t2, vos);
    for (int p = lane; p < N; p += kSrcLanes) {
      double q = fh[p];
      q *= h; f0[p] += q;
      q *= (1.0 / 3.0); fc[p] += q;
      fh[p] = q;
    }
    __syncthreads();
    ib_slope(a, f0, fh, red, t4, t2, vos);
    for (int p = lane; p < N; p += kSrcLanes) {
      double q = fh[p];
      q *= (h / 6.0); fc[p] += q;
      fh[p] = q;
    }
    __syncthreads();
  }
  ib_terms(a.tab + ST_U2P * N, a.tab + ST_U2PM1 * N, fc, t4, N);
  __syncthreads();
  if (lane == 0) red[0] = src_tsum(0.0, t4, 1, N);
  __syncthreads();
  const double r = I2_before / red[0];
  for (int p = lane; p < N; p += kSrcLanes) {
    double v = fc[p];
    v *= r;
    col[(long long)p * a.nx] = static_cast<T>(v - (a.bg ? a.bg[p] : 0.0));
  }
}

__global__ __launch_bounds__(256) void efld_kernel(SrcArgs<double> a) {
  const int nxy = a.nx * a.ny;
  const int cell = blockIdx.x * 256 + threadIdx.x;
  if (cell >= nxy) return;
  const int i = cell % a.nx, j = cell / a.nx;
  a.fld[cell] = a.vos * a.ILxe[i] * a.ILye[j] * a.env * sin(a.w0overwp * (a.tnew - a.xpos[i]));
}

}  // namespace oshun

namespace {

// the envelope times (t2 - t1) of PLaserSource / ExternalEfield (def-PLSource.cpp:259-289, :428-458)
double src_t_profile(const oshun_sources& s, double t2, double t1) {
  double t;
  if (t2 < s.rise_time) {
    t = t2 / s.rise_time;
    if (s.polynomial_Ivst) return (10 * std::pow(t, 3) - 15 * std::pow(t, 4) + 6 * std::pow(t, 5)) * (t2 - t1);
    if (s.linear_Ivst) return t * (t2 - t1);
    return 0;
  }
  if (t2 < s.rise_time + s.flat_time) return t2 - t1;
  if (t2 < s.rise_time + s.flat_time + s.fall_time) {
    t = (s.rise_time + s.flat_time + s.fall_time - t2) / s.fall_time;
    if (s.polynomial_Ivst) return (10 * std::pow(t, 3) - 15 * std::pow(t, 4) + 6 * std::pow(t, 5)) * (t2 - t1);
    if (s.linear_Ivst) return t * (t2 - t1);
    return 0;
  }
  return -1.0;
}

template <typename T>
oshun::SrcArgs<T> src_args(oshun_ctx* c) {
  oshun::SrcArgs<T> a{};
  a.f00 = static_cast<T*>(c->S) + (size_t)c->T.comp_re[0] * c->plane;
  a.bg = c->d_bg;
  const size_t nx = c->nx, ny = c->ny, np = c->np;
  a.tab = c->d_src;
  const double* b = c->d_src + oshun::ST_NROWS * np;
  a.ILx = b; a.ILy = b + nx; a.Ty = b + nx + ny; a.ILxe = b + nx + 2 * ny; a.ILye = b + 2 * nx + 2 * ny;
  a.xpos = b + 2 * nx + 3 * ny;
  a.flag = c->d_srcflag;
  a.np = c->np; a.nx = c->nx; a.ny = c->ny;
  a.cx0 = c->gn; a.cy0 = c->gn; a.cnx = c->nxi; a.cny = c->nyi;
  const double p0 = c->src_p0, p1 = c->src_p1;
  a.dp = p1 - p0;
  a.p0p1_sq = p0 * p0 / (p1 * p1);
  a.inv_mp0p1_sq = 1.0 / (1.0 - a.p0p1_sq);
  a.IL = c->src_IL; a.IL_sqrt = c->src_IL_sqrt; a.IL_min = c->src_IL_min; a.acoeff = c->src_acoeff; a.nmin = 0.00001;
  a.vos = c->src_vos; a.vw_coeff_cube = c->src_vw_coeff_cube; a.Qn_coeff = c->src_Qn_coeff;
  a.Inv_Uav6_nm1 = c->src_Inv_Uav6_nm1; a.Pnm1 = c->src_Pnm1;
  double q = p0 / p1;
  q *= q;
  a.p0overp1_sq = q;
  a.density_np = c->ph.density_np; a.zeta = c->ph.zeta;
  a.w0overwp = c->src_w0overwp;
  return a;
}

template <typename T>
int launch_pls(oshun_ctx* c, double tnew) {
  const double I0_t = src_t_profile(c->src, tnew, c->src_t_pls);
  const double dt = tnew - c->src_t_pls;
  const double ILt = I0_t / dt;
  c->src_t_pls = tnew;
  if (I0_t < 0) return OSHUN_OK;
  oshun::SrcArgs<T> a = src_args<T>(c);
  a.ILt = ILt;
  a.dt = dt;
  hipLaunchKernelGGL(oshun::pls_kernel<T>, dim3(c->nxy), dim3(oshun::kSrcLanes), 4 * c->np * sizeof(double), c->stream,
                     a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

template <typename T>
int launch_ib(oshun_ctx* c, double tnew) {
  oshun::SrcArgs<T> a = src_args<T>(c);
  const int numh = static_cast<int>(static_cast<size_t>(static_cast<int>((tnew - c->src_t_ib) / c->src.smaller_dt)) + 1);
  a.numh = numh;
  a.h = (tnew - c->src_t_ib) / static_cast<double>(static_cast<size_t>(numh));
  c->src_t_ib = tnew;
  hipLaunchKernelGGL(oshun::ib_kernel<T>, dim3(c->nxi * c->nyi), dim3(oshun::kSrcLanes),
                     6 * c->np * sizeof(double), c->stream, a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

// the sources of one step of main_loop (main.cpp:90-92), after the e-e collisions
int step_sources(oshun_ctx* c, double t) {
  int rc = OSHUN_OK;
  if (c->src_pls) rc = c->fp32 ? launch_pls<float>(c, t) : launch_pls<double>(c, t);
  if (!rc && c->src_ib == 1) rc = c->fp32 ? launch_ib<float>(c, t) : launch_ib<double>(c, t);
  return rc;
}

// a heating rate alpha outside [0, 0.99] since the flag was last cleared (the reference exits)
int check_source_flag(oshun_ctx* c) {
  if (!c->d_srcflag) return OSHUN_OK;
  int flag = 0;
  HIPCHK(c, hipMemcpyAsync(&flag, c->d_srcflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (flag) {
    HIPCHK(c, hipMemsetAsync(c->d_srcflag, 0, sizeof(int), c->stream));
    return fail(c, OSHUN_ERR_STATE, "laser source: a heating rate alpha left [0, 0.99] (def-PLSource.cpp:230-233 stops "
                                    "the run there); those cells were not heated");
  }
  return OSHUN_OK;
}

}  // namespace

extern "C" {

int oshun_set_sources(oshun_ctx* c, const oshun_sources* s, const double* x, const double* y, const double* xglob,
                      int nxglob) {
  if (!sane(c) || !s || !x || !y || !xglob || nxglob < 2) return OSHUN_ERR_ARG;
  if (!(s->lambda_0 > 0.0) || !(s->ab_depth != 0.0) || !(s->spot_w0 != 0.0) ||
      (s->inverse_bremsstrahlung == 1 && !(s->smaller_dt > 0.0)))
    return fail(c, OSHUN_ERR_ARG, "oshun_set_sources: lambda_0 > 0, ab_depth and spot_w0 nonzero, smaller_dt > 0");
  c->src = *s;
  const int nx = c->nx, ny = c->ny, np = c->np;
  const double dens = c->ph.density_np;
  // PLaserSource (def-PLSource.cpp:73-133)
  c->src_IL = s->I_0 * s->ab_eta;
  c->src_IL_sqrt = std::sqrt(c->src_IL);
  c->src_IL_min = 0.00001 * c->src_IL;
  const double T0 = 1000.0 * std::sqrt(1.0e-14 * s->lambda_0 * s->lambda_0);
  c->src_acoeff = 2.08050325 * 1.0e+12 / ((xglob[1] - xglob[0]) * dens * s->lambda_0);
  double sumI = 0.0;
  for (int i = 0; i < nxglob; ++i) sumI += std::exp((-1.0) * std::pow((xglob[i] - s->cntr_x0) / s->ab_depth, 2));
  std::vector<double> ILx(nx), ILy(ny), Ty(ny), ILxe(nx), ILye(ny);
  for (int i = 0; i < nx; ++i) {
    ILx[i] = std::exp((-1.0) * std::pow((x[i] - s->cntr_x0) / s->ab_depth, 2));
    ILxe[i] = ILx[i];  // inverse bremsstrahlung and the external field: unnormalised (:336-338, :891-893)
    ILx[i] *= 1.0 / sumI;
  }
  for (int j = 0; j < ny; ++j) {
    const double tmp2 = (y[j] - s->cntr_r0) / s->spot_w0;
    ILy[j] = std::exp(-2.0 * tmp2 * tmp2);
    ILye[j] = ILy[j];
    Ty[j] = T0 * std::sqrt(ILy[j] * c->src_IL);
  }
  // ExternalEfield and IB_Operator constants (:308-318, :476-545, :861-875)
  const double omega_0 = 3.0e+10 * 2.0 * M_PI / (1.0e-4 * s->lambda_0);
  const double omega_p = 5.64 * 1.0e+4 * std::sqrt(dens);
  c->src_w0overwp = omega_0 / omega_p;
  c->src_vos = s->lambda_0 * std::sqrt(7.3e-19 * s->I_0);
  const double re = 2.8179402894e-13;
  const double kp = std::sqrt(4.0 * M_PI * (dens) * re);
  c->src_vw_coeff_cube = omega_p / omega_0 * kp * re;
  c->src_Qn_coeff = kp * re / 6.0;
  c->src_pls = (s->linear_Ivst || s->polynomial_Ivst) ? 1 : 0;
  c->src_ib = std::abs(s->inverse_bremsstrahlung) > 1 ? 0 : s->inverse_bremsstrahlung;
  c->src_pol = std::abs(s->polarization_direction) > 3 ? 0 : s->polarization_direction;
  // per-p tables
  const std::vector<double>& pr = c->T.pd;
  c->src_p0 = pr[0];
  c->src_p1 = pr[1];
  std::vector<double> tab((size_t)oshun::ST_NROWS * np, 0.0);
  auto row = [&](int k) { return &tab[(size_t)k * np]; };
  std::vector<double> v(np);
  for (int i = 0; i < np; ++i) {
    const double g = std::sqrt(1.0 + pr[i] * pr[i]);
    row(oshun::ST_P)[i] = pr[i];
    row(oshun::ST_G)[i] = g;
    row(oshun::ST_P4)[i] = std::pow(pr[i], 4);
    row(oshun::ST_G2)[i] = std::pow(g, 2);
    row(oshun::ST_W4)[i] = std::pow(pr[i], 4) / std::pow(g, 2);
    v[i] = pr[i];
    v[i] = v[i] / (std::sqrt(1.0 + v[i] * v[i]));
  }
  for (int i = 1; i < np; ++i) {
    row(oshun::ST_U4)[i] = 0.5 * std::pow(v[i], 4) * (v[i] - v[i - 1]);
    row(oshun::ST_U4M1)[i] = 0.5 * std::pow(v[i - 1], 4) * (v[i] - v[i - 1]);
    row(oshun::ST_U2)[i] = 0.5 * v[i] * v[i] * (v[i] - v[i - 1]);
    row(oshun::ST_U2M1)[i] = 0.5 * v[i - 1] * v[i - 1] * (v[i] - v[i - 1]);
    row(oshun::ST_U2P)[i] = 0.5 * pr[i] * pr[i] * (pr[i] - pr[i - 1]);
    row(oshun::ST_U2PM1)[i] = 0.5 * pr[i - 1] * pr[i - 1] * (pr[i] - pr[i - 1]);
  }
  c->src_Inv_Uav6_nm1 = std::pow(2.0 / v[0], 6);
  for (int i = 0; i < np - 1; ++i) row(oshun::ST_IUAV6)[i] = std::pow(2.0 / (v[i + 1] + v[i]), 6);
  row(oshun::ST_QN)[0] = 1.0 / ((v[0] * v[0] * v[1]) / 2.0);
  for (int i = 1; i < np - 1; ++i) row(oshun::ST_QN)[i] = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]) / 2.0);
  c->src_Pnm1 = 2.0 / (v[0] * v[0]);
  for (int i = 0; i < np - 1; ++i) row(oshun::ST_PN)[i] = 1.0 / ((v[i + 1] - v[i]) / 2.0 * (v[i + 1] + v[i]));
  tab.insert(tab.end(), ILx.begin(), ILx.end());
  tab.insert(tab.end(), ILy.begin(), ILy.end());
  tab.insert(tab.end(), Ty.begin(), Ty.end());
  tab.insert(tab.end(), ILxe.begin(), ILxe.end());
  tab.insert(tab.end(), ILye.begin(), ILye.end());
  tab.insert(tab.end(), x, x + nx);
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->d_src) { (void)hipFree(c->d_src); c->d_src = nullptr; }
  if (upload_vec(c, &c->d_src, tab)) return fail(c, OSHUN_ERR_ALLOC, "oshun_set_sources: tables");
  if (!c->d_srcflag) {
    if (hipMalloc((void**)&c->d_srcflag, sizeof(int)) != hipSuccess)
      return fail(c, OSHUN_ERR_ALLOC, "oshun_set_sources: flag");
    HIPCHK(c, hipMemset(c->d_srcflag, 0, sizeof(int)));
  }
  c->src_t_pls = c->src_t_ib = c->src_t_efld = c->t_rk;
  c->src_on = true;
  return OSHUN_OK;
}

int oshun_laser_source(oshun_ctx* c, double tnew) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (!c->src_on) return fail(c, OSHUN_ERR_STATE, "oshun_laser_source: no oshun_set_sources on this context");
  int rc = c->fp32 ? launch_pls<float>(c, tnew) : launch_pls<double>(c, tnew);
  return rc ? rc : check_source_flag(c);
}

int oshun_inverse_bremsstrahlung(oshun_ctx* c, double tnew) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (!c->src_on) return fail(c, OSHUN_ERR_STATE, "oshun_inverse_bremsstrahlung: no oshun_set_sources on this context");
  return c->fp32 ? launch_ib<float>(c, tnew) : launch_ib<double>(c, tnew);
}

int oshun_laser_efield(oshun_ctx* c, double tnew) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (!c->src_on) return fail(c, OSHUN_ERR_STATE, "oshun_laser_efield: no oshun_set_sources on this context");
  const double dt = tnew - c->src_t_efld;
  const double env = src_t_profile(c->src, tnew, c->src_t_efld) / dt;
  if (c->src_pol >= 1 && c->src_pol <= 3) {
    oshun::SrcArgs<double> a = src_args<double>(c);
    a.fld = c->Sf + (size_t)(c->src_pol - 1) * c->nxy;
    a.env = env;
    a.tnew = tnew;
    hipLaunchKernelGGL(oshun::efld_kernel, dim3((c->nxy + 255) / 256), dim3(256), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
  }
  c->src_t_efld = tnew;
  return OSHUN_OK;
}

}  // extern "C"
