// implicit2d.hpp -- the implicit-E step of OSHUN-1 without if_implicit1D
// (the variant every production deck runs: implicit_E, RKLevel = 2,
// main.cpp:79-107 with Runge_Kutta_2_Imp and Implicit_E_Field) on the
// general (any l0 >= 3, m0 >= 2, 1D or 2D) FP64 path.  Included at the end of
// oshun_hip.hip.
//
// One step:
//   RK2 of Implicit_Actions::Adv1 (x/y-advection, B rotation, Faraday; the
//     last two not if_implicitES)          def-RK.cpp:470-498, def-Actions.cpp:118-132
//   e-e sub-cycles, sources                 main.cpp:87-92
//   Implicit_E_Field::advance               def-ImplicitE.cpp:472-606
//     probe fields DE = sqrt(|E|^2) sqrt(16 eps) + 16 eps on each component
//     (FindDE, :628-650); J0 = current after Euler_Backward::advance_1;
//     J_Ex, J_Ey, J_Ez = current after an RK2 of the probe couplings
//     (Implicit_Ex / Implicit_Eyz, def-VlasovMax.cpp:157-291) with one field
//     component set to its DE, and advance_1; JN from Ampere (:652-678), or,
//     if_implicitES, the fixed current of oshun_set_implicit_es_current;
//     per cell sgm(i,j) = (J_Ej - J0)_i / DE_j and E = sgm^-1 (JN - J0) by
//     Cramer's rule (Det33..Detz33, def-Nmethods.cpp:215-251).  Cells with
//     |det| <= 6 DBL_MIN get E = 0, DE.Ex *= 2, DE.Ey *= 1.3, and the probes
//     run again, at most 4 attempts in all (:486-489, 566-574).
//   RK2 of Adv2 (Exyz + CLEAN_LOW_PC)        def-RK.cpp:503-524, def-Actions.cpp:138-160
//   Euler_Backward::advance_1, advance_lm
//
// The reference runs its probes on the state and restores f00, f1m, f2m
// after each; here the state S is never touched by them: a probe's RK2 reads
// S and writes its planes l <= 2 to A (stage 1 writes Y0 of planes l <= 3 to
// B1), the l = 1 collisions then run on A, and the current is taken from A.
// The probe's J0 runs on a copy of S's planes l <= 1 in A.  A retry redoes
// the Ex and Ey probes only: J0 and J_Ez do not depend on what a retry
// changes.  Each current is Current_xyz::calculate_J (def-ImplicitE.cpp:
// 125-180), whose trapezoid end weight lands on a guard row, so interior
// cells sum p0 and pmax with weight 1.
//
// With mirror boundaries (whole arrays with guards) every kernel here runs on
// all cells, guards included, as the reference's loops do (the field solve
// over szx x szy, def-ImplicitE.cpp:536-577, counting zero determinants on
// guard cells too); the collisions stay on the interior; the magnetic-field
// guards are refreshed after the first RK2 half and all guards at the end.

namespace oshun {

struct ImpFarArgs {
  double* fld;  // state fields [6][nxy]
  double* y0;   // fields of the RK2's Y0 [6][nxy]
  int nx, ny;
  double midx, midy, h, h2;
  int ext;  // whole arrays with guards: flattened-array differences (ext_offsets)
};

// MaxwellEq::Faraday (def-VlasovMax.cpp:979-1005) for both RK2 stages of
// Adv1: E does not change in Adv1, so the two slopes are equal.  mode 0:
// Y0 fields (E, B + h k); mode 1: B = (B + (h k) / 2) + (h / 2) k in place.
__global__ __launch_bounds__(256) void imp_faraday_2d(ImpFarArgs a, int mode) {
  const int nx = a.nx, ny = a.ny;
  const long long nxy = (long long)nx * ny;
  const long long k = blockIdx.x * 256LL + threadIdx.x;
  if (k >= nxy) return;
  const int x = static_cast<int>(k % nx), y = static_cast<int>(k / nx);
  long long xm = (x == 0) ? k + nx - 1 : k - 1, xp = (x == nx - 1) ? k - (nx - 1) : k + 1;
  long long ym = (y == 0) ? k + (long long)(ny - 1) * nx : k - nx;
  long long yp = (y == ny - 1) ? k - (long long)(ny - 1) * nx : k + nx;
  bool xlast = false, ylast = false;
  if (a.ext) {
    long long oxm, oxp, oym, oyp;
    ext_offsets(x, y, nx, ny, nx, &oxm, &oxp, &oym, &oyp, &xlast, &ylast);
    xm = k + oxm; xp = k + oxp; ym = k + oym; yp = k + oyp;
  }
  const double* Ex = a.fld;
  const double* Ey = a.fld + nxy;
  const double* Ez = a.fld + 2 * nxy;
  auto X = [&](const double* F, long long i, bool zero) { return zero ? 0.0 : F[i]; };
  double k3 = 0.0, k4 = 0.0, k5 = 0.0;
  k3 += Ez[ym] * (-a.midy) - X(Ez, yp, ylast) * (-a.midy);
  k4 += Ez[xm] * a.midx - X(Ez, xp, xlast) * a.midx;
  k5 += Ex[ym] * a.midy - X(Ex, yp, ylast) * a.midy;
  k5 += Ey[xm] * (-a.midx) - X(Ey, xp, xlast) * (-a.midx);
  const double kc[3] = {k3, k4, k5};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double* B = a.fld + (3 + c) * nxy + k;
    const double kh = kc[c] * a.h;
    if (mode == 0) {
      a.y0[(3 + c) * nxy + k] = *B + kh;
      a.y0[c * nxy + k] = a.fld[c * nxy + k];
    } else {
      const double y1 = *B + kh * 0.5;
      *B = y1 + kc[c] * a.h2;
    }
  }
}

// FindDE (def-ImplicitE.cpp:628-650): one probe size per cell, written to the
// diagonal of the probe fields (probe d holds DE in component d, zero elsewhere)
__global__ __launch_bounds__(256) void imp_findde_2d(const double* fld, double* pf, long long nxy) {
  const long long k = blockIdx.x * 256LL + threadIdx.x;
  if (k >= nxy) return;
  const double Eps = 16.0 * 2.220446049250313e-16;  // 16 numeric_limits<double>::epsilon()
  const double LargeEps = sqrt(Eps);
  double ex = fld[k], ey = fld[nxy + k], ez = fld[2 * nxy + k];
  ex *= ex; ey *= ey; ez *= ez;
  ex += ey; ex += ez;
  ex = sqrt(ex);
  ex *= LargeEps;
  ex += Eps;
  pf[k] = ex;
  pf[7 * nxy + k] = ex;
  pf[14 * nxy + k] = ex;
}

struct ImpCurArgs {
  const double* f;  // state planes
  double* j;        // out [3][nxy]: Jx, Jy, Jz
  const double* p3og;
  long long plane;
  int c10, c11r, c11i, np, nx, ny;
  double delta_p, small;
  int ext;  // whole arrays: row 0 is the guard row the end-weight subscript lands on
};

// Current_xyz::calculate_J (def-ImplicitE.cpp:125-180): Jx from f10 with
// -4pi/3; Jy = Re and Jz = Re(i J) of the -8pi/3 sum over f11; each sum in
// the reference's order (p0 + pmax first, then p1..pmax-1).  The block's
// kImpCols2D columns x all p rows of the three planes are staged in LDS
// (3 np kImpCols2D doubles of dynamic shared memory) with every load in
// flight at once, then three threads per column run the three serial sums
// side by side.
constexpr int kImpCols2D = 8;
__global__ __launch_bounds__(256) void imp_current_2d(ImpCurArgs a) {
  extern __shared__ double chunk_lds[];
  const int tx = threadIdx.x % kImpCols2D, ty = threadIdx.x / kImpCols2D;
  const int x = blockIdx.x * kImpCols2D + tx, y = blockIdx.y;
  const int np = a.np, nx = a.nx;
  const bool ok = x < nx;
  const long long col = (long long)y * np * nx + (ok ? x : 0);
  const int comp[3] = {a.c10, a.c11r, a.c11i};
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double* f = a.f + (long long)comp[q] * a.plane + col;
#pragma unroll 8
    for (int ip = 1 + ty; ip < np - 1; ip += 256 / kImpCols2D)
      chunk_lds[(q * np + ip) * kImpCols2D + tx] = ok ? f[(long long)ip * nx] * a.p3og[ip] : 0.0;
  }
  __syncthreads();
  if (!ok || ty >= 3) return;
  const int q = ty;  // this thread's sum
  const double* f = a.f + (long long)comp[q] * a.plane + col;
  const double* chunk = chunk_lds + (long long)q * np * kImpCols2D + tx;  // row ip at chunk[ip * kImpCols2D]
  double j = f[0] * a.p3og[0];
  j += f[(long long)(np - 1) * nx] * a.p3og[np - 1];
  if (a.ext && y == 0) {
    // the guard row that jayx(ix) *= 0.5 addresses (an (ix, 0) subscript):
    // halved once before its p sum and once more for every later row
    j *= 0.5;
#pragma unroll 8
    for (int ip = 1; ip < np - 1; ++ip) j += chunk[ip * kImpCols2D];
    for (int r = 1; r < a.ny; ++r) j *= 0.5;
  } else {
#pragma unroll 8
    for (int ip = 1; ip < np - 1; ++ip) j += chunk[ip * kImpCols2D];  // eight LDS reads ahead of their adds
  }
  j *= a.delta_p;
  j += a.small * (f[nx] * a.p3og[1]);
  const long long nxy = (long long)nx * a.ny;
  const long long k = (long long)y * nx + x;
  const double c11 = (-8.0) * M_PI / 3.0;
  if (q == 0) a.j[k] = j * ((-4.0) * M_PI / 3.0);
  else if (q == 1) a.j[nxy + k] = j * c11;
  else a.j[2 * nxy + k] = -(j * c11);
}

struct ImpSolveArgs {
  double* fld;        // state fields [6][nxy]; E written
  double* pf;         // probe fields [3][6][nxy] (DE on the diagonal)
  const double* j;    // [4][3][nxy]: J0, J_Ex, J_Ey, J_Ez
  const double* jn;   // [3][nxy] electrostatic JN, or null: Ampere
  int* zeros;
  int nx, ny;
  double midx, midy;
  int ext;  // whole arrays with guards: flattened-array differences (ext_offsets)
};

// the 3x3 conductivity solve of Implicit_E_Field::advance (def-ImplicitE.cpp:
// 536-577), JN from Implicit_E_Field::Ampere (:652-678) unless given
__global__ __launch_bounds__(256) void imp_solve_2d(ImpSolveArgs a) {
  const int nx = a.nx, ny = a.ny;
  const long long nxy = (long long)nx * ny;
  const long long k = blockIdx.x * 256LL + threadIdx.x;
  if (k >= nxy) return;
  double de[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) de[d] = a.pf[(d * 7) * nxy + k];
  double jn[3];
  if (a.jn) {
#pragma unroll
    for (int i = 0; i < 3; ++i) jn[i] = a.jn[i * nxy + k];
  } else {
    const int x = static_cast<int>(k % nx), y = static_cast<int>(k / nx);
    long long xm = (x == 0) ? k + nx - 1 : k - 1, xp = (x == nx - 1) ? k - (nx - 1) : k + 1;
    long long ym = (y == 0) ? k + (long long)(ny - 1) * nx : k - nx;
    long long yp = (y == ny - 1) ? k - (long long)(ny - 1) * nx : k + nx;
    bool xlast = false, ylast = false;
    if (a.ext) {
      long long oxm, oxp, oym, oyp;
      ext_offsets(x, y, nx, ny, nx, &oxm, &oxp, &oym, &oyp, &xlast, &ylast);
      xm = k + oxm; xp = k + oxp; ym = k + oym; yp = k + oyp;
    }
    const double* Bx = a.fld + 3 * nxy;
    const double* By = a.fld + 4 * nxy;
    const double* Bz = a.fld + 5 * nxy;
    auto X = [&](const double* F, long long i, bool zero) { return zero ? 0.0 : F[i]; };
    jn[0] = Bz[ym] * a.midy - X(Bz, yp, ylast) * a.midy;
    jn[1] = Bz[xm] * (-a.midx) - X(Bz, xp, xlast) * (-a.midx);
    jn[2] = Bx[ym] * (-a.midy) - X(Bx, yp, ylast) * (-a.midy);
    jn[2] += By[xm] * a.midx - X(By, xp, xlast) * a.midx;
  }
  double A[3][3], D[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const double j0 = a.j[i * nxy + k];
#pragma unroll
    for (int jj = 0; jj < 3; ++jj) A[i][jj] = (a.j[((1 + jj) * 3 + i) * nxy + k] - j0) / de[jj];
    D[i] = jn[i] - j0;
  }
  const double det = A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                     A[1][0] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                     A[2][0] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]);
  double e0 = 0.0, e1 = 0.0, e2 = 0.0;
  if (fabs(det) > 6.0 * 2.2250738585072014e-308) {  // 6 DBL_MIN
    e0 = (D[0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
          D[1] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
          D[2] * (A[0][1] * A[1][2] - A[1][1] * A[0][2])) / det;
    e1 = (A[0][0] * (D[1] * A[2][2] - D[2] * A[1][2]) -
          A[1][0] * (D[0] * A[2][2] - D[2] * A[0][2]) +
          A[2][0] * (D[0] * A[1][2] - D[1] * A[0][2])) / det;
    e2 = (A[0][0] * (A[1][1] * D[2] - A[2][1] * D[1]) -
          A[1][0] * (A[0][1] * D[2] - A[2][1] * D[0]) +
          A[2][0] * (A[0][1] * D[1] - A[1][1] * D[0])) / det;
  } else {
    atomicAdd(a.zeros, 1);
    a.pf[k] = de[0] * 2;
    a.pf[7 * nxy + k] = de[1] * 1.3;
  }
  a.fld[k] = e0;
  a.fld[nxy + k] = e1;
  a.fld[2 * nxy + k] = e2;
}

}  // namespace oshun

namespace {

// planes with l <= lmax (a prefix: planes follow ind(l,m), which lists
// l <= 3 first whenever m0 >= 2)
int planes_upto(const oshun_ctx* c, int lmax) {
  int n = 0;
  while (n < c->nc && c->T.comp_l[n] <= lmax) ++n;
  return n;
}

int imp2d_setup(oshun_ctx* c) {
  if (c->d_imp) return OSHUN_OK;
  const bool es_of[4] = {false, true, false, false};
  const int kind_of[4] = {oshun::kProgAdv1, oshun::kProgAdv1, oshun::kProgAdv2, oshun::kProgProbe};
  for (int q = 0; q < 4; ++q) {
    const oshun::Program P = oshun::derive_program(c->T, kind_of[q], es_of[q]);
    int rc;
    if ((rc = upload_vec(c, &c->iprog[q].terms, P.terms))) return rc;
    if ((rc = upload_vec(c, &c->iprog[q].tstart, P.tstart))) return rc;
    if ((rc = upload_vec(c, &c->iprog[q].clean, P.clean))) return rc;
    if ((rc = lm_prepare(c, P.terms, P.tstart, P.clean, &c->iprog[q]))) return rc;
  }
  const size_t n = (size_t)(18 + 12 + 3) * c->nxy;
  HIPCHK(c, hipMalloc((void**)&c->d_imp, n * sizeof(double)));
  HIPCHK(c, hipMemset(c->d_imp, 0, n * sizeof(double)));
  HIPCHK(c, hipMalloc((void**)&c->d_imp_zero, sizeof(int)));
  HIPCHK(c, hipHostMalloc((void**)&c->h_imp_zero, sizeof(int), hipHostMallocDefault));
  return OSHUN_OK;
}
double* imp_pf(oshun_ctx* c) { return c->d_imp; }
double* imp_j(oshun_ctx* c, int slot) { return c->d_imp + (size_t)(18 + 3 * slot) * c->nxy; }
double* imp_jn(oshun_ctx* c) { return c->d_imp + (size_t)30 * c->nxy; }

oshun::ImpFarArgs imp_far_args(oshun_ctx* c, double h) {
  oshun::ImpFarArgs a;
  a.fld = c->Sf; a.y0 = c->B1f;
  a.nx = c->nx; a.ny = c->ny;
  a.midx = c->T.midx; a.midy = c->T.midy;
  a.h = h; a.h2 = 0.5 * h;
  a.ext = c->ext ? 1 : 0;
  return a;
}

// Current_xyz::calculate_J of the planes in buf into current slot `slot`
int imp_current(oshun_ctx* c, const void* buf, int slot) {
  oshun::ImpCurArgs a;
  a.f = static_cast<const double*>(buf);
  a.j = imp_j(c, slot);
  a.p3og = c->d_p3og;
  a.plane = c->plane;
  a.c10 = c->T.comp_re[oshun::hind(1, 0, c->T.m0)];
  a.c11r = c->T.comp_re[oshun::hind(1, 1, c->T.m0)];
  a.c11i = c->T.comp_im[oshun::hind(1, 1, c->T.m0)];
  a.np = c->np; a.nx = c->nx; a.ny = c->ny;
  a.delta_p = c->T.delta_p; a.small = c->T.small;
  a.ext = c->ext ? 1 : 0;
  ProfRec r;
  prof_begin(c, 1, &r);
  const size_t lds = 3 * static_cast<size_t>(c->np) * oshun::kImpCols2D * sizeof(double);  // <= 96 KB (Np <= 512)
  static bool lds_set = false;  // allow the largest staging area once
  if (!lds_set) {
    HIPCHK(c, hipFuncSetAttribute(reinterpret_cast<const void*>(oshun::imp_current_2d),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  3 * oshun::kImpRows * oshun::kImpCols2D * static_cast<int>(sizeof(double))));
    lds_set = true;
  }
  hipLaunchKernelGGL(oshun::imp_current_2d, dim3((c->nx + oshun::kImpCols2D - 1) / oshun::kImpCols2D, c->ny), dim3(256),
                     lds, c->stream, a);
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

// Euler_Backward::advance_1 on the planes in buf (f00 read from its plane 0)
int imp_collide_f1(oshun_ctx* c, void* buf, double dt_flm) {
  void* s = c->S;
  c->S = buf;
  const int rc = do_collide(c, 0, 1, 0, 0, 0.0, dt_flm);
  c->S = s;
  return rc;
}

// one probe: RK2 of the probe couplings with field component d at DE (S ->
// A), advance_1 on A, current into slot 1 + d
int imp_probe(oshun_ctx* c, int d, double h, double dt_flm, bool coll, int k2, int k3) {
  const double* pf = imp_pf(c) + (size_t)6 * d * c->nxy;
  const DevProg* pr = &c->iprog[3];
  int rc;
  if ((rc = launch_stage<double>(c, 5, c->S, c->S, c->B1, c->A, pf, h, pr, k3))) return rc;
  if ((rc = launch_stage<double>(c, 6, c->B1, nullptr, c->A, c->A, pf, h, pr, k2))) return rc;
  if (coll && (rc = imp_collide_f1(c, c->A, dt_flm))) return rc;
  return imp_current(c, c->A, 1 + d);
}

int implicit_step_2d(oshun_ctx* c, double h, int es) {
  int rc;
  const dim3 fb(256), fg(static_cast<unsigned>((c->nxy + 255) / 256));
  // RK2 of Adv1
  const DevProg* p1 = &c->iprog[es ? 1 : 0];
  if (!es) {
    hipLaunchKernelGGL(oshun::imp_faraday_2d, fg, fb, 0, c->stream, imp_far_args(c, h), 0);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = launch_stage<double>(c, 5, c->S, c->S, c->B1, c->A, c->Sf, h, p1, c->nc))) return rc;
  if ((rc = launch_stage<double>(c, 6, c->B1, nullptr, c->A, c->A, es ? c->Sf : c->B1f, h, p1, c->nc))) return rc;
  if (!es) {
    hipLaunchKernelGGL(oshun::imp_faraday_2d, fg, fb, 0, c->stream, imp_far_args(c, h), 1);
    HIPCHK(c, hipGetLastError());
  }
  std::swap(c->S, c->A);
  c->t_rk += h;
  const double t = c->t_rk;
  // CX.Neighbor_ImplicitE_Communications(Y) (main.cpp:84-86): the B guards
  if (c->ext && (rc = fill_guards(c, true))) return rc;
  const int coll = c->ph.collisions;
  const int do_ee = (coll & OSHUN_COLL_EE) ? 1 : 0;
  const int do_flm = (coll & OSHUN_COLL_FLM) ? 1 : 0;
  if (do_ee) {
    int numh;
    double hee;
    ee_substeps(c, t, &numh, &hee);
    if ((rc = do_collide(c, 1, 0, 0, numh, hee, 0.0))) return rc;
  }
  c->t_ee = t;
  if (sources_in_step(c) && (rc = step_sources(c, t))) return rc;  // main.cpp:90-92
  // Implicit_E_Field::advance
  const double dt_flm = t - c->t_eb;
  const int k1 = planes_upto(c, 1), k2 = planes_upto(c, 2), k3 = planes_upto(c, 3);
  hipLaunchKernelGGL(oshun::imp_findde_2d, fg, fb, 0, c->stream, c->Sf, imp_pf(c), (long long)c->nxy);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(c->A, c->S, (size_t)k1 * c->plane * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  if (do_flm && (rc = imp_collide_f1(c, c->A, dt_flm))) return rc;
  if ((rc = imp_current(c, c->A, 0))) return rc;
  oshun::ImpSolveArgs sa;
  sa.fld = c->Sf; sa.pf = imp_pf(c); sa.j = imp_j(c, 0); sa.jn = es ? imp_jn(c) : nullptr;
  sa.zeros = c->d_imp_zero; sa.nx = c->nx; sa.ny = c->ny; sa.midx = c->T.midx; sa.midy = c->T.midy;
  sa.ext = c->ext ? 1 : 0;
  int zeros = 0;
  for (int attempt = 0; attempt < 4; ++attempt) {
    for (int d = 0; d < (attempt ? 2 : 3); ++d)
      if ((rc = imp_probe(c, d, h, dt_flm, do_flm != 0, k2, k3))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_imp_zero, 0, sizeof(int), c->stream));
    ProfRec r;
    prof_begin(c, 1, &r);
    hipLaunchKernelGGL(oshun::imp_solve_2d, fg, fb, 0, c->stream, sa);
    prof_end(c, &r);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(c->h_imp_zero, c->d_imp_zero, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    zeros = *c->h_imp_zero;
    if (zeros == 0) break;
  }
  c->imp_zero_last = zeros;
  if (zeros > 8)
    return fail(c, OSHUN_ERR_STATE, "implicit field: zero determinant in " + std::to_string(zeros) +
                                        " cells after 4 attempts (the reference stops the run)");
  // RK2 of Adv2 with the solved field
  const DevProg* p2 = &c->iprog[2];
  if ((rc = launch_stage<double>(c, 5, c->S, c->S, c->B1, c->A, c->Sf, h, p2, c->nc))) return rc;
  if ((rc = launch_stage<double>(c, 6, c->B1, nullptr, c->A, c->A, c->Sf, h, p2, c->nc))) return rc;
  std::swap(c->S, c->A);
  if (do_flm && (rc = do_collide(c, 0, 1, 1, 0, 0.0, dt_flm))) return rc;
  c->t_eb = t;
  return c->ext ? fill_guards(c, false) : OSHUN_OK;  // CX.Neighbor_Communications(Y), main.cpp:104
}

}  // namespace

extern "C" {

int oshun_step_implicit_2d(oshun_ctx* c, double h, int nsteps, int if_implicitES) {
  if (!sane(c) || nsteps < 0 || !(h > 0.0)) return OSHUN_ERR_ARG;
  if (c->fp32 || c->T.l0 < 3 || c->T.m0 < 2)
    return fail(c, OSHUN_ERR_ARG, "oshun_step_implicit_2d: needs l0 >= 3, m0 >= 2, FP64");
  if (c->np > oshun::kImpRows)
    return fail(c, OSHUN_ERR_ARG, "oshun_step_implicit_2d: Np above the current kernel's 512 rows");
  int rc;
  if ((rc = imp2d_setup(c))) return rc;
  c->next1_ready = false;
  for (int s = 0; s < nsteps; ++s)
    if ((rc = implicit_step_2d(c, h, if_implicitES))) return rc;
  return OSHUN_OK;
}

int oshun_set_implicit_es_current(oshun_ctx* c, const double* jn, int nbc) {
  if (!sane(c) || !jn || nbc < 0) return OSHUN_ERR_ARG;
  int rc;
  if ((rc = imp2d_setup(c))) return rc;
  // a mirror-boundary context holds JN on its whole arrays: nbc must be its guard count
  if (c->ext && nbc != c->gn)
    return fail(c, OSHUN_ERR_ARG, "oshun_set_implicit_es_current: nbc must be the context's guard count");
  const int e = c->ext ? 0 : nbc;  // guards of the caller's arrays beyond the device's
  const int nxg = c->nx + 2 * e, nyg = c->ny + 2 * e;
  std::vector<double> h((size_t)3 * c->nxy);
  for (int k = 0; k < 3; ++k)
    for (int y = 0; y < c->ny; ++y)
      for (int x = 0; x < c->nx; ++x)
        h[((size_t)k * c->ny + y) * c->nx + x] = jn[2 * (((size_t)k * nyg + y + e) * nxg + x + e)];
  HIPCHK(c, hipMemcpyAsync(imp_jn(c), h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_implicit_zero_dets(const oshun_ctx* c) { return c ? c->imp_zero_last : OSHUN_ERR_ARG; }

}  // extern "C"
