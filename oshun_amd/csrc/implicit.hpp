// implicit.hpp -- the implicit-E step of OSHUN-1 in its 1D variant
// (implicit_E and if_implicit1D, main.cpp:79-107 with Runge_Kutta_2_Imp and
// Implicit_E_Field) for 1D runs with m0 = 0 in FP64.  Included at the end of
// oshun_hip.hip.
//
// One step:
//   RK2 of Implicit_Actions::Adv1 (x-advection; with Faraday on the fields
//     unless if_implicitES)                       def-RK.cpp:470-498, def-Actions.cpp:118-132
//   e-e sub-cycles                                 Explicit_EC_FP::loop
//   Implicit_E_Field::advance, 1D variant          def-ImplicitE.cpp:472-606
//     probe field DE = sqrt(|E|^2) sqrt(16 eps) + 16 eps (FindDE, :628-652);
//     J0 = current after Euler_Backward::advance_1 with E = 0;
//     J_DE = current after an RK2 of Implicit_Ex1D with Ex = DE and advance_1;
//     (the harmonics the probes touch are restored after each)
//     Ex = DE J0 / (J0 - J_DE), Ey = Ez = 0
//   RK2 of Implicit_Actions::Adv2 (Exyz1D + CLEAN_LOW_PC)   def-RK.cpp:503-524
//   Euler_Backward::advance_1, advance_lm
// The current of the field solve is Current_xyz::calculate_J
// (def-ImplicitE.cpp:125-156), whose trapezoid end weight lands on a guard row
// (a 1D-style subscript), so interior cells sum p0 and pmax with weight 1.
//
// Buffers: S holds the state.  The RK2 of Adv1 and of Adv2 are one fused
// stage-pair launch each (vlasov_pair_1d modes 2 and 3: S -> B1, then S and
// B1 swap); the field solve's probes run on a copy in B2 (mode 4, the RK2 of
// Implicit_Ex1D on planes 0..3, S -> B2); B2f holds the per-x probe field
// and currents.

namespace oshun {

struct ImpFieldArgs {
  double* fld;                  // [6][nx] state fields
  double* de;                   // [nx] probe field
  double* j0;                   // [nx] current without E
  double* jde;                  // [nx] current with the probe field
  const double* f10;            // l = 1 plane [p][x]
  const double* p3og;
  int np, nx;
  double delta_p, small, midx, midy, h, h2;
};

// both RK2 stages of Adv1's Faraday term on the fields (MaxwellEq::Faraday,
// def-VlasovMax.cpp:979-1005): E does not change during Adv1, so the two
// slopes are equal; B = (B + (h k) / 2) + (h / 2) k
__global__ __launch_bounds__(256) void imp_faraday_1d(ImpFieldArgs a) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = a.nx;
  if (x >= nx) return;
  const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
  const double* Ex = a.fld;
  const double* Ey = a.fld + nx;
  const double* Ez = a.fld + 2 * nx;
  double k3 = 0.0, k4 = 0.0, k5 = 0.0;
  k3 += Ez[x] * (-a.midy) - Ez[x] * (-a.midy);   // y differences of a y-uniform field
  k4 += Ez[xm] * a.midx - Ez[xp] * a.midx;
  k5 += Ex[x] * a.midy - Ex[x] * a.midy;
  k5 += Ey[xm] * (-a.midx) - Ey[xp] * (-a.midx);
  const double k[3] = {k3, k4, k5};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double* B = a.fld + (3 + c) * nx + x;
    const double kh = k[c] * a.h;
    const double y1 = *B + kh * 0.5;
    *B = y1 + k[c] * a.h2;
  }
}

// Current_xyz::calculate_J, x component (def-ImplicitE.cpp:125-156), with p0
// and pmax at weight 1 on interior cells (see above), in the reference's
// sequential order: a block of 16 columns x 16 rows stages f10 p3og of all
// p rows in LDS (every load of the block in flight at once) and the first row
// of threads adds them up in p order.
// stage 0: probe field from the state E, J0, and Ex = DE; stage 1: J_DE and
// the solved field.
constexpr int kImpCols = 16, kImpRows = 512;  // rows: the largest Np oshun_create accepts
__global__ __launch_bounds__(256) void imp_field_solve_1d(ImpFieldArgs a, int stage) {
  __shared__ double chunk[kImpRows][kImpCols];
  const int tx = threadIdx.x % kImpCols, ty = threadIdx.x / kImpCols;
  const int x = blockIdx.x * kImpCols + tx;
  const int np = a.np, nx = a.nx;
  const bool ok = x < nx;
  const double* f = a.f10 + (ok ? x : 0);
  double j = 0.0;
  if (ok && ty == 0) {
    j = f[0] * a.p3og[0];
    j += f[(long long)(np - 1) * nx] * a.p3og[np - 1];
  }
#pragma unroll 8
  for (int ip = 1 + ty; ip < np - 1; ip += 256 / kImpCols) chunk[ip][tx] = ok ? f[(long long)ip * nx] * a.p3og[ip] : 0.0;
  __syncthreads();
  if (ok && ty == 0) {
#pragma unroll 8
    for (int ip = 1; ip < np - 1; ++ip) j += chunk[ip][tx];  // eight LDS reads ahead of their adds
  }
  if (!ok || ty != 0) return;
  j *= a.delta_p;
  j += a.small * (f[nx] * a.p3og[1]);
  j *= (-4.0) * M_PI / 3.0;
  if (stage == 0) {
    const double Eps = 16.0 * 2.220446049250313e-16;  // 16 numeric_limits<double>::epsilon()
    const double LargeEps = sqrt(Eps);
    double ex = a.fld[x], ey = a.fld[nx + x], ez = a.fld[2 * nx + x];
    ex *= ex; ey *= ey; ez *= ez;
    ex += ey; ex += ez;
    ex = sqrt(ex);
    ex *= LargeEps;
    ex += Eps;
    a.de[x] = ex;
    a.j0[x] = j;
    a.fld[x] = ex;
  } else {
    a.jde[x] = j;
    a.fld[x] = a.de[x] * a.j0[x] / (a.j0[x] - j);
    a.fld[nx + x] = 0.0;
    a.fld[2 * nx + x] = 0.0;
  }
}

}  // namespace oshun

namespace {


// one RK2 of Adv1 (pair = 2: x-advection) or Adv2 (pair = 3: Exyz1D +
// CLEAN_LOW_PC) on every plane as one fused stage-pair launch: S -> B1, then
// S and B1 swap roles (neighbouring tiles still read S while one writes)
template <int L0>
void launch_imp_pair(oshun_ctx* c, int pair, const oshun::Stage1DArgs<double>& a, dim3 grid) {
  if (pair == 2)
    hipLaunchKernelGGL((oshun::vlasov_pair_1d<double, L0, pair_rows<double>(), 2>), grid, dim3(64, 4), 0, c->stream,
                       a);
  else
    hipLaunchKernelGGL((oshun::vlasov_pair_1d<double, L0, pair_rows<double>(), 3>), grid, dim3(64, 4), 0, c->stream,
                       a);
}
int imp_rk2_pair(oshun_ctx* c, double h, int pair) {
  const oshun::Stage1DArgs<double> a = pair_args<double>(c, c->S, c->S, c->B1, c->A, c->Sf, c->Sf, c->jp1, c->jp2, h);
  const int ntx = (c->nx + oshun::kPairX - 1) / oshun::kPairX;
  const dim3 grid((ntx * pair_tiles_r(c->np, pair_rows<double>()) + 7) / 8 * 8);  // the RK2 kernels keep R = 3
  switch (c->T.l0) {
    case 3: launch_imp_pair<3>(c, pair, a, grid); break;
    case 4: launch_imp_pair<4>(c, pair, a, grid); break;
    case 5: launch_imp_pair<5>(c, pair, a, grid); break;
    case 6: launch_imp_pair<6>(c, pair, a, grid); break;
    case 7: launch_imp_pair<7>(c, pair, a, grid); break;
    case 8: launch_imp_pair<8>(c, pair, a, grid); break;
    case 9: launch_imp_pair<9>(c, pair, a, grid); break;
    case 10: launch_imp_pair<10>(c, pair, a, grid); break;
    case 11: launch_imp_pair<11>(c, pair, a, grid); break;
    default: launch_imp_pair<12>(c, pair, a, grid); break;
  }
  HIPCHK(c, hipGetLastError());
  std::swap(c->S, c->B1);
  return OSHUN_OK;
}

oshun::ImpFieldArgs imp_field_args(oshun_ctx* c, double h) {
  oshun::ImpFieldArgs a;
  a.fld = c->Sf;
  a.de = c->B2f;
  a.j0 = c->B2f + c->nx;
  a.jde = c->B2f + 2 * c->nx;
  a.f10 = static_cast<const double*>(c->S) + c->plane;  // plane 1 = f10 (m0 = 0)
  a.p3og = c->d_p3og;
  a.np = c->np; a.nx = c->nx;
  a.delta_p = c->T.delta_p; a.small = c->T.small; a.midx = c->T.midx; a.midy = c->T.midy;
  a.h = h; a.h2 = 0.5 * h;
  return a;
}

// copies of the planes l <= lmax of S into / out of B2
int imp_planes(oshun_ctx* c, int lmax, bool save) {
  const size_t bytes = (size_t)(lmax + 1) * c->plane * sizeof(double);
  if (save) HIPCHK(c, hipMemcpyAsync(c->B2, c->S, bytes, hipMemcpyDeviceToDevice, c->stream));
  else HIPCHK(c, hipMemcpyAsync(c->S, c->B2, bytes, hipMemcpyDeviceToDevice, c->stream));
  return OSHUN_OK;
}

int implicit_step_1d(oshun_ctx* c, double h, int es) {
  int rc;
  const dim3 fb(256), fg((c->nx + 255) / 256);
  ProfRec r;
  // RK2 of Adv1: x-advection, and Faraday on the fields unless electrostatic
  prof_begin(c, 0, &r);
  rc = imp_rk2_pair(c, h, 2);
  if (!rc && !es) {
    hipLaunchKernelGGL(oshun::imp_faraday_1d, fg, fb, 0, c->stream, imp_field_args(c, h));
    HIPCHK(c, hipGetLastError());
  }
  prof_end(c, &r);
  if (rc) return rc;
  c->t_rk += h;
  const double t = c->t_rk;
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
  // Implicit_E_Field::advance, 1D variant.  The reference runs its two probes
  // on the state and restores f00, f1m, f2m afterwards (def-ImplicitE.cpp:
  // 491-520); here they run on a copy in B2 (S and B2 swap around them), so S
  // stays the state: probe 1 needs f00 and f10 (advance_1 reads f00, writes
  // f10), probe 2 writes its RK2 straight into B2.
  const double dt_flm = t - c->t_eb;
  const dim3 sg((c->nx + oshun::kImpCols - 1) / oshun::kImpCols);
  if ((rc = imp_planes(c, 1, true))) return rc;
  std::swap(c->S, c->B2);
  rc = do_flm ? do_collide(c, 0, 1, 0, 0, 0.0, dt_flm) : OSHUN_OK;
  if (!rc) {
    prof_begin(c, 1, &r);
    hipLaunchKernelGGL(oshun::imp_field_solve_1d, sg, fb, 0, c->stream, imp_field_args(c, h), 0);
    prof_end(c, &r);
  }
  std::swap(c->S, c->B2);
  if (rc) return rc;
  HIPCHK(c, hipGetLastError());
  prof_begin(c, 0, &r);
  // Implicit_Ex1D with E = DE: sources l <= 3, targets l <= 2; the planes
  // above l = 3 only receive +0, so both RK2 stages run on planes 0..3 alone,
  // as one stage-pair launch S -> B2
  {
    const oshun::Stage1DArgs<double> a = pair_args<double>(c, c->S, c->S, c->B2, c->A, c->B2f, c->B2f, c->jp1,
                                                           c->jp2, h);
    const int ntx = (c->nx + oshun::kPairX - 1) / oshun::kPairX;
    const dim3 grid((ntx * pair_tiles_r(c->np, pair_rows<double>()) + 7) / 8 * 8);
    hipLaunchKernelGGL((oshun::vlasov_pair_1d<double, 3, pair_rows<double>(), 4>), grid, dim3(64, 4), 0, c->stream, a);
  }
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  std::swap(c->S, c->B2);
  rc = do_flm ? do_collide(c, 0, 1, 0, 0, 0.0, dt_flm) : OSHUN_OK;
  if (!rc) {
    prof_begin(c, 1, &r);
    hipLaunchKernelGGL(oshun::imp_field_solve_1d, sg, fb, 0, c->stream, imp_field_args(c, h), 1);
    prof_end(c, &r);
  }
  std::swap(c->S, c->B2);
  if (rc) return rc;
  HIPCHK(c, hipGetLastError());
  // RK2 of Adv2 (Exyz1D + CLEAN_LOW_PC) with the solved field
  prof_begin(c, 0, &r);
  rc = imp_rk2_pair(c, h, 3);
  prof_end(c, &r);
  if (rc) return rc;
  if (do_flm && (rc = do_collide(c, 0, 1, 1, 0, 0.0, dt_flm))) return rc;
  c->t_eb = t;
  return OSHUN_OK;
}

}  // namespace

extern "C" {

int oshun_step_implicit_1d(oshun_ctx* c, double h, int nsteps, int if_implicitES) {
  if (!sane(c) || nsteps < 0 || !(h > 0.0)) return OSHUN_ERR_ARG;
  if (c->ext) return fail(c, OSHUN_ERR_ARG, "oshun_step_implicit_1d: periodic boundaries only");
  if (!c->use1d || c->fp32 || c->T.l0 < 3 || c->T.l0 > 12)
    return fail(c, OSHUN_ERR_ARG, "oshun_step_implicit_1d: needs a 1D run with m0 = 0, 3 <= l0 <= 12, FP64");
  if (!c->use_pair || !c->rows_npl)
    return fail(c, OSHUN_ERR_ARG, "oshun_step_implicit_1d: the state is too large for its 32-bit-offset kernels");
  c->next1_ready = false;
  for (int s = 0; s < nsteps; ++s) {
    const int rc = implicit_step_1d(c, h, if_implicitES);
    if (rc) return rc;
  }
  return OSHUN_OK;
}

}  // extern "C"
