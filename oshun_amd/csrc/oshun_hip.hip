// C ABI of the MI355X distribution-function update (declared in
// include/oshun_hip.h).  One oshun_ctx = one device-resident OSHUN state.
//
// Step structure per call of oshun_step (main.cpp:79-107):
//   4 x { vlasov_stage<T>, field_stage<T> }   Runge_Kutta_4::advance_p1
//   collide<T,NPL>                            Explicit_EC_FP::loop +
//                                             Euler_Backward::advance_1/_lm
// Guard cells do not exist on the device: stencils index periodically, which
// reproduces the reference's interior (its guards are refreshed only after a
// whole step, and every RK stage widens the stencil by one cell of four).

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "instances.hpp"
#include "oshun_hip.h"

// the heavy kernel templates are compiled in k_*.hip
OSHUN_EXTERN_KERNELS
#define OSHUN_SPEC_EXT extern
OSHUN_STAGE_SPECS(OSHUN_SPEC_F64)
OSHUN_STAGE_SPECS(OSHUN_SPEC_F32)
#undef OSHUN_SPEC_EXT

using oshun::HostTables;
using oshun::Term;

struct ProfRec {
  int cls;
  hipEvent_t e0, e1;
};

// a coupling program on the device (terms, per-plane starts, CLEAN_LOW_PC rows)
struct DevProg {
  Term* terms = nullptr;
  int* tstart = nullptr;
  int* clean = nullptr;
  int lm = -1;             // LmKind of the l-streaming kernel that runs it, -1: the interpreter
  int* gstart = nullptr;   // that kernel's coefficients and interior layout (LmPlan)
  double* coef = nullptr;
  int lm_istart = 0, lm_istride = 0, lm_ioff[5] = {0, 0, 0, 0, 0};
};

struct oshun_ctx {
  oshun_grid g;
  oshun_phys ph;
  HostTables T;
  int nc = 0, np = 0, nx = 0, ny = 0, nxy = 0, npb = 0;
  long long plane = 0;
  // mirror boundaries (ext): nx, ny, plane above count the whole arrays,
  // gn guard cells on each side; nxi, nyi the interior (= nx, ny otherwise)
  bool ext = false;
  int gn = 0, nxi = 0, nyi = 0;
  bool fp32 = false;
  size_t esz = 8;
  void *S = nullptr, *A = nullptr, *B1 = nullptr, *B2 = nullptr;
  double *Sf = nullptr, *Af = nullptr, *B1f = nullptr, *B2f = nullptr;
  double* Cf = nullptr;  // 1D pair path: fields of the fourth stage's input
  // 1D pair path with fused collisions: the collision launch does the last
  // field stage of its step (fuse_fields >= 1) and the first field stage of
  // the next step (fuse_fields = 2, its accumulator into Af2, then Af and Af2
  // swap); next1_ready says that the next rk4 finds its stage-1 fields there
  double* Af2 = nullptr;
  int fuse_fields = 0;
  double fuse_h = 0.0;
  bool next1_ready = false;
  double* jpart = nullptr;
  Term* d_terms = nullptr;
  int* d_tstart = nullptr;
  int* d_clean = nullptr;
  int* d_comp_l = nullptr;
  int* d_comp_m = nullptr;     // m of each plane (guard refresh)
  int* d_comp_isim = nullptr;  // 1 on imaginary planes (guard refresh)
  void* d_vt = nullptr;  // Vlasov tables in storage precision: invpax | vtab | cg | hp0
  void* d_rowc = nullptr;  // the marched kernels' coefficients by row (storage precision)
  void* d_eAT = nullptr;   // the 1D E coefficients eA1 | eA2 in the storage precision
  double* d_p3og = nullptr;
  double* d_e1d = nullptr;  // 1D m0 = 0 coefficient rows: eA1 | eA2 | xA1 | xA2
  double* d_mw = nullptr;   // moment weights w2 | w3 | w4 | w5
  double* d_mom = nullptr;    // moments output, kMomOut planes of nxy (first oshun_moments call)
  double* d_mpart = nullptr;  // partial p sums of the moments, moment_chunks x kMomSums planes
  bool use1d = false;       // specialised 1D m0 = 0 Vlasov kernel
  // general system: index into OSHUN_STAGE_SPECS of the compile-time stage
  // kernel for this (l0, m0, 2D), -1 for the term-list interpreter
  int spec = -1;
  bool force_general = false;
  DevProg full;  // lm and coef of the full program (the interpreter reads d_terms, d_tstart, d_clean)
  double* d_coef = nullptr;  // term coefficients in the spec kernel's visit order
  double* d_coll = nullptr;
  double* d_rt = nullptr;   // collide_rows tables of the e-e launch (interleaved for rows_npl/rows_segw)
  double* d_rt2 = nullptr;  // of the implicit l >= 1 launch (rows_npl2/rows_segw2)
  // FP32 split: plane 0 stores f00 - bg(p), bg an FP64 background per p
  // (the x,y mean of f00 at upload), so that FP32 rounding acts on the
  // deviation only; d_bgG holds G00(bg) for the Vlasov kernels.
  double* d_bg = nullptr;
  float* d_bgG = nullptr;
  // 1D m0 = 0: RK4 as two fused stage pairs (vlasov_pair_1d)
  bool use_pair = false;
  int pair_r = 3;              // rows per thread of the RK4 pair kernels (tile height 4 pair_r - 2);
                               // 0: the marched pair kernels (march_pw rows per wavefront)
  int march_pw = 8;
  int npt = 0;                 // p tiles of the pair kernels
  double* jp1 = nullptr;       // of the second stage input of a pair [npt][3][nx]
  double* jp2 = nullptr;       // of Y2 [npt][3][nx]
  double* f1y = nullptr;       // f10(p1) of the second stage input, per x
  std::vector<double> bg;
  int rows_npl = 0;
  int rows_segw = 16;       // lanes per column of collide_rows (16 or 32)
  int rows_npl2 = 0, rows_segw2 = 16;
  oshun::CollTables ct;
  // sources of def-PLSource.cpp (oshun_set_sources): switches, clocks, constants
  bool src_on = false;
  oshun_sources src = {};
  int src_pls = 0, src_ib = 0, src_pol = 0;
  double src_t_pls = 0.0, src_t_ib = 0.0, src_t_efld = 0.0;
  double src_IL = 0.0, src_IL_sqrt = 0.0, src_IL_min = 0.0, src_acoeff = 0.0, src_vos = 0.0, src_w0overwp = 0.0;
  double src_vw_coeff_cube = 0.0, src_Qn_coeff = 0.0, src_Inv_Uav6_nm1 = 0.0, src_Pnm1 = 0.0, src_p0 = 0.0, src_p1 = 0.0;
  double* d_src = nullptr;   // per-p and per-cell source tables
  int* d_srcflag = nullptr;  // a heating rate out of range
  // implicit-E step, 2D variant (implicit2d.hpp), set up by its first call:
  // the action programs (Adv1 with B, Adv1 electrostatic, Adv2, probes) and
  // d_imp = probe fields [3][6][nxy] | currents J0, J_Ex, J_Ey, J_Ez [4][3][nxy]
  // | electrostatic JN [3][nxy]; d_imp_zero counts zero determinants
  DevProg iprog[4];
  double* d_imp = nullptr;
  int* d_imp_zero = nullptr;
  int* h_imp_zero = nullptr;  // pinned
  int imp_zero_last = 0;
  hipStream_t stream = nullptr;
  int device = 0;  // the HIP device current at oshun_create
  double t_rk = 0.0, t_ee = 0.0, t_eb = 0.0;
  std::string err;
  bool prof = false;
  std::vector<ProfRec> prec;
  double prof_ms[4] = {0, 0, 0, 0};
  long long prof_n[4] = {0, 0, 0, 0};
};

namespace {

int fail(oshun_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(ctx, expr)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(ctx, OSHUN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename V>
int upload_vec(oshun_ctx* c, V** dst, const std::vector<V>& src) {
  size_t n = std::max<size_t>(src.size(), 1) * sizeof(V);
  HIPCHK(c, hipMalloc((void**)dst, n));
  if (!src.empty()) HIPCHK(c, hipMemcpy(*dst, src.data(), src.size() * sizeof(V), hipMemcpyHostToDevice));
  return OSHUN_OK;
}

// The l-streaming kernel's plan of a coupling program (terms, tstart over all
// planes): q->lm = the program kind it runs and q->gstart, q->coef its
// coefficients, or q->lm = -1 (the interpreter runs the program; FP32
// storage has the full program's kernel only)
int lm_prepare(oshun_ctx* c, const std::vector<Term>& terms, const std::vector<int>& tstart,
               const std::vector<int>& clean, DevProg* q) {
  q->lm = -1;
  if (c->use1d) return OSHUN_OK;
  oshun::LmPlan P;
  oshun::lm_build(c->T, terms, tstart, clean, &P);
  if (P.kind < 0 || (c->fp32 && P.kind != oshun::kLmFull)) return OSHUN_OK;
  if (int rc = upload_vec(c, &q->gstart, P.gstart)) return rc;
  if (int rc = upload_vec(c, &q->coef, P.coef)) return rc;
  q->lm_istart = P.istart;
  q->lm_istride = P.istride;
  for (int m = 0; m <= oshun::kLmMaxM0; ++m) q->lm_ioff[m] = P.ioff[m];
  q->lm = P.kind;
  return OSHUN_OK;
}

// the l-streaming kernel (one thread per plane element) and then the
// stage input's current partials on the stage grid (grid, block)
template <typename T, int M0, int KIND>
void launch_lm_td(oshun_ctx* c, const oshun::StageArgs<T>& a, dim3 grid, dim3 block) {
  const dim3 g1((unsigned)((c->plane + oshun::kLmBlock - 1) / oshun::kLmBlock));
  if (c->T.two_d)
    hipLaunchKernelGGL((oshun::vlasov_stage_lm<T, M0, KIND, true>), g1, dim3(oshun::kLmBlock), 0, c->stream, a);
  else
    hipLaunchKernelGGL((oshun::vlasov_stage_lm<T, M0, KIND, false>), g1, dim3(oshun::kLmBlock), 0, c->stream, a);
  hipLaunchKernelGGL(oshun::vlasov_current_tiles<T>, grid, block, 0, c->stream, a);
}
template <typename T, int M0>
int launch_lm_kind(oshun_ctx* c, const oshun::StageArgs<T>& a, int kind, dim3 grid, dim3 block) {
  if constexpr (std::is_same<T, float>::value) {
    if (kind != oshun::kLmFull) return OSHUN_ERR_ARG;
    launch_lm_td<T, M0, oshun::kLmFull>(c, a, grid, block);
  } else {
    switch (kind) {
      case oshun::kLmFull: launch_lm_td<T, M0, oshun::kLmFull>(c, a, grid, block); break;
      case oshun::kLmAdv1: launch_lm_td<T, M0, oshun::kLmAdv1>(c, a, grid, block); break;
      case oshun::kLmAdv1Es: launch_lm_td<T, M0, oshun::kLmAdv1Es>(c, a, grid, block); break;
      case oshun::kLmAdv2: launch_lm_td<T, M0, oshun::kLmAdv2>(c, a, grid, block); break;
      case oshun::kLmProbe: launch_lm_td<T, M0, oshun::kLmProbe>(c, a, grid, block); break;
      default: return OSHUN_ERR_ARG;
    }
  }
  return OSHUN_OK;
}

void prof_begin(oshun_ctx* c, int cls, ProfRec* r) {
  if (!c->prof) return;
  r->cls = cls;
  hipEventCreate(&r->e0);
  hipEventCreate(&r->e1);
  hipEventRecord(r->e0, c->stream);
}
void prof_end(oshun_ctx* c, ProfRec* r) {
  if (!c->prof) return;
  hipEventRecord(r->e1, c->stream);
  c->prec.push_back(*r);
}

template <typename T>
const T* vt_ptr(const oshun_ctx* c, int which) {
  // invpax (l0+1)*np | vtab np | cg l0+1 | hp0 l0+1 | xv1 (l0+1)*np | xv2 (l0+1)*np
  const T* b = static_cast<const T*>(c->d_vt);
  const size_t L = c->T.l0 + 1, np = c->np;
  switch (which) {
    case 0: return b;
    case 1: return b + L * np;
    case 2: return b + L * np + np;
    case 3: return b + L * np + np + L;
    case 4: return b + L * np + np + 2 * L;
    default: return b + 2 * L * np + np + 2 * L;
  }
}

constexpr int kMaxL0Special = 12;

template <typename T, int L0>
void launch_1d(oshun_ctx* c, const oshun::Stage1DArgs<T>& a, dim3 grid, dim3 block) {
  hipLaunchKernelGGL((oshun::vlasov_stage_1d<T, L0>), grid, block, 0, c->stream, a);
}

// G00 of the FP32 background table, null in FP64 (no background there)
template <typename T>
const T* bg_g00(const oshun_ctx* c);
template <>
const double* bg_g00<double>(const oshun_ctx*) { return nullptr; }
template <>
const float* bg_g00<float>(const oshun_ctx* c) { return c->d_bgG; }

template <typename T>
void fill_1d_args(oshun_ctx* c, oshun::Stage1DArgs<T>& a, int mode, const void* in, const void* y0, void* out,
                  void* acc, const double* fin, double h) {
  a.in = static_cast<const T*>(in);
  a.y0 = static_cast<const T*>(y0);
  a.out = static_cast<T*>(out);
  a.acc = static_cast<T*>(acc);
  a.fin = fin;
  a.jpart = c->jpart;
  a.invpax = vt_ptr<T>(c, 0);
  a.vtab = vt_ptr<T>(c, 1);
  a.cg = vt_ptr<T>(c, 2);
  a.hp0 = vt_ptr<T>(c, 3);
  a.xv1 = vt_ptr<T>(c, 4);
  a.xv2 = vt_ptr<T>(c, 5);
  a.p3og = c->d_p3og;
  const size_t L = c->T.l0 + 1;
  a.eA1 = c->d_e1d;
  a.eA2 = c->d_e1d + L;
  a.xA1 = c->d_e1d + 2 * L;
  a.xA2 = c->d_e1d + 3 * L;
  a.bgG = bg_g00<T>(c);
  a.p0p1_sq = static_cast<T>(c->T.p0p1_sq);
  a.inv_mp0p1_sq = static_cast<T>(c->T.inv_mp0p1_sq);
  a.g_r = static_cast<T>(c->T.g_r);
  a.c05h = static_cast<T>(0.5 * h);
  a.hh = static_cast<T>(h);
  a.h6 = static_cast<T>(h / 6.0);
  a.third = static_cast<T>(1.0 / 3.0);
  a.twothird = static_cast<T>(2.0 / 3.0);
  a.np = c->np; a.nx = c->nx; a.mode = mode;
  a.plane = c->plane;
  a.fin2 = nullptr; a.jpA = nullptr; a.jpB = nullptr; a.f1A = nullptr;
}

template <typename T>
int launch_stage_1d(oshun_ctx* c, int mode, const void* in, const void* y0, void* out, void* acc,
                    const double* fin, double h) {
  oshun::Stage1DArgs<T> a;
  fill_1d_args<T>(c, a, mode, in, y0, out, acc, fin, h);
  dim3 block(64, oshun::kRows1D, 1);
  dim3 grid((c->nx + 63) / 64, c->npb, 1);
  ProfRec r;
  prof_begin(c, 0, &r);
  switch (c->T.l0) {
    case 1: launch_1d<T, 1>(c, a, grid, block); break;
    case 2: launch_1d<T, 2>(c, a, grid, block); break;
    case 3: launch_1d<T, 3>(c, a, grid, block); break;
    case 4: launch_1d<T, 4>(c, a, grid, block); break;
    case 5: launch_1d<T, 5>(c, a, grid, block); break;
    case 6: launch_1d<T, 6>(c, a, grid, block); break;
    case 7: launch_1d<T, 7>(c, a, grid, block); break;
    case 8: launch_1d<T, 8>(c, a, grid, block); break;
    case 9: launch_1d<T, 9>(c, a, grid, block); break;
    case 10: launch_1d<T, 10>(c, a, grid, block); break;
    case 11: launch_1d<T, 11>(c, a, grid, block); break;
    default: launch_1d<T, 12>(c, a, grid, block); break;
  }
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

// prog: a coupling program other than the full one (term-list interpreter,
// the first nc planes only)
template <typename T>
int launch_stage(oshun_ctx* c, int mode, const void* in, const void* y0, void* out, void* acc, const double* fin,
                 double h, const DevProg* prog = nullptr, int nc = 0) {
  if (c->use1d && !prog) return launch_stage_1d<T>(c, mode, in, y0, out, acc, fin, h);
  oshun::StageArgs<T> a;
  a.in = static_cast<const T*>(in);
  a.y0 = static_cast<const T*>(y0);
  a.out = static_cast<T*>(out);
  a.acc = static_cast<T*>(acc);
  a.fin = fin;
  a.jpart = c->jpart;
  a.terms = c->d_terms;
  a.tstart = c->d_tstart;
  a.clean = c->d_clean;
  a.invpax = vt_ptr<T>(c, 0);
  a.vtab = vt_ptr<T>(c, 1);
  a.cg = vt_ptr<T>(c, 2);
  a.hp0 = vt_ptr<T>(c, 3);
  a.p3og = c->d_p3og;
  a.bgG = bg_g00<T>(c);
  a.p0p1_sq = static_cast<T>(c->T.p0p1_sq);
  a.inv_mp0p1_sq = static_cast<T>(c->T.inv_mp0p1_sq);
  a.g_r = static_cast<T>(c->T.g_r);
  a.c05h = static_cast<T>(0.5 * h);
  a.hh = static_cast<T>(h);
  a.h6 = static_cast<T>(h / 6.0);
  a.third = static_cast<T>(1.0 / 3.0);
  a.twothird = static_cast<T>(2.0 / 3.0);
  a.np = c->np; a.nx = c->nx; a.ny = c->ny; a.nc = c->nc;
  a.plane = c->plane;
  a.mode = mode;
  a.c10 = c->T.comp_re[oshun::hind(1, 0, c->T.m0)];
  a.c11r = c->T.m0 >= 1 ? c->T.comp_re[oshun::hind(1, 1, c->T.m0)] : -1;
  a.c11i = c->T.m0 >= 1 ? c->T.comp_im[oshun::hind(1, 1, c->T.m0)] : -1;
  a.coef = c->d_coef;
  a.ext = c->ext ? 1 : 0;
  a.l0 = c->T.l0;
  if (prog) {
    a.terms = prog->terms;
    a.tstart = prog->tstart;
    a.clean = prog->clean;
    a.nc = nc;
  }
  dim3 block(oshun::kTileX, oshun::kTileP, 1);
  dim3 grid((c->nx + oshun::kTileX - 1) / oshun::kTileX, c->npb, c->ny);
  // the l-streaming kernel: programs that it matches (over all planes, or
  // the probe over its plane prefix), the full program where no specialised
  // kernel is built for the shape
  const DevProg* lmp = prog ? prog : (c->spec < 0 ? &c->full : nullptr);
  int lm = (lmp && !c->force_general) ? lmp->lm : -1;
  // the layers the kernel reads (0..a.l0) and writes (0..ltop): all, or for
  // the probe sources up to l = 3 and the planes of l <= 2 or l <= 3
  int ltop = c->T.l0;
  if (lm == oshun::kLmProbe) {
    a.l0 = std::min(c->T.l0, oshun::kLmProbeTop);
    ltop = -1;
    for (int t = 0; t <= a.l0; ++t)
      if (nc == oshun::lm_layer_start(t + 1, c->T.m0)) ltop = t;
    if (ltop < 0) lm = -1;
  } else if (prog && nc != c->nc) {
    lm = -1;
  }
  a.lm_ltop = ltop;
  // the kernels run the full program in modes 0..4 (1 with y0 the stage
  // input) and the implicit ones in modes 0, 6 and 5 (y0 the stage input);
  // lane byte offsets within a plane are 32-bit
  if (lm == oshun::kLmFull && !(mode == 0 || mode == 2 || mode == 3 || mode == 4 || (mode == 1 && y0 == in))) lm = -1;
  if (lm > 0 && !(mode == 0 || mode == 6 || (mode == 5 && y0 == in))) lm = -1;
  if ((unsigned long long)c->plane * sizeof(T) >= (1ull << 32)) lm = -1;
  ProfRec r;
  prof_begin(c, 0, &r);
  if (lm >= 0) {
    a.coef = lmp->coef;
    a.tstart = lmp->gstart;
    a.lm_istart = lmp->lm_istart;
    a.lm_istride = lmp->lm_istride;
    for (int m = 0; m <= oshun::kLmMaxM0; ++m) a.lm_ioff[m] = lmp->lm_ioff[m];
    int rc = OSHUN_ERR_ARG;
    switch (c->T.m0) {
      case 1: rc = launch_lm_kind<T, 1>(c, a, lm, grid, block); break;
      case 2: rc = launch_lm_kind<T, 2>(c, a, lm, grid, block); break;
      case 3: rc = launch_lm_kind<T, 3>(c, a, lm, grid, block); break;
      case 4: rc = launch_lm_kind<T, 4>(c, a, lm, grid, block); break;
      default: break;
    }
    if (rc) return fail(c, OSHUN_ERR_ARG, "no l-streaming stage kernel for this m0 and program");
  } else if (c->spec >= 0 && !c->force_general && !prog) {
    switch (c->spec) {
#define OSHUN_SPEC_LAUNCH(ID, L, M, TD)                                                        \
      case ID:                                                                                 \
        hipLaunchKernelGGL((oshun::vlasov_stage_spec<T, L, M, TD, oshun::kSpecDepth>), grid, block, 0, c->stream, a); \
        break;
      OSHUN_STAGE_SPECS(OSHUN_SPEC_LAUNCH)
#undef OSHUN_SPEC_LAUNCH
      default:
        return fail(c, OSHUN_ERR_ARG, "no specialised stage kernel for this shape");
    }
  } else {
    hipLaunchKernelGGL(oshun::vlasov_stage<T>, grid, block, 0, c->stream, a);
  }
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

template <typename T>
oshun::FieldArgs<T> field_args(oshun_ctx* c, int mode, const double* fin, const double* f0, double* fout,
                               double* facc, const void* fpl, double h, const double* jp, int npb,
                               const double* f1p1) {
  oshun::FieldArgs<T> a;
  a.fin = fin; a.f0 = f0; a.fout = fout; a.facc = facc;
  a.jpart = jp ? jp : c->jpart;
  a.fpl = static_cast<const T*>(fpl);
  a.f1p1 = f1p1;
  a.p3og = nullptr;
  a.npb = jp ? npb : c->npb; a.np = c->np; a.nx = c->nx; a.ny = c->ny; a.mode = mode;
  a.c10 = c->T.comp_re[oshun::hind(1, 0, c->T.m0)];
  a.c11r = c->T.m0 >= 1 ? c->T.comp_re[oshun::hind(1, 1, c->T.m0)] : -1;
  a.c11i = c->T.m0 >= 1 ? c->T.comp_im[oshun::hind(1, 1, c->T.m0)] : -1;
  a.plane = c->plane;
  a.p3og1 = c->T.p3og[1];
  a.small = c->T.small;
  a.delta_p = c->T.delta_p;
  a.midx = c->T.midx; a.midy = c->T.midy;
  a.c05h = 0.5 * h; a.hh = h; a.h6 = h / 6.0; a.third = 1.0 / 3.0; a.twothird = 2.0 / 3.0;
  a.ext = c->ext ? 1 : 0;
  return a;
}

template <typename T>
int launch_field(oshun_ctx* c, int mode, const double* fin, const double* f0, double* fout, double* facc,
                 const void* fpl, double h, const double* jp = nullptr, int npb = 0, const double* f1p1 = nullptr) {
  const oshun::FieldArgs<T> a = field_args<T>(c, mode, fin, f0, fout, facc, fpl, h, jp, npb, f1p1);
  dim3 grid((c->nxy + oshun::kFieldX - 1) / oshun::kFieldX);
  ProfRec r;
  prof_begin(c, 1, &r);
  hipLaunchKernelGGL(oshun::field_stage<T>, grid, dim3(oshun::kFieldX, oshun::kFieldRows), 0, c->stream, a);
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

template <typename T>
int launch_stage_pair(oshun_ctx* c, int mode, void* in, void* y0, void* out, double* fin, double* f0, double* fout,
                      double h) {
  int rc = launch_stage<T>(c, mode, in, y0, out, c->A, fin, h);
  if (rc) return rc;
  return launch_field<T>(c, mode, fin, f0, fout, c->Af, in, h);
}

template <typename T>
int rk4(oshun_ctx* c, double h) {
  // Runge_Kutta_4::advance_p1, def-RK.cpp:110-145 (Y0 = S, Y1 = B1/B2 ping-pong, Y = A)
  int rc;
  if ((rc = launch_stage_pair<T>(c, 1, c->S, c->S, c->B1, c->Sf, c->Sf, c->B1f, h))) return rc;
  if ((rc = launch_stage_pair<T>(c, 2, c->B1, c->S, c->B2, c->B1f, c->Sf, c->B2f, h))) return rc;
  if ((rc = launch_stage_pair<T>(c, 3, c->B2, c->S, c->B1, c->B2f, c->Sf, c->B1f, h))) return rc;
  if ((rc = launch_stage_pair<T>(c, 4, c->B1, nullptr, c->S, c->B1f, nullptr, c->Sf, h))) return rc;
  c->t_rk += h;
  return OSHUN_OK;
}

// rows per thread of the pair kernels (tile height 4R-2 p rows): 3 (10 rows);
// measured at the target shape, 14-row float tiles were 15% slower (less halo
// re-read but fewer wavefronts), 6-row tiles re-read more halo.  Grids of
// fewer than two workgroups per CU at R = 3 take R = 2 (c->pair_r): there the
// launch is latency bound and more, shorter row marches finish sooner.
template <typename T>
constexpr int pair_rows() { return 3; }
constexpr int kPairTP = 4 * 3 - 2;
constexpr int kPairSmallL0 = 8;  // R = 2 single-run kernels exist for l0 <= 8
// p tiles of the pair kernels (r = 0: chunks of pw rows of the marched kernels)
inline int pair_tiles_r(int np, int r, int pw = 0) {
  return r == 0 ? (np + pw - 1) / pw : (np + 4 * r - 3) / (4 * r - 2);
}
template <typename T>
int pair_tiles(const oshun_ctx* c) { return pair_tiles_r(c->np, c->pair_r, c->march_pw); }
inline int march_strips(const oshun_ctx* c) { return (c->nx + oshun::kMarchX - 1) / oshun::kMarchX; }
inline int march_waves(const oshun_ctx* c) { return march_strips(c) * pair_tiles_r(c->np, 0, c->march_pw); }
// Wavefronts per SIMD of the marched pair kernels by l0: 512 VGPRs per SIMD
// lane over the larger VGPR count of the two pair kernels as compiled for
// gfx950 (FP64 from 72 at l0 = 1 to 360 at l0 = 12; FP32 from 49 to 226).
inline int march_waves_per_simd(bool fp32, int l0) {
  static const int kF64[13] = {0, 7, 5, 4, 3, 2, 2, 2, 2, 1, 1, 1, 1};
  static const int kF32[13] = {0, 9, 7, 5, 4, 4, 3, 3, 3, 2, 2, 2, 2};
  const int i = std::min(std::max(l0, 1), 12);
  return fp32 ? kF32[i] : kF64[i];
}
// rows per wavefront: the fewest that fit the whole grid in one round of
// wavefronts on the chip's 1024 SIMDs (4 at least: shorter marches re-read
// more halo rows than they save)
inline int march_rows(int np, int nstrip, int wps) {
  const int per_strip = std::max(1, (1024 * wps) / nstrip);
  return std::max(4, (np + per_strip - 1) / per_strip);
}
constexpr int kPairTPMin = 4 * 2 - 2;  // smallest tile height the batched path can choose

template <typename T, int L0>
void launch_pair_l0(oshun_ctx* c, int pair, const oshun::Stage1DArgs<T>& a, dim3 grid) {
  if (c->pair_r == 0) {
    if (pair == 0)
      hipLaunchKernelGGL((oshun::vlasov_march_1d<T, L0, 0>), grid, dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((oshun::vlasov_march_1d<T, L0, 1>), grid, dim3(256), 0, c->stream, a);
    return;
  }
  if constexpr (L0 <= kPairSmallL0) {
    if (c->pair_r == 2) {
      if (pair == 0)
        hipLaunchKernelGGL((oshun::vlasov_pair_1d<T, L0, 2, 0>), grid, dim3(64, 4), 0, c->stream, a);
      else
        hipLaunchKernelGGL((oshun::vlasov_pair_1d<T, L0, 2, 1>), grid, dim3(64, 4), 0, c->stream, a);
      return;
    }
  }
  if (pair == 0)
    hipLaunchKernelGGL((oshun::vlasov_pair_1d<T, L0, pair_rows<T>(), 0>), grid, dim3(64, 4), 0, c->stream, a);
  else
    hipLaunchKernelGGL((oshun::vlasov_pair_1d<T, L0, pair_rows<T>(), 1>), grid, dim3(64, 4), 0, c->stream, a);
}

// RK4 stages 1+2 (pair 0) or 3+4 (pair 1) of the 1D m0 = 0 system
template <typename T>
oshun::Stage1DArgs<T> pair_args(oshun_ctx* c, const void* in, const void* y0, void* out, void* acc,
                                const double* fin, const double* fin2, double* jpA, double* jpB, double h) {
  oshun::Stage1DArgs<T> a;
  fill_1d_args<T>(c, a, 0, in, y0, out, acc, fin, h);
  a.fin2 = fin2;
  a.jpA = jpA;
  a.jpB = jpB;
  a.f1A = c->f1y;
  a.pw = c->march_pw;
  a.rowc = static_cast<const T*>(c->d_rowc);
  a.eAT = static_cast<const T*>(c->d_eAT);
  return a;
}

template <typename T>
int launch_pair(oshun_ctx* c, int pair, const void* in, const void* y0, void* out, void* acc, const double* fin,
                const double* fin2, double* jpA, double* jpB, double h) {
  const oshun::Stage1DArgs<T> a = pair_args<T>(c, in, y0, out, acc, fin, fin2, jpA, jpB, h);
  const int ntx = (c->nx + oshun::kPairX - 1) / oshun::kPairX;
  // padded to a multiple of the 8 XCDs so that xcd_tile gives each XCD one
  // contiguous run of tiles (p-tiles of a strip share their halo rows in L2);
  // the marched kernels take four (strip, chunk) wavefronts per workgroup
  const int nwg = c->pair_r == 0 ? (march_waves(c) + 3) / 4 : ntx * pair_tiles<T>(c);
  dim3 grid((nwg + 7) / 8 * 8);
  ProfRec r;
  prof_begin(c, 0, &r);
  switch (c->T.l0) {
    case 1: launch_pair_l0<T, 1>(c, pair, a, grid); break;
    case 2: launch_pair_l0<T, 2>(c, pair, a, grid); break;
    case 3: launch_pair_l0<T, 3>(c, pair, a, grid); break;
    case 4: launch_pair_l0<T, 4>(c, pair, a, grid); break;
    case 5: launch_pair_l0<T, 5>(c, pair, a, grid); break;
    case 6: launch_pair_l0<T, 6>(c, pair, a, grid); break;
    case 7: launch_pair_l0<T, 7>(c, pair, a, grid); break;
    case 8: launch_pair_l0<T, 8>(c, pair, a, grid); break;
    case 9: launch_pair_l0<T, 9>(c, pair, a, grid); break;
    case 10: launch_pair_l0<T, 10>(c, pair, a, grid); break;
    case 11: launch_pair_l0<T, 11>(c, pair, a, grid); break;
    default: launch_pair_l0<T, 12>(c, pair, a, grid); break;
  }
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

// Runge_Kutta_4::advance_p1 (def-RK.cpp:110-145) for 1D m0 = 0 with the
// stages fused in pairs.  Y0 = S; Y2 -> B2; accumulator A; Y -> B1, then S
// and B1 swap (the second pair reads Y0 around its tile while other tiles
// write Y, so Y cannot go to S in place).  The field stages are unchanged:
// each uses the current of its stage input.
template <typename T>
int rk4_pair(oshun_ctx* c, double h) {
  int rc;
  const bool have1 = c->next1_ready;
  c->next1_ready = false;
  if (!have1) {  // first field stage: the current of Y0 summed directly over p (no partials)
    oshun::FieldArgs<T> a = field_args<T>(c, 1, c->Sf, c->Sf, c->B1f, c->Af, c->S, h, nullptr, 0, nullptr);
    a.p3og = c->d_p3og;
    ProfRec r;
    prof_begin(c, 1, &r);
    hipLaunchKernelGGL(oshun::field_stage<T>, dim3((c->nxy + oshun::kFieldX - 1) / oshun::kFieldX),
                       dim3(oshun::kFieldX, oshun::kFieldRows), 0, c->stream, a);
    prof_end(c, &r);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = launch_pair<T>(c, 0, c->S, nullptr, c->B2, c->A, c->Sf, c->B1f, c->jp1, c->jp2, h))) return rc;
  const int npt = pair_tiles<T>(c);
  {  // field stages 2 and 3 in one pass (E2 -> B2f, E3 -> Cf)
    oshun::Field23Args<T> f23;
    f23.a2 = field_args<T>(c, 2, c->B1f, c->Sf, c->B2f, c->Af, nullptr, h, c->jp1, npt, c->f1y);
    f23.a3 = field_args<T>(c, 3, c->B2f, c->Sf, c->Cf, c->Af, c->B2, h, c->jp2, npt, nullptr);
    ProfRec r;
    prof_begin(c, 1, &r);
    hipLaunchKernelGGL(oshun::field_stage23_1d<T>, dim3((c->nxy + oshun::kFieldX - 1) / oshun::kFieldX),
                       dim3(oshun::kFieldX, oshun::kFieldRows), 0, c->stream, f23);
    prof_end(c, &r);
    HIPCHK(c, hipGetLastError());
  }
  if ((rc = launch_pair<T>(c, 1, c->B2, c->S, c->B1, c->A, c->B2f, c->Cf, c->jp1, nullptr, h))) return rc;
  // the last field stage: here, or in the collision launch that follows
  if (!c->fuse_fields && (rc = launch_field<T>(c, 4, c->Cf, nullptr, c->Sf, c->Af, nullptr, h, c->jp1, npt, c->f1y)))
    return rc;
  std::swap(c->S, c->B1);
  c->t_rk += h;
  return OSHUN_OK;
}

// stage 4 of this step and stage 1 of the next in one field launch (steps
// without collisions; the caller swaps Af/Af2 and marks next1_ready)
template <typename T>
int launch_field41(oshun_ctx* c, double h) {
  oshun::Field41Args<T> A;
  A.a4 = field_args<T>(c, 4, c->Cf, nullptr, c->Sf, c->Af, nullptr, h, c->jp1, pair_tiles<T>(c), c->f1y);
  A.a1 = field_args<T>(c, 1, c->Sf, c->Sf, c->B1f, c->Af2, c->S, h, nullptr, 0, nullptr);
  A.a1.p3og = c->d_p3og;
  A.af2 = c->Af2;
  ProfRec r;
  prof_begin(c, 1, &r);
  hipLaunchKernelGGL(oshun::field_stage41_1d<T>, dim3((c->nxy + oshun::kFieldX - 1) / oshun::kFieldX),
                     dim3(oshun::kFieldX, oshun::kFieldRows), 0, c->stream, A);
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

int do_rk4(oshun_ctx* c, double h) {
  if (c->use_pair) return c->fp32 ? rk4_pair<float>(c, h) : rk4_pair<double>(c, h);
  return c->fp32 ? rk4<float>(c, h) : rk4<double>(c, h);
}

// workgroups of collide<T, NPL, COLS> resident per CU (registers and LDS)
template <typename T, int NPL, int COLS>
int collide_wgs_per_cu() {
  static int n = 0;  // the same on every MI355X
  if (n == 0) {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, oshun::collide<T, NPL, COLS>, 64 * COLS, 0) != hipSuccess ||
        b < 1)
      b = 1;
    n = b;
  }
  return n;
}

template <typename T, int NPL>
int launch_collide_npl(oshun_ctx* c, oshun::CollArgs<T>& a) {
  ProfRec r;
  prof_begin(c, 2, &r);
  auto groups = [c](int cols) { return ((c->nxi + cols - 1) / cols) * c->nyi; };
  // the widest column group that still gives a workgroup per CU
  int cols = groups(8) >= 256 ? 8 : groups(4) >= 256 ? 4 : groups(2) >= 256 ? 2 : 1;
  int split = 1;
  if (!a.do_ee && (a.do_f1 || a.do_flm)) {
    // Without the e-e step the planes of a column group can be split among
    // workgroups (grid.y): a column's planes are one serial chain per
    // wavefront, so the launch takes the (cols, split) with the least
    // rounds of resident workgroups x (planes per chain + the set-up, about
    // three planes' time).
    const int lmin = a.do_f1 ? 1 : 2, lmax = a.do_flm ? a.l0 : 1;
    int n = 0;
    for (int i = 1; i < a.nc; ++i) n += (c->T.comp_l[i] >= lmin && c->T.comp_l[i] <= lmax) ? 1 : 0;
    const int wpc[4] = {collide_wgs_per_cu<T, NPL, 8>(), collide_wgs_per_cu<T, NPL, 4>(),
                        collide_wgs_per_cu<T, NPL, 2>(), collide_wgs_per_cu<T, NPL, 1>()};
    const int colv[4] = {8, 4, 2, 1};
    double best = 1e300;
    for (int k = 0; k < 4; ++k) {
      const long long tiles = groups(colv[k]), cap = 256LL * wpc[k];
      for (int s = 1; s <= 16 && s * 4 <= n; ++s) {
        const double cost = static_cast<double>((tiles * s + cap - 1) / cap) * ((n + s - 1) / s + 3);
        if (cost < best * 0.999) { best = cost; cols = colv[k]; split = s; }
      }
    }
  }
  const dim3 grid(groups(cols), split), block(64 * cols);
  switch (cols) {
    case 8: hipLaunchKernelGGL((oshun::collide<T, NPL, 8>), grid, block, 0, c->stream, a); break;
    case 4: hipLaunchKernelGGL((oshun::collide<T, NPL, 4>), grid, block, 0, c->stream, a); break;
    case 2: hipLaunchKernelGGL((oshun::collide<T, NPL, 2>), grid, block, 0, c->stream, a); break;
    default: hipLaunchKernelGGL((oshun::collide<T, NPL, 1>), grid, block, 0, c->stream, a); break;
  }
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

template <typename T, int NPL, int SEGW, int MODE>
int launch_collide_rows(oshun_ctx* c, oshun::CollArgs<T>& a) {
  const int cols = 256 / SEGW;
  const int ntx = (c->nxi + cols - 1) / cols;
  hipLaunchKernelGGL((oshun::collide_rows<T, NPL, SEGW, MODE>), dim3(ntx * c->nyi), dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

template <typename T>
oshun::CollArgs<T> coll_args(oshun_ctx* c, int do_ee, int do_f1, int do_flm, int numh, double h_ee,
                             double dt_flm) {
  oshun::CollArgs<T> a;
  a.f = static_cast<T*>(c->S);
  a.t = c->ct;
  a.comp_l = c->d_comp_l;
  a.nc = c->nc; a.np = c->np; a.nx = c->nx; a.ny = c->ny; a.nb = c->ph.nb_algorithms; a.l0 = c->T.l0;
  a.plane = c->plane;
  a.do_ee = do_ee; a.do_f1 = do_f1; a.do_flm = do_flm; a.numh = numh;
  a.h_ee = h_ee; a.dt_flm = dt_flm;
  a.c_kpre = c->T.c_kpre; a.kpre = c->T.kpre; a.density_np = c->ph.density_np; a.zeta = c->ph.zeta;
  a.v0sq = c->T.v0sq; a.v1sq = c->T.v1sq; a.g_den = c->T.g_den;
  a.rt = c->d_rt;
  a.bg = c->d_bg;
  a.tclk = nullptr; a.hstep = 0.0; a.small_dt = c->ph.small_dt;
  a.r01 = c->T.v0sq / c->T.v1sq; a.inv_gden = 1.0 / c->T.g_den; a.c2v0 = 2.0 / c->T.v0sq;
  a.fuse = 0;
  a.fz = {};
  a.cx0 = c->gn; a.cy0 = c->gn; a.cnx = c->nxi; a.cny = c->nyi;  // interior columns (Euler_Backward
  return a;                                                       // and Explicit_EC_FP skip guards)
}

// the field stages a fused collision launch takes over (see oshun_ctx::fuse_fields)
template <typename T>
void set_fused_fields(oshun_ctx* c, oshun::CollArgs<T>& a, double h) {
  a.fuse = 1;
  a.fz.cf = c->Cf; a.fz.af = c->Af; a.fz.sf = c->Sf; a.fz.e1 = c->B1f; a.fz.af2 = c->Af2;
  a.fz.jp = c->jp1; a.fz.f1y = c->f1y;
  a.fz.npt = pair_tiles<T>(c);
  a.fz.next1 = c->fuse_fields == 2;
  a.fz.c10 = c->T.comp_re[oshun::hind(1, 0, c->T.m0)];
  a.fz.p3og1 = c->T.p3og[1]; a.fz.small = c->T.small; a.fz.delta_p = c->T.delta_p;
  a.fz.midx = c->T.midx; a.fz.midy = c->T.midy;
  a.fz.c05h = 0.5 * h; a.fz.h6 = h / 6.0; a.fz.third = 1.0 / 3.0;
}

template <typename T>
int collide(oshun_ctx* c, int do_ee, int do_f1, int do_flm, int numh, double h_ee, double dt_flm) {
  oshun::CollArgs<T> a = coll_args<T>(c, do_ee, do_f1, do_flm, numh, h_ee, dt_flm);
  if (c->rows_npl) {
    ProfRec r;
    prof_begin(c, 2, &r);
    int rc = OSHUN_OK;
    if (a.do_ee && (a.do_f1 || a.do_flm) && c->rows_segw == c->rows_segw2 && c->rows_npl == c->rows_npl2) {
      a.rt = c->d_rt;  // one launch for both (same lane layout, same tables)
      if (c->fuse_fields) set_fused_fields<T>(c, a, c->fuse_h);
      switch (c->rows_segw * 100 + c->rows_npl) {
        case 1604: rc = launch_collide_rows<T, 4, 16, 3>(c, a); break;
        case 1608: rc = launch_collide_rows<T, 8, 16, 3>(c, a); break;
        default: rc = launch_collide_rows<T, 8, 32, 3>(c, a); break;
      }
      prof_end(c, &r);
      return rc;
    }
    if (a.do_ee) {
      a.rt = c->d_rt;
      switch (c->rows_segw * 100 + c->rows_npl) {
        case 1604: rc = launch_collide_rows<T, 4, 16, 1>(c, a); break;
        case 1608: rc = launch_collide_rows<T, 8, 16, 1>(c, a); break;
        default: rc = launch_collide_rows<T, 8, 32, 1>(c, a); break;
      }
    }
    if (rc == OSHUN_OK && (a.do_f1 || a.do_flm)) {
      a.rt = c->d_rt2;
      switch (c->rows_segw2 * 100 + c->rows_npl2) {
        case 1604: rc = launch_collide_rows<T, 4, 16, 2>(c, a); break;
        case 1608: rc = launch_collide_rows<T, 8, 16, 2>(c, a); break;
        default: rc = launch_collide_rows<T, 8, 32, 2>(c, a); break;
      }
    }
    prof_end(c, &r);
    return rc;
  }
  if (c->np <= 64) return launch_collide_npl<T, 1>(c, a);
  if (c->np <= 128) return launch_collide_npl<T, 2>(c, a);
  if (c->np <= 256) return launch_collide_npl<T, 4>(c, a);
  return launch_collide_npl<T, 8>(c, a);
}

int do_collide(oshun_ctx* c, int do_ee, int do_f1, int do_flm, int numh, double h_ee, double dt_flm) {
  return c->fp32 ? collide<float>(c, do_ee, do_f1, do_flm, numh, h_ee, dt_flm)
                 : collide<double>(c, do_ee, do_f1, do_flm, numh, h_ee, dt_flm);
}

// sources of one step after the e-e collisions (main.cpp:90-92; sources.hpp)
int step_sources(oshun_ctx* c, double t);
int check_source_flag(oshun_ctx* c);
int fill_guards(oshun_ctx* c, bool b_only = false);  // mirror-boundary contexts (guards.hpp)
bool sources_in_step(const oshun_ctx* c) { return c->src_on && (c->src_pls || c->src_ib == 1); }

// numh and sub-step of Explicit_EC_FP::loop, def-FokkerPlanck.cpp:597-599
void ee_substeps(const oshun_ctx* c, double tnew, int* numh, double* h) {
  size_t n = static_cast<size_t>(static_cast<int>((tnew - c->t_ee) / c->ph.small_dt)) + 1;
  *numh = static_cast<int>(n);
  *h = (tnew - c->t_ee) / static_cast<double>(n);
}

// One step of main.cpp:79-107.  On the 1D pair path with both collision
// operators in one launch, that launch also does the step's last RK4 field
// stage and, if next1 (another step with the same h follows), the next step's
// first field stage, from the collided f10.
int full_step(oshun_ctx* c, double h, bool next1) {
  const int coll = c->ph.collisions;
  const int do_ee = (coll & OSHUN_COLL_EE) ? 1 : 0;
  const int do_flm = (coll & OSHUN_COLL_FLM) ? 1 : 0;
  // (the launch keeps up to oshun::kFuseTiles current partials per column in LDS)
  const int npt = c->fp32 ? pair_tiles<float>(c) : pair_tiles<double>(c);
  // with sources the e-e and implicit collisions are separate launches around them
  const bool srcs = sources_in_step(c);
  const bool fuse = c->use_pair && c->ny == 1 && do_ee && do_flm && c->rows_npl && c->rows_segw == c->rows_segw2 &&
                    c->rows_npl == c->rows_npl2 && npt <= oshun::kFuseTiles && !srcs;
  // without collisions the step's last field stage and the next step's first
  // run as one field launch after the RK4 (fuse_fields = 3)
  const bool f41 = !fuse && next1 && c->use_pair && c->ny == 1 && !do_ee && !do_flm;
  c->fuse_fields = fuse ? (next1 ? 2 : 1) : (f41 ? 3 : 0);
  c->fuse_h = h;
  int rc = do_rk4(c, h);
  if (!rc && f41) {
    rc = c->fp32 ? launch_field41<float>(c, h) : launch_field41<double>(c, h);
    if (!rc) {
      std::swap(c->Af, c->Af2);
      c->next1_ready = true;
    }
  }
  if (rc) { c->fuse_fields = 0; return rc; }
  const double t = c->t_rk;
  int numh = 0;
  double hee = 0.0;
  if (do_ee) ee_substeps(c, t, &numh, &hee);
  if (srcs) {
    if (do_ee) rc = do_collide(c, 1, 0, 0, numh, hee, 0.0);
    if (!rc) rc = step_sources(c, t);
    if (!rc && do_flm) rc = do_collide(c, 0, 1, 1, 0, 0.0, t - c->t_eb);
  } else if (do_ee || do_flm) {
    rc = do_collide(c, do_ee, do_flm, do_flm, numh, hee, t - c->t_eb);
  }
  const int ff = c->fuse_fields;
  c->fuse_fields = 0;
  if (rc) return rc;
  if (ff == 2) {
    std::swap(c->Af, c->Af2);
    c->next1_ready = true;
  }
  c->t_ee = t;
  c->t_eb = t;
  return c->ext ? fill_guards(c, false) : OSHUN_OK;  // CX.Neighbor_Communications(Y), main.cpp:104
}
int full_step(oshun_ctx* c, double h) { return full_step(c, h, false); }

// FP32 split background -------------------------------------------------------
// bg(p) = mean over the interior x,y of Re f00 of the harmonic array sh0
void set_background(oshun_ctx* c, const double* sh0, int nbc) {
  const int np = c->np, nx = c->nx, ny = c->ny, nxg = nx + 2 * nbc;
  c->bg.assign(np, 0.0);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      const double* col = sh0 + 2 * (((size_t)(y + nbc) * nxg + (x + nbc)) * np);
      for (int p = 0; p < np; ++p) c->bg[p] += col[2 * p];
    }
  for (int p = 0; p < np; ++p) c->bg[p] /= static_cast<double>(nx) * ny;
}
// G00 of the background with the formulas of the Vlasov kernels (FP64, then rounded)
void background_g00(const oshun_ctx* c, std::vector<float>& g) {
  const int np = c->np;
  const std::vector<double>& b = c->bg;
  g.assign(np, 0.0f);
  const double fe = (b[0] - b[1] * c->T.p0p1_sq) * c->T.inv_mp0p1_sq;  // def-VlasovMax.cpp:531-532
  g[0] = static_cast<float>((b[1] - fe) * c->T.g_r);
  for (int p = 1; p < np; ++p)
    g[p] = static_cast<float>(p == np - 1 ? b[p - 1] - b[p] : b[p - 1] - b[p + 1]);  // def-State.cpp:359-380
}

// host <-> device repacking ---------------------------------------------------
template <typename T>
void pack(const oshun_ctx* c, const double* const* sh, int nbc, std::vector<T>& out) {
  const int np = c->np, nx = c->nx, ny = c->ny;
  const int nxg = nx + 2 * nbc;
  out.assign((size_t)c->nc * c->plane, T(0));
  for (int h = 0; h < c->T.nh; ++h) {
    const double* src = sh[h];
    const int cr = c->T.comp_re[h], ci = c->T.comp_im[h];
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const double* col = src + 2 * (((size_t)(y + nbc) * nxg + (x + nbc)) * np);
        for (int p = 0; p < np; ++p) {
          const size_t o = ((size_t)y * np + p) * nx + x;
          const double b = (h == 0 && !c->bg.empty()) ? c->bg[p] : 0.0;
          out[(size_t)cr * c->plane + o] = static_cast<T>(col[2 * p] - b);
          if (ci >= 0) out[(size_t)ci * c->plane + o] = static_cast<T>(col[2 * p + 1]);
        }
      }
  }
}

template <typename T>
void unpack(const oshun_ctx* c, const std::vector<T>& in, double* const* sh, int nbc, bool state) {
  const int np = c->np, nx = c->nx, ny = c->ny;
  const int nxg = nx + 2 * nbc, nyg = ny + 2 * nbc;
  for (int h = 0; h < c->T.nh; ++h) {
    double* dst = sh[h];
    const int cr = c->T.comp_re[h], ci = c->T.comp_im[h];
    for (int yg = 0; yg < nyg; ++yg)
      for (int xg = 0; xg < nxg; ++xg) {
        const int y = (((yg - nbc) % ny) + ny) % ny, x = (((xg - nbc) % nx) + nx) % nx;
        double* col = dst + 2 * (((size_t)yg * nxg + xg) * np);
        for (int p = 0; p < np; ++p) {
          const size_t o = ((size_t)y * np + p) * nx + x;
          const double b = (state && h == 0 && !c->bg.empty()) ? c->bg[p] : 0.0;
          col[2 * p] = static_cast<double>(in[(size_t)cr * c->plane + o]) + b;
          col[2 * p + 1] = ci >= 0 ? static_cast<double>(in[(size_t)ci * c->plane + o]) : 0.0;
        }
      }
  }
}

void pack_fields(const oshun_ctx* c, const double* const* fld, int nbc, std::vector<double>& out) {
  const int nx = c->nx, ny = c->ny, nxg = nx + 2 * nbc;
  out.assign((size_t)6 * c->nxy, 0.0);
  for (int k = 0; k < 6; ++k)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        out[(size_t)k * c->nxy + (size_t)y * nx + x] = fld[k][2 * ((size_t)(y + nbc) * nxg + (x + nbc))];
}
void unpack_fields(const oshun_ctx* c, const std::vector<double>& in, double* const* fld, int nbc) {
  const int nx = c->nx, ny = c->ny, nxg = nx + 2 * nbc, nyg = ny + 2 * nbc;
  for (int k = 0; k < 6; ++k)
    for (int yg = 0; yg < nyg; ++yg)
      for (int xg = 0; xg < nxg; ++xg) {
        const int y = (((yg - nbc) % ny) + ny) % ny, x = (((xg - nbc) % nx) + nx) % nx;
        double* d = fld[k] + 2 * ((size_t)yg * nxg + xg);
        d[0] = in[(size_t)k * c->nxy + (size_t)y * nx + x];
        d[1] = 0.0;
      }
}

// state: fsrc holds a state (plane 0 relative to the background), not a slope
int download_raw(oshun_ctx* c, const void* fsrc, const double* fldsrc, double* const* sh, double* const* fld,
                 int nbc, bool state) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->nc * c->plane;
  if (sh) {
    if (c->fp32) {
      std::vector<float> h(n);
      HIPCHK(c, hipMemcpy(h.data(), fsrc, n * sizeof(float), hipMemcpyDeviceToHost));
      unpack<float>(c, h, sh, nbc, state);
    } else {
      std::vector<double> h(n);
      HIPCHK(c, hipMemcpy(h.data(), fsrc, n * sizeof(double), hipMemcpyDeviceToHost));
      unpack<double>(c, h, sh, nbc, state);
    }
  }
  if (fld) {
    std::vector<double> h((size_t)6 * c->nxy);
    HIPCHK(c, hipMemcpy(h.data(), fldsrc, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    unpack_fields(c, h, fld, nbc);
  }
  return OSHUN_OK;
}

void free_all(oshun_ctx* c) {
  void* ptrs[] = {c->S, c->A, c->B1, c->B2, c->Sf, c->Af, c->B1f, c->B2f, c->Cf, c->jpart, c->d_terms, c->d_tstart,
                  c->d_clean, c->d_comp_l, c->d_vt, c->d_rowc, c->d_eAT, c->d_p3og, c->d_coll, c->d_e1d, c->d_mw, c->d_mom, c->d_mpart, c->d_rt, c->d_rt2, c->d_bg, c->d_bgG, c->jp1, c->jp2, c->f1y, c->Af2, c->d_coef, c->d_src, c->d_srcflag, c->d_imp, c->d_imp_zero, c->d_comp_m, c->d_comp_isim};
  for (void* p : ptrs)
    if (p) hipFree(p);
  for (DevProg& q : c->iprog)
    for (void* p : {(void*)q.terms, (void*)q.tstart, (void*)q.clean, (void*)q.gstart, (void*)q.coef})
      if (p) hipFree(p);
  if (c->full.gstart) hipFree(c->full.gstart);
  if (c->full.coef) hipFree(c->full.coef);
  if (c->h_imp_zero) hipHostFree(c->h_imp_zero);
  for (auto& r : c->prec) { hipEventDestroy(r.e0); hipEventDestroy(r.e1); }
  c->prec.clear();
  if (c->stream) hipStreamDestroy(c->stream);
}

bool sane(const oshun_ctx* c) { return c && c->S; }

}  // namespace

// Kernel choice of a context, on the host only (oshun_create and
// oshun_kernel_plan): c->g, c->ph, c->T, nc, np, nx, ny, plane and esz set.
// The stage-pair kernels address the whole state with 32-bit byte offsets and
// collide_rows a plane; larger states take the 64-bit kernels (the 1D
// single-stage kernel, one wavefront per collision column).
static void plan_kernels(oshun_ctx* c, std::vector<double>* spec_coef) {
  const oshun_grid& g = c->g;
  const unsigned long long state_bytes = (unsigned long long)c->nc * c->plane * c->esz;
  // whole arrays with guards (mirror boundaries) run on the term-list interpreter
  c->use1d = (g.m0 == 0 && g.numy == 1 && g.l0 <= kMaxL0Special) && !c->ext;
  c->use_pair = c->use1d && state_bytes < (1ull << 31);
  c->pair_r = 3;
  if (c->use_pair) {
    const int ntx = (c->nx + oshun::kPairX - 1) / oshun::kPairX;
    const bool fusable = (c->ph.collisions & OSHUN_COLL_EE) && (c->ph.collisions & OSHUN_COLL_FLM);
    if (g.l0 <= kPairSmallL0 && ntx * pair_tiles_r(c->np, 3) < 2 * 256 &&
        (!fusable || pair_tiles_r(c->np, 3) > oshun::kFuseTiles || pair_tiles_r(c->np, 2) <= oshun::kFuseTiles))
      c->pair_r = 2;  // small grid (keeping the fused field stages where R = 3 has them)
  }
  if (c->use_pair && c->pair_r == 3) {
    // grids that fill the chip with marches of at least 4 rows take the
    // marched pair kernels (keeping the fused field stages where the tile
    // kernels have them)
    const bool fusable = (c->ph.collisions & OSHUN_COLL_EE) && (c->ph.collisions & OSHUN_COLL_FLM);
    const int wps = march_waves_per_simd(c->fp32, g.l0);
    const int pw = march_rows(c->np, march_strips(c), wps);
    const int nchunk = pair_tiles_r(c->np, 0, pw);
    if ((long long)march_strips(c) * nchunk >= 1024LL * wps * 3 / 4 && pw <= 32 &&
        (!fusable || nchunk <= oshun::kFuseTiles || pair_tiles_r(c->np, 3) > oshun::kFuseTiles)) {
      c->pair_r = 0;
      c->march_pw = pw;
    }
  }
  c->npt = c->use_pair ? pair_tiles_r(c->np, c->pair_r, c->march_pw) : 0;
  c->spec = -1;
  if (!c->use1d && !c->ext && (unsigned long long)c->plane * sizeof(double) < (1ull << 32)) {  // 32-bit offsets
    std::vector<double> coef;
#define OSHUN_SPEC_FIND(ID, L, M, TD) \
    if (c->spec < 0 && oshun::spec_coefficients<L, M, TD>(c->T, coef)) c->spec = ID;
    OSHUN_STAGE_SPECS(OSHUN_SPEC_FIND)
#undef OSHUN_SPEC_FIND
    if (spec_coef) *spec_coef = coef;
  }
  // collide_rows: SEGW lanes x NPL rows per column; needs NPL >= NB (the
  // parabolic low-p rows stay in lane 0) and SEGW*NPL >= np.  Above np = 128
  // both launches use 32 lanes (two DPP rows) per column: at Nx = 4096 that
  // keeps two wavefronts per SIMD in flight instead of one.  The two launches
  // keep separate tables so that their shapes can differ.
  c->rows_npl = c->rows_npl2 = 0;
  c->rows_segw = c->rows_segw2 = 16;
  const int nb = c->ph.nb_algorithms;
  auto pick = [&](const int (*cand)[2], int n, int* segw, int* npl) {
    for (int k = 0; k < n; ++k)
      if (cand[k][0] * cand[k][1] >= c->np && cand[k][1] >= nb && c->np >= 3) {
        *segw = cand[k][0];
        *npl = cand[k][1];
        return;
      }
  };
  const int ee[3][2] = {{16, 4}, {16, 8}, {32, 8}};
  const int im[3][2] = {{16, 4}, {16, 8}, {32, 8}};
  pick(ee, 3, &c->rows_segw, &c->rows_npl);
  pick(im, 3, &c->rows_segw2, &c->rows_npl2);
  const bool plane_fits = (unsigned long long)c->plane * c->esz <= 0x7ffffff0ull;  // collide_rows' kOut
  if (!c->rows_npl || !c->rows_npl2 || !plane_fits) c->rows_npl = c->rows_npl2 = 0;  // both or neither
}

// ---- scratch of the allocation-free moments (oshun_moments_noalloc) ---------
// The partial p sums live in this file-scope device array, one copy per GPU,
// held until the process ends.  Its name is a host-side stand-in: host code
// uses the device address that hipGetSymbolAddress returns for the device
// current at the time (kept per device below), never the name itself.
constexpr size_t kNoAllocScratchDoubles = (size_t)1 << 20;  // 8 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

// p chunks of the moment sums: kMomChunkMin rows each at least, and no more
// partial sums than the static scratch holds, so that oshun_moments and
// oshun_moments_noalloc always sum in the same order (bitwise equal results)
static int moment_chunks(const oshun_ctx* c) {
  const size_t per_chunk = (size_t)oshun::kMomSums * c->nxy;
  const int cap = static_cast<int>(std::min<size_t>(kNoAllocScratchDoubles / per_chunk, oshun::kMomChunkMax));
  return std::max(1, std::min(cap, c->np / oshun::kMomChunkMin));
}

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on device dev (resolved once per device, with dev current)
static int scratch_address(oshun_ctx* c, double** out) {
  if (c->device < 0 || c->device >= kMaxDevices) return fail(c, OSHUN_ERR_ARG, "moments_noalloc: device id");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[c->device]) {
    int prev = 0;
    HIPCHK(c, hipGetDevice(&prev));
    if (prev != c->device) HIPCHK(c, hipSetDevice(c->device));
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (prev != c->device) HIPCHK(c, hipSetDevice(prev));
    if (e != hipSuccess)
      return fail(c, OSHUN_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[c->device] = static_cast<double*>(ptr);
  }
  *out = g_scratch_ptr[c->device];
  return OSHUN_OK;
}

static_assert(oshun::kMomOut == OSHUN_NMOMENTS, "moment planes of the kernels and of the C ABI");


// the three moment launches on c's stream; part holds nchunk x kMomSums planes
// of nxy doubles, out kMomOut planes
template <typename T>
static int launch_moments(oshun_ctx* c, double* part, int nchunk, double* out) {
  oshun::MomArgs<T> a;
  a.f = static_cast<const T*>(c->S);
  a.part = part;
  a.out = out;
  const size_t n = c->np;
  a.w2 = c->d_mw; a.w3 = c->d_mw + n; a.w4 = c->d_mw + 2 * n; a.w5 = c->d_mw + 3 * n;
  a.bg = c->d_bg;
  a.np = c->np; a.nx = c->nx; a.ny = c->ny;
  a.nchunk = nchunk;
  a.plen = (c->np + nchunk - 1) / nchunk;
  const int l0 = c->T.l0, m0 = c->T.m0;
  auto re = [&](int l, int m) { return (l <= l0 && m <= m0) ? c->T.comp_re[oshun::hind(l, m, m0)] : -1; };
  auto im = [&](int l, int m) { return (l <= l0 && m <= m0 && m > 0) ? c->T.comp_im[oshun::hind(l, m, m0)] : -1; };
  a.c00 = re(0, 0); a.c10 = re(1, 0); a.c11r = re(1, 1); a.c20 = re(2, 0);
  a.c21r = re(2, 1); a.c21i = im(2, 1); a.c22i = im(2, 2);
  a.plane = c->plane;
  a.dp = c->T.dpd; a.r = c->T.m_r; a.inv = c->T.m_inv; a.c5 = c->T.m_c5; a.c7 = c->T.m_c7;
  a.c00k = c->T.m_c00; a.c04k = c->T.m_c04;
  a.density_np = c->ph.density_np; a.zeta = c->ph.zeta; a.dx = c->g.dx; a.dy = c->g.dy;
  const dim3 g1((c->nx + 63) / 64, nchunk, c->ny), g2((c->nxy + 255) / 256);
  hipLaunchKernelGGL(oshun::moments_partial<T>, g1, dim3(64), 0, c->stream, a);
  hipLaunchKernelGGL(oshun::moments_final<T>, g2, dim3(256), 0, c->stream, a);
  hipLaunchKernelGGL(oshun::moments_cond<T>, g2, dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}


extern "C" {

int oshun_create(const oshun_grid* grid, const oshun_phys* phys, oshun_ctx** out) {
  if (!grid || !phys || !out) return OSHUN_ERR_ARG;
  *out = nullptr;
  const oshun_grid& g = *grid;
  if (g.l0 < 1 || g.m0 < 0 || g.m0 > g.l0 || g.nump < 4 || g.nump > 512 || g.numx < 3 || g.numy < 1 ||
      !(g.pmax > g.pmin) || !(g.pmin > 0.0) || !(g.dx > 0.0) || !(g.dy > 0.0) ||
      (g.precision != OSHUN_FP64 && g.precision != OSHUN_FP32_SPLIT))
    return OSHUN_ERR_ARG;
  if (phys->nb_algorithms < 0 || phys->nb_algorithms > g.nump || !(phys->small_dt > 0.0) ||
      !(phys->density_np > 0.0))
    return OSHUN_ERR_ARG;
  const bool ext = g.bndx != 0 || g.bndy != 0;
  if (g.bndx < 0 || g.bndx > 1 || g.bndy < 0 || g.bndy > 1) return OSHUN_ERR_ARG;
  if (ext && (g.nbc < 1 || g.numx < g.nbc || g.numy < g.nbc || g.precision != OSHUN_FP64)) return OSHUN_ERR_ARG;
  oshun_ctx* c = new oshun_ctx();
  c->g = g;
  c->ph = *phys;
  c->fp32 = g.precision == OSHUN_FP32_SPLIT;
  c->esz = c->fp32 ? 4 : 8;
  c->ext = ext;
  c->gn = ext ? g.nbc : 0;
  c->nxi = g.numx; c->nyi = g.numy;
  const int nxa = g.numx + 2 * c->gn, nya = g.numy + 2 * c->gn;  // the arrays the device holds
  oshun::build_tables(c->T, g.l0, g.m0, g.nump, nya, g.pmin, g.pmax, g.dx, g.dy, phys->density_np,
                      phys->nb_algorithms, phys->ref_quirks);
  c->nc = c->T.nc; c->np = g.nump; c->nx = nxa; c->ny = nya; c->nxy = nxa * nya;
  c->plane = (long long)nxa * nya * g.nump;
  std::vector<double> spec_coef;
  plan_kernels(c, &spec_coef);
  const int prows = c->use1d ? oshun::kRows1D : oshun::kTileP;
  c->npb = (g.nump + prows - 1) / prows;
  auto bail = [&](int rc) { free_all(c); std::string m = c->err; delete c; (void)m; return rc; };
  if (hipGetDevice(&c->device) != hipSuccess) return bail(OSHUN_ERR_HIP);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return bail(OSHUN_ERR_HIP);
  const size_t fbytes = (size_t)c->nc * c->plane * c->esz;
  const size_t fldbytes = (size_t)6 * c->nxy * sizeof(double);
  void** fb[] = {&c->S, &c->A, &c->B1, &c->B2};
  for (void** p : fb)
    if (hipMalloc(p, fbytes) != hipSuccess) return bail(OSHUN_ERR_ALLOC);
  double** fl[] = {&c->Sf, &c->Af, &c->B1f, &c->B2f};
  for (double** p : fl)
    if (hipMalloc((void**)p, fldbytes) != hipSuccess) return bail(OSHUN_ERR_ALLOC);
  for (void** p : fb)
    if (hipMemset(*p, 0, fbytes) != hipSuccess) return bail(OSHUN_ERR_HIP);
  for (double** p : fl)
    if (hipMemset(*p, 0, fldbytes) != hipSuccess) return bail(OSHUN_ERR_HIP);
  if (hipMalloc((void**)&c->jpart, (size_t)c->npb * 3 * c->nxy * sizeof(double)) != hipSuccess)
    return bail(OSHUN_ERR_ALLOC);
  if (c->use_pair) {
    constexpr int kTileMin = kPairTPMin < 4 ? kPairTPMin : 4;  // tile kernels' 6 rows, marched kernels' 4
    const int npt_max = (c->np + kTileMin - 1) / kTileMin;      // partial buffers fit every tile height
    const size_t nx = c->nx;
    if (hipMalloc((void**)&c->jp1, (size_t)npt_max * 3 * nx * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->jp2, (size_t)npt_max * 3 * nx * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->f1y, nx * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->Cf, fldbytes) != hipSuccess || hipMalloc((void**)&c->Af2, fldbytes) != hipSuccess)
      return bail(OSHUN_ERR_ALLOC);
    if (hipMemset(c->Cf, 0, fldbytes) != hipSuccess || hipMemset(c->Af2, 0, fldbytes) != hipSuccess)
      return bail(OSHUN_ERR_HIP);
  }
  if (c->fp32) {
    c->bg.assign(c->np, 0.0);
    if (upload_vec(c, &c->d_bg, c->bg) || upload_vec(c, &c->d_bgG, std::vector<float>(c->np, 0.0f)))
      return bail(OSHUN_ERR_HIP);
  }
  if (upload_vec(c, &c->d_terms, c->T.terms) || upload_vec(c, &c->d_tstart, c->T.tstart) ||
      upload_vec(c, &c->d_clean, c->T.clean) || upload_vec(c, &c->d_comp_l, c->T.comp_l) ||
      upload_vec(c, &c->d_comp_m, c->T.comp_m) || upload_vec(c, &c->d_comp_isim, c->T.comp_isim) ||
      upload_vec(c, &c->d_p3og, c->T.p3og))
    return bail(OSHUN_ERR_HIP);
  if (c->spec >= 0 && upload_vec(c, &c->d_coef, spec_coef)) return bail(OSHUN_ERR_HIP);
  if (lm_prepare(c, c->T.terms, c->T.tstart, c->T.clean, &c->full)) return bail(OSHUN_ERR_HIP);
  {
    std::vector<double> e;
    e.insert(e.end(), c->T.e1d_A1.begin(), c->T.e1d_A1.end());
    e.insert(e.end(), c->T.e1d_A2.begin(), c->T.e1d_A2.end());
    e.insert(e.end(), c->T.x1d_A1.begin(), c->T.x1d_A1.end());
    e.insert(e.end(), c->T.x1d_A2.begin(), c->T.x1d_A2.end());
    if (upload_vec(c, &c->d_e1d, e)) return bail(OSHUN_ERR_HIP);
  }
  {
    std::vector<double> w;
    for (auto* t : {&c->T.mw2, &c->T.mw3, &c->T.mw4, &c->T.mw5}) w.insert(w.end(), t->begin(), t->end());
    if (upload_vec(c, &c->d_mw, w)) return bail(OSHUN_ERR_HIP);
  }
  {
    std::vector<double> vt;
    vt.insert(vt.end(), c->T.invpax.begin(), c->T.invpax.end());
    vt.insert(vt.end(), c->T.vtab.begin(), c->T.vtab.end());
    vt.insert(vt.end(), c->T.cg.begin(), c->T.cg.end());
    vt.insert(vt.end(), c->T.hp0.begin(), c->T.hp0.end());
    // x-advection multipliers v(p)*A of the 1D kernels, formed in the storage precision
    auto xv = [&](auto zero, std::vector<decltype(zero)>& out) {
      using V = decltype(zero);
      const size_t L = c->T.l0 + 1, np = c->np;
      out.assign(2 * L * np, V(0));
      for (size_t l = 0; l < L; ++l)
        for (size_t q = 0; q < np; ++q) {
          const V v = static_cast<V>(c->T.vtab[q]);
          out[l * np + q] = v * static_cast<V>(c->T.x1d_A1[l]);
          out[(L + l) * np + q] = v * static_cast<V>(c->T.x1d_A2[l]);
        }
    };
    // the same invpax and x-advection multipliers by row for the marched pair
    // kernels: [p][invpax l = 1..l0 | xv1 l = 0..l0-1 | xv2 l = 1..l0]
    auto rowc = [&](const auto& x2, auto zero, std::vector<decltype(zero)>& out) {
      using V = decltype(zero);
      const size_t L0 = c->T.l0, L = L0 + 1, np = c->np;
      out.assign(3 * L0 * np, V(0));
      for (size_t q = 0; q < np; ++q)
        for (size_t j = 0; j < L0; ++j) {
          out[q * 3 * L0 + j] = static_cast<V>(c->T.invpax[(j + 1) * np + q]);
          out[q * 3 * L0 + L0 + j] = x2[j * np + q];
          out[q * 3 * L0 + 2 * L0 + j] = x2[(L + j + 1) * np + q];
        }
    };
    {  // eA1 | eA2 of the 1D kernels in the storage precision (the marched kernels)
      std::vector<double> ea(c->T.e1d_A1.begin(), c->T.e1d_A1.end());
      ea.insert(ea.end(), c->T.e1d_A2.begin(), c->T.e1d_A2.end());
      if (c->fp32) {
        float* d = nullptr;
        if (upload_vec(c, &d, std::vector<float>(ea.begin(), ea.end()))) return bail(OSHUN_ERR_HIP);
        c->d_eAT = d;
      } else {
        double* d = nullptr;
        if (upload_vec(c, &d, ea)) return bail(OSHUN_ERR_HIP);
        c->d_eAT = d;
      }
    }
    if (c->fp32) {
      std::vector<float> vf(vt.begin(), vt.end()), x2, rc;
      xv(0.0f, x2);
      rowc(x2, 0.0f, rc);
      vf.insert(vf.end(), x2.begin(), x2.end());
      float* d = nullptr;
      if (upload_vec(c, &d, vf)) return bail(OSHUN_ERR_HIP);
      c->d_vt = d;
      float* dr = nullptr;
      if (upload_vec(c, &dr, rc)) return bail(OSHUN_ERR_HIP);
      c->d_rowc = dr;
    } else {
      std::vector<double> x2, rc;
      xv(0.0, x2);
      rowc(x2, 0.0, rc);
      vt.insert(vt.end(), x2.begin(), x2.end());
      double* d = nullptr;
      if (upload_vec(c, &d, vt)) return bail(OSHUN_ERR_HIP);
      c->d_vt = d;
      double* dr = nullptr;
      if (upload_vec(c, &dr, rc)) return bail(OSHUN_ERR_HIP);
      c->d_rowc = dr;
    }
  }
  {
    const HostTables& T = c->T;
    const std::vector<double>* tabs[] = {&T.U4, &T.U4m1, &T.U2, &T.U2m1, &T.U1s, &T.U1m1s, &T.U3, &T.Qn,
                                         &T.Pn, &T.vc,  &T.ivdm, &T.ivdp, &T.ivs2, &T.pw3, &T.pw5, &T.pw7};
    std::vector<double> all;
    for (auto* t : tabs) all.insert(all.end(), t->begin(), t->end());
    if (upload_vec(c, &c->d_coll, all)) return bail(OSHUN_ERR_HIP);
    const double* b = c->d_coll;
    const size_t n = c->np;
    c->ct.U4 = b; c->ct.U4m1 = b + n; c->ct.U2 = b + 2 * n; c->ct.U2m1 = b + 3 * n; c->ct.U1s = b + 4 * n;
    c->ct.U1m1s = b + 5 * n; c->ct.U3 = b + 6 * n; c->ct.Qn = b + 7 * n; c->ct.Pn = b + 8 * n;
    c->ct.vc = b + 9 * n; c->ct.ivdm = b + 10 * n; c->ct.ivdp = b + 11 * n; c->ct.ivs2 = b + 12 * n;
    c->ct.pw3 = b + 13 * n; c->ct.pw5 = b + 14 * n; c->ct.pw7 = b + 15 * n;
  }
  auto build_rt = [&](int segw, int npl, std::vector<double>& rt) {
    const HostTables& T = c->T;
    const int np = c->np, nb = std::min(c->ph.nb_algorithms, np);
    const double* v = T.vc.data();
    rt.assign((size_t)oshun::RT_COUNT * oshun::kRowP, 0.0);
    auto put = [&](int k, int p, double val) { rt[(size_t)k * oshun::kRowP + (p % npl) * segw + p / npl] = val; };
    for (int p = 0; p < segw * npl; ++p) {
      const bool in = p < np, mid = p >= 1 && p < np - 1;
      put(oshun::RT_U4, p, in ? T.U4[p] : 0.0);  // U4[0] = 0: the sums start at p = 1
      put(oshun::RT_U4M1, p, in ? T.U4m1[p] : 0.0);
      put(oshun::RT_U2, p, in ? T.U2[p] : 0.0);
      put(oshun::RT_U2M1, p, in ? T.U2m1[p] : 0.0);
      put(oshun::RT_U1S, p, p < np - 1 ? T.U1s[p] : 0.0);
      put(oshun::RT_U1M1S, p, p < np - 1 ? T.U1m1s[p] : 0.0);
      put(oshun::RT_U3, p, in ? T.U3[p] : 1.0);
      put(oshun::RT_PN, p, p < np - 1 ? T.Pn[p] : (p == np - 1 ? 1.0 : 0.0));
      put(oshun::RT_QN, p, p < np - 1 ? T.Qn[p] : 0.0);
      put(oshun::RT_VC, p, in ? v[p] : 1.0);
      put(oshun::RT_IV2, p, in ? 1.0 / (v[p] * v[p]) : 1.0);
      put(oshun::RT_IU3, p, in ? 1.0 / std::pow(v[p], 3) : 0.0);
      put(oshun::RT_IVDM, p, mid ? T.ivdm[p] : 0.0);
      put(oshun::RT_IVDP, p, mid ? T.ivdp[p] : 0.0);
      put(oshun::RT_IVS2, p, mid ? T.ivs2[p] : 0.0);
      put(oshun::RT_ROW, p, p < np - 1 ? 1.0 : 0.0);
      if (p < nb) {
        const double v1sq = T.v1sq;
        put(oshun::RT_P3, p, T.pw3[p]);
        put(oshun::RT_A2, p, T.pw3[p] / 3.0);
        put(oshun::RT_B2, p, T.pw5[p] / v1sq * 0.2);
        put(oshun::RT_A4, p, T.pw5[p] * 0.2);
        put(oshun::RT_B4, p, T.pw7[p] / (v1sq * 7.0));
      }
      // Current::Jxyz integrand weight: trapezoid w (1/2 at both ends, exact) times p^3/gamma
      put(oshun::RT_JW, p, in ? ((p == 0 || p == np - 1) ? 0.5 : 1.0) * T.p3og[p] : 0.0);
    }
  };
  if (c->rows_npl) {
    std::vector<double> rt;
    build_rt(c->rows_segw, c->rows_npl, rt);
    if (upload_vec(c, &c->d_rt, rt)) return bail(OSHUN_ERR_HIP);
    build_rt(c->rows_segw2, c->rows_npl2, rt);
    if (upload_vec(c, &c->d_rt2, rt)) return bail(OSHUN_ERR_HIP);
  }
  *out = c;
  return OSHUN_OK;
}

int oshun_kernel_plan(const oshun_grid* grid, const oshun_phys* phys, int* stage, int* collide) {
  if (!grid || !phys || !stage || !collide) return OSHUN_ERR_ARG;
  const oshun_grid& g = *grid;
  if (g.l0 < 1 || g.m0 < 0 || g.m0 > g.l0 || g.nump < 4 || g.nump > 512 || g.numx < 3 || g.numy < 1 ||
      !(g.pmax > g.pmin) || !(g.pmin > 0.0) || !(g.dx > 0.0) || !(g.dy > 0.0) ||
      (g.precision != OSHUN_FP64 && g.precision != OSHUN_FP32_SPLIT))
    return OSHUN_ERR_ARG;
  oshun_ctx c;  // host fields only: nothing is allocated on the device
  c.g = g;
  c.ph = *phys;
  c.fp32 = g.precision == OSHUN_FP32_SPLIT;
  c.esz = c.fp32 ? 4 : 8;
  c.ext = g.bndx != 0 || g.bndy != 0;
  c.gn = c.ext ? g.nbc : 0;
  c.nxi = g.numx; c.nyi = g.numy;
  const int nxa = g.numx + 2 * c.gn, nya = g.numy + 2 * c.gn;
  oshun::build_tables(c.T, g.l0, g.m0, g.nump, nya, g.pmin, g.pmax, g.dx, g.dy, phys->density_np,
                      phys->nb_algorithms, phys->ref_quirks);
  c.nc = c.T.nc; c.np = g.nump; c.nx = nxa; c.ny = nya; c.nxy = nxa * nya;
  c.plane = (long long)nxa * nya * g.nump;
  plan_kernels(&c, nullptr);
  oshun::LmPlan lp;
  if (!c.use1d && c.spec < 0) oshun::lm_build(c.T, c.T.terms, c.T.tstart, c.T.clean, &lp);
  const bool lm = lp.kind == oshun::kLmFull;
  *stage = c.use_pair ? 0 : (c.use1d ? 1 : (c.spec >= 0 ? 2 : (lm ? 4 : 3)));
  *collide = c.rows_npl ? 0 : 1;
  return OSHUN_OK;
}

int oshun_destroy(oshun_ctx* ctx) {
  if (!ctx) return OSHUN_ERR_ARG;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  oshun_adjoint_release(ctx);
  free_all(ctx);
  delete ctx;
  return OSHUN_OK;
}

const char* oshun_last_error(const oshun_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int oshun_num_harmonics(const oshun_ctx* ctx) { return ctx ? ctx->T.nh : OSHUN_ERR_ARG; }

static int upload_raw(oshun_ctx* c, const double* const* sh, const double* const* fld, int nbc);

// Whole arrays (nbc guards) of a mirror-boundary context, built on the host
// from interior-only ones (nbc = 0): guards zero until the device fills them.
static void ext_whole(const oshun_ctx* c, const double* const* in, int n, int per_cell,
                      std::vector<std::vector<double>>& buf, std::vector<const double*>& ptr) {
  buf.assign(n, std::vector<double>((size_t)2 * per_cell * c->nx * c->ny, 0.0));
  ptr.clear();
  for (int k = 0; k < n; ++k) {
    for (int y = 0; y < c->nyi; ++y)
      for (int x = 0; x < c->nxi; ++x)
        std::copy(in[k] + 2 * (size_t)per_cell * ((size_t)y * c->nxi + x),
                  in[k] + 2 * (size_t)per_cell * ((size_t)y * c->nxi + x + 1),
                  buf[k].begin() + 2 * (size_t)per_cell * ((size_t)(y + c->gn) * c->nx + x + c->gn));
    ptr.push_back(buf[k].data());
  }
}

int oshun_upload_harmonics(oshun_ctx* c, const double* const* sh, const double* const* fld, int nbc) {
  if (!sane(c) || nbc < 0) return OSHUN_ERR_ARG;
  if (c->ext) {  // the device holds whole arrays: nbc = its guards (copied as they are), or 0 (guards filled)
    if (nbc == c->gn) return upload_raw(c, sh, fld, 0);
    if (nbc != 0) return fail(c, OSHUN_ERR_ARG, "upload: nbc must be the context's guard count or 0");
    std::vector<std::vector<double>> bh, bf;
    std::vector<const double*> ph, pf;
    if (sh) ext_whole(c, sh, c->T.nh, c->np, bh, ph);
    if (fld) ext_whole(c, fld, 6, 1, bf, pf);
    int rc = upload_raw(c, sh ? ph.data() : nullptr, fld ? pf.data() : nullptr, 0);
    return rc ? rc : fill_guards(c, false);
  }
  return upload_raw(c, sh, fld, nbc);
}

// the device arrays from host arrays with nbc guard cells (skipped) -- or,
// in a mirror-boundary context, from the whole arrays (nbc = 0 on the
// device's own whole dimensions)
static int upload_raw(oshun_ctx* c, const double* const* sh, const double* const* fld, int nbc) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (sh) {
    for (int h = 0; h < c->T.nh; ++h)
      if (!sh[h]) return fail(c, OSHUN_ERR_ARG, "null harmonic pointer");
    if (c->fp32) {
      set_background(c, sh[0], nbc);
      std::vector<float> v;
      pack<float>(c, sh, nbc, v);
      HIPCHK(c, hipMemcpy(c->S, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice));
      HIPCHK(c, hipMemcpy(c->d_bg, c->bg.data(), c->np * sizeof(double), hipMemcpyHostToDevice));
      std::vector<float> g(c->np);
      background_g00(c, g);
      HIPCHK(c, hipMemcpy(c->d_bgG, g.data(), c->np * sizeof(float), hipMemcpyHostToDevice));
    } else {
      std::vector<double> v;
      pack<double>(c, sh, nbc, v);
      HIPCHK(c, hipMemcpy(c->S, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  if (fld) {
    for (int k = 0; k < 6; ++k)
      if (!fld[k]) return fail(c, OSHUN_ERR_ARG, "null field pointer");
    std::vector<double> v;
    pack_fields(c, fld, nbc, v);
    HIPCHK(c, hipMemcpy(c->Sf, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return OSHUN_OK;
}

// device arrays to host arrays with nbc guards; a mirror-boundary context
// gives its whole arrays (nbc = its guard count) or the interior (nbc = 0)
static int download_any(oshun_ctx* c, const void* fsrc, const double* fldsrc, double* const* sh,
                        double* const* fld, int nbc, bool state) {
  if (!c->ext) return download_raw(c, fsrc, fldsrc, sh, fld, nbc, state);
  if (nbc == c->gn) return download_raw(c, fsrc, fldsrc, sh, fld, 0, state);
  if (nbc != 0) return fail(c, OSHUN_ERR_ARG, "download: nbc must be the context's guard count or 0");
  std::vector<std::vector<double>> bh(sh ? c->T.nh : 0, std::vector<double>(2 * (size_t)c->np * c->nxy)),
      bf(fld ? 6 : 0, std::vector<double>(2 * (size_t)c->nxy));
  std::vector<double*> ph, pf;
  for (auto& b : bh) ph.push_back(b.data());
  for (auto& b : bf) pf.push_back(b.data());
  int rc = download_raw(c, fsrc, fldsrc, sh ? ph.data() : nullptr, fld ? pf.data() : nullptr, 0, state);
  if (rc) return rc;
  auto cut = [&](std::vector<std::vector<double>>& b, double* const* out, int per_cell) {
    for (size_t k = 0; k < b.size(); ++k)
      for (int y = 0; y < c->nyi; ++y)
        for (int x = 0; x < c->nxi; ++x)
          std::copy(b[k].begin() + 2 * (size_t)per_cell * ((size_t)(y + c->gn) * c->nx + x + c->gn),
                    b[k].begin() + 2 * (size_t)per_cell * ((size_t)(y + c->gn) * c->nx + x + c->gn + 1),
                    out[k] + 2 * (size_t)per_cell * ((size_t)y * c->nxi + x));
  };
  cut(bh, sh, c->np);
  cut(bf, fld, 1);
  return OSHUN_OK;
}

int oshun_download_harmonics(oshun_ctx* c, double* const* sh, double* const* fld, int nbc) {
  if (!sane(c) || nbc < 0) return OSHUN_ERR_ARG;
  return download_any(c, c->S, c->Sf, sh, fld, nbc, true);
}

// host arrays (interior plus nbc guards per side) of a state laid out harmonic after harmonic
static void split_ptrs(const oshun_ctx* c, const double* f, const double* fld, int nbc,
                       std::vector<const double*>& hp, std::vector<const double*>& fp) {
  const size_t hsz = 2 * (size_t)c->np * (c->nxi + 2 * nbc) * (c->nyi + 2 * nbc);
  const size_t fsz = 2 * (size_t)(c->nxi + 2 * nbc) * (c->nyi + 2 * nbc);
  hp.clear(); fp.clear();
  if (f) for (int h = 0; h < c->T.nh; ++h) hp.push_back(f + h * hsz);
  if (fld) for (int k = 0; k < 6; ++k) fp.push_back(fld + k * fsz);
}

int oshun_upload_state(oshun_ctx* c, const double* f, const double* fld, int nbc) {
  if (!sane(c) || nbc < 0) return OSHUN_ERR_ARG;
  std::vector<const double*> hp, fp;
  split_ptrs(c, f, fld, nbc, hp, fp);
  return oshun_upload_harmonics(c, f ? hp.data() : nullptr, fld ? fp.data() : nullptr, nbc);
}

int oshun_download_state(oshun_ctx* c, double* f, double* fld, int nbc) {
  if (!sane(c) || nbc < 0) return OSHUN_ERR_ARG;
  std::vector<const double*> hp, fp;
  split_ptrs(c, f, fld, nbc, hp, fp);
  return oshun_download_harmonics(c, f ? const_cast<double* const*>(hp.data()) : nullptr,
                                  fld ? const_cast<double* const*>(fp.data()) : nullptr, nbc);
}

int oshun_set_time(oshun_ctx* c, double t) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  c->t_rk = c->t_ee = c->t_eb = t;
  c->src_t_pls = c->src_t_ib = c->src_t_efld = t;
  return OSHUN_OK;
}

int oshun_get_time(const oshun_ctx* c, double* t) {
  if (!sane(c) || !t) return OSHUN_ERR_ARG;
  *t = c->t_rk;
  return OSHUN_OK;
}

int oshun_rhs(oshun_ctx* c, double* f_slope, double* fld_slope) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  int rc = c->fp32 ? launch_stage_pair<float>(c, 0, c->S, c->S, c->B1, c->Sf, c->Sf, c->B1f, 0.0)
                   : launch_stage_pair<double>(c, 0, c->S, c->S, c->B1, c->Sf, c->Sf, c->B1f, 0.0);
  if (rc) return rc;
  std::vector<const double*> hp, fp;
  split_ptrs(c, f_slope, fld_slope, 0, hp, fp);
  return download_any(c, c->B1, c->B1f, f_slope ? const_cast<double* const*>(hp.data()) : nullptr,
                      fld_slope ? const_cast<double* const*>(fp.data()) : nullptr, 0, false);
}

int oshun_rk4_step(oshun_ctx* c, double h) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  return do_rk4(c, h);
}

int oshun_collide_ee(oshun_ctx* c, double tnew) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  int numh;
  double h;
  ee_substeps(c, tnew, &numh, &h);
  int rc = do_collide(c, 1, 0, 0, numh, h, 0.0);
  if (rc) return rc;
  c->t_ee = tnew;
  return OSHUN_OK;
}

int oshun_collide_flm(oshun_ctx* c, double tnew, int which) {
  if (!sane(c) || (which != 1 && which != 2)) return OSHUN_ERR_ARG;
  int rc = do_collide(c, 0, which == 1, which == 2, 0, 0.0, tnew - c->t_eb);
  if (rc) return rc;
  if (which == 2) c->t_eb = tnew;
  return OSHUN_OK;
}

int oshun_step(oshun_ctx* c, double h, int nsteps) {
  if (!sane(c) || nsteps < 0) return OSHUN_ERR_ARG;
  c->next1_ready = false;
  for (int s = 0; s < nsteps; ++s) {
    int rc = full_step(c, h, s + 1 < nsteps);
    if (rc) { c->next1_ready = false; return rc; }
  }
  // a laser-source heating rate out of range stops the reference (def-PLSource.cpp:230-233)
  return (c->src_on && c->src_pls) ? check_source_flag(c) : OSHUN_OK;
}

int oshun_synchronize(oshun_ctx* c) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_get_stream(const oshun_ctx* c, void** s) {
  if (!sane(c) || !s) return OSHUN_ERR_ARG;
  *s = c->stream;
  return OSHUN_OK;
}

int oshun_device_state(oshun_ctx* c, void** f_dev, double** fld_dev) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (f_dev) *f_dev = c->S;
  if (fld_dev) *fld_dev = c->Sf;
  return OSHUN_OK;
}

int oshun_copy_device_state(oshun_ctx* c, void* f_dev, double* fld_dev, int into_ctx) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  const size_t fb = (size_t)c->nc * c->plane * c->esz, gb = (size_t)6 * c->nxy * sizeof(double);
  if (f_dev) {
    if (into_ctx) HIPCHK(c, hipMemcpyAsync(c->S, f_dev, fb, hipMemcpyDeviceToDevice, c->stream));
    else HIPCHK(c, hipMemcpyAsync(f_dev, c->S, fb, hipMemcpyDeviceToDevice, c->stream));
  }
  if (fld_dev) {
    if (into_ctx) HIPCHK(c, hipMemcpyAsync(c->Sf, fld_dev, gb, hipMemcpyDeviceToDevice, c->stream));
    else HIPCHK(c, hipMemcpyAsync(fld_dev, c->Sf, gb, hipMemcpyDeviceToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_plane_info(const oshun_ctx* c, int* ncomp, long long* plane_elems, int* comp_l, int* comp_m,
                     int* comp_im) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (ncomp) *ncomp = c->nc;
  if (plane_elems) *plane_elems = c->plane;
  for (int i = 0; i < c->nc; ++i) {
    if (comp_l) comp_l[i] = c->T.comp_l[i];
    if (comp_m) comp_m[i] = c->T.comp_m[i];
    if (comp_im) comp_im[i] = c->T.comp_isim[i];
  }
  return OSHUN_OK;
}

int oshun_profile_enable(oshun_ctx* c, int on) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  c->prof = on != 0;
  return OSHUN_OK;
}

int oshun_profile_read(oshun_ctx* c, double* ms_total, long long* launches, int nclasses) {
  if (!sane(c) || nclasses < 0) return OSHUN_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (auto& r : c->prec) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, r.e0, r.e1) == hipSuccess && r.cls >= 0 && r.cls < 4) {
      c->prof_ms[r.cls] += ms;
      c->prof_n[r.cls] += 1;
    }
    hipEventDestroy(r.e0);
    hipEventDestroy(r.e1);
  }
  c->prec.clear();
  for (int k = 0; k < nclasses && k < 4; ++k) {
    if (ms_total) ms_total[k] = c->prof_ms[k];
    if (launches) launches[k] = c->prof_n[k];
    c->prof_ms[k] = 0.0;
    c->prof_n[k] = 0;
  }
  return OSHUN_OK;
}

int oshun_moments(oshun_ctx* c, double* out) {
  if (!sane(c) || !out) return OSHUN_ERR_ARG;
  if (c->ext) return fail(c, OSHUN_ERR_ARG, "moments: not on a mirror-boundary context (whole arrays)");
  const int nchunk = moment_chunks(c);
  if (!c->d_mom) {  // output and partial-sum buffers on the first call
    if (hipMalloc((void**)&c->d_mom, (size_t)oshun::kMomOut * c->nxy * sizeof(double)) != hipSuccess ||
        hipMalloc((void**)&c->d_mpart, (size_t)nchunk * oshun::kMomSums * c->nxy * sizeof(double)) != hipSuccess)
      return fail(c, OSHUN_ERR_ALLOC, "moments: device buffers");
  }
  int rc = c->fp32 ? launch_moments<float>(c, c->d_mpart, nchunk, c->d_mom)
                   : launch_moments<double>(c, c->d_mpart, nchunk, c->d_mom);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(out, c->d_mom, (size_t)oshun::kMomOut * c->nxy * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return OSHUN_OK;
}

int oshun_moments_noalloc(oshun_ctx* c, double* out_dev) {
  if (!sane(c) || !out_dev) return OSHUN_ERR_ARG;
  if (c->ext) return fail(c, OSHUN_ERR_ARG, "moments: not on a mirror-boundary context (whole arrays)");
  const size_t per_chunk = (size_t)oshun::kMomSums * c->nxy;  // doubles of one chunk's partial sums
  if (per_chunk > kNoAllocScratchDoubles)
    return fail(c, OSHUN_ERR_ARG, "oshun_moments_noalloc: nx*ny = " + std::to_string(c->nxy) + " exceeds " +
                                      std::to_string(kNoAllocScratchDoubles / oshun::kMomSums) +
                                      " (the static scratch holds " + std::to_string(kNoAllocScratchDoubles) +
                                      " doubles of partial sums)");
  double* part = nullptr;
  int rc = scratch_address(c, &part);
  if (rc) return rc;
  // the chunking of oshun_moments (it fits the scratch by construction); moments_partial
  // writes every partial sum it forms, so the scratch needs no clearing
  const int nchunk = moment_chunks(c);
  return c->fp32 ? launch_moments<float>(c, part, nchunk, out_dev) : launch_moments<double>(c, part, nchunk, out_dev);
}

int oshun_pair_rows(oshun_ctx* c, int set, int* rows) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (set != 0) {
    if (!c->use_pair) return fail(c, OSHUN_ERR_ARG, "no stage-pair kernels on this context (1D, m0 = 0 only)");
    if (set == 10) c->pair_r = 3;
    else if (set == 6 && c->T.l0 <= kPairSmallL0) c->pair_r = 2;
    else if (set <= -4 && set >= -32) { c->pair_r = 0; c->march_pw = -set; }
    else return fail(c, OSHUN_ERR_ARG, "tile height must be 10, or 6 for l0 <= 8; -4 .. -32 for the marched kernels");
    c->npt = pair_tiles_r(c->np, c->pair_r, c->march_pw);
    c->next1_ready = false;
  }
  if (rows) *rows = c->use_pair ? (c->pair_r == 0 ? -c->march_pw : 4 * c->pair_r - 2) : 0;
  return OSHUN_OK;
}

int oshun_stage_kernel(oshun_ctx* c, int set, int* which) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (set == 0 || set == 1) c->force_general = (set == 1);
  else if (set != -1) return fail(c, OSHUN_ERR_ARG, "set must be -1, 0 or 1");
  if (which) *which = c->use1d ? 0 : (c->force_general ? 2 : (c->spec >= 0 ? 1 : (c->full.lm >= 0 ? 3 : 2)));
  return OSHUN_OK;
}

int oshun_diag_copy(oshun_ctx* c, long long nbytes) {
  if (!sane(c) || nbytes < 16) return OSHUN_ERR_ARG;
  const long long cap = (long long)c->nc * c->plane * (long long)c->esz;
  if (nbytes > cap) return fail(c, OSHUN_ERR_ARG, "diag copy larger than the state buffer");
  const long long n = nbytes / 8;
  ProfRec r;
  prof_begin(c, 3, &r);
  hipLaunchKernelGGL(oshun::diag_copy, dim3(2048), dim3(256), 0, c->stream, static_cast<const double*>(c->B1),
                     static_cast<double*>(c->B2), n);
  prof_end(c, &r);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

}  // extern "C"

// batched runs (oshun_batch_*)
#include "batch.hpp"
// adjoint of the step (oshun_adjoint_*)
#include "adjoint.hpp"
// laser heating, inverse bremsstrahlung, external field (oshun_set_sources, ...)
#include "sources.hpp"
// implicit-E step, 1D variant (oshun_step_implicit_1d)
#include "implicit.hpp"
#include "implicit2d.hpp"
#include "guards.hpp"
