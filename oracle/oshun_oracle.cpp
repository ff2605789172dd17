// OSHUN distribution-function update: CPU restatement used as the ORACLE.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load this library, and only as the checker
// (or the timed CPU baseline).  The product path (oshun_amd, liboshun_hip.so)
// never calls it.
//
// What it restates (reference = OSHUN-1, files cited as file:line):
//   * the Vlasov-Maxwell right-hand side  Actions::Adv   def-Actions.cpp:73-94
//       ElectricField::Exyz / MakeGH / G00  def-VlasovMax.cpp:76-152,346-535
//       MagneticField::Bxyz                 def-VlasovMax.cpp:547-644
//       Current::Jxyz                       def-VlasovMax.cpp:656-745
//       Spatial_Advection::Axy              def-VlasovMax.cpp:757-951
//       MaxwellEq::Ampere/Faraday           def-VlasovMax.cpp:963-1035
//       CLEAN_LOW_PC                        def-Actions.cpp:240-253
//   * classic RK4 Runge_Kutta_4::advance_p1 def-RK.cpp:49-62,110-145
//   * e-e f00 collisions (Bell's conservative explicit operator, RK4 sub-cycled)
//       EE_isotropic_conserva               def-FokkerPlanck.cpp:107-267
//       RungeKutta4_ee / Explicit_EC_FP     def-FokkerPlanck.cpp:484-625
//   * implicit l>=1 collisions (tridiagonal)
//       Implicit_FLM_step                   def-FokkerPlanck.cpp:757-1035
//       Anisotropic_Collisions/Euler_Backward def-FokkerPlanck.cpp:1049-1304
//       TridiagonalSolve/Thomas_Tridiagonal def-Nmethods.cpp:142-207
//   * per-step ordering                     main.cpp:79-107
//   * a tangent-linear model of that step for 1D runs with m0 = 0
//     (Step1D<R> below, R = double or a forward-mode dual number): the same
//     formulas written on real scalars, used to check the GPU adjoint
//
// Differences from the reference that are deliberate:
//   * periodic indexing on the interior instead of guard cells (the reference
//     refreshes guards only after the whole step; with a stencil of width one
//     per RK stage and four guard cells the interior is identical);
//   * the field multipliers of Exyz/Bxyz/Axy are formed directly from the
//     coefficient tables instead of as running products (def-VlasovMax.cpp
//     :359-484); the rounding differs by a few ulp;
//   * every (l,m) with l<=l0, m<=min(l,m0) is supported, including m0<2 and
//     l0<3, which the reference's hard-wired blocks cannot run
//     (def-Input.cpp:73).  For m0>=2, l0>=3 the set of couplings is the
//     reference's.
//   * the y-advection coefficient slip at def-VlasovMax.cpp:918 (divide by
//     C1[m+1] instead of C1[l-1]) is reproduced when ref_quirks != 0.
//
// Layout exchanged through the C ABI (same as the reference minus guards):
//   f   : complex<double>[nh][ny][nx][np], p fastest, harmonic order ind(l,m)
//   fld : complex<double>[6][ny][nx]  (Ex,Ey,Ez,Bx,By,Bz)

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <limits>
#include <vector>

typedef std::complex<double> cd;

namespace {

struct Cfg {
  int l0, m0, np, nx, ny;
  double pmin, pmax, dx, dy;
  int ref_quirks;
  double density_np, zeta, small_dt;
  int nb;
  int coll_ee, coll_flm;
};

// Harmonic index of (l,m) in the trapezoid, def-State.cpp:417-427.
inline int hind(int l, int m, int m0) {
  return (l < m0 + 1) ? (l * (l + 1)) / 2 + m : l * (m0 + 1) - (m0 * (m0 + 1)) / 2 + m;
}

struct State {
  int nh, n, nxy;
  std::vector<cd> f;    // nh * n
  std::vector<cd> fld;  // 6 * nxy
  void init(int nh_, int n_, int nxy_) {
    nh = nh_; n = n_; nxy = nxy_;
    f.assign((size_t)nh * n, cd(0.0, 0.0));
    fld.assign((size_t)6 * nxy, cd(0.0, 0.0));
  }
  cd* h(int i) { return f.data() + (size_t)i * n; }
  const cd* h(int i) const { return f.data() + (size_t)i * n; }
  cd* F(int i) { return fld.data() + (size_t)i * nxy; }
  const cd* F(int i) const { return fld.data() + (size_t)i * nxy; }
};

struct Oracle {
  Cfg c;
  int nh, np, nx, ny, n, nxy;
  // p axes
  std::vector<double> pd;   // Axis<double>(np,pmin,pmax)   matrices.h:860-866
  double dpd;               // Axis<double>::dx             matrices.h:880
  std::vector<cd> pc;       // Axis<complex>(np,p0,pN-1)    def-VlasovMax.cpp:88-90
  cd dpc;
  // ElectricField tables (def-VlasovMax.cpp:97-151)
  std::vector<cd> eA1, eA2, eC2, eC4;  // (l0+1)*(m0+1)
  std::vector<cd> eB1, eB2, eC1, eC3, Hp0;
  std::vector<cd> invpr;
  // MagneticField tables (def-VlasovMax.cpp:563-589)
  std::vector<cd> bA1, bA2, bB1;
  // Current (def-VlasovMax.cpp:664-674)
  std::vector<cd> p3og;
  cd Delta_p, small;
  // Spatial_Advection (def-VlasovMax.cpp:768-826)
  std::vector<cd> vr;
  cd idx, idy;
  std::vector<cd> aA1, aA2, aC2, aC4, aB1, aB2, aC1, aC3;
  // Maxwell (def-VlasovMax.cpp:967-973)
  cd midx, midy;

  // Collision tables
  std::vector<double> vr_c;                              // v = p/gamma on the double axis
  std::vector<double> U4, U4m1, U2, U2m1, U1, U1m1, U3, Qn, Pn;
  double c_kpre, kpre;

  // time bookkeeping (def-RK.cpp:53, def-FokkerPlanck.cpp:572,1262)
  double t_rk, t_ee, t_eb;

  State Y;

  int l2i(int l, int m) const { return l * (c.m0 + 1) + m; }
  bool valid(int l, int m) const { return l >= 0 && l <= c.l0 && m >= 0 && m <= c.m0 && m <= l; }
  int H(int l, int m) const { return hind(l, m, c.m0); }
  size_t at(int ip, int ix, int iy) const { return ((size_t)iy * nx + ix) * np + ip; }
  int xm(int ix) const { return ix == 0 ? nx - 1 : ix - 1; }
  int xp(int ix) const { return ix == nx - 1 ? 0 : ix + 1; }
  int ym(int iy) const { return iy == 0 ? ny - 1 : iy - 1; }
  int yp(int iy) const { return iy == ny - 1 ? 0 : iy + 1; }

  explicit Oracle(const Cfg& cfg) : c(cfg) {
    np = c.np; nx = c.nx; ny = c.ny;
    nh = ((c.m0 + 1) * (2 * c.l0 - c.m0 + 2)) / 2;
    n = np * nx * ny; nxy = nx * ny;
    Y.init(nh, n, nxy);
    t_rk = t_ee = t_eb = 0.0;
    build_axes();
    build_vlasov_tables();
    build_collision_tables();
  }

  void build_axes() {
    // Axis(n, min, max): ax[i] = i; ax *= (max-min)/(n-1); ax += min   (matrices.h:860-866)
    pd.resize(np);
    double d = (c.pmax - c.pmin) / static_cast<double>(np - 1);
    for (int i = 0; i < np; ++i) { double v = i; v *= d; v += c.pmin; pd[i] = v; }
    dpd = pd[1] - pd[0];
    pc.resize(np);
    cd p0(pd[0]), pN(pd[np - 1]);
    cd dc = (pN - p0) / static_cast<cd>(np - 1);
    for (int i = 0; i < np; ++i) { cd v(i); v *= dc; v += p0; pc[i] = v; }
    dpc = pc[1] - pc[0];
  }

  void build_vlasov_tables() {
    const int L = c.l0, M = c.m0;
    const size_t LM = (size_t)(L + 1) * (M + 1);
    eA1.assign(LM, 0.0); eA2.assign(LM, 0.0); eC2.assign(LM, 0.0); eC4.assign(LM, 0.0);
    eB1.assign(L + 1, 0.0); eB2.assign(L + 1, 0.0); eC1.assign(L + 1, 0.0); eC3.assign(L + 1, 0.0);
    Hp0.assign(L + 1, 0.0);
    cd lc, mc;
    // def-VlasovMax.cpp:97-105
    for (int l = 1; l <= L; ++l)
      for (int m = 0; m <= std::min(M, l); ++m) {
        lc = cd(l); mc = cd(m);
        eA1[l2i(l, m)] = (lc + 1.0 - mc) * lc / ((2.0 * lc + 1.0) * (lc + 1.0));
        eA2[l2i(l, m)] = -(lc + mc) / (2.0 * lc + 1.0);
      }
    eA1[l2i(0, 0)] = -1.0;
    {
      cd s = 0.5 / dpc;
      for (size_t i = 0; i < LM; ++i) { eA1[i] *= s; eA2[i] *= s; }
    }
    // def-VlasovMax.cpp:112-117
    for (int l = 1; l <= L; ++l) {
      lc = cd(l);
      eB1[l] = (-0.5 / dpc) * lc * lc / (2.0 * lc + 1.0);
      eB2[l] = (-0.5 / dpc) * lc * (lc + 1.0) / (2.0 * lc + 1.0);
    }
    eB2[0] = 1.0;
    // def-VlasovMax.cpp:122-127
    for (int l = 1; l <= L; ++l) {
      lc = cd(l);
      eC1[l] = (0.25 / dpc) * lc / ((2.0 * lc + 1.0) * (lc + 1.0));
      eC3[l] = (0.25 / dpc) / (2.0 * lc + 1.0);
    }
    eC1[0] = -0.25 / dpc;
    // def-VlasovMax.cpp:132-139
    for (int l = 2; l <= L; ++l)
      for (int m = 2; m <= std::min(M, l); ++m) {
        lc = cd(l); mc = cd(m);
        eC2[l2i(l, m)] = (-0.25 / dpc) * lc * (lc - mc + 2.0) * (lc - mc + 1.0) / ((2.0 * lc + 1.0) * (lc + 1.0));
        eC4[l2i(l, m)] = (-0.25 / dpc) * (lc + mc - 1.0) * (lc + mc) / (2.0 * lc + 1.0);
      }
    // def-VlasovMax.cpp:145-147
    for (int l = 0; l <= L; ++l) Hp0[l] = 1.0 / pc[0];
    for (int l = 1; l <= L; ++l) Hp0[l] = Hp0[l - 1] * (pc[0] / pc[1]) * (2.0 * l + 1.0) / (2.0 * l - 1);
    for (int l = 0; l <= L; ++l) Hp0[l] *= -2.0 * dpc;
    invpr.resize(np);
    for (int i = 0; i < np; ++i) invpr[i] = 1.0 / pc[i];

    // MagneticField, def-VlasovMax.cpp:563-589
    bA1.assign(M + 1, 0.0); bA2.assign(LM, 0.0); bB1.assign(L + 1, 0.0);
    cd c01(0.0, 1.0);
    for (int m = 0; m <= M; ++m) { mc = cd(m); bA1[m] = (-1.0) * c01 * mc; }
    for (int l = 1; l <= L; ++l)
      for (int m = 0; m <= std::min(M, l); ++m) {
        lc = cd(l); mc = cd(m);
        bA2[l2i(l, m)] = (-0.5) * (lc + 1.0 - mc) * (lc + mc);
      }
    for (int l = 0; l <= L; ++l) { lc = cd(l); bB1[l] = (-1.0) * lc * (lc + 1.0); }

    // Current, def-VlasovMax.cpp:664-674
    p3og.resize(np);
    for (int i = 0; i < np; ++i) {
      cd p = pc[i];
      p3og[i] = (p * p) * (p / std::sqrt(1.0 + p * p));
    }
    Delta_p = cd(dpd);
    small = cd(pd[0]);
    small *= small; small *= small; small *= pd[0];
    small *= 0.2; small *= 1.0 / pd[1];

    // Spatial_Advection, def-VlasovMax.cpp:768-826
    vr.resize(np);
    for (int i = 0; i < np; ++i) vr[i] = pc[i] / (std::sqrt(1.0 + pc[i] * pc[i]));
    idx = cd(c.dx); idy = cd(c.dy);
    idx = -1.0 / (2.0 * idx);
    idy = -1.0 / (2.0 * idy);
    aA1.assign(LM, 0.0); aA2.assign(LM, 0.0); aC2.assign(LM, 0.0); aC4.assign(LM, 0.0);
    aB1.assign(L + 1, 0.0); aB2.assign(L + 1, 0.0); aC1.assign(L + 1, 0.0); aC3.assign(L + 1, 0.0);
    for (int l = 0; l <= L; ++l)
      for (int m = 0; m <= std::min(M, l); ++m) {
        lc = cd(l); mc = cd(m);
        aA1[l2i(l, m)] = idx * (-1.0) * (lc - mc + 1.0) / (2.0 * lc + 1.0);
        aA2[l2i(l, m)] = idx * (-1.0) * (lc + mc) / (2.0 * lc + 1.0);
      }
    for (int l = 0; l <= L; ++l) {
      lc = cd(l);
      aB1[l] = idy * (lc + 1.0) * lc / (2.0 * lc + 1.0);
      aB2[l] = (-1.0) * aB1[l];
    }
    for (int l = 0; l <= L; ++l) {
      lc = cd(l);
      aC1[l] = (-0.5) * idy / (2.0 * lc + 1.0);
      aC3[l] = (-1.0) * aC1[l];
    }
    for (int l = 0; l <= L; ++l)
      for (int m = 0; m <= std::min(M, l); ++m) {
        lc = cd(l); mc = cd(m);
        aC2[l2i(l, m)] = idy * 0.5 * (lc + 2.0 - mc) * (lc - mc + 1.0) / (2.0 * lc + 1.0);
        aC4[l2i(l, m)] = idy * (-0.5) * (lc + mc - 1.0) * (lc + mc) / (2.0 * lc + 1.0);
      }
    // MaxwellEq, def-VlasovMax.cpp:967-973
    midx = cd(c.dx); midy = cd(c.dy);
    midx = -1.0 / (2.0 * midx);
    midy = -1.0 / (2.0 * midy);
  }

  void build_collision_tables() {
    // def-FokkerPlanck.cpp:135-173 and :785-803 (identical tables)
    const double re = 2.8179402894e-13;
    double kp = std::sqrt(4.0 * M_PI * (c.density_np) * re);
    c_kpre = 4.0 * M_PI / 3.0 * re * kp;
    kpre = re * kp;
    vr_c.assign(np, 0.0);
    for (int i = 0; i < np; ++i) { vr_c[i] = pd[i]; vr_c[i] = vr_c[i] / (std::sqrt(1.0 + vr_c[i] * vr_c[i])); }
    U4.assign(np, 0.0); U4m1.assign(np, 0.0); U2.assign(np, 0.0); U2m1.assign(np, 0.0);
    U1.assign(np, 0.0); U1m1.assign(np, 0.0); U3.assign(np, 0.0); Qn.assign(np, 0.0); Pn.assign(np, 0.0);
    const std::vector<double>& v = vr_c;
    for (int i = 1; i < np; ++i) {
      U4[i] = 0.5 * std::pow(v[i], 4) * (v[i] - v[i - 1]);
      U4m1[i] = 0.5 * std::pow(v[i - 1], 4) * (v[i] - v[i - 1]);
      U2[i] = 0.5 * v[i] * v[i] * (v[i] - v[i - 1]);
      U2m1[i] = 0.5 * v[i - 1] * v[i - 1] * (v[i] - v[i - 1]);
      U1[i] = 0.5 * v[i] * (v[i] - v[i - 1]);
      U1m1[i] = 0.5 * v[i - 1] * (v[i] - v[i - 1]);
    }
    for (int i = 0; i < np; ++i) U3[i] = std::pow(v[i], 3);
    Qn[0] = 1.0 / ((v[0] * v[0] * v[1]) / 2.0);
    for (int i = 1; i < np - 1; ++i) Qn[i] = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]) / 2.0);
    for (int i = 0; i < np - 1; ++i) Pn[i] = 1.0 / ((v[i + 1] - v[i]) / 2.0 * (v[i + 1] + v[i]));
  }

  // ---------------------------------------------------------------- RHS --
  // SHarmonic::Dp along p for one (x,y) column, def-State.cpp:359-380 with
  // Matrix3D::Dd1 (matrices.h:804-810): interior f(p-1)-f(p+1); p0 -> 0;
  // pmax -> f(N-2)-f(N-1).
  void dp_col(const cd* fcol, cd* out) const {
    out[0] = 0.0;
    for (int ip = 1; ip < np - 1; ++ip) out[ip] = fcol[ip - 1] - fcol[ip + 1];
    out[np - 1] = fcol[np - 2] - fcol[np - 1];
  }

  // MakeGH (def-VlasovMax.cpp:496-515) for one column of harmonic l.
  void make_gh_col(int l, const cd* fcol, cd* G, cd* Hh) const {
    cd lc(l);
    cd s = -2.0 * (lc + 1.0) * dpc;
    cd cg = -(2.0 * lc + 1.0) / lc;
    dp_col(fcol, G);
    for (int ip = 0; ip < np; ++ip) {
      cd invpax = invpr[ip] * s;
      Hh[ip] = fcol[ip] * invpax;
      Hh[ip] += G[ip];
      G[ip] *= cg;
      G[ip] += Hh[ip];
    }
    G[0] = 0.0;
    Hh[0] = fcol[1] * Hp0[l];
  }

  // G00 (def-VlasovMax.cpp:519-535) for one column of f00.
  void g00_col(const cd* fcol, cd* G) const {
    cd p0p1_sq(pc[0] * pc[0] / (pc[1] * pc[1]));
    cd inv_mp0p1_sq(1.0 / (1.0 - p0p1_sq));
    cd g_r = -4.0 * dpc * pc[0] / (pc[1] * pc[1]);
    dp_col(fcol, G);
    cd f00 = (fcol[0] - fcol[1] * p0p1_sq) * inv_mp0p1_sq;
    G[0] = (fcol[1] - f00) * g_r;
  }

  // Coupling list of ElectricField::Exyz (def-VlasovMax.cpp:346-491),
  // one entry per (source op -> target).  field: 0=Ex, 1=Em, 2=Ep.
  struct ETerm { int tl, tm; int field; cd coef; bool re; };

  void e_targets_G(int l, int m, std::vector<ETerm>& out) const {
    // G of source (l,m) feeds l+1 (G00 for l=0)
    if (l + 1 > c.l0) return;
    if (l == 0) {
      out.push_back({1, 0, 0, eA1[l2i(0, 0)], false});
      if (c.m0 >= 1) out.push_back({1, 1, 1, eC1[0], false});
      return;
    }
    out.push_back({l + 1, m, 0, eA1[l2i(l, m)], false});
    if (m + 1 <= c.m0) out.push_back({l + 1, m + 1, 1, eC1[l], false});
    if (m == 1) out.push_back({l + 1, 0, 2, eB1[l], true});
    else if (m >= 2) out.push_back({l + 1, m - 1, 2, eC2[l2i(l, m)], false});
  }
  void e_targets_H(int l, int m, std::vector<ETerm>& out) const {
    // H of source (l,m), l>=1, feeds l-1
    if (valid(l - 1, m)) out.push_back({l - 1, m, 0, eA2[l2i(l, m)], false});
    if (m + 1 <= c.m0 && valid(l - 1, m + 1)) out.push_back({l - 1, m + 1, 1, eC3[l], false});
    if (m == 1) out.push_back({l - 1, 0, 2, eB2[l], true});
    else if (m >= 2) out.push_back({l - 1, m - 1, 2, eC4[l2i(l, m)], false});
  }

  // probe = 0: ElectricField::Exyz.  probe = 1..3: the implicit field's
  // probe couplings, i.e. the terms of Exyz with sources l <= 3 and targets
  // l <= 2 --
  //   1  Implicit_Ex (def-VlasovMax.cpp:157-196): Ex only, sources m <= 1;
  //   2  Implicit_Eyz with Em = Ep = Ey (ImpEy, def-Actions.cpp:187-201);
  //   3  Implicit_Eyz with Em = -i Ez, Ep = i Ez (ImpEz, :207-223).
  // Implicit_Eyz (def-VlasovMax.cpp:231-291) is exactly the Em/Ep terms of
  // Exyz inside those bounds, the (3,3) -> (2,2) term only when m0 > 2.
  // Multipliers are running products as in the reference (:359-484): each of
  // the three field matrices Ex, Em, Ep is scaled by the first coefficient it
  // meets and then by the ratio of each coefficient to the one before it on
  // the same matrix (the "C4(l0,1) = B2[l0]" of :418 is that previous
  // coefficient), so the multipliers round as the reference's do.
  void exyz(const State& Yin, State& Yh, int probe = 0) const {
    std::vector<cd> G(np), Hh(np);
    const int lsrc = probe ? std::min(3, c.l0) : c.l0;
    const int ltgt = probe ? 2 : c.l0;
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        size_t ixy = (size_t)iy * nx + ix;
        cd Ex = Yin.F(0)[ixy];
        cd Ez = Yin.F(2)[ixy], Ey = Yin.F(1)[ixy];
        cd ii(0.0, 1.0);
        cd Em = Ez; Em *= (-1.0) * ii; Em += Ey;   // def-VlasovMax.cpp:352
        cd Ep = Ez; Ep *= ii; Ep += Ey;            // :353
        if (probe == 2) { Em = Ey; Ep = Ey; }
        if (probe == 3) { Em = Ez; Em *= (-1.0) * ii; Ep = Ez; Ep *= ii; }
        cd cur[3] = {Ex, Em, Ep}, prev[3];
        bool started[3] = {false, false, false};
        auto keep = [&](const ETerm& t, int m) {
          if (!probe) return true;
          if (t.tl > ltgt) return false;
          if (probe == 1) return t.field == 0 && m <= 1;
          return t.field != 0;
        };
        auto apply = [&](const ETerm& t, const std::vector<cd>& V) {
          const int s = t.field;
          if (!started[s]) cur[s] *= t.coef;
          else cur[s] *= t.coef / prev[s];
          started[s] = true;
          prev[s] = t.coef;
          cd* dst = Yh.h(H(t.tl, t.tm)) + at(0, ix, iy);
          for (int ip = 0; ip < np; ++ip) {
            cd v = V[ip] * cur[s];
            dst[ip] += t.re ? cd(v.real(), 0.0) : v;
          }
        };
        for (int m = 0; m <= c.m0; ++m)
          for (int l = m; l <= lsrc; ++l) {
            const cd* fcol = Yin.h(H(l, m)) + at(0, ix, iy);
            std::vector<ETerm> tg, th;
            if (l == 0) g00_col(fcol, G.data());
            else make_gh_col(l, fcol, G.data(), Hh.data());
            if (l >= 1) e_targets_H(l, m, th);
            e_targets_G(l, m, tg);
            for (const ETerm& t : th)
              if (keep(t, m)) apply(t, Hh);
            for (const ETerm& t : tg)
              if (keep(t, m)) apply(t, G);
          }
      }
  }

  // Current::Jxyz, def-VlasovMax.cpp:680-745
  void jxyz(const State& Yin, State& Yh) const {
    cd c01(0.0, 1.0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        size_t ixy = (size_t)iy * nx + ix;
        {
          const cd* f = Yin.h(H(1, 0)) + at(0, ix, iy);
          cd j = f[0] * p3og[0];
          j += f[np - 1] * p3og[np - 1];
          j *= 0.5;
          for (int ip = 1; ip < np - 1; ++ip) j += f[ip] * p3og[ip];
          j *= Delta_p;
          j += small * (f[1] * p3og[1]);
          j *= 4.0 * M_PI / 3.0;
          Yh.F(0)[ixy] += j;
        }
        if (c.m0 >= 1) {
          const cd* f = Yin.h(H(1, 1)) + at(0, ix, iy);
          cd j = f[0] * p3og[0];
          j += f[np - 1] * p3og[np - 1];
          j *= 0.5;
          for (int ip = 1; ip < np - 1; ++ip) j += f[ip] * p3og[ip];
          j *= Delta_p;
          j += small * (f[1] * p3og[1]);
          j *= 8.0 * M_PI / 3.0;
          cd jy = cd(j.real(), 0.0);
          cd jz = j; jz -= jy; jz *= c01;
          Yh.F(1)[ixy] += jy;
          Yh.F(2)[ixy] += jz;
        }
      }
  }

  // minus central differences of a field that was scaled first
  // (tmpEB *= s; tmpEB.Dx(): Matrix2D::Dd1/Dd2, matrices.h:511-531)
  cd fdx(const cd* g, cd s, int ix, int iy) const {
    return g[(size_t)iy * nx + xm(ix)] * s - g[(size_t)iy * nx + xp(ix)] * s;
  }
  cd fdy(const cd* g, cd s, int ix, int iy) const {
    return g[(size_t)ym(iy) * nx + ix] * s - g[(size_t)yp(iy) * nx + ix] * s;
  }

  // MaxwellEq::Ampere, def-VlasovMax.cpp:1009-1035
  void ampere(const State& Yin, State& Yh) const {
    const cd *Bx = Yin.F(3), *By = Yin.F(4), *Bz = Yin.F(5);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        size_t k = (size_t)iy * nx + ix;
        Yh.F(0)[k] += fdy(Bz, midy, ix, iy);
        Yh.F(1)[k] += fdx(Bz, (-1.0) * midx, ix, iy);
        Yh.F(2)[k] += fdy(Bx, (-1.0) * midy, ix, iy);
        Yh.F(2)[k] += fdx(By, midx, ix, iy);
      }
  }
  // MaxwellEq::Faraday, def-VlasovMax.cpp:979-1005
  void faraday(const State& Yin, State& Yh) const {
    const cd *Ex = Yin.F(0), *Ey = Yin.F(1), *Ez = Yin.F(2);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        size_t k = (size_t)iy * nx + ix;
        Yh.F(3)[k] += fdy(Ez, (-1.0) * midy, ix, iy);
        Yh.F(4)[k] += fdx(Ez, midx, ix, iy);
        Yh.F(5)[k] += fdy(Ex, midy, ix, iy);
        Yh.F(5)[k] += fdx(Ey, (-1.0) * midx, ix, iy);
      }
  }

  // Spatial_Advection::Axy, def-VlasovMax.cpp:834-951.
  struct ATerm { int tl, tm; cd coef; bool re; };
  void axy(const State& Yin, State& Yh) const {
    // x advection: (l,m) -> (l-1,m) with A2, (l+1,m) with A1    (:842-853)
    for (int m = 0; m <= c.m0; ++m)
      for (int l = m; l <= c.l0; ++l) {
        std::vector<ATerm> t;
        if (valid(l - 1, m)) t.push_back({l - 1, m, aA2[l2i(l, m)], false});
        if (l + 1 <= c.l0) t.push_back({l + 1, m, aA1[l2i(l, m)], false});
        if (t.empty()) continue;
        const cd* src = Yin.h(H(l, m));
        for (const ATerm& a : t) {
          cd* dst = Yh.h(H(a.tl, a.tm));
          for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
              const cd* fl = src + at(0, xm(ix), iy);
              const cd* fr = src + at(0, xp(ix), iy);
              cd* d = dst + at(0, ix, iy);
              for (int ip = 0; ip < np; ++ip) d[ip] += (fl[ip] - fr[ip]) * (vr[ip] * a.coef);
            }
        }
      }
    if (c.m0 < 1) return;  // y advection only couples m to m+-1
    // y advection (:857-947)
    cd errf(1.0);  // cumulative factor of the :918 coefficient slip
    for (int m = 0; m <= c.m0; ++m) {
      cd err_loop(1.0);
      for (int l = m; l <= c.l0; ++l) {
        std::vector<ATerm> t;
        cd q = errf;
        if (c.ref_quirks && m >= 2 && m < c.m0 && m + 1 < c.l0) {
          if (l >= m + 3 && l < c.l0) err_loop *= aC1[l - 1] / aC1[m + 1];
          if (l >= m + 3) q *= err_loop;
        }
        if (m + 1 <= c.m0 && valid(l - 1, m + 1)) t.push_back({l - 1, m + 1, aC3[l] * q, false});
        if (m + 1 <= c.m0 && l + 1 <= c.l0) t.push_back({l + 1, m + 1, aC1[l] * q, false});
        if (m == 1) {
          t.push_back({l - 1, 0, aB2[l] * q, true});
          if (l + 1 <= c.l0) t.push_back({l + 1, 0, aB1[l] * q, true});
        } else if (m >= 2) {
          t.push_back({l - 1, m - 1, aC4[l2i(l, m)] * q, false});
          if (l + 1 <= c.l0) t.push_back({l + 1, m - 1, aC2[l2i(l, m)] * q, false});
        }
        const cd* src = Yin.h(H(l, m));
        for (const ATerm& a : t) {
          if (!valid(a.tl, a.tm)) continue;
          cd* dst = Yh.h(H(a.tl, a.tm));
          for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
              const cd* fl = src + at(0, ix, ym(iy));
              const cd* fr = src + at(0, ix, yp(iy));
              cd* d = dst + at(0, ix, iy);
              for (int ip = 0; ip < np; ++ip) {
                cd v = (fl[ip] - fr[ip]) * (vr[ip] * a.coef);
                d[ip] += a.re ? cd(v.real(), 0.0) : v;
              }
            }
        }
      }
      if (c.ref_quirks && m >= 2 && m < c.m0 && m + 1 < c.l0) errf *= err_loop;
    }
  }

  // MagneticField::Bxyz, def-VlasovMax.cpp:597-644
  void bxyz(const State& Yin, State& Yh) const {
    if (c.m0 < 1) return;  // every coupling touches m>=1
    cd ii(0.0, 1.0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        size_t k = (size_t)iy * nx + ix;
        cd Bx = Yin.F(3)[k], By = Yin.F(4)[k], Bz = Yin.F(5)[k];
        cd Bm = By; Bm *= (-1.0) * ii; Bm += Bz;
        cd Bp = By; Bp *= ii; Bp += Bz;
        Bp *= 0.5;
        for (int m = 0; m <= c.m0; ++m)
          for (int l = std::max(m, 1); l <= c.l0; ++l) {
            const cd* f = Yin.h(H(l, m)) + at(0, ix, iy);
            if (m + 1 <= c.m0 && m + 1 <= l) {
              cd* d = Yh.h(H(l, m + 1)) + at(0, ix, iy);
              for (int ip = 0; ip < np; ++ip) d[ip] += f[ip] * Bp;
            }
            if (m >= 1) {
              cd mx = Bx * bA1[m];
              cd* d = Yh.h(H(l, m)) + at(0, ix, iy);
              for (int ip = 0; ip < np; ++ip) d[ip] += f[ip] * mx;
              if (m == 1) {
                cd mm = Bm * bB1[l];
                cd* d0 = Yh.h(H(l, 0)) + at(0, ix, iy);
                for (int ip = 0; ip < np; ++ip) d0[ip] += cd((f[ip] * mm).real(), 0.0);
              } else {
                cd mm = Bm * bA2[l2i(l, m)];
                cd* d1 = Yh.h(H(l, m - 1)) + at(0, ix, iy);
                for (int ip = 0; ip < np; ++ip) d1[ip] += f[ip] * mm;
              }
            }
          }
      }
  }

  // CLEAN_LOW_PC, def-Actions.cpp:240-253
  void clean_low_pc(State& Yh) const {
    for (int l = 1; l <= c.l0; ++l)
      for (int m = 0; m <= std::min(c.m0, l); ++m) {
        int lim = std::min(l - 1, 80);
        cd* d = Yh.h(H(l, m));
        for (int iy = 0; iy < ny; ++iy)
          for (int ix = 0; ix < nx; ++ix)
            for (int ip = 0; ip < lim; ++ip) d[at(ip, ix, iy)] = 0.0;
      }
  }

  // Actions::Adv, def-Actions.cpp:73-94
  void adv(const State& Yin, State& Yh) const {
    std::fill(Yh.f.begin(), Yh.f.end(), cd(0.0, 0.0));
    std::fill(Yh.fld.begin(), Yh.fld.end(), cd(0.0, 0.0));
    exyz(Yin, Yh);
    jxyz(Yin, Yh);
    ampere(Yin, Yh);
    axy(Yin, Yh);
    bxyz(Yin, Yh);
    faraday(Yin, Yh);
    clean_low_pc(Yh);
  }

  // ----------------------------------------------------------- RK4 ------
  static void axpy_state(State& y, const State& x) {  // y += x
    for (size_t i = 0; i < y.f.size(); ++i) y.f[i] += x.f[i];
    for (size_t i = 0; i < y.fld.size(); ++i) y.fld[i] += x.fld[i];
  }
  static void scale_state(State& y, double s) {
    cd cs(s);
    for (auto& v : y.f) v *= cs;
    for (auto& v : y.fld) v *= cs;
  }

  // Runge_Kutta_4::advance_p1, def-RK.cpp:110-145
  void rk4(double h) {
    State Y0 = Y, Y1 = Y, Yh;
    Yh.init(nh, n, nxy);
    adv(Y1, Yh);
    scale_state(Yh, 0.5 * h); axpy_state(Y1, Yh);
    scale_state(Yh, 1.0 / 3.0); axpy_state(Y, Yh);

    adv(Y1, Yh); Y1 = Y0;
    scale_state(Yh, 0.5 * h); axpy_state(Y1, Yh);
    scale_state(Yh, 2.0 / 3.0); axpy_state(Y, Yh);

    adv(Y1, Yh);
    scale_state(Yh, h); axpy_state(Y0, Yh);
    scale_state(Yh, 1.0 / 3.0); axpy_state(Y, Yh);

    adv(Y0, Yh);
    scale_state(Yh, h / 6.0); axpy_state(Y, Yh);
    t_rk += h;
  }

  // ------------------------------------------------------ collisions ----
  // LOGee / ZLOGei (NRL), def-FokkerPlanck.cpp:41-95
  double LOGee(double ne, double Te) const {
    if (ne > 0.000000001) {
      Te /= (3.0 * ne);
      Te *= 511000;
      ne *= c.density_np;
      Te = std::log(Te);
      ne = std::log(ne);
      double lnee = 23.5 - 0.5 * ne + 1.25 * Te - std::sqrt(0.00001 + 0.0625 * (Te - 2.0) * (Te - 2.0));
      if (lnee > 2.0) return lnee;
    }
    return 2.0;
  }
  double ZLOGei(double ne, double Te) const {
    if (ne > 0.0000001) {
      Te /= (3.0 * ne);
      Te *= 511000;
      ne *= c.density_np;
      Te = std::log(Te);
      ne = std::log(ne);
      double lnei = 24.0 - 0.5 * ne + Te;
      if (lnei > 2.0) return lnei * c.zeta;
    }
    return 2.0 * c.zeta;
  }

  // EE_isotropic_conserva::G, def-FokkerPlanck.cpp:178-188
  double Gpar(int n_, const double* fin, const double* J1) const {
    const std::vector<double>& v = vr_c;
    double f00 = (fin[0] - fin[1] * (v[0] * v[0]) / (v[1] * v[1])) / (1.0 - (v[0] * v[0]) / (v[1] * v[1]));
    double i2s = f00 * std::pow(v[n_], 3) / 3.0 + (fin[1] - f00) * std::pow(v[n_], 5) / (v[1] * v[1]) * 0.2;
    double i4s = f00 * std::pow(v[n_], 5) * 0.2 + (fin[1] - f00) * std::pow(v[n_], 7) / (v[1] * v[1] * 7.0);
    return fin[n_] * i4s + (std::pow(v[n_], 3) * fin[n_] - 3.0 * i2s) * J1[n_];
  }

  // EE_isotropic_conserva::Getslope, def-FokkerPlanck.cpp:192-267
  void getslope(const double* fin, double* fh) const {
    const int N = np;
    std::vector<double> I4(N), I2(N), J1(N);
    I4[0] = 0;
    for (int k = 1; k < N; ++k) { I4[k] = U4[k] * fin[k] + U4m1[k] * fin[k - 1]; I4[k] += I4[k - 1]; }
    I2[0] = 0;
    for (int k = 1; k < N; ++k) { I2[k] = U2[k] * fin[k] + U2m1[k] * fin[k - 1]; I2[k] += I2[k - 1]; }
    J1[N - 1] = 0;
    for (int k = N - 2; k > -1; --k) { J1[k] = U1[k + 1] * fin[k + 1] + U1m1[k + 1] * fin[k]; J1[k] += J1[k + 1]; }
    double Ln_ee = LOGee(4.0 * M_PI * I2[N - 1], 4.0 * M_PI * I4[N - 1]);
    for (int k = 0; k < N; ++k) {
      I2[k] *= J1[k];
      I2[k] *= -3.0;
      I4[k] += U3[k] * J1[k];
      I4[k] *= fin[k];
      I4[k] += I2[k];
    }
    for (int k = 0; k < c.nb && k < N; ++k) I4[k] = Gpar(k, fin, J1.data());
    fh[0] = (-1.0) * I4[0];
    for (int k = 0; k < N - 1; ++k) I4[k] -= I4[k + 1];
    fh[0] *= 2.0 / (vr_c[0] * vr_c[0]);
    for (int k = 0; k < N - 1; ++k) I4[k] *= Pn[k];
    fh[0] -= I4[0];
    for (int k = 0; k < N - 1; ++k) I4[k] -= I4[k + 1];
    fh[0] *= Qn[0];
    for (int k = 0; k < N - 1; ++k) fh[k + 1] = Qn[k + 1] * I4[k];
    for (int k = 0; k < N; ++k) fh[k] *= c_kpre * Ln_ee;
  }

  // RungeKutta4_ee::advance, def-FokkerPlanck.cpp:518-553
  void rk4_ee(double* f, double h) const {
    const int N = np;
    std::vector<double> f0(f, f + N), f1(f, f + N), fh(N);
    getslope(f1.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= (0.5 * h); f1[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (1.0 / 3.0); f[k] += fh[k]; }
    getslope(f1.data(), fh.data());
    f1 = f0;
    for (int k = 0; k < N; ++k) { fh[k] *= (0.5 * h); f1[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (2.0 / 3.0); f[k] += fh[k]; }
    getslope(f1.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= h; f0[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (1.0 / 3.0); f[k] += fh[k]; }
    getslope(f0.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= (h / 6.0); f[k] += fh[k]; }
  }

  // Explicit_EC_FP::loop, def-FokkerPlanck.cpp:590-625
  void collide_ee(double tnew) {
    int numh = static_cast<int>(static_cast<size_t>(static_cast<int>((tnew - t_ee) / c.small_dt)) + 1);
    double h = (tnew - t_ee) / static_cast<double>(static_cast<size_t>(numh));
    std::vector<double> fc(np);
    cd* f00 = Y.h(0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        cd* col = f00 + at(0, ix, iy);
        for (int ip = 0; ip < np; ++ip) fc[ip] = col[ip].real();
        for (int s = 0; s < numh; ++s) rk4_ee(fc.data(), h);
        for (int ip = 0; ip < np; ++ip) col[ip] = fc[ip];
      }
    t_ee = tnew;
  }

  // --------------------------------------------------------- sources ----
  // def-PLSource.cpp: PLaserSource::Laser_Source (:138-252), the
  // InverseBremsstrahlung operator (:476-777, :780-861) and
  // ExternalEfield::Laser_Efield (:368-418).  main.cpp:90-92 calls
  // Laser_Source and InverseBremsstrahlung::loop after the e-e collisions;
  // Laser_Efield is commented out there (:91) and is a separate call here.
  // Profiles and rates are built as the constructors build them
  // (:73-133, :308-353, :861-946) from the cell positions the caller passes
  // (the reference's x and y axes at the interior cells and its global x axis).
  struct Src {
    int linear = 0, poly = 0, ib = 0, pol = 0;
    double I_0 = 0, lambda_0 = 1, ab_eta = 0, ab_depth = 1, cntr_x0 = 0, cntr_r0 = 0, spot_w0 = 1;
    double rise = 0, flat = 0, fall = 0, smaller_dt = 1;
    std::vector<double> ILx, ILy, Ty, ILxe, ILye, xpos;  // PLS (normalised x), IB/EFLD profiles, x of cells
    std::vector<double> gam;                             // gamma(p)
    double nmin = 1e-5, IL = 0, IL_sqrt = 0, IL_min = 0, T0 = 0, acoeff = 0;
    double t_pls = 0, t_ib = 0, t_efld = 0;
    double vos = 0, w0overwp = 0;
    // IB_Operator tables (:476-545)
    std::vector<double> vr, U4, U4m1, U2, U2m1, Inv_Uav6, Qn, Pn, U2p, U2pm1;
    double Inv_Uav6_nm1 = 0, Pnm1 = 0, vw_coeff_cube = 0, Qn_coeff = 0;
    bool alpha_error = false;
  } src;

  void set_sources(const double* sv, const double* x, const double* y, const double* xglob, int nxg, double t0) {
    Src& S = src;
    S.linear = static_cast<int>(sv[0]); S.poly = static_cast<int>(sv[1]); S.ib = static_cast<int>(sv[2]);
    S.pol = static_cast<int>(sv[3]);
    S.I_0 = sv[4]; S.lambda_0 = sv[5]; S.ab_eta = sv[6]; S.ab_depth = sv[7]; S.cntr_x0 = sv[8]; S.cntr_r0 = sv[9];
    S.spot_w0 = sv[10]; S.rise = sv[11]; S.flat = sv[12]; S.fall = sv[13]; S.smaller_dt = sv[14];
    const double xgdx = xglob[1] - xglob[0];
    // PLaserSource (:73-133)
    S.IL = S.I_0 * S.ab_eta;
    S.IL_sqrt = std::sqrt(S.IL);
    S.IL_min = 0.00001 * S.IL;
    S.T0 = 1000.0 * std::sqrt(1.0e-14 * S.lambda_0 * S.lambda_0);
    S.acoeff = 2.08050325 * 1.0e+12 / (xgdx * c.density_np * S.lambda_0);
    double sumI = 0.0;
    for (int i = 0; i < nxg; ++i) sumI += std::exp((-1.0) * std::pow((xglob[i] - S.cntr_x0) / S.ab_depth, 2));
    S.ILx.resize(nx); S.ILxe.resize(nx); S.xpos.assign(x, x + nx);
    for (int i = 0; i < nx; ++i) {
      S.ILx[i] = std::exp((-1.0) * std::pow((x[i] - S.cntr_x0) / S.ab_depth, 2));
      S.ILxe[i] = S.ILx[i];  // IB and the external field keep the profile unnormalised (:336-338, :891-893)
      S.ILx[i] *= 1.0 / sumI;
    }
    S.ILy.resize(ny); S.ILye.resize(ny); S.Ty.resize(ny);
    for (int j = 0; j < ny; ++j) {
      const double tmp2 = (y[j] - S.cntr_r0) / S.spot_w0;
      S.ILy[j] = std::exp(-2.0 * tmp2 * tmp2);
      S.ILye[j] = S.ILy[j];
      S.Ty[j] = S.T0 * std::sqrt(S.ILy[j] * S.IL);
    }
    S.gam.resize(np);
    for (int i = 0; i < np; ++i) S.gam[i] = std::sqrt(1.0 + pd[i] * pd[i]);
    // ExternalEfield / InverseBremsstrahlung constants (:308-318, :861-875)
    const double omega_0 = 3.0e+10 * 2.0 * M_PI / (1.0e-4 * S.lambda_0);
    const double omega_p = 5.64 * 1.0e+4 * std::sqrt(c.density_np);
    S.w0overwp = omega_0 / omega_p;
    S.vos = S.lambda_0 * std::sqrt(7.3e-19 * S.I_0);
    if (std::abs(S.pol) > 3) S.pol = 0;
    if (std::abs(S.ib) > 1) S.ib = 0;
    // IB_Operator (:476-545)
    const double re = 2.8179402894e-13;
    const double kp = std::sqrt(4.0 * M_PI * (c.density_np) * re);
    S.vw_coeff_cube = omega_p / omega_0 * kp * re;
    S.Qn_coeff = kp * re / 6.0;
    const int N = np;
    S.vr.assign(N, 0.0); S.U4.assign(N, 0.0); S.U4m1.assign(N, 0.0); S.U2.assign(N, 0.0); S.U2m1.assign(N, 0.0);
    S.Inv_Uav6.assign(N, 0.0); S.Qn.assign(N, 0.0); S.Pn.assign(N, 0.0); S.U2p.assign(N, 0.0);
    S.U2pm1.assign(N, 0.0);
    for (int i = 0; i < N; ++i) { S.vr[i] = pd[i]; S.vr[i] = S.vr[i] / (std::sqrt(1.0 + S.vr[i] * S.vr[i])); }
    const std::vector<double>& v = S.vr;
    for (int i = 1; i < N; ++i) {
      S.U4[i] = 0.5 * std::pow(v[i], 4) * (v[i] - v[i - 1]);
      S.U4m1[i] = 0.5 * std::pow(v[i - 1], 4) * (v[i] - v[i - 1]);
      S.U2[i] = 0.5 * v[i] * v[i] * (v[i] - v[i - 1]);
      S.U2m1[i] = 0.5 * v[i - 1] * v[i - 1] * (v[i] - v[i - 1]);
    }
    S.Inv_Uav6_nm1 = std::pow(2.0 / v[0], 6);
    for (int i = 0; i < N - 1; ++i) S.Inv_Uav6[i] = std::pow(2.0 / (v[i + 1] + v[i]), 6);  // [N-1] is never used
    S.Qn[0] = 1.0 / ((v[0] * v[0] * v[1]) / 2.0);
    for (int i = 1; i < N - 1; ++i) S.Qn[i] = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]) / 2.0);
    S.Pnm1 = 2.0 / (v[0] * v[0]);
    for (int i = 0; i < N - 1; ++i) S.Pn[i] = 1.0 / ((v[i + 1] - v[i]) / 2.0 * (v[i + 1] + v[i]));
    // density weights of the IB renormalisation, on p (:902-905)
    for (int i = 1; i < N; ++i) {
      S.U2p[i] = 0.5 * pd[i] * pd[i] * (pd[i] - pd[i - 1]);
      S.U2pm1[i] = 0.5 * pd[i - 1] * pd[i - 1] * (pd[i] - pd[i - 1]);
    }
    S.t_pls = S.t_ib = S.t_efld = t0;
  }

  // the intensity envelope times (t2 - t1) of PLaserSource / ExternalEfield (:259-289, :428-458)
  double t_profile(double t2, double t1) const {
    const Src& S = src;
    double t;
    if (t2 < S.rise) {
      t = t2 / S.rise;
      if (S.poly) return (10 * std::pow(t, 3) - 15 * std::pow(t, 4) + 6 * std::pow(t, 5)) * (t2 - t1);
      if (S.linear) return t * (t2 - t1);
      return 0;
    }
    if (t2 < S.rise + S.flat) return t2 - t1;
    if (t2 < S.rise + S.flat + S.fall) {
      t = (S.rise + S.flat + S.fall - t2) / S.fall;
      if (S.poly) return (10 * std::pow(t, 3) - 15 * std::pow(t, 4) + 6 * std::pow(t, 5)) * (t2 - t1);
      if (S.linear) return t * (t2 - t1);
      return 0;
    }
    return -1.0;
  }

  // Make_Gaussian_f (:1057-1124): unit-density Maxwellian of temperature Thot [eV] in f; Thot
  // becomes the temperature of the discretised profile (-1 when it does not resolve p1)
  void make_gaussian(std::vector<double>& f, double& Thot) const {
    const std::vector<double>& p = pd;
    const std::vector<double>& g = src.gam;
    const int N = np;
    const double PTH = 4.19 / 3000.0 * std::sqrt(Thot);
    if (!(PTH > p[1])) { Thot = -1.0; return; }
    f = p;
    for (int i = 0; i < N; ++i) f[i] *= p[i];
    for (int i = 0; i < N; ++i) f[i] *= -0.5 / (PTH * PTH);
    for (int i = 0; i < N; ++i) f[i] = std::exp(f[i]);
    for (int i = 0; i < N; ++i) f[i] *= 1.0 / std::pow(std::sqrt(2 * M_PI) * PTH, 3);
    double n0 = f[0] * p[0] * p[0];
    n0 += f[N - 1] * p[N - 1] * p[N - 1];
    n0 *= 0.5;
    for (int i = 1; i < N - 1; ++i) n0 += f[i] * p[i] * p[i];
    const double dp = p[1] - p[0];
    const double p0p1_sq = p[0] * p[0] / (p[1] * p[1]), inv_mp0p1_sq = 1.0 / (1.0 - p0p1_sq);
    double f00 = f[1] / 5.0 * p0p1_sq;
    f00 += (f[0] - f[1] * p0p1_sq) * inv_mp0p1_sq * (1.0 / 3.0 - 1.0 / 5.0 * p0p1_sq);
    f00 *= p[0] * (p[0] / dp) * p[0];
    n0 += f00;
    n0 *= 4.0 * M_PI * dp;
    for (int i = 0; i < N; ++i) f[i] *= 1.0 / n0;
    double vtsq = f[0] * std::pow(p[0], 4) / std::pow(g[0], 2);
    vtsq += f[N - 1] * std::pow(p[N - 1], 4) / std::pow(g[N - 1], 2);
    vtsq *= 0.5;
    for (int i = 1; i < N - 1; ++i) vtsq += f[i] * std::pow(p[i], 4) / std::pow(g[i], 2);
    f00 = f[1] / 7.0 * p0p1_sq;
    f00 += (f[0] - f[1] * p0p1_sq) * inv_mp0p1_sq * (1.0 / 5.0 - 1.0 / 7.0 * p0p1_sq);
    f00 *= std::pow(p[0], 4) * (p[0] / dp);
    vtsq += f00;
    vtsq *= 4.0 * M_PI / 3.0 * dp;
    Thot = vtsq * (3000.0 / 4.19) * (3000.0 / 4.19);
  }

  // PLaserSource::Laser_Source (:138-252) on the interior cells (the
  // reference also heats guard cells, which its exchange then overwrites)
  void laser_source(double tnew) {
    Src& S = src;
    const double I0_t = t_profile(tnew, S.t_pls);
    const double dt = tnew - S.t_pls;
    const double ILt = I0_t / dt;
    S.t_pls = tnew;
    if (I0_t < 0) return;
    const std::vector<double>& p = pd;
    const std::vector<double>& g = S.gam;
    const int N = np;
    const double dp = p[1] - p[0];
    const double p0p1_sq = p[0] * p[0] / (p[1] * p[1]), inv_mp0p1_sq = 1.0 / (1.0 - p0p1_sq);
    std::vector<double> fp4(N);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        cd* col = Y.h(0) + at(0, ix, iy);
        double n0 = col[0].real() * p[0] * p[0];
        n0 += col[N - 1].real() * p[N - 1] * p[N - 1];
        n0 *= 0.5;
        for (int i = 1; i < N - 1; ++i) n0 += p[i] * p[i] * col[i].real();
        double f00 = col[1].real() / 5.0 * p0p1_sq;
        f00 += (col[0].real() - col[1].real() * p0p1_sq) * inv_mp0p1_sq * (1.0 / 3.0 - 1.0 / 5.0 * p0p1_sq);
        f00 *= p[0] * (p[0] / dp) * p[0];
        n0 += f00;
        n0 *= 4.0 * M_PI * dp;
        double vsq = col[0].real() * std::pow(p[0], 4) / std::pow(g[0], 2);
        vsq += col[N - 1].real() * std::pow(p[N - 1], 4) / std::pow(g[N - 1], 2);
        vsq *= 0.5;
        for (int i = 1; i < N - 1; ++i) vsq += std::pow(p[i], 4) / std::pow(g[i], 2) * col[i].real();
        f00 = col[1].real() / 7.0 * p0p1_sq;
        f00 += (col[0].real() - col[1].real() * p0p1_sq) * inv_mp0p1_sq * (1.0 / 5.0 - 1.0 / 7.0 * p0p1_sq);
        f00 *= std::pow(p[0], 4) * (p[0] / dp);
        vsq += f00;
        vsq *= 4.0 * M_PI / 3.0 * dp;
        vsq /= n0;
        vsq *= (3000.0 / 4.19) * (3000.0 / 4.19);  // temperature in eV
        double T_hot = S.Ty[iy] + vsq;
        make_gaussian(fp4, T_hot);
        if ((n0 > S.nmin) && (S.ILy[iy] * S.ILx[ix] * ILt * S.IL > S.IL_min) && (T_hot > vsq)) {
          double alpha = S.acoeff * S.IL_sqrt * dt / n0;
          alpha *= S.ILx[ix] * std::sqrt(S.ILy[iy] * ILt);
          if ((alpha < 0) || (alpha > 0.99)) { S.alpha_error = true; continue; }  // the reference exits (:230-233)
          for (int i = 0; i < N; ++i) {
            col[i] *= (1.0 - alpha);
            col[i] += alpha * fp4[i] * n0;
          }
        }
      }
  }

  // ZLOGei_v2 (:42-67)
  double zlogei_v2(double ne, double Te) const {
    if (ne > 0.0000001) {
      Te /= (3.0 * ne);
      Te *= 511000;
      ne *= c.density_np;
      Te = std::log(Te);
      ne = std::log(ne);
      const double lnei = 24.0 - 0.5 * ne + Te;
      if (lnei > 2.0) return lnei * c.zeta;
    }
    return 2.0 * c.zeta;
  }

  // IB_Operator::Getslope (:550-608); fh[N-1] is never written (stays 0)
  void ib_slope(const std::vector<double>& fin, double vos, std::vector<double>& fh) const {
    const Src& S = src;
    const int N = np;
    double p0overp1_sq = pd[0] / pd[1];
    p0overp1_sq *= p0overp1_sq;
    const double f00 = (fin[0] - fin[1] * p0overp1_sq) / (1.0 - p0overp1_sq);
    double I4 = 0;
    for (int n_ = 1; n_ < N; ++n_) I4 += S.U4[n_] * fin[n_] + S.U4m1[n_] * fin[n_ - 1];
    double I2 = 0;
    for (int n_ = 1; n_ < N; ++n_) I2 += S.U2[n_] * fin[n_] + S.U2m1[n_] * fin[n_ - 1];
    const double ZLn_ei = zlogei_v2(4.0 * M_PI * I2, 4.0 * M_PI * I4);
    const double vw_cube = S.vw_coeff_cube * ZLn_ei * 4.0 * M_PI * I2;
    std::vector<double> gn(N);
    for (int i = 0; i < N; ++i) {
      gn[i] = S.Inv_Uav6[i] * vw_cube * vw_cube;
      gn[i] = 1.0 / (1.0 + gn[i]);
    }
    const double qn_c = S.Qn_coeff * vos * vos * ZLn_ei * 4.0 * M_PI * I2;
    fh[0] = qn_c * S.Qn[0] *
            (S.Pn[0] * gn[0] * (fin[1] - fin[0]) -
             S.Pnm1 / (1.0 + S.Inv_Uav6_nm1 * vw_cube * vw_cube) * (fin[0] - f00));
    for (int i = 1; i < N - 1; ++i)
      fh[i] = qn_c * S.Qn[i] *
              (S.Pn[i] * gn[i] * (fin[i + 1] - fin[i]) - S.Pn[i - 1] * gn[i - 1] * (fin[i] - fin[i - 1]));
  }

  // InverseBremsstrahlung::loop (:780-861): RungeKutta4_ib sub-steps (:713-757) per cell, then the
  // density restored
  void inverse_bremsstrahlung(double tnew) {
    Src& S = src;
    const int N = np;
    const int numh = static_cast<int>(static_cast<size_t>(static_cast<int>((tnew - S.t_ib) / S.smaller_dt)) + 1);
    const double h = (tnew - S.t_ib) / static_cast<double>(static_cast<size_t>(numh));
    std::vector<double> fc(N), f0(N), f1(N), fh(N, 0.0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const double vos = S.vos * std::sqrt(S.ILxe[ix] * S.ILye[iy]);
        cd* col = Y.h(0) + at(0, ix, iy);
        for (int i = 0; i < N; ++i) fc[i] = col[i].real();
        double I2_before = 0.0;
        for (int n_ = 1; n_ < N; ++n_) I2_before += S.U2p[n_] * fc[n_] + S.U2pm1[n_] * fc[n_ - 1];
        for (int s = 0; s < numh; ++s) {
          f0 = fc; f1 = fc;
          ib_slope(f1, vos, fh);
          for (int i = 0; i < N; ++i) { fh[i] *= 0.5 * h; f1[i] += fh[i]; fh[i] *= 1.0 / 3.0; fc[i] += fh[i]; }
          ib_slope(f1, vos, fh);
          f1 = f0;
          for (int i = 0; i < N; ++i) { fh[i] *= 0.5 * h; f1[i] += fh[i]; fh[i] *= 2.0 / 3.0; fc[i] += fh[i]; }
          ib_slope(f1, vos, fh);
          for (int i = 0; i < N; ++i) { fh[i] *= h; f0[i] += fh[i]; fh[i] *= 1.0 / 3.0; fc[i] += fh[i]; }
          ib_slope(f0, vos, fh);
          for (int i = 0; i < N; ++i) { fh[i] *= h / 6.0; fc[i] += fh[i]; }
        }
        double I2_after = 0.0;
        for (int n_ = 1; n_ < N; ++n_) I2_after += S.U2p[n_] * fc[n_] + S.U2pm1[n_] * fc[n_ - 1];
        for (int i = 0; i < N; ++i) fc[i] *= I2_before / I2_after;
        for (int i = 0; i < N; ++i) col[i] = fc[i];
      }
    S.t_ib = tnew;
  }

  // ExternalEfield::Laser_Efield (:368-418): the chosen field component set on every cell
  void laser_efield(double tnew) {
    Src& S = src;
    const double dt = tnew - S.t_efld;
    const double env = t_profile(tnew, S.t_efld) / dt;
    if (S.pol >= 1 && S.pol <= 3) {
      cd* F = Y.fld.data() + (size_t)(S.pol - 1) * nxy;
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
          F[(size_t)j * nx + i] = S.vos * S.ILxe[i] * S.ILye[j] * env * std::sin(S.w0overwp * (tnew - S.xpos[i]));
    }
    S.t_efld = tnew;
  }

  // Implicit_FLM_step::reset_coeff, def-FokkerPlanck.cpp:808-939 (tridiagonal part)
  struct Coef {
    std::vector<double> sub, diag, sup, scatt;  // Alpha_Tri bands (already * -lnee kpre Dt), scattering term
  };
  void reset_coeff(const double* fin, double Dt, Coef& k) const {
    const int N = np;
    const std::vector<double>& v = vr_c;
    std::vector<double> I2(N), I0(N), J1m(N);
    I2[0] = 0;
    for (int q = 1; q < N; ++q) { I2[q] = U4[q] * fin[q] + U4m1[q] * fin[q - 1]; I2[q] += I2[q - 1]; }
    for (int q = 0; q < N; ++q) I2[q] *= 4.0 * M_PI;
    double I2_temperature = I2[N - 1];
    for (int q = 0; q < N; ++q) I2[q] /= v[q] * v[q];
    I0[0] = 0;
    for (int q = 1; q < N; ++q) { I0[q] = U2[q] * fin[q] + U2m1[q] * fin[q - 1]; I0[q] += I0[q - 1]; }
    for (int q = 0; q < N; ++q) I0[q] *= 4.0 * M_PI;
    double I0_density = I0[N - 1];
    J1m[N - 1] = 0;
    for (int q = N - 2; q > -1; --q) { J1m[q] = U1[q + 1] * fin[q + 1] + U1m1[q + 1] * fin[q]; J1m[q] += J1m[q + 1]; }
    for (int q = 0; q < N; ++q) J1m[q] *= 4.0 * M_PI * v[q];
    double lnee = LOGee(I0_density, I2_temperature);
    double zlnei = ZLOGei(I0_density, I2_temperature);
    std::vector<double> T1(N), T2(N);
    for (int i = 0; i < N; ++i) {
      T1[i] = I0[i] + (2.0 * J1m[i] - I2[i]) / 3.0;
      T2[i] = (I2[i] + J1m[i]) / 3.0;
    }
    k.scatt.assign(N, 0.0);
    for (int i = 0; i < N; ++i) k.scatt[i] = T1[i] * lnee;
    k.scatt[0] = 0.0;
    double zi = zlnei * I0_density;
    for (int i = 0; i < N; ++i) k.scatt[i] += zi;
    for (int i = 0; i < N; ++i) k.scatt[i] /= std::pow(v[i], 3);
    double kd = kpre * Dt;
    for (int i = 0; i < N; ++i) k.scatt[i] *= kd;
    k.sub.assign(N, 0.0); k.diag.assign(N, 0.0); k.sup.assign(N, 0.0);
    k.diag[0] = 8.0 * M_PI * fin[0];
    for (int i = 1; i < N - 1; ++i) {
      double IvDnDm1 = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i] - v[i - 1]));
      double IvDnDp1 = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i + 1] - v[i]));
      double Ivsq2Dn = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]));
      k.diag[i] = 8.0 * M_PI * fin[i] - T2[i] * (IvDnDm1 + IvDnDp1);
      k.sub[i] = T2[i] * IvDnDm1 - T1[i] * Ivsq2Dn;
      k.sup[i] = T2[i] * IvDnDp1 + T1[i] * Ivsq2Dn;
    }
    double s = (-1.0) * lnee * kpre * Dt;
    for (int i = 0; i < N; ++i) { k.sub[i] *= s; k.diag[i] *= s; k.sup[i] *= s; }
  }

  // Implicit_FLM_step::advance (tridiagonal) + TridiagonalSolve,
  // def-FokkerPlanck.cpp:943-1035, def-Nmethods.cpp:142-207
  void advance_flm(const Coef& k, cd* col, int el) const {
    const int N = np;
    std::vector<double> a(N, 0.0), b(N), cc(N, 0.0);
    for (int i = 1; i < N; ++i) a[i] = k.sub[i];
    for (int i = 0; i < N; ++i) b[i] = k.diag[i];
    for (int i = 0; i < N - 1; ++i) cc[i] = k.sup[i];
    if (el > 1) b[0] = 0.0;
    double ll1 = static_cast<double>(el);
    ll1 *= (-0.5) * (ll1 + 1.0);
    for (int i = 0; i < N; ++i) b[i] += 1.0 - ll1 * k.scatt[i];
    std::vector<cd> d(col, col + N), x(N);
    cc[0] /= b[0];
    d[0] /= b[0];
    for (int i = 1; i < N; ++i) {
      double id = 1.0 / (b[i] - cc[i - 1] * a[i]);
      cc[i] *= id;
      d[i] -= d[i - 1] * a[i];
      d[i] *= id;
    }
    x[N - 1] = d[N - 1];
    for (int i = N - 2; i > -1; --i) { x[i] = d[i]; x[i] -= cc[i] * x[i + 1]; }
    for (int i = 0; i < N; ++i) col[i] = x[i];
  }

  // Anisotropic_Collisions::f1_loop / flm_loop, def-FokkerPlanck.cpp:1112-1244
  // which = 1: l = 1 (m = 0,1);  which = 2: l = 2..l0.
  void collide_flm(double Dt, int which) {
    std::vector<double> f00(np);
    Coef k;
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const cd* c0 = Y.h(0) + at(0, ix, iy);
        for (int ip = 0; ip < np; ++ip) f00[ip] = c0[ip].real();
        reset_coeff(f00.data(), Dt, k);
        if (which == 1) {
          for (int m = 0; m <= std::min(1, c.m0); ++m) advance_flm(k, Y.h(H(1, m)) + at(0, ix, iy), 1);
        } else {
          for (int l = 2; l <= c.l0; ++l)
            for (int m = 0; m <= std::min(c.m0, l); ++m) advance_flm(k, Y.h(H(l, m)) + at(0, ix, iy), l);
        }
      }
  }

  // Nonrelativistic moments of Parallel_Output::nonrelativistic_moments,
  // def-Parallel.cpp:1677-2253, in double instead of the reference's float,
  // each in the reference's operation order.  out[k][y][x], k =
  //   0 density (:1699-1726)   1 Vx (:1731-1758)   2 Vy (:1770-1797)
  //   3 VxVx (:1805-1860)      4 VyVy (:1872-1897) 5 VzVz (:1903-1914)
  //   6 Qx (:2094-2134)        7 VxVy (:1920-1947) 8 VxVz (:1953-1980)
  //   9 VyVz (:1986-2013)     10 Vsq = <v^2> - Vx^2 - Vy^2 (:2019-2030)
  //  11 T [eV] (:2034-2040)   12 P [Mbar] (:2041-2051)  13 N_D (:2055-2065)
  //  14 nu (:2067-2080)       15 Qy (:2173-2213)  16 Kx (:2135-2165)
  //  17 Ky (:2214-2244)
  // VyVy takes its 16pi/5 term from f20 as the reference does (:1875-1886).
  // Harmonics the run does not have (f11, f21 for m0 < 1; f22 for m0 < 2;
  // l = 2 for l0 < 2) count as zero.  Kx and Ky difference <v^2>^2 over the
  // whole x (y) extent of one rank, as Matrix2D::Dd1/Dd2 do (matrices.h:
  // 511-533), and copy the first and last interior values to the two ends;
  // Ky needs ny >= 3 and is 0 otherwise (the reference reads past its array).
  static constexpr int kNMom = 18;
  void moments(double* out) const {
    const std::vector<double>& pr = pd;
    const int N = np;
    const double dp = dpd;
    const double fmin = 1.17549435082228750797e-38, feps = 1.1920928955078125e-07;
    std::vector<double> ig(N);
    for (int i = 0; i < N; ++i) ig[i] = 1.0 / std::sqrt(1.0 + pr[i] * pr[i]);
    const double r = pr[0] * pr[0] / (pr[1] * pr[1]);
    const double inv = 1.0 / (1.0 - r);
    // trapezoid of part(f) * w over p: ends first (f(p0) skipped when from1), halved, then the interior
    auto trap = [&](const cd* f, bool from1, double (*w)(double, double), bool imag) {
      auto v = [&](int i) { return imag ? f[i].imag() : f[i].real(); };
      double s = from1 ? 0.0 : v(0) * w(pr[0], ig[0]);
      s += v(N - 1) * w(pr[N - 1], ig[N - 1]);
      s *= 0.5;
      for (int i = 1; i < N - 1; ++i) s += w(pr[i], ig[i]) * v(i);
      return s;
    };
    auto w2 = [](double p, double g) { (void)g; return p * p; };
    auto w3 = [](double p, double g) { return std::pow(p, 3) * g; };
    auto w4 = [](double p, double g) { return std::pow(p, 4) * std::pow(g, 2); };
    auto w5 = [](double p, double g) { return std::pow(p, 5) * std::pow(g, 3); };
    const size_t nxy_ = (size_t)nx * ny;
    auto O = [&](int k, size_t i) -> double& { return out[k * nxy_ + i]; };
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const size_t k = (size_t)iy * nx + ix;
        auto harm = [&](int l, int m) -> const cd* {
          if (l > c.l0 || m > c.m0) return nullptr;
          return Y.h(H(l, m)) + at(0, ix, iy);
        };
        const cd *f0 = harm(0, 0), *f10 = harm(1, 0), *f11 = harm(1, 1), *f20 = harm(2, 0), *f21 = harm(2, 1),
                 *f22 = harm(2, 2);
        // density
        double n = trap(f0, false, w2, false);
        double c00 = f0[1].real() / 5.0 * r + (f0[0].real() - f0[1].real() * r) * inv * (1.0 / 3.0 - 1.0 / 5.0 * r);
        c00 *= pr[0] * (pr[0] / dp) * pr[0];
        n += c00;
        n *= 4.0 * M_PI * dp;
        if (n < 4.0 * feps) n = 1.0;
        // Vx, Vy
        const double c5 = std::pow(pr[0], 5) / (dp * pr[1] * 5.0), c7 = std::pow(pr[0], 7) / (dp * pr[1] * 7.0);
        double vx = trap(f10, false, w3, false) + f10[1].real() * c5;
        vx *= (4.0 * M_PI / 3.0) * dp;
        vx /= n;
        double vy = 0.0;
        if (f11) {
          vy = trap(f11, false, w3, false) + f11[1].real() * c5;
          vy *= (8.0 * M_PI / 3.0) * dp;
          vy /= n;
        }
        // VxVx, VyVy, VzVz
        double m1 = trap(f0, false, w4, false);
        double c01 = f0[1].real() / 7.0 * r + (f0[0].real() - f0[1].real() * r) * inv * (1.0 / 5.0 - 1.0 / 7.0 * r);
        c01 *= std::pow(pr[0], 4) * (pr[0] / dp);
        m1 += c01;
        m1 *= 4.0 * M_PI / 3.0 * dp;
        const double t20 = f20 ? trap(f20, true, w4, false) : 0.0;
        double m2 = t20 * (8.0 * M_PI / 15.0 * dp);
        const double vxvx = (m1 + m2) / n;
        m2 *= 0.5;
        const double m1b = m1 - m2;
        const double t22 = t20 * (16.0 * M_PI / 5.0 * dp);
        const double vyvy = (t22 + m1b) / n;
        const double vzvz = (m1b - t22) / n;
        // VxVy, VxVz (f21), VyVz (f22)
        const double vxvy = f21 ? trap(f21, true, w4, false) * (8.0 * M_PI / 5.0 * dp) / n : 0.0;
        const double vxvz = f21 ? trap(f21, true, w4, true) * (-(8.0 * M_PI / 5.0) * dp) / n : 0.0;
        const double vyvz = f22 ? trap(f22, true, w4, true) * (-(16.0 * M_PI / 5.0) * dp) / n : 0.0;
        // <v^2> about the mean, temperature, pressure, Debye number, collision frequency
        double vsq = 0.0;
        vsq += vxvx; vsq += vyvy; vsq += vzvz;
        vsq -= vx * vx; vsq -= vy * vy;
        const double teV = vsq * 170880.776;
        const double pre = vsq * (2.73785179e-19 * c.density_np) * n;
        const double nd = 1.72e+9 * std::sqrt(std::pow(vsq * 170880.776, 3) / (c.density_np * n));
        double lnei = 24.0 - 0.5 * std::log(c.density_np * n) + std::log(vsq * 170880.776);
        lnei *= c.zeta;
        const double nu = std::sqrt(2.0 / M_PI) / 9.0 * lnei / nd;
        // Qx, Qy
        double q = 0.0;
        q -= vx * (vxvx + vyvy + vzvz);
        q -= 2.0 * (vxvx * vx + vxvy * vy);
        q += 2.0 * (vx * vx + vy * vy) * vx;
        q *= n;
        double m5 = trap(f10, false, w5, false) + f10[1].real() * c7;
        m5 *= (4.0 * M_PI / 3.0) * dp;
        q += m5;
        q *= 0.5;
        double qy = 0.0;
        qy -= vy * (vxvx + vyvy + vzvz);
        qy -= 2.0 * (vxvy * vx + vyvy * vy);
        qy += 2.0 * (vx * vx + vy * vy) * vy;
        qy *= n;
        double m5y = f11 ? trap(f11, false, w5, false) + f11[1].real() * c7 : 0.0;
        m5y *= (8.0 * M_PI / 3.0) * dp;
        qy += m5y;
        qy *= 0.5;
        const double vals[16] = {n, vx, vy, vxvx, vyvy, vzvz, q, vxvy, vxvz, vyvz, vsq, teV, pre, nd, nu, qy};
        for (int j = 0; j < 16; ++j) O(j, k) = vals[j];
      }
    // heat conductivities, def-Parallel.cpp:2135-2165 and :2214-2244
    auto kond = [&](bool ydir, int kq, int kout) {
      const int len = ydir ? ny : nx;
      for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
          const size_t k = (size_t)iy * nx + ix;
          if (len < 3) { O(kout, k) = 0.0; continue; }
          const int i = ydir ? iy : ix;
          const int ii = i == 0 ? 1 : (i == len - 1 ? len - 2 : i);  // ends take their neighbour's value
          const size_t km = ydir ? (size_t)(ii - 1) * nx + ix : (size_t)iy * nx + ii - 1;
          const size_t kp = ydir ? (size_t)(ii + 1) * nx + ix : (size_t)iy * nx + ii + 1;
          double cond = O(10, km) * O(10, km) - O(10, kp) * O(10, kp);  // Dd: minus central difference
          cond *= 0.5 / (ydir ? -c.dy : -c.dx);                          // 0.5/(axis(0) - axis(1))
          cond *= O(0, k);  // density and nu of the cell itself (applied after the end copies)
          cond /= O(14, k);
          cond *= (-1.0) / 18.0;
          if (cond < 4.0 * fmin && cond > -4.0 * fmin) cond = 1.0;
          O(kout, k) = O(kq, k) / cond;
        }
    };
    kond(false, 6, 16);
    kond(true, 15, 17);
  }

  // main_loop step body, main.cpp:79-107
  void step(double h) {
    rk4(h);
    double t = t_rk;
    if (c.coll_ee) collide_ee(t); else t_ee = t;
    if (src.linear || src.poly) laser_source(t);  // main.cpp:90-92
    if (src.ib == 1) inverse_bremsstrahlung(t);
    if (c.coll_flm) {
      if (c.l0 >= 1) collide_flm(t - t_eb, 1);   // Euler_Backward::advance_1
      collide_flm(t - t_eb, 2);                  // Euler_Backward::advance_lm
    }
    t_eb = t;
  }

  // ------------------------------------------- implicit E, 1D variant ----
  // The step of main.cpp:79-107 with implicit_E and if_implicit1D:
  // Runge_Kutta_2_Imp (def-RK.cpp:400-588), Implicit_Actions (def-Actions.cpp:101-225),
  // Implicit_E_Field::advance (def-ImplicitE.cpp:472-606).  Only the m = 0
  // harmonics and Ex take part in the 1D variant's field solve.
  static void zero_state(State& y) {
    std::fill(y.f.begin(), y.f.end(), cd(0.0, 0.0));
    std::fill(y.fld.begin(), y.fld.end(), cd(0.0, 0.0));
  }
  // Runge_Kutta_2_Imp::advance_p1/p2/IEx (def-RK.cpp:470-498, 503-524, 529-550)
  template <class F>
  void rk2(F&& fn, double h) {
    State Y0 = Y, Yh;
    Yh.init(nh, n, nxy);
    fn(Y0, Yh);
    scale_state(Yh, h); axpy_state(Y0, Yh);
    scale_state(Yh, 0.5); axpy_state(Y, Yh);
    fn(Y0, Yh);
    scale_state(Yh, 0.5 * h); axpy_state(Y, Yh);
  }
  // The m = 0 E-field couplings with Ex of ElectricField::Exyz1D
  // (def-VlasovMax.cpp:291-341; sources l <= lsrc, targets l <= ltgt): per
  // source, H to l-1 then G to l+1.  Implicit_Ex1D (:200-226) is the same
  // with sources l <= 3 and targets l <= 2.
  void exyz_m0(const State& Yin, State& Yh, int lsrc, int ltgt) const {
    std::vector<cd> G(np), Hh(np);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const cd Ex = Yin.F(0)[(size_t)iy * nx + ix];
        for (int l = 0; l <= std::min(lsrc, c.l0); ++l) {
          const cd* fcol = Yin.h(H(l, 0)) + at(0, ix, iy);
          if (l == 0) g00_col(fcol, G.data());
          else make_gh_col(l, fcol, G.data(), Hh.data());
          if (l >= 1 && l - 1 <= ltgt) {
            const cd mult = Ex * eA2[l2i(l, 0)];
            cd* dst = Yh.h(H(l - 1, 0)) + at(0, ix, iy);
            for (int ip = 0; ip < np; ++ip) dst[ip] += Hh[ip] * mult;
          }
          if (l + 1 <= c.l0 && l + 1 <= ltgt) {
            const cd mult = Ex * eA1[l2i(l, 0)];
            cd* dst = Yh.h(H(l + 1, 0)) + at(0, ix, iy);
            for (int ip = 0; ip < np; ++ip) dst[ip] += G[ip] * mult;
          }
        }
      }
  }
  // Implicit_Actions::Adv1 (def-Actions.cpp:118-132)
  void adv1(const State& Yin, State& Yh, bool es) const {
    zero_state(Yh);
    axy(Yin, Yh);
    if (!es) {
      bxyz(Yin, Yh);
      faraday(Yin, Yh);
    }
  }
  // Implicit_Actions::Adv2 with if_implicit1D (def-Actions.cpp:138-160)
  void adv2_1d(const State& Yin, State& Yh) const {
    zero_state(Yh);
    exyz_m0(Yin, Yh, c.l0, c.l0);
    clean_low_pc(Yh);
  }
  // Electric_Field_Methods::Current_xyz::calculate_J, x component
  // (def-ImplicitE.cpp:125-156).  Its trapezoid end weight is applied as
  // jayx(ix) *= 0.5, a 1D-style subscript of the x-fastest field, i.e. to
  // element (ix, 0) -- a guard row -- so interior cells keep weight 1 at p0
  // and pmax.
  void current_imp_x(const State& Yin, std::vector<cd>& jx) const {
    jx.assign(nxy, cd(0.0, 0.0));
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const cd* f = Yin.h(H(1, 0)) + at(0, ix, iy);
        cd j = f[0] * p3og[0];
        j += f[np - 1] * p3og[np - 1];
        for (int ip = 1; ip < np - 1; ++ip) j += f[ip] * p3og[ip];
        j *= Delta_p;
        j += small * (f[1] * p3og[1]);
        j *= (-4.0) * M_PI / 3.0;
        jx[(size_t)iy * nx + ix] = j;
      }
  }
  // Implicit_E_Field::advance with if_implicit1D (def-ImplicitE.cpp:472-606):
  // the current without E (after the l = 1 collisions) and with the probe
  // field DE (FindDE, :628-652) after an RK2 of Implicit_Ex1D and the l = 1
  // collisions; Ex = DE J0 / (J0 - J_DE), Ey = Ez = 0.  The harmonics the
  // probes touch are restored after each.
  void implicit_e_1d(double h) {
    const double Eps = 16.0 * std::numeric_limits<double>::epsilon();
    const double LargeEps = std::sqrt(Eps);
    std::vector<cd> DE(nxy);
    for (int k = 0; k < nxy; ++k) {
      cd ex = Y.F(0)[k], ey = Y.F(1)[k], ez = Y.F(2)[k];
      ex *= ex; ey *= ey; ez *= ez;
      ex += ey; ex += ez;
      ex = std::sqrt(ex);
      ex *= LargeEps;
      ex += Eps;
      DE[k] = ex;
    }
    std::vector<int> saved;
    for (int l = 0; l <= std::min(2, c.l0); ++l)
      for (int m = 0; m <= std::min(l, c.m0); ++m) saved.push_back(H(l, m));
    std::vector<std::vector<cd>> keep;
    for (int i : saved) keep.emplace_back(Y.h(i), Y.h(i) + n);
    auto restore = [&](int lmax) {
      for (size_t k = 0; k < saved.size(); ++k) {
        bool in = false;
        for (int l = 0; l <= lmax; ++l)
          for (int m = 0; m <= std::min(l, c.m0); ++m) in = in || (saved[k] == H(l, m));
        if (in) std::copy(keep[k].begin(), keep[k].end(), Y.h(saved[k]));
      }
    };
    const double Dt = t_rk - t_eb;
    std::vector<cd> J0, JE;
    if (c.coll_flm) collide_flm(Dt, 1);
    current_imp_x(Y, J0);
    restore(1);
    for (int k = 0; k < nxy; ++k) Y.F(0)[k] = DE[k];
    rk2([&](const State& a, State& b) { zero_state(b); exyz_m0(a, b, 3, 2); }, h);
    if (c.coll_flm) collide_flm(Dt, 1);
    current_imp_x(Y, JE);
    restore(2);
    for (int k = 0; k < nxy; ++k) {
      Y.F(0)[k] = DE[k] * J0[k] / (J0[k] - JE[k]);
      Y.F(1)[k] = 0.0;
      Y.F(2)[k] = 0.0;
    }
  }
  void step_implicit_1d(double h, bool es) {
    rk2([&](const State& a, State& b) { adv1(a, b, es); }, h);
    t_rk += h;
    double t = t_rk;
    if (c.coll_ee) collide_ee(t); else t_ee = t;
    if (src.linear || src.poly) laser_source(t);  // main.cpp:90-92
    if (src.ib == 1) inverse_bremsstrahlung(t);
    implicit_e_1d(h);
    rk2([&](const State& a, State& b) { adv2_1d(a, b); }, h);
    if (c.coll_flm) {
      collide_flm(t - t_eb, 1);
      collide_flm(t - t_eb, 2);
    }
    t_eb = t;
  }

  // ------------------------------------------- implicit E, 2D variant ----
  // Current_xyz::calculate_J (def-ImplicitE.cpp:125-180): x from f10, y and z
  // from f11 (Jy = Re, Jz = Re(i J) of the -8pi/3 sum), p0 and pmax at
  // weight 1 on interior cells as in the x component above.
  void current_imp(const State& Yin, std::vector<cd>& jx, std::vector<cd>& jy, std::vector<cd>& jz) const {
    current_imp_x(Yin, jx);
    jy.assign(nxy, cd(0.0, 0.0));
    jz.assign(nxy, cd(0.0, 0.0));
    const cd c01(0.0, 1.0);
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) {
        const cd* f = Yin.h(H(1, 1)) + at(0, ix, iy);
        cd j = f[0] * p3og[0];
        j += f[np - 1] * p3og[np - 1];
        for (int ip = 1; ip < np - 1; ++ip) j += f[ip] * p3og[ip];
        j *= Delta_p;
        j += small * (f[1] * p3og[1]);
        j *= (-8.0) * M_PI / 3.0;
        const size_t k = (size_t)iy * nx + ix;
        jy[k] = cd(j.real(), 0.0);
        j *= c01;
        jz[k] = cd(j.real(), 0.0);
      }
  }
  // JN of Implicit_E_Field::Ampere (def-ImplicitE.cpp:652-678): the same
  // differences as MaxwellEq::Ampere, assigned instead of added
  void ampere_jn(std::vector<cd>* jn) const {
    State z;
    z.init(0, 0, nxy);
    for (int k = 3; k < 6; ++k) std::copy(Y.F(k), Y.F(k) + nxy, z.F(k));
    ampere(z, z);
    for (int k = 0; k < 3; ++k) jn[k].assign(z.F(k), z.F(k) + nxy);
  }
  // JN when if_implicitES: Ampere is skipped and JN keeps what Current_xyz's
  // constructor gave it, a copy of the Ex field at the solver's construction
  // in all three components (def-ImplicitE.cpp:56-58, 446); the caller sets it.
  std::vector<cd> es_jn[3];

  // Implicit_E_Field::advance, 2D variant (def-ImplicitE.cpp:472-606):
  // J0 = current after Euler_Backward::advance_1 with the fields untouched;
  // J_Ex, J_Ey, J_Ez = current after an RK2 of the probe couplings (exyz with
  // probe 1, 2, 3) with Ex = DE.Ex, Ey = DE.Ey, Ez = DE.Ez in turn and
  // advance_1; f00, f1m, f2m restored after each.  Then per cell the 3x3
  // conductivity sgm(i,j) = (J_Ej - J0)_i / DE_j and E = sgm^-1 (JN - J0) by
  // Cramer's rule (Det33, Detx33, Dety33, Detz33: def-Nmethods.cpp:215-251).
  // Cells whose |Re det| <= 6 DBL_MIN get E = 0, DE.Ex *= 2, DE.Ey *= 1.3,
  // and the whole probe sequence runs again, at most 4 attempts.  Returns the
  // number of such cells left after the last attempt (the reference warns,
  // and exits when there are more than 8).
  int implicit_e_2d(double h, bool es) {
    const double Eps = 16.0 * std::numeric_limits<double>::epsilon();
    const double LargeEps = std::sqrt(Eps);
    std::vector<cd> DE[3];
    for (int k = 0; k < 3; ++k) DE[k].assign(nxy, cd(0.0, 0.0));
    for (int k = 0; k < nxy; ++k) {  // FindDE, :628-650
      cd ex = Y.F(0)[k], ey = Y.F(1)[k], ez = Y.F(2)[k];
      ex *= ex; ey *= ey; ez *= ez;
      ex += ey; ex += ez;
      ex = std::sqrt(ex);
      ex *= LargeEps;
      ex += Eps;
      DE[0][k] = DE[1][k] = DE[2][k] = ex;
    }
    std::vector<int> saved;
    for (int l = 0; l <= 2; ++l)
      for (int m = 0; m <= std::min(l, c.m0); ++m) saved.push_back(H(l, m));
    std::vector<std::vector<cd>> keep;
    for (int i : saved) keep.emplace_back(Y.h(i), Y.h(i) + n);
    auto restore = [&](int lmax) {
      for (size_t k = 0; k < saved.size(); ++k) {
        bool in = false;
        for (int l = 0; l <= lmax; ++l)
          for (int m = 0; m <= std::min(l, c.m0); ++m) in = in || (saved[k] == H(l, m));
        if (in) std::copy(keep[k].begin(), keep[k].end(), Y.h(saved[k]));
      }
    };
    const double Dt = t_rk - t_eb;
    std::vector<cd> J0[3], JE[3][3], JN[3], EN[3];
    for (int k = 0; k < 3; ++k) EN[k].assign(nxy, cd(0.0, 0.0));
    int zeros = 1, attempt = 0;
    while (zeros > 0 && attempt < 4) {
      zeros = 0;
      ++attempt;
      if (c.coll_flm) collide_flm(Dt, 1);
      current_imp(Y, J0[0], J0[1], J0[2]);
      restore(1);
      for (int d = 0; d < 3; ++d) {
        std::copy(DE[d].begin(), DE[d].end(), Y.F(d));
        rk2([&](const State& a, State& b) { zero_state(b); exyz(a, b, d + 1); }, h);
        if (c.coll_flm) collide_flm(Dt, 1);
        current_imp(Y, JE[d][0], JE[d][1], JE[d][2]);
        restore(2);
      }
      if (!es) ampere_jn(JN);
      else for (int k = 0; k < 3; ++k) JN[k] = es_jn[k].empty() ? std::vector<cd>(nxy, cd(0.0, 0.0)) : es_jn[k];
      for (int k = 0; k < nxy; ++k) {
        cd A[3][3], D[3];
        for (int i = 0; i < 3; ++i) {
          for (int j = 0; j < 3; ++j) A[i][j] = (JE[j][i][k] - J0[i][k]) / DE[j][k];
          D[i] = JN[i][k] - J0[i][k];
        }
        const cd det = A[0][0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                       A[1][0] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                       A[2][0] * (A[0][1] * A[1][2] - A[1][1] * A[0][2]);
        if (std::abs(det.real()) > 6.0 * std::numeric_limits<double>::min()) {
          EN[0][k] = (D[0] * (A[1][1] * A[2][2] - A[2][1] * A[1][2]) -
                      D[1] * (A[0][1] * A[2][2] - A[2][1] * A[0][2]) +
                      D[2] * (A[0][1] * A[1][2] - A[1][1] * A[0][2])) / det;
          EN[1][k] = (A[0][0] * (D[1] * A[2][2] - D[2] * A[1][2]) -
                      A[1][0] * (D[0] * A[2][2] - D[2] * A[0][2]) +
                      A[2][0] * (D[0] * A[1][2] - D[1] * A[0][2])) / det;
          EN[2][k] = (A[0][0] * (A[1][1] * D[2] - A[2][1] * D[1]) -
                      A[1][0] * (A[0][1] * D[2] - A[2][1] * D[0]) +
                      A[2][0] * (A[0][1] * D[1] - A[1][1] * D[0])) / det;
        } else {
          ++zeros;
          DE[0][k] *= 2;
          DE[1][k] *= 1.3;
          EN[0][k] = EN[1][k] = EN[2][k] = 0.0;
        }
      }
    }
    for (int k = 0; k < 3; ++k) std::copy(EN[k].begin(), EN[k].end(), Y.F(k));
    return zeros;
  }
  // Implicit_Actions::Adv2 without if_implicit1D (def-Actions.cpp:138-160)
  void adv2_2d(const State& Yin, State& Yh) const {
    zero_state(Yh);
    exyz(Yin, Yh);
    clean_low_pc(Yh);
  }
  int step_implicit_2d(double h, bool es) {
    rk2([&](const State& a, State& b) { adv1(a, b, es); }, h);
    t_rk += h;
    double t = t_rk;
    if (c.coll_ee) collide_ee(t); else t_ee = t;
    if (src.linear || src.poly) laser_source(t);  // main.cpp:90-92
    if (src.ib == 1) inverse_bremsstrahlung(t);
    const int zeros = implicit_e_2d(h, es);
    rk2([&](const State& a, State& b) { adv2_2d(a, b); }, h);
    if (c.coll_flm) {
      collide_flm(t - t_eb, 1);
      collide_flm(t - t_eb, 2);
    }
    t_eb = t;
    return zeros;
  }
};


// ===================================================================== TLM ==
// Forward-mode dual number: value and directional derivative.
struct Dual {
  double v, d;
  Dual(double a = 0.0, double b = 0.0) : v(a), d(b) {}
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator-(Dual a) { return {-a.v, -a.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual& operator+=(Dual& a, Dual b) { return a = a + b; }
inline Dual& operator-=(Dual& a, Dual b) { return a = a - b; }
inline Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
inline Dual& operator/=(Dual& a, Dual b) { return a = a / b; }
inline Dual lg(Dual a) { return {std::log(a.v), a.d / a.v}; }
inline Dual sq(Dual a) { double s = std::sqrt(a.v); return {s, a.d / (2.0 * s)}; }
inline double lg(double a) { return std::log(a); }
inline double sq(double a) { return std::sqrt(a); }
inline double val(double a) { return a; }
inline double val(Dual a) { return a.v; }

// One oshun step (main.cpp:79-107) of a 1D run with m0 = 0 on real scalars R:
// the oracle's formulas above restricted to m = 0 (every coupling to m >= 1 and
// every y term vanishes).  State: f[l][x][p], fields[6][x].
template <class R>
struct Step1D {
  const Oracle& o;
  int L, np, nx;
  explicit Step1D(const Oracle& q) : o(q), L(q.c.l0), np(q.np), nx(q.nx) {}
  size_t ix(int l, int x, int p) const { return ((size_t)l * nx + x) * np + p; }

  // Actions::Adv (def-Actions.cpp:73-94) for m0 = 0, 1D
  void adv(const std::vector<R>& f, const std::vector<R>& fld, std::vector<R>& kf, std::vector<R>& kfld) const {
    const int N = np;
    kf.assign(f.size(), R(0.0));
    kfld.assign(fld.size(), R(0.0));
    std::vector<R> G(N), Hh(N), D(N);
    const double dpc = o.dpc.real();
    for (int x = 0; x < nx; ++x) {
      const R Ex = fld[x];
      for (int l = 0; l <= L; ++l) {  // ElectricField::Exyz for m = 0 (def-VlasovMax.cpp:346-491)
        const R* fc = &f[ix(l, x, 0)];
        D[0] = R(0.0);
        for (int p = 1; p < N - 1; ++p) D[p] = fc[p - 1] - fc[p + 1];
        D[N - 1] = fc[N - 2] - fc[N - 1];
        if (l == 0) {  // G00 (:519-535)
          const double p0p1_sq = (o.pc[0] * o.pc[0] / (o.pc[1] * o.pc[1])).real();
          const double inv = 1.0 / (1.0 - p0p1_sq);
          const double g_r = (-4.0 * o.dpc * o.pc[0] / (o.pc[1] * o.pc[1])).real();
          for (int p = 0; p < N; ++p) G[p] = D[p];
          const R f00 = (fc[0] - fc[1] * p0p1_sq) * inv;
          G[0] = (fc[1] - f00) * g_r;
        } else {  // MakeGH (:496-515)
          const double sg = (-2.0 * (l + 1.0)) * dpc, cg = -(2.0 * l + 1.0) / l;
          for (int p = 0; p < N; ++p) {
            const double invpax = o.invpr[p].real() * sg;
            Hh[p] = fc[p] * invpax + D[p];
            G[p] = D[p] * cg + Hh[p];
          }
          G[0] = R(0.0);
          Hh[0] = fc[1] * o.Hp0[l].real();
        }
        if (l >= 1) {  // H of l feeds l-1 with A2(l,0)
          const R mult = Ex * o.eA2[o.l2i(l, 0)].real();
          for (int p = 0; p < N; ++p) kf[ix(l - 1, x, p)] += Hh[p] * mult;
        }
        if (l + 1 <= L) {  // G of l feeds l+1 with A1(l,0)
          const R mult = Ex * o.eA1[o.l2i(l, 0)].real();
          for (int p = 0; p < N; ++p) kf[ix(l + 1, x, p)] += G[p] * mult;
        }
      }
      {  // Current::Jxyz into Ex (:680-745)
        const R* f1 = &f[ix(1, x, 0)];
        R j = f1[0] * o.p3og[0].real();
        j += f1[N - 1] * o.p3og[N - 1].real();
        j *= 0.5;
        for (int p = 1; p < N - 1; ++p) j += f1[p] * o.p3og[p].real();
        j *= o.Delta_p.real();
        j += o.small.real() * (f1[1] * o.p3og[1].real());
        j *= 4.0 * M_PI / 3.0;
        kfld[x] += j;
      }
      // Ampere and Faraday, x terms (:979-1035)
      const double mx = o.midx.real();
      const int xm = o.xm(x), xp = o.xp(x);
      auto F = [&](int c, int xx) { return fld[(size_t)c * nx + xx]; };
      kfld[(size_t)nx + x] += F(5, xm) * (-mx) - F(5, xp) * (-mx);
      kfld[(size_t)2 * nx + x] += F(4, xm) * mx - F(4, xp) * mx;
      kfld[(size_t)4 * nx + x] += F(2, xm) * mx - F(2, xp) * mx;
      kfld[(size_t)5 * nx + x] += F(1, xm) * (-mx) - F(1, xp) * (-mx);
    }
    // Spatial_Advection::Axy, x terms (:842-853)
    for (int l = 0; l <= L; ++l)
      for (int x = 0; x < nx; ++x) {
        const int xm = o.xm(x), xp = o.xp(x);
        for (int p = 0; p < N; ++p) {
          const R df = f[ix(l, xm, p)] - f[ix(l, xp, p)];
          if (l >= 1) kf[ix(l - 1, x, p)] += df * (o.vr[p] * o.aA2[o.l2i(l, 0)]).real();
          if (l + 1 <= L) kf[ix(l + 1, x, p)] += df * (o.vr[p] * o.aA1[o.l2i(l, 0)]).real();
        }
      }
    // CLEAN_LOW_PC (def-Actions.cpp:240-253)
    for (int l = 1; l <= L; ++l)
      for (int x = 0; x < nx; ++x)
        for (int p = 0; p < std::min(l - 1, 80); ++p) kf[ix(l, x, p)] = R(0.0);
  }

  // Runge_Kutta_4::advance_p1 (def-RK.cpp:110-145)
  void rk4(std::vector<R>& f, std::vector<R>& fld, double h) const {
    std::vector<R> f0 = f, g0 = fld, f1 = f, g1 = fld, kf, kg;
    auto axpy = [](std::vector<R>& y, const std::vector<R>& k, double a, double b, std::vector<R>* acc) {
      for (size_t i = 0; i < y.size(); ++i) {
        const R kh = k[i] * a;
        y[i] += kh;
        if (acc) (*acc)[i] += kh * b;
      }
    };
    adv(f1, g1, kf, kg);
    axpy(f1, kf, 0.5 * h, 1.0 / 3.0, &f); axpy(g1, kg, 0.5 * h, 1.0 / 3.0, &fld);
    adv(f1, g1, kf, kg);
    f1 = f0; g1 = g0;
    axpy(f1, kf, 0.5 * h, 2.0 / 3.0, &f); axpy(g1, kg, 0.5 * h, 2.0 / 3.0, &fld);
    adv(f1, g1, kf, kg);
    axpy(f0, kf, h, 1.0 / 3.0, &f); axpy(g0, kg, h, 1.0 / 3.0, &fld);
    adv(f0, g0, kf, kg);
    for (size_t i = 0; i < f.size(); ++i) f[i] += kf[i] * (h / 6.0);
    for (size_t i = 0; i < fld.size(); ++i) fld[i] += kg[i] * (h / 6.0);
  }

  // LOGee / ZLOGei (def-FokkerPlanck.cpp:41-95)
  R logee(R ne, R Te) const {
    if (val(ne) > 0.000000001) {
      Te /= (3.0 * ne);
      Te *= 511000.0;
      ne *= o.c.density_np;
      Te = lg(Te);
      ne = lg(ne);
      R lnee = 23.5 - 0.5 * ne + 1.25 * Te - sq(0.00001 + 0.0625 * (Te - 2.0) * (Te - 2.0));
      if (val(lnee) > 2.0) return lnee;
    }
    return R(2.0);
  }
  R zlogei(R ne, R Te) const {
    if (val(ne) > 0.0000001) {
      Te /= (3.0 * ne);
      Te *= 511000.0;
      ne *= o.c.density_np;
      Te = lg(Te);
      ne = lg(ne);
      R lnei = 24.0 - 0.5 * ne + Te;
      if (val(lnei) > 2.0) return lnei * o.c.zeta;
    }
    return R(2.0 * o.c.zeta);
  }

  // EE_isotropic_conserva::Getslope with G (def-FokkerPlanck.cpp:178-267)
  void getslope(const R* fin, R* fh) const {
    const int N = np;
    const std::vector<double>& v = o.vr_c;
    std::vector<R> I4(N), I2(N), J1(N);
    I4[0] = R(0.0);
    for (int k = 1; k < N; ++k) { I4[k] = o.U4[k] * fin[k] + o.U4m1[k] * fin[k - 1]; I4[k] += I4[k - 1]; }
    I2[0] = R(0.0);
    for (int k = 1; k < N; ++k) { I2[k] = o.U2[k] * fin[k] + o.U2m1[k] * fin[k - 1]; I2[k] += I2[k - 1]; }
    J1[N - 1] = R(0.0);
    for (int k = N - 2; k > -1; --k) { J1[k] = o.U1[k + 1] * fin[k + 1] + o.U1m1[k + 1] * fin[k]; J1[k] += J1[k + 1]; }
    const R Ln_ee = logee(4.0 * M_PI * I2[N - 1], 4.0 * M_PI * I4[N - 1]);
    std::vector<R> G(N);
    for (int k = 0; k < N; ++k) G[k] = (I4[k] + o.U3[k] * J1[k]) * fin[k] + (I2[k] * J1[k]) * (-3.0);
    const R f00 = (fin[0] - fin[1] * (v[0] * v[0]) / (v[1] * v[1])) / (1.0 - (v[0] * v[0]) / (v[1] * v[1]));
    for (int k = 0; k < o.c.nb && k < N; ++k) {
      const R i2s = f00 * std::pow(v[k], 3) / 3.0 + (fin[1] - f00) * (std::pow(v[k], 5) / (v[1] * v[1]) * 0.2);
      const R i4s = f00 * std::pow(v[k], 5) * 0.2 + (fin[1] - f00) * (std::pow(v[k], 7) / (v[1] * v[1] * 7.0));
      G[k] = fin[k] * i4s + (std::pow(v[k], 3) * fin[k] - 3.0 * i2s) * J1[k];
    }
    R g0 = (-1.0) * G[0];
    for (int k = 0; k < N - 1; ++k) G[k] -= G[k + 1];
    g0 *= 2.0 / (v[0] * v[0]);
    for (int k = 0; k < N - 1; ++k) G[k] *= o.Pn[k];
    g0 -= G[0];
    for (int k = 0; k < N - 1; ++k) G[k] -= G[k + 1];
    g0 *= o.Qn[0];
    fh[0] = g0;
    for (int k = 0; k < N - 1; ++k) fh[k + 1] = o.Qn[k + 1] * G[k];
    for (int k = 0; k < N; ++k) fh[k] *= o.c_kpre * Ln_ee;
  }
  // RungeKutta4_ee::advance (def-FokkerPlanck.cpp:518-553)
  void rk4_ee(R* f, double h) const {
    const int N = np;
    std::vector<R> f0(f, f + N), f1(f, f + N), fh(N);
    getslope(f1.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= (0.5 * h); f1[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (1.0 / 3.0); f[k] += fh[k]; }
    getslope(f1.data(), fh.data());
    f1 = f0;
    for (int k = 0; k < N; ++k) { fh[k] *= (0.5 * h); f1[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (2.0 / 3.0); f[k] += fh[k]; }
    getslope(f1.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= h; f0[k] += fh[k]; }
    for (int k = 0; k < N; ++k) { fh[k] *= (1.0 / 3.0); f[k] += fh[k]; }
    getslope(f0.data(), fh.data());
    for (int k = 0; k < N; ++k) { fh[k] *= (h / 6.0); f[k] += fh[k]; }
  }
  // reset_coeff + advance + TridiagonalSolve (def-FokkerPlanck.cpp:808-1035, def-Nmethods.cpp:142-167)
  void implicit(const R* fin, double Dt, int lmin, int lmax, std::vector<R>& f, int x) const {
    const int N = np;
    const std::vector<double>& v = o.vr_c;
    std::vector<R> I2(N), I0(N), J1m(N);
    I2[0] = R(0.0);
    for (int q = 1; q < N; ++q) { I2[q] = o.U4[q] * fin[q] + o.U4m1[q] * fin[q - 1]; I2[q] += I2[q - 1]; }
    for (int q = 0; q < N; ++q) I2[q] *= 4.0 * M_PI;
    const R I2_temperature = I2[N - 1];
    for (int q = 0; q < N; ++q) I2[q] /= v[q] * v[q];
    I0[0] = R(0.0);
    for (int q = 1; q < N; ++q) { I0[q] = o.U2[q] * fin[q] + o.U2m1[q] * fin[q - 1]; I0[q] += I0[q - 1]; }
    for (int q = 0; q < N; ++q) I0[q] *= 4.0 * M_PI;
    const R I0_density = I0[N - 1];
    J1m[N - 1] = R(0.0);
    for (int q = N - 2; q > -1; --q) { J1m[q] = o.U1[q + 1] * fin[q + 1] + o.U1m1[q + 1] * fin[q]; J1m[q] += J1m[q + 1]; }
    for (int q = 0; q < N; ++q) J1m[q] *= 4.0 * M_PI * v[q];
    const R lnee = logee(I0_density, I2_temperature);
    const R zlnei = zlogei(I0_density, I2_temperature);
    std::vector<R> T1(N), T2(N), sc(N), sub(N), dia(N), sup(N);
    for (int i = 0; i < N; ++i) {
      T1[i] = I0[i] + (2.0 * J1m[i] - I2[i]) / 3.0;
      T2[i] = (I2[i] + J1m[i]) / 3.0;
    }
    const R zi = zlnei * I0_density;
    const double kd = o.kpre * Dt;
    for (int i = 0; i < N; ++i) {
      R s_ = (i == 0) ? R(0.0) : T1[i] * lnee;
      s_ += zi;
      s_ /= std::pow(v[i], 3);
      sc[i] = s_ * kd;
    }
    dia[0] = 8.0 * M_PI * fin[0];
    for (int i = 1; i < N - 1; ++i) {
      const double IvDnDm1 = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i] - v[i - 1]));
      const double IvDnDp1 = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i + 1] - v[i]));
      const double Ivsq2Dn = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]));
      dia[i] = 8.0 * M_PI * fin[i] - T2[i] * (IvDnDm1 + IvDnDp1);
      sub[i] = T2[i] * IvDnDm1 - T1[i] * Ivsq2Dn;
      sup[i] = T2[i] * IvDnDp1 + T1[i] * Ivsq2Dn;
    }
    const R s = (-1.0) * lnee * o.kpre * Dt;
    for (int i = 0; i < N; ++i) { sub[i] *= s; dia[i] *= s; sup[i] *= s; }
    for (int l = lmin; l <= lmax; ++l) {
      std::vector<R> a(N, R(0.0)), b(N), cc(N, R(0.0)), d(N), xx(N);
      for (int i = 1; i < N; ++i) a[i] = sub[i];
      for (int i = 0; i < N; ++i) b[i] = dia[i];
      for (int i = 0; i < N - 1; ++i) cc[i] = sup[i];
      if (l > 1) b[0] = R(0.0);
      double ll1 = static_cast<double>(l);
      ll1 *= (-0.5) * (ll1 + 1.0);
      for (int i = 0; i < N; ++i) b[i] += 1.0 - ll1 * sc[i];
      for (int i = 0; i < N; ++i) d[i] = f[ix(l, x, i)];
      cc[0] /= b[0];
      d[0] /= b[0];
      for (int i = 1; i < N; ++i) {
        const R id = 1.0 / (b[i] - cc[i - 1] * a[i]);
        cc[i] *= id;
        d[i] -= d[i - 1] * a[i];
        d[i] *= id;
      }
      xx[N - 1] = d[N - 1];
      for (int i = N - 2; i > -1; --i) { xx[i] = d[i]; xx[i] -= cc[i] * xx[i + 1]; }
      for (int i = 0; i < N; ++i) f[ix(l, x, i)] = xx[i];
    }
  }
  // main.cpp:79-107 with the clocks of Oracle::step
  void step(std::vector<R>& f, std::vector<R>& fld, double h, double clk[3]) const {
    rk4(f, fld, h);
    clk[0] += h;
    const double t = clk[0];
    if (o.c.coll_ee) {
      int numh = static_cast<int>(static_cast<size_t>(static_cast<int>((t - clk[1]) / o.c.small_dt)) + 1);
      double hh = (t - clk[1]) / static_cast<double>(static_cast<size_t>(numh));
      std::vector<R> fc(np);
      for (int x = 0; x < nx; ++x) {
        for (int p = 0; p < np; ++p) fc[p] = f[ix(0, x, p)];
        for (int s = 0; s < numh; ++s) rk4_ee(fc.data(), hh);
        for (int p = 0; p < np; ++p) f[ix(0, x, p)] = fc[p];
      }
    }
    clk[1] = t;
    if (o.c.coll_flm) {
      const double Dt = t - clk[2];
      for (int x = 0; x < nx; ++x) {
        std::vector<R> f00(&f[ix(0, x, 0)], &f[ix(0, x, 0)] + np);
        implicit(f00.data(), Dt, 1, 1, f, x);         // Euler_Backward::advance_1
        implicit(f00.data(), Dt, 2, L, f, x);         // Euler_Backward::advance_lm
      }
    }
    clk[2] = t;
  }
};

void put_state(State& s, const double* f, const double* fld) {
  if (f) std::memcpy((void*)s.f.data(), f, s.f.size() * sizeof(cd));
  if (fld) std::memcpy((void*)s.fld.data(), fld, s.fld.size() * sizeof(cd));
}
void get_state(const State& s, double* f, double* fld) {
  if (f) std::memcpy(f, s.f.data(), s.f.size() * sizeof(cd));
  if (fld) std::memcpy(fld, s.fld.data(), s.fld.size() * sizeof(cd));
}

}  // namespace

extern "C" {

// config: ints {l0,m0,np,nx,ny,ref_quirks,nb,coll_ee,coll_flm}
//         doubles {pmin,pmax,dx,dy,density_np,zeta,small_dt}
void* oo_create(const int* iv, const double* dv) {
  Cfg c;
  c.l0 = iv[0]; c.m0 = iv[1]; c.np = iv[2]; c.nx = iv[3]; c.ny = iv[4];
  c.ref_quirks = iv[5]; c.nb = iv[6]; c.coll_ee = iv[7]; c.coll_flm = iv[8];
  c.pmin = dv[0]; c.pmax = dv[1]; c.dx = dv[2]; c.dy = dv[3];
  c.density_np = dv[4]; c.zeta = dv[5]; c.small_dt = dv[6];
  if (c.l0 < 1 || c.m0 < 0 || c.m0 > c.l0 || c.np < 4 || c.nx < 3 || c.ny < 1) return nullptr;
  return new Oracle(c);
}
void oo_destroy(void* o) { delete static_cast<Oracle*>(o); }
int oo_num_harmonics(void* o) { return static_cast<Oracle*>(o)->nh; }
void oo_set_state(void* o, const double* f, const double* fld) { put_state(static_cast<Oracle*>(o)->Y, f, fld); }
void oo_get_state(void* o, double* f, double* fld) { get_state(static_cast<Oracle*>(o)->Y, f, fld); }
void oo_set_time(void* o, double t) {
  Oracle* q = static_cast<Oracle*>(o);
  q->t_rk = q->t_ee = q->t_eb = t;
}
double oo_time(void* o) { return static_cast<Oracle*>(o)->t_rk; }

// Actions::Adv on the current state; writes the slope.
void oo_rhs(void* o, double* f_out, double* fld_out) {
  Oracle* q = static_cast<Oracle*>(o);
  State Yh; Yh.init(q->nh, q->n, q->nxy);
  q->adv(q->Y, Yh);
  get_state(Yh, f_out, fld_out);
}
void oo_rk4(void* o, double h) { static_cast<Oracle*>(o)->rk4(h); }
void oo_collide_ee(void* o, double tnew) { static_cast<Oracle*>(o)->collide_ee(tnew); }
void oo_collide_flm(void* o, double tnew, int which) {
  Oracle* q = static_cast<Oracle*>(o);
  q->collide_flm(tnew - q->t_eb, which);
  if (which == 2) q->t_eb = tnew;
}
void oo_step(void* o, double h, int nsteps) {
  Oracle* q = static_cast<Oracle*>(o);
  for (int s = 0; s < nsteps; ++s) q->step(h);
}
void oo_moments(void* o, double* out) { static_cast<Oracle*>(o)->moments(out); }
// sources: sv = {linear, poly, ib, pol, I_0, lambda_0, ab_eta, ab_depth, cntr_x0, cntr_r0, spot_w0,
// rise, flat, fall, smaller_dt}; x/y: positions of the interior cells; xglob: the global x axis
void oo_set_sources(void* o, const double* sv, const double* x, const double* y, const double* xglob, int nxg) {
  Oracle* q = static_cast<Oracle*>(o);
  q->set_sources(sv, x, y, xglob, nxg, q->t_rk);
}
int oo_laser_source(void* o, double tnew) {
  Oracle* q = static_cast<Oracle*>(o);
  q->src.alpha_error = false;
  q->laser_source(tnew);
  return q->src.alpha_error ? -1 : 0;
}
void oo_inverse_bremsstrahlung(void* o, double tnew) { static_cast<Oracle*>(o)->inverse_bremsstrahlung(tnew); }
void oo_laser_efield(void* o, double tnew) { static_cast<Oracle*>(o)->laser_efield(tnew); }
// nsteps implicit-E steps of the 1D variant (l0 >= 3); es = if_implicitES
int oo_step_implicit_1d(void* o, double h, int nsteps, int es) {
  Oracle* q = static_cast<Oracle*>(o);
  if (q->c.l0 < 3) return -1;
  for (int s = 0; s < nsteps; ++s) q->step_implicit_1d(h, es != 0);
  return 0;
}
// nsteps implicit-E steps of the 2D variant (l0 >= 3, m0 >= 2); es =
// if_implicitES.  Returns the largest count of zero-determinant cells left in
// a step, or -1 for a configuration the variant does not run.
int oo_step_implicit_2d(void* o, double h, int nsteps, int es) {
  Oracle* q = static_cast<Oracle*>(o);
  if (q->c.l0 < 3 || q->c.m0 < 2) return -1;
  int worst = 0;
  for (int s = 0; s < nsteps; ++s) worst = std::max(worst, q->step_implicit_2d(h, es != 0));
  return worst;
}
// JN of the electrostatic 2D variant: jn = complex [3][ny][nx] (x, y, z)
void oo_set_es_current(void* o, const double* jn) {
  Oracle* q = static_cast<Oracle*>(o);
  const cd* j = reinterpret_cast<const cd*>(jn);
  for (int k = 0; k < 3; ++k) q->es_jn[k].assign(j + (size_t)k * q->nxy, j + (size_t)(k + 1) * q->nxy);
}
// Tangent-linear model of nsteps steps of a 1D m0 = 0 run from t = 0.
// f, df: real [l][x][p]; fld, dfld: real [6][x]; outputs the final state and
// its tangent (df_out/dfld_out may be null: the double instantiation only).
int oo_tlm_1d(void* o, const double* f, const double* fld, const double* df, const double* dfld, double h,
              int nsteps, double* f_out, double* fld_out, double* df_out, double* dfld_out) {
  const Oracle* q = static_cast<Oracle*>(o);
  if (q->c.m0 != 0 || q->ny != 1 || q->c.l0 < 1) return -1;
  const size_t nf = (size_t)(q->c.l0 + 1) * q->nx * q->np, ng = (size_t)6 * q->nx;
  double clk[3] = {0.0, 0.0, 0.0};
  if (!df_out) {
    Step1D<double> S(*q);
    std::vector<double> F(f, f + nf), G(fld, fld + ng);
    for (int s = 0; s < nsteps; ++s) S.step(F, G, h, clk);
    std::copy(F.begin(), F.end(), f_out);
    std::copy(G.begin(), G.end(), fld_out);
    return 0;
  }
  Step1D<Dual> S(*q);
  std::vector<Dual> F(nf), G(ng);
  for (size_t i = 0; i < nf; ++i) F[i] = Dual(f[i], df[i]);
  for (size_t i = 0; i < ng; ++i) G[i] = Dual(fld[i], dfld[i]);
  for (int s = 0; s < nsteps; ++s) S.step(F, G, h, clk);
  for (size_t i = 0; i < nf; ++i) { f_out[i] = F[i].v; df_out[i] = F[i].d; }
  for (size_t i = 0; i < ng; ++i) { fld_out[i] = G[i].v; dfld_out[i] = G[i].d; }
  return 0;
}
// Single-column pieces for closed-form tests.
void oo_ee_getslope(void* o, const double* fin, double* fh) { static_cast<Oracle*>(o)->getslope(fin, fh); }
void oo_flm_solve(void* o, const double* f00, double Dt, int el, double* col_re_im) {
  Oracle* q = static_cast<Oracle*>(o);
  Oracle::Coef k;
  q->reset_coeff(f00, Dt, k);
  q->advance_flm(k, reinterpret_cast<cd*>(col_re_im), el);
}
void oo_flm_coeffs(void* o, const double* f00, double Dt, double* sub, double* diag, double* sup, double* scatt) {
  Oracle* q = static_cast<Oracle*>(o);
  Oracle::Coef k;
  q->reset_coeff(f00, Dt, k);
  for (int i = 0; i < q->np; ++i) { sub[i] = k.sub[i]; diag[i] = k.diag[i]; sup[i] = k.sup[i]; scatt[i] = k.scatt[i]; }
}
void oo_tridiag(int n, const double* a, const double* b, const double* c, const double* d, double* x) {
  std::vector<double> cc(c, c + n), dd(d, d + n);
  cc[0] /= b[0]; dd[0] /= b[0];
  for (int i = 1; i < n; ++i) {
    double id = 1.0 / (b[i] - cc[i - 1] * a[i]);
    cc[i] *= id; dd[i] -= dd[i - 1] * a[i]; dd[i] *= id;
  }
  x[n - 1] = dd[n - 1];
  for (int i = n - 2; i > -1; --i) x[i] = dd[i] - cc[i] * x[i + 1];
}
// Axes used by the restatement: pd (double axis), pc (complex axis, real part).
void oo_axes(void* o, double* pd, double* pc, double* vr_c) {
  Oracle* q = static_cast<Oracle*>(o);
  for (int i = 0; i < q->np; ++i) { pd[i] = q->pd[i]; pc[i] = q->pc[i].real(); vr_c[i] = q->vr_c[i]; }
}

}  // extern "C"
