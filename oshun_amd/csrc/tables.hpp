// Host-side construction of every coefficient table and of the per-target
// coupling program used by the Vlasov stage kernels.
//
// All tables are built in double on the host with the reference's own
// formulas and operation order (file:line cited per table), then uploaded;
// the kernels only multiply and add.  The coupling program lists, for each
// real target plane, the ordered terms that sum into its slope.  The order
// follows the reference's accumulation order into Yh (def-Actions.cpp:79-92):
// E-field terms by source (m outer, l inner), x-advection, y-advection, then
// B-field rotation.
#pragma once

#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>

namespace oshun {

typedef std::complex<double> cd;

enum Op { OP_G = 0, OP_H = 1, OP_G00 = 2, OP_DX = 3, OP_DY = 4, OP_ID = 5 };
enum Mul { MUL_EX = 0, MUL_EY, MUL_EZ, MUL_BX, MUL_BY, MUL_BZ, MUL_V = 6 };

struct Term {
  int src;   // source real plane
  int op;    // Op
  int l;     // harmonic l of the source (G/H tables)
  int mul;   // Mul: field component or v(p)
  double c;  // real coefficient (sign included)
};

inline int num_harmonics(int l0, int m0) { return ((m0 + 1) * (2 * l0 - m0 + 2)) / 2; }
// ind(l,m), def-State.cpp:422-426
inline int hind(int l, int m, int m0) {
  return (l < m0 + 1) ? (l * (l + 1)) / 2 + m : l * (m0 + 1) - (m0 * (m0 + 1)) / 2 + m;
}

struct HostTables {
  int l0, m0, np, nh, nc;
  bool two_d;
  // planes: component index of (h, re/im); -1 when absent
  std::vector<int> comp_re, comp_im;  // per harmonic
  std::vector<int> comp_l, comp_m, comp_isim;
  // axes
  std::vector<double> pd;  // Axis<double>
  double dpd;
  std::vector<cd> pc;      // complex axis of the Vlasov operators
  cd dpc;
  // Vlasov tables
  std::vector<double> invpax;  // (l0+1) x np : invpr * (-2(l+1) dp)
  std::vector<double> cg, hp0; // l0+1
  std::vector<double> vtab;    // v = p / sqrt(1+p^2) on the complex axis
  std::vector<double> p3og;    // p^3/gamma
  double p0p1_sq, inv_mp0p1_sq, g_r;
  double small, delta_p;
  double midx, midy;  // MaxwellEq -1/(2 dx), -1/(2 dy)
  // 1D m0 = 0 coefficient rows (per source l): E-field G / H, x-advection to l+1 / l-1
  std::vector<double> e1d_A1, e1d_A2, x1d_A1, x1d_A2;
  std::vector<Term> terms;
  std::vector<int> tstart;  // nc+1
  std::vector<int> clean;   // per plane: rows p < clean are zeroed (CLEAN_LOW_PC)
  // collision tables (double axis), def-FokkerPlanck.cpp:135-173,785-803
  std::vector<double> vc, U4, U4m1, U2, U2m1, U1s, U1m1s, U3, Qn, Pn;
  std::vector<double> ivdm, ivdp, ivs2;  // tridiagonal geometric factors
  std::vector<double> pw3, pw5, pw7;     // pow(v,3|5|7) for the parabolic G
  // moment weights on the double axis (def-Parallel.cpp:1699-2134): p^2, p^3/g, p^4/g^2, p^5/g^3
  std::vector<double> mw2, mw3, mw4, mw5;
  double m_r, m_inv, m_c5, m_c7, m_c04, m_c00;
  double c_kpre, kpre, v0sq, v1sq, g_den;
};

inline void build_tables(HostTables& T, int l0, int m0, int np, int ny, double pmin, double pmax, double dx,
                         double dy, double density_np, int nb, int ref_quirks) {
  T.l0 = l0; T.m0 = m0; T.np = np;
  T.nh = num_harmonics(l0, m0);
  T.two_d = ny > 1;
  // ---- planes
  T.comp_re.assign(T.nh, -1); T.comp_im.assign(T.nh, -1);
  T.comp_l.clear(); T.comp_m.clear(); T.comp_isim.clear();
  int nc = 0;
  for (int l = 0; l <= l0; ++l)
    for (int m = 0; m <= std::min(l, m0); ++m) {
      int h = hind(l, m, m0);
      (void)h;
    }
  // keep harmonic order ind(l,m) for planes
  std::vector<int> hl(T.nh), hm(T.nh);
  for (int l = 0; l <= l0; ++l)
    for (int m = 0; m <= std::min(l, m0); ++m) { hl[hind(l, m, m0)] = l; hm[hind(l, m, m0)] = m; }
  for (int h = 0; h < T.nh; ++h) {
    T.comp_re[h] = nc++;
    T.comp_l.push_back(hl[h]); T.comp_m.push_back(hm[h]); T.comp_isim.push_back(0);
    if (hm[h] > 0) {
      T.comp_im[h] = nc++;
      T.comp_l.push_back(hl[h]); T.comp_m.push_back(hm[h]); T.comp_isim.push_back(1);
    }
  }
  T.nc = nc;

  // ---- axes, matrices.h:860-866 and def-VlasovMax.cpp:88-90
  T.pd.resize(np);
  {
    double d = (pmax - pmin) / static_cast<double>(np - 1);
    for (int i = 0; i < np; ++i) { double v = i; v *= d; v += pmin; T.pd[i] = v; }
  }
  T.dpd = T.pd[1] - T.pd[0];
  T.pc.resize(np);
  {
    cd p0(T.pd[0]), pN(T.pd[np - 1]);
    cd dc = (pN - p0) / static_cast<cd>(np - 1);
    for (int i = 0; i < np; ++i) { cd v(i); v *= dc; v += p0; T.pc[i] = v; }
  }
  T.dpc = T.pc[1] - T.pc[0];
  const cd dpc = T.dpc;
  const auto& pc = T.pc;

  // ---- ElectricField tables, def-VlasovMax.cpp:97-151
  const int L = l0, M = m0;
  auto LI = [&](int l, int m) { return l * (M + 1) + m; };
  std::vector<cd> eA1((L + 1) * (M + 1)), eA2((L + 1) * (M + 1)), eC2((L + 1) * (M + 1)), eC4((L + 1) * (M + 1));
  std::vector<cd> eB1(L + 1), eB2(L + 1), eC1(L + 1), eC3(L + 1), Hp0(L + 1);
  cd lc, mc;
  for (int l = 1; l <= L; ++l)
    for (int m = 0; m <= std::min(M, l); ++m) {
      lc = cd(l); mc = cd(m);
      eA1[LI(l, m)] = (lc + 1.0 - mc) * lc / ((2.0 * lc + 1.0) * (lc + 1.0));
      eA2[LI(l, m)] = -(lc + mc) / (2.0 * lc + 1.0);
    }
  eA1[LI(0, 0)] = -1.0;
  {
    cd s = 0.5 / dpc;
    for (auto& v : eA1) v *= s;
    for (auto& v : eA2) v *= s;
  }
  for (int l = 1; l <= L; ++l) {
    lc = cd(l);
    eB1[l] = (-0.5 / dpc) * lc * lc / (2.0 * lc + 1.0);
    eB2[l] = (-0.5 / dpc) * lc * (lc + 1.0) / (2.0 * lc + 1.0);
    eC1[l] = (0.25 / dpc) * lc / ((2.0 * lc + 1.0) * (lc + 1.0));
    eC3[l] = (0.25 / dpc) / (2.0 * lc + 1.0);
  }
  eC1[0] = -0.25 / dpc;
  for (int l = 2; l <= L; ++l)
    for (int m = 2; m <= std::min(M, l); ++m) {
      lc = cd(l); mc = cd(m);
      eC2[LI(l, m)] = (-0.25 / dpc) * lc * (lc - mc + 2.0) * (lc - mc + 1.0) / ((2.0 * lc + 1.0) * (lc + 1.0));
      eC4[LI(l, m)] = (-0.25 / dpc) * (lc + mc - 1.0) * (lc + mc) / (2.0 * lc + 1.0);
    }
  for (int l = 0; l <= L; ++l) Hp0[l] = 1.0 / pc[0];
  for (int l = 1; l <= L; ++l) Hp0[l] = Hp0[l - 1] * (pc[0] / pc[1]) * (2.0 * l + 1.0) / (2.0 * l - 1);
  for (int l = 0; l <= L; ++l) Hp0[l] *= -2.0 * dpc;
  T.hp0.resize(L + 1);
  T.cg.assign(L + 1, 0.0);
  T.invpax.assign((size_t)(L + 1) * np, 0.0);
  for (int l = 0; l <= L; ++l) {
    T.hp0[l] = Hp0[l].real();
    lc = cd(l);
    if (l >= 1) T.cg[l] = (-(2.0 * lc + 1.0) / lc).real();
    cd s = -2.0 * (lc + 1.0) * dpc;  // MakeGH, def-VlasovMax.cpp:502
    for (int i = 0; i < np; ++i) T.invpax[(size_t)l * np + i] = ((1.0 / pc[i]) * s).real();
  }
  {  // G00, def-VlasovMax.cpp:523-526
    cd p0p1_sq(pc[0] * pc[0] / (pc[1] * pc[1]));
    cd inv(1.0 / (1.0 - p0p1_sq));
    cd g_r = -4.0 * dpc * pc[0] / (pc[1] * pc[1]);
    T.p0p1_sq = p0p1_sq.real(); T.inv_mp0p1_sq = inv.real(); T.g_r = g_r.real();
  }
  // ---- MagneticField tables, def-VlasovMax.cpp:563-589 (A1 = -i m handled below)
  std::vector<double> bA2((L + 1) * (M + 1), 0.0), bB1(L + 1);
  for (int l = 1; l <= L; ++l)
    for (int m = 0; m <= std::min(M, l); ++m) {
      lc = cd(l); mc = cd(m);
      bA2[LI(l, m)] = ((-0.5) * (lc + 1.0 - mc) * (lc + mc)).real();
    }
  for (int l = 0; l <= L; ++l) { lc = cd(l); bB1[l] = ((-1.0) * lc * (lc + 1.0)).real(); }
  // ---- Current, def-VlasovMax.cpp:664-674
  T.p3og.resize(np);
  for (int i = 0; i < np; ++i) T.p3og[i] = ((pc[i] * pc[i]) * (pc[i] / std::sqrt(1.0 + pc[i] * pc[i]))).real();
  T.delta_p = T.dpd;
  {
    double s = T.pd[0];
    s *= s; s *= s; s *= T.pd[0];
    s *= 0.2; s *= 1.0 / T.pd[1];
    T.small = s;
  }
  // ---- Spatial_Advection, def-VlasovMax.cpp:768-826
  T.vtab.resize(np);
  for (int i = 0; i < np; ++i) T.vtab[i] = (pc[i] / (std::sqrt(1.0 + pc[i] * pc[i]))).real();
  cd idx(dx), idy(dy);
  idx = -1.0 / (2.0 * idx);
  idy = -1.0 / (2.0 * idy);
  std::vector<cd> aA1((L + 1) * (M + 1)), aA2((L + 1) * (M + 1)), aC2((L + 1) * (M + 1)), aC4((L + 1) * (M + 1));
  std::vector<cd> aB1(L + 1), aB2(L + 1), aC1(L + 1), aC3(L + 1);
  for (int l = 0; l <= L; ++l)
    for (int m = 0; m <= std::min(M, l); ++m) {
      lc = cd(l); mc = cd(m);
      aA1[LI(l, m)] = idx * (-1.0) * (lc - mc + 1.0) / (2.0 * lc + 1.0);
      aA2[LI(l, m)] = idx * (-1.0) * (lc + mc) / (2.0 * lc + 1.0);
      aC2[LI(l, m)] = idy * 0.5 * (lc + 2.0 - mc) * (lc - mc + 1.0) / (2.0 * lc + 1.0);
      aC4[LI(l, m)] = idy * (-0.5) * (lc + mc - 1.0) * (lc + mc) / (2.0 * lc + 1.0);
    }
  for (int l = 0; l <= L; ++l) {
    lc = cd(l);
    aB1[l] = idy * (lc + 1.0) * lc / (2.0 * lc + 1.0);
    aB2[l] = (-1.0) * aB1[l];
    aC1[l] = (-0.5) * idy / (2.0 * lc + 1.0);
    aC3[l] = (-1.0) * aC1[l];
  }
  {  // MaxwellEq, def-VlasovMax.cpp:967-973
    cd a(dx), b(dy);
    a = -1.0 / (2.0 * a);
    b = -1.0 / (2.0 * b);
    T.midx = a.real(); T.midy = b.real();
  }

  T.e1d_A1.assign(L + 1, 0.0); T.e1d_A2.assign(L + 1, 0.0);
  T.x1d_A1.assign(L + 1, 0.0); T.x1d_A2.assign(L + 1, 0.0);
  for (int l = 0; l <= L; ++l) {
    T.e1d_A1[l] = eA1[LI(l, 0)].real();
    T.e1d_A2[l] = eA2[LI(l, 0)].real();
    T.x1d_A1[l] = aA1[LI(l, 0)].real();
    T.x1d_A2[l] = aA2[LI(l, 0)].real();
  }

  // ---- coupling program
  auto valid = [&](int l, int m) { return l >= 0 && l <= L && m >= 0 && m <= M && m <= l; };
  std::vector<std::vector<Term>> prog(nc);
  auto RE = [&](int l, int m) { return T.comp_re[hind(l, m, M)]; };
  auto IM = [&](int l, int m) { return T.comp_im[hind(l, m, M)]; };
  // target += OP(src) * M with M = sum_k coef_k * field_k (complex source a+ib,
  // complex multiplier Mr + i Mi):  re += OP(a) Mr - OP(b) Mi ; im += OP(a) Mi + OP(b) Mr
  struct FieldLin { int mul; double c; };
  auto add_complex = [&](int tl, int tm, int sl, int sm, int op, int lsrc, const std::vector<FieldLin>& Mr,
                         const std::vector<FieldLin>& Mi, bool re_only) {
    int ta = RE(tl, tm), tb = (tm > 0 && !re_only) ? IM(tl, tm) : -1;
    int sa = RE(sl, sm), sb = sm > 0 ? IM(sl, sm) : -1;
    for (const auto& q : Mr) if (q.c != 0.0) prog[ta].push_back({sa, op, lsrc, q.mul, q.c});
    if (sb >= 0) for (const auto& q : Mi) if (q.c != 0.0) prog[ta].push_back({sb, op, lsrc, q.mul, -q.c});
    if (tb >= 0) {
      for (const auto& q : Mi) if (q.c != 0.0) prog[tb].push_back({sa, op, lsrc, q.mul, q.c});
      if (sb >= 0) for (const auto& q : Mr) if (q.c != 0.0) prog[tb].push_back({sb, op, lsrc, q.mul, q.c});
    }
  };
  // field combinations (times a real coefficient c):
  //   Ex c        -> Mr = c Ex
  //   Em c        -> Mr = c Ey, Mi = -c Ez   (Em = Ey - i Ez, def-VlasovMax.cpp:352)
  //   Ep c        -> Mr = c Ey, Mi =  c Ez   (Ep = Ey + i Ez, :353)
  auto e_term = [&](int tl, int tm, int sl, int sm, int op, int which, double c, bool re_only) {
    if (!valid(tl, tm)) return;
    if (which == 0) add_complex(tl, tm, sl, sm, op, sl, {{MUL_EX, c}}, {}, re_only);
    else if (which == 1) add_complex(tl, tm, sl, sm, op, sl, {{MUL_EY, c}}, {{MUL_EZ, -c}}, re_only);
    else add_complex(tl, tm, sl, sm, op, sl, {{MUL_EY, c}}, {{MUL_EZ, c}}, re_only);
  };
  // ElectricField::Exyz, def-VlasovMax.cpp:346-491
  for (int m = 0; m <= M; ++m)
    for (int l = m; l <= L; ++l) {
      if (l == 0) {  // G00 feeds (1,0) and (1,1)
        if (L >= 1) {
          e_term(1, 0, 0, 0, OP_G00, 0, eA1[LI(0, 0)].real(), false);
          if (M >= 1) e_term(1, 1, 0, 0, OP_G00, 1, eC1[0].real(), false);
        }
        continue;
      }
      // H of (l,m) -> l-1
      e_term(l - 1, m, l, m, OP_H, 0, eA2[LI(l, m)].real(), false);
      if (m + 1 <= M) e_term(l - 1, m + 1, l, m, OP_H, 1, eC3[l].real(), false);
      if (m == 1) e_term(l - 1, 0, l, m, OP_H, 2, eB2[l].real(), true);
      else if (m >= 2) e_term(l - 1, m - 1, l, m, OP_H, 2, eC4[LI(l, m)].real(), false);
      // G of (l,m) -> l+1
      if (l + 1 <= L) {
        e_term(l + 1, m, l, m, OP_G, 0, eA1[LI(l, m)].real(), false);
        if (m + 1 <= M) e_term(l + 1, m + 1, l, m, OP_G, 1, eC1[l].real(), false);
        if (m == 1) e_term(l + 1, 0, l, m, OP_G, 2, eB1[l].real(), true);
        else if (m >= 2) e_term(l + 1, m - 1, l, m, OP_G, 2, eC2[LI(l, m)].real(), false);
      }
    }
  // Spatial_Advection::Axy, x part (def-VlasovMax.cpp:842-853): real v(p) multiplier
  auto a_term = [&](int tl, int tm, int sl, int sm, int op, double c, bool re_only) {
    if (!valid(tl, tm)) return;
    add_complex(tl, tm, sl, sm, op, sl, {{MUL_V, c}}, {}, re_only);
  };
  for (int m = 0; m <= M; ++m)
    for (int l = m; l <= L; ++l) {
      if (valid(l - 1, m)) a_term(l - 1, m, l, m, OP_DX, aA2[LI(l, m)].real(), false);
      if (l + 1 <= L) a_term(l + 1, m, l, m, OP_DX, aA1[LI(l, m)].real(), false);
    }
  // y part (def-VlasovMax.cpp:857-947); Dy of a 1D (ny = 1) state is exactly 0
  if (T.two_d && M >= 1) {
    cd errf(1.0);
    for (int m = 0; m <= M; ++m) {
      cd err_loop(1.0);
      for (int l = m; l <= L; ++l) {
        cd q = errf;
        if (ref_quirks && m >= 2 && m < M && m + 1 < L) {
          // def-VlasovMax.cpp:918 divides by C1[m+1] where C1[l-1] is meant
          if (l >= m + 3 && l < L) err_loop *= aC1[l - 1] / aC1[m + 1];
          if (l >= m + 3) q *= err_loop;
        }
        if (m + 1 <= M) {
          a_term(l - 1, m + 1, l, m, OP_DY, (aC3[l] * q).real(), false);
          if (l + 1 <= L) a_term(l + 1, m + 1, l, m, OP_DY, (aC1[l] * q).real(), false);
        }
        if (m == 1) {
          a_term(l - 1, 0, l, m, OP_DY, (aB2[l] * q).real(), true);
          if (l + 1 <= L) a_term(l + 1, 0, l, m, OP_DY, (aB1[l] * q).real(), true);
        } else if (m >= 2) {
          a_term(l - 1, m - 1, l, m, OP_DY, (aC4[LI(l, m)] * q).real(), false);
          if (l + 1 <= L) a_term(l + 1, m - 1, l, m, OP_DY, (aC2[LI(l, m)] * q).real(), false);
        }
      }
      if (ref_quirks && m >= 2 && m < M && m + 1 < L) errf *= err_loop;
    }
  }
  // MagneticField::Bxyz, def-VlasovMax.cpp:597-644 (pointwise)
  //   Bp/2 = (Bz + i By)/2, Bx (-i m), Bm c = (Bz - i By) c
  if (M >= 1) {
    for (int m = 0; m <= M; ++m)
      for (int l = std::max(m, 1); l <= L; ++l) {
        if (m + 1 <= M && m + 1 <= l)
          add_complex(l, m + 1, l, m, OP_ID, l, {{MUL_BZ, 0.5}}, {{MUL_BY, 0.5}}, false);
        if (m >= 1) {
          add_complex(l, m, l, m, OP_ID, l, {}, {{MUL_BX, -static_cast<double>(m)}}, false);
          if (m == 1) add_complex(l, 0, l, m, OP_ID, l, {{MUL_BZ, bB1[l]}}, {{MUL_BY, -bB1[l]}}, true);
          else add_complex(l, m - 1, l, m, OP_ID, l, {{MUL_BZ, bA2[LI(l, m)]}}, {{MUL_BY, -bA2[LI(l, m)]}}, false);
        }
      }
  }
  T.terms.clear();
  T.tstart.assign(nc + 1, 0);
  for (int c = 0; c < nc; ++c) {
    T.tstart[c] = static_cast<int>(T.terms.size());
    T.terms.insert(T.terms.end(), prog[c].begin(), prog[c].end());
  }
  T.tstart[nc] = static_cast<int>(T.terms.size());
  // CLEAN_LOW_PC, def-Actions.cpp:240-253
  T.clean.assign(nc, 0);
  for (int c = 0; c < nc; ++c) {
    int l = T.comp_l[c];
    T.clean[c] = l >= 1 ? std::min(l - 1, 80) : 0;
  }

  // ---- collision tables, def-FokkerPlanck.cpp:135-173, 785-803, 886-895
  const double re = 2.8179402894e-13;
  double kp = std::sqrt(4.0 * M_PI * (density_np) * re);
  T.c_kpre = 4.0 * M_PI / 3.0 * re * kp;
  T.kpre = re * kp;
  auto& v = T.vc;
  v.assign(np, 0.0);
  for (int i = 0; i < np; ++i) { v[i] = T.pd[i]; v[i] = v[i] / (std::sqrt(1.0 + v[i] * v[i])); }
  T.U4.assign(np, 0.0); T.U4m1.assign(np, 0.0); T.U2.assign(np, 0.0); T.U2m1.assign(np, 0.0);
  T.U1s.assign(np, 0.0); T.U1m1s.assign(np, 0.0); T.U3.assign(np, 0.0); T.Qn.assign(np, 0.0); T.Pn.assign(np, 0.0);
  std::vector<double> U1(np, 0.0), U1m1(np, 0.0);
  for (int i = 1; i < np; ++i) {
    T.U4[i] = 0.5 * std::pow(v[i], 4) * (v[i] - v[i - 1]);
    T.U4m1[i] = 0.5 * std::pow(v[i - 1], 4) * (v[i] - v[i - 1]);
    T.U2[i] = 0.5 * v[i] * v[i] * (v[i] - v[i - 1]);
    T.U2m1[i] = 0.5 * v[i - 1] * v[i - 1] * (v[i] - v[i - 1]);
    U1[i] = 0.5 * v[i] * (v[i] - v[i - 1]);
    U1m1[i] = 0.5 * v[i - 1] * (v[i] - v[i - 1]);
  }
  for (int i = 0; i + 1 < np; ++i) { T.U1s[i] = U1[i + 1]; T.U1m1s[i] = U1m1[i + 1]; }
  for (int i = 0; i < np; ++i) T.U3[i] = std::pow(v[i], 3);
  T.Qn[0] = 1.0 / ((v[0] * v[0] * v[1]) / 2.0);
  for (int i = 1; i < np - 1; ++i) T.Qn[i] = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]) / 2.0);
  for (int i = 0; i < np - 1; ++i) T.Pn[i] = 1.0 / ((v[i + 1] - v[i]) / 2.0 * (v[i + 1] + v[i]));
  T.ivdm.assign(np, 0.0); T.ivdp.assign(np, 0.0); T.ivs2.assign(np, 0.0);
  for (int i = 1; i < np - 1; ++i) {
    T.ivdm[i] = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i] - v[i - 1]));
    T.ivdp[i] = 1.0 / (v[i] * 0.5 * (v[i + 1] - v[i - 1]) * (v[i + 1] - v[i]));
    T.ivs2[i] = 1.0 / (v[i] * v[i] * (v[i + 1] - v[i - 1]));
  }
  T.pw3.assign(np, 0.0); T.pw5.assign(np, 0.0); T.pw7.assign(np, 0.0);
  for (int i = 0; i < std::min(nb, np); ++i) {
    T.pw3[i] = std::pow(v[i], 3); T.pw5[i] = std::pow(v[i], 5); T.pw7[i] = std::pow(v[i], 7);
  }
  {
    const auto& pr = T.pd;
    T.mw2.resize(np); T.mw3.resize(np); T.mw4.resize(np); T.mw5.resize(np);
    for (int i = 0; i < np; ++i) {
      const double ig = 1.0 / std::sqrt(1.0 + pr[i] * pr[i]);
      T.mw2[i] = pr[i] * pr[i];
      T.mw3[i] = std::pow(pr[i], 3) * ig;
      T.mw4[i] = std::pow(pr[i], 4) * std::pow(ig, 2);
      T.mw5[i] = std::pow(pr[i], 5) * std::pow(ig, 3);
    }
    const double dp = T.dpd;
    T.m_r = pr[0] * pr[0] / (pr[1] * pr[1]);
    T.m_inv = 1.0 / (1.0 - T.m_r);
    T.m_c5 = std::pow(pr[0], 5) / (dp * pr[1] * 5.0);
    T.m_c7 = std::pow(pr[0], 7) / (dp * pr[1] * 7.0);
    T.m_c00 = pr[0] * (pr[0] / dp) * pr[0];
    T.m_c04 = std::pow(pr[0], 4) * (pr[0] / dp);
  }
  T.v0sq = v[0] * v[0];
  T.v1sq = v[1] * v[1];
  T.g_den = 1.0 - (v[0] * v[0]) / (v[1] * v[1]);
}

// Coupling programs of the implicit-E step's actions, taken from the full
// program above (same per-target term order, same coefficients):
//   kProgAdv1    Implicit_Actions::Adv1 (def-Actions.cpp:118-132): x- and
//                y-advection, and the B-field rotation unless es; no
//                CLEAN_LOW_PC
//   kProgAdv2    Adv2 without if_implicit1D (:138-160): the E-field
//                couplings of Exyz, with CLEAN_LOW_PC
//   kProgProbe   Implicit_Ex and Implicit_Eyz (def-VlasovMax.cpp:157-196,
//                231-291): the E-field couplings with sources l <= 3 and
//                targets l <= 2, Ex only from sources m <= 1.  Run with one
//                field component nonzero (Ex, Ey or Ez), it is ImpEx, ImpEy
//                (Em = Ep = Ey) or ImpEz (Em = -i Ez, Ep = i Ez).
enum { kProgAdv1 = 0, kProgAdv2 = 1, kProgProbe = 2 };
struct Program {
  std::vector<Term> terms;
  std::vector<int> tstart, clean;
};
inline Program derive_program(const HostTables& T, int kind, bool es) {
  Program P;
  P.tstart.assign(T.nc + 1, 0);
  P.clean.assign(T.nc, 0);
  for (int tc = 0; tc < T.nc; ++tc) {
    P.tstart[tc] = static_cast<int>(P.terms.size());
    for (int k = T.tstart[tc]; k < T.tstart[tc + 1]; ++k) {
      const Term& t = T.terms[k];
      const bool efield = t.op == OP_G || t.op == OP_H || t.op == OP_G00;
      bool keep;
      if (kind == kProgAdv1) keep = t.op == OP_DX || t.op == OP_DY || (t.op == OP_ID && !es);
      else if (kind == kProgAdv2) keep = efield;
      else keep = efield && T.comp_l[t.src] <= 3 && T.comp_l[tc] <= 2 && (t.mul != MUL_EX || T.comp_m[t.src] <= 1);
      if (keep) P.terms.push_back(t);
    }
    if (kind == kProgAdv2) P.clean[tc] = T.clean[tc];
  }
  P.tstart[T.nc] = static_cast<int>(P.terms.size());
  return P;
}

}  // namespace oshun
