// Test driver for the UNMODIFIED reference classes of OSHUN-1.
//
// TEST INFRASTRUCTURE ONLY (oracle side).  It is compiled by oracle/Makefile
// together with nine reference translation units taken as they lie under
// the reference tree (def-Input, def-Nmethods, def-State, def-VlasovMax,
// def-Actions, def-RK, def-FokkerPlanck, def-ImplicitE, def-PLSource); none of
// them uses MPI.  Nothing of
// the reference is copied into this repository; the binary lands in
// oracle/_ref/, which git ignores.
//
// The reference is configured through its own input singleton: the fields
// of Inputdata::IN().list() are set at run time and the derived structs
// IN().inp() / IN().cont() are rebuilt from it (decl-input.h:56-190,
// def-Input.cpp:313-334), so no deck has to be edited or recompiled.
//
// Guard cells: the reference keeps RKLevel = 4 guard cells and refreshes them
// after every full step (main.cpp:107, def-Parallel.cpp:1138-1156,1229-1248).
// This driver fills the guards with periodic images before the first step
// and after each step, which is what the single-node periodic exchange does.
//
// Usage:  oshun_ref <in.bin> <out.bin>
//   in.bin : int32[13] {l0,m0,np,nx,ny,op,nsteps,nb,es,src,bndX,bndY,nbc}  (es: if_implicitES, op 5/9;
//            bndX/bndY: 0 periodic, 1 mirror guards, def-Parallel.cpp:1121-1307;
//            nbc: guard cells per side = RKLevel, 0 for the default 4)
//            double[9] {pmax,xmin,xmax,ymin,ymax,h,small_dt,density_np,zeta}
//            if src: double[15] {linear_Ivst, polynomial_Ivst, inverse_bremsstrahlung,
//                   polarization_direction, I_0, lambda_0, ab_eta, ab_depth, cntr_x0,
//                   cntr_r0, spot_w0, rise_time, flat_time, fall_time, smaller_dt}
//            complex<double>[nh][ny][nx][np]  (interior f, p fastest)
//            complex<double>[6][ny][nx]       (interior Ex..Bz)
//   op     : 0 = Actions::Adv slope of the input state
//            1 = nsteps x {RK4, e-e, f1, flm, halo}     (main.cpp:79-107)
//            2 = nsteps x {RK4, halo}
//            3 = nsteps x {e-e}       (tnew advanced by h each time)
//            4 = nsteps x {f1, flm}   (tnew advanced by h each time)
//            5 = nsteps x {RK2 advance_p1, e-e, implicit E, RK2 advance_p2,
//                f1, flm, halo} with implicit_E and if_implicit1D on
//                (main.cpp:79-107 with Runge_Kutta_2_Imp and Implicit_E_Field)
//            9 = the same with if_implicit1D off (the 2D variant: Implicit_Ex,
//                Implicit_Eyz probes and the 3x3 conductivity solve)
//            ops 1, 5 and 9 also run PLaserSource::Laser_Source and InverseBremsstrahlung::loop
//            after the e-e collisions when the deck turns them on (main.cpp:90-92)
//            6 = nsteps x {PLaserSource::Laser_Source}   (tnew advanced by h each time)
//            7 = nsteps x {InverseBremsstrahlung::loop}  (same)
//            8 = nsteps x {ExternalEfield::Laser_Efield} (same)
//   out.bin: double[4] {dx_used, dy_used, h_used, t_final},
//            double[np] p axis, then f and fields as in in.bin; with a mirror
//            boundary also the whole arrays, guards included, of the initial
//            state (after its first guard fill) and of the final state:
//            complex<double>[nh][ny+2g][nx+2g][np], complex<double>[6][ny+2g][nx+2g].

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <valarray>
#include <vector>

#include "matrices.h"
#include "decl-input.h"
#include "decl-state.h"
#include "decl-vlasovmax.h"
#include "decl-actions.h"
#include "decl-RK.h"
#include "decl-fokkerplanck.h"
#include "decl-implicitE.h"
#include "decl-plsource.h"

typedef std::complex<double> cd;

static void die(const char* m) {
  std::fprintf(stderr, "oshun_ref: %s\n", m);
  std::exit(2);
}

static void fill_guards_periodic(Stat& Y, int nbc) {
  // Guard g on the left holds interior cell (g - nbc) mod n, on the right
  // interior cell g mod n.  For n >= nbc this is exactly the slab copy of
  // sameNode_periodic_X/Y; for n < nbc that copy would read guard cells
  // (def-Parallel.cpp:1143-1153), so the images are taken modulo n here.
  const int nh = Y.DF().dim();
  const int npx = Y.SH(0).nump(), nxg = Y.SH(0).numx(), nyg = Y.SH(0).numy();
  const int nx = nxg - 2 * nbc, ny = nyg - 2 * nbc;
  auto lx = [&](int g) { return nbc + (((g - nbc) % nx) + nx) % nx; };
  auto rx = [&](int g) { return nbc + g % nx; };
  auto ly = [&](int g) { return nbc + (((g - nbc) % ny) + ny) % ny; };
  auto ry = [&](int g) { return nbc + g % ny; };
  for (int i = 0; i < nh; ++i) {
    SHarmonic& s = Y.SH(i);
    for (int iy = nbc; iy < nbc + ny; ++iy)
      for (int g = 0; g < nbc; ++g)
        for (int ip = 0; ip < npx; ++ip) {
          s(ip, g, iy) = s(ip, lx(g), iy);
          s(ip, nbc + nx + g, iy) = s(ip, rx(g), iy);
        }
    for (int ix = 0; ix < nxg; ++ix)
      for (int g = 0; g < nbc; ++g)
        for (int ip = 0; ip < npx; ++ip) {
          s(ip, ix, g) = s(ip, ix, ly(g));
          s(ip, ix, nbc + ny + g) = s(ip, ix, ry(g));
        }
  }
  for (int i = 0; i < 6; ++i) {
    Field2D& f = Y.FLD(i);
    for (int iy = nbc; iy < nbc + ny; ++iy)
      for (int g = 0; g < nbc; ++g) {
        f(g, iy) = f(lx(g), iy);
        f(nbc + nx + g, iy) = f(rx(g), iy);
      }
    for (int ix = 0; ix < nxg; ++ix)
      for (int g = 0; g < nbc; ++g) {
        f(ix, g) = f(ix, ly(g));
        f(ix, nbc + ny + g) = f(ix, ry(g));
      }
  }
}

// Guard fill of a single node with a mirror boundary on at least one side:
// Node_Communications::sameNode_bound_X then sameNode_bound_Y (the order of
// Parallel_Environment::Neighbor_Communications, def-Parallel.cpp:3040-3127).
// Restated from sameNode_periodic_X/Y and sameNode_mirror_X/Y
// (def-Parallel.cpp:1137-1307), which live with the MPI code and are not
// built here.  As in the reference, the x mirror refreshes harmonics l < l0
// only and the y mirror 1 <= l < l0 only; the others keep their guard values.
// x mirror: guard at distance d from the face = interior cell at distance d,
// times (-1)^(l+m); Ey, Ez, Bx change sign.  y mirror: the same copy, then
// v -> (-1)^m conj(v) (as v -= 2 Re v; v *= -(-1)^m); Ex, Ez, By change sign.
static void fill_guards_bnd(Stat& Y, int nbc, int bndX, int bndY) {
  const int l0 = Y.DF().l0(), m0 = Y.DF().m0();
  const int npx = Y.SH(0).nump(), nxg = Y.SH(0).numx(), nyg = Y.SH(0).numy();
  // x direction, whole columns (all rows)
  auto xcopy_h = [&](SHarmonic& s, int dst, int src, double sg) {
    for (int iy = 0; iy < nyg; ++iy)
      for (int ip = 0; ip < npx; ++ip) s(ip, dst, iy) = s(ip, src, iy) * sg;
  };
  auto xcopy_f = [&](Field2D& f, int dst, int src, double sg) {
    for (int iy = 0; iy < nyg; ++iy) f(dst, iy) = f(src, iy) * sg;
  };
  for (int l = 0; l <= l0; ++l)
    for (int m = 0; m <= std::min(l, m0); ++m) {
      SHarmonic& s = Y.SH(l, m);
      if (bndX == 0) {
        for (int g = 0; g < nbc; ++g) {
          xcopy_h(s, g, nxg - 2 * nbc + g, 1.0);  // right bound -> left guard
          xcopy_h(s, nxg - nbc + g, nbc + g, 1.0);  // left bound -> right guard
        }
      } else if (l < l0) {
        const double sg = ((l + m) % 2) ? -1.0 : 1.0;
        for (int c = 0; c < nbc; ++c) xcopy_h(s, nxg - c - 1, nxg - 2 * nbc + c, sg);
        for (int c = 0; c < nbc; ++c) xcopy_h(s, c, 2 * nbc - c - 1, sg);
      }
    }
  for (int k = 0; k < 6; ++k) {
    Field2D& f = Y.FLD(k);
    const double sg = (bndX != 0 && (k == 1 || k == 2 || k == 3)) ? -1.0 : 1.0;
    for (int g = 0; g < nbc; ++g) {
      if (bndX == 0) {
        xcopy_f(f, g, nxg - 2 * nbc + g, 1.0);
        xcopy_f(f, nxg - nbc + g, nbc + g, 1.0);
      } else {
        xcopy_f(f, nxg - g - 1, nxg - 2 * nbc + g, sg);
        xcopy_f(f, g, 2 * nbc - g - 1, sg);
      }
    }
  }
  // y direction, whole rows (all columns)
  for (int l = 0; l <= l0; ++l)
    for (int m = 0; m <= std::min(l, m0); ++m) {
      SHarmonic& s = Y.SH(l, m);
      auto ycopy = [&](int dst, int src, bool mir, double sg) {
        for (int ix = 0; ix < nxg; ++ix)
          for (int ip = 0; ip < npx; ++ip) {
            cd v = s(ip, ix, src);
            if (mir) {
              const cd w = v;
              v -= 2.0 * w.real();
              v *= sg;
            }
            s(ip, ix, dst) = v;
          }
      };
      if (bndY == 0) {
        for (int g = 0; g < nbc; ++g) {
          ycopy(g, nyg - 2 * nbc + g, false, 1.0);
          ycopy(nyg - nbc + g, nbc + g, false, 1.0);
        }
      } else if (l >= 1 && l < l0) {
        const double sg = -((m % 2) ? -1.0 : 1.0);
        for (int c = 0; c < nbc; ++c) ycopy(nyg - c - 1, nyg - 2 * nbc + c, true, sg);
        for (int c = 0; c < nbc; ++c) ycopy(c, 2 * nbc - c - 1, true, sg);
      }
    }
  for (int k = 0; k < 6; ++k) {
    Field2D& f = Y.FLD(k);
    const double sg = (bndY != 0 && (k == 0 || k == 2 || k == 4)) ? -1.0 : 1.0;
    auto ycopy = [&](int dst, int src) {
      for (int ix = 0; ix < nxg; ++ix) f(ix, dst) = f(ix, src) * sg;
    };
    for (int g = 0; g < nbc; ++g) {
      if (bndY == 0) {
        for (int ix = 0; ix < nxg; ++ix) {
          f(ix, g) = f(ix, nyg - 2 * nbc + g);
          f(ix, nyg - nbc + g) = f(ix, nbc + g);
        }
      } else {
        ycopy(nyg - g - 1, nyg - 2 * nbc + g);
        ycopy(g, 2 * nbc - g - 1);
      }
    }
  }
}

static void write_ext(FILE* fo, Stat& Y) {
  const int nh = Y.DF().dim();
  for (int h = 0; h < nh; ++h) {
    const SHarmonic& s = Y.SH(h);
    const size_t n = (size_t)s.nump() * s.numx() * s.numy();
    std::fwrite(&Y.SH(h)(0, 0, 0), sizeof(cd), n, fo);
  }
  for (int k = 0; k < 6; ++k) {
    const size_t n = (size_t)Y.FLD(k).numx() * Y.FLD(k).numy();
    std::fwrite(&Y.FLD(k)(0, 0), sizeof(cd), n, fo);
  }
}

int main(int argc, char** argv) {
  if (argc != 3) die("usage: oshun_ref in.bin out.bin");
  FILE* fi = std::fopen(argv[1], "rb");
  if (!fi) die("cannot open input");
  int iv[13];
  double dv[9];
  if (std::fread(iv, sizeof(int), 13, fi) != 13) die("short header");
  if (std::fread(dv, sizeof(double), 9, fi) != 9) die("short header");
  const int l0 = iv[0], m0 = iv[1], np = iv[2], nx = iv[3], ny = iv[4], op = iv[5], nsteps = iv[6], nb = iv[7];
  const int es = iv[8];
  const int bndX = iv[10], bndY = iv[11];
  const bool mirror = bndX != 0 || bndY != 0;
  double sv[15] = {0};
  if (iv[9] && std::fread(sv, sizeof(double), 15, fi) != 15) die("short source block");
  const int nh = ((m0 + 1) * (2 * l0 - m0 + 2)) / 2;
  const int nbc = iv[12] > 0 ? iv[12] : 4;
  if ((op == 1 || op == 2) && nbc < 4) die("RK4 steps need 4 guard cells");

  Inputdata::input_list& L = Inputdata::IN().list();
  L.NnodesX = 1; L.NnodesY = 1;
  L.l0 = l0; L.m0 = m0; L.nump = np;
  L.numx_glob = nx; L.numy_glob = ny;
  L.RKLevel = nbc;
  L.numx = nx + 2 * nbc; L.numy = ny + 2 * nbc;
  L.xmin = dv[1]; L.xmax = dv[2]; L.ymin = dv[3]; L.ymax = dv[4];
  L.pmax = dv[0]; L.pmin = L.pmax / (2.0 * np);
  L.clf_dp = 1.0e300;           // num_h = int(dt_out/CLF)+1 = 1  (def-RK.cpp:60)
  L.n_outsteps = 1; L.t_stop = dv[5];  // dt_out = h
  L.small_dt = dv[6];
  L.NB_algorithms = nb;
  L.if_tridiagonal = true; L.implicit_E = (op == 5 || op == 9); L.if_implicit1D = (op == 5);
  L.if_implicitES = ((op == 5 || op == 9) && es);
  L.density_np = dv[7]; L.Zeta = dv[8];
  L.bndX = bndX; L.bndY = bndY;
  L.linear_Ivst = sv[0] != 0; L.polynomial_Ivst = sv[1] != 0;
  L.inverse_bremsstrahlung = static_cast<int>(sv[2]); L.polarization_direction = static_cast<int>(sv[3]);
  L.I_0 = sv[4]; L.lambda_0 = sv[5]; L.ab_eta = sv[6]; L.ab_depth = sv[7]; L.cntr_x0 = sv[8];
  L.cntr_r0 = sv[9]; L.spot_w0 = sv[10]; L.rise_time = sv[11]; L.flat_time = sv[12]; L.fall_time = sv[13];
  L.smaller_dt = sv[14];
  Inputdata::IN().inp() = Inputdata::decl_input(L);
  Inputdata::IN().cont() = Inputdata::decl_control(L);
  if (ny == 1) {
    // Axis(1,ymin,ymax) divides by zero (matrices.h:864) and y.dx() then
    // reads past the one-element axis (def-Input.cpp:324-325), which makes
    // every y coefficient NaN.  A y-uniform 1D state has Dy f = 0 exactly in
    // the interior, so any finite spacing reproduces the 1D answer.
    Inputdata::IN().inp().y = Axis<double>(ny + 2 * nbc, L.ymin, L.ymax);
  }

  Stat Y;
  std::vector<cd> buf((size_t)nh * ny * nx * np), fb((size_t)6 * ny * nx);
  if (std::fread(buf.data(), sizeof(cd), buf.size(), fi) != buf.size()) die("short f");
  if (std::fread(fb.data(), sizeof(cd), fb.size(), fi) != fb.size()) die("short fields");
  std::fclose(fi);
  Y = cd(0.0, 0.0);
  for (int h = 0; h < nh; ++h)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix)
        for (int ip = 0; ip < np; ++ip)
          Y.SH(h)(ip, ix + nbc, iy + nbc) = buf[(((size_t)h * ny + iy) * nx + ix) * np + ip];
  for (int k = 0; k < 6; ++k)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) Y.FLD(k)(ix + nbc, iy + nbc) = fb[((size_t)k * ny + iy) * nx + ix];
  auto fill = [&](Stat& S) {
    if (mirror) fill_guards_bnd(S, nbc, bndX, bndY);
    else fill_guards_periodic(S, nbc);
  };
  fill(Y);
  Stat Y0(Y);  // the initial state with its guards

  double t_final = 0.0, h_used = dv[5];
  Stat Out(Y);
  if (op == 0) {
    Stat Yh(Y);
    Actions act(Yh);
    act.Adv(Y);
    Out = Yh;
  } else if (op == 1 || op == 2) {
    Runge_Kutta_4 rk(Y, 0);
    Explicit_EC_FP ee(Y, 0);
    Euler_Backward eb(Y, 0);
    PLaserSource PLS(Y, 0);
    InverseBremsstrahlung IBs(Y, 0);
    for (int s = 0; s < nsteps; ++s) {
      rk.advance_p1();
      if (op == 1) {
        ee.loop(rk.time());
        if (PLS.PLSOURCE()) PLS.Laser_Source(rk.time());
        if (IBs.IBSOURCE() == 1) IBs.loop(rk.time());
        eb.advance_1(rk.time());
        eb.advance_lm(rk.time());
      }
      fill(Y);
    }
    t_final = rk.time();
    Out = Y;
  } else if (op == 3 || op == 4) {
    Explicit_EC_FP ee(Y, 0);
    Euler_Backward eb(Y, 0);
    double t = 0.0;
    for (int s = 0; s < nsteps; ++s) {
      t += dv[5];
      if (op == 3) {
        ee.loop(t);
      } else {
        eb.advance_1(t);
        eb.advance_lm(t);
      }
      fill(Y);
    }
    t_final = t;
    Out = Y;
  } else if (op == 5 || op == 9) {
    Runge_Kutta_2_Imp rk(Y, 0);
    Explicit_EC_FP ee(Y, 0);
    Euler_Backward eb(Y, 0);
    Electric_Field_Methods::Implicit_E_Field eim(Y, 0);
    PLaserSource PLS(Y, 0);
    InverseBremsstrahlung IBs(Y, 0);
    for (int s = 0; s < nsteps; ++s) {
      rk.advance_p1();
      // Neighbor_ImplicitE_Communications moves the B guards only
      // (def-Parallel.cpp:2945-3033); f guards are not read before the
      // full exchange below (the operators in between are x-local)
      {
        Stat G(Y);
        fill(G);
        for (int k = 3; k < 6; ++k) Y.FLD(k) = G.FLD(k);
      }
      ee.loop(rk.time());
      if (PLS.PLSOURCE()) PLS.Laser_Source(rk.time());
      if (IBs.IBSOURCE() == 1) IBs.loop(rk.time());
      eim.advance(&rk, &eb);
      rk.advance_p2();
      eb.advance_1(rk.time());
      eb.advance_lm(rk.time());
      fill(Y);
    }
    t_final = rk.time();
    Out = Y;
  } else if (op >= 6 && op <= 8) {
    PLaserSource PLS(Y, 0);
    InverseBremsstrahlung IBs(Y, 0);
    ExternalEfield EFLD(Y, 0);
    double t = 0.0;
    for (int s = 0; s < nsteps; ++s) {
      t += dv[5];
      if (op == 6) PLS.Laser_Source(t);
      else if (op == 7) IBs.loop(t);
      else EFLD.Laser_Efield(t);
      fill(Y);
    }
    t_final = t;
    Out = Y;
  } else {
    die("bad op");
  }

  FILE* fo = std::fopen(argv[2], "wb");
  if (!fo) die("cannot open output");
  double hdr[4] = {Inputdata::IN().inp().x.dx(), Inputdata::IN().inp().y.dx(), h_used, t_final};
  std::fwrite(hdr, sizeof(double), 4, fo);
  std::vector<double> pax(np);
  for (int i = 0; i < np; ++i) pax[i] = Inputdata::IN().inp().pr(i);
  std::fwrite(pax.data(), sizeof(double), np, fo);
  for (int h = 0; h < nh; ++h)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix)
        for (int ip = 0; ip < np; ++ip) buf[(((size_t)h * ny + iy) * nx + ix) * np + ip] = Out.SH(h)(ip, ix + nbc, iy + nbc);
  for (int k = 0; k < 6; ++k)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < nx; ++ix) fb[((size_t)k * ny + iy) * nx + ix] = Out.FLD(k)(ix + nbc, iy + nbc);
  std::fwrite(buf.data(), sizeof(cd), buf.size(), fo);
  std::fwrite(fb.data(), sizeof(cd), fb.size(), fo);
  if (mirror) {
    write_ext(fo, Y0);
    write_ext(fo, Out);
  }
  std::fclose(fo);
  return 0;
}
