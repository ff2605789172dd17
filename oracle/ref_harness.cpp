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
//   in.bin : int32[10] {l0,m0,np,nx,ny,op,nsteps,nb,es,src}  (es: if_implicitES, op 5)
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
//            double[np] p axis, then f and fields as in in.bin.

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

int main(int argc, char** argv) {
  if (argc != 3) die("usage: oshun_ref in.bin out.bin");
  FILE* fi = std::fopen(argv[1], "rb");
  if (!fi) die("cannot open input");
  int iv[10];
  double dv[9];
  if (std::fread(iv, sizeof(int), 10, fi) != 10) die("short header");
  if (std::fread(dv, sizeof(double), 9, fi) != 9) die("short header");
  const int l0 = iv[0], m0 = iv[1], np = iv[2], nx = iv[3], ny = iv[4], op = iv[5], nsteps = iv[6], nb = iv[7];
  const int es = iv[8];
  double sv[15] = {0};
  if (iv[9] && std::fread(sv, sizeof(double), 15, fi) != 15) die("short source block");
  const int nh = ((m0 + 1) * (2 * l0 - m0 + 2)) / 2;
  const int nbc = 4;

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
  L.bndX = 0; L.bndY = 0;
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
  fill_guards_periodic(Y, nbc);

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
      fill_guards_periodic(Y, nbc);
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
      fill_guards_periodic(Y, nbc);
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
        fill_guards_periodic(G, nbc);
        for (int k = 3; k < 6; ++k) Y.FLD(k) = G.FLD(k);
      }
      ee.loop(rk.time());
      if (PLS.PLSOURCE()) PLS.Laser_Source(rk.time());
      if (IBs.IBSOURCE() == 1) IBs.loop(rk.time());
      eim.advance(&rk, &eb);
      rk.advance_p2();
      eb.advance_1(rk.time());
      eb.advance_lm(rk.time());
      fill_guards_periodic(Y, nbc);
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
      fill_guards_periodic(Y, nbc);
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
  std::fclose(fo);
  return 0;
}
