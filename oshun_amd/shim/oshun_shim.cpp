// oshun_shim.cpp -- OSHUN-1's RK4 driver, its implicit-E RK2 driver and the
// collision-step bodies, running on an MI355X through the C ABI of oshun_hip.h.
//
// Drop-in at link time: the class declarations stay the reference's own
// (decl-RK.h:48-79 and :169-203, decl-fokkerplanck.h:140-163 and :289-311,
// decl-implicitE.h:146-171); this file defines the member functions that do
// the work of one step of main_loop (main.cpp:74-107):
//   Runge_Kutta_4::Runge_Kutta_4, ::advance_p1             def-RK.cpp:49-62, 110-145
//   Runge_Kutta_2_Imp::Runge_Kutta_2_Imp, ::advance_p1/p2  def-RK.cpp:400-413, 470-524
//   Explicit_EC_FP::loop                                   def-FokkerPlanck.cpp:590-625
//   Euler_Backward::advance_1, ::advance_lm                def-FokkerPlanck.cpp:1270-1304
//   Electric_Field_Methods::Implicit_E_Field::advance      def-ImplicitE.cpp:472-606
//   PLaserSource::Laser_Source                             def-PLSource.cpp:138-252
//   InverseBremsstrahlung::loop                            def-PLSource.cpp:780-861
// With a mirror boundary (bndX or bndY = 1) the device keeps the whole arrays,
// guards included, and refreshes the guards at the end of each step
// (advance_lm) as Neighbor_Communications does (def-Parallel.cpp:1121-1307).
// Link this object ahead of def-RK.o, def-FokkerPlanck.o, def-ImplicitE.o and def-PLSource.o,
// either after deleting those bodies from the reference files or with
// -Wl,--allow-multiple-definition (then the first definition, this one, is
// the one every caller and every vtable binds to; the reference files stay
// untouched).  oracle/Makefile's `shim` target does the latter.
//
// One device state per process, as main.cpp builds one Stat Y: the first
// integrator call creates the device context from Inputdata::IN() and
// uploads the host Stat (after main_loop has read a restart, main.cpp:49-64);
// the host Stat is refreshed by a download at the end of every output
// interval -- every numh() steps of the driver, main.cpp:77-131 -- so that
// CX.Output and Write_Restart read current data.  The guard refresh
// CX.Neighbor_Communications then only rewrites host guard cells.  The laser
// and inverse-bremsstrahlung sources (def-PLSource.cpp) run on the device too,
// with the deck's parameters and axes.  Errors print and exit(1), as the
// reference does.
//
// Implicit-E decks: the whole step (RK2 halves, e-e, sources, field solve,
// implicit collisions) is one device call inside Runge_Kutta_2_Imp::
// advance_p1 -- oshun_step_implicit_1d for the 1D variant (if_implicit1D: the
// device context then holds the m = 0 harmonics only), oshun_step_implicit_2d
// otherwise; the later calls of the step only move their clocks.  With
// if_implicitES the 2D field solve's JN is the Ex the field solver was built
// with (def-ImplicitE.cpp:56-58, 446): the RK2 constructor, which main.cpp
// runs just before the solver's, keeps a copy of it for the device.

#include <cstdlib>
#include <iostream>
#include <vector>
#include <valarray>
#include <complex>
#include <limits>
#include <math.h>
#include <stdio.h>
#include <float.h>

#include "matrices.h"
#include "decl-nmethods.h"
#include "decl-input.h"
#include "decl-state.h"
#include "decl-vlasovmax.h"
#include "decl-actions.h"
#include "decl-RK.h"
#include "decl-fokkerplanck.h"
#include "decl-implicitE.h"
#include "decl-plsource.h"

#include "oshun_hip.h"

namespace {

struct Session {
  oshun_ctx* ctx = nullptr;
  Stat* Y = nullptr;     // the host Stat the device state mirrors
  bool implicit = false; // Runge_Kutta_2_Imp drives the step
  bool m0_only = false;  // the device holds the m = 0 harmonics only
  bool rk_driven = false;
  size_t numh = 1;       // steps per output interval of the driver
  size_t steps = 0;      // steps since the last download
  std::vector<double> es_ex;  // Ex at the implicit field solver's construction (2D, if_implicitES)
};

Session& session() {
  static Session s;
  return s;
}

[[noreturn]] void die(const char* where) {
  Session& s = session();
  std::cout << "oshun shim: " << where;
  if (s.ctx) std::cout << ": " << oshun_last_error(s.ctx);
  std::cout << "\n";
  std::exit(1);
}

void check(int rc, const char* where) {
  if (rc != OSHUN_OK) die(where);
}

// host pointers of the harmonics the device holds (ind(l,m) order) and of the six fields
void pointers(Stat& Y, std::vector<double*>& sh, std::vector<double*>& fld) {
  Session& s = session();
  const size_t l0 = Inputdata::IN().inp().l0;
  sh.clear();
  fld.clear();
  if (s.m0_only) {
    for (size_t l = 0; l <= l0; ++l) sh.push_back(reinterpret_cast<double*>(&Y.SH(l, 0)(0, 0, 0)));
  } else {
    for (size_t i = 0; i < Y.DF().dim(); ++i) sh.push_back(reinterpret_cast<double*>(&Y.SH(i)(0, 0, 0)));
  }
  for (size_t k = 0; k < 6; ++k) fld.push_back(reinterpret_cast<double*>(&Y.FLD(k)(0, 0)));
}

// the device context of this process, created and loaded from Y on first use
void attach(Stat& Y, double t) {
  Session& s = session();
  if (s.ctx) {
    if (s.Y != &Y) die("one state per process: the integrators were built on different Stat objects");
    return;
  }
  using Inputdata::IN;
  const int nbc = IN().list().RKLevel;
  s.m0_only = s.implicit && IN().list().if_implicit1D;
  oshun_grid g;
  g.l0 = static_cast<int>(IN().inp().l0);
  g.m0 = s.m0_only ? 0 : static_cast<int>(IN().inp().m0);
  g.nump = static_cast<int>(IN().inp().pr.dim());
  g.numx = static_cast<int>(IN().inp().x.dim()) - 2 * nbc;
  g.numy = static_cast<int>(IN().inp().y.dim()) - 2 * nbc;
  g.pmin = IN().inp().pr(0);
  g.pmax = IN().inp().pr(IN().inp().pr.dim() - 1);
  g.dx = IN().inp().x.dx();
  g.dy = IN().inp().y.dx();
  g.precision = OSHUN_FP64;
  g.bndx = IN().list().bndX;  // mirror boundaries: the device keeps the whole arrays
  g.bndy = IN().list().bndY;
  g.nbc = nbc;
  oshun_phys p;
  p.density_np = IN().list().density_np;
  p.zeta = IN().list().Zeta;
  p.small_dt = IN().list().small_dt;
  p.nb_algorithms = IN().list().NB_algorithms;
  p.collisions = OSHUN_COLL_EE | OSHUN_COLL_FLM;
  p.ref_quirks = 1;
  check(oshun_create(&g, &p, &s.ctx), "oshun_create");
  s.Y = &Y;
  std::vector<double*> sh, fld;
  pointers(Y, sh, fld);
  check(oshun_upload_harmonics(s.ctx, sh.data(), fld.data(), nbc), "oshun_upload_harmonics");
  check(oshun_set_time(s.ctx, t), "oshun_set_time");
  if (!s.es_ex.empty()) {
    std::vector<double> jn;
    for (int k = 0; k < 3; ++k) jn.insert(jn.end(), s.es_ex.begin(), s.es_ex.end());
    check(oshun_set_implicit_es_current(s.ctx, jn.data(), nbc), "oshun_set_implicit_es_current");
  }
  // the deck's laser sources, with the cell positions of its own axes (def-Input.cpp:313-325)
  const Inputdata::input_list& L = IN().list();
  if (L.linear_Ivst || L.polynomial_Ivst || L.inverse_bremsstrahlung == 1) {
    oshun_sources src;
    src.linear_Ivst = L.linear_Ivst; src.polynomial_Ivst = L.polynomial_Ivst;
    src.inverse_bremsstrahlung = L.inverse_bremsstrahlung; src.polarization_direction = L.polarization_direction;
    src.I_0 = L.I_0; src.lambda_0 = L.lambda_0; src.ab_eta = L.ab_eta; src.ab_depth = L.ab_depth;
    src.cntr_x0 = L.cntr_x0; src.cntr_r0 = L.cntr_r0; src.spot_w0 = L.spot_w0;
    src.rise_time = L.rise_time; src.flat_time = L.flat_time; src.fall_time = L.fall_time;
    src.smaller_dt = L.smaller_dt;
    // positions of the cells the device holds: the interior, or every cell with a mirror boundary
    const int gx = (g.bndx || g.bndy) ? nbc : 0;
    std::vector<double> xs(g.numx + 2 * gx), ys(g.numy + 2 * gx), xg(IN().inp().xglob.dim());
    for (size_t i = 0; i < xs.size(); ++i) xs[i] = IN().inp().x(i + nbc - gx);
    for (size_t j = 0; j < ys.size(); ++j) ys[j] = IN().inp().y(j + nbc - gx);
    for (size_t i = 0; i < xg.size(); ++i) xg[i] = IN().inp().xglob(i);
    check(oshun_set_sources(s.ctx, &src, xs.data(), ys.data(), xg.data(), static_cast<int>(xg.size())),
          "oshun_set_sources");
  }
}

void download() {
  Session& s = session();
  std::vector<double*> sh, fld;
  pointers(*s.Y, sh, fld);
  check(oshun_download_harmonics(s.ctx, sh.data(), fld.data(), Inputdata::IN().list().RKLevel),
        "oshun_download_harmonics");
}

// end of a step (after advance_lm): refresh the host Stat when an output interval ends
void step_done() {
  Session& s = session();
  if (++s.steps >= s.numh) {
    download();
    s.steps = 0;
  }
}

}  // namespace

// ---- Runge_Kutta_4 (def-RK.cpp:49-145) --------------------------------------
Runge_Kutta_4::Runge_Kutta_4(Stat& Yin, int tout_start)
    : Y(Yin), Y0(Yin), Y1(Yin), Yh(Yin), Act(Yh),
      t(static_cast<double>(tout_start) * Inputdata::IN().cont().dt_out),
      Tout(t + Inputdata::IN().cont().dt_out) {
  // the step count and size of the reference (def-RK.cpp:60-61)
  num_h = size_t(static_cast<int>(Inputdata::IN().cont().dt_out / Inputdata::IN().inp().CLF)) + 1;
  h = Inputdata::IN().cont().dt_out / static_cast<double>(num_h);
  Session& s = session();
  s.rk_driven = true;
  s.numh = num_h;
}

Runge_Kutta_4& Runge_Kutta_4::advance_p1() {
  attach(Y, t);
  check(oshun_rk4_step(session().ctx, h), "Runge_Kutta_4::advance_p1");
  t += h;  // the device clock advances by the same h
  return *this;
}

// ---- Runge_Kutta_2_Imp (def-RK.cpp:400-524) --------------------------------------
Runge_Kutta_2_Imp::Runge_Kutta_2_Imp(Stat& Yin, int tout_start)
    : Y(Yin), Y0(Yin), Yh(Yin), Act(Yh),
      t(static_cast<double>(tout_start) * Inputdata::IN().cont().dt_out),
      Tout(t + Inputdata::IN().cont().dt_out) {
  num_h = size_t(static_cast<int>(Inputdata::IN().cont().dt_out / Inputdata::IN().inp().CLF)) + 1;  // :411-412
  h = Inputdata::IN().cont().dt_out / static_cast<double>(num_h);
  Session& s = session();
  s.implicit = true;
  s.rk_driven = true;
  s.numh = num_h;
  const Inputdata::input_list& L = Inputdata::IN().list();
  if (!L.if_implicit1D && L.if_implicitES) {
    const Field2D& ex = Yin.EMF().Ex();
    const double* p = reinterpret_cast<const double*>(&Yin.FLD(0)(0, 0));
    s.es_ex.assign(p, p + 2 * ex.numx() * ex.numy());
  }
}

Runge_Kutta_2_Imp& Runge_Kutta_2_Imp::advance_p1() {
  const Inputdata::input_list& L = Inputdata::IN().list();
  attach(Y, t);
  Session& s = session();
  if (L.if_implicit1D) {
    check(oshun_step_implicit_1d(s.ctx, h, 1, L.if_implicitES ? 1 : 0), "Runge_Kutta_2_Imp::advance_p1");
  } else {
    check(oshun_step_implicit_2d(s.ctx, h, 1, L.if_implicitES ? 1 : 0), "Runge_Kutta_2_Imp::advance_p1");
    const int zeros = oshun_implicit_zero_dets(s.ctx);
    if (zeros > 0) std::cout << "WARNING, Det = 0 in " << zeros << "locations \n";  // def-ImplicitE.cpp:597-598
  }
  t += h;
  return *this;
}

Runge_Kutta_2_Imp& Runge_Kutta_2_Imp::advance_p2() { return *this; }  // done in advance_p1

void Electric_Field_Methods::Implicit_E_Field::advance(Explicit_Method*, Euler_Backward*) {
  // the field solve is part of oshun_step_implicit_1d/_2d (advance_p1)
}

// ---- collisions (def-FokkerPlanck.cpp:590-625, 1270-1304) ---------------------
void Explicit_EC_FP::loop(const double& tnew) {
  Session& s = session();
  if (!s.implicit) {
    attach(Y, t);
    check(oshun_collide_ee(s.ctx, tnew), "Explicit_EC_FP::loop");  // numh from tnew - t as at :597-599
    if (!s.rk_driven) download();  // used without a driver: the caller reads the Stat next
  }
  t = tnew;
}

// ---- sources (def-PLSource.cpp:138-252, 780-861); inside the implicit step's device call there
void PLaserSource::Laser_Source(const double& tnew) {
  Session& s = session();
  if (!s.implicit) {
    // the class holds f00 only, not the Stat: the device state comes from the step's integrators
    if (!s.ctx) die("PLaserSource::Laser_Source before the first integrator call");
    check(oshun_laser_source(s.ctx, tnew), "PLaserSource::Laser_Source");
  }
  t1 = tnew;
}

void InverseBremsstrahlung::loop(const double& tnew) {
  Session& s = session();
  if (!s.implicit) {
    attach(Y, t);
    check(oshun_inverse_bremsstrahlung(s.ctx, tnew), "InverseBremsstrahlung::loop");
    if (!s.rk_driven) download();
  }
  t = tnew;
}

void Euler_Backward::advance_1(const double& tnew) {
  Session& s = session();
  if (!s.implicit) {
    attach(Y, t);
    check(oshun_collide_flm(s.ctx, tnew, 1), "Euler_Backward::advance_1");
  }
}

void Euler_Backward::advance_lm(const double& tnew) {
  Session& s = session();
  if (!s.implicit) {
    attach(Y, t);
    check(oshun_collide_flm(s.ctx, tnew, 2), "Euler_Backward::advance_lm");
    // the step's guard refresh (CX.Neighbor_Communications, main.cpp:104) where the device keeps guards
    const Inputdata::input_list& L = Inputdata::IN().list();
    if (L.bndX || L.bndY) check(oshun_fill_guards(s.ctx), "Neighbor_Communications");
  }
  t = tnew;
  step_done();
}
