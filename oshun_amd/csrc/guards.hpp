// guards.hpp -- the guard-cell refresh of a context with a mirror boundary
// (oshun_grid bndx / bndy = 1), which keeps the reference's whole arrays,
// guards included, on the device.  Included at the end of oshun_hip.hip.
//
// Parallel_Environment::Neighbor_Communications of one node
// (def-Parallel.cpp:3040-3127): the x direction, then the y direction, each
// periodic (sameNode_periodic_X/Y, :1137-1156, :1229-1248) or mirror
// (sameNode_mirror_X/Y, :1160-1200, :1251-1307); and the implicit-E step's
// magnetic-field refresh after its first RK2 half (b_only).  As there:
//   * periodic: the last nbc interior columns (rows) go to the left guards,
//     the first nbc to the right ones, every harmonic and field;
//   * x mirror: guard at distance d from the face = interior cell at distance
//     d, times (-1)^(l+m), for the harmonics with l < l0 only (the loop bound
//     of :1168); fields copied likewise with Ey, Ez, Bx negated;
//   * y mirror: the same copy for 1 <= l < l0 only (:1258), then
//     v -> (v - 2 Re v) * (-(-1)^m), i.e. (-1)^m conj(v); fields with Ex, Ez,
//     By negated.
// Harmonics the mirror skips keep their guard values, which the Vlasov stages
// then advance like any other cell (the reference's behaviour, reproduced).
// The x pass covers whole columns (all rows, guards included) and the y pass
// whole rows, so the corners come from the y pass as in the reference.

namespace oshun {

struct GuardArgs {
  double* f;    // planes [plane][y][p][x]
  double* fld;  // fields [6][y][x]
  const int* comp_l;
  const int* comp_m;
  const int* comp_isim;  // 1 on the imaginary plane of a harmonic
  int nc, np, nx, ny, g, l0, mirror;
  int b_only;  // Neighbor_ImplicitE_Communications: the magnetic field only
  long long plane;
};

// x pass: one thread per (guard column slot, y, p); slot j < g is the left
// guard j, j >= g the right guard nx - 2g + j
__global__ __launch_bounds__(256) void guards_x(GuardArgs a) {
  const int g = a.g, np = a.np, nx = a.nx, ny = a.ny;
  const long long n = 2LL * g * ny * np;
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= n) return;
  const int j = static_cast<int>(t % (2 * g));
  const long long r = t / (2 * g);
  const int p = static_cast<int>(r % np), y = static_cast<int>(r / np);
  int dst, src;
  if (!a.mirror) {
    if (j < g) { dst = j; src = nx - 2 * g + j; }               // right bound -> left guard
    else { dst = nx - 2 * g + j; src = j; }                     // left bound -> right guard (g + (j - g))
  } else {
    if (j < g) { dst = j; src = 2 * g - j - 1; }                // left guard c <- column 2g - c - 1
    else { const int c = j - g; dst = nx - c - 1; src = nx - 2 * g + c; }  // right guard nx-c-1 <- nx-2g+c
  }
  const long long row = ((long long)y * np + p) * nx;
  for (int c = 0; c < (a.b_only ? 0 : a.nc); ++c) {
    const int l = a.comp_l[c];
    if (a.mirror && l >= a.l0) continue;
    const double sg = (a.mirror && ((l + a.comp_m[c]) & 1)) ? -1.0 : 1.0;
    double* pl = a.f + (long long)c * a.plane + row;
    pl[dst] = pl[src] * sg;
  }
  if (p == 0) {
    const long long frow = (long long)y * nx;
    for (int k = a.b_only ? 3 : 0; k < 6; ++k) {
      const double sg = (a.mirror && (k == 1 || k == 2 || k == 3)) ? -1.0 : 1.0;  // Ey, Ez, Bx
      double* F = a.fld + (long long)k * nx * ny + frow;
      F[dst] = F[src] * sg;
    }
  }
}

// y pass: one thread per (guard row slot, p, x)
__global__ __launch_bounds__(256) void guards_y(GuardArgs a) {
  const int g = a.g, np = a.np, nx = a.nx, ny = a.ny;
  const long long n = 2LL * g * np * nx;
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= n) return;
  const int x = static_cast<int>(t % nx);
  const long long r = t / nx;
  const int p = static_cast<int>(r % np), j = static_cast<int>(r / np);
  int dst, src;
  if (!a.mirror) {
    if (j < g) { dst = j; src = ny - 2 * g + j; }
    else { dst = ny - 2 * g + j; src = j; }
  } else {
    if (j < g) { dst = j; src = 2 * g - j - 1; }
    else { const int c = j - g; dst = ny - c - 1; src = ny - 2 * g + c; }
  }
  const long long od = ((long long)dst * np + p) * nx + x, os = ((long long)src * np + p) * nx + x;
  for (int c = 0; c < (a.b_only ? 0 : a.nc); ++c) {
    const int l = a.comp_l[c];
    double* pl = a.f + (long long)c * a.plane;
    if (!a.mirror) {
      pl[od] = pl[os];
    } else if (l >= 1 && l < a.l0) {
      const double s = (a.comp_m[c] & 1) ? 1.0 : -1.0;  // -(-1)^m
      double v = pl[os];
      if (!a.comp_isim[c]) {  // real part: v - 2 Re v
        const double w = v;
        v -= 2.0 * w;
      }
      pl[od] = v * s;
    }
  }
  if (p == 0) {
    for (int k = a.b_only ? 3 : 0; k < 6; ++k) {
      const double sg = (a.mirror && (k == 0 || k == 2 || k == 4)) ? -1.0 : 1.0;  // Ex, Ez, By
      double* F = a.fld + (long long)k * nx * ny;
      F[(long long)dst * nx + x] = F[(long long)src * nx + x] * sg;
    }
  }
}

}  // namespace oshun

namespace {

// b_only: Parallel_Environment::Neighbor_ImplicitE_Communications
// (def-Parallel.cpp:2945-3033 with Node_ImplicitE_Communications::
// sameNode_periodic/mirror_X/Y, :497-680): Bx, By, Bz only, with Bx negated
// across an x mirror and By across a y mirror
int fill_guards(oshun_ctx* c, bool b_only) {
  if (!c->ext) return OSHUN_OK;
  oshun::GuardArgs a;
  a.b_only = b_only ? 1 : 0;
  a.f = static_cast<double*>(c->S);
  a.fld = c->Sf;
  a.comp_l = c->d_comp_l;
  a.comp_m = c->d_comp_m;
  a.comp_isim = c->d_comp_isim;
  a.nc = c->nc; a.np = c->np; a.nx = c->nx; a.ny = c->ny; a.g = c->gn; a.l0 = c->T.l0;
  a.plane = c->plane;
  const long long nxg = 2LL * c->gn * c->ny * c->np, nyg = 2LL * c->gn * c->np * c->nx;
  a.mirror = c->g.bndx;
  hipLaunchKernelGGL(oshun::guards_x, dim3(static_cast<unsigned>((nxg + 255) / 256)), dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  a.mirror = c->g.bndy;
  hipLaunchKernelGGL(oshun::guards_y, dim3(static_cast<unsigned>((nyg + 255) / 256)), dim3(256), 0, c->stream, a);
  HIPCHK(c, hipGetLastError());
  return OSHUN_OK;
}

}  // namespace

extern "C" {

int oshun_fill_guards(oshun_ctx* c) {
  if (!sane(c)) return OSHUN_ERR_ARG;
  if (!c->ext) return fail(c, OSHUN_ERR_ARG, "oshun_fill_guards: the context keeps no guard cells");
  return fill_guards(c, false);
}

}  // extern "C"
