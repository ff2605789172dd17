// Batched independent runs: many 1D m0 = 0 runs of uneven Nx, Np and l0
// advanced together, one hipGraph launch per RK4 + collision step.
//
// Every launch of the step (four field stages, two stage pairs, the collision
// launches, clock) covers the tiles of all runs at once: the run of
// a workgroup comes from a prefix table of per-run tile counts and its
// arguments from a per-run array in device memory.  Kernels templated on l0
// (stage pairs) or on the collision layout get one launch per distinct value.
// The e-e sub-steps are derived on the device from per-run clocks, so the
// captured step needs no host input.  The pair kernels' tile height is chosen
// by timing the candidates on the actual batch (oshun_batch_autotune).
//
// Included at the end of oshun_hip.hip (same translation unit: it uses the
// context internals and the single-run argument builders).

struct oshun_batch {
  std::vector<oshun_ctx*> runs;
  std::vector<double> h;  // RK4 step per run
  int fp32 = 0;
  int pair_r = 3;         // rows per thread of the pair kernels (tile height 4R-2); 0: the marched kernels
  int march_pw = 8;       // p rows per wavefront of the marched kernels
  hipStream_t stream = nullptr;
  double* d_t = nullptr;  // run clocks
  double* d_h = nullptr;
  std::vector<void*> dev;  // argument and tile tables of the captured sequences
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  int parity = 0;         // number of steps taken, mod 2 (S and B1 of every run swap per step)
  std::string err;
};

namespace {

int bfail(oshun_batch* b, int code, const std::string& msg) {
  b->err = msg;
  return code;
}
#define BCHK(b, expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return bfail(b, OSHUN_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename V>
int bupload(oshun_batch* b, const std::vector<V>& v, const V** out) {
  void* d = nullptr;
  if (v.empty()) { *out = nullptr; return OSHUN_OK; }
  BCHK(b, hipMalloc(&d, v.size() * sizeof(V)));
  b->dev.push_back(d);
  BCHK(b, hipMemcpy(d, v.data(), v.size() * sizeof(V), hipMemcpyHostToDevice));
  *out = static_cast<const V*>(d);
  return OSHUN_OK;
}

// one batched launch: per-run arguments for the runs in `sel`, tiles per run
template <typename A>
struct BLaunch {
  std::vector<int> sel;
  std::vector<A> args;
  std::vector<int> ts;  // prefix of tiles
  const A* d_args = nullptr;
  const int* d_ts = nullptr;
  int total() const { return ts.empty() ? 0 : ts.back(); }
  void add(int run, const A& a, int tiles) {
    if (ts.empty()) ts.push_back(0);
    sel.push_back(run);
    args.push_back(a);
    ts.push_back(ts.back() + tiles);
  }
  int upload(oshun_batch* b) {
    int rc = bupload(b, args, &d_args);
    return rc ? rc : bupload(b, ts, &d_ts);
  }
};

// grids of the remapped batched kernels are padded to a multiple of the 8 XCDs
// (xcd_tile then gives each XCD one contiguous run of tiles; the padding
// workgroups return at once)
inline int pad8(int n) { return (n + 7) / 8 * 8; }

template <typename T, int R, int PAIR>
void launch_pair_batch(hipStream_t s, int l0, const BLaunch<oshun::Stage1DArgs<T>>& L) {
  const dim3 g(pad8(L.total())), blk(64, 4);
  const int n = static_cast<int>(L.sel.size());
  switch (l0) {
    case 1: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 1, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 2: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 2, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 3: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 3, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 4: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 4, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 5: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 5, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 6: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 6, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 7: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 7, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    default: hipLaunchKernelGGL((oshun::vlasov_pair_1d_batch<T, 8, R, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
  }
}
template <typename T, int PAIR>
void launch_march_batch(hipStream_t s, int l0, const BLaunch<oshun::Stage1DArgs<T>>& L) {
  const dim3 g(pad8((L.total() + 3) / 4)), blk(256);  // four wavefronts per workgroup
  const int n = static_cast<int>(L.sel.size());
  switch (l0) {
    case 1: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 1, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 2: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 2, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 3: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 3, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 4: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 4, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 5: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 5, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 6: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 6, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 7: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 7, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    default: hipLaunchKernelGGL((oshun::vlasov_march_1d_batch<T, 8, PAIR>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
  }
}
template <typename T, int MODE>
void launch_coll_batch(hipStream_t s, int key, const BLaunch<oshun::CollArgs<T>>& L) {
  const dim3 g(pad8(L.total())), blk(256);
  const int n = static_cast<int>(L.sel.size());
  switch (key) {
    case 1604: hipLaunchKernelGGL((oshun::collide_rows_batch<T, 4, 16, MODE>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    case 1608: hipLaunchKernelGGL((oshun::collide_rows_batch<T, 8, 16, MODE>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
    default: hipLaunchKernelGGL((oshun::collide_rows_batch<T, 8, 32, MODE>), g, blk, 0, s, L.d_args, L.d_ts, n); break;
  }
}

// Record one step (parity q: S/B1 of every run as after q steps) on b->stream.
// R > 0: the tile kernels with tiles of 4R-2 rows; R = 0: the marched kernels
// with b->march_pw rows per wavefront (the launch counts wavefronts).
template <typename T, int R>
int record_step(oshun_batch* b, int q) {
  const int nr = static_cast<int>(b->runs.size());
  BLaunch<oshun::FieldArgs<T>> f0, f2;
  BLaunch<oshun::Field23Args<T>> f23;
  std::map<int, BLaunch<oshun::Stage1DArgs<T>>> p0, p1;   // by l0
  std::map<int, BLaunch<oshun::CollArgs<T>>> c1, c2, c3;  // by collision layout (c3: both in one launch)
  for (int i = 0; i < nr; ++i) {
    oshun_ctx* c = b->runs[i];
    const double h = b->h[i];
    // buffers of this parity (the run's own S/B1 are those of the current parity)
    const int flip = (q != b->parity);
    void* S = flip ? c->B1 : c->S;
    void* B1 = flip ? c->S : c->B1;
    const int npt = pair_tiles_r(c->np, R, b->march_pw);
    const int units = R == 0 ? march_strips(c) * npt : ((c->nx + oshun::kPairX - 1) / oshun::kPairX) * npt;
    const int fblocks = (c->nxy + oshun::kFieldX - 1) / oshun::kFieldX;
    oshun::FieldArgs<T> fa0 = field_args<T>(c, 1, c->Sf, c->Sf, c->B1f, c->Af, S, h, nullptr, 0, nullptr);
    fa0.p3og = c->d_p3og;  // current of Y0 summed directly over p
    f0.add(i, fa0, fblocks);
    oshun::Stage1DArgs<T> pa0 = pair_args<T>(c, S, nullptr, c->B2, c->A, c->Sf, c->B1f, c->jp1, c->jp2, h);
    pa0.pw = b->march_pw;  // the batch's rows per wavefront (the run's own plan is left as it is)
    p0[c->T.l0].add(i, pa0, units);
    oshun::Field23Args<T> fa23;
    fa23.a2 = field_args<T>(c, 2, c->B1f, c->Sf, c->B2f, c->Af, nullptr, h, c->jp1, npt, c->f1y);
    fa23.a3 = field_args<T>(c, 3, c->B2f, c->Sf, c->Cf, c->Af, c->B2, h, c->jp2, npt, nullptr);
    f23.add(i, fa23, fblocks);
    oshun::Stage1DArgs<T> pa1 = pair_args<T>(c, c->B2, S, B1, c->A, c->B2f, c->Cf, c->jp1, nullptr, h);
    pa1.pw = b->march_pw;
    p1[c->T.l0].add(i, pa1, units);
    f2.add(i, field_args<T>(c, 4, c->Cf, nullptr, c->Sf, c->Af, nullptr, h, c->jp1, npt, c->f1y), fblocks);
    const int coll = c->ph.collisions;
    if (coll) {
      oshun::CollArgs<T> a = coll_args<T>(c, (coll & OSHUN_COLL_EE) ? 1 : 0, (coll & OSHUN_COLL_FLM) ? 1 : 0,
                                          (coll & OSHUN_COLL_FLM) ? 1 : 0, 1, 0.0, 0.0);
      a.f = static_cast<T*>(B1);  // the state after RK4
      a.tclk = b->d_t + i;
      a.hstep = h;
      const bool same = c->rows_segw == c->rows_segw2 && c->rows_npl == c->rows_npl2;
      if ((coll & OSHUN_COLL_EE) && (coll & OSHUN_COLL_FLM) && same) {
        a.rt = c->d_rt;
        c3[c->rows_segw * 100 + c->rows_npl].add(i, a, ((c->nx + 256 / c->rows_segw - 1) / (256 / c->rows_segw)));
      } else if (coll & OSHUN_COLL_EE) {
        a.rt = c->d_rt;
        c1[c->rows_segw * 100 + c->rows_npl].add(i, a, ((c->nx + 256 / c->rows_segw - 1) / (256 / c->rows_segw)));
      }
      if ((coll & OSHUN_COLL_FLM) && !((coll & OSHUN_COLL_EE) && same)) {
        a.rt = c->d_rt2;
        c2[c->rows_segw2 * 100 + c->rows_npl2].add(i, a, ((c->nx + 256 / c->rows_segw2 - 1) / (256 / c->rows_segw2)));
      }
    }
  }
  int rc;
  if ((rc = f0.upload(b)) || (rc = f23.upload(b)) ||
      (rc = f2.upload(b)))
    return rc;
  for (auto* m : {&p0, &p1})
    for (auto& kv : *m)
      if ((rc = kv.second.upload(b))) return rc;
  for (auto* m : {&c1, &c2, &c3})
    for (auto& kv : *m)
      if ((rc = kv.second.upload(b))) return rc;
  // capture the step
  hipStream_t s = b->stream;
  BCHK(b, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  const dim3 fblk(oshun::kFieldX, oshun::kFieldRows);
  hipLaunchKernelGGL(oshun::field_stage_batch<T>, dim3(f0.total()), fblk, 0, s, f0.d_args, f0.d_ts, nr);
  for (auto& kv : p0) {
    if (R == 0) launch_march_batch<T, 0>(s, kv.first, kv.second);
    else launch_pair_batch<T, (R > 0 ? R : 3), 0>(s, kv.first, kv.second);
  }
  hipLaunchKernelGGL(oshun::field_stage23_1d_batch<T>, dim3(f23.total()), fblk, 0, s, f23.d_args, f23.d_ts, nr);
  for (auto& kv : p1) {
    if (R == 0) launch_march_batch<T, 1>(s, kv.first, kv.second);
    else launch_pair_batch<T, (R > 0 ? R : 3), 1>(s, kv.first, kv.second);
  }
  hipLaunchKernelGGL(oshun::field_stage_batch<T>, dim3(f2.total()), fblk, 0, s, f2.d_args, f2.d_ts, nr);
  for (auto& kv : c1) launch_coll_batch<T, 1>(s, kv.first, kv.second);
  for (auto& kv : c2) launch_coll_batch<T, 2>(s, kv.first, kv.second);
  for (auto& kv : c3) launch_coll_batch<T, 3>(s, kv.first, kv.second);
  hipLaunchKernelGGL(oshun::batch_clock, dim3((nr + 255) / 256), dim3(256), 0, s, b->d_t, b->d_h, nr);
  hipGraph_t g = nullptr;
  BCHK(b, hipStreamEndCapture(s, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (ie != hipSuccess) return bfail(b, OSHUN_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
  b->exec[q] = e;
  return OSHUN_OK;
}

void drop_graphs(oshun_batch* b) {
  for (auto& e : b->exec)
    if (e) { hipGraphExecDestroy(e); e = nullptr; }
  for (void* p : b->dev) hipFree(p);
  b->dev.clear();
}

int build_graphs(oshun_batch* b) {
  drop_graphs(b);
  for (int q = 0; q < 2; ++q) {
    int rc;
    if (b->fp32)
      rc = b->pair_r == 0   ? record_step<float, 0>(b, q)
           : b->pair_r == 2 ? record_step<float, 2>(b, q)
                            : (b->pair_r == 3 ? record_step<float, 3>(b, q) : record_step<float, 4>(b, q));
    else rc = b->pair_r == 0   ? record_step<double, 0>(b, q)
              : b->pair_r == 2 ? record_step<double, 2>(b, q) : record_step<double, 3>(b, q);
    if (rc) return rc;
  }
  return OSHUN_OK;
}

// host side of one step: clocks and buffer roles follow the device
void after_step(oshun_batch* b) {
  for (size_t i = 0; i < b->runs.size(); ++i) {
    oshun_ctx* c = b->runs[i];
    std::swap(c->S, c->B1);
    c->t_rk += b->h[i];
    c->t_ee = c->t_eb = c->t_rk;
  }
  b->parity ^= 1;
}

int batch_steps(oshun_batch* b, int nsteps) {
  std::vector<double> t(b->runs.size());
  for (size_t i = 0; i < t.size(); ++i) t[i] = b->runs[i]->t_rk;
  BCHK(b, hipMemcpyAsync(b->d_t, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice, b->stream));
  BCHK(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k < nsteps; ++k) {
    BCHK(b, hipGraphLaunch(b->exec[b->parity], b->stream));
    after_step(b);
  }
  return OSHUN_OK;
}

}  // namespace

extern "C" {

int oshun_batch_create(const oshun_grid* grids, const oshun_phys* phys, const double* h, int nruns,
                       oshun_batch** out) {
  if (!grids || !phys || !h || !out || nruns < 1) return OSHUN_ERR_ARG;
  *out = nullptr;
  oshun_batch* b = new oshun_batch();
  auto bail = [&](int rc) { oshun_batch_destroy(b); return rc; };
  b->fp32 = grids[0].precision == OSHUN_FP32_SPLIT;
  for (int i = 0; i < nruns; ++i) {
    const oshun_grid& g = grids[i];
    if (g.m0 != 0 || g.numy != 1 || g.l0 > 8 || (g.precision == OSHUN_FP32_SPLIT) != (b->fp32 != 0) ||
        !(h[i] > 0.0))
      return bail(OSHUN_ERR_ARG);
    oshun_ctx* c = nullptr;
    const int rc = oshun_create(&grids[i], &phys[i], &c);
    if (rc) return bail(rc);
    b->runs.push_back(c);
    b->h.push_back(h[i]);
    if (!c->use_pair || (c->ph.collisions && !c->rows_npl)) return bail(OSHUN_ERR_ARG);
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return bail(OSHUN_ERR_HIP);
  // every run works on the batch's stream, so that uploads, downloads and
  // moments through a run's context are ordered with the batch's graphs
  for (oshun_ctx* c : b->runs) {
    if (c->stream) hipStreamDestroy(c->stream);
    c->stream = b->stream;
  }
  if (hipMalloc((void**)&b->d_t, nruns * sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&b->d_h, nruns * sizeof(double)) != hipSuccess ||
      hipMemcpy(b->d_h, h, nruns * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(OSHUN_ERR_HIP);
  b->pair_r = 3;  // the single-run tile height (oshun_step): a run is bitwise the same alone
  const int rc = build_graphs(b);
  if (rc) return bail(rc);
  *out = b;
  return OSHUN_OK;
}

int oshun_batch_destroy(oshun_batch* b) {
  if (!b) return OSHUN_ERR_ARG;
  if (b->stream) hipStreamSynchronize(b->stream);
  drop_graphs(b);
  if (b->d_t) hipFree(b->d_t);
  if (b->d_h) hipFree(b->d_h);
  for (oshun_ctx* c : b->runs) {
    if (c->stream == b->stream) c->stream = nullptr;  // owned by the batch
    oshun_destroy(c);
  }
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return OSHUN_OK;
}

const char* oshun_batch_last_error(const oshun_batch* b) { return b ? b->err.c_str() : "null batch"; }

int oshun_batch_size(const oshun_batch* b) { return b ? static_cast<int>(b->runs.size()) : OSHUN_ERR_ARG; }

oshun_ctx* oshun_batch_run(oshun_batch* b, int i) {
  if (!b || i < 0 || i >= static_cast<int>(b->runs.size())) return nullptr;
  return b->runs[i];
}

int oshun_batch_step(oshun_batch* b, int nsteps) {
  if (!b || nsteps < 0) return OSHUN_ERR_ARG;
  return batch_steps(b, nsteps);
}

int oshun_batch_synchronize(oshun_batch* b) {
  if (!b) return OSHUN_ERR_ARG;
  BCHK(b, hipStreamSynchronize(b->stream));
  return OSHUN_OK;
}

int oshun_batch_autotune(oshun_batch* b, int steps, int* chosen_rows, double* ms_per_step) {
  if (!b || steps < 1) return OSHUN_ERR_ARG;
  BCHK(b, hipStreamSynchronize(b->stream));
  // keep the state: every candidate starts from it and it is restored after
  const int nr = static_cast<int>(b->runs.size());
  std::vector<void*> save(nr, nullptr), savef(nr, nullptr);
  std::vector<double> t0(nr);
  auto release = [&]() {
    for (void* p : save) if (p) hipFree(p);
    for (void* p : savef) if (p) hipFree(p);
  };
  for (int i = 0; i < nr; ++i) {
    oshun_ctx* c = b->runs[i];
    const size_t fb = (size_t)c->nc * c->plane * c->esz, fl = (size_t)6 * c->nxy * sizeof(double);
    if (hipMalloc(&save[i], fb) != hipSuccess || hipMalloc(&savef[i], fl) != hipSuccess) {
      release();
      return bfail(b, OSHUN_ERR_ALLOC, "autotune: no memory for the state copy");
    }
    hipMemcpyAsync(save[i], c->S, fb, hipMemcpyDeviceToDevice, b->stream);
    hipMemcpyAsync(savef[i], c->Sf, fl, hipMemcpyDeviceToDevice, b->stream);
    t0[i] = c->t_rk;
  }
  auto restore = [&]() {
    for (int i = 0; i < nr; ++i) {
      oshun_ctx* c = b->runs[i];
      hipMemcpyAsync(c->S, save[i], (size_t)c->nc * c->plane * c->esz, hipMemcpyDeviceToDevice, b->stream);
      hipMemcpyAsync(c->Sf, savef[i], (size_t)6 * c->nxy * sizeof(double), hipMemcpyDeviceToDevice, b->stream);
      c->t_rk = c->t_ee = c->t_eb = t0[i];
    }
    hipStreamSynchronize(b->stream);
  };
  // candidates: tile kernels with 6, 10 (and in FP32 14) p rows per tile;
  // marched kernels with 6, 8 or 12 p rows per wavefront (negative)
  std::vector<int> cand = {2, 3};
  if (b->fp32) cand.push_back(4);   // 14 rows (float only: double would exceed 256 VGPRs)
  for (int pw : {6, 8, 12}) cand.push_back(-pw);
  double best = 1e300;
  int best_r = b->pair_r > 0 ? b->pair_r : -b->march_pw;
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  auto choose = [b](int r) {
    if (r > 0) b->pair_r = r;
    else { b->pair_r = 0; b->march_pw = -r; }
  };
  for (int r : cand) {
    choose(r);
    int rc = build_graphs(b);
    if (!rc) rc = batch_steps(b, 1);  // warm-up
    if (!rc) {
      hipEventRecord(e0, b->stream);
      rc = batch_steps(b, steps);
      hipEventRecord(e1, b->stream);
      hipEventSynchronize(e1);
    }
    restore();
    if (rc) { release(); hipEventDestroy(e0); hipEventDestroy(e1); return rc; }
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    if (ms / steps < best) { best = ms / steps; best_r = r; }
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  release();
  choose(best_r);
  const int rc = build_graphs(b);
  if (chosen_rows) *chosen_rows = best_r > 0 ? 4 * best_r - 2 : best_r;
  if (ms_per_step) *ms_per_step = best;
  return rc;
}

int oshun_batch_pair_rows(oshun_batch* b, int set, int* rows) {
  if (!b) return OSHUN_ERR_ARG;
  if (set != 0) {
    BCHK(b, hipStreamSynchronize(b->stream));
    if (set == 6 || set == 10 || (set == 14 && b->fp32)) b->pair_r = (set + 2) / 4;
    else if (set <= -4 && set >= -32) { b->pair_r = 0; b->march_pw = -set; }
    else return bfail(b, OSHUN_ERR_ARG, "tile height must be 6, 10 (14 in FP32); -4 .. -32 for the marched kernels");
    const int rc = build_graphs(b);
    if (rc) return rc;
  }
  if (rows) *rows = b->pair_r > 0 ? 4 * b->pair_r - 2 : -b->march_pw;
  return OSHUN_OK;
}

}  // extern "C"
