// C ABI of the batched-window solver: many independent BA problems (uneven
// camera and point counts) solved by Levenberg-Marquardt in one persistent
// launch (batch_kernels.hip), with the launch shape picked by an autotuner
// that times the candidates on the batch itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "batch.h"
#include "internal.h"
#include "nelson_amd.h"
#include "structure.h"

using namespace nsn;

#define HIP_TRY(expr)                                                                         \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail(NSN_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                                  \
  } while (0)

namespace {

constexpr int kMaxWinS = 256;  // largest S order of a window (6 variable cameras each)
const int kNbCand[3] = {8, 16, 64};
const int kGridCand[3] = {256, 512, 1024};

}  // namespace

struct nsn_batch {
  int device = 0, prec = NSN_FP64, nwin = 0;
  hipStream_t st = nullptr;
  std::vector<void*> allocs;
  std::vector<int> C, P, E;
  long long totC = 0, totP = 0;
  std::vector<BWin> hw;
  BWin* dw = nullptr;
  DevCtl* ctls = nullptr;
  DevCtl* hctls = nullptr;  // pinned
  double *cams = nullptr, *pts = nullptr, *cams_save = nullptr, *pts_save = nullptr;
  int* counter = nullptr;
  int nb = 0, grid = 0;
  int ir_steps = 0;  // refinement steps per increment (mixed precision)
  bool tuned = false;
  double tune_ms[9] = {};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  // one instantiated graph (counter reset + persistent kernel) per launch
  // shape (variant, grid, refined or not)
  std::map<std::tuple<int, int, bool>, hipGraphExec_t> graphs;

  ~nsn_batch() {
    if (st) (void)hipStreamSynchronize(st);
    for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (void* p : allocs) (void)hipFree(p);
    if (hctls) (void)hipHostFree(hctls);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
  }
  void* alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    allocs.push_back(p);
    return p;
  }
};

namespace {

size_t esz(int prec) { return prec == NSN_MIXED ? sizeof(float) : sizeof(double); }

// the default refinement of a mixed-precision batch: none.  Each step
// evaluates the FP64 residual from every edge twice (point rows, camera
// rows), and on the 512-window local-BA workload 1.55 steps per increment
// (at most 2, tol 1e-10) took a 10-iteration solve from 19.8 ms to 45.5 ms,
// beyond the FP64 batch's 28.9 ms; nsn_batch_set_refinement turns it on
constexpr int kBatchIrMaxSteps = 0;
constexpr double kBatchIrTol = 1e-10;

// the S assembly's work list: every block's terms cut into chunks of at most
// kSChunk, so that one lane sums one chunk whatever the block sizes
// (MatrixWWtMultiplier.hpp:124-153 sums the same terms block by block)
struct SChunks {
  std::vector<int> terms, ch, blk_ch_ptr, sb_diag;
};

void s_chunks(const HostStructure& s, SChunks& o) {
  std::vector<int> beg, end;
  o.terms.clear();
  o.blk_ch_ptr.assign(s.nsb + 1, 0);
  o.sb_diag.assign(s.NU, 0);
  std::vector<int> diag;
  for (int b = 0; b < s.nsb; b++) {
    const int t0 = (int)o.terms.size() / 2;
    const bool dg = s.sb_i[b] == s.sb_j[b];
    if (dg) {
      const int u = s.sb_i[b];
      o.sb_diag[u] = b;
      for (int kk = s.cw_ptr[u]; kk < s.cw_ptr[u + 1]; kk++) {
        o.terms.push_back(s.cw_idx[kk]);
        o.terms.push_back(s.cw_idx[kk]);
      }
    } else {
      o.terms.insert(o.terms.end(), s.sb_pairs.begin() + 2 * (size_t)s.sb_ptr[b],
                     s.sb_pairs.begin() + 2 * (size_t)s.sb_ptr[b + 1]);
    }
    const int t1 = (int)o.terms.size() / 2;
    for (int t = t0; t < t1; t += kSChunk) {
      beg.push_back(t);
      end.push_back(std::min(t1, t + kSChunk));
      diag.push_back(dg ? 1 : 0);
    }
    o.blk_ch_ptr[b + 1] = (int)beg.size();
  }
  const int nch = (int)beg.size();
  std::vector<int> ord(nch);
  for (int c = 0; c < nch; c++) ord[c] = c;
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return end[x] - beg[x] > end[y] - beg[y]; });
  o.ch.assign(4 * (size_t)nch, 0);
  for (int r = 0; r < nch; r++) {
    o.ch[4 * r] = beg[ord[r]];
    o.ch[4 * r + 1] = end[ord[r]];
    o.ch[4 * r + 2] = ord[r];
    o.ch[4 * r + 3] = diag[ord[r]];
  }
}

// the batch solve as a graph: the counter reset and the persistent kernel,
// captured once per (variant, grid) and replayed by every later solve
int batch_graph(nsn_batch* b, int nb, int grid, hipGraphExec_t* out) {
  const bool refine = b->ir_steps > 0;
  auto it = b->graphs.find({nb, grid, refine});
  if (it != b->graphs.end()) {
    *out = it->second;
    return NSN_OK;
  }
  hipGraph_t g = nullptr;
  HIP_TRY(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  launch_batch_reset(b->counter, b->st);
  const int lrc = b->prec == NSN_MIXED
                      ? launch_batch_lm<float>(b->dw, b->nwin, b->ctls, b->counter, nb, grid, refine, b->st)
                      : launch_batch_lm<double>(b->dw, b->nwin, b->ctls, b->counter, nb, grid, refine, b->st);
  const hipError_t ec = hipStreamEndCapture(b->st, &g);
  if (lrc || ec != hipSuccess || !g) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    if (lrc) return fail(NSN_EINVAL, "batch: unsupported factorisation variant %d", nb);
    return fail(NSN_EHIP, "batch: graph capture failed: %s", hipGetErrorString(ec));
  }
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return fail(NSN_EHIP, "batch: graph instantiation failed: %s", hipGetErrorString(ei));
  b->graphs[{nb, grid, refine}] = x;
  *out = x;
  return NSN_OK;
}

int run_batch(nsn_batch* b, const nsn_lm_settings* s, int nb, int grid, float* ms) {
  for (int w = 0; w < b->nwin; w++) {
    DevCtl& c = b->hctls[w];
    std::memset(&c, 0, sizeof(c));
    c.mode = 1;
    c.need_lin = 1;
    c.maxNumIt = s->maxNumIt; c.maxNumSubIt = s->maxNumSubIt; c.minNumIt = s->minNumIt;
    c.epsBVector = s->epsBVector; c.epsIncSquare = s->epsIncSquare; c.epsChi2 = s->epsChi2;
    c.tau = s->tau; c.penaltyV = s->penaltyV; c.initV = s->initV;
    c.maxScaleFactor = s->maxScaleFactor; c.minScaleFactor = s->minScaleFactor;
  }
  hipGraphExec_t graph = nullptr;
  int rc = batch_graph(b, nb, grid, &graph);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(b->ctls, b->hctls, sizeof(DevCtl) * b->nwin, hipMemcpyHostToDevice, b->st));
  HIP_TRY(hipEventRecord(b->e0, b->st));
  HIP_TRY(hipGraphLaunch(graph, b->st));
  HIP_TRY(hipEventRecord(b->e1, b->st));
  HIP_TRY(hipMemcpyAsync(b->hctls, b->ctls, sizeof(DevCtl) * b->nwin, hipMemcpyDeviceToHost, b->st));
  HIP_TRY(hipStreamSynchronize(b->st));
  HIP_TRY(hipEventElapsedTime(ms, b->e0, b->e1));
  return NSN_OK;
}

int save_state(nsn_batch* b, bool restore) {
  const size_t cb = sizeof(double) * 12 * (size_t)b->totC, pb = sizeof(double) * 3 * (size_t)b->totP;
  if (restore) {
    HIP_TRY(hipMemcpyAsync(b->cams, b->cams_save, cb, hipMemcpyDeviceToDevice, b->st));
    HIP_TRY(hipMemcpyAsync(b->pts, b->pts_save, pb, hipMemcpyDeviceToDevice, b->st));
  } else {
    HIP_TRY(hipMemcpyAsync(b->cams_save, b->cams, cb, hipMemcpyDeviceToDevice, b->st));
    HIP_TRY(hipMemcpyAsync(b->pts_save, b->pts, pb, hipMemcpyDeviceToDevice, b->st));
  }
  HIP_TRY(hipStreamSynchronize(b->st));
  return NSN_OK;
}

// times every (factorisation variant, grid) candidate on the batch itself for a few LM
// iterations from the current state (restored after each run) and keeps the
// fastest
int autotune(nsn_batch* b, const nsn_lm_settings* s) {
  nsn_lm_settings t = *s;
  t.maxNumIt = std::min(std::max(1, s->maxNumIt), 2);
  int rc;
  if ((rc = save_state(b, false))) return rc;
  double best = 1e300;
  int k = 0;
  for (int nb : kNbCand)
    for (int g : kGridCand) {
      float ms = 0.f;
      if ((rc = run_batch(b, &t, nb, g, &ms))) return rc;
      if ((rc = save_state(b, true))) return rc;
      b->tune_ms[k++] = ms;
      if (ms < best) {
        best = ms;
        b->nb = nb;
        b->grid = g;
      }
    }
  b->tuned = true;
  return NSN_OK;
}

}  // namespace

extern "C" int nsn_batch_create(const nsn_problem_desc* wd, int n_windows, nsn_batch** out) {
  TraceRange trace_range("nsn_batch_create");
  clear_error();
  if (!out) return fail(NSN_EINVAL, "nsn_batch_create: null argument");
  *out = nullptr;
  if (!wd || n_windows < 1) return fail(NSN_EINVAL, "nsn_batch_create: no windows");
  const int prec = wd[0].precision, dev = wd[0].device;
  if (prec != NSN_FP64 && prec != NSN_MIXED) return fail(NSN_EINVAL, "nsn_batch_create: unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NSN_EHIP, "nsn_batch_create: no HIP device");
  if (dev < 0 || dev >= ndev) return fail(NSN_EINVAL, "nsn_batch_create: bad device %d", dev);
  HIP_TRY(hipSetDevice(dev));
  std::unique_ptr<nsn_batch> b(new nsn_batch());
  b->device = dev;
  b->prec = prec;
  b->ir_steps = prec == NSN_MIXED ? kBatchIrMaxSteps : 0;
  b->nwin = n_windows;
  std::vector<HostStructure> hs(n_windows);
  for (int w = 0; w < n_windows; w++) {
    const nsn_problem_desc* d = wd + w;
    if (d->precision != prec || d->device != dev)
      return fail(NSN_EINVAL, "nsn_batch_create: window %d differs in precision or device", w);
    if (d->n_cameras < 1 || d->n_points < 1 || d->n_edges < 0 ||
        (d->n_edges > 0 && (!d->cam_idx || !d->pt_idx || !d->obs)))
      return fail(NSN_EINVAL, "nsn_batch_create: window %d: bad sizes or missing edge arrays", w);
    if (d->intrinsics)
      return fail(NSN_EINVAL, "nsn_batch_create: window %d: batched windows use the shared pinhole model only", w);
    int rc = host_structure(d, hs[w]);
    if (rc) return rc;
    if (6 * hs[w].NU > kMaxWinS)
      return fail(NSN_EINVAL, "nsn_batch_create: window %d has %d variable cameras (at most %d)", w, hs[w].NU,
                  kMaxWinS / 6);
    b->C.push_back(d->n_cameras);
    b->P.push_back(d->n_points);
    b->E.push_back(d->n_edges);
  }
  // one pool per element type, windows one after the other
  std::vector<int> ipool;
  std::vector<double> dpool;
  size_t tcount = 0;
  struct WinOff {
    size_t cam_var, pt_var, vcu, vpu, pt_ptr, pe_cam, pe_wblk, pe_edge, cam_ptr, ce_pt, wptr, w_cam, w_pt, cw_ptr,
        cw_idx, sb_i, sb_j, terms, ch, blk_ch_ptr, sb_diag;
    int nch;
    size_t pe_obs, pe_w, ce_obs, ce_w, xU, xV, rU, rV, bU64, bV64, rS, zU;
    size_t U, bU, V, bV, W, Vinv, Y, tW, S, Linv, part;
    size_t cams, pts;
  };
  std::vector<WinOff> wo(n_windows);
  auto push_i = [&](const std::vector<int>& v) {
    const size_t o = ipool.size();
    ipool.insert(ipool.end(), v.begin(), v.end());
    return o;
  };
  auto push_d = [&](const std::vector<double>& v) {
    const size_t o = dpool.size();
    dpool.insert(dpool.end(), v.begin(), v.end());
    return o;
  };
  auto take_d = [&](size_t n) {
    const size_t o = dpool.size();
    dpool.resize(dpool.size() + n, 0.0);
    return o;
  };
  auto take_t = [&](size_t n) {
    const size_t o = tcount;
    tcount += (n + 63) / 64 * 64;  // 256-byte aligned sub-arrays
    return o;
  };
  for (int w = 0; w < n_windows; w++) {
    const HostStructure& s = hs[w];
    WinOff& o = wo[w];
    o.cam_var = push_i(s.cam_var); o.pt_var = push_i(s.pt_var);
    o.vcu = push_i(s.var_cam_user); o.vpu = push_i(s.var_pt_user);
    o.pt_ptr = push_i(s.pt_ptr); o.pe_cam = push_i(s.pe_cam); o.pe_wblk = push_i(s.pe_wblk); o.pe_edge = push_i(s.pe);
    o.cam_ptr = push_i(s.cam_ptr); o.ce_pt = push_i(s.ce_pt);
    o.wptr = push_i(s.wptr); o.w_cam = push_i(s.w_cam); o.w_pt = push_i(s.w_pt);
    o.cw_ptr = push_i(s.cw_ptr); o.cw_idx = push_i(s.cw_idx);
    o.sb_i = push_i(s.sb_i); o.sb_j = push_i(s.sb_j);
    {
      SChunks sc;
      s_chunks(s, sc);
      ipool.resize((ipool.size() + 3) / 4 * 4, 0);  // int2 / int4 loads
      o.terms = push_i(sc.terms);
      ipool.resize((ipool.size() + 3) / 4 * 4, 0);
      o.ch = push_i(sc.ch);
      o.blk_ch_ptr = push_i(sc.blk_ch_ptr);
      o.sb_diag = push_i(sc.sb_diag);
      o.nch = (int)sc.ch.size() / 4;
    }
    o.pe_obs = push_d(s.pe_obs); o.pe_w = push_d(s.pe_w); o.ce_obs = push_d(s.ce_obs); o.ce_w = push_d(s.ce_w);
    const int n = 6 * s.NU, nP = (n + 63) / 64 * 64;
    o.xU = take_d(nP); o.xV = take_d(3 * (size_t)s.NV);
    o.rU = take_d(6 * (size_t)s.NU); o.rV = take_d(3 * (size_t)s.NV);
    o.bU64 = take_d(6 * (size_t)s.NU); o.bV64 = take_d(3 * (size_t)s.NV);
    o.rS = take_d(nP); o.zU = take_d(nP);
    o.U = take_t(21 * (size_t)s.NU); o.bU = take_t(6 * (size_t)s.NU);
    o.V = take_t(6 * (size_t)s.NV); o.bV = take_t(3 * (size_t)s.NV);
    o.W = take_t(18 * (size_t)s.nW); o.Vinv = take_t(6 * (size_t)s.NV);
    o.Y = take_t(18 * (size_t)s.nW); o.tW = take_t(6 * (size_t)s.nW);
    o.S = take_t((size_t)nP * nP);
    o.Linv = take_t((size_t)(nP / 64) * 4096);
    o.part = take_t(kPart * (size_t)o.nch);
    o.cams = 12 * (size_t)b->totC;
    o.pts = 3 * (size_t)b->totP;
    b->totC += s.C;
    b->totP += s.P;
  }
  const size_t es = esz(prec);
  int* di = (int*)b->alloc(sizeof(int) * ipool.size());
  double* dd = (double*)b->alloc(sizeof(double) * dpool.size());
  char* dt = (char*)b->alloc(es * tcount);
  b->cams = (double*)b->alloc(sizeof(double) * 12 * (size_t)b->totC);
  b->pts = (double*)b->alloc(sizeof(double) * 3 * (size_t)b->totP);
  b->cams_save = (double*)b->alloc(sizeof(double) * 12 * (size_t)b->totC);
  b->pts_save = (double*)b->alloc(sizeof(double) * 3 * (size_t)b->totP);
  double* bak_c = (double*)b->alloc(sizeof(double) * 12 * (size_t)b->totC);
  double* bak_p = (double*)b->alloc(sizeof(double) * 3 * (size_t)b->totP);
  b->ctls = (DevCtl*)b->alloc(sizeof(DevCtl) * n_windows);
  b->dw = (BWin*)b->alloc(sizeof(BWin) * n_windows);
  b->counter = (int*)b->alloc(sizeof(int) * (1 + (size_t)n_windows));
  if (!di || !dd || !dt || !b->cams || !b->pts || !b->cams_save || !b->pts_save || !bak_c || !bak_p || !b->ctls ||
      !b->dw || !b->counter)
    return fail(NSN_ENOMEM, "nsn_batch_create: device allocation failed (%.1f MB)",
                (ipool.size() * 4.0 + dpool.size() * 8.0 + tcount * es) / 1e6);
  HIP_TRY(hipHostMalloc((void**)&b->hctls, sizeof(DevCtl) * n_windows, hipHostMallocDefault));
  HIP_TRY(hipMemcpy(di, ipool.data(), sizeof(int) * ipool.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dd, dpool.data(), sizeof(double) * dpool.size(), hipMemcpyHostToDevice));
  b->hw.resize(n_windows);
  for (int w = 0; w < n_windows; w++) {
    const HostStructure& s = hs[w];
    const WinOff& o = wo[w];
    const nsn_problem_desc* d = wd + w;
    BWin& x = b->hw[w];
    std::memset(&x, 0, sizeof(x));
    BAView& v = x.v;
    v.C = s.C; v.P = s.P; v.E = s.E; v.NU = s.NU; v.NV = s.NV; v.nW = s.nW;
    v.f = d->focal; v.cx = d->cx; v.cy = d->cy; v.huber = d->huber_delta;
    v.cam_var = di + o.cam_var; v.pt_var = di + o.pt_var; v.var_cam_user = di + o.vcu; v.var_pt_user = di + o.vpu;
    v.pt_ptr = di + o.pt_ptr; v.pe_cam = di + o.pe_cam; v.pe_obs = dd + o.pe_obs; v.pe_w = dd + o.pe_w;
    v.pe_wblk = di + o.pe_wblk; v.pe_edge = di + o.pe_edge;
    v.cam_ptr = di + o.cam_ptr; v.ce_pt = di + o.ce_pt; v.ce_obs = dd + o.ce_obs; v.ce_w = dd + o.ce_w;
    v.wptr = di + o.wptr; v.w_cam = di + o.w_cam; v.w_pt = di + o.w_pt; v.cw_ptr = di + o.cw_ptr;
    v.cw_idx = di + o.cw_idx;
    v.cams = b->cams + o.cams; v.pts = b->pts + o.pts;
    x.n = 6 * s.NU;
    x.nP = (x.n + 63) / 64 * 64;
    x.nsb = s.nsb;
    x.sb_i = di + o.sb_i; x.sb_j = di + o.sb_j;
    x.nch = o.nch;
    x.terms = di + o.terms; x.ch = di + o.ch; x.blk_ch_ptr = di + o.blk_ch_ptr; x.sb_diag = di + o.sb_diag;
    x.part = dt + es * o.part;
    x.U = dt + es * o.U; x.bU = dt + es * o.bU; x.V = dt + es * o.V; x.bV = dt + es * o.bV;
    x.W = dt + es * o.W; x.Vinv = dt + es * o.Vinv; x.Y = dt + es * o.Y; x.tW = dt + es * o.tW;
    x.S = dt + es * o.S;
    x.Linv = dt + es * o.Linv;
    x.xU = dd + o.xU; x.xV = dd + o.xV;
    x.rU = dd + o.rU; x.rV = dd + o.rV; x.bU64 = dd + o.bU64; x.bV64 = dd + o.bV64;
    x.rS = dd + o.rS; x.zU = dd + o.zU;
    x.ir_steps = prec == NSN_MIXED ? kBatchIrMaxSteps : 0;
    x.ir_tol = kBatchIrTol;
    x.cams_bak = bak_c + o.cams; x.pts_bak = bak_p + o.pts;
  }
  HIP_TRY(hipMemcpy(b->dw, b->hw.data(), sizeof(BWin) * n_windows, hipMemcpyHostToDevice));
  {
    // windows in decreasing order of work (edges and S-assembly terms)
    std::vector<long long> work(n_windows);
    for (int w = 0; w < n_windows; w++) work[w] = (long long)hs[w].E + hs[w].nW + hs[w].npairs;
    std::vector<int> order(1 + (size_t)n_windows, 0);
    for (int w = 0; w < n_windows; w++) order[1 + w] = w;
    std::stable_sort(order.begin() + 1, order.end(), [&](int x, int y) { return work[x] > work[y]; });
    HIP_TRY(hipMemcpy(b->counter, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&b->e0));
  HIP_TRY(hipEventCreate(&b->e1));
  *out = b.release();
  return NSN_OK;
}

extern "C" int nsn_batch_destroy(nsn_batch* b) {
  delete b;
  return NSN_OK;
}

extern "C" int nsn_batch_set_state(nsn_batch* b, const double* cams, const double* pts) {
  if (!b || !cams || !pts) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(b->device));
  std::vector<double> rt(12 * (size_t)b->totC);
  for (long long c = 0; c < b->totC; c++) {
    rot_exp(cams + 6 * c, &rt[12 * c]);
    for (int i = 0; i < 3; i++) rt[12 * c + 9 + i] = cams[6 * c + 3 + i];
  }
  HIP_TRY(hipMemcpy(b->cams, rt.data(), sizeof(double) * rt.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->pts, pts, sizeof(double) * 3 * (size_t)b->totP, hipMemcpyHostToDevice));
  return NSN_OK;
}

extern "C" int nsn_batch_get_state(nsn_batch* b, double* cams, double* pts) {
  if (!b) return fail(NSN_EINVAL, "null argument");
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipStreamSynchronize(b->st));
  if (cams) {
    std::vector<double> rt(12 * (size_t)b->totC);
    HIP_TRY(hipMemcpy(rt.data(), b->cams, sizeof(double) * rt.size(), hipMemcpyDeviceToHost));
    for (long long c = 0; c < b->totC; c++) {
      rot_log(&rt[12 * c], cams + 6 * c);
      for (int i = 0; i < 3; i++) cams[6 * c + 3 + i] = rt[12 * c + 9 + i];
    }
  }
  if (pts) HIP_TRY(hipMemcpy(pts, b->pts, sizeof(double) * 3 * (size_t)b->totP, hipMemcpyDeviceToHost));
  return NSN_OK;
}

extern "C" int nsn_batch_tune(nsn_batch* b, const nsn_lm_settings* s, int* nb, int* grid) {
  TraceRange trace_range("nsn_batch_tune");
  if (!b || !s) return fail(NSN_EINVAL, "null argument");
  if (s->maxNumSubIt < 1) return fail(NSN_EINVAL, "maxNumSubIt must be >= 1");
  HIP_TRY(hipSetDevice(b->device));
  int rc = autotune(b, s);
  if (rc) return rc;
  if (nb) *nb = b->nb;
  if (grid) *grid = b->grid;
  return NSN_OK;
}

extern "C" int nsn_batch_set_refinement(nsn_batch* b, int max_steps, double tol) {
  if (!b) return fail(NSN_EINVAL, "null argument");
  if (max_steps < 0 || max_steps > 8) return fail(NSN_EINVAL, "%s: steps must be in 0..8, got %d", __func__, max_steps);
  if (!(tol >= 0.0) || !std::isfinite(tol)) return fail(NSN_EINVAL, "%s: tol must be finite and >= 0", __func__);
  if (b->prec != NSN_MIXED) return NSN_OK;  // FP64 windows are not refined
  b->ir_steps = max_steps;
  for (BWin& x : b->hw) {
    x.ir_steps = max_steps;
    x.ir_tol = tol;
  }
  HIP_TRY(hipSetDevice(b->device));
  HIP_TRY(hipMemcpy(b->dw, b->hw.data(), sizeof(BWin) * b->hw.size(), hipMemcpyHostToDevice));
  return NSN_OK;
}

extern "C" int nsn_batch_set_config(nsn_batch* b, int nb, int grid) {
  if (!b) return fail(NSN_EINVAL, "null argument");
  if ((nb != 8 && nb != 16 && nb != 64) || grid < 1)
    return fail(NSN_EINVAL, "nsn_batch_set_config: variant 8 or 16 (panels of that width) or 64 (tiles), grid >= 1");
  b->nb = nb;
  b->grid = grid;
  b->tuned = true;
  return NSN_OK;
}

extern "C" int nsn_batch_lm_solve(nsn_batch* b, const nsn_lm_settings* s, double* chi2_out, double* elapsed_ms) {
  TraceRange trace_range("nsn_batch_lm_solve");
  if (!b || !s) return fail(NSN_EINVAL, "null argument");
  if (s->maxNumSubIt < 1) return fail(NSN_EINVAL, "maxNumSubIt must be >= 1");
  HIP_TRY(hipSetDevice(b->device));
  int rc;
  if (!b->tuned && (rc = autotune(b, s))) return rc;
  float ms = 0.f;
  if ((rc = run_batch(b, s, b->nb, b->grid, &ms))) return rc;
  if (chi2_out)
    for (int w = 0; w < b->nwin; w++) chi2_out[w] = b->hctls[w].chi2;
  if (elapsed_ms) *elapsed_ms = ms;
  return NSN_OK;
}

extern "C" int nsn_batch_results(nsn_batch* b, int* reasons, int* iterations, int* sub_iterations) {
  if (!b) return fail(NSN_EINVAL, "null argument");
  for (int w = 0; w < b->nwin; w++) {
    const DevCtl& c = b->hctls[w];
    if (reasons) reasons[w] = c.done ? c.reason : NSN_MAX_ITERATION_REACHED;
    if (iterations) iterations[w] = c.iter;
    if (sub_iterations) sub_iterations[w] = c.total_solves;
  }
  return NSN_OK;
}

extern "C" int nsn_batch_refinement_counts(nsn_batch* b, int* steps) {
  if (!b || !steps) return fail(NSN_EINVAL, "null argument");
  for (int w = 0; w < b->nwin; w++) steps[w] = b->hctls[w].ir_count;
  return NSN_OK;
}

extern "C" int nsn_batch_config(nsn_batch* b, int* nb, int* grid, double* tune_ms, int n_tune) {
  if (!b) return fail(NSN_EINVAL, "null argument");
  if (nb) *nb = b->nb;
  if (grid) *grid = b->grid;
  if (tune_ms)
    for (int i = 0; i < std::min(n_tune, 9); i++) tune_ms[i] = b->tune_ms[i];
  return NSN_OK;
}
