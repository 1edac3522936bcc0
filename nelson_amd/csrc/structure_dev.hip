// Structure stage on the device: the edge-proportional part of the BA
// section's structureReady (nelson/DoubleSection.hpp:83-246; the (W, W)
// pairs of S as MatrixWWtMultiplier::prepare builds them,
// nelson/MatrixWWtMultiplier.hpp:18-120) as radix sorts, scans and
// gathers on the GPU.  The result is host_structure's, field by field:
//   * point-major edges: a stable sort of the edges by (point, camera);
//   * W blocks: the first edge of every run of equal (point, camera) keys
//     whose camera and point are variable starts a block (an inclusive scan
//     numbers them);
//   * camera-major edges of the variable cameras: a stable sort by
//     (variable camera, point);
//   * W blocks grouped by camera: a stable sort of the block numbers by
//     camera;
//   * S pattern: every pair of W blocks of one point, the lower camera
//     first, in (point, a, b) order, stably sorted by (camera i, camera j);
//     the first pair of every run of equal keys starts an S block.
// Stable radix sorts keep equal keys in input order, which is the order the
// host's std::stable_sort calls keep.  The small per-camera and per-block
// lists (S blocks, diagonal blocks first within a camera) are formed on the
// host from the run starts.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>
#include <vector>

#include "internal.h"
#include "structure_dev.h"

namespace nsn {
namespace {

#define DS_TRY(expr)                                                                                          \
  do {                                                                                                        \
    hipError_t e_ = (expr);                                                                                   \
    if (e_ != hipSuccess) return fail(NSN_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                                      __LINE__);                                                              \
  } while (0)

typedef unsigned long long u64;
constexpr int kB = 256;

int nblk(long long n) { return (int)std::max<long long>(1, (n + kB - 1) / kB); }

// bits needed for keys below n
int key_bits(u64 n) {
  int b = 1;
  while (b < 64 && (n >> b) != 0) b++;
  return b;
}

// the device allocations of one structure build, freed together
struct Arena {
  std::vector<void*> ptrs;
  bool failed = false;
  ~Arena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
      (void)hipGetLastError();
      failed = true;
      return nullptr;
    }
    ptrs.push_back(p);
    return (T*)p;
  }
};

__global__ void k_pt_keys(int E, const int* __restrict__ cam, const int* __restrict__ pt, u64 C,
                          u64* __restrict__ key, int* __restrict__ val) {
  const int e = blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  key[e] = (u64)pt[e] * C + (u64)cam[e];
  val[e] = e;
}

// cnt[idx[i] + 1] += 1 for idx[i] >= 0 (a count per index, read by a scan)
__global__ void k_count(int n, const int* __restrict__ idx, int* __restrict__ cnt) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const int v = idx[i];
  if (v >= 0) atomicAdd(&cnt[v + 1], 1);
}

// cnt[map[idx[i]] + 1] += 1 for map[idx[i]] >= 0
__global__ void k_count_mapped(int n, const int* __restrict__ idx, const int* __restrict__ map,
                               int* __restrict__ cnt) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const int v = map[idx[i]];
  if (v >= 0) atomicAdd(&cnt[v + 1], 1);
}

// per-edge data in the order of `perm` (first n entries)
__global__ void k_gather_edges(int n, const int* __restrict__ perm, const int* __restrict__ idx,
                               const double* __restrict__ obs, const double* __restrict__ w,
                               int* __restrict__ idx_out, double* __restrict__ obs_out, double* __restrict__ w_out) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= n) return;
  const int e = perm[k];
  idx_out[k] = idx[e];
  obs_out[2 * (size_t)k] = obs[2 * (size_t)e];
  obs_out[2 * (size_t)k + 1] = obs[2 * (size_t)e + 1];
  w_out[k] = w ? w[e] : 1.0;
}

// a W block starts at the first edge of a run of equal (point, camera) keys
// whose camera and point are variable
__global__ void k_w_flags(int E, const u64* __restrict__ skey, u64 C, const int* __restrict__ cam_var,
                          const int* __restrict__ pt_var, int* __restrict__ newf, int* __restrict__ dup) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= E) return;
  const u64 key = skey[k];
  const bool valid = cam_var[(int)(key % C)] >= 0 && pt_var[(int)(key / C)] >= 0;
  const bool first = k == 0 || skey[k - 1] != key;
  newf[k] = valid && first ? 1 : 0;
  if (valid && !first) *dup = 1;
}

__global__ void k_w_blocks(int E, const u64* __restrict__ skey, u64 C, const int* __restrict__ cam_var,
                           const int* __restrict__ pt_var, const int* __restrict__ newf,
                           const int* __restrict__ incl, int* __restrict__ pe_wblk, int* __restrict__ w_cam,
                           int* __restrict__ w_pt) {
  const int k = blockIdx.x * kB + threadIdx.x;
  if (k >= E) return;
  const u64 key = skey[k];
  const int cv = cam_var[(int)(key % C)], pv = pt_var[(int)(key / C)];
  pe_wblk[k] = (cv >= 0 && pv >= 0) ? incl[k] - 1 : -1;
  if (newf[k]) {
    w_cam[incl[k] - 1] = cv;
    w_pt[incl[k] - 1] = pv;
  }
}

// camera-major keys: (variable camera, point); fixed cameras sort last
__global__ void k_cam_keys(int E, const int* __restrict__ cam, const int* __restrict__ pt,
                           const int* __restrict__ cam_var, u64 P, u64 NU, u64* __restrict__ key,
                           int* __restrict__ val) {
  const int e = blockIdx.x * kB + threadIdx.x;
  if (e >= E) return;
  const int cv = cam_var[cam[e]];
  key[e] = cv >= 0 ? (u64)cv * P + (u64)pt[e] : NU * P;
  val[e] = e;
}

__global__ void k_iota(int n, int* __restrict__ v) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) v[i] = i;
}

__global__ void k_pair_counts(int NV, const int* __restrict__ wptr, long long* __restrict__ cnt) {
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v > NV) return;
  const long long n = v < NV ? wptr[v + 1] - wptr[v] : 0;
  cnt[v] = n * (n - 1) / 2;
}

// the pairs (a, b), a < b, of every point's W blocks, the block of the lower
// camera first, at the point's offset
__global__ void k_pairs(int NV, const int* __restrict__ wptr, const int* __restrict__ w_cam,
                        const long long* __restrict__ off, u64 NU, u64* __restrict__ key, u64* __restrict__ val) {
  const int v = blockIdx.x * kB + threadIdx.x;
  if (v >= NV) return;
  long long q = off[v];
  const int b0 = wptr[v], b1 = wptr[v + 1];
  for (int a = b0; a < b1; a++)
    for (int b = a + 1; b < b1; b++) {
      const int ca = w_cam[a], cb = w_cam[b];
      if (ca < cb) {
        key[q] = (u64)ca * NU + (u64)cb;
        val[q] = ((u64)(unsigned)a << 32) | (unsigned)b;
      } else {
        key[q] = (u64)cb * NU + (u64)ca;
        val[q] = ((u64)(unsigned)b << 32) | (unsigned)a;
      }
      q++;
    }
}

__global__ void k_run_starts(long long n, const u64* __restrict__ key, unsigned char* __restrict__ flag) {
  const long long q = (long long)blockIdx.x * kB + threadIdx.x;
  if (q >= n) return;
  flag[q] = (q == 0 || key[q - 1] != key[q]) ? 1 : 0;
}

// temporary storage of the library calls, grown as needed
struct Temp {
  void* p = nullptr;
  size_t n = 0;
  ~Temp() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return NSN_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return fail(NSN_ENOMEM, "device structure: %zu bytes of sort storage", bytes);
    }
    n = bytes;
    return NSN_OK;
  }
};

template <typename K, typename V>
int sort_pairs(Temp& t, const K* kin, K* kout, const V* vin, V* vout, long long n, int bits, hipStream_t st) {
  if (n <= 0) return NSN_OK;
  if (n > 0x7fffffffLL) return fail(NSN_EINVAL, "device structure: %lld items to sort", n);
  size_t bytes = 0;
  DS_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, bits, st));
  if (int rc = t.ensure(bytes)) return rc;
  DS_TRY(hipcub::DeviceRadixSort::SortPairs(t.p, bytes, kin, kout, vin, vout, (int)n, 0, bits, st));
  return NSN_OK;
}

template <typename T>
int inclusive_sum(Temp& t, const T* in, T* out, long long n, hipStream_t st) {
  if (n <= 0) return NSN_OK;
  size_t bytes = 0;
  DS_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, (int)n, st));
  if (int rc = t.ensure(bytes)) return rc;
  DS_TRY(hipcub::DeviceScan::InclusiveSum(t.p, bytes, in, out, (int)n, st));
  return NSN_OK;
}

template <typename T>
int exclusive_sum(Temp& t, const T* in, T* out, long long n, hipStream_t st) {
  if (n <= 0) return NSN_OK;
  size_t bytes = 0;
  DS_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
  if (int rc = t.ensure(bytes)) return rc;
  DS_TRY(hipcub::DeviceScan::ExclusiveSum(t.p, bytes, in, out, (int)n, st));
  return NSN_OK;
}

template <typename T>
int to_host(std::vector<T>& v, const T* d, size_t n, hipStream_t st) {
  v.resize(n);
  if (n) DS_TRY(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st));
  return NSN_OK;
}

// counts per index (idx[i] >= 0, or map[idx[i]] >= 0 with a map) -> the
// pointer array ptr[0..N] (ptr[0] = 0), on the host and (d_out) the device
int count_ptr(Arena& ar, Temp& t, const int* idx, const int* map, int n, int N, std::vector<int>& ptr,
              hipStream_t st, int** d_out = nullptr) {
  int* cnt = ar.get<int>(N + 1);
  int* d_ptr = ar.get<int>(N + 1);
  if (d_out) *d_out = d_ptr;
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: counts of %d entries", N + 1);
  DS_TRY(hipMemsetAsync(cnt, 0, (N + 1) * sizeof(int), st));
  if (n > 0) {
    if (map) hipLaunchKernelGGL(k_count_mapped, dim3(nblk(n)), dim3(kB), 0, st, n, idx, map, cnt);
    else hipLaunchKernelGGL(k_count, dim3(nblk(n)), dim3(kB), 0, st, n, idx, cnt);
    DS_TRY(hipGetLastError());
  }
  if (int rc = inclusive_sum(t, cnt, d_ptr, N + 1, st)) return rc;
  return to_host(ptr, d_ptr, N + 1, st);
}

// the inputs on the device
struct DevInput {
  int *cam = nullptr, *pt = nullptr;
  double *obs = nullptr, *w = nullptr;
};

// the edge-proportional fields of s for its current numbering
int edge_pass(const DevInput& in, const nsn_problem_desc* d, HostStructure& s, hipStream_t st) {
  const int C = s.C, P = s.P, E = s.E, NU = s.NU, NV = s.NV;
  Arena ar;
  Temp t;
  int* cam_var = ar.get<int>(C);
  int* pt_var = ar.get<int>(P);
  u64* key = ar.get<u64>(E);
  u64* skey = ar.get<u64>(E);
  int* val = ar.get<int>(E);
  int* pe = ar.get<int>(E);
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: edge arrays of %d edges", E);
  if (C) DS_TRY(hipMemcpyAsync(cam_var, s.cam_var.data(), C * sizeof(int), hipMemcpyHostToDevice, st));
  if (P) DS_TRY(hipMemcpyAsync(pt_var, s.pt_var.data(), P * sizeof(int), hipMemcpyHostToDevice, st));
  // ---- point-major edges: stable by (point, camera) ----
  if (E) {
    hipLaunchKernelGGL(k_pt_keys, dim3(nblk(E)), dim3(kB), 0, st, E, in.cam, in.pt, (u64)C, key, val);
    DS_TRY(hipGetLastError());
  }
  if (int rc = sort_pairs(t, key, skey, val, pe, E, key_bits((u64)P * (u64)C), st)) return rc;
  if (int rc = count_ptr(ar, t, in.pt, nullptr, E, P, s.pt_ptr, st)) return rc;
  int* pe_cam = ar.get<int>(E);
  double* pe_obs = ar.get<double>(2 * (size_t)E);
  double* pe_w = ar.get<double>(E);
  int* newf = ar.get<int>(E);
  int* incl = ar.get<int>(E);
  int* pe_wblk = ar.get<int>(E);
  int* w_cam = ar.get<int>(E);
  int* w_pt = ar.get<int>(E);
  int* dup = ar.get<int>(1);
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: edge arrays of %d edges", E);
  DS_TRY(hipMemsetAsync(dup, 0, sizeof(int), st));
  if (E) {
    hipLaunchKernelGGL(k_gather_edges, dim3(nblk(E)), dim3(kB), 0, st, E, pe, in.cam, in.obs, in.w, pe_cam, pe_obs,
                       pe_w);
    hipLaunchKernelGGL(k_w_flags, dim3(nblk(E)), dim3(kB), 0, st, E, skey, (u64)C, cam_var, pt_var, newf, dup);
    DS_TRY(hipGetLastError());
  }
  if (int rc = inclusive_sum(t, newf, incl, E, st)) return rc;
  if (E) {
    hipLaunchKernelGGL(k_w_blocks, dim3(nblk(E)), dim3(kB), 0, st, E, skey, (u64)C, cam_var, pt_var, newf, incl,
                       pe_wblk, w_cam, w_pt);
    DS_TRY(hipGetLastError());
  }
  int nW = 0, hdup = 0;
  if (E) DS_TRY(hipMemcpyAsync(&nW, incl + (E - 1), sizeof(int), hipMemcpyDeviceToHost, st));
  DS_TRY(hipMemcpyAsync(&hdup, dup, sizeof(int), hipMemcpyDeviceToHost, st));
  DS_TRY(hipStreamSynchronize(st));
  s.nW = nW;
  s.w_dup = hdup != 0;
  int* wptr = nullptr;
  if (int rc = count_ptr(ar, t, w_pt, nullptr, nW, NV, s.wptr, st, &wptr)) return rc;
  if (int rc = to_host(s.pe, pe, E, st)) return rc;
  if (int rc = to_host(s.pe_cam, pe_cam, E, st)) return rc;
  if (int rc = to_host(s.pe_wblk, pe_wblk, E, st)) return rc;
  if (int rc = to_host(s.pe_obs, pe_obs, 2 * (size_t)E, st)) return rc;
  if (int rc = to_host(s.pe_w, pe_w, E, st)) return rc;
  if (int rc = to_host(s.w_cam, w_cam, nW, st)) return rc;
  if (int rc = to_host(s.w_pt, w_pt, nW, st)) return rc;
  // ---- camera-major edges of the variable cameras: stable by (camera, point) ----
  if (E) {
    hipLaunchKernelGGL(k_cam_keys, dim3(nblk(E)), dim3(kB), 0, st, E, in.cam, in.pt, cam_var, (u64)P, (u64)NU, key,
                       val);
    DS_TRY(hipGetLastError());
  }
  int* ce = newf;  // free again: reused for the camera-major order
  if (int rc = sort_pairs(t, key, skey, val, ce, E, key_bits((u64)NU * (u64)P + 1), st)) return rc;
  if (int rc = count_ptr(ar, t, in.cam, cam_var, E, NU, s.cam_ptr, st)) return rc;
  DS_TRY(hipStreamSynchronize(st));
  s.Ec = s.cam_ptr[NU];
  const int Ec = s.Ec;
  if (Ec) {
    hipLaunchKernelGGL(k_gather_edges, dim3(nblk(Ec)), dim3(kB), 0, st, Ec, ce, in.pt, in.obs, in.w, pe_cam, pe_obs,
                       pe_w);
    DS_TRY(hipGetLastError());
  }
  if (int rc = to_host(s.ce, ce, Ec, st)) return rc;
  if (int rc = to_host(s.ce_pt, pe_cam, Ec, st)) return rc;
  if (int rc = to_host(s.ce_obs, pe_obs, 2 * (size_t)Ec, st)) return rc;
  if (int rc = to_host(s.ce_w, pe_w, Ec, st)) return rc;
  // ---- W blocks grouped by camera: stable by camera ----
  int* cw_idx = pe_wblk;  // free again
  int* kidx = incl;
  if (nW) {
    hipLaunchKernelGGL(k_iota, dim3(nblk(nW)), dim3(kB), 0, st, nW, kidx);
    DS_TRY(hipGetLastError());
  }
  // (the camera numbers are non-negative: sorted as unsigned keys)
  unsigned* wcam_sorted = (unsigned*)val;
  if (int rc = sort_pairs(t, (const unsigned*)w_cam, wcam_sorted, kidx, cw_idx, nW, key_bits((u64)NU), st)) return rc;
  if (int rc = count_ptr(ar, t, w_cam, nullptr, nW, NU, s.cw_ptr, st)) return rc;
  if (int rc = to_host(s.cw_idx, cw_idx, nW, st)) return rc;
  // ---- S pattern: pairs of W blocks sharing a point, grouped by camera pair ----
  long long* pcnt = ar.get<long long>(NV + 1);
  long long* poff = ar.get<long long>(NV + 1);
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: pair counts of %d points", NV);
  hipLaunchKernelGGL(k_pair_counts, dim3(nblk(NV + 1)), dim3(kB), 0, st, NV, wptr, pcnt);
  DS_TRY(hipGetLastError());
  if (int rc = exclusive_sum(t, pcnt, poff, NV + 1, st)) return rc;
  long long npairs = 0;
  DS_TRY(hipMemcpyAsync(&npairs, poff + NV, sizeof(long long), hipMemcpyDeviceToHost, st));
  DS_TRY(hipStreamSynchronize(st));
  if (npairs > 0x7fffffffLL) return fail(NSN_EINVAL, "%lld (W, W) pairs: more than the pair lists index", npairs);
  s.npairs = npairs;
  u64* pkey = ar.get<u64>(npairs);
  u64* pkey_s = ar.get<u64>(npairs);
  u64* pval = ar.get<u64>(npairs);
  u64* pval_s = ar.get<u64>(npairs);
  unsigned char* pflag = ar.get<unsigned char>(npairs);
  u64* ukey = ar.get<u64>(npairs);
  int* ustart = ar.get<int>(npairs);
  int* nsel = ar.get<int>(1);
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: %lld (W, W) pairs", npairs);
  int nu = 0;
  if (npairs > 0) {
    hipLaunchKernelGGL(k_pairs, dim3(nblk(NV)), dim3(kB), 0, st, NV, wptr, w_cam, poff, (u64)NU, pkey, pval);
    DS_TRY(hipGetLastError());
    if (int rc = sort_pairs(t, pkey, pkey_s, pval, pval_s, npairs, key_bits((u64)NU * (u64)NU), st)) return rc;
    hipLaunchKernelGGL(k_run_starts, dim3(nblk(npairs)), dim3(kB), 0, st, npairs, pkey_s, pflag);
    DS_TRY(hipGetLastError());
    size_t bytes = 0;
    DS_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, pkey_s, pflag, ukey, nsel, (int)npairs, st));
    if (int rc = t.ensure(bytes)) return rc;
    DS_TRY(hipcub::DeviceSelect::Flagged(t.p, bytes, pkey_s, pflag, ukey, nsel, (int)npairs, st));
    hipcub::CountingInputIterator<int> pos(0);
    bytes = 0;
    DS_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, pos, pflag, ustart, nsel, (int)npairs, st));
    if (int rc = t.ensure(bytes)) return rc;
    DS_TRY(hipcub::DeviceSelect::Flagged(t.p, bytes, pos, pflag, ustart, nsel, (int)npairs, st));
    DS_TRY(hipMemcpyAsync(&nu, nsel, sizeof(int), hipMemcpyDeviceToHost, st));
    DS_TRY(hipStreamSynchronize(st));
  }
  std::vector<u64> hkey, hval;
  std::vector<int> hstart;
  if (int rc = to_host(hkey, ukey, nu, st)) return rc;
  if (int rc = to_host(hstart, ustart, nu, st)) return rc;
  if (int rc = to_host(hval, pval_s, npairs, st)) return rc;
  DS_TRY(hipStreamSynchronize(st));
  // S blocks: per camera its diagonal block (no pairs), then its
  // off-diagonal blocks by column
  s.sb_i.clear();
  s.sb_j.clear();
  s.sb_ptr.clear();
  int uq = 0;
  for (int u = 0; u < NU; u++) {
    s.sb_i.push_back(u);
    s.sb_j.push_back(u);
    s.sb_ptr.push_back(uq < nu ? hstart[uq] : (int)npairs);
    while (uq < nu && (int)(hkey[uq] / (u64)NU) == u) {
      s.sb_i.push_back(u);
      s.sb_j.push_back((int)(hkey[uq] % (u64)NU));
      s.sb_ptr.push_back(hstart[uq]);
      uq++;
    }
  }
  s.sb_ptr.push_back((int)npairs);
  s.nsb = (int)s.sb_i.size();
  s.sb_pairs.resize(2 * (size_t)npairs);
  for (long long q = 0; q < npairs; q++) {
    s.sb_pairs[2 * q] = (int)(hval[q] >> 32);
    s.sb_pairs[2 * q + 1] = (int)(hval[q] & 0xffffffffu);
  }
  (void)d;
  return NSN_OK;
}

}  // namespace

int device_structure(const nsn_problem_desc* d, HostStructure& s, bool allow_order, hipStream_t st) {
  if (int rc = number_variables(d, s)) return rc;
  const int E = s.E;
  Arena ar;
  DevInput in;
  in.cam = ar.get<int>(E);
  in.pt = ar.get<int>(E);
  in.obs = ar.get<double>(2 * (size_t)E);
  if (d->weight) in.w = ar.get<double>(E);
  if (ar.failed) return fail(NSN_ENOMEM, "device structure: inputs of %d edges", E);
  if (E) {
    DS_TRY(hipMemcpyAsync(in.cam, d->cam_idx, E * sizeof(int), hipMemcpyHostToDevice, st));
    DS_TRY(hipMemcpyAsync(in.pt, d->pt_idx, E * sizeof(int), hipMemcpyHostToDevice, st));
    DS_TRY(hipMemcpyAsync(in.obs, d->obs, 2 * (size_t)E * sizeof(double), hipMemcpyHostToDevice, st));
    if (d->weight) DS_TRY(hipMemcpyAsync(in.w, d->weight, E * sizeof(double), hipMemcpyHostToDevice, st));
  }
  if (allow_order && wants_camera_order(d, s)) {
    // a first pass in the given order gives the camera graph: the
    // off-diagonal S blocks are the pairs of variable cameras sharing a
    // variable point
    if (int rc = edge_pass(in, d, s, st)) return rc;
    apply_camera_order(d, s, camera_graph_from_blocks(s));
  }
  if (int rc = edge_pass(in, d, s, st)) return rc;
  w_ref_order(s);
  DS_TRY(hipStreamSynchronize(st));
  return NSN_OK;
}

int structure_diff(const HostStructure& a, const HostStructure& b, const char** what) {
  int n = 0;
  const char* first = nullptr;
  auto note = [&](bool same, const char* name) {
    if (!same) {
      n++;
      if (!first) first = name;
    }
  };
#define SD(f) note(a.f == b.f, #f)
  SD(C); SD(P); SD(E); SD(NU); SD(NV); SD(nW); SD(Ec); SD(nsb); SD(npairs);
  SD(cam_var); SD(pt_var); SD(var_cam_user); SD(var_pt_user); SD(var_cam_api);
  SD(pe); SD(pt_ptr); SD(pe_cam); SD(pe_wblk); SD(pe_w); SD(wptr); SD(w_cam); SD(w_pt); SD(w_dup);
  SD(ce); SD(cam_ptr); SD(ce_pt); SD(ce_w); SD(cw_ptr); SD(cw_idx); SD(w_ref);
  SD(sb_i); SD(sb_j); SD(sb_ptr); SD(sb_pairs);
#undef SD
  // observations compared by bit pattern
  note(a.pe_obs.size() == b.pe_obs.size() &&
           std::memcmp(a.pe_obs.data(), b.pe_obs.data(), a.pe_obs.size() * sizeof(double)) == 0,
       "pe_obs");
  note(a.ce_obs.size() == b.ce_obs.size() &&
           std::memcmp(a.ce_obs.data(), b.ce_obs.data(), a.ce_obs.size() * sizeof(double)) == 0,
       "ce_obs");
  if (what) *what = first;
  return n;
}

}  // namespace nsn
