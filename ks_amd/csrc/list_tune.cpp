// Tile-shape autotuner of the batched summation (dgsks_list, ks_list_plan_*).
//
// A task of the persistent launch is cut into tiles of 64 target rows by
// `slice` sources, run either one row per lane on the vector units or as four
// 16 x 16 blocks on the FP64 matrix cores.  Which shape is fastest depends on
// the kernel type (the cost of one evaluation) and on k (the cost of the cross
// term and of the loads), so the shape is chosen per (kernel type, k,
// precision) on each GPU architecture by timing every candidate on a synthetic
// batch of that type --
// 256 tasks with m, n log-uniform in [64, 8192] over a pool of 262144 points
// uniform in [0,1)^k, i.e. enough tiles to fill the GPU several times and a pool
// larger than L2 -- and keeping the fastest.  Results are cached for the
// process and, when KS_TUNE_CACHE names a file, in that file (one line
// "arch type k precision slice tile" per key, arch being the device's gfx
// name, e.g. gfx950), which later processes read first; lines of another
// architecture, or in another format, are ignored.
// KS_AUTOTUNE=0 switches tuning off (the default shape is then used).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "host_util.h"
#include "list_launch.h"

namespace gsks {

TileChoice list_default_tile(int type, int k) {
  return TileChoice{1024, list_mfma_eligible(type, k) ? 1 : 0};
}

namespace {

using Key = std::tuple<std::string, int, int, int>;
std::mutex g_mu;
std::map<Key, TileChoice> g_tuned;
bool g_loaded = false;

const char *cache_path() {
  const char *e = getenv("KS_TUNE_CACHE");
  return (e && *e) ? e : nullptr;
}

bool autotune_enabled() {
  const char *e = getenv("KS_AUTOTUNE");
  return !(e && strcmp(e, "0") == 0);
}

// The current device's architecture without its feature suffix ("gfx950"
// from "gfx950:sramecc+:xnack-"); "unknown" if the runtime cannot say.
// Looked up once per device.
std::string device_arch() {
  static std::mutex mu;
  static std::map<int, std::string> by_device;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return "unknown";
  std::lock_guard<std::mutex> lk(mu);
  auto it = by_device.find(dev);
  if (it != by_device.end()) return it->second;
  hipDeviceProp_t prop;
  std::string a = "unknown";
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    a = prop.gcnArchName;
    const size_t colon = a.find(':');
    if (colon != std::string::npos) a.resize(colon);
    if (a.empty()) a = "unknown";
  }
  by_device[dev] = a;
  return a;
}

void load_cache_locked() {
  if (g_loaded) return;
  g_loaded = true;
  const char *path = cache_path();
  if (!path) return;
  FILE *f = fopen(path, "r");
  if (!f) return;
  char line[256], arch[64];
  int type, k, prec, slice, tile;
  while (fgets(line, sizeof(line), f)) {
    if (sscanf(line, "%63s %d %d %d %d %d", arch, &type, &k, &prec, &slice, &tile) != 6) continue;
    if (strncmp(arch, "gfx", 3) != 0) continue;
    if (slice > 0 && (tile == 0 || tile == 1))
      g_tuned[Key(std::string(arch), type, k, prec)] = TileChoice{slice, tile};
  }
  fclose(f);
}

void save_entry(const Key &key, const TileChoice &tc) {
  const char *path = cache_path();
  if (!path) return;
  FILE *f = fopen(path, "a");
  if (!f) return;
  fprintf(f, "%s %d %d %d %d %d\n", std::get<0>(key).c_str(), std::get<1>(key), std::get<2>(key),
          std::get<3>(key), tc.slice, tc.tile);
  fclose(f);
}

// Deterministic 64-bit LCG stream mapped to [0, 1).
struct Lcg {
  unsigned long long x;
  double next() {
    x = x * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  }
};

// Device copies of the synthetic batch of one kernel type.
struct Synthetic {
  std::vector<TaskDesc> tasks;
  long total_a = 0, total_b = 0;
  int *umap = nullptr, *amap = nullptr, *bmap = nullptr, *wmap = nullptr;
  double *X = nullptr, *X2 = nullptr, *w = nullptr, *h = nullptr, *u = nullptr;
  ~Synthetic() {
    for (void *p : {(void *)umap, (void *)amap, (void *)bmap, (void *)wmap, (void *)X, (void *)X2,
                    (void *)w, (void *)h, (void *)u})
      if (p) (void)hipFree(p);
  }
};

template <typename T>
bool upload(T **dst, const std::vector<T> &src) {
  if (hipMalloc((void **)dst, std::max<size_t>(src.size(), 1) * sizeof(T)) != hipSuccess) {
    *dst = nullptr;
    return false;
  }
  return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

int make_synthetic(int type, int k, Synthetic &S) {
  const int N = 1 << 18, ntask = 256;
  Lcg r{0x6b73616d64ull + (unsigned long long)(type * 131 + k)};
  std::vector<double> X((size_t)N * k), X2(N), w(N), h(N);
  for (auto &x : X) x = r.next();
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
    for (int p = 0; p < k; ++p) s += X[(size_t)i * k + p] * X[(size_t)i * k + p];
    X2[i] = s;
    w[i] = 2.0 * r.next() - 1.0;
    h[i] = 1.0 / (0.2 + 0.2 * r.next());
  }
  std::vector<int> amap, bmap;
  ks_t kd;
  std::memset(&kd, 0, sizeof(kd));
  set_ker_type(&kd, type);
  kd.scal = (type == KS_POLYNOMIAL || type == KS_TANH) ? 1.0 / k : -12.5;
  kd.cons = (type == KS_TANH) ? 0.1 : 1.0;
  kd.powe = 2.0;
  KerParams P;
  int rc = make_params(&kd, k, P, true);
  if (rc != KS_OK && rc != KS_ERR_LAPLACE_K) return rc;
  for (int t = 0; t < ntask; ++t) {
    const int m = (int)std::floor(std::exp(std::log(64.0) + r.next() * std::log(128.0)));
    const int n = (int)std::floor(std::exp(std::log(64.0) + r.next() * std::log(128.0)));
    TaskDesc td;
    td.type = type;
    td.P = P;
    td.a0 = (long)amap.size();
    td.b0 = (long)bmap.size();
    td.m = m;
    td.n = n;
    td.hi = td.hj = nullptr;
    const int sa = (int)(r.next() * N), sb = (int)(r.next() * N);
    for (int i = 0; i < m; ++i) amap.push_back((int)(((long)sa + (long)i * 7919) % N));
    for (int j = 0; j < n; ++j) bmap.push_back((int)(((long)sb + (long)j * 104729) % N));
    S.tasks.push_back(td);
  }
  S.total_a = (long)amap.size();
  S.total_b = (long)bmap.size();
  std::vector<double> u0(N, 0.0);
  bool ok = upload(&S.umap, amap) && upload(&S.amap, amap) && upload(&S.bmap, bmap) &&
            upload(&S.wmap, bmap) && upload(&S.X, X) && upload(&S.X2, X2) && upload(&S.w, w) &&
            upload(&S.h, h) && upload(&S.u, u0);
  if (!ok) return KS_ERR_ALLOC;
  for (auto &td : S.tasks)
    if (type == KS_GAUSSIAN_VAR_BANDWIDTH) td.hi = td.hj = S.h;
  return KS_OK;
}

// Median time of one run of the batch with shape tc (ms), or a negative value on failure.
float time_shape(const Synthetic &S, int k, const TileChoice &tc, hipStream_t s) {
  ListPlan *pl = nullptr;
  if (list_plan_build(&pl, S.tasks, k, S.umap, S.amap, S.bmap, S.wmap, S.total_a, KS_PREC_FP64,
                      &tc, s, 0) != KS_OK)
    return -1.0f;
  ListArrays L;
  L.u = S.u;
  L.XA = L.XB = S.X;
  L.XA2 = L.XB2 = S.X2;
  L.w = S.w;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  std::vector<float> ms;
  bool ok = list_plan_exec(*pl, L, s) == KS_OK;  // warm-up
  for (int rep = 0; ok && rep < 3; ++rep) {
    (void)hipEventRecord(e0, s);
    ok = list_plan_exec(*pl, L, s) == KS_OK;
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) ok = false;
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, e0, e1);
    ms.push_back(t);
  }
  (void)hipStreamSynchronize(s);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  list_plan_free(pl);
  if (!ok || ms.empty()) return -1.0f;
  std::sort(ms.begin(), ms.end());
  return ms[ms.size() / 2];
}

TileChoice tune(int type, int k) {
  const TileChoice dflt = list_default_tile(type, k);
  hipStream_t s = thread_stream();
  if (!s) return dflt;
  Synthetic S;
  if (make_synthetic(type, k, S) != KS_OK) return dflt;
  TileChoice best = dflt;
  float best_ms = -1.0f;
  const int slices[] = {256, 512, 1024, 2048, 4096};
  for (int tile = 0; tile < 2; ++tile) {
    if (tile == 1 && !list_mfma_eligible(type, k)) continue;
    for (int sl : slices) {
      const TileChoice tc{sl, tile};
      const float t = time_shape(S, k, tc, s);
      if (t > 0.0f && (best_ms < 0.0f || t < best_ms)) {
        best_ms = t;
        best = tc;
      }
    }
  }
  return best;
}

}  // namespace

TileChoice list_tile_choice(int type, int k, int precision) {
  const Key key(device_arch(), type, k, precision);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    load_cache_locked();
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) return it->second;
  }
  if (!autotune_enabled()) return list_default_tile(type, k);
  const TileChoice tc = tune(type, k);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_tuned.find(key);
  if (it != g_tuned.end()) return it->second;  // another thread tuned it meanwhile
  g_tuned[key] = tc;
  save_entry(key, tc);
  return tc;
}

}  // namespace gsks

extern "C" int ks_list_autotune(int type, int k, int precision, int *slice, int *tile) {
  using namespace gsks;
  if (type < 0 || type > KS_EPANECHNIKOV || k <= 0)
    return raise_error(KS_ERR_INVALID_ARG, false, "ks_list_autotune(): invalid argument\n");
  if (ensure_device() != KS_OK)
    return raise_error(KS_ERR_DEVICE, true, "ks_list_autotune(): no GPU available\n");
  const TileChoice tc = list_tile_choice(type, k, precision);
  if (slice) *slice = tc.slice;
  if (tile) *tile = tc.tile;
  return KS_OK;
}
