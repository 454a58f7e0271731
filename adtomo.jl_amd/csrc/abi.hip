// C ABI of libadtomo_hip.so (declared in include/adtomo_hip.h).
//
// Host-side orchestration only: argument checks, workspace, descriptor
// tables, launches, diagnostics.  All arithmetic runs in the HIP kernels.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <limits>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define ADT_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail(e_ == hipErrorOutOfMemory ? ADT_ENOMEM : ADT_EHIP,                     \
                  std::string(#call) + ": " + hipGetErrorString(e_));                    \
  } while (0)

constexpr int kMaxIter3 = 20;         // Eikonal3D.cpp:74
constexpr int kMaxIter2 = 100;        // Eikonal.h:73
constexpr int kMaxAdjSweeps = 8 * 64;  // cap of the adjoint fixed-point sweeps
constexpr int kWG = 1024;
constexpr int kStreamWG = 512;  // fwd3_stream_kernel: 8 waves
constexpr int kDiagWG = 512;    // fwd3_diag_kernel: 8 waves
// The forward kernels address each array with 32-bit buffer offsets.
constexpr size_t kMaxArrayBytes = 0xFFFFFFF0ull;

// Stream-ordered device allocation that is released when the scope ends.
struct DevBuf {
  void* p = nullptr;
  hipStream_t s = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  int alloc(size_t bytes, hipStream_t st) {
    s = st;
    hipError_t e = hipMallocAsync(&p, bytes ? bytes : 1, st);
    if (e != hipSuccess) {
      p = nullptr;
      return fail(ADT_ENOMEM, std::string("hipMallocAsync: ") + hipGetErrorString(e));
    }
    return ADT_OK;
  }
};

// ------------------------------------------------------------ tuning ---
// Launch settings of the 3D kernels, one set per precision ([0] FP64,
// [1] FP32), read at each launch / graph creation and written by
// adt_set_tune / adt_autotune under a mutex.
std::mutex g_tune_mu;
constexpr int32_t kFuseMaxDefault = 1 << 21;
constexpr int kSegsMax = 8;  // c segments of the diagonal forward (its waves)
adt_tune g_tune[2] = {{adt::kDiagAll, 1024, 1, kFuseMaxDefault, 0.f, 0.f, 0, 0},
                      {adt::kDiagAll, 1024, 1, kFuseMaxDefault, 0.f, 0.f, 0, 0}};

// Box-shape classes: events whose k extent needs the same number of
// 8-node segments, whose shorter horizontal side needs the same number of
// 64-row bands, and whose node count has the same binary order share a
// class.  A class may carry its own setting of the per-event fields
// (diag_max_l, fwd_segs, adj_algo); the launch-wide fields (adj_wg, order,
// fuse_max_n) always come from the global setting.
struct ClassKey {
  int l_blocks, bands, log2n;
  bool operator<(const ClassKey& o) const {
    return l_blocks != o.l_blocks ? l_blocks < o.l_blocks
                                  : (bands != o.bands ? bands < o.bands : log2n < o.log2n);
  }
  bool operator==(const ClassKey& o) const {
    return l_blocks == o.l_blocks && bands == o.bands && log2n == o.log2n;
  }
};
ClassKey class_of(int64_t m, int64_t n, int64_t l) {
  ClassKey k;
  k.l_blocks = (int)((l + 7) / 8);
  k.bands = (int)((std::min(m, n) + 63) / 64);
  int64_t N = m * n * l;
  int lg = 0;
  while (N > 1) {
    N >>= 1;
    ++lg;
  }
  k.log2n = lg;
  return k;
}
std::map<ClassKey, adt_tune> g_classes[2];

template <typename T>
constexpr int tune_index() {
  return std::is_same<T, float>::value ? 1 : 0;
}
// The settings a launch uses, taken together under the lock.
struct TuneSnap {
  adt_tune global;
  std::map<ClassKey, adt_tune> classes;
  // the setting of an event of box (m, n, l): its class's per-event fields
  // over the global setting
  adt_tune for_box(int64_t m, int64_t n, int64_t l) const {
    adt_tune t = global;
    auto it = classes.find(class_of(m, n, l));
    if (it != classes.end()) {
      t.diag_max_l = it->second.diag_max_l;
      t.fwd_segs = it->second.fwd_segs;
      t.adj_algo = it->second.adj_algo;
    }
    return t;
  }
};
template <typename T>
TuneSnap current_snap() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return TuneSnap{g_tune[tune_index<T>()], g_classes[tune_index<T>()]};
}
template <typename T>
adt_tune current_tune() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tune[tune_index<T>()];
}
bool tune_valid(const adt_tune& t) {
  return t.diag_max_l >= 0 && t.diag_max_l <= adt::kDiagAll &&
         (t.adj_wg == 256 || t.adj_wg == 512 || t.adj_wg == 1024) && (t.order == 0 || t.order == 1) &&
         t.fuse_max_n >= 0 && (t.adj_algo == 0 || t.adj_algo == 1) && t.fwd_segs >= 0 &&
         t.fwd_segs <= kSegsMax;
}

// ---- the settings as JSON text (adt_tune_save / adt_tune_load): one
// object per setting, {"dtype": d, <field>: value, ...}, global ones in
// "global" and class ones (with "l_blocks", "bands", "log2n") in "classes".
const char* const kTuneFields[] = {"diag_max_l", "adj_wg", "order", "fuse_max_n",
                                   "fwd_ms", "adj_ms", "adj_algo", "fwd_segs"};

std::string tune_json(int dtype, const adt_tune& t, const ClassKey* k) {
  char buf[512];
  int o = snprintf(buf, sizeof buf, "{\"dtype\": %d, ", dtype);
  if (k)
    o += snprintf(buf + o, sizeof buf - o, "\"l_blocks\": %d, \"bands\": %d, \"log2n\": %d, ",
                  k->l_blocks, k->bands, k->log2n);
  snprintf(buf + o, sizeof buf - o,
           "\"diag_max_l\": %d, \"adj_wg\": %d, \"order\": %d, \"fuse_max_n\": %d, "
           "\"fwd_ms\": %.9g, \"adj_ms\": %.9g, \"adj_algo\": %d, \"fwd_segs\": %d}",
           t.diag_max_l, t.adj_wg, t.order, t.fuse_max_n, (double)t.fwd_ms, (double)t.adj_ms,
           t.adj_algo, t.fwd_segs);
  return buf;
}

// Reads the objects of the text written by tune_json (keys in any order,
// numbers only); every field must be present.
int parse_tune_json(const std::string& txt, adt_tune glob[2], bool have_glob[2],
                    std::map<ClassKey, adt_tune> cls[2]) {
  size_t i = 0;
  std::string section;
  while (i < txt.size()) {
    const char ch = txt[i];
    if (ch == '"') {  // a key outside the setting objects names the section
      const size_t e = txt.find('"', i + 1);
      if (e == std::string::npos) return fail(ADT_EINVAL, "tune file: unterminated string");
      section = txt.substr(i + 1, e - i - 1);
      i = e + 1;
      continue;
    }
    if (ch != '{' || (section != "global" && section != "classes")) {
      ++i;
      continue;
    }
    const size_t e = txt.find('}', i);
    if (e == std::string::npos) return fail(ADT_EINVAL, "tune file: unterminated object");
    const std::string obj = txt.substr(i + 1, e - i - 1);
    i = e + 1;
    std::map<std::string, double> kv;
    size_t j = 0;
    while ((j = obj.find('"', j)) != std::string::npos) {
      const size_t k = obj.find('"', j + 1);
      const size_t c = k == std::string::npos ? k : obj.find(':', k);
      if (c == std::string::npos) return fail(ADT_EINVAL, "tune file: malformed object");
      char* end = nullptr;
      const double v = strtod(obj.c_str() + c + 1, &end);
      if (end == obj.c_str() + c + 1) return fail(ADT_EINVAL, "tune file: value is not a number");
      kv[obj.substr(j + 1, k - j - 1)] = v;
      j = (size_t)(end - obj.c_str());
    }
    auto get = [&](const char* key, double* v) {
      auto it = kv.find(key);
      if (it == kv.end()) return false;
      *v = it->second;
      return true;
    };
    double d = 0, f[8];
    if (!get("dtype", &d) || (d != 0 && d != 1)) return fail(ADT_EINVAL, "tune file: bad dtype");
    for (int q = 0; q < 8; ++q)
      if (!get(kTuneFields[q], &f[q]))
        return fail(ADT_EINVAL, std::string("tune file: missing ") + kTuneFields[q]);
    adt_tune t{(int32_t)f[0], (int32_t)f[1], (int32_t)f[2], (int32_t)f[3],
               (float)f[4], (float)f[5], (int32_t)f[6], (int32_t)f[7]};
    if (!tune_valid(t)) return fail(ADT_EINVAL, "tune file: setting out of range");
    const int di = (int)d;
    if (section == "global") {
      glob[di] = t;
      have_glob[di] = true;
    } else {
      double lb, bd, lg;
      if (!get("l_blocks", &lb) || !get("bands", &bd) || !get("log2n", &lg))
        return fail(ADT_EINVAL, "tune file: class without l_blocks / bands / log2n");
      cls[di][ClassKey{(int)lb, (int)bd, (int)lg}] = t;
    }
  }
  return ADT_OK;
}

// ------------------------------------------------------------- adjoint ---
// Two adjoint traversals (adt_tune.adj_algo): over 8x8x8 tiles (0; events
// of at most kTileLargeMax tiles) or node by node over the event (1).

int64_t tiles_of(int64_t m, int64_t n, int64_t l) {
  const int64_t e = adt::kTileEdge;
  return ((m + e - 1) / e) * ((n + e - 1) / e) * ((l + e - 1) / e);
}

// The traversal an event set takes: the requested one, or the node-level
// one when some event has more tiles than the tile kernels hold.
int adj_algo_for(int requested, int64_t max_tiles) {
  return (requested == 0 && max_tiles <= adt::kTileLargeMax) ? 0 : 1;
}

// Adjoint workspace of one event (bytes): the EventDesc lam, code and queue
// buffers of the traversal.
template <typename T>
void adj_ws_bytes(int algo, int64_t N, int64_t tiles, size_t* lam, size_t* code, size_t* queue) {
  if (algo == 0) {
    *lam = sizeof(T) * size_t(N);  // multipliers
    *code = size_t(N);             // upwind codes
    *queue = 16 * size_t(tiles);   // per tile: external-edge count, rank, inverse, key
  } else {
    *lam = adt::kAdjRecBytes<T> * size_t(N);  // per-node records
    *code = 2 * size_t(N);                    // 16-bit node states
    *queue = 4 * size_t(N);                   // topological queue
  }
}

template <typename T>
void launch_adj(int algo, int64_t max_tiles, int nev, const adt::EventDesc* d, int adj_wg,
                hipStream_t s) {
  if (algo != 0) {
    adt::adj3_kahn_kernel<T><<<nev, adj_wg, 0, s>>>(d);
    return;
  }
  if (max_tiles <= adt::kTileSmallMax) {
    adt::adj3_tile_kernel<T, adt::kTileWavesSmall, adt::kTileSmallMax>
        <<<nev, adt::kTileWavesSmall * 64, 0, s>>>(d);
  } else if (std::is_same<T, double>::value) {
    adt::adj3_tile_kernel<T, adt::kTileWavesLarge64, adt::kTileLargeMax>
        <<<nev, adt::kTileWavesLarge64 * 64, 0, s>>>(d);
  } else {
    adt::adj3_tile_kernel<T, adt::kTileWavesLarge32, adt::kTileLargeMax>
        <<<nev, adt::kTileWavesLarge32 * 64, 0, s>>>(d);
  }
}

int check_dims3(int64_t m, int64_t n, int64_t l) {
  if (m < 2 || n < 2 || l < 2)
    return fail(ADT_EINVAL, "eikonal3d: every dimension must be >= 2 (m=" + std::to_string(m) +
                                ", n=" + std::to_string(n) + ", l=" + std::to_string(l) + ")");
  if (m > (1 << 30) || n > (1 << 30) || l > (1 << 30) || m * n * l > (int64_t(1) << 40))
    return fail(ADT_EINVAL, "eikonal3d: dimensions too large");
  return ADT_OK;
}

int check_dims2(int64_t m, int64_t n, int64_t srcx, int64_t srcy) {
  if (m < 1 || n < 1) return fail(ADT_EINVAL, "eikonal2d: m and n must be >= 1");
  if (m >= (1 << 30) || n >= (1 << 30)) return fail(ADT_EINVAL, "eikonal2d: dimensions too large");
  if (srcx < 1 || srcx > m + 1 || srcy < 1 || srcy > n + 1)
    return fail(ADT_EINVAL, "eikonal2d: source (" + std::to_string(srcx) + "," +
                                std::to_string(srcy) + ") outside the 1-based grid");
  return ADT_OK;
}

void print_verbose(const adt_info& inf) {
  for (int i = 0; i < inf.iterations && i < ADT_MAX_LOGGED_ITERS; ++i)
    printf("Iteration %d, Error = %0.6e\n", i, inf.err_log[i]);
}

// Copies host input arrays in, or aliases device pointers.  std::list keeps
// the buffers in place (they free themselves on destruction).
struct Staged {
  std::list<DevBuf> bufs;
  int in(const void* host, size_t bytes, bool host_ptrs, const void** dev, hipStream_t s) {
    if (!host_ptrs) {
      *dev = host;
      return ADT_OK;
    }
    bufs.emplace_back();
    int rc = bufs.back().alloc(bytes, s);
    if (rc) return rc;
    if (host) ADT_HIP(hipMemcpyAsync(bufs.back().p, host, bytes, hipMemcpyHostToDevice, s));
    *dev = bufs.back().p;
    return ADT_OK;
  }
  int out(void* host, size_t bytes, bool host_ptrs, void** dev, hipStream_t s) {
    if (!host_ptrs || !host) {
      *dev = host;
      return ADT_OK;
    }
    bufs.emplace_back();
    int rc = bufs.back().alloc(bytes, s);
    if (rc) return rc;
    *dev = bufs.back().p;
    return ADT_OK;
  }
};

// ------------------------------------------------------------ 3D forward ---

// Forward launches; mix (FP32 only): FP64 local solves on FP32 storage.
template <typename T>
void launch_fwd_diag(bool mix, int grid, const adt::EventDesc* d, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (mix) {
      adt::fwd3_diag_kernel<float, true><<<grid, kDiagWG, 0, s>>>(d, kMaxIter3);
      return;
    }
  }
  adt::fwd3_diag_kernel<T, false><<<grid, kDiagWG, 0, s>>>(d, kMaxIter3);
}
template <typename T>
void launch_fwd_stream(bool mix, int grid, const adt::EventDesc* d, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (mix) {
      adt::fwd3_stream_kernel<float, true><<<grid, kStreamWG, 0, s>>>(d, kMaxIter3);
      return;
    }
  }
  adt::fwd3_stream_kernel<T, false><<<grid, kStreamWG, 0, s>>>(d, kMaxIter3);
}

// A single event's descriptor takes its box class's setting.
void set_event_setting(adt::EventDesc& d, const adt_tune& t) {
  d.fwd_diag = d.l <= t.diag_max_l ? 1 : 0;
  d.fwd_segs = (int8_t)(t.fwd_segs > 0 ? t.fwd_segs : kSegsMax);
  d.adj_algo = (int8_t)adj_algo_for(t.adj_algo, tiles_of(d.m, d.n, d.l));
}

template <typename T>
int fwd3_impl(const T* u0, const T* f, T h, int64_t m, int64_t n, int64_t l, T tol, int verbose,
              T* u, adt_info* info, int flags, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!u0 || !f || !u) return fail(ADT_EINVAL, "eikonal3d: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const bool hp = flags & ADT_HOST_POINTERS;
  const int64_t N = m * n * l;
  const size_t bytes = sizeof(T) * N;
  if (bytes >= kMaxArrayBytes) return fail(ADT_EINVAL, "eikonal3d: arrays of 4 GiB or more are not supported");
  Staged st;
  const void *du0, *df;
  void* du;
  if ((rc = st.in(u0, bytes, hp, &du0, s)) || (rc = st.in(f, bytes, hp, &df, s)) ||
      (rc = st.out(u, bytes, hp, &du, s)))
    return rc;
  DevBuf ws;
  const size_t off_desc = ((bytes + 255) / 256) * 256;
  const size_t off_info = off_desc + 256;
  if ((rc = ws.alloc(off_info + sizeof(adt_info), s))) return rc;
  char* base = (char*)ws.p;
  adt::EventDesc d{};
  d.m = (int)m;
  d.n = (int)n;
  d.l = (int)l;
  d.h = (double)h;
  d.tol = (double)tol;
  d.u0 = du0;
  d.f = df;
  d.u = du;
  d.uold = base;
  d.info = (adt_info*)(base + off_info);
  set_event_setting(d, current_snap<T>().for_box(m, n, l));
  ADT_HIP(hipMemcpyAsync(base + off_desc, &d, sizeof(d), hipMemcpyHostToDevice, s));
  adt::info_clear_kernel<<<1, 64, 0, s>>>(d.info, 1);
  const bool mix = flags & ADT_F32_COMPUTE_F64;
  if (d.fwd_diag)
    launch_fwd_diag<T>(mix, 1, (const adt::EventDesc*)(base + off_desc), s);
  else
    launch_fwd_stream<T>(mix, 1, (const adt::EventDesc*)(base + off_desc), s);
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  ADT_HIP(hipMemcpyAsync(&hinf, d.info, sizeof(adt_info), hipMemcpyDeviceToHost, s));
  if (hp) ADT_HIP(hipMemcpyAsync(u, du, bytes, hipMemcpyDeviceToHost, s));
  ADT_HIP(hipStreamSynchronize(s));
  if (verbose) print_verbose(hinf);
  if (info) *info = hinf;
  if (hinf.fault) return fail(ADT_EHIP, "eikonal3d: wave synchronisation timed out in the kernel");
  return ADT_OK;
}

// ------------------------------------------- 3D forward, no allocation ---
//
// adt_eikonal3d_fwd_noalloc_f64: the forward of adt_eikonal3d_fwd_f64 with its
// workspace (iteration snapshot, event descriptor, diagnostics) in a fixed
// device array instead of a stream-ordered allocation, for callers that
// cannot allocate.  Every call on a device shares that device's array, so
// calls that overlap on one device race.

constexpr size_t kNoAllocSnapBytes = size_t(16) << 20;  // snapshot: 2^21 FP64 nodes
constexpr size_t kNoAllocDescOff = kNoAllocSnapBytes;
constexpr size_t kNoAllocInfoOff = kNoAllocDescOff + 256;
constexpr size_t kNoAllocBytes = kNoAllocInfoOff + ((sizeof(adt_info) + 255) / 256) * 256;
static_assert(sizeof(adt::EventDesc) <= 256, "descriptor slot");

// Held by each device the library runs on until the process ends.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch array, looked up once per
// device with that device current (the array's name is a host-side stand-in).
int scratch_ptr(void** out) {
  int dev = -1;
  ADT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(ADT_ENODEV, "noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(ADT_EHIP, std::string("noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return ADT_OK;
}

int fwd3_noalloc_impl(const double* u0, const double* f, double h, int64_t m, int64_t n, int64_t l,
                      double tol, int verbose, double* u, adt_info* info, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!u0 || !f || !u) return fail(ADT_EINVAL, "eikonal3d noalloc: NULL buffer");
  const int64_t N = m * n * l;
  if ((size_t)N * sizeof(double) > kNoAllocSnapBytes)
    return fail(ADT_EINVAL, "eikonal3d noalloc: " + std::to_string(N) + " nodes exceed the " +
                                std::to_string(kNoAllocSnapBytes / sizeof(double)) +
                                "-node limit of the fixed workspace");
  void* base = nullptr;
  if ((rc = scratch_ptr(&base))) return rc;
  hipStream_t s = (hipStream_t)stream;
  char* b = (char*)base;
  adt::EventDesc d{};
  d.m = (int)m;
  d.n = (int)n;
  d.l = (int)l;
  d.h = h;
  d.tol = tol;
  d.u0 = u0;
  d.f = f;
  d.u = u;
  d.uold = b;  // written before it is read (u0 is the first snapshot)
  d.info = (adt_info*)(b + kNoAllocInfoOff);
  set_event_setting(d, current_snap<double>().for_box(m, n, l));
  ADT_HIP(hipMemcpyAsync(b + kNoAllocDescOff, &d, sizeof(d), hipMemcpyHostToDevice, s));
  adt::info_clear_kernel<<<1, 64, 0, s>>>(d.info, 1);
  if (d.fwd_diag)
    launch_fwd_diag<double>(false, 1, (const adt::EventDesc*)(b + kNoAllocDescOff), s);
  else
    launch_fwd_stream<double>(false, 1, (const adt::EventDesc*)(b + kNoAllocDescOff), s);
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  ADT_HIP(hipMemcpyAsync(&hinf, d.info, sizeof(adt_info), hipMemcpyDeviceToHost, s));
  ADT_HIP(hipStreamSynchronize(s));
  if (verbose) print_verbose(hinf);
  if (info) *info = hinf;
  if (hinf.fault) return fail(ADT_EHIP, "eikonal3d: wave synchronisation timed out in the kernel");
  return ADT_OK;
}

// ------------------------------------------------------------ 3D adjoint ---

template <typename T>
int bwd3_impl(const T* g, const T* u, const T* u0, const T* f, T h, int64_t m, int64_t n,
              int64_t l, T* grad_u0, T* grad_f, adt_info* info, int flags, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!g || !u || !u0 || !f || !grad_f) return fail(ADT_EINVAL, "eikonal3d grad: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const bool hp = flags & ADT_HOST_POINTERS;
  const int64_t N = m * n * l;
  const size_t bytes = sizeof(T) * N;
  Staged st;
  const void *dg, *du, *du0, *df;
  void *dgu0, *dgf;
  if ((rc = st.in(g, bytes, hp, &dg, s)) || (rc = st.in(u, bytes, hp, &du, s)) ||
      (rc = st.in(u0, bytes, hp, &du0, s)) || (rc = st.in(f, bytes, hp, &df, s)) ||
      (rc = st.out(grad_u0, bytes, hp, &dgu0, s)) || (rc = st.out(grad_f, bytes, hp, &dgf, s)))
    return rc;
  DevBuf ws;
  const int64_t tiles = tiles_of(m, n, l);
  const TuneSnap snap = current_snap<T>();
  const adt_tune tn = snap.for_box(m, n, l);
  const int algo = adj_algo_for(tn.adj_algo, tiles);
  size_t a, c, q;
  adj_ws_bytes<T>(algo, N, tiles, &a, &c, &q);
  a = ((a + 255) / 256) * 256;
  c = ((c + 255) / 256) * 256;
  q = ((q + 255) / 256) * 256;
  if ((rc = ws.alloc(a + c + q + 256 + sizeof(adt_info), s))) return rc;
  char* base = (char*)ws.p;
  adt::EventDesc d{};
  d.m = (int)m;
  d.n = (int)n;
  d.l = (int)l;
  d.h = (double)h;
  d.u0 = du0;
  d.f = df;
  d.grad_u = dg;
  d.u = (void*)du;
  d.lam = base;
  d.code = (uint8_t*)(base + a);
  d.indeg = nullptr;
  d.queue = (int32_t*)(base + a + c);
  d.grad_u0 = dgu0;
  d.grad_f = dgf;
  set_event_setting(d, tn);
  const size_t od = a + c + q;
  d.info = (adt_info*)(base + od + 256);
  ADT_HIP(hipMemcpyAsync(base + od, &d, sizeof(d), hipMemcpyHostToDevice, s));
  adt::info_clear_kernel<<<1, 64, 0, s>>>(d.info, 1);
  launch_adj<T>(algo, tiles, 1, (const adt::EventDesc*)(base + od), tn.adj_wg, s);
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  ADT_HIP(hipMemcpyAsync(&hinf, d.info, sizeof(adt_info), hipMemcpyDeviceToHost, s));
  if (hp) {
    if (grad_u0) ADT_HIP(hipMemcpyAsync(grad_u0, dgu0, bytes, hipMemcpyDeviceToHost, s));
    ADT_HIP(hipMemcpyAsync(grad_f, dgf, bytes, hipMemcpyDeviceToHost, s));
  }
  ADT_HIP(hipStreamSynchronize(s));
  if (info) {
    hinf.iterations = 0;
    *info = hinf;
  }
  if (!hinf.adj_converged) return fail(ADT_EINVAL, "eikonal3d grad: adjoint did not reach its fixed point");
  return ADT_OK;
}

// ------------------------------------------------------------------- 2D ---

template <typename T>
int fwd2_impl(const T* f, int64_t m, int64_t n, int64_t srcx, int64_t srcy, T h, T* u,
              adt_info* info, int flags, void* stream) {
  int rc = check_dims2(m, n, srcx, srcy);
  if (rc) return rc;
  if (!f || !u) return fail(ADT_EINVAL, "eikonal2d: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const bool hp = flags & ADT_HOST_POINTERS;
  const int64_t N = (m + 1) * (n + 1);
  const size_t bytes = sizeof(T) * N;
  Staged st;
  const void* df;
  void* du;
  if ((rc = st.in(f, bytes, hp, &df, s)) || (rc = st.out(u, bytes, hp, &du, s))) return rc;
  DevBuf ws;
  const size_t a = ((bytes + 255) / 256) * 256;
  if ((rc = ws.alloc(a + 256 + sizeof(adt_info), s))) return rc;
  char* base = (char*)ws.p;
  adt::Event2Desc d{};
  d.m = (int)m;
  d.n = (int)n;
  d.ix = (int)(srcx - 1);
  d.jx = (int)(srcy - 1);
  d.h = (double)h;
  d.f = df;
  d.u = du;
  d.uold = base;
  d.info = (adt_info*)(base + a + 256);
  ADT_HIP(hipMemcpyAsync(base + a, &d, sizeof(d), hipMemcpyHostToDevice, s));
  adt::info_clear_kernel<<<1, 64, 0, s>>>(d.info, 1);
  adt::fwd2_event_kernel<T><<<1, kWG, 0, s>>>((const adt::Event2Desc*)(base + a), kMaxIter2);
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  ADT_HIP(hipMemcpyAsync(&hinf, d.info, sizeof(adt_info), hipMemcpyDeviceToHost, s));
  if (hp) ADT_HIP(hipMemcpyAsync(u, du, bytes, hipMemcpyDeviceToHost, s));
  ADT_HIP(hipStreamSynchronize(s));
  if (!hinf.converged) fprintf(stderr, "ERROR: Eikonal does not converge!\n");
  if (info) *info = hinf;
  return ADT_OK;
}

template <typename T>
int bwd2_impl(const T* g, const T* u, const T* f, int64_t m, int64_t n, int64_t srcx,
              int64_t srcy, T h, T* grad_f, adt_info* info, int flags, void* stream) {
  int rc = check_dims2(m, n, srcx, srcy);
  if (rc) return rc;
  if (!g || !u || !f || !grad_f) return fail(ADT_EINVAL, "eikonal2d grad: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const bool hp = flags & ADT_HOST_POINTERS;
  const int64_t N = (m + 1) * (n + 1);
  const size_t bytes = sizeof(T) * N;
  Staged st;
  const void *dg, *du, *df;
  void* dgf;
  if ((rc = st.in(g, bytes, hp, &dg, s)) || (rc = st.in(u, bytes, hp, &du, s)) ||
      (rc = st.in(f, bytes, hp, &df, s)) || (rc = st.out(grad_f, bytes, hp, &dgf, s)))
    return rc;
  DevBuf ws;
  const size_t a = ((bytes + 255) / 256) * 256;
  const size_t c = ((size_t(N) + 255) / 256) * 256;
  const size_t q = ((size_t(N) * 4 + 255) / 256) * 256;
  if ((rc = ws.alloc(a + 2 * c + q + 256 + sizeof(adt_info), s))) return rc;
  char* base = (char*)ws.p;
  adt::Event2Desc d{};
  d.m = (int)m;
  d.n = (int)n;
  d.ix = (int)(srcx - 1);
  d.jx = (int)(srcy - 1);
  d.h = (double)h;
  d.f = df;
  d.grad_u = dg;
  d.u = (void*)du;
  d.lam = base;
  d.code = (uint8_t*)(base + a);
  d.indeg = (uint8_t*)(base + a + c);
  d.queue = (int32_t*)(base + a + 2 * c);
  d.grad_f = dgf;
  const size_t od = a + 2 * c + q;
  d.info = (adt_info*)(base + od + 256);
  ADT_HIP(hipMemcpyAsync(base + od, &d, sizeof(d), hipMemcpyHostToDevice, s));
  adt::info_clear_kernel<<<1, 64, 0, s>>>(d.info, 1);
  adt::adj2_kahn_kernel<T><<<1, kWG, 0, s>>>((const adt::Event2Desc*)(base + od));
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  ADT_HIP(hipMemcpyAsync(&hinf, d.info, sizeof(adt_info), hipMemcpyDeviceToHost, s));
  if (hp) ADT_HIP(hipMemcpyAsync(grad_f, dgf, bytes, hipMemcpyDeviceToHost, s));
  ADT_HIP(hipStreamSynchronize(s));
  if (info) {
    hinf.iterations = 0;
    *info = hinf;
  }
  if (!hinf.adj_converged) return fail(ADT_EINVAL, "eikonal2d grad: adjoint did not reach its fixed point");
  return ADT_OK;
}

// -------------------------------------------------------------- strided ---
//
// The _strided entry points: every array of a call is gathered on the
// device into a row-major scratch twin, the row-major path runs on the
// twins, and the outputs are scattered back.  Host arrays (ADT_HOST_POINTERS)
// travel as their whole span, so positions between elements keep their
// values.

// A strided layout as three row-major axes (2D: a unit axis, then j, i).
struct Layout {
  int64_t dim[3], st[3];
  int64_t N, span;  // elements, and the span of addresses they cover
};

int make_layout(const int64_t (&dim)[3], const int64_t (&st)[3], size_t elem, Layout* L) {
  int idx[3], na = 0;
  int64_t span = 1;
  for (int a = 0; a < 3; ++a) {
    L->dim[a] = dim[a];
    L->st[a] = dim[a] > 1 ? st[a] : 0;  // an axis of one element has no stride
    if (dim[a] <= 1) continue;
    if (st[a] <= 0) return fail(ADT_EINVAL, "strides must be positive");
    int64_t ext;
    if (__builtin_mul_overflow(dim[a] - 1, st[a], &ext) || __builtin_add_overflow(span, ext, &span))
      return fail(ADT_EINVAL, "strides: span overflows");
    idx[na++] = a;
  }
  int64_t bytes;
  if (__builtin_mul_overflow(span, (int64_t)elem, &bytes) || bytes > (int64_t(1) << 46))
    return fail(ADT_EINVAL, "strides: span too large");
  // no two elements share an address: with the axes in increasing stride,
  // each stride must reach past the whole extent of the axes below it
  // (st * dim <= span + st stays far from overflow after the check above)
  std::sort(idx, idx + na, [&](int x, int y) { return L->st[x] < L->st[y]; });
  int64_t reach = 1;
  for (int t = 0; t < na; ++t) {
    const int a = idx[t];
    if (L->st[a] < reach) return fail(ADT_EINVAL, "strides map two elements to one address");
    reach = L->st[a] * L->dim[a];
  }
  L->N = dim[0] * dim[1] * dim[2];
  L->span = span;
  return ADT_OK;
}

template <typename T>
struct StridedIO {
  const Layout& L;
  bool hp;
  hipStream_t s;
  std::list<DevBuf> bufs;
  StridedIO(const Layout& lay, bool host_ptrs, hipStream_t st) : L(lay), hp(host_ptrs), s(st) {}

  int scratch(size_t n, T** out) {
    bufs.emplace_back();
    int rc = bufs.back().alloc(n * sizeof(T), s);
    *out = (T*)bufs.back().p;
    return rc;
  }
  void copy(const T* src, T* dst, bool to_rowmajor) {
    const int64_t b = (L.N + 255) / 256;
    const int grid = (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
    adt::strided_copy_kernel<T><<<grid, 256, 0, s>>>(src, dst, L.N, L.dim[1], L.dim[2], L.st[0],
                                                       L.st[1], L.st[2], to_rowmajor ? 1 : 0);
  }
  // row-major device twin of a strided input
  int in(const T* a, const T** out) {
    int rc;
    const T* src = a;
    if (hp) {
      T* span;
      if ((rc = scratch(L.span, &span))) return rc;
      ADT_HIP(hipMemcpyAsync(span, a, L.span * sizeof(T), hipMemcpyHostToDevice, s));
      src = span;
    }
    T* twin;
    if ((rc = scratch(L.N, &twin))) return rc;
    copy(src, twin, true);
    ADT_HIP(hipGetLastError());
    *out = twin;
    return ADT_OK;
  }
  // row-major device buffer for an output (NULL stays NULL)
  int out_twin(const T* a, T** out) {
    if (!a) {
      *out = nullptr;
      return ADT_OK;
    }
    return scratch(L.N, out);
  }
  // scatters a row-major result into the caller's strided output
  int out(const T* twin, T* a) {
    if (!a) return ADT_OK;
    int rc;
    T* dst = a;
    if (hp) {
      if ((rc = scratch(L.span, &dst))) return rc;
      ADT_HIP(hipMemcpyAsync(dst, a, L.span * sizeof(T), hipMemcpyHostToDevice, s));
    }
    copy(twin, dst, false);
    ADT_HIP(hipGetLastError());
    if (hp) ADT_HIP(hipMemcpyAsync(a, dst, L.span * sizeof(T), hipMemcpyDeviceToHost, s));
    return ADT_OK;
  }
  int finish() {
    ADT_HIP(hipStreamSynchronize(s));
    return ADT_OK;
  }
};

int layout3(int64_t m, int64_t n, int64_t l, const int64_t* strides, size_t elem, Layout* L) {
  if (!strides) return fail(ADT_EINVAL, "strides: NULL");
  return make_layout({m, n, l}, {strides[0], strides[1], strides[2]}, elem, L);
}
// 2D: flat j*(m+1) + i is row-major over (j, i)
int layout2(int64_t m, int64_t n, const int64_t* strides, size_t elem, Layout* L) {
  if (!strides) return fail(ADT_EINVAL, "strides: NULL");
  return make_layout({1, n + 1, m + 1}, {0, strides[1], strides[0]}, elem, L);
}

template <typename T>
int fwd3_strided_impl(const T* u0, const T* f, T h, int64_t m, int64_t n, int64_t l, T tol,
                      int verbose, const int64_t* strides, T* u, adt_info* info, int flags,
                      void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!u0 || !f || !u) return fail(ADT_EINVAL, "eikonal3d: NULL buffer");
  Layout L;
  if ((rc = layout3(m, n, l, strides, sizeof(T), &L))) return rc;
  StridedIO<T> io(L, flags & ADT_HOST_POINTERS, (hipStream_t)stream);
  const T *r_u0, *r_f;
  T* r_u;
  if ((rc = io.in(u0, &r_u0)) || (rc = io.in(f, &r_f)) || (rc = io.out_twin(u, &r_u))) return rc;
  rc = fwd3_impl<T>(r_u0, r_f, h, m, n, l, tol, verbose, r_u, info, flags & ~ADT_HOST_POINTERS,
                    stream);
  if (rc) return rc;
  if ((rc = io.out(r_u, u))) return rc;
  return io.finish();
}

template <typename T>
int bwd3_strided_impl(const T* g, const T* u, const T* u0, const T* f, T h, int64_t m,
                      int64_t n, int64_t l, const int64_t* strides, T* grad_u0, T* grad_f,
                      adt_info* info, int flags, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!g || !u || !u0 || !f || !grad_f) return fail(ADT_EINVAL, "eikonal3d grad: NULL buffer");
  Layout L;
  if ((rc = layout3(m, n, l, strides, sizeof(T), &L))) return rc;
  StridedIO<T> io(L, flags & ADT_HOST_POINTERS, (hipStream_t)stream);
  const T *r_g, *r_u, *r_u0, *r_f;
  T *r_gu0, *r_gf;
  if ((rc = io.in(g, &r_g)) || (rc = io.in(u, &r_u)) || (rc = io.in(u0, &r_u0)) ||
      (rc = io.in(f, &r_f)) || (rc = io.out_twin(grad_u0, &r_gu0)) ||
      (rc = io.out_twin(grad_f, &r_gf)))
    return rc;
  rc = bwd3_impl<T>(r_g, r_u, r_u0, r_f, h, m, n, l, r_gu0, r_gf, info,
                    flags & ~ADT_HOST_POINTERS, stream);
  if (rc) return rc;
  if ((rc = io.out(r_gu0, grad_u0)) || (rc = io.out(r_gf, grad_f))) return rc;
  return io.finish();
}

template <typename T>
int fwd2_strided_impl(const T* f, int64_t m, int64_t n, int64_t srcx, int64_t srcy, T h,
                      const int64_t* strides, T* u, adt_info* info, int flags, void* stream) {
  int rc = check_dims2(m, n, srcx, srcy);
  if (rc) return rc;
  if (!f || !u) return fail(ADT_EINVAL, "eikonal2d: NULL buffer");
  Layout L;
  if ((rc = layout2(m, n, strides, sizeof(T), &L))) return rc;
  StridedIO<T> io(L, flags & ADT_HOST_POINTERS, (hipStream_t)stream);
  const T* r_f;
  T* r_u;
  if ((rc = io.in(f, &r_f)) || (rc = io.out_twin(u, &r_u))) return rc;
  rc = fwd2_impl<T>(r_f, m, n, srcx, srcy, h, r_u, info, flags & ~ADT_HOST_POINTERS, stream);
  if (rc) return rc;
  if ((rc = io.out(r_u, u))) return rc;
  return io.finish();
}

template <typename T>
int bwd2_strided_impl(const T* g, const T* u, const T* f, int64_t m, int64_t n, int64_t srcx,
                      int64_t srcy, T h, const int64_t* strides, T* grad_f, adt_info* info,
                      int flags, void* stream) {
  int rc = check_dims2(m, n, srcx, srcy);
  if (rc) return rc;
  if (!g || !u || !f || !grad_f) return fail(ADT_EINVAL, "eikonal2d grad: NULL buffer");
  Layout L;
  if ((rc = layout2(m, n, strides, sizeof(T), &L))) return rc;
  StridedIO<T> io(L, flags & ADT_HOST_POINTERS, (hipStream_t)stream);
  const T *r_g, *r_u, *r_f;
  T* r_gf;
  if ((rc = io.in(g, &r_g)) || (rc = io.in(u, &r_u)) || (rc = io.in(f, &r_f)) ||
      (rc = io.out_twin(grad_f, &r_gf)))
    return rc;
  rc = bwd2_impl<T>(r_g, r_u, r_f, m, n, srcx, srcy, h, r_gf, info, flags & ~ADT_HOST_POINTERS,
                    stream);
  if (rc) return rc;
  if ((rc = io.out(r_gf, grad_f))) return rc;
  return io.finish();
}

// -------------------------------------------------------------- batched ---

// Builds the device descriptor table and workspace for `nev` events.
// Receiver misfit attached to a batched forward + adjoint (host view).
struct MisfitArgs {
  const adt_misfit3d* items;
  const double *rx, *ry, *rz;
  int64_t nrec;
};

// Device view of the same, for the separate-kernel path.
struct MisfitRun {
  const adt_misfit3d* table;
  const adt::MisfitNodes* nodes;
  const double *rx, *ry, *rz;
  int nrec;
};

// The receivers of a misfit grouped by interpolation node, for each box
// among its items (host side of the deterministic misfit gradient, see
// adt::MisfitNodes), and each item's seed scratch.  build() reserves
// workspace through take(bytes) -> offset; upload() copies the tables
// (the host vectors must outlive the copies: callers synchronise).
template <typename T>
struct RecvNodes {
  struct Shape {
    int64_t n = 0, l = 0;
    std::vector<int64_t> node, off, ent;
    size_t o_node = 0, o_off = 0, o_ent = 0;
  };
  std::vector<Shape> shapes;
  std::vector<int> shape_of;             // per item
  std::vector<adt::MisfitNodes> descs;  // per item (upload)
  size_t o_tab = 0, o_gj = 0, gj_bytes = 0;

  // Corner nodes exactly as the device's trilinear_setup takes them: floor
  // and ceil of each 1-based coordinate, one corner where they coincide.
  template <class Take>
  void build(const adt_misfit3d* items, int nitem, const double* rx, const double* ry,
             const double* rz, int64_t nrec, Take take) {
    shape_of.assign(nitem, 0);
    for (int e = 0; e < nitem; ++e) {
      size_t k = 0;
      while (k < shapes.size() && !(shapes[k].n == items[e].n && shapes[k].l == items[e].l)) ++k;
      if (k == shapes.size()) {
        shapes.emplace_back();
        shapes.back().n = items[e].n;
        shapes.back().l = items[e].l;
      }
      shape_of[e] = (int)k;
    }
    for (auto& sh : shapes) {
      std::vector<std::pair<int64_t, int64_t>> pr;  // (node, entry), receivers in order
      pr.reserve((size_t)nrec * 8);
      for (int64_t j = 0; j < nrec; ++j) {
        const int xs[2] = {(int)std::floor(rx[j]), (int)std::ceil(rx[j])};
        const int ys[2] = {(int)std::floor(ry[j]), (int)std::ceil(ry[j])};
        const int zs[2] = {(int)std::floor(rz[j]), (int)std::ceil(rz[j])};
        const int na = xs[0] != xs[1] ? 2 : 1, nb = ys[0] != ys[1] ? 2 : 1,
                  nc = zs[0] != zs[1] ? 2 : 1;
        for (int a = 0; a < na; ++a)
          for (int b = 0; b < nb; ++b)
            for (int c = 0; c < nc; ++c)
              pr.emplace_back((int64_t)(xs[a] - 1) * sh.n * sh.l + (int64_t)(ys[b] - 1) * sh.l +
                                  (zs[c] - 1),
                              j * 8 + (a << 2) + (b << 1) + c);
      }
      std::stable_sort(pr.begin(), pr.end(),
                       [](const std::pair<int64_t, int64_t>& x, const std::pair<int64_t, int64_t>& y) {
                         return x.first < y.first;
                       });
      sh.off.push_back(0);
      for (size_t i = 0; i < pr.size(); ++i) {
        if (i == 0 || pr[i].first != pr[i - 1].first) {
          if (i) sh.off.push_back((int64_t)i);
          sh.node.push_back(pr[i].first);
        }
        sh.ent.push_back(pr[i].second);
      }
      sh.off.push_back((int64_t)pr.size());
      if (sh.node.empty()) sh.off.assign(1, 0);
      sh.o_node = take(sizeof(int64_t) * (sh.node.size() + 1));
      sh.o_off = take(sizeof(int64_t) * sh.off.size());
      sh.o_ent = take(sizeof(int64_t) * (sh.ent.size() + 1));
    }
    o_tab = take(sizeof(adt::MisfitNodes) * (nitem ? nitem : 1));
    gj_bytes = ((sizeof(T) * (nrec ? nrec : 1) + 255) / 256) * 256;
    o_gj = take(gj_bytes * (nitem ? nitem : 1));
  }
  int upload(char* base, int nitem, hipStream_t s) {
    for (const auto& sh : shapes) {
      if (!sh.node.empty()) {
        ADT_HIP(hipMemcpyAsync(base + sh.o_node, sh.node.data(), sizeof(int64_t) * sh.node.size(),
                               hipMemcpyHostToDevice, s));
        ADT_HIP(hipMemcpyAsync(base + sh.o_ent, sh.ent.data(), sizeof(int64_t) * sh.ent.size(),
                               hipMemcpyHostToDevice, s));
      }
      ADT_HIP(hipMemcpyAsync(base + sh.o_off, sh.off.data(), sizeof(int64_t) * sh.off.size(),
                             hipMemcpyHostToDevice, s));
    }
    descs.resize(nitem);
    for (int e = 0; e < nitem; ++e) {
      const Shape& sh = shapes[shape_of[e]];
      descs[e] = adt::MisfitNodes{(const int64_t*)(base + sh.o_node), (const int64_t*)(base + sh.o_off),
                                  (const int64_t*)(base + sh.o_ent), (int64_t)sh.node.size(),
                                  base + o_gj + gj_bytes * e};
    }
    if (nitem)
      ADT_HIP(hipMemcpyAsync(base + o_tab, descs.data(), sizeof(adt::MisfitNodes) * nitem,
                             hipMemcpyHostToDevice, s));
    return ADT_OK;
  }
  const adt::MisfitNodes* table(char* base) const { return (const adt::MisfitNodes*)(base + o_tab); }
};

// Checks one misfit item against its event and the receivers against the
// event's box (1-based coordinates within [1, extent]).
int check_misfit_item(const adt_misfit3d& it, int i, const double* rx, const double* ry,
                      const double* rz, int64_t nrec) {
  int rc = check_dims3(it.m, it.n, it.l);
  if (rc) return rc;
  if (!it.u || !it.loss || (nrec > 0 && (!it.tobs || !it.weight)))
    return fail(ADT_EINVAL, "misfit: NULL buffer for solve " + std::to_string(i));
  for (int64_t j = 0; j < nrec; ++j) {
    const double c[3] = {rx[j], ry[j], rz[j]};
    const double ext[3] = {(double)it.m, (double)it.n, (double)it.l};
    for (int d = 0; d < 3; ++d)
      if (!(c[d] >= 1.0 && c[d] <= ext[d]))
        return fail(ADT_EINVAL, "misfit: receiver " + std::to_string(j) +
                                    " outside the 1-based grid of solve " + std::to_string(i));
  }
  return ADT_OK;
}

// Which kernels a set of events needs (from each event's setting).
struct KernelMix {
  bool diag = false, stream = false;  // forward kernels
  bool tile = false, node = false;    // adjoint traversals
  int64_t tile_max = 0;  // largest tile count among the tile-traversal events
  int64_t max_n = 0;     // largest node count
};

template <typename T>
struct BatchPlan {
  std::vector<adt::EventDesc> descs;
  size_t ws_bytes = 0;
  size_t off_desc = 0, off_info = 0;
  KernelMix mix;
  // attached receiver misfit (attach_misfit): table and coordinates in the
  // workspace
  const MisfitArgs* mis = nullptr;
  size_t off_mis = 0, off_coords = 0;
  RecvNodes<T> rnodes;

  // Each event's receiver misfit runs between its forward and its adjoint
  // and writes its grad_u: the item must name the event's u and grad_u.
  int attach_misfit(const adt_event3d* ev, int nev, const MisfitArgs* m) {
    if (!m || (nev > 0 && !m->items) || m->nrec < 0 || m->nrec > (int64_t(1) << 30) ||
        (m->nrec > 0 && (!m->rx || !m->ry || !m->rz)))
      return fail(ADT_EINVAL, "misfit: bad arguments");
    for (int e = 0; e < nev; ++e) {
      const adt_misfit3d& it = m->items[e];
      int rc = check_misfit_item(it, e, m->rx, m->ry, m->rz, m->nrec);
      if (rc) return rc;
      if (it.m != ev[e].m || it.n != ev[e].n || it.l != ev[e].l || it.u != ev[e].u ||
          it.grad_u != ev[e].grad_u)
        return fail(ADT_EINVAL, "misfit: item " + std::to_string(e) +
                                    " must have its event's box, u and grad_u");
    }
    // each event's misfit zeroes and accumulates its grad_u and reads its u
    // while other workgroups run: two events must not share either buffer
    for (int which = 0; which < 2; ++which) {
      std::vector<const void*> ptrs(nev);
      for (int e = 0; e < nev; ++e) ptrs[e] = which ? (const void*)ev[e].grad_u : ev[e].u;
      std::sort(ptrs.begin(), ptrs.end());
      if (std::adjacent_find(ptrs.begin(), ptrs.end()) != ptrs.end())
        return fail(ADT_EINVAL, std::string("misfit: two events share one ") +
                                    (which ? "grad_u" : "u") + " buffer");
    }
    mis = m;
    auto take = [&](size_t b) {
      size_t o = ws_bytes;
      ws_bytes += ((b + 255) / 256) * 256;
      return o;
    };
    off_mis = take(sizeof(adt_misfit3d) * (nev ? nev : 1));
    off_coords = take(sizeof(double) * 3 * (m->nrec ? m->nrec : 1));
    rnodes.build(m->items, nev, m->rx, m->ry, m->rz, m->nrec, take);
    for (int e = 0; e < nev; ++e) {
      // relative until bind(); descs are still in input order here
      descs[e].mis = (const adt_misfit3d*)(off_mis + sizeof(adt_misfit3d) * e);
      descs[e].mnodes = (const adt::MisfitNodes*)(rnodes.o_tab + sizeof(adt::MisfitNodes) * e);
      descs[e].nrec = m->nrec;
    }
    return ADT_OK;
  }
  // Device copies of the misfit table and coordinates (after bind()).
  int upload_misfit(char* base, int nev, hipStream_t s) {
    if (!mis) return ADT_OK;
    ADT_HIP(hipMemcpyAsync(base + off_mis, mis->items, sizeof(adt_misfit3d) * nev,
                           hipMemcpyHostToDevice, s));
    double* dc = (double*)(base + off_coords);
    const size_t nb = sizeof(double) * mis->nrec;
    if (nb) {
      ADT_HIP(hipMemcpyAsync(dc, mis->rx, nb, hipMemcpyHostToDevice, s));
      ADT_HIP(hipMemcpyAsync(dc + mis->nrec, mis->ry, nb, hipMemcpyHostToDevice, s));
      ADT_HIP(hipMemcpyAsync(dc + 2 * mis->nrec, mis->rz, nb, hipMemcpyHostToDevice, s));
    }
    return rnodes.upload(base, nev, s);
  }
  MisfitRun misfit_run(char* base) const {
    if (!mis) return MisfitRun{nullptr, nullptr, nullptr, nullptr, nullptr, 0};
    const double* dc = (const double*)(base + off_coords);
    return MisfitRun{(const adt_misfit3d*)(base + off_mis), rnodes.table(base), dc,
                     dc + mis->nrec, dc + 2 * mis->nrec, (int)mis->nrec};
  }

  // Each event takes the setting of its box class (or the global one).
  int build(const adt_event3d* ev, int nev, int flags, const TuneSnap& snap) {
    if (nev < 0) return fail(ADT_EINVAL, "batched: nev < 0");
    if (nev > 0 && !ev) return fail(ADT_EINVAL, "batched: NULL events");
    descs.resize(nev);
    size_t off = 0;
    auto take = [&](size_t b) {
      size_t o = off;
      off += ((b + 255) / 256) * 256;
      return o;
    };
    std::vector<size_t> o_uold(nev), o_lam(nev), o_code(nev), o_indeg(nev), o_queue(nev);
    for (int e = 0; e < nev; ++e) {
      const adt_event3d& x = ev[e];
      int rc = check_dims3(x.m, x.n, x.l);
      if (rc) return fail(rc, "event " + std::to_string(e) + ": " + g_err);
      if (!x.f || !x.u || ((flags & ADT_FORWARD) && !x.u0))
        return fail(ADT_EINVAL, "event " + std::to_string(e) + ": NULL buffer");
      if ((flags & ADT_ADJOINT) && (!x.grad_u || !x.grad_f || !x.u0))
        return fail(ADT_EINVAL, "event " + std::to_string(e) + ": NULL adjoint buffer");
      const size_t N = size_t(x.m * x.n * x.l);
      if (N * sizeof(T) >= kMaxArrayBytes)
        return fail(ADT_EINVAL, "event " + std::to_string(e) + ": arrays of 4 GiB or more are not supported");
      const adt_tune t = snap.for_box(x.m, x.n, x.l);
      const int64_t tiles = tiles_of(x.m, x.n, x.l);
      descs[e] = adt::EventDesc{};
      descs[e].fwd_diag = x.l <= t.diag_max_l ? 1 : 0;
      descs[e].fwd_segs = (int8_t)(t.fwd_segs > 0 ? t.fwd_segs : kSegsMax);
      descs[e].adj_algo = (int8_t)adj_algo_for(t.adj_algo, tiles);
      if (flags & ADT_FORWARD) (descs[e].fwd_diag ? mix.diag : mix.stream) = true;
      if (flags & ADT_ADJOINT) {
        if (descs[e].adj_algo == 0) {
          mix.tile = true;
          mix.tile_max = std::max(mix.tile_max, tiles);
        } else {
          mix.node = true;
        }
      }
      if ((int64_t)N > mix.max_n) mix.max_n = (int64_t)N;
      o_uold[e] = (flags & ADT_FORWARD) ? take(N * sizeof(T)) : 0;
      size_t wl, wc, wq;
      adj_ws_bytes<T>(descs[e].adj_algo, (int64_t)N, tiles, &wl, &wc, &wq);
      o_lam[e] = (flags & ADT_ADJOINT) ? take(wl) : 0;
      o_code[e] = (flags & ADT_ADJOINT) ? take(wc) : 0;
      o_indeg[e] = 0;
      o_queue[e] = (flags & ADT_ADJOINT) ? take(wq) : 0;
    }
    off_desc = take(sizeof(adt::EventDesc) * (nev ? nev : 1));
    off_info = take(sizeof(adt_info) * (nev ? nev : 1));
    ws_bytes = off;
    for (int e = 0; e < nev; ++e) {
      const adt_event3d& x = ev[e];
      adt::EventDesc& d = descs[e];
      d.m = (int)x.m;
      d.n = (int)x.n;
      d.l = (int)x.l;
      d.h = x.h;
      d.tol = x.tol;
      d.u0 = x.u0;
      d.f = x.f;
      d.grad_u = x.grad_u;
      d.u = x.u;
      d.grad_u0 = x.grad_u0;
      d.grad_f = x.grad_f;
      // workspace pointers are relative offsets until bind()
      d.uold = (void*)o_uold[e];
      d.lam = (void*)o_lam[e];
      d.code = (uint8_t*)o_code[e];
      d.indeg = (uint8_t*)o_indeg[e];
      d.queue = (int32_t*)o_queue[e];
      d.info = (adt_info*)(off_info + sizeof(adt_info) * e);
    }
    return ADT_OK;
  }
  // Largest events first (by node count), so that the workgroups of the
  // long solves start first and short ones fill in behind them.  The
  // per-event info slots stay with their events.
  void order_largest_first() {
    std::stable_sort(descs.begin(), descs.end(), [](const adt::EventDesc& a, const adt::EventDesc& b) {
      return (int64_t)a.m * a.n * a.l > (int64_t)b.m * b.n * b.l;
    });
  }
  void bind(char* base) {
    for (auto& d : descs) {
      d.uold = base + (size_t)d.uold;
      d.lam = base + (size_t)d.lam;
      d.code = (uint8_t*)(base + (size_t)d.code);
      d.indeg = nullptr;
      d.queue = (int32_t*)(base + (size_t)d.queue);
      d.info = (adt_info*)(base + (size_t)d.info);
      if (mis) {
        const double* dc = (const double*)(base + off_coords);
        d.mis = (const adt_misfit3d*)(base + (size_t)d.mis);
        d.mnodes = (const adt::MisfitNodes*)(base + (size_t)d.mnodes);
        d.rx = dc;
        d.ry = dc + mis->nrec;
        d.rz = dc + 2 * mis->nrec;
      }
    }
  }
};

// The kernels of a batch: each event's forward on the kernel its setting
// names (each kernel skips the other's events), the misfit, then each
// event's adjoint traversal likewise.  tn: the launch-wide fields.
template <typename T>
int launch_batch(const adt::EventDesc* ddesc, int nev, int flags, const KernelMix& km,
                 const adt_tune& tn, adt_info* infos, hipStream_t s,
                 const MisfitRun& mr = MisfitRun{}) {
  if (nev == 0) return ADT_OK;
  adt::info_clear_kernel<<<1, 256, 0, s>>>(infos, nev);
  ADT_HIP(hipGetLastError());
  const bool mix = flags & ADT_F32_COMPUTE_F64;
  // forward and adjoint of each event in one kernel (every event on the
  // diagonal forward and the node-level adjoint)
  if ((flags & ADT_FORWARD) && (flags & ADT_ADJOINT) && km.diag && !km.stream && km.node &&
      !km.tile && km.max_n <= tn.fuse_max_n) {
    if constexpr (std::is_same<T, float>::value) {
      if (mix) {
        adt::fwd_adj3_diag_kernel<float, true><<<nev, kDiagWG, 0, s>>>(ddesc, kMaxIter3);
        ADT_HIP(hipGetLastError());
        return ADT_OK;
      }
    }
    adt::fwd_adj3_diag_kernel<T, false><<<nev, kDiagWG, 0, s>>>(ddesc, kMaxIter3);
    ADT_HIP(hipGetLastError());
    return ADT_OK;
  }
  if (flags & ADT_FORWARD) {
    if (km.diag) {
      launch_fwd_diag<T>(mix, nev, ddesc, s);
      ADT_HIP(hipGetLastError());
    }
    if (km.stream) {
      launch_fwd_stream<T>(mix, nev, ddesc, s);
      ADT_HIP(hipGetLastError());
    }
  }
  if (mr.table) {  // the events' receiver misfit: its gradient is the adjoint source
    adt::misfit_kernel<T><<<nev, 256, 0, s>>>(mr.table, mr.nodes, mr.rx, mr.ry, mr.rz, mr.nrec);
    ADT_HIP(hipGetLastError());
  }
  if (flags & ADT_ADJOINT) {
    if (km.tile) {
      launch_adj<T>(0, km.tile_max, nev, ddesc, tn.adj_wg, s);
      ADT_HIP(hipGetLastError());
    }
    if (km.node) {
      launch_adj<T>(1, 0, nev, ddesc, tn.adj_wg, s);
      ADT_HIP(hipGetLastError());
    }
  }
  return ADT_OK;
}

template <typename T>
int batched_impl(const adt_event3d* ev, int nev, int flags, adt_info* infos, void* stream,
                 const MisfitArgs* mis = nullptr) {
  if (!(flags & (ADT_FORWARD | ADT_ADJOINT))) return fail(ADT_EINVAL, "batched: no ADT_FORWARD/ADT_ADJOINT flag");
  if (flags & ADT_HOST_POINTERS) return fail(ADT_EINVAL, "batched: device pointers only");
  if (mis && !((flags & ADT_FORWARD) && (flags & ADT_ADJOINT)))
    return fail(ADT_EINVAL, "misfit batched: needs ADT_FORWARD and ADT_ADJOINT");
  hipStream_t s = (hipStream_t)stream;
  const TuneSnap snap = current_snap<T>();
  const adt_tune& tn = snap.global;
  BatchPlan<T> plan;
  int rc = plan.build(ev, nev, flags, snap);
  if (rc) return rc;
  if (mis && (rc = plan.attach_misfit(ev, nev, mis))) return rc;
  if (nev == 0) return ADT_OK;
  if (tn.order) plan.order_largest_first();
  DevBuf ws;
  if ((rc = ws.alloc(plan.ws_bytes, s))) return rc;
  plan.bind((char*)ws.p);
  char* base = (char*)ws.p;
  ADT_HIP(hipMemcpyAsync(base + plan.off_desc, plan.descs.data(), sizeof(adt::EventDesc) * nev,
                         hipMemcpyHostToDevice, s));
  if ((rc = plan.upload_misfit(base, nev, s))) return rc;
  if ((rc = launch_batch<T>((const adt::EventDesc*)(base + plan.off_desc), nev, flags, plan.mix,
                            tn, (adt_info*)(base + plan.off_info), s, plan.misfit_run(base))))
    return rc;
  // the misfit table and coordinates were copied from the caller's host
  // memory: wait before returning
  if (mis && !infos) ADT_HIP(hipStreamSynchronize(s));
  if (infos) {
    ADT_HIP(hipMemcpyAsync(infos, base + plan.off_info, sizeof(adt_info) * nev,
                           hipMemcpyDeviceToHost, s));
    ADT_HIP(hipStreamSynchronize(s));
    for (int e = 0; e < nev; ++e)
      if (infos[e].fault)
        return fail(ADT_EHIP, "batched: wave synchronisation timed out in event " + std::to_string(e));
  }
  return ADT_OK;
}

// ------------------------------------------------------- plain baseline ---

template <typename T>
int fwd3_plain_impl(const T* u0, const T* f, T h, int64_t m, int64_t n, int64_t l, T tol, T* u,
                    adt_info* info, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!u0 || !f || !u) return fail(ADT_EINVAL, "eikonal3d plain: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = m * n * l;
  DevBuf ws;
  const size_t a = ((sizeof(T) * N + 255) / 256) * 256;
  if ((rc = ws.alloc(a + 256, s))) return rc;
  T* uold = (T*)ws.p;
  unsigned long long* dkey = (unsigned long long*)((char*)ws.p + a);
  const int grid = 1024;
  adt::copy_kernel<T><<<grid, 256, 0, s>>>(u, u0, N);
  ADT_HIP(hipGetLastError());
  adt_info hinf{};
  const int S = int(m + n + l - 2);
  const int dirs[8] = {0, 1, 3, 2, 6, 4, 5, 7};
  int it;
  T err = 0;
  for (it = 0; it < kMaxIter3; ++it) {
    adt::copy_kernel<T><<<grid, 256, 0, s>>>(uold, u, N);
    for (int sw = 0; sw < 8; ++sw)
      for (int sp = 0; sp < S; ++sp) {
        int lo = sp - int(n - 1) - int(l - 1);
        lo = lo > 0 ? lo : 0;
        int hi = sp < int(m - 1) ? sp : int(m - 1);
        int cnt = hi - lo + 1;
        int64_t threads = int64_t(cnt) * n;
        adt::fwd3_plane_kernel<T><<<int((threads + 255) / 256), 256, 0, s>>>(
            u, f, h, (int)m, (int)n, (int)l, dirs[sw], sp, lo, cnt);
      }
    ADT_HIP(hipGetLastError());
    // err in the reference's memory order: the maximum after the last NaN
    // difference (dkey[1]: 1 + its index; NaN if it is the last node)
    ADT_HIP(hipMemsetAsync(dkey, 0, 2 * sizeof(unsigned long long), s));
    adt::lastnan_kernel<T><<<grid, 256, 0, s>>>(u, uold, N, dkey + 1);
    adt::maxdiff_kernel<T><<<grid, 256, 0, s>>>(u, uold, N, dkey + 1, dkey);
    unsigned long long hk[2] = {0, 0};
    ADT_HIP(hipMemcpyAsync(hk, dkey, sizeof(hk), hipMemcpyDeviceToHost, s));
    ADT_HIP(hipStreamSynchronize(s));
    if (hk[1] == (unsigned long long)N) {
      err = std::numeric_limits<T>::quiet_NaN();
    } else if (sizeof(T) == 8) {
      double e;
      memcpy(&e, &hk[0], 8);
      err = (T)e;
    } else {
      uint32_t k32 = (uint32_t)hk[0];
      float e;
      memcpy(&e, &k32, 4);
      err = (T)e;
    }
    if (it < ADT_MAX_LOGGED_ITERS) hinf.err_log[it] = (double)err;
    if (err < tol) {
      ++it;
      break;
    }
  }
  hinf.iterations = it > kMaxIter3 ? kMaxIter3 : it;
  hinf.converged = err < tol;
  hinf.err = (double)err;
  if (info) *info = hinf;
  return ADT_OK;
}

template <typename T>
int bwd3_plain_impl(const T* g, const T* u, const T* u0, const T* f, T h, int64_t m, int64_t n,
                    int64_t l, T* grad_u0, T* grad_f, adt_info* info, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!g || !u || !u0 || !f || !grad_f) return fail(ADT_EINVAL, "eikonal3d plain grad: NULL buffer");
  hipStream_t s = (hipStream_t)stream;
  const int64_t N = m * n * l;
  DevBuf ws;
  const size_t a = ((sizeof(T) * N + 255) / 256) * 256;
  const size_t c = ((size_t(N) + 255) / 256) * 256;
  if ((rc = ws.alloc(a + c + 256, s))) return rc;
  T* lam = (T*)ws.p;
  uint8_t* code = (uint8_t*)ws.p + a;
  int* dchg = (int*)((char*)ws.p + a + c);
  adt::adj3_prep_kernel<T><<<1024, 256, 0, s>>>(u, u0, g, (int)m, (int)n, (int)l, code, lam, grad_u0);
  ADT_HIP(hipGetLastError());
  const int S = int(m + n + l - 2);
  const int dirs[8] = {0, 1, 3, 2, 6, 4, 5, 7};
  int sweeps;
  bool conv = false;
  for (sweeps = 0; sweeps < kMaxAdjSweeps;) {
    const int dir = dirs[7 - (sweeps & 7)] ^ 7;
    ADT_HIP(hipMemsetAsync(dchg, 0, sizeof(int), s));
    for (int sp = 0; sp < S; ++sp) {
      int lo = sp - int(n - 1) - int(l - 1);
      lo = lo > 0 ? lo : 0;
      int hi = sp < int(m - 1) ? sp : int(m - 1);
      int cnt = hi - lo + 1;
      int64_t threads = int64_t(cnt) * n;
      adt::adj3_plane_kernel<T><<<int((threads + 255) / 256), 256, 0, s>>>(
          u, g, code, lam, (int)m, (int)n, (int)l, dir, sp, lo, cnt, dchg);
    }
    ADT_HIP(hipGetLastError());
    int hc = 0;
    ADT_HIP(hipMemcpyAsync(&hc, dchg, sizeof(int), hipMemcpyDeviceToHost, s));
    ADT_HIP(hipStreamSynchronize(s));
    ++sweeps;
    if (!hc) {
      conv = true;
      break;
    }
  }
  adt::adj3_finish_kernel<T><<<1024, 256, 0, s>>>(lam, f, h, N, grad_f);
  ADT_HIP(hipGetLastError());
  ADT_HIP(hipStreamSynchronize(s));
  adt_info hinf{};
  hinf.adj_sweeps = sweeps;
  hinf.adj_converged = conv;
  if (info) *info = hinf;
  if (!conv) return fail(ADT_EINVAL, "eikonal3d plain grad: adjoint did not reach its fixed point");
  return ADT_OK;
}

// -------------------------------------------------------------- autotune ---

// Times launch settings on the caller's events and keeps the fastest.
// First, for each box-shape class among the events and on that class's
// events alone: the forward kernel (k-streaming, or diagonal streaming with
// 8, 4 or ceil(l / 8) c segments) and then the adjoint traversal (tiles or
// node by node), stored as the class's setting.  Then, on all events with
// those class settings: the event order, the node-level adjoint's
// workgroup size, and whether the fused forward + adjoint kernel beats the
// separate ones (which moves fuse_max_n; a loss lowers it only when the
// batch fills the device, one event per compute unit, since a small batch
// says little about a full one), stored as the global setting.  Every
// setting computes the same results (the forward bit for bit, the adjoint
// to rounding), so the caller's outputs hold a valid solve afterwards.
template <typename T>
int autotune_impl(const adt_event3d* ev, int nev, int flags, int reps, adt_tune* best,
                  hipStream_t s) {
  if (!(flags & (ADT_FORWARD | ADT_ADJOINT)))
    return fail(ADT_EINVAL, "autotune: no ADT_FORWARD/ADT_ADJOINT flag");
  if (flags & ADT_HOST_POINTERS) return fail(ADT_EINVAL, "autotune: device pointers only");
  if (nev < 1) return fail(ADT_EINVAL, "autotune: no events");
  if (!ev) return fail(ADT_EINVAL, "autotune: NULL events");
  reps = reps < 1 ? 1 : (reps > 50 ? 50 : reps);
  const TuneSnap cur = current_snap<T>();
  const int fmix = flags & ADT_F32_COMPUTE_F64;
  std::map<ClassKey, std::vector<adt_event3d>> groups;
  for (int e = 0; e < nev; ++e) {
    int rc = check_dims3(ev[e].m, ev[e].n, ev[e].l);
    if (rc) return fail(rc, "autotune: event " + std::to_string(e) + ": " + g_err);
    groups[class_of(ev[e].m, ev[e].n, ev[e].l)].push_back(ev[e]);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ADT_HIP(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(ADT_EHIP, std::string("autotune: hipEventCreate: ") + hipGetErrorString(he));
  }
  // mean time of `reps` launches of `evs` under snapshot `sn` with flags
  // `fl`, after one untimed launch
  auto time_set = [&](const std::vector<adt_event3d>& evs, const TuneSnap& sn, int fl,
                      float* ms) -> int {
    const int n = (int)evs.size();
    BatchPlan<T> plan;
    int r = plan.build(evs.data(), n, fl, sn);
    if (r) return r;
    if (sn.global.order) plan.order_largest_first();
    DevBuf ws;
    if ((r = ws.alloc(plan.ws_bytes, s))) return r;
    plan.bind((char*)ws.p);
    char* base = (char*)ws.p;
    const adt::EventDesc* dd = (const adt::EventDesc*)(base + plan.off_desc);
    adt_info* inf = (adt_info*)(base + plan.off_info);
    ADT_HIP(hipMemcpyAsync(base + plan.off_desc, plan.descs.data(), sizeof(adt::EventDesc) * n,
                           hipMemcpyHostToDevice, s));
    if ((r = launch_batch<T>(dd, n, fl, plan.mix, sn.global, inf, s))) return r;
    ADT_HIP(hipEventRecord(e0, s));
    for (int k = 0; k < reps; ++k)
      if ((r = launch_batch<T>(dd, n, fl, plan.mix, sn.global, inf, s))) return r;
    ADT_HIP(hipEventRecord(e1, s));
    ADT_HIP(hipEventSynchronize(e1));
    ADT_HIP(hipEventElapsedTime(ms, e0, e1));
    *ms /= (float)reps;
    return ADT_OK;
  };
  int rc = ADT_OK;
  std::map<ClassKey, adt_tune> classes = cur.classes;
  // --- per class: forward kernel and segments, then adjoint traversal
  for (auto it = groups.begin(); it != groups.end() && rc == ADT_OK; ++it) {
    const std::vector<adt_event3d>& evs = it->second;
    const adt_event3d& x0 = evs.front();
    TuneSnap sc;
    sc.global = cur.for_box(x0.m, x0.n, x0.l);
    sc.global.fuse_max_n = 0;  // the class's kernels are timed apart
    float fbest = 0.f, abest = 0.f;
    if (flags & ADT_FORWARD) {
      std::vector<adt_tune> cands;
      adt_tune c = sc.global;
      c.diag_max_l = 0;  // k-streaming
      cands.push_back(c);
      std::vector<int> segs = {kSegsMax, 4};
      if (it->first.l_blocks >= 1 && it->first.l_blocks < kSegsMax) segs.push_back(it->first.l_blocks);
      std::sort(segs.begin(), segs.end());
      segs.erase(std::unique(segs.begin(), segs.end()), segs.end());
      for (int k : segs) {
        c = sc.global;
        c.diag_max_l = adt::kDiagAll;
        c.fwd_segs = k;
        cands.push_back(c);
      }
      fbest = 3.4e38f;
      for (const adt_tune& cand : cands) {
        TuneSnap t1 = sc;
        t1.global = cand;
        float ms = 0.f;
        if ((rc = time_set(evs, t1, ADT_FORWARD | fmix, &ms))) break;
        if (ms < fbest) {
          fbest = ms;
          sc.global.diag_max_l = cand.diag_max_l;
          sc.global.fwd_segs = cand.fwd_segs;
        }
      }
      if (rc) break;
    }
    if (flags & ADT_ADJOINT) {
      abest = 3.4e38f;
      for (int algo = 0; algo < 2; ++algo) {
        if (algo == 0 && tiles_of(x0.m, x0.n, x0.l) > adt::kTileLargeMax) continue;
        TuneSnap t1 = sc;
        t1.global.adj_algo = algo;
        float ms = 0.f;
        if ((rc = time_set(evs, t1, ADT_ADJOINT, &ms))) break;
        if (ms < abest) {
          abest = ms;
          sc.global.adj_algo = algo;
        }
      }
      if (rc) break;
    }
    adt_tune ct = sc.global;
    ct.fuse_max_n = cur.global.fuse_max_n;
    ct.fwd_ms = fbest;
    ct.adj_ms = abest;
    classes[it->first] = ct;
  }
  // --- all events: order, node-level workgroup size, fusion
  adt_tune win = cur.global;
  float win_total = 3.4e38f;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) ncu = 0;
  std::vector<adt_event3d> all(ev, ev + nev);
  const int wgs[3] = {256, 512, 1024};
  for (int order = 0; order < 2 && rc == ADT_OK; ++order) {
    TuneSnap sn{cur.global, classes};
    sn.global.order = order;
    sn.global.fuse_max_n = 0;
    float fbest = 0.f, abest = 0.f;
    if (flags & ADT_FORWARD) {
      if ((rc = time_set(all, sn, ADT_FORWARD | fmix, &fbest))) break;
    }
    bool any_node = false, all_fusable = (flags & ADT_FORWARD) && (flags & ADT_ADJOINT);
    int64_t max_n = 0;
    for (const adt_event3d& x : all) {
      const adt_tune t = sn.for_box(x.m, x.n, x.l);
      const bool node = adj_algo_for(t.adj_algo, tiles_of(x.m, x.n, x.l)) == 1;
      any_node = any_node || node;
      all_fusable = all_fusable && node && x.l <= t.diag_max_l;
      max_n = std::max<int64_t>(max_n, x.m * x.n * x.l);
    }
    if (flags & ADT_ADJOINT) {
      abest = 3.4e38f;
      for (int wg : wgs) {
        if (!any_node && wg != wgs[0]) break;  // the tile traversal has one shape
        TuneSnap t1 = sn;
        t1.global.adj_wg = wg;
        float ms = 0.f;
        if ((rc = time_set(all, t1, ADT_ADJOINT, &ms))) break;
        if (ms < abest) {
          abest = ms;
          sn.global.adj_wg = wg;
        }
      }
      if (rc) break;
    }
    float total = fbest + abest;
    sn.global.fuse_max_n = cur.global.fuse_max_n;
    if (all_fusable) {
      TuneSnap t1 = sn;
      t1.global.fuse_max_n = (int32_t)std::min<int64_t>(max_n, INT32_MAX);
      float ms = 0.f;
      if ((rc = time_set(all, t1, flags & (ADT_FORWARD | ADT_ADJOINT | ADT_F32_COMPUTE_F64), &ms)))
        break;
      if (ms < total) {
        total = ms;
        sn.global.fuse_max_n = std::max(cur.global.fuse_max_n, t1.global.fuse_max_n);
      } else if (ncu > 0 && nev >= ncu) {
        sn.global.fuse_max_n = std::min(cur.global.fuse_max_n, (int32_t)(t1.global.fuse_max_n - 1));
      }
    }
    if (total < win_total) {
      win_total = total;
      win = sn.global;
      win.fwd_ms = fbest;
      win.adj_ms = abest;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[tune_index<T>()] = win;
    g_classes[tune_index<T>()] = classes;
  }
  if (best) *best = win;
  return ADT_OK;
}

// ----------------------------------------------------------------- graph ---

struct GraphBase {
  virtual ~GraphBase() {}
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void* ws = nullptr;
  int nev = 0;
  size_t off_info = 0;
};

template <typename T>
int graph_create_impl(const adt_event3d* ev, int nev, int flags, GraphBase** out, hipStream_t s,
                      const MisfitArgs* mis = nullptr) {
  if (!(flags & (ADT_FORWARD | ADT_ADJOINT))) return fail(ADT_EINVAL, "graph: no ADT_FORWARD/ADT_ADJOINT flag");
  if (mis && !((flags & ADT_FORWARD) && (flags & ADT_ADJOINT)))
    return fail(ADT_EINVAL, "misfit graph: needs ADT_FORWARD and ADT_ADJOINT");
  const TuneSnap snap = current_snap<T>();
  const adt_tune& tn = snap.global;
  BatchPlan<T> plan;
  int rc = plan.build(ev, nev, flags, snap);
  if (rc) return rc;
  if (mis && (rc = plan.attach_misfit(ev, nev, mis))) return rc;
  if (tn.order) plan.order_largest_first();
  GraphBase* g = new GraphBase();
  g->nev = nev;
  g->off_info = plan.off_info;
  hipError_t e = hipMalloc(&g->ws, plan.ws_bytes ? plan.ws_bytes : 1);
  if (e != hipSuccess) {
    delete g;
    return fail(ADT_ENOMEM, std::string("graph: hipMalloc: ") + hipGetErrorString(e));
  }
  plan.bind((char*)g->ws);
  char* base = (char*)g->ws;
  e = hipMemcpy(base + plan.off_desc, plan.descs.data(), sizeof(adt::EventDesc) * nev,
                hipMemcpyHostToDevice);
  if (e == hipSuccess && plan.mis) {
    rc = plan.upload_misfit(base, nev, nullptr);
    if (rc == ADT_OK) e = hipDeviceSynchronize();
    else e = hipErrorUnknown;
  }
  if (e != hipSuccess) {
    (void)hipFree(g->ws);
    delete g;
    return fail(ADT_EHIP, std::string("graph: hipMemcpy: ") + hipGetErrorString(e));
  }
  hipStream_t cs;
  e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    rc = launch_batch<T>((const adt::EventDesc*)(base + plan.off_desc), nev, flags, plan.mix, tn,
                         (adt_info*)(base + plan.off_info), cs, plan.misfit_run(base));
    hipError_t e2 = hipStreamEndCapture(cs, &g->graph);
    if (rc == ADT_OK && e2 != hipSuccess) e = e2;
  }
  if (e == hipSuccess && rc == ADT_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  if (e != hipSuccess || rc != ADT_OK) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    (void)hipFree(g->ws);
    delete g;
    return rc ? rc : fail(ADT_EHIP, std::string("graph capture: ") + hipGetErrorString(e));
  }
  (void)s;
  *out = g;
  return ADT_OK;
}

}  // namespace

struct adt_graph {
  GraphBase* g;
};

extern "C" {

const char* adt_version(void) { return "adtomo_hip 0.1.0 (gfx950)"; }
const char* adt_last_error(void) { return g_err.c_str(); }

int adt_get_tune(int32_t dtype, adt_tune* out) {
  if ((dtype != 0 && dtype != 1) || !out) return fail(ADT_EINVAL, "get_tune: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  *out = g_tune[dtype];
  return ADT_OK;
}
int adt_set_tune(int32_t dtype, const adt_tune* t) {
  if ((dtype != 0 && dtype != 1) || !t) return fail(ADT_EINVAL, "set_tune: bad arguments");
  if (!tune_valid(*t))
    return fail(ADT_EINVAL, "set_tune: diag_max_l must be in [0, 2^30], adj_wg 256/512/1024, order 0/1, fuse_max_n >= 0, adj_algo 0/1, fwd_segs 0..8");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[dtype] = *t;
  return ADT_OK;
}
int adt_autotune(const adt_event3d* events, int32_t nev, int32_t flags, int32_t dtype,
                 int32_t reps, adt_tune* best, void* stream) {
  if (dtype == 0) return autotune_impl<double>(events, nev, flags, reps, best, (hipStream_t)stream);
  if (dtype == 1) return autotune_impl<float>(events, nev, flags, reps, best, (hipStream_t)stream);
  return fail(ADT_EINVAL, "autotune: dtype must be 0 (FP64) or 1 (FP32)");
}

int adt_tune_class_set(int32_t dtype, int64_t m, int64_t n, int64_t l, const adt_tune* t) {
  if (dtype != 0 && dtype != 1) return fail(ADT_EINVAL, "tune_class_set: dtype must be 0 or 1");
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (t && !tune_valid(*t)) return fail(ADT_EINVAL, "tune_class_set: setting out of range");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (t)
    g_classes[dtype][class_of(m, n, l)] = *t;
  else
    g_classes[dtype].erase(class_of(m, n, l));
  return ADT_OK;
}
int adt_tune_class_count(int32_t dtype) {
  if (dtype != 0 && dtype != 1) return fail(ADT_EINVAL, "tune_class_count: dtype must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (int)g_classes[dtype].size();
}
int adt_tune_class_get(int32_t dtype, int32_t i, adt_tune_class* out) {
  if ((dtype != 0 && dtype != 1) || !out) return fail(ADT_EINVAL, "tune_class_get: bad arguments");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  if (i < 0 || i >= (int)g_classes[dtype].size())
    return fail(ADT_EINVAL, "tune_class_get: index out of range");
  auto it = g_classes[dtype].begin();
  std::advance(it, i);
  out->l_blocks = it->first.l_blocks;
  out->bands = it->first.bands;
  out->log2n = it->first.log2n;
  out->tune = it->second;
  return ADT_OK;
}
int adt_tune_class_clear(int32_t dtype) {
  if (dtype != 0 && dtype != 1) return fail(ADT_EINVAL, "tune_class_clear: dtype must be 0 or 1");
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_classes[dtype].clear();
  return ADT_OK;
}
int adt_tune_save(const char* path) {
  if (!path) return fail(ADT_EINVAL, "tune_save: NULL path");
  std::string txt = "{\"format\": \"adtomo_hip tune 1\",\n \"global\": [";
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    for (int d = 0; d < 2; ++d) txt += std::string(d ? ",\n  " : "\n  ") + tune_json(d, g_tune[d], nullptr);
    txt += "],\n \"classes\": [";
    bool first = true;
    for (int d = 0; d < 2; ++d)
      for (const auto& kv : g_classes[d]) {
        txt += std::string(first ? "\n  " : ",\n  ") + tune_json(d, kv.second, &kv.first);
        first = false;
      }
  }
  txt += "]}\n";
  FILE* fp = fopen(path, "w");
  if (!fp) return fail(ADT_EINVAL, std::string("tune_save: cannot open ") + path);
  const bool ok = fwrite(txt.data(), 1, txt.size(), fp) == txt.size();
  if (fclose(fp) != 0 || !ok) return fail(ADT_EINVAL, std::string("tune_save: cannot write ") + path);
  return ADT_OK;
}
int adt_tune_load(const char* path) {
  if (!path) return fail(ADT_EINVAL, "tune_load: NULL path");
  FILE* fp = fopen(path, "r");
  if (!fp) return fail(ADT_EINVAL, std::string("tune_load: cannot open ") + path);
  std::string txt;
  char buf[4096];
  size_t k;
  while ((k = fread(buf, 1, sizeof buf, fp)) > 0) txt.append(buf, k);
  fclose(fp);
  adt_tune glob[2];
  bool have[2] = {false, false};
  std::map<ClassKey, adt_tune> cls[2];
  int rc = parse_tune_json(txt, glob, have, cls);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int d = 0; d < 2; ++d) {
    if (have[d]) g_tune[d] = glob[d];
    g_classes[d] = cls[d];
  }
  return ADT_OK;
}

int adt_eikonal2d_fwd_f64(const double* f, int64_t m, int64_t n, int64_t srcx, int64_t srcy,
                          double h, double* u, adt_info* info, int32_t flags, void* stream) {
  return fwd2_impl<double>(f, m, n, srcx, srcy, h, u, info, flags, stream);
}
int adt_eikonal2d_fwd_f32(const float* f, int64_t m, int64_t n, int64_t srcx, int64_t srcy,
                          float h, float* u, adt_info* info, int32_t flags, void* stream) {
  return fwd2_impl<float>(f, m, n, srcx, srcy, h, u, info, flags, stream);
}
int adt_eikonal2d_bwd_f64(const double* grad_u, const double* u, const double* f, int64_t m,
                          int64_t n, int64_t srcx, int64_t srcy, double h, double* grad_f,
                          adt_info* info, int32_t flags, void* stream) {
  return bwd2_impl<double>(grad_u, u, f, m, n, srcx, srcy, h, grad_f, info, flags, stream);
}
int adt_eikonal2d_bwd_f32(const float* grad_u, const float* u, const float* f, int64_t m,
                          int64_t n, int64_t srcx, int64_t srcy, float h, float* grad_f,
                          adt_info* info, int32_t flags, void* stream) {
  return bwd2_impl<float>(grad_u, u, f, m, n, srcx, srcy, h, grad_f, info, flags, stream);
}
int adt_eikonal3d_fwd_f64(const double* u0, const double* f, double h, int64_t m, int64_t n,
                          int64_t l, double tol, int32_t verbose, double* u, adt_info* info,
                          int32_t flags, void* stream) {
  return fwd3_impl<double>(u0, f, h, m, n, l, tol, verbose, u, info, flags, stream);
}
int adt_eikonal3d_fwd_f32(const float* u0, const float* f, float h, int64_t m, int64_t n,
                          int64_t l, float tol, int32_t verbose, float* u, adt_info* info,
                          int32_t flags, void* stream) {
  return fwd3_impl<float>(u0, f, h, m, n, l, tol, verbose, u, info, flags, stream);
}
int adt_eikonal3d_fwd_noalloc_f64(const double* u0, const double* f, double h, int64_t m,
                                  int64_t n, int64_t l, double tol, int32_t verbose, double* u,
                                  adt_info* info, void* stream) {
  return fwd3_noalloc_impl(u0, f, h, m, n, l, tol, verbose, u, info, stream);
}
int adt_eikonal3d_bwd_f64(const double* grad_u, const double* u, const double* u0,
                          const double* f, double h, int64_t m, int64_t n, int64_t l,
                          double* grad_u0, double* grad_f, adt_info* info, int32_t flags,
                          void* stream) {
  return bwd3_impl<double>(grad_u, u, u0, f, h, m, n, l, grad_u0, grad_f, info, flags, stream);
}
int adt_eikonal3d_bwd_f32(const float* grad_u, const float* u, const float* u0, const float* f,
                          float h, int64_t m, int64_t n, int64_t l, float* grad_u0,
                          float* grad_f, adt_info* info, int32_t flags, void* stream) {
  return bwd3_impl<float>(grad_u, u, u0, f, h, m, n, l, grad_u0, grad_f, info, flags, stream);
}
int adt_eikonal2d_fwd_strided_f64(const double* f, int64_t m, int64_t n, int64_t srcx,
                                  int64_t srcy, double h, const int64_t strides[2], double* u,
                                  adt_info* info, int32_t flags, void* stream) {
  return fwd2_strided_impl<double>(f, m, n, srcx, srcy, h, strides, u, info, flags, stream);
}
int adt_eikonal2d_fwd_strided_f32(const float* f, int64_t m, int64_t n, int64_t srcx,
                                  int64_t srcy, float h, const int64_t strides[2], float* u,
                                  adt_info* info, int32_t flags, void* stream) {
  return fwd2_strided_impl<float>(f, m, n, srcx, srcy, h, strides, u, info, flags, stream);
}
int adt_eikonal2d_bwd_strided_f64(const double* grad_u, const double* u, const double* f,
                                  int64_t m, int64_t n, int64_t srcx, int64_t srcy, double h,
                                  const int64_t strides[2], double* grad_f, adt_info* info,
                                  int32_t flags, void* stream) {
  return bwd2_strided_impl<double>(grad_u, u, f, m, n, srcx, srcy, h, strides, grad_f, info,
                                   flags, stream);
}
int adt_eikonal2d_bwd_strided_f32(const float* grad_u, const float* u, const float* f,
                                  int64_t m, int64_t n, int64_t srcx, int64_t srcy, float h,
                                  const int64_t strides[2], float* grad_f, adt_info* info,
                                  int32_t flags, void* stream) {
  return bwd2_strided_impl<float>(grad_u, u, f, m, n, srcx, srcy, h, strides, grad_f, info,
                                  flags, stream);
}
int adt_eikonal3d_fwd_strided_f64(const double* u0, const double* f, double h, int64_t m,
                                  int64_t n, int64_t l, double tol, int32_t verbose,
                                  const int64_t strides[3], double* u, adt_info* info,
                                  int32_t flags, void* stream) {
  return fwd3_strided_impl<double>(u0, f, h, m, n, l, tol, verbose, strides, u, info, flags,
                                   stream);
}
int adt_eikonal3d_fwd_strided_f32(const float* u0, const float* f, float h, int64_t m,
                                  int64_t n, int64_t l, float tol, int32_t verbose,
                                  const int64_t strides[3], float* u, adt_info* info,
                                  int32_t flags, void* stream) {
  return fwd3_strided_impl<float>(u0, f, h, m, n, l, tol, verbose, strides, u, info, flags,
                                  stream);
}
int adt_eikonal3d_bwd_strided_f64(const double* grad_u, const double* u, const double* u0,
                                  const double* f, double h, int64_t m, int64_t n, int64_t l,
                                  const int64_t strides[3], double* grad_u0, double* grad_f,
                                  adt_info* info, int32_t flags, void* stream) {
  return bwd3_strided_impl<double>(grad_u, u, u0, f, h, m, n, l, strides, grad_u0, grad_f, info,
                                   flags, stream);
}
int adt_eikonal3d_bwd_strided_f32(const float* grad_u, const float* u, const float* u0,
                                  const float* f, float h, int64_t m, int64_t n, int64_t l,
                                  const int64_t strides[3], float* grad_u0, float* grad_f,
                                  adt_info* info, int32_t flags, void* stream) {
  return bwd3_strided_impl<float>(grad_u, u, u0, f, h, m, n, l, strides, grad_u0, grad_f, info,
                                  flags, stream);
}
int adt_eikonal3d_batched_f64(const adt_event3d* events, int32_t nev, int32_t flags,
                              adt_info* infos, void* stream) {
  return batched_impl<double>(events, nev, flags, infos, stream);
}
int adt_eikonal3d_batched_f32(const adt_event3d* events, int32_t nev, int32_t flags,
                              adt_info* infos, void* stream) {
  return batched_impl<float>(events, nev, flags, infos, stream);
}
int adt_eikonal3d_misfit_batched_f64(const adt_event3d* events, const adt_misfit3d* items,
                                     int32_t nev, const double* rx, const double* ry,
                                     const double* rz, int64_t nrec, int32_t flags,
                                     adt_info* infos, void* stream) {
  const MisfitArgs m{items, rx, ry, rz, nrec};
  return batched_impl<double>(events, nev, flags, infos, stream, &m);
}
int adt_eikonal3d_misfit_batched_f32(const adt_event3d* events, const adt_misfit3d* items,
                                     int32_t nev, const double* rx, const double* ry,
                                     const double* rz, int64_t nrec, int32_t flags,
                                     adt_info* infos, void* stream) {
  const MisfitArgs m{items, rx, ry, rz, nrec};
  return batched_impl<float>(events, nev, flags, infos, stream, &m);
}
int adt_eikonal3d_fwd_plain_f64(const double* u0, const double* f, double h, int64_t m,
                                int64_t n, int64_t l, double tol, double* u, adt_info* info,
                                void* stream) {
  return fwd3_plain_impl<double>(u0, f, h, m, n, l, tol, u, info, stream);
}
int adt_eikonal3d_bwd_plain_f64(const double* grad_u, const double* u, const double* u0,
                                const double* f, double h, int64_t m, int64_t n, int64_t l,
                                double* grad_u0, double* grad_f, adt_info* info, void* stream) {
  return bwd3_plain_impl<double>(grad_u, u, u0, f, h, m, n, l, grad_u0, grad_f, info, stream);
}

int adt_graph_create(const adt_event3d* events, int32_t nev, int32_t flags, int32_t dtype,
                     adt_graph** out, void* stream) {
  if (!out) return fail(ADT_EINVAL, "graph: NULL out");
  if (flags & ADT_HOST_POINTERS) return fail(ADT_EINVAL, "graph: device pointers only");
  GraphBase* g = nullptr;
  int rc = dtype == 0 ? graph_create_impl<double>(events, nev, flags, &g, (hipStream_t)stream)
                      : graph_create_impl<float>(events, nev, flags, &g, (hipStream_t)stream);
  if (rc) return rc;
  *out = new adt_graph{g};
  return ADT_OK;
}
int adt_graph_create_misfit(const adt_event3d* events, const adt_misfit3d* items, int32_t nev,
                            const double* rx, const double* ry, const double* rz, int64_t nrec,
                            int32_t flags, int32_t dtype, adt_graph** out, void* stream) {
  if (!out) return fail(ADT_EINVAL, "graph: NULL out");
  if (flags & ADT_HOST_POINTERS) return fail(ADT_EINVAL, "graph: device pointers only");
  if (dtype != 0 && dtype != 1) return fail(ADT_EINVAL, "graph: dtype must be 0 (FP64) or 1 (FP32)");
  const MisfitArgs m{items, rx, ry, rz, nrec};
  GraphBase* g = nullptr;
  int rc = dtype == 0 ? graph_create_impl<double>(events, nev, flags, &g, (hipStream_t)stream, &m)
                      : graph_create_impl<float>(events, nev, flags, &g, (hipStream_t)stream, &m);
  if (rc) return rc;
  *out = new adt_graph{g};
  return ADT_OK;
}
int adt_graph_launch(adt_graph* g, void* stream) {
  if (!g) return fail(ADT_EINVAL, "graph: NULL");
  ADT_HIP(hipGraphLaunch(g->g->exec, (hipStream_t)stream));
  return ADT_OK;
}
int adt_graph_infos(adt_graph* g, adt_info* infos, void* stream) {
  if (!g || !infos) return fail(ADT_EINVAL, "graph: NULL");
  hipStream_t s = (hipStream_t)stream;
  ADT_HIP(hipMemcpyAsync(infos, (char*)g->g->ws + g->g->off_info, sizeof(adt_info) * g->g->nev,
                         hipMemcpyDeviceToHost, s));
  ADT_HIP(hipStreamSynchronize(s));
  for (int e = 0; e < g->g->nev; ++e)
    if (infos[e].fault)
      return fail(ADT_EHIP, "graph: wave synchronisation timed out in event " + std::to_string(e));
  return ADT_OK;
}
int adt_graph_destroy(adt_graph* g) {
  if (!g) return ADT_OK;
  if (g->g->exec) (void)hipGraphExecDestroy(g->g->exec);
  if (g->g->graph) (void)hipGraphDestroy(g->g->graph);
  if (g->g->ws) (void)hipFree(g->g->ws);
  delete g->g;
  delete g;
  return ADT_OK;
}

}  // extern "C"

// ----------------------------------------------------------- objective ---

namespace {

template <typename T>
int source_u0_impl(const adt_source3d* src, int32_t nsrc, const T* vel0, int64_t m, int64_t n,
                   int64_t l, T h, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (nsrc < 0 || (nsrc > 0 && (!src || !vel0))) return fail(ADT_EINVAL, "source_u0: bad arguments");
  if (nsrc == 0) return ADT_OK;
  const double ext[3] = {(double)m, (double)n, (double)l};
  for (int s = 0; s < nsrc; ++s) {
    const double c[3] = {src[s].x, src[s].y, src[s].z};
    if (!src[s].u0) return fail(ADT_EINVAL, "source_u0: NULL u0 for station " + std::to_string(s));
    for (int d = 0; d < 3; ++d)
      if (!(c[d] >= 1.0 && c[d] <= ext[d]))
        return fail(ADT_EINVAL, "source_u0: station " + std::to_string(s) +
                                    " outside the 1-based grid");
  }
  hipStream_t st = (hipStream_t)stream;
  DevBuf tab;
  if ((rc = tab.alloc(sizeof(adt_source3d) * nsrc, st))) return rc;
  ADT_HIP(hipMemcpyAsync(tab.p, src, sizeof(adt_source3d) * nsrc, hipMemcpyHostToDevice, st));
  ADT_HIP(hipStreamSynchronize(st));  // the table is the caller's host memory
  adt::source_u0_kernel<T><<<nsrc, 256, 0, st>>>((const adt_source3d*)tab.p, vel0, (int)m, (int)n,
                                                 (int)l, h);
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

template <typename T>
int misfit_impl(const adt_misfit3d* items, int32_t nitem, const double* rx, const double* ry,
                const double* rz, int64_t nrec, void* stream) {
  if (nitem < 0 || nrec < 0 || nrec > (int64_t(1) << 30) ||
      (nitem > 0 && nrec > 0 && (!items || !rx || !ry || !rz)))
    return fail(ADT_EINVAL, "misfit: bad arguments");
  if (nitem == 0) return ADT_OK;
  for (int i = 0; i < nitem; ++i) {
    int rc = check_misfit_item(items[i], i, rx, ry, rz, nrec);
    if (rc) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  size_t used = 0;
  auto take = [&](size_t b) {
    size_t o = used;
    used += ((b + 255) / 256) * 256;
    return o;
  };
  const size_t o_tab = take(sizeof(adt_misfit3d) * nitem);
  const size_t o_coords = take(sizeof(double) * 3 * (nrec ? nrec : 1));
  RecvNodes<T> rn;
  rn.build(items, nitem, rx, ry, rz, nrec, take);
  DevBuf ws;
  int rc;
  if ((rc = ws.alloc(used, st))) return rc;
  char* base = (char*)ws.p;
  double* dc = (double*)(base + o_coords);
  ADT_HIP(hipMemcpyAsync(base + o_tab, items, sizeof(adt_misfit3d) * nitem, hipMemcpyHostToDevice, st));
  if (nrec) {
    ADT_HIP(hipMemcpyAsync(dc, rx, sizeof(double) * nrec, hipMemcpyHostToDevice, st));
    ADT_HIP(hipMemcpyAsync(dc + nrec, ry, sizeof(double) * nrec, hipMemcpyHostToDevice, st));
    ADT_HIP(hipMemcpyAsync(dc + 2 * nrec, rz, sizeof(double) * nrec, hipMemcpyHostToDevice, st));
  }
  if ((rc = rn.upload(base, nitem, st))) return rc;
  adt::misfit_kernel<T><<<nitem, 256, 0, st>>>((const adt_misfit3d*)(base + o_tab), rn.table(base),
                                                dc, dc + nrec, dc + 2 * nrec, (int)nrec);
  ADT_HIP(hipGetLastError());
  // the staged coordinates are host copies: wait before they go out of scope
  ADT_HIP(hipStreamSynchronize(st));
  return ADT_OK;
}

int grid_for(int64_t N) {
  const int64_t b = (N + 255) / 256;
  return (int)(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

template <typename T>
int model_impl(const T* x, const T* vel0, int64_t N, T* fvar, T* f, void* stream) {
  if (N < 0 || (N > 0 && (!x || !vel0 || !fvar))) return fail(ADT_EINVAL, "model3d: bad arguments");
  if (N == 0) return ADT_OK;
  adt::model_kernel<T><<<grid_for(N), 256, 0, (hipStream_t)stream>>>(x, vel0, N, fvar, f);
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

template <typename T>
int model_grad_impl(const T* x, const T* fvar, const T* grad_f, const T* grad_fvar, int64_t N,
                    T* grad_x, void* stream) {
  if (N < 0 || (N > 0 && (!x || !fvar || !grad_x)))
    return fail(ADT_EINVAL, "model3d_grad: bad arguments");
  if (N == 0) return ADT_OK;
  adt::model_grad_kernel<T><<<grid_for(N), 256, 0, (hipStream_t)stream>>>(x, fvar, grad_f,
                                                                          grad_fvar, N, grad_x);
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

template <typename T>
int smooth_impl(const T* fvar, int64_t m, int64_t n, int64_t l, int32_t sh, int32_t sv, T lambda,
                T* grad_fvar, T* value, void* stream) {
  int rc = check_dims3(m, n, l);
  if (rc) return rc;
  if (!fvar || !value) return fail(ADT_EINVAL, "smooth3d: NULL buffer");
  if (sh < 1 || sv < 1 || !(sh & 1) || !(sv & 1) || sh > m || sh > n || sv > l)
    return fail(ADT_EINVAL, "smooth3d: window sizes must be odd and fit the grid");
  hipStream_t st = (hipStream_t)stream;
  const int64_t N = m * n * l;
  const int nb = grid_for(N);
  DevBuf ws;
  if ((rc = ws.alloc(sizeof(T) * (N + nb), st))) return rc;
  T* sgn = (T*)ws.p;
  T* partial = sgn + N;
  const T wgt = T(1) / T((double)sh * sh * sv);
  adt::smooth_pass1_kernel<T><<<nb, 256, 0, st>>>(fvar, (int)m, (int)n, (int)l, sh / 2, sv / 2, wgt,
                                                  sgn, partial);
  ADT_HIP(hipGetLastError());
  if (grad_fvar) {
    adt::smooth_pass2_kernel<T><<<nb, 256, 0, st>>>(sgn, (int)m, (int)n, (int)l, sh / 2, sv / 2,
                                                    wgt, lambda, grad_fvar, partial, nb, value);
  } else {
    // value only: one thread sums the block partials
    adt::smooth_pass2_kernel<T><<<1, 1, 0, st>>>(sgn, (int)m, (int)n, (int)l, sh / 2, sv / 2, wgt,
                                                 lambda, nullptr, partial, nb, value);
  }
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

template <typename T>
int slowness_impl(const T* fvar, int64_t N, T scale, T* f_s, void* stream) {
  if (N < 0 || (N > 0 && (!fvar || !f_s))) return fail(ADT_EINVAL, "slowness3d: bad arguments");
  if (N == 0) return ADT_OK;
  adt::slowness_kernel<T><<<grid_for(N), 256, 0, (hipStream_t)stream>>>(fvar, N, scale, f_s);
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

template <typename T>
int slowness_grad_impl(const T* fvar, int64_t N, T scale, const T* grad_fs, T* grad_fvar,
                       T* grad_scale, void* stream) {
  if (N < 0 || (N > 0 && (!fvar || !grad_fs)))
    return fail(ADT_EINVAL, "slowness3d_grad: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  const int nb = grid_for(N);
  DevBuf ws;
  int rc;
  if ((rc = ws.alloc(sizeof(T) * nb, st))) return rc;
  if (N > 0) {
    adt::slowness_grad_kernel<T><<<nb, 256, 0, st>>>(fvar, N, scale, grad_fs, grad_fvar, (T*)ws.p);
    ADT_HIP(hipGetLastError());
  }
  if (grad_scale) {
    adt::sum_partials_kernel<T><<<1, 64, 0, st>>>((const T*)ws.p, N > 0 ? nb : 0, grad_scale);
    ADT_HIP(hipGetLastError());
  }
  return ADT_OK;
}

template <typename T>
int sum_arrays_impl(const T* const* arrays, int32_t narr, int64_t N, T* out, void* stream) {
  if (narr < 1 || N < 0 || !arrays || !out) return fail(ADT_EINVAL, "sum_arrays: bad arguments");
  for (int k = 0; k < narr; ++k)
    if (!arrays[k]) return fail(ADT_EINVAL, "sum_arrays: NULL array");
  if (N == 0) return ADT_OK;
  hipStream_t st = (hipStream_t)stream;
  DevBuf tab;
  int rc;
  if ((rc = tab.alloc(sizeof(T*) * narr, st))) return rc;
  ADT_HIP(hipMemcpyAsync(tab.p, arrays, sizeof(T*) * narr, hipMemcpyHostToDevice, st));
  ADT_HIP(hipStreamSynchronize(st));  // the pointer table is the caller's host memory
  adt::sum_arrays_kernel<T><<<grid_for(N), 256, 0, st>>>((const T* const*)tab.p, narr, N, out);
  ADT_HIP(hipGetLastError());
  return ADT_OK;
}

}  // namespace

extern "C" {

int adt_source_u0_3d_batched_f64(const adt_source3d* src, int32_t nsrc, const double* vel0,
                                 int64_t m, int64_t n, int64_t l, double h, void* stream) {
  return source_u0_impl<double>(src, nsrc, vel0, m, n, l, h, stream);
}
int adt_source_u0_3d_batched_f32(const adt_source3d* src, int32_t nsrc, const float* vel0,
                                 int64_t m, int64_t n, int64_t l, float h, void* stream) {
  return source_u0_impl<float>(src, nsrc, vel0, m, n, l, h, stream);
}
int adt_misfit3d_batched_f64(const adt_misfit3d* items, int32_t nitem, const double* rx,
                             const double* ry, const double* rz, int64_t nrec, void* stream) {
  return misfit_impl<double>(items, nitem, rx, ry, rz, nrec, stream);
}
int adt_misfit3d_batched_f32(const adt_misfit3d* items, int32_t nitem, const double* rx,
                             const double* ry, const double* rz, int64_t nrec, void* stream) {
  return misfit_impl<float>(items, nitem, rx, ry, rz, nrec, stream);
}
int adt_model3d_f64(const double* x, const double* vel0, int64_t N, double* fvar, double* f,
                    void* stream) {
  return model_impl<double>(x, vel0, N, fvar, f, stream);
}
int adt_model3d_f32(const float* x, const float* vel0, int64_t N, float* fvar, float* f,
                    void* stream) {
  return model_impl<float>(x, vel0, N, fvar, f, stream);
}
int adt_model3d_grad_f64(const double* x, const double* fvar, const double* grad_f,
                         const double* grad_fvar, int64_t N, double* grad_x, void* stream) {
  return model_grad_impl<double>(x, fvar, grad_f, grad_fvar, N, grad_x, stream);
}
int adt_model3d_grad_f32(const float* x, const float* fvar, const float* grad_f,
                         const float* grad_fvar, int64_t N, float* grad_x, void* stream) {
  return model_grad_impl<float>(x, fvar, grad_f, grad_fvar, N, grad_x, stream);
}
int adt_smooth3d_l1_f64(const double* fvar, int64_t m, int64_t n, int64_t l, int32_t sh,
                        int32_t sv, double lambda, double* grad_fvar, double* value,
                        void* stream) {
  return smooth_impl<double>(fvar, m, n, l, sh, sv, lambda, grad_fvar, value, stream);
}
int adt_smooth3d_l1_f32(const float* fvar, int64_t m, int64_t n, int64_t l, int32_t sh,
                        int32_t sv, float lambda, float* grad_fvar, float* value, void* stream) {
  return smooth_impl<float>(fvar, m, n, l, sh, sv, lambda, grad_fvar, value, stream);
}
int adt_slowness3d_f64(const double* fvar, int64_t N, double scale, double* f_s, void* stream) {
  return slowness_impl<double>(fvar, N, scale, f_s, stream);
}
int adt_slowness3d_f32(const float* fvar, int64_t N, float scale, float* f_s, void* stream) {
  return slowness_impl<float>(fvar, N, scale, f_s, stream);
}
int adt_slowness3d_grad_f64(const double* fvar, int64_t N, double scale, const double* grad_fs,
                            double* grad_fvar, double* grad_scale, void* stream) {
  return slowness_grad_impl<double>(fvar, N, scale, grad_fs, grad_fvar, grad_scale, stream);
}
int adt_slowness3d_grad_f32(const float* fvar, int64_t N, float scale, const float* grad_fs,
                            float* grad_fvar, float* grad_scale, void* stream) {
  return slowness_grad_impl<float>(fvar, N, scale, grad_fs, grad_fvar, grad_scale, stream);
}
int adt_sum_arrays_f64(const double* const* arrays, int32_t narr, int64_t N, double* out,
                       void* stream) {
  return sum_arrays_impl<double>(arrays, narr, N, out, stream);
}
int adt_sum_arrays_f32(const float* const* arrays, int32_t narr, int64_t N, float* out,
                       void* stream) {
  return sum_arrays_impl<float>(arrays, narr, N, out, stream);
}

int adt_selftest_arith_f64(int32_t op, const double* a, const double* b, double* out, int64_t n,
                           void* stream) {
  if (n < 0 || !a || !out || op < 0 || op > 8 || ((op == 3 || op == 4 || op == 5 || op == 8) && !b))
    return fail(ADT_EINVAL, "selftest: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  adt::arith_kernel<double><<<op == 6 ? 4096 : 256, 256, 0, s>>>(op, a, b, out, n);
  ADT_HIP(hipGetLastError());
  ADT_HIP(hipStreamSynchronize(s));
  return ADT_OK;
}
int adt_selftest_arith_f32(int32_t op, const float* a, const float* b, float* out, int64_t n,
                           void* stream) {
  if (n < 0 || !a || !out || op < 0 || op > 8 || ((op == 3 || op == 4 || op == 5 || op == 8) && !b))
    return fail(ADT_EINVAL, "selftest: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  adt::arith_kernel<float><<<op == 6 ? 4096 : 256, 256, 0, s>>>(op, a, b, out, n);
  ADT_HIP(hipGetLastError());
  ADT_HIP(hipStreamSynchronize(s));
  return ADT_OK;
}

}  // extern "C"
