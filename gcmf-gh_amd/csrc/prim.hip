// prim.hip -- device primitives for gfx950: exclusive scan, stream
// compaction and a stable LSD radix sort.  These replace the reference's
// multi-level Hillis-Steele scans (GPU_PrefixSum.h, GPU_Utility.h:62-219),
// Make_Filtered_Index_List (GPU_Vector.h:108-120) and the one-hot decimal
// radix sort (GPU_RadixSort.h, GPU_Utility.h:224-289).
//
// Scan: reduce-then-scan over 4096-element tiles (256 threads x 16 items),
// wave64 shuffles inside a wave, LDS across the 4 waves of a block.
// Radix sort: 8-bit digits, LDS digit histograms per 4096-key tile, a
// digit-major scan of the histograms, and a stable scatter ranked with
// 8 wave ballots per 64 keys (peer mask = keys with the same digit).  The
// scatter and the compaction's write pass take their tiles through
// xcd_tile, so that neighbouring tiles, whose outputs share cache lines at
// every digit run's ends, run on one XCD (measured on MI355X at 2^27 keys:
// 1.57 -> 0.75 ms per scatter pass).
#include "gcmf_device.hpp"
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/gcmf_hip.h"
#include "gcmf_internal.hpp"

namespace gcmf {

// ------------------------------------------------------------ context ----
void *Workspace::get(const std::string &name, size_t bytes) {
  if (bytes == 0) bytes = 256;
  Slot &s = slots[name];
  if (s.bytes < bytes) {
    if (no_alloc)
      throw Error{GCMF_ERR_ARG, "capture would allocate workspace '" + name + "'"};
    if (s.ptr) HIP_CHECK(hipFree(s.ptr));
    s.ptr = nullptr;
    size_t want = bytes + bytes / 4;  // headroom against small growth
    HIP_CHECK(hipMalloc(&s.ptr, want));
    s.bytes = want;
  }
  return s.ptr;
}

void Workspace::release_all() {
  for (auto &kv : slots)
    if (kv.second.ptr) (void)hipFree(kv.second.ptr);
  slots.clear();
}

hipEvent_t Context::take_event() {
  if (!event_pool.empty()) {
    hipEvent_t e = event_pool.back();
    event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_CHECK(hipEventCreate(&e));
  return e;
}

void Context::flush_timers() {
  if (pending.empty()) return;
  HIP_CHECK(hipStreamSynchronize(stream));
  for (auto &p : pending) {
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    KernelTimer &t = timers[p.first];
    t.total_ms += ms;
    t.launches += 1;
    event_pool.push_back(p.second.first);
    event_pool.push_back(p.second.second);
  }
  pending.clear();
}

// true when launches named n are timed: every name when the filter is
// empty, else the names listed in it (comma-separated)
static bool timed_kernel(const std::string &only, const char *n) {
  if (only.empty()) return true;
  const size_t len = std::strlen(n);
  size_t a = 0;
  while (a <= only.size()) {
    size_t z = only.find(',', a);
    if (z == std::string::npos) z = only.size();
    if (z - a == len && only.compare(a, len, n) == 0) return true;
    a = z + 1;
  }
  return false;
}

LaunchTimer::LaunchTimer(Context &ctx, const char *n) : c(ctx), name(n) {
  if (c.timing && timed_kernel(c.timing_only, n)) {
    a = c.take_event();
    b = c.take_event();
    HIP_CHECK(hipEventRecord(a, c.stream));
  }
}

LaunchTimer::~LaunchTimer() {
  if (c.timing && a) {
    if (hipEventRecord(b, c.stream) == hipSuccess) c.pending.push_back({name, {a, b}});
  }
}

// --------------------------------------------------------- read-backs ----
struct ScalarSrc {
  const long long *p[kHostSlots];
};

__global__ void k_gather_scalars(ScalarSrc s, int n, long long *dst) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = s.p[i] ? *s.p[i] : 0;
}

struct ScalarVals {
  long long v[kHostSlots];
};

// the pinned staging buffer for small read-backs, mapped into the device's
// address space and coherent, so that a kernel stores the values straight
// into host memory (no copy launch); returns its device address
static long long *host_slots(Context &c) {
  if (!c.hbuf) {
    HIP_CHECK(hipHostMalloc((void **)&c.hbuf, sizeof(long long) * kHostSlots,
                            hipHostMallocMapped | hipHostMallocCoherent));
    HIP_CHECK(hipHostGetDevicePointer((void **)&c.hbuf_dev, c.hbuf, 0));
  }
  return c.hbuf_dev;
}

__global__ void k_copy_small(const unsigned char *src, int bytes, unsigned char *dst) {
  for (int i = threadIdx.x; i < bytes; i += blockDim.x) dst[i] = src[i];
}

// sets *flag when a source's value differs from the recorded one
__global__ void k_check_scalars(ScalarSrc s, int n, ScalarVals want, int *flag) {
  const int i = threadIdx.x;
  if (i < n && (s.p[i] ? *s.p[i] : 0) != want.v[i]) *flag = 1;
}

// sets *flag when the n bytes at a differ from those at b
__global__ void k_check_bytes(const unsigned char *a, const unsigned char *b, long n, int *flag) {
  bool diff = false;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x)
    diff |= a[i] != b[i];
  if (diff) *flag = 1;
}

void host_read(Context &c, void *dst, const void *src, size_t bytes) {
  if (bytes == 0) return;
  if (c.replay == 2) {
    if (c.rec_bytes_pos >= c.rec_bytes.size() || c.rec_bytes[c.rec_bytes_pos].size() != bytes)
      throw Error{GCMF_ERR_ARG, "capture: read-back sequence differs from the recorded run"};
    const size_t k = c.rec_bytes_pos++;
    std::memcpy(dst, c.rec_bytes[k].data(), bytes);
    const void *snap = c.ws.get("replay_snap_" + std::to_string(k), bytes);
    hipLaunchKernelGGL(k_check_bytes, dim3(std::min<unsigned>(grid_for((long)bytes, 256), 64)),
                       dim3(256), 0, c.stream, (const unsigned char *)src,
                       (const unsigned char *)snap, (long)bytes, c.graph_flag);
    HIP_CHECK(hipGetLastError());
    return;
  }
  if (c.replay == 1) {
    void *snap = c.ws.get("replay_snap_" + std::to_string(c.rec_bytes.size()), bytes);
    HIP_CHECK(hipMemcpyAsync(snap, src, bytes, hipMemcpyDeviceToDevice, c.stream));
  }
  if (bytes <= sizeof(long long) * kHostSlots) {
    // a few bytes: stored into the mapped staging buffer by a kernel
    hipLaunchKernelGGL(k_copy_small, dim3(1), dim3(64), 0, c.stream,
                       (const unsigned char *)src, (int)bytes, (unsigned char *)host_slots(c));
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(c.stream));
    std::memcpy(dst, c.hbuf, bytes);
  } else {
    HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
  }
  if (c.replay == 1) {
    const char *p = (const char *)dst;
    c.rec_bytes.emplace_back(p, p + bytes);
  }
}

// the look-back error word rides along as one more slot
void read_scalars(Context &c, std::initializer_list<const long long *> src, long long *out) {
  const int n = (int)src.size();
  if (n + 1 > kHostSlots) throw Error{GCMF_ERR_ARG, "too many scalars in one read-back"};
  ScalarSrc s{};
  int i = 0;
  for (const long long *p : src) s.p[i++] = p;
  s.p[n] = c.lb_err;
  if (c.replay == 2) {
    if (c.rec_pos + n > c.rec.size())
      throw Error{GCMF_ERR_ARG, "capture: read-back sequence differs from the recorded run"};
    ScalarVals want{};
    for (int j = 0; j < n; j++) want.v[j] = out[j] = c.rec[c.rec_pos++];
    want.v[n] = 0;  // no look-back overrun
    hipLaunchKernelGGL(k_check_scalars, dim3(1), dim3(64), 0, c.stream, s, n + 1, want,
                       c.graph_flag);
    HIP_CHECK(hipGetLastError());
    return;
  }
  // gathered straight into the mapped staging buffer
  hipLaunchKernelGGL(k_gather_scalars, dim3(1), dim3(64), 0, c.stream, s, n + 1, host_slots(c));
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c.stream));
  for (int j = 0; j < n; j++) out[j] = c.hbuf[j];
  if (c.replay == 1) c.rec.insert(c.rec.end(), out, out + n);
  if (c.hbuf[n] != 0) {
    fill_i32(c, (int *)c.lb_err, 0, 2);
    throw Error{GCMF_ERR_HIP, "a scan's look-back wait exceeded its bound"};
  }
}

void check_lookback(Context &c) {
  long long v;
  read_scalars(c, {}, &v);
}

// --------------------------------------------------------------- fill ----
__global__ void k_fill_i32(int *p, int v, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_i32(Context &c, int *p, int v, long n) {
  if (n <= 0) return;
  LaunchTimer t(c, "fill_i32");
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, p, v, n);
  HIP_CHECK(hipGetLastError());
}

// --------------------------------------------------------------- scan ----
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;
// up to this many tiles every tile sums its predecessors' partials itself
constexpr long kScanPrefixTiles = 2048;

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    T u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T *lds_waves, T &total) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  T inc = wave_incl_scan(v);
  if (lane == 63) lds_waves[wave] = inc;
  __syncthreads();
  T wbase = 0, tot = 0;
  const int nw = blockDim.x >> 6;
  for (int w = 0; w < nw; w++) {
    T s = lds_waves[w];
    if (w < wave) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + inc - v;
}

template <typename TI, typename TA>
__global__ void __launch_bounds__(kScanThreads) k_tile_reduce(const TI *in, long n, TA *partials) {
  __shared__ TA lds[kScanThreads / 64];
  const long base = (long)blockIdx.x * kScanTile;
  TA s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    long i = base + (long)k * kScanThreads + threadIdx.x;
    if (i < n) s += (TA)in[i];
  }
  TA tot;
  block_excl_scan<TA>(s, lds, tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// Scans one tile; adds offsets[blockIdx.x] (if given).  When total_out is
// given the block writes its grand total there (single-tile scans).
template <typename TI, typename TA, typename TO>
__global__ void __launch_bounds__(kScanThreads) k_tile_scan(const TI *in, long n, TO *out,
                                                            const TA *offsets, TO *total_out) {
  __shared__ TA tile[kScanTile];
  __shared__ TA lds[kScanThreads / 64];
  const long base = (long)blockIdx.x * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    long i = base + j;
    tile[j] = i < n ? (TA)in[i] : (TA)0;
  }
  __syncthreads();
  TA local[kScanItems];
  TA s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    local[k] = tile[threadIdx.x * kScanItems + k];
    s += local[k];
  }
  TA tot;
  TA run = block_excl_scan<TA>(s, lds, tot);
  TA off = offsets ? offsets[blockIdx.x] : (TA)0;
  run += off;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    tile[threadIdx.x * kScanItems + k] = run;
    run += local[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    long i = base + j;
    if (i < n) out[i] = (TO)tile[j];
  }
  if (total_out && threadIdx.x == 0) *total_out = (TO)(off + tot);
}

// Scans tile b after summing the partials of tiles [0, b) itself; the last
// tile writes the grand total to out[n].  Used while the partials are few
// enough for every block to read them.
template <typename TI, typename TO>
__global__ void __launch_bounds__(kScanThreads)
    k_tile_scan_prefix(const TI *in, long n, TO *out, const long long *partials, long tiles) {
  __shared__ long long lds[kScanThreads / 64];
  long long pre = 0;
  for (long t = threadIdx.x; t < blockIdx.x; t += kScanThreads) pre += partials[t];
  long long ptot;
  block_excl_scan<long long>(pre, lds, ptot);
  __shared__ long long tile[kScanTile];
  const long base = (long)blockIdx.x * kScanTile;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    long i = base + j;
    tile[j] = i < n ? (long long)in[i] : 0LL;
  }
  __syncthreads();
  long long local[kScanItems];
  long long s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    local[k] = tile[threadIdx.x * kScanItems + k];
    s += local[k];
  }
  long long tot;
  long long run = block_excl_scan<long long>(s, lds, tot) + ptot;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    tile[threadIdx.x * kScanItems + k] = run;
    run += local[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    int j = k * kScanThreads + threadIdx.x;
    long i = base + j;
    if (i < n) out[i] = (TO)tile[j];
  }
  if (blockIdx.x == tiles - 1 && threadIdx.x == 0) out[n] = (TO)(ptot + tot);
}

template <typename TO>
__global__ void k_copy_total(TO *dst, const long long *src) {
  if (threadIdx.x == 0) *dst = (TO)*src;
}

// ---- i32 -> i64 scans of large arrays: 16-byte loads and stores.
// Thread t of a tile holds the four int4 groups at tile offsets k * 1024 +
// 4 t (k = 0..3), so each load instruction of a wave reads 1 KiB
// contiguously; the tile is scanned group by group with a carry, and every
// thread writes its four results of a group as two 16-byte stores.
constexpr int kVecTile = 4096;

__device__ __forceinline__ int4 ld_i4(const int *in, long long base, int k) {
  return *(const int4 *)(in + base + k * 1024 + 4 * threadIdx.x);
}

__global__ void __launch_bounds__(256) k_vscan_reduce(const int *in, long n, long long *partials) {
  __shared__ long long lds[4];
  const long long base = (long long)blockIdx.x * kVecTile;
  long long s = 0;
  if (base + kVecTile <= n) {
    int4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = ld_i4(in, base, k);
#pragma unroll
    for (int k = 0; k < 4; k++) s += (long long)v[k].x + v[k].y + v[k].z + v[k].w;
  } else {
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) {
        const long long i = base + k * 1024 + 4 * threadIdx.x + j;
        if (i < n) s += in[i];
      }
  }
  long long tot;
  block_excl_scan<long long>(s, lds, tot);
  if (threadIdx.x == 0) partials[blockIdx.x] = tot;
}

// offsets[b] = the sum of tiles before b
__global__ void __launch_bounds__(256)
    k_vscan_down(const int *in, long n, long long *out, const long long *offsets) {
  __shared__ long long lds[4];
  const long long base = (long long)blockIdx.x * kVecTile;
  long long carry = offsets[blockIdx.x];
  const bool full = base + kVecTile <= n;
  int4 v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (full) {
      v[k] = ld_i4(in, base, k);
    } else {
      const long long i = base + k * 1024 + 4 * threadIdx.x;
      v[k].x = i < n ? in[i] : 0;
      v[k].y = i + 1 < n ? in[i + 1] : 0;
      v[k].z = i + 2 < n ? in[i + 2] : 0;
      v[k].w = i + 3 < n ? in[i + 3] : 0;
    }
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const long long s = (long long)v[k].x + v[k].y + v[k].z + v[k].w;
    long long tot;
    const long long e = carry + block_excl_scan<long long>(s, lds, tot);
    const long long o0 = e, o1 = o0 + v[k].x, o2 = o1 + v[k].y, o3 = o2 + v[k].z;
    const long long i = base + k * 1024 + 4 * threadIdx.x;
    if (full) {
      longlong2 *d = (longlong2 *)(out + i);
      d[0] = make_longlong2(o0, o1);
      d[1] = make_longlong2(o2, o3);
    } else {
      if (i < n) out[i] = o0;
      if (i + 1 < n) out[i + 1] = o1;
      if (i + 2 < n) out[i + 2] = o2;
      if (i + 3 < n) out[i + 3] = o3;
    }
    carry += tot;
  }
}

// ---- single-pass scans with decoupled look-back (Merrill and Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", NVIDIA
// technical report NVR-2016-002), over the same 4096-element tiles.
//
// A block takes its tile from an atomic counter, so every tile it waits for
// belongs to a block that is already running, and that block publishes its
// tile's sum before it waits for anything: every wait ends.  The wait is
// still bounded (kLbSpinLimit sleeps); past the bound the block records an
// error in the context's error word, which the next read-back raises.
//
// A tile descriptor is one 64-bit word, value * 4 + status (kLbAgg: the
// tile's own sum; kLbIncl: the sum of all tiles through this one), stored
// and loaded with device-scope atomics.  The value is the only data passed
// between blocks, so no other memory needs ordering across the XCDs' L2s.
// Values must stay inside [-2^61, 2^61), which holds for fewer than 2^30
// int32 inputs.
//
// Each poll is a device-scope load that bypasses the XCD's L2, and tile t
// reaches the nearest published inclusive prefix only after the tiles
// before it have published theirs, so the chain of look-backs, not HBM,
// bounds this path on long arrays: measured on MI355X, 0.67 ms for 2^27
// elements against 0.44 ms for the reduce-then-scan path.  It is used up
// to 2^22 elements (1024 tiles), where it replaces that path's four or more
// launches by one.  One launch may scan several arrays (segments, up to
// kLbSegs): each segment's first tile publishes an inclusive prefix, so no
// look-back crosses into the segment before it.
constexpr unsigned long long kLbAgg = 1, kLbIncl = 2;
constexpr long kLbMaxN = 1L << 22;
constexpr int kLbSpinLimit = 1 << 22;
// descriptors per half: the tiles of one launch and the tile counter
constexpr long kLbSlots = 4096;
static_assert(kLbMaxN / 4096 + 1 <= kLbSlots, "descriptor half too small");

// every launch clears the other half, which the next launch uses
__device__ __forceinline__ void lb_clear_next(unsigned long long *next) {
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < kLbSlots;
       j += (long)gridDim.x * blockDim.x)
    next[j] = 0;
}

__device__ __forceinline__ void lb_publish(unsigned long long *d, long long v,
                                           unsigned long long st) {
  __hip_atomic_store(d, ((unsigned long long)v << 2) | st, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// The sum of all tiles before tile t (t > 0), by one wave: lane l reads the
// descriptor of tile j - l, 64 tiles a window, until the nearest inclusive
// descriptor with every descriptor after it published.
__device__ long long lb_lookback(const unsigned long long *desc, long t, long long *err) {
  const int lane = __lane_id();
  long long prefix = 0;
  long j = t - 1;
  int spins = 0;
  while (true) {
    const long idx = j - lane;
    unsigned long long d;
    unsigned long long ready, incl;
    while (true) {
      d = idx >= 0 ? __hip_atomic_load(desc + idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                   : kLbIncl;
      const unsigned long long notReady = __ballot((d & 3) == 0);
      incl = __ballot((d & 3) == kLbIncl);
      // lanes before the first unpublished one can be used
      ready = notReady ? ((1ULL << (__ffsll((long long)notReady) - 1)) - 1ULL) : ~0ULL;
      if ((incl & ready) || !notReady) break;
      if (++spins > kLbSpinLimit) {
        if (lane == 0) atomicExch((unsigned long long *)err, 1ULL);
        incl = 1;  // give up: treat the nearest tile as the end
        ready = 1;
        d = lane == 0 ? (d & ~3ULL) | kLbIncl : 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const unsigned long long hit = incl & ready;
    const int stop = hit ? __ffsll((long long)hit) - 1 : 63;
    long long v = lane <= stop ? ((long long)d >> 2) : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    prefix += v;
    if (hit) return prefix;
    j -= 64;
  }
}

// Exclusive scan of four values per thread over a 256-thread block, in
// group-major order (every thread's s[0], then every thread's s[1], ...).
__device__ __forceinline__ void block_excl_scan4(const long long s[4], long long e[4],
                                                 long long &total, long long (*lds)[4]) {
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  long long inc[4];
#pragma unroll
  for (int k = 0; k < 4; k++) inc[k] = wave_incl_scan(s[k]);
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 4; k++) lds[wave][k] = inc[k];
  __syncthreads();
  long long run = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    long long wb = 0, gt = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      const long long x = lds[w][k];
      wb += w < wave ? x : 0;
      gt += x;
    }
    e[k] = run + wb + inc[k] - s[k];
    run += gt;
  }
  total = run;
}

__device__ __forceinline__ void ld_tile4(const int *in, long n, long long base, bool full,
                                         int4 v[4]) {
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (full) {
      v[k] = ld_i4(in, base, k);
    } else {
      const long long i = base + k * 1024 + 4 * threadIdx.x;
      v[k].x = i < n ? in[i] : 0;
      v[k].y = i + 1 < n ? in[i + 1] : 0;
      v[k].z = i + 2 < n ? in[i + 2] : 0;
      v[k].w = i + 3 < n ? in[i + 3] : 0;
    }
  }
}

// Takes the block's tile, publishes its sum, and returns the sum of the
// tiles before it (every thread).
__device__ __forceinline__ long long lb_tile_prefix(unsigned long long *desc, long t,
                                                    long long tot, long long *err,
                                                    long long *s_pre, bool first) {
  if (threadIdx.x == 0) lb_publish(desc + t, tot, first ? kLbIncl : kLbAgg);
  if (!first) {
    if (threadIdx.x < 64) {
      const long long pre = lb_lookback(desc, t, err);
      if (threadIdx.x == 0) {
        lb_publish(desc + t, pre + tot, kLbIncl);
        *s_pre = pre;
      }
    }
  } else if (threadIdx.x == 0) {
    *s_pre = 0;
  }
  __syncthreads();
  return *s_pre;
}

__global__ void k_lb_clear_all(unsigned long long *desc) {
  for (long j = threadIdx.x; j < 2 * kLbSlots; j += blockDim.x) desc[j] = 0;
}

// up to kLbSegs arrays scanned by one launch: segment j takes the tiles
// [t0[j], t0[j + 1]) (at least one, so that out[0] = 0 of an empty array
// is written too)
constexpr int kLbSegs = 4;
struct LbSegs {
  const int *in[kLbSegs];
  long long *out[kLbSegs];
  long n[kLbSegs];
  long t0[kLbSegs + 1];
  int k;
};

// desc[tiles] holds the tile counter
__global__ void __launch_bounds__(256)
    k_lb_scan(LbSegs S, unsigned long long *desc, unsigned long long *next, long tiles,
              long long *err) {
  __shared__ long long lds[4][4];
  __shared__ long long s_pre;
  __shared__ long s_tile;
  lb_clear_next(next);
  if (threadIdx.x == 0) s_tile = (long)atomicAdd(desc + tiles, 1ULL);
  __syncthreads();
  const long t = s_tile;
  int j = 0;
  while (j + 1 < S.k && t >= S.t0[j + 1]) j++;
  const int *in = S.in[j];
  long long *out = S.out[j];
  const long n = S.n[j], lt = t - S.t0[j];
  const long long base = (long long)lt * kVecTile;
  const bool full = base + kVecTile <= n;
  int4 v[4];
  ld_tile4(in, n, base, full, v);
  long long s[4], e[4], tot;
#pragma unroll
  for (int k = 0; k < 4; k++) s[k] = (long long)v[k].x + v[k].y + v[k].z + v[k].w;
  block_excl_scan4(s, e, tot, lds);
  const long long pre = lb_tile_prefix(desc, t, tot, err, &s_pre, lt == 0);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const long long o0 = pre + e[k], o1 = o0 + v[k].x, o2 = o1 + v[k].y, o3 = o2 + v[k].z;
    const long long i = base + k * 1024 + 4 * threadIdx.x;
    if (full) {
      longlong2 *d = (longlong2 *)(out + i);
      d[0] = make_longlong2(o0, o1);
      d[1] = make_longlong2(o2, o3);
    } else {
      if (i < n) out[i] = o0;
      if (i + 1 < n) out[i + 1] = o1;
      if (i + 2 < n) out[i + 2] = o2;
      if (i + 3 < n) out[i + 3] = o3;
    }
  }
  if (t == S.t0[j + 1] - 1 && threadIdx.x == 0) out[n] = pre + tot;
}

__global__ void __launch_bounds__(256)
    k_lb_compact(const int *flag, long n, int *out, long long *count, unsigned long long *desc,
                 unsigned long long *next, long tiles, long long *err) {
  __shared__ long long lds[4][4];
  __shared__ long long s_pre;
  __shared__ long s_tile;
  lb_clear_next(next);
  if (threadIdx.x == 0) s_tile = (long)atomicAdd(desc + tiles, 1ULL);
  __syncthreads();
  const long t = s_tile;
  const long long base = (long long)t * kVecTile;
  const bool full = base + kVecTile <= n;
  int4 v[4];
  ld_tile4(flag, n, base, full, v);
  long long s[4], e[4], tot;
#pragma unroll
  for (int k = 0; k < 4; k++)
    s[k] = (v[k].x != 0) + (v[k].y != 0) + (v[k].z != 0) + (v[k].w != 0);
  block_excl_scan4(s, e, tot, lds);
  const long long pre = lb_tile_prefix(desc, t, tot, err, &s_pre, t == 0);
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const long long i = base + k * 1024 + 4 * threadIdx.x;
    long long p = pre + e[k];
    if (v[k].x != 0) out[p++] = (int)i;
    if (v[k].y != 0) out[p++] = (int)(i + 1);
    if (v[k].z != 0) out[p++] = (int)(i + 2);
    if (v[k].w != 0) out[p++] = (int)(i + 3);
  }
  if (t == tiles - 1 && threadIdx.x == 0) *count = pre + tot;
}

static long long *lb_err_word(Context &c) {
  if (!c.lb_err) {
    c.lb_err = (long long *)c.ws.get("lookback_err", 256);
    fill_i32(c, (int *)c.lb_err, 0, 2);
  }
  return c.lb_err;
}

// the cleared descriptor half for this launch and the half it clears for
// the next; lb_flip switches them once the launch is queued
static unsigned long long *lb_halves(Context &c, unsigned long long **next) {
  if (!c.lb_desc) {
    unsigned long long *d = (unsigned long long *)c.ws.get(
        "lookback_desc", sizeof(unsigned long long) * 2 * kLbSlots);
    hipLaunchKernelGGL(k_lb_clear_all, dim3(1), dim3(256), 0, c.stream, d);
    HIP_CHECK(hipGetLastError());
    c.lb_desc = d;
    c.lb_half = 0;
  }
  *next = c.lb_desc + (1 - c.lb_half) * kLbSlots;
  return c.lb_desc + c.lb_half * kLbSlots;
}

static void lb_flip(Context &c) { c.lb_half = 1 - c.lb_half; }

void lb_reset(Context &c) {
  if (c.lb_desc) {
    hipLaunchKernelGGL(k_lb_clear_all, dim3(1), dim3(256), 0, c.stream, c.lb_desc);
    HIP_CHECK(hipGetLastError());
  }
  if (c.lb_err) fill_i32(c, (int *)c.lb_err, 0, 2);
}

static bool lb_fits(const void *a, const void *b, long n) {
  return n > 0 && n <= kLbMaxN && !((uintptr_t)a & 15) && !((uintptr_t)b & 15);
}

// one look-back launch over the given arrays, when every one suits it and
// their tiles fit one descriptor half
static bool lb_scan_segs(Context &c, std::initializer_list<ScanJob> jobs) {
  LbSegs S{};
  long tiles = 0;
  if (jobs.size() == 0 || jobs.size() > (size_t)kLbSegs) return false;
  for (const ScanJob &j : jobs) {
    if (j.n < 0 || j.n > kLbMaxN || ((uintptr_t)j.in & 15) || ((uintptr_t)j.out & 15))
      return false;
    S.in[S.k] = j.in;
    S.out[S.k] = j.out;
    S.n[S.k] = j.n;
    S.t0[S.k] = tiles;
    tiles += j.n > 0 ? (j.n + kVecTile - 1) / kVecTile : 1;
    S.k++;
  }
  S.t0[S.k] = tiles;
  if (tiles + 1 > kLbSlots) return false;
  long long *err = lb_err_word(c);
  unsigned long long *next;
  unsigned long long *desc = lb_halves(c, &next);
  LaunchTimer t(c, "scan_lookback");
  hipLaunchKernelGGL(k_lb_scan, dim3((unsigned)tiles), dim3(256), 0, c.stream, S, desc, next,
                     tiles, err);
  HIP_CHECK(hipGetLastError());
  lb_flip(c);
  return true;
}

static bool lb_scan(Context &c, const int *in, long long *out, long n) {
  return n > 0 && lb_scan_segs(c, {ScanJob{in, out, n}});
}

// returns the device address of the count, or nullptr when the arrays do
// not suit the single-pass path
static const long long *lb_compact(Context &c, const int *flag, long n, int *out_idx,
                                   const std::string &t) {
  if (!lb_fits(flag, flag, n)) return nullptr;
  const long tiles = (n + kVecTile - 1) / kVecTile;
  long long *count = (long long *)c.ws.get(t + "_n", sizeof(long long));
  long long *err = lb_err_word(c);
  unsigned long long *next;
  unsigned long long *desc = lb_halves(c, &next);
  LaunchTimer tm(c, "compact_lookback");
  hipLaunchKernelGGL(k_lb_compact, dim3((unsigned)tiles), dim3(256), 0, c.stream, flag, n,
                     out_idx, count, desc, next, tiles, err);
  HIP_CHECK(hipGetLastError());
  lb_flip(c);
  return count;
}

template <typename TI, typename TO>
static void scan_impl(Context &c, const TI *in, TO *out, long n, int depth);

// the vector path: both pointers 16-byte aligned and n of at least a few
// tiles
static bool vscan(Context &c, const int *in, long long *out, long n) {
  if (n < 4L * kVecTile || ((uintptr_t)in & 15) || ((uintptr_t)out & 15)) return false;
  const long tiles = (n + kVecTile - 1) / kVecTile;
  long long *partials = (long long *)c.ws.get("vscan_partials", sizeof(long long) * (tiles + 1));
  {
    LaunchTimer t(c, "scan_reduce");
    hipLaunchKernelGGL(k_vscan_reduce, dim3((unsigned)tiles), dim3(256), 0, c.stream, in, n,
                       partials);
    HIP_CHECK(hipGetLastError());
  }
  // exclusive scan of the tile sums in place, the total at partials[tiles]
  scan_impl<long long, long long>(c, partials, partials, tiles, 1);
  {
    LaunchTimer t(c, "scan_downsweep");
    hipLaunchKernelGGL(k_vscan_down, dim3((unsigned)tiles), dim3(256), 0, c.stream, in, n, out,
                       (const long long *)partials);
    HIP_CHECK(hipGetLastError());
  }
  hipLaunchKernelGGL((k_copy_total<long long>), dim3(1), dim3(64), 0, c.stream, out + n,
                     (const long long *)(partials + tiles));
  HIP_CHECK(hipGetLastError());
  return true;
}

template <typename TI, typename TO>
static void scan_impl(Context &c, const TI *in, TO *out, long n, int depth) {
  // out has n+1 entries; out[n] = total.
  if (n <= kScanTile) {
    LaunchTimer t(c, "scan_tile");
    hipLaunchKernelGGL((k_tile_scan<TI, long long, TO>), dim3(1), dim3(kScanThreads), 0,
                       c.stream, in, n, out, (const long long *)nullptr, out + n);
    HIP_CHECK(hipGetLastError());
    return;
  }
  long tiles = (n + kScanTile - 1) / kScanTile;
  std::string tag = "scan_partials_" + std::to_string(depth);
  long long *partials = (long long *)c.ws.get(tag, sizeof(long long) * (size_t)(tiles + 1));
  {
    LaunchTimer t(c, "scan_reduce");
    hipLaunchKernelGGL((k_tile_reduce<TI, long long>), dim3((unsigned)tiles), dim3(kScanThreads),
                       0, c.stream, in, n, partials);
    HIP_CHECK(hipGetLastError());
  }
  if (tiles <= kScanPrefixTiles) {
    LaunchTimer t(c, "scan_downsweep");
    hipLaunchKernelGGL((k_tile_scan_prefix<TI, TO>), dim3((unsigned)tiles), dim3(kScanThreads), 0,
                       c.stream, in, n, out, (const long long *)partials, tiles);
    HIP_CHECK(hipGetLastError());
    return;
  }
  scan_impl<long long, long long>(c, partials, partials, tiles, depth + 1);
  {
    LaunchTimer t(c, "scan_downsweep");
    hipLaunchKernelGGL((k_tile_scan<TI, long long, TO>), dim3((unsigned)tiles),
                       dim3(kScanThreads), 0, c.stream, in, n, out, (const long long *)partials,
                       (TO *)nullptr);
    HIP_CHECK(hipGetLastError());
  }
  // out[n] = partials[tiles] (the grand total)
  hipLaunchKernelGGL((k_copy_total<TO>), dim3(1), dim3(64), 0, c.stream, out + n,
                     (const long long *)(partials + tiles));
  HIP_CHECK(hipGetLastError());
}

// The in-place recursive call (partials -> partials) is safe because every
// tile is read into LDS before any element of it is written.

__global__ void k_zero_total(void *p, int bytes) {
  if (threadIdx.x == 0) {
    if (bytes == 8) *(long long *)p = 0;
    else *(int *)p = 0;
  }
}

void scan_i32_to_i64(Context &c, const int *in, long long *out, long n) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(64), 0, c.stream, (void *)out, 8);
    HIP_CHECK(hipGetLastError());
    return;
  }
  if (lb_scan(c, in, out, n)) return;
  if (vscan(c, in, out, n)) return;
  scan_impl<int, long long>(c, in, out, n, 0);
}

void scan_i32_to_i64_multi(Context &c, std::initializer_list<ScanJob> jobs) {
  if (lb_scan_segs(c, jobs)) return;
  for (const ScanJob &j : jobs) scan_i32_to_i64(c, j.in, j.out, j.n);
}

void scan_i32_to_i32(Context &c, const int *in, int *out, long n) {
  if (n <= 0) {
    hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(64), 0, c.stream, (void *)out, 4);
    HIP_CHECK(hipGetLastError());
    return;
  }
  scan_impl<int, int>(c, in, out, n, 0);
}

// --------------------------------------------------------- compaction ----
__global__ void k_flags01(const int *flag, int *f01, long n) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f01[i] = flag[i] != 0;
}

__global__ void k_compact(const int *flag, const long long *pos, long n, int *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flag[i] != 0) out[pos[i]] = (int)i;
}

// Two-pass compaction of large flag arrays, with the vector scan's tile
// layout: nonzero counts per tile, their scan, then every tile rescans its
// flags and writes the indices of the nonzero ones.  The write pass takes
// the four groups' counts through ONE block scan, as 16-bit fields of one
// 64-bit word (a group holds 1024 flags, so no field carries into the next),
// places the tile's indices in LDS in index order and stores them as one
// contiguous run.
__global__ void __launch_bounds__(256) k_vcompact_count(const int *flag, long n, long long *cnt) {
  __shared__ long long lds[4];
  const long long base = (long long)blockIdx.x * kVecTile;
  long long s = 0;
  if (base + kVecTile <= n) {
    int4 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = ld_i4(flag, base, k);
#pragma unroll
    for (int k = 0; k < 4; k++) s += (v[k].x != 0) + (v[k].y != 0) + (v[k].z != 0) + (v[k].w != 0);
  } else {
    for (int k = 0; k < 4; k++)
      for (int j = 0; j < 4; j++) {
        const long long i = base + k * 1024 + 4 * threadIdx.x + j;
        if (i < n) s += flag[i] != 0;
      }
  }
  long long tot;
  block_excl_scan<long long>(s, lds, tot);
  if (threadIdx.x == 0) cnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(256)
    k_vcompact_write(const int *flag, long n, const long long *offsets, long tiles, int *out) {
  __shared__ long long lds[4];
  __shared__ int stage[kVecTile];
  const long T = xcd_tile(blockIdx.x, tiles);
  const long long base = (long long)T * kVecTile;
  const long long carry = offsets[T];
  const bool full = base + kVecTile <= n;
  int4 v[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (full) {
      v[k] = ld_i4(flag, base, k);
    } else {
      const long long i = base + k * 1024 + 4 * threadIdx.x;
      v[k].x = i < n ? flag[i] : 0;
      v[k].y = i + 1 < n ? flag[i + 1] : 0;
      v[k].z = i + 2 < n ? flag[i + 2] : 0;
      v[k].w = i + 3 < n ? flag[i + 3] : 0;
    }
  }
  long long packed = 0;
#pragma unroll
  for (int k = 0; k < 4; k++)
    packed |= (long long)((v[k].x != 0) + (v[k].y != 0) + (v[k].z != 0) + (v[k].w != 0))
              << (16 * k);
  long long tot;
  const long long pre = block_excl_scan<long long>(packed, lds, tot);
  int start = 0;  // LDS position of group k's first index
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int p = start + (int)((pre >> (16 * k)) & 0xffff);
    const int i = k * 1024 + 4 * threadIdx.x;  // offset in the tile
    if (v[k].x != 0) stage[p++] = i;
    if (v[k].y != 0) stage[p++] = i + 1;
    if (v[k].z != 0) stage[p++] = i + 2;
    if (v[k].w != 0) stage[p++] = i + 3;
    start += (int)((tot >> (16 * k)) & 0xffff);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < start; j += 256) out[carry + j] = (int)base + stage[j];
}

// the vector path of compaction; returns the device address of the count, or
// nullptr when the arrays do not suit it
static const long long *vcompact(Context &c, const int *flag, long n, int *out_idx,
                                 const std::string &t) {
  if (n < 4L * kVecTile || ((uintptr_t)flag & 15)) return nullptr;
  const long tiles = (n + kVecTile - 1) / kVecTile;
  long long *cnt = (long long *)c.ws.get(t + "_vc", sizeof(long long) * (tiles + 1));
  hipLaunchKernelGGL(k_vcompact_count, dim3((unsigned)tiles), dim3(256), 0, c.stream, flag, n,
                     cnt);
  HIP_CHECK(hipGetLastError());
  scan_impl<long long, long long>(c, cnt, cnt, tiles, 1);
  hipLaunchKernelGGL(k_vcompact_write, dim3((unsigned)tiles), dim3(256), 0, c.stream, flag, n,
                     (const long long *)cnt, tiles, out_idx);
  HIP_CHECK(hipGetLastError());
  return cnt + tiles;
}

long compact_flags(Context &c, const int *flag, long n, int *out_idx, const char *tag) {
  if (n <= 0) return 0;
  if (const long long *d = lb_compact(c, flag, n, out_idx, std::string("compact_") + tag))
    return (long)read_scalar(c, d);
  if (const long long *d = vcompact(c, flag, n, out_idx, std::string("compact_") + tag))
    return (long)read_scalar(c, d);
  std::string t = std::string("compact_") + tag;
  int *f01 = (int *)c.ws.get(t + "_f", sizeof(int) * (size_t)n);
  long long *pos = (long long *)c.ws.get(t + "_p", sizeof(long long) * (size_t)(n + 1));
  hipLaunchKernelGGL(k_flags01, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, flag, f01, n);
  HIP_CHECK(hipGetLastError());
  scan_i32_to_i64(c, f01, pos, n);
  hipLaunchKernelGGL(k_compact, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, flag, pos, n,
                     out_idx);
  HIP_CHECK(hipGetLastError());
  return (long)read_scalar(c, pos + n);
}

const long long *compact_flags_dev(Context &c, const int *flag, long n, int *out_idx,
                                   const char *tag) {
  std::string t = std::string("compact_") + tag;
  long long *pos = (long long *)c.ws.get(t + "_p", sizeof(long long) * (size_t)(n > 0 ? n + 1 : 1));
  if (n <= 0) {
    hipLaunchKernelGGL(k_zero_total, dim3(1), dim3(64), 0, c.stream, (void *)pos, 8);
    HIP_CHECK(hipGetLastError());
    return pos;
  }
  if (const long long *d = lb_compact(c, flag, n, out_idx, t)) return d;
  if (const long long *d = vcompact(c, flag, n, out_idx, t)) return d;
  int *f01 = (int *)c.ws.get(t + "_f", sizeof(int) * (size_t)n);
  hipLaunchKernelGGL(k_flags01, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, flag, f01, n);
  HIP_CHECK(hipGetLastError());
  scan_i32_to_i64(c, f01, pos, n);
  hipLaunchKernelGGL(k_compact, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, flag, pos, n,
                     out_idx);
  HIP_CHECK(hipGetLastError());
  return pos + n;
}

// --------------------------------------------------------- radix sort ----
constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;  // per lane
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kWaveKeys = 64 * kSortItems;

// Per-tile digit histogram, written digit-major: hist[d * tiles + tile].
__global__ void __launch_bounds__(kSortThreads)
    k_sort_hist(const uint64_t *keys, long n, int shift, int *hist, long tiles) {
  __shared__ int h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  // tiles in XCD order, so that the neighbouring tiles whose counts share a
  // cache line of every digit's row are counted under one L2
  const long T = xcd_tile(blockIdx.x, tiles);
  const long base = T * kSortTile;
  // every load issued before the first LDS update; a full tile of 16-byte
  // aligned keys is read as 16-byte pairs (the counts do not depend on which
  // thread holds which key)
  uint64_t kk[kSortItems];
  const bool vec = base + kSortTile <= n && ((uintptr_t)keys & 15) == 0;
  if (vec) {
    const ulonglong2 *k2 = (const ulonglong2 *)(keys + base);
#pragma unroll
    for (int k = 0; k < kSortItems / 2; k++) {
      const ulonglong2 v = k2[k * kSortThreads + threadIdx.x];
      kk[2 * k] = v.x;
      kk[2 * k + 1] = v.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kSortItems; k++) {
      const long i = base + (long)k * kSortThreads + threadIdx.x;
      kk[k] = i < n ? keys[i] : 0;
    }
  }
#pragma unroll
  for (int k = 0; k < kSortItems; k++) {
    const long i = base + (long)k * kSortThreads + threadIdx.x;
    if (vec || i < n) atomicAdd(&h[(int)((kk[k] >> shift) & 255)], 1);
  }
  __syncthreads();
  hist[(long)threadIdx.x * tiles + T] = h[threadIdx.x];
}

// Stable scatter: wave w ranks its 1024 keys in 16 rounds of 64 in order;
// position = tile digit offset + earlier waves' counts + rank within wave.
// The tile digit offsets come from off (the exclusive scan of the
// digit-major histograms) or, for sorts of at most kSortFuseTiles tiles,
// from the histograms themselves (hist): every block adds up its digit's
// row and scans the digits' totals, which saves the scan's launch.
constexpr long kSortFuseTiles = 64;
__global__ void __launch_bounds__(kSortThreads)
    k_sort_scatter(const uint64_t *keys, const int *vals, long n, int shift,
                   const long long *off, const int *hist, long tiles, uint64_t *keys_out,
                   int *vals_out) {
  __shared__ int wcnt[kSortThreads / 64][256];
  __shared__ long long lds_hist[kSortThreads / 64];
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  const long T = xcd_tile(blockIdx.x, tiles);
  // this tile's output start for digit threadIdx.x (256 threads, 256
  // digits), fetched now so that its latency hides behind the ranking
  long long dOff;
  if (hist) {
    const int *row = hist + (long)threadIdx.x * tiles;
    long long tot = 0, pre = 0;
    for (long t = 0; t < tiles; t++) {
      const int h = row[t];
      tot += h;
      pre += t < T ? h : 0;
    }
    long long all;
    dOff = block_excl_scan<long long>(tot, lds_hist, all) + pre;
  } else {
    dOff = off[(long)threadIdx.x * tiles + T];
  }
  for (int d = threadIdx.x; d < 256 * (kSortThreads / 64); d += kSortThreads)
    (&wcnt[0][0])[d] = 0;
  const long wbase = T * kSortTile + (long)wave * kWaveKeys;
  uint64_t kk[kSortItems];
  int vv[kSortItems];
  int pos[kSortItems];
  // all of the wave's loads in flight at once, before the ranking rounds
#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    const long i = wbase + (long)r * 64 + lane;
    kk[r] = i < n ? keys[i] : 0;
    vv[r] = i < n ? vals[i] : 0;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    long i = wbase + (long)r * 64 + lane;
    bool valid = i < n;
    uint64_t key = kk[r];
    int d = (int)((key >> shift) & 255);
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      unsigned long long m = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    int before = valid ? wcnt[wave][d] : 0;
    int rk = __popcll(peers & ((1ULL << lane) - 1ULL));
    pos[r] = valid ? before + rk : -1;
    // the lowest lane of each peer group advances the wave's digit counter
    bool leader = valid && rk == 0;
    // every lane has read wcnt for this round before any leader updates it
    __builtin_amdgcn_wave_barrier();
    if (leader) wcnt[wave][d] = before + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // exclusive over waves, per digit, and the tile's own digit starts
  __shared__ int dstart[256];
  __shared__ long long lds_scan[kSortThreads / 64];
  {
    int d = threadIdx.x;  // 256 threads == 256 digits
    int s = 0;
    for (int w = 0; w < kSortThreads / 64; w++) {
      int v = wcnt[w][d];
      wcnt[w][d] = s;
      s += v;
    }
    long long tot;
    dstart[d] = (int)block_excl_scan<long long>(s, lds_scan, tot);
  }
  __syncthreads();
  // the tile sorted by digit in LDS, stably ...
  __shared__ uint64_t sk[kSortTile];
  __shared__ int sv[kSortTile];
#pragma unroll
  for (int r = 0; r < kSortItems; r++) {
    if (pos[r] < 0) continue;
    const int d = (int)((kk[r] >> shift) & 255);
    const int lp = dstart[d] + wcnt[wave][d] + pos[r];
    sk[lp] = kk[r];
    sv[lp] = vv[r];
  }
  __syncthreads();
  // the per-wave counters are dead now: their LDS holds each digit's
  // output start minus its start inside the tile
  static_assert(sizeof(wcnt) >= 256 * sizeof(long long), "digit bases do not fit");
  long long *gbase = (long long *)&wcnt[0][0];
  gbase[threadIdx.x] = dOff - dstart[threadIdx.x];
  __syncthreads();
  // ... then written out in that order: consecutive threads store
  // consecutive addresses inside each digit's run
  const long tbase = T * kSortTile;
  const int tn = (int)min((long)kSortTile, n - tbase);
  for (int i = threadIdx.x; i < tn; i += kSortThreads) {
    const uint64_t key = sk[i];
    const int d = (int)((key >> shift) & 255);
    const long long dst = gbase[d] + i;
    keys_out[dst] = key;
    vals_out[dst] = sv[i];
  }
}

void radix_sort_u64(Context &c, const uint64_t *keys_in, const int *vals_in, uint64_t *keys_out,
                    int *vals_out, long n, int bits) {
  if (n <= 0) return;
  int passes = (bits + 7) / 8;
  if (passes == 0) {
    HIP_CHECK(hipMemcpyAsync(keys_out, keys_in, sizeof(uint64_t) * (size_t)n,
                             hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(vals_out, vals_in, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice,
                             c.stream));
    return;
  }
  long tiles = (n + kSortTile - 1) / kSortTile;
  uint64_t *kb = (uint64_t *)c.ws.get("sort_kb", sizeof(uint64_t) * (size_t)n);
  int *vb = (int *)c.ws.get("sort_vb", sizeof(int) * (size_t)n);
  int *hist = (int *)c.ws.get("sort_hist", sizeof(int) * (size_t)(256 * tiles));
  long long *off = (long long *)c.ws.get("sort_off", sizeof(long long) * (size_t)(256 * tiles + 1));
  // ping-pong so that the last pass lands in *_out
  const uint64_t *src_k = keys_in;
  const int *src_v = vals_in;
  for (int p = 0; p < passes; p++) {
    bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t *dk = to_out ? keys_out : kb;
    int *dv = to_out ? vals_out : vb;
    {
      LaunchTimer t(c, "sort_hist");
      hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)tiles), dim3(kSortThreads), 0, c.stream,
                         src_k, n, 8 * p, hist, tiles);
      HIP_CHECK(hipGetLastError());
    }
    const bool fuse = tiles <= kSortFuseTiles;
    if (!fuse) scan_i32_to_i64(c, hist, off, 256 * tiles);
    {
      LaunchTimer t(c, "sort_scatter");
      hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)tiles), dim3(kSortThreads), 0, c.stream,
                         src_k, src_v, n, 8 * p, (const long long *)off,
                         fuse ? (const int *)hist : nullptr, tiles, dk, dv);
      HIP_CHECK(hipGetLastError());
    }
    src_k = dk;
    src_v = dv;
  }
}

// ---------------------------------------------- sort without allocation ----
// gcmf_radix_sort_u64_noalloc keeps its scratch -- the ping-pong key and value
// buffers, the per-tile digit histograms, their scanned offsets and the
// scan's tile partials -- in one static device array per GPU, so that the
// call allocates nothing.  Every region is fully written before it is read
// in every pass, so no clearing is needed; calls that overlap on one GPU
// race on it.
constexpr long kNoAllocSortMax = 1L << 21;                    // keys per call
constexpr long kNoAllocTiles = kNoAllocSortMax / kSortTile;  // 512
constexpr long kNoAllocHist = 256 * kNoAllocTiles;
constexpr long kNoAllocParts = kNoAllocHist / kScanTile + 1;
constexpr size_t noalloc_round(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNoAllocKeys = 0;
constexpr size_t kNoAllocVals = kNoAllocKeys + noalloc_round(8 * kNoAllocSortMax);
constexpr size_t kNoAllocHistAt = kNoAllocVals + noalloc_round(4 * kNoAllocSortMax);
constexpr size_t kNoAllocOffAt = kNoAllocHistAt + noalloc_round(4 * kNoAllocHist);
constexpr size_t kNoAllocPartAt = kNoAllocOffAt + noalloc_round(8 * (kNoAllocHist + 1));
constexpr size_t kNoAllocBytes = kNoAllocPartAt + noalloc_round(8 * kNoAllocParts);
static_assert(kNoAllocHist / kScanTile + 1 <= kScanPrefixTiles, "histogram scan too long");

static_assert(kNoAllocBytes == (size_t)GCMF_NOALLOC_SCRATCH_BYTES, "header disagrees");
static_assert(kNoAllocSortMax == GCMF_NOALLOC_SORT_MAX, "header disagrees");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

long noalloc_sort_max() { return kNoAllocSortMax; }
size_t noalloc_scratch_bytes() { return kNoAllocBytes; }

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is no device address).
void *noalloc_scratch_address() {
  static std::mutex mu;
  static void *addr[64] = {};
  int dev = 0;
  HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) throw Error{GCMF_ERR_ARG, "device index beyond 63"};
  std::lock_guard<std::mutex> lk(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    addr[dev] = p;
  }
  return addr[dev];
}

void radix_sort_u64_noalloc(Context &c, const uint64_t *keys_in, const int *vals_in,
                            uint64_t *keys_out, int *vals_out, long n, int bits) {
  if (n > kNoAllocSortMax)
    throw Error{GCMF_ERR_ARG, "gcmf_radix_sort_u64_noalloc takes at most 2^21 keys"};
  if (!c.scratch) throw Error{GCMF_ERR_ARG, "context has no scratch address"};
  if (n <= 0) return;
  const int passes = (bits + 7) / 8;
  if (passes == 0) {
    HIP_CHECK(hipMemcpyAsync(keys_out, keys_in, sizeof(uint64_t) * (size_t)n,
                             hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipMemcpyAsync(vals_out, vals_in, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice,
                             c.stream));
    return;
  }
  unsigned char *s = (unsigned char *)c.scratch;
  uint64_t *kb = (uint64_t *)(s + kNoAllocKeys);
  int *vb = (int *)(s + kNoAllocVals);
  int *hist = (int *)(s + kNoAllocHistAt);
  long long *off = (long long *)(s + kNoAllocOffAt);
  long long *parts = (long long *)(s + kNoAllocPartAt);
  const long tiles = (n + kSortTile - 1) / kSortTile;
  const long nh = 256 * tiles, st = (nh + kScanTile - 1) / kScanTile;
  const uint64_t *src_k = keys_in;
  const int *src_v = vals_in;
  for (int p = 0; p < passes; p++) {
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint64_t *dk = to_out ? keys_out : kb;
    int *dv = to_out ? vals_out : vb;
    hipLaunchKernelGGL(k_sort_hist, dim3((unsigned)tiles), dim3(kSortThreads), 0, c.stream,
                       src_k, n, 8 * p, hist, tiles);
    // exclusive scan of the histograms into off (the total at off[nh])
    if (st == 1) {
      hipLaunchKernelGGL((k_tile_scan<int, long long, long long>), dim3(1), dim3(kScanThreads),
                         0, c.stream, (const int *)hist, nh, off, (const long long *)nullptr,
                         off + nh);
    } else {
      hipLaunchKernelGGL((k_tile_reduce<int, long long>), dim3((unsigned)st), dim3(kScanThreads),
                         0, c.stream, (const int *)hist, nh, parts);
      hipLaunchKernelGGL((k_tile_scan_prefix<int, long long>), dim3((unsigned)st),
                         dim3(kScanThreads), 0, c.stream, (const int *)hist, nh, off,
                         (const long long *)parts, st);
    }
    hipLaunchKernelGGL(k_sort_scatter, dim3((unsigned)tiles), dim3(kSortThreads), 0, c.stream,
                       src_k, src_v, n, 8 * p, (const long long *)off, (const int *)nullptr,
                       tiles, dk, dv);
    HIP_CHECK(hipGetLastError());
    src_k = dk;
    src_v = dv;
  }
}

}  // namespace gcmf
