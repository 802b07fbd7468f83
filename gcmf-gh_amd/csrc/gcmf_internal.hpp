// gcmf_internal.hpp -- host-side internals shared by the HIP translation
// units of libgcmf_hip.so: the context (stream, workspace arena, kernel
// timers) and the device primitives' host drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <initializer_list>
#include <map>
#include <string>
#include <vector>

#define GCMF_OK 0
#define GCMF_ERR_HIP -1
#define GCMF_ERR_ALLOC -2
#define GCMF_ERR_ARG -3
#define GCMF_ERR_NODEV -4
#define GCMF_ERR_OVERFLOW -5

namespace gcmf {

// Raised by HIP_CHECK inside library code; converted to a status code at the
// C-ABI boundary.
struct Error {
  int code;
  std::string what;
};

#define HIP_CHECK(expr)                                                        \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess)                                                      \
      throw ::gcmf::Error{_e == hipErrorOutOfMemory ? GCMF_ERR_ALLOC           \
                                                    : GCMF_ERR_HIP,            \
                          std::string(#expr) + ": " + hipGetErrorString(_e)};  \
  } while (0)

// Grow-only device workspace: named slots are reused across calls so the
// steady-state path performs no hipMalloc.
struct Workspace {
  struct Slot {
    void *ptr = nullptr;
    size_t bytes = 0;
  };
  std::map<std::string, Slot> slots;
  bool no_alloc = false;  // set while a launch sequence is captured: get() may not allocate
  void *get(const std::string &name, size_t bytes);
  void release_all();
};

struct KernelTimer {
  double total_ms = 0.0;
  long launches = 0;
};

// Launch shapes and thresholds of the batched clip that the tuner
// (gcmf_tune_clip) chooses by measurement; the defaults suit the join
// workload's pairs.
struct Tuning {
  long cullMin = 512;       // rings longer than this are walked through run boxes
  long long bigNodes = 8192;  // pairs with this many nodes are labelled by a block
  long long gridNodes = 1LL << 20;  // ... and with this many by the whole GPU
  int gridClean = 16384;    // there, components longer than this are cleaned by the grid
  long long stripMin = 1LL << 17;  // joins of this many polygons sweep the MBRs in y strips
  int cmfGroup = 0;         // lanes per pair in the CMF edge filter (16, 64; 0: by polygon size)
  int labelSteal = 0;       // 1: labelling waves are persistent and take pairs from a counter
  int blockCap = 128;       // blocks of the block labelling (each takes pairs from a counter)
  int countStage = 0;       // 1: each count lane walks the whole Q ring (a wave's copy in
                            // LDS); 0: a wave's lanes of one pair walk it together by runs
  int countBlock = 128;     // threads per block of the count kernel (64, 128, 256)
  int labelLds = 20480;     // LDS bytes per labelling wave for a pair's arrays (0: none)
  int ldsMaxQ = 1024;       // pairs with Q rings above cullMin and up to this size are
                            // counted with Q binned in LDS (at most 1024; 0: none)
};

struct Context {
  int device = 0;
  Tuning tune;
  hipStream_t stream = nullptr;
  Workspace ws;
  bool timing = false;
  std::string timing_only;  // when set, only launches of this name are timed
  std::map<std::string, KernelTimer> timers;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> event_pool;
  std::string last_error;
  int cu_count = 256;
  long long *hbuf = nullptr;  // pinned host staging for small read-backs (kHostSlots)
  long long *hbuf_dev = nullptr;  // its device address (mapped, coherent)
  // device address of the allocation-free sort's static scratch on this
  // context's device (noalloc_scratch_address, looked up by gcmf_init)
  void *scratch = nullptr;
  // a second stream (default priority) for work that overlaps the main one,
  // with the events that fork it from and join it to `stream`
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // device word set when a single-pass scan's look-back wait ran past its
  // bound (never in a correct run); read back and cleared by read_scalars
  // and check_lookback
  long long *lb_err = nullptr;
  // the two halves of the single-pass scans' tile descriptors: each launch
  // uses one and clears the other for the next launch
  unsigned long long *lb_desc = nullptr;
  int lb_half = 0;
  // Capture of a launch sequence into a HIP graph: replay 0 runs eagerly;
  // 1 runs eagerly and records every value read back (rec, rec_bytes) and a
  // device copy of every buffer read back whole; 2 captures, answering each
  // read-back from the record and adding a kernel that sets *graph_flag when
  // the value on the device differs from the recorded one.
  int replay = 0;
  std::vector<long long> rec;
  size_t rec_pos = 0;
  std::vector<std::vector<char>> rec_bytes;
  size_t rec_bytes_pos = 0;
  int *graph_flag = nullptr;
  // mean vertex count per polygon of the last join's layers (-1: unknown)
  double avgVerts = -1.0;
  // the persistent labelling launch's grid (waves the GPU holds at once) for
  // the LDS arena size it was measured with
  long label_steal_blocks = 0;
  int label_steal_lds = -1;

  hipEvent_t take_event();
  void flush_timers();  // synchronises on pending events and accumulates
};

// RAII timer around one launch (records events on ctx.stream when enabled).
struct LaunchTimer {
  Context &c;
  const char *name;
  hipEvent_t a = nullptr, b = nullptr;
  LaunchTimer(Context &ctx, const char *n);
  ~LaunchTimer();
};

// ---------------------------------------------------------------- prims ----
// Exclusive scan of n int32 values into out[0..n] (out[n] = total).
void scan_i32_to_i64(Context &c, const int *in, long long *out, long n);
// the same over up to four arrays, in one launch where they suit the
// single-pass scan
struct ScanJob {
  const int *in;
  long long *out;
  long n;
};
void scan_i32_to_i64_multi(Context &c, std::initializer_list<ScanJob> jobs);
void scan_i32_to_i32(Context &c, const int *in, int *out, long n);
// out[k] = i for the k-th i with flag[i] != 0; returns the count (synchronises).
long compact_flags(Context &c, const int *flag, long n, int *out_idx, const char *tag);
// the same without the read-back: returns the device address of the count,
// valid until the next call with the same tag
const long long *compact_flags_dev(Context &c, const int *flag, long n, int *out_idx,
                                   const char *tag);
// Stable LSD radix sort of keys (u64) carrying int32 values; bits = number of
// significant low bits of every key.  Results land in keys_out / vals_out.
void radix_sort_u64(Context &c, const uint64_t *keys_in, const int *vals_in, uint64_t *keys_out,
                    int *vals_out, long n, int bits);
// The same sort with its scratch in a static device array (no allocation;
// n <= noalloc_sort_max(); stream-ordered, not synchronised).
void radix_sort_u64_noalloc(Context &c, const uint64_t *keys_in, const int *vals_in,
                            uint64_t *keys_out, int *vals_out, long n, int bits);
long noalloc_sort_max();
size_t noalloc_scratch_bytes();
// the scratch array's device address on the current device
void *noalloc_scratch_address();
// fills
void fill_i32(Context &c, int *p, int v, long n);
// raises an Error when a single-pass scan's look-back wait ran past its
// bound since the last check (synchronises)
void check_lookback(Context &c);
// copies `bytes` of device memory back to the host (synchronises); under
// capture, answered from the recorded run and checked on the device
void host_read(Context &c, void *dst, const void *src, size_t bytes);
// clears the single-pass scans' descriptors and error word (the first step
// of a captured sequence, so that every replay starts as the capture did)
void lb_reset(Context &c);
// copies one value back to the host (synchronises)
template <typename T>
T read_scalar(Context &c, const T *dptr) {
  T v;
  host_read(c, &v, dptr, sizeof(T));
  return v;
}

// copies up to kHostSlots int64 values back to the host in one transfer
// (synchronises); null sources read as 0
constexpr int kHostSlots = 24;
void read_scalars(Context &c, std::initializer_list<const long long *> src, long long *out);

inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace gcmf
