// join.hip -- the GCMF filter-and-refine spatial join on gfx950.
//
// Stages (reference semantics in brackets):
//   1. MBR filter as sort-and-sweep [Join.h:5-153, GPU_MBR.h:339-539]: the
//      2(B+O) x-endpoint keys (mod 1e16, the reference sorts 16 decimal
//      digits) are radix sorted with the endpoint index as a stable
//      tie-break; each polygon then scans the contiguous rank range between
//      its own two endpoints in a structure-of-arrays record stream
//      (endpoint id, partner y-interval), one wave64 per polygon, ballots
//      for counting and ordered output.
//   2. CMBR + PiP candidate flags [Join.h:397-432, GPU_Join.h:461-508].
//   3. PiP "within" refine, one wave per candidate, crossing parity from
//      ballot popcounts [Join.h:437-467, GPU_Join.h:839-943].
//   4. CMF edge filter count + apply, one wave per surviving pair, ordered
//      edge lists from ballot ranks [Join.h:157-287, GPU_Join.h:511-735].
//   5. Segment-intersection refine, one wave per CMF survivor, wave-uniform
//      early exit [Join.h:356-393, GPU_Join.h:326-456].
#include <algorithm>
#include <vector>

#include "gcmf_device.hpp"
#include "gcmf_internal.hpp"
#include "gcmf_join.hpp"

namespace gcmf {

constexpr uint64_t kKeyMod = 10000000000000000ULL;  // 16 decimal digits

// ------------------------------------------------------------ stage 1 ----
// mm: x key min, x key max, y key min, y key max (over both corners) and,
// when the vertex counts are given, their sum; a grid of at most
// kEndpointBlocks blocks striding over the polygons
constexpr unsigned kEndpointBlocks = 1024;
__global__ void k_endpoints(long B, long O, const uint64_t *bMBR, const uint64_t *oMBR,
                            const int *bVNum, const int *oVNum, uint64_t *keys, int *idx,
                            uint64_t *Y, unsigned long long *mm) {
  uint64_t lo = ~0ULL, hi = 0, ylo = ~0ULL, yhi = 0, vs = 0;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < B + O;
       e += (long)gridDim.x * blockDim.x) {
    if (bVNum) vs += (uint64_t)(e < B ? bVNum[e] : oVNum[e - B]);
    const uint64_t *m = e < B ? bMBR + 4 * e : oMBR + 4 * (e - B);
    const uint64_t x1 = m[0] % kKeyMod, x2 = m[2] % kKeyMod, y1 = m[1], y2 = m[3];
    keys[2 * e] = x1;
    keys[2 * e + 1] = x2;
    idx[2 * e] = (int)(2 * e);
    idx[2 * e + 1] = (int)(2 * e + 1);
    Y[2 * e] = y1;
    Y[2 * e + 1] = y2;
    lo = min(lo, min(x1, x2));
    hi = max(hi, max(x1, x2));
    ylo = min(ylo, min(y1, y2));
    yhi = max(yhi, max(y1, y2));
  }
  // wave min/max, block min/max, then one atomic per block and value
  // (atomics on one cache line serialise across the chip, so the grid is
  // capped)
  __shared__ uint64_t slo[4], shi[4], sylo[4], syhi[4], svs[4];
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t l2 = __shfl_xor(lo, off, 64), h2 = __shfl_xor(hi, off, 64);
    uint64_t yl2 = __shfl_xor(ylo, off, 64), yh2 = __shfl_xor(yhi, off, 64);
    lo = l2 < lo ? l2 : lo;
    hi = h2 > hi ? h2 : hi;
    ylo = yl2 < ylo ? yl2 : ylo;
    yhi = yh2 > yhi ? yh2 : yhi;
    vs += __shfl_xor(vs, off, 64);
  }
  if (__lane_id() == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
    sylo[threadIdx.x >> 6] = ylo;
    syhi[threadIdx.x >> 6] = yhi;
    svs[threadIdx.x >> 6] = vs;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      lo = slo[w] < lo ? slo[w] : lo;
      hi = shi[w] > hi ? shi[w] : hi;
      ylo = sylo[w] < ylo ? sylo[w] : ylo;
      yhi = syhi[w] > yhi ? syhi[w] : yhi;
      vs += svs[w];
    }
    atomicMin(mm, (unsigned long long)lo);
    atomicMax(mm + 1, (unsigned long long)hi);
    atomicMin(mm + 2, (unsigned long long)ylo);
    atomicMax(mm + 3, (unsigned long long)yhi);
    if (vs) atomicAdd(mm + 4, (unsigned long long)vs);
  }
}

__global__ void k_init_minmax(unsigned long long *mm) {
  if (threadIdx.x < 5) mm[threadIdx.x] = threadIdx.x == 4 || (threadIdx.x & 1) ? 0ULL : ~0ULL;
}

__global__ void k_sub_key(uint64_t *keys, long n, const unsigned long long *mn) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] -= *mn;
}

// rank[perm[i]] = i; sweep records in sorted order
__global__ void k_sweep_prep(const int *perm, long n, const uint64_t *Y, int *rank, int *rec_id,
                             uint64_t *rec_ylo, uint64_t *rec_yhi) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int e = perm[i];
  rank[e] = (int)i;
  rec_id[i] = e;
  int m = e & ~1;
  rec_ylo[i] = Y[m];
  rec_yhi[i] = Y[m + 1];
}

// one wave per polygon element (or per polygon of list[0, nList)); pairs ==
// nullptr -> count only
__global__ void __launch_bounds__(256)
    k_sweep(long B, long O, const int *rank, const int *rec_id, const uint64_t *rec_ylo,
            const uint64_t *rec_yhi, const uint64_t *Y, int *counts, const long long *offs,
            int *pairs, const int *list, long nList) {
  long e = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (list) {
    if (e >= nList) return;
    e = list[e];
  } else if (e >= B + O) {
    return;
  }
  const int lane = __lane_id();
  const int lo = rank[2 * e], hi = rank[2 * e + 1];
  const uint64_t ey1 = Y[2 * e], ey2 = Y[2 * e + 1];
  const bool overlay = e >= B;
  const int lim = (int)(2 * B);
  long long pos = pairs ? offs[e] : 0;
  int cnt = 0;
  for (int i0 = lo + 1; i0 < hi; i0 += 64) {
    int i = i0 + lane;
    bool ok = false;
    int m = 0;
    if (i < hi) {
      m = rec_id[i];
      bool other = overlay ? (m < lim) : (m >= lim);
      ok = other && !(m & 1) && !(ey1 > rec_yhi[i]) && !(rec_ylo[i] > ey2);
    }
    unsigned long long msk = __ballot(ok);
    if (pairs && ok) {
      long long p = pos + __popcll(msk & ((1ULL << lane) - 1ULL));
      if (!overlay) {
        pairs[2 * p] = (int)e;
        pairs[2 * p + 1] = m / 2 - (int)B;
      } else {
        pairs[2 * p] = m / 2;
        pairs[2 * p + 1] = (int)(e - B);
      }
    }
    pos += __popcll(msk);
    cnt += __popcll(msk);
  }
  if (!pairs && lane == 0) counts[e] = cnt;
}

// ------------------------------------------------ stage 1, y strips ----
// The sweep above walks a polygon's whole rank range, so its work grows with
// the sum of the ranges (about N^1.5 for N evenly spread polygons), not with
// the candidates.  For layers of Tuning::stripMin polygons or more the y
// range of the MBR keys is cut into S strips of height 2^shift (the power of
// two at or above twice the median MBR height of 1024 sampled polygons), and
// every left-endpoint record is listed, per layer and in rank order, in each
// strip its y interval touches.  A polygon then visits, in each strip its
// own y interval touches, the other layer's records whose rank lies
// strictly between its endpoints' ranks (a binary search for the first,
// then a walk), applies the sweep's y test, and takes a record only in the
// strip of the larger of the two intervals' lower ends, where both
// intervals are listed, so each candidate is found once.  A polygon's
// candidates are merged over its strips in rank order, the sweep's order.
// Polygons touching more than kStripSpan strips, or with more than
// kStripScan records to visit, are counted and written by the sweep itself
// (one wave each), so the candidate list is the sweep's, entry for entry.
constexpr int kStripSpan = 4;
constexpr int kStripScan = 256;
constexpr int kStripSample = 1024;
constexpr long kStripMax = 1L << 22;

struct Strips {
  uint64_t y0;
  int shift, S;
};
__device__ __forceinline__ int strip_of(const Strips &g, uint64_t y) {
  const uint64_t s = (y - g.y0) >> g.shift;
  return s >= (uint64_t)g.S ? g.S - 1 : (int)s;
}

// |ymax - ymin| of polygons j N / ns, j < ns
__global__ void k_height_sample(long N, const uint64_t *Y, int ns, uint64_t *out) {
  const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (j >= ns) return;
  const long e = (long)((long long)j * N / ns);
  const uint64_t a = Y[2 * e], b = Y[2 * e + 1];
  out[j] = a > b ? a - b : b - a;
}

// the number of strips each record (left endpoint at rank i) is listed in
__global__ void k_strip_rec_count(long n, const int *rec_id, const uint64_t *rec_ylo,
                                  const uint64_t *rec_yhi, Strips g, int *cnt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = 0;
  if (!(rec_id[i] & 1)) {
    const uint64_t a = rec_ylo[i], b = rec_yhi[i];
    c = strip_of(g, a < b ? b : a) - strip_of(g, a < b ? a : b) + 1;
  }
  cnt[i] = c;
}

// (layer * S + strip, rank) for every listing, in rank order; layer 1 is
// the overlay (endpoint ids from 2B)
__global__ void k_strip_rec_emit(long n, long B, const int *rec_id, const uint64_t *rec_ylo,
                                 const uint64_t *rec_yhi, Strips g, const long long *off,
                                 uint64_t *keys, int *vals) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (rec_id[i] & 1)) return;
  const uint64_t a = rec_ylo[i], b = rec_yhi[i];
  const int s0 = strip_of(g, a < b ? a : b), s1 = strip_of(g, a < b ? b : a);
  const uint64_t side = rec_id[i] < 2 * B ? 0 : 1;
  long long o = off[i];
  for (int s = s0; s <= s1; s++, o++) {
    keys[o] = side * (uint64_t)g.S + (uint64_t)s;
    vals[o] = (int)i;
  }
}

// sOff[k] = the first listing of key k in the sorted keys, k <= 2S (each
// entry written once: the keys in (key[i - 1], key[i]] start at i)
__global__ void k_strip_bounds(long R, const uint64_t *skeys, long nk, long long *sOff) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i > R) return;
  const long cur = i < R ? (long)skeys[i] : nk;
  const long prev = i > 0 ? (long)skeys[i - 1] : -1;
  for (long k = prev + 1; k <= cur; k++) sOff[k] = i;
}

// what a search needs of the left-endpoint record at rank r: its endpoint id
// (2 e), the rank of polygon e's right endpoint and the y interval, in one
// record so that a random visit touches one cache line
struct StripRec {
  int id, hi;
  uint64_t y1, y2;
};
__global__ void k_strip_recs(long n, const int *rec_id, const uint64_t *rec_ylo,
                             const uint64_t *rec_yhi, const int *rank, StripRec *qrec) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int m = rec_id[i];
  if (!(m & 1)) qrec[i] = StripRec{m, rank[m + 1], rec_ylo[i], rec_yhi[i]};
}

// first index of [l, r) whose rank is above v
__device__ __forceinline__ long first_above(const int *a, long l, long r, int v) {
  while (l < r) {
    const long m = (l + r) >> 1;
    if (a[m] > v) r = m;
    else l = m + 1;
  }
  return l;
}

// One polygon's strip ranges: returns false when the sweep takes it (more
// than kStripSpan strips or kStripScan records).
struct StripQuery {
  int sa, ns;
  long p0[kStripSpan], p1[kStripSpan];
};
__device__ __forceinline__ bool strip_query(const Strips &g, const long long *sOff,
                                            const int *list, int side, int lo, int hi,
                                            uint64_t ya, uint64_t yb, StripQuery &q) {
  q.sa = strip_of(g, ya);
  q.ns = strip_of(g, yb) - q.sa + 1;
  if (q.ns > kStripSpan) return false;
  long left = kStripScan;
#pragma unroll
  for (int t = 0; t < kStripSpan; t++) {
    q.p0[t] = q.p1[t] = 0;
    if (t < q.ns) {
      const long k = (long)side * g.S + q.sa + t;
      const long l = sOff[k], r = sOff[k + 1];
      long p = first_above(list, l, r, lo);
      q.p0[t] = p;
      while (p < r && list[p] < hi) {
        if (left-- == 0) return false;
        p++;
      }
      q.p1[t] = p;
    }
  }
  return true;
}

// the sweep's y test and the strip that owns the pair
__device__ __forceinline__ bool strip_hit(const Strips &g, int s, uint64_t ey1, uint64_t ey2,
                                          uint64_t ya, uint64_t ry1, uint64_t ry2) {
  if (ey1 > ry2 || ry1 > ey2) return false;
  const uint64_t rlo = ry1 < ry2 ? ry1 : ry2;
  return strip_of(g, ya > rlo ? ya : rlo) == s;
}

// Count (pairs == nullptr) or write.  The count pass runs one thread per
// listing (R of them, skeys / list): a polygon is taken at the listing of
// its own record in the first strip it touches, so every polygon once, in
// (layer, strip, rank) order -- neighbouring threads then search the same
// strip's list near each other.  It keeps the first kStripInline
// candidates' endpoint ids in inl (rank order), so the write pass (one
// thread per polygon) copies them and searches again only for polygons with
// more; a polygon the sweep takes is marked inl[e, 0] = -1 and appended to
// heavy (heavy[0] counts them, the list follows).
constexpr int kStripInline = 8;
__global__ void __launch_bounds__(256)
    k_strip_query(long B, long O, Strips g, const long long *sOff, const int *list,
                  const uint64_t *skeys, long R, const int *rank, const StripRec *qrec,
                  const uint64_t *Y, int *counts, const long long *offs, int *pairs,
                  int *heavy, int *inl) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long e = t;
  int lo, hi;
  uint64_t ey1, ey2;
  if (pairs) {
    if (e >= B + O) return;
    lo = rank[2 * e];
    hi = rank[2 * e + 1];
    ey1 = Y[2 * e];
    ey2 = Y[2 * e + 1];
  } else {
    if (t >= R) return;
    lo = list[t];
    const StripRec me = qrec[lo];
    e = me.id / 2;
    hi = me.hi;
    ey1 = me.y1;
    ey2 = me.y2;
    const long key = (long)skeys[t];
    if ((key >= g.S ? key - g.S : key) != strip_of(g, ey1 < ey2 ? ey1 : ey2)) return;
  }
  const bool overlay = e >= B;
  int *mine = inl + (long long)kStripInline * e;
  long long pos = 0;
  int cnt = 0;
  if (pairs) {
    cnt = counts[e];
    if (cnt == 0 || mine[0] < 0) return;
    pos = offs[e];
    if (cnt <= kStripInline) {
      for (int j = 0; j < cnt; j++) {
        const int m = mine[j];
        pairs[2 * (pos + j)] = overlay ? m / 2 : (int)e;
        pairs[2 * (pos + j) + 1] = overlay ? (int)(e - B) : m / 2 - (int)B;
      }
      return;
    }
  }
  const uint64_t ya = ey1 < ey2 ? ey1 : ey2, yb = ey1 < ey2 ? ey2 : ey1;
  StripQuery q;
  if (!strip_query(g, sOff, list, overlay ? 0 : 1, lo, hi, ya, yb, q)) {
    counts[e] = 0;
    mine[0] = -1;
    heavy[1 + atomicAdd(heavy, 1)] = (int)e;
    return;
  }
  // merge the strips' lists by rank
  cnt = 0;
  for (;;) {
    int best = -1, br = 0x7fffffff;
#pragma unroll
    for (int u = 0; u < kStripSpan; u++)
      if (q.p0[u] < q.p1[u]) {
        const int r = list[q.p0[u]];
        if (r < br) {
          br = r;
          best = u;
        }
      }
    if (best < 0) break;
#pragma unroll
    for (int u = 0; u < kStripSpan; u++)
      if (u == best) q.p0[u]++;
    const StripRec x = qrec[br];
    if (!strip_hit(g, q.sa + best, ey1, ey2, ya, x.y1, x.y2)) continue;
    const int m = x.id;
    if (pairs) {
      pairs[2 * pos] = overlay ? m / 2 : (int)e;
      pairs[2 * pos + 1] = overlay ? (int)(e - B) : m / 2 - (int)B;
      pos++;
    } else {
      if (cnt < kStripInline) mine[cnt] = m;
      cnt++;
    }
  }
  if (!pairs) counts[e] = cnt;
}

// The strip form of the sweep: fills counts[N] and offs[N + 1] and returns
// the candidate count, with the pairs written to *pairs_out; returns -1
// (nothing written) when the records would be listed more than four times
// each on average, where the plain sweep is used instead.
static long strip_sweep(Context &c, long B, long O, const int *rank, const int *rec_id,
                        const uint64_t *rec_ylo, const uint64_t *rec_yhi, const uint64_t *Y,
                        uint64_t y0, uint64_t y1, int *counts, long long *offs,
                        int **pairs_out) {
  const long N = B + O, n = 2 * N;
  // strip height: the power of two at or above twice the median sampled MBR
  // height, at most kStripMax strips
  const int ns = (int)std::min<long>(kStripSample, N);
  uint64_t *samp = (uint64_t *)c.ws.get("strip_sample", 8 * (size_t)ns);
  hipLaunchKernelGGL(k_height_sample, dim3(grid_for(ns, 256)), dim3(256), 0, c.stream, N, Y, ns,
                     samp);
  HIP_CHECK(hipGetLastError());
  std::vector<uint64_t> hs((size_t)ns);
  host_read(c, hs.data(), samp, 8 * (size_t)ns);
  std::nth_element(hs.begin(), hs.begin() + ns / 2, hs.end());
  const uint64_t range = y1 - y0, h2 = 2 * hs[(size_t)ns / 2];
  int shift = 0;
  while (shift < 63 && ((1ULL << shift) < h2 || (range >> shift) + 1 > (uint64_t)kStripMax))
    shift++;
  Strips g{y0, shift, (int)((range >> shift) + 1)};
  // listings, in rank order, then stably sorted by (layer, strip)
  int *cnt = (int *)c.ws.get("strip_cnt", 4 * (size_t)n);
  long long *off = (long long *)c.ws.get("strip_off", 8 * (size_t)(n + 1));
  {
    LaunchTimer t(c, "mbr_strip_build");
    hipLaunchKernelGGL(k_strip_rec_count, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n,
                       rec_id, rec_ylo, rec_yhi, g, cnt);
    HIP_CHECK(hipGetLastError());
  }
  scan_i32_to_i64(c, cnt, off, n);
  const long R = (long)read_scalar(c, off + n);
  if (R > 4 * N) return -1;
  const long nk = 2L * g.S;
  const size_t R1 = (size_t)std::max(R, 1L);
  uint64_t *keys = (uint64_t *)c.ws.get("strip_keys", 8 * R1);
  int *vals = (int *)c.ws.get("strip_vals", 4 * R1);
  uint64_t *skeys = (uint64_t *)c.ws.get("strip_skeys", 8 * R1);
  int *list = (int *)c.ws.get("strip_list", 4 * R1);
  long long *sOff = (long long *)c.ws.get("strip_soff", 8 * (size_t)(nk + 1));
  StripRec *qrec = (StripRec *)c.ws.get("strip_recs", sizeof(StripRec) * (size_t)n);
  {
    LaunchTimer t(c, "mbr_strip_build");
    hipLaunchKernelGGL(k_strip_rec_emit, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, B,
                       rec_id, rec_ylo, rec_yhi, g, (const long long *)off, keys, vals);
    HIP_CHECK(hipGetLastError());
  }
  int kbits = 0;
  while (kbits < 64 && ((uint64_t)(nk - 1) >> kbits) != 0) kbits++;
  radix_sort_u64(c, keys, vals, skeys, list, R, kbits);
  {
    LaunchTimer t(c, "mbr_strip_build");
    hipLaunchKernelGGL(k_strip_bounds, dim3(grid_for(R + 1, 256)), dim3(256), 0, c.stream, R,
                       (const uint64_t *)skeys, nk, sOff);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_strip_recs, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, n, rec_id,
                       rec_ylo, rec_yhi, rank, qrec);
    HIP_CHECK(hipGetLastError());
  }
  // counts; the polygons the strips leave to the sweep
  int *heavy = (int *)c.ws.get("strip_heavy", 4 * (size_t)(N + 1));
  int *inl = (int *)c.ws.get("strip_inline", 4 * (size_t)kStripInline * (size_t)N);
  fill_i32(c, heavy, 0, 1);
  {
    LaunchTimer t(c, "mbr_strip_count");
    hipLaunchKernelGGL(k_strip_query, dim3(grid_for(R, 256)), dim3(256), 0, c.stream, B, O, g,
                       (const long long *)sOff, (const int *)list, (const uint64_t *)skeys, R,
                       rank, (const StripRec *)qrec, Y, counts, (const long long *)nullptr,
                       (int *)nullptr, heavy, inl);
    HIP_CHECK(hipGetLastError());
  }
  const int nHeavy = read_scalar(c, (const int *)heavy);
  if (nHeavy > 0) {
    LaunchTimer t(c, "mbr_sweep_count");
    hipLaunchKernelGGL(k_sweep, dim3(grid_for(nHeavy, 4)), dim3(256), 0, c.stream, B, O, rank,
                       rec_id, rec_ylo, rec_yhi, Y, counts, (const long long *)nullptr,
                       (int *)nullptr, (const int *)(heavy + 1), (long)nHeavy);
    HIP_CHECK(hipGetLastError());
  }
  scan_i32_to_i64(c, counts, offs, N);
  const long K = (long)read_scalar(c, offs + N);
  int *pairs = (int *)c.ws.get("mbr_pairs", 8 * (size_t)(K > 0 ? K : 1));
  if (K > 0) {
    {
      LaunchTimer t(c, "mbr_strip_write");
      hipLaunchKernelGGL(k_strip_query, dim3(grid_for(N, 256)), dim3(256), 0, c.stream, B, O, g,
                         (const long long *)sOff, (const int *)list, (const uint64_t *)skeys, R,
                         rank, (const StripRec *)qrec, Y, counts, (const long long *)offs,
                         pairs, heavy, inl);
      HIP_CHECK(hipGetLastError());
    }
    if (nHeavy > 0) {
      LaunchTimer t(c, "mbr_sweep_write");
      hipLaunchKernelGGL(k_sweep, dim3(grid_for(nHeavy, 4)), dim3(256), 0, c.stream, B, O, rank,
                         rec_id, rec_ylo, rec_yhi, Y, counts, (const long long *)offs, pairs,
                         (const int *)(heavy + 1), (long)nHeavy);
      HIP_CHECK(hipGetLastError());
    }
  }
  *pairs_out = pairs;
  return K;
}

long mbr_filter(Context &c, long B, long O, const uint64_t *bMBR, const uint64_t *oMBR,
                int **counts_out, int **pairs_out, MbrSortView *sv, const int *bVNum,
                const int *oVNum) {
  const long N = B + O, n = 2 * N;
  if (N <= 0) return 0;
  uint64_t *keys = (uint64_t *)c.ws.get("mbr_keys", 8 * (size_t)n);
  int *idx = (int *)c.ws.get("mbr_idx", 4 * (size_t)n);
  uint64_t *Y = (uint64_t *)c.ws.get("mbr_Y", 8 * (size_t)n);
  unsigned long long *mm = (unsigned long long *)c.ws.get("mbr_minmax", 40);
  uint64_t *skeys = (uint64_t *)c.ws.get("mbr_skeys", 8 * (size_t)n);
  int *perm = (int *)c.ws.get("mbr_perm", 4 * (size_t)n);
  int *rank = (int *)c.ws.get("mbr_rank", 4 * (size_t)n);
  int *rec_id = (int *)c.ws.get("mbr_rec_id", 4 * (size_t)n);
  uint64_t *rec_ylo = (uint64_t *)c.ws.get("mbr_rec_ylo", 8 * (size_t)n);
  uint64_t *rec_yhi = (uint64_t *)c.ws.get("mbr_rec_yhi", 8 * (size_t)n);
  int *counts = (int *)c.ws.get("mbr_counts", 4 * (size_t)N);
  long long *offs = (long long *)c.ws.get("mbr_offs", 8 * (size_t)(N + 1));
  hipLaunchKernelGGL(k_init_minmax, dim3(1), dim3(64), 0, c.stream, mm);
  {
    LaunchTimer t(c, "mbr_endpoints");
    hipLaunchKernelGGL(k_endpoints, dim3(std::min(grid_for(N, 256), kEndpointBlocks)), dim3(256),
                       0, c.stream, B, O, bMBR, oMBR, bVNum && oVNum ? bVNum : nullptr, oVNum,
                       keys, idx, Y, mm);
    HIP_CHECK(hipGetLastError());
  }
  unsigned long long h[5];
  host_read(c, h, mm, 40);
  c.avgVerts = bVNum && oVNum ? (double)h[4] / (double)N : -1.0;
  unsigned long long range = h[1] - h[0];
  int bits = 0;
  while (bits < 64 && (range >> bits) != 0) bits++;
  hipLaunchKernelGGL(k_sub_key, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, keys, n, mm);
  HIP_CHECK(hipGetLastError());
  radix_sort_u64(c, keys, idx, skeys, perm, n, bits);
  {
    LaunchTimer t(c, "mbr_sweep_prep");
    hipLaunchKernelGGL(k_sweep_prep, dim3(grid_for(n, 256)), dim3(256), 0, c.stream, perm, n, Y,
                       rank, rec_id, rec_ylo, rec_yhi);
    HIP_CHECK(hipGetLastError());
  }
  long K = -1;
  int *pairs = nullptr;
  if (N >= c.tune.stripMin)
    K = strip_sweep(c, B, O, rank, rec_id, rec_ylo, rec_yhi, Y, h[2], h[3], counts, offs, &pairs);
  if (K < 0) {
    {
      LaunchTimer t(c, "mbr_sweep_count");
      hipLaunchKernelGGL(k_sweep, dim3(grid_for(N, 4)), dim3(256), 0, c.stream, B, O, rank,
                         rec_id, rec_ylo, rec_yhi, Y, counts, (const long long *)nullptr,
                         (int *)nullptr, (const int *)nullptr, 0L);
      HIP_CHECK(hipGetLastError());
    }
    scan_i32_to_i64(c, counts, offs, N);
    K = (long)read_scalar(c, offs + N);
    pairs = (int *)c.ws.get("mbr_pairs", 8 * (size_t)(K > 0 ? K : 1));
    if (K > 0) {
      LaunchTimer t(c, "mbr_sweep_write");
      hipLaunchKernelGGL(k_sweep, dim3(grid_for(N, 4)), dim3(256), 0, c.stream, B, O, rank,
                         rec_id, rec_ylo, rec_yhi, Y, counts, (const long long *)offs, pairs,
                         (const int *)nullptr, 0L);
      HIP_CHECK(hipGetLastError());
    }
  }
  if (counts_out) *counts_out = counts;
  if (pairs_out) *pairs_out = pairs;
  if (sv) {
    sv->perm = perm;
    sv->rank = rank;
    sv->n = n;
  }
  return K;
}

// ------------------------------------------------------------ stage 2 ----
// zero0..zero2 (optional): per-candidate arrays of the later passes,
// cleared here rather than by fills of their own
__global__ void k_get_cmbr(long K, const int *pairs, const double *bMBR2, const double *oMBR2,
                           double *cmbr, int *pipFlag, signed char *pipType, int *joinFlag,
                           int *zero0, int *zero1, int *zero2) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  if (zero0) {
    zero0[k] = 0;
    zero1[k] = 0;
    zero2[k] = 0;
  }
  int b = pairs[2 * k], o = pairs[2 * k + 1];
  double x1 = bMBR2[4 * b], x2 = bMBR2[4 * b + 2], y1 = bMBR2[4 * b + 1], y2 = bMBR2[4 * b + 3];
  double a1 = oMBR2[4 * o], a2 = oMBR2[4 * o + 2], b1 = oMBR2[4 * o + 1], b2 = oMBR2[4 * o + 3];
  double cX1 = a1 > x1 ? a1 : x1, cX2 = a2 < x2 ? a2 : x2;
  double cY1 = b1 > y1 ? b1 : y1, cY2 = b2 < y2 ? b2 : y2;
  cmbr[4 * k] = cX1;
  cmbr[4 * k + 1] = cY1;
  cmbr[4 * k + 2] = cX2;
  cmbr[4 * k + 3] = cY2;
  joinFlag[k] = 1;
  int f = 0, t = 0;
  if (x1 == cX1 && x2 == cX2 && y1 == cY1 && y2 == cY2) {
    f = 1;
    t = 1;
  } else if (a1 == cX1 && a2 == cX2 && b1 == cY1 && b2 == cY2) {
    f = 1;
    t = 2;
  }
  pipFlag[k] = f;
  pipType[k] = (signed char)t;
}

// ------------------------------------------------------------ stage 3 ----
__global__ void __launch_bounds__(256)
    k_pip(long nCand, const int *cand, const int *pairs, const signed char *pipType,
          const double *bCoords, const long *bVPSNum, const double *oCoords, const long *oVPSNum,
          int *pipFlag, int *joinFlag, const long long *nDev) {
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= (nDev ? (long)*nDev : nCand)) return;
  const int lane = __lane_id();
  const int k = cand[w];
  const int b = pairs[2 * k], o = pairs[2 * k + 1];
  const long bS = b > 0 ? bVPSNum[b - 1] : 0, oS = o > 0 ? oVPSNum[o - 1] : 0;
  double pX, pY;
  const double *poly;
  long vNum;
  if (pipType[k] == 1) {
    pX = bCoords[2 * bS];
    pY = bCoords[2 * bS + 1];
    vNum = oVPSNum[o] - oS;
    poly = oCoords + 2 * oS;
  } else {
    pX = oCoords[2 * oS];
    pY = oCoords[2 * oS + 1];
    vNum = bVPSNum[b] - bS;
    poly = bCoords + 2 * bS;
  }
  int parity = 0;
  for (long i0 = 0; i0 < vNum; i0 += 64) {
    long i = i0 + lane;
    bool tog = false;
    if (i < vNum) {
      long j = i == 0 ? vNum - 1 : i - 1;
      double X1 = poly[2 * i], Y1 = poly[2 * i + 1], X2 = poly[2 * j], Y2 = poly[2 * j + 1];
      if ((Y1 > pY) != (Y2 > pY)) tog = pX < (X2 - X1) * (pY - Y1) / (Y2 - Y1) + X1;
    }
    parity ^= __popcll(__ballot(tog)) & 1;
  }
  if (lane == 0) {
    pipFlag[k] = parity;
    if (parity == 1) joinFlag[k] = 0;
  }
}

// ------------------------------------------------------------ stage 4 ----
__device__ __forceinline__ bool cmf_edge(const double *c, int vNum, int i, double cX1, double cY1,
                                         double cX2, double cY2) {
  double x1 = c[2 * i], y1 = c[2 * i + 1], x2, y2;
  if (i < vNum - 1) {
    x2 = c[2 * (i + 1)];
    y2 = c[2 * (i + 1) + 1];
  } else {
    x2 = c[0];
    y2 = c[1];
  }
  if ((cX1 > x1 && cX1 > x2) || (cX2 < x1 && cX2 < x2)) return false;
  if ((cY1 > y1 && cY1 > y2) || (cY2 < y1 && cY2 < y2)) return false;
  return true;
}

__device__ __forceinline__ long long list_base(const long long *off, int k, bool inclusive) {
  return inclusive ? (k > 0 ? off[k - 1] : 0) : off[k];
}

// polygons of at most this many vertices on average take the 16-lane CMF
constexpr double kCmfSmall = 32.0;

// count mode (bEdge == nullptr): counts and clears joinFlag when a side has
// no edge; apply mode: writes the ordered edge lists at the pair offsets
// (exclusive offsets, or the reference's inclusive prefix when inclusive).
// One group of G lanes per pair (G = 16: four pairs per wave, for layers of
// small polygons; G = 64: one wave per pair); ballots are masked to the
// group, whose edges are ranked in ring order.
template <typename EdgeT, int G>
__global__ void __launch_bounds__(256)
    k_cmf(long nList, const int *list, const int *pairs, const double *cmbr,
          const double *bCoords, const int *bVNum, const long *bVPSNum, const double *oCoords,
          const int *oVNum, const long *oVPSNum, int *joinFlag, int *bCount, int *oCount,
          const long long *bOff, const long long *oOff, EdgeT *bEdge, EdgeT *oEdge,
          bool inclusive, const long long *nDev) {
  const long w = (long)blockIdx.x * (256 / G) + (threadIdx.x / G);
  if (w >= (nDev ? (long)*nDev : nList)) return;
  const int lane = __lane_id(), g = lane & (G - 1);
  const unsigned long long gm = G == 64 ? ~0ULL : (((1ULL << G) - 1ULL) << (lane & ~(G - 1)));
  const int k = list[w];
  const int b = pairs[2 * k], o = pairs[2 * k + 1];
  const double cX1 = cmbr[4 * k], cY1 = cmbr[4 * k + 1], cX2 = cmbr[4 * k + 2],
               cY2 = cmbr[4 * k + 3];
  int cnt[2];
  for (int side = 0; side < 2; side++) {
    const double *c = side == 0 ? bCoords + 2 * (b > 0 ? bVPSNum[b - 1] : 0)
                                : oCoords + 2 * (o > 0 ? oVPSNum[o - 1] : 0);
    const int vNum = side == 0 ? bVNum[b] : oVNum[o];
    EdgeT *outE = side == 0 ? bEdge : oEdge;
    long long pos = bEdge ? list_base(side == 0 ? bOff : oOff, k, inclusive) : 0;
    int n = 0;
    for (int i0 = 0; i0 < vNum; i0 += G) {
      int i = i0 + g;
      bool ok = i < vNum && cmf_edge(c, vNum, i, cX1, cY1, cX2, cY2);
      unsigned long long m = __ballot(ok) & gm;
      if (outE && ok) outE[pos + n + __popcll(m & ((1ULL << lane) - 1ULL))] = (EdgeT)i;
      n += __popcll(m);
    }
    cnt[side] = n;
  }
  if (!bEdge && g == 0) {
    if (cnt[0] == 0 || cnt[1] == 0) {
      joinFlag[k] = 0;
      bCount[k] = 0;
      oCount[k] = 0;
    } else {
      bCount[k] = cnt[0];
      oCount[k] = cnt[1];
    }
  }
}

// ------------------------------------------------------------ stage 5 ----
__device__ __forceinline__ bool seg_hit(double bX1, double bY1, double bX2, double bY2,
                                        double oX1, double oY1, double oX2, double oY2) {
  if (((bX2 < oX1) && (bX2 < oX2)) && ((bX1 < oX1) && (bX1 < oX2))) return false;
  if (((bX2 > oX1) && (bX2 > oX2)) && ((bX1 > oX1) && (bX1 > oX2))) return false;
  if (((bY2 < oY1) && (bY2 < oY2)) && ((bY1 < oY1) && (bY1 < oY2))) return false;
  if (((bY2 > oY1) && (bY2 > oY2)) && ((bY1 > oY1) && (bY1 > oY2))) return false;
  double denom = ((bX1 - bX2) * (oY1 - oY2)) - ((bY1 - bY2) * (oX1 - oX2));
  if (denom == 0) {
    if ((oX1 == oX2 && oX1 == bX1) || (bX1 == bX2 && bX1 == oX1)) return true;
    double oM = (oY2 - oY1) / (oX2 - oX1);
    double bM = (bY2 - bY1) / (bX2 - bX1);
    double oB = oY1 - oM * oX1;
    double bB = bY1 - bM * bX1;
    return oB == bB;
  }
  double nomX = (((bX1 * bY2) - (bY1 * bX2)) * (oX1 - oX2)) -
                ((bX1 - bX2) * ((oX1 * oY2) - (oY1 * oX2)));
  double nomY = ((((bX1 * bY2) - (bY1 * bX2)) * (oY1 - oY2)) -
                 (((bY1 - bY2) * ((oX1 * oY2) - (oY1 * oX2)))));
  double x0 = nomX / denom, y0 = nomY / denom;
  if ((x0 > bX1 && x0 > bX2) || (x0 < bX1 && x0 < bX2)) return false;
  if ((x0 > oX1 && x0 > oX2) || (x0 < oX1 && x0 < oX2)) return false;
  if ((y0 > bY1 && y0 > bY2) || (y0 < bY1 && y0 < bY2)) return false;
  if ((y0 > oY1 && y0 > oY2) || (y0 < oY1 && y0 < oY2)) return false;
  return true;
}

constexpr int kSegStage = 4;

__device__ __forceinline__ int gcd_int(int a, int b) {
  while (b) {
    const int t = a % b;
    a = b;
    b = t;
  }
  return a;
}

// The order in which a pair's base edges are taken: i -> (i * S) mod bE,
// S coprime to bE and near 0.618 bE, a permutation that spreads successive
// rounds over the whole base list.  The flag is "some edge pair hits", so
// the order changes only how soon a pair that hits stops: its hits cluster
// where the rings cross, which index order reaches after half the list on
// average and this order after about bE / (number of base edges with hits).
__device__ __forceinline__ int seg_stride(int bE) {
  if (bE <= 2) return 1;
  int S = (int)(0.6180339887498949 * bE);
  if (S < 1) S = 1;
  while (gcd_int(S, bE) != 1) S++;
  return S;
}

// counts come from bCount/oCount, or (when null) from the offsets; the flag
// goes to segFlag[k] or, with flagByList, to segFlag[w] (the reference's
// per-survivor layout).
template <typename EdgeT>
__global__ void __launch_bounds__(256)
    k_segjoin(long nList, const int *list, const int *pairs, const double *bCoords,
              const long *bVPSNum, const double *oCoords, const long *oVPSNum, const int *bCount,
              const int *oCount, const long long *bOff, const long long *oOff,
              const EdgeT *bEdge, const EdgeT *oEdge, int *segFlag, bool inclusive,
              bool flagByList, const long long *nDev) {
  const long w = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (w >= (nDev ? (long)*nDev : nList)) return;
  const int lane = __lane_id();
  const int k = list[w];
  const int b = pairs[2 * k], o = pairs[2 * k + 1];
  const long bBase = b > 0 ? bVPSNum[b - 1] : 0, oBase = o > 0 ? oVPSNum[o - 1] : 0;
  const int bV = (int)(bVPSNum[b] - bBase), oV = (int)(oVPSNum[o] - oBase);
  const long long b0 = list_base(bOff, k, inclusive), o0 = list_base(oOff, k, inclusive);
  const int bE = bCount ? bCount[k] : (int)(bOff[k] - b0);
  const int oE = oCount ? oCount[k] : (int)(oOff[k] - o0);
  const EdgeT *be = bEdge + b0, *oe = oEdge + o0;
  const double *bc = bCoords + 2 * bBase, *oc = oCoords + 2 * oBase;
  const long long total = (long long)bE * oE;
  int flag = 0;
  // edge pair t's four vertices; false when the base edge is degenerate
  auto pair_at = [&](long long t, double *v) {
    const int bi = be[t / oE], oi = oe[t % oE];
    const int b2 = bi < bV - 1 ? bi + 1 : 0, o2 = oi < oV - 1 ? oi + 1 : 0;
    v[0] = bc[2 * bi], v[1] = bc[2 * bi + 1], v[2] = bc[2 * b2], v[3] = bc[2 * b2 + 1];
    v[4] = oc[2 * oi], v[5] = oc[2 * oi + 1], v[6] = oc[2 * o2], v[7] = oc[2 * o2 + 1];
    return !(v[0] == v[2] && v[1] == v[3]);
  };
  if (total >= 0x7fffffffLL) {
    // lists of 2^31 pairs or more (rings of tens of thousands of listed
    // edges): one pair per lane a round, 64-bit indices
    for (long long t0 = 0; t0 < total && !flag; t0 += 64) {
      const long long t = t0 + lane;
      double v[8];
      const bool hit = t < total && pair_at(t, v) &&
                       seg_hit(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7]);
      flag = __ballot(hit) != 0ULL;
    }
    if (lane == 0) segFlag[flagByList ? w : k] = flag;
    return;
  }
  // kSegStage edge pairs per lane a round, all their loads issued before
  // the tests; the flag is "some pair hits", which testing pairs past the
  // first hit cannot change
  const unsigned tot = (unsigned)total, uoE = (unsigned)(oE > 0 ? oE : 1);
  const unsigned long long S = (unsigned long long)seg_stride(bE);
  for (unsigned t0 = 0; t0 < tot && !flag; t0 += 64 * kSegStage) {
    int bi[kSegStage], oi[kSegStage];
#pragma unroll
    for (int u = 0; u < kSegStage; u++) {
      const unsigned t = t0 + 64 * u + lane;
      bi[u] = t < tot ? (int)be[(unsigned)((t / uoE) * S % (unsigned long long)bE)] : -1;
      oi[u] = t < tot ? (int)oe[t % uoE] : -1;
    }
    double s[kSegStage][8];
#pragma unroll
    for (int u = 0; u < kSegStage; u++) {
      if (bi[u] < 0) continue;
      const int b2 = bi[u] < bV - 1 ? bi[u] + 1 : 0, o2 = oi[u] < oV - 1 ? oi[u] + 1 : 0;
      s[u][0] = bc[2 * bi[u]];
      s[u][1] = bc[2 * bi[u] + 1];
      s[u][2] = bc[2 * b2];
      s[u][3] = bc[2 * b2 + 1];
      s[u][4] = oc[2 * oi[u]];
      s[u][5] = oc[2 * oi[u] + 1];
      s[u][6] = oc[2 * o2];
      s[u][7] = oc[2 * o2 + 1];
    }
    bool hit = false;
#pragma unroll
    for (int u = 0; u < kSegStage; u++) {
      if (bi[u] < 0) continue;
      if (!(s[u][0] == s[u][2] && s[u][1] == s[u][3]))
        hit |= seg_hit(s[u][0], s[u][1], s[u][2], s[u][3], s[u][4], s[u][5], s[u][6], s[u][7]);
    }
    flag = __ballot(hit) != 0ULL;
  }
  if (lane == 0) segFlag[flagByList ? w : k] = flag;
}

__global__ void k_join_final(long K, const int *pipFlag, const int *segFlag, int *joined) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < K) joined[k] = (pipFlag[k] == 1 || segFlag[k] == 1) ? 1 : 0;
}

__global__ void k_gather_pairs(long n, const int *idx, const int *pairs, int *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = idx[i];
  out[2 * i] = pairs[2 * k];
  out[2 * i + 1] = pairs[2 * k + 1];
}

void join_refine(Context &c, const JoinInputs &in, long K, const int *pairs, JoinRefine &r) {
  r.K = K;
  long Kc = K > 0 ? K : 1;
  r.cmbr = (double *)c.ws.get("jr_cmbr", 32 * (size_t)Kc);
  r.pipFlag = (int *)c.ws.get("jr_pipFlag", 4 * (size_t)Kc);
  r.pipType = (signed char *)c.ws.get("jr_pipType", (size_t)Kc);
  r.joinFlag = (int *)c.ws.get("jr_joinFlag", 4 * (size_t)Kc);
  r.pipCand = (int *)c.ws.get("jr_pipCand", 4 * (size_t)Kc);
  r.bCount = (int *)c.ws.get("jr_bCount", 4 * (size_t)Kc);
  r.oCount = (int *)c.ws.get("jr_oCount", 4 * (size_t)Kc);
  r.bOff = (long long *)c.ws.get("jr_bOff", 8 * (size_t)(Kc + 1));
  r.oOff = (long long *)c.ws.get("jr_oOff", 8 * (size_t)(Kc + 1));
  r.segFlag = (int *)c.ws.get("jr_segFlag", 4 * (size_t)Kc);
  r.list = (int *)c.ws.get("jr_list", 4 * (size_t)Kc);
  r.joined = (int *)c.ws.get("jr_joined", 4 * (size_t)Kc);
  if (K <= 0) {
    r.pipNum = r.within = r.cmfCand = r.E = r.hits = r.nJoined = 0;
    r.joinedPairs = (int *)c.ws.get("jr_joinedPairs", 8);
    return;
  }
  {
    LaunchTimer t(c, "get_cmbr");
    hipLaunchKernelGGL(k_get_cmbr, dim3(grid_for(K, 256)), dim3(256), 0, c.stream, K, pairs,
                       in.bMBR2, in.oMBR2, r.cmbr, r.pipFlag, r.pipType, r.joinFlag, r.bCount,
                       r.oCount, r.segFlag);
    HIP_CHECK(hipGetLastError());
  }
  // list sizes stay on the device (grids are sized by K); they come back in
  // one read at the end
  const long long *pipN = compact_flags_dev(c, r.pipFlag, K, r.pipCand, "pip");
  {
    LaunchTimer t(c, "pip");
    hipLaunchKernelGGL(k_pip, dim3(grid_for(K, 4)), dim3(256), 0, c.stream, K, r.pipCand, pairs,
                       r.pipType, in.bCoords, in.bVPSNum, in.oCoords, in.oVPSNum, r.pipFlag,
                       r.joinFlag, pipN);
    HIP_CHECK(hipGetLastError());
  }
  // within count = number of pipFlag==1 after the test (bCount, oCount and
  // segFlag were cleared by k_get_cmbr)
  const long long *withinN = compact_flags_dev(c, r.pipFlag, K, r.list, "within");
  const long long *cmfN = compact_flags_dev(c, r.joinFlag, K, r.list, "cmf");
  // four pairs per wave where the layers' polygons are small on average
  const bool cmf16 = c.tune.cmfGroup == 16 ||
                     (c.tune.cmfGroup == 0 && c.avgVerts >= 0 && c.avgVerts <= kCmfSmall);
  const int cmfPer = cmf16 ? 16 : 4;
  {
    LaunchTimer t(c, "cmf_count");
    hipLaunchKernelGGL((cmf16 ? k_cmf<int, 16> : k_cmf<int, 64>), dim3(grid_for(K, cmfPer)),
                       dim3(256), 0, c.stream, K, r.list,
                       pairs, r.cmbr, in.bCoords, in.bVNum, in.bVPSNum, in.oCoords, in.oVNum,
                       in.oVPSNum, r.joinFlag, r.bCount, r.oCount, (const long long *)nullptr,
                       (const long long *)nullptr, (int *)nullptr, (int *)nullptr, false, cmfN);
    HIP_CHECK(hipGetLastError());
  }
  scan_i32_to_i64_multi(c, {{r.bCount, r.bOff, K}, {r.oCount, r.oOff, K}});
  long long rb[kHostSlots];
  read_scalars(c, {r.bOff + K, r.oOff + K}, rb);
  r.nbEdges = (long)rb[0];
  r.noEdges = (long)rb[1];
  r.bEdge = (int *)c.ws.get("jr_bEdge", 4 * (size_t)(r.nbEdges + 1));
  r.oEdge = (int *)c.ws.get("jr_oEdge", 4 * (size_t)(r.noEdges + 1));
  const long long *EN = compact_flags_dev(c, r.joinFlag, K, r.list, "cmf2");
  {
    LaunchTimer t(c, "cmf_apply");
    hipLaunchKernelGGL((cmf16 ? k_cmf<int, 16> : k_cmf<int, 64>), dim3(grid_for(K, cmfPer)),
                       dim3(256), 0, c.stream, K, r.list,
                       pairs, r.cmbr, in.bCoords, in.bVNum, in.bVPSNum, in.oCoords, in.oVNum,
                       in.oVPSNum, r.joinFlag, r.bCount, r.oCount, (const long long *)r.bOff,
                       (const long long *)r.oOff, r.bEdge, r.oEdge, false, EN);
    HIP_CHECK(hipGetLastError());
  }
  {
    LaunchTimer t(c, "segment_join");
    hipLaunchKernelGGL(k_segjoin<int>, dim3(grid_for(K, 4)), dim3(256), 0, c.stream, K, r.list,
                       pairs, in.bCoords, in.bVPSNum, in.oCoords, in.oVPSNum, r.bCount, r.oCount,
                       (const long long *)r.bOff, (const long long *)r.oOff, r.bEdge, r.oEdge,
                       r.segFlag, false, false, EN);
    HIP_CHECK(hipGetLastError());
  }
  int *tmp = (int *)c.ws.get("jr_hits_idx", 4 * (size_t)Kc);
  const long long *hitsN = compact_flags_dev(c, r.segFlag, K, tmp, "hits");
  hipLaunchKernelGGL(k_join_final, dim3(grid_for(K, 256)), dim3(256), 0, c.stream, K, r.pipFlag,
                     r.segFlag, r.joined);
  HIP_CHECK(hipGetLastError());
  int *jidx = (int *)c.ws.get("jr_joined_idx", 4 * (size_t)Kc);
  const long long *joinedN = compact_flags_dev(c, r.joined, K, jidx, "joined");
  read_scalars(c, {pipN, withinN, cmfN, EN, hitsN, joinedN}, rb);
  r.pipNum = (long)rb[0];
  r.within = (long)rb[1];
  r.cmfCand = (long)rb[2];
  r.E = (long)rb[3];
  r.hits = (long)rb[4];
  r.nJoined = (long)rb[5];
  r.joinedIdx = jidx;
  r.joinedPairs = (int *)c.ws.get("jr_joinedPairs", 8 * (size_t)(r.nJoined > 0 ? r.nJoined : 1));
  if (r.nJoined > 0) {
    hipLaunchKernelGGL(k_gather_pairs, dim3(grid_for(r.nJoined, 256)), dim3(256), 0, c.stream,
                       r.nJoined, jidx, pairs, r.joinedPairs);
    HIP_CHECK(hipGetLastError());
  }
}

// ------------------------------------------- single-stage launchers ----
void launch_get_cmbr(Context &c, long K, const int *pairs, const double *bMBR2,
                     const double *oMBR2, double *cmbr, int *pipFlag, signed char *pipType,
                     int *joinFlag) {
  if (K <= 0) return;
  LaunchTimer t(c, "get_cmbr");
  hipLaunchKernelGGL(k_get_cmbr, dim3(grid_for(K, 256)), dim3(256), 0, c.stream, K, pairs, bMBR2,
                     oMBR2, cmbr, pipFlag, pipType, joinFlag, (int *)nullptr, (int *)nullptr,
                     (int *)nullptr);
  HIP_CHECK(hipGetLastError());
}

void launch_pip(Context &c, long nCand, const int *cand, const int *pairs,
                const signed char *pipType, const double *bCoords, const long *bVPSNum,
                const double *oCoords, const long *oVPSNum, int *pipFlag, int *joinFlag) {
  if (nCand <= 0) return;
  LaunchTimer t(c, "pip");
  hipLaunchKernelGGL(k_pip, dim3(grid_for(nCand, 4)), dim3(256), 0, c.stream, nCand, cand, pairs,
                     pipType, bCoords, bVPSNum, oCoords, oVPSNum, pipFlag, joinFlag,
                     (const long long *)nullptr);
  HIP_CHECK(hipGetLastError());
}

void launch_cmf_count(Context &c, long nList, const int *list, const int *pairs,
                      const double *cmbr, const JoinInputs &in, int *joinFlag, int *bCount,
                      int *oCount) {
  if (nList <= 0) return;
  LaunchTimer t(c, "cmf_count");
  hipLaunchKernelGGL((k_cmf<int, 64>), dim3(grid_for(nList, 4)), dim3(256), 0, c.stream, nList,
                     list,
                     pairs, cmbr, in.bCoords, in.bVNum, in.bVPSNum, in.oCoords, in.oVNum,
                     in.oVPSNum, joinFlag, bCount, oCount, (const long long *)nullptr,
                     (const long long *)nullptr, (int *)nullptr, (int *)nullptr, false,
                     (const long long *)nullptr);
  HIP_CHECK(hipGetLastError());
}

void launch_cmf_apply_short(Context &c, long nList, const int *list, const int *pairs,
                            const double *cmbr, const JoinInputs &in, const long long *bIncl,
                            const long long *oIncl, short *bEdge, short *oEdge) {
  if (nList <= 0) return;
  LaunchTimer t(c, "cmf_apply");
  hipLaunchKernelGGL((k_cmf<short, 64>), dim3(grid_for(nList, 4)), dim3(256), 0, c.stream,
                     nList, list,
                     pairs, cmbr, in.bCoords, in.bVNum, in.bVPSNum, in.oCoords, in.oVNum,
                     in.oVPSNum, (int *)nullptr, (int *)nullptr, (int *)nullptr, bIncl, oIncl,
                     bEdge, oEdge, true, (const long long *)nullptr);
  HIP_CHECK(hipGetLastError());
}

void launch_segjoin_short(Context &c, long nList, const int *list, const int *pairs,
                          const JoinInputs &in, const long long *bIncl, const long long *oIncl,
                          const short *bEdge, const short *oEdge, int *flagByList) {
  if (nList <= 0) return;
  LaunchTimer t(c, "segment_join");
  hipLaunchKernelGGL(k_segjoin<short>, dim3(grid_for(nList, 4)), dim3(256), 0, c.stream, nList,
                     list, pairs, in.bCoords, in.bVPSNum, in.oCoords, in.oVPSNum,
                     (const int *)nullptr, (const int *)nullptr, bIncl, oIncl, bEdge, oEdge,
                     flagByList, true, true, (const long long *)nullptr);
  HIP_CHECK(hipGetLastError());
}

}  // namespace gcmf
