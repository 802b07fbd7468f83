// gh.hip -- batched Greiner-Hormann (Foster-Hormann-Popa) clipping of many
// polygon pairs in one pass on gfx950.
//
// Phase 1 reproduces the reference GPU phase (gppolyclip_mDIV512.cu:382-1223)
// for every pair of the batch at once: all edges of all pairs form one index
// space (edge -> pair map), counts of both views are scanned globally so each
// pair's slot range falls out of the scan, and the intersection records,
// neighbour links, per-edge alpha order and initial labels are written at
// those global slots.  The reference's host round trip (D2H counts, host
// thrust scans, H2D) and its per-pair allocations disappear.
//
// Phases 2-3 (labelling, tracing, clean-up; polyclip.h:335-807 with the
// commented step 1 restored, fed as ghcuda.cpp:183-339 feeds it) run on the
// GPU over array-linked node lists: one wave per pair (k_label_wave), one
// 1024-thread block per pair of Tuning::bigNodes nodes or more
// (k_label_block), and the sequential steps on one lane for the degenerate
// pairs the parallel steps do not cover.
//
// Every predicate runs on per-pair normalised coordinates
// ((x-cx)*s, (y-cy)*s), (cx,cy) = CMBR corner, s = 1/max CMBR extent, and
// every output vertex is produced in original coordinates from the same alpha.
#include <functional>

#include <hip/hip_cooperative_groups.h>

#include "gcmf_device.hpp"
#include "gcmf_internal.hpp"
#include "gcmf_join.hpp"

namespace gcmf {

__device__ __forceinline__ pt ld(const double *c, long i) { return pt{c[2 * i], c[2 * i + 1]}; }
__device__ __forceinline__ void st(double *c, long i, pt p) {
  c[2 * i] = p.x;
  c[2 * i + 1] = p.y;
}

// Appends the lanes with pred to a list counted by *cnt with one atomic per
// wave; returns each such lane's index (lane order within the wave).
__device__ __forceinline__ long wave_append(unsigned long long *cnt, bool pred) {
  const unsigned long long m = __ballot(pred);
  if (!m) return -1;
  const int leader = __ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (__lane_id() == leader) base = atomicAdd(cnt, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  return pred ? (long)base + __popcll(m & ((1ULL << __lane_id()) - 1ULL)) : -1;
}

// The same with one atomic per block (blockDim.x <= 1024); every thread of
// the block must call it.  lds: 17 values of scratch.
__device__ __forceinline__ long block_append(unsigned long long *cnt, bool pred,
                                             unsigned long long *lds) {
  const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  const unsigned long long m = __ballot(pred);
  if (__lane_id() == 0) lds[wave] = (unsigned long long)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < nw; w++) {
      const unsigned long long c = lds[w];
      lds[w] = t;
      t += c;
    }
    lds[16] = t ? atomicAdd(cnt, t) : 0ULL;
  }
  __syncthreads();
  const long r = pred ? (long)(lds[16] + lds[wave]) +
                            __popcll(m & ((1ULL << __lane_id()) - 1ULL))
                      : -1;
  __syncthreads();
  return r;
}

// block_append for four lists at once, one block scan for all four: the
// per-wave counts travel as 16-bit fields of one word (a block holds at most
// 1024 threads).  Returns in r[i] the slot in list i, or -1.
__device__ __forceinline__ void block_append4(unsigned long long *const cnt[4], const bool pred[4],
                                              unsigned long long *lds, long r[4]) {
  const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  unsigned long long m[4], packed = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    m[i] = __ballot(pred[i]);
    packed |= (unsigned long long)__popcll(m[i]) << (16 * i);
  }
  if (__lane_id() == 0) lds[wave] = packed;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < nw; w++) {
      const unsigned long long c = lds[w];
      lds[w] = t;
      t += c;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const unsigned long long ti = (t >> (16 * i)) & 0xffffULL;
      lds[16 + i] = ti ? atomicAdd(cnt[i], ti) : 0ULL;
    }
  }
  __syncthreads();
  const unsigned long long below = (1ULL << __lane_id()) - 1ULL, before = lds[wave];
#pragma unroll
  for (int i = 0; i < 4; i++)
    r[i] = pred[i] ? (long)(lds[16 + i] + ((before >> (16 * i)) & 0xffffULL)) +
                         __popcll(m[i] & below)
                   : -1;
  __syncthreads();
}

// ------------------------------------------------------------- setup ----
struct PairMeta {
  long *startP, *startQ;  // ring start (vertex index) inside the source layer
  int *nP, *nQ, *swapped, *status;
  double *norm;   // [3M] cx, cy, s
  double *cmbrN;  // [4M] normalised CMBR
};

// The per-pair setup kernels run one group of W threads per pair: four waves
// of a 256-thread block (W = 64) for batches, one 1024-thread block (W =
// 1024) when a few pairs are large.  Only exact min / max reductions cross
// lanes, so both widths give identical results.
template <int W>
__device__ __forceinline__ long pgrp_pair() {
  return W == 64 ? (long)blockIdx.x * 4 + (threadIdx.x >> 6) : (long)blockIdx.x;
}
template <int W>
__device__ __forceinline__ int pgrp_lane() {
  return W == 64 ? (int)__lane_id() : (int)threadIdx.x;
}
// (min of a, max of b) over the group
template <int W>
__device__ __forceinline__ void pgrp_minmax(double &a, double &b) {
  for (int off = 32; off > 0; off >>= 1) {
    a = fmin(a, __shfl_xor(a, off, 64));
    b = fmax(b, __shfl_xor(b, off, 64));
  }
  if (W == 64) return;
  __shared__ double sa[W / 64], sb[W / 64];
  __syncthreads();
  if (__lane_id() == 0) {
    sa[threadIdx.x >> 6] = a;
    sb[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  a = sa[0];
  b = sb[0];
  for (int i = 1; i < W / 64; i++) {
    a = fmin(a, sa[i]);
    b = fmax(b, sb[i]);
  }
}
// threads per block and blocks for M pairs at group width W
inline unsigned pgrp_grid(long M, int W) { return W == 64 ? grid_for(M, 4) : (unsigned)M; }
inline int pgrp_block(int W) { return W == 64 ? 256 : W; }
// Kernels without cross-lane work (copies, per-vertex maps) split each pair
// of the 1024-thread width over `split` blocks, so that a few large pairs
// still fill the GPU: block b takes pair b / split, part b % split.
inline int pgrp_split(long M, int W) { return W == 64 ? 1 : (int)std::max<long>(1, 512 / M); }
template <int W>
__device__ __forceinline__ long pgrp_split_pair(int split) {
  return W == 64 ? pgrp_pair<W>() : (long)(blockIdx.x / (unsigned)split);
}
template <int W>
__device__ __forceinline__ int pgrp_split_part(int split) {
  return W == 64 ? 0 : (int)(blockIdx.x % (unsigned)split);
}

// one group per pair: sizes, swap (|P| >= |Q|, ghcuda.cpp:64-91), ring MBRs
// (readShapefile.h:304-334), CMBR and disjointness (ghcuda.cpp:94-125)
// out[0] = max of bm[2b], out[1] = max of bm[2b + 1] over b < n (out zeroed
// beforehand); a grid of up to 64 blocks, one atomic per block and value
__global__ void __launch_bounds__(256)
    k_max_block_pairs(const unsigned long long *bm, long n, unsigned long long *out) {
  __shared__ unsigned long long s0[4], s1[4];
  unsigned long long a = 0, q = 0;
  for (long b = (long)blockIdx.x * blockDim.x + threadIdx.x; b < n;
       b += (long)gridDim.x * blockDim.x) {
    a = max(a, bm[2 * b]);
    q = max(q, bm[2 * b + 1]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    a = max(a, (unsigned long long)__shfl_xor(a, off, 64));
    q = max(q, (unsigned long long)__shfl_xor(q, off, 64));
  }
  if (__lane_id() == 0) {
    s0[threadIdx.x >> 6] = a;
    s1[threadIdx.x >> 6] = q;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = max(max(s0[0], s0[1]), max(s0[2], s0[3]));
    q = max(max(s1[0], s1[1]), max(s1[2], s1[3]));
    if (a) atomicMax(out, a);
    if (q) atomicMax(out + 1, q);
  }
}

// The ring boxes of a few large pairs over `split` 1024-thread blocks per
// pair: block b takes a strided part of pair b / split's rings and leaves
// its (min x, min y, max x, max y) of P and of Q in part[8 b..8 b + 8), as
// the pairs are oriented by k_pair_setup (a ring's vertex 0 seeds every
// part, which leaves the minima and maxima unchanged).
__global__ void __launch_bounds__(1024)
    k_pair_box_parts(long M, int split, const int *pairs, const double *bCoords,
                     const int *bVNum, const long *bVPSNum, const double *oCoords,
                     const int *oVNum, const long *oVPSNum, int noswap, double *part) {
  const long k = (long)(blockIdx.x / (unsigned)split);
  const int j = (int)(blockIdx.x % (unsigned)split);
  if (k >= M) return;
  const int b = pairs[2 * k], o = pairs[2 * k + 1];
  const int nb = bVNum[b] - 1, no = oVNum[o] - 1;
  const long sb = bVPSNum[b] - bVNum[b], so = oVPSNum[o] - oVNum[o];
  const int sw = noswap ? 0 : (nb < no);
  for (int side = 0; side < 2; side++) {
    const double *c = side == 0 ? (sw ? oCoords + 2 * so : bCoords + 2 * sb)
                                : (sw ? bCoords + 2 * sb : oCoords + 2 * so);
    const int n = side == 0 ? (sw ? no : nb) : (sw ? nb : no);
    double mnx = c[0], mny = c[1], mxx = c[0], mxy = c[1];
    for (long i = (long)j * 1024 + threadIdx.x; i < n; i += (long)split * 1024) {
      const double x = c[2 * i], y = c[2 * i + 1];
      mnx = fmin(mnx, x);
      mny = fmin(mny, y);
      mxx = fmax(mxx, x);
      mxy = fmax(mxy, y);
    }
    pgrp_minmax<1024>(mnx, mxx);
    pgrp_minmax<1024>(mny, mxy);
    if (threadIdx.x == 0) {
      double *d = part + 8 * (long)blockIdx.x + 4 * side;
      d[0] = mnx;
      d[1] = mny;
      d[2] = mxx;
      d[3] = mxy;
    }
  }
}

template <int W>
__global__ void __launch_bounds__(W == 64 ? 256 : W)
    k_pair_setup(long M, const int *pairs, const double *bCoords, const int *bVNum,
                 const long *bVPSNum, const double *oCoords, const int *oVNum, const long *oVPSNum,
                 int normalize, int noswap, PairMeta pm, int *eP, int *eQ,
                 unsigned long long *blockMax, const double *part, int split) {
  const long k = pgrp_pair<W>();
  if (W == 64) {
    __shared__ int lmax[2][4];
    if (threadIdx.x < 8) lmax[threadIdx.x >> 2][threadIdx.x & 3] = 0;
    __syncthreads();
    if (k < M) {
      // the largest P and Q rings as the pairs are oriented below (P the
      // longer ring unless noswap keeps P the base ring)
      const int b = pairs[2 * k], o = pairs[2 * k + 1];
      const int nb = bVNum[b] - 1, no = oVNum[o] - 1;
      const bool sw = !noswap && nb < no;
      if (__lane_id() == 0) {
        lmax[0][threadIdx.x >> 6] = max(sw ? no : nb, 0);
        lmax[1][threadIdx.x >> 6] = max(sw ? nb : no, 0);
      }
    }
    __syncthreads();
    // the block's largest rings, stored plainly (k_max_block_pairs folds
    // them): an atomic per block on one address serialises across the chip
    if (threadIdx.x == 0) {
      int a = 0, q = 0;
      for (int w = 0; w < 4; w++) {
        a = max(a, lmax[0][w]);
        q = max(q, lmax[1][w]);
      }
      blockMax[2 * blockIdx.x] = (unsigned long long)a;
      blockMax[2 * blockIdx.x + 1] = (unsigned long long)q;
    }
  } else if (threadIdx.x == 0) {
    int a = 0, q = 0;
    if (k < M) {
      const int nb = bVNum[pairs[2 * k]] - 1, no = oVNum[pairs[2 * k + 1]] - 1;
      const bool sw = !noswap && nb < no;
      a = max(sw ? no : nb, 0);
      q = max(sw ? nb : no, 0);
    }
    blockMax[2 * blockIdx.x] = (unsigned long long)a;
    blockMax[2 * blockIdx.x + 1] = (unsigned long long)q;
  }
  if (k >= M) return;
  const int lane = pgrp_lane<W>();
  const int b = pairs[2 * k], o = pairs[2 * k + 1];
  const int nb = bVNum[b] - 1, no = oVNum[o] - 1;
  const long sb = bVPSNum[b] - bVNum[b], so = oVPSNum[o] - oVNum[o];
  const int sw = noswap ? 0 : (nb < no);
  const int nP = sw ? no : nb, nQ = sw ? nb : no;
  const long sP = sw ? so : sb, sQ = sw ? sb : so;
  const double *cP = sw ? oCoords : bCoords, *cQ = sw ? bCoords : oCoords;
  double m[2][4];
  for (int side = 0; side < 2; side++) {
    const double *c = side == 0 ? cP + 2 * sP : cQ + 2 * sQ;
    const int n = side == 0 ? nP : nQ;
    double mnx = c[0], mny = c[1], mxx = c[0], mxy = c[1];
    if (part) {
      // the parts k_pair_box_parts left (W = 1024)
      for (int j = lane; j < split; j += W) {
        const double *d = part + 8 * (k * split + j) + 4 * side;
        mnx = fmin(mnx, d[0]);
        mny = fmin(mny, d[1]);
        mxx = fmax(mxx, d[2]);
        mxy = fmax(mxy, d[3]);
      }
    } else {
      for (int i = lane; i < n; i += W) {
        double x = c[2 * i], y = c[2 * i + 1];
        mnx = fmin(mnx, x);
        mny = fmin(mny, y);
        mxx = fmax(mxx, x);
        mxy = fmax(mxy, y);
      }
    }
    pgrp_minmax<W>(mnx, mxx);
    pgrp_minmax<W>(mny, mxy);
    m[side][0] = mnx;
    m[side][1] = mny;
    m[side][2] = mxx;
    m[side][3] = mxy;
  }
  if (lane != 0) return;
  int status = GH_OK;
  if (nP < 1 || nQ < 1) status = GH_DISJOINT;
  if (m[0][0] > m[1][2] || m[0][2] < m[1][0]) status = GH_DISJOINT;
  if (m[0][1] > m[1][3] || m[0][3] < m[1][1]) status = GH_DISJOINT;
  double cm[4];
  cm[0] = m[0][0] < m[1][0] ? m[1][0] : m[0][0];
  cm[1] = m[0][1] < m[1][1] ? m[1][1] : m[0][1];
  cm[2] = m[1][2] < m[0][2] ? m[1][2] : m[0][2];
  cm[3] = m[1][3] < m[0][3] ? m[1][3] : m[0][3];
  double cx = 0.0, cy = 0.0, s = 1.0;
  if (normalize & 1) {
    double w = cm[2] - cm[0], h = cm[3] - cm[1];
    double e = w > h ? w : h;
    cx = cm[0];
    cy = cm[1];
    s = e > 0.0 ? 1.0 / e : 1.0;
    if (normalize & 4) {
      // edge-scaled: times the power of two f >= sqrt(nP nQ) / 1024, so that
      // fine rings keep edge lengths well above the absolute epsilon
      const double r = sqrt((double)nP * (double)nQ) / 1024.0;
      double f = 1.0;
      while (f < r) f *= 2.0;
      s *= f;
    }
  }
  pm.startP[k] = sP;
  pm.startQ[k] = sQ;
  pm.nP[k] = nP;
  pm.nQ[k] = nQ;
  pm.swapped[k] = sw;
  pm.status[k] = status;
  pm.norm[3 * k] = cx;
  pm.norm[3 * k + 1] = cy;
  pm.norm[3 * k + 2] = s;
  pm.cmbrN[4 * k] = (cm[0] - cx) * s;
  pm.cmbrN[4 * k + 1] = (cm[1] - cy) * s;
  pm.cmbrN[4 * k + 2] = (cm[2] - cx) * s;
  pm.cmbrN[4 * k + 3] = (cm[3] - cy) * s;
  eP[k] = status == GH_OK ? nP : 0;
  eQ[k] = status == GH_OK ? nQ : 0;
}

// one group per pair (split over several blocks at width 1024): edge ->
// pair maps and normalised rings
template <int W>
__global__ void __launch_bounds__(W == 64 ? 256 : W)
    k_pair_fill(long M, int split, const int *pairs, const double *bCoords,
                const double *oCoords, PairMeta pm, const long long *offP, const long long *offQ,
                int *mapP, int *mapQ, double *nP_, double *nQ_) {
  const long k = pgrp_split_pair<W>(split);
  if (k >= M) return;
  const int lane = pgrp_lane<W>() + W * pgrp_split_part<W>(split);
  const long long stride = (long long)W * split;
  const int sw = pm.swapped[k];
  const double cx = pm.norm[3 * k], cy = pm.norm[3 * k + 1], s = pm.norm[3 * k + 2];
  for (int side = 0; side < 2; side++) {
    const long long base = side == 0 ? offP[k] : offQ[k];
    const long long n = (side == 0 ? offP[k + 1] : offQ[k + 1]) - base;
    const double *c = (side == 0) == (sw == 0) ? bCoords : oCoords;
    const long start = side == 0 ? pm.startP[k] : pm.startQ[k];
    int *map = side == 0 ? mapP : mapQ;
    double *dst = side == 0 ? nP_ : nQ_;
    for (long long i = lane; i < n; i += stride) {
      double x = c[2 * (start + i)], y = c[2 * (start + i) + 1];
      dst[2 * (base + i)] = (x - cx) * s;
      dst[2 * (base + i) + 1] = (y - cy) * s;
      map[base + i] = (int)k;
    }
  }
}

// ------------------------------------------------- edge-pair prefilter ----
// The edge-pair loops walk the other ring with the next vertex loaded one
// iteration ahead, and reject an edge first by plain comparisons of its two
// endpoints against this edge's box: both endpoints on one side of the box
// means lsmf (which takes the boxes' min / max) fails as well, for finite
// coordinates and NaN alike, so only pairs lsmf could accept go on to it.
struct EdgeBox {
  double x0, y0, x1, y1;
};
__device__ __forceinline__ EdgeBox edge_box(pt a, pt b) {
  return EdgeBox{fmin(a.x, b.x), fmin(a.y, b.y), fmax(a.x, b.x), fmax(a.y, b.y)};
}
__device__ __forceinline__ bool box_misses(const EdgeBox &e, pt b1, pt b2) {
  // bitwise, so that the eight compares combine without branches
  return ((b1.x < e.x0) & (b2.x < e.x0)) | ((b1.x > e.x1) & (b2.x > e.x1)) |
         ((b1.y < e.y0) & (b2.y < e.y0)) | ((b1.y > e.y1) & (b2.y > e.y1));
}
// f(q, B_q, B_q+1) for the n edges of ring V in order, vertex q + 2 loaded
// while edge q is tested
template <class F>
__device__ __forceinline__ void ring_edges(const double *V, int n, F &&f) {
  if (n <= 0) return;
  pt B1 = ld(V, 0), B2 = ld(V, n == 1 ? 0 : 1);
  int q = 0;
  for (; q + 2 < n; q++) {
    const pt B3 = ld(V, q + 2);
    f(q, B1, B2);
    B1 = B2;
    B2 = B3;
  }
  for (; q < n; q++) {
    const pt B3 = ld(V, (q + 2) % n);
    f(q, B1, B2);
    B1 = B2;
    B2 = B3;
  }
}

// ----------------------------------------------------------- run boxes ----
// Boxes of runs of kRun1, kRun2 = 16 kRun1 and kRun3 = 16 kRun2 consecutive
// edges in global edge order (a run spanning two pairs covers both, which
// only loosens it).  Rings longer than the culling threshold (RunBoxes::
// cullMin, set from Tuning::cullMin) are walked through them: a
// run whose box misses the edge's box holds no edge that lsmf could accept,
// and the edges of the other runs are still visited in index order.
constexpr int kRun1 = 16, kRun2 = 256, kRun3 = 4096;

struct RunBoxes {
  const double *b1, *b2, *b3;  // [4 * runs] x0, y0, x1, y1
  const double *l1, *l2, *l3;  // [runs] shortest edge of the run
  long cullMin;                // rings longer than this are walked through the runs
};

// one thread per edge; the kRun1 = 16 lanes of a run combine their boxes
__global__ void __launch_bounds__(256)
    k_run_boxes(long edges, const int *map, const long long *off, const double *ring,
                double *box1, double *len1) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, lmin = INFINITY;
  bool nan = false;
  if (e < edges) {
    const int k = map[e];
    const long long b = off[k], n = off[k + 1] - b;
    const pt p = ld(ring, e), q = ld(ring, e - b == n - 1 ? b : e + 1);
    nan = (p.x != p.x) | (p.y != p.y) | (q.x != q.x) | (q.y != q.y);
    x0 = fmin(p.x, q.x);
    y0 = fmin(p.y, q.y);
    x1 = fmax(p.x, q.x);
    y1 = fmax(p.y, q.y);
    const double dx = q.x - p.x, dy = q.y - p.y;
    lmin = sqrt(dx * dx + dy * dy);
  }
  for (int off2 = 1; off2 < kRun1; off2 <<= 1) {
    x0 = fmin(x0, __shfl_xor(x0, off2, 64));
    y0 = fmin(y0, __shfl_xor(y0, off2, 64));
    x1 = fmax(x1, __shfl_xor(x1, off2, 64));
    y1 = fmax(y1, __shfl_xor(y1, off2, 64));
    lmin = fmin(lmin, __shfl_xor(lmin, off2, 64));
    nan |= __shfl_xor((int)nan, off2, 64) != 0;
  }
  const long G = e / kRun1;
  if ((e & (kRun1 - 1)) != 0 || e >= edges) return;
  if (nan) {
    // a run with a NaN coordinate is never left out
    x0 = y0 = -INFINITY;
    x1 = y1 = INFINITY;
    lmin = 0.0;
  }
  box1[4 * G] = x0;
  box1[4 * G + 1] = y0;
  box1[4 * G + 2] = x1;
  box1[4 * G + 3] = y1;
  len1[G] = lmin;
}

// boxes of the next level: 16 boxes merged
__global__ void k_run_boxes_merge(long runsIn, const double *in, const double *lenIn,
                                  double *out, double *lenOut) {
  const long S = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long g0 = S * 16;
  if (g0 >= runsIn) return;
  double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY, lmin = INFINITY;
  for (long g = g0; g < g0 + 16 && g < runsIn; g++) {
    x0 = fmin(x0, in[4 * g]);
    y0 = fmin(y0, in[4 * g + 1]);
    x1 = fmax(x1, in[4 * g + 2]);
    y1 = fmax(y1, in[4 * g + 3]);
    lmin = fmin(lmin, lenIn[g]);
  }
  out[4 * S] = x0;
  out[4 * S + 1] = y0;
  out[4 * S + 2] = x1;
  out[4 * S + 3] = y1;
  lenOut[S] = lmin;
}

__device__ __forceinline__ bool box_hit(const double *b, const EdgeBox &e) {
  return !(e.x0 > b[2] || e.x1 < b[0] || e.y0 > b[3] || e.y1 < b[1]);
}

// A run is visited when its box meets ab, or when one of its edges may form
// a pair shorter than the tiny threshold T with an edge of length la (T = 0
// for box culling alone)
__device__ __forceinline__ bool run_hit(const double *b, const double *lmin, const EdgeBox &e,
                                        double la, double T) {
  return box_hit(b, e) || (T > 0.0 && la * *lmin < T);
}

// f(q, B_q, B_q+1) over the edges of the ring at global edges [e0, e1)
// (vertices V) whose runs pass run_hit, in index order
template <class F>
__device__ __forceinline__ long long culled_edges(const double *V, long long e0, long long e1,
                                                  const RunBoxes &rb, const EdgeBox &ab,
                                                  double la, double T, F &&f) {
  const long long n = e1 - e0;
  long long visited = 0;
  for (long long S3 = e0 / kRun3; S3 * kRun3 < e1; S3++) {
    if (!run_hit(rb.b3 + 4 * S3, rb.l3 + S3, ab, la, T)) continue;
    const long long a2 = max(S3 * 16, e0 / kRun2), z2 = min(S3 * 16 + 16, (e1 - 1) / kRun2 + 1);
    for (long long S2 = a2; S2 < z2; S2++) {
      if (!run_hit(rb.b2 + 4 * S2, rb.l2 + S2, ab, la, T)) continue;
      const long long a1 = max(S2 * 16, e0 / kRun1), z1 = min(S2 * 16 + 16, (e1 - 1) / kRun1 + 1);
      for (long long S1 = a1; S1 < z1; S1++) {
        if (!run_hit(rb.b1 + 4 * S1, rb.l1 + S1, ab, la, T)) continue;
        const long long a = max(S1 * kRun1, e0), z = min(S1 * kRun1 + kRun1, e1);
        visited += z - a;
        for (long long e = a; e < z; e++) {
          const long long q = e - e0;
          f((int)q, ld(V, q), ld(V, q + 1 == n ? 0 : q + 1));
        }
      }
    }
  }
  return visited;
}

// The count pass records, for every edge, the first kPartners edges of the
// other ring with a nonzero intersection type, in index order; the neighbour
// map and intersection passes act only on those edges, so they walk that
// list instead of the whole ring when the edge has no more than kPartners
// of them (its count1).
constexpr int kPartners = 16;

// f(q, B_q, B_q+1) for the listed edges lst[0, c) of ring V (n vertices)
constexpr int kListStage = 4;
template <class F>
__device__ __forceinline__ void listed_edges(const double *V, int n, const int *lst, int c,
                                             F &&f) {
  // kListStage entries a round: their indices, then their vertices, are
  // loaded before the round's visits (which write neither V nor lst), so a
  // round waits on memory twice rather than twice per entry; the visits
  // keep the list's order
  constexpr int U = kListStage;
  for (int i0 = 0; i0 < c; i0 += U) {
    int q[U];
    pt b1[U], b2[U];
#pragma unroll
    for (int u = 0; u < U; u++) q[u] = i0 + u < c ? lst[i0 + u] : 0;
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (i0 + u < c) {
        b1[u] = ld(V, q[u]);
        b2[u] = ld(V, q[u] + 1 == n ? 0 : q[u] + 1);
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (i0 + u < c) f(q[u], b1[u], b2[u]);
  }
}

// ------------------------------------------------------------- count ----
// The count kernels' statistics (edge pairs classified, edge pairs tested):
// lane 0 of wave w stores its wave's sums at wstat[2w], wstat[2w + 1] with
// plain stores, and k_sum_wave_stats adds them up.  (One device-scope
// atomic per wave on one address serialised at the memory side: measured
// on MI355X at about 0.4 ms for the 36 000 waves of the join workload's
// count launch.)
__device__ __forceinline__ void wave_stats(unsigned long long *wstat, long g,
                                           unsigned long long ev, unsigned long long tv) {
  for (int off = 32; off > 0; off >>= 1) {
    ev += __shfl_xor(ev, off, 64);
    tv += __shfl_xor(tv, off, 64);
  }
  if (__lane_id() == 0) {
    wstat[2 * (g >> 6)] = ev;
    wstat[2 * (g >> 6) + 1] = tv;
  }
}

// stats[0] += sum of wstat[2w], stats[2] += sum of wstat[2w + 1], w < waves;
// a grid of up to kStatBlocks blocks, one atomic per block and statistic
constexpr int kStatBlocks = 128;
__global__ void __launch_bounds__(256)
    k_sum_wave_stats(const unsigned long long *wstat, long waves, unsigned long long *stats) {
  __shared__ unsigned long long se[4], st[4];
  unsigned long long e = 0, t = 0;
  for (long w = (long)blockIdx.x * blockDim.x + threadIdx.x; w < waves;
       w += (long)gridDim.x * blockDim.x) {
    e += wstat[2 * w];
    t += wstat[2 * w + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off, 64);
    t += __shfl_xor(t, off, 64);
  }
  if (__lane_id() == 0) {
    se[threadIdx.x >> 6] = e;
    st[threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long be = se[0] + se[1] + se[2] + se[3];
    const unsigned long long bt = st[0] + st[1] + st[2] + st[3];
    if (be) atomicAdd(stats, be);
    if (bt) atomicAdd(stats + 2, bt);
  }
}

static void sum_wave_stats(Context &c, const unsigned long long *wstat, long waves,
                           unsigned long long *stats) {
  const long blocks = std::min<long>(kStatBlocks, (waves + 255) / 256);
  hipLaunchKernelGGL(k_sum_wave_stats, dim3((unsigned)blocks), dim3(256), 0, c.stream, wstat,
                     waves, stats);
  HIP_CHECK(hipGetLastError());
}

// gpuCMBRFilter (:382-415) + gpuCountIntersections (:463-548), both views in
// one launch: threads [0, edgesP) count P edges against Q, the rest count Q
// edges against P.  CULL selects the run-box traversal for rings longer than
// the culling threshold; the plain variant keeps the register budget of the
// short-ring walk.
template <bool CULL>
__device__ __forceinline__ void count_edge(long g, long edgesP, long edgesQ, const int *mapP,
                                           const int *mapQ, const long long *offP,
                                           const long long *offQ, const double *ringP,
                                           const double *ringQ, const double *cmbrN, int *cnt1P,
                                           int *cnt2P, int *cnt1Q, int *cnt2Q, int *boolP,
                                           int *boolQ, int *lstP, int *lstQ, const RunBoxes &boxP,
                                           const RunBoxes &boxQ, unsigned long long &ev,
                                           unsigned long long &tv) {
  const bool sideP = g < edgesP;
  const long ga = sideP ? g : g - edgesP;
  const int k = sideP ? mapP[ga] : mapQ[ga];
  const long long aBase = sideP ? offP[k] : offQ[k];
  const long long aN = (sideP ? offP[k + 1] : offQ[k + 1]) - aBase;
  const long long bBase = sideP ? offQ[k] : offP[k];
  const long long bN = (sideP ? offQ[k + 1] : offP[k + 1]) - bBase;
  const double *A = sideP ? ringP : ringQ;
  const double *Bv = (sideP ? ringQ : ringP) + 2 * bBase;
  const long long i = ga - aBase;
  const pt A1 = ld(A, ga), A2 = ld(A, i == aN - 1 ? aBase : ga + 1);
  const double cx1 = cmbrN[4 * k], cy1 = cmbrN[4 * k + 1], cx2 = cmbrN[4 * k + 2],
               cy2 = cmbrN[4 * k + 3];
  const double mnX = fmin(A1.x, A2.x), mxX = fmax(A1.x, A2.x);
  const double mnY = fmin(A1.y, A2.y), mxY = fmax(A1.y, A2.y);
  bool flag = !(mnX > cx2 || mxX < cx1) && !(mnY > cy2 || mxY < cy1);
  int c1 = 0, c2 = 0;
  if (flag) {
    const EdgeBox ab = edge_box(A1, A2);
    int *lst = (sideP ? lstP : lstQ) + (long long)ga * kPartners;
    auto visit = [&](int q, pt B1, pt B2) {
      if (box_misses(ab, B1, B2) || !lsmf(A1, A2, B1, B2)) return;
      double al, be;
      ev++;
      const int t = itype(A1, A2, B1, B2, al, be);
      if (t != 0) {
        if (c1 < kPartners) lst[c1] = q;
        c1++;
      }
      c2 += adds_vertex(t);
    };
    if (CULL && bN > boxP.cullMin) {
      tv = culled_edges(Bv, bBase, bBase + bN, sideP ? boxQ : boxP, ab, 0.0, 0.0, visit);
    } else {
      ring_edges(Bv, (int)bN, visit);
      tv = bN;
    }
  }
  if (sideP) {
    cnt1P[ga] = c1;
    cnt2P[ga] = c2 + 1;
    boolP[ga] = flag;
  } else {
    cnt1Q[ga] = c1;
    cnt2Q[ga] = c2 + 1;
    boolQ[ga] = flag;
  }
}

template <bool CULL>
__global__ void __launch_bounds__(256)
    k_count(long edgesP, long edgesQ, const int *mapP, const int *mapQ, const long long *offP,
            const long long *offQ, const double *ringP, const double *ringQ,
            const double *cmbrN, int *cnt1P, int *cnt2P, int *cnt1Q, int *cnt2Q, int *boolP,
            int *boolQ, unsigned long long *wstat, int *lstP, int *lstQ, RunBoxes boxP,
            RunBoxes boxQ) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long ev = 0, tv = 0;
  if (g < edgesP + edgesQ)
    count_edge<CULL>(g, edgesP, edgesQ, mapP, mapQ, offP, offQ, ringP, ringQ, cmbrN, cnt1P, cnt2P,
                     cnt1Q, cnt2Q, boolP, boolQ, lstP, lstQ, boxP, boxQ, ev, tv);
  // every lane of every wave reaches the statistics
  wave_stats(wstat, g, ev, tv);
}

// Partner lists at the count1 offsets for the active edges with at most
// kPartners partners: the entries the count pass recorded (the others are
// rescanned by k_partner_rescan*, in the same order).
__global__ void k_partner_lists(long nAct, const int *act, const long long *psA1,
                                const int *lstFixed, int *lst) {
  const long ia = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= nAct) return;
  const long g = act[ia];
  const long long c1 = psA1[g + 1] - psA1[g];
  int *dst = lst + psA1[g];
  if (c1 <= kPartners)
    for (int i = 0; i < (int)c1; i++) dst[i] = lstFixed[(long long)g * kPartners + i];
}

// Partner lists of the edges with more than kPartners partners: one wave
// per edge rescans the other ring with the count pass's tests, 64 edges at a
// time, and appends the hits in ring order.
__global__ void __launch_bounds__(64)
    k_partner_rescan(long nOver, const int *listO, const int *mapA, const long long *offA,
                     const long long *offB, const double *ringA, const double *ringB,
                     const long long *psA1, int *lst, RunBoxes rb) {
  for (long w = blockIdx.x; w < nOver; w += gridDim.x) {
    const long g = listO[w];
    const int k = mapA[g];
    const long long c1 = psA1[g + 1] - psA1[g];
    int *dst = lst + psA1[g];
    const long long aBase = offA[k], aN = offA[k + 1] - aBase;
    const long long bBase = offB[k], bN = offB[k + 1] - bBase;
    const long long i = g - aBase;
    const pt A1 = ld(ringA, g), A2 = ld(ringA, i == aN - 1 ? aBase : g + 1);
    const EdgeBox ab = edge_box(A1, A2);
    const double *V = ringB + 2 * bBase;
    long long n = 0;
    // edges [a, z) of the ring, 64 at a time, hits appended in order
    auto scan = [&](long long a, long long z) {
      for (long long q0 = a; q0 < z; q0 += 64) {
        const long long q = q0 + __lane_id();
        bool hit = false;
        if (q < z) {
          const pt B1 = ld(V, q), B2 = ld(V, q + 1 == bN ? 0 : q + 1);
          if (!box_misses(ab, B1, B2) && lsmf(A1, A2, B1, B2)) {
            double al, be;
            hit = itype(A1, A2, B1, B2, al, be) != 0;
          }
        }
        const unsigned long long m = __ballot(hit);
        const long long r = n + __popcll(m & ((1ULL << __lane_id()) - 1ULL));
        if (hit && r < c1) dst[r] = (int)q;
        n += __popcll(m);
      }
    };
    if (bN <= rb.cullMin) {
      scan(0, bN);
    } else {
      const long long e0 = bBase, e1 = bBase + bN;
      for (long long S3 = e0 / kRun3; S3 * kRun3 < e1; S3++) {
        if (!box_hit(rb.b3 + 4 * S3, ab)) continue;
        const long long a2 = max(S3 * 16, e0 / kRun2),
                        z2 = min(S3 * 16 + 16, (e1 - 1) / kRun2 + 1);
        for (long long S2 = a2; S2 < z2; S2++)
          if (box_hit(rb.b2 + 4 * S2, ab))
            scan(max(S2 * kRun2, e0) - e0, min(S2 * kRun2 + kRun2, e1) - e0);
      }
    }
    // entries the rescan did not fill (none when it agrees with the count
    // pass) name edge 0, so that no later pass reads an unset index
    for (long long r = n + __lane_id(); r < c1; r += 64) dst[r] = 0;
  }
}

// per pair: invalid check (:1460-1476) and slot bases
__global__ void k_pair_check(long M, const long long *offP, const long long *offQ,
                             const long long *psP1, const long long *psQ1,
                             const long long *psP2, const long long *psQ2, int *status,
                             unsigned long long *nInter) {
  __shared__ unsigned long long part[4];
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (k < M && status[k] == GH_OK) {
    const long long c1P = psP1[offP[k + 1]] - psP1[offP[k]];
    const long long c1Q = psQ1[offQ[k + 1]] - psQ1[offQ[k]];
    if (c1P != c1Q) status[k] = GH_INVALID;
    else n = (unsigned long long)c1P;
  }
  // one atomic per block (256 threads)
  for (int off = 32; off > 0; off >>= 1) n += __shfl_xor(n, off, 64);
  if (__lane_id() == 0) part[threadIdx.x >> 6] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(nInter, t);
  }
}

// ----------------------------------------------------- neighbour map ----
// gpuNeighborMap (:549-624): for each Q edge, the P edge of its k-th
// vertex-adding intersection (degenerate types use the parent slot).
__global__ void __launch_bounds__(256)
    k_nmap(long edgesQ, const int *act, const int *mapQ, const long long *offP,
           const long long *offQ,
           const double *ringP, const double *ringQ, const long long *psP1,
           const long long *psQ1, const long long *psQ2, const int *status, int *nmap,
           const int *lstQ) {
  // edgesQ: number of active Q edges, listed in act
  const long ia = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= edgesQ) return;
  const long g = act[ia];
  const int k = mapQ[g];
  if (status[k] != GH_OK) return;
  if (psQ1[g + 1] == psQ1[g]) return;
  const long long qBase = offQ[k], qN = offQ[k + 1] - qBase;
  const long long pBase = offP[k], pN = offP[k + 1] - pBase;
  const long long i = g - qBase;
  const pt A1 = ld(ringQ, g), A2 = ld(ringQ, i == qN - 1 ? qBase : g + 1);
  const long long slot = psQ2[g];
  int c2 = 0, nd = 0;
  const EdgeBox ab = edge_box(A1, A2);
  auto visit = [&](int p, pt b1, pt b2) {
    const long long pg = pBase + p;
    if (box_misses(ab, b1, b2) || psP1[pg + 1] == psP1[pg]) return;
    if (!lsmf(A1, A2, b1, b2)) return;
    double al, be;
    const int t = itype(A1, A2, b1, b2, al, be);
    if (t) {
      if (adds_vertex(t)) c2 = ++nd;
      else c2 = 0;
      nmap[slot + c2] = p;
    }
  };
  listed_edges(ringP + 2 * pBase, (int)pN, lstQ + psQ1[g], (int)(psQ1[g + 1] - psQ1[g]), visit);
}

// ----------------------------------------------------- intersections ----
struct SlotArrays {
  double *iP, *iQ, *oP, *oQ;      // normalised / original coordinates
  double *iP2, *iQ2, *oP2, *oQ2;  // unsorted copies (sort sources)
  int *aP, *aQ, *aP2, *aQ2;       // trunc(1e7 alpha), -100 none
  int *nbP, *nbQ, *nbP2, *nbQ2;   // global slot + 1, 0 none
  int *lP, *lQ;                   // initial labels
};

// gpuRoute: also the arrays only the GPU route reads (its sort sources aP2,
// aQ2, nbP2, nbQ2; the neighbour map is exported with the phase-1 arrays,
// so it is cleared on both routes)
__global__ void k_init_slots(long slotsP, long slotsQ, SlotArrays s, int *nmap, int gpuRoute) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < slotsP) {
    s.aP[i] = kNone;
    s.nbP[i] = 0;
    s.lP[i] = kNone;
    if (gpuRoute) {
      s.aP2[i] = kNone;
      s.nbP2[i] = 0;
    }
  }
  if (i < slotsQ) {
    s.aQ[i] = kNone;
    s.nbQ[i] = 0;
    s.lQ[i] = kNone;
    nmap[i] = kNone;
    if (gpuRoute) {
      s.aQ2[i] = kNone;
      s.nbQ2[i] = 0;
    }
  }
}

// Q parents: slot psQ2[g] holds Q vertex g (:930-935)
__global__ void k_q_parents(long edgesQ, const int *mapQ, const long long *offQ,
                            const double *ringQ, const long long *psQ2, const int *status,
                            PairMeta pm, const double *bCoords, const double *oCoords,
                            SlotArrays s) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= edgesQ) return;
  const int k = mapQ[g];
  if (status[k] != GH_OK) return;
  const long long i = g - offQ[k];
  const double *cQ = pm.swapped[k] ? bCoords : oCoords;
  const long long sl = psQ2[g];
  // (the sort sources iQ2 / oQ2 are read for intersection slots only)
  st(s.iQ, sl, ld(ringQ, g));
  st(s.oQ, sl, ld(cQ, pm.startQ[k] + i));
}

// k_p_parents and k_q_parents in one launch: threads [0, edgesP) take the P
// edges, the rest the Q edges
__global__ void k_parents(long edgesP, long edgesQ, const int *mapP, const int *mapQ,
                          const long long *offP, const long long *offQ, const double *ringP,
                          const double *ringQ, const long long *psP2, const long long *psQ2,
                          const int *status, PairMeta pm, const double *bCoords,
                          const double *oCoords, SlotArrays s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= edgesP + edgesQ) return;
  const bool sideP = t < edgesP;
  const long g = sideP ? t : t - edgesP;
  const int k = (sideP ? mapP : mapQ)[g];
  if (status[k] != GH_OK) return;
  const long long i = g - (sideP ? offP : offQ)[k];
  // P's coordinates are the overlay's when the pair is swapped, Q's the base's
  const double *cA = (pm.swapped[k] != 0) == sideP ? oCoords : bCoords;
  const long long sl = (sideP ? psP2 : psQ2)[g];
  st(sideP ? s.iP : s.iQ, sl, ld(sideP ? ringP : ringQ, g));
  st(sideP ? s.oP : s.oQ, sl, ld(cA, (sideP ? pm.startP[k] : pm.startQ[k]) + i));
}

// Stable order of keys[lo..hi) by value: dest[j] = #{i : k_i < k_j} +
// #{i < j : k_i == k_j}, which equals the reference's 7-digit decimal LSD
// counting sort (:263-305) on these non-negative keys.
__device__ __forceinline__ long long stable_dest(const int *keys, long long lo, long long hi,
                                                 long long j) {
  const int kj = keys[j];
  long long d = 0;
  for (long long i = lo; i < hi; i++) {
    int ki = keys[i];
    d += (ki < kj) || (ki == kj && i < j);
  }
  return d;
}

// First slot of [start, end) (a Q edge's slots) whose neighbour-map entry is
// P edge pid, or -1.  After the parent slot the map holds the P edges of the
// vertex-adding intersections in increasing order, one slot each, written
// from start + 1 on, so the search is a lower bound with unwritten (kNone)
// slots, which can only trail, taken as +infinity.
__device__ __forceinline__ long long find_slot(const int *nmap, long long start, long long end,
                                               int pid) {
  if (nmap[start] == pid) return start;
  long long lo = start + 1, hi = end;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    const int v = nmap[mid];
    if ((v == kNone ? 0x7fffffff : v) < pid) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && nmap[lo] == pid) ? lo : -1;
}

// Several P edges may write the same Q slot (degenerate configurations, and
// the reference's neighbour index that persists across Q edges when a search
// fails).  The reference resolves these races arbitrarily; here the writes of
// the highest P edge win, which is the order of a sequential loop over P
// edges.  OWN pass: record that writer per Q slot and field group with
// atomicMax; WRITE pass: only the recorded writer stores.
struct Owners {
  int *a;  // aQ / aQ2
  int *c;  // iQ, oQ and copies
  int *n;  // nbQ / nbQ2
};

// P parents: slot psP2[g] holds P vertex g (:897-1123, the parent store of
// gpuCalculateIntersections)
__global__ void k_p_parents(long edgesP, const int *mapP, const long long *offP,
                            const double *ringP, const long long *psP2, const int *status,
                            PairMeta pm, const double *bCoords, const double *oCoords,
                            SlotArrays s) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= edgesP) return;
  const int k = mapP[g];
  if (status[k] != GH_OK) return;
  const double *cP = pm.swapped[k] ? oCoords : bCoords;
  const long long sl = psP2[g];
  // (the sort sources iP2 / oP2 are read for intersection slots only)
  st(s.iP, sl, ld(ringP, g));
  st(s.oP, sl, ld(cP, pm.startP[k] + (g - offP[k])));
}

// gpuCalculateIntersections (:897-1123) without the per-edge sort, over the
// P edges with intersections
template <bool OWN>
__global__ void __launch_bounds__(256)
    k_calc(long edgesP, const int *act, const int *mapP, const long long *offP,
           const long long *offQ,
           const double *ringP, const double *ringQ, const long long *psP1,
           const long long *psP2, const long long *psQ1, const long long *psQ2,
           const int *status, PairMeta pm, const double *bCoords, const double *oCoords,
           const int *nmap, SlotArrays s, Owners own, const int *lstP) {
  // edgesP: number of active P edges, listed in act
  const long ia = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= edgesP) return;
  const long g = act[ia];
  const int k = mapP[g];
  if (status[k] != GH_OK) return;
  const long long pBase = offP[k], pN = offP[k + 1] - pBase;
  const long long qBase = offQ[k], qN = offQ[k + 1] - qBase;
  const int pid = (int)(g - pBase);
  const int sw = pm.swapped[k];
  const double *cP = sw ? oCoords : bCoords, *cQ = sw ? bCoords : oCoords;
  const long sP = pm.startP[k], sQ = pm.startQ[k];
  const long long base = psP2[g];
  const pt P1 = ld(ringP, g), P2 = ld(ringP, pid == pN - 1 ? pBase : g + 1);
  const pt P1o = ld(cP, sP + pid), P2o = ld(cP, sP + (pid == pN - 1 ? 0 : pid + 1));
  if (psP1[g + 1] == psP1[g]) return;
  // Q-slot stores (conditional on ownership in the WRITE pass)
  auto setA = [&](long long sl, int v) {
    if (OWN) atomicMax(&own.a[sl], pid);
    else if (own.a[sl] == pid) {
      s.aQ[sl] = v;
      s.aQ2[sl] = v;
    }
  };
  auto setC = [&](long long sl, pt v, pt vo) {
    if (OWN) atomicMax(&own.c[sl], pid);
    else if (own.c[sl] == pid) {
      st(s.iQ, sl, v);
      st(s.iQ2, sl, v);
      st(s.oQ, sl, vo);
      st(s.oQ2, sl, vo);
    }
  };
  auto setN = [&](long long sl, int v) {
    if (OWN) atomicMax(&own.n[sl], pid);
    else if (own.n[sl] == pid) {
      s.nbQ[sl] = v;
      s.nbQ2[sl] = v;
    }
  };
  auto setP = [&](long long sl, int nb) {
    if (!OWN) {
      s.nbP[sl] = nb;
      s.nbP2[sl] = nb;
    }
  };
  int c2 = 0, nd = 0;
  long long nQId = -1;
  const EdgeBox ab = edge_box(P1, P2);
  auto visit = [&](int q, pt Q1, pt Q2) {
    const long long qg = qBase + q;
    if (box_misses(ab, Q1, Q2) || psQ1[qg + 1] == psQ1[qg]) return;
    if (!lsmf(P1, P2, Q1, Q2)) return;
    double alpha, beta;
    const int t = itype(P1, P2, Q1, Q2, alpha, beta);
    if (!t) return;
    if (adds_vertex(t)) c2 = ++nd;
    else c2 = 0;
    const long long start = psQ2[qg], end = psQ2[qg + 1];
    const long long ps = base + c2;
    const long long l = find_slot(nmap, start, end, pid);
    if (t != 5) {
      if (l >= 0) {
        nQId = l;
        setP(ps, (int)(l + 1));
        setN(l, (int)(ps + 1));
      }
    } else {
      nQId = start;
      setP(ps, (int)(start + 1));
      setN(start, (int)(ps + 1));
      if (l >= 0) {
        nQId = l;
        setP(base, (int)(l + 1));
        setN(l, (int)(base + 1));
      }
    }
    const int ka = trunc_sat((double)kAlphaScale * alpha);
    const int kb = trunc_sat((double)kAlphaScale * beta);
    const pt Q1o = ld(cQ, sQ + q);
    // P-side values of this record
    bool pAdds = t == 1 || t == 5 || t == 3 || t == 7;
    if (!OWN) {
      if (pAdds) {
        pt v = Q1, vo = Q1o;
        if (t == 1) {
          v = pt{(1.0 - alpha) * P1.x + alpha * P2.x, (1.0 - alpha) * P1.y + alpha * P2.y};
          vo = pt{(1.0 - alpha) * P1o.x + alpha * P2o.x, (1.0 - alpha) * P1o.y + alpha * P2o.y};
        }
        st(s.iP, ps, v);
        st(s.iP2, ps, v);
        st(s.oP, ps, vo);
        st(s.oP2, ps, vo);
        s.aP[ps] = ka;
        s.aP2[ps] = ka;
      } else {
        s.aP[base] = ka;
        s.aP2[base] = ka;
      }
    }
    // Q-side values
    if (t == 1 || t == 5 || t == 2 || t == 6) {
      if (nQId >= 0) {
        if (t == 1) {
          const pt I = {(1.0 - alpha) * P1.x + alpha * P2.x, (1.0 - alpha) * P1.y + alpha * P2.y};
          const pt Io = {(1.0 - alpha) * P1o.x + alpha * P2o.x,
                         (1.0 - alpha) * P1o.y + alpha * P2o.y};
          setC(nQId, I, Io);
        } else {
          setC(nQId, P1, P1o);
        }
        setA(nQId, kb);
      }
    } else {
      setA(start, kb);
    }
  };
  listed_edges(ringQ + 2 * qBase, (int)qN, lstP + psP1[g], (int)(psP1[g + 1] - psP1[g]), visit);
}

// ------------------------------------------------ per-edge alpha sorts ----
// The intersection slots after each edge's parent slot are put in order of
// their truncated alpha, stably (the order stable_dest defines).  Runs of up
// to kSmallSort slots are ranked by a 32-lane group, runs up to kBlockSort by
// one block, longer runs by one segmented radix sort keyed (run, alpha) in
// slot order, which gives the same stable order.
constexpr int kSmallSort = 32;

// move of P slot j to sorted position d (calculateIntersections :1103-1120);
// the Q slot's back link follows when edge pid owns it
__device__ __forceinline__ void move_p(SlotArrays &s, const Owners &own, long long j,
                                       long long d, int pid) {
  const int nb = s.nbP2[j];
  s.aP[d] = s.aP2[j];
  st(s.iP, d, ld(s.iP2, j));
  st(s.oP, d, ld(s.oP2, j));
  s.nbP[d] = nb;
  if (nb > 0 && own.n[nb - 1] == pid) {
    s.nbQ[nb - 1] = (int)(d + 1);
    s.nbQ2[nb - 1] = (int)(d + 1);
  }
}

// move of Q slot j to sorted position d (gpuSortPolyQ :1132-1160); the P
// slot's back link follows when Q edge g owns it (the highest Q edge writing
// it, the order of a sequential loop over Q edges)
__device__ __forceinline__ void move_q(SlotArrays &s, const int *ownP, long long j, long long d,
                                       long g) {
  const int nb = s.nbQ2[j];
  s.aQ[d] = s.aQ2[j];
  st(s.iQ, d, ld(s.iQ2, j));
  st(s.oQ, d, ld(s.oQ2, j));
  s.nbQ[d] = nb;
  if (nb > 0 && ownP[nb - 1] == (int)g) s.nbP[nb - 1] = (int)(d + 1);
}

// Runs of kSmallSort < m <= kBlockSort slots are ranked by one block each
// from a copy of their keys in LDS; longer ones go through the radix sort.
constexpr int kBlockSort = 2048;

// One side's edges for k_list_long_runs: the short, block-sorted and long
// runs, and the edges with more intersecting partners than the count pass
// recorded, appended to listS / list / listL (+ lenL) / listO with the
// counts at cnt[1], cnt[0], cnt[2] (+ the long runs' total at cnt[3]),
// cnt[4].
struct RunListSide {
  long edges;
  const int *map;
  const int *cnt2, *cnt1;  // the count pass's count2 (1 + inserts) and count1
  int *listS, *list, *listL, *lenL, *listO;
  unsigned long long *cnt;
};
constexpr int kRunListEdges = 8;  // edges per thread of k_list_long_runs

// Both sides in one launch (blocks [0, blocksA) take side A), each block
// kRunListEdges x 1024 edges: one block-wide scan of the four counts packed
// in 16-bit fields and at most four atomics per block (one per non-empty
// list), so that the atomics on the shared counters stay few.  The lists'
// order is not used: every listed edge is worked on its own.
__global__ void __launch_bounds__(1024)
    k_list_long_runs(RunListSide A, RunListSide B, long blocksA, const int *status) {
  __shared__ unsigned long long lds[20];
  const bool sa = (long)blockIdx.x < blocksA;
  const RunListSide &S = sa ? A : B;
  const long blk = sa ? (long)blockIdx.x : (long)blockIdx.x - blocksA;
  const long g0 = blk * (1024L * kRunListEdges) + threadIdx.x;
  unsigned bits = 0;
  unsigned long long packed = 0;
#pragma unroll
  for (int r = 0; r < kRunListEdges; r++) {
    const long g = g0 + 1024L * r;
    const bool ok = g < S.edges && status[S.map[g]] == GH_OK;
    const long long m = ok ? S.cnt2[g] - 1 : 0;
    const bool pr[4] = {m >= 2 && m <= kSmallSort, m > kSmallSort && m <= kBlockSort,
                        m > kBlockSort, ok && S.cnt1[g] > kPartners};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      bits |= (unsigned)pr[i] << (4 * r + i);
      packed += (unsigned long long)pr[i] << (16 * i);
    }
  }
  // block-wide exclusive scan of the packed counts (a block's total per
  // field is at most 8192)
  const int lane = __lane_id(), wave = threadIdx.x >> 6;
  unsigned long long inc = packed;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const unsigned long long v = __shfl_up(inc, off, 64);
    if (lane >= off) inc += v;
  }
  if (lane == 63) lds[wave] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < 16; w++) {
      const unsigned long long c = lds[w];
      lds[w] = t;
      t += c;
    }
    unsigned long long *const cs[4] = {S.cnt + 1, S.cnt, S.cnt + 2, S.cnt + 4};
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const unsigned long long ti = (t >> (16 * i)) & 0xffffULL;
      lds[16 + i] = ti ? atomicAdd(cs[i], ti) : 0ULL;
    }
  }
  __syncthreads();
  unsigned long long run = lds[wave] + inc - packed;
  if (!bits) return;
  // edges by run length m (inserted vertices): short runs have 2 <= m <=
  // kSmallSort, block-sorted ones up to kBlockSort, long ones more
  int *const out[4] = {S.listS, S.list, S.listL, S.listO};
#pragma unroll
  for (int r = 0; r < kRunListEdges; r++) {
    const long g = g0 + 1024L * r;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (!((bits >> (4 * r + i)) & 1u)) continue;
      const long long pos = (long long)lds[16 + i] + (long long)((run >> (16 * i)) & 0xffffULL);
      run += 1ULL << (16 * i);
      out[i][pos] = (int)g;
      if (i == 2) {
        const long long m = S.cnt2[g] - 1;
        S.lenL[pos] = (int)m;
        atomicAdd(S.cnt + 3, (unsigned long long)m);
      }
    }
  }
}

// k_list_long_runs over both sides (a side with no edges takes no block)
static void list_runs(Context &c, const RunListSide &A, const RunListSide &B,
                      const int *status) {
  const long per = 1024L * kRunListEdges;
  const long bA = (A.edges + per - 1) / per, bB = (B.edges + per - 1) / per;
  if (bA + bB == 0) return;
  hipLaunchKernelGGL(k_list_long_runs, dim3((unsigned)(bA + bB)), dim3(1024), 0, c.stream, A, B,
                     bA, status);
  HIP_CHECK(hipGetLastError());
}

// OWN pass of the Q sort over listed runs, one block per run
__global__ void k_run_own(int nRuns, const int *list, const long long *psQ2, const int *nbQ2,
                          int *ownP) {
  for (int r = blockIdx.x; r < nRuns; r += gridDim.x) {
    const long g = list[r];
    const long long lo = psQ2[g] + 1, hi = psQ2[g + 1];
    for (long long j = lo + threadIdx.x; j < hi; j += blockDim.x) {
      const int nb = nbQ2[j];
      if (nb > 0) atomicMax(&ownP[nb - 1], (int)g);
    }
  }
}

// stable ranks of a medium run from its keys in LDS, then the moves
template <bool QSIDE>
__global__ void __launch_bounds__(256)
    k_run_sort_block(int nRuns, const int *list, const long long *ps2, const int *map,
                     const long long *off, SlotArrays s, Owners own, const int *ownP) {
  __shared__ int key[kBlockSort];
  const int *a2 = QSIDE ? s.aQ2 : s.aP2;
  for (int r = blockIdx.x; r < nRuns; r += gridDim.x) {
    const long g = list[r];
    const long long lo = ps2[g] + 1;
    const int m = (int)(ps2[g + 1] - lo);
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += blockDim.x) key[t] = a2[lo + t];
    __syncthreads();
    for (int t = threadIdx.x; t < m; t += blockDim.x) {
      const int kt = key[t];
      int d = 0;
      for (int i = 0; i < m; i++) {
        const int ki = key[i];
        d += (ki < kt) || (ki == kt && i < t);
      }
      if (QSIDE) move_q(s, ownP, lo + t, lo + d, g);
      else move_p(s, own, lo + t, lo + d, (int)(g - off[map[g]]));
    }
  }
}

// Runs of up to kSmallSort slots: 32 lanes per active edge, lane l ranking
// slot lo + l against the run's keys broadcast by shuffles (the rank of
// stable_dest).
__device__ __forceinline__ int short_rank(int key, int l, int m) {
  int d = 0;
  for (int i = 0; i < m; i++) {
    const int ki = __shfl(key, i, 32);
    d += (ki < key) || (ki == key && i < l);
  }
  return d;
}

__global__ void __launch_bounds__(256)
    k_sort_p_short(long nAct, const int *act, const int *mapP, const long long *offP,
                   const long long *psP2, const int *status, SlotArrays s, Owners own) {
  const long ia = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int l = threadIdx.x & 31;
  if (ia >= nAct) return;
  const long g = act[ia];
  const int k = mapP[g];
  if (status[k] != GH_OK) return;
  const long long lo = psP2[g] + 1;
  const int m = (int)(psP2[g + 1] - lo);
  if (m < 2 || m > kSmallSort) return;
  const int d = short_rank(l < m ? s.aP2[lo + l] : 0, l, m);
  if (l < m) move_p(s, own, lo + l, lo + d, (int)(g - offP[k]));
}

// OWN pass: record the highest Q edge writing each P slot's back link
template <bool OWN>
__global__ void __launch_bounds__(256)
    k_sort_q_short(long nAct, const int *act, const int *mapQ, const long long *psQ2,
                   const int *status, SlotArrays s, int *ownP) {
  const long ia = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int l = threadIdx.x & 31;
  if (ia >= nAct) return;
  const long g = act[ia];
  if (status[mapQ[g]] != GH_OK) return;
  const long long lo = psQ2[g] + 1;
  const int m = (int)(psQ2[g + 1] - lo);
  if (m < 2 || m > kSmallSort) return;
  if (OWN) {
    if (l < m) {
      const int nb = s.nbQ2[lo + l];
      if (nb > 0) atomicMax(&ownP[nb - 1], (int)g);
    }
    return;
  }
  const int d = short_rank(l < m ? s.aQ2[lo + l] : 0, l, m);
  if (l < m) move_q(s, ownP, lo + l, lo + d, g);
}

// keys of the long runs: (run index, alpha as an order-preserving u32), the
// slot as value; for Q runs also the OWN pass of their slots
template <bool QOWN>
__global__ void k_long_run_keys(int nRuns, const int *list, const long long *runOff,
                                const long long *ps2, const int *a2, const int *nb2, int *ownP,
                                uint64_t *keys, int *vals) {
  for (int r = blockIdx.x; r < nRuns; r += gridDim.x) {
    const long g = list[r];
    const long long lo = ps2[g] + 1, m = ps2[g + 1] - lo, o = runOff[r];
    for (long long t = threadIdx.x; t < m; t += blockDim.x) {
      const long long j = lo + t;
      keys[o + t] = ((uint64_t)r << 32) | (uint64_t)((unsigned)a2[j] ^ 0x80000000u);
      vals[o + t] = (int)j;
      if (QOWN) {
        const int nb = nb2[j];
        if (nb > 0) atomicMax(&ownP[nb - 1], (int)g);
      }
    }
  }
}

template <bool QSIDE>
__global__ void k_long_run_apply(long n, const uint64_t *keys, const int *vals, const int *list,
                                 const long long *runOff, const long long *ps2, const int *map,
                                 const long long *off, SlotArrays s, Owners own,
                                 const int *ownP) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = (int)(keys[i] >> 32);
  const long g = list[r];
  const long long d = ps2[g] + 1 + (i - runOff[r]);
  if (QSIDE) move_q(s, ownP, vals[i], d, g);
  else move_p(s, own, vals[i], d, (int)(g - off[map[g]]));
}

// gpuCalculateInitLabel (:1171-1223); circular ids within the pair's slots.
// OWN pass: the highest P slot writing a Q label owns it.
template <bool OWN>
__global__ void __launch_bounds__(256)
    k_init_label(long nAct, const int *act, const int *mapP, const long long *offP,
                 const long long *offQ, const long long *psP2, const long long *psQ2,
                 const int *status, SlotArrays s, int *ownL) {
  // 4 lanes per active P edge, one slot each in turn; the slots of inactive
  // edges keep their initial kNone
  const long ia = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 2;
  if (ia >= nAct) return;
  const long g = act[ia];
  const int k = mapP[g];
  if (status[k] != GH_OK) return;
  const long long sbP = psP2[offP[k]], cP = psP2[offP[k + 1]] - sbP;
  const long long sbQ = psQ2[offQ[k]], cQ = psQ2[offQ[k + 1]] - sbQ;
  for (long long i = psP2[g] + (threadIdx.x & 3); i < psP2[g + 1]; i += 4) {
    if (!OWN) s.lP[i] = kNone;
    if (s.aP[i] == kNone) continue;
    const long long nId = (long long)s.nbP[i] - 1;
    if (nId < sbQ || nId >= sbQ + cQ) continue;
    if (OWN) {
      atomicMax(&ownL[nId], (int)i);
      continue;
    }
    const pt cur = ld(s.iP, i);
    long long t = i - 1 < sbP ? sbP + cP - 1 : i - 1;
    const pt pM = ld(s.iP, t);
    const long long pMN = s.aP[t] != kNone ? (long long)s.nbP[t] - 1 : -100;
    t = i + 1 == sbP + cP ? sbP : i + 1;
    const pt pP = ld(s.iP, t);
    const long long pPN = s.aP[t] != kNone ? (long long)s.nbP[t] - 1 : -100;
    t = nId - 1 < sbQ ? sbQ + cQ - 1 : nId - 1;
    const int tM = rel_pos(pMN != -100 && pMN == t, pPN != -100 && pPN == t, ld(s.iQ, t), pM,
                           cur, pP);
    t = nId + 1 == sbQ + cQ ? sbQ : nId + 1;
    const int tP = rel_pos(pMN != -100 && pMN == t, pPN != -100 && pPN == t, ld(s.iQ, t), pM,
                           cur, pP);
    const int lab = initial_label(tM, tP);
    s.lP[i] = lab;
    if (ownL[nId] == (int)i) s.lQ[nId] = lab;
  }
}

// ================================================= CPU-route phase 1 ====
// computeIntersections (polyclip.h:188-281) classifies every pair (P edge p,
// Q edge q) once, as intersect(P_p, Q_q) (polyclip.h:60-179), and inserts
// the vertices the type adds with insertVertex (polygon.h:101-113): on a P
// edge in increasing alpha, on a Q edge in increasing beta, a later insertion
// before an earlier one of equal value.  The loop runs over P edges outside
// and Q edges inside, so ties come in decreasing partner order on both
// sides.  link() (polygon.h:89-94) sets both ends, so a source vertex keeps
// the partner of its last link in loop order: the highest Q edge for a P
// source, the highest P edge for a Q source.
//
// Edge-pair prefilter.  A pair is skipped only when intersect() returns 0 for
// it.  With L the edge lengths, u the unit roundoff, D the diameter of the
// pair's joint box (normalised coordinates) and eps = 1e-7:
//  - types 1-4 put the lines' intersection within (eps + eA) L of each
//    segment, where eA = 3e-8 D^2 + 1e-15 bounds the rounding of alpha, beta;
//  - types 5-8 need |A(P1,Q1,Q2)| < eps and |A(P1,Q1,Q2) - A(P2,Q1,Q2)| <=
//    eps, so P1 lies within (eps + dl) / L_Q and P2 within (2 eps + 3 dl) /
//    L_Q of Q's line (dl = 8 u D^2 bounds the rounding of A).  Types 5/6 put
//    P1's foot inside segment Q; types 7/8 put Q1's foot on P's line inside
//    segment P, and when L_P L_Q >= T = 4.5 (eps + 2 dl) the two lines meet at
//    a sine <= 1/4, which puts Q1 within 2.2 (eps + 2 dl) / L_Q of segment P.
// Hence the Chebyshev gap of the closed edge boxes of a nonzero pair is at
// most mP + mQ, mP = kA L_P + c0, mQ = kA L_Q + kB / L_Q + c0, with kA = 1.5
// (eps + eA), kB = 1.5 * 2.2 (eps + 2 dl), c0 = 16 u D, unless L_P L_Q < T,
// in which case the pair is classified unless |A(P1,Q1,Q2)| >= eps rules the
// collinear branch out.  Every reject, including the cheaper ring-level ones
// below, is of a pair whose type is provably 0, so both views find exactly
// the pairs with a nonzero type and their partner sets agree.

// per-pair filter constants (cf[16 k + i])
enum { CF_BOXP = 0, CF_BOXQ = 4, CF_MAXMP = 8, CF_MAXMQ = 9, CF_MINLP = 10, CF_MINLQ = 11,
       CF_TINY = 12, CF_KA = 13, CF_KB = 14, CF_C0 = 15, CF_N = 16 };

// The filter's constants from the two rings' boxes (normalised frame):
// D is the diagonal of their union, dl and eA bound the rounding of the
// predicates, kA / kB / c0 build an edge's margin from its length.
struct FilterK {
  double kA, kB, c0, dl;
};
__device__ __forceinline__ FilterK filter_consts(const double bx[2][4]) {
  const double w = fmax(bx[0][2], bx[1][2]) - fmin(bx[0][0], bx[1][0]);
  const double h = fmax(bx[0][3], bx[1][3]) - fmin(bx[0][1], bx[1][1]);
  const double D2 = w * w + h * h, D = sqrt(D2);
  const double u = 1.1102230246251565e-16;
  const double dl = 8.0 * u * D2, eA = 3e-8 * D2 + 1e-15;
  return FilterK{1.5 * (kGhEps + eA), 1.5 * 2.2 * (kGhEps + 2.0 * dl), 16.0 * u * D, dl};
}
// an edge's length and margin (Q edges carry the 1/L term)
__device__ __forceinline__ void edge_margin(const FilterK &K, pt p, pt q, bool sideQ, double &L,
                                            double &m) {
  const double dx = q.x - p.x, dy = q.y - p.y;
  L = sqrt(dx * dx + dy * dy);
  m = sideQ ? K.kA * L + K.kB / L + K.c0 : K.kA * L + K.c0;
}
__device__ __forceinline__ void write_cf(double *c, const double bx[2][4], const double minL[2],
                                         const double maxM[2], const FilterK &K) {
  for (int i = 0; i < 4; i++) {
    c[CF_BOXP + i] = bx[0][i];
    c[CF_BOXQ + i] = bx[1][i];
  }
  c[CF_MAXMP] = maxM[0];
  c[CF_MAXMQ] = maxM[1];
  c[CF_MINLP] = minL[0];
  c[CF_MINLQ] = minL[1];
  c[CF_TINY] = 4.5 * (kGhEps + 2.0 * K.dl);
  c[CF_KA] = K.kA;
  c[CF_KB] = K.kB;
  c[CF_C0] = K.c0;
}

// one group per pair: ring boxes, then the shortest edge and the largest
// margin of each ring (the edges' own lengths and margins are recomputed
// where they are needed)
template <int W>
__global__ void __launch_bounds__(W == 64 ? 256 : W)
    k_cpu_filter_setup(long M, const int *status, const long long *offP, const long long *offQ,
                       const double *ringP, const double *ringQ, double *cf) {
  const long k = pgrp_pair<W>();
  if (k >= M || status[k] != GH_OK) return;
  const int lane = pgrp_lane<W>();
  double bx[2][4];
  for (int side = 0; side < 2; side++) {
    const long long b = side ? offQ[k] : offP[k], n = (side ? offQ[k + 1] : offP[k + 1]) - b;
    const double *V = side ? ringQ : ringP;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (long long i = lane; i < n; i += W) {
      const pt p = ld(V, b + i);
      x0 = fmin(x0, p.x);
      y0 = fmin(y0, p.y);
      x1 = fmax(x1, p.x);
      y1 = fmax(y1, p.y);
    }
    pgrp_minmax<W>(x0, x1);
    pgrp_minmax<W>(y0, y1);
    bx[side][0] = x0;
    bx[side][1] = y0;
    bx[side][2] = x1;
    bx[side][3] = y1;
  }
  const FilterK K = filter_consts(bx);
  double minL[2], maxM[2];
  for (int side = 0; side < 2; side++) {
    const long long b = side ? offQ[k] : offP[k], n = (side ? offQ[k + 1] : offP[k + 1]) - b;
    const double *V = side ? ringQ : ringP;
    double mn = INFINITY, mx = 0.0;
    for (long long i = lane; i < n; i += W) {
      double L, m;
      edge_margin(K, ld(V, b + i), ld(V, b + (i + 1 == n ? 0 : i + 1)), side != 0, L, m);
      mn = fmin(mn, L);
      mx = fmax(mx, m);
    }
    pgrp_minmax<W>(mn, mx);
    minL[side] = mn;
    maxM[side] = mx;
  }
  if (lane == 0) write_cf(cf + (long long)CF_N * k, bx, minL, maxM, K);
}

// The same over `split` 1024-thread blocks per pair, for a few large pairs:
// the ring boxes' parts (k_filter_box_parts), combined into the constants'
// box fields (k_filter_box_combine); then each block takes its share of
// the edges' lengths and margins to parts of their minimum and maximum
// (k_filter_edge_parts), which k_filter_finish combines into the remaining
// constants.  Minima and maxima are exact in any order, so the constants
// are those of k_cpu_filter_setup.
__global__ void __launch_bounds__(1024)
    k_filter_box_parts(long M, int split, const int *status, const long long *offP,
                       const long long *offQ, const double *ringP, const double *ringQ,
                       double *part) {
  const long k = (long)(blockIdx.x / (unsigned)split);
  const int j = (int)(blockIdx.x % (unsigned)split);
  if (k >= M || status[k] != GH_OK) return;
  for (int side = 0; side < 2; side++) {
    const long long b = side ? offQ[k] : offP[k], n = (side ? offQ[k + 1] : offP[k + 1]) - b;
    const double *V = side ? ringQ : ringP;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (long long i = (long long)j * 1024 + threadIdx.x; i < n; i += (long long)split * 1024) {
      const pt p = ld(V, b + i);
      x0 = fmin(x0, p.x);
      y0 = fmin(y0, p.y);
      x1 = fmax(x1, p.x);
      y1 = fmax(y1, p.y);
    }
    pgrp_minmax<1024>(x0, x1);
    pgrp_minmax<1024>(y0, y1);
    if (threadIdx.x == 0) {
      double *d = part + 8 * (long)blockIdx.x + 4 * side;
      d[0] = x0;
      d[1] = y0;
      d[2] = x1;
      d[3] = y1;
    }
  }
}

// part 1's boxes of pair blockIdx.x combined by one 256-thread block into
// the pair's constants' box fields
__global__ void __launch_bounds__(256)
    k_filter_box_combine(long M, int split, const int *status, const double *part, double *cf) {
  const long k = blockIdx.x;
  if (k >= M || status[k] != GH_OK) return;
  double *c = cf + (long long)CF_N * k;
  for (int side = 0; side < 2; side++) {
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int j = threadIdx.x; j < split; j += 256) {
      const double *d = part + 8 * (k * split + j) + 4 * side;
      x0 = fmin(x0, d[0]);
      y0 = fmin(y0, d[1]);
      x1 = fmax(x1, d[2]);
      y1 = fmax(y1, d[3]);
    }
    pgrp_minmax<256>(x0, x1);
    pgrp_minmax<256>(y0, y1);
    if (threadIdx.x == 0) {
      const int o = side ? CF_BOXQ : CF_BOXP;
      c[o + 0] = x0;
      c[o + 1] = y0;
      c[o + 2] = x1;
      c[o + 3] = y1;
    }
  }
}

__device__ __forceinline__ void cf_boxes(const double *c, double bx[2][4]) {
  for (int i = 0; i < 4; i++) {
    bx[0][i] = c[CF_BOXP + i];
    bx[1][i] = c[CF_BOXQ + i];
  }
}

__global__ void __launch_bounds__(1024)
    k_filter_edge_parts(long M, int split, const int *status, const long long *offP,
                        const long long *offQ, const double *ringP, const double *ringQ,
                        const double *cf, double *lenPart) {
  const long k = (long)(blockIdx.x / (unsigned)split);
  const int j = (int)(blockIdx.x % (unsigned)split);
  if (k >= M || status[k] != GH_OK) return;
  double bx[2][4];
  cf_boxes(cf + (long long)CF_N * k, bx);
  const FilterK K = filter_consts(bx);
  for (int side = 0; side < 2; side++) {
    const long long b = side ? offQ[k] : offP[k], n = (side ? offQ[k + 1] : offP[k + 1]) - b;
    const double *V = side ? ringQ : ringP;
    double mn = INFINITY, mx = 0.0;
    for (long long i = (long long)j * 1024 + threadIdx.x; i < n; i += (long long)split * 1024) {
      double L, m;
      edge_margin(K, ld(V, b + i), ld(V, b + (i + 1 == n ? 0 : i + 1)), side != 0, L, m);
      mn = fmin(mn, L);
      mx = fmax(mx, m);
    }
    pgrp_minmax<1024>(mn, mx);
    if (threadIdx.x == 0) {
      lenPart[4 * (long)blockIdx.x + 2 * side] = mn;
      lenPart[4 * (long)blockIdx.x + 2 * side + 1] = mx;
    }
  }
}

// part 2's minima and maxima combined by one 256-thread block per pair
__global__ void __launch_bounds__(256)
    k_filter_finish(long M, int split, const int *status, const double *lenPart, double *cf) {
  const long k = blockIdx.x;
  if (k >= M || status[k] != GH_OK) return;
  double *c = cf + (long long)CF_N * k;
  double minL[2], maxM[2];
  for (int side = 0; side < 2; side++) {
    double mn = INFINITY, mx = 0.0;
    for (int j = threadIdx.x; j < split; j += 256) {
      mn = fmin(mn, lenPart[4 * (k * split + j) + 2 * side]);
      mx = fmax(mx, lenPart[4 * (k * split + j) + 2 * side + 1]);
    }
    pgrp_minmax<256>(mn, mx);
    minL[side] = mn;
    maxM[side] = mx;
  }
  if (threadIdx.x != 0) return;
  double bx[2][4];
  cf_boxes(c, bx);
  write_cf(c, bx, minL, maxM, filter_consts(bx));
}

// Chebyshev gap of two closed edge boxes (negative when they overlap);
// symmetric in the two edges
__device__ __forceinline__ double edge_gap(pt a1, pt a2, pt b1, pt b2) {
  const double gx = fmax(fmin(b1.x, b2.x) - fmax(a1.x, a2.x), fmin(a1.x, a2.x) - fmax(b1.x, b2.x));
  const double gy = fmax(fmin(b1.y, b2.y) - fmax(a1.y, a2.y), fmin(a1.y, a2.y) - fmax(b1.y, b2.y));
  return fmax(gx, gy);
}

// Per-edge view of the filter: an edge A of one ring against the edges of
// the other ring (B).  skip() is the pair decision; the ring-level bound G
// (A's margin plus the largest margin of B) rejects without computing B's
// length and margin when no pair of A can be tiny.  An edge's length and
// margin are computed from its vertices where they are needed, with the
// same operations as the set-up kernels that bound them per ring.  A tiny pair whose boxes
// are apart is still skipped when |A(P1, Q1, Q2)| >= eps: intersect() then
// cannot take its collinear branch (the value is computed as intersect()
// computes it), and the other types are bounded by the margins.
struct CpuFilter {
  double la, ma, G, T;
  bool tiny;   // some pair of A may have L_A L_B < T
  bool sideP;  // A is a P edge
  EdgeBox wide;  // A's box widened by G (outward), for the plain-compare reject
  const double *c;  // the pair's constants
  __device__ __forceinline__ bool skip(int q, pt A1, pt A2, pt B1, pt B2) const {
    // both of B's endpoints beyond one side of the widened box: the gap
    // exceeds G, and no pair of A is tiny
    if (!tiny && box_misses(wide, B1, B2)) return true;
    const double gap = edge_gap(A1, A2, B1, B2);
    if (!(gap > 0.0)) return false;
    if (!tiny && gap > G) return true;
    double lb, mb;
    edge_margin(FilterK{c[CF_KA], c[CF_KB], c[CF_C0], 0.0}, B1, B2, sideP, lb, mb);
    if (!(gap > ma + mb)) return false;
    if (!(la * lb < T)) return true;
    const double AP1 = sideP ? triA(A1, B1, B2) : triA(B1, A1, A2);
    return !(fabs(AP1) < kGhEps);
  }
};

// the filter of edge A1 A2 of pair k (a P edge when sideP)
__device__ __forceinline__ CpuFilter cpu_filter(const double *cf, int k, bool sideP, pt A1,
                                                pt A2) {
  const double *c = cf + (long long)CF_N * k;
  CpuFilter f;
  edge_margin(FilterK{c[CF_KA], c[CF_KB], c[CF_C0], 0.0}, A1, A2, !sideP, f.la, f.ma);
  f.T = c[CF_TINY];
  f.G = f.ma + (sideP ? c[CF_MAXMQ] : c[CF_MAXMP]);
  f.tiny = f.la * (sideP ? c[CF_MINLQ] : c[CF_MINLP]) < f.T;
  f.sideP = sideP;
  f.wide = EdgeBox{0.0, 0.0, 0.0, 0.0};
  f.c = c;
  return f;
}

__device__ __forceinline__ bool ins_on_p(int t) { return t == 1 || t == 3 || t == 5 || t == 7; }
__device__ __forceinline__ bool ins_on_q(int t) { return t == 1 || t == 2 || t == 5 || t == 6; }
// types whose link() touches the P source (one) / the Q source (q1)
__device__ __forceinline__ bool src_on_p(int t) {
  return t == 2 || t == 4 || t == 5 || t == 6 || t == 8;
}
__device__ __forceinline__ bool src_on_q(int t) {
  return t == 3 || t == 4 || t == 5 || t == 7 || t == 8;
}

// Count pass of the CPU route: count1 = partners with a nonzero type,
// count2 = 1 + the vertices the pairs add on this edge's ring.  Threads
// [0, edgesP) take P edges, the rest Q edges; the Q view evaluates
// intersect(P edge, Q edge) with the same argument order as the P view.
// clip_pairs launches it over P's edges only (edgesQ = 0) and builds the Q
// side by transposition (k_cpu_transpose).
// A wave whose P edges all belong to one pair, with a Q ring of at most
// kStageQ vertices walked in full, first copies that ring into its slice of
// LDS (count_stage, kStageQ points per wave) and walks it there: the walk's
// loads, which every lane shares, then come from LDS.
constexpr int kStageQ = 512;
extern __shared__ __align__(16) char count_stage[];

// One lane's edge of the count pass: its pair, its ring neighbours, the
// filter and whether the edge is walked at all (flag).  live is false for
// lanes past the edges and for P edges of pairs the LDS pass takes.
struct CountLane {
  bool live, sideP, flag;
  long ga;
  int k;
  long long bBase, bN;
  pt A1, A2;
  EdgeBox ab;
  CpuFilter f;
};

__device__ __forceinline__ CountLane count_prep(long g, long edgesP, long edgesQ,
                                                const int *mapP, const int *mapQ,
                                                const long long *offP, const long long *offQ,
                                                const double *ringP, const double *ringQ,
                                                const double *cf, long cullMin, int ldsQ) {
  CountLane L;
  L.live = false;
  L.flag = false;
  L.sideP = g < edgesP;
  if (g >= edgesP + edgesQ) return L;
  const bool sideP = L.sideP;
  const long ga = sideP ? g : g - edgesP;
  const int k = sideP ? mapP[ga] : mapQ[ga];
  const long long aBase = sideP ? offP[k] : offQ[k];
  const long long aN = (sideP ? offP[k + 1] : offQ[k + 1]) - aBase;
  const long long bBase = sideP ? offQ[k] : offP[k];
  const long long bN = (sideP ? offQ[k + 1] : offP[k + 1]) - bBase;
  L.ga = ga;
  L.k = k;
  L.bBase = bBase;
  L.bN = bN;
  // P edges of pairs whose Q ring the LDS pass (k_count_cpu_lds) takes
  if (sideP && bN > cullMin && bN <= ldsQ) return L;
  L.live = true;
  const double *A = sideP ? ringP : ringQ;
  const long long i = ga - aBase;
  L.A1 = ld(A, ga);
  L.A2 = ld(A, i == aN - 1 ? aBase : ga + 1);
  L.f = cpu_filter(cf, k, sideP, L.A1, L.A2);
  L.ab = edge_box(L.A1, L.A2);
  {
    // G widened by a relative 1e-12 so that rounding of the box sums cannot
    // turn a gap of exactly G into a reject
    const double Gw = L.f.G * (1.0 + 1e-12) + 1e-300;
    L.f.wide = EdgeBox{L.ab.x0 - Gw, L.ab.y0 - Gw, L.ab.x1 + Gw, L.ab.y1 + Gw};
  }
  const double *ob = cf + (long long)CF_N * k + (sideP ? CF_BOXQ : CF_BOXP);
  const EdgeBox &ab = L.ab;
  const double ogap = fmax(fmax(ob[0] - ab.x1, ab.x0 - ob[2]), fmax(ob[1] - ab.y1, ab.y0 - ob[3]));
  L.flag = L.f.tiny || !(ogap > L.f.G);
  return L;
}

// a value every lane of the wave holds, moved to scalar registers
__device__ __forceinline__ double wave_uniform(double v) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)b);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The next group of a wave's lanes for the wave walk: the pending lanes
// (P edges walked over their whole Q ring, not yet counted) of the pair of
// the first pending lane, with that ring's first global edge e0 and length
// n, and U, the union of the group's widened boxes over the lanes that walk
// (all of the plane when one of them may form a tiny pair).  False when no
// lane is pending.  Every lane of the wave calls it.
__device__ __forceinline__ bool wave_group(const CountLane &L, bool pending, bool &group,
                                           EdgeBox &U, long long &e0, int &n) {
  const unsigned long long pm = __ballot(pending);
  if (!pm) return false;
  const int first = __ffsll((long long)pm) - 1;
  const int k0 = __shfl(L.k, first, 64);
  group = pending && L.k == k0;
  const bool walks = group && L.flag;
  const bool anyTiny = __ballot(walks && L.f.tiny) != 0;
  double x0 = walks ? L.f.wide.x0 : INFINITY, y0 = walks ? L.f.wide.y0 : INFINITY;
  double x1 = walks ? L.f.wide.x1 : -INFINITY, y1 = walks ? L.f.wide.y1 : -INFINITY;
  for (int off = 32; off > 0; off >>= 1) {
    x0 = fmin(x0, __shfl_xor(x0, off, 64));
    y0 = fmin(y0, __shfl_xor(y0, off, 64));
    x1 = fmax(x1, __shfl_xor(x1, off, 64));
    y1 = fmax(y1, __shfl_xor(y1, off, 64));
  }
  if (anyTiny) {
    x0 = y0 = -INFINITY;
    x1 = y1 = INFINITY;
  }
  // the same in every lane: held in scalar registers
  U = EdgeBox{wave_uniform(x0), wave_uniform(y0), wave_uniform(x1), wave_uniform(y1)};
  const long long b0 = __shfl(L.bBase, first, 64);
  const int n0 = (int)__shfl(L.bN, first, 64);
  e0 = (long long)__builtin_amdgcn_readfirstlane((unsigned)b0) |
       ((long long)__builtin_amdgcn_readfirstlane((unsigned)((unsigned long long)b0 >> 32)) << 32);
  n = __builtin_amdgcn_readfirstlane(n0);
  return true;
}

// lane l's double, read by every lane (l the same in every lane)
__device__ __forceinline__ double lane_d(double v, int l) {
  const unsigned long long b = __double_as_longlong(v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The count pass of a group of a wave's lanes that share one Q ring (global
// edges [e0, e0 + n)), walked by every lane of the wave in step: the ring's
// runs of kRun1 edges (boxes b1, global edge order) whose box misses U are
// left out, since U holds the widened boxes of every walking lane of the
// group, none with a possible tiny pair (U is the whole plane otherwise), so
// each walking lane's filter skips every edge of such a run (its endpoints
// lie beyond one side of the lane's widened box); of each other run, lanes
// 0..kRun1 load its vertices once and every lane reads them from there (in
// scalar registers), in index order.  Lanes that do not walk take part
// without visiting.
template <class F>
__device__ __forceinline__ int wave_ring_edges(const double *V, long long e0, int n,
                                               const double *b1, const EdgeBox &U, F &&f) {
  // run r covers the ring's edges [r kRun1 - o, (r + 1) kRun1 - o)
  const double *b = b1 + 4 * (e0 / kRun1);
  const int o = (int)(e0 % kRun1);
  const int lane = (int)__lane_id();
  int visited = 0;
  for (int qa = 0, qz = kRun1 - o; qa < n; qa = qz, qz += kRun1, b += 4) {
    if (U.x0 > b[2] || U.x1 < b[0] || U.y0 > b[3] || U.y1 < b[1]) continue;
    const int z = min(qz, n);
    visited += z - qa;
    // vertices qa .. z (the last one wraps to 0 at the ring's end)
    pt v{0.0, 0.0};
    if (lane <= z - qa) v = ld(V, qa + lane == n ? 0 : qa + lane);
    for (int q = qa; q < z; q++) {
      const int j = q - qa;
      f(q, pt{lane_d(v.x, j), lane_d(v.y, j)}, pt{lane_d(v.x, j + 1), lane_d(v.y, j + 1)});
    }
  }
  return visited;
}

template <bool CULL>
__device__ __forceinline__ void count_edge_cpu(const CountLane &L, const double *ringP,
                                               const double *ringQ, int *cnt1P, int *cnt2P,
                                               int *cnt1Q, int *cnt2Q, int *boolP, int *boolQ,
                                               int *lstP, int *lstQ, const RunBoxes &boxP,
                                               const RunBoxes &boxQ, int stage,
                                               unsigned long long &ev, unsigned long long &tv) {
  const bool sideP = L.sideP;
  const long ga = L.ga;
  const long long bBase = L.bBase, bN = L.bN;
  const double *Bv = (sideP ? ringQ : ringP) + 2 * bBase;
  const pt A1 = L.A1, A2 = L.A2;
  const CpuFilter &f = L.f;
  const EdgeBox &ab = L.ab;
  int c1 = 0, c2 = 0;
  if (L.flag) {
    int *lst = (sideP ? lstP : lstQ) + (long long)ga * kPartners;
    auto visit = [&](int q, pt B1, pt B2) {
      if (f.skip(q, A1, A2, B1, B2)) return;
      double al, be;
      ev++;
      const int t = sideP ? itype(A1, A2, B1, B2, al, be) : itype(B1, B2, A1, A2, al, be);
      if (t != 0) {
        if (c1 < kPartners) lst[c1] = q;
        c1++;
      }
      c2 += (sideP ? ins_on_p(t) : ins_on_q(t)) ? 1 : 0;
    };
    const double *Vs =
        sideP && stage >= 0 ? (const double *)count_stage + 2 * kStageQ * stage : Bv;
    if (CULL && bN > boxP.cullMin) {
      // runs apart from the box widened by G and without a possible tiny
      // pair hold no pair the filter keeps
      const double G = f.G * (1.0 + 1e-12) + 1e-300;
      const EdgeBox wide{ab.x0 - G, ab.y0 - G, ab.x1 + G, ab.y1 + G};
      tv = culled_edges(Bv, bBase, bBase + bN, sideP ? boxQ : boxP, wide, f.la, f.T, visit);
    } else {
      ring_edges(Vs, (int)bN, visit);
      tv = bN;
    }
  }
  if (sideP) {
    cnt1P[ga] = c1;
    cnt2P[ga] = c2 + 1;
    boolP[ga] = L.flag;
  } else {
    cnt1Q[ga] = c1;
    cnt2Q[ga] = c2 + 1;
    boolQ[ga] = L.flag;
  }
}

// count_edge_cpu for a group of a wave's lanes that walks one Q ring
// together (wave_ring_edges); every lane of the wave calls it
__device__ __forceinline__ void count_edge_wave(const CountLane &L, bool group,
                                                const double *ringQ, int *cnt1P, int *cnt2P,
                                                int *boolP, int *lstP, long long e0, int n,
                                                const double *b1, const EdgeBox &U,
                                                unsigned long long &ev, unsigned long long &tv) {
  const bool walks = group && L.flag;
  const pt A1 = L.A1, A2 = L.A2;
  const CpuFilter &f = L.f;
  int c1 = 0, c2 = 0;
  int *lst = lstP + (long long)L.ga * kPartners;
  auto visit = [&](int q, pt B1, pt B2) {
    if (!walks || f.skip(q, A1, A2, B1, B2)) return;
    double al, be;
    ev++;
    const int t = itype(A1, A2, B1, B2, al, be);
    if (t != 0) {
      if (c1 < kPartners) lst[c1] = q;
      c1++;
    }
    c2 += ins_on_p(t) ? 1 : 0;
  };
  if (__ballot(walks)) {
    const int vis = wave_ring_edges(ringQ + 2 * e0, e0, n, b1, U, visit);
    if (walks) tv = vis;
  }
  if (group) {
    cnt1P[L.ga] = c1;
    cnt2P[L.ga] = c2 + 1;
    boolP[L.ga] = L.flag;
  }
}

// five waves per SIMD for the plain walk (at most 102 VGPRs, no spills): it
// is bound by latency, and 5 waves run it 6 % faster than the 4 that its
// register count would otherwise allow (0.42 vs 0.45 ms on the join
// workload); the culled walk keeps the registers it needs to run without spills
// skipQ (optional): the Q edges' kRun1-edge run boxes, with which the P
// edges of a wave walked over their whole Q ring walk it together, one
// group per pair, leaving out the runs no lane of the group can pair with
// (count_edge_wave).  WAVE: skipQ is given and every edge is a P edge
// (edgesQ = 0), so that without culling no lane walks on its own and the
// lane walk is left out of the kernel.
template <bool CULL, bool WAVE>
__global__ void __launch_bounds__(256, CULL ? 1 : 5)
    k_count_cpu(long edgesP, long edgesQ, const int *mapP, const int *mapQ,
                const long long *offP, const long long *offQ, const double *ringP,
                const double *ringQ, const double *cf, int *cnt1P, int *cnt2P, int *cnt1Q,
                int *cnt2Q, int *boolP, int *boolQ, unsigned long long *wstat, int *lstP,
                int *lstQ, RunBoxes boxP, RunBoxes boxQ, int ldsQ, int stageOk,
                const double *skipQ) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long ev = 0, tv = 0;
  // stage the wave's Q ring when every P edge of the wave is of one pair
  // whose ring is walked in full and fits (the same in every lane)
  int stage = -1;
  if (!WAVE && stageOk) {
    const bool pv = g < edgesP;
    const int k = pv ? mapP[g] : -1;
    const unsigned long long vm = __ballot(pv);
    if (vm) {
      const int k0 = __shfl(k, __ffsll((long long)vm) - 1, 64);
      if (__ballot(pv && k != k0) == 0) {
        const long long b0 = offQ[k0], bn = offQ[k0 + 1] - b0;
        const bool walked = !(bn > boxP.cullMin && (CULL || bn <= ldsQ));
        if (bn > 0 && bn <= kStageQ && walked) {
          double *st = (double *)count_stage + 2 * kStageQ * (threadIdx.x >> 6);
          for (long j = __lane_id(); j < bn; j += 64) {
            const pt v = ld(ringQ, b0 + j);
            st[2 * j] = v.x;
            st[2 * j + 1] = v.y;
          }
          stage = (int)(threadIdx.x >> 6);
        }
      }
    }
    __syncthreads();
  }
  const CountLane L = count_prep(g, edgesP, edgesQ, mapP, mapQ, offP, offQ, ringP, ringQ, cf,
                                 boxP.cullMin, ldsQ);
  // P edges walked over their whole Q ring: one group per pair of the wave
  bool pending = WAVE && L.live && L.sideP && !(CULL && L.bN > boxP.cullMin);
  if (WAVE) {
    EdgeBox U;
    long long e0;
    int n;
    bool group;
    while (wave_group(L, pending, group, U, e0, n)) {
      count_edge_wave(L, group, ringQ, cnt1P, cnt2P, boolP, lstP, e0, n, skipQ, U, ev, tv);
      if (group) pending = false;
    }
  }
  if ((CULL || !WAVE) && L.live && !(WAVE && L.sideP && !(CULL && L.bN > boxP.cullMin)))
    count_edge_cpu<CULL>(L, ringP, ringQ, cnt1P, cnt2P, cnt1Q, cnt2Q, boolP, boolQ, lstP, lstQ,
                         boxP, boxQ, stage, ev, tv);
  // every lane of every wave reaches the statistics
  wave_stats(wstat, g, ev, tv);
}

// ---------------------------------------- count pass, Q binned in LDS ----
// The CPU route's count pass for the pairs whose Q ring is longer than the
// culling threshold and fits in LDS (Tuning::cullMin < nQ <= Tuning::ldsMaxQ
// <= kLdsQ): one block per pair.  (Shorter rings keep the lockstep walk of
// k_count_cpu, whose loads every lane of a wave shares: on the join
// workload, rings of at most 512 vertices, this pass visited 9.4 times
// fewer edge pairs but took 0.81 ms against the walk's 0.43 ms on MI355X,
// the per-pair staging and binning and the lanes' divergent walks costing
// more than the visits saved.)  The block stages
// the Q ring in LDS and bins its edges in a g x g grid over Q's box
// (g = ceil(sqrt(nQ / 2)), at most kLdsG); an edge whose box spans more
// than kLdsSpan cells goes to a list that every P edge walks.  A P edge then
// visits only the edges binned in the cells its box widened by G covers,
// each edge once (from the first covered cell of its own cell range),
// instead of the whole ring.  Cell indices are a monotone function of the
// coordinate, so an edge outside those cells lies beyond the widened box in
// x or y, and CpuFilter::skip rejects it by its first test: counts, flags and
// partner sets are the ring walk's.  Partners come out of index order; a
// list of at most kPartners is sorted before it is stored (longer lists are
// rebuilt in index order by k_partner_rescan_cpu).  A P edge whose filter
// admits tiny pairs walks the whole ring from LDS, in index order.
constexpr int kLdsQ = 1024;   // largest Q ring the LDS pass takes
constexpr int kLdsG = 23;     // largest grid side: ceil(sqrt(kLdsQ / 2))
constexpr int kLdsSpan = 4;   // edges spanning more cells go to the long list
constexpr int kLdsBlock = 128;
static_assert(kLdsBlock == 128, "the cell scan below adds the totals of exactly two waves");

struct LdsGrid {
  int g;
  double x0, y0, sx, sy;  // grid corner and cells per unit
  __device__ __forceinline__ int cx(double x) const {
    // (x - x0) * sx is monotone in x; the clamp maps NaN to 0
    return (int)fmin(fmax((x - x0) * sx, 0.0), (double)(g - 1));
  }
  __device__ __forceinline__ int cy(double y) const {
    return (int)fmin(fmax((y - y0) * sy, 0.0), (double)(g - 1));
  }
};

// dynamic LDS of the pass for Q rings of up to capQ vertices
__host__ __device__ constexpr size_t lds_count_bytes(int capQ) {
  return 16 * (size_t)capQ                      // Q vertices
         + 4 * (size_t)(kLdsG * kLdsG + 1)      // cell offsets
         + 4 * (size_t)(kLdsG * kLdsG)          // cell cursors
         + 2 * (size_t)kLdsSpan * capQ          // binned edge indices
         + 2 * (size_t)capQ                     // long edges
         + 2 * (size_t)kLdsBlock * kPartners    // each thread's partners
         + 64;                                  // counters and scan scratch
}

__global__ void __launch_bounds__(kLdsBlock)
    k_count_cpu_lds(long M, int minQ, int capQ, const long long *offP, const long long *offQ,
                    const double *ringP, const double *ringQ, const double *cf,
                    int *cnt1P, int *cnt2P, int *boolP, int *lstP, unsigned long long *bstat) {
  extern __shared__ double lds_count[];
  pt *sq = (pt *)lds_count;
  int *cellOff = (int *)(sq + capQ);
  int *cellCur = cellOff + (kLdsG * kLdsG + 1);
  short *ent = (short *)(cellCur + kLdsG * kLdsG);
  short *longL = ent + kLdsSpan * capQ;
  short *hl = longL + capQ;
  int *misc = (int *)(((uintptr_t)(hl + kLdsBlock * kPartners) + 15) & ~(uintptr_t)15);
  // misc[0] long-list length, misc[2..3] the two waves' sums of the cell scan
  const long k = blockIdx.x;
  unsigned long long ev = 0, tv = 0;
  const long long pBase = offP[k], nP = offP[k + 1] - pBase;
  const long long qBase = offQ[k], nQ = offQ[k + 1] - qBase;
  const bool mine = nP > 0 && nQ > minQ && nQ <= capQ;
  if (mine) {
    // -- stage Q and bin its edges
    const double *c = cf + (long long)CF_N * k + CF_BOXQ;
    LdsGrid gr;
    gr.g = min(kLdsG, max(1, (int)ceil(sqrt(0.5 * (double)nQ))));
    gr.x0 = c[0];
    gr.y0 = c[1];
    const double w = c[2] - c[0], h = c[3] - c[1];
    gr.sx = w > 0.0 ? (double)gr.g / w : 0.0;
    gr.sy = h > 0.0 ? (double)gr.g / h : 0.0;
    const int cells = gr.g * gr.g;
    for (long long q = threadIdx.x; q < nQ; q += kLdsBlock) sq[q] = ld(ringQ, qBase + q);
    for (int i = threadIdx.x; i < cells; i += kLdsBlock) cellCur[i] = 0;
    if (threadIdx.x == 0) misc[0] = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < (int)nQ; q += kLdsBlock) {
      const pt B1 = sq[q], B2 = sq[q + 1 == nQ ? 0 : q + 1];
      const int i0 = gr.cx(fmin(B1.x, B2.x)), i1 = gr.cx(fmax(B1.x, B2.x));
      const int j0 = gr.cy(fmin(B1.y, B2.y)), j1 = gr.cy(fmax(B1.y, B2.y));
      if ((i1 - i0 + 1) * (j1 - j0 + 1) > kLdsSpan) {
        longL[atomicAdd(&misc[0], 1)] = (short)q;
      } else {
        for (int i = i0; i <= i1; i++)
          for (int j = j0; j <= j1; j++) atomicAdd(&cellCur[i * gr.g + j], 1);
      }
    }
    __syncthreads();
    // exclusive scan of the cell counts: each thread sums a run of cells
    {
      const int per = (cells + kLdsBlock - 1) / kLdsBlock;
      const int a = min(cells, (int)threadIdx.x * per), z = min(cells, a + per);
      int sum = 0;
      for (int i = a; i < z; i++) sum += cellCur[i];
      int x = sum;  // inclusive scan over the wave
      for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (__lane_id() >= off) x += y;
      }
      if (__lane_id() == 63) misc[2 + (threadIdx.x >> 6)] = x;
      __syncthreads();
      int run = x - sum + ((threadIdx.x >> 6) ? misc[2] : 0);
      for (int i = a; i < z; i++) {
        const int v = cellCur[i];
        cellOff[i] = run;
        cellCur[i] = run;
        run += v;
      }
      if (threadIdx.x == kLdsBlock - 1) cellOff[cells] = run;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < (int)nQ; q += kLdsBlock) {
      const pt B1 = sq[q], B2 = sq[q + 1 == nQ ? 0 : q + 1];
      const int i0 = gr.cx(fmin(B1.x, B2.x)), i1 = gr.cx(fmax(B1.x, B2.x));
      const int j0 = gr.cy(fmin(B1.y, B2.y)), j1 = gr.cy(fmax(B1.y, B2.y));
      if ((i1 - i0 + 1) * (j1 - j0 + 1) <= kLdsSpan)
        for (int i = i0; i <= i1; i++)
          for (int j = j0; j <= j1; j++) ent[atomicAdd(&cellCur[i * gr.g + j], 1)] = (short)q;
    }
    __syncthreads();
    const int nLong = misc[0];
    // -- the P edges
    short *mine_hl = hl + threadIdx.x * kPartners;
    for (long long i = threadIdx.x; i < nP; i += kLdsBlock) {
      const long long ga = pBase + i;
      const pt A1 = ld(ringP, ga), A2 = ld(ringP, i == nP - 1 ? pBase : ga + 1);
      CpuFilter f = cpu_filter(cf, (int)k, true, A1, A2);
      const EdgeBox ab = edge_box(A1, A2);
      {
        const double Gw = f.G * (1.0 + 1e-12) + 1e-300;
        f.wide = EdgeBox{ab.x0 - Gw, ab.y0 - Gw, ab.x1 + Gw, ab.y1 + Gw};
      }
      const double ogap = fmax(fmax(c[0] - ab.x1, ab.x0 - c[2]), fmax(c[1] - ab.y1, ab.y0 - c[3]));
      const bool flag = f.tiny || !(ogap > f.G);
      int c1 = 0, c2 = 0;
      if (flag) {
        auto visit = [&](int q, pt B1, pt B2) {
          if (f.skip(q, A1, A2, B1, B2)) return;
          double al, be;
          ev++;
          const int t = itype(A1, A2, B1, B2, al, be);
          if (t != 0) {
            if (c1 < kPartners) mine_hl[c1] = (short)q;
            c1++;
          }
          c2 += ins_on_p(t) ? 1 : 0;
        };
        if (f.tiny) {
          for (int q = 0; q < (int)nQ; q++) visit(q, sq[q], sq[q + 1 == nQ ? 0 : q + 1]);
          tv += nQ;
        } else {
          const int qi0 = gr.cx(f.wide.x0), qi1 = gr.cx(f.wide.x1);
          const int qj0 = gr.cy(f.wide.y0), qj1 = gr.cy(f.wide.y1);
          for (int ci = qi0; ci <= qi1; ci++) {
            for (int cj = qj0; cj <= qj1; cj++) {
              const int cell = ci * gr.g + cj;
              for (int e = cellOff[cell]; e < cellOff[cell + 1]; e++) {
                const int q = ent[e];
                const pt B1 = sq[q], B2 = sq[q + 1 == nQ ? 0 : q + 1];
                // visited from the first cell its range shares with the query
                if (max(gr.cx(fmin(B1.x, B2.x)), qi0) != ci ||
                    max(gr.cy(fmin(B1.y, B2.y)), qj0) != cj)
                  continue;
                tv++;
                visit(q, B1, B2);
              }
            }
          }
          for (int e = 0; e < nLong; e++) {
            const int q = longL[e];
            tv++;
            visit(q, sq[q], sq[q + 1 == nQ ? 0 : q + 1]);
          }
        }
      }
      cnt1P[ga] = c1;
      cnt2P[ga] = c2 + 1;
      boolP[ga] = flag;
      if (c1 <= kPartners) {
        // the partners in index order (insertion sort of at most 16)
        for (int a = 1; a < c1; a++) {
          const short v = mine_hl[a];
          int b = a - 1;
          while (b >= 0 && mine_hl[b] > v) {
            mine_hl[b + 1] = mine_hl[b];
            b--;
          }
          mine_hl[b + 1] = v;
        }
        int *lst = lstP + ga * kPartners;
        for (int a = 0; a < c1; a++) lst[a] = mine_hl[a];
      }
    }
  }
  // the block's statistics, one plain store per block (k_sum_wave_stats)
  for (int off = 32; off > 0; off >>= 1) {
    ev += __shfl_xor(ev, off, 64);
    tv += __shfl_xor(tv, off, 64);
  }
  __shared__ unsigned long long wsum[2][kLdsBlock / 64];
  if (__lane_id() == 0) {
    wsum[0][threadIdx.x >> 6] = ev;
    wsum[1][threadIdx.x >> 6] = tv;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long e = 0, t = 0;
    for (int w = 0; w < kLdsBlock / 64; w++) {
      e += wsum[0][w];
      t += wsum[1][w];
    }
    bstat[2 * k] = e;
    bstat[2 * k + 1] = t;
  }
}

// Partner lists of the edges with more than kPartners partners, CPU route:
// one wave per edge, 64 edges of the other ring at a time, hits in order.
__global__ void __launch_bounds__(64)
    k_partner_rescan_cpu(long nOver, const int *listO, bool sideP, const int *mapA,
                         const long long *offA, const long long *offB, const double *ringA,
                         const double *ringB, const long long *psA1, int *lst,
                         const double *cf, RunBoxes rb) {
  for (long w = blockIdx.x; w < nOver; w += gridDim.x) {
    const long g = listO[w];
    const int k = mapA[g];
    const long long c1 = psA1[g + 1] - psA1[g];
    int *dst = lst + psA1[g];
    const long long aBase = offA[k], aN = offA[k + 1] - aBase;
    const long long bBase = offB[k], bN = offB[k + 1] - bBase;
    const long long i = g - aBase;
    const pt A1 = ld(ringA, g), A2 = ld(ringA, i == aN - 1 ? aBase : g + 1);
    CpuFilter f = cpu_filter(cf, k, sideP, A1, A2);
    {
      const EdgeBox ab = edge_box(A1, A2);
      const double Gw = f.G * (1.0 + 1e-12) + 1e-300;
      f.wide = EdgeBox{ab.x0 - Gw, ab.y0 - Gw, ab.x1 + Gw, ab.y1 + Gw};
    }
    const double *V = ringB + 2 * bBase;
    long long n = 0;
    auto scan = [&](long long a, long long z) {
      for (long long q0 = a; q0 < z; q0 += 64) {
        const long long q = q0 + __lane_id();
        bool hit = false;
        if (q < z) {
          const pt B1 = ld(V, q), B2 = ld(V, q + 1 == bN ? 0 : q + 1);
          if (!f.skip((int)q, A1, A2, B1, B2)) {
            double al, be;
            hit = (sideP ? itype(A1, A2, B1, B2, al, be) : itype(B1, B2, A1, A2, al, be)) != 0;
          }
        }
        const unsigned long long m = __ballot(hit);
        const long long r = n + __popcll(m & ((1ULL << __lane_id()) - 1ULL));
        if (hit && r < c1) dst[r] = (int)q;
        n += __popcll(m);
      }
    };
    if (rb.b1 == nullptr || bN <= rb.cullMin) {
      scan(0, bN);
    } else {
      const double G = f.G * (1.0 + 1e-12) + 1e-300;
      const EdgeBox ab0 = edge_box(A1, A2);
      const EdgeBox ab{ab0.x0 - G, ab0.y0 - G, ab0.x1 + G, ab0.y1 + G};
      const long long e0 = bBase, e1 = bBase + bN;
      for (long long S3 = e0 / kRun3; S3 * kRun3 < e1; S3++) {
        if (!run_hit(rb.b3 + 4 * S3, rb.l3 + S3, ab, f.la, f.T)) continue;
        const long long a2 = max(S3 * 16, e0 / kRun2),
                        z2 = min(S3 * 16 + 16, (e1 - 1) / kRun2 + 1);
        for (long long S2 = a2; S2 < z2; S2++)
          if (run_hit(rb.b2 + 4 * S2, rb.l2 + S2, ab, f.la, f.T))
            scan(max(S2 * kRun2, e0) - e0, min(S2 * kRun2 + kRun2, e1) - e0);
      }
    }
    for (long long r = n + __lane_id(); r < c1; r += 64) dst[r] = 0;
  }
}

// The Q side of the CPU route from the P side's complete partner lists.
// Every pair (P edge, Q edge) with a nonzero type, taken in P-edge order,
// becomes a record keyed by its global Q edge, and the Q edge's counts are
// added from the same pass (the type recomputed as intersect(P edge, Q
// edge), as every pass computes it).  A stable sort of the records by Q edge
// then lists each Q edge's partners in increasing P order: exactly the list
// a walk of that Q edge over the P ring would build.
__global__ void __launch_bounds__(256)
    k_cpu_transpose(long nActP, const int *actP, const int *mapP, const long long *offP,
                    const long long *offQ, const double *ringP, const double *ringQ,
                    const long long *psP1, const int *lstAllP, int *cnt1Q, int *cnt2Q,
                    uint64_t *keys, int *vals) {
  const long ia = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ia >= nActP) return;
  const long g = actP[ia];
  const int k = mapP[g];
  const long long pBase = offP[k], pN = offP[k + 1] - pBase;
  const long long qBase = offQ[k], qN = offQ[k + 1] - qBase;
  const int pid = (int)(g - pBase);
  const pt A1 = ld(ringP, g), A2 = ld(ringP, pid == pN - 1 ? pBase : g + 1);
  const long long j0 = psP1[g];
  long long j = j0;
  auto visit = [&](int q, pt B1, pt B2) {
    const long long qg = qBase + q;
    double al, be;
    const int t = itype(A1, A2, B1, B2, al, be);
    atomicAdd(&cnt1Q[qg], 1);
    if (ins_on_q(t)) atomicAdd(&cnt2Q[qg], 1);
    keys[j] = (uint64_t)qg;
    vals[j] = pid;
    j++;
  };
  listed_edges(ringQ + 2 * qBase, (int)qN, lstAllP + j0, (int)(psP1[g + 1] - j0), visit);
}

// The Q edges' prefilter flags of the CPU route (the flag the count kernel
// records for a P edge: a possible tiny pair, or the edge's box within G of
// the P ring's box), for a batch whose Q side is built by transposition.
__global__ void __launch_bounds__(256)
    k_cpu_flags_q(long edgesQ, const int *mapQ, const long long *offP, const long long *offQ,
                  const double *ringQ, const double *cf, int *boolQ, int *cnt1Q, int *cnt2Q) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= edgesQ) return;
  // the transposition's counters start here: count1 0, count2 1 (the parent)
  cnt1Q[g] = 0;
  cnt2Q[g] = 1;
  const int k = mapQ[g];
  const long long aBase = offQ[k], aN = offQ[k + 1] - aBase;
  const pt A1 = ld(ringQ, g), A2 = ld(ringQ, g - aBase == aN - 1 ? aBase : g + 1);
  const CpuFilter f = cpu_filter(cf, k, false, A1, A2);
  const double *ob = cf + (long long)CF_N * k + CF_BOXP;
  const EdgeBox ab = edge_box(A1, A2);
  const double ogap = fmax(fmax(ob[0] - ab.x1, ab.x0 - ob[2]), fmax(ob[1] - ab.y1, ab.y0 - ob[3]));
  boolQ[g] = (f.tiny || !(ogap > f.G)) ? 1 : 0;
}

// CPU-route slot records
struct CpuSlots {
  uint64_t *kP2, *kQ2;       // exact alpha / beta (IEEE bits) of the unsorted slots
  int *rP2, *rQ2, *rP, *rQ;  // (partner index in its ring << 4) | type; unsorted, sorted
  int *dP, *dQ;              // sorted position of each unsorted slot
  int *srcP, *srcQ;          // per edge: its source vertex's last link, (partner << 4) | type
};

__device__ __forceinline__ uint64_t dbits(double v) {
  union {
    double d;
    uint64_t u;
  } c;
  c.d = v;
  return c.u;
}

// Edges with more than kSmallSort inserted vertices (the medium and long
// runs' lists) are filled and linked by one wave each, 64 list entries or
// slots a round (k_cpu_fill_runs, k_cpu_link_runs); the per-edge threads
// leave them out.  A few such edges can hold most of a large pair's
// intersections.
__device__ __forceinline__ bool wave_run(long long nIns) { return nIns > kSmallSort; }

// the wave-filled edges: wave w takes entry w of [medium P | long P |
// medium Q | long Q]
struct RunLists {
  const int *l[4];
  int n[4];
  __device__ __forceinline__ bool pick(long w, bool &sideP, long &g) const {
    for (int j = 0; j < 4; j++) {
      if (w < n[j]) {
        sideP = j < 2;
        g = l[j][w];
        return true;
      }
      w -= n[j];
    }
    return false;
  }
};

// The inserted vertices of every active edge, in decreasing partner order
// after the parent slot (so that a stable sort by alpha gives insertVertex's
// order), and the source vertex's last link.  Coordinates as
// computeIntersections makes them: I = (1 - alpha) P1 + alpha P2 for type 1
// on both rings, Q1 on P for types 3/5/7, P1 on Q for types 2/5/6.
__global__ void __launch_bounds__(256)
    k_cpu_fill(long nActP, long nActQ, const int *actP, const int *actQ, const int *mapP,
               const int *mapQ, const long long *offP, const long long *offQ,
               const double *ringP, const double *ringQ, const long long *psP1,
               const long long *psQ1, const long long *psP2, const long long *psQ2,
               const int *status, PairMeta pm, const double *bCoords, const double *oCoords,
               const int *lstP, const int *lstQ, SlotArrays s, CpuSlots cs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nActP + nActQ) return;
  const bool sideP = i < nActP;
  const long g = sideP ? actP[i] : actQ[i - nActP];
  const int k = sideP ? mapP[g] : mapQ[g];
  if (status[k] != GH_OK) return;
  const int sw = pm.swapped[k];
  const double *cP = sw ? oCoords : bCoords, *cQ = sw ? bCoords : oCoords;
  const long long aBase = sideP ? offP[k] : offQ[k];
  const long long aN = (sideP ? offP[k + 1] : offQ[k + 1]) - aBase;
  const long long bBase = sideP ? offQ[k] : offP[k];
  const long long bN = (sideP ? offQ[k + 1] : offP[k + 1]) - bBase;
  const double *cA = sideP ? cP : cQ, *cB = sideP ? cQ : cP;
  const long sA = sideP ? pm.startP[k] : pm.startQ[k], sB = sideP ? pm.startQ[k] : pm.startP[k];
  const int ia = (int)(g - aBase);
  const int ia2 = ia == aN - 1 ? 0 : ia + 1;
  const double *ringA = sideP ? ringP : ringQ, *ringB = sideP ? ringQ : ringP;
  const pt A1 = ld(ringA, g), A2 = ld(ringA, aBase + ia2);
  const long long *ps2 = sideP ? psP2 : psQ2, *ps1 = sideP ? psP1 : psQ1;
  const long long base = ps2[g], nIns = ps2[g + 1] - base - 1;
  if (wave_run(nIns)) return;  // k_cpu_fill_runs
  double *i2 = sideP ? s.iP2 : s.iQ2, *o2 = sideP ? s.oP2 : s.oQ2;
  uint64_t *k2 = sideP ? cs.kP2 : cs.kQ2;
  int *r2 = sideP ? cs.rP2 : cs.rQ2;
  long long r = 0;
  int src = -1;
  auto visit = [&](int q, pt B1, pt B2) {
    double al, be;
    const int t = sideP ? itype(A1, A2, B1, B2, al, be) : itype(B1, B2, A1, A2, al, be);
    if (t == 0) return;
    if (sideP ? src_on_p(t) : src_on_q(t)) src = (q << 4) | t;
    if (!(sideP ? ins_on_p(t) : ins_on_q(t))) return;
    const long long u = base + nIns - r;
    r++;
    pt v, vo;
    if (t == 1) {
      const int q2 = q + 1 == bN ? 0 : q + 1;
      const pt P1 = sideP ? A1 : B1, P2 = sideP ? A2 : B2;
      const pt P1o = sideP ? ld(cA, sA + ia) : ld(cB, sB + q);
      const pt P2o = sideP ? ld(cA, sA + ia2) : ld(cB, sB + q2);
      v = pt{(1.0 - al) * P1.x + al * P2.x, (1.0 - al) * P1.y + al * P2.y};
      vo = pt{(1.0 - al) * P1o.x + al * P2o.x, (1.0 - al) * P1o.y + al * P2o.y};
    } else {
      // the other ring's first vertex: Q1 on P, P1 on Q
      v = B1;
      vo = ld(cB, sB + q);
    }
    if (u > base && u <= base + nIns) {
      r2[u] = (q << 4) | t;
      if (nIns == 1) {
        // a run of one is in place: no sort reads its key or staged copy
        st(sideP ? s.iP : s.iQ, u, v);
        st(sideP ? s.oP : s.oQ, u, vo);
        (sideP ? cs.rP : cs.rQ)[u] = (q << 4) | t;
        (sideP ? cs.dP : cs.dQ)[u] = (int)u;
      } else {
        st(i2, u, v);
        st(o2, u, vo);
        k2[u] = dbits(sideP ? al : be);
      }
    }
  };
  listed_edges(ringB + 2 * bBase, (int)bN, (sideP ? lstP : lstQ) + ps1[g],
               (int)(ps1[g + 1] - ps1[g]), visit);
  (sideP ? cs.srcP : cs.srcQ)[g] = src;
}

template <bool QSIDE>
__device__ __forceinline__ void cpu_move(SlotArrays &s, CpuSlots &cs, long long j, long long d) {
  st(QSIDE ? s.iQ : s.iP, d, ld(QSIDE ? s.iQ2 : s.iP2, j));
  st(QSIDE ? s.oQ : s.oP, d, ld(QSIDE ? s.oQ2 : s.oP2, j));
  (QSIDE ? cs.rQ : cs.rP)[d] = (QSIDE ? cs.rQ2 : cs.rP2)[j];
  (QSIDE ? cs.dQ : cs.dP)[j] = (int)d;
}

// runs of 2..kSmallSort slots: 32 lanes per edge rank by (key, slot)
template <bool QSIDE>
__global__ void __launch_bounds__(256)
    k_cpu_sort_short(long n, const int *list, const long long *ps2, SlotArrays s, CpuSlots cs) {
  const long ia = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 5;
  const int l = threadIdx.x & 31;
  if (ia >= n) return;
  const long g = list[ia];
  const long long lo = ps2[g] + 1;
  const int m = (int)(ps2[g + 1] - lo);
  const uint64_t *k2 = QSIDE ? cs.kQ2 : cs.kP2;
  const uint64_t key = l < m ? k2[lo + l] : 0ULL;
  int d = 0;
  for (int i = 0; i < m; i++) {
    const uint64_t ki = __shfl(key, i, 32);
    d += (ki < key) || (ki == key && i < l);
  }
  if (l < m) cpu_move<QSIDE>(s, cs, lo + l, lo + d);
}

// runs of kSmallSort < m <= kBlockSort slots: one block per run from LDS.
// Up to 128 slots each slot's rank is counted directly; longer runs take a
// bitonic sort of (key, slot) pairs padded to a power of two (pads take the
// largest key and slots past the run, so they sort last); the pairs are
// distinct, so the order is the stable one by key either way
template <bool QSIDE>
__global__ void __launch_bounds__(256)
    k_cpu_sort_block(int nRuns, const int *list, const long long *ps2, SlotArrays s,
                     CpuSlots cs) {
  __shared__ uint64_t key[kBlockSort];
  __shared__ int slot[kBlockSort];
  const uint64_t *k2 = QSIDE ? cs.kQ2 : cs.kP2;
  for (int r = blockIdx.x; r < nRuns; r += gridDim.x) {
    const long g = list[r];
    const long long lo = ps2[g] + 1;
    const int m = (int)(ps2[g + 1] - lo);
    __syncthreads();
    if (m <= 128) {
      for (int t = threadIdx.x; t < m; t += blockDim.x) key[t] = k2[lo + t];
      __syncthreads();
      for (int t = threadIdx.x; t < m; t += blockDim.x) {
        const uint64_t kt = key[t];
        int d = 0;
        for (int i = 0; i < m; i++) {
          const uint64_t ki = key[i];
          d += (ki < kt) || (ki == kt && i < t);
        }
        cpu_move<QSIDE>(s, cs, lo + t, lo + d);
      }
      continue;
    }
    int P = 256;
    while (P < m) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) {
      key[t] = t < m ? k2[lo + t] : ~0ULL;
      slot[t] = t;
    }
    __syncthreads();
    for (int size = 2; size <= P; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int t = threadIdx.x; t < P / 2; t += blockDim.x) {
          const int a = 2 * t - (t & (stride - 1)), b = a + stride;
          const uint64_t ka = key[a], kb = key[b];
          const int sa = slot[a], sb = slot[b];
          const bool gt = ka > kb || (ka == kb && sa > sb);
          if (gt == ((a & size) == 0)) {
            key[a] = kb;
            key[b] = ka;
            slot[a] = sb;
            slot[b] = sa;
          }
        }
        __syncthreads();
      }
    }
    for (int d = threadIdx.x; d < m; d += blockDim.x) cpu_move<QSIDE>(s, cs, lo + slot[d], lo + d);
  }
}

// longer runs: the keys of all long runs concatenated (run order, slot order
// inside a run), sorted by key and then stably by run
__global__ void k_cpu_long_keys(int nRuns, const int *list, const long long *runOff,
                                const long long *ps2, const uint64_t *k2, uint64_t *keys,
                                int *vals, int *runOf) {
  for (int r = blockIdx.x; r < nRuns; r += gridDim.x) {
    const long g = list[r];
    const long long lo = ps2[g] + 1, m = ps2[g + 1] - lo, o = runOff[r];
    for (long long t = threadIdx.x; t < m; t += blockDim.x) {
      keys[o + t] = k2[lo + t];
      vals[o + t] = (int)(o + t);
      runOf[o + t] = r;
    }
  }
}

__global__ void k_cpu_long_runkey(long n, const int *vals, const int *runOf, uint64_t *keys) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keys[i] = (uint64_t)runOf[vals[i]];
}

template <bool QSIDE>
__global__ void k_cpu_long_apply(long n, const int *vals, const int *runOf, const int *list,
                                 const long long *runOff, const long long *ps2, SlotArrays s,
                                 CpuSlots cs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = vals[i];
  const int r = runOf[c];
  const long long lo = ps2[list[r]] + 1;
  cpu_move<QSIDE>(s, cs, lo + (c - runOff[r]), lo + (i - runOff[r]));
}

// the unsorted slot of [lo, hi) (partners strictly decreasing) whose partner
// is p, or -1
__device__ __forceinline__ long long find_partner(const int *r2, long long lo, long long hi,
                                                  int p) {
  const long long end = hi;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if ((r2[mid] >> 4) > p) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && (r2[lo] >> 4) == p) ? lo : -1;
}

// Links of the CPU route (global slot + 1 in nbP / nbQ), one thread per
// active edge writing its own slots: an inserted vertex links to its
// counterpart (type 1) or to the other ring's source vertex of the pair; a
// source vertex to the partner of its last link.  A counterpart that is
// missing marks the pair invalid.
__global__ void __launch_bounds__(256)
    k_cpu_link(long nActP, long nActQ, const int *actP, const int *actQ, const int *mapP,
               const int *mapQ, const long long *offP, const long long *offQ,
               const long long *psP2, const long long *psQ2, int *status, SlotArrays s,
               CpuSlots cs) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nActP + nActQ) return;
  const bool sideP = i < nActP;
  const long g = sideP ? actP[i] : actQ[i - nActP];
  const int k = sideP ? mapP[g] : mapQ[g];
  if (status[k] != GH_OK) return;
  const long long aBase = sideP ? offP[k] : offQ[k], bBase = sideP ? offQ[k] : offP[k];
  const int ia = (int)(g - aBase);
  const long long *ps2A = sideP ? psP2 : psQ2, *ps2B = sideP ? psQ2 : psP2;
  const int *rA = sideP ? cs.rP : cs.rQ, *rB2 = sideP ? cs.rQ2 : cs.rP2;
  const int *dB = sideP ? cs.dQ : cs.dP;
  int *nbA = sideP ? s.nbP : s.nbQ;
  bool bad = false;
  // the counterpart on the other ring: the inserted vertex of this pair
  auto counterpart = [&](int q) -> long long {
    const long long qg = bBase + q;
    const long long u = find_partner(rB2, ps2B[qg] + 1, ps2B[qg + 1], ia);
    return u >= 0 ? (long long)dB[u] : -1;
  };
  const long long base = ps2A[g];
  // the slots of wave-filled edges are linked by k_cpu_link_runs
  const long long hi = wave_run(ps2A[g + 1] - base - 1) ? base + 1 : ps2A[g + 1];
  // kListStage slots a round: their records, then the partner edges' slot
  // ranges, are loaded before the round's searches and stores (which write
  // only nbA)
  for (long long d0 = base + 1; d0 < hi; d0 += kListStage) {
    int rec[kListStage];
    long long lo[kListStage], up[kListStage];
#pragma unroll
    for (int u = 0; u < kListStage; u++) rec[u] = d0 + u < hi ? rA[d0 + u] : 0;
#pragma unroll
    for (int u = 0; u < kListStage; u++) {
      if (d0 + u >= hi) continue;
      const long long qg = bBase + (rec[u] >> 4);
      lo[u] = ps2B[qg];
      up[u] = (rec[u] & 15) == 1 ? ps2B[qg + 1] : 0;
    }
#pragma unroll
    for (int u = 0; u < kListStage; u++) {
      if (d0 + u >= hi) continue;
      long long nb = lo[u];
      if ((rec[u] & 15) == 1) {
        const long long w = find_partner(rB2, lo[u] + 1, up[u], ia);
        nb = w >= 0 ? (long long)dB[w] : -1;
      }
      if (nb < 0) bad = true;
      else nbA[d0 + u] = (int)(nb + 1);
    }
  }
  const int src = (sideP ? cs.srcP : cs.srcQ)[g];
  if (src >= 0) {
    const int q = src >> 4, t = src & 15;
    const long long nb = (t == 4 || t == 8) ? ps2B[bBase + q] : counterpart(q);
    if (nb < 0) bad = true;
    else nbA[base] = (int)(nb + 1);
  }
  if (bad) status[k] = GH_INVALID;
}

// k_cpu_fill for the wave-filled edges: lane l visits list entry i0 + l of
// each round; ballots give an inserted vertex its rank among the edge's
// inserts in list order (so its slot is the one the per-edge loop gives it)
// and the source link its last entry in list order.
__global__ void __launch_bounds__(256)
    k_cpu_fill_runs(RunLists L, const int *mapP, const int *mapQ, const long long *offP,
                    const long long *offQ, const double *ringP, const double *ringQ,
                    const long long *psP1, const long long *psQ1, const long long *psP2,
                    const long long *psQ2, const int *status, PairMeta pm,
                    const double *bCoords, const double *oCoords, const int *lstP,
                    const int *lstQ, SlotArrays s, CpuSlots cs) {
  const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  bool sideP;
  long g;
  if (!L.pick(w, sideP, g)) return;
  const int k = sideP ? mapP[g] : mapQ[g];
  if (status[k] != GH_OK) return;
  const int lane = (int)__lane_id();
  const int sw = pm.swapped[k];
  const double *cP = sw ? oCoords : bCoords, *cQ = sw ? bCoords : oCoords;
  const long long aBase = sideP ? offP[k] : offQ[k];
  const long long aN = (sideP ? offP[k + 1] : offQ[k + 1]) - aBase;
  const long long bBase = sideP ? offQ[k] : offP[k];
  const long long bN = (sideP ? offQ[k + 1] : offP[k + 1]) - bBase;
  const double *cA = sideP ? cP : cQ, *cB = sideP ? cQ : cP;
  const long sA = sideP ? pm.startP[k] : pm.startQ[k], sB = sideP ? pm.startQ[k] : pm.startP[k];
  const int ia = (int)(g - aBase);
  const int ia2 = ia == aN - 1 ? 0 : ia + 1;
  const double *ringA = sideP ? ringP : ringQ, *VB = (sideP ? ringQ : ringP) + 2 * bBase;
  const pt A1 = ld(ringA, g), A2 = ld(ringA, aBase + ia2);
  const long long *ps2 = sideP ? psP2 : psQ2, *ps1 = sideP ? psP1 : psQ1;
  const long long base = ps2[g], nIns = ps2[g + 1] - base - 1;
  double *i2 = sideP ? s.iP2 : s.iQ2, *o2 = sideP ? s.oP2 : s.oQ2;
  uint64_t *k2 = sideP ? cs.kP2 : cs.kQ2;
  int *r2 = sideP ? cs.rP2 : cs.rQ2;
  const int *lst = (sideP ? lstP : lstQ) + ps1[g];
  const int cN = (int)(ps1[g + 1] - ps1[g]);
  long long r = 0;
  int src = -1;
  for (int i0 = 0; i0 < cN; i0 += 64) {
    const int i = i0 + lane;
    int q = 0, t = 0;
    double al = 0.0, be = 0.0;
    pt B1{0.0, 0.0}, B2{0.0, 0.0};
    if (i < cN) {
      q = lst[i];
      B1 = ld(VB, q);
      B2 = ld(VB, q + 1 == bN ? 0 : q + 1);
      t = sideP ? itype(A1, A2, B1, B2, al, be) : itype(B1, B2, A1, A2, al, be);
    }
    const bool isSrc = t != 0 && (sideP ? src_on_p(t) : src_on_q(t));
    const bool isIns = t != 0 && (sideP ? ins_on_p(t) : ins_on_q(t));
    const unsigned long long sb = __ballot(isSrc), ib = __ballot(isIns);
    if (sb) src = __shfl(isSrc ? ((q << 4) | t) : 0, 63 - __clzll(sb), 64);
    if (isIns) {
      const long long u = base + nIns - (r + __popcll(ib & ((1ULL << lane) - 1)));
      pt v, vo;
      if (t == 1) {
        const int q2 = q + 1 == bN ? 0 : q + 1;
        const pt P1 = sideP ? A1 : B1, P2 = sideP ? A2 : B2;
        const pt P1o = sideP ? ld(cA, sA + ia) : ld(cB, sB + q);
        const pt P2o = sideP ? ld(cA, sA + ia2) : ld(cB, sB + q2);
        v = pt{(1.0 - al) * P1.x + al * P2.x, (1.0 - al) * P1.y + al * P2.y};
        vo = pt{(1.0 - al) * P1o.x + al * P2o.x, (1.0 - al) * P1o.y + al * P2o.y};
      } else {
        v = B1;
        vo = ld(cB, sB + q);
      }
      if (u > base && u <= base + nIns) {
        st(i2, u, v);
        st(o2, u, vo);
        k2[u] = dbits(sideP ? al : be);
        r2[u] = (q << 4) | t;
      }
    }
    r += __popcll(ib);
  }
  if (lane == 0) (sideP ? cs.srcP : cs.srcQ)[g] = src;
}

// k_cpu_link's slot links for the wave-filled edges, one slot per lane
__global__ void __launch_bounds__(256)
    k_cpu_link_runs(RunLists L, const int *mapP, const int *mapQ, const long long *offP,
                    const long long *offQ, const long long *psP2, const long long *psQ2,
                    int *status, SlotArrays s, CpuSlots cs) {
  const long w = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  bool sideP;
  long g;
  if (!L.pick(w, sideP, g)) return;
  const int k = sideP ? mapP[g] : mapQ[g];
  if (status[k] != GH_OK) return;
  const long long aBase = sideP ? offP[k] : offQ[k], bBase = sideP ? offQ[k] : offP[k];
  const int ia = (int)(g - aBase);
  const long long *ps2A = sideP ? psP2 : psQ2, *ps2B = sideP ? psQ2 : psP2;
  const int *rA = sideP ? cs.rP : cs.rQ, *rB2 = sideP ? cs.rQ2 : cs.rP2;
  const int *dB = sideP ? cs.dQ : cs.dP;
  int *nbA = sideP ? s.nbP : s.nbQ;
  bool bad = false;
  for (long long d = ps2A[g] + 1 + __lane_id(), hi = ps2A[g + 1]; d < hi; d += 64) {
    const int rec = rA[d];
    const long long qg = bBase + (rec >> 4);
    long long nb = ps2B[qg];
    if ((rec & 15) == 1) {
      const long long u = find_partner(rB2, nb + 1, ps2B[qg + 1], ia);
      nb = u >= 0 ? (long long)dB[u] : -1;
    }
    if (nb < 0) bad = true;
    else nbA[d] = (int)(nb + 1);
  }
  if (bad) status[k] = GH_INVALID;
}

// ------------------------------------------------ labelling and tracing ----
enum { L_NONE = 0, L_CROSSING, L_BOUNCING, L_LEFT_ON, L_RIGHT_ON, L_ON_ON, L_ON_LEFT,
       L_ON_RIGHT, L_DELAYED_CROSSING, L_DELAYED_BOUNCING };
enum { EXIT = 0, ENTRY = 1, NEITHER = 2 };
// F_LINK: set by the wave build on every slot whose phase-1 link is set
// (only the parents' is read; the bit means nothing afterwards)
enum { F_SRC = 1, F_INT = 2, F_SPLIT = 4, F_CROSS = 8, F_LINK = 16 };

struct Nodes {  // one pair's node region, local indices
  double *pn, *po;
  int *next, *prev, *nb;
  unsigned char *fl;
  signed char *lab, *enex;
  int n, cap;
  // po == nullptr: the original coordinates are read from the pair's slots,
  // nodes [0, nP0) at oP, the others at oQ (the parallel steps, which keep
  // the nodes in slot order)
  const double *oP = nullptr, *oQ = nullptr;
  int nP0 = 0;
};

__device__ __forceinline__ bool isint(const Nodes &N, int v) { return N.fl[v] & F_INT; }
__device__ __forceinline__ pt P_(const Nodes &N, int v) { return ld(N.pn, v); }
__device__ __forceinline__ pt P_orig(const Nodes &N, int v) {
  return N.po ? ld(N.po, v) : (v < N.nP0 ? ld(N.oP, v) : ld(N.oQ, v - N.nP0));
}

__device__ int node_new(Nodes &N, pt p, pt o) {
  if (N.n >= N.cap) return -1;
  int v = N.n++;
  st(N.pn, v, p);
  st(N.po, v, o);
  N.next[v] = N.prev[v] = N.nb[v] = -1;
  N.fl[v] = 0;
  N.lab[v] = L_NONE;
  N.enex[v] = NEITHER;
  return v;
}

// insertVertex(V, curr) without alpha (polygon.h:101-113): after curr
__device__ void node_insert_after(Nodes &N, int V, int curr) {
  int c = N.next[curr];
  N.next[N.prev[c]] = V;
  N.prev[V] = N.prev[c];
  N.next[V] = c;
  N.prev[c] = V;
}

__device__ void node_link(Nodes &N, int a, int b) {
  N.nb[a] = b;
  N.nb[b] = a;
  N.fl[a] |= F_INT;
  N.fl[b] |= F_INT;
}

// polygon::pointInPoly over edges(ALL), polygon.h:427-448
__device__ bool node_pip(const Nodes &N, int root, pt R) {
  int w = 0, a = root;
  do {
    int b = N.next[a];
    pt P0 = P_(N, a), P1 = P_(N, b);
    if ((P0.y < R.y) != (P1.y < R.y)) {
      if (P0.x >= R.x) {
        if (P1.x > R.x) w = w + 2 * (P1.y > P0.y) - 1;
        else if ((triA(P0, P1, R) > 0) == (P1.y > P0.y)) w = w + 2 * (P1.y > P0.y) - 1;
      } else if (P1.x > R.x) {
        if ((triA(P0, P1, R) > 0) == (P1.y > P0.y)) w = w + 2 * (P1.y > P0.y) - 1;
      }
    }
    a = b;
  } while (a != root);
  return (w % 2) != 0;
}

__device__ __forceinline__ void toggle(int &s) {
  if (s == ENTRY) s = EXIT;
  else if (s == EXIT) s = ENTRY;
}

__device__ int node_rel(const Nodes &N, int Q, int P1, int P2, int P3) {
  return rel_pos(isint(N, P1) && N.nb[P1] == Q, isint(N, P3) && N.nb[P3] == Q, P_(N, Q),
                 P_(N, P1), P_(N, P2), P_(N, P3));
}

struct OutBuf {  // one pair's result region
  double *pn, *po;
  int *next, *prev;
  int n, cap;
  int *compRoot;  // local roots
  int ncomp, capComp;
};

__device__ int out_new(OutBuf &O, pt p, pt o) {
  if (O.n >= O.cap) return -1;
  int v = O.n++;
  st(O.pn, v, p);
  st(O.po, v, o);
  O.next[v] = O.prev[v] = v;
  return v;
}

__device__ void out_append(OutBuf &O, int root, int v) {
  O.prev[v] = O.prev[root];
  O.next[v] = root;
  O.next[O.prev[root]] = v;
  O.prev[root] = v;
}

// copies a whole node ring into the output as one component
__device__ bool out_copy_ring(OutBuf &O, const Nodes &N, int root) {
  if (O.ncomp >= O.capComp) return false;
  int r = -1, v = root;
  do {
    int w = out_new(O, P_(N, v), ld(N.po, v));
    if (w < 0) return false;
    if (r < 0) r = w;
    else out_append(O, r, w);
    v = N.next[v];
  } while (v != root);
  O.compRoot[O.ncomp++] = r;
  return true;
}

// steps 1..6 of labelIntersections, createResult; returns status
// labelsDone: step 1 and step 3 have been applied already (by the group
// path, which computes the same labels in parallel before it falls back).
// pinHint (or null): {V_P, pip(V_P in Q), V_Q, pip(V_Q in P)}, step 4's
// point-in-polygon results for the vertices it names, computed in parallel;
// step 4 uses one only when its own start vertex is the one named.
__device__ int label_and_trace(Nodes &N, int rootP, int rootQ, OutBuf &O, bool labelsDone,
                               const int *pinHint) {
  // step 1 (polyclip.h:343-397, restored) then step 3 (copy labels to Q)
  if (!labelsDone) {
    int I = rootP;
    do {
      const int In = N.next[I];  // loaded with I's fields (no step moves nodes)
      if (isint(N, I) && N.nb[I] >= 0) {
        int Pm = N.prev[I], Pp = In, Qm = N.prev[N.nb[I]], Qp = N.next[N.nb[I]];
        int tm = node_rel(N, Qm, Pm, I, Pp), tp = node_rel(N, Qp, Pm, I, Pp);
        int l = N.lab[I];
        if ((tm == 0 && tp == 1) || (tm == 1 && tp == 0)) l = L_CROSSING;
        if ((tm == 0 && tp == 0) || (tm == 1 && tp == 1)) l = L_BOUNCING;
        if ((tp == 3 && tm == 1) || (tm == 3 && tp == 1)) l = L_LEFT_ON;
        if ((tp == 3 && tm == 0) || (tm == 3 && tp == 0)) l = L_RIGHT_ON;
        if ((tp == 3 && tm == 2) || (tm == 3 && tp == 2)) l = L_ON_ON;
        if ((tm == 2 && tp == 1) || (tp == 2 && tm == 1)) l = L_ON_LEFT;
        if ((tm == 2 && tp == 0) || (tp == 2 && tm == 0)) l = L_ON_RIGHT;
        N.lab[I] = (signed char)l;
      }
      I = In;
    } while (I != rootP);
    I = rootP;
    do {
      const int In = N.next[I];
      if (isint(N, I) && N.nb[I] >= 0) N.lab[N.nb[I]] = N.lab[I];
      I = In;
    } while (I != rootP);
  }
  // step 2: intersection chains (:398-451)
  {
    int it = rootP;
    do {
      const int itn = N.next[it];  // the chain below changes labels only
      if (isint(N, it) && (N.lab[it] == L_LEFT_ON || N.lab[it] == L_RIGHT_ON)) {
        int I = it;
        int x = N.lab[I] == L_LEFT_ON ? 0 : 1;
        int X = I;
        do {
          N.lab[I] = L_NONE;
          if (N.nb[I] >= 0) N.lab[N.nb[I]] = L_NONE;
          I = N.next[I];
        } while (N.lab[I] == L_ON_ON);
        int y = N.lab[I] == L_ON_LEFT ? 0 : 1;
        signed char ct = (x != y) ? L_DELAYED_CROSSING : L_DELAYED_BOUNCING;
        N.lab[X] = ct;
        N.lab[I] = ct;
        if (N.nb[X] >= 0) N.lab[N.nb[X]] = ct;
        if (N.nb[I] >= 0) N.lab[N.nb[I]] = ct;
      }
      it = itn;
    } while (it != rootP);
  }
  // step 3.5: components without crossings (:474-543)
  bool noInter[2] = {false, false}, ident[2] = {false, false};
  for (int i = 0; i < 2; i++) {
    int Pr = i == 0 ? rootP : rootQ, Qr = i == 0 ? rootQ : rootP;
    bool nc = true;
    int V = Pr;
    do {
      if (isint(N, V) && (N.lab[V] == L_CROSSING || N.lab[V] == L_DELAYED_CROSSING)) {
        nc = false;
        break;
      }
      V = N.next[V];
    } while (V != Pr);
    if (!nc) continue;
    noInter[i] = true;
    bool oo = true;
    V = Pr;
    do {
      if (N.lab[V] != L_ON_ON) {
        oo = false;
        break;
      }
      V = N.next[V];
    } while (V != Pr);
    if (oo) {
      ident[i] = true;
      continue;
    }
    // getNonIntersectionPoint (polygon.h:480-491)
    pt p;
    bool have = false;
    V = Pr;
    do {
      if (!isint(N, V)) {
        p = P_(N, V);
        have = true;
        break;
      }
      V = N.next[V];
    } while (V != Pr);
    if (!have) {
      V = Pr;
      do {
        int n = N.next[V], nb = N.nb[V], nnb = N.nb[n];
        if (nb >= 0 && nnb != N.prev[nb] && nnb != N.next[nb]) {
          pt a = P_(N, V), b = P_(N, n);
          pt sm = {a.x + b.x, a.y + b.y};
          p = pt{0.5 * sm.x, 0.5 * sm.y};
          have = true;
          break;
        }
        V = n;
      } while (V != Pr);
    }
    if (!have) return GH_TRACE_FAIL;
    if (node_pip(N, Qr, p))
      if (!out_copy_ring(O, N, Pr)) return GH_OVERFLOW;
  }
  if (ident[0] && ident[1]) {
    int V = rootQ, target = N.nb[rootP];
    bool found = false;
    do {
      if (V == target) {
        found = true;
        break;
      }
      V = N.next[V];
    } while (V != rootQ);
    if (found && !out_copy_ring(O, N, rootP)) return GH_OVERFLOW;
  }
  // step 4: entry/exit (:549-655)
  for (int i = 0; i < 2; i++) {
    if (noInter[i]) continue;
    int Pr = i == 0 ? rootP : rootQ, Qr = i == 0 ? rootQ : rootP;
    int V = -1;
    {
      int W = Pr;
      do {
        if (!isint(N, W)) {
          V = W;
          break;
        }
        W = N.next[W];
      } while (W != Pr);
      if (V < 0) {  // getNonIntersectionVertex: temporary midpoint vertex
        W = Pr;
        do {
          int n = N.next[W], nb = N.nb[W], nnb = N.nb[n];
          if (nb >= 0 && nnb != N.prev[nb] && nnb != N.next[nb]) {
            pt a = P_(N, W), b = P_(N, n), ao = ld(N.po, W), bo = ld(N.po, n);
            pt sm = {a.x + b.x, a.y + b.y}, so = {ao.x + bo.x, ao.y + bo.y};
            int T = node_new(N, pt{0.5 * sm.x, 0.5 * sm.y}, pt{0.5 * so.x, 0.5 * so.y});
            if (T < 0) return GH_OVERFLOW;
            node_insert_after(N, T, W);
            V = T;
            break;
          }
          W = n;
        } while (W != Pr);
      }
      if (V < 0) return GH_TRACE_FAIL;
    }
    int status = ENTRY;
    const bool pin = (pinHint && pinHint[2 * i] == V) ? pinHint[2 * i + 1] != 0
                                                      : node_pip(N, Qr, P_(N, V));
    if (pin) toggle(status);
    bool first = true;
    int I = V;
    do {
      I = N.next[I];
    } while (!isint(N, I) && I != V);
    // the walk loads each vertex's flags, label and successor together (one
    // wait per vertex); the stores touch only the current vertex's enex,
    // lab and fl
    int lI = N.lab[I], In = N.next[I];
    while (I != V) {
      int l = lI;
      if (l == L_CROSSING) {
        N.enex[I] = (signed char)status;
        toggle(status);
      }
      if (N.lab[I] == L_BOUNCING && status == EXIT) N.fl[I] |= F_SPLIT;
      if (N.lab[I] == L_DELAYED_CROSSING) {
        N.enex[I] = (signed char)status;
        if (first) {
          if (status == EXIT) N.lab[I] = L_CROSSING;
          first = false;
        } else {
          if (status == ENTRY) N.lab[I] = L_CROSSING;
          first = true;
          toggle(status);
        }
      }
      if (N.lab[I] == L_DELAYED_BOUNCING) {
        N.enex[I] = (signed char)status;
        if (first) {
          if (status == EXIT) N.fl[I] |= F_CROSS;
          first = false;
        } else {
          if (status == ENTRY) N.fl[I] |= F_CROSS;
          first = true;
        }
        toggle(status);
      }
      I = In;
      unsigned char fI = N.fl[I];
      lI = N.lab[I];
      In = N.next[I];
      while (!(fI & F_INT) && I != V) {
        I = In;
        fI = N.fl[I];
        lI = N.lab[I];
        In = N.next[I];
      }
    }
  }
  // step 5: split bouncing vertex pairs (:661-714); step 6 (:722-734)
  {
    const int nNodes = N.n;
    int I = rootP;
    do {
      int nextI = N.next[I];
      if ((N.fl[I] & F_SPLIT) && N.nb[I] >= 0 && (N.fl[N.nb[I]] & F_SPLIT)) {
        int IQ = N.nb[I];
        int VP = node_new(N, P_(N, I), ld(N.po, I));
        int VQ = node_new(N, P_(N, IQ), ld(N.po, IQ));
        if (VP < 0 || VQ < 0) return GH_OVERFLOW;
        double sP = triA(P_(N, N.prev[I]), P_(N, I), P_(N, N.next[I]));
        double sQ = triA(P_(N, N.prev[IQ]), P_(N, IQ), P_(N, N.next[IQ]));
        if (sP * sQ > 0) {
          node_link(N, I, VQ);
          node_link(N, IQ, VP);
        } else {
          node_link(N, VP, VQ);
        }
        node_insert_after(N, VP, I);
        node_insert_after(N, VQ, IQ);
        N.enex[I] = EXIT;
        N.enex[VP] = ENTRY;
        N.enex[IQ] = EXIT;
        N.enex[VQ] = ENTRY;
        N.lab[I] = N.lab[VP] = N.lab[IQ] = N.lab[VQ] = L_CROSSING;
        N.fl[I] &= ~F_SPLIT;  // handled
      }
      I = nextI;
    } while (I != rootP);
    (void)nNodes;
    I = rootP;
    do {
      const int In = N.next[I];  // labels only change below
      if ((N.fl[I] & F_CROSS) && N.nb[I] >= 0 && (N.fl[N.nb[I]] & F_CROSS)) {
        N.lab[I] = L_CROSSING;
        N.lab[N.nb[I]] = L_CROSSING;
      }
      I = In;
    } while (I != rootP);
  }
  // createResult (:743-783)
  {
    const long long guard = 4LL * N.n + 64;
#define IS_CI(v) (isint(N, (v)) && N.lab[(v)] == L_CROSSING)
    int I = rootP;
    bool have = IS_CI(I);
    if (!have) {
      do {
        I = N.next[I];
      } while (!IS_CI(I) && I != rootP);
      have = I != rootP;
    }
    // a trace that outgrows the output region keeps walking without writing
    // so that a runaway trace still ends in the guard's failure status
    bool full = false;
    while (have) {
      if (O.ncomp >= O.capComp) return GH_OVERFLOW;
      // tail: the component's last vertex (O.prev[R]), kept in a register
      int R = -1, tail = -1, V = I;
      N.fl[V] &= ~F_INT;
      long long steps = 0;
      // (V, status) after a move fixes the rest of the walk, so a repeated
      // pair means the walk never returns to I and the guard will fail it:
      // Brent's cycle detection ends such a walk early with the same status
      int savedV = -1, savedS = -1;
      long long power = 1, lam = 0;
      do {
        int status = N.enex[V];
        toggle(status);
        // the walk is a chain of dependent loads: every load that depends on
        // the current vertex alone (its status, its move, its coordinates,
        // its flags) is issued before the step's stores, so that a step
        // waits on memory once
        const int *mv = status == EXIT ? N.next : N.prev;
        int Vn = mv[V], enV;
        do {
          V = Vn;
          enV = N.enex[V];
          Vn = mv[V];
          const pt pv = P_(N, V), ov = ld(N.po, V);
          const unsigned char fv = N.fl[V];
          N.fl[V] = (unsigned char)(fv & ~F_INT);
          if (!full) {
            int w = out_new(O, pv, ov);
            if (w < 0) {
              full = true;
            } else if (R < 0) {
              R = tail = w;
            } else {
              // out_append(O, R, w) with O.prev[R] == tail
              O.prev[w] = tail;
              O.next[w] = R;
              O.next[tail] = w;
              O.prev[R] = w;
              tail = w;
            }
          }
          if (++steps > guard) return GH_TRACE_FAIL;
          if (V == savedV && status == savedS) return GH_TRACE_FAIL;
          if (lam == power) {
            savedV = V;
            savedS = status;
            power <<= 1;
            lam = 0;
          }
          lam++;
        } while (!(enV == status) && V != I);
        if (V != I) {
          V = N.nb[V];
          if (V < 0) return GH_TRACE_FAIL;
          N.fl[V] &= ~F_INT;
        }
      } while (V != I);
      if (full) return GH_OVERFLOW;
      O.compRoot[O.ncomp++] = R;
      do {
        I = N.next[I];
      } while (!IS_CI(I) && I != rootP);
      have = I != rootP;
    }
#undef IS_CI
  }
  return GH_OK;
}

// cleanUpResult (:791-807) on one output component; returns the new root
__device__ int out_cleanup(OutBuf &O, int r) {
  while (r >= 0 &&
         fabs(triA(ld(O.pn, O.prev[r]), ld(O.pn, r), ld(O.pn, O.next[r]))) < kGhEps) {
    int V = r;
    r = O.next[V];
    if (O.next[r] == r) r = -1;
    O.next[O.prev[V]] = O.next[V];
    O.prev[O.next[V]] = O.prev[V];
  }
  if (r < 0) return -1;
  int it = r;
  do {
    int nx = O.next[it];
    if (fabs(triA(ld(O.pn, O.prev[it]), ld(O.pn, it), ld(O.pn, O.next[it]))) < kGhEps) {
      O.next[O.prev[it]] = O.next[it];
      O.prev[O.next[it]] = O.prev[it];
    }
    it = nx;
  } while (it != r);
  return r;
}

struct TraceArgs {
  const long long *offP, *offQ, *psP2, *psQ2;
  const long long *nodeOff;  // [M+1] node / output capacity prefix
  const long long *compOff;  // [M+1] component capacity prefix
  double *npn, *npo;
  int *nnext, *nprev, *nnb;
  unsigned char *nfl;
  signed char *nlab, *nenex;
  double *opn, *opo;
  int *onext, *oprev;
  int *ocompRoot;
  int *outNComp, *outNVerts;
  int *compLen;  // [comp capacity] final component sizes
  double *xyTmp; // [node capacity] compacted result vertices per pair
  int cleanup;
  long long bigNodes;  // pairs with this many nodes go to k_label_block
  int gridClean;       // components longer than this in a grid-labelled pair: grid clean-up
  // links given on both rings (nbP and nbQ, the CPU route); otherwise only
  // the Q slots' links are read and applied with link() (the GPU route)
  int explicitLinks;
};

__global__ void __launch_bounds__(64)
    k_label_trace(long M, int *status, SlotArrays s, TraceArgs a) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  a.outNComp[k] = 0;
  a.outNVerts[k] = 0;
  if (status[k] != GH_OK) return;
  const long long sbP = a.psP2[a.offP[k]], cP = a.psP2[a.offP[k + 1]] - sbP;
  const long long sbQ = a.psQ2[a.offQ[k]], cQ = a.psQ2[a.offQ[k + 1]] - sbQ;
  const long long nb0 = a.nodeOff[k], cap = a.nodeOff[k + 1] - nb0;
  Nodes N;
  N.pn = a.npn + 2 * nb0;
  N.po = a.npo + 2 * nb0;
  N.next = a.nnext + nb0;
  N.prev = a.nprev + nb0;
  N.nb = a.nnb + nb0;
  N.fl = a.nfl + nb0;
  N.lab = a.nlab + nb0;
  N.enex = a.nenex + nb0;
  N.n = 0;
  N.cap = (int)cap;
  // P list from the P slots, parents flagged by their alpha (ghcuda.cpp:246-273)
  const long long pE0 = a.offP[k], pE1 = a.offP[k + 1];
  long long pe = pE0;
  for (long long j = 0; j < cP; j++) {
    const long long sl = sbP + j;
    int v = node_new(N, ld(s.iP, sl), ld(s.oP, sl));
    N.prev[v] = v == 0 ? (int)(cP - 1) : v - 1;
    N.next[v] = v == cP - 1 ? 0 : v + 1;
    if (pe < pE1 && a.psP2[pe] == sl) {
      const bool linked = a.explicitLinks ? s.nbP[sl] > 0 : s.aP[sl] != kNone;
      N.fl[v] = F_SRC | (linked ? F_INT : 0);
      pe++;
    } else {
      N.fl[v] = F_INT;
    }
    if (a.explicitLinks) {
      const long long kq = (long long)s.nbP[sl] - 1 - sbQ;
      N.nb[v] = (kq >= 0 && kq < cQ) ? (int)(cP + kq) : -1;
    }
  }
  // Q list; parents are linked when phase 1 recorded a neighbour
  // (ghcuda.cpp:277-322, with X-overlap parents linked as link() does)
  const long long qE0 = a.offQ[k], qE1 = a.offQ[k + 1];
  long long qe = qE0;
  for (long long j = 0; j < cQ; j++) {
    const long long sl = sbQ + j;
    int v = node_new(N, ld(s.iQ, sl), ld(s.oQ, sl));
    N.prev[v] = j == 0 ? (int)(cP + cQ - 1) : v - 1;
    N.next[v] = j == cQ - 1 ? (int)cP : v + 1;
    bool linked;
    if (qe < qE1 && a.psQ2[qe] == sl) {
      N.fl[v] = F_SRC;
      linked = s.aQ[sl] != kNone || s.nbQ[sl] > 0;
      qe++;
    } else {
      N.fl[v] = 0;
      linked = true;
    }
    if (a.explicitLinks) {
      const long long kp = (long long)s.nbQ[sl] - 1 - sbP;
      N.nb[v] = (kp >= 0 && kp < cP) ? (int)kp : -1;
      if (linked) N.fl[v] |= F_INT;
    } else if (linked) {
      long long kp = (long long)s.nbQ[sl] - 1 - sbP;
      if (kp >= 0 && kp < cP) node_link(N, v, (int)kp);
      else N.fl[v] |= F_INT;
    }
  }
  OutBuf O;
  O.pn = a.opn + 2 * nb0;
  O.po = a.opo + 2 * nb0;
  O.next = a.onext + nb0;
  O.prev = a.oprev + nb0;
  O.n = 0;
  O.cap = (int)cap;
  O.compRoot = a.ocompRoot + a.compOff[k];
  O.ncomp = 0;
  O.capComp = (int)(a.compOff[k + 1] - a.compOff[k]);
  int rc = label_and_trace(N, 0, (int)cP, O, false, nullptr);
  if (rc != GH_OK) {
    status[k] = rc;
    return;
  }
  // clean-up and compaction into xyTmp / compLen (per pair regions)
  int nc = 0, nv = 0;
  double *xy = a.xyTmp + 2 * nb0;
  int *clen = a.compLen + a.compOff[k];
  for (int c = 0; c < O.ncomp; c++) {
    int r = O.compRoot[c];
    if (a.cleanup) r = out_cleanup(O, r);
    if (r < 0) continue;
    int len = 0, v = r;
    do {
      st(xy, nv + len, ld(O.po, v));
      len++;
      v = O.next[v];
    } while (v != r);
    clen[nc++] = len;
    nv += len;
  }
  a.outNComp[k] = nc;
  a.outNVerts[k] = nv;
}

// ---------------------------------------- wave-parallel labelling path ----
// nodes per lane whose loads a staged pass issues before its stores
constexpr int kStage = 4;
constexpr int kStageWave = 8;  // the list build's, for one-wave groups
constexpr int kRingStage = 2;  // the same for the ring writes of segment_trace
constexpr int kRingStageWave = 2;
// One wave64 per pair.  The lists are built in parallel; labelling step 1/3
// runs per vertex; when every intersection vertex is a plain CROSSING (no
// bouncing, chain or overlap labels, mutual links, a non-intersection vertex
// on both rings) the remaining steps reduce to prefix parities and list
// ranking over crossing segments:
//   step 4: enex of the i-th intersection vertex after the start vertex V is
//           the start status toggled by the parity of the intersections
//           before it (ballot prefix counts);
//   createResult: every crossing vertex c departs along its ring to the
//           adjacent crossing A(c) in its direction of travel (one segment)
//           and continues from the neighbour of A(c).  Those successor links
//           form cycles, each result component and its reverse; pointer
//           jumping finds each cycle's smallest P crossing (where the
//           sequential loop over P's crossing vertices starts it) and the
//           output offset of every segment, and the ring vertices are then
//           written in one flat pass.
// Every other pair (degenerate labels, inconsistent links, a point met twice
// by one cycle) runs the sequential code above on lane 0, so results are
// identical either way.

__device__ __forceinline__ int wave_sum(int v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int wave_min(int v) {
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ int lt_count(unsigned long long m) {
  return __popcll(m & ((1ULL << __lane_id()) - 1ULL));
}

// Group-wide reductions and scans.  A group is the W threads of one block,
// one wave (W = 64, pairs of ordinary size) or sixteen (W = 1024, large
// pairs), or, for W = kGrid, every thread of a cooperative launch (pairs of
// a million nodes and more): grp_rank / grp_size / grp_sync stand for the
// thread's index, the group's size and its barrier.  Every thread of the
// group must call them, with control flow uniform over the group; the LDS
// slots are guarded by a barrier on entry.  A grid-wide reduction leaves
// one partial per block in the partials buffer (grid_part, two halves used
// in turn, so a block never overwrites a half another block still reads),
// syncs the grid, and lets every block combine the partials it needs.
constexpr int kGrid = 0;

__shared__ long long *grid_part;  // [2 * gridDim.x], set by the grid kernel
__shared__ int grid_half;         // the half the next grid reduction uses

template <int W>
__device__ __forceinline__ int grp_rank() {
  return W == kGrid ? (int)(blockIdx.x * blockDim.x + threadIdx.x) : (int)threadIdx.x;
}
template <int W>
__device__ __forceinline__ int grp_size() {
  return W == kGrid ? (int)(gridDim.x * blockDim.x) : W;
}
template <int W>
__device__ __forceinline__ void grp_sync() {
  if constexpr (W == kGrid) cooperative_groups::this_grid().sync();
  else __syncthreads();
}

// block-wide sum of v (blockDim.x <= 1024); every thread of the block calls it
__device__ __forceinline__ long long block_sum_ll(long long v) {
  __shared__ long long red[16];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  long long t = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) t += red[i];
  return t;
}

// the grid-wide combination of one value per block (its block total b):
// returns the sum over blocks before this one and sets total; op 0 sums,
// op 1 takes the minimum (before then unused)
__device__ long long grid_combine(long long b, int op, long long &total) {
  long long *part = grid_part + (long long)grid_half * gridDim.x;
  if (threadIdx.x == 0) part[blockIdx.x] = b;
  cooperative_groups::this_grid().sync();
  long long before = 0, all = op == 0 ? 0 : 0x7fffffffffffffffLL;
  for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) {
    const long long x = part[i];
    if (op == 0) {
      all += x;
      if (i < (int)blockIdx.x) before += x;
    } else {
      all = min(all, x);
    }
  }
  if (op == 0) {
    before = block_sum_ll(before);
    all = block_sum_ll(all);
  } else {
    __shared__ long long red[16];
    for (int off = 32; off > 0; off >>= 1) all = min(all, __shfl_xor(all, off, 64));
    __syncthreads();
    if (__lane_id() == 0) red[threadIdx.x >> 6] = all;
    __syncthreads();
    all = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) all = min(all, red[i]);
  }
  __syncthreads();
  if (threadIdx.x == 0) grid_half ^= 1;
  __syncthreads();
  total = all;
  return before;
}

template <int W>
__device__ __forceinline__ int grp_sum(int v) {
  v = wave_sum(v);
  if constexpr (W == 64) {
    return v;
  } else if constexpr (W == kGrid) {
    long long total;
    grid_combine(block_sum_ll(__lane_id() == 0 ? v : 0), 0, total);
    return (int)total;
  } else {
    __shared__ int red[W / 64];
    __syncthreads();
    if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = 0;
    for (int i = 0; i < W / 64; i++) t += red[i];
    return t;
  }
}
template <int W>
__device__ __forceinline__ int grp_min(int v) {
  v = wave_min(v);
  if constexpr (W == 64) {
    return v;
  } else if constexpr (W == kGrid) {
    __shared__ int red[16];
    __syncthreads();
    if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); i++) t = min(t, red[i]);
    long long total;
    grid_combine(t, 1, total);
    return (int)total;
  } else {
    __shared__ int red[W / 64];
    __syncthreads();
    if (__lane_id() == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    int t = red[0];
    for (int i = 1; i < W / 64; i++) t = min(t, red[i]);
    return t;
  }
}
template <int W>
__device__ __forceinline__ bool grp_any(bool b) {
  if (W == 64) return __ballot(b) != 0ULL;
  return grp_sum<W>(__ballot(b) != 0ULL && __lane_id() == 0 ? 1 : 0) != 0;
}
// exclusive prefix of v over the group in rank order; total: the sum
template <int W>
__device__ __forceinline__ int grp_scan(int v, int &total) {
  const int lane = __lane_id();
  int incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const int t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  if constexpr (W == 64) {
    total = __shfl(incl, 63, 64);
    return incl - v;
  } else {
    constexpr int kParts = W == kGrid ? 16 : W / 64;
    const int nw = W == kGrid ? (int)(blockDim.x >> 6) : W / 64;
    __shared__ int part[kParts];
    __syncthreads();
    if (lane == 63) part[threadIdx.x >> 6] = incl;
    __syncthreads();
    const int w = threadIdx.x >> 6;
    int before = 0, tot = 0;
    for (int i = 0; i < nw; i++) {
      if (i < w) before += part[i];
      tot += part[i];
    }
    if constexpr (W == kGrid) {
      long long all;
      const long long pre = grid_combine(tot, 0, all);
      total = (int)all;
      return (int)pre + before + incl - v;
    } else {
      total = tot;
      return before + incl - v;
    }
  }
}

// exclusive prefix of v over the block in thread order; total: the block's sum
__device__ __forceinline__ long long block_scan_ll(long long v, long long &total) {
  const int lane = __lane_id();
  long long incl = v;
  for (int o = 1; o < 64; o <<= 1) {
    const long long t = __shfl_up(incl, o, 64);
    if (lane >= o) incl += t;
  }
  __shared__ long long part[16];
  __syncthreads();
  if (lane == 63) part[threadIdx.x >> 6] = incl;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  long long before = 0, tot = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); i++) {
    if (i < w) before += part[i];
    tot += part[i];
  }
  total = tot;
  return before + incl - v;
}

// grp_scan for int or long long values (a wave or one block)
template <int W, typename T>
__device__ __forceinline__ T grp_scan_t(T v, T &total) {
  if constexpr (sizeof(T) == sizeof(int)) {
    int tot;
    const int pre = grp_scan<W>((int)v, tot);
    total = tot;
    return pre;
  } else if constexpr (W == 64) {
    const int lane = __lane_id();
    long long incl = v;
    for (int o = 1; o < 64; o <<= 1) {
      const long long t = __shfl_up(incl, o, 64);
      if (lane >= o) incl += t;
    }
    total = __shfl(incl, 63, 64);
    return incl - v;
  } else {
    long long tot;
    const long long pre = block_scan_ll(v, tot);
    total = tot;
    return pre;
  }
}

// The range [0, n) in rank order with the exclusive running sum of v: calls
// f(j, prefix, x) for every j, where x = v(j) and prefix is the sum of v over
// [0, j); returns the sum over [0, n).  A wave or block walks the range in
// steps of its size.  The whole grid gives each block one contiguous tile:
// v is summed over every tile first (one grid-wide combination), then
// evaluated again as each block walks its tile in blockDim steps with
// block-wide scans, so the grid needs one grid barrier where stepping by
// its size needs one per step.  v must not read what f writes in another
// tile, and f's stores must not reach another tile's reads.
template <int W, typename T, class V, class F>
__device__ T grp_scan_range(int n, V v, F f) {
  T carry = 0;
  if constexpr (W == kGrid) {
    const int tile = (int)(((long long)n + gridDim.x - 1) / gridDim.x);
    const int t0 = (int)min((long long)blockIdx.x * tile, (long long)n), t1 = min(t0 + tile, n);
    long long s = 0;
    for (int j = t0 + (int)threadIdx.x; j < t1; j += blockDim.x) s += (long long)v(j);
    long long total;
    carry = (T)grid_combine(block_sum_ll(s), 0, total);
    for (int b = t0; b < t1; b += blockDim.x) {
      const int j = b + (int)threadIdx.x;
      const T x = j < t1 ? v(j) : T(0);
      long long tot;
      const T pre = (T)block_scan_ll((long long)x, tot);
      if (j < t1) f(j, (T)(carry + pre), x);
      carry += (T)tot;
    }
    return (T)total;
  } else {
    for (int b = 0; b < n; b += W) {
      const int j = b + grp_rank<W>();
      const T x = j < n ? v(j) : T(0);
      T tot;
      const T pre = grp_scan_t<W, T>(x, tot);
      if (j < n) f(j, (T)(carry + pre), x);
      carry += tot;
    }
    return carry;
  }
}

// parity of the winding number of R over the ring [r0, r0+n) in index order
template <int W>
__device__ bool grp_pip(const Nodes &N, int r0, int n, pt R) {
  int w = 0;
  for (int a = grp_rank<W>(); a < n; a += grp_size<W>()) {
    int ia = r0 + a, ib = r0 + (a + 1 == n ? 0 : a + 1);
    pt P0 = P_(N, ia), P1 = P_(N, ib);
    if ((P0.y < R.y) != (P1.y < R.y)) {
      if (P0.x >= R.x) {
        if (P1.x > R.x) w += 2 * (P1.y > P0.y) - 1;
        else if ((triA(P0, P1, R) > 0) == (P1.y > P0.y)) w += 2 * (P1.y > P0.y) - 1;
      } else if (P1.x > R.x) {
        if ((triA(P0, P1, R) > 0) == (P1.y > P0.y)) w += 2 * (P1.y > P0.y) - 1;
      }
    }
  }
  return (grp_sum<W>(w) & 1) != 0;
}

// index of the first node of [r0, r0+n) without the intersection flag, or -1
template <int W>
__device__ int grp_first_nonint(const Nodes &N, int r0, int n) {
  int best = 0x7fffffff;
  for (int a = grp_rank<W>(); a < n; a += grp_size<W>())
    if (!isint(N, r0 + a)) best = min(best, a);
  best = grp_min<W>(best);
  return best == 0x7fffffff ? -1 : r0 + best;
}

__device__ __forceinline__ bool iscross(const Nodes &N, int v) {
  return isint(N, v) && N.lab[v] == L_CROSSING;
}

// step 4 for one ring whose intersection vertices are CROSSING or BOUNCING:
// the status at a vertex is the start status toggled once per crossing
// before it; crossings take it as enex, bouncing vertices met with status
// EXIT are flagged for splitting
template <int W>
__device__ void grp_enex(Nodes &N, int r0, int n, int V, int status0) {
  // i-th vertex after V in ring order, i = 1..n-1
  const int v0 = V - r0;
  auto node = [&](int i) {
    const int u = v0 + i;
    return r0 + (u >= n ? u - n : u);
  };
  grp_scan_range<W, int>(
      n - 1,
      [&](int j) {
        const int u = node(j + 1);
        return isint(N, u) && N.lab[u] == L_CROSSING ? 1 : 0;
      },
      [&](int j, int pre, int cr) {
        const int u = node(j + 1);
        if (!isint(N, u)) return;
        const int status = (pre & 1) ? (status0 == ENTRY ? EXIT : ENTRY) : status0;
        if (cr) N.enex[u] = (signed char)status;
        else if (status == EXIT) N.fl[u] |= F_SPLIT;
      });
}

// Departures of createResult (polyclip.h:743-783).  Every CROSSING vertex
// is one.  A BOUNCING pair flagged for splitting on both rings gives two at
// its vertex u: u itself (enex EXIT) and the copy that step 5
// (polyclip.h:661-714) inserts after u (enex ENTRY).  Departures 0..XP-1 are
// P's in ring order, XP..XP+XQ-1 Q's; dn[d] = 2 u + (second at u), rk[u] =
// the first departure at u.  A departure moves to next (enex ENTRY: status
// EXIT after the toggle) or prev and stops at the adjacent departure of its
// ring, `dist` vertices on (arrival included).
__device__ __forceinline__ bool issplit(const Nodes &N, int u) {
  return isint(N, u) && N.lab[u] == L_BOUNCING && (N.fl[u] & F_SPLIT) && N.nb[u] >= 0 &&
         (N.fl[N.nb[u]] & F_SPLIT);
}
__device__ __forceinline__ int dmult(const Nodes &N, int u) {
  return iscross(N, u) ? 1 : (issplit(N, u) ? 2 : 0);
}

// step 5's link choice for the split pair at P vertex I: sP * sQ > 0 links
// each vertex to the other ring's copy.  Copies inserted by earlier splits
// repeat their vertex's coordinates, so the original neighbours give the
// same areas.
__device__ __forceinline__ bool split_cross(const Nodes &N, int I, int nP, int nQ) {
  const int IQ = N.nb[I];
  const int Pm = I == 0 ? nP - 1 : I - 1, Pp = I == nP - 1 ? 0 : I + 1;
  const int Qm = IQ == nP ? nP + nQ - 1 : IQ - 1, Qp = IQ == nP + nQ - 1 ? nP : IQ + 1;
  const double sP = triA(P_(N, Pm), P_(N, I), P_(N, Pp));
  const double sQ = triA(P_(N, Qm), P_(N, IQ), P_(N, Qp));
  return sP * sQ > 0;
}

__device__ __forceinline__ int dep_enex(const Nodes &N, int code) {
  const int u = code >> 1;
  return N.lab[u] == L_CROSSING ? N.enex[u] : ((code & 1) ? ENTRY : EXIT);
}

// departure of the neighbour of departure `code`
__device__ __forceinline__ int dep_partner(const Nodes &N, const int *rk, int code, int nP,
                                           int nQ) {
  const int u = code >> 1, q = N.nb[u];
  if (N.lab[u] == L_CROSSING) return rk[q];
  const bool cross = u < nP ? split_cross(N, u, nP, nQ) : split_cross(N, q, nP, nQ);
  return rk[q] + (((code & 1) == 0) == cross ? 1 : 0);
}

struct Departure {
  int code, arr, dist;
  bool fwd;
};

__device__ __forceinline__ Departure departure(const Nodes &N, const int *dn, int d, int nP,
                                               int nQ, int XP, int XQ) {
  const bool onP = d < XP;
  const int base = onP ? 0 : XP, XR = onP ? XP : XQ, nR = onP ? nP : nQ, i = d - base;
  Departure r;
  r.code = dn[d];
  r.fwd = dep_enex(N, r.code) == ENTRY;
  r.arr = base + (r.fwd ? (i + 1 == XR ? 0 : i + 1) : (i == 0 ? XR - 1 : i - 1));
  const int u = r.code >> 1, ua = dn[r.arr] >> 1;
  int dist = r.fwd ? ua - u : u - ua;
  if (dist <= 0) dist += nR;
  r.dist = dist;
  return r;
}

enum { SEG_DONE = 0, SEG_SLOW = 2, SEG_OVERFLOW = 3 };

__host__ __device__ constexpr int even_up(int v) { return (v + 1) & ~1; }
// ints of segment scratch for a pair of n nodes and at most Dc departures
// (Dc even): rg[0, n) and Dc-sized departure arrays for the trace, and for
// the clean-up after it the position map rg[0, n) plus 4 ints per component
// (at most Dc / 2 + 2) and one per output position (at most n).  Within the
// 10 n + 20 ints of a pair's global region whenever Dc <= even_up(n).
__host__ __device__ constexpr int seg_scratch_ints(int n, int Dc) {
  return even_up(n) + (9 * Dc > even_up(n) + 2 * Dc + 16 ? 9 * Dc : even_up(n) + 2 * Dc + 16);
}

// A bump allocator over a wave's LDS: the arrays a pair's passes revisit
// are placed in it while they fit, 16-byte aligned; take() returns the
// array's place in the pair's global region once the LDS is full.  Every
// lane makes the same calls, so the placement is uniform.
struct Arena {
  char *p;
  int left;
  // reserve: bytes that must stay free for later arrays
  template <typename T>
  __device__ T *take(long count, T *global, long reserve = 0) {
    const long b = ((long)sizeof(T) * count + 15) & ~15L;
    if (!p || b + reserve > left) return global;
    T *r = (T *)p;
    p += b;
    left -= (int)b;
    return r;
  }
};

// createResult for a pair whose intersection vertices are CROSSING or
// BOUNCING (passed over like ordinary vertices unless split), with mutual
// links and enex alternating over the departures of both rings.  Dc: an
// even bound on the number of departures (2 per intersection vertex at
// most); `rg` holds seg_scratch_ints(n, Dc) ints, 16-byte aligned.  Writes
// the components (roots in O.compRoot, lengths in clen, result vertices in
// xy; omap and rg[0, n) map output positions to their nodes and
// components) in the sequential order.  Returns SEG_DONE, SEG_SLOW when the
// pair needs the sequential trace, or SEG_OVERFLOW.
template <int W>
__device__ int segment_trace(const Nodes &N, OutBuf &O, double *xy, int *clen, int *rg, int nP,
                             int nQ, int Dc, int *omap) {
  const int lane = grp_rank<W>();
  const int n = nP + nQ, nE = even_up(n);
  // rk: per node; the rest per departure
  int *rk = rg, *dn = rg + nE;
  int2 *vc[2] = {(int2 *)(rg + nE + Dc), (int2 *)(rg + nE + 3 * Dc)};
  int *nx[2] = {rg + nE + 5 * Dc, rg + nE + 6 * Dc};
  int *cbase = rg + nE + 7 * Dc, *off = rg + nE + 8 * Dc;
  // departures by ring order
  int XP = 0, XQ = 0;
  for (int ring = 0; ring < 2; ring++) {
    const int rb = ring ? nP : 0, nR = ring ? nQ : nP, r0 = ring ? XP : 0;
    const int tot = grp_scan_range<W, int>(
        nR, [&](int j) { return dmult(N, rb + j); },
        [&](int j, int pre, int mu) {
          const int u = rb + j, r = r0 + pre;
          if (mu > 0 && r + mu <= Dc) {
            rk[u] = r;
            dn[r] = 2 * u;
            if (mu == 2) dn[r + 1] = 2 * u + 1;
          }
        });
    if (ring == 0) XP = tot;
    else XQ = tot;
  }
  const int D = XP + XQ;
  if (XP != XQ || XP == 0 || D > Dc) return SEG_SLOW;
  grp_sync<W>();
  // successor of each departure: the departure from the neighbour of its
  // arrival; the arrival must carry the status the inner loop stops on.
  // Key .x: the P departure of the departure's intersection point.
  bool bad = false;
  for (int d = lane; d < D; d += grp_size<W>()) {
    const Departure p = departure(N, dn, d, nP, nQ, XP, XQ);
    const bool onP = d < XP;
    const int ua = dn[p.arr] >> 1, u = p.code >> 1;
    const int q = N.nb[ua], pn = N.nb[u];
    if (dep_enex(N, dn[p.arr]) != (p.fwd ? EXIT : ENTRY) || q < 0 || pn < 0 ||
        (q < nP) == onP || (pn < nP) == onP || dmult(N, q) == 0 || dmult(N, pn) == 0) {
      bad = true;
      continue;
    }
    const int s = dep_partner(N, rk, dn[p.arr], nP, nQ);
    const int key = onP ? d : dep_partner(N, rk, p.code, nP, nQ);
    if (s < 0 || s >= D || key < 0 || key >= XP) {
      bad = true;
      continue;
    }
    nx[0][d] = s;
    vc[0][d] = make_int2(key, d);
    off[d] = p.dist;
  }
  if (grp_any<W>(bad)) return SEG_SLOW;
  // pointer jumping: smallest P key (.x) and smallest departure (.y) of
  // every cycle; four independent chains per lane in flight.  After round t
  // every value is the minimum over the next 2^t departures of its cycle; a
  // round that changes no value means every window already spans its cycle's
  // minimum, so the loop ends there instead of after ceil(log2 D) rounds.
  int rounds = 0;
  while ((1 << rounds) < D) rounds++;
  int cur = 0;
  for (int t = 0; t < rounds; t++) {
    // for the whole grid, the grid barrier of the grp_any before orders the
    // last round's stores
    if constexpr (W != kGrid) grp_sync<W>();
    const int2 *__restrict__ vs = vc[cur];
    const int *__restrict__ ns = nx[cur];
    int2 *__restrict__ vd = vc[cur ^ 1];
    int *__restrict__ nd = nx[cur ^ 1];
    bool changed = false;
    for (int d0 = lane; d0 < D; d0 += 4 * grp_size<W>()) {
      int x[4];
      int2 a[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int d = d0 + grp_size<W>() * k;
        x[k] = d < D ? ns[d] : 0;
        a[k] = d < D ? vs[d] : make_int2(0, 0);
      }
      int2 b[4];
      int y[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        b[k] = vs[x[k]];
        y[k] = ns[x[k]];
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int d = d0 + grp_size<W>() * k;
        if (d < D) {
          const int2 m = make_int2(min(a[k].x, b[k].x), min(a[k].y, b[k].y));
          changed |= m.x != a[k].x || m.y != a[k].y;
          vd[d] = m;
          nd[d] = y[k];
        }
      }
    }
    cur ^= 1;
    if (!grp_any<W>(changed)) break;
  }
  grp_sync<W>();
  const int2 *mc = vc[cur];
  // a cycle is traced when it holds the departure of its smallest P key; its
  // list starts there.  S: segment lengths, then suffix sums
  int *S[2] = {(int *)vc[cur ^ 1], (int *)vc[cur ^ 1] + Dc};
  for (int d = lane; d < D; d += grp_size<W>()) {
    const int2 c = mc[d];
    const int sd = c.x;
    const bool sel = mc[sd].y == c.y;
    int nxt = -1;
    if (sel) {
      const Departure p = departure(N, dn, d, nP, nQ, XP, XQ);
      const int s = dep_partner(N, rk, dn[p.arr], nP, nQ);
      nxt = s == sd ? -1 : s;
    }
    S[1][d] = sel ? 1 : 0;
    S[0][d] = off[d];
    nx[0][d] = nxt;
  }
  grp_sync<W>();
  // each intersection point departs exactly once over the traced cycles
  for (int d = lane; d < XP; d += grp_size<W>())
    if (S[1][d] == S[1][dep_partner(N, rk, dn[d], nP, nQ)]) bad = true;
  if (grp_any<W>(bad)) return SEG_SLOW;
  // list ranking of the traced lists (nxt = -1 ends a list): the loop ends
  // once no element still points anywhere
  cur = 0;
  for (int t = 0; t < rounds; t++) {
    if constexpr (W != kGrid) grp_sync<W>();
    const int *__restrict__ ss = S[cur];
    const int *__restrict__ ns = nx[cur];
    int *__restrict__ sd = S[cur ^ 1];
    int *__restrict__ nd = nx[cur ^ 1];
    bool live = false;
    for (int d0 = lane; d0 < D; d0 += 4 * grp_size<W>()) {
      int x[4], a[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int d = d0 + grp_size<W>() * k;
        x[k] = d < D ? ns[d] : -1;
        a[k] = d < D ? ss[d] : 0;
      }
      int b[4], y[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        b[k] = x[k] >= 0 ? ss[x[k]] : 0;
        y[k] = x[k] >= 0 ? ns[x[k]] : -1;
      }
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int d = d0 + grp_size<W>() * k;
        if (d < D) {
          sd[d] = a[k] + b[k];
          nd[d] = y[k];
          live |= y[k] >= 0;
        }
      }
    }
    cur ^= 1;
    if (!grp_any<W>(live)) break;
  }
  grp_sync<W>();
  const int *Sf = S[cur];
  // components in order of their start departure
  int *cix = nx[0], *ciD = nx[1];
  // one scan of (components << 32) + vertices
  const long long vc_all = grp_scan_range<W, long long>(
      XP, [&](int d) { return mc[d].x == d ? (1LL << 32) + Sf[d] : 0LL; },
      [&](int d, long long pre, long long x) {
        if (x == 0) return;
        const int ci = (int)(pre >> 32), base = (int)(pre & 0xffffffffLL);
        cbase[d] = base;
        cix[d] = ci;
        if (ci < O.capComp) {
          O.compRoot[ci] = base;
          clen[ci] = (int)(x & 0xffffffffLL);
        }
      });
  const int vcarry = (int)(vc_all & 0xffffffffLL), ccarry = (int)(vc_all >> 32);
  if (vcarry > O.cap || ccarry > O.capComp) return SEG_OVERFLOW;
  // the output position map holds n entries
  if (vcarry > n) return SEG_SLOW;
  O.n = vcarry;
  O.ncomp = ccarry;
  grp_sync<W>();
  for (int d = lane; d < D; d += grp_size<W>()) {
    const int2 c = mc[d];
    const int sd = c.x;
    if (mc[sd].y == c.y) {
      off[d] = cbase[sd] + Sf[sd] - Sf[d];
      ciD[d] = cix[sd];
    } else {
      off[d] = -1;
    }
  }
  grp_sync<W>();
  // a ring vertex lies on the forward segment of the departure before it
  // and on the backward segment of the departure after it; of these two,
  // which are reverses of each other, at most one is traced, except at a
  // split vertex, where the two arrive at its two copies
  // (kRS chunks of W vertices at a time: each level of loads, the
  // vertices' multiplicities, then their two departures, then those
  // departures' statuses, components and the vertices' coordinates, is
  // issued for all the chunks before it is used; the stores come last)
  int *cw = rk;  // output position -> component
  if constexpr (W == kGrid) {
    // the whole grid: one vertex per thread at a time, in block tiles
    for (int ring = 0; ring < 2; ring++) {
      const int rb = ring ? nP : 0, nR = ring ? nQ : nP, base = ring ? XP : 0;
      const int XR = ring ? XQ : XP;
      grp_scan_range<W, int>(
          nR, [&](int j) { return dmult(N, rb + j); },
          [&](int j, int r, int mu) {
            const int u = rb + j, ri = r + mu;
            const pt ou = P_orig(N, u);
            for (int side = 0; side < 2; side++) {
              const int dd = base + (side ? (ri == XR ? 0 : ri) : (r == 0 ? XR - 1 : r - 1));
              const int of = off[dd], cd = dn[dd];
              if (of < 0 || dep_enex(N, cd) != (side ? EXIT : ENTRY)) continue;
              const int ud = cd >> 1;
              int dist = side ? ud - u : u - ud;
              if (dist <= 0) dist += nR;
              const int w = of + dist - 1;
              if (w >= 0 && w < O.n) {
                st(xy, w, ou);
                cw[w] = ciD[dd];
                omap[w] = u;
              }
            }
          });
    }
    grp_sync<W>();
    return SEG_DONE;
  }
  constexpr int kRS = W == 64 ? kRingStageWave : kRingStage;
  for (int ring = 0; ring < 2; ring++) {
    const int rb = ring ? nP : 0, nR = ring ? nQ : nP, base = ring ? XP : 0;
    const int XR = ring ? XQ : XP;
    int carry = 0;
    for (int b0 = 0; b0 < nR; b0 += kRS * grp_size<W>()) {
      int mu[kRS], r[kRS];
#pragma unroll
      for (int t = 0; t < kRS; t++) {
        const int j = b0 + t * grp_size<W>() + lane;
        mu[t] = j < nR ? dmult(N, rb + j) : 0;
      }
#pragma unroll
      for (int t = 0; t < kRS; t++) {
        int tot;
        r[t] = carry + grp_scan<W>(mu[t], tot);
        carry += tot;
      }
      // the departures before and after each vertex: [t][0] forward, [t][1]
      // backward
      int dd[kRS][2], of[kRS][2], cd[kRS][2];
#pragma unroll
      for (int t = 0; t < kRS; t++) {
        const int j = b0 + t * grp_size<W>() + lane, ri = r[t] + mu[t];
        dd[t][0] = base + (r[t] == 0 ? XR - 1 : r[t] - 1);
        dd[t][1] = base + (ri == XR ? 0 : ri);
#pragma unroll
        for (int side = 0; side < 2; side++) {
          of[t][side] = j < nR ? off[dd[t][side]] : -1;
          cd[t][side] = j < nR ? dn[dd[t][side]] : 0;
        }
      }
      bool on[kRS][2];
      int ci[kRS][2];
      pt ou[kRS];
#pragma unroll
      for (int t = 0; t < kRS; t++) {
        const int j = b0 + t * grp_size<W>() + lane;
        if (j < nR) ou[t] = P_orig(N, rb + j);
#pragma unroll
        for (int side = 0; side < 2; side++) {
          on[t][side] = of[t][side] >= 0 && dep_enex(N, cd[t][side]) == (side ? EXIT : ENTRY);
          ci[t][side] = on[t][side] ? ciD[dd[t][side]] : 0;
        }
      }
#pragma unroll
      for (int t = 0; t < kRS; t++) {
        const int u = rb + b0 + t * grp_size<W>() + lane;
#pragma unroll
        for (int side = 0; side < 2; side++) {
          if (!on[t][side]) continue;
          const int ud = cd[t][side] >> 1;
          int dist = side ? ud - u : u - ud;
          if (dist <= 0) dist += nR;
          const int w = of[t][side] + dist - 1;
          if (w >= 0 && w < O.n) {
            st(xy, w, ou[t]);
            cw[w] = ci[t][side];
            omap[w] = u;
          }
        }
      }
    }
  }
  grp_sync<W>();
  return SEG_DONE;
}

// a whole node ring as output component c at [off, off + n); omap: output
// position -> node
template <int W>
__device__ void copy_ring_out(double *xy, int *cw, int *omap, const Nodes &N, int r0, int n,
                              int off, int c) {
  for (int j = grp_rank<W>(); j < n; j += grp_size<W>()) {
    st(xy, off + j, P_orig(N, r0 + j));
    cw[off + j] = c;
    omap[off + j] = r0 + j;
  }
}

// ------------------------------------------------- parallel clean-up ----
// cleanUpResult (polyclip.h:791-807) on one component stored in list order
// at [b, b + L).  Sequentially it (A) drops the root while
// |A(prev, root, next)| < eps, the prev of every later root being the last
// vertex; then (B) makes one pass from the root in which a vertex is dropped
// when |A(last kept, it, next)| < eps.  (A) tests against a fixed vertex, so
// it is one parallel search for the first vertex that stays.  In (B) a
// vertex's outcome depends on the distance d back to the last kept vertex;
// each lane evaluates its vertex for d = 1..kCleanW, which makes it a map on
// the states {1..kCleanW, overflow}, and a wave prefix of map compositions
// gives every lane its state.  A run of more than kCleanW dropped vertices
// ends in the overflow state: that chunk of 64 vertices then runs on one
// lane with the last kept vertex itself, and the map prefix resumes with the
// first chunk that starts within kCleanW of a kept vertex.
constexpr int kCleanW = 7;

__device__ __forceinline__ unsigned nib(unsigned f, unsigned s) { return (f >> (4 * s)) & 15u; }
// (g o f)(s) = g(f(s)) over the 8 states held as nibbles
__device__ __forceinline__ unsigned compose(unsigned g, unsigned f) {
  unsigned h = 0;
  for (unsigned s = 0; s < 8; s++) h |= nib(g, nib(f, s)) << (4 * s);
  return h;
}

// the normalised point at output position w: node omap[w] of pn
__device__ __forceinline__ pt out_pt(const double *pn, const int *omap, int w) {
  return ld(pn, omap[w]);
}

// writes keep[b..b+L) (0/1) and returns the number of vertices kept
__device__ int clean_component(const double *pn, const int *omap, int *keep, int b, int L) {
  const int lane = __lane_id();
  int r = -1;
  if (L >= 3) {
    const pt Pl = out_pt(pn, omap, b + L - 1);
    for (int base = 0; base <= L - 3 && r < 0; base += 64) {
      const int i = base + lane;
      bool stays = false;
      if (i <= L - 3) stays = !(fabs(triA(Pl, out_pt(pn, omap, b + i), out_pt(pn, omap, b + i + 1))) < kGhEps);
      const unsigned long long m = __ballot(stays);
      if (m) r = base + __ffsll((long long)m) - 1;
    }
  }
  for (int j = lane; j < (r < 0 ? L : r); j += 64) keep[b + j] = 0;
  if (r < 0) return 0;
  if (lane == 0) keep[b + r] = 1;
  const int x0 = b + r, m = L - 1 - r;
  // chunks of 64 vertices in order, the state carried between them as the
  // position of the last kept vertex (x0 + jl); a chunk whose first vertex
  // lies within kCleanW of it runs the map prefix, and a chunk that reaches
  // the overflow state, or starts too far from it, runs on lane 0 with the
  // kept vertex itself (a long run of dropped vertices costs one chunk at a
  // time, not the component)
  int kept = 1, jl = 0;
  for (int base = 1; base <= m; base += 64) {
    const int j = base + lane, s = base - jl;
    bool seq = s > kCleanW;
    if (!seq) {
      unsigned K = 0;
      if (j <= m) {
        const pt cur = out_pt(pn, omap, x0 + j), nxt = out_pt(pn, omap, j == m ? x0 : x0 + j + 1);
        for (int d = 1; d <= kCleanW && d <= j; d++)
          if (!(fabs(triA(out_pt(pn, omap, x0 + j - d), cur, nxt)) < kGhEps)) K |= 1u << (d - 1);
      }
      unsigned f = 0x76543210u;  // identity past the end
      if (j <= m) {
        f = 0;  // state 0 (overflow) stays 0
        for (unsigned t = 1; t <= kCleanW; t++) {
          const unsigned to = ((K >> (t - 1)) & 1u) ? 1u : (t < kCleanW ? t + 1 : 0u);
          f |= to << (4 * t);
        }
      }
      unsigned F = f;
      for (int o = 1; o < 64; o <<= 1) {
        const unsigned G = __shfl_up(F, o, 64);
        if (lane >= o) F = compose(F, G);
      }
      const unsigned Fex = __shfl_up(F, 1, 64);
      const unsigned sj = lane == 0 ? (unsigned)s : nib(Fex, (unsigned)s);
      bool kp = false, ovf = false;
      if (j <= m) {
        if (sj == 0) ovf = true;
        else kp = ((K >> (sj - 1)) & 1u) != 0;
      }
      if (__ballot(ovf)) {
        seq = true;
      } else {
        if (j <= m) keep[x0 + j] = kp ? 1 : 0;
        const unsigned long long km = __ballot(kp);
        kept += __popcll(km);
        if (km) jl = base + 63 - __clzll((long long)km);
      }
    }
    if (seq) {
      if (lane == 0) {
        pt last = out_pt(pn, omap, x0 + jl);
        const int e = base + 63 < m ? base + 63 : m;
        for (int jj = base; jj <= e; jj++) {
          const pt cur = out_pt(pn, omap, x0 + jj),
                   nxt = out_pt(pn, omap, jj == m ? x0 : x0 + jj + 1);
          const bool kp = !(fabs(triA(last, cur, nxt)) < kGhEps);
          keep[x0 + jj] = kp ? 1 : 0;
          if (kp) {
            last = cur;
            jl = jj;
            kept++;
          }
        }
      }
      kept = __shfl(kept, 0, 64);
      jl = __shfl(jl, 0, 64);
    }
  }
  return kept;
}

// the same clean-up of one component by one thread, in order: the root
// search (A), then the pass (B)
__device__ int lane_clean(const double *pn, const int *omap, int *keep, int b, int L) {
  int r = -1;
  if (L >= 3) {
    const pt Pl = out_pt(pn, omap, b + L - 1);
    for (int i = 0; i <= L - 3; i++)
      if (!(fabs(triA(Pl, out_pt(pn, omap, b + i), out_pt(pn, omap, b + i + 1))) < kGhEps)) {
        r = i;
        break;
      }
  }
  for (int j = 0; j < (r < 0 ? L : r); j++) keep[b + j] = 0;
  if (r < 0) return 0;
  keep[b + r] = 1;
  const int x0 = b + r, m = L - 1 - r;
  pt last = out_pt(pn, omap, x0);
  int kept = 1;
  for (int j = 1; j <= m; j++) {
    const pt cur = out_pt(pn, omap, x0 + j), nxt = out_pt(pn, omap, j == m ? x0 : x0 + j + 1);
    const bool kp = !(fabs(triA(last, cur, nxt)) < kGhEps);
    keep[x0 + j] = kp ? 1 : 0;
    if (kp) {
      last = cur;
      kept++;
    }
  }
  return kept;
}

// components up to this length are cleaned by one thread each, longer ones
// by one wave each, and, in a pair labelled by the whole grid, those longer
// than TraceArgs::gridClean by the whole grid
constexpr int kLaneClean = 48;

constexpr unsigned kMapId = 0x76543210u;  // the identity on the 8 states

// the map on the states {overflow, 1..kCleanW} of a vertex whose
// |A(vertex d back, it, next)| >= eps for the d with bit d - 1 of K set
__device__ __forceinline__ unsigned clean_map(unsigned K) {
  unsigned f = 0;  // state 0 (overflow) stays 0
  for (unsigned t = 1; t <= kCleanW; t++) {
    const unsigned to = ((K >> (t - 1)) & 1u) ? 1u : (t < kCleanW ? t + 1 : 0u);
    f |= to << (4 * t);
  }
  return f;
}

// Exclusive composition, over every thread of a cooperative launch in rank
// order, of one map per thread: returns the composition of the maps of the
// threads before this one (the earliest applied first).  Within a wave by
// shuffles, over the block's waves in LDS, over the blocks through the
// partials buffer: each block composes the aggregates of the blocks before
// it, a contiguous run per thread, then an ordered tree over the block.
__device__ unsigned grid_compose_scan(unsigned f) {
  const int lane = __lane_id(), w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned F = f;
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned G = __shfl_up(F, o, 64);
    if (lane >= o) F = compose(F, G);
  }
  unsigned Fex = __shfl_up(F, 1, 64);
  if (lane == 0) Fex = kMapId;
  __shared__ unsigned wagg[16], wred[16];
  __syncthreads();
  if (lane == 63) wagg[w] = F;
  __syncthreads();
  unsigned before = kMapId, blk = kMapId;
  for (int i = 0; i < nw; i++) {
    if (i < w) before = compose(wagg[i], before);
    blk = compose(wagg[i], blk);
  }
  long long *part = grid_part + (long long)grid_half * gridDim.x;
  if (threadIdx.x == 0) part[blockIdx.x] = (long long)blk;
  cooperative_groups::this_grid().sync();
  const int nb = (int)blockIdx.x, per = (nb + (int)blockDim.x - 1) / (int)blockDim.x;
  const int i0 = (int)threadIdx.x * per, i1 = min(nb, i0 + per);
  unsigned g = kMapId;
  for (int i = i0; i < i1; i++) g = compose((unsigned)part[i], g);
  // lane l of a wave then holds the runs of lanes [l, l + 2o) after step o
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned h = __shfl_down(g, o, 64);
    if (lane + o < 64) g = compose(h, g);
  }
  if (lane == 0) wred[w] = g;
  __syncthreads();
  unsigned gb = kMapId;
  for (int i = 0; i < nw; i++) gb = compose(wred[i], gb);
  if (threadIdx.x == 0) grid_half ^= 1;
  __syncthreads();
  return compose(Fex, compose(before, gb));
}

// clean_component by every thread of a cooperative launch, for the long
// components of a pair labelled grid-wide (L >= 3): the root search a grid-size slice at a time; in
// the pass each thread takes a contiguous run of vertices, stores each one's
// bits K in keep, composes their maps, and after the grid-wide exclusive
// composition walks its run from its own starting state.  A run that
// reaches the overflow state sends the component to clean_component on the
// first wave.  Writes keep[b..b+L) and returns the number of vertices kept
// (the same on every thread); the caller syncs the grid before keep is read.
__device__ int grid_clean_component(const double *pn, const int *omap, int *keep, int b, int L) {
  const int g = grp_rank<kGrid>(), G = grp_size<kGrid>();
  constexpr int kNone = 0x7fffffff;
  int r = -1;
  const pt Pl = out_pt(pn, omap, b + L - 1);
  for (int base = 0; base <= L - 3; base += G) {
    const int i = base + g;
    int cand = kNone;
    if (i <= L - 3 &&
        !(fabs(triA(Pl, out_pt(pn, omap, b + i), out_pt(pn, omap, b + i + 1))) < kGhEps))
      cand = i;
    const int mn = grp_min<kGrid>(cand);
    if (mn != kNone) {
      r = mn;
      break;
    }
  }
  for (int j = g; j < (r < 0 ? L : r); j += G) keep[b + j] = 0;
  if (r < 0) return 0;
  if (g == 0) keep[b + r] = 1;
  const int x0 = b + r, m = L - 1 - r;
  const int per = (m + G - 1) / G;
  const int j0 = 1 + (int)min((long long)g * per, (long long)m), j1 = min(j0 + per, m + 1);
  unsigned agg = kMapId;
  for (int j = j0; j < j1; j++) {
    const pt cur = out_pt(pn, omap, x0 + j), nxt = out_pt(pn, omap, j == m ? x0 : x0 + j + 1);
    unsigned K = 0;
    for (int d = 1; d <= kCleanW && d <= j; d++)
      if (!(fabs(triA(out_pt(pn, omap, x0 + j - d), cur, nxt)) < kGhEps)) K |= 1u << (d - 1);
    keep[x0 + j] = (int)K;
    agg = compose(clean_map(K), agg);
  }
  unsigned s = nib(grid_compose_scan(agg), 1);
  long long kept = 0;
  bool ovf = false;
  for (int j = j0; j < j1; j++) {
    if (s == 0) {
      ovf = true;
      break;
    }
    const bool kp = (((unsigned)keep[x0 + j] >> (s - 1)) & 1u) != 0;
    keep[x0 + j] = kp ? 1 : 0;
    kept += kp ? 1 : 0;
    s = kp ? 1u : (s < kCleanW ? s + 1 : 0u);
  }
  // the kept count and, above bit 40, the threads that overflowed
  long long total;
  grid_combine(block_sum_ll(kept + (ovf ? 1LL << 40 : 0LL)), 0, total);
  if ((total >> 40) == 0) return 1 + (int)total;
  int k2 = 0;
  if (g < 64) {
    k2 = clean_component(pn, omap, keep, b, L);
    if (g != 0) k2 = 0;
  }
  return grp_sum<kGrid>(k2);
}

// clean-up of every component of O (component c at [O.compRoot[c], +clen[c])
// in list order, cw: position -> component), then compaction of xy and clen
// over the kept vertices and non-empty components; the normalised point at
// output position w is node omap[w] of pn.  Scratch: 3 O.ncomp + O.n ints
// at sc (4 O.ncomp + O.n for the whole grid).  Returns true when some vertex was dropped; O.n / O.ncomp are
// updated.
template <int W>
__device__ bool grp_cleanup_compact(OutBuf &O, double *xy, int *clen, const int *cw, int *sc,
                                    const double *pn, const int *omap, int gridClean) {
  const int lane = grp_rank<W>();
  const int T = O.n, NC = O.ncomp;
  int *dirty = sc, *dl = sc + NC, *nlen = sc + 2 * NC, *keep = sc + 3 * NC;
  for (int c = lane; c < NC; c += grp_size<W>()) dirty[c] = 0;
  grp_sync<W>();
  for (int w = lane; w < T; w += grp_size<W>()) {
    const int c = min(max(cw[w], 0), NC - 1), b = O.compRoot[c], L = clen[c];
    const int pw = w == b ? b + L - 1 : w - 1, nw = w == b + L - 1 ? b : w + 1;
    if (fabs(triA(out_pt(pn, omap, pw), out_pt(pn, omap, w), out_pt(pn, omap, nw))) < kGhEps)
      dirty[c] = 1;
  }
  grp_sync<W>();
  const int nd = grp_scan_range<W, int>(
      NC, [&](int c) { return dirty[c] != 0 ? 1 : 0; },
      [&](int c, int pre, int f) {
        if (f) dl[pre] = c;
      });
  if (nd == 0) return false;
  grp_sync<W>();
  for (int i = lane; i < nd; i += grp_size<W>()) {
    const int c = dl[i];
    if (clen[c] <= kLaneClean) nlen[c] = lane_clean(pn, omap, keep, O.compRoot[c], clen[c]);
  }
  for (int i = lane >> 6; i < nd; i += grp_size<W>() / 64) {
    const int c = dl[i];
    if (clen[c] > kLaneClean && (W != kGrid || clen[c] <= gridClean)) {
      const int kept = clean_component(pn, omap, keep, O.compRoot[c], clen[c]);
      if (__lane_id() == 0) nlen[c] = kept;
    }
  }
  if constexpr (W == kGrid) {
    // the components longer than gridClean (> kLaneClean), listed in order at
    // gl (at most NC ints past keep), one after another by the whole grid
    int *gl = keep + T;
    const int ng = grp_scan_range<W, int>(
        nd, [&](int i) { return clen[dl[i]] > gridClean ? 1 : 0; },
        [&](int i, int pre, int f) {
          if (f) gl[pre] = dl[i];
        });
    grp_sync<W>();
    for (int i = 0; i < ng; i++) {
      const int c = gl[i];
      const int kept = grid_clean_component(pn, omap, keep, O.compRoot[c], clen[c]);
      if (lane == 0) nlen[c] = kept;
    }
  }
  grp_sync<W>();
  if constexpr (W == kGrid) {
    // compaction out of place, since a block's stores may reach tiles other
    // blocks have yet to read: the kept vertices to O.po (unused by the
    // segment path) and the lengths to dl, then back
    double *xo = O.po;
    const int vout = grp_scan_range<W, int>(
        T,
        [&](int w) {
          const int c = min(max(cw[w], 0), NC - 1);
          return dirty[c] ? (keep[w] != 0 ? 1 : 0) : 1;
        },
        [&](int w, int pre, int kp) {
          if (kp) st(xo, pre, ld(xy, w));
        });
    const int cout = grp_scan_range<W, int>(
        NC, [&](int c) { return (dirty[c] ? nlen[c] : clen[c]) > 0 ? 1 : 0; },
        [&](int c, int pre, int f) {
          if (f) dl[pre] = dirty[c] ? nlen[c] : clen[c];
        });
    grp_sync<W>();
    for (int i = lane; i < vout; i += grp_size<W>()) st(xy, i, ld(xo, i));
    for (int i = lane; i < cout; i += grp_size<W>()) clen[i] = dl[i];
    O.n = vout;
    O.ncomp = cout;
    return true;
  }
  int vout = 0;
  for (int base = 0; base < T; base += grp_size<W>()) {
    const int w = base + lane;
    bool kp = false;
    pt v;
    if (w < T) {
      const int c = min(max(cw[w], 0), NC - 1);
      kp = dirty[c] ? keep[w] != 0 : true;
      if (kp) v = ld(xy, w);
    }
    int tot;
    const int pre = grp_scan<W>(kp ? 1 : 0, tot);
    if (kp) st(xy, vout + pre, v);
    vout += tot;
  }
  int cout = 0;
  for (int base = 0; base < NC; base += grp_size<W>()) {
    const int c = base + lane;
    int len = 0;
    if (c < NC) len = dirty[c] ? nlen[c] : clen[c];
    int tot;
    const int pre = grp_scan<W>(len > 0 ? 1 : 0, tot);
    if (len > 0) clen[cout + pre] = len;
    cout += tot;
  }
  O.n = vout;
  O.ncomp = cout;
  return true;
}

// which way a pair went through the wave labelling kernel
enum {
  LP_SKIPPED = 0,      // status was not OK on entry
  LP_SEGMENTS = 1,     // segment-parallel trace, nothing to clean up
  LP_SEG_CLEANUP = 2,  // segment-parallel trace, then the per-component clean-up
  LP_NO_CROSSING = 3,  // no intersection vertices: containment tests
  LP_SEQUENTIAL = 4,   // the sequential steps on lane 0
  LP_FAILED = 5        // overflow or failed trace
};

// gives the sequential steps, which insert nodes, the pair's nodes in its
// global region G: copies the arrays of N that lie in the LDS arena there,
// writes the ring links and the original coordinates (which the parallel
// steps derive from the slot order) and makes N the global region; every
// lane must call it
template <int W>
__device__ void spill_nodes(Nodes &N, const Nodes &G, int nP, int nQ) {
  const int n = nP + nQ;
  grp_sync<W>();
  for (int j = grp_rank<W>(); j < n; j += grp_size<W>()) {
    if (N.pn != G.pn) st(G.pn, j, ld(N.pn, j));
    st(G.po, j, P_orig(N, j));
    G.next[j] = j == nP - 1 ? 0 : (j == n - 1 ? nP : j + 1);
    G.prev[j] = j == 0 ? nP - 1 : (j == nP ? n - 1 : j - 1);
    if (N.nb != G.nb) G.nb[j] = N.nb[j];
    if (N.fl != G.fl) G.fl[j] = N.fl[j];
    if (N.lab != G.lab) G.lab[j] = N.lab[j];
    if (N.enex != G.enex) G.enex[j] = N.enex[j];
  }
  grp_sync<W>();
  const int nn = N.n, cap = N.cap;
  N = G;
  N.n = nn;
  N.cap = cap;
}

// arena / arenaBytes: the group's LDS for the pair's arrays (null: none)
template <int W>
__device__ int label_pair(long k, int *status, const SlotArrays &s, const TraceArgs &a,
                          int *aux, char *arena, int arenaBytes) {
  const int lane = grp_rank<W>();
  if (status[k] != GH_OK) {
    if (lane == 0) {
      a.outNComp[k] = 0;
      a.outNVerts[k] = 0;
    }
    return LP_SKIPPED;
  }
  const long long sbP = a.psP2[a.offP[k]], cP = a.psP2[a.offP[k + 1]] - sbP;
  const long long sbQ = a.psQ2[a.offQ[k]], cQ = a.psQ2[a.offQ[k + 1]] - sbQ;
  const long long nb0 = a.nodeOff[k], cap = a.nodeOff[k + 1] - nb0;
  // the pair's global region (node capacity cap >= 2 (nP + nQ) + 4)
  Nodes G;
  G.pn = a.npn + 2 * nb0;
  G.po = a.npo + 2 * nb0;
  G.next = a.nnext + nb0;
  G.prev = a.nprev + nb0;
  G.nb = a.nnb + nb0;
  G.fl = a.nfl + nb0;
  G.lab = a.nlab + nb0;
  G.enex = a.nenex + nb0;
  G.n = (int)(cP + cQ);
  G.cap = (int)cap;
  const long long pE0 = a.offP[k], pE1 = a.offP[k + 1];
  const long long qE0 = a.offQ[k], qE1 = a.offQ[k + 1];
  const int nP = (int)cP, nQ = (int)cQ, n = nP + nQ;
  // the parallel steps keep the n nodes in LDS as far as they fit, the
  // arrays read most often first; the sequential steps, which insert
  // nodes, run on the global region (spill_nodes)
  Arena ar{arena, arenaBytes};
  // Ring links follow the slot order (node j + 1 after node j on each
  // ring) and the original coordinates stay in the slots, so neither is
  // materialised here.
  Nodes N = G;
  N.pn = ar.take<double>(2 * n, G.pn);
  N.nb = ar.take<int>(n, G.nb);
  N.fl = ar.take<unsigned char>(n, G.fl);
  N.lab = ar.take<signed char>(n, G.lab);
  N.enex = ar.take<signed char>(n, G.enex);
  N.next = N.prev = nullptr;
  N.po = nullptr;
  N.oP = s.oP + 2 * sbP;
  N.oQ = s.oQ + 2 * sbQ;
  N.nP0 = nP;
  int *rg = aux + 5 * nb0;  // segment scratch: 10 n + 20 ints in the global region
  // -------- build (ghcuda.cpp:239-322 semantics, see k_label_trace): every
  // slot is first taken as an intersection vertex, with its link as loaded
  // kept in F_LINK; then the first slot of each edge is re-flagged as its
  // parent vertex, linked when F_LINK is set.  Each pass takes kS nodes per
  // lane at a time and issues all their loads before the first store, so
  // that it waits on memory once per kS * W nodes.
  const bool xl = a.explicitLinks != 0;
  constexpr int kS = W == 64 ? kStageWave : kStage;
  for (int j0 = lane; j0 < nP; j0 += kS * grp_size<W>()) {
    pt ip[kS];
    int nbv[kS];
    bool lk[kS];
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const int j = j0 + u * grp_size<W>();
      if (j < nP) {
        const long long sl = sbP + j;
        ip[u] = ld(s.iP, sl);
        nbv[u] = xl ? s.nbP[sl] : 0;
        lk[u] = xl ? nbv[u] > 0 : s.aP[sl] != kNone;
      }
    }
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const int j = j0 + u * grp_size<W>();
      if (j < nP) {
        st(N.pn, j, ip[u]);
        if (xl) {
          const long long kq = (long long)nbv[u] - 1 - sbQ;
          N.nb[j] = (kq >= 0 && kq < cQ) ? nP + (int)kq : -1;
        } else {
          N.nb[j] = -1;
        }
        N.lab[j] = L_NONE;
        N.enex[j] = NEITHER;
        N.fl[j] = (unsigned char)(F_INT | (lk[u] ? F_LINK : 0));
      }
    }
  }
  for (int j0 = lane; j0 < nQ; j0 += kS * grp_size<W>()) {
    pt ip[kS];
    int nbv[kS];
    bool lk[kS];
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const int j = j0 + u * grp_size<W>();
      if (j < nQ) {
        const long long sl = sbQ + j;
        ip[u] = ld(s.iQ, sl);
        nbv[u] = s.nbQ[sl];
        lk[u] = s.aQ[sl] != kNone || nbv[u] > 0;
      }
    }
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const int j = j0 + u * grp_size<W>();
      if (j < nQ) {
        const int v = nP + j;
        st(N.pn, v, ip[u]);
        N.lab[v] = L_NONE;
        N.enex[v] = NEITHER;
        // an unlinked parent has nbQ <= 0, which lands outside [0, cP) as well
        const long long kp = (long long)nbv[u] - 1 - sbP;
        N.nb[v] = (kp >= 0 && kp < cP) ? (int)kp : -1;
        N.fl[v] = (unsigned char)(F_INT | (lk[u] ? F_LINK : 0));
      }
    }
  }
  grp_sync<W>();
  // parents: the edges' first slots (one edge each)
  for (long long e0 = pE0 + lane; e0 < pE1; e0 += kS * grp_size<W>()) {
    long long sl[kS];
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const long long e = e0 + u * grp_size<W>();
      sl[u] = e < pE1 ? a.psP2[e] - sbP : -1;
    }
#pragma unroll
    for (int u = 0; u < kS; u++)
      if (sl[u] >= 0)
        N.fl[sl[u]] = (unsigned char)(F_SRC | ((N.fl[sl[u]] & F_LINK) ? F_INT : 0));
  }
  for (long long e0 = qE0 + lane; e0 < qE1; e0 += kS * grp_size<W>()) {
    long long sl[kS];
#pragma unroll
    for (int u = 0; u < kS; u++) {
      const long long e = e0 + u * grp_size<W>();
      sl[u] = e < qE1 ? nP + (a.psQ2[e] - sbQ) : -1;
    }
#pragma unroll
    for (int u = 0; u < kS; u++)
      if (sl[u] >= 0)
        N.fl[sl[u]] = (unsigned char)(F_SRC | ((N.fl[sl[u]] & F_LINK) ? F_INT : 0));
  }
  grp_sync<W>();
  if (!xl) {
    // a P node linked from several Q slots keeps the last one (link() order)
    for (int j = lane; j < nQ; j += grp_size<W>())
      if (N.nb[nP + j] >= 0) atomicMax(&N.nb[N.nb[nP + j]], nP + j);
    grp_sync<W>();
    for (int j = lane; j < nP; j += grp_size<W>())
      if (N.nb[j] >= 0) N.fl[j] |= F_INT;
    grp_sync<W>();
  }
  // -------- step 1 (restored) + step 3, one lane per P intersection vertex
  // The labels are computed for kStage vertices per lane from loads only
  // and stored afterwards.  The P ring's mutual-link check rides along: a
  // P intersection vertex I must have nb[nb[I]] == I.
  bool slow = false;
  for (int I0 = lane; I0 < nP; I0 += kStage * grp_size<W>()) {
    int q[kStage], l[kStage];
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      const int I = I0 + u * grp_size<W>();
      q[u] = I < nP && isint(N, I) ? N.nb[I] : -2;
      l[u] = L_NONE;
    }
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      const int I = I0 + u * grp_size<W>();
      if (q[u] == -2) continue;
      if (q[u] < 0) {
        slow = true;
        continue;
      }
      if (N.nb[q[u]] != I) slow = true;
      const int Pm = I == 0 ? nP - 1 : I - 1, Pp = I == nP - 1 ? 0 : I + 1;
      const int Qm = q[u] == nP ? n - 1 : q[u] - 1, Qp = q[u] == n - 1 ? nP : q[u] + 1;
      const int tm = node_rel(N, Qm, Pm, I, Pp), tp = node_rel(N, Qp, Pm, I, Pp);
      int lb = L_NONE;
      if ((tm == 0 && tp == 1) || (tm == 1 && tp == 0)) lb = L_CROSSING;
      if ((tm == 0 && tp == 0) || (tm == 1 && tp == 1)) lb = L_BOUNCING;
      if ((tp == 3 && tm == 1) || (tm == 3 && tp == 1)) lb = L_LEFT_ON;
      if ((tp == 3 && tm == 0) || (tm == 3 && tp == 0)) lb = L_RIGHT_ON;
      if ((tp == 3 && tm == 2) || (tm == 3 && tp == 2)) lb = L_ON_ON;
      if ((tm == 2 && tp == 1) || (tp == 2 && tm == 1)) lb = L_ON_LEFT;
      if ((tm == 2 && tp == 0) || (tp == 2 && tm == 0)) lb = L_ON_RIGHT;
      l[u] = lb;
      if (lb != L_CROSSING && lb != L_BOUNCING) slow = true;
    }
#pragma unroll
    for (int u = 0; u < kStage; u++) {
      if (q[u] < 0) continue;
      N.lab[I0 + u * grp_size<W>()] = (signed char)l[u];
      N.lab[q[u]] = (signed char)l[u];
    }
  }
  grp_sync<W>();
  // links must be mutual on the Q ring too (several P sources may point at
  // one Q vertex when links are explicit)
#pragma unroll 4
  for (int j = lane; j < nQ; j += grp_size<W>()) {
    const int v = nP + j;
    if (isint(N, v)) {
      const int q = N.nb[v];
      if (q < 0 || N.nb[q] != v) slow = true;
    }
  }
  slow = grp_any<W>(slow);
  // -------- output regions
  OutBuf O;
  O.pn = a.opn + 2 * nb0;
  O.po = a.opo + 2 * nb0;
  O.next = a.onext + nb0;
  O.prev = a.oprev + nb0;
  O.n = 0;
  O.cap = (int)cap;
  O.compRoot = a.ocompRoot + a.compOff[k];
  O.ncomp = 0;
  O.capComp = (int)(a.compOff[k + 1] - a.compOff[k]);
  // crossings on P (the links are mutual here, so Q has as many), and the
  // intersection vertices of both rings, each of which departs at most twice
  int nCrossP = 0, nInt = 0;
  for (int j = lane; j < n; j += grp_size<W>()) {
    nInt += isint(N, j) ? 1 : 0;
    if (j < nP) nCrossP += iscross(N, j) ? 1 : 0;
  }
  nCrossP = grp_sum<W>(nCrossP);
  nInt = grp_sum<W>(nInt);
  const int Dc = even_up(min(2 * nInt, n));
  rg = ar.take<int>(seg_scratch_ints(n, Dc), rg);
  // output position -> node (the clean-up reads the node's normalised point)
  int *omap = ar.take<int>(n, (int *)O.pn);
  int VP = -1, VQ = -1;
  if (!slow) {
    VP = grp_first_nonint<W>(N, 0, nP);
    VQ = grp_first_nonint<W>(N, nP, nQ);
    if (VP < 0 || VQ < 0) slow = true;
  }
  double *xy = a.xyTmp + 2 * nb0;
  int *clen = a.compLen + a.compOff[k];
  int path = LP_SEGMENTS;
  if (!slow) {
    if (nCrossP == 0) {
      // step 3.5: no crossings on either ring -> containment by PiP of
      // each ring's first non-intersection vertex
      path = LP_NO_CROSSING;
      const bool pin = grp_pip<W>(N, nP, nQ, P_(N, VP)), qin = grp_pip<W>(N, 0, nP, P_(N, VQ));
      if ((pin ? nP : 0) + (qin ? nQ : 0) > O.cap || (pin ? 1 : 0) + (qin ? 1 : 0) > O.capComp) {
        slow = true;
      } else {
        if (pin) {
          copy_ring_out<W>(xy, rg, omap, N, 0, nP, 0, 0);
          if (lane == 0) {
            O.compRoot[0] = 0;
            clen[0] = nP;
          }
          O.ncomp = 1;
          O.n = nP;
        }
        if (qin) {
          copy_ring_out<W>(xy, rg, omap, N, nP, nQ, O.n, O.ncomp);
          if (lane == 0) {
            O.compRoot[O.ncomp] = O.n;
            clen[O.ncomp] = nQ;
          }
          O.ncomp++;
          O.n += nQ;
        }
      }
    } else {
      // step 4
      const bool pinP = grp_pip<W>(N, nP, nQ, P_(N, VP));
      grp_enex<W>(N, 0, nP, VP, pinP ? EXIT : ENTRY);
      const bool pinQ = grp_pip<W>(N, 0, nP, P_(N, VQ));
      grp_enex<W>(N, nP, nQ, VQ, pinQ ? EXIT : ENTRY);
      grp_sync<W>();
      const int rc = segment_trace<W>(N, O, xy, clen, rg, nP, nQ, Dc, omap);
      if (rc == SEG_OVERFLOW) {
        if (lane == 0) {
          status[k] = GH_OVERFLOW;
          a.outNComp[k] = 0;
          a.outNVerts[k] = 0;
        }
        return LP_FAILED;
      }
      if (rc == SEG_SLOW) slow = true;
    }
  }
  if (slow) {
    // sequential path for degenerate pairs, from step 2 on.  Step 1's P
    // labels above are the sequential step 1's (the same function of the
    // same nodes, the P ring still in index order).  Its step 3 copies each
    // P intersection vertex's label to its neighbour in ring order, so a Q
    // vertex linked from several P vertices ends with the last one's label:
    // the largest such P index, found by an atomic maximum.
    spill_nodes<W>(N, G, nP, nQ);
    O.pn = a.opn + 2 * nb0;
    int *owner = rg;  // nQ entries of the segment scratch
    for (int j = lane; j < nQ; j += grp_size<W>()) owner[j] = -1;
    grp_sync<W>();
    for (int I = lane; I < nP; I += grp_size<W>()) {
      const int q = isint(N, I) ? N.nb[I] : -1;
      if (q >= nP && q < nP + nQ) atomicMax(&owner[q - nP], I);
    }
    grp_sync<W>();
    for (int j = lane; j < nQ; j += grp_size<W>())
      if (owner[j] >= 0) N.lab[nP + j] = N.lab[owner[j]];
    grp_sync<W>();
    // step 4's point-in-polygon tests of each ring's first non-intersection
    // vertex (the rings are still in index order; the flags do not change
    // before step 4), in parallel
    int pinHint[4] = {-1, 0, -1, 0};
    {
      const int vp = grp_first_nonint<W>(N, 0, nP), vq = grp_first_nonint<W>(N, nP, nQ);
      if (vp >= 0 && vq >= 0) {
        pinHint[0] = vp;
        pinHint[1] = grp_pip<W>(N, nP, nQ, P_(N, vp)) ? 1 : 0;
        pinHint[2] = vq;
        pinHint[3] = grp_pip<W>(N, 0, nP, P_(N, vq)) ? 1 : 0;
      }
    }
    if (lane != 0) return LP_SEQUENTIAL;
    O.n = 0;
    O.ncomp = 0;
    int rc = label_and_trace(N, 0, nP, O, true, pinHint);
    if (rc != GH_OK) {
      status[k] = rc;
      a.outNComp[k] = 0;
      a.outNVerts[k] = 0;
      return LP_FAILED;
    }
    int nc = 0, nv = 0;
    for (int c = 0; c < O.ncomp; c++) {
      int r = O.compRoot[c];
      if (a.cleanup) r = out_cleanup(O, r);
      if (r < 0) continue;
      int len = 0, v = r;
      do {
        st(xy, nv + len, ld(O.po, v));
        len++;
        v = O.next[v];
      } while (v != r);
      clen[nc++] = len;
      nv += len;
    }
    a.outNComp[k] = nc;
    a.outNVerts[k] = nv;
    return LP_SEQUENTIAL;
  }
  grp_sync<W>();
  if (a.cleanup && grp_cleanup_compact<W>(O, xy, clen, rg, rg + even_up(n), N.pn, omap, a.gridClean) &&
      path == LP_SEGMENTS)
    path = LP_SEG_CLEANUP;
  if (lane == 0) {
    a.outNComp[k] = O.ncomp;
    a.outNVerts[k] = O.n;
  }
  return path;
}

// Pairs with at least TraceArgs::bigNodes nodes (Tuning::bigNodes) are
// labelled by a block of kBigGroup
// threads (k_label_block), the others by one wave each (k_label_wave).  At
// 1024 threads the sequential fallback spills a few registers; the extra
// parallelism of the block is worth more (the 1e6 x 1e5 pair labels in 113
// ms at 1024 threads, 167 ms at 512).
constexpr int kBigGroup = 1024;
// Pairs with at least Tuning::gridNodes nodes are labelled by the whole GPU,
// one cooperative launch of kGridBlock-thread blocks each (k_label_grid).
constexpr int kGridBlock = 256;
// batches of at least this many pairs are labelled largest first
constexpr long kLptMinPairs = 1024;

__device__ __forceinline__ long long pair_nodes(const long long *offP, const long long *offQ,
                                                const long long *psP2, const long long *psQ2,
                                                long k) {
  return psP2[offP[k + 1]] - psP2[offP[k]] + psQ2[offQ[k + 1]] - psQ2[offQ[k]];
}
__device__ __forceinline__ long long pair_nodes(const TraceArgs &a, long k) {
  return pair_nodes(a.offP, a.offQ, a.psP2, a.psQ2, k);
}

// Largest pairs first (longest-processing-time order, by power-of-two size
// class), values the pair indices, for one stable 8-bit radix pass: key
// 128 * (pair at or above gridNodes) + 64 * (pair below bigNodes) + 63 -
// floor(log2 nodes).  The block path's pairs (bigNodes <= nodes <
// gridNodes) are then exactly the first nBig entries, for any thresholds,
// and the grid path's come last (the waves skip them).  Blocks are
// dispatched in index order, so the pairs that take longest start first
// instead of forming the launch's tail.
__global__ void k_label_order_keys(long M, const long long *offP, const long long *offQ,
                                   const long long *psP2, const long long *psQ2,
                                   long long bigNodes, long long gridNodes, uint64_t *keys,
                                   int *vals) {
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  const long long n = pair_nodes(offP, offQ, psP2, psQ2, k);
  const int cls = n > 0 ? 63 - __clzll((unsigned long long)n) : 0;
  keys[k] = (uint64_t)((n >= gridNodes ? 128 : 0) + (n >= bigNodes ? 0 : 64) + 63 - cls);
  vals[k] = (int)k;
}

// One wave per pair, the pairs taken in `order` (null: index order);
// records the path each pair took and its duration in ticks of the
// constant-rate wall clock (both optional).
__global__ void __launch_bounds__(64)
    k_label_wave(long M, const int *order, int *status, SlotArrays s, TraceArgs a, int *aux,
                 int *lpath, long long *lticks, int arenaBytes) {
  if ((long)blockIdx.x >= M) return;
  const long k = order ? order[blockIdx.x] : (long)blockIdx.x;
  if (pair_nodes(a, k) >= a.bigNodes) return;  // k_label_block's
  const long long t0 = wall_clock64();
  extern __shared__ __align__(16) char label_lds[];
  const int path = label_pair<64>(k, status, s, a, aux, arenaBytes > 0 ? label_lds : nullptr,
                                  arenaBytes);
  if (threadIdx.x == 0 && lpath) {
    lpath[k] = path;
    lticks[k] = wall_clock64() - t0;
  }
}

// The same labelling as a persistent, work-stealing launch: as many waves as
// the GPU holds at once, each taking the next pair of the order from a
// device counter until the order is exhausted (every wave reaches the exit:
// the counter only grows); the pairs of k_label_block are skipped as above.
__global__ void __launch_bounds__(64)
    k_label_wave_steal(long M, const int *order, unsigned long long *next, int *status,
                       SlotArrays s, TraceArgs a, int *aux, int *lpath, long long *lticks,
                       int arenaBytes) {
  extern __shared__ __align__(16) char label_lds[];
  for (;;) {
    unsigned long long idx = 0;
    if (threadIdx.x == 0) idx = atomicAdd(next, 1ULL);
    idx = __shfl(idx, 0, 64);
    if (idx >= (unsigned long long)M) return;
    const long k = order ? order[idx] : (long)idx;
    if (pair_nodes(a, k) >= a.bigNodes) continue;  // k_label_block's
    const long long t0 = wall_clock64();
    const int path = label_pair<64>(k, status, s, a, aux, arenaBytes > 0 ? label_lds : nullptr,
                                    arenaBytes);
    if (threadIdx.x == 0 && lpath) {
      lpath[k] = path;
      lticks[k] = wall_clock64() - t0;
    }
    __syncthreads();  // the arena is reused by the next pair
  }
}

// the pairs k_label_block takes (bigNodes <= nodes < gridNodes) appended to
// list, cnt[0] their number, and those k_label_grid takes (nodes >=
// gridNodes) to glist, cnt[1] their number; by size alone, whatever their
// status, so that each pair is labelled by exactly one of the three kernels
__global__ void __launch_bounds__(1024)
    k_big_pairs(long M, const long long *offP, const long long *offQ, const long long *psP2,
                const long long *psQ2, long long bigNodes, long long gridNodes, int *list,
                int *glist, unsigned long long *cnt) {
  __shared__ unsigned long long lds[17];
  const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long n = k < M ? pair_nodes(offP, offQ, psP2, psQ2, k) : 0;
  const bool big = k < M && n >= bigNodes && n < gridNodes;
  const long i = block_append(cnt, big, lds);
  if (i >= 0) list[i] = (int)k;
  __syncthreads();
  const bool huge = k < M && n >= gridNodes;
  const long j = block_append(cnt + 1, huge, lds);
  if (j >= 0) glist[j] = (int)k;
}

// One pair labelled by every thread of a cooperative launch (the pairs of
// Tuning::gridNodes nodes or more, one launch each); part: the grid
// reductions' partials, 2 * gridDim.x entries
__global__ void __launch_bounds__(kGridBlock)
    k_label_grid(long k, int *status, SlotArrays s, TraceArgs a, int *aux, long long *part,
                 int *lpath, long long *lticks) {
  if (threadIdx.x == 0) {
    grid_part = part;
    grid_half = 0;
  }
  __syncthreads();
  const long long t0 = wall_clock64();
  const int path = label_pair<kGrid>(k, status, s, a, aux, nullptr, 0);
  if (blockIdx.x == 0 && threadIdx.x == 0 && lpath) {
    lpath[k] = path;
    lticks[k] = wall_clock64() - t0;
  }
}

// One block of kBigGroup threads per listed pair, the blocks taking the list
// in turn
__global__ void __launch_bounds__(kBigGroup)
    k_label_block(const int *list, const unsigned long long *cnt, unsigned long long *next,
                  int *status, SlotArrays s, TraceArgs a, int *aux, int *lpath,
                  long long *lticks) {
  const long n = (long)*cnt;
  __shared__ long s_i;
  for (;;) {
    // the next listed pair from a device counter: a block held up by a slow
    // pair (a sequential fallback) takes no further pairs meanwhile
    if (threadIdx.x == 0) s_i = (long)atomicAdd(next, 1ULL);
    __syncthreads();
    const long i = s_i;
    if (i >= n) return;
    const long k = list[i];
    const long long t0 = wall_clock64();
    const int path = label_pair<kBigGroup>(k, status, s, a, aux, nullptr, 0);
    if (threadIdx.x == 0 && lpath) {
      lpath[k] = path;
      lticks[k] = wall_clock64() - t0;
    }
    __syncthreads();
  }
}

// pairs by final status: OK, invalid, disjoint, failed
// (a grid of at most kStatBlocks 256-thread blocks striding over the pairs,
// one atomic per block and status: atomics per wave on one cache line
// serialise across the chip)
__global__ void __launch_bounds__(256)
    k_status_counts(long M, const int *status, unsigned long long *cnt) {
  __shared__ unsigned long long part[4][4];
  unsigned long long c[4] = {0, 0, 0, 0};
  for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < M;
       k += (long)gridDim.x * blockDim.x) {
    const int s = status[k];
    c[s == GH_OK ? 0 : s == GH_INVALID ? 1 : s == GH_DISJOINT ? 2 : 3]++;
  }
  for (int w = 0; w < 4; w++) {
    unsigned long long v = c[w];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (__lane_id() == 0) part[threadIdx.x >> 6][w] = v;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    const unsigned long long v = part[0][threadIdx.x] + part[1][threadIdx.x] +
                                 part[2][threadIdx.x] + part[3][threadIdx.x];
    if (v) atomicAdd(cnt + threadIdx.x, v);
  }
}

// node / component capacity per pair
__global__ void k_capacity(long M, const long long *offP, const long long *offQ,
                           const long long *psP2, const long long *psQ2, const int *status,
                           int *nodeCap, int *compCap) {
  long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  if (status[k] != GH_OK) {
    nodeCap[k] = 0;
    compCap[k] = 0;
    return;
  }
  long long cP = psP2[offP[k + 1]] - psP2[offP[k]];
  long long cQ = psQ2[offQ[k + 1]] - psQ2[offQ[k]];
  long long n = 2 * (cP + cQ) + 4;
  nodeCap[k] = (int)(n > 0x7fffffff ? 0x7fffffff : n);
  compCap[k] = (int)(n / 2 + 4);
}

// final dense output: result vertices and component sizes in pair order
// (each pair split over several blocks at width 1024)
template <int W>
__global__ void __launch_bounds__(W == 64 ? 256 : W)
    k_gather_out(long M, int split, const long long *nodeOff, const long long *compCapOff,
                 const long long *vOff, const long long *cOff, const double *xyTmp,
                 const int *compLen, double *xy, int *compSizes) {
  const long k = pgrp_split_pair<W>(split);
  if (k >= M) return;
  const int lane = pgrp_lane<W>() + W * pgrp_split_part<W>(split);
  const long long stride = (long long)W * split;
  const long long nv = vOff[k + 1] - vOff[k], nc = cOff[k + 1] - cOff[k];
  for (long long i = lane; i < nv; i += stride) {
    xy[2 * (vOff[k] + i)] = xyTmp[2 * (nodeOff[k] + i)];
    xy[2 * (vOff[k] + i) + 1] = xyTmp[2 * (nodeOff[k] + i) + 1];
  }
  for (long long i = lane; i < nc; i += stride)
    compSizes[cOff[k] + i] = compLen[compCapOff[k] + i];
}

template <typename T>
static T *wsT(Context &c, const char *name, long long count) {
  return (T *)c.ws.get(name, sizeof(T) * (size_t)(count > 0 ? count : 1));
}

// Signed area of each result component, one thread per component: the fan
// from the component's first vertex, summed in vertex order exactly as the
// host's gcmf_ring_area (both built without contraction), so the two agree
// bit for bit.
__global__ void __launch_bounds__(256)
    k_comp_area(long nc, const int *sizes, const long long *voff, const double *xy,
                double *out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const long long b = voff[c];
  const int n = sizes[c];
  double a = 0.0;
  if (n >= 3) {
    const double x0 = xy[2 * b], y0 = xy[2 * b + 1];
    double s = 0.0;
    for (int i = 1; i + 1 < n; i++) {
      const double ax = xy[2 * (b + i)] - x0, ay = xy[2 * (b + i) + 1] - y0;
      const double bx = xy[2 * (b + i + 1)] - x0, by = xy[2 * (b + i + 1) + 1] - y0;
      s += ax * by - bx * ay;
    }
    a = 0.5 * s;
  }
  out[c] = a;
}

void comp_areas(Context &c, const ClipResult &res, double *out) {
  const long nc = res.totalComps;
  if (nc <= 0) return;
  long long *voff = wsT<long long>(c, "gh_area_voff", nc + 1);
  scan_i32_to_i64(c, res.compSizes, voff, nc);
  hipLaunchKernelGGL(k_comp_area, dim3(grid_for(nc, 256)), dim3(256), 0, c.stream, nc,
                     (const int *)res.compSizes, (const long long *)voff,
                     (const double *)res.xy, out);
  HIP_CHECK(hipGetLastError());
}

void clip_pairs(Context &c, const JoinInputs &in, const int *pairs, long M,
                const ClipOptions &opt, ClipResult &res, Phase1View *p1) {
  res = ClipResult();
  res.M = M;
  if (M <= 0) return;
  // group width of the per-pair setup kernels: a 1024-thread block per pair
  // for a few pairs, four waves per block otherwise
  const int pw = M <= 32 ? 1024 : 64;
  PairMeta pm;
  pm.startP = wsT<long>(c, "gh_startP", M);
  pm.startQ = wsT<long>(c, "gh_startQ", M);
  pm.nP = wsT<int>(c, "gh_nP", M);
  pm.nQ = wsT<int>(c, "gh_nQ", M);
  pm.swapped = wsT<int>(c, "gh_swapped", M);
  pm.status = wsT<int>(c, "gh_status", M);
  pm.norm = wsT<double>(c, "gh_norm", 3 * M);
  pm.cmbrN = wsT<double>(c, "gh_cmbrN", 4 * M);
  int *eP = wsT<int>(c, "gh_eP", M), *eQ = wsT<int>(c, "gh_eQ", M);
  long long *offP = wsT<long long>(c, "gh_offP", M + 1);
  long long *offQ = wsT<long long>(c, "gh_offQ", M + 1);
  // the call's counters, cleared by one fill: [0, 2) largest rings, [2, 5)
  // count statistics, [8, 24) run lists, [24, 26) large pairs, [26] and
  // [27] labelling work counters, [28, 32) status counts
  unsigned long long *ctr = wsT<unsigned long long>(c, "gh_counters", 32);
  fill_i32(c, (int *)ctr, 0, 64);
  unsigned long long *maxRing = ctr;
  {
    const long nb = pgrp_grid(M, pw);
    unsigned long long *blockMax = wsT<unsigned long long>(c, "gh_setup_bmax", 2 * nb);
    {
      LaunchTimer t(c, "gh_pair_setup");
      // few large pairs: the ring boxes over several blocks per pair first
      const int split = pgrp_split(M, pw);
      double *part = nullptr;
      if (split > 1) {
        part = wsT<double>(c, "gh_setup_parts", 8 * M * (long)split);
        hipLaunchKernelGGL(k_pair_box_parts, dim3((unsigned)(M * split)), dim3(1024), 0,
                           c.stream, M, split, pairs, in.bCoords, in.bVNum, in.bVPSNum,
                           in.oCoords, in.oVNum, in.oVPSNum, opt.noswap, part);
        HIP_CHECK(hipGetLastError());
      }
      hipLaunchKernelGGL(pw == 64 ? k_pair_setup<64> : k_pair_setup<1024>, dim3((unsigned)nb),
                         dim3(pgrp_block(pw)), 0, c.stream, M, pairs, in.bCoords, in.bVNum,
                         in.bVPSNum, in.oCoords, in.oVNum, in.oVPSNum, opt.normalize,
                         opt.noswap, pm, eP, eQ, blockMax, (const double *)part, split);
      HIP_CHECK(hipGetLastError());
    }
    const long rblocks = std::min<long>(64, (nb + 255) / 256);
    hipLaunchKernelGGL(k_max_block_pairs, dim3((unsigned)rblocks), dim3(256), 0, c.stream,
                       (const unsigned long long *)blockMax, nb, maxRing);
    HIP_CHECK(hipGetLastError());
  }
  scan_i32_to_i64_multi(c, {{eP, offP, M}, {eQ, offQ, M}});
  long long rb[kHostSlots];
  read_scalars(c, {offP + M, offQ + M, (const long long *)maxRing,
                   (const long long *)maxRing + 1}, rb);
  const long edgesP = (long)rb[0], edgesQ = (long)rb[1];
  const long maxRingP = (long)rb[2], maxRingQ = (long)rb[3];
  int *mapP = wsT<int>(c, "gh_mapP", edgesP), *mapQ = wsT<int>(c, "gh_mapQ", edgesQ);
  double *ringP = wsT<double>(c, "gh_ringP", 2 * edgesP);
  double *ringQ = wsT<double>(c, "gh_ringQ", 2 * edgesQ);
  {
    LaunchTimer t(c, "gh_pair_fill");
    const int split = pgrp_split(M, pw);
    hipLaunchKernelGGL(pw == 64 ? k_pair_fill<64> : k_pair_fill<1024>,
                       dim3(pgrp_grid(M, pw) * (unsigned)split), dim3(pgrp_block(pw)), 0,
                       c.stream, M, split, pairs,
                       in.bCoords, in.oCoords, pm, (const long long *)offP,
                       (const long long *)offQ, mapP, mapQ, ringP, ringQ);
    HIP_CHECK(hipGetLastError());
  }
  // phase 1 of the CPU route (computeIntersections) or of the GPU route
  // (calculateIntersections, whose phase-1 arrays the view then holds)
  const bool cpuRoute = opt.route != 0;
  double *cf = nullptr;
  if (cpuRoute) {
    if (maxRingP >= (1L << 27) || maxRingQ >= (1L << 27))
      throw Error{GCMF_ERR_OVERFLOW, "CPU-route clip takes rings of fewer than 2^27 vertices"};
    cf = wsT<double>(c, "gh_cf", (long long)CF_N * M);
    const int split = pgrp_split(M, pw);
    if (split > 1) {
      // few large pairs: over several blocks per pair
      double *boxPart = wsT<double>(c, "gh_filter_box_parts", 8 * M * (long)split);
      double *lenPart = wsT<double>(c, "gh_filter_len_parts", 4 * M * (long)split);
      hipLaunchKernelGGL(k_filter_box_parts, dim3((unsigned)(M * split)), dim3(1024), 0,
                         c.stream, M, split, (const int *)pm.status, (const long long *)offP,
                         (const long long *)offQ, (const double *)ringP, (const double *)ringQ,
                         boxPart);
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_filter_box_combine, dim3((unsigned)M), dim3(256), 0, c.stream, M,
                         split, (const int *)pm.status, (const double *)boxPart, cf);
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_filter_edge_parts, dim3((unsigned)(M * split)), dim3(1024), 0,
                         c.stream, M, split, (const int *)pm.status, (const long long *)offP,
                         (const long long *)offQ, (const double *)ringP, (const double *)ringQ,
                         (const double *)cf, lenPart);
      HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(k_filter_finish, dim3((unsigned)M), dim3(256), 0, c.stream, M, split,
                         (const int *)pm.status, (const double *)lenPart, cf);
    } else {
      hipLaunchKernelGGL(pw == 64 ? k_cpu_filter_setup<64> : k_cpu_filter_setup<1024>,
                         dim3(pgrp_grid(M, pw)), dim3(pgrp_block(pw)), 0, c.stream, M,
                         (const int *)pm.status, (const long long *)offP,
                         (const long long *)offQ, (const double *)ringP, (const double *)ringQ,
                         cf);
    }
    HIP_CHECK(hipGetLastError());
  }
  int *cnt1P = wsT<int>(c, "gh_cnt1P", edgesP), *cnt2P = wsT<int>(c, "gh_cnt2P", edgesP);
  int *cnt1Q = wsT<int>(c, "gh_cnt1Q", edgesQ), *cnt2Q = wsT<int>(c, "gh_cnt2Q", edgesQ);
  int *boolP = wsT<int>(c, "gh_boolP", edgesP), *boolQ = wsT<int>(c, "gh_boolQ", edgesQ);
  int *lstP = wsT<int>(c, "gh_lstP", (long)kPartners * edgesP);
  int *lstQ = wsT<int>(c, "gh_lstQ", (long)kPartners * edgesQ);
  // [0] edge pairs classified, [1] intersection records, [2] edge pairs tested
  unsigned long long *stats = ctr + 2;
  // run boxes of a side, when some ring of it is long enough to be culled
  auto run_boxes = [&](bool qside, long edges, const int *map, const long long *off,
                       const double *ring) {
    RunBoxes r{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c.tune.cullMin};
    const long r1 = (edges + kRun1 - 1) / kRun1, r2 = (r1 + 15) / 16, r3 = (r2 + 15) / 16;
    double *b1 = wsT<double>(c, qside ? "gh_rb1Q" : "gh_rb1P", 4 * r1);
    double *b2 = wsT<double>(c, qside ? "gh_rb2Q" : "gh_rb2P", 4 * r2);
    double *b3 = wsT<double>(c, qside ? "gh_rb3Q" : "gh_rb3P", 4 * r3);
    double *l1 = wsT<double>(c, qside ? "gh_rl1Q" : "gh_rl1P", r1);
    double *l2 = wsT<double>(c, qside ? "gh_rl2Q" : "gh_rl2P", r2);
    double *l3 = wsT<double>(c, qside ? "gh_rl3Q" : "gh_rl3P", r3);
    hipLaunchKernelGGL(k_run_boxes, dim3(grid_for(edges, 256)), dim3(256), 0, c.stream, edges,
                       map, off, ring, b1, l1);
    hipLaunchKernelGGL(k_run_boxes_merge, dim3(grid_for(r2, 256)), dim3(256), 0, c.stream, r1,
                       (const double *)b1, (const double *)l1, b2, l2);
    hipLaunchKernelGGL(k_run_boxes_merge, dim3(grid_for(r3, 256)), dim3(256), 0, c.stream, r2,
                       (const double *)b2, (const double *)l2, b3, l3);
    HIP_CHECK(hipGetLastError());
    r.b1 = b1;
    r.b2 = b2;
    r.b3 = b3;
    r.l1 = l1;
    r.l2 = l2;
    r.l3 = l3;
    return r;
  };
  RunBoxes boxP{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c.tune.cullMin};
  RunBoxes boxQ{nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, c.tune.cullMin};
  if (maxRingP > c.tune.cullMin && edgesP > 0)
    boxP = run_boxes(false, edgesP, mapP, (const long long *)offP, (const double *)ringP);
  if (maxRingQ > c.tune.cullMin && edgesQ > 0)
    boxQ = run_boxes(true, edgesQ, mapQ, (const long long *)offQ, (const double *)ringQ);
  long long *psP1 = wsT<long long>(c, "gh_psP1", edgesP + 1);
  long long *psP2 = wsT<long long>(c, "gh_psP2", edgesP + 1);
  long long *psQ1 = wsT<long long>(c, "gh_psQ1", edgesQ + 1);
  long long *psQ2 = wsT<long long>(c, "gh_psQ2", edgesQ + 1);
  unsigned long long *runCnt = ctr + 8;  // 8 counters per side
  int *listMP = wsT<int>(c, "gh_runListMP", edgesP), *listMQ = wsT<int>(c, "gh_runListMQ", edgesQ);
  int *listSP = wsT<int>(c, "gh_runListSP", edgesP), *listSQ = wsT<int>(c, "gh_runListSQ", edgesQ);
  int *listOP = wsT<int>(c, "gh_runListOP", edgesP), *listOQ = wsT<int>(c, "gh_runListOQ", edgesQ);
  int *listP = wsT<int>(c, "gh_runListP", edgesP), *lenP = wsT<int>(c, "gh_runLenP", edgesP);
  int *listQ = wsT<int>(c, "gh_runListQ", edgesQ), *lenQ = wsT<int>(c, "gh_runLenQ", edgesQ);
  int *actP = wsT<int>(c, "gh_actP", edgesP), *actQ = wsT<int>(c, "gh_actQ", edgesQ);
  int *bigList = wsT<int>(c, "gh_bigList", M);
  unsigned long long *bigCnt = ctr + 24;
  int *gridList = wsT<int>(c, "gh_gridList", M);
  int *lstAllP = nullptr, *lstAllQ = nullptr;
  if (cpuRoute) {
    // -- the P view counts, lists and partner lists: pairs whose Q ring
    // fits in LDS by the binned pass, the others by the ring walk
    if (edgesP > 0) {
      const int capQ = (int)std::min<long>(maxRingQ, (long)std::min(c.tune.ldsMaxQ, kLdsQ));
      const int minQ = (int)std::min<long>(c.tune.cullMin, 0x7fffffffL);
      if (capQ > minQ) {
        unsigned long long *bstat = wsT<unsigned long long>(c, "gh_lds_stats", 2 * M);
        {
          LaunchTimer t(c, "gh_count");
          hipLaunchKernelGGL(k_count_cpu_lds, dim3((unsigned)M), dim3(kLdsBlock),
                             lds_count_bytes(capQ), c.stream, M, minQ, capQ,
                             (const long long *)offP,
                             (const long long *)offQ, (const double *)ringP,
                             (const double *)ringQ, (const double *)cf,
                             cnt1P, cnt2P, boolP, lstP, bstat);
          HIP_CHECK(hipGetLastError());
        }
        sum_wave_stats(c, bstat, M, stats);
      }
      {  // the ring walk: every pair the LDS pass does not take
        const bool cull = boxQ.b1 != nullptr;
        // the Q edges' kRun1-edge run boxes for the waves' walks (the
        // staged walk, Tuning::countStage, walks lane by lane instead)
        const double *skipQ = c.tune.countStage ? nullptr : boxQ.b1;
        if (!skipQ && !c.tune.countStage && edgesQ > 0) {
          const long r1 = (edgesQ + kRun1 - 1) / kRun1;
          double *b1 = wsT<double>(c, "gh_rb1Q", 4 * r1);
          double *l1 = wsT<double>(c, "gh_rl1Q", r1);
          hipLaunchKernelGGL(k_run_boxes, dim3(grid_for(edgesQ, 256)), dim3(256), 0, c.stream,
                             edgesQ, mapQ, (const long long *)offQ, (const double *)ringQ, b1,
                             l1);
          HIP_CHECK(hipGetLastError());
          skipQ = b1;
        }
        const int cb = c.tune.countBlock;
        const unsigned grid = grid_for(edgesP, cb);
        const long waves = (long)grid * cb / 64;
        unsigned long long *wstat = wsT<unsigned long long>(c, "gh_wave_stats", 2 * waves);
        {
          LaunchTimer t(c, "gh_count");
          // the Q-ring staging slices: kStageQ points per wave
          const size_t stageBytes = c.tune.countStage ? (size_t)(cb / 64) * kStageQ * 16 : 0;
          // (P edges only, so the wave walk's form applies whenever skipQ is given)
          auto kern = cull ? (skipQ ? k_count_cpu<true, true> : k_count_cpu<true, false>)
                           : (skipQ ? k_count_cpu<false, true> : k_count_cpu<false, false>);
          hipLaunchKernelGGL(kern, dim3(grid), dim3(cb),
                             stageBytes, c.stream, edgesP, 0L, mapP, mapQ,
                             (const long long *)offP, (const long long *)offQ, ringP, ringQ,
                             (const double *)cf,
                             cnt1P, cnt2P, cnt1Q, cnt2Q, boolP, boolQ, wstat, lstP, lstQ, boxP,
                             boxQ, capQ > minQ ? capQ : 0, c.tune.countStage, skipQ);
          HIP_CHECK(hipGetLastError());
        }
        sum_wave_stats(c, wstat, waves, stats);
      }
    }
    scan_i32_to_i64_multi(c, {{cnt1P, psP1, edgesP}, {cnt2P, psP2, edgesP}});
    if (edgesP > 0)
      list_runs(c, RunListSide{edgesP, mapP, cnt2P, cnt1P, listSP, listMP, listP, lenP, listOP,
                               runCnt},
                RunListSide{0, mapQ, nullptr, nullptr, listSQ, listMQ, listQ, lenQ, listOQ,
                            runCnt + 8},
                pm.status);
    HIP_CHECK(hipGetLastError());
    const long long *nActPd = compact_flags_dev(c, cnt1P, edgesP, actP, "actP");
    const long long *rc = (const long long *)runCnt;
    long long ra[8];
    read_scalars(c, {psP2 + edgesP, psP1 + edgesP, rc, rc + 2, rc + 3, nActPd, rc + 1, rc + 4},
                 ra);
    const long H = (long)ra[1];  // pairs with a nonzero type
    if (H >= 0x7fffffffL)
      throw Error{GCMF_ERR_OVERFLOW, "clip batch exceeds 2^31 intersecting edge pairs"};
    rb[0] = ra[0];
    rb[2] = ra[2];
    rb[3] = ra[3];
    rb[4] = ra[4];
    rb[8] = ra[5];
    rb[10] = H;
    rb[11] = H;
    rb[12] = ra[6];
    rb[14] = ra[7];
    rb[15] = 0;  // the Q lists come complete from the transposition
    const long nActP0 = (long)ra[5], nOverP0 = (long)ra[7];
    lstAllP = wsT<int>(c, "gh_lstAllP", H);
    if (nActP0 > 0)
      hipLaunchKernelGGL(k_partner_lists, dim3(grid_for(nActP0, 256)), dim3(256), 0, c.stream,
                         nActP0, (const int *)actP, (const long long *)psP1, (const int *)lstP,
                         lstAllP);
    if (nOverP0 > 0)
      hipLaunchKernelGGL(k_partner_rescan_cpu, dim3(nOverP0 < 65536 ? nOverP0 : 65536),
                         dim3(64), 0, c.stream, nOverP0, (const int *)listOP, true, mapP,
                         (const long long *)offP, (const long long *)offQ,
                         (const double *)ringP, (const double *)ringQ, (const long long *)psP1,
                         lstAllP, (const double *)cf,
                         boxQ);
    HIP_CHECK(hipGetLastError());
    // -- the Q side by transposition
    if (edgesQ > 0)
      hipLaunchKernelGGL(k_cpu_flags_q, dim3(grid_for(edgesQ, 256)), dim3(256), 0, c.stream,
                         edgesQ, mapQ, (const long long *)offP, (const long long *)offQ,
                         (const double *)ringQ, (const double *)cf, boolQ, cnt1Q, cnt2Q);
    uint64_t *tk = wsT<uint64_t>(c, "gh_trK", H), *tk2 = wsT<uint64_t>(c, "gh_trK2", H);
    int *tv = wsT<int>(c, "gh_trV", H);
    lstAllQ = wsT<int>(c, "gh_lstAllQ", H);
    if (nActP0 > 0) {
      LaunchTimer t(c, "gh_transpose");
      hipLaunchKernelGGL(k_cpu_transpose, dim3(grid_for(nActP0, 256)), dim3(256), 0, c.stream,
                         nActP0, (const int *)actP, mapP, (const long long *)offP,
                         (const long long *)offQ, (const double *)ringP, (const double *)ringQ,
                         (const long long *)psP1, (const int *)lstAllP, cnt1Q, cnt2Q, tk, tv);
      HIP_CHECK(hipGetLastError());
    }
    if (H > 0) {
      int bits = 0;
      for (long x = edgesQ; x > 0; x >>= 1) bits++;
      radix_sort_u64(c, tk, tv, tk2, lstAllQ, H, bits);
    }
    scan_i32_to_i64_multi(c, {{cnt1Q, psQ1, edgesQ}, {cnt2Q, psQ2, edgesQ}});
    hipLaunchKernelGGL(k_pair_check, dim3(grid_for(M, 256)), dim3(256), 0, c.stream, M,
                       (const long long *)offP, (const long long *)offQ,
                       (const long long *)psP1, (const long long *)psQ1,
                       (const long long *)psP2, (const long long *)psQ2, pm.status, stats + 1);
    if (edgesQ > 0)
      list_runs(c, RunListSide{0, mapP, nullptr, nullptr, listSP, listMP, listP, lenP, listOP,
                               runCnt},
                RunListSide{edgesQ, mapQ, cnt2Q, cnt1Q, listSQ, listMQ, listQ, lenQ, listOQ,
                            runCnt + 8},
                pm.status);
    HIP_CHECK(hipGetLastError());
    const long long *nActQd = compact_flags_dev(c, cnt1Q, edgesQ, actQ, "actQ");
    hipLaunchKernelGGL(k_big_pairs, dim3(grid_for(M, 1024)), dim3(1024), 0, c.stream, M,
                       (const long long *)offP, (const long long *)offQ, (const long long *)psP2,
                       (const long long *)psQ2, (long long)c.tune.bigNodes,
                       (long long)c.tune.gridNodes, bigList, gridList, bigCnt);
    HIP_CHECK(hipGetLastError());
    long long rq[8];
    read_scalars(c, {psQ2 + edgesQ, rc + 8, rc + 10, rc + 11, nActQd, rc + 9,
                     (const long long *)bigCnt, (const long long *)bigCnt + 1},
                 rq);
    rb[1] = rq[0];
    rb[5] = rq[1];
    rb[6] = rq[2];
    rb[7] = rq[3];
    rb[9] = rq[4];
    rb[13] = rq[5];
    rb[16] = rq[6];
    rb[17] = rq[7];
  } else {
    if (edgesP + edgesQ > 0) {
      const bool cull = boxP.b1 != nullptr || boxQ.b1 != nullptr;
      const int cb = c.tune.countBlock;
      const unsigned grid = grid_for(edgesP + edgesQ, cb);
      const long waves = (long)grid * cb / 64;
      unsigned long long *wstat = wsT<unsigned long long>(c, "gh_wave_stats", 2 * waves);
      {
        LaunchTimer t(c, "gh_count");
        hipLaunchKernelGGL(cull ? k_count<true> : k_count<false>, dim3(grid), dim3(cb), 0,
                           c.stream, edgesP, edgesQ, mapP, mapQ, (const long long *)offP,
                           (const long long *)offQ, ringP, ringQ, pm.cmbrN, cnt1P, cnt2P, cnt1Q,
                           cnt2Q, boolP, boolQ, wstat, lstP, lstQ, boxP, boxQ);
        HIP_CHECK(hipGetLastError());
      }
      sum_wave_stats(c, wstat, waves, stats);
    }
    scan_i32_to_i64_multi(c, {{cnt1P, psP1, edgesP},
                              {cnt2P, psP2, edgesP},
                              {cnt1Q, psQ1, edgesQ},
                              {cnt2Q, psQ2, edgesQ}});
    hipLaunchKernelGGL(k_pair_check, dim3(grid_for(M, 256)), dim3(256), 0, c.stream, M,
                       (const long long *)offP, (const long long *)offQ, (const long long *)psP1,
                       (const long long *)psQ1, (const long long *)psP2, (const long long *)psQ2,
                       pm.status, stats + 1);
    HIP_CHECK(hipGetLastError());
    // edges whose alpha runs go through the segmented sort (statuses are final
    // here), read back with the slot totals
    list_runs(c, RunListSide{edgesP, mapP, cnt2P, cnt1P, listSP, listMP, listP, lenP, listOP,
                             runCnt},
              RunListSide{edgesQ, mapQ, cnt2Q, cnt1Q, listSQ, listMQ, listQ, lenQ, listOQ,
                          runCnt + 8},
              pm.status);
    HIP_CHECK(hipGetLastError());
    // edges with intersections (count1 > 0): the later passes run over these
    const long long *nActPd = compact_flags_dev(c, cnt1P, edgesP, actP, "actP");
    const long long *nActQd = compact_flags_dev(c, cnt1Q, edgesQ, actQ, "actQ");
    // the pairs labelled by a block (their number comes back with the totals)
    hipLaunchKernelGGL(k_big_pairs, dim3(grid_for(M, 1024)), dim3(1024), 0, c.stream, M,
                       (const long long *)offP, (const long long *)offQ, (const long long *)psP2,
                       (const long long *)psQ2, (long long)c.tune.bigNodes,
                       (long long)c.tune.gridNodes, bigList, gridList, bigCnt);
    HIP_CHECK(hipGetLastError());
    {
      const long long *rc = (const long long *)runCnt;
      read_scalars(c, {psP2 + edgesP, psQ2 + edgesQ, rc, rc + 2, rc + 3, rc + 8, rc + 10, rc + 11,
                       nActPd, nActQd, psP1 + edgesP, psQ1 + edgesQ, rc + 1, rc + 9, rc + 4,
                       rc + 12, (const long long *)bigCnt, (const long long *)bigCnt + 1},
                   rb);
    }
    const long nActP = (long)rb[8], nActQ = (long)rb[9];
    const long nOverP = (long)rb[14], nOverQ = (long)rb[15];
    // every active edge's partners at its count1 offset
    lstAllP = wsT<int>(c, "gh_lstAllP", rb[10]);
    lstAllQ = wsT<int>(c, "gh_lstAllQ", rb[11]);
    if (nActP > 0)
      hipLaunchKernelGGL(k_partner_lists, dim3(grid_for(nActP, 256)), dim3(256), 0, c.stream, nActP,
                         (const int *)actP, (const long long *)psP1, (const int *)lstP, lstAllP);
    if (nActQ > 0)
      hipLaunchKernelGGL(k_partner_lists, dim3(grid_for(nActQ, 256)), dim3(256), 0, c.stream, nActQ,
                         (const int *)actQ, (const long long *)psQ1, (const int *)lstQ, lstAllQ);
    if (nOverP > 0 && cpuRoute)
      hipLaunchKernelGGL(k_partner_rescan_cpu, dim3(nOverP < 65536 ? nOverP : 65536), dim3(64), 0,
                         c.stream, nOverP, (const int *)listOP, true, mapP,
                         (const long long *)offP, (const long long *)offQ, (const double *)ringP,
                         (const double *)ringQ, (const long long *)psP1, lstAllP,
                         (const double *)cf, boxQ);
    else if (nOverP > 0)
      hipLaunchKernelGGL(k_partner_rescan, dim3(nOverP < 65536 ? nOverP : 65536), dim3(64), 0,
                         c.stream, nOverP, (const int *)listOP, mapP, (const long long *)offP,
                         (const long long *)offQ, (const double *)ringP, (const double *)ringQ,
                         (const long long *)psP1, lstAllP, boxQ);
    if (nOverQ > 0 && cpuRoute)
      hipLaunchKernelGGL(k_partner_rescan_cpu, dim3(nOverQ < 65536 ? nOverQ : 65536), dim3(64), 0,
                         c.stream, nOverQ, (const int *)listOQ, false, mapQ,
                         (const long long *)offQ, (const long long *)offP, (const double *)ringQ,
                         (const double *)ringP, (const long long *)psQ1, lstAllQ,
                         (const double *)cf, boxP);
    else if (nOverQ > 0)
      hipLaunchKernelGGL(k_partner_rescan, dim3(nOverQ < 65536 ? nOverQ : 65536), dim3(64), 0,
                         c.stream, nOverQ, (const int *)listOQ, mapQ, (const long long *)offQ,
                         (const long long *)offP, (const double *)ringQ, (const double *)ringP,
                         (const long long *)psQ1, lstAllQ, boxP);
    HIP_CHECK(hipGetLastError());
  }
  const long nBig = (long)rb[16], nGrid = (long)rb[17];
  const long nActP = (long)rb[8], nActQ = (long)rb[9];
  const long nShortP = (long)rb[12], nShortQ = (long)rb[13];
  const long slotsP = (long)rb[0], slotsQ = (long)rb[1];
  const int nMedP = (int)rb[2], nRunsP = (int)rb[3], nMedQ = (int)rb[5], nRunsQ = (int)rb[6];
  const long runTotP = (long)rb[4], runTotQ = (long)rb[7];
  if (slotsP >= 0x7fffffffL || slotsQ >= 0x7fffffffL)
    throw Error{GCMF_ERR_OVERFLOW, "clip batch exceeds 2^31 intersection slots"};
  SlotArrays s;
  s.iP = wsT<double>(c, "gh_iP", 2 * slotsP);
  s.iP2 = wsT<double>(c, "gh_iP2", 2 * slotsP);
  s.oP = wsT<double>(c, "gh_oP", 2 * slotsP);
  s.oP2 = wsT<double>(c, "gh_oP2", 2 * slotsP);
  s.iQ = wsT<double>(c, "gh_iQ", 2 * slotsQ);
  s.iQ2 = wsT<double>(c, "gh_iQ2", 2 * slotsQ);
  s.oQ = wsT<double>(c, "gh_oQ", 2 * slotsQ);
  s.oQ2 = wsT<double>(c, "gh_oQ2", 2 * slotsQ);
  s.aP = wsT<int>(c, "gh_aP", slotsP);
  s.aP2 = wsT<int>(c, "gh_aP2", slotsP);
  s.aQ = wsT<int>(c, "gh_aQ", slotsQ);
  s.aQ2 = wsT<int>(c, "gh_aQ2", slotsQ);
  s.nbP = wsT<int>(c, "gh_nbP", slotsP);
  s.nbP2 = wsT<int>(c, "gh_nbP2", slotsP);
  s.nbQ = wsT<int>(c, "gh_nbQ", slotsQ);
  s.nbQ2 = wsT<int>(c, "gh_nbQ2", slotsQ);
  s.lP = wsT<int>(c, "gh_lP", slotsP);
  s.lQ = wsT<int>(c, "gh_lQ", slotsQ);
  int *nmap = wsT<int>(c, "gh_nmap", slotsQ);
  const long slotsMax = slotsP > slotsQ ? slotsP : slotsQ;
  if (slotsMax > 0) {
    hipLaunchKernelGGL(k_init_slots, dim3(grid_for(slotsMax, 256)), dim3(256), 0, c.stream,
                       slotsP, slotsQ, s, nmap, cpuRoute ? 0 : 1);
    HIP_CHECK(hipGetLastError());
  }
  if (cpuRoute) {
    CpuSlots cs;
    cs.kP2 = wsT<uint64_t>(c, "gh_kP2", slotsP);
    cs.kQ2 = wsT<uint64_t>(c, "gh_kQ2", slotsQ);
    cs.rP2 = wsT<int>(c, "gh_rP2", slotsP);
    cs.rQ2 = wsT<int>(c, "gh_rQ2", slotsQ);
    cs.rP = wsT<int>(c, "gh_rP", slotsP);
    cs.rQ = wsT<int>(c, "gh_rQ", slotsQ);
    cs.dP = wsT<int>(c, "gh_dP", slotsP);
    cs.dQ = wsT<int>(c, "gh_dQ", slotsQ);
    cs.srcP = wsT<int>(c, "gh_srcP", edgesP);
    cs.srcQ = wsT<int>(c, "gh_srcQ", edgesQ);
    if (edgesP + edgesQ > 0)
      hipLaunchKernelGGL(k_parents, dim3(grid_for(edgesP + edgesQ, 256)), dim3(256), 0,
                         c.stream, edgesP, edgesQ, mapP, mapQ, (const long long *)offP,
                         (const long long *)offQ, ringP, ringQ, (const long long *)psP2,
                         (const long long *)psQ2, pm.status, pm, in.bCoords, in.oCoords, s);
    HIP_CHECK(hipGetLastError());
    const long nAct = nActP + nActQ;
    RunLists runLists{{listMP, listP, listMQ, listQ}, {nMedP, nRunsP, nMedQ, nRunsQ}};
    const long nWaveRuns = (long)nMedP + nRunsP + nMedQ + nRunsQ;
    if (nAct > 0) {
      LaunchTimer t(c, "gh_cpu_fill");
      hipLaunchKernelGGL(k_cpu_fill, dim3(grid_for(nAct, 256)), dim3(256), 0, c.stream, nActP,
                         nActQ, (const int *)actP, (const int *)actQ, mapP, mapQ,
                         (const long long *)offP, (const long long *)offQ, ringP, ringQ,
                         (const long long *)psP1, (const long long *)psQ1,
                         (const long long *)psP2, (const long long *)psQ2, pm.status, pm,
                         in.bCoords, in.oCoords, (const int *)lstAllP, (const int *)lstAllQ, s,
                         cs);
      HIP_CHECK(hipGetLastError());
      if (nWaveRuns > 0)
        hipLaunchKernelGGL(k_cpu_fill_runs, dim3(grid_for(64 * nWaveRuns, 256)), dim3(256), 0,
                           c.stream, runLists, mapP, mapQ, (const long long *)offP,
                           (const long long *)offQ, ringP, ringQ, (const long long *)psP1,
                           (const long long *)psQ1, (const long long *)psP2,
                           (const long long *)psQ2, pm.status, pm, in.bCoords, in.oCoords,
                           (const int *)lstAllP, (const int *)lstAllQ, s, cs);
      HIP_CHECK(hipGetLastError());
    }
    // per-edge sorts by exact alpha (beta), stable in slot order
    auto sort_side = [&](bool qside, long nShort, const int *lS, long nMed, const int *lM,
                         long nRuns, long runTot, const int *lL, const int *lenL,
                         const long long *ps2) {
      if (nShort > 0) {
        LaunchTimer t(c, "gh_cpu_sort");
        hipLaunchKernelGGL(qside ? k_cpu_sort_short<true> : k_cpu_sort_short<false>,
                           dim3(grid_for(32 * nShort, 256)), dim3(256), 0, c.stream, nShort, lS,
                           ps2, s, cs);
        HIP_CHECK(hipGetLastError());
      }
      if (nMed > 0) {
        LaunchTimer t(c, "gh_sort_runs");
        hipLaunchKernelGGL(qside ? k_cpu_sort_block<true> : k_cpu_sort_block<false>,
                           dim3(nMed < 8192 ? nMed : 8192), dim3(256), 0, c.stream, (int)nMed,
                           lM, ps2, s, cs);
        HIP_CHECK(hipGetLastError());
      }
      if (nRuns > 0) {
        long long *runOff = wsT<long long>(c, qside ? "gh_runOffQ" : "gh_runOffP", nRuns + 1);
        scan_i32_to_i64(c, lenL, runOff, nRuns);
        uint64_t *k0 = wsT<uint64_t>(c, "gh_runK0", runTot);
        uint64_t *k1 = wsT<uint64_t>(c, "gh_runK1", runTot);
        int *v0 = wsT<int>(c, "gh_runV0", runTot), *v1 = wsT<int>(c, "gh_runV1", runTot);
        int *runOf = wsT<int>(c, "gh_runOf", runTot);
        const unsigned blocks = (unsigned)(nRuns < 4096 ? nRuns : 4096);
        hipLaunchKernelGGL(k_cpu_long_keys, dim3(blocks), dim3(256), 0, c.stream, (int)nRuns, lL,
                           (const long long *)runOff, ps2,
                           (const uint64_t *)(qside ? cs.kQ2 : cs.kP2), k0, v0, runOf);
        radix_sort_u64(c, k0, v0, k1, v1, runTot, 64);
        hipLaunchKernelGGL(k_cpu_long_runkey, dim3(grid_for(runTot, 256)), dim3(256), 0,
                           c.stream, runTot, (const int *)v1, (const int *)runOf, k0);
        int bits = 0;
        for (long x = nRuns; x > 0; x >>= 1) bits++;
        radix_sort_u64(c, k0, v1, k1, v0, runTot, bits);
        hipLaunchKernelGGL(qside ? k_cpu_long_apply<true> : k_cpu_long_apply<false>,
                           dim3(grid_for(runTot, 256)), dim3(256), 0, c.stream, runTot,
                           (const int *)v0, (const int *)runOf, lL, (const long long *)runOff,
                           ps2, s, cs);
        HIP_CHECK(hipGetLastError());
      }
    };
    sort_side(false, nShortP, listSP, nMedP, listMP, nRunsP, runTotP, listP, lenP,
              (const long long *)psP2);
    sort_side(true, nShortQ, listSQ, nMedQ, listMQ, nRunsQ, runTotQ, listQ, lenQ,
              (const long long *)psQ2);
    if (nAct > 0) {
      LaunchTimer t(c, "gh_cpu_link");
      hipLaunchKernelGGL(k_cpu_link, dim3(grid_for(nAct, 256)), dim3(256), 0, c.stream, nActP,
                         nActQ, (const int *)actP, (const int *)actQ, mapP, mapQ,
                         (const long long *)offP, (const long long *)offQ,
                         (const long long *)psP2, (const long long *)psQ2, pm.status, s, cs);
      HIP_CHECK(hipGetLastError());
      if (nWaveRuns > 0)
        hipLaunchKernelGGL(k_cpu_link_runs, dim3(grid_for(64 * nWaveRuns, 256)), dim3(256), 0,
                           c.stream, runLists, mapP, mapQ, (const long long *)offP,
                           (const long long *)offQ, (const long long *)psP2,
                           (const long long *)psQ2, pm.status, s, cs);
      HIP_CHECK(hipGetLastError());
    }
  } else {
    if (edgesQ > 0) {
      {
        LaunchTimer t(c, "gh_nmap");
        if (nActQ > 0)
        hipLaunchKernelGGL(k_nmap, dim3(grid_for(nActQ, 256)), dim3(256), 0, c.stream, nActQ,
                           (const int *)actQ, mapQ, (const long long *)offP, (const long long *)offQ,
                           ringP, ringQ, (const long long *)psP1, (const long long *)psQ1,
                           (const long long *)psQ2, pm.status, nmap, (const int *)lstAllQ);
        HIP_CHECK(hipGetLastError());
      }
      hipLaunchKernelGGL(k_q_parents, dim3(grid_for(edgesQ, 256)), dim3(256), 0, c.stream, edgesQ,
                         mapQ, (const long long *)offQ, ringQ, (const long long *)psQ2, pm.status,
                         pm, in.bCoords, in.oCoords, s);
      HIP_CHECK(hipGetLastError());
    }
    Owners own;
    own.a = wsT<int>(c, "gh_ownA", slotsQ);
    own.c = wsT<int>(c, "gh_ownC", slotsQ);
    own.n = wsT<int>(c, "gh_ownN", slotsQ);
    int *ownP = wsT<int>(c, "gh_ownP", slotsP);
    int *ownL = wsT<int>(c, "gh_ownL", slotsQ);
    fill_i32(c, own.a, -1, slotsQ);
    fill_i32(c, own.c, -1, slotsQ);
    fill_i32(c, own.n, -1, slotsQ);
    fill_i32(c, ownP, -1, slotsP);
    fill_i32(c, ownL, -1, slotsQ);
    if (edgesP > 0) {
      hipLaunchKernelGGL(k_p_parents, dim3(grid_for(edgesP, 256)), dim3(256), 0, c.stream, edgesP,
                         mapP, (const long long *)offP, ringP, (const long long *)psP2, pm.status,
                         pm, in.bCoords, in.oCoords, s);
      HIP_CHECK(hipGetLastError());
    }
    if (nActP > 0) {
      {
        LaunchTimer t(c, "gh_calc_own");
        hipLaunchKernelGGL(k_calc<true>, dim3(grid_for(nActP, 256)), dim3(256), 0, c.stream,
                           nActP, (const int *)actP, mapP, (const long long *)offP,
                           (const long long *)offQ, ringP, ringQ, (const long long *)psP1,
                           (const long long *)psP2, (const long long *)psQ1,
                           (const long long *)psQ2, pm.status, pm, in.bCoords, in.oCoords, nmap, s,
                           own, (const int *)lstAllP);
        HIP_CHECK(hipGetLastError());
      }
      LaunchTimer t(c, "gh_calc");
      hipLaunchKernelGGL(k_calc<false>, dim3(grid_for(nActP, 256)), dim3(256), 0, c.stream, nActP,
                         (const int *)actP, mapP, (const long long *)offP, (const long long *)offQ,
                         ringP, ringQ, (const long long *)psP1, (const long long *)psP2,
                         (const long long *)psQ1, (const long long *)psQ2, pm.status, pm,
                         in.bCoords, in.oCoords, nmap, s, own, (const int *)lstAllP);
      HIP_CHECK(hipGetLastError());
    }
    // per-edge alpha sorts, all P edges before the Q edges (the Q sort reads
    // the back links the P sort moves)
    {
      // keys, segmented sort and moves of one side's long runs
      auto long_runs = [&](bool qside, int n, long total, const int *list, const int *len,
                           const long long *ps2, const int *map, const long long *off) {
        long long *runOff = wsT<long long>(c, qside ? "gh_runOffQ" : "gh_runOffP", n + 1);
        scan_i32_to_i64(c, len, runOff, n);
        uint64_t *k0 = wsT<uint64_t>(c, "gh_runK0", total), *k1 = wsT<uint64_t>(c, "gh_runK1", total);
        int *v0 = wsT<int>(c, "gh_runV0", total), *v1 = wsT<int>(c, "gh_runV1", total);
        const unsigned blocks = (unsigned)(n < 4096 ? n : 4096);
        hipLaunchKernelGGL(k_long_run_keys<false>, dim3(blocks), dim3(256), 0, c.stream, n, list,
                           (const long long *)runOff, ps2,
                           (const int *)(qside ? s.aQ2 : s.aP2),
                           (const int *)(qside ? s.nbQ2 : s.nbP2), ownP, k0, v0);
        HIP_CHECK(hipGetLastError());
        return [=, &c, &s, &own](void) {
          int bits = 32;
          for (int x = n - 1; x > 0; x >>= 1) bits++;
          radix_sort_u64(c, k0, v0, k1, v1, total, bits);
          if (qside)
            hipLaunchKernelGGL(k_long_run_apply<true>, dim3(grid_for(total, 256)), dim3(256), 0,
                               c.stream, total, (const uint64_t *)k1, (const int *)v1, list,
                               (const long long *)runOff, ps2, map, off, s, own,
                               (const int *)ownP);
          else
            hipLaunchKernelGGL(k_long_run_apply<false>, dim3(grid_for(total, 256)), dim3(256), 0,
                               c.stream, total, (const uint64_t *)k1, (const int *)v1, list,
                               (const long long *)runOff, ps2, map, off, s, own,
                               (const int *)ownP);
          HIP_CHECK(hipGetLastError());
        };
      };
      if (nShortP > 0) {
        LaunchTimer t(c, "gh_sort_p");
        hipLaunchKernelGGL(k_sort_p_short, dim3(grid_for(32 * nShortP, 256)), dim3(256), 0,
                           c.stream, nShortP, (const int *)listSP, mapP, (const long long *)offP,
                           (const long long *)psP2, pm.status, s, own);
        HIP_CHECK(hipGetLastError());
      }
      if (nMedP > 0) {
        LaunchTimer t(c, "gh_sort_runs");
        hipLaunchKernelGGL(k_run_sort_block<false>, dim3(nMedP < 8192 ? nMedP : 8192), dim3(256),
                           0, c.stream, nMedP, (const int *)listMP, (const long long *)psP2, mapP,
                           (const long long *)offP, s, own, (const int *)ownP);
        HIP_CHECK(hipGetLastError());
      }
      if (nRunsP > 0)
        long_runs(false, nRunsP, runTotP, listP, lenP, (const long long *)psP2, mapP,
                  (const long long *)offP)();
      if (edgesQ > 0) {
        if (nShortQ > 0)
          hipLaunchKernelGGL(k_sort_q_short<true>, dim3(grid_for(32 * nShortQ, 256)), dim3(256), 0,
                             c.stream, nShortQ, (const int *)listSQ, mapQ,
                             (const long long *)psQ2, pm.status, s, ownP);
        if (nMedQ > 0)
          hipLaunchKernelGGL(k_run_own, dim3(nMedQ < 8192 ? nMedQ : 8192), dim3(256), 0, c.stream,
                             nMedQ, (const int *)listMQ, (const long long *)psQ2,
                             (const int *)s.nbQ2, ownP);
        if (nRunsQ > 0)
          hipLaunchKernelGGL(k_run_own, dim3(nRunsQ < 8192 ? nRunsQ : 8192), dim3(256), 0,
                             c.stream, nRunsQ, (const int *)listQ, (const long long *)psQ2,
                             (const int *)s.nbQ2, ownP);
        HIP_CHECK(hipGetLastError());
        auto finish = nRunsQ > 0 ? long_runs(true, nRunsQ, runTotQ, listQ, lenQ,
                                             (const long long *)psQ2, mapQ,
                                             (const long long *)offQ)
                                 : std::function<void()>();
        {
          LaunchTimer t(c, "gh_sort_q");
          if (nShortQ > 0)
            hipLaunchKernelGGL(k_sort_q_short<false>, dim3(grid_for(32 * nShortQ, 256)),
                               dim3(256), 0, c.stream, nShortQ, (const int *)listSQ, mapQ,
                               (const long long *)psQ2, pm.status, s, ownP);
          HIP_CHECK(hipGetLastError());
        }
        if (nMedQ > 0) {
          LaunchTimer t(c, "gh_sort_runs");
          hipLaunchKernelGGL(k_run_sort_block<true>, dim3(nMedQ < 8192 ? nMedQ : 8192), dim3(256),
                             0, c.stream, nMedQ, (const int *)listMQ, (const long long *)psQ2,
                             mapQ, (const long long *)offQ, s, own, (const int *)ownP);
          HIP_CHECK(hipGetLastError());
        }
        if (finish) finish();
      }
    }
    if (nActP > 0) {
      hipLaunchKernelGGL(k_init_label<true>, dim3(grid_for(4 * nActP, 256)), dim3(256), 0, c.stream,
                         nActP, (const int *)actP, mapP, (const long long *)offP,
                         (const long long *)offQ, (const long long *)psP2, (const long long *)psQ2,
                         pm.status, s, ownL);
      HIP_CHECK(hipGetLastError());
      LaunchTimer t(c, "gh_init_label");
      hipLaunchKernelGGL(k_init_label<false>, dim3(grid_for(4 * nActP, 256)), dim3(256), 0, c.stream,
                         nActP, (const int *)actP, mapP, (const long long *)offP,
                         (const long long *)offQ, (const long long *)psP2, (const long long *)psQ2,
                         pm.status, s, ownL);
      HIP_CHECK(hipGetLastError());
    }
  }

  // capacities of the per-pair node, output and component regions
  int *nodeCap = wsT<int>(c, "gh_nodeCap", M), *compCap = wsT<int>(c, "gh_compCap", M);
  long long *nodeOff = wsT<long long>(c, "gh_nodeOff", M + 1);
  long long *compCapOff = wsT<long long>(c, "gh_compCapOff", M + 1);
  hipLaunchKernelGGL(k_capacity, dim3(grid_for(M, 256)), dim3(256), 0, c.stream, M,
                     (const long long *)offP, (const long long *)offQ, (const long long *)psP2,
                     (const long long *)psQ2, pm.status, nodeCap, compCap);
  HIP_CHECK(hipGetLastError());
  scan_i32_to_i64_multi(c, {{nodeCap, nodeOff, M}, {compCap, compCapOff, M}});
  // bounds of the capacity sums (k_capacity gives an OK pair 2 (cP + cQ) + 4
  // nodes and cP + cQ + 6 components), so no read-back is needed
  const long nodes = 2 * (slotsP + slotsQ) + 4 * M;
  const long comps = slotsP + slotsQ + 6 * M;
  TraceArgs ta;
  ta.offP = offP;
  ta.offQ = offQ;
  ta.psP2 = psP2;
  ta.psQ2 = psQ2;
  ta.nodeOff = nodeOff;
  ta.compOff = compCapOff;
  ta.npn = wsT<double>(c, "gh_npn", 2 * nodes);
  ta.npo = wsT<double>(c, "gh_npo", 2 * nodes);
  ta.nnext = wsT<int>(c, "gh_nnext", nodes);
  ta.nprev = wsT<int>(c, "gh_nprev", nodes);
  ta.nnb = wsT<int>(c, "gh_nnb", nodes);
  ta.nfl = wsT<unsigned char>(c, "gh_nfl", nodes);
  ta.nlab = wsT<signed char>(c, "gh_nlab", nodes);
  ta.nenex = wsT<signed char>(c, "gh_nenex", nodes);
  ta.opn = wsT<double>(c, "gh_opn", 2 * nodes);
  ta.opo = wsT<double>(c, "gh_opo", 2 * nodes);
  ta.onext = wsT<int>(c, "gh_onext", nodes);
  ta.oprev = wsT<int>(c, "gh_oprev", nodes);
  ta.ocompRoot = wsT<int>(c, "gh_ocompRoot", comps);
  ta.outNComp = wsT<int>(c, "gh_outNComp", M);
  ta.outNVerts = wsT<int>(c, "gh_outNVerts", M);
  ta.compLen = wsT<int>(c, "gh_compLen", comps);
  ta.xyTmp = wsT<double>(c, "gh_xyTmp", 2 * nodes);
  ta.cleanup = opt.cleanup;
  ta.explicitLinks = cpuRoute ? 1 : 0;
  ta.bigNodes = c.tune.bigNodes;
  ta.gridClean = c.tune.gridClean;
  if (opt.label_mode == 1) {
    LaunchTimer t(c, "gh_label_trace");
    hipLaunchKernelGGL(k_label_trace, dim3(grid_for(M, 64)), dim3(64), 0, c.stream, M,
                       pm.status, s, ta);
    HIP_CHECK(hipGetLastError());
  } else {
    // node capacity per pair is >= 2 (nP + nQ), so 5 ints per node give each
    // pair the 10 (nP + nQ) ints of segment scratch
    int *aux = wsT<int>(c, "gh_aux", 5 * nodes);
    res.labelPath = wsT<int>(c, "gh_lpath", M);
    res.labelTicks = wsT<long long>(c, "gh_lticks", M);
    if (nBig > 0 && !c.side) {
      HIP_CHECK(hipStreamCreate(&c.side));
      HIP_CHECK(hipEventCreateWithFlags(&c.ev_fork, hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&c.ev_join, hipEventDisableTiming));
    }
    // batches of many pairs are labelled largest first; the sorted order's
    // first nBig entries are then the block path's pairs, largest first
    int *order = nullptr;
    if (M >= kLptMinPairs) {
      LaunchTimer t(c, "gh_label_order");
      uint64_t *ok = wsT<uint64_t>(c, "gh_lpt_keys", M), *ok2 = wsT<uint64_t>(c, "gh_lpt_keys2", M);
      int *ov = wsT<int>(c, "gh_lpt_vals", M);
      order = wsT<int>(c, "gh_lpt_order", M);
      hipLaunchKernelGGL(k_label_order_keys, dim3(grid_for(M, 256)), dim3(256), 0, c.stream, M,
                         (const long long *)offP, (const long long *)offQ,
                         (const long long *)psP2, (const long long *)psQ2,
                         (long long)c.tune.bigNodes, (long long)c.tune.gridNodes, ok, ov);
      HIP_CHECK(hipGetLastError());
      radix_sort_u64(c, ok, ov, ok2, order, M, 8);
    }
    LaunchTimer t(c, "gh_label_trace");
    // the large pairs on the side stream, concurrently with one wave per
    // remaining pair on the main one (disjoint pairs, disjoint regions)
    if (nBig > 0) {
      unsigned long long *bigNext = ctr + 26;
      HIP_CHECK(hipEventRecord(c.ev_fork, c.stream));
      HIP_CHECK(hipStreamWaitEvent(c.side, c.ev_fork, 0));
      const long blocks = std::min<long>(nBig, c.tune.blockCap);
      hipLaunchKernelGGL(k_label_block, dim3((unsigned)blocks), dim3(kBigGroup), 0, c.side,
                         (const int *)(order ? order : bigList),
                         (const unsigned long long *)bigCnt, bigNext, pm.status, s, ta, aux,
                         res.labelPath, res.labelTicks);
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipEventRecord(c.ev_join, c.side));
    }
    if (c.tune.labelSteal) {
      // persistent waves taking the pairs in turn
      if (c.label_steal_blocks <= 0) {
        int perCu = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &perCu, k_label_wave_steal, 64, (size_t)c.tune.labelLds));
        c.label_steal_blocks = (perCu > 0 ? perCu : 1) * c.cu_count;
        c.label_steal_lds = c.tune.labelLds;
      }
      if (c.label_steal_lds != c.tune.labelLds) {
        int perCu = 0;
        HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &perCu, k_label_wave_steal, 64, (size_t)c.tune.labelLds));
        c.label_steal_blocks = (perCu > 0 ? perCu : 1) * c.cu_count;
        c.label_steal_lds = c.tune.labelLds;
      }
      unsigned long long *next = ctr + 27;
      const long blocks = std::min<long>(M, c.label_steal_blocks);
      hipLaunchKernelGGL(k_label_wave_steal, dim3((unsigned)blocks), dim3(64),
                         (size_t)c.tune.labelLds, c.stream, M, (const int *)order, next,
                         pm.status, s, ta, aux, res.labelPath, res.labelTicks, c.tune.labelLds);
    } else {
      hipLaunchKernelGGL(k_label_wave, dim3((unsigned)M), dim3(64), (size_t)c.tune.labelLds,
                         c.stream, M, (const int *)order, pm.status, s, ta, aux, res.labelPath,
                         res.labelTicks, c.tune.labelLds);
    }
    HIP_CHECK(hipGetLastError());
    if (nBig > 0) HIP_CHECK(hipStreamWaitEvent(c.stream, c.ev_join, 0));
    if (nGrid > 0) {
      // the pairs of gridNodes nodes or more, one cooperative launch each over
      // as many blocks as the GPU holds at once, after the other labelling
      if (c.replay == 2)
        throw Error{GCMF_ERR_ARG, "pairs labelled by the whole GPU cannot be captured"};
      std::vector<int> gl((size_t)nGrid);
      host_read(c, gl.data(), gridList, sizeof(int) * (size_t)nGrid);
      int perCu = 0;
      HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k_label_grid, kGridBlock,
                                                             0));
      const int blocks = (perCu > 0 ? perCu : 1) * c.cu_count;
      long long *part = wsT<long long>(c, "gh_grid_part", 2L * blocks);
      for (int kp : gl) {
        long kk = kp;
        int *st = pm.status, *lp = res.labelPath;
        long long *lt = res.labelTicks;
        void *args[] = {&kk, &st, &s, &ta, &aux, &part, &lp, &lt};
        HIP_CHECK(hipLaunchCooperativeKernel((const void *)k_label_grid, dim3((unsigned)blocks),
                                             dim3(kGridBlock), args, 0, c.stream));
      }
    }
  }
  res.vertOff = wsT<long long>(c, "gh_vOff", M + 1);
  res.compOff = wsT<long long>(c, "gh_cOff", M + 1);
  scan_i32_to_i64_multi(c, {{ta.outNVerts, res.vertOff, M}, {ta.outNComp, res.compOff, M}});
  // a pair's result fits its node / component region, so the region totals
  // bound the outputs
  res.xy = wsT<double>(c, "gh_out_xy", 2 * nodes);
  res.compSizes = wsT<int>(c, "gh_out_comp", comps);
  {
    LaunchTimer t(c, "gh_gather_out");
    const int split = pgrp_split(M, pw);
    hipLaunchKernelGGL(pw == 64 ? k_gather_out<64> : k_gather_out<1024>,
                       dim3(pgrp_grid(M, pw) * (unsigned)split), dim3(pgrp_block(pw)), 0,
                       c.stream, M, split, (const long long *)nodeOff, (const long long *)compCapOff,
                       (const long long *)res.vertOff, (const long long *)res.compOff, ta.xyTmp,
                       ta.compLen, res.xy, res.compSizes);
    HIP_CHECK(hipGetLastError());
  }
  res.status = pm.status;
  res.swapped = pm.swapped;
  res.ncomp = ta.outNComp;
  res.nverts = ta.outNVerts;
  unsigned long long *sc = ctr + 28;
  hipLaunchKernelGGL(k_status_counts, dim3(std::min<unsigned>(kStatBlocks, grid_for(M, 256))),
                     dim3(256), 0, c.stream, M, (const int *)pm.status, sc);
  HIP_CHECK(hipGetLastError());
  {
    const long long *st = (const long long *)stats, *sl = (const long long *)sc;
    read_scalars(c, {st, st + 1, res.vertOff + M, res.compOff + M, sl, sl + 1, sl + 2, sl + 3,
                     st + 2},
                 rb);
  }
  res.nTested = rb[8];
  res.nEvalP = rb[0];
  res.totalIntersections = rb[1];
  res.totalVerts = (long)rb[2];
  res.totalComps = (long)rb[3];
  res.nOk = rb[4];
  res.nInvalid = rb[5];
  res.nDisjoint = rb[6];
  res.nFailed = rb[7];
  if (p1) {
    p1->M = M;
    p1->edgesP = edgesP;
    p1->edgesQ = edgesQ;
    p1->slotsP = slotsP;
    p1->slotsQ = slotsQ;
    p1->offP = offP;
    p1->offQ = offQ;
    p1->psP1 = psP1;
    p1->psP2 = psP2;
    p1->psQ1 = psQ1;
    p1->psQ2 = psQ2;
    p1->boolP = boolP;
    p1->boolQ = boolQ;
    p1->nmapQ = nmap;
    p1->iP = s.iP;
    p1->iQ = s.iQ;
    p1->oP = s.oP;
    p1->oQ = s.oQ;
    p1->aP = s.aP;
    p1->aQ = s.aQ;
    p1->lP = s.lP;
    p1->lQ = s.lQ;
    p1->nbP = s.nbP;
    p1->nbQ = s.nbQ;
    p1->status = pm.status;
    p1->norm = pm.norm;
  }
}

// complex polygons (several rings per side) and the union
#include "gh_complex.inl"

}  // namespace gcmf
