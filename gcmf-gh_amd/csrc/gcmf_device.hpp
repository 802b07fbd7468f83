// gcmf_device.hpp -- device-side geometric predicates shared by the join and
// the clip kernels.  All arithmetic is binary64 with contraction disabled so
// every predicate rounds exactly as the reference source writes it; the
// library is also compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace gcmf {

constexpr double kGhEps = 0.0000001;   // GH-CUDA/lib/constants.h:4
constexpr int kAlphaScale = 10000000;  // (int)pow(10, EPSILON_POSITIONS), :1034
constexpr int kAlphaDigits = 7;        // EPSILON_POSITIONS, constants.h:3
constexpr int kNone = -100;            // "no intersection" marker, :928

struct pt {
  double x, y;
};

__device__ __forceinline__ double triA(pt P, pt Q, pt R) {  // point2D.h:78-80
  return (Q.x - P.x) * (R.y - P.y) - (Q.y - P.y) * (R.x - P.x);
}

// Saturating truncation toward zero of a double to int32, NaN -> 0.
__device__ __forceinline__ int trunc_sat(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return 2147483647;
  if (v <= -2147483648.0) return (int)0x80000000;
  return (int)v;
}

// Intersection type of edges [P1,P2] and [Q1,Q2] with parameters alpha on P
// and beta on Q; polyclip.h:60-179, gppolyclip_mDIV512.cu:95-169 (including
// the reference's use of |alpha| when classifying beta on collinear edges).
__device__ __forceinline__ int itype(pt P1, pt P2, pt Q1, pt Q2, double &alpha,
                                     double &beta) {
  double AP1 = triA(P1, Q1, Q2);
  double AP2 = triA(P2, Q1, Q2);
  if (fabs(AP1 - AP2) > kGhEps) {
    double AQ1 = triA(Q1, P1, P2);
    double AQ2 = triA(Q2, P1, P2);
    double a = AP1 / (AP1 - AP2);
    double b = AQ1 / (AQ1 - AQ2);
    alpha = a;
    beta = b;
    bool a01 = (a > kGhEps) && (a < 1.0 - kGhEps);
    bool a0 = !a01 && fabs(a) <= kGhEps;
    bool b01 = (b > kGhEps) && (b < 1.0 - kGhEps);
    bool b0 = !b01 && fabs(b) <= kGhEps;
    if (a01 && b01) return 1;
    if (a0 && b01) return 2;
    if (b0 && a01) return 3;
    if (a0 && b0) return 4;
  } else if (fabs(AP1) < kGhEps) {
    pt dP = {P2.x - P1.x, P2.y - P1.y};
    pt dQ = {Q2.x - Q1.x, Q2.y - Q1.y};
    pt PQ = {Q1.x - P1.x, Q1.y - P1.y};
    double a = (PQ.x * dP.x + PQ.y * dP.y) / (dP.x * dP.x + dP.y * dP.y);
    double b = -(PQ.x * dQ.x + PQ.y * dQ.y) / (dQ.x * dQ.x + dQ.y * dQ.y);
    alpha = a;
    beta = b;
    bool a01 = (a > kGhEps) && (a < 1.0 - kGhEps);
    bool a0 = !a01 && fabs(a) <= kGhEps;
    bool an = !a01 && !a0;
    bool b01 = (b > kGhEps) && (b < 1.0 - kGhEps);
    bool b0 = !b01 && fabs(a) <= kGhEps;
    bool bn = !b01 && !b0;
    if (a01 && b01) return 5;
    if (an && b01) return 6;
    if (bn && a01) return 7;
    if (a0 && b0) return 8;
  }
  return 0;
}

__device__ __forceinline__ bool adds_vertex(int t) {  // types that insert into the first edge
  return t == 1 || t == 3 || t == 5 || t == 7;
}

// Edge bounding-box overlap, gppolyclip_mDIV512.cu:345-371 (closed).
__device__ __forceinline__ bool lsmf(pt P1, pt P2, pt Q1, pt Q2) {
  double minPX = fmin(P1.x, P2.x), maxPX = fmax(P1.x, P2.x);
  double minPY = fmin(P1.y, P2.y), maxPY = fmax(P1.y, P2.y);
  double minQX = fmin(Q1.x, Q2.x), maxQX = fmax(Q1.x, Q2.x);
  double minQY = fmin(Q1.y, Q2.y), maxQY = fmax(Q1.y, Q2.y);
  if (minPX > maxQX || maxPX < minQX) return false;
  if (minPY > maxQY || maxPY < minQY) return false;
  return true;
}

// Relative position of Q w.r.t. the chain (P1,P2,P3), polyclip.h:297-326 /
// gppolyclip_mDIV512.cu:195-217.  0 LEFT, 1 RIGHT, 2 IS_P_m, 3 IS_P_p.
__device__ __forceinline__ int rel_pos(bool qIsPmNb, bool qIsPpNb, pt Q, pt P1, pt P2, pt P3) {
  if (qIsPmNb) return 2;
  if (qIsPpNb) return 3;
  double s1 = triA(Q, P1, P2), s2 = triA(Q, P2, P3), s3 = triA(P1, P2, P3);
  if (s3 > 0) return (s1 > 0 && s2 > 0) ? 0 : 1;
  return (s1 < 0 && s2 < 0) ? 1 : 0;
}

// getInitialLabel, gppolyclip_mDIV512.cu:237-252.
__device__ __forceinline__ int initial_label(int m, int p) {
  if ((m == 0 && p == 1) || (m == 1 && p == 0)) return 1;
  if ((m == 0 && p == 0) || (m == 1 && p == 1)) return 2;
  if ((p == 3 && m == 1) || (m == 3 && p == 1)) return 3;
  if ((p == 3 && m == 0) || (m == 3 && p == 0)) return 4;
  if ((p == 3 && m == 2) || (m == 3 && p == 2)) return 5;
  if ((m == 2 && p == 1) || (p == 2 && m == 1)) return 6;
  if ((m == 2 && p == 0) || (p == 2 && m == 0)) return 7;
  return -102;
}

// wavefront helpers (wave64)
// MI355X deals the blocks of a launch to its 8 XCDs round-robin (block b to
// XCD b % 8), and each XCD has its own L2.  Kernels whose neighbouring tiles
// write into shared cache lines (scatters, compactions) process tile
// xcd_tile(b) instead of tile b: every XCD then owns one contiguous range of
// tiles, and the partly written lines at tile boundaries are merged in one
// L2 instead of being written back piecewise from two.  The map is a
// bijection of [0, tiles) for any XCD count; only its speed assumes 8.
__device__ __forceinline__ long xcd_tile(long b, long tiles) {
  const long q = tiles >> 3, r = tiles & 7, x = b & 7, k = b >> 3;
  return x * q + (x < r ? x : r) + k;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ unsigned long long ballot(bool b) { return __ballot(b); }
__device__ __forceinline__ int popc64(unsigned long long m) { return __popcll(m); }
__device__ __forceinline__ int rank_lt(unsigned long long m) {  // set bits below this lane
  return __popcll(m & ((1ULL << __lane_id()) - 1ULL));
}

}  // namespace gcmf
