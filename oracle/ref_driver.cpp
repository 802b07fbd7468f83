// ref_driver.cpp -- a thin C-ABI driver around the reference's own CPU
// Greiner-Hormann code (GH-CUDA/lib/polyclip.h, polygon.h, point2D.h), which
// the recipe in oracle/Makefile compiles where those headers lie.
//
// TEST INFRASTRUCTURE ONLY (checker and CPU baseline, never shipped).
//
// The reference's CPU route cannot run as shipped: step 1 of the labelling
// (initial classification) and step 3 (copy labels to Q) are commented out
// (polyclip.h:343-397, :461-465) because the GPU supplies them.  This driver
// restores both in its own code, written against the reference's types, and
// then calls the reference's labelIntersections (steps 2, 3.5-6),
// createResult and cleanUpResult unchanged.
#include <iostream>
#include <sstream>
#include <streambuf>
#include "polyclip.h"

namespace {
class NullBuf : public std::streambuf {
 protected:
  int overflow(int c) override { return c; }
};

// polyclip.h:343-397 (restored step 1)
void restored_step1() {
  for (polygon &P : pPolygon)
    for (vertex *I : P.vertices(INTERSECTION)) {
      vertex *P_m = I->prev;
      vertex *P_p = I->next;
      vertex *Q_m = I->neighbour->prev;
      vertex *Q_p = I->neighbour->next;
      RelativePositionType tm = oracle(Q_m, P_m, I, P_p);
      RelativePositionType tp = oracle(Q_p, P_m, I, P_p);
      if ((tm == LEFT && tp == RIGHT) || (tm == RIGHT && tp == LEFT)) I->label = CROSSING;
      if ((tm == LEFT && tp == LEFT) || (tm == RIGHT && tp == RIGHT)) I->label = BOUNCING;
      if ((tp == IS_P_p && tm == RIGHT) || (tm == IS_P_p && tp == RIGHT)) I->label = LEFT_ON;
      if ((tp == IS_P_p && tm == LEFT) || (tm == IS_P_p && tp == LEFT)) I->label = RIGHT_ON;
      if ((tp == IS_P_p && tm == IS_P_m) || (tm == IS_P_p && tp == IS_P_m)) I->label = ON_ON;
      if ((tm == IS_P_m && tp == RIGHT) || (tp == IS_P_m && tm == RIGHT)) I->label = ON_LEFT;
      if ((tm == IS_P_m && tp == LEFT) || (tp == IS_P_m && tm == LEFT)) I->label = ON_RIGHT;
    }
}

// polyclip.h:461-465 (restored step 3)
void restored_step3() {
  for (polygon &P : pPolygon)
    for (vertex *I : P.vertices(INTERSECTION)) I->neighbour->label = I->label;
}
}  // namespace

// Complex polygons: P has nPR components of pSizes[r] vertices each (packed
// in P), Q likewise; union_op selects the union (UNION, polyclip.h:42).
extern "C" int ref_clip_multi(const double *P, const int *pSizes, int nPR, const double *Q,
                              const int *qSizes, int nQR, int union_op, double *out_xy,
                              int out_cap, int *comp_sizes, int comp_cap, int *ncomp,
                              int *nverts) {
  NullBuf nb;
  std::streambuf *old = std::cout.rdbuf(&nb);
  pPolygon.clear();
  qPolygon.clear();
  resultPolygon.clear();
  UNION = union_op != 0;
  const double *src[2] = {P, Q};
  const int *sizes[2] = {pSizes, qSizes};
  const int nr[2] = {nPR, nQR};
  for (int side = 0; side < 2; side++) {
    int off = 0;
    for (int r = 0; r < nr[side]; r++) {
      polygon p;
      for (int i = 0; i < sizes[side][r]; i++) {
        point2D v(src[side][2 * (off + i)], src[side][2 * (off + i) + 1]);
        p.newVertex(v, true);
      }
      off += sizes[side][r];
      (side == 0 ? pPolygon : qPolygon).push_back(p);
    }
  }
  computeIntersections();
  restored_step1();
  restored_step3();
  labelIntersections();
  createResult();
  cleanUpResult();
  int c = 0, w = 0, rc = 0;
  for (polygon &R : resultPolygon) {
    if (R.root == NULL) continue;
    int s = 0;
    for (vertex *V : R.vertices(ALL)) {
      if (w < out_cap) {
        out_xy[2 * w] = V->p.x;
        out_xy[2 * w + 1] = V->p.y;
      } else {
        rc = -1;
      }
      w++;
      s++;
    }
    if (c < comp_cap) comp_sizes[c] = s;
    else rc = -1;
    c++;
  }
  *ncomp = c;
  *nverts = w;
  resultPolygon.clear();
  pPolygon.clear();
  qPolygon.clear();
  UNION = false;
  std::cout.rdbuf(old);
  return rc;
}

extern "C" int ref_clip(const double *P, int nP, const double *Q, int nQ,
                        double *out_xy, int out_cap, int *comp_sizes,
                        int comp_cap, int *ncomp, int *nverts) {
  return ref_clip_multi(P, &nP, 1, Q, &nQ, 1, 0, out_xy, out_cap, comp_sizes, comp_cap, ncomp,
                        nverts);
}
