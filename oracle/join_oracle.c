/*
 * join_oracle.c -- CPU restatement of GCMF's filter-and-refine spatial join.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product library never links or calls it.
 *
 * Layer layout (Appendix of the reference reader, SEQ_Overlay.h:285-389):
 * coords AoS (x,y) with the ring's closing vertex repeated, VNum[i] = ring
 * length including the closing vertex, VPSNum[i] inclusive prefix of VNum,
 * MBR u64 keys (x1,y1,x2,y2), MBR2 the same as doubles.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

/* Planted mistakes for the tests that show the oracle's pins are not
 * vacuous (each pin must fail under its mutant).  0, the default, is the
 * oracle itself; nothing but those tests sets it. */
int orc_mutant = 0;

/* CoordToMBR(type 1), SEQ_Overlay.h:69-97: keep at most 12 fractional
 * digits of the decimal token, drop '.', zero-pad, atoll, add 9e15. */
uint64_t orc_coord_to_mbr(const char *ct) {
  char buff[64];
  int i = 0, fCount = 0, fDot = 0;
  const int fracNum = 12;
  while (*ct && i < 48) {
    if (*ct != '.') {
      if (fCount < fracNum) {
        buff[i++] = *ct;
        if (fDot) fCount++;
      } else
        break;
    } else
      fDot = 1;
    ct++;
  }
  for (; fCount < fracNum; fCount++) buff[i++] = '0';
  buff[i] = '\0';
  uint64_t r = (uint64_t)atoll(buff);
  r += 9000000000000000ULL;
  return r;
}

/* The reference's LSD sort looks at 16 decimal digits (MAX_DIGITS,
 * Constants.h:3; GPU_Utility.h:242-289), i.e. it sorts by key mod 1e16,
 * stably, starting from the identity permutation. */
#define KEY_MOD 10000000000000000ULL

typedef struct { uint64_t k; int i; } kv;
static int cmp_kv(const void *a, const void *b) {
  const kv *x = (const kv *)a, *y = (const kv *)b;
  if (x->k != y->k) return x->k < y->k ? -1 : 1;
  return (x->i > y->i) - (x->i < y->i);
}

void orc_endpoint_sort(long long B, long long O, const uint64_t *bMBR,
                       const uint64_t *oMBR, int *perm, int *rank) {
  long long N = B + O, n = 2 * N;
  kv *a = (kv *)malloc(sizeof(kv) * (size_t)(n ? n : 1));
  /* InitializeData, GPU_RadixSort.h:4-22 */
  for (long long e = 0; e < N; e++) {
    const uint64_t *m = e < B ? bMBR + 4 * e : oMBR + 4 * (e - B);
    a[2 * e].k = m[0] % KEY_MOD;     a[2 * e].i = (int)(2 * e);
    a[2 * e + 1].k = m[2] % KEY_MOD; a[2 * e + 1].i = (int)(2 * e + 1);
  }
  qsort(a, (size_t)n, sizeof(kv), cmp_kv);
  for (long long r = 0; r < n; r++) {
    perm[r] = a[r].i;
    rank[a[r].i] = (int)r;
  }
  free(a);
}

/* CountSortBaseMBROverlapLoad / SortBaseMBROverlapLoadCalculated,
 * GPU_MBR.h:339-539.  Element e's partners are the opposite-layer left
 * endpoints ranked strictly between e's own two endpoints, with closed
 * overlap on y, emitted in rank order. */
long long orc_mbr_sweep(long long B, long long O, const uint64_t *bMBR,
                        const uint64_t *oMBR, int *counts, int *pairs) {
  long long N = B + O, n = 2 * N;
  int *perm = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
  int *rank = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
  uint64_t *Y = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
  orc_endpoint_sort(B, O, bMBR, oMBR, perm, rank);
  for (long long e = 0; e < N; e++) {
    const uint64_t *m = e < B ? bMBR + 4 * e : oMBR + 4 * (e - B);
    Y[2 * e] = m[1];
    Y[2 * e + 1] = m[3];
  }
  long long K = 0;
  for (long long e = 0; e < N; e++) {
    int lo = rank[2 * e], hi = rank[2 * e + 1];
    int cnt = 0;
    for (int i = lo + 1; i < hi; i++) {
      int m = perm[i];
      if (e >= B) {
        if (m >= 2 * B) continue;
      } else {
        if (m < 2 * B) continue;
      }
      if (m % 2 == 1) continue;
      if (Y[2 * e] > Y[m + 1]) continue;
      if (Y[m] > Y[2 * e + 1]) continue;
      if (pairs) {
        if (e < B) {
          pairs[2 * (K + cnt)] = (int)e;
          pairs[2 * (K + cnt) + 1] = m / 2 - (int)B;
        } else {
          pairs[2 * (K + cnt)] = m / 2;
          pairs[2 * (K + cnt) + 1] = (int)(e - B);
        }
      }
      cnt++;
    }
    if (counts) counts[e] = cnt;
    K += cnt;
  }
  free(perm); free(rank); free(Y);
  return K;
}

/* Get_CMBR, GPU_Join.h:461-508 */
void orc_get_cmbr(long long K, const int *pairs, const double *bMBR2,
                  const double *oMBR2, double *cmbr, int *pipFlag,
                  signed char *pipType, int *joinFlag) {
  for (long long k = 0; k < K; k++) {
    int b = pairs[2 * k], o = pairs[2 * k + 1];
    double x1 = bMBR2[4 * b], x2 = bMBR2[4 * b + 2], y1 = bMBR2[4 * b + 1], y2 = bMBR2[4 * b + 3];
    double a1 = oMBR2[4 * o], a2 = oMBR2[4 * o + 2], b1 = oMBR2[4 * o + 1], b2 = oMBR2[4 * o + 3];
    double cX1 = a1 > x1 ? a1 : x1;
    double cX2 = a2 < x2 ? a2 : x2;
    double cY1 = b1 > y1 ? b1 : y1;
    double cY2 = b2 < y2 ? b2 : y2;
    cmbr[4 * k] = cX1; cmbr[4 * k + 1] = cY1; cmbr[4 * k + 2] = cX2; cmbr[4 * k + 3] = cY2;
    joinFlag[k] = 1;
    if (x1 == cX1 && x2 == cX2 && y1 == cY1 && y2 == cY2) {
      pipFlag[k] = 1; pipType[k] = 1;
    } else if (a1 == cX1 && a2 == cX2 && b1 == cY1 && b2 == cY2) {
      pipFlag[k] = 1; pipType[k] = 2;
    } else {
      pipFlag[k] = 0; pipType[k] = 0;
    }
  }
}

/* Point_In_Polygon_Test, GPU_Join.h:839-943: crossing parity of the inner
 * polygon's first vertex over the other ring (VNum vertices, closing vertex
 * included), predicate pX < (X2-X1)*(pY-Y1)/(Y2-Y1)+X1 (:912). */
long long orc_pip(long long K, const int *pairs, const signed char *pipType,
                  const double *bCoords, const long long *bVPSNum,
                  const double *oCoords, const long long *oVPSNum,
                  int *pipFlag, int *joinFlag) {
  long long within = 0;
  for (long long k = 0; k < K; k++) {
    if (pipType[k] == 0) continue;
    int b = pairs[2 * k], o = pairs[2 * k + 1];
    long long bS = b > 0 ? bVPSNum[b - 1] : 0, oS = o > 0 ? oVPSNum[o - 1] : 0;
    double pX, pY;
    const double *poly;
    long long vNum;
    if (pipType[k] == 1) {
      pX = bCoords[2 * bS]; pY = bCoords[2 * bS + 1];
      vNum = oVPSNum[o] - oS; poly = oCoords + 2 * oS;
    } else {
      pX = oCoords[2 * oS]; pY = oCoords[2 * oS + 1];
      vNum = bVPSNum[b] - bS; poly = bCoords + 2 * bS;
    }
    int c = 0;
    for (long long i = 0, j = vNum - 1; i < vNum; j = i++) {
      double Y1 = poly[2 * i + 1], Y2 = poly[2 * j + 1], X1 = poly[2 * i], X2 = poly[2 * j];
      if (orc_mutant == ORC_MUT_PIP_BOTH_ENDS) {
        /* counts an edge with an endpoint on the ray from both sides */
        if ((Y1 > pY && Y2 > pY) || (Y1 < pY && Y2 < pY) || Y1 == Y2) continue;
      } else if (orc_mutant == ORC_MUT_PIP_CLOSED_Y) {
        if ((Y1 >= pY) == (Y2 >= pY)) continue;
      } else if ((Y1 > pY) == (Y2 > pY)) {
        continue;
      }
      double xc = (X2 - X1) * (pY - Y1) / (Y2 - Y1) + X1;
      if (orc_mutant == ORC_MUT_PIP_LE ? pX <= xc : pX < xc) c = !c;
    }
    pipFlag[k] = c;
    if (c == 1) { joinFlag[k] = 0; within++; }
  }
  return within;
}

/* Count_CMF / Apply_CMF, GPU_Join.h:511-735: edge i joins v[i] and v[i+1]
 * (the last joins v[VNum-1] and v[0], the degenerate closing edge). */
static int cmf_edge(const double *c, int vNum, int i, const double *cm) {
  double x1 = c[2 * i], y1 = c[2 * i + 1], x2, y2;
  if (i < vNum - 1) { x2 = c[2 * (i + 1)]; y2 = c[2 * (i + 1) + 1]; }
  else { x2 = c[0]; y2 = c[1]; }
  double cX1 = cm[0], cY1 = cm[1], cX2 = cm[2], cY2 = cm[3];
  if (orc_mutant == ORC_MUT_CMF_NO_CLOSING && i == vNum - 1) return 0;
  if (orc_mutant == ORC_MUT_CMF_OPEN) {
    if ((cX1 >= x1 && cX1 >= x2) || (cX2 <= x1 && cX2 <= x2)) return 0;
    if ((cY1 >= y1 && cY1 >= y2) || (cY2 <= y1 && cY2 <= y2)) return 0;
    return 1;
  }
  if ((cX1 > x1 && cX1 > x2) || (cX2 < x1 && cX2 < x2)) return 0;
  if ((cY1 > y1 && cY1 > y2) || (cY2 < y1 && cY2 < y2)) return 0;
  return 1;
}

long long orc_cmf(long long K, const int *pairs, const double *cmbr,
                  const double *bCoords, const int *bVNum, const long long *bVPSNum,
                  const double *oCoords, const int *oVNum, const long long *oVPSNum,
                  int *joinFlag, int *bCount, int *oCount, short *bEdge, short *oEdge) {
  long long E = 0, bw = 0, ow = 0;
  for (long long k = 0; k < K; k++) {
    bCount[k] = 0; oCount[k] = 0;
    if (joinFlag[k] != 1) continue;
    int b = pairs[2 * k], o = pairs[2 * k + 1];
    const double *bc = bCoords + 2 * (b > 0 ? bVPSNum[b - 1] : 0);
    const double *oc = oCoords + 2 * (o > 0 ? oVPSNum[o - 1] : 0);
    int nb = 0, no = 0;
    for (int i = 0; i < bVNum[b]; i++) nb += cmf_edge(bc, bVNum[b], i, cmbr + 4 * k);
    for (int i = 0; i < oVNum[o]; i++) no += cmf_edge(oc, oVNum[o], i, cmbr + 4 * k);
    if (nb == 0 || no == 0) { joinFlag[k] = 0; continue; }
    bCount[k] = nb; oCount[k] = no;
    E++;
    if (bEdge) {
      for (int i = 0; i < bVNum[b]; i++) if (cmf_edge(bc, bVNum[b], i, cmbr + 4 * k)) bEdge[bw++] = (short)i;
      for (int i = 0; i < oVNum[o]; i++) if (cmf_edge(oc, oVNum[o], i, cmbr + 4 * k)) oEdge[ow++] = (short)i;
    }
  }
  return E;
}

/* one edge-pair test of Segment_Intersect_Join, GPU_Join.h:376-445;
 * returns 1 when the pair sets the join flag */
static int seg_hit(double bX1, double bY1, double bX2, double bY2,
                   double oX1, double oY1, double oX2, double oY2) {
  if (((bX2 < oX1) && (bX2 < oX2)) && ((bX1 < oX1) && (bX1 < oX2))) return 0;
  if (((bX2 > oX1) && (bX2 > oX2)) && ((bX1 > oX1) && (bX1 > oX2))) return 0;
  if (((bY2 < oY1) && (bY2 < oY2)) && ((bY1 < oY1) && (bY1 < oY2))) return 0;
  if (((bY2 > oY1) && (bY2 > oY2)) && ((bY1 > oY1) && (bY1 > oY2))) return 0;
  double denom = ((bX1 - bX2) * (oY1 - oY2)) - ((bY1 - bY2) * (oX1 - oX2));
  if (denom == 0) {
    if (orc_mutant == ORC_MUT_SEG_NO_PARALLEL) return 0;
    if (orc_mutant != ORC_MUT_SEG_NO_VERTICAL &&
        ((oX1 == oX2 && oX1 == bX1) || (bX1 == bX2 && bX1 == oX1)))
      return 1;
    double oM = (oY2 - oY1) / (oX2 - oX1);
    double bM = (bY2 - bY1) / (bX2 - bX1);
    double oB = oY1 - oM * oX1;
    double bB = bY1 - bM * bX1;
    return oB == bB;
  }
  double nomX = (((bX1 * bY2) - (bY1 * bX2)) * (oX1 - oX2)) - ((bX1 - bX2) * ((oX1 * oY2) - (oY1 * oX2)));
  double nomY = ((((bX1 * bY2) - (bY1 * bX2)) * (oY1 - oY2)) - (((bY1 - bY2) * ((oX1 * oY2) - (oY1 * oX2)))));
  double x0 = nomX / denom, y0 = nomY / denom;
  if (orc_mutant == ORC_MUT_SEG_OPEN &&
      ((x0 == bX1 && y0 == bY1) || (x0 == bX2 && y0 == bY2) || (x0 == oX1 && y0 == oY1) ||
       (x0 == oX2 && y0 == oY2)))
    return 0; /* touching at an endpoint does not count */
  if ((x0 > bX1 && x0 > bX2) || (x0 < bX1 && x0 < bX2)) return 0;
  if ((x0 > oX1 && x0 > oX2) || (x0 < oX1 && x0 < oX2)) return 0;
  if ((y0 > bY1 && y0 > bY2) || (y0 < bY1 && y0 < bY2)) return 0;
  if ((y0 > oY1 && y0 > oY2) || (y0 < oY1 && y0 < oY2)) return 0;
  return 1;
}

long long orc_segment_join(long long K, const int *pairs, const int *joinFlag,
                           const double *bCoords, const long long *bVPSNum,
                           const double *oCoords, const long long *oVPSNum,
                           const int *bCount, const int *oCount,
                           const short *bEdge, const short *oEdge, int *segFlag) {
  long long bw = 0, ow = 0, hits = 0;
  for (long long k = 0; k < K; k++) {
    segFlag[k] = 0;
    if (joinFlag[k] != 1) continue;
    int b = pairs[2 * k], o = pairs[2 * k + 1];
    long long bBase = b > 0 ? bVPSNum[b - 1] : 0, oBase = o > 0 ? oVPSNum[o - 1] : 0;
    int bV = (int)(bVPSNum[b] - bBase), oV = (int)(oVPSNum[o] - oBase);
    int bE = bCount[k], oE = oCount[k];
    int flag = 0;
    for (int i = 0; i < bE && !flag; i++) {
      int bi = bEdge[bw + i];
      double bX1 = bCoords[2 * (bBase + bi)], bY1 = bCoords[2 * (bBase + bi) + 1], bX2, bY2;
      if (bi < bV - 1) { bX2 = bCoords[2 * (bBase + bi + 1)]; bY2 = bCoords[2 * (bBase + bi + 1) + 1]; }
      else { bX2 = bCoords[2 * bBase]; bY2 = bCoords[2 * bBase + 1]; }
      if (bX1 == bX2 && bY1 == bY2) continue;
      for (int j = 0; j < oE && !flag; j++) {
        int oi = oEdge[ow + j];
        double oX1 = oCoords[2 * (oBase + oi)], oY1 = oCoords[2 * (oBase + oi) + 1], oX2, oY2;
        if (oi < oV - 1) { oX2 = oCoords[2 * (oBase + oi + 1)]; oY2 = oCoords[2 * (oBase + oi + 1) + 1]; }
        else { oX2 = oCoords[2 * oBase]; oY2 = oCoords[2 * oBase + 1]; }
        flag = seg_hit(bX1, bY1, bX2, bY2, oX1, oY1, oX2, oY2);
      }
    }
    bw += bE; ow += oE;
    segFlag[k] = flag;
    hits += flag;
  }
  return hits;
}
