/*
 * gh_oracle.c -- CPU restatement of GCMF-GH's Greiner-Hormann clip
 * (Foster-Hormann-Popa variant with degeneracy handling).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker / CPU baseline.  The product
 * library never links or calls it.
 *
 * Two routes are restated:
 *   - the CPU route, GH-CUDA/lib/polyclip.h:60-807 (computeIntersections,
 *     labelIntersections with the commented step 1 (:343-397) and step 3
 *     (:461-465) restored, createResult, cleanUpResult) over doubly linked
 *     vertex lists (GH-CUDA/lib/polygon.h);
 *   - the reference GPU phase 1, GH-CUDA/gppolyclip_mDIV512.cu:95-1715, as
 *     sequential loops producing the same arrays (counts, exclusive scans,
 *     neighbour map, alpha-sorted intersection arrays, initial labels).
 *
 * Deliberate, documented deviations (identical in the HIP implementation):
 *   - per-pair normalisation (normalize=1): every predicate runs on
 *     ((x-cx)*s, (y-cy)*s) with (cx,cy) the CMBR minimum corner and
 *     s = 1/max(CMBR width, height); output vertices are produced in the
 *     original coordinates from the same alpha.  The reference's epsilons are
 *     absolute (GH-CUDA/lib/constants.h:4), which breaks at small scales.
 *   - arrays the reference leaves uninitialised are initialised
 *     (neighbour map -100, alpha -100, neighbour links 0) and never-found
 *     neighbour searches do not write through a stale index.
 *   - double->int conversion of 1e7*alpha saturates (the CUDA conversion
 *     the reference relies on saturates; C's cast is undefined there).
 */
#include "oracle.h"
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GH_EPS 0.0000001        /* GH-CUDA/lib/constants.h:4 */
#define ALPHA_SCALE 10000000    /* (int)pow(10,EPSILON_POSITIONS), :1034 */
#define ALPHA_DIGITS 7          /* EPSILON_POSITIONS, constants.h:3 */

typedef struct { double x, y; } pt;

/* twice the signed area of [P,Q,R], point2D.h:78-80 */
static double triA(pt P, pt Q, pt R) {
  return (Q.x - P.x) * (R.y - P.y) - (Q.y - P.y) * (R.x - P.x);
}

/* saturating trunc toward zero (cvt.rzi.s32.f64 semantics) */
static int trunc_sat(double v) {
  if (v != v) return 0;
  if (v >= 2147483647.0) return INT_MAX;
  if (v <= -2147483648.0) return INT_MIN;
  return (int)v;
}

/* ----------------------------------------------------------------------
 * intersection classification, polyclip.h:60-179 == gppolyclip:95-169
 * (note the reference tests |alpha| when classifying beta in the collinear
 * branch, :156 / :157; kept as is)
 * -------------------------------------------------------------------- */
static int itype(pt P1, pt P2, pt Q1, pt Q2, double *alpha, double *beta) {
  double AP1 = triA(P1, Q1, Q2);
  double AP2 = triA(P2, Q1, Q2);
  if (fabs(AP1 - AP2) > GH_EPS) {
    double AQ1 = triA(Q1, P1, P2);
    double AQ2 = triA(Q2, P1, P2);
    double a = AP1 / (AP1 - AP2);
    double b = AQ1 / (AQ1 - AQ2);
    *alpha = a;
    *beta = b;
    int a0 = 0, a01 = 0, b0 = 0, b01 = 0;
    if ((a > GH_EPS) && (a < 1.0 - GH_EPS)) a01 = 1;
    else if (fabs(a) <= GH_EPS) a0 = 1;
    if ((b > GH_EPS) && (b < 1.0 - GH_EPS)) b01 = 1;
    else if (fabs(b) <= GH_EPS) b0 = 1;
    if (a01 && b01) return 1;
    if (a0 && b01) return 2;
    if (b0 && a01) return 3;
    if (a0 && b0) return 4;
  } else if (fabs(AP1) < GH_EPS) {
    pt dP = {P2.x - P1.x, P2.y - P1.y};
    pt dQ = {Q2.x - Q1.x, Q2.y - Q1.y};
    pt PQ = {Q1.x - P1.x, Q1.y - P1.y};
    double a = (PQ.x * dP.x + PQ.y * dP.y) / (dP.x * dP.x + dP.y * dP.y);
    double b = -(PQ.x * dQ.x + PQ.y * dQ.y) / (dQ.x * dQ.x + dQ.y * dQ.y);
    *alpha = a;
    *beta = b;
    int a0 = 0, a01 = 0, an = 0, b0 = 0, b01 = 0, bn = 0;
    if ((a > GH_EPS) && (a < 1.0 - GH_EPS)) a01 = 1;
    else if (fabs(a) <= GH_EPS) a0 = 1;
    else an = 1;
    if ((b > GH_EPS) && (b < 1.0 - GH_EPS)) b01 = 1;
    else if (fabs(a) <= GH_EPS) b0 = 1;
    else bn = 1;
    if (a01 && b01) return 5;
    if (an && b01) return 6;
    if (bn && a01) return 7;
    if (a0 && b0) return 8;
  }
  return 0;
}

int orc_intersect_type(const double *P1, const double *P2, const double *Q1,
                       const double *Q2, double *alpha, double *beta) {
  pt a = {P1[0], P1[1]}, b = {P2[0], P2[1]}, c = {Q1[0], Q1[1]}, d = {Q2[0], Q2[1]};
  double al = 0, be = 0;
  int t = itype(a, b, c, d, &al, &be);
  if (alpha) *alpha = al;
  if (beta) *beta = be;
  return t;
}

/* edge-bbox overlap, gppolyclip:345-371 */
static int lsmf(pt P1, pt P2, pt Q1, pt Q2) {
  double minPX = P1.x, minPY = P1.y, maxPX = P2.x, maxPY = P2.y;
  double minQX = Q1.x, minQY = Q1.y, maxQX = Q2.x, maxQY = Q2.y;
  if (minPX > P2.x) { minPX = P2.x; maxPX = P1.x; }
  if (minPY > P2.y) { minPY = P2.y; maxPY = P1.y; }
  if (minQX > Q2.x) { minQX = Q2.x; maxQX = Q1.x; }
  if (minQY > Q2.y) { minQY = Q2.y; maxQY = Q1.y; }
  if (minPX > maxQX || maxPX < minQX) return 0;
  if (minPY > maxQY || maxPY < minQY) return 0;
  return 1;
}

/* ----------------------------------------------------------------------
 * per-pair CMBR and normalisation (ghcuda.cpp:94-125, readShapefile.h:304)
 * -------------------------------------------------------------------- */
static void ring_mbr(const double *c, int n, double *m) {
  double minX = c[0], minY = c[1], maxX = c[0], maxY = c[1];
  for (int i = 0; i < n; i++) {
    double x = c[2 * i], y = c[2 * i + 1];
    if (maxX < x) maxX = x;
    if (minX > x) minX = x;
    if (maxY < y) maxY = y;
    if (minY > y) minY = y;
  }
  m[0] = minX; m[1] = minY; m[2] = maxX; m[3] = maxY;
}

/* the power of two f >= sqrt(nP nQ) / 1024 (1 for rings of up to 1024
 * vertices): the extra scale of the edge-scaled normalisation */
static double edge_scale(int nP, int nQ) {
  double r = sqrt((double)nP * (double)nQ) / 1024.0, f = 1.0;
  while (f < r) f *= 2.0;
  return f;
}

/* returns 0 when the MBRs are disjoint (the reference exits the process);
 * normalize & 1: per-pair CMBR normalisation, & 4: edge-scaled */
static int pair_frame(const double *P, int nP, const double *Q, int nQ,
                      int normalize, double *cmbr, double *norm) {
  double a[4], b[4];
  ring_mbr(P, nP, a);
  ring_mbr(Q, nQ, b);
  if (a[0] > b[2] || a[2] < b[0]) return 0;
  if (a[1] > b[3] || a[3] < b[1]) return 0;
  cmbr[0] = a[0] < b[0] ? b[0] : a[0];
  cmbr[1] = a[1] < b[1] ? b[1] : a[1];
  cmbr[2] = b[2] < a[2] ? b[2] : a[2];
  cmbr[3] = b[3] < a[3] ? b[3] : a[3];
  if (normalize & 1) {
    double w = cmbr[2] - cmbr[0], h = cmbr[3] - cmbr[1];
    double e = w > h ? w : h;
    norm[0] = cmbr[0];
    norm[1] = cmbr[1];
    norm[2] = e > 0.0 ? 1.0 / e : 1.0;
    if (normalize & 4) norm[2] *= edge_scale(nP, nQ);
  } else {
    norm[0] = 0.0; norm[1] = 0.0; norm[2] = 1.0;
  }
  return 1;
}

static pt tf(double x, double y, const double *norm) {
  pt r = {(x - norm[0]) * norm[2], (y - norm[1]) * norm[2]};
  return r;
}

/* ----------------------------------------------------------------------
 * linked vertex lists, polygon.h:40-138,377-544
 * -------------------------------------------------------------------- */
enum { L_NONE, L_CROSSING, L_BOUNCING, L_LEFT_ON, L_RIGHT_ON, L_ON_ON,
       L_ON_LEFT, L_ON_RIGHT, L_DELAYED_CROSSING, L_DELAYED_BOUNCING };
enum { EXIT, ENTRY, NEITHER };
enum { R_LEFT, R_RIGHT, R_IS_P_m, R_IS_P_p };

typedef struct {
  pt p, o;              /* predicate and original coordinates */
  int prev, next, nb;
  unsigned char source, inter, split, cross;
  int label, enex;
  double alpha;
} vtx;

typedef struct { vtx *v; int n, cap; } vpool;

static int vnew(vpool *pl, pt p, pt o, double alpha) {
  if (pl->n == pl->cap) {
    pl->cap = pl->cap ? pl->cap * 2 : 1024;
    pl->v = (vtx *)realloc(pl->v, sizeof(vtx) * (size_t)pl->cap);
  }
  vtx *v = &pl->v[pl->n];
  memset(v, 0, sizeof(*v));
  v->p = p; v->o = o;
  v->prev = v->next = v->nb = -1;
  v->alpha = alpha;
  v->inter = alpha > -1.0;
  v->label = L_NONE;
  v->enex = NEITHER;
  return pl->n++;
}

/* polygon::newVertex, polygon.h:388-407 */
static void poly_append(vpool *pl, int *root, int V) {
  vtx *v = pl->v;
  if (*root < 0) {
    v[V].next = v[V].prev = V;
    *root = V;
  } else {
    int r = *root;
    v[V].prev = v[r].prev;
    v[V].next = r;
    v[v[r].prev].next = V;
    v[r].prev = V;
  }
}

/* insertVertex, polygon.h:101-113 */
static void insert_vertex(vpool *pl, int V, int curr, double alpha) {
  vtx *v = pl->v;
  if (alpha > -1.0) {
    do { curr = v[curr].next; } while (!v[curr].source && v[curr].alpha < alpha);
  } else {
    curr = v[curr].next;
  }
  v[v[curr].prev].next = V;
  v[V].prev = v[curr].prev;
  v[V].next = curr;
  v[curr].prev = V;
}

/* link, polygon.h:89-94 */
static void vlink(vpool *pl, int P, int Q) {
  pl->v[P].nb = Q; pl->v[Q].nb = P;
  pl->v[P].inter = 1; pl->v[Q].inter = 1;
}

/* polygon::removeVertex, polygon.h:412-421 (vertex is not freed) */
static void remove_vertex(vpool *pl, int *root, int V) {
  vtx *v = pl->v;
  if (*root == V) {
    *root = v[V].next;
    if (v[*root].next == *root) *root = -1;
  }
  v[v[V].prev].next = v[V].next;
  v[v[V].next].prev = v[V].prev;
}

/* polygon::pointInPoly, polygon.h:427-448 (edges(ALL)) */
static int point_in_poly(const vpool *pl, int root, pt R) {
  const vtx *v = pl->v;
  int w = 0;
  if (root < 0) return 0;
  int a = root;
  do {
    int b = v[a].next;
    pt P0 = v[a].p, P1 = v[b].p;
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

static void toggle(int *s) {
  if (*s == ENTRY) { *s = EXIT; return; }
  if (*s == EXIT) { *s = ENTRY; return; }
}

/* oracle(), polyclip.h:297-326 */
static int rel_oracle(const vpool *pl, int Q, int P1, int P2, int P3) {
  const vtx *v = pl->v;
  if (v[P1].inter && v[P1].nb == Q) return R_IS_P_m;
  if (v[P3].inter && v[P3].nb == Q) return R_IS_P_p;
  double s1 = triA(v[Q].p, v[P1].p, v[P2].p);
  double s2 = triA(v[Q].p, v[P2].p, v[P3].p);
  double s3 = triA(v[P1].p, v[P2].p, v[P3].p);
  if (s3 > 0) {
    if (s1 > 0 && s2 > 0) return R_LEFT;
    return R_RIGHT;
  }
  if (s1 < 0 && s2 < 0) return R_RIGHT;
  return R_LEFT;
}

typedef struct {
  vpool pl;
  int rootP, rootQ;       /* single-ring callers: the one component of each */
  int *rP, *rQ, nRP, nRQ; /* the components (rings) of P and Q */
  int unionOp;            /* UNION instead of intersection (polyclip.h:42) */
  int *res_roots; int nres, capres;
} ghstate;

/* a single ring on each side: its root is the only component */
static void single_roots(ghstate *g) {
  g->rP = &g->rootP; g->nRP = 1;
  g->rQ = &g->rootQ; g->nRQ = 1;
}

static void res_push(ghstate *g, int root) {
  if (g->nres == g->capres) {
    g->capres = g->capres ? 2 * g->capres : 16;
    g->res_roots = (int *)realloc(g->res_roots, sizeof(int) * (size_t)g->capres);
  }
  g->res_roots[g->nres++] = root;
}

/* step 1 (restored, polyclip.h:343-397): initial classification of P's
 * intersection vertices, then step 3 (polyclip.h:461-465): copy to Q */
static void label_step1_and_3(ghstate *g) {
  vtx *v = g->pl.v;
  for (int c = 0; c < g->nRP; c++) {
  int root = g->rP[c], I = root;
  do {
    if (v[I].inter && v[I].nb >= 0) {
      int P_m = v[I].prev, P_p = v[I].next;
      int Q_m = v[v[I].nb].prev, Q_p = v[v[I].nb].next;
      int tm = rel_oracle(&g->pl, Q_m, P_m, I, P_p);
      int tp = rel_oracle(&g->pl, Q_p, P_m, I, P_p);
      if ((tm == R_LEFT && tp == R_RIGHT) || (tm == R_RIGHT && tp == R_LEFT)) v[I].label = L_CROSSING;
      if ((tm == R_LEFT && tp == R_LEFT) || (tm == R_RIGHT && tp == R_RIGHT)) v[I].label = L_BOUNCING;
      if ((tp == R_IS_P_p && tm == R_RIGHT) || (tm == R_IS_P_p && tp == R_RIGHT)) v[I].label = L_LEFT_ON;
      if ((tp == R_IS_P_p && tm == R_LEFT) || (tm == R_IS_P_p && tp == R_LEFT)) v[I].label = L_RIGHT_ON;
      if ((tp == R_IS_P_p && tm == R_IS_P_m) || (tm == R_IS_P_p && tp == R_IS_P_m)) v[I].label = L_ON_ON;
      if ((tm == R_IS_P_m && tp == R_RIGHT) || (tp == R_IS_P_m && tm == R_RIGHT)) v[I].label = L_ON_LEFT;
      if ((tm == R_IS_P_m && tp == R_LEFT) || (tp == R_IS_P_m && tm == R_LEFT)) v[I].label = L_ON_RIGHT;
    }
    I = v[I].next;
  } while (I != root);
  }
  for (int c = 0; c < g->nRP; c++) {
    int root = g->rP[c], I = root;
    do {
      if (v[I].inter && v[I].nb >= 0) v[v[I].nb].label = v[I].label;
      I = v[I].next;
    } while (I != root);
  }
}

/* polygon::noCrossingVertex, polygon.h:465-475 (bounces count as
 * crossings for the union) */
static int no_crossing_vertex(const vpool *pl, int root, int union_case) {
  const vtx *v = pl->v;
  int V = root;
  do {
    if (v[V].inter) {
      if (v[V].label == L_CROSSING || v[V].label == L_DELAYED_CROSSING) return 0;
      if (union_case && (v[V].label == L_BOUNCING || v[V].label == L_DELAYED_BOUNCING)) return 0;
    }
    V = v[V].next;
  } while (V != root);
  return 1;
}

/* even-odd containment in a multi-component polygon: toggled once per
 * component that contains R (polyclip.h:505-507, 563-565) */
static int point_in_comps(const vpool *pl, const int *roots, int n, pt R) {
  int in = 0;
  for (int c = 0; c < n; c++)
    if (point_in_poly(pl, roots[c], R)) in = !in;
  return in;
}

/* is vertex V on the ring rooted at root? */
static int on_ring(const vpool *pl, int root, int V) {
  int W = root;
  do {
    if (W == V) return 1;
    W = pl->v[W].next;
  } while (W != root);
  return 0;
}

static int all_on_on(const vpool *pl, int root) {
  const vtx *v = pl->v;
  int V = root;
  do {
    if (v[V].label != L_ON_ON) return 0;
    V = v[V].next;
  } while (V != root);
  return 1;
}

/* polygon::getNonIntersectionPoint, polygon.h:480-491 */
static int non_intersection_point(const vpool *pl, int root, pt *out) {
  const vtx *v = pl->v;
  int V = root;
  do {
    if (!v[V].inter) { *out = v[V].p; return 1; }
    V = v[V].next;
  } while (V != root);
  V = root;
  do {
    int n = v[V].next, nb = v[V].nb, nnb = v[n].nb;
    if (nb >= 0 && nnb != v[nb].prev && nnb != v[nb].next) {
      pt s = {v[V].p.x + v[n].p.x, v[V].p.y + v[n].p.y};
      out->x = 0.5 * s.x; out->y = 0.5 * s.y;
      return 1;
    }
    V = n;
  } while (V != root);
  return 0;
}

/* polygon::getNonIntersectionVertex, polygon.h:496-512 */
static int non_intersection_vertex(vpool *pl, int root) {
  int V = root;
  do {
    if (!pl->v[V].inter) return V;
    V = pl->v[V].next;
  } while (V != root);
  V = root;
  do {
    vtx *v = pl->v;
    int n = v[V].next, nb = v[V].nb, nnb = v[n].nb;
    if (nb >= 0 && nnb != v[nb].prev && nnb != v[nb].next) {
      pt s = {v[V].p.x + v[n].p.x, v[V].p.y + v[n].p.y};
      pt so = {v[V].o.x + v[n].o.x, v[V].o.y + v[n].o.y};
      pt p = {0.5 * s.x, 0.5 * s.y}, o = {0.5 * so.x, 0.5 * so.y};
      int T = vnew(pl, p, o, -1.0);
      insert_vertex(pl, T, V, -1.0);
      return T;
    }
    V = n;
  } while (V != root);
  return -1;
}

/* steps 2, 3.5, 4, 5, 6 of labelIntersections (polyclip.h:398-735),
 * createResult (:743-783), cleanUpResult (:791-807) */
static int label_trace_cleanup(ghstate *g, int do_cleanup) {
  vpool *pl = &g->pl;
  const int U = g->unionOp ? 1 : 0;
  int root;
  /* 2) intersection chains, over the components of P */
  for (int c = 0; c < g->nRP; c++) {
    root = g->rP[c];
    int it = root;
    do {
      vtx *v = pl->v;
      if (v[it].inter && (v[it].label == L_LEFT_ON || v[it].label == L_RIGHT_ON)) {
        int I = it;
        int x = v[I].label == L_LEFT_ON ? R_LEFT : R_RIGHT;
        int X = I;
        do {
          v[I].label = L_NONE;
          if (v[I].nb >= 0) v[v[I].nb].label = L_NONE;
          I = v[I].next;
        } while (v[I].label == L_ON_ON);
        int y = v[I].label == L_ON_LEFT ? R_LEFT : R_RIGHT;
        int ct = (x != y) ? L_DELAYED_CROSSING : L_DELAYED_BOUNCING;
        v[X].label = ct;
        v[I].label = ct;
        if (v[X].nb >= 0) v[v[X].nb].label = ct;
        if (v[I].nb >= 0) v[v[I].nb].label = ct;
      }
      it = pl->v[it].next;
    } while (it != root);
  }

  /* 3.5) components without crossings (polyclip.h:478-541) */
  const int *roots[2] = {g->rP, g->rQ};
  const int nr[2] = {g->nRP, g->nRQ};
  int *noInter[2], *ident[2];
  int rc = 0;
  int *splitList[2] = {NULL, NULL}, nsplit[2] = {0, 0};
  int *crossList[2] = {NULL, NULL}, ncross[2] = {0, 0};
  for (int i = 0; i < 2; i++) {
    noInter[i] = (int *)calloc((size_t)nr[i] + 1, sizeof(int));
    ident[i] = (int *)calloc((size_t)nr[i] + 1, sizeof(int));
  }
  for (int i = 0; i < 2; i++) {
    for (int c = 0; c < nr[i]; c++) {
      int Pr = roots[i][c];
      if (no_crossing_vertex(pl, Pr, U)) {
        noInter[i][c] = 1;
        if (all_on_on(pl, Pr)) {
          ident[i][c] = 1;
        } else {
          pt p;
          if (!non_intersection_point(pl, Pr, &p)) { rc = -1; goto done; }
          int inside = point_in_comps(pl, roots[1 - i], nr[1 - i], p);
          if (inside ^ U) res_push(g, Pr);
        }
      }
    }
  }
  /* components of P identical to one of Q: added when both or neither are
   * holes (polyclip.h:517-541) */
  for (int c = 0; c < nr[0]; c++) {
    if (!ident[0][c]) continue;
    int Pr = roots[0][c];
    int P_isHole = 0;
    for (int c2 = 0; c2 < nr[0]; c2++)
      if (roots[0][c2] != Pr && point_in_poly(pl, roots[0][c2], pl->v[Pr].p)) P_isHole = !P_isHole;
    for (int q = 0; q < nr[1]; q++) {
      if (!ident[1][q] || !on_ring(pl, roots[1][q], pl->v[Pr].nb)) continue;
      int Qr = roots[1][q], Q_isHole = 0;
      for (int q2 = 0; q2 < nr[1]; q2++)
        if (roots[1][q2] != Qr && point_in_poly(pl, roots[1][q2], pl->v[Qr].p)) Q_isHole = !Q_isHole;
      if (P_isHole == Q_isHole) res_push(g, Pr);
      break;
    }
  }

  /* 4) entry/exit flags (polyclip.h:546-660) */
  for (int i = 0; i < 2; i++) {
    for (int c = 0; c < nr[i]; c++) {
    if (noInter[i][c]) continue;
    int Pr = roots[i][c];
    int V = non_intersection_vertex(pl, Pr);
    if (V < 0) { rc = -1; goto done; }
    int status = ENTRY;
    if (point_in_comps(pl, roots[1 - i], nr[1 - i], pl->v[V].p)) toggle(&status);
    int first_chain_vertex = 1;
    int I = V;
    int started = 0;
    for (;;) {
      if (!started) {
        started = 1;
        if (!pl->v[I].inter) {
          do { I = pl->v[I].next; } while (!pl->v[I].inter && I != V);
          if (I == V) break;
        }
      }
      vtx *v = pl->v;
      if (v[I].label == L_CROSSING) {
        v[I].enex = status;
        toggle(&status);
      }
      if (v[I].label == L_BOUNCING && ((status == EXIT) ^ U)) {
        if (!v[I].split) {
          v[I].split = 1;
          splitList[i] = (int *)realloc(splitList[i], sizeof(int) * (size_t)(nsplit[i] + 1));
          splitList[i][nsplit[i]++] = I;
        }
      }
      if (v[I].label == L_DELAYED_CROSSING) {
        v[I].enex = status;
        if (first_chain_vertex) {
          if ((status == EXIT) ^ U) v[I].label = L_CROSSING;
          first_chain_vertex = 0;
        } else {
          if ((status == ENTRY) ^ U) v[I].label = L_CROSSING;
          first_chain_vertex = 1;
          toggle(&status);
        }
      }
      if (v[I].label == L_DELAYED_BOUNCING) {
        v[I].enex = status;
        if (first_chain_vertex) {
          if (((status == EXIT) ^ U) && !v[I].cross) {
            v[I].cross = 1;
            crossList[i] = (int *)realloc(crossList[i], sizeof(int) * (size_t)(ncross[i] + 1));
            crossList[i][ncross[i]++] = I;
          }
          first_chain_vertex = 0;
        } else {
          if (((status == ENTRY) ^ U) && !v[I].cross) {
            v[I].cross = 1;
            crossList[i] = (int *)realloc(crossList[i], sizeof(int) * (size_t)(ncross[i] + 1));
            crossList[i][ncross[i]++] = I;
          }
          first_chain_vertex = 1;
        }
        toggle(&status);
      }
      do { I = pl->v[I].next; } while (!pl->v[I].inter && I != V);
      if (I == V) break;
    }
    }
  }

  /* 5) split vertex pairs */
  for (int k = 0; k < nsplit[0]; k++) {
    int I_P = splitList[0][k];
    int I_Q = pl->v[I_P].nb;
    /* membership of split[1]: the split flag is only set while walking Q */
    if (I_Q < 0 || !pl->v[I_Q].split) continue;
    pt pP = pl->v[I_P].p, oP = pl->v[I_P].o, pQ = pl->v[I_Q].p, oQ = pl->v[I_Q].o;
    int V_P = vnew(pl, pP, oP, -1.0);
    int V_Q = vnew(pl, pQ, oQ, -1.0);
    vtx *v = pl->v;
    double sP = triA(v[v[I_P].prev].p, v[I_P].p, v[v[I_P].next].p);
    double sQ = triA(v[v[I_Q].prev].p, v[I_Q].p, v[v[I_Q].next].p);
    if (sP * sQ > 0) {
      vlink(pl, I_P, V_Q);
      vlink(pl, I_Q, V_P);
    } else {
      vlink(pl, V_P, V_Q);
    }
    insert_vertex(pl, V_P, I_P, -1.0);
    insert_vertex(pl, V_Q, I_Q, -1.0);
    v = pl->v;
    const int ex = U ? ENTRY : EXIT, en = U ? EXIT : ENTRY;
    v[I_P].enex = ex; v[V_P].enex = en;
    v[I_Q].enex = ex; v[V_Q].enex = en;
    v[I_P].label = L_CROSSING; v[V_P].label = L_CROSSING;
    v[I_Q].label = L_CROSSING; v[V_Q].label = L_CROSSING;
  }

  /* 6) crossing candidates */
  for (int k = 0; k < ncross[0]; k++) {
    int I_P = crossList[0][k];
    int I_Q = pl->v[I_P].nb;
    if (I_Q < 0) continue;
    if (pl->v[I_Q].cross) {
      pl->v[I_P].label = L_CROSSING;
      pl->v[I_Q].label = L_CROSSING;
    }
  }

  /* createResult, polyclip.h:743-783, over the components of P */
  long long guard = 4LL * (long long)pl->n + 64;
  for (int c = 0; c < nr[0]; c++) {
    root = roots[0][c];
#define IS_CI(k) (pl->v[(k)].inter && pl->v[(k)].label == L_CROSSING)
    int I = root;
    int have = IS_CI(I);
    if (!have) {
      do { I = pl->v[I].next; } while (!IS_CI(I) && I != root);
      have = (I != root);
    }
    while (have) {
      int R = -1;
      int V = I;
      pl->v[V].inter = 0;
      long long steps = 0;
      do {
        int status = pl->v[V].enex;
        toggle(&status);
        do {
          if ((status == EXIT) ^ U) V = pl->v[V].next;
          else V = pl->v[V].prev;
          pl->v[V].inter = 0;
          pt p = pl->v[V].p, o = pl->v[V].o;
          int nv = vnew(pl, p, o, -1.0);
          poly_append(pl, &R, nv);
          if (++steps > guard) { rc = -1; goto done; }
        } while (!(pl->v[V].enex == status) && (V != I));
        if (V != I) {
          V = pl->v[V].nb;
          if (V < 0) { rc = -1; goto done; }
          pl->v[V].inter = 0;
        }
      } while (V != I);
      res_push(g, R);
      do { I = pl->v[I].next; } while (!IS_CI(I) && I != root);
      have = (I != root);
    }
#undef IS_CI
  }
done:
  free(splitList[0]); free(splitList[1]);
  free(crossList[0]); free(crossList[1]);
  for (int i = 0; i < 2; i++) { free(noInter[i]); free(ident[i]); }
  if (rc < 0) return rc;

  /* cleanUpResult, polyclip.h:791-807 */
  for (int k = 0; do_cleanup && k < g->nres; k++) {
    int *r = &g->res_roots[k];
    while (*r >= 0) {
      vtx *v = pl->v;
      int V = *r;
      if (fabs(triA(v[v[V].prev].p, v[V].p, v[v[V].next].p)) < GH_EPS) remove_vertex(pl, r, V);
      else break;
    }
    if (*r >= 0) {
      int it = *r, rt = *r;
      do {
        vtx *v = pl->v;
        int nx = v[it].next;
        if (fabs(triA(v[v[it].prev].p, v[it].p, v[v[it].next].p)) < GH_EPS) remove_vertex(pl, r, it);
        it = nx;
      } while (it != rt);
    }
  }
  return 0;
}

static void collect(ghstate *g, orc_result_t *out) {
  vpool *pl = &g->pl;
  int nc = 0, nv = 0;
  for (int k = 0; k < g->nres; k++) {
    int r = g->res_roots[k];
    if (r < 0) continue;
    nc++;
    int V = r;
    do { nv++; V = pl->v[V].next; } while (V != r);
  }
  out->ncomp = nc;
  out->nverts = nv;
  out->comp_sizes = (int *)malloc(sizeof(int) * (size_t)(nc ? nc : 1));
  out->xy = (double *)malloc(sizeof(double) * 2 * (size_t)(nv ? nv : 1));
  int c = 0, w = 0;
  for (int k = 0; k < g->nres; k++) {
    int r = g->res_roots[k];
    if (r < 0) continue;
    int s = 0, V = r;
    do {
      out->xy[2 * w] = pl->v[V].o.x;
      out->xy[2 * w + 1] = pl->v[V].o.y;
      w++; s++;
      V = pl->v[V].next;
    } while (V != r);
    out->comp_sizes[c++] = s;
  }
}

static void gh_free(ghstate *g) {
  free(g->pl.v);
  free(g->res_roots);
}

void orc_result_free(orc_result_t *r) {
  if (!r) return;
  free(r->comp_sizes);
  free(r->xy);
  r->comp_sizes = NULL;
  r->xy = NULL;
}

static void result_empty(orc_result_t *out, int status) {
  memset(out, 0, sizeof(*out));
  out->status = status;
  out->comp_sizes = (int *)malloc(sizeof(int));
  out->xy = (double *)malloc(2 * sizeof(double));
}

/* ----------------------------------------------------------------------
 * CPU route: computeIntersections, polyclip.h:188-281
 * -------------------------------------------------------------------- */
/* the frame of a union whose MBRs are disjoint: the joint MBR's corner and
 * extent take the CMBR's place */
static void joint_frame(const double *P, int nP, const double *Q, int nQ, int normalize,
                        double *norm) {
  double a[4], b[4];
  ring_mbr(P, nP, a);
  ring_mbr(Q, nQ, b);
  double x0 = a[0] < b[0] ? a[0] : b[0], y0 = a[1] < b[1] ? a[1] : b[1];
  double x1 = a[2] > b[2] ? a[2] : b[2], y1 = a[3] > b[3] ? a[3] : b[3];
  norm[0] = 0.0; norm[1] = 0.0; norm[2] = 1.0;
  if (normalize & 1) {
    double e = (x1 - x0) > (y1 - y0) ? (x1 - x0) : (y1 - y0);
    norm[0] = x0;
    norm[1] = y0;
    norm[2] = e > 0.0 ? 1.0 / e : 1.0;
    if (normalize & 4) norm[2] *= edge_scale(nP, nQ);
  }
}

/* CPU route over complex polygons: computeIntersections (polyclip.h:188-281,
 * every source edge of every component of P against every source edge of
 * every component of Q), labelIntersections with steps 1 and 3 restored,
 * createResult, cleanUpResult.  mode: 1 normalise, 2 no clean-up, 4 edge
 * scale, 8 union (polyclip.h:42). */
int orc_clip_multi(const double *P, const int *pSizes, int nPR, const double *Q,
                   const int *qSizes, int nQR, int mode, orc_result_t *out) {
  int nP = 0, nQ = 0;
  for (int r = 0; r < nPR; r++) nP += pSizes[r];
  for (int r = 0; r < nQR; r++) nQ += qSizes[r];
  double cmbr[4], norm[3];
  if (nP < 1 || nQ < 1) {
    result_empty(out, 2);
    return 2;
  }
  if (!pair_frame(P, nP, Q, nQ, mode & 5, cmbr, norm)) {
    if (!(mode & 8)) {
      result_empty(out, 2);
      return 2;
    }
    joint_frame(P, nP, Q, nQ, mode & 5, norm);
  }
  ghstate g;
  memset(&g, 0, sizeof(g));
  g.rootP = g.rootQ = -1;
  g.unionOp = (mode & 8) ? 1 : 0;
  g.rP = (int *)malloc(sizeof(int) * (size_t)(nPR + 1));
  g.rQ = (int *)malloc(sizeof(int) * (size_t)(nQR + 1));
  const double *src[2] = {P, Q};
  const int *sizes[2] = {pSizes, qSizes}, nrs[2] = {nPR, nQR};
  int *rts[2] = {g.rP, g.rQ}, *cnt[2] = {&g.nRP, &g.nRQ};
  for (int side = 0; side < 2; side++) {
    int off = 0;
    for (int r = 0; r < nrs[side]; r++) {
      int root = -1;
      for (int i = 0; i < sizes[side][r]; i++) {
        pt o = {src[side][2 * (off + i)], src[side][2 * (off + i) + 1]};
        int V = vnew(&g.pl, tf(o.x, o.y, norm), o, -1.0);
        g.pl.v[V].source = 1;
        poly_append(&g.pl, &root, V);
      }
      off += sizes[side][r];
      if (root >= 0) rts[side][(*cnt[side])++] = root;
    }
  }
  long long n_eval = 0, n_inter = 0;
  /* source edges of P x source edges of Q; inserted vertices are skipped by
   * walking source vertices only (edgeIterator with SOURCE) */
  for (int cp = 0; cp < g.nRP; cp++) {
  int rootP = g.rP[cp];
  int one = rootP;
  do {
    int two = g.pl.v[one].next;
    while (!g.pl.v[two].source) two = g.pl.v[two].next;
    for (int cq = 0; cq < g.nRQ; cq++) {
    int rootQ = g.rQ[cq];
    int q1 = rootQ;
    do {
      int q2 = g.pl.v[q1].next;
      while (!g.pl.v[q2].source) q2 = g.pl.v[q2].next;
      double alpha = 0, beta = 0;
      vtx *v = g.pl.v;
      pt P1 = v[one].p, P2 = v[two].p, Q1 = v[q1].p, Q2 = v[q2].p;
      pt P1o = v[one].o, P2o = v[two].o, Q1o = v[q1].o;
      n_eval++;
      int t = itype(P1, P2, Q1, Q2, &alpha, &beta);
      if (t) n_inter++;
      int IP, IQ;
      switch (t) {
      case 1: {
        pt I = {(1.0 - alpha) * P1.x + alpha * P2.x, (1.0 - alpha) * P1.y + alpha * P2.y};
        pt Io = {(1.0 - alpha) * P1o.x + alpha * P2o.x, (1.0 - alpha) * P1o.y + alpha * P2o.y};
        IP = vnew(&g.pl, I, Io, alpha);
        IQ = vnew(&g.pl, I, Io, beta);
        insert_vertex(&g.pl, IP, one, alpha);
        insert_vertex(&g.pl, IQ, q1, beta);
        vlink(&g.pl, IP, IQ);
        break;
      }
      case 5:
        IQ = vnew(&g.pl, P1, P1o, beta);
        insert_vertex(&g.pl, IQ, q1, beta);
        vlink(&g.pl, one, IQ);
        IP = vnew(&g.pl, Q1, Q1o, alpha);
        insert_vertex(&g.pl, IP, one, alpha);
        vlink(&g.pl, IP, q1);
        break;
      case 2:
      case 6:
        IQ = vnew(&g.pl, P1, P1o, beta);
        insert_vertex(&g.pl, IQ, q1, beta);
        vlink(&g.pl, one, IQ);
        break;
      case 3:
      case 7:
        IP = vnew(&g.pl, Q1, Q1o, alpha);
        insert_vertex(&g.pl, IP, one, alpha);
        vlink(&g.pl, IP, q1);
        break;
      case 4:
      case 8:
        vlink(&g.pl, one, q1);
        break;
      }
      q1 = q2;
    } while (q1 != rootQ);
    }
    one = two;
  } while (one != rootP);
  }

  label_step1_and_3(&g);
  int rc = label_trace_cleanup(&g, !(mode & 2));
  memset(out, 0, sizeof(*out));
  if (rc < 0) {
    result_empty(out, -1);
  } else {
    collect(&g, out);
  }
  out->n_eval = n_eval;
  out->n_inter = n_inter;
  free(g.rP);
  free(g.rQ);
  gh_free(&g);
  return out->status;
}

int orc_clip_cpu(const double *P, int nP, const double *Q, int nQ, int normalize,
                 orc_result_t *out) {
  return orc_clip_multi(P, &nP, 1, Q, &nQ, 1, normalize & 7, out);
}

/* ----------------------------------------------------------------------
 * reference GPU phase 1, gppolyclip_mDIV512.cu:382-1223, 1298-1715
 * -------------------------------------------------------------------- */
void orc_phase1_free(orc_phase1_t *t) {
  if (!t) return;
  free(t->psP1); free(t->psP2); free(t->psQ1); free(t->psQ2);
  free(t->boolP); free(t->boolQ);
  free(t->neighborMapQ);
  free(t->intersectionsP); free(t->intersectionsQ);
  free(t->origP); free(t->origQ);
  free(t->alphaP); free(t->alphaQ);
  free(t->initLabelsP); free(t->initLabelsQ);
  free(t->neighborP); free(t->neighborQ);
  memset(t, 0, sizeof(*t));
}

static pt edge_end(const pt *c, int n, int i) { return c[i == n - 1 ? 0 : i + 1]; }

/* gpuCMBRFilter, :382-415 */
static void cmbr_filter(const pt *c, int n, const pt *cmin, const pt *cmax, int *flag) {
  for (int id = 0; id < n; id++) {
    pt P1 = c[id], P2 = edge_end(c, n, id);
    double minX = P1.x < P2.x ? P1.x : P2.x, minY = P1.y < P2.y ? P1.y : P2.y;
    double maxX = P1.x < P2.x ? P2.x : P1.x, maxY = P1.y < P2.y ? P2.y : P1.y;
    int b = 1;
    if (minX > cmax->x || maxX < cmin->x) b = 0;
    if (minY > cmax->y || maxY < cmin->y) b = 0;
    flag[id] = b;
  }
}

/* gpuCountIntersections, :463-548 (first argument's edges are counted) */
static void count_view(const pt *a, int na, const pt *b, int nb, const int *flagA,
                       int *c1, int *c2, long long *n_eval) {
  for (int id = 0; id < na; id++) {
    int k1 = 0, k2 = 0;
    if (flagA[id]) {
      pt P1 = a[id], P2 = edge_end(a, na, id);
      for (int q = 0; q < nb; q++) {
        pt Q1 = b[q], Q2 = edge_end(b, nb, q);
        if (lsmf(P1, P2, Q1, Q2)) {
          double al, be;
          if (n_eval) (*n_eval)++;
          int t = itype(P1, P2, Q1, Q2, &al, &be);
          if (t != 0) {
            k1++;
            if (t == 1 || t == 3 || t == 5 || t == 7) k2++;
          }
        }
      }
    }
    c1[id] = k1;
    c2[id] = k2 + 1;
  }
}

static void exscan(int *a, int n) { /* n+1 entries, a[n] gets the total */
  int s = 0;
  for (int i = 0; i < n; i++) { int v = a[i]; a[i] = s; s += v; }
  a[n] = s;
}

/* stable decimal LSD sort of keys[start..end) (gpuRadixsort / gpuCountSort,
 * :263-305); writes the permutation into idx[start..end) */
static void alpha_sort(const int *keys, int *idx, int *tmp, int start, int end) {
  for (int i = start; i < end; i++) idx[i] = i;
  int exp = 1;
  for (int d = 0; d < ALPHA_DIGITS; d++) {
    int count[10] = {0};
    for (int i = start; i < end; i++) {
      tmp[i] = idx[i];
      int k = keys[tmp[i]];
      int dg = k < 0 ? 0 : (k / exp) % 10;
      count[dg]++;
    }
    for (int i = 1; i < 10; i++) count[i] += count[i - 1];
    for (int i = end - 1; i >= start; i--) {
      int k = keys[tmp[i]];
      int dg = k < 0 ? 0 : (k / exp) % 10;
      idx[start + (count[dg] - 1)] = tmp[i];
      count[dg]--;
    }
    exp *= 10;
  }
}

static int circ(int id, int max) {
  if (max == id) return 0;
  if (id == -1) return max - 1;
  return id;
}

/* oracle(), :195-217 */
static int rel_idx(int pMNId, int pPNId, int qId, pt Q, pt P1, pt P2, pt P3) {
  if (pMNId != -100 && pMNId == qId) return 2;
  else if (pPNId != -100 && pPNId == qId) return 3;
  double s1 = triA(Q, P1, P2), s2 = triA(Q, P2, P3), s3 = triA(P1, P2, P3);
  if (s3 > 0) {
    if (s1 > 0 && s2 > 0) return 0;
    return 1;
  }
  if (s1 < 0 && s2 < 0) return 1;
  return 0;
}

/* getInitialLabel, :237-252 */
int orc_initial_label(int qMType, int qPType) {
  if ((qMType == 0 && qPType == 1) || (qMType == 1 && qPType == 0)) return 1;
  if ((qMType == 0 && qPType == 0) || (qMType == 1 && qPType == 1)) return 2;
  if (((qPType == 3) && (qMType == 1)) || ((qMType == 3) && (qPType == 1))) return 3;
  if (((qPType == 3) && (qMType == 0)) || ((qMType == 3) && (qPType == 0))) return 4;
  if (((qPType == 3) && (qMType == 2)) || ((qMType == 3) && (qPType == 2))) return 5;
  if (((qMType == 2) && (qPType == 1)) || ((qPType == 2) && (qMType == 1))) return 6;
  if (((qMType == 2) && (qPType == 0)) || ((qPType == 2) && (qMType == 0))) return 7;
  return -102;
}

int orc_phase1(const double *Pc, int nP, const double *Qc, int nQ, int normalize,
               orc_phase1_t *t) {
  memset(t, 0, sizeof(*t));
  t->nP = nP; t->nQ = nQ;
  if (!pair_frame(Pc, nP, Qc, nQ, normalize, t->cmbr, t->norm)) return 2;
  pt *P = (pt *)malloc(sizeof(pt) * (size_t)nP), *Q = (pt *)malloc(sizeof(pt) * (size_t)nQ);
  pt *Po = (pt *)malloc(sizeof(pt) * (size_t)nP), *Qo = (pt *)malloc(sizeof(pt) * (size_t)nQ);
  for (int i = 0; i < nP; i++) { Po[i].x = Pc[2 * i]; Po[i].y = Pc[2 * i + 1]; P[i] = tf(Po[i].x, Po[i].y, t->norm); }
  for (int i = 0; i < nQ; i++) { Qo[i].x = Qc[2 * i]; Qo[i].y = Qc[2 * i + 1]; Q[i] = tf(Qo[i].x, Qo[i].y, t->norm); }
  pt cmin = tf(t->cmbr[0], t->cmbr[1], t->norm), cmax = tf(t->cmbr[2], t->cmbr[3], t->norm);

  t->psP1 = (int *)calloc((size_t)nP + 1, sizeof(int));
  t->psP2 = (int *)calloc((size_t)nP + 1, sizeof(int));
  t->psQ1 = (int *)calloc((size_t)nQ + 1, sizeof(int));
  t->psQ2 = (int *)calloc((size_t)nQ + 1, sizeof(int));
  t->boolP = (int *)calloc((size_t)nP + 1, sizeof(int));
  t->boolQ = (int *)calloc((size_t)nQ + 1, sizeof(int));
  cmbr_filter(P, nP, &cmin, &cmax, t->boolP);
  cmbr_filter(Q, nQ, &cmin, &cmax, t->boolQ);
  count_view(Q, nQ, P, nP, t->boolQ, t->psQ1, t->psQ2, NULL);
  count_view(P, nP, Q, nQ, t->boolP, t->psP1, t->psP2, NULL);
  exscan(t->psP1, nP); exscan(t->psP2, nP);
  exscan(t->psQ1, nQ); exscan(t->psQ2, nQ);
  t->cP = t->psP2[nP];
  t->cQ = t->psQ2[nQ];
  if (t->psP1[nP] != t->psQ1[nQ]) {
    t->invalid = 1;
    free(P); free(Q); free(Po); free(Qo);
    return 1;
  }
  int cP = t->cP, cQ = t->cQ;
  const int *psP1 = t->psP1, *psP2 = t->psP2, *psQ1 = t->psQ1, *psQ2 = t->psQ2;

  /* gpuNeighborMap, :549-624 */
  long long *nmap = (long long *)malloc(sizeof(long long) * (size_t)(cQ ? cQ : 1));
  for (int i = 0; i < cQ; i++) nmap[i] = -100;
  for (int id = 0; id < nQ; id++) {
    if (psQ1[id + 1] == psQ1[id]) continue;
    pt A1 = Q[id], A2 = edge_end(Q, nQ, id);
    int c2 = 0, nd = 0;
    for (int q = 0; q < nP; q++) {
      if (psP1[q + 1] == psP1[q]) continue;
      pt B1 = P[q], B2 = edge_end(P, nP, q);
      if (!lsmf(A1, A2, B1, B2)) continue;
      double al, be;
      int i = itype(A1, A2, B1, B2, &al, &be);
      if (i != 0) {
        if (i == 1 || i == 3 || i == 5 || i == 7) { nd++; c2 = nd; }
        else c2 = 0;
        nmap[psQ2[id] + c2] = q;
      }
    }
  }
  t->neighborMapQ = nmap;

  /* gpuCalculateIntersections, :897-1123 */
  double *iP = (double *)calloc(2 * (size_t)(cP ? cP : 1), sizeof(double));
  double *iP2 = (double *)calloc(2 * (size_t)(cP ? cP : 1), sizeof(double));
  double *oP = (double *)calloc(2 * (size_t)(cP ? cP : 1), sizeof(double));
  double *oP2 = (double *)calloc(2 * (size_t)(cP ? cP : 1), sizeof(double));
  double *iQ = (double *)calloc(2 * (size_t)(cQ ? cQ : 1), sizeof(double));
  double *iQ2 = (double *)calloc(2 * (size_t)(cQ ? cQ : 1), sizeof(double));
  double *oQ = (double *)calloc(2 * (size_t)(cQ ? cQ : 1), sizeof(double));
  double *oQ2 = (double *)calloc(2 * (size_t)(cQ ? cQ : 1), sizeof(double));
  int *aP = (int *)malloc(sizeof(int) * (size_t)(cP ? cP : 1));
  int *aQ = (int *)malloc(sizeof(int) * (size_t)(cQ ? cQ : 1));
  int *nP_ = (int *)calloc((size_t)(cP ? cP : 1), sizeof(int));
  int *nP2 = (int *)calloc((size_t)(cP ? cP : 1), sizeof(int));
  int *nQ_ = (int *)calloc((size_t)(cQ ? cQ : 1), sizeof(int));
  int *nQ2 = (int *)calloc((size_t)(cQ ? cQ : 1), sizeof(int));
  int *tmpB = (int *)malloc(sizeof(int) * (size_t)((cP > cQ ? cP : cQ) + 1));
  int *sidx = (int *)malloc(sizeof(int) * (size_t)((cP > cQ ? cP : cQ) + 1));
  for (int i = 0; i < cP; i++) aP[i] = -100;
  for (int i = 0; i < cQ; i++) aQ[i] = -100;
#define SETP(arr, s, v) do { (arr)[2 * (s)] = (v).x; (arr)[2 * (s) + 1] = (v).y; } while (0)
  for (int q = 0; q < nQ; q++) {
    SETP(iQ, psQ2[q], Q[q]); SETP(iQ2, psQ2[q], Q[q]);
    SETP(oQ, psQ2[q], Qo[q]); SETP(oQ2, psQ2[q], Qo[q]);
  }
  for (int pid = 0; pid < nP; pid++) {
    int base = psP2[pid];
    SETP(iP, base, P[pid]); SETP(iP2, base, P[pid]);
    SETP(oP, base, Po[pid]); SETP(oP2, base, Po[pid]);
    aP[base] = -100;
    if (psP1[pid + 1] == psP1[pid]) continue;
    pt P1 = P[pid], P2 = edge_end(P, nP, pid);
    pt P1o = Po[pid], P2o = edge_end(Po, nP, pid);
    int c2 = 0, nd = 0, nQId = -1;
    for (int qid = 0; qid < nQ; qid++) {
      if (psQ1[qid + 1] == psQ1[qid]) continue;
      pt Q1 = Q[qid], Q2 = edge_end(Q, nQ, qid);
      pt Q1o = Qo[qid];
      if (!lsmf(P1, P2, Q1, Q2)) continue;
      double alpha, beta;
      int i = itype(P1, P2, Q1, Q2, &alpha, &beta);
      if (!i) continue;
      if (i == 1 || i == 3 || i == 5 || i == 7) { nd++; c2 = nd; }
      else c2 = 0;
      int start = psQ2[qid], end = psQ2[qid + 1];
      if (i != 5) {
        for (int l = start; l < end; l++) {
          if (nmap[l] == pid) {
            nQId = l;
            nP_[base + c2] = l + 1; nP2[base + c2] = l + 1;
            nQ_[l] = base + c2 + 1; nQ2[l] = base + c2 + 1;
            break;
          }
        }
      } else {
        nQId = start;
        nP_[base + c2] = start + 1; nP2[base + c2] = start + 1;
        nQ_[start] = base + c2 + 1; nQ2[start] = base + c2 + 1;
        for (int l = start; l < end; l++) {
          if (nmap[l] == pid) {
            nQId = l;
            nP_[base] = l + 1; nP2[base] = l + 1;
            nQ_[l] = base + 1; nQ2[l] = base + 1;
            break;
          }
        }
      }
      int ka = trunc_sat((double)ALPHA_SCALE * alpha);
      int kb = trunc_sat((double)ALPHA_SCALE * beta);
      switch (i) {
      case 1: {
        pt I = {(1.0 - alpha) * P1.x + alpha * P2.x, (1.0 - alpha) * P1.y + alpha * P2.y};
        pt Io = {(1.0 - alpha) * P1o.x + alpha * P2o.x, (1.0 - alpha) * P1o.y + alpha * P2o.y};
        SETP(iP, base + c2, I); SETP(iP2, base + c2, I);
        SETP(oP, base + c2, Io); SETP(oP2, base + c2, Io);
        aP[base + c2] = ka;
        if (nQId >= 0) {
          SETP(iQ, nQId, I); SETP(iQ2, nQId, I);
          SETP(oQ, nQId, Io); SETP(oQ2, nQId, Io);
          aQ[nQId] = kb;
        }
        break;
      }
      case 5:
        SETP(iP, base + c2, Q1); SETP(iP2, base + c2, Q1);
        SETP(oP, base + c2, Q1o); SETP(oP2, base + c2, Q1o);
        aP[base + c2] = ka;
        if (nQId >= 0) {
          SETP(iQ, nQId, P1); SETP(iQ2, nQId, P1);
          SETP(oQ, nQId, P1o); SETP(oQ2, nQId, P1o);
          aQ[nQId] = kb;
        }
        break;
      case 2:
      case 6:
        aP[base] = ka;
        if (nQId >= 0) {
          SETP(iQ, nQId, P1); SETP(iQ2, nQId, P1);
          SETP(oQ, nQId, P1o); SETP(oQ2, nQId, P1o);
          aQ[nQId] = kb;
        }
        break;
      case 3:
      case 7:
        SETP(iP, base + c2, Q1); SETP(iP2, base + c2, Q1);
        SETP(oP, base + c2, Q1o); SETP(oP2, base + c2, Q1o);
        aP[base + c2] = ka;
        aQ[psQ2[qid]] = kb;
        break;
      case 4:
      case 8:
        aP[base] = ka;
        aQ[psQ2[qid]] = kb;
        break;
      }
    }
    /* per-edge alpha sort, :1103-1120 */
    int start = psP2[pid], end = psP2[pid + 1];
    if (end - start > 2) {
      alpha_sort(aP, sidx, tmpB, start + 1, end);
      for (int k = start + 1; k < end; k++) tmpB[k] = aP[sidx[k]];
      for (int k = start + 1; k < end; k++) {
        int s = sidx[k];
        aP[k] = tmpB[k];
        iP[2 * k] = iP2[2 * s]; iP[2 * k + 1] = iP2[2 * s + 1];
        oP[2 * k] = oP2[2 * s]; oP[2 * k + 1] = oP2[2 * s + 1];
        nP_[k] = nP2[s];
        if (nP2[s] > 0) { nQ_[nP2[s] - 1] = k + 1; nQ2[nP2[s] - 1] = k + 1; }
      }
    }
  }
  /* gpuSortPolyQ, :1132-1160 */
  for (int id = 0; id < nQ; id++) {
    int start = psQ2[id], end = psQ2[id + 1];
    if (end - start > 2) {
      alpha_sort(aQ, sidx, tmpB, start + 1, end);
      for (int k = start + 1; k < end; k++) tmpB[k] = aQ[sidx[k]];
      for (int k = start + 1; k < end; k++) {
        int s = sidx[k];
        aQ[k] = tmpB[k];
        iQ[2 * k] = iQ2[2 * s]; iQ[2 * k + 1] = iQ2[2 * s + 1];
        oQ[2 * k] = oQ2[2 * s]; oQ[2 * k + 1] = oQ2[2 * s + 1];
        nQ_[k] = nQ2[s];
        if (nQ2[s] > 0) nP_[nQ2[s] - 1] = k + 1;
      }
    }
  }
  /* gpuCalculateInitLabel, :1171-1223 */
  int *lP = (int *)malloc(sizeof(int) * (size_t)(cP ? cP : 1));
  int *lQ = (int *)malloc(sizeof(int) * (size_t)(cQ ? cQ : 1));
  for (int i = 0; i < cQ; i++) lQ[i] = -100;
  for (int i = 0; i < cP; i++) {
    lP[i] = -100;
    if (aP[i] == -100) continue;
    pt cur = {iP[2 * i], iP[2 * i + 1]};
    int tmp = circ(i - 1, cP);
    pt pM = {iP[2 * tmp], iP[2 * tmp + 1]};
    int pMN = aP[tmp] != -100 ? nP_[tmp] - 1 : -100;
    tmp = circ(i + 1, cP);
    pt pP = {iP[2 * tmp], iP[2 * tmp + 1]};
    int pPN = aP[tmp] != -100 ? nP_[tmp] - 1 : -100;
    int nId = nP_[i] - 1;
    if (nId < 0 || nId >= cQ) continue;
    tmp = circ(nId - 1, cQ);
    pt qM = {iQ[2 * tmp], iQ[2 * tmp + 1]};
    int tM = rel_idx(pMN, pPN, tmp, qM, pM, cur, pP);
    tmp = circ(nId + 1, cQ);
    pt qP = {iQ[2 * tmp], iQ[2 * tmp + 1]};
    int tP = rel_idx(pMN, pPN, tmp, qP, pM, cur, pP);
    int lab = orc_initial_label(tM, tP);
    lP[i] = lab;
    lQ[nId] = lab;
  }
#undef SETP
  t->intersectionsP = iP; t->intersectionsQ = iQ;
  t->origP = oP; t->origQ = oQ;
  t->alphaP = aP; t->alphaQ = aQ;
  t->neighborP = nP_; t->neighborQ = nQ_;
  t->initLabelsP = lP; t->initLabelsQ = lQ;
  free(iP2); free(iQ2); free(oP2); free(oQ2); free(nP2); free(nQ2);
  free(tmpB); free(sidx);
  free(P); free(Q); free(Po); free(Qo);
  return 0;
}

/* ----------------------------------------------------------------------
 * GPU route: regularPolygonHandler splice (ghcuda.cpp:239-322) by array
 * index (parent slot of edge p is psP2[p]), neighbour links with link()
 * semantics, then the CPU labelling from step 1.
 * -------------------------------------------------------------------- */
int orc_clip_gpu_route(const double *P, int nP, const double *Q, int nQ,
                       int normalize, orc_result_t *out) {
  orc_phase1_t t;
  int rc = orc_phase1(P, nP, Q, nQ, normalize & 5, &t);
  if (rc == 2 || rc == 1) {
    result_empty(out, rc);
    out->n_inter = t.psP1 ? t.psP1[nP] : 0;
    orc_phase1_free(&t);
    return rc;
  }
  ghstate g;
  memset(&g, 0, sizeof(g));
  g.rootP = g.rootQ = -1;
  single_roots(&g);
  int *mapP = (int *)malloc(sizeof(int) * (size_t)(t.cP ? t.cP : 1));
  int pe = 0;
  for (int s = 0; s < t.cP; s++) {
    pt p = {t.intersectionsP[2 * s], t.intersectionsP[2 * s + 1]};
    pt o = {t.origP[2 * s], t.origP[2 * s + 1]};
    int V = vnew(&g.pl, p, o, -1.0);
    int parent = (pe < nP && t.psP2[pe] == s);
    if (parent) { g.pl.v[V].source = 1; g.pl.v[V].inter = t.alphaP[s] != -100; pe++; }
    else g.pl.v[V].inter = 1;
    mapP[s] = V;
    poly_append(&g.pl, &g.rootP, V);
  }
  int qe = 0;
  for (int s = 0; s < t.cQ; s++) {
    pt p = {t.intersectionsQ[2 * s], t.intersectionsQ[2 * s + 1]};
    pt o = {t.origQ[2 * s], t.origQ[2 * s + 1]};
    int V = vnew(&g.pl, p, o, -1.0);
    int parent = (qe < nQ && t.psQ2[qe] == s);
    int linked;
    /* ghcuda.cpp:310 links a Q parent only when its alpha is set; an
     * X-overlap (type 5) links Q's parent without setting it, so a parent is
     * linked whenever phase 1 recorded a neighbour for it */
    if (parent) { g.pl.v[V].source = 1; linked = t.alphaQ[s] != -100 || t.neighborQ[s] > 0; qe++; }
    else linked = 1;
    poly_append(&g.pl, &g.rootQ, V);
    if (linked) {
      int k = t.neighborQ[s] - 1;
      if (k >= 0 && k < t.cP) vlink(&g.pl, V, mapP[k]);
      else g.pl.v[V].inter = 1;
    }
  }
  free(mapP);
  label_step1_and_3(&g);
  rc = label_trace_cleanup(&g, !(normalize & 2));
  memset(out, 0, sizeof(*out));
  if (rc < 0) result_empty(out, -1);
  else collect(&g, out);
  out->n_inter = t.psP1[nP];
  gh_free(&g);
  orc_phase1_free(&t);
  return out->status;
}

/* ----------------------------------------------------------------------
 * FP64 area oracle, independent of labelling and tracing
 * -------------------------------------------------------------------- */
typedef struct { double s, c; } ksum;
static void kadd(ksum *k, double v) { /* Neumaier */
  double t = k->s + v;
  if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v;
  else k->c += (v - t) + k->s;
  k->s = t;
}

static int pip_plain(const pt *c, int n, pt R) {
  int in = 0;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    if ((c[i].y > R.y) != (c[j].y > R.y)) {
      double xi = c[j].x + (R.y - c[j].y) * (c[i].x - c[j].x) / (c[i].y - c[j].y);
      if (R.x < xi) in = !in;
    }
  }
  return in;
}

static int cmp_d(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return (x > y) - (x < y);
}

static double pieces_inside(const pt *a, int na, const pt *b, int nb) {
  ksum acc = {0, 0};
  double *ts = (double *)malloc(sizeof(double) * (size_t)(nb + 2));
  for (int i = 0; i < na; i++) {
    pt A1 = a[i], A2 = a[(i + 1) % na];
    double dx = A2.x - A1.x, dy = A2.y - A1.y;
    int nt = 0;
    ts[nt++] = 0.0;
    for (int j = 0; j < nb; j++) {
      pt B1 = b[j], B2 = b[(j + 1) % nb];
      double ex = B2.x - B1.x, ey = B2.y - B1.y;
      double den = dx * ey - dy * ex;
      if (den == 0.0) continue;
      double wx = B1.x - A1.x, wy = B1.y - A1.y;
      double tA = (wx * ey - wy * ex) / den;
      double tB = (wx * dy - wy * dx) / den;
      if (tA > 0.0 && tA < 1.0 && tB >= 0.0 && tB < 1.0) ts[nt++] = tA;
    }
    ts[nt++] = 1.0;
    qsort(ts, (size_t)nt, sizeof(double), cmp_d);
    for (int k = 0; k + 1 < nt; k++) {
      double t0 = ts[k], t1 = ts[k + 1];
      if (t1 <= t0) continue;
      double tm = 0.5 * (t0 + t1);
      pt M = {A1.x + tm * dx, A1.y + tm * dy};
      if (!pip_plain(b, nb, M)) continue;
      pt S = {A1.x + t0 * dx, A1.y + t0 * dy};
      pt E = {A1.x + t1 * dx, A1.y + t1 * dy};
      if (k == 0) S = A1;
      if (k + 2 == nt) E = A2;
      kadd(&acc, 0.5 * (S.x * E.y - E.x * S.y));
    }
  }
  free(ts);
  return acc.s + acc.c;
}

static double signed_area(const pt *c, int n) {
  ksum k = {0, 0};
  for (int i = 0; i < n; i++) {
    pt a = c[i], b = c[(i + 1) % n];
    kadd(&k, 0.5 * (a.x * b.y - b.x * a.y));
  }
  return k.s + k.c;
}

double orc_area_boundary(const double *Pc, int nP, const double *Qc, int nQ) {
  double cm[4], nm[3];
  if (!pair_frame(Pc, nP, Qc, nQ, 0, cm, nm)) return 0.0;
  /* evaluate in a frame centred on the CMBR to limit cancellation */
  double ox = 0.5 * (cm[0] + cm[2]), oy = 0.5 * (cm[1] + cm[3]);
  pt *P = (pt *)malloc(sizeof(pt) * (size_t)nP), *Q = (pt *)malloc(sizeof(pt) * (size_t)nQ);
  for (int i = 0; i < nP; i++) { P[i].x = Pc[2 * i] - ox; P[i].y = Pc[2 * i + 1] - oy; }
  for (int i = 0; i < nQ; i++) { Q[i].x = Qc[2 * i] - ox; Q[i].y = Qc[2 * i + 1] - oy; }
  double sp = signed_area(P, nP) >= 0 ? 1.0 : -1.0;
  double sq = signed_area(Q, nQ) >= 0 ? 1.0 : -1.0;
  double a = sp * pieces_inside(P, nP, Q, nQ) + sq * pieces_inside(Q, nQ, P, nP);
  free(P); free(Q);
  return a;
}

double orc_result_area(const orc_result_t *r, double ox, double oy) {
  double total = 0.0;
  int off = 0;
  for (int c = 0; c < r->ncomp; c++) {
    int n = r->comp_sizes[c];
    ksum k = {0, 0};
    for (int i = 0; i < n; i++) {
      int j = (i + 1) % n;
      double ax = r->xy[2 * (off + i)] - ox, ay = r->xy[2 * (off + i) + 1] - oy;
      double bx = r->xy[2 * (off + j)] - ox, by = r->xy[2 * (off + j) + 1] - oy;
      kadd(&k, 0.5 * (ax * by - bx * ay));
    }
    total += fabs(k.s + k.c);
    off += n;
  }
  return total;
}

/* debugging aid: the augmented lists (after phase 1 and labelling step 1/3)
 * of either route.  out rows: x, y, inter, label, neighbour position. */
static int list_pos(const vpool *pl, int root, int V) {
  int k = 0, it = root;
  do { if (it == V) return k; k++; it = pl->v[it].next; } while (it != root);
  return -1;
}
int orc_debug_lists(const double *P, int nP, const double *Q, int nQ, int normalize,
                    int route, double *outP, int capP, double *outQ, int capQ,
                    int *nOutP, int *nOutQ) {
  ghstate g;
  memset(&g, 0, sizeof(g));
  g.rootP = g.rootQ = -1;
  single_roots(&g);
  double cmbr[4], norm[3];
  if (!pair_frame(P, nP, Q, nQ, normalize, cmbr, norm)) return 2;
  if (route == 0) {
    for (int i = 0; i < nP; i++) { pt o = {P[2*i], P[2*i+1]}; int V = vnew(&g.pl, tf(o.x,o.y,norm), o, -1.0); g.pl.v[V].source = 1; poly_append(&g.pl, &g.rootP, V); }
    for (int i = 0; i < nQ; i++) { pt o = {Q[2*i], Q[2*i+1]}; int V = vnew(&g.pl, tf(o.x,o.y,norm), o, -1.0); g.pl.v[V].source = 1; poly_append(&g.pl, &g.rootQ, V); }
    int one = g.rootP;
    do {
      int two = g.pl.v[one].next; while (!g.pl.v[two].source) two = g.pl.v[two].next;
      int q1 = g.rootQ;
      do {
        int q2 = g.pl.v[q1].next; while (!g.pl.v[q2].source) q2 = g.pl.v[q2].next;
        double alpha = 0, beta = 0; vtx *v = g.pl.v;
        pt P1 = v[one].p, P2 = v[two].p, Q1 = v[q1].p, Q2 = v[q2].p;
        pt P1o = v[one].o, P2o = v[two].o, Q1o = v[q1].o;
        int t = itype(P1, P2, Q1, Q2, &alpha, &beta); int IP, IQ;
        switch (t) {
        case 1: { pt I = {(1.0-alpha)*P1.x+alpha*P2.x, (1.0-alpha)*P1.y+alpha*P2.y};
          pt Io = {(1.0-alpha)*P1o.x+alpha*P2o.x, (1.0-alpha)*P1o.y+alpha*P2o.y};
          IP = vnew(&g.pl, I, Io, alpha); IQ = vnew(&g.pl, I, Io, beta);
          insert_vertex(&g.pl, IP, one, alpha); insert_vertex(&g.pl, IQ, q1, beta); vlink(&g.pl, IP, IQ); break; }
        case 5: IQ = vnew(&g.pl, P1, P1o, beta); insert_vertex(&g.pl, IQ, q1, beta); vlink(&g.pl, one, IQ);
          IP = vnew(&g.pl, Q1, Q1o, alpha); insert_vertex(&g.pl, IP, one, alpha); vlink(&g.pl, IP, q1); break;
        case 2: case 6: IQ = vnew(&g.pl, P1, P1o, beta); insert_vertex(&g.pl, IQ, q1, beta); vlink(&g.pl, one, IQ); break;
        case 3: case 7: IP = vnew(&g.pl, Q1, Q1o, alpha); insert_vertex(&g.pl, IP, one, alpha); vlink(&g.pl, IP, q1); break;
        case 4: case 8: vlink(&g.pl, one, q1); break;
        }
        q1 = q2;
      } while (q1 != g.rootQ);
      one = two;
    } while (one != g.rootP);
  } else {
    orc_phase1_t t;
    int rc = orc_phase1(P, nP, Q, nQ, normalize, &t);
    if (rc) { orc_phase1_free(&t); return rc; }
    int *mapP = (int *)malloc(sizeof(int) * (size_t)(t.cP ? t.cP : 1)); int pe = 0, qe = 0;
    for (int s = 0; s < t.cP; s++) {
      pt p = {t.intersectionsP[2*s], t.intersectionsP[2*s+1]}, o = {t.origP[2*s], t.origP[2*s+1]};
      int V = vnew(&g.pl, p, o, -1.0);
      int parent = (pe < nP && t.psP2[pe] == s);
      if (parent) { g.pl.v[V].source = 1; g.pl.v[V].inter = t.alphaP[s] != -100; pe++; } else g.pl.v[V].inter = 1;
      mapP[s] = V; poly_append(&g.pl, &g.rootP, V);
    }
    for (int s = 0; s < t.cQ; s++) {
      pt p = {t.intersectionsQ[2*s], t.intersectionsQ[2*s+1]}, o = {t.origQ[2*s], t.origQ[2*s+1]};
      int V = vnew(&g.pl, p, o, -1.0);
      int parent = (qe < nQ && t.psQ2[qe] == s), linked;
      if (parent) { g.pl.v[V].source = 1; linked = t.alphaQ[s] != -100 || t.neighborQ[s] > 0; qe++; } else linked = 1;
      poly_append(&g.pl, &g.rootQ, V);
      if (linked) { int k = t.neighborQ[s] - 1; if (k >= 0 && k < t.cP) vlink(&g.pl, V, mapP[k]); else g.pl.v[V].inter = 1; }
    }
    free(mapP); orc_phase1_free(&t);
  }
  label_step1_and_3(&g);
  for (int side = 0; side < 2; side++) {
    int root = side == 0 ? g.rootP : g.rootQ, other = side == 0 ? g.rootQ : g.rootP;
    double *out = side == 0 ? outP : outQ; int cap = side == 0 ? capP : capQ; int k = 0, it = root;
    do {
      if (k < cap) { vtx *v = &g.pl.v[it]; out[5*k] = v->o.x; out[5*k+1] = v->o.y; out[5*k+2] = v->inter;
        out[5*k+3] = v->label; out[5*k+4] = v->nb >= 0 ? list_pos(&g.pl, other, v->nb) : -1; }
      k++; it = g.pl.v[it].next;
    } while (it != root);
    if (side == 0) *nOutP = k; else *nOutQ = k;
  }
  gh_free(&g);
  return 0;
}

/* ----------------------------------------------------------------------
 * Soundness check of the batched clip's CPU-route edge-pair prefilter (the
 * skip rule of k_count_cpu and k_partner_rescan_cpu, and the edge relevance
 * test of k_count_cpu in gcmf-gh_amd/csrc/gh.hip).  The rule is restated
 * with the same constants and evaluated on every pair of edges next to
 * intersect(); out[0] pairs, out[1] pairs skipped, out[2] pairs skipped
 * although intersect() is nonzero (must be 0), out[3] edges dropped by the
 * relevance test, out[4] dropped edges that have a nonzero pair (must be 0).
 * -------------------------------------------------------------------- */
static double edge_gap_c(pt a1, pt a2, pt b1, pt b2) {
  double gx = fmax(fmin(b1.x, b2.x) - fmax(a1.x, a2.x), fmin(a1.x, a2.x) - fmax(b1.x, b2.x));
  double gy = fmax(fmin(b1.y, b2.y) - fmax(a1.y, a2.y), fmin(a1.y, a2.y) - fmax(b1.y, b2.y));
  return fmax(gx, gy);
}

int orc_prefilter_check(const double *P, int nP, const double *Q, int nQ, int normalize,
                        long long *out) {
  double cmbr[4], norm[3];
  for (int i = 0; i < 5; i++) out[i] = 0;
  if (!pair_frame(P, nP, Q, nQ, normalize, cmbr, norm)) return 2;
  int n[2] = {nP, nQ};
  pt *r[2];
  double *L[2], *m[2], bx[2][4], minL[2], maxM[2];
  for (int s = 0; s < 2; s++) {
    const double *c = s ? Q : P;
    r[s] = (pt *)malloc(sizeof(pt) * (size_t)n[s]);
    L[s] = (double *)malloc(sizeof(double) * (size_t)n[s]);
    m[s] = (double *)malloc(sizeof(double) * (size_t)n[s]);
    bx[s][0] = bx[s][1] = INFINITY;
    bx[s][2] = bx[s][3] = -INFINITY;
    for (int i = 0; i < n[s]; i++) {
      r[s][i] = tf(c[2 * i], c[2 * i + 1], norm);
      bx[s][0] = fmin(bx[s][0], r[s][i].x);
      bx[s][1] = fmin(bx[s][1], r[s][i].y);
      bx[s][2] = fmax(bx[s][2], r[s][i].x);
      bx[s][3] = fmax(bx[s][3], r[s][i].y);
    }
  }
  double w = fmax(bx[0][2], bx[1][2]) - fmin(bx[0][0], bx[1][0]);
  double h = fmax(bx[0][3], bx[1][3]) - fmin(bx[0][1], bx[1][1]);
  double D2 = w * w + h * h, D = sqrt(D2), u = 1.1102230246251565e-16;
  double dl = 8.0 * u * D2, eA = 3e-8 * D2 + 1e-15;
  double kA = 1.5 * (GH_EPS + eA), kB = 1.5 * 2.2 * (GH_EPS + 2.0 * dl), c0 = 16.0 * u * D;
  double T = 4.5 * (GH_EPS + 2.0 * dl);
  for (int s = 0; s < 2; s++) {
    minL[s] = INFINITY;
    maxM[s] = 0.0;
    for (int i = 0; i < n[s]; i++) {
      pt a = r[s][i], b = r[s][i + 1 == n[s] ? 0 : i + 1];
      double dx = b.x - a.x, dy = b.y - a.y;
      L[s][i] = sqrt(dx * dx + dy * dy);
      m[s][i] = s ? kA * L[s][i] + kB / L[s][i] + c0 : kA * L[s][i] + c0;
      minL[s] = fmin(minL[s], L[s][i]);
      maxM[s] = fmax(maxM[s], m[s][i]);
    }
  }
  char *hasP = (char *)calloc((size_t)nP, 1), *hasQ = (char *)calloc((size_t)nQ, 1);
  for (int p = 0; p < nP; p++) {
    pt P1 = r[0][p], P2 = r[0][p + 1 == nP ? 0 : p + 1];
    for (int q = 0; q < nQ; q++) {
      pt Q1 = r[1][q], Q2 = r[1][q + 1 == nQ ? 0 : q + 1];
      double gap = edge_gap_c(P1, P2, Q1, Q2);
      int skip = gap > 0.0 && gap > m[0][p] + m[1][q] &&
                 (!(L[0][p] * L[1][q] < T) || !(fabs(triA(P1, Q1, Q2)) < GH_EPS));
      double al, be;
      int t = itype(P1, P2, Q1, Q2, &al, &be);
      out[0]++;
      if (skip) out[1]++;
      if (skip && t) out[2]++;
      if (t) hasP[p] = hasQ[q] = 1;
    }
  }
  for (int s = 0; s < 2; s++) {
    const double *ob = bx[1 - s];
    for (int i = 0; i < n[s]; i++) {
      pt a = r[s][i], b = r[s][i + 1 == n[s] ? 0 : i + 1];
      double x0 = fmin(a.x, b.x), y0 = fmin(a.y, b.y), x1 = fmax(a.x, b.x), y1 = fmax(a.y, b.y);
      double ogap = fmax(fmax(ob[0] - x1, x0 - ob[2]), fmax(ob[1] - y1, y0 - ob[3]));
      int tiny = L[s][i] * minL[1 - s] < T;
      double G = m[s][i] + maxM[1 - s];
      int flag = tiny || !(ogap > G);
      if (!flag) {
        out[3]++;
        if ((s ? hasQ : hasP)[i]) out[4]++;
      }
    }
  }
  for (int s = 0; s < 2; s++) {
    free(r[s]);
    free(L[s]);
    free(m[s]);
  }
  free(hasP);
  free(hasQ);
  return 0;
}
