// gh_complex.inl -- Greiner-Hormann clipping of complex polygons (several
// rings per side: outer rings, holes, islands) and their union, on gfx950.
// Included at the end of gh.hip (same translation unit: it shares the node
// lists, the output buffers and the predicates).
//
// The reference clips complex polygons with its multi-polygon program
// (GH-CUDA/ghcuda-multiPolygon.cpp:259-330 with
// gppolyclip_mDIV512-multiPolygon.cu for phase 1) and the CPU labelling of
// GH-CUDA/lib/polyclip.h, whose loops run over every component of P and Q
// (vector<polygon>) and whose global UNION switch (polyclip.h:42) turns the
// intersection into the union.
//
// Here one wave takes one pair of complex polygons:
//   1. the lanes build the node lists (every ring its own cycle), classify
//      every source edge of every ring of P against every source edge of
//      every ring of Q (intersect(), polyclip.h:60-179) and write the
//      non-zero records, each P edge's in Q order, at scanned offsets;
//   2. lane 0 replays the records in the reference's loop order (P ring, P
//      edge, Q ring, Q edge; polyclip.h:200-277), inserting vertices by alpha
//      with the reference's tie rule (polygon.h:101-113), so the lists are
//      the reference's;
//   3. lane 0 labels and traces over all components (label_trace_multi) and
//      cleans the result up (polyclip.h:791-807).
// Pairs of this kind are few and small next to the single-ring batches, so
// the sequential part runs on one lane per pair.

constexpr int kComplexW = 64;

// ---------------------------------------------------- multi-ring labelling
__device__ bool node_pip_comps(const Nodes &N, const int *roots, int n, pt R) {
  bool in = false;  // even-odd over the components (polyclip.h:505-507)
  for (int c = 0; c < n; c++)
    if (node_pip(N, roots[c], R)) in = !in;
  return in;
}

__device__ bool node_on_ring(const Nodes &N, int root, int V) {
  int W = root;
  do {
    if (W == V) return true;
    W = N.next[W];
  } while (W != root);
  return false;
}

// polygon::noCrossingVertex(union_case), polygon.h:465-475
__device__ bool node_no_crossing(const Nodes &N, int root, bool U) {
  int V = root;
  do {
    if (isint(N, V)) {
      const int l = N.lab[V];
      if (l == L_CROSSING || l == L_DELAYED_CROSSING) return false;
      if (U && (l == L_BOUNCING || l == L_DELAYED_BOUNCING)) return false;
    }
    V = N.next[V];
  } while (V != root);
  return true;
}

__device__ bool node_all_on_on(const Nodes &N, int root) {
  int V = root;
  do {
    if (N.lab[V] != L_ON_ON) return false;
    V = N.next[V];
  } while (V != root);
  return true;
}

// polygon::getNonIntersectionPoint, polygon.h:480-491
__device__ bool node_nonint_point(const Nodes &N, int root, pt &p) {
  int V = root;
  do {
    if (!isint(N, V)) {
      p = P_(N, V);
      return true;
    }
    V = N.next[V];
  } while (V != root);
  V = root;
  do {
    const int n = N.next[V], nb = N.nb[V], nnb = N.nb[n];
    if (nb >= 0 && nnb != N.prev[nb] && nnb != N.next[nb]) {
      const pt a = P_(N, V), b = P_(N, n);
      p = pt{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)};
      return true;
    }
    V = n;
  } while (V != root);
  return false;
}

// polygon::getNonIntersectionVertex, polygon.h:496-512 (a midpoint vertex
// is inserted when every vertex is an intersection); -1 none, -2 overflow
__device__ int node_nonint_vertex(Nodes &N, int root) {
  int V = root;
  do {
    if (!isint(N, V)) return V;
    V = N.next[V];
  } while (V != root);
  V = root;
  do {
    const int n = N.next[V], nb = N.nb[V], nnb = N.nb[n];
    if (nb >= 0 && nnb != N.prev[nb] && nnb != N.next[nb]) {
      const pt a = P_(N, V), b = P_(N, n), ao = ld(N.po, V), bo = ld(N.po, n);
      const int T = node_new(N, pt{0.5 * (a.x + b.x), 0.5 * (a.y + b.y)},
                             pt{0.5 * (ao.x + bo.x), 0.5 * (ao.y + bo.y)});
      if (T < 0) return -2;
      node_insert_after(N, T, V);
      return T;
    }
    V = n;
  } while (V != root);
  return -1;
}

// labelIntersections (polyclip.h:335-735, steps 1 and 3 restored),
// createResult (:743-783) over every component of P and Q, for the
// intersection or (U) the union.  roots[0] / roots[1] hold the rings' start
// nodes; the result components go to O in the reference's order.
__device__ int label_trace_multi(Nodes &N, const int *const *roots, const int *nr, OutBuf &O,
                                 bool U) {
  // step 1 (restored, :343-397) and step 3 (:461-465) over P's components
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
    int I = r;
    do {
      if (isint(N, I) && N.nb[I] >= 0) {
        const int Pm = N.prev[I], Pp = N.next[I], Qm = N.prev[N.nb[I]], Qp = N.next[N.nb[I]];
        const int tm = node_rel(N, Qm, Pm, I, Pp), tp = node_rel(N, Qp, Pm, I, Pp);
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
      I = N.next[I];
    } while (I != r);
  }
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
    int I = r;
    do {
      if (isint(N, I) && N.nb[I] >= 0) N.lab[N.nb[I]] = N.lab[I];
      I = N.next[I];
    } while (I != r);
  }
  // step 2: intersection chains (:398-451)
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
    int it = r;
    do {
      if (isint(N, it) && (N.lab[it] == L_LEFT_ON || N.lab[it] == L_RIGHT_ON)) {
        int I = it;
        const int x = N.lab[I] == L_LEFT_ON ? 0 : 1;
        const int X = I;
        do {
          N.lab[I] = L_NONE;
          if (N.nb[I] >= 0) N.lab[N.nb[I]] = L_NONE;
          I = N.next[I];
        } while (N.lab[I] == L_ON_ON);
        const int y = N.lab[I] == L_ON_LEFT ? 0 : 1;
        const signed char ct = (x != y) ? L_DELAYED_CROSSING : L_DELAYED_BOUNCING;
        N.lab[X] = ct;
        N.lab[I] = ct;
        if (N.nb[X] >= 0) N.lab[N.nb[X]] = ct;
        if (N.nb[I] >= 0) N.lab[N.nb[I]] = ct;
      }
      it = N.next[it];
    } while (it != r);
  }
  // step 3.5: components without crossings (:474-543); the per-component
  // flags are kept in the split bit of the root node (free until step 4)
  // as: bit F_SPLIT set = no intersection, bit F_CROSS set = identical
  for (int i = 0; i < 2; i++) {
    for (int c = 0; c < nr[i]; c++) {
      const int Pr = roots[i][c];
      if (!node_no_crossing(N, Pr, U)) continue;
      N.fl[Pr] |= F_SPLIT;
      if (node_all_on_on(N, Pr)) {
        N.fl[Pr] |= F_CROSS;
        continue;
      }
      pt p;
      if (!node_nonint_point(N, Pr, p)) return GH_TRACE_FAIL;
      if (node_pip_comps(N, roots[1 - i], nr[1 - i], p) != U)
        if (!out_copy_ring(O, N, Pr)) return GH_OVERFLOW;
    }
  }
  // components of P identical to one of Q: added when both or neither are
  // holes (:517-541)
  for (int c = 0; c < nr[0]; c++) {
    const int Pr = roots[0][c];
    if (!(N.fl[Pr] & F_SPLIT) || !(N.fl[Pr] & F_CROSS)) continue;
    bool pHole = false;
    for (int c2 = 0; c2 < nr[0]; c2++)
      if (roots[0][c2] != Pr && node_pip(N, roots[0][c2], P_(N, Pr))) pHole = !pHole;
    for (int q = 0; q < nr[1]; q++) {
      const int Qr = roots[1][q];
      if (!(N.fl[Qr] & F_SPLIT) || !(N.fl[Qr] & F_CROSS) || !node_on_ring(N, Qr, N.nb[Pr]))
        continue;
      bool qHole = false;
      for (int q2 = 0; q2 < nr[1]; q2++)
        if (roots[1][q2] != Qr && node_pip(N, roots[1][q2], P_(N, Qr))) qHole = !qHole;
      if (pHole == qHole && !out_copy_ring(O, N, Pr)) return GH_OVERFLOW;
      break;
    }
  }
  // the component flags back to per-node split / crossing candidates
  bool skip[2][64];  // components without intersections (at most 64 per side here)
  for (int i = 0; i < 2; i++)
    for (int c = 0; c < nr[i]; c++) {
      const int Pr = roots[i][c];
      skip[i][c] = (N.fl[Pr] & F_SPLIT) != 0;
      N.fl[Pr] &= (unsigned char)~(F_SPLIT | F_CROSS);
    }
  // step 4: entry/exit flags (:549-655)
  for (int i = 0; i < 2; i++) {
    for (int c = 0; c < nr[i]; c++) {
      if (skip[i][c]) continue;
      const int Pr = roots[i][c];
      const int V = node_nonint_vertex(N, Pr);
      if (V == -2) return GH_OVERFLOW;
      if (V < 0) return GH_TRACE_FAIL;
      int status = ENTRY;
      if (node_pip_comps(N, roots[1 - i], nr[1 - i], P_(N, V))) toggle(status);
      bool first = true;
      int I = V;
      do {
        I = N.next[I];
      } while (!isint(N, I) && I != V);
      while (I != V) {
        if (N.lab[I] == L_CROSSING) {
          N.enex[I] = (signed char)status;
          toggle(status);
        }
        if (N.lab[I] == L_BOUNCING && ((status == EXIT) != U)) N.fl[I] |= F_SPLIT;
        if (N.lab[I] == L_DELAYED_CROSSING) {
          N.enex[I] = (signed char)status;
          if (first) {
            if ((status == EXIT) != U) N.lab[I] = L_CROSSING;
            first = false;
          } else {
            if ((status == ENTRY) != U) N.lab[I] = L_CROSSING;
            first = true;
            toggle(status);
          }
        }
        if (N.lab[I] == L_DELAYED_BOUNCING) {
          N.enex[I] = (signed char)status;
          if (first) {
            if ((status == EXIT) != U) N.fl[I] |= F_CROSS;
            first = false;
          } else {
            if ((status == ENTRY) != U) N.fl[I] |= F_CROSS;
            first = true;
          }
          toggle(status);
        }
        do {
          I = N.next[I];
        } while (!isint(N, I) && I != V);
      }
    }
  }
  // step 5: split bouncing vertex pairs (:661-714), P's candidates in list
  // order (each pair is handled on its own); step 6 (:722-734)
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
    int I = r;
    do {
      const int nextI = N.next[I];
      if ((N.fl[I] & F_SPLIT) && N.nb[I] >= 0 && (N.fl[N.nb[I]] & F_SPLIT)) {
        const int IQ = N.nb[I];
        const int VP = node_new(N, P_(N, I), ld(N.po, I));
        const int VQ = node_new(N, P_(N, IQ), ld(N.po, IQ));
        if (VP < 0 || VQ < 0) return GH_OVERFLOW;
        const double sP = triA(P_(N, N.prev[I]), P_(N, I), P_(N, N.next[I]));
        const double sQ = triA(P_(N, N.prev[IQ]), P_(N, IQ), P_(N, N.next[IQ]));
        if (sP * sQ > 0) {
          node_link(N, I, VQ);
          node_link(N, IQ, VP);
        } else {
          node_link(N, VP, VQ);
        }
        node_insert_after(N, VP, I);
        node_insert_after(N, VQ, IQ);
        const signed char ex = U ? ENTRY : EXIT, en = U ? EXIT : ENTRY;
        N.enex[I] = ex;
        N.enex[VP] = en;
        N.enex[IQ] = ex;
        N.enex[VQ] = en;
        N.lab[I] = N.lab[VP] = N.lab[IQ] = N.lab[VQ] = L_CROSSING;
        N.fl[I] &= (unsigned char)~F_SPLIT;
      }
      I = nextI;
    } while (I != r);
  }
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
    int I = r;
    do {
      if ((N.fl[I] & F_CROSS) && N.nb[I] >= 0 && (N.fl[N.nb[I]] & F_CROSS)) {
        N.lab[I] = L_CROSSING;
        N.lab[N.nb[I]] = L_CROSSING;
      }
      I = N.next[I];
    } while (I != r);
  }
  // createResult (:743-783) over P's components
  const long long guard = 4LL * N.n + 64;
  for (int c = 0; c < nr[0]; c++) {
    const int r = roots[0][c];
#define IS_CI(v) (isint(N, (v)) && N.lab[(v)] == L_CROSSING)
    int I = r;
    bool have = IS_CI(I);
    if (!have) {
      do {
        I = N.next[I];
      } while (!IS_CI(I) && I != r);
      have = I != r;
    }
    while (have) {
      if (O.ncomp >= O.capComp) return GH_OVERFLOW;
      int R = -1, V = I;
      N.fl[V] &= (unsigned char)~F_INT;
      long long steps = 0;
      do {
        int status = N.enex[V];
        toggle(status);
        do {
          V = ((status == EXIT) != U) ? N.next[V] : N.prev[V];
          N.fl[V] &= (unsigned char)~F_INT;
          const int w = out_new(O, P_(N, V), ld(N.po, V));
          if (w < 0) return GH_OVERFLOW;
          if (R < 0) R = w;
          else out_append(O, R, w);
          if (++steps > guard) return GH_TRACE_FAIL;
        } while (!(N.enex[V] == status) && V != I);
        if (V != I) {
          V = N.nb[V];
          if (V < 0) return GH_TRACE_FAIL;
          N.fl[V] &= (unsigned char)~F_INT;
        }
      } while (V != I);
      O.compRoot[O.ncomp++] = R;
      do {
        I = N.next[I];
      } while (!IS_CI(I) && I != r);
      have = I != r;
    }
#undef IS_CI
  }
  return GH_OK;
}

// ------------------------------------------------------------- kernels ----
// The pair's parts: base rings [b0, b0 + bn), overlay rings [o0, o0 + on).
struct ComplexPair {
  int b0, bn, o0, on;
};

__device__ __forceinline__ ComplexPair complex_pair(const int *parts, long k) {
  return ComplexPair{parts[4 * k], parts[4 * k + 1], parts[4 * k + 2], parts[4 * k + 3]};
}

// ring r of a side: its first vertex in the layer and its length without the
// closing vertex
__device__ __forceinline__ void ring_of(const int *vnum, const long *vpsnum, int r, long &s,
                                        int &n) {
  n = vnum[r] - 1;
  s = vpsnum[r] - vnum[r];
}

// The pair's predicate frame (ComplexArgs::norm): the MBRs over every ring,
// the CMBR, then as the single-ring clip (k_pair_setup) with the vertex
// totals; a union of polygons with disjoint MBRs uses their joint MBR.
// Status GH_DISJOINT for an intersection of disjoint MBRs (or an empty side).
struct ComplexArgs {
  const int *parts;
  long B, O;  // rings in the base and overlay layers
  const double *bCoords, *oCoords;
  const int *bVNum, *oVNum;
  const long *bVPSNum, *oVPSNum;
  int normalize, unionOp;
  int *status;
  double *norm;          // [3M]
  long long *nodeOff;    // [M+1] node / output capacity prefix
  long long *compOff;    // [M+1] component capacity prefix
  long long *recOff;     // [M+1] record capacity prefix
  long long *ringOff;    // [M+1] ring-root prefix
  long long *edgeOff;    // [M+1] P edge prefix (record offsets per P edge)
  unsigned long long *bad;  // pairs whose parts lie outside the layers
};

__global__ void __launch_bounds__(kComplexW)
    k_complex_setup(long M, ComplexArgs a, int *nodeCap, int *compCap, int *rings,
                    int *edgesP) {
  const long k = blockIdx.x;
  if (k >= M) return;
  const int lane = threadIdx.x;
  const ComplexPair cp = complex_pair(a.parts, k);
  // parts outside the layers: counted and nothing read (the call fails)
  if (cp.b0 < 0 || cp.bn < 0 || cp.o0 < 0 || cp.on < 0 || (long)cp.b0 + cp.bn > a.B ||
      (long)cp.o0 + cp.on > a.O) {
    if (lane == 0) {
      a.status[k] = GH_DISJOINT;
      atomicAdd(a.bad, 1ULL);
      rings[k] = edgesP[k] = nodeCap[k] = compCap[k] = 0;
    }
    return;
  }
  double m[2][4];
  long long nv[2] = {0, 0};
  bool emptyRing = false;
  for (int side = 0; side < 2; side++) {
    const int r0 = side ? cp.o0 : cp.b0, rn = side ? cp.on : cp.bn;
    const double *C = side ? a.oCoords : a.bCoords;
    const int *vn = side ? a.oVNum : a.bVNum;
    const long *vp = side ? a.oVPSNum : a.bVPSNum;
    double x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
    for (int r = r0; r < r0 + rn; r++) {
      long s;
      int n;
      ring_of(vn, vp, r, s, n);
      emptyRing |= n < 1;
      nv[side] += n > 0 ? n : 0;
      for (int i = lane; i < n; i += kComplexW) {
        const double x = C[2 * (s + i)], y = C[2 * (s + i) + 1];
        x0 = fmin(x0, x);
        y0 = fmin(y0, y);
        x1 = fmax(x1, x);
        y1 = fmax(y1, y);
      }
    }
    pgrp_minmax<kComplexW>(x0, x1);
    pgrp_minmax<kComplexW>(y0, y1);
    m[side][0] = x0;
    m[side][1] = y0;
    m[side][2] = x1;
    m[side][3] = y1;
  }
  if (lane != 0) return;
  int st = GH_OK;
  const bool apart = m[0][0] > m[1][2] || m[0][2] < m[1][0] || m[0][1] > m[1][3] ||
                     m[0][3] < m[1][1];
  if (nv[0] < 1 || nv[1] < 1 || cp.bn > 64 || cp.on > 64 || emptyRing) st = GH_DISJOINT;
  if (apart && !a.unionOp) st = GH_DISJOINT;
  double cx = 0.0, cy = 0.0, s = 1.0;
  if (a.normalize & 1) {
    double b[4];
    if (apart) {  // the joint MBR (a union only)
      b[0] = fmin(m[0][0], m[1][0]);
      b[1] = fmin(m[0][1], m[1][1]);
      b[2] = fmax(m[0][2], m[1][2]);
      b[3] = fmax(m[0][3], m[1][3]);
    } else {  // the CMBR, as getCMBR (ghcuda.cpp:94-125)
      b[0] = m[0][0] < m[1][0] ? m[1][0] : m[0][0];
      b[1] = m[0][1] < m[1][1] ? m[1][1] : m[0][1];
      b[2] = m[1][2] < m[0][2] ? m[1][2] : m[0][2];
      b[3] = m[1][3] < m[0][3] ? m[1][3] : m[0][3];
    }
    const double w = b[2] - b[0], h = b[3] - b[1], e = w > h ? w : h;
    cx = b[0];
    cy = b[1];
    s = e > 0.0 ? 1.0 / e : 1.0;
    if (a.normalize & 4) {
      const double rr = sqrt((double)nv[0] * (double)nv[1]) / 1024.0;
      double f = 1.0;
      while (f < rr) f *= 2.0;
      s *= f;
    }
  }
  a.status[k] = st;
  a.norm[3 * k] = cx;
  a.norm[3 * k + 1] = cy;
  a.norm[3 * k + 2] = s;
  rings[k] = st == GH_OK ? cp.bn + cp.on : 0;
  edgesP[k] = st == GH_OK ? (int)nv[0] : 0;
  // the node and component capacities need the intersection count
  // (k_complex_count adds it)
  nodeCap[k] = st == GH_OK ? (int)(nv[0] + nv[1]) + 4 * (cp.bn + cp.on) + 16 : 0;
  compCap[k] = st == GH_OK ? cp.bn + cp.on + 8 : 0;
}

// the side's vertex i (over its rings in order) in the pair's frame and
// original coordinates; sEdge lists the rings' first global vertex and
// ring lengths are read from VNum
struct RingWalk {
  const double *C;
  const int *vn;
  const long *vp;
  int r0, rn;
};

// records of P edge e (global over P's rings) against every Q edge, in Q
// order; f(qEdge, type, alpha, beta) for the non-zero types
template <class F>
__device__ void complex_edge_records(const RingWalk &P, int pr, int pi, const RingWalk &Q,
                                     const double *nrm, F &&f) {
  long ps;
  int pn;
  ring_of(P.vn, P.vp, pr, ps, pn);
  const double cx = nrm[0], cy = nrm[1], s = nrm[2];
  auto tf = [&](double x, double y) { return pt{(x - cx) * s, (y - cy) * s}; };
  const pt P1 = tf(P.C[2 * (ps + pi)], P.C[2 * (ps + pi) + 1]);
  const int pj = pi + 1 == pn ? 0 : pi + 1;
  const pt P2 = tf(P.C[2 * (ps + pj)], P.C[2 * (ps + pj) + 1]);
  int qe = 0;
  for (int qr = Q.r0; qr < Q.r0 + Q.rn; qr++) {
    long qs;
    int qn;
    ring_of(Q.vn, Q.vp, qr, qs, qn);
    for (int qi = 0; qi < qn; qi++, qe++) {
      const int qj = qi + 1 == qn ? 0 : qi + 1;
      const pt Q1 = tf(Q.C[2 * (qs + qi)], Q.C[2 * (qs + qi) + 1]);
      const pt Q2 = tf(Q.C[2 * (qs + qj)], Q.C[2 * (qs + qj) + 1]);
      double al, be;
      const int t = itype(P1, P2, Q1, Q2, al, be);
      if (t != 0) f(qe, t, al, be);
    }
  }
}

__device__ __forceinline__ void complex_walks(const ComplexArgs &a, const ComplexPair &cp,
                                              RingWalk &P, RingWalk &Q) {
  P = RingWalk{a.bCoords, a.bVNum, a.bVPSNum, cp.b0, cp.bn};
  Q = RingWalk{a.oCoords, a.oVNum, a.oVPSNum, cp.o0, cp.on};
}

// intersection records per pair (for the capacities)
__global__ void __launch_bounds__(kComplexW)
    k_complex_count(long M, ComplexArgs a, int *recCnt, int *nodeCap, int *compCap) {
  const long k = blockIdx.x;
  if (k >= M) return;
  if (a.status[k] != GH_OK) {
    if (threadIdx.x == 0) recCnt[k] = 0;
    return;
  }
  const ComplexPair cp = complex_pair(a.parts, k);
  RingWalk P, Q;
  complex_walks(a, cp, P, Q);
  int cnt = 0;
  for (int pr = P.r0; pr < P.r0 + P.rn; pr++) {
    const int pn = P.vn[pr] - 1;
    for (int pi = threadIdx.x; pi < pn; pi += kComplexW)
      complex_edge_records(P, pr, pi, Q, a.norm + 3 * k,
                           [&](int, int, double, double) { cnt++; });
  }
  cnt = grp_sum<kComplexW>(cnt);
  if (threadIdx.x == 0) {
    recCnt[k] = cnt;
    // two nodes per record, two more per split pair, and the result
    nodeCap[k] += 4 * cnt;
    compCap[k] += cnt;
  }
}

struct ComplexWork {
  double *npn, *npo, *nalpha;
  int *nnext, *nprev, *nnb;
  unsigned char *nfl;
  signed char *nlab, *nenex;
  double *opn, *opo;
  int *onext, *oprev, *ocompRoot;
  int *roots;             // [ring total] ring start nodes
  int *recQ, *recT;       // [records] Q edge and type
  double *recA, *recB;    // [records] alpha, beta
  int *edgeRec;           // [P edges + M] per-P-edge record offsets within the pair
  int *outNComp, *outNVerts, *compLen;
  double *xyTmp;
  int cleanup;
};

__global__ void __launch_bounds__(kComplexW)
    k_complex_clip(long M, ComplexArgs a, ComplexWork w) {
  const long k = blockIdx.x;
  if (k >= M) return;
  const int lane = threadIdx.x;
  if (a.status[k] != GH_OK) {
    if (lane == 0) {
      w.outNComp[k] = 0;
      w.outNVerts[k] = 0;
    }
    return;
  }
  const ComplexPair cp = complex_pair(a.parts, k);
  RingWalk P, Q;
  complex_walks(a, cp, P, Q);
  const long long nb0 = a.nodeOff[k], cap = a.nodeOff[k + 1] - nb0;
  Nodes N;
  N.pn = w.npn + 2 * nb0;
  N.po = w.npo + 2 * nb0;
  N.next = w.nnext + nb0;
  N.prev = w.nprev + nb0;
  N.nb = w.nnb + nb0;
  N.fl = w.nfl + nb0;
  N.lab = w.nlab + nb0;
  N.enex = w.nenex + nb0;
  N.cap = (int)cap;
  double *alpha = w.nalpha + nb0;
  int *roots = w.roots + a.ringOff[k];
  const double *nrm = a.norm + 3 * k;
  // -- source nodes, ring by ring (P's rings, then Q's)
  int base = 0;
  for (int side = 0; side < 2; side++) {
    const RingWalk &R = side ? Q : P;
    for (int r = R.r0; r < R.r0 + R.rn; r++) {
      long s;
      int n;
      ring_of(R.vn, R.vp, r, s, n);
      for (int i = lane; i < n; i += kComplexW) {
        const double x = R.C[2 * (s + i)], y = R.C[2 * (s + i) + 1];
        const int v = base + i;
        st(N.pn, v, pt{(x - nrm[0]) * nrm[2], (y - nrm[1]) * nrm[2]});
        st(N.po, v, pt{x, y});
        N.next[v] = base + (i + 1 == n ? 0 : i + 1);
        N.prev[v] = base + (i == 0 ? n - 1 : i - 1);
        N.nb[v] = -1;
        N.fl[v] = F_SRC;
        N.lab[v] = L_NONE;
        N.enex[v] = NEITHER;
        alpha[v] = -2.0;
      }
      if (lane == 0) roots[(side ? cp.bn : 0) + (r - R.r0)] = base;
      base += n;
    }
  }
  const int nSrcP = (int)(a.edgeOff[k + 1] - a.edgeOff[k]);
  N.n = base;
  // -- records of every P edge, in Q order, at the edge's scanned offset
  int *er = w.edgeRec + a.edgeOff[k] + k;  // nSrcP + 1 entries
  const long long rb = a.recOff[k];
  {
    int e0 = 0;
    for (int pr = P.r0; pr < P.r0 + P.rn; pr++) {
      const int pn = P.vn[pr] - 1;
      for (int pi = lane; pi < pn; pi += kComplexW) {
        int cnt = 0;
        complex_edge_records(P, pr, pi, Q, nrm, [&](int, int, double, double) { cnt++; });
        er[e0 + pi] = cnt;
      }
      e0 += pn;
    }
    __syncthreads();
    if (lane == 0) {  // exclusive scan of the per-edge counts
      int run = 0;
      for (int e = 0; e < nSrcP; e++) {
        const int c = er[e];
        er[e] = run;
        run += c;
      }
      er[nSrcP] = run;
    }
    __syncthreads();
    e0 = 0;
    for (int pr = P.r0; pr < P.r0 + P.rn; pr++) {
      const int pn = P.vn[pr] - 1;
      for (int pi = lane; pi < pn; pi += kComplexW) {
        long long o = rb + er[e0 + pi];
        complex_edge_records(P, pr, pi, Q, nrm, [&](int qe, int t, double al, double be) {
          w.recQ[o] = qe;
          w.recT[o] = t;
          w.recA[o] = al;
          w.recB[o] = be;
          o++;
        });
      }
      e0 += pn;
    }
    __syncthreads();
  }
  if (lane != 0) return;
  // -- the insertions in the reference's loop order (polyclip.h:200-277)
  const int nSrcQ = base - nSrcP;
  (void)nSrcQ;
  auto insert = [&](int V, int curr, double al) {  // insertVertex, polygon.h:101-113
    do {
      curr = N.next[curr];
    } while (!(N.fl[curr] & F_SRC) && alpha[curr] < al);
    N.next[N.prev[curr]] = V;
    N.prev[V] = N.prev[curr];
    N.next[V] = curr;
    N.prev[curr] = V;
  };
  auto fresh = [&](pt p, pt o, double al) {
    const int v = node_new(N, p, o);
    if (v >= 0) alpha[v] = al;
    return v;
  };
  int rc = GH_OK;
  for (int e = 0; e < nSrcP && rc == GH_OK; e++) {
    const int one = e;  // P's source nodes are 0 .. nSrcP - 1 in edge order
    int two = N.next[one];
    while (!(N.fl[two] & F_SRC)) two = N.next[two];
    const pt P1 = P_(N, one), P2 = P_(N, two), P1o = ld(N.po, one), P2o = ld(N.po, two);
    for (int j = er[e]; j < er[e + 1]; j++) {
      const long long o = rb + j;
      const int q1 = nSrcP + w.recQ[o], t = w.recT[o];
      const double al = w.recA[o], be = w.recB[o];
      const pt Q1 = P_(N, q1), Q1o = ld(N.po, q1);
      int IP, IQ;
      switch (t) {
        case 1: {
          const pt I = {(1.0 - al) * P1.x + al * P2.x, (1.0 - al) * P1.y + al * P2.y};
          const pt Io = {(1.0 - al) * P1o.x + al * P2o.x, (1.0 - al) * P1o.y + al * P2o.y};
          IP = fresh(I, Io, al);
          IQ = fresh(I, Io, be);
          if (IP < 0 || IQ < 0) {
            rc = GH_OVERFLOW;
            break;
          }
          insert(IP, one, al);
          insert(IQ, q1, be);
          node_link(N, IP, IQ);
          break;
        }
        case 5:
          IQ = fresh(P1, P1o, be);
          if (IQ < 0) {
            rc = GH_OVERFLOW;
            break;
          }
          insert(IQ, q1, be);
          node_link(N, one, IQ);
          IP = fresh(Q1, Q1o, al);
          if (IP < 0) {
            rc = GH_OVERFLOW;
            break;
          }
          insert(IP, one, al);
          node_link(N, IP, q1);
          break;
        case 2:
        case 6:
          IQ = fresh(P1, P1o, be);
          if (IQ < 0) {
            rc = GH_OVERFLOW;
            break;
          }
          insert(IQ, q1, be);
          node_link(N, one, IQ);
          break;
        case 3:
        case 7:
          IP = fresh(Q1, Q1o, al);
          if (IP < 0) {
            rc = GH_OVERFLOW;
            break;
          }
          insert(IP, one, al);
          node_link(N, IP, q1);
          break;
        default:  // 4, 8
          node_link(N, one, q1);
          break;
      }
      if (rc != GH_OK) break;
    }
  }
  // -- labelling, tracing, clean-up
  OutBuf O;
  O.pn = w.opn + 2 * nb0;
  O.po = w.opo + 2 * nb0;
  O.next = w.onext + nb0;
  O.prev = w.oprev + nb0;
  O.n = 0;
  O.cap = (int)cap;
  O.compRoot = w.ocompRoot + a.compOff[k];
  O.ncomp = 0;
  O.capComp = (int)(a.compOff[k + 1] - a.compOff[k]);
  if (rc == GH_OK) {
    const int *rts[2] = {roots, roots + cp.bn};
    const int nrs[2] = {cp.bn, cp.on};
    rc = label_trace_multi(N, rts, nrs, O, a.unionOp != 0);
  }
  if (rc != GH_OK) {
    a.status[k] = rc;
    w.outNComp[k] = 0;
    w.outNVerts[k] = 0;
    return;
  }
  double *xy = w.xyTmp + 2 * nb0;
  int *clen = w.compLen + a.compOff[k];
  int nc = 0, nv = 0;
  for (int c = 0; c < O.ncomp; c++) {
    int r = O.compRoot[c];
    if (w.cleanup) r = out_cleanup(O, r);
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
  w.outNComp[k] = nc;
  w.outNVerts[k] = nv;
}

// ----------------------------------------------------------------- host ----
void clip_complex(Context &c, const JoinInputs &in, const int *parts, long M,
                  const ClipOptions &opt, bool unionOp, ClipResult &res) {
  res = ClipResult();
  res.M = M;
  if (M <= 0) return;
  ComplexArgs a;
  a.parts = parts;
  a.bCoords = in.bCoords;
  a.oCoords = in.oCoords;
  a.bVNum = in.bVNum;
  a.oVNum = in.oVNum;
  a.bVPSNum = in.bVPSNum;
  a.oVPSNum = in.oVPSNum;
  a.normalize = opt.normalize;
  a.unionOp = unionOp ? 1 : 0;
  a.B = in.B;
  a.O = in.O;
  a.bad = wsT<unsigned long long>(c, "cx_bad", 1);
  fill_i32(c, (int *)a.bad, 0, 2);
  a.status = wsT<int>(c, "cx_status", M);
  a.norm = wsT<double>(c, "cx_norm", 3 * M);
  int *nodeCap = wsT<int>(c, "cx_nodeCap", M), *compCap = wsT<int>(c, "cx_compCap", M);
  int *rings = wsT<int>(c, "cx_rings", M), *edgesP = wsT<int>(c, "cx_edgesP", M);
  int *recCnt = wsT<int>(c, "cx_recCnt", M);
  a.nodeOff = wsT<long long>(c, "cx_nodeOff", M + 1);
  a.compOff = wsT<long long>(c, "cx_compOff", M + 1);
  a.recOff = wsT<long long>(c, "cx_recOff", M + 1);
  a.ringOff = wsT<long long>(c, "cx_ringOff", M + 1);
  a.edgeOff = wsT<long long>(c, "cx_edgeOff", M + 1);
  hipLaunchKernelGGL(k_complex_setup, dim3((unsigned)M), dim3(kComplexW), 0, c.stream, M, a,
                     nodeCap, compCap, rings, edgesP);
  HIP_CHECK(hipGetLastError());
  {
    LaunchTimer t(c, "gh_complex_count");
    hipLaunchKernelGGL(k_complex_count, dim3((unsigned)M), dim3(kComplexW), 0, c.stream, M, a,
                       recCnt, nodeCap, compCap);
    HIP_CHECK(hipGetLastError());
  }
  scan_i32_to_i64(c, nodeCap, a.nodeOff, M);
  scan_i32_to_i64(c, compCap, a.compOff, M);
  scan_i32_to_i64(c, recCnt, a.recOff, M);
  scan_i32_to_i64(c, rings, a.ringOff, M);
  scan_i32_to_i64(c, edgesP, a.edgeOff, M);
  long long tot[6];
  read_scalars(c, {a.nodeOff + M, a.compOff + M, a.recOff + M, a.ringOff + M, a.edgeOff + M,
                   (const long long *)a.bad},
               tot);
  if (tot[5] != 0)
    throw Error{GCMF_ERR_ARG, "complex clip: a pair's rings lie outside its layer"};
  const long long nodes = tot[0], comps = tot[1], recs = tot[2], nrings = tot[3],
                  nedges = tot[4];
  if (nodes >= 0x7fffffffLL || recs >= 0x7fffffffLL)
    throw Error{GCMF_ERR_OVERFLOW, "complex clip batch exceeds 2^31 nodes or records"};
  ComplexWork w;
  w.npn = wsT<double>(c, "cx_npn", 2 * nodes);
  w.npo = wsT<double>(c, "cx_npo", 2 * nodes);
  w.nalpha = wsT<double>(c, "cx_nalpha", nodes);
  w.nnext = wsT<int>(c, "cx_nnext", nodes);
  w.nprev = wsT<int>(c, "cx_nprev", nodes);
  w.nnb = wsT<int>(c, "cx_nnb", nodes);
  w.nfl = wsT<unsigned char>(c, "cx_nfl", nodes);
  w.nlab = wsT<signed char>(c, "cx_nlab", nodes);
  w.nenex = wsT<signed char>(c, "cx_nenex", nodes);
  w.opn = wsT<double>(c, "cx_opn", 2 * nodes);
  w.opo = wsT<double>(c, "cx_opo", 2 * nodes);
  w.onext = wsT<int>(c, "cx_onext", nodes);
  w.oprev = wsT<int>(c, "cx_oprev", nodes);
  w.ocompRoot = wsT<int>(c, "cx_ocompRoot", comps);
  w.roots = wsT<int>(c, "cx_roots", nrings);
  w.recQ = wsT<int>(c, "cx_recQ", recs);
  w.recT = wsT<int>(c, "cx_recT", recs);
  w.recA = wsT<double>(c, "cx_recA", recs);
  w.recB = wsT<double>(c, "cx_recB", recs);
  w.edgeRec = wsT<int>(c, "cx_edgeRec", nedges + M);
  w.outNComp = wsT<int>(c, "cx_outNComp", M);
  w.outNVerts = wsT<int>(c, "cx_outNVerts", M);
  w.compLen = wsT<int>(c, "cx_compLen", comps);
  w.xyTmp = wsT<double>(c, "cx_xyTmp", 2 * nodes);
  w.cleanup = opt.cleanup;
  {
    LaunchTimer t(c, "gh_complex_clip");
    hipLaunchKernelGGL(k_complex_clip, dim3((unsigned)M), dim3(kComplexW), 0, c.stream, M, a,
                       w);
    HIP_CHECK(hipGetLastError());
  }
  res.vertOff = wsT<long long>(c, "cx_vOff", M + 1);
  res.compOff = wsT<long long>(c, "cx_cOff", M + 1);
  scan_i32_to_i64(c, w.outNVerts, res.vertOff, M);
  scan_i32_to_i64(c, w.outNComp, res.compOff, M);
  res.xy = wsT<double>(c, "cx_out_xy", 2 * nodes);
  res.compSizes = wsT<int>(c, "cx_out_comp", comps);
  hipLaunchKernelGGL(k_gather_out<64>, dim3(pgrp_grid(M, 64)), dim3(pgrp_block(64)), 0, c.stream,
                     M, 1, (const long long *)a.nodeOff, (const long long *)a.compOff,
                     (const long long *)res.vertOff, (const long long *)res.compOff, w.xyTmp,
                     w.compLen, res.xy, res.compSizes);
  HIP_CHECK(hipGetLastError());
  res.status = a.status;
  res.swapped = wsT<int>(c, "cx_swapped", M);
  fill_i32(c, res.swapped, 0, M);
  res.ncomp = w.outNComp;
  res.nverts = w.outNVerts;
  unsigned long long *sc = wsT<unsigned long long>(c, "cx_status_counts", 4);
  fill_i32(c, (int *)sc, 0, 8);
  hipLaunchKernelGGL(k_status_counts, dim3(std::min<unsigned>(kStatBlocks, grid_for(M, 256))),
                     dim3(256), 0, c.stream, M, (const int *)a.status, sc);
  HIP_CHECK(hipGetLastError());
  long long rb[6];
  const long long *sl = (const long long *)sc;
  read_scalars(c, {res.vertOff + M, res.compOff + M, sl, sl + 1, sl + 2, sl + 3}, rb);
  res.totalVerts = (long)rb[0];
  res.totalComps = (long)rb[1];
  res.nOk = rb[2];
  res.nInvalid = rb[3];
  res.nDisjoint = rb[4];
  res.nFailed = rb[5];
  res.totalIntersections = recs;
}
