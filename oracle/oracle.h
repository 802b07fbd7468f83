/*
 * oracle.h -- CPU restatement of the GCMF-GH join-and-clip path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or the timed CPU baseline.
 *
 * Every function states the reference file:line it restates.  All arithmetic
 * is IEEE binary64 compiled with -ffp-contract=off so that each predicate is
 * evaluated with exactly the operation order of the reference source.
 */
#ifndef GCMF_ORACLE_H
#define GCMF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- GH ---- */

/* Result of one clipped pair: non-empty components only, vertices in
 * original coordinates, interleaved x,y. */
typedef struct {
  int status;        /* 0 ok, 1 invalid pair (views disagree), 2 MBRs disjoint,
                        -1 traversal did not close (malformed labels) */
  int ncomp;         /* number of non-empty result components */
  int nverts;        /* total vertices over all components */
  int *comp_sizes;   /* [ncomp] */
  double *xy;        /* [2*nverts] */
  /* statistics */
  long long n_eval;  /* edge pairs classified by the intersection test */
  long long n_inter; /* intersection records (count1 total) */
} orc_result_t;

void orc_result_free(orc_result_t *r);

/* Reference GPU phase-1 outputs (gppolyclip_mDIV512.cu:1298-1715) for one
 * pair, sizes as the reference: psX1/psX2 have n+1 entries and hold the
 * exclusive scans (last entry = total). */
typedef struct {
  int nP, nQ, invalid, cP, cQ;
  int *psP1, *psP2, *psQ1, *psQ2;
  int *boolP, *boolQ;                 /* CMBR edge flags */
  long long *neighborMapQ;            /* [cQ] */
  double *intersectionsP, *intersectionsQ; /* [2cP], [2cQ] predicate coords */
  double *origP, *origQ;              /* [2cP], [2cQ] original coords */
  int *alphaP, *alphaQ;               /* [cP], [cQ] trunc(1e7*alpha), -100 none */
  int *initLabelsP, *initLabelsQ;     /* [cP], [cQ] */
  int *neighborP, *neighborQ;         /* [cP], [cQ] 1-based, 0 = none */
  double cmbr[4];                     /* CMBR in original coordinates */
  double norm[3];                     /* cx, cy, scale */
} orc_phase1_t;

void orc_phase1_free(orc_phase1_t *t);

/* Per-pair normalisation (see gh_oracle.c): 1 = translate to the CMBR corner
 * and scale by 1/max extent before every predicate, 0 = raw coordinates.
 * For the clip entry points the argument is a mode mask: bit 0 normalise,
 * bit 1 skip cleanUpResult (to compare traced areas with the area oracle). */
int orc_phase1(const double *P, int nP, const double *Q, int nQ, int normalize,
               orc_phase1_t *out);

/* CPU route: computeIntersections + labelling (steps 1..6) + createResult +
 * cleanUpResult, GH-CUDA/lib/polyclip.h. */
int orc_clip_cpu(const double *P, int nP, const double *Q, int nQ, int normalize,
                 orc_result_t *out);

/* CPU route over complex polygons (several rings per side, polyclip.h's
 * vector<polygon>): P packs nPR rings of pSizes[r] vertices, Q likewise.
 * mode: bit 0 normalise, bit 1 skip cleanUpResult, bit 2 edge scale,
 * bit 3 union instead of intersection (UNION, polyclip.h:42).  A union of
 * polygons with disjoint MBRs is normalised by the joint MBR. */
int orc_clip_multi(const double *P, const int *pSizes, int nPR, const double *Q,
                   const int *qSizes, int nQR, int mode, orc_result_t *out);

/* GPU route: orc_phase1 arrays spliced into lists (ghcuda.cpp:183-339), then
 * the same labelling / tracing / clean-up as the CPU route. */
int orc_clip_gpu_route(const double *P, int nP, const double *Q, int nQ,
                       int normalize, orc_result_t *out);

/* Independent area of P intersect Q: boundary integral over the pieces of
 * each boundary inside the other polygon (general position only). */
double orc_area_boundary(const double *P, int nP, const double *Q, int nQ);
/* soundness of the batched clip's CPU-route edge-pair prefilter: out[5] =
 * pairs, skipped, skipped with a nonzero type, edges dropped, dropped edges
 * with a nonzero pair; returns 2 for disjoint MBRs */
int orc_prefilter_check(const double *P, int nP, const double *Q, int nQ, int normalize,
                        long long *out);

/* Shoelace area of the components in r about the origin (ox, oy),
 * compensated summation.  Signed per component, summed as absolute values. */
double orc_result_area(const orc_result_t *r, double ox, double oy);

/* Intersection type of two edges, polyclip.h:60-179. */
int orc_intersect_type(const double *P1, const double *P2, const double *Q1,
                       const double *Q2, double *alpha, double *beta);

/* Initial label from the two relative-position types,
 * gppolyclip_mDIV512.cu:237-252. */
int orc_initial_label(int qMType, int qPType);

/* ------------------------------------------------------------- join ---- */

/* Planted mistakes (test infrastructure: the tests that show each pin of
 * the join oracle fails when the oracle is wrong).  0 = the oracle. */
extern int orc_mutant;
enum {
  ORC_MUT_SEG_OPEN = 1,        /* segment join: touching endpoints miss      */
  ORC_MUT_SEG_NO_PARALLEL = 2, /* segment join: parallel branch always 0     */
  ORC_MUT_SEG_NO_VERTICAL = 3, /* segment join: vertical collinear case lost */
  ORC_MUT_PIP_BOTH_ENDS = 4,   /* PiP: edge ending on the ray counted twice  */
  ORC_MUT_PIP_CLOSED_Y = 5,    /* PiP: y test (y0, y1] instead of [y0, y1)   */
  ORC_MUT_PIP_LE = 6,          /* PiP: pX <= crossing instead of <           */
  ORC_MUT_CMF_OPEN = 7,        /* CMF: open box test, touching edges miss    */
  ORC_MUT_CMF_NO_CLOSING = 8   /* CMF: the closing edge v[n-1] -> v[0] lost  */
};

/* CoordToMBR type 1 (SEQ_Overlay.h:69-97) on one decimal token. */
uint64_t orc_coord_to_mbr(const char *tok);

/* MBR sort-and-sweep filter (Join.h:5-153, GPU_MBR.h:339-539).  mbr is the
 * u64 key layout (x1,y1,x2,y2) per polygon.  Writes counts[B+O] and, when
 * pairs != NULL, the (base, overlay) pairs in reference order.  Returns K. */
long long orc_mbr_sweep(long long B, long long O, const uint64_t *bMBR,
                        const uint64_t *oMBR, int *counts, int *pairs);

/* Stable sort permutation and rank of the 2(B+O) x endpoints
 * (GPU_Utility.h:242-289 with dimSort=1, dimSelect=0). */
void orc_endpoint_sort(long long B, long long O, const uint64_t *bMBR,
                       const uint64_t *oMBR, int *perm, int *rank);

/* GetCMBR (Join.h:397-432, GPU_Join.h:461-508). */
void orc_get_cmbr(long long K, const int *pairs, const double *bMBR2,
                  const double *oMBR2, double *cmbr, int *pipFlag,
                  signed char *pipType, int *joinFlag);

/* PointInPolygonTest (Join.h:437-467, GPU_Join.h:839-943): updates pipFlag
 * and joinFlag in place for pairs with pipType != 0.  Returns #within. */
long long orc_pip(long long K, const int *pairs, const signed char *pipType,
                  const double *bCoords, const long long *bVPSNum,
                  const double *oCoords, const long long *oVPSNum,
                  int *pipFlag, int *joinFlag);

/* CountCMF + ApplyCMF (Join.h:157-287, GPU_Join.h:511-735).  For every pair
 * with joinFlag==1 counts / lists the edges meeting the CMBR; clears
 * joinFlag when either side has none.  bEdge/oEdge may be NULL (count only).
 * Edge lists are concatenated in pair order.  Returns E (# survivors). */
long long orc_cmf(long long K, const int *pairs, const double *cmbr,
                  const double *bCoords, const int *bVNum, const long long *bVPSNum,
                  const double *oCoords, const int *oVNum, const long long *oVPSNum,
                  int *joinFlag, int *bCount, int *oCount,
                  short *bEdge, short *oEdge);

/* Segment_Intersect_Join (Join.h:356-393, GPU_Join.h:326-456) for every
 * pair with joinFlag==1 after orc_cmf; writes segFlag[K] (0 elsewhere). */
long long orc_segment_join(long long K, const int *pairs, const int *joinFlag,
                           const double *bCoords, const long long *bVPSNum,
                           const double *oCoords, const long long *oVPSNum,
                           const int *bCount, const int *oCount,
                           const short *bEdge, const short *oEdge, int *segFlag);

#ifdef __cplusplus
}
#endif
#endif
