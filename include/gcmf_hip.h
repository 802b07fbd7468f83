/*
 * gcmf_hip.h -- C ABI of libgcmf_hip.so, the MI355X (gfx950) join-and-clip
 * path of GCMF-GH: MBR sort-and-sweep filter, CMBR/PiP refine, CMF edge
 * filter, segment-intersection refine and batched Greiner-Hormann clipping.
 *
 * Plain pointers and sizes only.  "d_" pointers are device memory (HBM),
 * everything else is host memory.  Every entry point returns 0 on success or
 * a negative status (GCMF_ERR_*); gcmf_last_error() describes the failure.
 * Calls on one context are serialised on that context's HIP stream.
 *
 * Layer layout (the reference's, SEQ_Overlay.h:285-389 / Types.h:1-8):
 *   coords  f64 [2*sum(VNum)]  AoS (x,y), each ring's closing vertex repeated
 *   vnum    i32 [npoly]        ring length including the closing vertex
 *   vpsnum  i64 [npoly]        inclusive prefix sum of vnum
 *   mbr     u64 [4*npoly]      CoordToMBR keys (minx, miny, maxx, maxy)
 *   mbr2    f64 [4*npoly]      the same as doubles
 */
#ifndef GCMF_HIP_H
#define GCMF_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GCMF_OK 0
#define GCMF_ERR_HIP -1
#define GCMF_ERR_ALLOC -2
#define GCMF_ERR_ARG -3
#define GCMF_ERR_NODEV -4
#define GCMF_ERR_OVERFLOW -5

typedef struct gcmf_ctx gcmf_ctx;

typedef struct {
  long npoly;
  const double *d_coords;
  const int *d_vnum;
  const long *d_vpsnum;
  const uint64_t *d_mbr;
  const double *d_mbr2;
} gcmf_dlayer;

/* ------------------------------------------------ context and memory ---- */
/* No reference counterpart: the reference keeps global state and the legacy
 * default stream (GPU_Manage.h).  A context owns one stream, a grow-only
 * device workspace and optional per-kernel HIP-event timers. */
int gcmf_device_count(void);
int gcmf_init(int device, gcmf_ctx **out);
void gcmf_shutdown(gcmf_ctx *ctx);
const char *gcmf_last_error(const gcmf_ctx *ctx);
int gcmf_malloc(gcmf_ctx *ctx, void **d_ptr, size_t bytes);
int gcmf_free(gcmf_ctx *ctx, void *d_ptr);
int gcmf_memcpy_h2d(gcmf_ctx *ctx, void *d_dst, const void *src, size_t bytes);
int gcmf_memcpy_d2h(gcmf_ctx *ctx, void *dst, const void *d_src, size_t bytes);
int gcmf_synchronize(gcmf_ctx *ctx);
/* Free and total device memory of the context's GPU (hipMemGetInfo) and the
 * bytes the context's grow-only workspace holds (reused by later calls);
 * any pointer may be null. */
int gcmf_mem_info(gcmf_ctx *ctx, size_t *free_bytes, size_t *total_bytes,
                  size_t *workspace_bytes);
/* Kernel timing with HIP events recorded on the context stream. */
int gcmf_set_timing(gcmf_ctx *ctx, int on);
/* Restricts timing to the launches named `kernel`, or to a comma-separated
 * list of names (NULL or "" times all), so that a few kernels can be timed
 * inside an otherwise untimed region. */
int gcmf_set_timing_only(gcmf_ctx *ctx, const char *kernel);
int gcmf_timer_reset(gcmf_ctx *ctx);
int gcmf_timer_get(gcmf_ctx *ctx, const char *kernel, double *total_ms, long *launches);

/* ------------------------------------------------------- primitives ---- */
/* Device primitives on caller-owned device buffers, ordered on the context's
 * stream and finished on return. */

/* Stable LSD radix sort of n u64 keys carrying i32 values over their low
 * `bits` bits (8 bits per pass, bits <= 64); replaces RadixSort
 * (GPU_Utility.h:242-289) with its digit-bucket kernels
 * (GPU_RadixSort.h:27-146, 16 decimal-digit passes there). */
int gcmf_radix_sort_u64(gcmf_ctx *ctx, const uint64_t *d_keys_in, const int *d_vals_in,
                        uint64_t *d_keys_out, int *d_vals_out, long n, int bits);
/* No reference counterpart.  gcmf_radix_sort_u64 for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops): the same
 * sort, bit for bit, with its scratch (ping-pong buffers, digit histograms,
 * offsets) in one static device array of GCMF_NOALLOC_SCRATCH_BYTES per GPU,
 * held until the process ends; it calls no allocating routine.
 * n <= GCMF_NOALLOC_SORT_MAX, else GCMF_ERR_ARG.  Stream-ordered on the
 * context's stream and not synchronised (gcmf_synchronize waits).  Calls
 * that overlap on one GPU, from any context, race on the shared array. */
#define GCMF_NOALLOC_SORT_MAX (1L << 21)
#define GCMF_NOALLOC_SCRATCH_BYTES 26739456L
int gcmf_radix_sort_u64_noalloc(gcmf_ctx *ctx, const uint64_t *d_keys_in, const int *d_vals_in,
                                uint64_t *d_keys_out, int *d_vals_out, long n, int bits);
/* Exclusive prefix sum of n i32 into n + 1 i64 (d_out[n] = the total, so
 * d_out[1..n] is the inclusive scan); replaces PrefixSum int->long
 * (GPU_Utility.h:128-154, GPU_PrefixSum.h:355-458). */
int gcmf_scan_i32_i64(gcmf_ctx *ctx, const int *d_in, long long *d_out, long n);
/* Indices of the nonzero flags in increasing order and their number;
 * replaces Make_Filtered_Index_List (GPU_Vector.h:108-120). */
int gcmf_compact_flags(gcmf_ctx *ctx, const int *d_flags, long n, int *d_idx, long *count);

/* ------------------------------------------------------------ join ----- */
typedef struct {
  long candidates;      /* K: MBR-filter pairs (SortBaseMBROverlap return) */
  long pip_candidates;  /* pipNum (GetCMBR) */
  long within;          /* pairs inside by PiP (PointInPolygonTest return) */
  long cmf_candidates;  /* pairs entering CountCMF */
  long survivors;       /* eiNum: pairs with edges on both sides after CMF */
  long b_edges, o_edges;/* CMF edge list lengths */
  long intersecting;    /* SegmentIntersectJoin return */
  long joined;          /* within + intersecting (the fixed hand-off) */
} gcmf_join_stats;

/* The whole join chain of spatialJoin (spatialJoin3.cu:40-412) on
 * device-resident layers: SortBaseMBROverlap (Join.h:5) -> GetCMBR
 * (Join.h:397) -> PointInPolygonTest (Join.h:437) -> CountCMF / ApplyCMF
 * (Join.h:157,257) -> SegmentIntersectJoin (Join.h:356).  Results stay on
 * the device until the next call; read them with gcmf_join_get. */
int gcmf_join(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
              gcmf_join_stats *stats);

enum {
  GCMF_J_PAIRS = 0,     /* i32 [2K] (base, overlay), reference order        */
  GCMF_J_COUNTS = 1,    /* i32 [B+O] per-polygon candidate counts           */
  GCMF_J_PERM = 2,      /* i32 [2(B+O)] endpoint sort permutation           */
  GCMF_J_RANK = 3,      /* i32 [2(B+O)] endpoint rank                        */
  GCMF_J_CMBR = 4,      /* f64 [4K]                                          */
  GCMF_J_PIPTYPE = 5,   /* i8  [K]                                           */
  GCMF_J_PIPFLAG = 6,   /* i32 [K]                                           */
  GCMF_J_JOINFLAG = 7,  /* i32 [K] after CMF                                 */
  GCMF_J_BCOUNT = 8,    /* i32 [K] CMF base edge counts                      */
  GCMF_J_OCOUNT = 9,    /* i32 [K]                                           */
  GCMF_J_BEDGE = 10,    /* i32 [b_edges] CMF base edge lists, pair order     */
  GCMF_J_OEDGE = 11,    /* i32 [o_edges]                                     */
  GCMF_J_SEGFLAG = 12,  /* i32 [K] segment-intersection flags                */
  GCMF_J_JOINED = 13    /* i32 [2*joined] within U intersecting pairs        */
};
int gcmf_join_get(gcmf_ctx *ctx, int field, void *dst, long count);

/* ------------------------------------------------------------ clip ----- */
/* Per-pair predicate frame.  The reference evaluates every predicate on raw
 * coordinates against an absolute epsilon (1e-7, constants.h:4), which
 * classifies nearly every edge pair of small polygons as collinear.  With
 * GCMF_CLIP_NORMALIZE each pair's predicates run on ((x - cx) s, (y - cy) s),
 * (cx, cy) the CMBR corner and s = f / max(CMBR width, height), where f is
 * the power of two f >= sqrt(nP nQ) / 1024 (1 for rings of up to 1024
 * vertices), so that the edges of long rings stay long against the epsilon.
 * Output vertices are computed in the original coordinates.  Without the
 * flag the predicates use raw coordinates, as the reference does. */
#define GCMF_CLIP_NORMALIZE 1
#define GCMF_CLIP_NO_CLEANUP 2 /* skip cleanUpResult (polyclip.h:791)      */
#define GCMF_CLIP_SEQ_LABEL 4  /* label/trace one thread per pair (checking) */
/* Phase 1 as the reference's GPU route (calculateIntersections,
 * gppolyclip_mDIV512.cu:1298) instead of its CPU route (computeIntersections,
 * polyclip.h:188-281).  The two agree in general position; on degenerate
 * inputs only the CPU route is well defined. */
#define GCMF_CLIP_GPU_ROUTE 8
/* Accepted for compatibility: the edge scale f is now part of
 * GCMF_CLIP_NORMALIZE's frame, so this flag changes nothing. */
#define GCMF_CLIP_SCALE_EDGES 16
/* With GCMF_CLIP_NORMALIZE: keep f = 1 (the plain unit CMBR frame).  Rings
 * of 10^4-scale integer coordinates with 10^5 vertices (the reference's
 * s.txt x c.txt pair) fail their trace in that frame. */
#define GCMF_CLIP_UNIT_FRAME 32

typedef struct {
  long pairs, ok, invalid, disjoint, failed;
  long total_comps, total_verts;
  long edges_p, edges_q, slots_p, slots_q;
  long long n_eval;          /* edge pairs classified by the count kernel */
  long long n_intersections; /* intersection records over valid pairs */
  long long n_tested;        /* edge pairs the count kernel visited and box-tested */
} gcmf_clip_stats;

/* Batched Greiner-Hormann intersection of M pairs (d_pairs = 2M i32 of
 * (base index, overlay index)); replaces the per-pair loop of ghcuda
 * (ghcuda.cpp:385-425) with one pass over all pairs.  Each pair is swapped so
 * |P| >= |Q| as loadPolygonDataFromLayer does. */
int gcmf_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
              const int *d_pairs, long M, int flags, gcmf_clip_stats *stats);

/* Complex polygons: each side of a pair is a run of consecutive rings of
 * its layer (outer rings, holes and islands, even-odd), as the reference's
 * multi-polygon program clips them (GH-CUDA/ghcuda-multiPolygon.cpp:259-330,
 * gppolyclip_mDIV512-multiPolygon.cu, the component loops of
 * GH-CUDA/lib/polyclip.h).  d_parts = 4M i32: (first base ring, base ring
 * count, first overlay ring, overlay ring count), at most 64 rings a side
 * (pairs with more, or with a ring of no vertex, get status 2); parts
 * outside the layers fail the call with GCMF_ERR_ARG.  P is the base side (no swap).  With
 * GCMF_CLIP_UNION the result is the union (polyclip.h:42, UNION); a union of
 * polygons with disjoint MBRs is normalised by their joint MBR.  Results
 * are read with gcmf_clip_get (the phase-1 fields are not available). */
#define GCMF_CLIP_UNION 64
int gcmf_clip_complex(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                      const int *d_parts, long M, int flags, gcmf_clip_stats *stats);

/* Results of the context's last clip (gcmf_clip, gcmf_clip_complex):
 * the signed area of every component (counter-clockwise positive; the sum of
 * gcmf_ring_area, bit for bit) into d_out[total_comps] on the device, and
 * the results as a file: GCMF_WRITE_POLY, the .poly text of savePolygon
 * (polyclip.h:859-865) at %.17g, or GCMF_WRITE_WKT, one line per pair
 * "pair<TAB>status<TAB>MULTIPOLYGON (...)", the components of a complex
 * clip or a union nested by the even-odd rule (gcmf_result_write_wkt in
 * gcmf_host.h).  The reference has neither area nor WKT. */
#define GCMF_WRITE_POLY 0
#define GCMF_WRITE_WKT 1
int gcmf_clip_areas(gcmf_ctx *ctx, double *d_out);
int gcmf_clip_write(gcmf_ctx *ctx, const char *path, int format);

/* Launch shapes and thresholds of the batched clip, per context. */
typedef struct {
  long cull_min;       /* rings longer than this are walked through run boxes (512) */
  long long big_nodes; /* pairs with this many nodes are labelled by a
                          1024-thread block instead of one wave (8192) */
  int count_block;     /* threads per block of the count kernel: 64, 128, 256 (128) */
  int lds_max_q;       /* pairs whose Q ring is longer than cull_min and has at
                          most this many vertices are counted with Q binned in
                          LDS, one block per pair; the others walk Q's ring
                          (0..1024, 1024; 0 or <= cull_min: no LDS pass) */
  int label_lds;       /* bytes of LDS each labelling wave keeps a pair's node
                          arrays and scratch in; what does not fit stays in
                          HBM (0..65536, a multiple of 16; 20480) */
  long long grid_nodes; /* pairs with this many nodes are labelled by the whole
                          GPU, one cooperative launch each (2^20) */
  int grid_clean;       /* in such a pair, result components longer than this
                          are cleaned up by the whole GPU, the others by one
                          wave or thread each (49.., 16384) */
  long long strip_min;  /* joins of at least this many polygons (both layers)
                          run the MBR sweep over y strips (2^17; 0: always) */
  int cmf_group;        /* lanes per candidate pair in the CMF edge filter:
                          16 (four pairs per wave), 64, or 0: 16 when the
                          layers average at most 32 vertices per polygon (0) */
  int label_steal;      /* 1: the labelling waves are a persistent launch that
                          takes the pairs from a device counter in turn
                          (work stealing); 0: one wave per pair (0) */
  int block_cap;        /* blocks of the 1024-thread labelling launch, each
                          taking the next listed pair from a device counter
                          (1..65536, 128) */
  int count_stage;      /* 0: the edges of one pair in a count wave walk the
                          other ring together, 16-edge runs at a time, leaving
                          out the runs none of them can pair with; 1: every
                          edge walks the whole ring, which a wave whose edges
                          are of one pair first copies into LDS (up to 512
                          vertices) (0) */
} gcmf_tuning;
int gcmf_get_tuning(gcmf_ctx *ctx, gcmf_tuning *out);
int gcmf_set_tuning(gcmf_ctx *ctx, const gcmf_tuning *t);
/* Chooses the tuning by measurement on one batch (the autotuner): clips the
 * pairs once per candidate value of each parameter in turn, times the
 * launches the parameter affects with HIP events, and keeps the fastest in
 * the context.  The chosen values are returned in *out.  No reference
 * counterpart. */
int gcmf_tune_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                   const int *d_pairs, long M, int flags, gcmf_tuning *out);

/* Join then clip the joined pairs without leaving the device. */
int gcmf_join_clip(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                   int flags, gcmf_join_stats *jstats, gcmf_clip_stats *cstats);

/* HIP-graph capture of a batch, for callers that run the same batch again
 * (the launch sequence of ghcuda.cpp:393-422's per-pair loop, captured once
 * for the whole batch; no reference counterpart).  gcmf_clip_capture (and
 * gcmf_join_clip_capture) first runs gcmf_clip (gcmf_join_clip) eagerly --
 * stats and results are that run's -- then captures the same launches on
 * the context's stream as one graph, each size the eager run read back to
 * the host becoming a constant that a kernel of the graph checks against the
 * device's value.  gcmf_graph_launch replays the graph: every kernel runs
 * again on the current contents of the layers and pairs, and the results
 * are read with gcmf_clip_get / gcmf_join_get as after the eager call.  With
 * valid non-null it waits for the replay and sets *valid to 0 when some
 * size differed from the captured one (the inputs' shapes changed: the
 * results are then not to be used); with valid null it returns at once.
 * Replays are for inputs whose sizes are those captured: the buffers are
 * the capture's.  While a graph exists, the calls that use the context's
 * workspace (join, clip, tuning, sort, scan, compaction) return
 * GCMF_ERR_ARG; gcmf_graph_destroy releases it.  Pairs labelled by the
 * whole GPU (tuning grid_nodes) cannot be captured (GCMF_ERR_ARG). */
int gcmf_clip_capture(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                      const int *d_pairs, long M, int flags, gcmf_clip_stats *stats);
int gcmf_join_clip_capture(gcmf_ctx *ctx, const gcmf_dlayer *base, const gcmf_dlayer *over,
                           int flags, gcmf_join_stats *jstats, gcmf_clip_stats *cstats);
int gcmf_graph_launch(gcmf_ctx *ctx, int *valid);
int gcmf_graph_destroy(gcmf_ctx *ctx);

enum {
  GCMF_C_STATUS = 0,    /* i32 [M] 0 ok, 1 invalid, 2 disjoint, <0 failure */
  GCMF_C_SWAPPED = 1,   /* i32 [M] */
  GCMF_C_NCOMP = 2,     /* i32 [M] */
  GCMF_C_NVERTS = 3,    /* i32 [M] */
  GCMF_C_COMPOFF = 4,   /* i64 [M+1] first component of each pair */
  GCMF_C_VERTOFF = 5,   /* i64 [M+1] first vertex of each pair */
  GCMF_C_COMPSIZES = 6, /* i32 [total_comps] */
  GCMF_C_XY = 7,        /* f64 [2*total_verts] original coordinates */
  GCMF_C_LABEL_PATH = 8,  /* i32 [M] labelling path: 0 skipped, 1 segments, 2 segments +
                             clean-up, 3 no crossing, 4 sequential, 5 failed */
  GCMF_C_LABEL_TICKS = 9, /* i64 [M] labelling time per pair, device wall-clock ticks */
  /* phase-1 arrays (gppolyclip_mDIV512.cu outputs), global over the batch */
  GCMF_C_OFFP = 20,     /* i64 [M+1] first P edge of each pair */
  GCMF_C_OFFQ = 21,     /* i64 [M+1] */
  GCMF_C_PSP1 = 22,     /* i64 [edges_p+1] exclusive scan of count1 (P view) */
  GCMF_C_PSP2 = 23,     /* i64 [edges_p+1] exclusive scan of count2 */
  GCMF_C_PSQ1 = 24,     /* i64 [edges_q+1] */
  GCMF_C_PSQ2 = 25,     /* i64 [edges_q+1] */
  GCMF_C_BOOLP = 26,    /* i32 [edges_p] CMBR edge flags */
  GCMF_C_BOOLQ = 27,    /* i32 [edges_q] */
  GCMF_C_NMAPQ = 28,    /* i32 [slots_q] neighbour map (local P edge) */
  GCMF_C_IP = 29,       /* f64 [2*slots_p] predicate coordinates */
  GCMF_C_IQ = 30,       /* f64 [2*slots_q] */
  GCMF_C_OP = 31,       /* f64 [2*slots_p] original coordinates */
  GCMF_C_OQ = 32,       /* f64 [2*slots_q] */
  GCMF_C_AP = 33,       /* i32 [slots_p] trunc(1e7 alpha) or -100 */
  GCMF_C_AQ = 34,       /* i32 [slots_q] */
  GCMF_C_LP = 35,       /* i32 [slots_p] initial labels */
  GCMF_C_LQ = 36,       /* i32 [slots_q] */
  GCMF_C_NBP = 37,      /* i32 [slots_p] global Q slot + 1, 0 none */
  GCMF_C_NBQ = 38,      /* i32 [slots_q] global P slot + 1, 0 none */
  GCMF_C_NORM = 39      /* f64 [3M] cx, cy, scale */
};
int gcmf_clip_get(gcmf_ctx *ctx, int field, void *dst, long count);

/* ----------------------------------- reference call surface (drop-in) ---- */
typedef unsigned long long mbr_t;
typedef double coord_t;
typedef short poly_size_t;

/* Join.h:5 -- returns K; *djxyCounter (i32 [B+O]) and *djxyVector (i32 [2K])
 * are hipMalloc'ed and owned by the caller (hipFree).  dimSort 1 or 2 and
 * dimSelect 0 sort on x; dimSort 1 with dimSelect 1 sorts on y. */
long SortBaseMBROverlap(long bPolNum, long oPolNum, mbr_t *dbMBR, mbr_t *doMBR,
                        int **djxyCounter, int **djxyVector, char dimSort, char dimSelect);

/* Join.h:397 -- device outputs hipMalloc'ed, caller frees. */
void GetCMBR(long pairNum, int *jxyVector, coord_t *bMBR, coord_t *oMBR, coord_t **cMBR,
             int **jpipIndexList, int **pipFlag, char **pipType, int **joinFlag, long *pipNum);

/* Join.h:437 -- returns the number of pairs with pipFlag==1. */
long PointInPolygonTest(coord_t *bCoords, coord_t *oCoords, long pairNum, long pipNum,
                        int *jxyVector, int *jPiPIndexList, char *pipType, long *bVPSNum,
                        long *oVPSNum, int *pipFlag, int *joinFlag);

/* Join.h:157 -- CMF counts, prefix sums, workload and survivor list. */
void CountCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, int *joinFlag,
              int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, coord_t *cMBR,
              long **bEdgePSCounter, long **oEdgePSCounter, long **workLoadPSCounter,
              int **jxy2IndexList, poly_size_t **bEdgeList, poly_size_t **oEdgeList,
              long *eiNum, long *workLoadNum);

/* Join.h:257 -- fills the CMF edge lists allocated by CountCMF. */
void ApplyCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, long eiNum,
              int *jxy2IndexList, int *bVNum, int *oVNum, long *bPFVNum, long *oPFVNum,
              coord_t *cMBR, long *bEdgePSCounter, long *oEdgePSCounter,
              poly_size_t *bEdgeList, poly_size_t *oEdgeList);

/* Join.h:356 -- returns #intersecting; *segIntersectJoinFlag (i32 [eiNum])
 * hipMalloc'ed, caller frees. */
long SegmentIntersectJoin(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jCompactVector,
                          int *jxy2IndexList, long *bPFVNum, long *oPFVNum,
                          long *bEdgePSCounter, long *oEdgePSCounter, poly_size_t *bEdgeList,
                          poly_size_t *oEdgeList, int **segIntersectJoinFlag);

/* ghcuda.h:6-14 / gppolyclip_mDIV512.cu:1298 -- GH phase 1 for one pair
 * (P = base ring bPID of baseCoord, Q = overlay ring oPID), outputs malloc'ed
 * host arrays the caller frees; *invalid = 1 leaves them unallocated.
 * Predicates use raw coordinates as the reference does. */
void calculateIntersections(coord_t *baseCoord, coord_t *overlayCoords, int sizeP, int sizeQ,
                            coord_t *cmbr, int *dbVNum, long *dbVPSNum, int *doVNum,
                            long *doVPSNum, int bPID, int oPID, int *countNonDegenIntP,
                            int *countNonDegenIntQ, coord_t **intersectionsP,
                            coord_t **intersectionsQ, int **alphaValuesP, int **alphaValuesQ,
                            int **initLabelsP, int **initLabelsQ, int **neighborP,
                            int **neighborQ, int *invalid);

/* ghcuda.cpp:341 -- one pair; returns invalid (0/1). */
int GH_CUDA(coord_t *bCoords, coord_t *oCoords, int *pBVNum, long *pBVPSNum, int *pOVNum,
            long *pOVPSNum, int bPID, int oPID, int swapped);

/* main.h:8-11 / ghcuda.cpp:385 -- clips every listed pair of the resident
 * layers (each swapped so |P| >= |Q|, ghcuda.cpp:64-91); returns 0.  The
 * results stay on the process-default context: read them with
 * gcmf_clip_last_stats / gcmf_clip_get on gcmf_ref_context(), or set
 * GCMF_RESULT_FILE to have them written in savePolygon's format.
 * GCMF_NORMALIZE=0 selects the raw predicate frame (the reference's). */
int ghcuda(int pIDList[], int qIDList[], int totalNumPairs, coord_t *bCoords, coord_t *oCoords,
           int *pBVNum, long *pBVPSNum, int *pOVNum, long *pOVPSNum);

/* main.h:5-7 / spatialJoin3.cu:40-412 -- reads the two layer files named by
 * GCMF_BASE_FILE and GCMF_OVERLAY_FILE (the reference compiles its dataset
 * paths in, spatialJoin3.cu:88-117; argv picks the sort dimension as there),
 * joins them on the GPU and returns: the joined pairs as host malloc'ed
 * lists (*pIDList base ids, *qIDList overlay ids, *totalPairNum of them;
 * within U segment-intersecting, where the reference hands over the within
 * pairs only, spatialJoin3.cu:407), and the layers resident in device memory
 * (hipMalloc'ed, owned by the caller), ready for ghcuda.  Returns 0, or -1
 * with every output untouched. */
int spatialJoin(int argc, char *argv[], int **pIDList, int **qIDList, long *totalPairNum,
                coord_t **pBaseCoords, coord_t **pOverlayCoords, int **pBVNum, long **pBVPSNum,
                int **pOVNum, long **pOVPSNum);

/* ST_Intersect.cu:18-19 -- the reference's prototype: host layers in;
 * returns #within + #intersecting.  jPairs is passed by value there and the
 * pair list never reaches the caller, so this entry point does not write
 * through it either; use gcmf_st_intersect_pairs for the pairs.  Returns -1
 * on failure. */
int ST_Intersect(long bPolNum, long oPolNum, coord_t *baseCoords, coord_t *overlayCoords,
                 int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, mbr_t *seqMBR,
                 mbr_t *seqOMBR, coord_t *seqMBR2, coord_t *seqOMBR2, int *jPairs);

/* ST_Intersect with its output fixed (ST_Intersect.cu:186-212): *jPairs
 * (host, malloc'ed, caller frees) receives the 2n joined (base, overlay)
 * pairs, within pairs and intersecting pairs in candidate order; returns n,
 * or a negative status with *jPairs NULL. */
int gcmf_st_intersect_pairs(long bPolNum, long oPolNum, const double *baseCoords,
                            const double *overlayCoords, const int *bVNum, const int *oVNum,
                            const long *bVPSNum, const long *oVPSNum, const uint64_t *seqMBR,
                            const uint64_t *seqOMBR, const double *seqMBR2,
                            const double *seqOMBR2, int **jPairs);

/* The process-default context the reference-named entry points run on
 * (created on first use on the current device), so that their device
 * results can be read with gcmf_clip_get / gcmf_join_get; NULL when no
 * device is usable.  No reference counterpart (it keeps global state, F7). */
gcmf_ctx *gcmf_ref_context(void);
/* Releases the process-default context and its device memory; the next
 * reference-named call creates a new one. */
void gcmf_ref_release(void);
/* Statistics of the last clip on a context. */
int gcmf_clip_last_stats(gcmf_ctx *ctx, gcmf_clip_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
