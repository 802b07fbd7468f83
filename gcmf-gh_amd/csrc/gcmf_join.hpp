// gcmf_join.hpp -- host drivers of the spatial-join stages (join.hip) and of
// the batched Greiner-Hormann clip (gh.hip).
#pragma once
#include "gcmf_internal.hpp"

namespace gcmf {

struct JoinInputs {
  long B = 0, O = 0;
  const double *bCoords = nullptr, *oCoords = nullptr;
  const int *bVNum = nullptr, *oVNum = nullptr;
  const long *bVPSNum = nullptr, *oVPSNum = nullptr;
  const uint64_t *bMBR = nullptr, *oMBR = nullptr;
  const double *bMBR2 = nullptr, *oMBR2 = nullptr;
};

struct MbrSortView {
  const int *perm = nullptr;
  const int *rank = nullptr;
  long n = 0;
};

// Stage 1.  Returns K; *pairs_out = workspace array of 2K int32 (b, o).
// With the vertex counts, their mean per polygon is kept in c.avgVerts.
long mbr_filter(Context &c, long B, long O, const uint64_t *bMBR, const uint64_t *oMBR,
                int **counts_out, int **pairs_out, MbrSortView *sv,
                const int *bVNum = nullptr, const int *oVNum = nullptr);

// Stages 2-5, all device arrays in the context workspace.
struct JoinRefine {
  long K = 0, pipNum = 0, within = 0, cmfCand = 0, E = 0, hits = 0, nJoined = 0;
  long nbEdges = 0, noEdges = 0;
  double *cmbr = nullptr;
  int *pipFlag = nullptr, *joinFlag = nullptr, *pipCand = nullptr;
  signed char *pipType = nullptr;
  int *bCount = nullptr, *oCount = nullptr;
  long long *bOff = nullptr, *oOff = nullptr;
  int *bEdge = nullptr, *oEdge = nullptr;
  int *segFlag = nullptr, *list = nullptr, *joined = nullptr;
  int *joinedIdx = nullptr, *joinedPairs = nullptr;
};
void join_refine(Context &c, const JoinInputs &in, long K, const int *pairs, JoinRefine &r);

// single-stage launchers used by the reference-surface entry points
void launch_get_cmbr(Context &c, long K, const int *pairs, const double *bMBR2,
                     const double *oMBR2, double *cmbr, int *pipFlag, signed char *pipType,
                     int *joinFlag);
void launch_pip(Context &c, long nCand, const int *cand, const int *pairs,
                const signed char *pipType, const double *bCoords, const long *bVPSNum,
                const double *oCoords, const long *oVPSNum, int *pipFlag, int *joinFlag);
void launch_cmf_count(Context &c, long nList, const int *list, const int *pairs,
                      const double *cmbr, const JoinInputs &in, int *joinFlag, int *bCount,
                      int *oCount);
void launch_cmf_apply_short(Context &c, long nList, const int *list, const int *pairs,
                            const double *cmbr, const JoinInputs &in, const long long *bIncl,
                            const long long *oIncl, short *bEdge, short *oEdge);
void launch_segjoin_short(Context &c, long nList, const int *list, const int *pairs,
                          const JoinInputs &in, const long long *bIncl, const long long *oIncl,
                          const short *bEdge, const short *oEdge, int *flagByList);

// ------------------------------------------------------------ GH clip ----
// Per-pair status codes of the batched clip.
enum { GH_OK = 0, GH_INVALID = 1, GH_DISJOINT = 2, GH_TRACE_FAIL = -1, GH_OVERFLOW = -5 };

struct ClipOptions {
  int normalize = 1;  // 1: per-pair CMBR normalisation of every predicate, | 4: edge-scaled
  int cleanup = 1;    // run cleanUpResult
  int noswap = 0;     // keep P = base ring even when it is the smaller one
  int label_mode = 0; // 0: wave per pair (sequential fallback), 1: thread per pair
  int route = 1;      // phase 1: 1 the CPU route (computeIntersections), 0 the GPU route
};

struct ClipResult {
  long M = 0;
  // per pair (device): status, swapped, ncomp, nverts, comp offset, vertex offset
  int *status = nullptr, *swapped = nullptr, *ncomp = nullptr, *nverts = nullptr;
  long long *compOff = nullptr, *vertOff = nullptr;  // [M+1]
  long totalComps = 0, totalVerts = 0;
  int *compSizes = nullptr;  // [totalComps]
  double *xy = nullptr;      // [2*totalVerts] original coordinates
  // labelling path of each pair (LP_* in gh.hip) and its duration in ticks
  // of the device's constant-rate wall clock; null for the one-thread path
  int *labelPath = nullptr;
  long long *labelTicks = nullptr;
  // phase-1 statistics
  long long nEvalP = 0, totalIntersections = 0;
  long long nTested = 0;  // edge pairs the count kernel tested (box compare at least)
  long long nOk = 0, nInvalid = 0, nDisjoint = 0, nFailed = 0;  // pairs by status
};

// Per-pair phase-1 arrays for one clip launch (device, workspace-owned); slot
// indices are global over the batch.
struct Phase1View {
  long M = 0, edgesP = 0, edgesQ = 0, slotsP = 0, slotsQ = 0;
  const long long *offP = nullptr, *offQ = nullptr;       // edge offsets [M+1]
  const long long *psP1 = nullptr, *psP2 = nullptr;       // [edgesP+1]
  const long long *psQ1 = nullptr, *psQ2 = nullptr;       // [edgesQ+1]
  const int *boolP = nullptr, *boolQ = nullptr;           // [edges]
  const int *nmapQ = nullptr;                              // [slotsQ]
  const double *iP = nullptr, *iQ = nullptr, *oP = nullptr, *oQ = nullptr;  // [2 slots]
  const int *aP = nullptr, *aQ = nullptr, *lP = nullptr, *lQ = nullptr;
  const int *nbP = nullptr, *nbQ = nullptr;                // global slot + 1, 0 none
  const int *status = nullptr;
  const double *norm = nullptr;                            // [3M]
};

// Clips M pairs (bIdx[k] of layer b, oIdx[k] of layer o; pairs interleaved).
void clip_pairs(Context &c, const JoinInputs &in, const int *pairs, long M,
                const ClipOptions &opt, ClipResult &res, Phase1View *p1);

// Clips M pairs of complex polygons: parts[4k..4k+3] = first base ring,
// base ring count, first overlay ring, overlay ring count (at most 64 rings
// a side); the union instead of the intersection when unionOp.
void clip_complex(Context &c, const JoinInputs &in, const int *parts, long M,
                  const ClipOptions &opt, bool unionOp, ClipResult &res);

// Signed area of every component of res into out[res.totalComps] (device),
// summed as gcmf_ring_area sums it on the host.
void comp_areas(Context &c, const ClipResult &res, double *out);

}  // namespace gcmf

// The state behind one gcmf_ctx handle of the C ABI: the context and the
// results of its last join and clip, kept on the device for the *_get calls.
struct gcmf_ctx {
  gcmf::Context c;
  bool join_valid = false;
  gcmf::JoinInputs jin;
  long K = 0;
  int *pairs = nullptr, *counts = nullptr;
  gcmf::MbrSortView sv;
  gcmf::JoinRefine jr;
  bool clip_valid = false;
  bool clip_nested = false;  // the last clip was of complex polygons (or the union)
  gcmf::ClipResult cr;
  gcmf::Phase1View p1;
  // a captured clip or join-and-clip (gcmf_*_capture), replayed by
  // gcmf_graph_launch; while it exists the calls that use the workspace
  // are refused, since the graph holds the workspace's buffers
  hipGraphExec_t graph = nullptr;
};
