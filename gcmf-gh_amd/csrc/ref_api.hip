// ref_api.hip -- the reference's own call surface (Join.h, ghcuda.h,
// ghcuda.cpp, ST_Intersect.cu) as extern "C" entry points over the gfx950
// kernels, so a caller of the reference links against libgcmf_hip.so
// unchanged.  Semantics follow the cited reference functions; device outputs
// are hipMalloc'ed and owned by the caller exactly where the reference
// cudaMalloc's them.  Failures print to stderr and return -1 / leave the
// outputs NULL instead of calling exit().
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/gcmf_hip.h"
#include "../../include/gcmf_host.h"
#include "gcmf_internal.hpp"
#include "gcmf_join.hpp"

namespace {

std::mutex g_mu;
// The process-default context of the reference-named entry points (the
// reference keeps global state instead, F7); created on first use on the
// device current at that time.
gcmf_ctx *g_ctx = nullptr;

gcmf_ctx &dctx() {
  if (!g_ctx) {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    gcmf_ctx *x = nullptr;
    const int rc = gcmf_init(dev, &x);
    if (rc != GCMF_OK) throw gcmf::Error{rc, "cannot create the default context"};
    g_ctx = x;
  }
  HIP_CHECK(hipSetDevice(g_ctx->c.device));
  return *g_ctx;
}

// Runs f on the default context; library errors (and allocation failures of
// the host containers) are printed and reported as false, never thrown
// across the C boundary.
template <typename F>
bool shim(const char *name, F &&f) {
  std::lock_guard<std::mutex> lk(g_mu);
  std::string msg;
  try {
    f(dctx());
    return true;
  } catch (const gcmf::Error &e) {
    msg = e.what;
  } catch (const std::exception &e) {
    msg = e.what();
  }
  std::fprintf(stderr, "%s: %s\n", name, msg.c_str());
  if (g_ctx) g_ctx->c.last_error = msg;
  return false;
}

template <typename T>
T *dalloc(long count) {
  void *p = nullptr;
  HIP_CHECK(hipMalloc(&p, sizeof(T) * (size_t)(count > 0 ? count : 1)));
  return (T *)p;
}

__global__ void k_swap_xy(long n, const uint64_t *in, uint64_t *out) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[4 * i] = in[4 * i + 1];
  out[4 * i + 1] = in[4 * i];
  out[4 * i + 2] = in[4 * i + 3];
  out[4 * i + 3] = in[4 * i + 2];
}

__global__ void k_incl_from_excl(long n, const long long *ex, long long *in) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) in[i] = ex[i + 1];
}

__global__ void k_mul(long n, const int *a, const int *b, int *c) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) c[i] = a[i] * b[i];
}

__global__ void k_eq1(long n, const int *f, int *o) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) o[i] = f[i] == 1;
}

// Device arrays freed at scope exit unless `keep` is set (they reached the
// caller).
struct OwnedDev {
  std::vector<void *> p;
  bool keep = false;
  ~OwnedDev() {
    if (!keep)
      for (void *q : p) (void)hipFree(q);
  }
  template <typename T>
  T *add(T *q) {
    p.push_back(q);
    return q;
  }
};

// savePolygon's format (polygon.h:575-584) at full precision: one vertex
// per line, "x y," inside a component and "x y;" at its end.
void write_results(gcmf::Context &c, const gcmf::ClipResult &r) {
  const char *path = std::getenv("GCMF_RESULT_FILE");
  if (!path || r.totalVerts <= 0) return;
  std::vector<double> xy(2 * (size_t)r.totalVerts);
  std::vector<int> cs((size_t)r.totalComps);
  HIP_CHECK(hipMemcpyAsync(xy.data(), r.xy, 16 * (size_t)r.totalVerts, hipMemcpyDeviceToHost,
                           c.stream));
  HIP_CHECK(hipMemcpyAsync(cs.data(), r.compSizes, 4 * (size_t)r.totalComps,
                           hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
  if (gcmf_result_write_poly(path, xy.data(), cs.data(), r.totalComps) != 0)
    throw gcmf::Error{GCMF_ERR_ARG, std::string("cannot write ") + path};
}

// Predicate frame of ghcuda / GH_CUDA: the normalised, edge-scaled frame
// (frame bits 1 | 4 of ClipOptions::normalize) unless GCMF_NORMALIZE=0 asks
// for the reference's raw coordinates.
int normalize_default() {
  const char *e = std::getenv("GCMF_NORMALIZE");
  return (e && e[0] == '0') ? 0 : 5;
}

// Clips M pairs on the default context and keeps the results there, where
// gcmf_clip_get / gcmf_clip_last_stats on gcmf_ref_context() read them.
void ref_clip(gcmf_ctx &x, const gcmf::JoinInputs &in, const int *d_pairs, long M,
              const gcmf::ClipOptions &opt) {
  x.clip_valid = false;
  gcmf::clip_pairs(x.c, in, d_pairs, M, opt, x.cr, &x.p1);
  x.c.flush_timers();
  x.clip_valid = true;
}

// Host layers to the device: the join of ST_Intersect (ST_Intersect.cu:18-212)
// with the fixed hand-off; the joined pairs stay in the context workspace.
gcmf::JoinRefine &st_intersect(gcmf_ctx &x, long bPolNum, long oPolNum, coord_t *baseCoords,
                               coord_t *overlayCoords, int *bVNum, int *oVNum, long *bVPSNum,
                               long *oVPSNum, mbr_t *seqMBR, mbr_t *seqOMBR, coord_t *seqMBR2,
                               coord_t *seqOMBR2) {
  gcmf::Context &c = x.c;
  if (bPolNum < 0 || oPolNum < 0) throw gcmf::Error{GCMF_ERR_ARG, "negative polygon count"};
  const long nb = bPolNum > 0 ? bVPSNum[bPolNum - 1] : 0;
  const long no = oPolNum > 0 ? oVPSNum[oPolNum - 1] : 0;
  auto up = [&](const char *tag, const void *src, size_t bytes) {
    void *d = c.ws.get(std::string("ref_st_") + tag, bytes);
    if (bytes) HIP_CHECK(hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, c.stream));
    return d;
  };
  gcmf::JoinInputs in;
  in.B = bPolNum;
  in.O = oPolNum;
  in.bCoords = (const double *)up("bc", baseCoords, 16 * (size_t)nb);
  in.oCoords = (const double *)up("oc", overlayCoords, 16 * (size_t)no);
  in.bVNum = (const int *)up("bv", bVNum, 4 * (size_t)bPolNum);
  in.oVNum = (const int *)up("ov", oVNum, 4 * (size_t)oPolNum);
  in.bVPSNum = (const long *)up("bp", bVPSNum, 8 * (size_t)bPolNum);
  in.oVPSNum = (const long *)up("op", oVPSNum, 8 * (size_t)oPolNum);
  in.bMBR = (const uint64_t *)up("bm", seqMBR, 32 * (size_t)bPolNum);
  in.oMBR = (const uint64_t *)up("om", seqOMBR, 32 * (size_t)oPolNum);
  in.bMBR2 = (const double *)up("bm2", seqMBR2, 32 * (size_t)bPolNum);
  in.oMBR2 = (const double *)up("om2", seqOMBR2, 32 * (size_t)oPolNum);
  x.join_valid = false;
  x.jin = in;
  x.K = gcmf::mbr_filter(c, bPolNum, oPolNum, in.bMBR, in.oMBR, &x.counts, &x.pairs, &x.sv,
                         in.bVNum, in.oVNum);
  gcmf::join_refine(c, in, x.K, x.pairs, x.jr);
  x.join_valid = true;
  return x.jr;
}

}  // namespace

extern "C" {

long SortBaseMBROverlap(long bPolNum, long oPolNum, mbr_t *dbMBR, mbr_t *doMBR, int **djxyCounter,
                        int **djxyVector, char dimSort, char dimSelect) {
  long K = -1;
  *djxyCounter = nullptr;
  *djxyVector = nullptr;
  shim("SortBaseMBROverlap", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    const uint64_t *b = (const uint64_t *)dbMBR, *o = (const uint64_t *)doMBR;
    if (dimSort == 1 && dimSelect == 1) {  // InitializeData with x/y exchanged (Join.h:34)
      uint64_t *sb = (uint64_t *)c.ws.get("ref_swap_b", 32 * (size_t)(bPolNum + 1));
      uint64_t *so = (uint64_t *)c.ws.get("ref_swap_o", 32 * (size_t)(oPolNum + 1));
      if (bPolNum > 0)
        hipLaunchKernelGGL(k_swap_xy, dim3(gcmf::grid_for(bPolNum, 256)), dim3(256), 0, c.stream,
                           bPolNum, (const uint64_t *)dbMBR, sb);
      if (oPolNum > 0)
        hipLaunchKernelGGL(k_swap_xy, dim3(gcmf::grid_for(oPolNum, 256)), dim3(256), 0, c.stream,
                           oPolNum, (const uint64_t *)doMBR, so);
      b = sb;
      o = so;
    }
    int *counts = nullptr, *pairs = nullptr;
    long k = gcmf::mbr_filter(c, bPolNum, oPolNum, b, o, &counts, &pairs, nullptr);
    int *dc = dalloc<int>(bPolNum + oPolNum), *dv = dalloc<int>(2 * k);
    HIP_CHECK(hipMemcpyAsync(dc, counts, 4 * (size_t)(bPolNum + oPolNum), hipMemcpyDeviceToDevice,
                             c.stream));
    if (k > 0)
      HIP_CHECK(hipMemcpyAsync(dv, pairs, 8 * (size_t)k, hipMemcpyDeviceToDevice, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    *djxyCounter = dc;
    *djxyVector = dv;
    K = k;
  });
  return K;
}

void GetCMBR(long pairNum, int *jxyVector, coord_t *bMBR, coord_t *oMBR, coord_t **cMBR,
             int **jpipIndexList, int **pipFlag, char **pipType, int **joinFlag, long *pipNum) {
  shim("GetCMBR", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    *cMBR = dalloc<double>(4 * pairNum);
    *pipFlag = dalloc<int>(pairNum);
    *pipType = dalloc<char>(pairNum);
    *joinFlag = dalloc<int>(pairNum);
    gcmf::launch_get_cmbr(c, pairNum, jxyVector, bMBR, oMBR, *cMBR, *pipFlag,
                          (signed char *)*pipType, *joinFlag);
    int *tmp = (int *)c.ws.get("ref_pip_idx", 4 * (size_t)(pairNum + 1));
    long n = gcmf::compact_flags(c, *pipFlag, pairNum, tmp, "ref_pip");
    *jpipIndexList = dalloc<int>(n);
    if (n > 0)
      HIP_CHECK(hipMemcpyAsync(*jpipIndexList, tmp, 4 * (size_t)n, hipMemcpyDeviceToDevice,
                               c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    *pipNum = n;
  });
}

long PointInPolygonTest(coord_t *bCoords, coord_t *oCoords, long pairNum, long pipNum,
                        int *jxyVector, int *jPiPIndexList, char *pipType, long *bVPSNum,
                        long *oVPSNum, int *pipFlag, int *joinFlag) {
  long within = -1;
  shim("PointInPolygonTest", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::launch_pip(c, pipNum, jPiPIndexList, jxyVector, (const signed char *)pipType, bCoords,
                     bVPSNum, oCoords, oVPSNum, pipFlag, joinFlag);
    int *tmp = (int *)c.ws.get("ref_within_idx", 4 * (size_t)(pairNum + 1));
    within = gcmf::compact_flags(c, pipFlag, pairNum, tmp, "ref_within");
  });
  return within;
}

void CountCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, int *joinFlag,
              int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, coord_t *cMBR,
              long **bEdgePSCounter, long **oEdgePSCounter, long **workLoadPSCounter,
              int **jxy2IndexList, poly_size_t **bEdgeList, poly_size_t **oEdgeList, long *eiNum,
              long *workLoadNum) {
  shim("CountCMF", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::JoinInputs in;
    in.bCoords = bCoords;
    in.oCoords = oCoords;
    in.bVNum = bVNum;
    in.oVNum = oVNum;
    in.bVPSNum = bVPSNum;
    in.oVPSNum = oVPSNum;
    const long n = pairNum;
    int *list = (int *)c.ws.get("ref_cmf_list", 4 * (size_t)(n + 1));
    int *bc = (int *)c.ws.get("ref_cmf_bc", 4 * (size_t)(n + 1));
    int *oc = (int *)c.ws.get("ref_cmf_oc", 4 * (size_t)(n + 1));
    int *wl = (int *)c.ws.get("ref_cmf_wl", 4 * (size_t)(n + 1));
    int *f01 = (int *)c.ws.get("ref_cmf_f01", 4 * (size_t)(n + 1));
    long long *ex = (long long *)c.ws.get("ref_cmf_ex", 8 * (size_t)(n + 1));
    hipLaunchKernelGGL(k_eq1, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n, joinFlag,
                       f01);
    long n0 = gcmf::compact_flags(c, f01, n, list, "ref_cmf0");
    gcmf::fill_i32(c, bc, 0, n);
    gcmf::fill_i32(c, oc, 0, n);
    gcmf::launch_cmf_count(c, n0, list, jxyVector, cMBR, in, joinFlag, bc, oc);
    long long *bps = dalloc<long long>(n), *ops = dalloc<long long>(n), *wps = dalloc<long long>(n);
    gcmf::scan_i32_to_i64(c, bc, ex, n);
    hipLaunchKernelGGL(k_incl_from_excl, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n,
                       ex, bps);
    long nb = (long)gcmf::read_scalar(c, ex + n);
    gcmf::scan_i32_to_i64(c, oc, ex, n);
    hipLaunchKernelGGL(k_incl_from_excl, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n,
                       ex, ops);
    long no = (long)gcmf::read_scalar(c, ex + n);
    hipLaunchKernelGGL(k_mul, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n, bc, oc,
                       wl);
    gcmf::scan_i32_to_i64(c, wl, ex, n);
    hipLaunchKernelGGL(k_incl_from_excl, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n,
                       ex, wps);
    *workLoadNum = (long)gcmf::read_scalar(c, ex + n);
    hipLaunchKernelGGL(k_eq1, dim3(gcmf::grid_for(n, 256)), dim3(256), 0, c.stream, n, joinFlag,
                       f01);
    long E = gcmf::compact_flags(c, f01, n, list, "ref_cmf1");
    *jxy2IndexList = dalloc<int>(E);
    if (E > 0)
      HIP_CHECK(hipMemcpyAsync(*jxy2IndexList, list, 4 * (size_t)E, hipMemcpyDeviceToDevice,
                               c.stream));
    *bEdgeList = dalloc<short>(nb);
    *oEdgeList = dalloc<short>(no);
    HIP_CHECK(hipStreamSynchronize(c.stream));
    *bEdgePSCounter = (long *)bps;
    *oEdgePSCounter = (long *)ops;
    *workLoadPSCounter = (long *)wps;
    *eiNum = E;
  });
}

void ApplyCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, long eiNum,
              int *jxy2IndexList, int *bVNum, int *oVNum, long *bPFVNum, long *oPFVNum,
              coord_t *cMBR, long *bEdgePSCounter, long *oEdgePSCounter, poly_size_t *bEdgeList,
              poly_size_t *oEdgeList) {
  (void)pairNum;
  shim("ApplyCMF", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::JoinInputs in;
    in.bCoords = bCoords;
    in.oCoords = oCoords;
    in.bVNum = bVNum;
    in.oVNum = oVNum;
    in.bVPSNum = bPFVNum;
    in.oVPSNum = oPFVNum;
    gcmf::launch_cmf_apply_short(c, eiNum, jxy2IndexList, jxyVector, cMBR, in,
                                 (const long long *)bEdgePSCounter,
                                 (const long long *)oEdgePSCounter, bEdgeList, oEdgeList);
    HIP_CHECK(hipStreamSynchronize(c.stream));
  });
}

long SegmentIntersectJoin(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jCompactVector,
                          int *jxy2IndexList, long *bPFVNum, long *oPFVNum,
                          long *bEdgePSCounter, long *oEdgePSCounter, poly_size_t *bEdgeList,
                          poly_size_t *oEdgeList, int **segIntersectJoinFlag) {
  long hits = -1;
  *segIntersectJoinFlag = nullptr;
  shim("SegmentIntersectJoin", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::JoinInputs in;
    in.bCoords = bCoords;
    in.oCoords = oCoords;
    in.bVPSNum = bPFVNum;
    in.oVPSNum = oPFVNum;
    int *flag = dalloc<int>(pairNum);
    gcmf::launch_segjoin_short(c, pairNum, jxy2IndexList, jCompactVector, in,
                               (const long long *)bEdgePSCounter,
                               (const long long *)oEdgePSCounter, bEdgeList, oEdgeList, flag);
    int *tmp = (int *)c.ws.get("ref_seg_idx", 4 * (size_t)(pairNum + 1));
    hits = gcmf::compact_flags(c, flag, pairNum, tmp, "ref_seg");
    *segIntersectJoinFlag = flag;
  });
  return hits;
}

void calculateIntersections(coord_t *baseCoord, coord_t *overlayCoords, int sizeP, int sizeQ,
                            coord_t *cmbr, int *dbVNum, long *dbVPSNum, int *doVNum,
                            long *doVPSNum, int bPID, int oPID, int *countNonDegenIntP,
                            int *countNonDegenIntQ, coord_t **intersectionsP,
                            coord_t **intersectionsQ, int **alphaValuesP, int **alphaValuesQ,
                            int **initLabelsP, int **initLabelsQ, int **neighborP,
                            int **neighborQ, int *invalid) {
  (void)sizeP;
  (void)sizeQ;
  (void)cmbr;  // recomputed from the rings exactly as getCMBR does
  *invalid = 1;
  shim("calculateIntersections", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::JoinInputs in;
    in.bCoords = baseCoord;
    in.oCoords = overlayCoords;
    in.bVNum = dbVNum;
    in.oVNum = doVNum;
    in.bVPSNum = dbVPSNum;
    in.oVPSNum = doVPSNum;
    int *dp = (int *)c.ws.get("ref_ci_pair", 8);
    int hp[2] = {bPID, oPID};
    HIP_CHECK(hipMemcpyAsync(dp, hp, 8, hipMemcpyHostToDevice, c.stream));
    gcmf::ClipOptions opt;
    opt.normalize = 0;
    opt.noswap = 1;
    opt.route = 0;  // the phase-1 arrays of calculateIntersections
    ref_clip(x, in, dp, 1, opt);
    const gcmf::Phase1View &p = x.p1;
    int st;
    HIP_CHECK(hipMemcpy(&st, p.status, 4, hipMemcpyDeviceToHost));
    if (st != gcmf::GH_OK && st < 0) st = gcmf::GH_OK;  // phase 1 itself succeeded
    if (st != gcmf::GH_OK) return;
    const long cP = p.slotsP, cQ = p.slotsQ;
    *countNonDegenIntP = (int)cP;
    *countNonDegenIntQ = (int)cQ;
    *intersectionsP = (double *)std::malloc(16 * (size_t)(cP ? cP : 1));
    *intersectionsQ = (double *)std::malloc(16 * (size_t)(cQ ? cQ : 1));
    *alphaValuesP = (int *)std::malloc(4 * (size_t)(cP ? cP : 1));
    *alphaValuesQ = (int *)std::malloc(4 * (size_t)(cQ ? cQ : 1));
    *initLabelsP = (int *)std::malloc(4 * (size_t)(cP ? cP : 1));
    *initLabelsQ = (int *)std::malloc(4 * (size_t)(cQ ? cQ : 1));
    *neighborP = (int *)std::malloc(4 * (size_t)(cP ? cP : 1));
    *neighborQ = (int *)std::malloc(4 * (size_t)(cQ ? cQ : 1));
    HIP_CHECK(hipMemcpy(*intersectionsP, p.iP, 16 * (size_t)cP, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*intersectionsQ, p.iQ, 16 * (size_t)cQ, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*alphaValuesP, p.aP, 4 * (size_t)cP, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*alphaValuesQ, p.aQ, 4 * (size_t)cQ, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*initLabelsP, p.lP, 4 * (size_t)cP, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*initLabelsQ, p.lQ, 4 * (size_t)cQ, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*neighborP, p.nbP, 4 * (size_t)cP, hipMemcpyDeviceToHost));
    HIP_CHECK(hipMemcpy(*neighborQ, p.nbQ, 4 * (size_t)cQ, hipMemcpyDeviceToHost));
    *invalid = 0;
  });
}

int GH_CUDA(coord_t *bCoords, coord_t *oCoords, int *pBVNum, long *pBVPSNum, int *pOVNum,
            long *pOVPSNum, int bPID, int oPID, int swapped) {
  int invalid = 1;
  shim("GH_CUDA", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    // swapped = 1: P is overlay ring oPID and Q base ring bPID
    // (regularPolygonHandler, ghcuda.cpp:196-219); the caller decides, as
    // loadPolygonDataFromLayer does for ghcuda (ghcuda.cpp:64-91)
    gcmf::JoinInputs in;
    in.bCoords = swapped ? oCoords : bCoords;
    in.oCoords = swapped ? bCoords : oCoords;
    in.bVNum = swapped ? pOVNum : pBVNum;
    in.oVNum = swapped ? pBVNum : pOVNum;
    in.bVPSNum = swapped ? pOVPSNum : pBVPSNum;
    in.oVPSNum = swapped ? pBVPSNum : pOVPSNum;
    int *dp = (int *)c.ws.get("ref_gh_pair", 8);
    int hp[2] = {swapped ? oPID : bPID, swapped ? bPID : oPID};
    HIP_CHECK(hipMemcpyAsync(dp, hp, 8, hipMemcpyHostToDevice, c.stream));
    gcmf::ClipOptions opt;
    opt.normalize = normalize_default();
    opt.noswap = 1;
    ref_clip(x, in, dp, 1, opt);
    int st;
    HIP_CHECK(hipMemcpyAsync(&st, x.cr.status, 4, hipMemcpyDeviceToHost, c.stream));
    HIP_CHECK(hipStreamSynchronize(c.stream));
    invalid = st == gcmf::GH_INVALID ? 1 : 0;
    write_results(c, x.cr);
  });
  return invalid;
}

int ghcuda(int pIDList[], int qIDList[], int totalNumPairs, coord_t *bCoords, coord_t *oCoords,
           int *pBVNum, long *pBVPSNum, int *pOVNum, long *pOVPSNum) {
  shim("ghcuda", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    gcmf::JoinInputs in;
    in.bCoords = bCoords;
    in.oCoords = oCoords;
    in.bVNum = pBVNum;
    in.oVNum = pOVNum;
    in.bVPSNum = pBVPSNum;
    in.oVPSNum = pOVPSNum;
    const long M = totalNumPairs > 0 ? totalNumPairs : 0;
    std::vector<int> hp(2 * (size_t)(M ? M : 1));
    for (long k = 0; k < M; k++) {
      hp[2 * k] = pIDList[k];
      hp[2 * k + 1] = qIDList[k];
    }
    int *dp = (int *)c.ws.get("ref_ghcuda_pairs", 8 * (size_t)(M ? M : 1));
    HIP_CHECK(hipMemcpyAsync(dp, hp.data(), 8 * (size_t)M, hipMemcpyHostToDevice, c.stream));
    gcmf::ClipOptions opt;
    opt.normalize = normalize_default();
    ref_clip(x, in, dp, M, opt);
    write_results(c, x.cr);
  });
  return 0;
}

int ST_Intersect(long bPolNum, long oPolNum, coord_t *baseCoords, coord_t *overlayCoords,
                 int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, mbr_t *seqMBR,
                 mbr_t *seqOMBR, coord_t *seqMBR2, coord_t *seqOMBR2, int *jPairs) {
  // jPairs is passed by value in the reference and never written through
  // (ST_Intersect.cu:191 assigns a fresh malloc to the local copy), so the
  // caller's buffer is left untouched here too
  (void)jPairs;
  int ret = -1;
  shim("ST_Intersect", [&](gcmf_ctx &x) {
    gcmf::JoinRefine &r = st_intersect(x, bPolNum, oPolNum, baseCoords, overlayCoords, bVNum,
                                       oVNum, bVPSNum, oVPSNum, seqMBR, seqOMBR, seqMBR2,
                                       seqOMBR2);
    ret = (int)r.nJoined;
  });
  return ret;
}

int gcmf_st_intersect_pairs(long bPolNum, long oPolNum, const double *baseCoords,
                            const double *overlayCoords, const int *bVNum, const int *oVNum,
                            const long *bVPSNum, const long *oVPSNum, const uint64_t *seqMBR,
                            const uint64_t *seqOMBR, const double *seqMBR2,
                            const double *seqOMBR2, int **jPairs) {
  if (!jPairs) return GCMF_ERR_ARG;
  *jPairs = nullptr;
  int ret = GCMF_ERR_HIP;
  shim("gcmf_st_intersect_pairs", [&](gcmf_ctx &x) {
    gcmf::JoinRefine &r = st_intersect(
        x, bPolNum, oPolNum, const_cast<double *>(baseCoords), const_cast<double *>(overlayCoords),
        const_cast<int *>(bVNum), const_cast<int *>(oVNum), const_cast<long *>(bVPSNum),
        const_cast<long *>(oVPSNum), (mbr_t *)seqMBR, (mbr_t *)seqOMBR,
        const_cast<double *>(seqMBR2), const_cast<double *>(seqOMBR2));
    int *h = (int *)std::malloc(8 * (size_t)(r.nJoined ? r.nJoined : 1));
    if (!h) throw gcmf::Error{GCMF_ERR_ALLOC, "host allocation of the joined pairs"};
    if (r.nJoined > 0) {
      HIP_CHECK(hipMemcpyAsync(h, r.joinedPairs, 8 * (size_t)r.nJoined, hipMemcpyDeviceToHost,
                               x.c.stream));
      HIP_CHECK(hipStreamSynchronize(x.c.stream));
    }
    *jPairs = h;
    ret = (int)r.nJoined;
  });
  return ret;
}

int spatialJoin(int argc, char *argv[], int **pIDList, int **qIDList, long *totalPairNum,
                coord_t **pBaseCoords, coord_t **pOverlayCoords, int **pBVNum, long **pBVPSNum,
                int **pOVNum, long **pOVPSNum) {
  int rc = -1;
  shim("spatialJoin", [&](gcmf_ctx &x) {
    gcmf::Context &c = x.c;
    // console arguments as spatialJoin3.cu:52-69: dimSort 1 sorts one
    // dimension (dimSelect 'y' picks y), 2 both (here: x)
    int dimSort = 1, dimSelect = 1;
    if (argc < 2) {
      dimSelect = 0;
    } else if (argc < 3) {
      dimSort = argv[1][0] == '2' ? 2 : 1;
      dimSelect = 0;
    } else if (argc < 4) {
      dimSelect = argv[2][0] == 'y' ? 1 : 0;
      if (argv[1][0] == '2') {
        dimSort = 2;
        dimSelect = 0;
      }
    }
    // layer files: the reference compiles their names in (DATASET,
    // spatialJoin3.cu:88-117); here they come from the environment
    const char *bf = std::getenv("GCMF_BASE_FILE"), *of = std::getenv("GCMF_OVERLAY_FILE");
    if (!bf || !of)
      throw gcmf::Error{GCMF_ERR_ARG, "set GCMF_BASE_FILE and GCMF_OVERLAY_FILE to the layers"};
    std::unique_ptr<gcmf_layer, void (*)(gcmf_layer *)> bl(gcmf_layer_read_text(bf, -1),
                                                           gcmf_layer_free);
    std::unique_ptr<gcmf_layer, void (*)(gcmf_layer *)> ol(gcmf_layer_read_text(of, -1),
                                                           gcmf_layer_free);
    if (!bl || !ol) throw gcmf::Error{GCMF_ERR_ARG, "cannot read the layer files"};
    const long B = gcmf_layer_npoly(bl.get()), O = gcmf_layer_npoly(ol.get());
    const long nb = gcmf_layer_nverts(bl.get()), no = gcmf_layer_nverts(ol.get());
    // resident layers, handed to the caller as the reference hands its
    // device arrays to ghcuda (spatialJoin3.cu:397-402)
    OwnedDev own;
    double *bc = own.add(dalloc<double>(2 * nb)), *oc = own.add(dalloc<double>(2 * no));
    int *bv = own.add(dalloc<int>(B)), *ov = own.add(dalloc<int>(O));
    long *bp = own.add(dalloc<long>(B)), *op = own.add(dalloc<long>(O));
    auto up = [&](void *d, const void *h, size_t bytes) {
      if (bytes) HIP_CHECK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c.stream));
    };
    up(bc, gcmf_layer_coords(bl.get()), 16 * (size_t)nb);
    up(oc, gcmf_layer_coords(ol.get()), 16 * (size_t)no);
    up(bv, gcmf_layer_vnum(bl.get()), 4 * (size_t)B);
    up(ov, gcmf_layer_vnum(ol.get()), 4 * (size_t)O);
    up(bp, gcmf_layer_vpsnum(bl.get()), 8 * (size_t)B);
    up(op, gcmf_layer_vpsnum(ol.get()), 8 * (size_t)O);
    gcmf::JoinInputs in;
    in.B = B;
    in.O = O;
    in.bCoords = bc;
    in.oCoords = oc;
    in.bVNum = bv;
    in.oVNum = ov;
    in.bVPSNum = bp;
    in.oVPSNum = op;
    in.bMBR2 = (const double *)c.ws.get("ref_sj_bm2", 32 * (size_t)(B + 1));
    in.oMBR2 = (const double *)c.ws.get("ref_sj_om2", 32 * (size_t)(O + 1));
    up((void *)in.bMBR2, gcmf_layer_mbr2(bl.get()), 32 * (size_t)B);
    up((void *)in.oMBR2, gcmf_layer_mbr2(ol.get()), 32 * (size_t)O);
    // MBR keys in the dimension the arguments select
    std::vector<uint64_t> km[2];
    const gcmf_layer *ls[2] = {bl.get(), ol.get()};
    const long ns[2] = {B, O};
    const char *tags[2] = {"ref_sj_bm", "ref_sj_om"};
    const uint64_t *dk[2];
    for (int s = 0; s < 2; s++) {
      const uint64_t *m = gcmf_layer_mbr(ls[s]);
      km[s].assign(m, m + 4 * ns[s]);
      if (dimSort == 1 && dimSelect == 1)
        for (long i = 0; i < ns[s]; i++) {
          std::swap(km[s][4 * i], km[s][4 * i + 1]);
          std::swap(km[s][4 * i + 2], km[s][4 * i + 3]);
        }
      uint64_t *d = (uint64_t *)c.ws.get(tags[s], 32 * (size_t)(ns[s] + 1));
      up(d, km[s].data(), 32 * (size_t)ns[s]);
      dk[s] = d;
    }
    in.bMBR = dk[0];
    in.oMBR = dk[1];
    x.join_valid = false;
    x.jin = in;
    x.K = gcmf::mbr_filter(c, B, O, in.bMBR, in.oMBR, &x.counts, &x.pairs, &x.sv, in.bVNum,
                           in.oVNum);
    gcmf::join_refine(c, in, x.K, x.pairs, x.jr);
    x.join_valid = true;
    // the joined pairs (within U intersecting; the reference hands over the
    // within pairs only, spatialJoin3.cu:407)
    const long n = x.jr.nJoined;
    std::vector<int> jp(2 * (size_t)(n ? n : 1));
    if (n > 0) {
      HIP_CHECK(hipMemcpyAsync(jp.data(), x.jr.joinedPairs, 8 * (size_t)n, hipMemcpyDeviceToHost,
                               c.stream));
    }
    HIP_CHECK(hipStreamSynchronize(c.stream));
    int *pl = (int *)std::malloc(4 * (size_t)(n ? n : 1));
    int *ql = (int *)std::malloc(4 * (size_t)(n ? n : 1));
    if (!pl || !ql) {
      std::free(pl);
      std::free(ql);
      throw gcmf::Error{GCMF_ERR_ALLOC, "host allocation of the pair lists"};
    }
    for (long k = 0; k < n; k++) {
      pl[k] = jp[2 * k];
      ql[k] = jp[2 * k + 1];
    }
    *pIDList = pl;
    *qIDList = ql;
    *totalPairNum = n;
    *pBaseCoords = bc;
    *pOverlayCoords = oc;
    *pBVNum = bv;
    *pBVPSNum = bp;
    *pOVNum = ov;
    *pOVPSNum = op;
    own.keep = true;
    rc = 0;
  });
  return rc;
}

gcmf_ctx *gcmf_ref_context(void) {
  gcmf_ctx *r = nullptr;
  shim("gcmf_ref_context", [&](gcmf_ctx &x) { r = &x; });
  return r;
}

void gcmf_ref_release(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_ctx) gcmf_shutdown(g_ctx);
  g_ctx = nullptr;
}

}  // extern "C"
