// ref_api_cxx.cpp -- the reference-named entry points once more, with C++
// linkage.  The reference declares them in C++ headers without extern "C"
// (main.h:5-11, Join.h, GH-CUDA/ghcuda.h:6-14) and compiles its callers as
// C++ (main.cpp, spatialJoin3.cu), so an unchanged caller refers to their
// mangled names.  Each function here forwards to the extern "C" entry point
// of the same name (ref_api.hip), bound by its symbol name: the C++ name of
// the declaration (gcmf_c_*) differs, so both linkages coexist.
//
// Parameter types are the reference's (Types.h:1-8: mbr_t unsigned long
// long, coord_t double, poly_size_t short), which fixes the mangled names.

typedef unsigned long long mbr_t;
typedef double coord_t;
typedef short int poly_size_t;

#define C_NAME(sym) __asm__(#sym)

extern "C" {
long gcmf_c_SortBaseMBROverlap(long, long, mbr_t *, mbr_t *, int **, int **, char,
                               char) C_NAME(SortBaseMBROverlap);
void gcmf_c_GetCMBR(long, int *, coord_t *, coord_t *, coord_t **, int **, int **, char **,
                    int **, long *) C_NAME(GetCMBR);
long gcmf_c_PointInPolygonTest(coord_t *, coord_t *, long, long, int *, int *, char *, long *,
                               long *, int *, int *) C_NAME(PointInPolygonTest);
void gcmf_c_CountCMF(coord_t *, coord_t *, long, int *, int *, int *, int *, long *, long *,
                     coord_t *, long **, long **, long **, int **, poly_size_t **,
                     poly_size_t **, long *, long *) C_NAME(CountCMF);
void gcmf_c_ApplyCMF(coord_t *, coord_t *, long, int *, long, int *, int *, int *, long *, long *,
                     coord_t *, long *, long *, poly_size_t *, poly_size_t *) C_NAME(ApplyCMF);
long gcmf_c_SegmentIntersectJoin(coord_t *, coord_t *, long, int *, int *, long *, long *, long *,
                                 long *, poly_size_t *, poly_size_t *,
                                 int **) C_NAME(SegmentIntersectJoin);
void gcmf_c_calculateIntersections(coord_t *, coord_t *, int, int, coord_t *, int *, long *, int *,
                                   long *, int, int, int *, int *, coord_t **, coord_t **, int **,
                                   int **, int **, int **, int **, int **,
                                   int *) C_NAME(calculateIntersections);
int gcmf_c_GH_CUDA(coord_t *, coord_t *, int *, long *, int *, long *, int, int,
                   int) C_NAME(GH_CUDA);
int gcmf_c_ghcuda(int *, int *, int, coord_t *, coord_t *, int *, long *, int *,
                  long *) C_NAME(ghcuda);
int gcmf_c_ST_Intersect(long, long, coord_t *, coord_t *, int *, int *, long *, long *, mbr_t *,
                        mbr_t *, coord_t *, coord_t *, int *) C_NAME(ST_Intersect);
int gcmf_c_spatialJoin(int, char **, int **, int **, long *, coord_t **, coord_t **, int **,
                       long **, int **, long **) C_NAME(spatialJoin);
}

// Join.h:5
long SortBaseMBROverlap(long bPolNum, long oPolNum, mbr_t *dbMBR, mbr_t *doMBR, int **djxyCounter,
                        int **djxyVector, char dimSort, char dimSelect) {
  return gcmf_c_SortBaseMBROverlap(bPolNum, oPolNum, dbMBR, doMBR, djxyCounter, djxyVector,
                                   dimSort, dimSelect);
}

// Join.h:397
void GetCMBR(long pairNum, int *jxyVector, coord_t *bMBR, coord_t *oMBR, coord_t **cMBR,
             int **jpipIndexList, int **pipFlag, char **pipType, int **joinFlag, long *pipNum) {
  gcmf_c_GetCMBR(pairNum, jxyVector, bMBR, oMBR, cMBR, jpipIndexList, pipFlag, pipType, joinFlag,
                 pipNum);
}

// Join.h:437
long PointInPolygonTest(coord_t *bCoords, coord_t *oCoords, long pairNum, long pipNum,
                        int *jxyVector, int *jPiPIndexList, char *pipType, long *bVPSNum,
                        long *oVPSNum, int *pipFlag, int *joinFlag) {
  return gcmf_c_PointInPolygonTest(bCoords, oCoords, pairNum, pipNum, jxyVector, jPiPIndexList,
                                   pipType, bVPSNum, oVPSNum, pipFlag, joinFlag);
}

// Join.h:157
void CountCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, int *joinFlag,
              int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, coord_t *cMBR,
              long **bEdgePSCounter, long **oEdgePSCounter, long **workLoadPSCounter,
              int **jxy2IndexList, poly_size_t **bEdgeList, poly_size_t **oEdgeList, long *eiNum,
              long *workLoadNum) {
  gcmf_c_CountCMF(bCoords, oCoords, pairNum, jxyVector, joinFlag, bVNum, oVNum, bVPSNum, oVPSNum,
                  cMBR, bEdgePSCounter, oEdgePSCounter, workLoadPSCounter, jxy2IndexList,
                  bEdgeList, oEdgeList, eiNum, workLoadNum);
}

// Join.h:257
void ApplyCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, long eiNum,
              int *jxy2IndexList, int *bVNum, int *oVNum, long *bPFVNum, long *oPFVNum,
              coord_t *cMBR, long *bEdgePSCounter, long *oEdgePSCounter, poly_size_t *bEdgeList,
              poly_size_t *oEdgeList) {
  gcmf_c_ApplyCMF(bCoords, oCoords, pairNum, jxyVector, eiNum, jxy2IndexList, bVNum, oVNum,
                  bPFVNum, oPFVNum, cMBR, bEdgePSCounter, oEdgePSCounter, bEdgeList, oEdgeList);
}

// Join.h:356
long SegmentIntersectJoin(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jCompactVector,
                          int *jxy2IndexList, long *bPFVNum, long *oPFVNum, long *bEdgePSCounter,
                          long *oEdgePSCounter, poly_size_t *bEdgeList, poly_size_t *oEdgeList,
                          int **segIntersectJoinFlag) {
  return gcmf_c_SegmentIntersectJoin(bCoords, oCoords, pairNum, jCompactVector, jxy2IndexList,
                                     bPFVNum, oPFVNum, bEdgePSCounter, oEdgePSCounter, bEdgeList,
                                     oEdgeList, segIntersectJoinFlag);
}

// GH-CUDA/ghcuda.h:6-14
void calculateIntersections(coord_t *baseCoord, coord_t *overlayCoords, int sizeP, int sizeQ,
                            coord_t *cmbr, int *dbVNum, long *dbVPSNum, int *doVNum,
                            long *doVPSNum, int bPID, int oPID, int *countNonDegenIntP,
                            int *countNonDegenIntQ, coord_t **intersectionsP,
                            coord_t **intersectionsQ, int **alphaValuesP, int **alphaValuesQ,
                            int **initLabelsP, int **initLabelsQ, int **neighborP,
                            int **neighborQ, int *invalid) {
  gcmf_c_calculateIntersections(baseCoord, overlayCoords, sizeP, sizeQ, cmbr, dbVNum, dbVPSNum,
                                doVNum, doVPSNum, bPID, oPID, countNonDegenIntP,
                                countNonDegenIntQ, intersectionsP, intersectionsQ, alphaValuesP,
                                alphaValuesQ, initLabelsP, initLabelsQ, neighborP, neighborQ,
                                invalid);
}

// GH-CUDA/ghcuda.cpp:341
int GH_CUDA(coord_t *bCoords, coord_t *oCoords, int *pBVNum, long *pBVPSNum, int *pOVNum,
            long *pOVPSNum, int bPID, int oPID, int swapped) {
  return gcmf_c_GH_CUDA(bCoords, oCoords, pBVNum, pBVPSNum, pOVNum, pOVPSNum, bPID, oPID, swapped);
}

// main.h:8-11
int ghcuda(int pIDList[], int qIDList[], int totalNumPairs, coord_t *bCoords, coord_t *oCoords,
           int *pBVNum, long *pBVPSNum, int *pOVNum, long *pOVPSNum) {
  return gcmf_c_ghcuda(pIDList, qIDList, totalNumPairs, bCoords, oCoords, pBVNum, pBVPSNum,
                       pOVNum, pOVPSNum);
}

// ST_Intersect.cu:18-19
int ST_Intersect(long bPolNum, long oPolNum, coord_t *baseCoords, coord_t *overlayCoords,
                 int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, mbr_t *seqMBR,
                 mbr_t *seqOMBR, coord_t *seqMBR2, coord_t *seqOMBR2, int *jPairs) {
  return gcmf_c_ST_Intersect(bPolNum, oPolNum, baseCoords, overlayCoords, bVNum, oVNum, bVPSNum,
                             oVPSNum, seqMBR, seqOMBR, seqMBR2, seqOMBR2, jPairs);
}

// main.h:5-7
int spatialJoin(int argc, char *argv[], int **pIDList, int **qIDList, long *totalPairNum,
                coord_t **pBaseCoords, coord_t **pOverlayCoords, int **pBVNum, long **pBVPSNum,
                int **pOVNum, long **pOVPSNum) {
  return gcmf_c_spatialJoin(argc, argv, pIDList, qIDList, totalPairNum, pBaseCoords,
                            pOverlayCoords, pBVNum, pBVPSNum, pOVNum, pOVPSNum);
}
