// A C++ caller of the reference's join-and-clip call surface, declared as
// the reference's headers declare it (main.h:5-11, Join.h:5-437,
// GH-CUDA/ghcuda.h:6-14, GH-CUDA/ghcuda.cpp:341, ST_Intersect.cu:18-19): C++
// linkage, the reference's types (Types.h:1-8).  tests/test_host.py links it
// against libgcmf_hip.so; the link succeeds only when every mangled name is
// exported.  It is never run.
typedef unsigned long long mbr_t;
typedef double coord_t;
typedef short int poly_size_t;

int spatialJoin(int argc, char *argv[], int **pIDList, int **qIDList, long *totalPairNum,
                coord_t **pBaseCoords, coord_t **pOverlayCoords, int **pBVNum, long **pBVPSNum,
                int **pOVNum, long **pOVPSNum);
int ghcuda(int pIDList[], int qIDList[], int totalNumPairs, coord_t *bCoords, coord_t *oCoords,
           int *pBVNum, long *pBVPSNum, int *pOVNum, long *pOVPSNum);
long SortBaseMBROverlap(long bPolNum, long oPolNum, mbr_t *dbMBR, mbr_t *doMBR, int **djxyCounter,
                        int **djxyVector, char dimSort, char dimSelect);
void CountCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, int *joinFlag,
              int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, coord_t *cMBR,
              long **bEdgePSCounter, long **oEdgePSCounter, long **workLoadPSCounter,
              int **jxy2IndexList, poly_size_t **bEdgeList, poly_size_t **oEdgeList, long *eiNum,
              long *workLoadNum);
void ApplyCMF(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jxyVector, long eiNum,
              int *jxy2IndexList, int *bVNum, int *oVNum, long *bPFVNum, long *oPFVNum,
              coord_t *cMBR, long *bEdgePSCounter, long *oEdgePSCounter, poly_size_t *bEdgeList,
              poly_size_t *oEdgeList);
long SegmentIntersectJoin(coord_t *bCoords, coord_t *oCoords, long pairNum, int *jCompactVector,
                          int *jxy2IndexList, long *bPFVNum, long *oPFVNum, long *bEdgePSCounter,
                          long *oEdgePSCounter, poly_size_t *bEdgeList, poly_size_t *oEdgeList,
                          int **segIntersectJoinFlag);
void GetCMBR(long pairNum, int *jxyVector, coord_t *bMBR, coord_t *oMBR, coord_t **cMBR,
             int **jpipIndexList, int **pipFlag, char **pipType, int **joinFlag, long *pipNum);
long PointInPolygonTest(coord_t *bCoords, coord_t *oCoords, long pairNum, long pipNum,
                        int *jxyVector, int *jPiPIndexList, char *pipType, long *bVPSNum,
                        long *oVPSNum, int *pipFlag, int *joinFlag);
void calculateIntersections(coord_t *baseCoord, coord_t *overlayCoords, int sizeP, int sizeQ,
                            coord_t *cmbr, int *dbVNum, long *dbVPSNum, int *doVNum,
                            long *doVPSNum, int bPID, int oPID, int *countNonDegenIntP,
                            int *countNonDegenIntQ, coord_t **intersectionsP,
                            coord_t **intersectionsQ, int **alphaValuesP, int **alphaValuesQ,
                            int **initLabelsP, int **initLabelsQ, int **neighborP,
                            int **neighborQ, int *invalid);
int GH_CUDA(coord_t *bCoords, coord_t *oCoords, int *pBVNum, long *pBVPSNum, int *pOVNum,
            long *pOVPSNum, int bPID, int oPID, int swapped);
int ST_Intersect(long bPolNum, long oPolNum, coord_t *baseCoords, coord_t *overlayCoords,
                 int *bVNum, int *oVNum, long *bVPSNum, long *oVPSNum, mbr_t *seqMBR,
                 mbr_t *seqOMBR, coord_t *seqMBR2, coord_t *seqOMBR2, int *jPairs);

int main() {
  void *volatile f[] = {(void *)&spatialJoin,          (void *)&ghcuda,
                        (void *)&SortBaseMBROverlap,   (void *)&CountCMF,
                        (void *)&ApplyCMF,             (void *)&SegmentIntersectJoin,
                        (void *)&GetCMBR,              (void *)&PointInPolygonTest,
                        (void *)&calculateIntersections, (void *)&GH_CUDA,
                        (void *)&ST_Intersect};
  return f[0] == 0;
}
