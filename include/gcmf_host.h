/*
 * gcmf_host.h -- host-side (CPU) part of the join-and-clip path: the layer
 * layout, CoordToMBR keys, the reference's text layer format and seeded
 * synthetic generators.  Library: libgcmf_host.so (no GPU needed).
 *
 * Layer layout kept from the reference (SEQ_Overlay.h:285-389, Types.h:1-8):
 *   coords  f64 [2*sum(VNum)]  AoS (x,y), each ring's closing vertex repeated
 *   VNum    i32 [N]            ring length including the closing vertex
 *   VPSNum  i64 [N]            inclusive prefix sum of VNum
 *   MBR     u64 [4N]           CoordToMBR keys of (minx, miny, maxx, maxy)
 *   MBR2    f64 [4N]           the same coordinates as doubles
 */
#ifndef GCMF_HOST_H
#define GCMF_HOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gcmf_layer gcmf_layer;

/* Replaces CoordToMBR(token, 1), SEQ_Overlay.h:69-97. */
uint64_t gcmf_coord_to_mbr(const char *token);

long gcmf_layer_npoly(const gcmf_layer *L);
long gcmf_layer_nverts(const gcmf_layer *L);
const double *gcmf_layer_coords(const gcmf_layer *L);
const int *gcmf_layer_vnum(const gcmf_layer *L);
const long *gcmf_layer_vpsnum(const gcmf_layer *L);
const uint64_t *gcmf_layer_mbr(const gcmf_layer *L);
const double *gcmf_layer_mbr2(const gcmf_layer *L);
void gcmf_layer_free(gcmf_layer *L);

/* Builds a layer from rings given without closing vertices (xy packed, n[i]
 * vertices each).  text_round_trip: print each coordinate with %.15f and
 * parse it back, keys from those tokens (the reference's text path). */
gcmf_layer *gcmf_layer_from_rings(const double *xy, const int *n, long npoly,
                                  int text_round_trip);

/* Replaces ReadTextFormatPolygon2WithVector, SEQ_Overlay.h:285-389 (without
 * the linked-list side output).  max_poly < 0 reads every line. */
gcmf_layer *gcmf_layer_read_text(const char *path, long max_poly);
int gcmf_layer_write_text(const gcmf_layer *L, const char *path);

/* Two join layers of star-shaped polygons: centres U[0,1]^2, radius
 * log-uniform in [rmin, rmax], vertex count floor(exp(U(ln nmin, ln(nmax+1)))),
 * coordinates through the %.15f text round trip.  Base layer drawn first. */
int gcmf_gen_join_layers(uint64_t seed, long nBase, long nOver, double rmin,
                         double rmax, int nmin, int nmax, gcmf_layer **base,
                         gcmf_layer **over);

/* nPairs independent pairs (P[i], Q[i]): vertex counts
 * round(lognormal(ln median, sigma)) clamped to [nmin, nmax], unit-radius
 * stars; a seeded shuffle picks floor(overlap_frac*nPairs) pairs whose Q is
 * offset by U(0,1.4) (overlapping), the others by 3 (disjoint). */
int gcmf_gen_pair_layers(uint64_t seed, long nPairs, double median, double sigma,
                         int nmin, int nmax, double overlap_frac, gcmf_layer **P,
                         gcmf_layer **Q);

/* One pair as two one-polygon layers.  kind 0: star rings (radius 1 +- 30%)
 * about (0,0) and (0.3,0.2); kind 1: six-sinusoid rings.  snap_grid > 0
 * snaps to that grid and drops consecutive duplicate vertices. */
int gcmf_gen_single_pair(int kind, uint64_t seed, int nP, int nQ, double snap_grid,
                         gcmf_layer **P, gcmf_layer **Q);

/* ------------------------------------------------------------ results ----
 * A batch's results as host arrays: xy (x, y of every result vertex, the
 * components one after another, each vertex once), comp_sizes (vertices per
 * component), pair_ncomp / pair_status (per pair, in pair order). */
/* Signed area of one ring of n vertices (counter-clockwise positive):
 * 0.5 * sum over the fan from vertex 0 of the cross products of
 * (v_i - v_0, v_i+1 - v_0).  The reference has no area routine. */
double gcmf_ring_area(const double *xy, long n);
/* gcmf_ring_area of every component into out[ncomps]. */
int gcmf_result_areas(const double *xy, const int *comp_sizes, long ncomps, double *out);
/* savePolygon's .poly text (polyclip.h:859-865, polygon.h:575-584): one
 * "x y," line per vertex, the component's last vertex ending in ";", at
 * %.17g, so that every double round-trips. */
int gcmf_result_write_poly(const char *path, const double *xy, const int *comp_sizes,
                           long ncomps);
/* One line per pair: "pair<TAB>status<TAB>" and the pair's result as a WKT
 * MULTIPOLYGON (each ring closed by repeating its first vertex, %.17g;
 * MULTIPOLYGON EMPTY when there is none).  nest = 0: every component is a
 * polygon of its own (the intersection of two simple rings has no holes);
 * nest = 1: components are nested by the even-odd rule, a ring inside an
 * odd number of others becoming a hole of the innermost ring holding it
 * (complex polygons, union). */
int gcmf_result_write_wkt(const char *path, const double *xy, const int *comp_sizes,
                          const int *pair_ncomp, const int *pair_status, long npairs,
                          int nest);

#ifdef __cplusplus
}
#endif
#endif
