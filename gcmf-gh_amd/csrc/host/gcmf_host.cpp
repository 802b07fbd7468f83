// gcmf_host.cpp -- host side of the join-and-clip path: the layer layout,
// CoordToMBR integer keys, the layer text format, and seeded synthetic
// generators for the benchmark workloads.  Plain C++17, no GPU.
//
// Layer layout (SEQ_Overlay.h:285-389, Types.h:1-8): coords AoS (x,y) f64
// with each ring's closing vertex repeated; VNum[i] = ring length including
// the closing vertex (int32); VPSNum[i] = inclusive prefix sum of VNum
// (int64); MBR[4i..4i+3] = u64 keys of (minx, miny, maxx, maxy) taken from
// the decimal tokens; MBR2 = the same tokens as doubles.
#include "../../../include/gcmf_host.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

struct gcmf_layer {
  std::vector<double> coords;
  std::vector<int> vnum;
  std::vector<long> vpsnum;
  std::vector<uint64_t> mbr;
  std::vector<double> mbr2;
};

namespace {

// splitmix64-seeded xoshiro256** (public-domain algorithms by Vigna).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (int i = 0; i < 4; i++) {
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {  // Box-Muller, one value per call
    double u1 = uniform(), u2 = uniform();
    if (u1 < 1e-300) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

// CoordToMBR type 1 (SEQ_Overlay.h:69-97): keep at most 12 fractional digits
// of the decimal token, drop '.', zero-pad to 12 fractional digits, atoll,
// add 9e15 (two's-complement wrap for negative tokens, as the reference).
uint64_t coord_to_mbr(const char *ct) {
  char buff[64];
  int i = 0, fCount = 0, fDot = 0;
  const int fracNum = 12;
  while (*ct && i < 48) {
    if (*ct != '.') {
      if (fCount < fracNum) {
        buff[i++] = *ct;
        if (fDot) fCount++;
      } else {
        break;
      }
    } else {
      fDot = 1;
    }
    ct++;
  }
  for (; fCount < fracNum; fCount++) buff[i++] = '0';
  buff[i] = '\0';
  uint64_t r = (uint64_t)std::atoll(buff);
  return r + 9000000000000000ULL;
}

// Appends one ring (n vertices, no closing vertex) to the layer.  When
// text_round_trip is set every coordinate is printed with %.15f and parsed
// back, and the MBR keys come from those tokens, as the reference's reader
// does for its text layers; otherwise keys come from %.15f of the doubles.
void append_ring(gcmf_layer *L, const double *xy, int n, bool text_round_trip) {
  char tok[64];
  double mn[2] = {0, 0}, mx[2] = {0, 0};
  std::string tmin[2], tmax[2];
  for (int k = 0; k < n; k++) {
    double v[2];
    std::string t[2];
    for (int d = 0; d < 2; d++) {
      std::snprintf(tok, sizeof(tok), "%.15f", xy[2 * k + d]);
      t[d] = tok;
      v[d] = text_round_trip ? std::strtod(tok, nullptr) : xy[2 * k + d];
    }
    L->coords.push_back(v[0]);
    L->coords.push_back(v[1]);
    for (int d = 0; d < 2; d++) {
      if (k == 0 || v[d] < mn[d]) { mn[d] = v[d]; tmin[d] = t[d]; }
      if (k == 0 || v[d] > mx[d]) { mx[d] = v[d]; tmax[d] = t[d]; }
    }
  }
  size_t first = L->coords.size() - 2 * (size_t)n;
  double x0 = L->coords[first], y0 = L->coords[first + 1];
  L->coords.push_back(x0);  // closing vertex
  L->coords.push_back(y0);
  long prev = L->vpsnum.empty() ? 0 : L->vpsnum.back();
  L->vnum.push_back(n + 1);
  L->vpsnum.push_back(prev + n + 1);
  L->mbr.push_back(coord_to_mbr(tmin[0].c_str()));
  L->mbr.push_back(coord_to_mbr(tmin[1].c_str()));
  L->mbr.push_back(coord_to_mbr(tmax[0].c_str()));
  L->mbr.push_back(coord_to_mbr(tmax[1].c_str()));
  L->mbr2.push_back(mn[0]);
  L->mbr2.push_back(mn[1]);
  L->mbr2.push_back(mx[0]);
  L->mbr2.push_back(mx[1]);
}

// Star-shaped simple ring: vertex k at angle 2*pi*k/n + U(-0.4,0.4)*2*pi/n,
// radius R*(1 + amp*U(-1,1)) about (cx, cy).  Draw order: angle jitter then
// radius jitter, per vertex.
void star(Rng &r, int n, double cx, double cy, double R, double amp, std::vector<double> &out) {
  out.resize(2 * (size_t)n);
  const double step = 2.0 * M_PI / n;
  for (int k = 0; k < n; k++) {
    double th = step * k + r.uniform(-0.4, 0.4) * step;
    double rad = R * (1.0 + amp * r.uniform(-1.0, 1.0));
    out[2 * k] = cx + rad * std::cos(th);
    out[2 * k + 1] = cy + rad * std::sin(th);
  }
}

// Large ring: radius 1 + 0.2*(1/6)*sum_{k<6} sin(2^k th + phi_k) + 0.02*U(-1,1).
void wavy(Rng &r, int n, double cx, double cy, std::vector<double> &out) {
  double phi[6];
  for (int k = 0; k < 6; k++) phi[k] = r.uniform(0.0, 2.0 * M_PI);
  out.resize(2 * (size_t)n);
  const double step = 2.0 * M_PI / n;
  for (int k = 0; k < n; k++) {
    double th = step * k + r.uniform(-0.4, 0.4) * step;
    double s = 0.0;
    for (int j = 0; j < 6; j++) s += std::sin((double)(1 << j) * th + phi[j]);
    double rad = 1.0 + 0.2 * (s / 6.0) + 0.02 * r.uniform(-1.0, 1.0);
    out[2 * k] = cx + rad * std::cos(th);
    out[2 * k + 1] = cy + rad * std::sin(th);
  }
}

// Snap to a grid and drop consecutive duplicates (cyclically), so no edge
// has zero length.
void snap(std::vector<double> &xy, double g) {
  std::vector<double> o;
  size_t n = xy.size() / 2;
  for (size_t k = 0; k < n; k++) {
    double x = std::round(xy[2 * k] / g) * g, y = std::round(xy[2 * k + 1] / g) * g;
    if (!o.empty() && o[o.size() - 2] == x && o[o.size() - 1] == y) continue;
    o.push_back(x);
    o.push_back(y);
  }
  while (o.size() >= 4 && o[0] == o[o.size() - 2] && o[1] == o[o.size() - 1]) o.resize(o.size() - 2);
  xy.swap(o);
}

}  // namespace

extern "C" {

uint64_t gcmf_coord_to_mbr(const char *token) { return coord_to_mbr(token); }

long gcmf_layer_npoly(const gcmf_layer *L) { return (long)L->vnum.size(); }
long gcmf_layer_nverts(const gcmf_layer *L) { return (long)(L->coords.size() / 2); }
const double *gcmf_layer_coords(const gcmf_layer *L) { return L->coords.data(); }
const int *gcmf_layer_vnum(const gcmf_layer *L) { return L->vnum.data(); }
const long *gcmf_layer_vpsnum(const gcmf_layer *L) { return L->vpsnum.data(); }
const uint64_t *gcmf_layer_mbr(const gcmf_layer *L) { return L->mbr.data(); }
const double *gcmf_layer_mbr2(const gcmf_layer *L) { return L->mbr2.data(); }
void gcmf_layer_free(gcmf_layer *L) { delete L; }

gcmf_layer *gcmf_layer_from_rings(const double *xy, const int *n, long npoly, int text_round_trip) {
  gcmf_layer *L = new gcmf_layer();
  long off = 0;
  for (long i = 0; i < npoly; i++) {
    append_ring(L, xy + 2 * off, n[i], text_round_trip != 0);
    off += n[i];
  }
  return L;
}

int gcmf_gen_join_layers(uint64_t seed, long nBase, long nOver, double rmin, double rmax,
                         int nmin, int nmax, gcmf_layer **base, gcmf_layer **over) {
  Rng r(seed);
  std::vector<double> ring;
  for (int side = 0; side < 2; side++) {
    long N = side == 0 ? nBase : nOver;
    gcmf_layer *L = new gcmf_layer();
    L->coords.reserve((size_t)N * 2 * 160);
    for (long i = 0; i < N; i++) {
      double cx = r.uniform(), cy = r.uniform();
      double R = std::exp(r.uniform(std::log(rmin), std::log(rmax)));
      int n = (int)std::floor(std::exp(r.uniform(std::log((double)nmin), std::log((double)nmax + 1.0))));
      if (n < nmin) n = nmin;
      if (n > nmax) n = nmax;
      star(r, n, cx, cy, R, 0.3, ring);
      append_ring(L, ring.data(), n, true);
    }
    (side == 0 ? *base : *over) = L;
  }
  return 0;
}

int gcmf_gen_pair_layers(uint64_t seed, long nPairs, double median, double sigma, int nmin,
                         int nmax, double overlap_frac, gcmf_layer **P, gcmf_layer **Q) {
  Rng r(seed);
  std::vector<long> order((size_t)nPairs);
  for (long i = 0; i < nPairs; i++) order[(size_t)i] = i;
  for (long i = nPairs - 1; i > 0; i--) {  // Fisher-Yates
    long j = (long)(r.next() % (uint64_t)(i + 1));
    std::swap(order[(size_t)i], order[(size_t)j]);
  }
  std::vector<char> overlap((size_t)nPairs, 0);
  long nov = (long)std::floor(overlap_frac * (double)nPairs);
  for (long i = 0; i < nov; i++) overlap[(size_t)order[(size_t)i]] = 1;
  gcmf_layer *LP = new gcmf_layer(), *LQ = new gcmf_layer();
  std::vector<double> ring;
  for (long i = 0; i < nPairs; i++) {
    int n[2];
    for (int s = 0; s < 2; s++) {
      double v = std::round(std::exp(std::log(median) + sigma * r.normal()));
      if (v < nmin) v = nmin;
      if (v > nmax) v = nmax;
      n[s] = (int)v;
    }
    double d = overlap[(size_t)i] ? r.uniform(0.0, 1.4) : 3.0;
    double ph = r.uniform(0.0, 2.0 * M_PI);
    star(r, n[0], 0.0, 0.0, 1.0, 0.3, ring);
    append_ring(LP, ring.data(), n[0], false);
    star(r, n[1], d * std::cos(ph), d * std::sin(ph), 1.0, 0.3, ring);
    append_ring(LQ, ring.data(), n[1], false);
  }
  *P = LP;
  *Q = LQ;
  return 0;
}

int gcmf_gen_single_pair(int kind, uint64_t seed, int nP, int nQ, double snap_grid,
                         gcmf_layer **P, gcmf_layer **Q) {
  Rng r(seed);
  std::vector<double> a, b;
  if (kind == 0) {  // star pair about (0,0) and (0.3,0.2), radius 1 +- 30%
    star(r, nP, 0.0, 0.0, 1.0, 0.3, a);
    star(r, nQ, 0.3, 0.2, 1.0, 0.3, b);
  } else {  // six-sinusoid rings about (0,0) and (0.3,0.2)
    wavy(r, nP, 0.0, 0.0, a);
    wavy(r, nQ, 0.3, 0.2, b);
  }
  if (snap_grid > 0.0) {
    snap(a, snap_grid);
    snap(b, snap_grid);
  }
  int na = (int)(a.size() / 2), nb = (int)(b.size() / 2);
  *P = gcmf_layer_from_rings(a.data(), &na, 1, 0);
  *Q = gcmf_layer_from_rings(b.data(), &nb, 1, 0);
  return 0;
}

// Writes the reference's text layer format (dataConverter.py:94-104 as read by
// SEQ_Overlay.h:285-389): a header line, then per polygon
// "id minx miny maxx maxy x0 y0 ... x0 y0" with %.15f tokens.
int gcmf_layer_write_text(const gcmf_layer *L, const char *path) {
  FILE *f = std::fopen(path, "w");
  if (!f) return -1;
  std::fprintf(f, "id,geometry\n");
  long off = 0;
  for (size_t i = 0; i < L->vnum.size(); i++) {
    std::fprintf(f, "%zu %.15f %.15f %.15f %.15f", i, L->mbr2[4 * i], L->mbr2[4 * i + 1],
                 L->mbr2[4 * i + 2], L->mbr2[4 * i + 3]);
    for (int k = 0; k < L->vnum[i]; k++)
      std::fprintf(f, " %.15f %.15f", L->coords[2 * (off + k)], L->coords[2 * (off + k) + 1]);
    std::fprintf(f, "\n");
    off += L->vnum[i];
  }
  std::fclose(f);
  return 0;
}

// Reads the text layer format (SEQ_Overlay.h:285-389): skips the header
// line; per line an id token then coordinate tokens made of [0-9.-]; the MBR
// is taken over every (x,y) of the line (first strict min/max token wins),
// the first two pairs (stored MBR corners) are dropped and the rest is the
// ring including its closing vertex.  Lines whose MBR has zero width or
// height are skipped ("MBR Error").  Returns NULL on I/O failure.
gcmf_layer *gcmf_layer_read_text(const char *path, long max_poly) {
  FILE *f = std::fopen(path, "r");
  if (!f) return nullptr;
  gcmf_layer *L = new gcmf_layer();
  std::string line;
  std::vector<char> buf(1 << 16);
  bool header = true;
  auto is_tok = [](char c) { return c == '.' || c == '-' || (c >= '0' && c <= '9'); };
  for (;;) {
    line.clear();
    bool got = false;
    while (std::fgets(buf.data(), (int)buf.size(), f)) {
      got = true;
      line += buf.data();
      if (!line.empty() && line.back() == '\n') break;
    }
    if (!got) break;
    if (header) { header = false; continue; }
    if (max_poly >= 0 && (long)L->vnum.size() >= max_poly) break;
    std::vector<std::string> toks;
    size_t i = 0;
    while (i < line.size()) {
      while (i < line.size() && !is_tok(line[i])) i++;
      size_t j = i;
      while (j < line.size() && is_tok(line[j])) j++;
      if (j > i) toks.emplace_back(line.substr(i, j - i));
      i = j;
    }
    if (toks.size() < 3) continue;
    size_t npairs = (toks.size() - 1) / 2;
    double mn[2] = {1000000, 1000000}, mx[2] = {-1000000, -1000000};
    std::string tmn[2], tmx[2];
    std::vector<double> v(2 * npairs);
    for (size_t k = 0; k < npairs; k++) {
      for (int d = 0; d < 2; d++) {
        const std::string &t = toks[1 + 2 * k + d];
        double x = std::atof(t.c_str());
        v[2 * k + d] = x;
        if (x < mn[d]) { mn[d] = x; tmn[d] = t; }
        if (x > mx[d]) { mx[d] = x; tmx[d] = t; }
      }
    }
    uint64_t key[4] = {coord_to_mbr(tmn[0].c_str()), coord_to_mbr(tmn[1].c_str()),
                       coord_to_mbr(tmx[0].c_str()), coord_to_mbr(tmx[1].c_str())};
    if (key[0] == key[2] || key[1] == key[3] || mn[0] == mx[0] || mn[1] == mx[1]) continue;
    if (npairs < 3) continue;
    long prev = L->vpsnum.empty() ? 0 : L->vpsnum.back();
    int ring = (int)(npairs - 2);
    for (size_t k = 2; k < npairs; k++) {
      L->coords.push_back(v[2 * k]);
      L->coords.push_back(v[2 * k + 1]);
    }
    L->vnum.push_back(ring);
    L->vpsnum.push_back(prev + ring);
    for (int d = 0; d < 4; d++) L->mbr.push_back(key[d]);
    L->mbr2.push_back(mn[0]);
    L->mbr2.push_back(mn[1]);
    L->mbr2.push_back(mx[0]);
    L->mbr2.push_back(mx[1]);
  }
  std::fclose(f);
  return L;
}

// ------------------------------------------------------------ results ----
double gcmf_ring_area(const double *xy, long n) {
  if (n < 3) return 0.0;
  const double x0 = xy[0], y0 = xy[1];
  double s = 0.0;
  for (long i = 1; i + 1 < n; i++) {
    const double ax = xy[2 * i] - x0, ay = xy[2 * i + 1] - y0;
    const double bx = xy[2 * i + 2] - x0, by = xy[2 * i + 3] - y0;
    s += ax * by - bx * ay;
  }
  return 0.5 * s;
}

int gcmf_result_areas(const double *xy, const int *comp_sizes, long ncomps, double *out) {
  if (ncomps < 0 || (ncomps > 0 && (!xy || !comp_sizes || !out))) return -1;
  long v = 0;
  for (long c = 0; c < ncomps; c++) {
    if (comp_sizes[c] < 0) return -1;
    out[c] = gcmf_ring_area(xy + 2 * v, comp_sizes[c]);
    v += comp_sizes[c];
  }
  return 0;
}

int gcmf_result_write_poly(const char *path, const double *xy, const int *comp_sizes,
                           long ncomps) {
  if (!path || ncomps < 0 || (ncomps > 0 && (!xy || !comp_sizes))) return -1;
  FILE *f = std::fopen(path, "w");
  if (!f) return -1;
  long v = 0;
  for (long c = 0; c < ncomps; c++)
    for (int i = 0; i < comp_sizes[c]; i++, v++)
      std::fprintf(f, "%.17g %.17g%s\n", xy[2 * v], xy[2 * v + 1],
                   i + 1 < comp_sizes[c] ? "," : ";");
  return std::fclose(f) == 0 ? 0 : -1;
}

namespace {
// crossing-number point-in-ring test (the rule of polygon::pointInPoly's
// parity without its on-edge cases)
bool ring_contains(const double *r, int n, double x, double y) {
  bool in = false;
  for (int i = 0, j = n - 1; i < n; j = i++) {
    const double xi = r[2 * i], yi = r[2 * i + 1], xj = r[2 * j], yj = r[2 * j + 1];
    if ((yi > y) != (yj > y) && x < (xj - xi) * (y - yi) / (yj - yi) + xi) in = !in;
  }
  return in;
}

void wkt_ring(FILE *f, const double *r, int n) {
  std::fputc('(', f);
  for (int i = 0; i <= n; i++) {
    const int k = i < n ? i : 0;  // WKT repeats the first vertex
    std::fprintf(f, "%s%.17g %.17g", i ? ", " : "", r[2 * k], r[2 * k + 1]);
  }
  std::fputc(')', f);
}
}  // namespace

int gcmf_result_write_wkt(const char *path, const double *xy, const int *comp_sizes,
                          const int *pair_ncomp, const int *pair_status, long npairs,
                          int nest) {
  if (!path || npairs < 0 || (npairs > 0 && (!pair_ncomp || !pair_status))) return -1;
  FILE *f = std::fopen(path, "w");
  if (!f) return -1;
  long v = 0, c = 0;
  std::vector<const double *> ring;
  std::vector<int> len, depth, parent;
  for (long k = 0; k < npairs; k++) {
    const int nc = pair_ncomp[k];
    ring.assign(nc, nullptr);
    len.assign(nc, 0);
    for (int i = 0; i < nc; i++, c++) {
      ring[i] = xy + 2 * v;
      len[i] = comp_sizes[c];
      v += comp_sizes[c];
    }
    // even-odd nesting: a ring inside an even number of the others bounds
    // a polygon, one inside an odd number is a hole of the innermost ring
    // of one less depth that holds it (the depth is taken at the midpoint
    // of the ring's first edge)
    depth.assign(nc, 0);
    parent.assign(nc, -1);
    if (nest && nc > 1) {
      for (int i = 0; i < nc; i++) {
        if (len[i] < 2) continue;
        const double mx = 0.5 * (ring[i][0] + ring[i][2]), my = 0.5 * (ring[i][1] + ring[i][3]);
        std::vector<int> in;
        for (int j = 0; j < nc; j++)
          if (j != i && len[j] >= 3 && ring_contains(ring[j], len[j], mx, my)) in.push_back(j);
        depth[i] = (int)in.size();
        if (depth[i] % 2 == 1) {
          for (int j : in) {
            int dj = 0;  // depth of j: the rings of `in` that hold j's point
            const double px = 0.5 * (ring[j][0] + ring[j][2]), py = 0.5 * (ring[j][1] + ring[j][3]);
            for (int h : in)
              if (h != j && ring_contains(ring[h], len[h], px, py)) dj++;
            if (dj == depth[i] - 1) parent[i] = j;
          }
        }
      }
    }
    std::fprintf(f, "%ld\t%d\t", k, pair_status[k]);
    int npoly = 0;
    for (int i = 0; i < nc; i++)
      if (parent[i] < 0 && depth[i] % 2 == 0) npoly++;
    if (npoly == 0) {
      std::fputs("MULTIPOLYGON EMPTY\n", f);
      continue;
    }
    std::fputs("MULTIPOLYGON (", f);
    int w = 0;
    for (int i = 0; i < nc; i++) {
      if (parent[i] >= 0 || depth[i] % 2 != 0) continue;
      std::fputs(w++ ? ", (" : "(", f);
      wkt_ring(f, ring[i], len[i]);
      for (int h = 0; h < nc; h++)
        if (parent[h] == i) {
          std::fputs(", ", f);
          wkt_ring(f, ring[h], len[h]);
        }
      std::fputc(')', f);
    }
    std::fputs(")\n", f);
  }
  return std::fclose(f) == 0 ? 0 : -1;
}

}  // extern "C"
