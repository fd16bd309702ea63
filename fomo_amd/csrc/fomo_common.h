// fomo_common.h — device-side types and helpers shared by the engine's translation units, and
// the fused NN + integration kernel k_nn_los. fomo_device.hip (every other kernel and the host
// code) and fomo_nnlos_part.hip (k_nn_los instantiated for a group of lambda chunk sizes, compiled
// once per group so that the instantiations build in parallel) include it.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "lookup_geom.h"
#include "octree.h"
#include "rect_geom.h"
#include "spectral_exp.h"

// Physical constants (fomo-CHIANTI.cpp:16-19, fomo-thomson.cpp:31-34, FoMo-internal.h:4; GSL
// CODATA-2006 values)
#define FOMO_C_LIGHT 2.99792458e8
#define FOMO_K_BOLTZ 1.3806504e-23
#define FOMO_M_PROTON 1.672621637e-27
#define FOMO_ALPHA 1.0645
#define FOMO_MM_PER_ARCSEC 0.715
#define FOMO_RSUN 695.7
#define FOMO_U_LIMB 0.58
#define FOMO_RE 2.8179e-13
#define FOMO_LUM 2.8e10

namespace fomo_dev {

// ------------------------------------------------------------------------------------
// shared structs (host + device)
// ------------------------------------------------------------------------------------
struct GridInfo {
  double o[3];      // origin (data bbox minimum)
  double h;         // cell edge
  double invh;      // 1/h
  double slack;     // conservative margin on cell-boundary distances
  float bmin[3], bmax[3];
  int dims[3];
  int pad;
  double sc[3];     // bounding sphere of the points: centre (the bbox centre) and squared radius,
  double sr2;       // rounded up (every point p has |p - sc|^2 <= sr2)
};

// A search box [lo, hi] (float bounds) that holds no data point: it misses the data's bounding
// box, or lies wholly outside its bounding sphere (the exact query would return -1 for it).
__device__ __forceinline__ bool box_misses_data(const GridInfo& g, float lo0, float lo1, float lo2, float hi0,
                                                float hi1, float hi2) {
  if (!(hi0 > g.bmin[0] && lo0 < g.bmax[0] && hi1 > g.bmin[1] && lo1 < g.bmax[1] && hi2 > g.bmin[2] &&
        lo2 < g.bmax[2]))
    return true;
  const double c0 = fmin(fmax(g.sc[0], (double)lo0), (double)hi0) - g.sc[0];
  const double c1 = fmin(fmax(g.sc[1], (double)lo1), (double)hi1) - g.sc[1];
  const double c2 = fmin(fmax(g.sc[2], (double)lo2), (double)hi2) - g.sc[2];
  return (c0 * c0 + c1 * c1) + c2 * c2 > g.sr2;
}

// Candidate lookup grid: a finer grid over the data bbox whose cells list the only points that
// can be the (unconstrained) nearest point of any target inside the cell.
struct LookupInfo {
  double o[3];               // origin (shifted off the cell-list lattice)
  double h[3], invh[3];      // lookup cell edges (the x edge is up to twice the others)
  double eps;                // box expansion covering the rounding of a target's cell assignment
  int dims[3];
  int enabled;
  // outer groups: candidate lists for targets beyond the grid (ohdr = nullptr: none)
  double oo[3], oinvh[3];    // origin and inverse edges of the outer group lattice
  int odims[3];
  const int2* ohdr;          // per outer group: (start in lcand, count), count 0 for none
  // far groups whose cells all use the group's list (ghdr = nullptr: none): read before the record
  int gcell[3], gdims[3];    // lookup cells per group edge, groups per axis
  const int2* ghdr;          // per group: (start in lcand, count), count 0 when a cell has its own
  // rectilinear datasets (rect != 0; no candidate grid, enabled == 0): the points are the product
  // grid of the axis values rv[a], and a point's sorted position is its grid position
  // i + n0 (j + n1 k) (rect_geom.h)
  int rect;
  // masked grids (mrect != 0): the points are a subset of the product grid of rv; gpos gives the
  // sorted position of the point on each grid node (-1: none). The candidate grid is built too,
  // for the targets whose nearest grid node holds no point.
  int mrect;
  const int32_t* gpos;
  int rn[3];                 // values per axis (bins: kRectBinsPerValue times as many)
  float ro[3], rinv[3];      // bin origin and inverse bin width per axis
  double rdmax;              // largest nearest distance without a possible tie two values away
  const float* rv;           // axis values, ascending, the three axes one after another
  const float4* rbin;        // bin tables {index as int bits, v[index], v[index + 1], v[index + 2]},
                             // the three axes one after another
};

// Start grid of the tetra walk (a finite tet near each cell centre) and the points' bounding box,
// which contains the convex hull: a sample outside the box is outside the hull without a walk.
struct TetGrid {
  double o[3], h, invh;
  double lo[3], hi[3];
  int dims[3];
};

// Per-view geometry, filled on the device by k_view_finalize.
struct ViewGeom {
  double cb, sb, cl, sl;          // trig of b and l for the LOS samples (host glibc)
  double ext[6];                  // minx,maxx,miny,maxy,minz,maxz of rotated points
  double md;                      // search-box half width
  double deltaz;                  // LOS sample spacing
  double pathlength, apix, scale; // output scaling
};

enum { MODE_NN = 0, MODE_THOMSON = 1, MODE_TET = 2, MODE_LINEAR = 3, MODE_PROJ = 4 };

// Spectral lines rendered together from one NN result (blockIdx.y selects the line).
constexpr int kMaxLines = 8;
// Per-point records for the NN passes: {vx, vy, vz, T} then the peak of every line. A line's
// fwhm is not stored: it is recomputed exactly as the emission kernel computed it,
// (float)(K * sqrt((double)T)), or 1 for imaging cubes (K == 0) (fomo-CHIANTI.cpp:220-225, :284-285);
// cubes made from a given GoftCube (K < 0) store the fwhm itself in the T slot.
struct RecInfo {
  double K[kMaxLines];
};

// The square-root factor a record's lines share: sqrt(T) for cubes computed from T, the T slot
// itself for cubes made from a given GoftCube (K < 0: the slot holds the fwhm). K0 is the first
// line's K; the lines of one record set are all of one kind.
__host__ __device__ __forceinline__ double rec_st(double K0, float tslot) {
  return K0 < 0 ? (double)tslot : sqrt((double)tslot);
}
// A line's fwhm from that factor: (float)(K sqrt T) as the emission kernel computed it, the given
// fwhm (K < 0), or 1 for imaging cubes (K == 0, fomo-CHIANTI.cpp:284-285).
__host__ __device__ __forceinline__ float rec_fw(double K, double st) {
  return K > 0 ? (float)(K * st) : (K < 0 ? (float)st : 1.f);
}

struct LineSet {
  const float2* em[kMaxLines];  // (peak, fwhm): per sorted point (record build) or per sample (K5)
  float* out[kMaxLines];        // R*nl output per line
  double wla[kMaxLines];        // lambda window width in Angstrom
  double l0c[kMaxLines];        // lambda0 / c
  int n;
};

constexpr int kBlock = 256;
constexpr int kMaxPartials = 2048;  // rotate kernel grid cap (grid-stride beyond)

// ------------------------------------------------------------------------------------
// small device helpers
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int cell_coord(const GridInfo& g, int a, double v) {
  double c = floor((v - g.o[a]) * g.invh);
  c = fmin(fmax(c, 0.0), (double)(g.dims[a] - 1));
  return (int)c;
}

template <class T> __device__ __forceinline__ T wave_min(T v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
template <class T> __device__ __forceinline__ T wave_max(T v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}


constexpr int kLookupCap = 24;
// Lookup cell record: kRecSlots float4 (128 bytes, one L2 line: an L2 miss fetches an aligned
// 128-byte line) holding up to kRecSlots candidates (x, y, z, sorted position as int bits; unused
// slots have w = -1). Slot 0 with w = -1: no candidates (the query falls back to the exact
// search); w = -2: more than kRecSlots candidates, at lcand[x] .. (y of them, x and y as int bits).
// Queries load the first kSlots slots ahead of time and the rest, from the same line, only when
// slot kSlots - 1 says that more follow (its w = -3 - sorted position).
constexpr int kSlots = 4;
constexpr int kRecSlots = 8;
constexpr int kLookupMargin = 6;  // cell-list cells the lookup grid extends beyond the data bbox
// cells farther than this from the nearest non-empty cell take their far group's list instead
// of a list of their own (their own search box would span too many cells)
constexpr int kLookupNearRings = 2;
// cap of the cell list's distance transform: the ring search of a cell at distance d visits rings
// d and d + 1
constexpr int kLookupDistCap = kLookupNearRings + 1;

using fomo::box_maxdist2;
using fomo::box_mindist2;
using fomo::dominated;
using fomo::node_dominated;


// Lookup cell of a target (-1 outside the lookup grid).
__device__ __forceinline__ int64_t lookup_cell(const LookupInfo& lk, float t0, float t1, float t2) {
  const double v0 = ((double)t0 - lk.o[0]) * lk.invh[0];
  const double v1 = ((double)t1 - lk.o[1]) * lk.invh[1];
  const double v2 = ((double)t2 - lk.o[2]) * lk.invh[2];
  if (!(v0 >= 0.0 && v0 < (double)lk.dims[0] && v1 >= 0.0 && v1 < (double)lk.dims[1] && v2 >= 0.0 &&
        v2 < (double)lk.dims[2]))
    return -1;
  return ((int64_t)(int)v2 * lk.dims[1] + (int)v1) * lk.dims[0] + (int)v0;
}

// The nearest candidate of a lookup cell record sl (read from the overflow list for long lists),
// if it lies strictly inside the search box; else -2.
__device__ __forceinline__ int lookup_pick(const float4* __restrict__ pts, const float4* __restrict__ lcand,
                                           const float4* __restrict__ rp, const float4 (&sl)[kSlots], float t0,
                                           float t1, float t2, float lo0,
                                           float lo1, float lo2, float hi0, float hi1, float hi2) {
  const int w0 = __float_as_int(sl[0].w);
  if (w0 == -1) return -2;
  // candidates hold their coordinates inline (x, y, z, sorted position); the original index
  // (tie-break) is read from pts only on an exact distance tie
  double best = INFINITY;
  int bpos = -1;
  float4 bq = make_float4(0.f, 0.f, 0.f, 0.f);
  const auto consider = [&](const float4 q) {
    const double d0 = (double)q.x - (double)t0;
    const double d1 = (double)q.y - (double)t1;
    const double d2 = (double)q.z - (double)t2;
    const double d = (d0 * d0 + d1 * d1) + d2 * d2;
    const int p = __float_as_int(q.w);
    bool take = d < best;
    if (d == best && bpos >= 0) take = __float_as_int(pts[p].w) < __float_as_int(pts[bpos].w);
    if (take) {
      best = d;
      bpos = p;
      bq = q;
    }
  };
  if (w0 != -2) {
    const int wl = __float_as_int(sl[kSlots - 1].w);
#pragma unroll
    for (int q = 0; q < kSlots - 1; ++q)
      if (__float_as_int(sl[q].w) >= 0) consider(sl[q]);
    if (wl >= 0) consider(sl[kSlots - 1]);
    if (wl <= -3) {
      consider(make_float4(sl[kSlots - 1].x, sl[kSlots - 1].y, sl[kSlots - 1].z, __int_as_float(-3 - wl)));
      // the rest of the record: the same 128-byte line as sl
      float4 s2[kRecSlots - kSlots];
#pragma unroll
      for (int q = 0; q < kRecSlots - kSlots; ++q) s2[q] = rp[kSlots + q];
#pragma unroll
      for (int q = 0; q < kRecSlots - kSlots; ++q)
        if (__float_as_int(s2[q].w) >= 0) consider(s2[q]);
    }
  } else {
    // four candidates per step: their loads are in flight together
    const float4* ov = lcand + __float_as_int(sl[0].x);
    const int n = __float_as_int(sl[0].y);
    for (int i = 0; i < n; i += 4) {
      float4 c4[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i + u < n) c4[u] = ov[i + u];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (i + u < n) consider(c4[u]);
    }
  }
  if (lo0 < bq.x && bq.x < hi0 && lo1 < bq.y && bq.y < hi1 && lo2 < bq.z && bq.z < hi2) return bpos;
  return -2;
}

// Outer group of a target beyond the lookup grid (-1 outside the outer lattice).
__device__ __forceinline__ int64_t outer_group(const LookupInfo& lk, float t0, float t1, float t2) {
  const double v0 = ((double)t0 - lk.oo[0]) * lk.oinvh[0];
  const double v1 = ((double)t1 - lk.oo[1]) * lk.oinvh[1];
  const double v2 = ((double)t2 - lk.oo[2]) * lk.oinvh[2];
  if (!(v0 >= 0.0 && v0 < (double)lk.odims[0] && v1 >= 0.0 && v1 < (double)lk.odims[1] && v2 >= 0.0 &&
        v2 < (double)lk.odims[2]))
    return -1;
  return ((int64_t)(int)v2 * lk.odims[1] + (int)v1) * lk.odims[0] + (int)v0;
}

// Targets beyond the lookup grid: the list of their outer group, if it has one (else -2).
__device__ __forceinline__ int outer_lookup(const float4* __restrict__ pts, const float4* __restrict__ lcand,
                                            const LookupInfo& lk, float t0, float t1, float t2, float lo0,
                                            float lo1, float lo2, float hi0, float hi1, float hi2) {
  const int64_t og = outer_group(lk, t0, t1, t2);
  if (og < 0) return -2;
  const int2 hd = lk.ohdr[og];
  if (hd.y <= 0) return -2;
  const float4 h4 = make_float4(__int_as_float(hd.x), __int_as_float(hd.y), 0.f, __int_as_float(-2));
  const float4 sl[kSlots] = {h4, h4, h4, h4};  // an overflow header: only slot 0 is read
  return lookup_pick(pts, lcand, nullptr, sl, t0, t1, t2, lo0, lo1, lo2, hi0, hi1, hi2);
}

// Rectilinear datasets, first step: the bin table entries of the target (one 16-byte read per
// axis, from tables small enough to stay in the caches).
__device__ __forceinline__ void rect_issue(const LookupInfo& lk, float t0, float t1, float t2, float4 (&e)[kSlots]) {
  constexpr int B = fomo::kRectBinsPerValue;
  const int nb0 = B * lk.rn[0], nb1 = B * lk.rn[1], nb2 = B * lk.rn[2];
  e[0] = lk.rbin[fomo::rect_bin(lk.ro[0], lk.rinv[0], nb0, t0)];
  e[1] = lk.rbin[nb0 + fomo::rect_bin(lk.ro[1], lk.rinv[1], nb1, t1)];
  e[2] = lk.rbin[nb0 + nb1 + fomo::rect_bin(lk.ro[2], lk.rinv[2], nb2, t2)];
}

// Rectilinear datasets, second step: the nearest point's sorted position if it lies strictly
// inside the search box, else -2 (rect_pick).
__device__ __forceinline__ int rect_resolve(const float4* __restrict__ pts, const LookupInfo& lk,
                                            const float4 (&e)[kSlots], const fomo::SearchBox& box) {
  return fomo::rect_pick(
      e[0], e[1], e[2], lk.rn[0], lk.rn[1], lk.rn[2], lk.rdmax, box,
      [&](int a, int k) { return lk.rv[(a > 0 ? lk.rn[0] : 0) + (a > 1 ? lk.rn[1] : 0) + k]; },
      [&](int64_t gp) { return __float_as_int(pts[gp].w); });
}

// Masked grids: the nearest point from the grid, or -2 (nearest point outside the search box) or
// -3 (no point on the nearest grid node: the caller uses the candidate grid).
__device__ __forceinline__ int mrect_resolve(const float4* __restrict__ pts, const LookupInfo& lk,
                                             const float4 (&e)[kSlots], const fomo::SearchBox& box) {
  return fomo::mrect_pick(
      e[0], e[1], e[2], lk.rn[0], lk.rn[1], lk.rn[2], lk.rdmax, box,
      [&](int a, int k) { return lk.rv[(a > 0 ? lk.rn[0] : 0) + (a > 1 ? lk.rn[1] : 0) + k]; },
      [&](int64_t g) { return lk.gpos[g]; }, [&](int sp) { return __float_as_int(pts[sp].w); });
}

// Fast path: the unconstrained nearest point from the target's lookup cell. Returns its sorted
// position if it lies strictly inside the search box, else -2 (the caller runs the exact query).
__device__ __forceinline__ int nn_lookup(const float4* __restrict__ pts, const float4* __restrict__ lslot,
                                         const float4* __restrict__ lcand, const LookupInfo& lk, float t0, float t1,
                                         float t2, float lo0, float lo1, float lo2, float hi0, float hi1, float hi2) {
  if (lk.rect || lk.mrect) {
    float4 e[kSlots];
    rect_issue(lk, t0, t1, t2, e);
    fomo::SearchBox box;
    box.t0 = t0, box.t1 = t1, box.t2 = t2;
    box.lo0 = lo0, box.lo1 = lo1, box.lo2 = lo2, box.hi0 = hi0, box.hi1 = hi1, box.hi2 = hi2;
    if (lk.rect) return rect_resolve(pts, lk, e, box);
    const int r = mrect_resolve(pts, lk, e, box);
    if (r != -3 || !lk.enabled) return r == -3 ? -2 : r;
  }
  const int64_t L = lookup_cell(lk, t0, t1, t2);
  if (L < 0) return lk.ohdr ? outer_lookup(pts, lcand, lk, t0, t1, t2, lo0, lo1, lo2, hi0, hi1, hi2) : -2;
  if (lk.ghdr) {
    // a group whose cells all point at its list: its 8-byte header instead of the cell's record
    const int lx = (int)(L % lk.dims[0]);
    const int ly = (int)((L / lk.dims[0]) % lk.dims[1]);
    const int lz = (int)(L / ((int64_t)lk.dims[0] * lk.dims[1]));
    const int2 hd = lk.ghdr[((int64_t)(lz / lk.gcell[2]) * lk.gdims[1] + ly / lk.gcell[1]) * lk.gdims[0] +
                            lx / lk.gcell[0]];
    if (hd.y > 0) {
      const float4 h4 = make_float4(__int_as_float(hd.x), __int_as_float(hd.y), 0.f, __int_as_float(-2));
      const float4 sg[kSlots] = {h4, h4, h4, h4};  // an overflow header: only slot 0 is read
      return lookup_pick(pts, lcand, nullptr, sg, t0, t1, t2, lo0, lo1, lo2, hi0, hi1, hi2);
    }
  }
  // the first half of the 128-byte record; the rest and the overflow list only when taken
  const float4* rp = lslot + (int64_t)kRecSlots * L;
  float4 sl[kSlots];
#pragma unroll
  for (int q = 0; q < kSlots; ++q) sl[q] = rp[q];
  return lookup_pick(pts, lcand, rp, sl, t0, t1, t2, lo0, lo1, lo2, hi0, hi1, hi2);
}


struct BvhInfo {
  int64_t nodes;  // 0: no tree
  int64_t n;      // points
  // 256-byte records of the internal nodes, or null (wide_walk_wave): per node, its children in
  // slots 0..7, slot s being wide[16 r + 2 s] = (box min, ref) and wide[16 r + 2 s + 1] =
  // (box max, count) with ref and count as int bits; count 0: internal child whose record is ref,
  // count 1..8: leaf of mpts[ref .. ref + count); an empty slot has an empty box (min +inf,
  // max -inf). Record 0 is the root.
  const float4* wide;
};


using fomo::SearchBox;

__device__ __forceinline__ int bvh_query(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                                         const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                         const BvhInfo& bv, const SearchBox& sb, double best = INFINITY,
                                         int bpos = -1, int borig = 0x7fffffff) {
  return fomo::oct_query(pts, mpts, nlo, nhi, bv.nodes, sb, best, bpos, borig);
}

// Whether the box [l, h] can hold a point strictly inside the search box at a squared distance
// <= bestf (the running best rounded up, bvh_bound). The test is in float: the squared box
// distance has a relative error below 1e-6 (exact differences of nearby floats, then three
// rounded squares and two rounded sums), so with a 1e-5 margin and the bound rounded up it
// admits every box oct_enter admits. Admitting more boxes cannot change an answer: their extra
// leaves only add points that lose to the nearest one.
__device__ __forceinline__ bool box_may_improve(const float4& l, const float4& h, const SearchBox& sb, float bestf) {
  const float d0 = fmaxf(fmaxf(l.x - sb.t0, sb.t0 - h.x), 0.f);
  const float d1 = fmaxf(fmaxf(l.y - sb.t1, sb.t1 - h.y), 0.f);
  const float d2 = fmaxf(fmaxf(l.z - sb.t2, sb.t2 - h.z), 0.f);
  const float md = (d0 * d0 + d1 * d1) + d2 * d2;
  return sb.node_overlaps(l, h) && md * 0.99999f <= bestf;
}
__device__ __forceinline__ float bvh_bound(double best) { return (float)best * 1.000001f; }

// One point of mpts (x, y, z, sorted position) against a lane's running best: kept when strictly
// inside the box and nearer, or as near with a lower original index.
__device__ __forceinline__ void wave_scan_point(const float4* __restrict__ pts, const float4& q, const SearchBox& sb,
                                                double& best, int& bpos, int& borig) {
  if (!sb.holds(q)) return;
  const double e0 = (double)q.x - (double)sb.t0;
  const double e1 = (double)q.y - (double)sb.t1;
  const double e2 = (double)q.z - (double)sb.t2;
  const double d = (e0 * e0 + e1 * e1) + e2 * e2;
  const int p = __float_as_int(q.w);
  if (d < best) {
    best = d;
    bpos = p;
    borig = -1;
  } else if (d == best && bpos >= 0) {
    // borig of the running best is read from pts only when a tie needs it (-1: unread)
    if (borig < 0) borig = __float_as_int(pts[bpos].w);
    const int oi = __float_as_int(pts[p].w);
    if (oi < borig) {
      bpos = p;
      borig = oi;
    }
  }
}

// The points of one leaf, mpts[first .. first + cnt) with cnt <= 8, against the running best of
// every lane with in set: one load, lane u & 7 holding point u, then every such lane scans them in
// leaf order, as leaf_scan does. Every lane of the wave must call it.
__device__ __forceinline__ void wave_leaf_scan(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                                               int first, int cnt, const SearchBox& sb, bool in, double& best,
                                               int& bpos, int& borig) {
  const auto pick = [](float v, int o) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), o));
  };
  const int lane = (int)(threadIdx.x & 63);
  const float4 q8 = mpts[first + min(lane & 7, cnt - 1)];
  for (int u = 0; u < cnt; ++u) {
    const float4 q = make_float4(pick(q8.x, u), pick(q8.y, u), pick(q8.z, u), pick(q8.w, u));
    if (in) wave_scan_point(pts, q, sb, best, bpos, borig);
  }
}

// Wave-cooperative walk of the preorder node range [e0, e1) of an n-node tree, from the running
// best (best, bpos, borig; borig < 0: not read yet). Every lane with act set ends with the nearest
// point strictly inside its box among the leaves of the range, the running best included: a leaf
// in the range is either reached or inside the subtree of a node the walk tested and rejected, so
// any contiguous range is a complete walk of its own leaves. A node is entered when any active
// lane could find a better point in it; a lane's own test fails on every node inside a subtree
// it would have skipped (a child's box lies inside its parent's), so each lane scans exactly the
// leaves of its own walk, in the same order. The node index is wave-uniform, so the node records
// are read once per wave instead of once per lane; the 64 queries of a wave are neighbouring
// samples and enter nearly the same nodes. Every lane of the wave must call it (the node window
// and the leaf points are loaded by all 64 lanes and read back with lane reads).
// budget > 0: the walk stops after that many node steps and returns true (wave-uniform) with
// the running best so far; a restart from node e0 with it finishes the query.
__device__ bool bvh_walk_wave(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                              const float4* __restrict__ nlo, const float4* __restrict__ nhi, int n, int e0, int e1,
                              const SearchBox& sb, bool act, double& best, int& bpos, int& borig, int budget) {
  float bestf = bvh_bound(best);
  const int lane = (int)(threadIdx.x & 63);
  // The walk only moves forward (e + 1 or a skip pointer past the subtree), so node records are
  // read 64 at a time, one per lane, into a window; a step reads its node from the window with
  // lane reads instead of a dependent load. The window moves to e when e leaves it.
  int base = e0 - 64, steps = 0;
  float4 wl = make_float4(0.f, 0.f, 0.f, 0.f), wh = wl;
  const auto pick = [](float v, int o) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), o));
  };
  for (int e = e0; e < e1;) {
    e = __builtin_amdgcn_readfirstlane(e);
    if (budget > 0 && ++steps > budget) return true;
    if (e - base >= 64) {
      base = e;
      const int me = min(e + lane, n - 1);
      wl = nlo[me];
      wh = nhi[me];
    }
    const int o = e - base;
    const float4 l = make_float4(pick(wl.x, o), pick(wl.y, o), pick(wl.z, o), pick(wl.w, o));
    const float4 h = make_float4(pick(wh.x, o), pick(wh.y, o), pick(wh.z, o), pick(wh.w, o));
    const int w = __float_as_int(l.w);
    const bool in = act && box_may_improve(l, h, sb, bestf);
    const bool any = __ballot(in) != 0;
    if (w < 0) {
      if (any) {
        wave_leaf_scan(pts, mpts, __float_as_int(h.w), -w, sb, in, best, bpos, borig);
        bestf = bvh_bound(best);
      }
      ++e;
    } else {
      e = any ? e + 1 : w;
    }
  }
  return false;
}

// Entries of a wave's wide_walk_wave stack (two float4 each). Expanding a node pops one entry
// and pushes at most 8, so a tree whose internal nodes lie at depth <= kWideDepth needs at most
// 7 kWideDepth + 8 entries; the builder makes wide records only for such trees.
constexpr int kWideStack = 128;
constexpr int kWideDepth = 17;

// Wave-cooperative exact query over the wide records of the octree (BvhInfo::wide), from the
// running best (best, bpos, borig; borig < 0: not read yet): every lane with act set ends with the
// nearest point strictly inside its box (distance, then lowest original index), oct_query's
// answer. Depth first with a stack in LDS (stk: kWideStack entries of this wave). Expanding a node
// takes all its children's boxes in one 256-byte load (lane l holding slot l & 7), and every lane
// tests the 8 children. The entered leaves are scanned at once: their points, contiguous per leaf,
// come in one more load (lane l holding point l & 7 of slot l >> 3). The entered internal children
// are pushed, the farthest from the first active lane's target first, so the nearest is expanded
// next, and an entry is tested again against the bounds of the moment when it is popped. A lane's
// answer is found because every node on the path to it passes that lane's own test at any time
// (its box holds the answer, which is no farther than the lane's running best). Against the
// preorder walk, an expansion costs two dependent loads where that walk reloads its window of
// records at every skip past it and loads each leaf on its own. budget > 0: stops after that many
// expansions and returns true (wave-uniform) with the running best so far, as bvh_walk_wave does.
// Every lane of the wave must call it.
__device__ __forceinline__ bool wide_walk_wave(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                                               const float4* __restrict__ wide, const SearchBox& sb, bool act,
                                               double& best, int& bpos, int& borig, int budget,
                                               float4* __restrict__ stk) {
  const int lane = (int)(threadIdx.x & 63);
  const int slot = lane & 7;
  const unsigned long long am = __ballot(act);
  if (am == 0) return false;
  // the first active lane's target: the children are expanded nearest first for it
  const int leader = (int)__builtin_ctzll(am);
  const auto bcast = [](float v, int o) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), o));
  };
  const float lt0 = bcast(sb.t0, leader), lt1 = bcast(sb.t1, leader), lt2 = bcast(sb.t2, leader);
  float bestf = bvh_bound(best);
  int sp = 0, steps = 0, r = 0;  // r: record to expand (-1: pop the next entry)
  for (;;) {
    if (r >= 0) {
      if (budget > 0 && ++steps > budget) return true;
      const float4 a = wide[(size_t)r * 16 + 2 * slot], b = wide[(size_t)r * 16 + 2 * slot + 1];
      unsigned mask = 0, mine = 0;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float4 l = make_float4(bcast(a.x, c), bcast(a.y, c), bcast(a.z, c), 0.f);
        const float4 h = make_float4(bcast(b.x, c), bcast(b.y, c), bcast(b.z, c), 0.f);
        const bool in = act && box_may_improve(l, h, sb, bestf);
        mine |= (unsigned)in << c;
        if (__ballot(in) != 0) mask |= 1u << c;
      }
      const unsigned leaves = (unsigned)__ballot(lane < 8 && __float_as_int(b.w) > 0) & 0xffu;
      const unsigned lmask = mask & leaves, imask = mask & ~leaves;
      if (lmask != 0) {
        const int c = lane >> 3, u = lane & 7;
        const int lref = __shfl(__float_as_int(a.w), c, 64), lcnt = __shfl(__float_as_int(b.w), c, 64);
        float4 q = make_float4(0.f, 0.f, 0.f, 0.f);
        if (((lmask >> c) & 1u) && u < lcnt) q = mpts[lref + u];
#pragma unroll
        for (int c2 = 0; c2 < 8; ++c2) {
          if (!((lmask >> c2) & 1u)) continue;
          const int n2 = __builtin_amdgcn_readlane(__float_as_int(b.w), c2);
          const bool in = (mine >> c2) & 1u;
          for (int u2 = 0; u2 < n2; ++u2) {
            const int o = 8 * c2 + u2;
            const float4 p = make_float4(bcast(q.x, o), bcast(q.y, o), bcast(q.z, o), bcast(q.w, o));
            if (in) wave_scan_point(pts, p, sb, best, bpos, borig);
          }
        }
        bestf = bvh_bound(best);
      }
      const int cnt = __builtin_popcount(imask);
      if (sp + cnt > kWideStack) return true;  // not reached for trees of depth <= kWideDepth
      // lane l < 8 holds slot l: its distance from the leader's target orders the pushes
      const float d0 = fmaxf(fmaxf(a.x - lt0, lt0 - b.x), 0.f);
      const float d1 = fmaxf(fmaxf(a.y - lt1, lt1 - b.y), 0.f);
      const float d2 = fmaxf(fmaxf(a.z - lt2, lt2 - b.z), 0.f);
      const float key = (d0 * d0 + d1 * d1) + d2 * d2;
      int rank = 0;  // entered internal slots pushed before this one: farther, or as far with a higher slot
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const float kc = bcast(key, c);
        rank += ((imask >> c) & 1u) && (kc > key || (kc == key && c > slot));
      }
      if (lane < 8 && ((imask >> lane) & 1u)) {
        stk[2 * (sp + rank)] = a;
        stk[2 * (sp + rank) + 1] = b;
      }
      sp += cnt;
      r = -1;
    }
    if (sp == 0) break;
    --sp;
    const float4 l = stk[2 * sp], h = stk[2 * sp + 1];
    if (__ballot(act && box_may_improve(l, h, sb, bestf)) == 0) continue;
    r = __builtin_amdgcn_readfirstlane(__float_as_int(l.w));  // an internal child's record
  }
  return false;
}

// Wave-cooperative oct_query: every lane with act set gets exactly oct_query's answer (see
// bvh_walk_wave; all lanes of the wave must call it).
__device__ int bvh_query_wave(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                              const float4* __restrict__ nlo, const float4* __restrict__ nhi, const BvhInfo& bv,
                              const SearchBox& sb, bool act, double best, int bpos, int borig) {
  if (bv.nodes <= 0) return bpos;
  if (act && bpos < 0) fomo::oct_descend(pts, mpts, nlo, nhi, sb, best, bpos, borig);
  bvh_walk_wave(pts, mpts, nlo, nhi, (int)bv.nodes, 0, (int)bv.nodes, sb, act, best, bpos, borig, 0);
  return bpos;
}


// The Gaussian line profile exp(-kap (d0 + c dl)^2) over lambda pixels c = 0, 1, ... by ratio
// recurrence. profile_args: xa = {aq, ar, a0}, the exponents of the ratio of ratios, of the first
// ratio (pixel 1 over pixel 0) and of pixel 0; false when the recurrence could leave the double
// range (the caller then evaluates every pixel directly).
__device__ __forceinline__ bool profile_args(double kap, double d0, double dl, double (&xa)[3]) {
  const double a0 = -kap * d0 * d0;
  const double ar = -kap * (2.0 * d0 * dl + dl * dl);
  const double aq = -2.0 * kap * dl * dl;
  xa[0] = aq, xa[1] = ar, xa[2] = a0;
  return a0 > -600.0 && fabs(ar) < 100.0 && aq > -30.0;
}

// Stride-4 chains of the profile from ex = {q, r, g_0} = exp(xa): chain j holds the pixels
// j, j+4, j+8, ..., starting at g[j] with ratio R[j]; one step is g *= R, R *= q16.
//   g_{c+4} = g_c R_c,  R_{c+4} = R_c q^16,  R_0 = r_0 r_1 r_2 r_3 = r^4 q^6
// Every kernel that integrates the profile builds its chains here, so that the same pixel gets
// the same value in all of them.
__device__ __forceinline__ void profile_chains(const double (&ex)[3], double (&g)[4], double (&R)[4], double& q16) {
  const double q = ex[0], ra = ex[1];
  double r = ra;
  g[0] = ex[2];
  g[1] = g[0] * r;
  r *= q;
  g[2] = g[1] * r;
  r *= q;
  g[3] = g[2] * r;
  const double q2 = q * q, q4 = q2 * q2;
  q16 = (q4 * q4) * (q4 * q4);
  const double ra2 = ra * ra;
  R[0] = (ra2 * ra2) * (q4 * q2);
  R[1] = R[0] * q4;
  R[2] = R[1] * q4;
  R[3] = R[2] * q4;
}


// One term of a spectral sum: peak * g added to the float sum of a (ray, line, pixel), rounded
// once (the reference sums its profiles in float: intens, fomo-nearestneighbour.cpp:128, 213).
// Every kernel that integrates a profile adds through here, so the fused and the separate passes
// give the same bits.
__device__ __forceinline__ float spectral_add(float acc, float peak, double g) { return fmaf(peak, (float)g, acc); }

// Orders the LDS accesses of one wave around it: what its lanes wrote before is visible to the
// reads of any lane after it (no code motion across, LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// f(integral_constant<int, I>) for I = 0 .. N-1, unrolled at compile time
template <int N, int I = 0, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<N, I + 1>(f);
  }
}

// Lane O of each aligned group of G lanes, to every lane of the group. G = 2, 4: a DPP quad
// permutation (a VALU operand modifier, no LDS round trip; every lane of the quad must be
// active); G = 8: ds_bpermute.
template <int G, int O>
__device__ __forceinline__ int group_bcast(int x) {
  static_assert(O >= 0 && O < G, "lane outside the group");
  if constexpr (G == 4) {
    return __builtin_amdgcn_update_dpp(0, x, O * 0x55, 0xF, 0xF, false);
  } else if constexpr (G == 2) {
    return __builtin_amdgcn_update_dpp(0, x, O | (O << 2) | ((2 + O) << 4) | ((2 + O) << 6), 0xF, 0xF, false);
  } else {
    return __shfl(x, (int)(threadIdx.x & 63 & ~(G - 1)) + O, 64);
  }
}
template <int G, int O>
__device__ __forceinline__ float group_bcast(float x) {
  return __int_as_float(group_bcast<G, O>(__float_as_int(x)));
}
template <int G, int O>
__device__ __forceinline__ double group_bcast(double x) {
  const int lo = group_bcast<G, O>(__double2loint(x)), hi = group_bcast<G, O>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}

// ------------------------------------------------------------------------------------
// Fused NN query + spectral integration (k_nn, k_nn_far and k_los_spectral in one pass, for
// 2 <= nl <= 64: several lines, or one line on views of at least kFusedMinRays rays). G lanes per ray query G consecutive samples at once (lane q:
// sample kb + q; lookup grid, then the exact octree query) and hand each sample's values to the
// other lanes of the ray by cross-lane broadcasts, so the per-sample values never go through HBM
// and the memory-bound queries of some waves overlap the FP64 integration of others. Two ways
// to split the integration over the G lanes:
//   SL == 0: lane q integrates line q (G = lines rounded up to 2, 4, 8) with the step of
//            k_los_spectral, from the broadcast (LOS velocity, sqrt T, peak).
//   SL > 0:  the lines share one profile (equal lambda0, lambda window and thermal-width factor,
//            so their Gaussians differ only in the peak; lines beyond ls.n have peak 0), or there
//            is one line (SL = 1). G = 4:
//            lane q evaluates the anchors of its own sample, the ray's 4 lanes exchange them,
//            and lane q runs the profile's chain q (pixels q, q+4, ...) of every sample, adding
//            peak_l * g to line l's sums. The profile is evaluated once for all lines.
// Both add, for each (ray, line, pixel), the same values in the same order as k_nn +
// k_los_spectral (profile_chains, then spectral_add), so a batch of L lines and L
// single-line renders give identical images. As in k_los_spectral, a (ray, line) with a sample
// whose recurrence anchors would under- or overflow is flagged (flags[line * R + ray]) and redone
// by the DIRECT instantiation, launched right after, with one exp per sample and lambda pixel.
// ------------------------------------------------------------------------------------
// One wave of the pass: rays wave * RPW .. of view `view`. The main pass (DIRECT = false) lists
// the waves that flagged a ray in dlist (dlist[0]: count, dlist[2 ..]: view * waves per view +
// wave), for the direct pass.
template <int RS4, int CH, int G, bool DIRECT, int SL, int NP, bool RECT>
__device__ __forceinline__ void nn_los_wave(const int view, const int64_t wave, const float4* __restrict__ pts,
                                            const GridInfo& g, const float4* __restrict__ lslot,
                                            const float4* __restrict__ lcand, const LookupInfo& lk,
                                            const float4* __restrict__ mpts, const float4* __restrict__ nlo,
                                            const float4* __restrict__ nhi, const BvhInfo& bv,
                                            const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                            const float4* __restrict__ rec, const RecInfo& ri, const LineSet& ls,
                                            int nl, int32_t* __restrict__ flags, int64_t out_vs, int64_t flag_vs,
                                            int32_t* __restrict__ dlist) {
  constexpr int NL = 4 * RS4 - 4;  // lines a record holds
  constexpr int RPW = 64 / G;      // rays per wave
  static_assert(SL == 0 || (G == 4 && SL <= NL && CH % 4 == 0), "a shared profile runs 4 chains per ray");
  static_assert(NP == 1 || (SL > 0 && NP * SL <= 8), "several lambda passes only with a shared profile");
  constexpr int CJ = CH / 4;                  // pixels of one chain (SL > 0)
  // SL > 0: acc[(p * SL + l) * CJ + m] is pass p, line l, pixel p * CH + 4 m + q
  constexpr int NACC = SL ? NP * SL * CJ : CH;
  const int lane = threadIdx.x & 63;
  const int q = lane & (G - 1);
  const int64_t R = (int64_t)nx * ny;
  // SL > 0 (not DIRECT): every lane writes, for its own sample, the starts of the profile's 4
  // chains (g, R) and q16 of each lambda pass, the lines' peaks and the pass bits to its LDS slot;
  // lane q of the ray then reads chain q of each of the ray's 4 samples. These are the values of
  // profile_chains at the owner's anchors, as every lane computing them itself would get.
  constexpr bool kXch = SL > 0 && !DIRECT;
  constexpr int kSlot = NP * 80 + 32;  // per pass {(g, R) x 4, q16, pad}, then peaks (float4), bits
  __shared__ __align__(16) unsigned char xch[kXch ? 256 * kSlot : 16];
  const int64_t ray = wave * RPW + lane / G;
  const bool live = ray < R;
  // the view's geometry vgp[view], outputs ls.out[l] + view * out_vs, ray flags flags + view *
  // flag_vs
  const int64_t ooff = (int64_t)view * out_vs;
  int32_t* const flag = flags + (int64_t)view * flag_vs + (size_t)(SL > 0 ? 0 : q) * R + (live ? ray : 0);
  // flagged (ray, line, lambda pass)s of this lane: bit l * NP + p (shared profile; flags[l * R +
  // ray] holds the passes' bits), bit 0 (line q). The direct pass reads them first: usually none
  // is set, and whole waves leave before any other work.
  unsigned need = 0;
  if (DIRECT) {
    if (live && (SL > 0 || q < ls.n)) {
      if constexpr (SL > 0) {
#pragma unroll
        for (int l = 0; l < SL; ++l)
          if (l < ls.n) need |= ((unsigned)flag[(size_t)l * R] & ((1u << NP) - 1)) << (l * NP);
      } else {
        need = *flag != 0 ? 1u : 0u;
      }
    }
    if (__ballot(need != 0) == 0) return;
  }
  const int i = live ? (int)(ray / nx) : 0, j = live ? (int)(ray % nx) : 0;
  const ViewGeom vg = vgp[view];
  const double minx = vg.ext[0], maxx = vg.ext[1], miny = vg.ext[2], maxy = vg.ext[3], minz = vg.ext[4];
  const double xx = double(j) / (nx - 1) * (maxx - minx) + minx;  // fomo-nearestneighbour.cpp:152
  const double yy = double(i) / (ny - 1) * (maxy - miny) + miny;  // :153
  const double cb = vg.cb, sb = vg.sb, cl = vg.cl, sl_ = vg.sl, md = vg.md;
  // the profile constants: line q's, selected without lane-dependent indexing of the kernel
  // arguments, or (shared profile) line 0's
  double Kq = 0.0, wla = 0.0, l0c = 0.0;
  const double K0 = ri.K[0];
  float* out = nullptr;
  if constexpr (SL > 0) {
    Kq = ri.K[0], wla = ls.wla[0], l0c = ls.l0c[0];
  } else {
#pragma unroll
    for (int l = 0; l < kMaxLines; ++l)
      if (l == q) {
        Kq = ri.K[l];
        wla = ls.wla[l];
        l0c = ls.l0c[l];
        out = ls.out[l] + ooff;
      }
  }
  bool integ = live && (SL > 0 || q < ls.n);
  if (DIRECT) integ = integ && need != 0;
  unsigned bad = 0;  // bits as in need: recurrence anchors out of range
  const double dl = wla / (nl - 1);
  const double lamc = double(0) / (nl - 1) * wla - wla / 2.;  // lambdaval of the first pixel (:200)
  // lambdaval of the first pixel of pass p (k_los_spectral's l0 = p * CH)
  const auto lamc_of = [&](int p) { return double(p * CH) / (nl - 1) * wla - wla / 2.; };
  float acc[NACC];
#pragma unroll
  for (int c = 0; c < NACC; ++c) acc[c] = 0.f;
  int prev = -1;  // this lane's last resolved point: the warm start of the exact query
  // Pipeline over groups of G samples. While group kb is integrated, with a lookup grid: the point
  // record of group kb + G (winner picked from its lookup cell record) and the lookup cell record
  // of group kb + 2G are in flight; on a rectilinear dataset: the point records of groups kb + G
  // and kb + 2G and the bin entries of group kb + 3G.
  // stage 1: target, search box and lookup cell record of this lane's sample k
  const auto query_issue = [&](int k, SearchBox& box, const float4*& rp, float4 (&sl)[kSlots], bool& act,
                               bool& cell) {
    act = live && k < nz;
    cell = false;
    if (!act) return;
    const double zz = double(k) * vg.deltaz + minz;             // :162
    const double p0 = xx * cb * cl + yy * sl_ + zz * sb * cl;   // :165
    const double p1 = -xx * cb * sl_ + yy * cl - zz * sb * sl_;
    const double p2 = -xx * sb + zz * cb;
    box.t0 = (float)p0, box.t1 = (float)p1, box.t2 = (float)p2;  // :171
    box.lo0 = (float)(p0 - md), box.lo1 = (float)(p1 - md), box.lo2 = (float)(p2 - md);  // :179
    box.hi0 = (float)(p0 + md), box.hi1 = (float)(p1 + md), box.hi2 = (float)(p2 + md);
    if constexpr (RECT) {
      // rectilinear dataset: the bin table entries of the target's three axes
      cell = true;
      rect_issue(lk, box.t0, box.t1, box.t2, sl);
      sl[3] = make_float4(0.f, 0.f, 0.f, 0.f);
      return;
    }
    if (!lk.enabled) return;
    const int64_t L = lookup_cell(lk, box.t0, box.t1, box.t2);
    if (L < 0) return;
    cell = true;
    rp = lslot + (int64_t)kRecSlots * L;
#pragma unroll
    for (int c = 0; c < kSlots; ++c) sl[c] = rp[c];
  };
  // stage 2: the winner (lookup, else the exact octree query), -1 for none
  const auto query_resolve = [&](const SearchBox& box, const float4* rp, const float4 (&sl)[kSlots], bool act,
                                 bool cell) -> int {
    if (!act) return -1;
    int r = -2;
    if (cell) {
      if constexpr (RECT)
        r = rect_resolve(pts, lk, sl, box);
      else
        r = lookup_pick(pts, lcand, rp, sl, box.t0, box.t1, box.t2, box.lo0, box.lo1, box.lo2, box.hi0, box.hi1,
                        box.hi2);
    }
    if (r == -2) {
      if (box_misses_data(g, box.lo0, box.lo1, box.lo2, box.hi0, box.hi1, box.hi2)) {
        r = -1;  // no data point can be inside the box
      } else {
        double best = INFINITY;
        int bpos = -1, borig = 0x7fffffff;
        if (prev >= 0) {
          const float4 pq = pts[prev];
          if (box.holds(pq)) {
            const double d0 = (double)pq.x - (double)box.t0, d1 = (double)pq.y - (double)box.t1,
                         d2 = (double)pq.z - (double)box.t2;
            best = (d0 * d0 + d1 * d1) + d2 * d2;
            bpos = prev;
            borig = __float_as_int(pq.w);
          }
        }
        r = bvh_query(pts, mpts, nlo, nhi, bv, box, best, bpos, borig);
      }
    }
    if (r >= 0) prev = r;
    return r;
  };
  const auto record_issue = [&](int r, float (&f)[4 * RS4]) {
#pragma unroll
    for (int c = 0; c < 4 * RS4; ++c) f[c] = 0.f;
    if (r < 0) return;
    const float4* rp = rec + (size_t)r * RS4;
#pragma unroll
    for (int c = 0; c < RS4; ++c) {
      const float4 a = rp[c];
      f[4 * c] = a.x, f[4 * c + 1] = a.y, f[4 * c + 2] = a.z, f[4 * c + 3] = a.w;
    }
  };
  // Rectilinear datasets (bin entries in cache, a small query state) keep two point records of
  // 32 bytes in flight; with a lookup grid the larger query state (the cell record) leaves room
  // for one, and so do larger records and the direct pass (rare rays, one exp per pixel), whose
  // code is kept compact.
  constexpr bool kDeep = RECT && !DIRECT && RS4 == 2;
  // the integration of group kb from its point records fC (winner rC)
  const auto integrate = [&](const int kb, const float (&fC)[4 * RS4], const int rC) {
    // group kb: this lane's sample, the values write_sample would store
    const float vls = rC >= 0 ? (float)(((0.0 + (sb * cl) * (double)fC[0]) + (-sb * sl_) * (double)fC[1]) +
                                        cb * (double)fC[2])
                              : 0.f;
    const double st = rC >= 0 ? rec_st(K0, fC[3]) : 0.0;
    // whole waves skip groups without a data point
    if (__ballot(rC >= 0) == 0) {
    } else if constexpr (SL == 0) {
      // the G samples of the group in LOS order, each with k_los_spectral's step for line q
      static_for<G>([&](auto oc) {
        constexpr int o = decltype(oc)::value;
        const float v = group_bcast<G, o>(vls);
        const double sto = group_bcast<G, o>(st);
        float pk = 0.f;
#pragma unroll
        for (int l = 0; l < (G < NL ? G : NL); ++l) {
          const float t = group_bcast<G, o>(fC[4 + l]);
          if (l == q) pk = t;
        }
        const bool on = integ && pk != 0.f && kb + o < nz;
        const double peak = pk, w = (double)rec_fw(Kq, sto);
        const double mu = (double)v * l0c;
        const double kap = fomo::spectral_kappa(w);
        const double d0 = lamc - mu;
        if constexpr (DIRECT) {
          if (on) {
#pragma unroll
            for (int c = 0; c < CH; ++c) {
              const double d = d0 + (double)c * dl;
              acc[c] = spectral_add(acc[c], (float)peak, exp(-kap * d * d));
            }
          }
          return;
        }
        if (!on) return;
        double xa[3];
        if (!profile_args(kap, d0, dl, xa)) {
          bad = 1u;
          return;
        }
        double ex[3], gc[4], RR[4], q16;
        fomo::exp3(xa, ex);
        profile_chains(ex, gc, RR, q16);
#pragma unroll
        for (int c0 = 0; c0 < CH; c0 += 4) {
#pragma unroll
          for (int jj = 0; jj < 4; ++jj) {
            acc[c0 + jj] = spectral_add(acc[c0 + jj], (float)peak, gc[jj]);
            gc[jj] *= RR[jj];
            RR[jj] *= q16;
          }
        }
      });
    } else {
      // the shared profile at this lane's own sample, for each lambda pass p: its anchors (ex, ok
      // bit p), or for the direct pass its (kap, d0)
      const double w = (double)rec_fw(Kq, st);
      const double mu = (double)vls * l0c;
      const double kapS = fomo::spectral_kappa(w);
      double d0S[NP], exS[NP][3];
      int okS = 0;
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        d0S[p] = lamc_of(p) - mu;
        exS[p][0] = exS[p][1] = exS[p][2] = 1.0;
        if constexpr (!DIRECT) {
          double xa[3];
          const bool ok = profile_args(kapS, d0S[p], dl, xa);
          okS |= ok ? 1 << p : 0;
          if (!ok) xa[0] = xa[1] = xa[2] = 0.0;  // finite chains; nothing is added from them
          fomo::exp3(xa, exS[p]);
        }
      }
      if constexpr (kXch) {
        unsigned char* my = xch + threadIdx.x * kSlot;
        wave_sync();  // the reads of the previous group are done
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          double gc[4], RR[4], q16;
          profile_chains(exS[p], gc, RR, q16);
          double2* gr = reinterpret_cast<double2*>(my + p * 80);
#pragma unroll
          for (int c = 0; c < 4; ++c) gr[c] = make_double2(gc[c], RR[c]);
          *reinterpret_cast<double*>(my + p * 80 + 64) = q16;
        }
        float pw[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int l = 0; l < SL; ++l) pw[l] = fC[4 + l];
        *reinterpret_cast<float4*>(my + NP * 80) = make_float4(pw[0], pw[1], pw[2], pw[3]);
        *reinterpret_cast<int*>(my + NP * 80 + 16) = okS;
        wave_sync();
      }
      // the 4 samples of the group in LOS order; lane q runs chain q of each
      static_for<4>([&](auto oc) {
        constexpr int o = decltype(oc)::value;
        // lane o of the ray's quad: the owner of sample kb + o
        const unsigned char* src = xch + (kXch ? (((unsigned)threadIdx.x & ~3u) | (unsigned)o) * kSlot : 0);
        float pk[SL];
        unsigned on = 0;
        if constexpr (kXch) {
          const float4 p4 = *reinterpret_cast<const float4*>(src + NP * 80);
          const float pw[4] = {p4.x, p4.y, p4.z, p4.w};
#pragma unroll
          for (int l = 0; l < SL; ++l) pk[l] = pw[l];
        } else {
#pragma unroll
          for (int l = 0; l < SL; ++l) pk[l] = group_bcast<4, o>(fC[4 + l]);
        }
#pragma unroll
        for (int l = 0; l < SL; ++l)
          if (pk[l] != 0.f && l < ls.n) on |= 1u << l;
        if (!integ || kb + o >= nz) on = 0;
        if constexpr (DIRECT) {
          const double kap = group_bcast<4, o>(kapS);
#pragma unroll
          for (int p = 0; p < NP; ++p) {
            const double d0 = group_bcast<4, o>(d0S[p]);
            unsigned use = 0;
#pragma unroll
            for (int l = 0; l < SL; ++l)
              if ((on >> l) & (need >> (l * NP + p)) & 1u) use |= 1u << l;
            if (use == 0) continue;
#pragma unroll
            for (int m = 0; m < CJ; ++m) {
              const double d = d0 + (double)(4 * m + q) * dl;
              const double e = exp(-kap * d * d);
#pragma unroll
              for (int l = 0; l < SL; ++l)
                if (use & (1u << l)) acc[(p * SL + l) * CJ + m] = spectral_add(acc[(p * SL + l) * CJ + m], pk[l], e);
            }
          }
          return;
        }
        const int okb = *reinterpret_cast<const int*>(src + NP * 80 + 16);
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          if (!((okb >> p) & 1)) {
#pragma unroll
            for (int l = 0; l < SL; ++l)
              if (on & (1u << l)) bad |= 1u << (l * NP + p);
            continue;
          }
          if (on == 0) continue;
          const double2 gr = *reinterpret_cast<const double2*>(src + p * 80 + 16 * q);
          const double q16 = *reinterpret_cast<const double*>(src + p * 80 + 64);
          double gq = gr.x, Rq = gr.y;
          // a line with peak 0 adds fma(0, g, sum) = sum: g is finite here
#pragma unroll
          for (int m = 0; m < CJ; ++m) {
#pragma unroll
            for (int l = 0; l < SL; ++l) acc[(p * SL + l) * CJ + m] = spectral_add(acc[(p * SL + l) * CJ + m], pk[l], gq);
            gq *= Rq;
            Rq *= q16;
          }
        }
      });
    }
  };
  SearchBox boxN;
  const float4* rpN = lslot;
  float4 slN[kSlots];
  bool actN, cellN;
  if constexpr (kDeep) {
    // Three record sets in fixed roles, unrolled by three so that no set is copied while its load
    // is in flight: step g integrates group g from `cur` and loads group g + 2G's record into
    // `ahead` (the set group g - G used), then issues group g + 3G's bin entries.
    float fa[4 * RS4], fb[4 * RS4], fc[4 * RS4];
    int ra, rb, rc = -1;
    query_issue(q, boxN, rpN, slN, actN, cellN);
    ra = query_resolve(boxN, rpN, slN, actN, cellN);
    record_issue(ra, fa);
    query_issue(G + q, boxN, rpN, slN, actN, cellN);
    rb = query_resolve(boxN, rpN, slN, actN, cellN);
    record_issue(rb, fb);
    query_issue(2 * G + q, boxN, rpN, slN, actN, cellN);
    const auto step = [&](const int kb, const float (&cur)[4 * RS4], const int rcur, float (&ahead)[4 * RS4],
                          int& rahead) {
      rahead = query_resolve(boxN, rpN, slN, actN, cellN);
      record_issue(rahead, ahead);
      integrate(kb, cur, rcur);
      query_issue(kb + 3 * G + q, boxN, rpN, slN, actN, cellN);
    };
    for (int kb = 0; kb < nz; kb += 3 * G) {
      step(kb, fa, ra, fc, rc);
      if (kb + G >= nz) break;
      step(kb + G, fb, rb, fa, ra);
      if (kb + 2 * G >= nz) break;
      step(kb + 2 * G, fc, rc, fb, rb);
    }
  } else {
    // group kb + 2G's lookup cell record and group kb + G's point record in flight while group kb
    // is integrated
    float fC[4 * RS4], fB[4 * RS4];
    query_issue(q, boxN, rpN, slN, actN, cellN);
    int rC = query_resolve(boxN, rpN, slN, actN, cellN);
    record_issue(rC, fC);
    query_issue(G + q, boxN, rpN, slN, actN, cellN);
    for (int kb = 0; kb < nz; kb += G) {
      SearchBox boxA;
      const float4* rpA = lslot;
      float4 slA[kSlots];
      bool actA, cellA;
      query_issue(kb + 2 * G + q, boxA, rpA, slA, actA, cellA);
      const int rB = query_resolve(boxN, rpN, slN, actN, cellN);
      record_issue(rB, fB);
      integrate(kb, fC, rC);
      rC = rB;
#pragma unroll
      for (int c = 0; c < 4 * RS4; ++c) fC[c] = fB[c];
      boxN = boxA;
      rpN = rpA;
#pragma unroll
      for (int c = 0; c < kSlots; ++c) slN[c] = slA[c];
      actN = actA;
      cellN = cellA;
    }
  }
  if constexpr (!DIRECT) {
    // a wave with a flagged ray goes on the direct pass's list
    if (__ballot(bad != 0) != 0 && lane == 0) {
      const int64_t wpv = (R + RPW - 1) / RPW;
      dlist[2 + atomicAdd(&dlist[0], 1)] = (int32_t)(view * wpv + wave);
    }
  }
  if (!integ) return;
  const double sc = vg.scale;
  if constexpr (SL > 0) {
    if (!DIRECT && q == 0) {
#pragma unroll
      for (int l = 0; l < SL; ++l)
        if (l < ls.n) flag[(size_t)l * R] = (int)((bad >> (l * NP)) & ((1u << NP) - 1));
    }
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int l = 0; l < SL; ++l) {
        const unsigned bit = 1u << (l * NP + p);
        if (l >= ls.n || (DIRECT ? !(need & bit) : (bad & bit) != 0)) continue;
        float* o = ls.out[l] + ooff + (size_t)ray * nl + p * CH;
#pragma unroll
        for (int m = 0; m < CJ; ++m)
          if (p * CH + 4 * m + q < nl) o[4 * m + q] = (float)(sc * (double)acc[(p * SL + l) * CJ + m]);
      }
  } else {
    if (!DIRECT) {
      *flag = bad ? 1 : 0;
      if (bad) return;
    }
    float* o = out + (size_t)ray * nl;
#pragma unroll
    for (int c = 0; c < CH; ++c)
      if (c < nl) o[c] = (float)(sc * (double)acc[c]);
  }
}


// The fused pass. Main pass (DIRECT = false): grid (waves / 4, views), one wave per 64 / G rays.
// Direct pass: any grid; its waves take the listed waves of the main pass in turn, and the last
// block to finish clears the list (dlist[0], dlist[1]) for the next main pass. The list starts
// cleared: the view finalize kernels clear it before the first main pass of a buffer.
// The main pass of the four-line shared profile on a rectilinear dataset asks for 3 waves per SIMD
// (at most 168 VGPRs; a few values spill): its FP64 chains and float sums otherwise take about
// 185 registers and 2 waves (C2 11.64 -> 11.22 ms per step; the scattered index gained nothing).
template <int RS4, int CH, int G, bool DIRECT, int SL, int NP = 1, bool RECT = false>
__global__ void __launch_bounds__(256, (RECT && !DIRECT && SL == 4) ? 3 : 1) k_nn_los(const float4* __restrict__ pts, GridInfo g,
                                                const float4* __restrict__ lslot, const float4* __restrict__ lcand,
                                                LookupInfo lk, const float4* __restrict__ mpts,
                                                const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                BvhInfo bv, const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                                const float4* __restrict__ rec, RecInfo ri, const LineSet ls, int nl,
                                                int32_t* __restrict__ flags, int64_t out_vs, int64_t flag_vs,
                                                int32_t* __restrict__ dlist) {
  // the wave's index is uniform: kept in scalar registers
  const int w4 = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if constexpr (!DIRECT) {
    nn_los_wave<RS4, CH, G, false, SL, NP, RECT>((int)blockIdx.y, (int64_t)blockIdx.x * 4 + w4, pts,
                                                 g, lslot, lcand, lk, mpts, nlo, nhi, bv, vgp, nx, ny, nz, rec, ri,
                                                 ls, nl, flags, out_vs, flag_vs, dlist);
  } else {
    constexpr int RPW = 64 / G;
    const int64_t wpv = ((int64_t)nx * ny + RPW - 1) / RPW;
    const int n = dlist[0];
    for (int it = (int)blockIdx.x * 4 + w4; it < n; it += (int)gridDim.x * 4) {
      const int item = __builtin_amdgcn_readfirstlane(dlist[2 + it]);
      nn_los_wave<RS4, CH, G, true, SL, NP, RECT>((int)(item / wpv), item % wpv, pts, g, lslot, lcand, lk, mpts, nlo,
                                                  nhi, bv, vgp, nx, ny, nz, rec, ri, ls, nl, flags, out_vs, flag_vs,
                                                  dlist);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      __threadfence();
      if (atomicAdd(&dlist[1], 1) == (int)gridDim.x - 1) {
        atomicExch(&dlist[0], 0);
        atomicExch(&dlist[1], 0);
      }
    }
  }
}

// Ints of a direct-pass list for nviews views of R rays (at least 8 rays per wave).
__host__ __device__ constexpr int64_t nn_los_list_ints(int64_t R, int64_t nviews) {
  return 2 + nviews * ((R + 7) / 8);
}

// ------------------------------------------------------------------------------------
// Host launch of k_nn_los. The host code builds the arguments and picks the variant; the
// instantiations live in fomo_nnlos_part.hip.
// ------------------------------------------------------------------------------------
struct NnLosArgs {
  const float4* pts;
  GridInfo g;
  const float4* lslot;
  const float4* lcand;
  LookupInfo lk;
  const float4* mpts;
  const float4* nlo;
  const float4* nhi;
  BvhInfo bvh;
  const ViewGeom* geom;
  int nx, ny, nz;
  const float4* rec;
  RecInfo ri;
  LineSet ls;
  int nl;
  int32_t* flags;
  int nviews;      // views of one launch (blockIdx.y): geom[v], outputs + v * out_vs, flags + v * flag_vs
  int64_t out_vs;  // floats between two views' outputs of one line
  int64_t flag_vs; // ints between two views' ray flags
  int32_t* dlist;  // direct-pass list, nn_los_list_ints(R, nviews) ints, cleared before the main pass
};

// k_nn_los<RS4, CH, G, DIRECT, SL, NP> variants the host launches, as (RS4, G, SL, NP).
enum NnLosVariant {
  kNnLosSingle = 0,   // (2, 4, 1, 1): one line, the 4 lanes of a ray on its profile's chains
  kNnLosShared = 1,   // (2, 4, 4, 1): up to 4 lines sharing one profile
  kNnLosLines2 = 2,   // (2, 2, 0, 1): 2 lines, lane q integrates line q
  kNnLosLines4 = 3,   // (2, 4, 0, 1): 3-4 lines
  kNnLosLines8 = 4,   // (4, 8, 0, 1): 5-8 lines
  kNnLosSingle2 = 5,  // (2, 4, 1, 2): one line in two lambda passes (CH >= 36)
};

// Launches the variant with lambda chunk CH, first with DIRECT = false, then the DIRECT pass for
// the flagged rays; RECT: the dataset is rectilinear (lk.rect), else it has a candidate lookup
// grid or none. Explicitly instantiated in fomo_nnlos_part.hip for CH = 4, 8, ..., 64.
template <int CH, bool RECT>
void nn_los_launch(int variant, unsigned blocks, hipStream_t s, const NnLosArgs& a);

}  // namespace fomo_dev
