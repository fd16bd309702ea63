// Octree over Morton-ordered points for the exact nearest-neighbour query inside a search box.
// Shared by the device kernels (fomo_device.hip) and a host unit test: the query is a template
// over the array accessor type, so the same code runs on raw device pointers and on bounds-checked
// host arrays.
//
// Layout: nodes in depth-first preorder; every node covers a contiguous range of the
// Morton-sorted points. Node i is two 4-float records:
//   lo[i] = (box min, w), hi[i] = (box max, first)
//   w >= 0: internal node, first child at i+1, w = index just past its subtree (skip pointer);
//   w <  0: leaf of -w points mpts[first .. first-w).
// mpts[m] = (x, y, z, sorted position as int bits); pts[p].w = original index as int bits.
//
// The query returns the nearest point strictly inside the open search box, comparing squared
// distances of float coordinates in double ((d0^2 + d1^2) + d2^2, no FMA), ties to the lowest
// original index: the reference's R-tree query (fomo-nearestneighbour.cpp:179-183,
// fomo-thomson.cpp:310) with its tie order fixed.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#if defined(__HIPCC__)
#define FOMO_HD __host__ __device__
#else
#define FOMO_HD
#endif

namespace fomo {

FOMO_HD inline int f2i(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __float_as_int(f);
#else
  int i;
  std::memcpy(&i, &f, sizeof i);
  return i;
#endif
}

inline float i2f(int32_t v) {
  float f;
  std::memcpy(&f, &v, sizeof f);
  return f;
}

struct SearchBox {
  float t0, t1, t2;                    // target
  float lo0, lo1, lo2, hi0, hi1, hi2;  // open search box
  template <class V> FOMO_HD bool node_overlaps(const V& l, const V& h) const {
    return h.x > lo0 && l.x < hi0 && h.y > lo1 && l.y < hi1 && h.z > lo2 && l.z < hi2;
  }
  template <class V> FOMO_HD bool holds(const V& q) const {
    return lo0 < q.x && q.x < hi0 && lo1 < q.y && q.y < hi1 && lo2 < q.z && q.z < hi2;
  }
};

// Squared distance from the target to the box [l, h] (l <= h on every axis). Per axis
// max(l - t, t - h, 0): at most one of the two differences is positive, and it equals the
// distance to the nearer face; written without branches, so a wave's lanes stay together.
template <class V>
FOMO_HD inline double node_mindist2(const V& l, const V& h, float t0, float t1, float t2) {
  double s = 0.0, d;
  d = std::fmax(std::fmax((double)l.x - (double)t0, (double)t0 - (double)h.x), 0.0);
  s = s + d * d;
  d = std::fmax(std::fmax((double)l.y - (double)t1, (double)t1 - (double)h.y), 0.0);
  s = s + d * d;
  d = std::fmax(std::fmax((double)l.z - (double)t2, (double)t2 - (double)h.z), 0.0);
  s = s + d * d;
  return s;
}

// Points of one leaf against the running best (sorted position bpos, original index borig).
template <class A>
FOMO_HD inline void leaf_scan(const A& pts, const A& mpts, int first, int count, const SearchBox& sb, double& best,
                              int& bpos, int& borig) {
  for (int i = first; i < first + count; ++i) {
    const auto q = mpts[i];
    if (!sb.holds(q)) continue;
    const double d0 = (double)q.x - (double)sb.t0;
    const double d1 = (double)q.y - (double)sb.t1;
    const double d2 = (double)q.z - (double)sb.t2;
    const double d = (d0 * d0 + d1 * d1) + d2 * d2;
    if (d <= best) {
      const int p = f2i(q.w);
      const int oi = f2i(pts[p].w);
      if (d < best || oi < borig) {
        best = d;
        bpos = p;
        borig = oi;
      }
    }
  }
}

// First bound of a query without one (bpos < 0): a greedy descent to the leaf whose box is
// nearest at every level (among the children that overlap the search box), and that leaf's scan.
template <class A>
FOMO_HD inline void oct_descend(const A& pts, const A& mpts, const A& nlo, const A& nhi, const SearchBox& sb,
                                double& best, int& bpos, int& borig) {
  int e = 0;
  int w = f2i(nlo[0].w);
  while (w >= 0) {
    int bc = -1;
    double bm = INFINITY;
    for (int c = e + 1; c < w;) {
      const auto cl = nlo[c];
      const auto ch = nhi[c];
      if (sb.node_overlaps(cl, ch)) {
        const double md = node_mindist2(cl, ch, sb.t0, sb.t1, sb.t2);
        if (md < bm) {
          bm = md;
          bc = c;
        }
      }
      const int cw = f2i(cl.w);
      c = cw < 0 ? c + 1 : cw;
    }
    if (bc < 0) break;
    e = bc;
    w = f2i(nlo[e].w);
  }
  if (w < 0) leaf_scan(pts, mpts, f2i(nhi[e].w), -w, sb, best, bpos, borig);
}

// Whether node (l, h) can hold a point strictly inside the search box at distance <= best.
template <class V>
FOMO_HD inline bool oct_enter(const V& l, const V& h, const SearchBox& sb, double best) {
  return sb.node_overlaps(l, h) && node_mindist2(l, h, sb.t0, sb.t1, sb.t2) <= best;
}

// Exact NN strictly inside the search box. (best, bpos, borig) may carry a known point strictly
// inside the box as the starting bound; without one (bpos < 0) oct_descend supplies the first
// bound. Then a stackless preorder walk: a node that oct_enter rejects is skipped with its
// subtree. Returns the sorted position of the answer or -1.
template <class A>
FOMO_HD inline int oct_query(const A& pts, const A& mpts, const A& nlo, const A& nhi, int64_t nodes,
                             const SearchBox& sb, double best, int bpos, int borig) {
  if (nodes <= 0) return bpos;
  if (bpos < 0) oct_descend(pts, mpts, nlo, nhi, sb, best, bpos, borig);
  const int n = (int)nodes;
  for (int e = 0; e < n;) {
    const auto l = nlo[e];
    const auto h = nhi[e];
    const int w = f2i(l.w);
    const bool in = oct_enter(l, h, sb, best);
    if (w < 0) {
      if (in) leaf_scan(pts, mpts, f2i(h.w), -w, sb, best, bpos, borig);
      ++e;
    } else {
      e = in ? e + 1 : w;
    }
  }
  return bpos;
}

// Host side: nodes in depth-first preorder from the sorted 30-bit Morton keys (a node is a
// contiguous key range, so only the keys are needed). The children of a range are the runs of its
// highest differing 3-bit group (2..8 runs); a range of identical keys splits into up to 8 equal
// parts; a range of at most 8 points is a leaf.
struct OctBuilder {
  const uint32_t* key = nullptr;
  std::vector<int32_t> w, first;              // per node: skip (internal) or -count (leaf); range start
  std::vector<std::vector<int32_t>> bydepth;  // node ids per depth (boxes are filled deepest first)

  void build(const uint32_t* sorted_keys, int64_t n) {
    key = sorted_keys;
    w.clear();
    first.clear();
    bydepth.clear();
    if (n > 0) node(0, n, 27, 0);
  }

 private:
  void node(int64_t s, int64_t e, int shift, int depth) {
    const int32_t me = (int32_t)w.size();
    w.push_back(0);
    first.push_back((int32_t)s);
    if ((int)bydepth.size() <= depth) bydepth.resize(depth + 1);
    bydepth[depth].push_back(me);
    if (e - s <= 8) {
      w[me] = -(int32_t)(e - s);
      return;
    }
    while (shift >= 0 && (key[s] >> shift) == (key[e - 1] >> shift)) shift -= 3;
    if (shift < 0) {
      const int64_t part = (e - s + 7) / 8;
      for (int64_t a = s; a < e; a += part) node(a, std::min(e, a + part), -3, depth + 1);
    } else {
      for (int64_t a = s; a < e;) {
        const uint32_t pre = key[a] >> shift;
        const int64_t b =
            std::upper_bound(key + a, key + e, pre, [shift](uint32_t v, uint32_t k) { return v < (k >> shift); }) - key;
        node(a, b, shift - 3, depth + 1);
        a = b;
      }
    }
    w[me] = (int32_t)w.size();
  }
};

// Wide records: per internal node, in preorder rank order (the root is rank 0), 8 slots of two
// 4-float records, slot s of rank r at [16 r + 2 s] = (box min, ref) and [16 r + 2 s + 1] =
// (box max, count), ref and count as int bits: count 0 for an internal child (ref: its rank),
// 1..8 for a leaf (ref: its first point in mpts), and an empty slot with an empty box (min +inf,
// max -inf, never overlapping a search box). The device walks them depth first, a node's 8
// children in one load.
//
// The children of every internal node: kid[8 r + s] = node index of child s of rank r (-1: empty
// slot), kref[8 r + s] = the child's rank when internal (-1: leaf). Returns the number of internal
// nodes; 0 when the tree has none or a node has more than 8 children (the builder gives at most 8).
inline int64_t wide_children(const OctBuilder& ob, std::vector<int32_t>& kid, std::vector<int32_t>& kref) {
  const int64_t nodes = (int64_t)ob.w.size();
  std::vector<int32_t> rank(nodes, -1);
  int64_t nint = 0;
  for (int64_t i = 0; i < nodes; ++i)
    if (ob.w[i] >= 0) rank[i] = (int32_t)nint++;
  kid.assign((size_t)nint * 8, -1);
  kref.assign((size_t)nint * 8, -1);
  for (int64_t i = 0; i < nodes; ++i) {
    if (ob.w[i] < 0) continue;
    int slot = 0;
    for (int64_t c = i + 1; c < ob.w[i]; c = ob.w[c] < 0 ? c + 1 : ob.w[c]) {
      if (slot == 8) {
        kid.clear();
        kref.clear();
        return 0;
      }
      kid[(size_t)rank[i] * 8 + slot] = (int32_t)c;
      kref[(size_t)rank[i] * 8 + slot] = rank[c];
      ++slot;
    }
  }
  return nint;
}

FOMO_HD inline float i2f_hd(int32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __int_as_float(v);
#else
  return i2f(v);
#endif
}

// The slot (a, b) of child c (node index, -1: empty slot) with internal rank ref (-1: leaf) from
// the preorder node records l = lo[c], h = hi[c].
template <class V>
FOMO_HD inline void wide_slot(int c, int ref, const V& l, const V& h, V& a, V& b) {
  if (c < 0) {
    a.x = a.y = a.z = INFINITY;
    b.x = b.y = b.z = -INFINITY;
    a.w = b.w = i2f_hd(0);
    return;
  }
  const int w = f2i(l.w);
  a.x = l.x, a.y = l.y, a.z = l.z;
  b.x = h.x, b.y = h.y, b.z = h.z;
  a.w = i2f_hd(w < 0 ? f2i(h.w) : ref);
  b.w = i2f_hd(w < 0 ? -w : 0);
}

// One query over the wide records, depth first (the order the device's wave walk takes is
// another; the answer is the same): the host model of the device walk, checked against brute
// force by the octree test.
template <class A>
FOMO_HD inline int wide_query(const A& pts, const A& mpts, const A& wide, const SearchBox& sb, double best, int bpos,
                              int borig) {
  constexpr int kStack = 128;
  int stk[kStack];
  int sp = 0;
  stk[sp++] = 0;
  while (sp > 0) {
    const int r = stk[--sp];
    for (int s = 0; s < 8; ++s) {
      const auto a = wide[(int64_t)r * 16 + 2 * s];
      const auto b = wide[(int64_t)r * 16 + 2 * s + 1];
      if (!oct_enter(a, b, sb, best)) continue;
      const int ref = f2i(a.w), cnt = f2i(b.w);
      if (cnt > 0) {
        leaf_scan(pts, mpts, ref, cnt, sb, best, bpos, borig);
      } else {
        if (sp == kStack) return -3;  // deeper than the device's stack allows
        stk[sp++] = ref;
      }
    }
  }
  return bpos;
}

// 30-bit Morton code of a point in the box [bmin, bmax] (10 bits per axis, x highest).
FOMO_HD inline uint32_t morton30(float x, float y, float z, const float bmin[3], const float bmax[3]) {
  const float c[3] = {x, y, z};
  uint32_t u[3];
  for (int a = 0; a < 3; ++a) {
    const double ext = (double)bmax[a] - (double)bmin[a];
    double t = ext > 0 ? ((double)c[a] - (double)bmin[a]) / ext : 0.0;
    t = fmin(fmax(t, 0.0), 1.0);
    u[a] = (uint32_t)fmin(t * 1024.0, 1023.0);
  }
  uint32_t code = 0;
  for (int bit = 9; bit >= 0; --bit)
    code = (code << 3) | (((u[0] >> bit) & 1u) << 2) | (((u[1] >> bit) & 1u) << 1) | ((u[2] >> bit) & 1u);
  return code;
}

}  // namespace fomo
