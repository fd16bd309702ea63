// 3D Delaunay triangulation (host), replacing the CGAL triangulation of the reference's CGAL
// render path (fomo-CGAL.cpp:33-73, Delaunay_triangulation_3 of the data points).
//
// Incremental Bowyer-Watson insertion with an infinite vertex (index -1): every hull face has an
// infinite tet on its outer side, so points outside the current hull are inserted like the
// others. Points are inserted in Morton order; each is located by a visibility walk from the last
// created tet, its conflict region (tets whose circumsphere holds it, infinite tets whose hull
// face it sees) is removed and the region's boundary is connected to the point.
// Predicates are exact (predicates.h); cospherical configurations are resolved by a symbolic
// perturbation that depends only on the coordinates, so the result is unique and independent of
// the insertion order. Duplicate points are inserted once: the vertex keeps the index of the LAST
// input point with those coordinates (the reference reads the values through std::map keyed by
// the point, where later insertions overwrite earlier ones, fomo-CGAL.cpp:128-131).
//
// Tet layout: v[4] vertex ids (point indices, -1 = infinite), n[4] neighbour across the face
// opposite v[i]. Finite tets are positively oriented: orient(v0,v1,v2,v3) > 0. The face opposite
// v[i] is kFace[i], ordered so that orient(face, v[i]) > 0.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "predicates.h"

namespace fomo {

constexpr int kFace[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};

struct Tet {
  int32_t v[4];
  int32_t n[4];
};

class Delaunay {
 public:
  std::vector<pred::P3> P;   // input points
  std::vector<Tet> T;        // tets (dead ones have v[0] == kDead)
  std::vector<int32_t> rep;  // input point -> vertex point carrying its location (duplicates)
  int64_t ninserted = 0, nduplicates = 0;
  std::string error;

  static constexpr int32_t kInf = -1;
  static constexpr int32_t kDead = -2;

  bool build(const float* x, const float* y, const float* z, int64_t n) {
    P.resize(n);
    for (int64_t i = 0; i < n; ++i) P[i] = {(double)x[i], (double)y[i], (double)z[i]};
    rep.assign(n, -1);
    T.clear();
    free_.clear();
    if (n < 4) return fail("fewer than 4 points");
    // Morton order of the points (insertion order)
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (auto& p : P) {
      lo[0] = std::min(lo[0], p.x), lo[1] = std::min(lo[1], p.y), lo[2] = std::min(lo[2], p.z);
      hi[0] = std::max(hi[0], p.x), hi[1] = std::max(hi[1], p.y), hi[2] = std::max(hi[2], p.z);
    }
    std::vector<std::pair<uint64_t, int32_t>> ord(n);
    for (int64_t i = 0; i < n; ++i) {
      const double c[3] = {P[i].x, P[i].y, P[i].z};
      uint64_t code = 0;
      uint32_t u[3];
      for (int a = 0; a < 3; ++a) {
        const double e = hi[a] - lo[a];
        const double t = e > 0 ? (c[a] - lo[a]) / e : 0.0;
        u[a] = (uint32_t)std::min(2097151.0, std::max(0.0, t * 2097152.0));
      }
      for (int bit = 20; bit >= 0; --bit)
        code = (code << 3) | (((u[0] >> bit) & 1u) << 2) | (((u[1] >> bit) & 1u) << 1) | ((u[2] >> bit) & 1u);
      ord[i] = {code, (int32_t)i};
    }
    std::sort(ord.begin(), ord.end());
    // duplicates: equal coordinates; the vertex is the last input index of the group
    std::vector<int32_t> seq;
    seq.reserve(n);
    {
      std::vector<int32_t> byc(n);
      for (int64_t i = 0; i < n; ++i) byc[i] = (int32_t)i;
      std::sort(byc.begin(), byc.end(), [&](int32_t a, int32_t b) {
        if (P[a].x != P[b].x) return P[a].x < P[b].x;
        if (P[a].y != P[b].y) return P[a].y < P[b].y;
        if (P[a].z != P[b].z) return P[a].z < P[b].z;
        return a < b;
      });
      for (int64_t s = 0; s < n;) {
        int64_t e = s + 1;
        while (e < n && P[byc[e]].x == P[byc[s]].x && P[byc[e]].y == P[byc[s]].y && P[byc[e]].z == P[byc[s]].z) ++e;
        const int32_t last = byc[e - 1];
        for (int64_t k = s; k < e; ++k) rep[byc[k]] = last;
        nduplicates += e - s - 1;
        s = e;
      }
      for (auto& o : ord)
        if (rep[o.second] == o.second) seq.push_back(o.second);
    }
    // work on the points renumbered in insertion order (memory locality), map back at the end
    std::vector<pred::P3> input;
    input.swap(P);
    P.resize(seq.size());
    for (size_t i = 0; i < seq.size(); ++i) P[i] = input[seq[i]];
    std::vector<int32_t> order(seq.size());
    for (size_t i = 0; i < seq.size(); ++i) order[i] = (int32_t)i;
    bool ok = init(order);
    // optional progress lines on stderr (FOMO_PROGRESS=1), as the reference prints its own
    // "Doing Delaunay triangulation" progress (fomo-CGAL.cpp:55)
    const char* pe = std::getenv("FOMO_PROGRESS");
    const bool progress = pe && pe[0] == '1';
    for (size_t i = 0; ok && i < order.size(); ++i) {
      if (!inserted_[order[i]]) ok = insert(order[i]);
      if (progress && (i + 1) % 1000000 == 0)
        std::fprintf(stderr, "Delaunay: %zu of %zu points inserted, %zu tets\n", i + 1, order.size(), T.size());
    }
    for (auto& t : T)
      if (t.v[0] != kDead)
        for (int k = 0; k < 4; ++k)
          if (t.v[k] >= 0) t.v[k] = seq[t.v[k]];
    P.swap(input);
    return ok;
  }

  // number of live tets (finite + infinite) and compaction into a dense array
  std::vector<Tet> compact(int64_t* nfinite) const {
    std::vector<int32_t> remap(T.size(), -1);
    int64_t m = 0;
    for (size_t i = 0; i < T.size(); ++i)
      if (T[i].v[0] != kDead) remap[i] = (int32_t)m++;
    std::vector<Tet> out;
    out.reserve(m);
    int64_t nf = 0;
    for (size_t i = 0; i < T.size(); ++i) {
      if (T[i].v[0] == kDead) continue;
      Tet t = T[i];
      for (int k = 0; k < 4; ++k) t.n[k] = remap[t.n[k]];
      if (t.v[0] >= 0 && t.v[1] >= 0 && t.v[2] >= 0 && t.v[3] >= 0) ++nf;
      out.push_back(t);
    }
    if (nfinite) *nfinite = nf;
    return out;
  }

  bool finite(const Tet& t) const { return t.v[0] >= 0 && t.v[1] >= 0 && t.v[2] >= 0 && t.v[3] >= 0; }

 private:
  std::vector<int32_t> free_;
  std::vector<char> inserted_;
  int32_t last_ = 0;
  uint64_t rng_ = 0x9e3779b97f4a7c15ull;
  std::vector<int32_t> stack_, cavity_, tested_, created_;
  std::vector<char> mark_;  // per tet: 1 in conflict, 2 tested not in conflict
  // edge -> (new tet, face) table for connecting the tets of one insertion (stamped entries)
  std::vector<uint64_t> hkey_;
  std::vector<int64_t> hval_;
  std::vector<uint32_t> hstamp_;
  uint32_t stamp_ = 0;

  bool fail(const std::string& m) {
    error = m;
    return false;
  }

  int32_t new_tet(int32_t a, int32_t b, int32_t c, int32_t d) {
    int32_t id;
    if (!free_.empty()) {
      id = free_.back();
      free_.pop_back();
    } else {
      id = (int32_t)T.size();
      T.push_back(Tet{});
      mark_.push_back(0);
    }
    T[id].v[0] = a, T[id].v[1] = b, T[id].v[2] = c, T[id].v[3] = d;
    T[id].n[0] = T[id].n[1] = T[id].n[2] = T[id].n[3] = -1;
    mark_[id] = 0;
    return id;
  }

  uint32_t rnd() {
    rng_ ^= rng_ << 13;
    rng_ ^= rng_ >> 7;
    rng_ ^= rng_ << 17;
    return (uint32_t)rng_;
  }

  const pred::P3& pt(int32_t v) const { return P[v]; }

  // first tet from 4 affinely independent points, plus the 4 infinite tets on its faces
  bool init(const std::vector<int32_t>& seq) {
    inserted_.assign(P.size(), 0);
    if (seq.size() < 4) return fail("fewer than 4 distinct points");
    const int32_t a = seq[0];
    int32_t b = -1, c = -1, d = -1;
    size_t k = 1;
    for (; k < seq.size(); ++k)
      if (seq[k] != a) {
        b = seq[k];
        break;
      }
    // c: not collinear with a,b
    for (++k; k < seq.size() && c < 0; ++k) {
      const pred::P3 &A = pt(a), &B = pt(b), &C = pt(seq[k]);
      // collinear iff all 2x2 minors of (B-A, C-A) vanish: test with orient against 3 axis points
      const pred::P3 e1{A.x + 1, A.y, A.z}, e2{A.x, A.y + 1, A.z}, e3{A.x, A.y, A.z + 1};
      if (pred::orient(A, B, C, e1) || pred::orient(A, B, C, e2) || pred::orient(A, B, C, e3)) c = seq[k];
    }
    for (; k < seq.size() && d < 0; ++k)
      if (pred::orient(pt(a), pt(b), pt(c), pt(seq[k])) != 0) d = seq[k];
    if (c < 0 || d < 0) return fail("all points are coplanar: no 3D triangulation");
    if (pred::orient(pt(a), pt(b), pt(c), pt(d)) < 0) std::swap(c, d);
    const int32_t t0 = new_tet(a, b, c, d);
    int32_t inf[4];
    for (int i = 0; i < 4; ++i) {
      // outer tet on face i: the face vertices plus the infinite vertex in place of v[i]
      int32_t vv[4] = {T[t0].v[0], T[t0].v[1], T[t0].v[2], T[t0].v[3]};
      // the infinite tet shares face i but lies on its other side: swap two face vertices
      const int f0 = kFace[i][0], f1 = kFace[i][1];
      std::swap(vv[f0], vv[f1]);
      vv[i] = kInf;
      inf[i] = new_tet(vv[0], vv[1], vv[2], vv[3]);
    }
    for (int i = 0; i < 4; ++i) T[t0].n[i] = inf[i];
    link_all({t0, inf[0], inf[1], inf[2], inf[3]});
    for (int32_t p : {a, b, c, d}) inserted_[p] = 1;
    last_ = t0;
    ninserted = 4;
    return true;
  }

  // connect the faces of a small set of tets that share 3 vertices
  void link_all(const std::vector<int32_t>& ids) {
    for (size_t x = 0; x < ids.size(); ++x)
      for (int i = 0; i < 4; ++i) {
        const Tet& A = T[ids[x]];
        int32_t fa[3] = {A.v[kFace[i][0]], A.v[kFace[i][1]], A.v[kFace[i][2]]};
        std::sort(fa, fa + 3);
        for (size_t y = 0; y < ids.size(); ++y) {
          if (y == x) continue;
          const Tet& B = T[ids[y]];
          for (int j = 0; j < 4; ++j) {
            int32_t fb[3] = {B.v[kFace[j][0]], B.v[kFace[j][1]], B.v[kFace[j][2]]};
            std::sort(fb, fb + 3);
            if (fa[0] == fb[0] && fa[1] == fb[1] && fa[2] == fb[2]) T[ids[x]].n[i] = ids[y];
          }
        }
      }
  }

  // orientation of point q against face i of tet t (positive: same side as v[i]); for faces
  // containing the infinite vertex the finite face is the only meaningful one, callers check
  int face_orient(const Tet& t, int i, const pred::P3& q) const {
    return pred::orient(pt(t.v[kFace[i][0]]), pt(t.v[kFace[i][1]]), pt(t.v[kFace[i][2]]), q);
  }

  // is tet t in conflict with point p (its circumsphere strictly contains p, perturbed)? For a
  // positively oriented tet the lifted determinant is negative exactly when p is inside.
  bool conflict(int32_t id, int32_t p) {
    const Tet& t = T[id];
    int j = -1;
    for (int k = 0; k < 4; ++k)
      if (t.v[k] == kInf) j = k;
    if (j < 0) return pred::insphere_perturbed(pt(t.v[0]), pt(t.v[1]), pt(t.v[2]), pt(t.v[3]), pt(p)) < 0;
    // infinite tet: hull face kFace[j] seen from outside (orient(face, inf side) > 0)
    const int o = face_orient(t, j, pt(p));
    if (o > 0) return true;
    if (o < 0) return false;
    // p on the hull plane: the circumcircle test, as the circumsphere of the finite tet on the
    // other side of the face (same circle in the plane)
    const Tet& u = T[t.n[j]];
    int q = -1;
    for (int k = 0; k < 4; ++k) {
      const int32_t v = u.v[k];
      if (v != t.v[kFace[j][0]] && v != t.v[kFace[j][1]] && v != t.v[kFace[j][2]]) q = k;
    }
    return pred::insphere_perturbed(pt(u.v[0]), pt(u.v[1]), pt(u.v[2]), pt(u.v[3]), pt(p)) < 0 && q >= 0;
  }

  // tet containing p (finite), or an infinite tet whose hull face p sees
  int32_t locate(int32_t p) {
    int32_t t = last_;
    const pred::P3& q = pt(p);
    for (int64_t steps = 0; steps < 100000000; ++steps) {
      const Tet& c = T[t];
      int j = -1;
      for (int k = 0; k < 4; ++k)
        if (c.v[k] == kInf) j = k;
      if (j >= 0) {
        // infinite tet: if p is beyond its hull face we are done, else step inside
        const int o = face_orient(c, j, q);
        if (o >= 0) return t;
        t = c.n[j];
        continue;
      }
      const int s = (int)(rnd() & 3);
      bool moved = false;
      for (int k = 0; k < 4; ++k) {
        const int i = (s + k) & 3;
        if (face_orient(c, i, q) < 0) {
          t = c.n[i];
          moved = true;
          break;
        }
      }
      if (!moved) return t;
    }
    return -1;
  }

  bool insert(int32_t p) {
    const int32_t t0 = locate(p);
    if (t0 < 0) return fail("point location did not terminate");
    // the located tet may fail the strict conflict test only in degenerate hull cases: then search
    // its neighbours (a point on the hull plane outside the circumcircle belongs to another face)
    int32_t seed = -1;
    if (conflict(t0, p)) seed = t0;
    else {
      for (int k = 0; k < 4 && seed < 0; ++k)
        if (T[t0].n[k] >= 0 && conflict(T[t0].n[k], p)) seed = T[t0].n[k];
      if (seed < 0) {
        // fall back: scan the tets around (rare); flood over infinite tets
        std::vector<int32_t> st{t0};
        std::vector<char> seen(T.size(), 0);
        seen[t0] = 1;
        while (!st.empty() && seed < 0) {
          const int32_t c = st.back();
          st.pop_back();
          if (T[c].v[0] != kDead && conflict(c, p)) seed = c;
          for (int k = 0; k < 4; ++k) {
            const int32_t nb = T[c].n[k];
            if (nb >= 0 && !seen[nb] && (st.size() < 4096)) {
              seen[nb] = 1;
              st.push_back(nb);
            }
          }
        }
      }
      if (seed < 0) return fail("no conflict region for a point (degenerate input)");
    }
    // conflict region by flood fill
    cavity_.clear();
    stack_.clear();
    tested_.clear();
    stack_.push_back(seed);
    mark_[seed] = 1;
    while (!stack_.empty()) {
      const int32_t c = stack_.back();
      stack_.pop_back();
      cavity_.push_back(c);
      for (int k = 0; k < 4; ++k) {
        const int32_t nb = T[c].n[k];
        if (mark_[nb]) continue;
        if (conflict(nb, p)) {
          mark_[nb] = 1;
          stack_.push_back(nb);
        } else {
          mark_[nb] = 2;
          tested_.push_back(nb);
        }
      }
    }
    // boundary faces: (cavity tet c, face k) with neighbour outside; the new tet is c with its
    // vertex k replaced by p (same orientation: p sees the face from c's side)
    created_.clear();
    for (int32_t c : cavity_) {
      for (int k = 0; k < 4; ++k) {
        const int32_t nb = T[c].n[k];
        if (mark_[nb] == 1) continue;
        int32_t vv[4] = {T[c].v[0], T[c].v[1], T[c].v[2], T[c].v[3]};
        vv[k] = p;
        const int32_t id = new_tet(vv[0], vv[1], vv[2], vv[3]);
        T[id].n[k] = nb;
        for (int m = 0; m < 4; ++m)
          if (T[nb].n[m] == c) T[nb].n[m] = id;
        created_.push_back(id);
      }
    }
    // connect the new tets among themselves through their faces that contain p; such a face is
    // identified by its other two vertices (an edge of the cavity boundary, shared by 2 faces)
    if (++stamp_ == 0) {
      std::fill(hstamp_.begin(), hstamp_.end(), 0u);
      stamp_ = 1;
    }
    const size_t need = created_.size() * 8;
    if (hkey_.size() < need) {
      size_t sz = 1024;
      while (sz < need) sz <<= 1;
      hkey_.assign(sz, 0);
      hval_.assign(sz, 0);
      hstamp_.assign(sz, 0u);
    }
    const size_t mask = hkey_.size() - 1;
    for (int32_t id : created_) {
      const Tet& t = T[id];
      for (int k = 0; k < 4; ++k) {
        if (t.v[k] == p) continue;
        int32_t o[2], m = 0;
        for (int f = 0; f < 3; ++f) {
          const int32_t v = t.v[kFace[k][f]];
          if (v != p) o[m++] = v;
        }
        if (o[0] > o[1]) std::swap(o[0], o[1]);
        const uint64_t key = ((uint64_t)(uint32_t)(o[0] + 2) << 32) | (uint32_t)(o[1] + 2);
        size_t h = (size_t)((key * 0x9E3779B97F4A7C15ull) >> 20) & mask;
        for (;;) {
          if (hstamp_[h] != stamp_) {
            hstamp_[h] = stamp_;
            hkey_[h] = key;
            hval_[h] = ((int64_t)id << 2) | k;
            break;
          }
          if (hkey_[h] == key) {
            const int32_t ot = (int32_t)(hval_[h] >> 2);
            const int of = (int)(hval_[h] & 3);
            T[id].n[k] = ot;
            T[ot].n[of] = id;
            hkey_[h] = ~0ull;  // consumed
            break;
          }
          h = (h + 1) & mask;
        }
      }
    }
    for (int32_t c : cavity_) {
      T[c].v[0] = kDead;
      mark_[c] = 0;
      free_.push_back(c);
    }
    for (int32_t c : tested_) mark_[c] = 0;
    last_ = created_.empty() ? last_ : created_[0];
    inserted_[p] = 1;
    ++ninserted;
    return true;
  }
};

}  // namespace fomo
