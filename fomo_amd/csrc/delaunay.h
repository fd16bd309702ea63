// 3D Delaunay triangulation (host), replacing the CGAL triangulation of the reference's CGAL
// render path (fomo-CGAL.cpp:33-73, Delaunay_triangulation_3 of the data points).
//
// Incremental Bowyer-Watson insertion with an infinite vertex (index -1): every hull face has an
// infinite tet on its outer side, so points outside the current hull are inserted like the
// others. Each point is located by a visibility walk from a recently created tet, its conflict
// region (tets whose circumsphere holds it, infinite tets whose hull face it sees) is removed and
// the region's boundary is connected to the point.
// Predicates are exact (predicates.h); cospherical configurations are resolved by a symbolic
// perturbation that depends only on the coordinates, so the triangulation is unique and
// independent of the insertion order. Duplicate points are inserted once: the vertex keeps the
// index of the LAST input point with those coordinates (the reference reads the values through
// std::map keyed by the point, where later insertions overwrite earlier ones,
// fomo-CGAL.cpp:128-131).
//
// Insertion order and threads. With one thread the points go in Morton order. With several, the
// points are split into levels of growing density (each point joins a coarser level with
// probability 1/8, the biased randomized insertion order of Amenta, Choi and Rote); the coarsest
// level goes in serially, each finer one in parallel passes. In a pass every vertex belongs to one
// of T chunks (equal-count ranges of a Morton order of the coordinates, shifted differently in
// each pass) and worker t inserts the points of chunk t in Morton order, under two rules:
//   * it changes (kills, or relinks as a boundary tet) only tets whose finite vertices all belong
//     to chunk t, and its new tets join a point of chunk t to faces of such tets;
//   * it reads only tets that hold a vertex of chunk t, plus the vertex ids of tets reached from
//     those across a face.
// A tet with a vertex of chunk t is never changed by another worker, and no other worker reads a
// tet that worker t changes, so the workers share the tet array without locks. Killing a tet also
// relinks its neighbours, all of which worker t owns, so no other worker can hold a pointer to it.
// A point whose walk, conflict region or region boundary leaves chunk t is deferred to the next
// pass (shifted chunks put it in the interior), and what the passes leave goes in serially.
// Because the triangulation is unique the result equals the serial one; build() stores it in a
// canonical order (below), so the output does not depend on the thread count either.
//
// Tet layout: v[4] vertex ids (point indices, -1 = infinite), n[4] neighbour across the face
// opposite v[i]. Finite tets are positively oriented: orient(v0,v1,v2,v3) > 0. The face opposite
// v[i] is kFace[i], ordered so that orient(face, v[i]) > 0. Canonical order: each tet is rotated
// by an even permutation (orientation kept) to put its smallest vertex first and the smallest of
// the rest second (the infinite vertex counts as the largest), and the tets are sorted by their
// vertices in the Morton order of the points, which also keeps neighbouring tets close in memory.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "predicates.h"

namespace fomo {

constexpr int kFace[4][3] = {{2, 1, 3}, {0, 2, 3}, {1, 0, 3}, {0, 1, 2}};

struct Tet {
  int32_t v[4];
  int32_t n[4];
};

// vector storage that leaves new elements uninitialised (large arrays filled right after)
template <class X>
struct NoInitAlloc : std::allocator<X> {
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) {}
  template <class U>
  void construct(U* p) {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
using TetVec = std::vector<Tet, NoInitAlloc<Tet>>;

// spread the low 21 bits of x to every third bit (Morton interleave)
inline uint64_t morton_spread(uint32_t x) {
  uint64_t v = x & 0x1fffffu;
  v = (v | v << 32) & 0x1f00000000ffffull;
  v = (v | v << 16) & 0x1f0000ff0000ffull;
  v = (v | v << 8) & 0x100f00f00f00f00full;
  v = (v | v << 4) & 0x10c30c30c30c30c3ull;
  v = (v | v << 2) & 0x1249249249249249ull;
  return v;
}

// f(0..nt-1) on nt threads (the caller's thread runs f(0))
template <class F>
void run_threads(int nt, F f) {
  if (nt <= 1) {
    f(0);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt - 1);
  for (int t = 1; t < nt; ++t) th.emplace_back(f, t);
  f(0);
  for (auto& x : th) x.join();
}

class Delaunay {
 public:
  std::vector<pred::P3> P;   // input points
  TetVec T;                  // after build: the live tets (finite and infinite), canonical order
  std::vector<int32_t> rep;  // input point -> vertex point carrying its location (duplicates)
  int64_t ninserted = 0, nduplicates = 0, nfinite = 0;
  // worker threads: 0 = FOMO_DELAUNAY_THREADS, else OMP_NUM_THREADS, else the hardware threads
  int threads = 0;
  int nthreads_used = 1;
  int64_t nparallel = 0;  // points inserted by the parallel passes (the rest went in serially)
  std::string error;

  static constexpr int32_t kInf = -1;
  static constexpr int32_t kDead = -2;

  Delaunay() = default;
  Delaunay(const Delaunay&) = delete;
  Delaunay& operator=(const Delaunay&) = delete;
  ~Delaunay() { release(); }

  bool finite(const Tet& t) const { return t.v[0] >= 0 && t.v[1] >= 0 && t.v[2] >= 0 && t.v[3] >= 0; }

  bool build(const float* x, const float* y, const float* z, int64_t n) {
    release();
    T.clear();
    error.clear();
    ninserted = nduplicates = nfinite = nparallel = 0;
    P.resize(n);
    for (int64_t i = 0; i < n; ++i) P[i] = {(double)x[i], (double)y[i], (double)z[i]};
    rep.assign(n, -1);
    if (n < 4) return fail("fewer than 4 points");
    if (n > (int64_t)INT32_MAX / 8) return fail("too many points");
    const int nt = resolve_threads();
    nthreads_used = nt;
    const auto t0 = std::chrono::steady_clock::now();
    // optional progress lines on stderr (FOMO_PROGRESS=1), as the reference prints its own
    // "Doing Delaunay triangulation" progress (fomo-CGAL.cpp:55)
    const char* pe = std::getenv("FOMO_PROGRESS");
    progress_ = pe && pe[0] == '1';
    // Morton order of the points
    for (int a = 0; a < 3; ++a) lo_[a] = 1e300, hi_[a] = -1e300;
    for (auto& p : P) {
      lo_[0] = std::min(lo_[0], p.x), lo_[1] = std::min(lo_[1], p.y), lo_[2] = std::min(lo_[2], p.z);
      hi_[0] = std::max(hi_[0], p.x), hi_[1] = std::max(hi_[1], p.y), hi_[2] = std::max(hi_[2], p.z);
    }
    std::vector<std::pair<uint64_t, int32_t>> ord(n);
    run_threads(nt, [&](int t) {
      for (int64_t i = n * t / nt; i < n * (t + 1) / nt; ++i) ord[i] = {morton(P[i], 0), (int32_t)i};
    });
    par_sort(ord, nt, [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
      return a < b;
    });
    // duplicates: equal coordinates give equal codes, so they are neighbours within a run of equal
    // codes of the sorted order; the vertex of a group is its last input index
    std::vector<int32_t> seq;
    seq.reserve(n);
    {
      const auto same = [&](int32_t a, int32_t b) { return P[a].x == P[b].x && P[a].y == P[b].y && P[a].z == P[b].z; };
      std::vector<int32_t> run;
      for (int64_t s = 0; s < n;) {
        int64_t e = s + 1;
        while (e < n && ord[e].first == ord[s].first) ++e;
        if (e - s == 1) {
          rep[ord[s].second] = ord[s].second;
        } else {
          // a run of equal codes (a cell of 2^-21 of the box): group its points by coordinates
          run.clear();
          for (int64_t k = s; k < e; ++k) run.push_back(ord[k].second);
          std::sort(run.begin(), run.end(), [&](int32_t a, int32_t b) {
            if (P[a].x != P[b].x) return P[a].x < P[b].x;
            if (P[a].y != P[b].y) return P[a].y < P[b].y;
            if (P[a].z != P[b].z) return P[a].z < P[b].z;
            return a < b;
          });
          for (size_t g = 0; g < run.size();) {
            size_t h = g + 1;
            while (h < run.size() && same(run[h], run[g])) ++h;
            for (size_t k = g; k < h; ++k) rep[run[k]] = run[h - 1];
            nduplicates += (int64_t)(h - g - 1);
            g = h;
          }
        }
        s = e;
      }
      for (auto& o : ord)
        if (rep[o.second] == o.second) seq.push_back(o.second);
    }
    std::vector<std::pair<uint64_t, int32_t>>().swap(ord);
    // work on the points renumbered in Morton order (memory locality), map back at the end
    std::vector<pred::P3> input;
    input.swap(P);
    const int64_t m = (int64_t)seq.size();
    P.resize(m);
    for (int64_t i = 0; i < m; ++i) P[i] = input[seq[i]];
    const auto t1 = std::chrono::steady_clock::now();
    bool ok = insert_all(nt);
    const auto t2 = std::chrono::steady_clock::now();
    if (ok) canonical(nt, seq);
    if (progress_) {
      const auto t3 = std::chrono::steady_clock::now();
      std::fprintf(stderr, "Delaunay: %d threads, order %.2f s, insertion %.2f s, canonical order %.2f s\n", nt,
                   std::chrono::duration<double>(t1 - t0).count(), std::chrono::duration<double>(t2 - t1).count(),
                   std::chrono::duration<double>(t3 - t2).count());
    }
    release();
    P.swap(input);
    return ok;
  }

 private:
  struct Worker {
    int id = 0;
    std::vector<int32_t> free, stack, cavity, tested, created, deferred;
    // edge -> (new tet, face) table for connecting the tets of one insertion (stamped entries)
    std::vector<uint64_t> hkey;
    std::vector<int64_t> hval;
    std::vector<uint32_t> hstamp;
    uint32_t stamp = 0;
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    int32_t last = -1;          // walk start: a tet created by the last insertion
    // in a parallel pass: per cell of a coarse grid over the points, a tet this worker created or
    // was given as a start there (always safe for it to read), the fallback start of a walk
    std::vector<int32_t> hint;
    int64_t cnext = 0, cend = 0;  // unused slots of the chunk taken from the slot array
    int64_t ninserted = 0;
    uint32_t rnd() {
      rng ^= rng << 13;
      rng ^= rng >> 7;
      rng ^= rng << 17;
      return (uint32_t)rng;
    }
  };

  static constexpr int64_t kChunk = 4096;    // tet slots taken at a time by a worker
  static constexpr int64_t kLevelBase = 1024;  // least size of the coarsest level (several threads)
  static constexpr int kPasses = 5;            // parallel passes per level
  static constexpr int kMaxThreads = 64;
  static constexpr int kHintGrid = 32;         // cells per axis of the walk start grids
  static constexpr int kSeeds = 64;            // walk starts located per chunk before a pass

  Tet* S_ = nullptr;     // tet slots (dead ones have v[0] == kDead)
  char* mark_ = nullptr;  // per slot: 1 in conflict, 2 tested not in conflict
  int64_t cap_ = 0;
  std::atomic<int64_t> top_{0};
  std::vector<char> inserted_;
  std::vector<std::unique_ptr<Worker>> W_;
  const uint8_t* own_ = nullptr;  // per vertex: its chunk in the current parallel pass
  std::vector<std::vector<uint8_t>> owners_;
  double lo_[3] = {1e300, 1e300, 1e300}, hi_[3] = {-1e300, -1e300, -1e300};
  bool progress_ = false;

  bool fail(const std::string& m) {
    error = m;
    return false;
  }

  void release() {
    std::free(S_);
    std::free(mark_);
    S_ = nullptr;
    mark_ = nullptr;
    cap_ = 0;
    top_ = 0;
    W_.clear();
    owners_.clear();
    own_ = nullptr;
    std::vector<char>().swap(inserted_);
  }

  int resolve_threads() const {
    int t = threads;
    for (const char* env : {"FOMO_DELAUNAY_THREADS", "OMP_NUM_THREADS"}) {
      if (t > 0) break;
      const char* e = std::getenv(env);
      if (e && std::atoi(e) > 0) t = std::atoi(e);
    }
    if (t <= 0) t = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(t, kMaxThreads));
  }

  // Morton code of a point over the bounding box: 21-bit cell coordinates (the insertion order),
  // or with sh, 20-bit coordinates offset by sh[] < 2^20 cells (the chunks of the passes)
  uint64_t morton(const pred::P3& p, const uint32_t* sh) const {
    const double c[3] = {p.x, p.y, p.z};
    const double s = sh ? 1048576.0 : 2097152.0;
    uint32_t u[3];
    for (int a = 0; a < 3; ++a) {
      const double e = hi_[a] - lo_[a];
      const double t = e > 0 ? (c[a] - lo_[a]) / e : 0.0;
      u[a] = (uint32_t)std::min(s - 1.0, std::max(0.0, t * s));
      if (sh) u[a] += sh[a];
    }
    return morton_spread(u[0]) << 2 | morton_spread(u[1]) << 1 | morton_spread(u[2]);
  }

  // cell of the walk start grids
  int hint_cell(const pred::P3& p) const {
    const double c[3] = {p.x, p.y, p.z};
    int k[3];
    for (int a = 0; a < 3; ++a) {
      const double e = hi_[a] - lo_[a];
      const double t = e > 0 ? (c[a] - lo_[a]) / e : 0.0;
      k[a] = (int)std::min((double)(kHintGrid - 1), std::max(0.0, t * kHintGrid));
    }
    return (k[2] * kHintGrid + k[1]) * kHintGrid + k[0];
  }

  // sort: nt sorted ranges, merged pairwise in parallel
  template <class V, class C>
  static void par_sort(V& a, int nt, C cmp) {
    const int64_t n = (int64_t)a.size();
    if (nt <= 1 || n < 65536) {
      std::sort(a.begin(), a.end(), cmp);
      return;
    }
    std::vector<int64_t> cut(nt + 1);
    for (int t = 0; t <= nt; ++t) cut[t] = n * t / nt;
    run_threads(nt, [&](int t) { std::sort(a.begin() + cut[t], a.begin() + cut[t + 1], cmp); });
    for (int w = 1; w < nt; w *= 2) {
      std::vector<int> starts;
      for (int t = 0; t + w < nt; t += 2 * w) starts.push_back(t);
      run_threads((int)starts.size(), [&](int k) {
        const int t = starts[k];
        std::inplace_merge(a.begin() + cut[t], a.begin() + cut[t + w], a.begin() + cut[std::min(nt, t + 2 * w)], cmp);
      });
    }
  }

  // ---- slots ----

  bool grow(int64_t need) {
    int64_t nc = std::max<int64_t>(need, cap_ + cap_ / 2 + kChunk);
    Tet* s = (Tet*)std::realloc(S_, sizeof(Tet) * nc);
    if (!s) return false;
    S_ = s;
    char* mk = (char*)std::realloc(mark_, nc);
    if (!mk) return false;
    std::memset(mk + cap_, 0, nc - cap_);
    mark_ = mk;
    cap_ = nc;
    return true;
  }

  // make k slots available to w (its free list and its chunk); in a parallel pass the slot array
  // cannot grow and the caller defers the point instead
  bool reserve(Worker& w, size_t k, bool par) {
    while (w.free.size() + (size_t)(w.cend - w.cnext) < k) {
      for (int64_t s = w.cnext; s < w.cend; ++s) {
        S_[s].v[0] = kDead;
        w.free.push_back((int32_t)s);
      }
      w.cnext = w.cend;
      int64_t t = top_.load();
      for (;;) {
        if (t + kChunk > cap_) {
          if (par || t + kChunk > (int64_t)INT32_MAX || !grow(t + kChunk)) return false;
          continue;
        }
        if (top_.compare_exchange_weak(t, t + kChunk)) break;
      }
      w.cnext = t;
      w.cend = t + kChunk;
    }
    return true;
  }

  int32_t new_tet(Worker& w, int32_t a, int32_t b, int32_t c, int32_t d) {
    int32_t id;
    if (!w.free.empty()) {
      id = w.free.back();
      w.free.pop_back();
    } else {
      id = (int32_t)w.cnext++;
    }
    Tet& t = S_[id];
    t.v[0] = a, t.v[1] = b, t.v[2] = c, t.v[3] = d;
    t.n[0] = t.n[1] = t.n[2] = t.n[3] = -1;
    mark_[id] = 0;
    return id;
  }

  const pred::P3& pt(int32_t v) const { return P[v]; }

  // ---- ownership in a parallel pass ----

  // worker `me` may read the tet: one of its finite vertices is in chunk me
  bool readable(const Tet& t, int me) const {
    for (int k = 0; k < 4; ++k)
      if (t.v[k] >= 0 && own_[t.v[k]] == me) return true;
    return false;
  }
  // worker `me` may change the tet: all its finite vertices are in chunk me
  bool owned(const Tet& t, int me) const {
    for (int k = 0; k < 4; ++k)
      if (t.v[k] >= 0 && own_[t.v[k]] != me) return false;
    return true;
  }

  // ---- insertion ----

  // first tet from 4 affinely independent points, plus the 4 infinite tets on its faces
  bool init(Worker& w, const std::vector<int32_t>& seq) {
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
    if (!reserve(w, 5, false)) return fail("out of memory (tets)");
    const int32_t t0 = new_tet(w, a, b, c, d);
    int32_t inf[4];
    for (int i = 0; i < 4; ++i) {
      // outer tet on face i: the face vertices plus the infinite vertex in place of v[i]
      int32_t vv[4] = {S_[t0].v[0], S_[t0].v[1], S_[t0].v[2], S_[t0].v[3]};
      // the infinite tet shares face i but lies on its other side: swap two face vertices
      const int f0 = kFace[i][0], f1 = kFace[i][1];
      std::swap(vv[f0], vv[f1]);
      vv[i] = kInf;
      inf[i] = new_tet(w, vv[0], vv[1], vv[2], vv[3]);
    }
    for (int i = 0; i < 4; ++i) S_[t0].n[i] = inf[i];
    const int32_t ids[5] = {t0, inf[0], inf[1], inf[2], inf[3]};
    link_all(ids, 5);
    for (int32_t p : {a, b, c, d}) inserted_[p] = 1;
    w.last = t0;
    w.ninserted += 4;
    return true;
  }

  // connect the faces of a small set of tets that share 3 vertices
  void link_all(const int32_t* ids, int n) {
    for (int x = 0; x < n; ++x)
      for (int i = 0; i < 4; ++i) {
        const Tet& A = S_[ids[x]];
        int32_t fa[3] = {A.v[kFace[i][0]], A.v[kFace[i][1]], A.v[kFace[i][2]]};
        std::sort(fa, fa + 3);
        for (int y = 0; y < n; ++y) {
          if (y == x) continue;
          const Tet& B = S_[ids[y]];
          for (int j = 0; j < 4; ++j) {
            int32_t fb[3] = {B.v[kFace[j][0]], B.v[kFace[j][1]], B.v[kFace[j][2]]};
            std::sort(fb, fb + 3);
            if (fa[0] == fb[0] && fa[1] == fb[1] && fa[2] == fb[2]) S_[ids[x]].n[i] = ids[y];
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
  bool conflict(int32_t id, int32_t p) const {
    const Tet& t = S_[id];
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
    const Tet& u = S_[t.n[j]];
    return pred::insphere_perturbed(pt(u.v[0]), pt(u.v[1]), pt(u.v[2]), pt(u.v[3]), pt(p)) < 0;
  }

  // tet containing p (finite), or an infinite tet whose hull face p sees; -1: no end, -2 (parallel
  // pass): the walk reached a tet that worker w may not read
  template <bool PAR>
  int32_t locate(Worker& w, int32_t p, int32_t start) {
    int32_t t = start;
    const pred::P3& q = pt(p);
    for (int64_t steps = 0; steps < 100000000; ++steps) {
      const Tet& c = S_[t];
      if (PAR && !readable(c, w.id)) return -2;
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
      const int s = (int)(w.rnd() & 3);
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

  void clear_marks(Worker& w) {
    for (int32_t c : w.cavity) mark_[c] = 0;
    for (int32_t c : w.stack) mark_[c] = 0;
    for (int32_t c : w.tested) mark_[c] = 0;
  }

  // insert point p, walking from tet `start`: 1 inserted, 0 deferred (parallel pass: the region
  // or its boundary leaves the chunk), 2 deferred (the walk left the chunk), -1 failed (serial,
  // error set)
  template <bool PAR>
  int insert(Worker& w, int32_t p, int32_t start) {
    const int32_t t0 = locate<PAR>(w, p, start);
    if (t0 == -2) return 2;
    if (t0 < 0) return PAR ? 0 : (fail("point location did not terminate"), -1);
    if (PAR && !owned(S_[t0], w.id)) return 0;
    // the located tet may fail the strict conflict test only in degenerate hull cases: then search
    // its neighbours (a point on the hull plane outside the circumcircle belongs to another face)
    int32_t seed = -1;
    if (conflict(t0, p)) {
      seed = t0;
    } else {
      for (int k = 0; k < 4 && seed < 0; ++k) {
        const int32_t nb = S_[t0].n[k];
        if (PAR && !owned(S_[nb], w.id)) return 0;
        if (conflict(nb, p)) seed = nb;
      }
      if (seed < 0) {
        if (PAR) return 0;
        // fall back: scan the tets around (rare); flood over infinite tets
        std::vector<int32_t> st{t0};
        std::vector<char> seen(top_.load(), 0);
        seen[t0] = 1;
        while (!st.empty() && seed < 0) {
          const int32_t c = st.back();
          st.pop_back();
          if (S_[c].v[0] != kDead && conflict(c, p)) seed = c;
          for (int k = 0; k < 4; ++k) {
            const int32_t nb = S_[c].n[k];
            if (nb >= 0 && !seen[nb] && (st.size() < 4096)) {
              seen[nb] = 1;
              st.push_back(nb);
            }
          }
        }
      }
      if (seed < 0) return fail("no conflict region for a point (degenerate input)"), -1;
    }
    // conflict region by flood fill; in a parallel pass every tet tested must be owned
    w.cavity.clear();
    w.stack.clear();
    w.tested.clear();
    w.stack.push_back(seed);
    mark_[seed] = 1;
    size_t nbound = 0;
    while (!w.stack.empty()) {
      const int32_t c = w.stack.back();
      w.stack.pop_back();
      w.cavity.push_back(c);
      for (int k = 0; k < 4; ++k) {
        const int32_t nb = S_[c].n[k];
        // (a tet this insertion marked is owned; the check comes first so that a parallel pass
        // never reads the mark of a tet that another worker may be marking)
        if (PAR && !owned(S_[nb], w.id)) {
          clear_marks(w);
          return 0;
        }
        if (mark_[nb] == 1) continue;
        if (mark_[nb] == 2) {
          ++nbound;
          continue;
        }
        if (conflict(nb, p)) {
          mark_[nb] = 1;
          w.stack.push_back(nb);
        } else {
          mark_[nb] = 2;
          w.tested.push_back(nb);
          ++nbound;
        }
      }
    }
    if (!reserve(w, nbound, PAR)) {
      clear_marks(w);
      if (PAR) return 0;
      return fail("out of memory (tets)"), -1;
    }
    // boundary faces: (cavity tet c, face k) with neighbour outside; the new tet is c with its
    // vertex k replaced by p (same orientation: p sees the face from c's side)
    w.created.clear();
    for (int32_t c : w.cavity) {
      for (int k = 0; k < 4; ++k) {
        const int32_t nb = S_[c].n[k];
        if (mark_[nb] == 1) continue;
        int32_t vv[4] = {S_[c].v[0], S_[c].v[1], S_[c].v[2], S_[c].v[3]};
        vv[k] = p;
        const int32_t id = new_tet(w, vv[0], vv[1], vv[2], vv[3]);
        S_[id].n[k] = nb;
        for (int m = 0; m < 4; ++m)
          if (S_[nb].n[m] == c) S_[nb].n[m] = id;
        w.created.push_back(id);
      }
    }
    link_created(w, p);
    for (int32_t c : w.cavity) {
      S_[c].v[0] = kDead;
      mark_[c] = 0;
      w.free.push_back(c);
    }
    for (int32_t c : w.tested) mark_[c] = 0;
    if (!w.created.empty()) w.last = w.created[0];
    inserted_[p] = 1;
    ++w.ninserted;
    return 1;
  }

  // connect the new tets of one insertion among themselves through their faces that contain p;
  // such a face is identified by its other two vertices (an edge of the cavity boundary, shared
  // by 2 faces)
  void link_created(Worker& w, int32_t p) {
    if (++w.stamp == 0) {
      std::fill(w.hstamp.begin(), w.hstamp.end(), 0u);
      w.stamp = 1;
    }
    const size_t need = w.created.size() * 8;
    if (w.hkey.size() < need) {
      size_t sz = 1024;
      while (sz < need) sz <<= 1;
      w.hkey.assign(sz, 0);
      w.hval.assign(sz, 0);
      w.hstamp.assign(sz, 0u);
    }
    const size_t mask = w.hkey.size() - 1;
    for (int32_t id : w.created) {
      const Tet& t = S_[id];
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
          if (w.hstamp[h] != w.stamp) {
            w.hstamp[h] = w.stamp;
            w.hkey[h] = key;
            w.hval[h] = ((int64_t)id << 2) | k;
            break;
          }
          if (w.hkey[h] == key) {
            const int32_t ot = (int32_t)(w.hval[h] >> 2);
            const int of = (int)(w.hval[h] & 3);
            S_[id].n[k] = ot;
            S_[ot].n[of] = id;
            w.hkey[h] = ~0ull;  // consumed
            break;
          }
          h = (h + 1) & mask;
        }
      }
    }
  }

  // ---- drivers ----

  bool insert_all(int nt) {
    const int64_t m = (int64_t)P.size();
    inserted_.assign(m, 0);
    // slots: about 6.6 live tets per point; the serial phases grow the array if needed
    if (!grow(7 * m + 2 * kChunk * (nt + 1))) return fail("out of memory (tets)");
    for (int t = 0; t < nt; ++t) {
      W_.push_back(std::make_unique<Worker>());
      W_.back()->id = t;
      W_.back()->rng ^= 0x632be59bd9b4e019ull * (uint64_t)(t + 1);
    }
    Worker& w0 = *W_[0];
    // levels: with one thread a single level in Morton order
    int L = 0;
    if (nt > 1)
      while (L < 8 && (m >> (3 * (L + 1))) >= kLevelBase) ++L;
    std::vector<std::vector<int32_t>> lev(L + 1);
    for (int64_t i = 0; i < m; ++i) {
      uint64_t h = (uint64_t)i * 0x9E3779B97F4A7C15ull;
      h ^= h >> 29;
      h *= 0xBF58476D1CE4E5B9ull;
      h ^= h >> 32;
      int lv = L;
      while (lv > 0 && (h & 7) == 0) {
        --lv;
        h >>= 3;
      }
      lev[lv].push_back((int32_t)i);
    }
    {
      std::vector<int32_t> order;
      order.reserve(m);
      for (auto& l : lev) order.insert(order.end(), l.begin(), l.end());
      if (!init(w0, order)) return false;
    }
    const bool progress = progress_;
    for (int k = 0; k <= L; ++k) {
      if (k == 0) {
        if (!insert_serial(lev[0], progress)) return false;
      } else if (!insert_level(lev[k], nt)) {
        return false;
      }
      if (progress)
        std::fprintf(stderr, "Delaunay: level %d of %d done, %zu points, %lld of them in parallel\n", k, L,
                     lev[k].size(), (long long)nparallel);
    }
    ninserted = 0;
    for (auto& w : W_) ninserted += w->ninserted;
    // the unused rest of every chunk is dead
    for (auto& w : W_) {
      for (int64_t s = w->cnext; s < w->cend; ++s) S_[s].v[0] = kDead;
      w->cnext = w->cend;
    }
    return true;
  }

  // an alive tet to start walks from
  int32_t anchor() const {
    for (auto& w : W_)
      if (w->last >= 0 && S_[w->last].v[0] != kDead) return w->last;
    for (int64_t s = 0; s < top_.load(); ++s)
      if (S_[s].v[0] != kDead) return (int32_t)s;
    return -1;
  }

  bool insert_serial(const std::vector<int32_t>& pts, bool progress) {
    Worker& w = *W_[0];
    if (w.last < 0 || S_[w.last].v[0] == kDead) w.last = anchor();
    for (size_t i = 0; i < pts.size(); ++i) {
      if (!inserted_[pts[i]] && insert<false>(w, pts[i], w.last) < 0) return false;
      if (progress && (i + 1) % 1000000 == 0)
        std::fprintf(stderr, "Delaunay: %zu of %zu points inserted serially\n", i + 1, pts.size());
    }
    return true;
  }

  // chunk of every vertex for the parallel passes with shift s: equal-count ranges of the shifted
  // Morton codes (quantiles of a sample)
  const uint8_t* owners(int s, int nt) {
    if ((int)owners_.size() <= s) owners_.resize(s + 1);
    auto& o = owners_[s];
    if (!o.empty()) return o.data();
    // offsets of the grid under the Morton order, in cells of 2^-20 of the box: different in each
    // pass, so that the chunk borders of one pass cross the interiors of the next pass's chunks
    static constexpr double kShift[kPasses][3] = {{0.0, 0.0, 0.0}, {0.31, 0.19, 0.27}, {0.13, 0.41, 0.57},
                                                   {0.47, 0.29, 0.11}, {0.23, 0.53, 0.37}};
    const uint32_t sh[3] = {(uint32_t)(kShift[s][0] * 1048576.0), (uint32_t)(kShift[s][1] * 1048576.0),
                            (uint32_t)(kShift[s][2] * 1048576.0)};
    const int64_t m = (int64_t)P.size();
    const int64_t step = std::max<int64_t>(1, m / 65536);
    std::vector<uint64_t> smp;
    for (int64_t i = 0; i < m; i += step) smp.push_back(morton(P[i], sh));
    std::sort(smp.begin(), smp.end());
    std::vector<uint64_t> thr(nt - 1);
    for (int t = 1; t < nt; ++t) thr[t - 1] = smp[smp.size() * t / nt];
    o.resize(m);
    run_threads(nt, [&](int t) {
      for (int64_t i = m * t / nt; i < m * (t + 1) / nt; ++i)
        o[i] = (uint8_t)(std::upper_bound(thr.begin(), thr.end(), morton(P[i], sh)) - thr.begin());
    });
    return o.data();
  }

  bool insert_level(const std::vector<int32_t>& pts, int nt) {
    std::vector<int32_t> work;
    for (int32_t p : pts)
      if (!inserted_[p]) work.push_back(p);
    int32_t anc = anchor();
    for (int pass = 0; pass < kPasses && (int64_t)work.size() >= 64 * (int64_t)nt; ++pass) {
      own_ = owners(pass, nt);
      std::vector<std::vector<int32_t>> mine(nt);
      for (int32_t p : work) mine[own_[p]].push_back(p);
      // walk starts (serial, nothing changes meanwhile): the tets holding points spread over each
      // chunk's list, where the chunk's worker may read them
      for (int t = 0; t < nt; ++t) {
        Worker& w = *W_[t];
        w.last = -1;
        w.hint.assign((size_t)kHintGrid * kHintGrid * kHintGrid, -1);
        int32_t at = anc;
        const size_t np = mine[t].size();
        for (size_t j = 0; j < (size_t)kSeeds && j < np; ++j) {
          const int32_t q = mine[t][j * np / std::min(np, (size_t)kSeeds)];
          const int32_t f = locate<false>(*W_[0], q, at);
          if (f < 0) break;
          at = f;
          if (!readable(S_[f], t)) continue;
          w.hint[hint_cell(pt(q))] = f;
          if (w.last < 0) w.last = f;
        }
      }
      const auto tp0 = std::chrono::steady_clock::now();
      std::vector<int64_t> before(nt);
      for (int t = 0; t < nt; ++t) before[t] = W_[t]->ninserted;
      run_threads(nt, [&](int t) {
        Worker& w = *W_[t];
        w.deferred.clear();
        for (int32_t p : mine[t]) {
          int r = w.last < 0 ? 2 : insert<true>(w, p, w.last);
          const int cell = hint_cell(pt(p));
          if (r == 2) {
            // the walk from the last insertion left the chunk: try the start recorded near p
            const int32_t h = w.hint[cell];
            r = h >= 0 && h != w.last && S_[h].v[0] != kDead ? insert<true>(w, p, h) : 0;
          }
          if (r == 1)
            w.hint[cell] = w.last;
          else
            w.deferred.push_back(p);
        }
      });
      own_ = nullptr;
      work.clear();
      for (int t = 0; t < nt; ++t) {
        Worker& w = *W_[t];
        work.insert(work.end(), w.deferred.begin(), w.deferred.end());
        nparallel += w.ninserted - before[t];
        // a worker that inserted holds a live tet (only it could change its own new tets)
        if (w.ninserted > before[t]) anc = w.last;
      }
      std::sort(work.begin(), work.end());
      if (progress_)
        std::fprintf(stderr, "Delaunay: pass %d, %zu points deferred, %.3f s\n", pass, work.size(),
                     std::chrono::duration<double>(std::chrono::steady_clock::now() - tp0).count());
    }
    W_[0]->last = anc;
    return insert_serial(work, false);
  }

  // ---- canonical output ----

  static void canon(Tet& t) {
    static constexpr int kEven[4][4] = {{0, 1, 2, 3}, {1, 0, 3, 2}, {2, 3, 0, 1}, {3, 2, 1, 0}};
    int i = 0;
    for (int k = 1; k < 4; ++k)
      if ((uint32_t)t.v[k] < (uint32_t)t.v[i]) i = k;
    Tet r;
    for (int k = 0; k < 4; ++k) r.v[k] = t.v[kEven[i][k]], r.n[k] = t.n[kEven[i][k]];
    int j = 1;
    for (int k = 2; k < 4; ++k)
      if ((uint32_t)r.v[k] < (uint32_t)r.v[j]) j = k;
    // rotate positions 1..3 (a 3-cycle, even) to bring j to position 1
    for (int rot = 0; rot < j - 1; ++rot) {
      const int32_t v1 = r.v[1], n1 = r.n[1];
      r.v[1] = r.v[2], r.v[2] = r.v[3], r.v[3] = v1;
      r.n[1] = r.n[2], r.n[2] = r.n[3], r.n[3] = n1;
    }
    t = r;
  }

  static bool tet_less(const Tet& a, const Tet& b) {
    for (int k = 0; k < 4; ++k)
      if (a.v[k] != b.v[k]) return (uint32_t)a.v[k] < (uint32_t)b.v[k];
    return false;
  }

  // live tets into T in canonical order (vertex ids still in Morton numbering while sorting), then
  // neighbours renumbered and vertices mapped to input indices
  void canonical(int nt, const std::vector<int32_t>& seq) {
    const int64_t ns = top_.load(), m = (int64_t)P.size();
    std::vector<int32_t> cnt(m + 1, 0);
    run_threads(nt, [&](int t) {
      for (int64_t s = ns * t / nt; s < ns * (t + 1) / nt; ++s)
        if (S_[s].v[0] != kDead) {
          canon(S_[s]);
          __atomic_fetch_add(&cnt[S_[s].v[0]], 1, __ATOMIC_RELAXED);
        }
    });
    int64_t tot = 0;
    for (int64_t v = 0; v <= m; ++v) {
      const int32_t c = cnt[v];
      cnt[v] = (int32_t)tot;
      tot += c;
    }
    std::vector<int32_t> start(cnt);  // bucket starts (cnt becomes the cursors)
    T.resize(tot);
    std::vector<int32_t, NoInitAlloc<int32_t>> old(tot), newid(ns);
    run_threads(nt, [&](int t) {
      for (int64_t s = ns * t / nt; s < ns * (t + 1) / nt; ++s)
        if (S_[s].v[0] != kDead) {
          const int32_t pos = __atomic_fetch_add(&cnt[S_[s].v[0]], 1, __ATOMIC_RELAXED);
          T[pos] = S_[s];
          old[pos] = (int32_t)s;
        }
    });
    // within a bucket (same first vertex): insertion sort by the other vertices
    run_threads(nt, [&](int t) {
      for (int64_t v = m * t / nt; v < m * (t + 1) / nt; ++v) {
        const int64_t b = start[v], e = start[v + 1];
        for (int64_t i = b + 1; i < e; ++i) {
          const Tet x = T[i];
          const int32_t ox = old[i];
          int64_t j = i;
          while (j > b && tet_less(x, T[j - 1])) {
            T[j] = T[j - 1];
            old[j] = old[j - 1];
            --j;
          }
          T[j] = x;
          old[j] = ox;
        }
        for (int64_t i = b; i < e; ++i) newid[old[i]] = (int32_t)i;
      }
    });
    std::vector<int64_t> nf(nt, 0);
    run_threads(nt, [&](int t) {
      for (int64_t i = tot * t / nt; i < tot * (t + 1) / nt; ++i) {
        Tet& x = T[i];
        for (int k = 0; k < 4; ++k) {
          x.n[k] = newid[x.n[k]];
          if (x.v[k] >= 0) x.v[k] = seq[x.v[k]];
        }
        if (finite(x)) ++nf[t];
      }
    });
    nfinite = 0;
    for (int64_t f : nf) nfinite += f;
  }
};

}  // namespace fomo
