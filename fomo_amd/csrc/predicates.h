// Exact geometric predicates for the Delaunay builder (host only).
//
// orient(a,b,c,d) = sign det[b-a; c-a; d-a]; insphere5(a,b,c,d,e) = sign of the 5x5 determinant
// with rows (x, y, z, x^2+y^2+z^2, 1), equal to the 4x4 determinant with rows
// (p - e, |p - e|^2) for p = a,b,c,d. For orient(a,b,c,d) > 0, insphere5 < 0 exactly when e is
// strictly inside the circumsphere of a,b,c,d. Both are evaluated in double with a conservative error
// filter; when the filter cannot decide, the determinant is recomputed exactly with floating-point
// expansion arithmetic (sums and products of non-overlapping double components, the method of
// Priest and Shewchuk: two-sum, Dekker's two-product, expansion sum, scaling and compression).
//
// Degenerate (cospherical) cases get a symbolic perturbation of the lifted coordinate:
// w_i -> w_i + eps^rank(i), ranks by descending lexicographic (x, y, z) order, which makes every
// 5-point test non-zero and independent of insertion order (see insphere_perturbed).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace fomo {
namespace pred {

// Expansion: non-overlapping double components in increasing magnitude. Inline storage for the
// common short expansions, heap beyond that (no allocation on the usual exact paths).
class Exp {
 public:
  Exp() = default;
  Exp(const Exp& o) { assign(o); }
  Exp& operator=(const Exp& o) {
    if (this != &o) assign(o);
    return *this;
  }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  double operator[](size_t i) const { return data()[i]; }
  double& operator[](size_t i) { return data()[i]; }
  void reserve(size_t m) {
    if (m <= cap()) return;
    std::vector<double> h(m);
    for (size_t i = 0; i < n_; ++i) h[i] = data()[i];
    heap_.swap(h);
  }
  void push_back(double x) {
    if (n_ == cap()) reserve(2 * cap());
    data()[n_++] = x;
  }

 private:
  static constexpr size_t kInline = 96;
  double inl_[kInline];
  std::vector<double> heap_;
  size_t n_ = 0;
  size_t cap() const { return heap_.empty() ? kInline : heap_.size(); }
  double* data() { return heap_.empty() ? inl_ : heap_.data(); }
  const double* data() const { return heap_.empty() ? inl_ : heap_.data(); }
  void assign(const Exp& o) {
    n_ = 0;
    heap_.clear();
    reserve(o.n_);
    for (size_t i = 0; i < o.n_; ++i) data()[i] = o[i];
    n_ = o.n_;
  }
};

inline void two_sum(double a, double b, double& x, double& y) {
  x = a + b;
  const double bv = x - a;
  const double av = x - bv;
  y = (a - av) + (b - bv);
}

// Veltkamp split of a double into two 26-bit halves (hi + lo == a exactly)
inline void split(double a, double& hi, double& lo) {
  const double c = 134217729.0 * a;  // 2^27 + 1
  const double big = c - a;
  hi = c - big;
  lo = a - hi;
}

// x + y == a * b exactly (Dekker's product; no FMA instruction or library call needed)
inline void two_prod(double a, double b, double& x, double& y) {
  x = a * b;
  double ah, al, bh, bl;
  split(a, ah, al);
  split(b, bh, bl);
  const double e1 = x - ah * bh;
  const double e2 = e1 - al * bh;
  const double e3 = e2 - ah * bl;
  y = al * bl - e3;
}

// h = e + (sign) f (both expansions), zero components dropped: merge by magnitude, then
// accumulate with two-sum
inline Exp add_signed(const Exp& e, const Exp& f, double fs) {
  Exp h;
  const size_t ne = e.size(), nf = f.size();
  if (ne + nf == 0) return h;
  h.reserve(ne + nf);
  size_t i = 0, j = 0;
  auto next = [&]() -> double {
    if (i < ne && (j >= nf || std::fabs(e[i]) < std::fabs(f[j]))) return e[i++];
    return fs * f[j++];
  };
  double q = next();
  for (size_t k = 1; k < ne + nf; ++k) {
    double sm, t;
    two_sum(q, next(), sm, t);
    if (t != 0.0) h.push_back(t);
    q = sm;
  }
  if (q != 0.0 || h.empty()) h.push_back(q);
  return h;
}

inline Exp add(const Exp& e, const Exp& f) { return add_signed(e, f, 1.0); }
inline Exp sub(const Exp& e, const Exp& f) { return add_signed(e, f, -1.0); }

// h = e * b
inline Exp scale(const Exp& e, double b) {
  Exp h;
  if (e.empty() || b == 0.0) return h;
  h.reserve(2 * e.size());
  double q, t;
  two_prod(e[0], b, q, t);
  if (t != 0.0) h.push_back(t);
  for (size_t i = 1; i < e.size(); ++i) {
    double p1, p0, sm, u;
    two_prod(e[i], b, p1, p0);
    two_sum(q, p0, sm, u);
    if (u != 0.0) h.push_back(u);
    two_sum(p1, sm, q, u);
    if (u != 0.0) h.push_back(u);
  }
  if (q != 0.0 || h.empty()) h.push_back(q);
  return h;
}

// Shortest equivalent expansion (Priest/Shewchuk compression): a top-down pass gathers the
// components into non-adjacent ones, a bottom-up pass renormalises. The value is unchanged.
inline Exp compress(const Exp& e) {
  const size_t m = e.size();
  if (m < 3) return e;
  Exp g;
  g.reserve(m);
  for (size_t i = 0; i < m; ++i) g.push_back(0.0);
  size_t bottom = m - 1;
  double q = e[m - 1];
  for (size_t i = m - 1; i-- > 0;) {
    double qn, r;
    two_sum(q, e[i], qn, r);
    if (r != 0.0) {
      g[bottom--] = qn;
      q = r;
    } else {
      q = qn;
    }
  }
  g[bottom] = q;
  Exp h;
  h.reserve(m - bottom);
  for (size_t i = bottom + 1; i < m; ++i) {
    double qn, r;
    two_sum(g[i], q, qn, r);
    if (r != 0.0) h.push_back(r);
    q = qn;
  }
  h.push_back(q);
  return h;
}

inline Exp mul(const Exp& e, const Exp& f) {
  Exp h;
  for (size_t i = 0; i < f.size(); ++i) {
    h = add(h, scale(e, f[i]));
    if (h.size() > 16) h = compress(h);
  }
  return h.size() > 4 ? compress(h) : h;
}

inline Exp diff(double a, double b) {
  double x, y;
  two_sum(a, -b, x, y);
  Exp e;
  if (y != 0.0) e.push_back(y);
  e.push_back(x);
  return e;
}

inline int sign(const Exp& e) {
  for (size_t i = e.size(); i-- > 0;)
    if (e[i] != 0.0) return e[i] > 0 ? 1 : -1;
  return 0;
}

inline Exp det3(const Exp m[3][3]) {
  const Exp a = mul(m[0][0], sub(mul(m[1][1], m[2][2]), mul(m[1][2], m[2][1])));
  const Exp b = mul(m[0][1], sub(mul(m[1][0], m[2][2]), mul(m[1][2], m[2][0])));
  const Exp c = mul(m[0][2], sub(mul(m[1][0], m[2][1]), mul(m[1][1], m[2][0])));
  return add(sub(a, b), c);
}

struct P3 {
  double x, y, z;
};

inline int orient_exact(const P3& a, const P3& b, const P3& c, const P3& d) {
  const Exp m[3][3] = {{diff(b.x, a.x), diff(b.y, a.y), diff(b.z, a.z)},
                       {diff(c.x, a.x), diff(c.y, a.y), diff(c.z, a.z)},
                       {diff(d.x, a.x), diff(d.y, a.y), diff(d.z, a.z)}};
  return sign(det3(m));
}

// sign det[b-a; c-a; d-a]
inline int orient(const P3& a, const P3& b, const P3& c, const P3& d) {
  const double bx = b.x - a.x, by = b.y - a.y, bz = b.z - a.z;
  const double cx = c.x - a.x, cy = c.y - a.y, cz = c.z - a.z;
  const double dx = d.x - a.x, dy = d.y - a.y, dz = d.z - a.z;
  const double m1 = cy * dz - cz * dy, m2 = cx * dz - cz * dx, m3 = cx * dy - cy * dx;
  const double det = bx * m1 - by * m2 + bz * m3;
  const double perm = std::fabs(bx) * (std::fabs(cy * dz) + std::fabs(cz * dy)) +
                      std::fabs(by) * (std::fabs(cx * dz) + std::fabs(cz * dx)) +
                      std::fabs(bz) * (std::fabs(cx * dy) + std::fabs(cy * dx));
  // rounding of the differences and of the products: far below 1e-12 of the permanent
  const double bound = 1e-12 * perm;
  if (det > bound) return 1;
  if (det < -bound) return -1;
  return orient_exact(a, b, c, d);
}

// 4x4 determinant with rows (p - e, |p - e|^2), exactly
inline int insphere_exact(const P3* p4, const P3& e) {
  Exp d[4][3], w[4];
  for (int i = 0; i < 4; ++i) {
    d[i][0] = diff(p4[i].x, e.x);
    d[i][1] = diff(p4[i].y, e.y);
    d[i][2] = diff(p4[i].z, e.z);
    w[i] = add(add(mul(d[i][0], d[i][0]), mul(d[i][1], d[i][1])), mul(d[i][2], d[i][2]));
  }
  // Laplace expansion along the lifted column (column 3): sum_i (-1)^(i+3) w_i M_i
  Exp tot;
  for (int i = 0; i < 4; ++i) {
    Exp m[3][3];
    int r = 0;
    for (int j = 0; j < 4; ++j) {
      if (j == i) continue;
      for (int c = 0; c < 3; ++c) m[r][c] = d[j][c];
      ++r;
    }
    const Exp t = mul(w[i], det3(m));
    tot = ((i + 3) & 1) ? sub(tot, t) : add(tot, t);
  }
  return sign(tot);
}

// sign of the 5x5 lifted determinant (a,b,c,d,e); 0 when cospherical (or degenerate).
// With p' = p - e and w = |p'|^2: D = -w_a det(b',c',d') + w_b det(a',c',d') - w_c det(a',b',d')
// + w_d det(a',b',c'), each 3x3 determinant from the six 2x2 minors of the (x, y) columns.
inline int insphere5(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e) {
  const double ax = a.x - e.x, ay = a.y - e.y, az = a.z - e.z;
  const double bx = b.x - e.x, by = b.y - e.y, bz = b.z - e.z;
  const double cx = c.x - e.x, cy = c.y - e.y, cz = c.z - e.z;
  const double dx = d.x - e.x, dy = d.y - e.y, dz = d.z - e.z;
  const double axby = ax * by, bxay = bx * ay, axcy = ax * cy, cxay = cx * ay, axdy = ax * dy, dxay = dx * ay;
  const double bxcy = bx * cy, cxby = cx * by, bxdy = bx * dy, dxby = dx * by, cxdy = cx * dy, dxcy = dx * cy;
  const double ab = axby - bxay, ac = axcy - cxay, ad = axdy - dxay;
  const double bc = bxcy - cxby, bd = bxdy - dxby, cd = cxdy - dxcy;
  const double abc = az * bc - bz * ac + cz * ab;
  const double abd = az * bd - bz * ad + dz * ab;
  const double acd = az * cd - cz * ad + dz * ac;
  const double bcd = bz * cd - cz * bd + dz * bc;
  const double wa = ax * ax + ay * ay + az * az, wb = bx * bx + by * by + bz * bz;
  const double wc = cx * cx + cy * cy + cz * cz, wd = dx * dx + dy * dy + dz * dz;
  const double det = ((wb * acd - wa * bcd) + (wd * abc - wc * abd));
  const double pab = std::fabs(axby) + std::fabs(bxay), pac = std::fabs(axcy) + std::fabs(cxay);
  const double pad = std::fabs(axdy) + std::fabs(dxay), pbc = std::fabs(bxcy) + std::fabs(cxby);
  const double pbd = std::fabs(bxdy) + std::fabs(dxby), pcd = std::fabs(cxdy) + std::fabs(dxcy);
  const double aaz = std::fabs(az), abz = std::fabs(bz), acz = std::fabs(cz), adz = std::fabs(dz);
  const double perm = wa * (abz * pcd + acz * pbd + adz * pbc) + wb * (aaz * pcd + acz * pad + adz * pac) +
                      wc * (aaz * pbd + abz * pad + adz * pab) + wd * (aaz * pbc + abz * pac + acz * pab);
  // rounding of the differences, minors, lifts and sums: far below 1e-11 of the permanent
  const double bound = 1e-11 * perm;
  if (det > bound) return 1;
  if (det < -bound) return -1;
  const P3 p[4] = {a, b, c, d};
  return insphere_exact(p, e);
}

inline bool lex_greater(const P3& a, const P3& b) {
  if (a.x != b.x) return a.x > b.x;
  if (a.y != b.y) return a.y > b.y;
  return a.z > b.z;
}

// insphere5 with the symbolic perturbation w_i += eps^rank(i): rank 0 (the dominant term) is the
// lexicographically largest point. dD5/dw_i = (-1)^i orient(the other four, in order), so the
// perturbed sign is that of the first non-zero such cofactor in rank order. For a non-degenerate
// tet (a,b,c,d) the cofactor of e is orient(a,b,c,d) != 0, so the result is never 0.
inline int insphere_perturbed(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e) {
  const int s = insphere5(a, b, c, d, e);
  if (s != 0) return s;
  const P3 p[5] = {a, b, c, d, e};
  int ord[5] = {0, 1, 2, 3, 4};
  std::sort(ord, ord + 5, [&](int i, int j) { return lex_greater(p[i], p[j]); });
  for (int k = 0; k < 5; ++k) {
    const int i = ord[k];
    P3 q[4];
    int r = 0;
    for (int j = 0; j < 5; ++j)
      if (j != i) q[r++] = p[j];
    const int o = orient(q[0], q[1], q[2], q[3]);
    if (o != 0) return (i & 1) ? -o : o;
  }
  return 0;  // only for a degenerate (flat) tet
}

}  // namespace pred
}  // namespace fomo
