// FoMo-C render path: CPU oracle.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP path and the
// CPU baseline in bench.py. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it. The product (fomo_amd/) never links or calls it.
//
// It restates the reference algorithm in two modes:
//   mode 0 "ref"  : float storage and float accumulation in the same operation
//                   order as the reference (what the reference binary outputs);
//   mode 1 "fp64" : the same algorithm with double arithmetic everywhere, starting
//                   from the same float32-stored inputs (the accuracy target).
// Geometry (rotation, LOS samples, box, nearest-neighbour choice) is FP64 in both
// modes, exactly as in the reference. Build with -ffp-contract=off (no FMA).
//
// Citations are file:line in the reference tree (fomo-c/src/...).
//
// Parity pinning (see tests/test_oracle_*.py):
//   table parse      : round trip of the synthetic table writer; shipped DN tables.
//   goft             : numpy.interp along log T at the nearest log n row; brute force.
//   linefwhm/emission: closed forms in FP64.
//   thomson factors  : mpmath evaluation of the closed form.
//   rotation/samples : identity view (image = grid); at oblique views, scipy's elementary
//                      rotations, the transposed matrix for the samples, Doppler centroids,
//                      Thomson's matrix blend and sin^2 chi limits, planted-mistake rebuilds.
//   NN query         : brute-force argmin on small inputs; scipy cKDTree fixtures.
//   integration      : uniform slab, line integral and Doppler shift known answers.
//   projection       : pixels of grid points in an axis view (exact), half-away-from-zero
//                      rounding cases, point-order float sums, one-point Doppler profile.
// Unpinned (the reference cannot be built here: its Boost.Geometry, GSL and CGAL dependencies
// are absent): Boost R-tree tie order
// (lowest index is used here), GSL constant versions (CODATA 2006 values used).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// GSL CODATA-2006 constants used by fomo-CHIANTI.cpp:17-19, fomo-nearestneighbour.cpp:28.
const double kSpeedOfLight = 2.99792458e8;
const double kBoltzmann = 1.3806504e-23;
const double kMassProton = 1.672621637e-27;
const double kAlpha = 1.0645;        // fomo-CHIANTI.cpp:16
const double kMmPerArcsec = 0.715;   // FoMo-internal.h:4
const double kPi = M_PI;
// fomo-thomson.cpp:31-34
const double kRsun = 695.7;
const double kU = 0.58;
const double kRe = 2.8179e-13;
const double kL = 2.8e10;

// ---------------------------------------------------------------------------
// kd-tree over the UNROTATED float coordinates (the reference's R-tree also indexes
// unrotated points, fomo-nearestneighbour.cpp:93-100). Query = nearest point that
// lies strictly inside an axis-aligned float box, by double squared distance of
// float coordinates; ties go to the lowest original index.
// ---------------------------------------------------------------------------
struct KdTree {
  struct Node {
    float lo[3], hi[3];
    int32_t begin, end, left, right;
  };
  std::vector<float> p[3];
  std::vector<int32_t> id;
  std::vector<Node> nodes;

  void build(int64_t n, const float* x, const float* y, const float* z) {
    p[0].assign(x, x + n);
    p[1].assign(y, y + n);
    p[2].assign(z, z + n);
    id.resize(n);
    for (int64_t i = 0; i < n; ++i) id[i] = (int32_t)i;
    nodes.clear();
    nodes.reserve(n / 4 + 16);
    if (n > 0) build_node(0, (int32_t)n);
  }

  int32_t build_node(int32_t b, int32_t e) {
    Node nd;
    for (int a = 0; a < 3; ++a) {
      nd.lo[a] = std::numeric_limits<float>::infinity();
      nd.hi[a] = -std::numeric_limits<float>::infinity();
    }
    for (int32_t i = b; i < e; ++i)
      for (int a = 0; a < 3; ++a) {
        float v = p[a][id[i]];
        nd.lo[a] = std::min(nd.lo[a], v);
        nd.hi[a] = std::max(nd.hi[a], v);
      }
    nd.begin = b;
    nd.end = e;
    nd.left = nd.right = -1;
    int32_t me = (int32_t)nodes.size();
    nodes.push_back(nd);
    if (e - b > 8) {
      int ax = 0;
      float ext = nd.hi[0] - nd.lo[0];
      for (int a = 1; a < 3; ++a)
        if (nd.hi[a] - nd.lo[a] > ext) { ext = nd.hi[a] - nd.lo[a]; ax = a; }
      int32_t m = b + (e - b) / 2;
      const std::vector<float>& pa = p[ax];
      std::nth_element(id.begin() + b, id.begin() + m, id.begin() + e,
                       [&](int32_t u, int32_t v) { return pa[u] < pa[v]; });
      int32_t l = build_node(b, m);
      int32_t r = build_node(m, e);
      nodes[me].left = l;
      nodes[me].right = r;
    }
    return me;
  }

  // Distance as Boost.Geometry's comparable pythagoras: ((0+d0^2)+d1^2)+d2^2 in double.
  static inline double dist2(const float q[3], const float t[3]) {
    double d0 = (double)q[0] - (double)t[0];
    double d1 = (double)q[1] - (double)t[1];
    double d2 = (double)q[2] - (double)t[2];
    return (d0 * d0 + d1 * d1) + d2 * d2;
  }
  static inline double mindist2(const Node& nd, const float t[3]) {
    double s = 0.0;
    for (int a = 0; a < 3; ++a) {
      double d = 0.0;
      if (t[a] < nd.lo[a]) d = (double)nd.lo[a] - (double)t[a];
      else if (t[a] > nd.hi[a]) d = (double)t[a] - (double)nd.hi[a];
      s = s + d * d;
    }
    return s;
  }

  // tie (optional): set when another point in the box lies at exactly the nearest distance.
  int32_t query(const float t[3], const float blo[3], const float bhi[3], double* outd = nullptr,
                bool* tie = nullptr) const {
    double best = std::numeric_limits<double>::infinity();
    int32_t bi = -1;
    bool tied = false;
    if (nodes.empty()) return -1;
    int32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
      const Node& nd = nodes[stack[--sp]];
      // no point of this node can be strictly inside the box
      bool out = false;
      for (int a = 0; a < 3; ++a)
        if (!(nd.hi[a] > blo[a] && nd.lo[a] < bhi[a])) out = true;
      if (out) continue;
      if (mindist2(nd, t) > best) continue;
      if (nd.left < 0) {
        for (int32_t i = nd.begin; i < nd.end; ++i) {
          int32_t k = id[i];
          float q[3] = {p[0][k], p[1][k], p[2][k]};
          if (!(blo[0] < q[0] && q[0] < bhi[0] && blo[1] < q[1] && q[1] < bhi[1] &&
                blo[2] < q[2] && q[2] < bhi[2]))
            continue;
          double d = dist2(q, t);
          if (d < best) { best = d; bi = k; tied = false; }
          else if (d == best) { tied = true; if (k < bi) bi = k; }
        }
      } else {
        const Node& L = nodes[nd.left];
        const Node& R = nodes[nd.right];
        double dl = mindist2(L, t), dr = mindist2(R, t);
        if (dl <= dr) { stack[sp++] = nd.right; stack[sp++] = nd.left; }
        else { stack[sp++] = nd.left; stack[sp++] = nd.right; }
      }
    }
    if (outd) *outd = best;
    if (tie) *tie = tied;
    return bi;
  }
};

struct Extents {
  double minx, maxx, miny, maxy, minz, maxz;
};

// ---------------------------------------------------------------------------
// View geometry. Every render function below takes its rotation, LOS unit vector, sample
// positions and search half-width from these helpers, and orc_rotate, orc_los_velocity,
// orc_view_samples and orc_maxdistance export them, so the
// independent checks of tests/test_oracle_geometry.py pin exactly the code the renders run.
// All of it is left-associative FP64 (no FMA), as the reference evaluates it.
// ---------------------------------------------------------------------------

// fomo-nearestneighbour.cpp:83-85 rotation of a data point (FP64 from float coordinates).
inline void rotate_point(double g0, double g1, double g2, double l, double b, double* xr, double* yr, double* zr) {
  *xr = g0 * cos(b) * cos(l) - g1 * cos(b) * sin(l) - g2 * sin(b);
  *yr = g0 * sin(l) + g1 * cos(l);
  *zr = g0 * sin(b) * cos(l) - g1 * sin(b) * sin(l) + g2 * cos(b);
}

// fomo-nearestneighbour.cpp:87-90 LOS unit vector and inner_product (starting from 0.0).
inline void los_unit(double l, double b, double u[3]) {
  u[0] = sin(b) * cos(l);
  u[1] = -sin(b) * sin(l);
  u[2] = cos(b);
}
inline double los_velocity(const double u[3], double vx, double vy, double vz) {
  return ((0.0 + u[0] * vx) + u[1] * vy) + u[2] * vz;
}

// De-rotation of the LOS sample (x, y, z) of the image frame back to the data frame:
// fomo-nearestneighbour.cpp:165, fomo-thomson.cpp:295 (unclamped cos_b/sin_b, :275-276),
// fomo-CGAL.cpp:192.
inline void derotate_sample(double x, double y, double z, double l, double b, double* p0, double* p1, double* p2) {
  *p0 = x * cos(b) * cos(l) + y * sin(l) + z * sin(b) * cos(l);
  *p1 = -x * cos(b) * sin(l) + y * cos(l) - z * sin(b) * sin(l);
  *p2 = -x * sin(b) + z * cos(b);
}

// Image-frame coordinates of ray (i, j) and sample k. kind 0: NearestNeighbour / Thomson,
// z = k*deltaz + minz with deltaz = (maxz-minz)/(nz-1) unless nz == 1
// (fomo-nearestneighbour.cpp:141-162, fomo-thomson.cpp:264-288); kind 2: CGAL,
// z = k/(nz-1)*(maxz-minz) + minz (fomo-CGAL.cpp:187-189; minz when nz == 1, where the
// reference's 0/0 is undefined).
inline void sample_xyz(int kind, const double* e, int nx, int ny, int nz, int i, int j, int k,
                       double* x, double* y, double* z) {
  *x = double(j) / (nx - 1) * (e[1] - e[0]) + e[0];
  *y = double(i) / (ny - 1) * (e[3] - e[2]) + e[2];
  if (kind == 2) {
    *z = nz > 1 ? double(k) / (nz - 1) * (e[5] - e[4]) + e[4] : e[4];
  } else {
    double deltaz = e[5] - e[4];
    if (nz != 1) deltaz /= (nz - 1);
    *z = double(k) * deltaz + e[4];
  }
}

// fomo-nearestneighbour.cpp:138: search half-width of the NearestNeighbour render.
inline double nn_maxdistance(const double* e, int nx, int ny) {
  return std::max((e[1] - e[0]) / (nx - 1), (e[3] - e[2]) / (ny - 1)) / .3;
}

// fomo-thomson.cpp:139-199: blended rotation of a data point (full rotation with eps-clamped
// cos_b/sin_b, sigmoid-blended into the pure z rotation near the pole) and sin^2 chi from the
// blended position and the clamped LOS unit vector.
inline void thomson_rotate(double g0, double g1, double g2, double l, double b, double* xa, double* ya,
                           double* za, double* s2) {
  const double epsilon = 1e-6;  // :139-152
  double cos_b = cos(b), sin_b = sin(b), cos_l = cos(l), sin_l = sin(l);
  if (std::abs(cos_b) < epsilon) cos_b = epsilon;
  if (std::abs(sin_b) < epsilon) sin_b = epsilon;
  double threshold = 1.309, steepness = 20.0;  // :173-176
  double blend = 1.0 / (1.0 + exp(-steepness * (std::fabs(b) - threshold)));
  double full_x = g0 * cos_b * cos_l - g1 * cos_b * sin_l - g2 * sin_b;  // :179-181
  double full_y = g0 * sin_l + g1 * cos_l;
  double full_z = g0 * sin_b * cos_l - g1 * sin_b * sin_l + g2 * cos_b;
  double pole_x = g0 * cos_l - g1 * sin_l;  // :184-186
  double pole_y = g0 * sin_l + g1 * cos_l;
  double pole_z = g2;
  *xa = full_x * (1.0 - blend) + pole_x * blend;  // :189-191
  *ya = full_y * (1.0 - blend) + pole_y * blend;
  *za = full_z * (1.0 - blend) + pole_z * blend;
  double unit[3] = {sin_b * cos(l), -sin_b * sin(l), cos_b};  // :194
  double ru = ((0.0 + unit[0] * *xa) + unit[1] * *ya) + unit[2] * *za;
  double rr = ((0.0 + *xa * *xa) + *ya * *ya) + *za * *za;
  *s2 = 1. - std::pow(ru, 2) / rr;  // :198
}

// fomo-thomson.cpp:249-260: search half-width of the Thomson render, a sigmoid blend in |b|.
inline double thomson_maxdistance(const double* e, int nx, int ny, double b) {
  const double minx = e[0], maxx = e[1], miny = e[2], maxy = e[3];
  double blending_factor = 1.0 / (1.0 + exp(-10 * (std::fabs(b) - 0.1745)));
  double equator_formula = std::max(0.5 * std::sqrt(std::pow((maxx - minx) / (nx - 1), 2) + std::pow((maxy - miny) / (ny - 1), 2)),
                                    0.1 * (maxx - minx) / (nx - 1));
  double away_from_equator_formula = std::max((maxx - minx) / (nx - 1), (maxy - miny) / (ny - 1)) / 0.005;
  return equator_formula * (1.0 - blending_factor) + away_from_equator_formula * blending_factor;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A1: table reader, fomo-CHIANTI.cpp:143-218 (header per docfiles; rows of logn + nt G).
// Call with logt==NULL to get sizes. Returns 0 on success, <0 on error.
// ---------------------------------------------------------------------------
int orc_table_read(const char* path, char* ion, int ionlen, double* lambda0, double* atweight,
                   char* unit, int unitlen, int* nt_out, int* nrho_out,
                   float* logt, float* logn, float* g) {
  std::ifstream in(path);
  if (!in) return -1;
  std::string sion, unitstr, version;
  double minlevel, maxlevel;
  int nrho, nt;
  in >> sion >> *lambda0 >> *atweight >> minlevel >> maxlevel;
  in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  std::getline(in, unitstr);
  std::getline(in, version);
  in >> nrho >> nt;
  if (!in || nrho <= 0 || nt <= 0) return -2;
  std::vector<float> tvec(nt);
  for (int i = 0; i < nt; ++i) { double f; in >> f; tvec[i] = (float)f; }
  std::vector<float> rows, vals;
  double templogrho;
  while (in >> templogrho) {
    rows.push_back((float)templogrho);
    for (int i = 0; i < nt; ++i) { double f; in >> f; vals.push_back((float)f); }
  }
  if ((long)vals.size() != (long)nt * nrho || (long)rows.size() != nrho) return -3;  // :192-197
  if (ion) { std::snprintf(ion, ionlen, "%s", sion.c_str()); }
  if (unit) { std::snprintf(unit, unitlen, "%s", unitstr.c_str()); }
  *nt_out = nt;
  *nrho_out = nrho;
  if (logt) {
    std::copy(tvec.begin(), tvec.end(), logt);
    std::copy(rows.begin(), rows.end(), logn);
    std::copy(vals.begin(), vals.end(), g);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A2-A4: emissionfromdatacube, fomo-CHIANTI.cpp:228-322, with goft :64-121,
// linefwhm :220-225 and the element-wise ops of fomo-operations.cpp:8-114.
// Outputs peak and fwhm per point as double (mode 0 values are float-exact).
// imaging!=0 selects the Imaging branch (fwhm=1, normalisation 1), :275-286.
// abundratio multiplies the spectroscopic normalisation (1 for "/empty"), :257-274.
// Returns the number of points with density <= 0 (reference prints a warning).
// ---------------------------------------------------------------------------
int64_t orc_emission(int mode, int64_t n, const float* rho, const float* T,
                     int nt, int nrho, const float* tlog, const float* nlog, const float* gtab,
                     double lambda0, double atweight, int imaging, double abundratio,
                     double* peak, double* fwhm) {
  const double K = 2 * sqrt(2 * log(2)) * sqrt(kBoltzmann / kMassProton / atweight) * lambda0 / kSpeedOfLight;
  const double normc = imaging ? 1.0 : (1. / kAlpha) * abundratio;
  int64_t nonpos = 0;
#pragma omp parallel for schedule(static) reduction(+ : nonpos)
  for (int64_t i = 0; i < n; ++i) {
    if (rho[i] <= 0) nonpos++;
    if (mode == 0) {
      // float pipeline
      float logrho = rho[i] <= 0 ? 0.f : std::log10(rho[i]);
      float logT = T[i] <= 0 ? 0.f : std::log10(T[i]);
      float fl = std::floor((logT - tlog[0]) / (tlog[1] - tlog[0]));
      float rr = std::round((logrho - nlog[0]) / (nlog[1] - nlog[0]));
      int floortemp = (int)fl, ceiltemp = floortemp + 1, rhoindex = (int)rr;
      double res;
      if (floortemp < 0 || rhoindex < 0 || ceiltemp > nt - 1 || rhoindex > nrho - 1) {
        res = 0;
      } else {
        int gi = rhoindex * nt + ceiltemp;
        double x1 = tlog[ceiltemp], y1 = gtab[gi], x2 = tlog[floortemp], y2 = gtab[gi - 1];
        res = ((double)logT - x1) * (y2 - y1) / (x2 - x1) + y1;
      }
      float gf = (float)res;
      float w = imaging ? 1.f : (float)(K * (double)std::sqrt(T[i]));
      float t1 = (float)(2. * (double)logrho);
      float t2 = (float)std::pow(10., (double)t1);
      float t3 = t2 * gf;
      float t4 = (float)(normc * (double)t3);
      float pk = t4 / w;
      peak[i] = pk;
      fwhm[i] = w;
    } else {
      double logrho = rho[i] <= 0 ? 0. : std::log10((double)rho[i]);
      double logT = T[i] <= 0 ? 0. : std::log10((double)T[i]);
      double fl = std::floor((logT - (double)tlog[0]) / ((double)tlog[1] - (double)tlog[0]));
      double rr = std::round((logrho - (double)nlog[0]) / ((double)nlog[1] - (double)nlog[0]));
      double res;
      if (fl < 0 || rr < 0 || fl + 1 > nt - 1 || rr > nrho - 1) {
        res = 0;
      } else {
        int floortemp = (int)fl, ceiltemp = floortemp + 1, rhoindex = (int)rr;
        int gi = rhoindex * nt + ceiltemp;
        double x1 = tlog[ceiltemp], y1 = gtab[gi], x2 = tlog[floortemp], y2 = gtab[gi - 1];
        res = (logT - x1) * (y2 - y1) / (x2 - x1) + y1;
      }
      double w = imaging ? 1.0 : K * std::sqrt((double)T[i]);
      double pk = normc * (std::pow(10., 2. * logrho) * res) / w;
      peak[i] = pk;
      fwhm[i] = w;
    }
  }
  return nonpos;
}

// ---------------------------------------------------------------------------
// A5: thomsonfromdatacube + getThomsonGeomFactor, fomo-thomson.cpp:39-92.
// Outputs n*tot and n*pol per point.
// ---------------------------------------------------------------------------
void orc_thomson_factor(int mode, double r, double* tot, double* pol) {
  if (mode == 0) {
    float rf = (float)r;
    float sinomega = kRsun / rf;
    float Coeff = 0.5 * kPi * std::pow(kRe, 2) * kL;
    float s2 = sinomega * sinomega;
    float c2 = 1 - s2;
    float cosomega = std::sqrt(c2);
    float logterm = std::log((1. + sinomega) / cosomega);
    float a = cosomega * s2;
    float b = (-1.0 / 8.0) * (1.0 - 3.0 * s2 - c2 * ((1.0 + 3.0 * s2) / sinomega) * logterm);
    float c = (4.0 / 3.0) - cosomega - (cosomega * c2) / 3.0;
    float d = (1.0 / 8.0) * (5.0 + s2 - c2 * ((5.0 - s2) / sinomega) * logterm);
    float polterm = Coeff * (1. / (1. - kU / 3.)) * (a + kU * (b - a));
    float totterm = Coeff * (1. / (1. - kU / 3.)) * (2. * (c + kU * (d - c)));
    *tot = totterm;
    *pol = polterm;
  } else {
    double sinomega = kRsun / r;
    double Coeff = 0.5 * kPi * kRe * kRe * kL;
    double s2 = sinomega * sinomega;
    double c2 = 1 - s2;
    double cosomega = std::sqrt(c2);
    double logterm = std::log((1. + sinomega) / cosomega);
    double a = cosomega * s2;
    double b = (-1.0 / 8.0) * (1.0 - 3.0 * s2 - c2 * ((1.0 + 3.0 * s2) / sinomega) * logterm);
    double c = (4.0 / 3.0) - cosomega - (cosomega * c2) / 3.0;
    double d = (1.0 / 8.0) * (5.0 + s2 - c2 * ((5.0 - s2) / sinomega) * logterm);
    *pol = Coeff * (1. / (1. - kU / 3.)) * (a + kU * (b - a));
    *tot = Coeff * (1. / (1. - kU / 3.)) * (2. * (c + kU * (d - c)));
  }
}

void orc_thomson_emission(int mode, int64_t n, const float* x, const float* y, const float* z,
                          const float* dens, double* totn, double* poln) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) {
    double r = std::sqrt(std::pow((double)x[i], 2) + std::pow((double)y[i], 2) + std::pow((double)z[i], 2));
    double tot, pol;
    if (mode == 0) {
      orc_thomson_factor(0, (float)r, &tot, &pol);
      totn[i] = dens[i] * (float)tot;
      poln[i] = dens[i] * (float)pol;
    } else {
      orc_thomson_factor(1, r, &tot, &pol);
      totn[i] = (double)dens[i] * tot;
      poln[i] = (double)dens[i] * pol;
    }
  }
}

// ---------------------------------------------------------------------------
// A7: per-view rotation extents (fomo-nearestneighbour.cpp:76-108).
// ---------------------------------------------------------------------------
void orc_extents(int64_t n, const float* x, const float* y, const float* z, double l, double b, double* ext6) {
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, mnz = INFINITY, mxz = -INFINITY;
#pragma omp parallel for reduction(min : mnx, mny, mnz) reduction(max : mxx, mxy, mxz)
  for (int64_t i = 0; i < n; ++i) {
    double xr, yr, zr;
    rotate_point(x[i], y[i], z[i], l, b, &xr, &yr, &zr);
    mnx = std::min(mnx, xr); mxx = std::max(mxx, xr);
    mny = std::min(mny, yr); mxy = std::max(mxy, yr);
    mnz = std::min(mnz, zr); mxz = std::max(mxz, zr);
  }
  ext6[0] = mnx; ext6[1] = mxx; ext6[2] = mny; ext6[3] = mxy; ext6[4] = mnz; ext6[5] = mxz;
}

// Brute-force box-constrained nearest neighbour for a list of double targets (tiny inputs).
void orc_nn_bruteforce(int64_t n, const float* x, const float* y, const float* z, int64_t m,
                       const double* px, const double* py, const double* pz, double md, int32_t* out) {
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < m; ++s) {
    float t[3] = {(float)px[s], (float)py[s], (float)pz[s]};
    float blo[3] = {(float)(px[s] - md), (float)(py[s] - md), (float)(pz[s] - md)};
    float bhi[3] = {(float)(px[s] + md), (float)(py[s] + md), (float)(pz[s] + md)};
    double best = INFINITY;
    int32_t bi = -1;
    for (int64_t i = 0; i < n; ++i) {
      float q[3] = {x[i], y[i], z[i]};
      if (!(blo[0] < q[0] && q[0] < bhi[0] && blo[1] < q[1] && q[1] < bhi[1] && blo[2] < q[2] && q[2] < bhi[2]))
        continue;
      double d = KdTree::dist2(q, t);
      if (d < best || (d == best && (int32_t)i < bi)) { best = d; bi = (int32_t)i; }
    }
    out[s] = bi;
  }
}

// kd-tree variant of the same query (for cross-checking the tree against brute force).
void orc_nn_kdtree(int64_t n, const float* x, const float* y, const float* z, int64_t m,
                   const double* px, const double* py, const double* pz, double md, int32_t* out) {
  KdTree kd;
  kd.build(n, x, y, z);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t s = 0; s < m; ++s) {
    float t[3] = {(float)px[s], (float)py[s], (float)pz[s]};
    float blo[3] = {(float)(px[s] - md), (float)(py[s] - md), (float)(pz[s] - md)};
    float bhi[3] = {(float)(px[s] + md), (float)(py[s] + md), (float)(pz[s] + md)};
    out[s] = kd.query(t, blo, bhi);
  }
}

// ---------------------------------------------------------------------------
// A7-A11 + A14: nearestneighbourinterpolation, fomo-nearestneighbour.cpp:31-281, for nlines
// emission cubes of one dataset at once (the reference renders each in its own call; the NN
// query does not depend on the line, so one query per sample serves every line).
// peak, fwhm: nlines * n values (line-major). out: nlines * ny*nx*max(nl,1) intensities (line
// major, index (i*nx+j)*nl+il within a line), already scaled.
// out_idx (optional): ny*nx*nz nearest original indices (-1 = none), index (i*nx+j)*nz+k.
// out_tie (optional): ny*nx*nz flags, 1 where another point in the box lies at exactly the
// nearest distance (the chosen index then rests on the tie rule, lowest index here).
// meta (optional, 10 doubles): minx,maxx,miny,maxy,minz,maxz,maxdistance,pathlength,scale,apix.
// dn!=0 applies the DN/arcsec branch (:252-265). Returns 0, or <0 on bad sizes.
// timing (optional, 3 doubles): seconds for rotate+extents, index build, ray loop.
// row0,row1: render only image rows [row0,row1) (row1<=0: all rows); other rows are untouched.
// ---------------------------------------------------------------------------
int orc_render_nn_multi(int mode, int64_t n, const float* x, const float* y, const float* z, int nlines,
                        const double* peak, const double* fwhm, const float* vx, const float* vy, const float* vz,
                        double l, double b, int nx, int ny, int nz, int nl, double lambda_width, double lambda0,
                        int dn, double* out, int32_t* out_idx, uint8_t* out_tie, double* meta, double* timing,
                        int row0, int row1) {
  if (nx < 2 || ny < 2 || nz < 1 || nl < 1 || n < 1 || nlines < 1) return -1;
  if (row1 <= 0 || row1 > ny) row1 = ny;
  if (row0 < 0) row0 = 0;
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#else
  double t0 = 0;
#endif
  // :76-96 rotation and LOS velocity (double); :103-108 extents
  std::vector<double> losvel(n);
  double ext[6];
  orc_extents(n, x, y, z, l, b, ext);
  double unit[3];
  los_unit(l, b, unit);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; ++i) losvel[i] = los_velocity(unit, vx[i], vy[i], vz[i]);
#ifdef _OPENMP
  double t1 = omp_get_wtime();
#else
  double t1 = 0;
#endif
  KdTree kd;
  kd.build(n, x, y, z);
#ifdef _OPENMP
  double t2 = omp_get_wtime();
#else
  double t2 = 0;
#endif
  const double minx = ext[0], maxx = ext[1], miny = ext[2], maxy = ext[3], minz = ext[4], maxz = ext[5];
  const double lambda_width_in_A = lambda_width * lambda0 / kSpeedOfLight;  // :116
  const double maxdistance = nn_maxdistance(ext, nx, ny);  // :138
  const int nlo = nl > 1 ? nl : 1;
  const int64_t img = (int64_t)nx * ny * nlo;
  std::vector<float> peakf, fwhmf;
  if (mode == 0) {
    peakf.resize((size_t)n * nlines);
    fwhmf.resize((size_t)n * nlines);
    for (int64_t i = 0; i < n * nlines; ++i) { peakf[i] = (float)peak[i]; fwhmf[i] = (float)fwhm[i]; }
  }
  double pathlength = (maxz - minz) / (nz - 1);  // :235-267 scaling
  if ((maxz == minz) || (nz == 1)) pathlength = 1.;
  double apix = 1.;
  if (dn) {
    float dx = (maxx - minx) / (nx - 1), dy = (maxy - miny) / (ny - 1);
    apix = (dx / kMmPerArcsec) * (dy / kMmPerArcsec) * pow(kPi / 180. / 3600., 2);
  }
  const double sc = pathlength * 1e8 * apix;
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int i = row0; i < row1; ++i)
    for (int j = 0; j < nx; ++j) {
      std::vector<float> accf((size_t)nlo * nlines, 0.f);
      std::vector<double> accd((size_t)nlo * nlines, 0.0);
      std::vector<double> intpolpeak(nlines, 0.0), intpolfwhm(nlines, 0.0);
      double intpollosvel = 0;
      for (int k = 0; k < nz; ++k) {
        double xx, yy, zz, p0, p1, p2;
        sample_xyz(0, ext, nx, ny, nz, i, j, k, &xx, &yy, &zz);  // :152-162
        derotate_sample(xx, yy, zz, l, b, &p0, &p1, &p2);        // :165
        float t[3] = {(float)p0, (float)p1, (float)p2};  // :171
        float blo[3] = {(float)(p0 - maxdistance), (float)(p1 - maxdistance), (float)(p2 - maxdistance)};  // :179
        float bhi[3] = {(float)(p0 + maxdistance), (float)(p1 + maxdistance), (float)(p2 + maxdistance)};
        bool tie = false;
        int32_t nearest = kd.query(t, blo, bhi, nullptr, out_tie ? &tie : nullptr);  // :181
        const int64_t sidx = ((int64_t)i * nx + j) * nz + k;
        if (out_idx) out_idx[sidx] = nearest;
        if (out_tie) out_tie[sidx] = tie ? 1 : 0;
        for (int g = 0; g < nlines; ++g) {
          const int64_t q = (int64_t)g * n + nearest;
          if (nearest >= 0) {  // :183-193
            intpolpeak[g] = mode == 0 ? (double)peakf[q] : peak[q];
            intpolfwhm[g] = mode == 0 ? (double)fwhmf[q] : fwhm[q];
          } else {
            intpolpeak[g] = 0;
          }
        }
        if (nearest >= 0) intpollosvel = losvel[nearest];
        for (int g = 0; g < nlines; ++g) {
          float* af = mode == 0 ? &accf[(size_t)g * nlo] : nullptr;
          double* ad = mode == 0 ? nullptr : &accd[(size_t)g * nlo];
          if (nl > 1) {  // :195-215
            for (int il = 0; il < nl; ++il) {
              double lambdaval = double(il) / (nl - 1) * lambda_width_in_A - lambda_width_in_A / 2.;
              double tempintens = intpolpeak[g]
                  ? intpolpeak[g] * exp(-pow(lambdaval - intpollosvel / kSpeedOfLight * lambda0, 2) /
                                        pow(intpolfwhm[g], 2) * 4. * log(2.))
                  : 0;
              if (mode == 0) af[il] = (float)((double)af[il] + tempintens);
              else ad[il] += tempintens;
            }
          } else {  // :217-225
            if (mode == 0) af[0] = (float)((double)af[0] + intpolpeak[g]);
            else ad[0] += intpolpeak[g];
          }
        }
      }
      for (int g = 0; g < nlines; ++g)
        for (int il = 0; il < nlo; ++il) {
          const int64_t ind = (int64_t)g * img + ((int64_t)i * nx + j) * nlo + il;
          const size_t a = (size_t)g * nlo + il;
          out[ind] = mode == 0 ? (double)(float)(sc * (double)accf[a]) : sc * accd[a];
        }
    }
#ifdef _OPENMP
  double t3 = omp_get_wtime();
#else
  double t3 = 0;
#endif
  if (meta) {
    double m[10] = {minx, maxx, miny, maxy, minz, maxz, maxdistance, pathlength, sc, apix};
    std::memcpy(meta, m, sizeof(m));
  }
  if (timing) { timing[0] = t1 - t0; timing[1] = t2 - t1; timing[2] = t3 - t2; }
  return 0;
}

// One emission cube: orc_render_nn_multi with nlines = 1.
int orc_render_nn(int mode, int64_t n, const float* x, const float* y, const float* z,
                  const double* peak, const double* fwhm, const float* vx, const float* vy, const float* vz,
                  double l, double b, int nx, int ny, int nz, int nl, double lambda_width, double lambda0,
                  int dn, double* out, int32_t* out_idx, double* meta, double* timing, int row0, int row1) {
  return orc_render_nn_multi(mode, n, x, y, z, 1, peak, fwhm, vx, vy, vz, l, b, nx, ny, nz, nl, lambda_width, lambda0,
                             dn, out, out_idx, nullptr, meta, timing, row0, row1);
}

// ---------------------------------------------------------------------------
// Projection method: projectioninterpolation, fomo-projection.cpp:17-214. Every point adds its
// line profile (nl > 1, :133-146) or its peak (nl == 1, :148-155) to the pixel its rotated
// (x', y') rounds to (:126-131, std::round). The reference adds with OpenMP atomics in no fixed
// order; this restatement adds in point order (a one-thread run of the reference): mode 0 in
// float as intens.at(ind) += tempintens, mode 1 in double. A degenerate extent (maxx == minx or
// maxy == miny: 0/0 in the reference's pixel formula) puts every point in column (row) 0.
// out: ny*nx*max(nl,1) scaled intensities; out_pix (optional): pixel i*nx+j per point;
// meta as in orc_render_nn with meta[6] = 0.
// ---------------------------------------------------------------------------
static int proj_index(double v, double mn, double mx, int n) {
  if (!(mx > mn)) return 0;
  return (int)std::round((v - mn) * (n - 1) / (mx - mn));  // :129, :131
}

int orc_render_projection(int mode, int64_t n, const float* x, const float* y, const float* z,
                          const double* peak, const double* fwhm, const float* vx, const float* vy,
                          const float* vz, double l, double b, int nx, int ny, int nz, int nl,
                          double lambda_width, double lambda0, int dn, double* out, int32_t* out_pix,
                          double* meta) {
  if (nx < 2 || ny < 2 || nz < 1 || nl < 1 || n < 1) return -1;
  double ext[6];
  orc_extents(n, x, y, z, l, b, ext);  // :45-64 (the same rotation as the NN path)
  const double minx = ext[0], maxx = ext[1], miny = ext[2], maxy = ext[3], minz = ext[4], maxz = ext[5];
  const double lambda_width_in_A = lambda_width * lambda0 / kSpeedOfLight;  // :69
  const int nlo = nl > 1 ? nl : 1;
  const size_t ntot = (size_t)nx * ny * nlo;
  std::vector<float> accf(mode == 0 ? ntot : 0, 0.f);
  std::vector<double> accd(mode == 0 ? 0 : ntot, 0.0);
  double unit[3];
  los_unit(l, b, unit);  // :34
  for (int64_t k = 0; k < n; ++k) {
    double xr, yr, zr;
    rotate_point(x[k], y[k], z[k], l, b, &xr, &yr, &zr);  // :50-52
    const int j = proj_index(xr, minx, maxx, nx);
    const int i = proj_index(yr, miny, maxy, ny);
    if (out_pix) out_pix[k] = i * nx + j;
    if (nl > 1) {
      const double losvel = los_velocity(unit, vx[k], vy[k], vz[k]);  // :53-55
      const double pk = mode == 0 ? (double)(float)peak[k] : peak[k];
      const double fw = mode == 0 ? (double)(float)fwhm[k] : fwhm[k];
      for (int il = 0; il < nl; ++il) {
        const double lambdaval = static_cast<double>(il) / (nl - 1) * lambda_width_in_A - lambda_width_in_A / 2.;
        const double t = pk * exp(-pow(lambdaval - losvel / kSpeedOfLight * lambda0, 2) / pow(fw, 2) * 4. * log(2.));
        const size_t ind = ((size_t)i * nx + j) * nl + il;  // :140
        if (mode == 0) accf[ind] = (float)((double)accf[ind] + t);
        else accd[ind] += t;
      }
    } else {
      const size_t ind = (size_t)i * nx + j;  // :150
      if (mode == 0) accf[ind] = accf[ind] + (float)peak[k];
      else accd[ind] += peak[k];
    }
  }
  // :163-176 path length, :179-200 DN branch and scaling
  double pathlength = (maxz - minz) / (nz - 1);
  if ((maxz == minz) || (nz == 1)) pathlength = 1.;
  double apix = 1.;
  if (dn) {
    float dx = (maxx - minx) / (nx - 1), dy = (maxy - miny) / (ny - 1);
    apix = (dx / kMmPerArcsec) * (dy / kMmPerArcsec) * pow(kPi / 180. / 3600., 2);
  }
  const double sc = pathlength * 1e8 * apix;
  for (size_t q = 0; q < ntot; ++q) out[q] = mode == 0 ? (double)(float)(sc * (double)accf[q]) : sc * accd[q];
  if (meta) {
    double m[10] = {minx, maxx, miny, maxy, minz, maxz, 0.0, pathlength, sc, apix};
    std::memcpy(meta, m, sizeof(m));
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A12: thomsoninterpolation, fomo-thomson.cpp:94-367 (pB, nearest of a k=5 query).
// out: ny*nx pB image (scaled). out_idx, out_tie, meta and rows as in orc_render_nn_multi.
// ---------------------------------------------------------------------------
int orc_render_thomson(int mode, int64_t n, const float* x, const float* y, const float* z,
                       const double* totn, const double* poln, double l, double b, int nx, int ny, int nz,
                       double* out, int32_t* out_idx, uint8_t* out_tie, double* meta, double* timing, int row0,
                       int row1) {
  if (nx < 2 || ny < 2 || nz < 1 || n < 1) return -1;
  if (row1 <= 0 || row1 > ny) row1 = ny;
  if (row0 < 0) row0 = 0;
#ifdef _OPENMP
  double t0 = omp_get_wtime();
#else
  double t0 = 0;
#endif
  std::vector<double> sin2chi(n);
  double mnx = INFINITY, mxx = -INFINITY, mny = INFINITY, mxy = -INFINITY, mnz = INFINITY, mxz = -INFINITY;
#pragma omp parallel for reduction(min : mnx, mny, mnz) reduction(max : mxx, mxy, mxz)
  for (int64_t i = 0; i < n; ++i) {
    double xa, ya, za;
    thomson_rotate(x[i], y[i], z[i], l, b, &xa, &ya, &za, &sin2chi[i]);  // :136-199
    mnx = std::min(mnx, xa); mxx = std::max(mxx, xa);
    mny = std::min(mny, ya); mxy = std::max(mxy, ya);
    mnz = std::min(mnz, za); mxz = std::max(mxz, za);
  }
#ifdef _OPENMP
  double t1 = omp_get_wtime();
#else
  double t1 = 0;
#endif
  KdTree kd;
  kd.build(n, x, y, z);
#ifdef _OPENMP
  double t2 = omp_get_wtime();
#else
  double t2 = 0;
#endif
  const double minx = mnx, maxx = mxx, miny = mny, maxy = mxy, minz = mnz, maxz = mxz;
  const double ext[6] = {mnx, mxx, mny, mxy, mnz, mxz};
  const double maxdistance = thomson_maxdistance(ext, nx, ny, b);  // :249-260
  std::vector<float> totf, polf;
  if (mode == 0) {
    totf.resize(n);
    polf.resize(n);
    for (int64_t i = 0; i < n; ++i) { totf[i] = (float)totn[i]; polf[i] = (float)poln[i]; }
  }
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
  for (int i = row0; i < row1; ++i)
    for (int j = 0; j < nx; ++j) {
      float accf = 0.f;
      double accd = 0.0;
      for (int k = 0; k < nz; ++k) {
        double xx, yy, zz, p0, p1, p2;
        sample_xyz(0, ext, nx, ny, nz, i, j, k, &xx, &yy, &zz);  // :282-288
        derotate_sample(xx, yy, zz, l, b, &p0, &p1, &p2);        // :295 (cos_b, sin_b not clamped, :275-276)
        float t[3] = {(float)p0, (float)p1, (float)p2};
        float blo[3] = {(float)(p0 - maxdistance), (float)(p1 - maxdistance), (float)(p2 - maxdistance)};  // :308
        float bhi[3] = {(float)(p0 + maxdistance), (float)(p1 + maxdistance), (float)(p2 + maxdistance)};
        bool tie = false;
        int32_t nearest = kd.query(t, blo, bhi, nullptr, out_tie ? &tie : nullptr);
        if (out_idx) out_idx[((int64_t)i * nx + j) * nz + k] = nearest;
        if (out_tie) out_tie[((int64_t)i * nx + j) * nz + k] = tie ? 1 : 0;
        double intpoltot = 0, intpolpol = 0, intpollos = 0;
        if (nearest >= 0) {
          intpoltot = mode == 0 ? (double)totf[nearest] : totn[nearest];
          intpolpol = mode == 0 ? (double)polf[nearest] : poln[nearest];
          intpollos = sin2chi[nearest];
        }
        double tempintens2 = intpolpol * intpollos;  // :326
        if (tempintens2 < 0.) tempintens2 = 0;
        (void)intpoltot;
        if (mode == 0) accf = (float)((double)accf + tempintens2);
        else accd += tempintens2;
      }
      double pathlength = (maxz - minz) / (nz - 1);  // :350-357
      if ((maxz == minz) || (nz == 1)) pathlength = 1.;
      double sc = pathlength * 1e8;
      int64_t ind = (int64_t)i * nx + j;
      out[ind] = mode == 0 ? (double)(float)(sc * (double)accf) : sc * accd;
    }
#ifdef _OPENMP
  double t3 = omp_get_wtime();
#else
  double t3 = 0;
#endif
  if (meta) {
    double pathlength = (maxz - minz) / (nz - 1);
    if ((maxz == minz) || (nz == 1)) pathlength = 1.;
    double m[10] = {minx, maxx, miny, maxy, minz, maxz, maxdistance, pathlength, pathlength * 1e8, 1.0};
    std::memcpy(meta, m, sizeof(m));
  }
  if (timing) { timing[0] = t1 - t0; timing[1] = t2 - t1; timing[2] = t3 - t2; }
  return 0;
}

// ---------------------------------------------------------------------------
// View geometry, exported so that tests can check the helpers every render above runs.
// orc_rotate: kind 0 = the rotation of fomo-nearestneighbour.cpp:83-85 (also CGAL :118-121 and
//   Projection :50-52); kind 1 = Thomson's blended rotation (fomo-thomson.cpp:136-191), with
//   s2 (optional) receiving sin^2 chi (:193-198).
// orc_los_velocity: LOS velocity of fomo-nearestneighbour.cpp:87-90.
// orc_view_samples: data-frame position of every LOS sample of a view with extents ext6,
//   layout ((i*nx + j)*nz + k); kind 0 = NearestNeighbour/Thomson z spacing, 2 = CGAL.
// orc_maxdistance: kind 0 = fomo-nearestneighbour.cpp:138, 1 = fomo-thomson.cpp:249-260.
// ---------------------------------------------------------------------------
void orc_rotate(int kind, int64_t n, const float* x, const float* y, const float* z, double l, double b,
                double* xr, double* yr, double* zr, double* s2) {
  for (int64_t i = 0; i < n; ++i) {
    if (kind == 1) {
      double s;
      thomson_rotate(x[i], y[i], z[i], l, b, &xr[i], &yr[i], &zr[i], &s);
      if (s2) s2[i] = s;
    } else {
      rotate_point(x[i], y[i], z[i], l, b, &xr[i], &yr[i], &zr[i]);
    }
  }
}

void orc_rotate_d(int64_t n, const double* x, const double* y, const double* z, double l, double b, double* xr,
                  double* yr, double* zr) {
  for (int64_t i = 0; i < n; ++i) rotate_point(x[i], y[i], z[i], l, b, &xr[i], &yr[i], &zr[i]);
}

void orc_los_velocity(int64_t n, const float* vx, const float* vy, const float* vz, double l, double b,
                      double* out) {
  double unit[3];
  los_unit(l, b, unit);
  for (int64_t i = 0; i < n; ++i) out[i] = los_velocity(unit, vx[i], vy[i], vz[i]);
}

void orc_view_samples(int kind, const double* ext6, double l, double b, int nx, int ny, int nz, double* p0,
                      double* p1, double* p2) {
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j)
      for (int k = 0; k < nz; ++k) {
        double xx, yy, zz;
        sample_xyz(kind, ext6, nx, ny, nz, i, j, k, &xx, &yy, &zz);
        const int64_t s = ((int64_t)i * nx + j) * nz + k;
        derotate_sample(xx, yy, zz, l, b, &p0[s], &p1[s], &p2[s]);
      }
}

double orc_maxdistance(int kind, const double* ext6, int nx, int ny, double b) {
  return kind == 1 ? thomson_maxdistance(ext6, nx, ny, b) : nn_maxdistance(ext6, nx, ny);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
