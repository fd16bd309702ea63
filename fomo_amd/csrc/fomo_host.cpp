// fomo_host.cpp — host-side parts of the engine's C ABI: the CHIANTI G(T,n) table reader and
// the deterministic synthetic-input generators used by the tests and the benchmark.

#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fomo_hip.h"
#include "delaunay.h"

namespace {

thread_local std::string g_host_err;

// counter-based generator: splitmix64 of (seed, stream, counter) -> 64 random bits
inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
inline uint64_t rbits(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return mix64(mix64(seed ^ (stream * 0xd1b54a32d192ed03ULL)) ^ (ctr * 0x9e3779b97f4a7c15ULL));
}
// uniform in (0,1)
inline double runif(uint64_t seed, uint64_t stream, uint64_t ctr) {
  return ((double)(rbits(seed, stream, ctr) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}
// standard normal (Box-Muller on two counter draws)
inline double rnorm(uint64_t seed, uint64_t stream, uint64_t ctr) {
  const double u1 = runif(seed, stream, 2 * ctr), u2 = runif(seed, stream, 2 * ctr + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

template <class F> void parallel_for(int64_t n, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 4;
  if (nt > 16) nt = 16;
  if (n < 100000) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=]() {
      const int64_t b = n * t / nt, e = n * (t + 1) / nt;
      for (int64_t i = b; i < e; ++i) f(i);
    });
  for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// CHIANTI table text format (fomo-CHIANTI.cpp:143-218): ion, lambda0, atomic weight, min/max
// level, unit line, version line, "nrho nt", nt log T values, then nrho rows of log n + nt G.
int fomo_table_read(const char* path, char* ion, int ion_len, double* lambda0, double* atweight, char* unit,
                    int unit_len, int* nt_out, int* nrho_out, float* logt, float* logn, float* g) {
  if (!path || !lambda0 || !atweight || !nt_out || !nrho_out) return FOMO_ERR_ARG;
  std::ifstream in(path);
  if (!in) return FOMO_ERR_IO;
  std::string sion, unitstr, version;
  double lmin = 0, lmax = 0;
  int nrho = 0, nt = 0;
  in >> sion >> *lambda0 >> *atweight >> lmin >> lmax;
  in.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
  std::getline(in, unitstr);
  std::getline(in, version);
  in >> nrho >> nt;
  if (!in || nrho < 1 || nt < 1) return FOMO_ERR_IO;
  std::vector<float> tv(nt), rows, vals;
  rows.reserve(nrho);
  vals.reserve((size_t)nrho * nt);
  for (int i = 0; i < nt; ++i) {
    double f;
    if (!(in >> f)) return FOMO_ERR_IO;
    tv[i] = (float)f;
  }
  // rows as the reference reads them (fomo-CHIANTI.cpp:178-189): a row cut short at the end of
  // the file keeps its missing values as 0 (a failed extraction stores 0); only a row count that
  // does not match the header is an error (:192-197)
  double r;
  while (in >> r) {
    rows.push_back((float)r);
    for (int i = 0; i < nt; ++i) {
      double f = 0;
      in >> f;
      vals.push_back((float)f);
    }
  }
  if ((int64_t)rows.size() != nrho || (int64_t)vals.size() != (int64_t)nrho * nt) return FOMO_ERR_IO;
  if (ion && ion_len > 0) std::snprintf(ion, ion_len, "%s", sion.c_str());
  if (unit && unit_len > 0) std::snprintf(unit, unit_len, "%s", unitstr.c_str());
  *nt_out = nt;
  *nrho_out = nrho;
  if (logt) {
    if (!logn || !g) return FOMO_ERR_ARG;
    std::memcpy(logt, tv.data(), sizeof(float) * nt);
    std::memcpy(logn, rows.data(), sizeof(float) * nrho);
    std::memcpy(g, vals.data(), sizeof(float) * vals.size());
  }
  return FOMO_OK;
}

int fomo_gen_goft_table(const char* path, const char* ion, double lambda0, double atweight, int nt, int nrho,
                        double peak_logt, double sigma) {
  if (!path || !ion || nt < 2 || nrho < 2 || !(sigma > 0)) return FOMO_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return FOMO_ERR_IO;
  std::fprintf(f, "%s\n", ion);
  std::fprintf(f, "%.6f\n", lambda0);
  std::fprintf(f, "%.6f\n", atweight);
  std::fprintf(f, "%.4f %.4f\n", lambda0 - 0.5, lambda0 + 0.5);
  std::fprintf(f, "erg cm^3 s^{-1} sr^{-1}\n");
  std::fprintf(f, "synthetic-gaussian\n");
  std::fprintf(f, "%d %d\n", nrho, nt);
  std::vector<double> lt(nt), ln(nrho);
  for (int i = 0; i < nt; ++i) lt[i] = 4.0 + 4.0 * i / (nt - 1);
  for (int j = 0; j < nrho; ++j) ln[j] = 8.0 + 4.0 * j / (nrho - 1);
  for (int i = 0; i < nt; ++i) std::fprintf(f, "%.9f%s", lt[i], (i % 5 == 4 || i == nt - 1) ? "\n" : " ");
  for (int j = 0; j < nrho; ++j) {
    std::fprintf(f, "%.9f\n", ln[j]);
    for (int i = 0; i < nt; ++i) {
      const double d = lt[i] - peak_logt;
      const double gv = 1e-24 * std::exp(-d * d / (2 * sigma * sigma)) * (1 + 0.05 * (ln[j] - 10));
      std::fprintf(f, "%.9e%s", gv, (i % 5 == 4 || i == nt - 1) ? "\n" : " ");
    }
  }
  std::fclose(f);
  return FOMO_OK;
}

int fomo_gen_loop_cube(int m, uint64_t seed, double t_in, double t_out, double noise_sigma, double vamp,
                       double jitter, double sausage, float* x, float* y, float* z, float* rho, float* T, float* vx,
                       float* vy, float* vz) {
  if (m < 2 || !x || !y || !z || !rho || !T || !vx || !vy || !vz) return FOMO_ERR_ARG;
  const int64_t n = (int64_t)m * m * m;
  const double L = 10.0, h = L / (m - 1);
  const double nin = std::pow(10.0, 9.5), nout = 1e9;
  const double pi = 3.141592653589793;
  parallel_for(n, [=](int64_t p) {
    const int64_t i = p % m, j = (p / m) % m, k = p / ((int64_t)m * m);
    double xc = -5.0 + h * i, yc = -5.0 + h * j, zc = -5.0 + h * k;
    if (jitter > 0) {
      xc += jitter * h * (2 * runif(seed, 1, p) - 1);
      yc += jitter * h * (2 * runif(seed, 2, p) - 1);
      zc += jitter * h * (2 * runif(seed, 3, p) - 1);
    }
    const double r = std::sqrt(xc * xc + yc * yc);
    const double s = 0.5 * (1.0 - std::tanh((r - 1.0) / 0.1));  // 1 inside the loop, 0 outside
    double ni = nin;
    if (sausage != 0) ni *= 1.0 + sausage * std::cos(2 * pi * zc / 5.0);
    double nn = nout + (ni - nout) * s;
    double tt = t_out + (t_in - t_out) * s;
    // fundamental kink: transverse displacement velocity along x, node-free between footpoints
    double ux = vamp * std::exp(-r * r / 2.0) * std::cos(pi * zc / L);
    double uy = 0.0, uz = 0.0;
    if (noise_sigma > 0) {
      nn *= std::exp(noise_sigma * rnorm(seed, 4, p));
      tt *= std::exp(noise_sigma * rnorm(seed, 5, p));
      ux *= std::exp(noise_sigma * rnorm(seed, 6, p));
    }
    x[p] = (float)xc;
    y[p] = (float)yc;
    z[p] = (float)zc;
    rho[p] = (float)nn;
    T[p] = (float)tt;
    vx[p] = (float)ux;
    vy[p] = (float)uy;
    vz[p] = (float)uz;
  });
  return FOMO_OK;
}

// Corona for Thomson scattering on a Cartesian m^3 grid over [-3,3] Rsun, keeping the points
// with 1 < r/Rsun <= 3 (r <= Rsun has no finite scattering factor). With x==NULL only the
// point count is returned in *count. Density n = 1e8 exp(-(r/Rsun-1)/0.07) (1 + delta), delta a
// sum of nmodes random-phase plane waves with amplitude ~ k^(-11/6) (power ~ k^(-11/3)),
// normalised to rms sigma and clipped at -0.9.
int fomo_gen_corona(int m, uint64_t seed, double sigma, int64_t* count, float* x, float* y, float* z, float* rho) {
  if (m < 2 || !count) return FOMO_ERR_ARG;
  const double rs = 695.7, L = 6.0 * rs, h = L / (m - 1);
  const int64_t n = (int64_t)m * m * m;
  auto inside = [=](int64_t p, double* xc, double* yc, double* zc, double* rr) {
    const int64_t i = p % m, j = (p / m) % m, k = p / ((int64_t)m * m);
    *xc = -3 * rs + h * i;
    *yc = -3 * rs + h * j;
    *zc = -3 * rs + h * k;
    *rr = std::sqrt(*xc * *xc + *yc * *yc + *zc * *zc) / rs;
    return *rr > 1.0 && *rr <= 3.0;
  };
  // prefix counts per z-slab so that points are written in grid order in parallel
  std::vector<int64_t> slab(m + 1, 0);
  {
    std::vector<int64_t> cnt(m, 0);
    parallel_for(m, [&](int64_t k) {
      int64_t c = 0;
      for (int64_t q = 0; q < (int64_t)m * m; ++q) {
        double a, b, cc, r;
        if (inside(k * m * m + q, &a, &b, &cc, &r)) c++;
      }
      cnt[k] = c;
    });
    for (int k = 0; k < m; ++k) slab[k + 1] = slab[k] + cnt[k];
  }
  *count = slab[m];
  if (!x) return FOMO_OK;
  if (!y || !z || !rho) return FOMO_ERR_ARG;
  const int nmodes = 48;
  std::vector<double> kx(nmodes), ky(nmodes), kz(nmodes), ph(nmodes), amp(nmodes);
  double asum = 0;
  const double kmin = 2 * 3.141592653589793 / L, kmax = 2 * 3.141592653589793 / (8 * h);
  for (int q = 0; q < nmodes; ++q) {
    const double u = runif(seed, 10, q);
    const double kk = kmin * std::pow(kmax / kmin, u);
    const double ct = 2 * runif(seed, 11, q) - 1, st = std::sqrt(1 - ct * ct);
    const double pp = 2 * 3.141592653589793 * runif(seed, 12, q);
    kx[q] = kk * st * std::cos(pp);
    ky[q] = kk * st * std::sin(pp);
    kz[q] = kk * ct;
    ph[q] = 2 * 3.141592653589793 * runif(seed, 13, q);
    // log-uniform k sampling: weight k^3 P(k) ~ k^(-2/3) per mode in power
    amp[q] = std::pow(kk / kmin, -1.0 / 3.0);
    asum += amp[q] * amp[q] / 2;
  }
  const double norm = sigma / std::sqrt(asum);
  // plane waves are separable: exp(i k.x) = exp(i kx x) exp(i ky y) exp(i kz z); tabulate the
  // factors per mode along each axis so a point costs two complex products per mode
  typedef std::complex<double> cd;
  std::vector<cd> ex((size_t)nmodes * m), ey((size_t)nmodes * m), ez((size_t)nmodes * m);
  for (int q = 0; q < nmodes; ++q)
    for (int i = 0; i < m; ++i) {
      const double c = -3 * rs + h * i;
      ex[(size_t)q * m + i] = std::polar(1.0, kx[q] * c);
      ey[(size_t)q * m + i] = std::polar(1.0, ky[q] * c);
      ez[(size_t)q * m + i] = std::polar(amp[q] * norm, kz[q] * c + ph[q]);
    }
  parallel_for(m, [&](int64_t k) {
    int64_t w = slab[k];
    for (int64_t q2 = 0; q2 < (int64_t)m * m; ++q2) {
      double xc, yc, zc, r;
      if (!inside(k * m * m + q2, &xc, &yc, &zc, &r)) continue;
      const int64_t i = q2 % m, j = q2 / m;
      double d = 0;
      for (int t = 0; t < nmodes; ++t)
        d += (ex[(size_t)t * m + i] * ey[(size_t)t * m + j] * ez[(size_t)t * m + k]).real();
      if (d < -0.9) d = -0.9;
      x[w] = (float)xc;
      y[w] = (float)yc;
      z[w] = (float)zc;
      rho[w] = (float)(1e8 * std::exp(-(r - 1.0) / 0.07) * (1.0 + d));
      w++;
    }
  });
  return FOMO_OK;
}

// Worker threads of the host Delaunay builder (0: from the environment or the hardware).
static std::atomic<int> g_delaunay_threads{0};

int fomo_set_delaunay_threads(int threads) {
  if (threads < 0) return FOMO_ERR_ARG;
  g_delaunay_threads = threads;
  return FOMO_OK;
}

int fomo_delaunay_threads(void) { return g_delaunay_threads.load(); }

// Delaunay triangulation of a point set on the host (the structure the tetra render walks).
int fomo_delaunay(int64_t n, const float* x, const float* y, const float* z, int64_t cap, int64_t* ntets,
                  int32_t* tets) {
  if (n < 4 || !x || !y || !z || !ntets) return FOMO_ERR_ARG;
  fomo::Delaunay dt;
  dt.threads = fomo_delaunay_threads();
  if (!dt.build(x, y, z, n)) return FOMO_ERR_ARG;
  int64_t m = 0;
  for (const auto& t : dt.T) {
    if (!dt.finite(t)) continue;
    if (tets && m < cap)
      for (int k = 0; k < 4; ++k) tets[4 * m + k] = t.v[k];
    ++m;
  }
  *ntets = m;
  return FOMO_OK;
}

}  // extern "C"
