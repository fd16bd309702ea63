// fomo_device.hip — MI355X (gfx950) render engine for the FoMo-C render path.
//
// Device-resident objects (datasets with their cell-list index, G(T,n) tables, emission
// cubes, batches) live in HBM and are addressed by ids through the C ABI in fomo_hip.h.
//
// Numerics:
//  * geometry (rotation, LOS sample de-rotation, search box, distances) is FP64 with no FMA
//    contraction (this file is compiled with -ffp-contract=off) and host-computed glibc trig,
//    so nearest-neighbour choices are bit-identical to the reference's R-tree query;
//  * emissivity is computed in FP64 and stored float32 (the reference stores float32);
//  * the Gaussian line profile uses an FP64 ratio recurrence and FP64 accumulation along
//    each ray (FP64 VALU runs at the non-packed FP32 rate on CDNA4, so this is both the
//    cheapest and the most accurate option).
//
// Reference citations are fomo-c/src/<file>:<line>.

#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/fomo_hip.h"
#include "arena.h"
#include "delaunay.h"
#include "host_copy.h"
#include "lookup_geom.h"
#include "octree.h"
#include "orient_exact.h"
#include "spectral_exp.h"
#include "fomo_common.h"

// ------------------------------------------------------------------------------------
// constants (GSL CODATA-2006 values used by fomo-CHIANTI.cpp:17-19)
// ------------------------------------------------------------------------------------

// Scratch of the allocation-free NN render (fomo_render_nn_noalloc): one array per GPU, held
// until the process ends. In host code `scratch` names a host-side stand-in, not device memory;
// fomo_init takes the device address with hipGetSymbolAddress (Ctx::noalloc) and host code uses
// only that address.
static __device__ __align__(256) unsigned char scratch[FOMO_NOALLOC_SCRATCH_BYTES];

namespace {

using namespace fomo_dev;

thread_local std::string g_err;
std::recursive_mutex g_mu;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPOK(expr)                                                                      \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(FOMO_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(_e));        \
  } while (0)

// ------------------------------------------------------------------------------------
// device buffers
// ------------------------------------------------------------------------------------
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) bytes = b;
    else p = nullptr;
    return e;
  }
  // grow-only reuse
  hipError_t ensure(size_t b) {
    if (b <= bytes && p) return hipSuccess;
    return alloc(b);
  }
  // grow-only reuse for buffers that must read zero whenever they are idle
  hipError_t ensure_zeroed(size_t b, hipStream_t s) {
    if (b <= bytes && p) return hipSuccess;
    hipError_t e = alloc(b);
    return e == hipSuccess ? hipMemsetAsync(p, 0, bytes, s) : e;
  }
  template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Block-wide min/max of 6 doubles (vals[0,2,4] minima, [1,3,5] maxima) -> out (6 doubles).
__device__ void block_minmax6(double v[6], double* out) {
  __shared__ double red[kBlock / 64][6];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < 6; ++a) v[a] = (a & 1) ? wave_max(v[a]) : wave_min(v[a]);
  if (lane == 0)
#pragma unroll
    for (int a = 0; a < 6; ++a) red[wid][a] = v[a];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int a = 0; a < 6; ++a) {
      double r = red[0][a];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) r = (a & 1) ? fmax(r, red[w][a]) : fmin(r, red[w][a]);
      out[a] = r;
    }
  }
}

// ------------------------------------------------------------------------------------
// K3: spatial index build (once per dataset; view-invariant)
// ------------------------------------------------------------------------------------
__global__ void k_bbox(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                       int64_t n, double* __restrict__ partial) {
  double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double a = x[i], b = y[i], c = z[i];
    v[0] = fmin(v[0], a); v[1] = fmax(v[1], a);
    v[2] = fmin(v[2], b); v[3] = fmax(v[3], b);
    v[4] = fmin(v[4], c); v[5] = fmax(v[5], c);
  }
  block_minmax6(v, partial + 6 * blockIdx.x);
}

// Largest squared distance of the points from c (in slot 1 of the block's 6 partials).
__global__ void k_sphere_r2(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                            int64_t n, double c0, double c1, double c2, double* __restrict__ partial) {
  double v[6] = {INFINITY, 0.0, INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double a = (double)x[i] - c0, b = (double)y[i] - c1, c = (double)z[i] - c2;
    v[1] = fmax(v[1], (a * a + b * b) + c * c);
  }
  block_minmax6(v, partial + 6 * blockIdx.x);
}

__global__ void k_reduce_partials(const double* __restrict__ partial, int np, double* __restrict__ out6) {
  double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < np; i += blockDim.x)
#pragma unroll
    for (int a = 0; a < 6; ++a) v[a] = (a & 1) ? fmax(v[a], partial[6 * i + a]) : fmin(v[a], partial[6 * i + a]);
  block_minmax6(v, out6);
}

// A product grid may have holes (a masked grid) when it has at most this many nodes per point.
constexpr int64_t kMaskedGridMaxRatio = 4;

// Grid position of every point on a candidate product grid (axis values v0, v1, v2 ascending):
// perm[gid] = original index. Sets *bad when a coordinate is not an axis value or two points
// share a grid position (perm must be all -1 on entry).
__device__ __forceinline__ int axis_index(const float* __restrict__ v, int n, float c) {
  int a = 0, b = n;  // first index with v[idx] >= c
  while (a < b) {
    const int m = (a + b) >> 1;
    if (v[m] < c) a = m + 1;
    else b = m;
  }
  return a < n && v[a] == c ? a : -1;
}
__global__ void k_rect_map(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                           int64_t n, const float* __restrict__ v0, const float* __restrict__ v1,
                           const float* __restrict__ v2, int n0, int n1, int n2, int32_t* __restrict__ perm,
                           int* __restrict__ bad) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const int i = axis_index(v0, n0, x[p]), j = axis_index(v1, n1, y[p]), k = axis_index(v2, n2, z[p]);
    if (i < 0 || j < 0 || k < 0) {
      atomicOr(bad, 1);
      continue;
    }
    const int64_t gid = i + (int64_t)n0 * (j + (int64_t)n1 * k);
    if (atomicCAS(&perm[gid], -1, (int32_t)p) != -1) atomicOr(bad, 1);
  }
}

__global__ void k_cell_keys(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                            int64_t n, GridInfo g, uint32_t* __restrict__ key, int32_t* __restrict__ val) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int cx = cell_coord(g, 0, x[i]), cy = cell_coord(g, 1, y[i]), cz = cell_coord(g, 2, z[i]);
  key[i] = (uint32_t)(((int64_t)cz * g.dims[1] + cy) * g.dims[0] + cx);
  val[i] = (int32_t)i;
}

// cstart[c] = first sorted position with key >= c, for c in [0, ncells]
__global__ void k_cell_start(const uint32_t* __restrict__ skey, int64_t n, int64_t ncells, int32_t* __restrict__ cstart) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p > n) return;
  int64_t lo = (p == 0) ? 0 : (int64_t)skey[p - 1] + 1;
  int64_t hi = (p == n) ? ncells : (int64_t)skey[p];
  for (int64_t c = lo; c <= hi; ++c) cstart[c] = (int32_t)p;
}

__global__ void k_cell_maxcount(const int32_t* __restrict__ cstart, int64_t ncells, int* __restrict__ mx) {
  int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int v = 0;
  if (c < ncells) v = cstart[c + 1] - cstart[c];
  v = wave_max(v);
  __shared__ int red[kBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = max(m, red[w]);
    if (m > 0) atomicMax(mx, m);
  }
}

__global__ void k_permute_points(const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
                                 const int32_t* __restrict__ perm, int64_t n, float4* __restrict__ pts) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t i = perm[p];
  pts[p] = make_float4(x[i], y[i], z[i], __int_as_float(i));
}

__global__ void k_permute_var(const float* __restrict__ src, const int32_t* __restrict__ perm, int64_t n,
                              float* __restrict__ dst) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  dst[p] = src[perm[p]];
}

// inv[perm[p]] = p
__global__ void k_invert_perm(const int32_t* __restrict__ perm, int64_t n, int32_t* __restrict__ inv) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < n) inv[perm[p]] = (int32_t)p;
}

// Move per-point data of `width` float4 from the previous sorted order to the current one:
// point at old position p (original index old_perm[p]) goes to new position inv_new[old_perm[p]].
__global__ void k_reorder_f4(const float4* __restrict__ src, const int32_t* __restrict__ old_perm,
                             const int32_t* __restrict__ inv_new, int64_t n, int width, float4* __restrict__ dst) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int64_t q = inv_new[old_perm[p]];
  for (int k = 0; k < width; ++k) dst[q * width + k] = src[p * width + k];
}
__global__ void k_reorder_f2(const float2* __restrict__ src, const int32_t* __restrict__ old_perm,
                             const int32_t* __restrict__ inv_new, int64_t n, float2* __restrict__ dst) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  dst[inv_new[old_perm[p]]] = src[p];
}

__global__ void k_pack_vel(const float* __restrict__ vx, const float* __restrict__ vy, const float* __restrict__ vz,
                           const int32_t* __restrict__ perm, int64_t n, float4* __restrict__ vel) {
  int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  int32_t i = perm[p];
  vel[p] = make_float4(vx[i], vy[i], vz[i], 0.f);
}

// ------------------------------------------------------------------------------------
// K1: emissivity (emissionfromdatacube, fomo-CHIANTI.cpp:228-322; goft :64-121;
// linefwhm :220-225), FP64 arithmetic from float inputs, output float2(peak, fwhm)
// per sorted point. The table is staged in LDS when it fits. G(T, n) as the reference reads it
// (linear in log T, nearest row in log n), or with bilinear = 1 (an extension) linear in both,
// zero outside the table's cells as the reference's temperature rule has it (:100).
// ------------------------------------------------------------------------------------
constexpr int kLdsTable = 12288;  // floats of G staged in LDS (48 KB)

__global__ void k_emission_chianti(const float* __restrict__ rho, const float* __restrict__ T, int64_t n,
                                   const float* __restrict__ tlog, const float* __restrict__ nlog,
                                   const float* __restrict__ gtab, int nt, int nrho, double K, double normc,
                                   int imaging, int bilinear, float2* __restrict__ em) {
  __shared__ float sg[kLdsTable];
  const bool staged = (int64_t)nt * nrho <= kLdsTable;
  if (staged) {
    for (int i = threadIdx.x; i < nt * nrho; i += blockDim.x) sg[i] = gtab[i];
    __syncthreads();
  }
  const float* G = staged ? sg : gtab;
  const double t0 = tlog[0], dt = (double)tlog[1] - (double)tlog[0];
  const double n0 = nlog[0], dn = (double)nlog[1] - (double)nlog[0];
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const double r = rho[p], tt = T[p];
    const double logrho = r <= 0 ? 0.0 : log10(r);  // fomo-operations.cpp:73-82
    const double logT = tt <= 0 ? 0.0 : log10(tt);
    const double fl = floor((logT - t0) / dt);     // fomo-CHIANTI.cpp:95
    const double rr = round((logrho - n0) / dn);   // :97
    double res = 0.0;
    if (bilinear) {
      const double fn = floor((logrho - n0) / dn);
      if (!(fl < 0 || fn < 0 || fl + 1 > nt - 1 || fn + 1 > nrho - 1)) {
        const int ft = (int)fl, ct = ft + 1, fr = (int)fn, cr = fr + 1;
        const double x1 = tlog[ct], x2 = tlog[ft];
        const double ga = (logT - x1) * ((double)G[fr * nt + ft] - G[fr * nt + ct]) / (x2 - x1) + G[fr * nt + ct];
        const double gb = (logT - x1) * ((double)G[cr * nt + ft] - G[cr * nt + ct]) / (x2 - x1) + G[cr * nt + ct];
        const double r1 = nlog[cr], r2 = nlog[fr];
        res = (logrho - r1) * (ga - gb) / (r2 - r1) + gb;
      }
    } else if (!(fl < 0 || rr < 0 || fl + 1 > nt - 1 || rr > nrho - 1)) {  // :100
      const int ft = (int)fl, ct = ft + 1, ri = (int)rr;
      const int gi = ri * nt + ct;
      const double x1 = tlog[ct], y1 = G[gi], x2 = tlog[ft], y2 = G[gi - 1];
      res = (logT - x1) * (y2 - y1) / (x2 - x1) + y1;  // :109-113
    }
    const double w = imaging ? 1.0 : K * sqrt(tt);                   // :224, :284-285
    const double pk = normc * (pow(10.0, 2.0 * logrho) * res) / w;   // :292-293
    em[p] = make_float2((float)pk, (float)w);
  }
}

// thomsonfromdatacube + getThomsonGeomFactor (fomo-thomson.cpp:39-92), FP64.
__global__ void k_emission_thomson(const float4* __restrict__ pts, const float* __restrict__ dens, int64_t n,
                                   float2* __restrict__ em) {
  const double pi = 3.14159265358979323846;
  const double Coeff = 0.5 * pi * FOMO_RE * FOMO_RE * FOMO_LUM;
  const double U = FOMO_U_LIMB;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = pts[p];
    const double x = q.x, y = q.y, z = q.z;
    const double r = sqrt((x * x + y * y) + z * z);
    const double so = FOMO_RSUN / r;
    const double s2 = so * so, c2 = 1 - s2, co = sqrt(c2);
    const double lt = log((1. + so) / co);
    const double a = co * s2;
    const double b = (-1.0 / 8.0) * (1.0 - 3.0 * s2 - c2 * ((1.0 + 3.0 * s2) / so) * lt);
    const double c = (4.0 / 3.0) - co - (co * c2) / 3.0;
    const double d = (1.0 / 8.0) * (5.0 + s2 - c2 * ((5.0 - s2) / so) * lt);
    const double pol = Coeff * (1. / (1. - U / 3.)) * (a + U * (b - a));
    const double tot = Coeff * (1. / (1. - U / 3.)) * (2. * (c + U * (d - c)));
    const double nn = dens[p];
    em[p] = make_float2((float)(nn * tot), (float)(nn * pol));
  }
}

// ------------------------------------------------------------------------------------
// K2: per-view rotation, LOS velocity / sin^2(chi), extents
// NN: fomo-nearestneighbour.cpp:76-108. Thomson: fomo-thomson.cpp:136-217.
// ------------------------------------------------------------------------------------
__global__ void k_view_rotate_nn(const float4* __restrict__ pts, const float4* __restrict__ vel, int64_t n,
                                 double cb, double sb, double cl, double sl, float* __restrict__ vlos,
                                 double* __restrict__ partial) {
  double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  const double u0 = sb * cl, u1 = -sb * sl, u2 = cb;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = pts[p];
    const double g0 = q.x, g1 = q.y, g2 = q.z;
    const double xr = g0 * cb * cl - g1 * cb * sl - g2 * sb;  // :83
    const double yr = g0 * sl + g1 * cl;                      // :84
    const double zr = g0 * sb * cl - g1 * sb * sl + g2 * cb;  // :85
    v[0] = fmin(v[0], xr); v[1] = fmax(v[1], xr);
    v[2] = fmin(v[2], yr); v[3] = fmax(v[3], yr);
    v[4] = fmin(v[4], zr); v[5] = fmax(v[5], zr);
    if (vlos) {
      const float4 w = vel[p];
      vlos[p] = (float)(((0.0 + u0 * (double)w.x) + u1 * (double)w.y) + u2 * (double)w.z);  // :87-89
    }
  }
  block_minmax6(v, partial + 6 * blockIdx.x);
}

// Extents of up to kViewsPerPass NN views in one pass over the points (batches): the same
// rotation and min/max as k_view_rotate_nn per view (fomo-nearestneighbour.cpp:83-85, 95-100).
// partial[v * vstride + 6 * block + a], the layout k_view_finalize reads.
constexpr int kViewsPerPass = 8;
struct ViewTrigs {
  double cb[kViewsPerPass], sb[kViewsPerPass], cl[kViewsPerPass], sl[kViewsPerPass];
};
// SAMEB: every view of the launch has the same b, so the six products with cos b and sin b are
// formed once per point (the same values and the same left-to-right order,
// fomo-nearestneighbour.cpp:83-85).
template <bool SAMEB>
__global__ void __launch_bounds__(256) k_view_extents_multi(const float4* __restrict__ pts, int64_t n, ViewTrigs t,
                                                            int nv, double* __restrict__ partial, int64_t vstride) {
  double v[kViewsPerPass][6];
#pragma unroll
  for (int w = 0; w < kViewsPerPass; ++w)
#pragma unroll
    for (int a = 0; a < 6; ++a) v[w][a] = (a & 1) ? -INFINITY : INFINITY;
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = pts[p];
    const double g0 = q.x, g1 = q.y, g2 = q.z;
    const double gcb0 = g0 * t.cb[0], gcb1 = g1 * t.cb[0], gsb2 = g2 * t.sb[0];
    const double gsb0 = g0 * t.sb[0], gsb1 = g1 * t.sb[0], gcb2 = g2 * t.cb[0];
#pragma unroll
    for (int w = 0; w < kViewsPerPass; ++w) {
      if (w >= nv) break;
      const double cb = t.cb[w], sb = t.sb[w], cl = t.cl[w], sl = t.sl[w];
      const double xr = SAMEB ? gcb0 * cl - gcb1 * sl - gsb2 : g0 * cb * cl - g1 * cb * sl - g2 * sb;
      const double yr = g0 * sl + g1 * cl;
      const double zr = SAMEB ? gsb0 * cl - gsb1 * sl + gcb2 : g0 * sb * cl - g1 * sb * sl + g2 * cb;
      v[w][0] = fmin(v[w][0], xr); v[w][1] = fmax(v[w][1], xr);
      v[w][2] = fmin(v[w][2], yr); v[w][3] = fmax(v[w][3], yr);
      v[w][4] = fmin(v[w][4], zr); v[w][5] = fmax(v[w][5], zr);
    }
  }
  __shared__ double red[kBlock / 64][kViewsPerPass * 6];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < kViewsPerPass; ++w)
#pragma unroll
    for (int a = 0; a < 6; ++a) {
      const double r = (a & 1) ? wave_max(v[w][a]) : wave_min(v[w][a]);
      if (lane == 0) red[wid][w * 6 + a] = r;
    }
  __syncthreads();
  const int e = threadIdx.x;
  if (e < kViewsPerPass * 6 && e / 6 < nv) {
    const int w = e / 6, a = e % 6;
    double r = red[0][e];
    for (int k = 1; k < (int)(blockDim.x >> 6); ++k) r = (a & 1) ? fmax(r, red[k][e]) : fmin(r, red[k][e]);
    partial[(int64_t)w * vstride + 6 * blockIdx.x + a] = r;
  }
}

// With em (the Thomson emission: n tot, n pol per point), writes each point's pB term
// max(0, pol * sin^2 chi) in FP64 to term instead of sin^2 chi to sin2chi: the product of the two
// floats is exact in FP64, so k_los_thomson adds the same values it would form per sample, with
// one gather per sample instead of two.
__global__ void k_view_rotate_thomson(const float4* __restrict__ pts, int64_t n, double cbc, double sbc, double cl,
                                      double sl, double blend, const float2* __restrict__ em,
                                      float* __restrict__ sin2chi, double* __restrict__ term,
                                      double* __restrict__ partial) {
  double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  const double u0 = sbc * cl, u1 = -sbc * sl, u2 = cbc;  // :194 (clamped sin_b, cos_b)
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
    const float4 q = pts[p];
    const double g0 = q.x, g1 = q.y, g2 = q.z;
    const double fx = g0 * cbc * cl - g1 * cbc * sl - g2 * sbc;  // :179-181
    const double fy = g0 * sl + g1 * cl;
    const double fz = g0 * sbc * cl - g1 * sbc * sl + g2 * cbc;
    const double px = g0 * cl - g1 * sl, py = g0 * sl + g1 * cl, pz = g2;  // :184-186
    const double xa = fx * (1.0 - blend) + px * blend;  // :189-191
    const double ya = fy * (1.0 - blend) + py * blend;
    const double za = fz * (1.0 - blend) + pz * blend;
    const double ru = ((0.0 + u0 * xa) + u1 * ya) + u2 * za;
    const double rr = ((0.0 + xa * xa) + ya * ya) + za * za;
    const float s2 = (float)(1. - ru * ru / rr);  // :198
    if (em) {
      const double t = (double)em[p].y * (double)s2;  // :327-334
      term[p] = t < 0. ? 0. : t;
    } else {
      sin2chi[p] = s2;
    }
    v[0] = fmin(v[0], xa); v[1] = fmax(v[1], xa);
    v[2] = fmin(v[2], ya); v[3] = fmax(v[3], ya);
    v[4] = fmin(v[4], za); v[5] = fmax(v[5], za);
  }
  block_minmax6(v, partial + 6 * blockIdx.x);
}

// Reduce the per-block extents and derive the view geometry (one block).
// NN: maxdistance fomo-nearestneighbour.cpp:138, deltaz :142-143, scale :235-267.
// Thomson: maxdistance fomo-thomson.cpp:249-260, scale :350-359.
__device__ void view_finalize(const double* __restrict__ partial, int np, int mode, int nx, int ny, int nz, int dn,
                              double cb, double sb, double cl, double sl, double bf, ViewGeom* __restrict__ out,
                              int32_t* __restrict__ clear2) {
  __shared__ double ext[6];
  // the fused pass's direct-pass list (count, finished blocks) starts cleared
  if (clear2 && threadIdx.x < 2) clear2[threadIdx.x] = 0;
  double v[6] = {INFINITY, -INFINITY, INFINITY, -INFINITY, INFINITY, -INFINITY};
  for (int i = threadIdx.x; i < np; i += blockDim.x)
#pragma unroll
    for (int a = 0; a < 6; ++a) v[a] = (a & 1) ? fmax(v[a], partial[6 * i + a]) : fmin(v[a], partial[6 * i + a]);
  block_minmax6(v, ext);
  __syncthreads();
  if (threadIdx.x != 0) return;
  ViewGeom g;
  g.cb = cb; g.sb = sb; g.cl = cl; g.sl = sl;
  for (int a = 0; a < 6; ++a) g.ext[a] = ext[a];
  const double minx = ext[0], maxx = ext[1], miny = ext[2], maxy = ext[3], minz = ext[4], maxz = ext[5];
  const double dxp = (maxx - minx) / (nx - 1), dyp = (maxy - miny) / (ny - 1);
  if (mode == MODE_NN) {
    g.md = ((dxp < dyp) ? dyp : dxp) / .3;  // std::max(a,b) returns b when a<b, else a
  } else {
    // bf = 1/(1+exp(-10(|b|-0.1745))) is evaluated on the host with the C library exp
    const double e1 = 0.5 * sqrt(dxp * dxp + dyp * dyp);
    const double e2 = 0.1 * (maxx - minx) / (nx - 1);
    const double eq = (e1 < e2) ? e2 : e1;
    const double aw = ((dxp < dyp) ? dyp : dxp) / 0.005;
    g.md = eq * (1.0 - bf) + aw * bf;
  }
  double dz = maxz - minz;
  if (nz != 1) dz /= (nz - 1);
  g.deltaz = dz;
  double pl = (maxz - minz) / (nz - 1);
  if ((maxz == minz) || (nz == 1)) pl = 1.;
  g.pathlength = pl;
  double apix = 1.;
  if (dn && mode == MODE_NN) {
    const double pi = 3.14159265358979323846;
    const float dxf = (float)dxp, dyf = (float)dyp;
    const double c = pi / 180. / 3600.;
    apix = ((double)dxf / FOMO_MM_PER_ARCSEC) * ((double)dyf / FOMO_MM_PER_ARCSEC) * (c * c);
  }
  g.apix = apix;
  g.scale = pl * 1e8 * apix;
  *out = g;
}

__global__ void k_view_finalize(const double* __restrict__ partial, int np, int mode, int nx, int ny, int nz, int dn,
                                double cb, double sb, double cl, double sl, double bf, ViewGeom* __restrict__ out,
                                int32_t* __restrict__ clear2) {
  view_finalize(partial, np, mode, nx, ny, nz, dn, cb, sb, cl, sl, bf, out, clear2);
}

// The views of a batch in one launch: block v finalizes view v from its partials
// (partial + v * pstride) and its host trig (trig[5 v ..]: cos b, sin b, cos l, sin l, blend).
__global__ void k_view_finalize_views(const double* __restrict__ partial, int64_t pstride, int np, int mode, int nx,
                                      int ny, int nz, int dn, const double* __restrict__ trig,
                                      ViewGeom* __restrict__ out, int32_t* __restrict__ clear2) {
  const int v = (int)blockIdx.x;
  const double* t = trig + 5 * v;
  view_finalize(partial + v * pstride, np, mode, nx, ny, nz, dn, t[0], t[1], t[2], t[3], t[4], out + v,
                v == 0 ? clear2 : nullptr);
}

// ------------------------------------------------------------------------------------
// K3b: candidate lookup grid (once per dataset). For lookup cell B (expanded by eps):
//  U   = min over nearby points of maxdist(B, q)           (any target in B has a point within U)
//  C(B) = { q : mindist(B,q) <= U' and q not dominated over B by one of the 4 points with the
//           smallest maxdist },  q dominated by p over B  <=>  min_{t in B}(|t-q|^2-|t-p|^2) > margin.
// Margins are relative 1e-9 / 1e-10, far above FP64 rounding of the query distances, so the
// query's argmin over C(B) (by double distance, then lowest index) is the argmin over all points.
// Cells with no point within 3 coarse rings, or more than cap candidates, are marked -1 and their
// queries use the exact ring search.
// ------------------------------------------------------------------------------------
// MODE 0: count the candidates of every cell (cnt, -1 for cells left to the exact search);
// MODE 1: write them, given those counts and the overflow lists' offsets (off);
// MODE 2: one pass: write each record directly, cells with more than kRecSlots candidates taking
// a block of kLookupCap entries of the overflow pool cand (pool_used: entries requested; a cell
// whose block does not fit is left to the exact search).
template <int MODE>
__global__ void k_lookup_build(const float4* __restrict__ pts, const int32_t* __restrict__ cstart, GridInfo g,
                               LookupInfo lk, int64_t ncl, const uint8_t* __restrict__ dist,
                               int32_t* __restrict__ cnt, const int64_t* __restrict__ off, float4* __restrict__ cand,
                               float4* __restrict__ slot, unsigned long long* __restrict__ pool_used,
                               int64_t pool_cap) {
  constexpr bool FILL = MODE == 1;
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl) return;
  if (FILL && cnt[L] <= 0) return;
  float4* const rec = slot + (int64_t)kRecSlots * L;  // MODE 2
  const float4 unused = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
  const auto leave = [&]() {  // left to the exact search
    if (MODE != 1) cnt[L] = -1;
    if (MODE == 2) rec[0] = unused;
  };
  const int lx = (int)(L % lk.dims[0]);
  const int ly = (int)((L / lk.dims[0]) % lk.dims[1]);
  const int lz = (int)(L / ((int64_t)lk.dims[0] * lk.dims[1]));
  const int li[3] = {lx, ly, lz};
  double lo[3], hi[3], c[3], hw[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = lk.o[a] + (double)li[a] * lk.h[a] - lk.eps;
    hi[a] = lk.o[a] + (double)(li[a] + 1) * lk.h[a] + lk.eps;
    c[a] = 0.5 * (lo[a] + hi[a]);
    hw[a] = 0.5 * (hi[a] - lo[a]);
  }
  const int cc0 = cell_coord(g, 0, c[0]), cc1 = cell_coord(g, 1, c[1]), cc2 = cell_coord(g, 2, c[2]);
  const int dx = g.dims[0], dxy = g.dims[0] * g.dims[1];
  // no point within kLookupNearRings cells: leave the cell to its far group or the exact query;
  // rings closer than the distance to the nearest non-empty cell are empty and not visited
  const int rfirst = dist[(int64_t)cc2 * dxy + (int64_t)cc1 * dx + cc0];
  if (rfirst > kLookupNearRings) {
    leave();
    return;
  }
  // step 1: upper bound U and dominators
  double U2 = INFINITY;
  float4 dom[4];
  int dpos[4];
  double dkey[4];
  int nd = 0;
  int found = -1;
  for (int r = rfirst; r <= kLookupDistCap; ++r) {
    for (int z = max(cc2 - r, 0); z <= min(cc2 + r, g.dims[2] - 1); ++z)
      for (int y = max(cc1 - r, 0); y <= min(cc1 + r, g.dims[1] - 1); ++y) {
        // cells at Chebyshev distance exactly r: the whole x-range on an edge row, else x = cc0 -/+ r
        const bool edge = (abs(z - cc2) == r) || (abs(y - cc1) == r);
        const int xs = max(cc0 - r, 0), xe = min(cc0 + r, g.dims[0] - 1);
        const int nvis = edge ? (xe - xs + 1) : 2;
        for (int t = 0; t < nvis; ++t) {
          const int x = edge ? xs + t : (t == 0 ? cc0 - r : cc0 + r);
          if (x < 0 || x >= g.dims[0]) continue;
          const int cell = z * dxy + y * dx + x;
          for (int p = cstart[cell]; p < cstart[cell + 1]; ++p) {
            const float4 q = pts[p];
            const double m2 = box_maxdist2(lo, hi, q);
            U2 = fmin(U2, m2);
            // keep the 4 smallest m2
            int k = nd < 4 ? nd++ : 4;
            if (k == 4) {
              if (m2 >= dkey[3]) continue;
              k = 3;
            }
            while (k > 0 && dkey[k - 1] > m2) {
              dkey[k] = dkey[k - 1];
              dom[k] = dom[k - 1];
              dpos[k] = dpos[k - 1];
              --k;
            }
            dkey[k] = m2;
            dom[k] = q;
            dpos[k] = p;
          }
        }
      }
    if (found < 0 && nd > 0) found = r;
    if (found >= 0 && r >= found + 1) break;
  }
  if (nd == 0) {
    leave();
    return;
  }
  const double Up = sqrt(U2) * (1.0 + 1e-9) + 1e-9 * lk.h[0];
  const double Up2 = Up * Up;
  int r0[3], r1[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    r0[a] = cell_coord(g, a, lo[a] - Up);
    r1[a] = cell_coord(g, a, hi[a] + Up);
  }
  if ((int64_t)(r1[0] - r0[0] + 1) * (r1[1] - r0[1] + 1) * (r1[2] - r0[2] + 1) > 2197) {
    leave();
    return;
  }
  int n = 0;
  int64_t pbase = -1;  // MODE 2: this cell's block of the overflow pool
  // small lists go to the cell's own record, longer ones to the overflow list
  float4* const dst = FILL ? (cnt[L] <= kRecSlots ? slot + (int64_t)kRecSlots * L : cand + off[L]) : nullptr;
  for (int z = r0[2]; z <= r1[2]; ++z)
    for (int y = r0[1]; y <= r1[1]; ++y) {
      const int row = z * dxy + y * dx;
      for (int p = cstart[row + r0[0]]; p < cstart[row + r1[0] + 1]; ++p) {
        const float4 q = pts[p];
        if (box_mindist2(lo, hi, q) > Up2) continue;
        bool dm = false;
        for (int k = 0; k < nd && !dm; ++k)
          if (dpos[k] != p && dominated(c, hw, q, dom[k])) dm = true;
        if (dm) continue;
        // in a record with more than kSlots candidates, slot kSlots - 1 says that more follow
        // (w = -3 - position)
        if (FILL) {
          const bool more = n == kSlots - 1 && cnt[L] > kSlots && cnt[L] <= kRecSlots;
          dst[n] = make_float4(q.x, q.y, q.z, __int_as_float(more ? -3 - p : p));
        }
        if (MODE == 2) {
          const float4 cq = make_float4(q.x, q.y, q.z, __int_as_float(p));
          if (n < kRecSlots) {
            rec[n] = cq;
          } else {
            if (n == kRecSlots) {
              // the record overflows: a block of the pool (kLookupCap entries, whole 128-byte
              // lines), starting with the candidates written so far
              pbase = (int64_t)atomicAdd(pool_used, (unsigned long long)kLookupCap);
              if (pbase + kLookupCap > pool_cap) {
                leave();
                return;
              }
              for (int i = 0; i < kRecSlots; ++i) cand[pbase + i] = rec[i];
            }
            if (n < kLookupCap) cand[pbase + n] = cq;
          }
        }
        ++n;
        if (!FILL && n > kLookupCap) {
          leave();
          return;
        }
      }
    }
  if (MODE == 2) {
    if (n == 0) {
      rec[0] = unused;
    } else if (n <= kRecSlots) {
      for (int i = n; i < kRecSlots; ++i) rec[i] = unused;
      if (n > kSlots) {
        const float4 c3 = rec[kSlots - 1];
        rec[kSlots - 1] = make_float4(c3.x, c3.y, c3.z, __int_as_float(-3 - __float_as_int(c3.w)));
      }
    } else {
      rec[0] = make_float4(__int_as_float((int)pbase), __int_as_float(n), 0.f, __int_as_float(-2));
    }
  }
  if (!FILL) cnt[L] = n;
}

// dist[c] = Chebyshev distance (in cells) from cell c to the nearest non-empty cell, capped at
// R + 1: a separable transform, per axis dist(c) = min over d of max(|d|, dist_prev(c + d)).
__global__ void k_occupancy(const int32_t* __restrict__ cstart, int64_t ncells, int R, uint8_t* __restrict__ occ) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < ncells) occ[c] = cstart[c + 1] > cstart[c] ? 0 : (uint8_t)(R + 1);
}
__global__ void k_dist_axis(const uint8_t* __restrict__ in, uint8_t* __restrict__ out, GridInfo g, int axis, int R) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
  if (c >= n) return;
  const int64_t stride = axis == 0 ? 1 : (axis == 1 ? (int64_t)g.dims[0] : (int64_t)g.dims[0] * g.dims[1]);
  const int coord = axis == 0 ? (int)(c % g.dims[0])
                              : (axis == 1 ? (int)((c / g.dims[0]) % g.dims[1]) : (int)(c / ((int64_t)g.dims[0] * g.dims[1])));
  const int lo = max(coord - R, 0) - coord, hi = min(coord + R, g.dims[axis] - 1) - coord;
  int v = in[c];
  for (int d = lo; d <= hi; ++d) v = min(v, max(abs(d), (int)in[c + d * stride]));
  out[c] = (uint8_t)v;
}

// Headers of the lookup cell records: the no-candidate and overflow markers, and w = -1 in the
// slots a short list leaves unused (the filling pass of the two-pass build wrote the others).
__global__ void k_lookup_headers(const int32_t* __restrict__ cnt, const int64_t* __restrict__ off, int64_t ncl,
                                 float4* __restrict__ slot) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl) return;
  const int c = cnt[L];
  float4* r = slot + (int64_t)kRecSlots * L;
  const float4 unused = make_float4(0.f, 0.f, 0.f, __int_as_float(-1));
  if (c <= 0) {
    r[0] = unused;
  } else if (c > kRecSlots) {
    r[0] = make_float4(__int_as_float((int)off[L]), __int_as_float(c), 0.f, __int_as_float(-2));
  } else {
    for (int q = c; q < kRecSlots; ++q) r[q] = unused;
  }
}

// pos = the room of each cell's overflow list (cells with more than kRecSlots candidates): its
// length rounded up to whole 128-byte lines (8 candidates), so that a list of up to 8 candidates
// is one line fetch (an L2 miss fetches an aligned 128-byte line)
__global__ void k_overflow_counts(const int32_t* __restrict__ cnt, int64_t ncl, int32_t* __restrict__ pos) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl) return;
  pos[L] = cnt[L] > kRecSlots ? (cnt[L] + 7) & ~7 : 0;
}

__global__ void k_count_negative(const int32_t* __restrict__ cnt, int64_t ncl, int* __restrict__ out) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int v = (L < ncl && cnt[L] < 0) ? 1 : 0;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __shared__ int red[kBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
    if (t) atomicAdd(out, t);
  }
}

__global__ void k_clamp_counts(const int32_t* __restrict__ cnt, int64_t ncl, int32_t* __restrict__ pos) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl) return;
  pos[L] = max(cnt[L], 0);
}

// ------------------------------------------------------------------------------------
// Octree over the points in Morton order (exact fallback of the NN query), in depth-first
// preorder. Every node covers a contiguous range of the Morton-sorted points; an internal node
// splits its range by the highest 3-bit Morton group in which its codes differ (2..8 children),
// a range of at most 8 points is a leaf. Layout per node i (two float4):
//   nlo[i] = (box min, w), nhi[i] = (box max, first)
//   w >= 0: internal node, first child at i+1, w = index just past its subtree (skip pointer);
//   w <  0: leaf of -w points mpts[first .. first-w).
// The query walks the array without a stack: a node whose box cannot hold a better point
// strictly inside the search box is skipped with its subtree. Result: the nearest point strictly
// inside the box, ties to the lowest original index (the reference's R-tree query,
// fomo-nearestneighbour.cpp:179-183, fomo-thomson.cpp:310).
// ------------------------------------------------------------------------------------
__global__ void k_morton_keys(const float4* __restrict__ pts, int64_t n, GridInfo g, uint32_t* __restrict__ key,
                              int32_t* __restrict__ val) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const float4 q = pts[p];
  key[p] = fomo::morton30(q.x, q.y, q.z, g.bmin, g.bmax);
  val[p] = (int32_t)p;
}

// Morton-ordered points (x, y, z, sorted position)
__global__ void k_morton_points(const float4* __restrict__ pts, const int32_t* __restrict__ order, int64_t n,
                                float4* __restrict__ mpts) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t p = order[i];
  const float4 q = pts[p];
  mpts[i] = make_float4(q.x, q.y, q.z, __int_as_float(p));
}

// Boxes of the nodes listed in ids (all of one depth, deepest first): leaves from their points,
// internal nodes from their children (already done).
__global__ void k_oct_boxes(const int32_t* __restrict__ ids, int64_t cnt, const float4* __restrict__ mpts,
                            float4* __restrict__ nlo, float4* __restrict__ nhi) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int32_t i = ids[t];
  const float4 a0 = nlo[i], b0 = nhi[i];
  const int w = __float_as_int(a0.w);
  float3 l = make_float3(INFINITY, INFINITY, INFINITY), h = make_float3(-INFINITY, -INFINITY, -INFINITY);
  if (w < 0) {
    const int first = __float_as_int(b0.w);
    for (int q = first; q < first - w; ++q) {
      const float4 v = mpts[q];
      l.x = fminf(l.x, v.x); l.y = fminf(l.y, v.y); l.z = fminf(l.z, v.z);
      h.x = fmaxf(h.x, v.x); h.y = fmaxf(h.y, v.y); h.z = fmaxf(h.z, v.z);
    }
  } else {
    for (int c = i + 1; c < w;) {
      const float4 cl = nlo[c], ch = nhi[c];
      l.x = fminf(l.x, cl.x); l.y = fminf(l.y, cl.y); l.z = fminf(l.z, cl.z);
      h.x = fmaxf(h.x, ch.x); h.y = fmaxf(h.y, ch.y); h.z = fmaxf(h.z, ch.z);
      const int cw = __float_as_int(cl.w);
      c = cw < 0 ? c + 1 : cw;
    }
  }
  nlo[i] = make_float4(l.x, l.y, l.z, a0.w);
  nhi[i] = make_float4(h.x, h.y, h.z, b0.w);
}

// Wide records of the internal nodes (BvhInfo::wide, fomo::wide_slot), one thread per slot: kid =
// the child's node index (-1: empty slot), kref = the child's record when it is internal (-1: leaf)
// (fomo::wide_children).
__global__ void k_wide_records(const int32_t* __restrict__ kid, const int32_t* __restrict__ kref, int64_t nslots,
                               const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                               float4* __restrict__ wide) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nslots) return;
  const int c = kid[t];
  const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 a, b;
  fomo::wide_slot(c, kref[t], c >= 0 ? nlo[c] : zero, c >= 0 ? nhi[c] : zero, a, b);
  wide[2 * t] = a;
  wide[2 * t + 1] = b;
}

// Per-sample outputs of the NN pass for the point at sorted position r (-1: none):
// svlos[s] (LOS velocity, fomo-nearestneighbour.cpp:87-89) and sem[l*S + s] = (peak_l, fwhm_l),
// zero when no point is inside the search box (:190-193).
template <int RS4>
__device__ __forceinline__ void write_sample(int r, const float4* __restrict__ rec, const RecInfo& ri, int nlines,
                                             size_t S, size_t sidx, double cb, double sb, double cl, double sl,
                                             float* __restrict__ svlos, float2* __restrict__ sem) {
  float f[4 * RS4];
#pragma unroll
  for (int q = 0; q < 4 * RS4; ++q) f[q] = 0.f;
  if (r >= 0) {
    const float4* rp = rec + (size_t)r * RS4;
#pragma unroll
    for (int q = 0; q < RS4; ++q) {
      const float4 a = rp[q];
      f[4 * q] = a.x;
      f[4 * q + 1] = a.y;
      f[4 * q + 2] = a.z;
      f[4 * q + 3] = a.w;
    }
  }
  // unit vector (sin b cos l, -sin b sin l, cos b)
  svlos[sidx] = r >= 0 ? (float)(((0.0 + (sb * cl) * (double)f[0]) + (-sb * sl) * (double)f[1]) + cb * (double)f[2])
                       : 0.f;
  const double st = r >= 0 ? rec_st(ri.K[0], f[3]) : 0.0;
  // a record of RS4 float4 holds 4*RS4-4 lines after {v, T}
#pragma unroll
  for (int l = 0; l < 4 * RS4 - 4; ++l)
    if (l < nlines) {
      const float fw = r < 0 ? 0.f : rec_fw(ri.K[l], st);
      sem[(size_t)l * S + sidx] = make_float2(f[4 + l], fw);
    }
}

// One thread per LOS sample. Block = one ray tile at one k (blockIdx.y = k).
// RS4 > 0: per-point records of RS4 float4 {vx,vy,vz, peak_0,fwhm_0, ..., peak_{L-1},fwhm_{L-1}}
// are gathered once for the chosen point, and the kernel writes coalesced per-sample arrays:
// svlos[s] (LOS velocity, fomo-nearestneighbour.cpp:87-89) and sem[l*S + s] = (peak_l, fwhm_l),
// zero when no point is inside the search box (:190-193). idx (optional) gets the sorted position.
template <int RS4>
__device__ __forceinline__ void nn_sample(const float4* __restrict__ pts, const GridInfo& g,
                                          const float4* __restrict__ lslot, const float4* __restrict__ lcand,
                                          const LookupInfo& lk, const ViewGeom& vg, int nx, int ny, int nz, int i,
                                          int j, int k, double xx, double yy, int32_t* __restrict__ idx,
                                          const float4* __restrict__ rec,
                                          const RecInfo& ri, int nlines, float* __restrict__ svlos,
                                          float2* __restrict__ sem, uint8_t* __restrict__ rayflag) {
  const double minz = vg.ext[4];
  const double zz = double(k) * vg.deltaz + minz;                 // :162
  const double cb = vg.cb, sb = vg.sb, cl = vg.cl, sl = vg.sl;
  const double p0 = xx * cb * cl + yy * sl + zz * sb * cl;        // :165
  const double p1 = -xx * cb * sl + yy * cl - zz * sb * sl;
  const double p2 = -xx * sb + zz * cb;
  int r = -2;
  const double md0 = vg.md;
  // a box that misses the data's bounding box holds no point: the lookup is skipped (r = -1 below)
  const bool may_hold = (float)(p0 + md0) > g.bmin[0] && (float)(p0 - md0) < g.bmax[0] &&
                        (float)(p1 + md0) > g.bmin[1] && (float)(p1 - md0) < g.bmax[1] &&
                        (float)(p2 + md0) > g.bmin[2] && (float)(p2 - md0) < g.bmax[2];
  if ((lk.enabled || lk.rect || lk.mrect) && may_hold) {
    const double md = vg.md;
    r = nn_lookup(pts, lslot, lcand, lk, (float)p0, (float)p1, (float)p2, (float)(p0 - md), (float)(p1 - md),
                  (float)(p2 - md), (float)(p0 + md), (float)(p1 + md), (float)(p2 + md));
  }
  if (r == -2) {
    const double md = vg.md;
    const float lo0 = (float)(p0 - md), lo1 = (float)(p1 - md), lo2 = (float)(p2 - md);
    const float hi0 = (float)(p0 + md), hi1 = (float)(p1 + md), hi2 = (float)(p2 + md);
    if (box_misses_data(g, lo0, lo1, lo2, hi0, hi1, hi2))
      r = -1;  // no data point can be inside the box
    else
      rayflag[(size_t)i * nx + j] = 1;  // left as -2 for k_nn_far (exact octree query)
  }
  const size_t S = (size_t)nx * ny * nz;
  const size_t sidx = (size_t)k * ((size_t)nx * ny) + (size_t)i * nx + j;
  idx[sidx] = r;
  if constexpr (RS4 > 0) {
    if (r != -2) write_sample<RS4>(r, rec, ri, nlines, S, sidx, cb, sb, cl, sl, svlos, sem);
  }
}


// Ray tile of a k_nn block: 256 rays of one LOS plane, 2^tlog wide and 256 / 2^tlog
// high (tlog 3..6: 8x32 .. 64x4), so each wave of 64 lanes covers 2^tlog x 64 / 2^tlog rays.
__device__ __forceinline__ void ray_of_thread(int nx, int tlog, int* i, int* j) {
  const int tw = 1 << tlog, th = 256 >> tlog;
  const int tilesx = (nx + tw - 1) >> tlog;
  *j = (int)(blockIdx.x % tilesx) * tw + (int)(threadIdx.x & (tw - 1));
  *i = (int)(blockIdx.x / tilesx) * th + (int)(threadIdx.x >> tlog);
}

// One thread per LOS sample: block = one ray tile of one LOS plane (blockIdx.y = k).
template <int RS4>
__global__ void __launch_bounds__(256) k_nn(const float4* __restrict__ pts, GridInfo g,
                                            const float4* __restrict__ lslot, const float4* __restrict__ lcand,
                                            LookupInfo lk, const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                            int32_t* __restrict__ idx, const float4* __restrict__ rec, RecInfo ri,
                                            int nlines, float* __restrict__ svlos, float2* __restrict__ sem,
                                            uint8_t* __restrict__ rayflag, int tlog, unsigned* __restrict__ dcount) {
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) *dcount = 0;  // k_nn_far's deep queue
  // the tile's lattice x of each column and y of each row, once per block
  // (fomo-nearestneighbour.cpp:152-153)
  __shared__ double tx[64], ty[32];
  const ViewGeom& vg = *vgp;
  const int tw = 1 << tlog, th = 256 >> tlog;
  const int tilesx = (nx + tw - 1) >> tlog;
  const int j0 = (int)(blockIdx.x % tilesx) * tw, i0 = (int)(blockIdx.x / tilesx) * th;
  if ((int)threadIdx.x < tw) {
    const int j = j0 + (int)threadIdx.x;
    tx[threadIdx.x] = double(j) / (nx - 1) * (vg.ext[1] - vg.ext[0]) + vg.ext[0];
  } else if ((int)threadIdx.x < tw + th) {
    const int i = i0 + (int)threadIdx.x - tw;
    ty[threadIdx.x - tw] = double(i) / (ny - 1) * (vg.ext[3] - vg.ext[2]) + vg.ext[2];
  }
  __syncthreads();
  int i, j;
  ray_of_thread(nx, tlog, &i, &j);
  if (j >= nx || i >= ny) return;
  nn_sample<RS4>(pts, g, lslot, lcand, lk, vg, nx, ny, nz, i, j, (int)blockIdx.y, tx[j - j0], ty[i - i0], idx, rec,
                 ri, nlines, svlos, sem, rayflag);
}

// Search box of LOS sample (i, j, k) of a view (fomo-nearestneighbour.cpp:152-183).
__device__ __forceinline__ SearchBox sample_box(const ViewGeom& vg, int nx, int ny, int i, int j, int k) {
  const double minx = vg.ext[0], maxx = vg.ext[1], miny = vg.ext[2], maxy = vg.ext[3], minz = vg.ext[4];
  const double xx = double(j) / (nx - 1) * (maxx - minx) + minx;  // fomo-nearestneighbour.cpp:152
  const double yy = double(i) / (ny - 1) * (maxy - miny) + miny;  // :153
  const double zz = double(k) * vg.deltaz + minz;                 // :162
  const double cb = vg.cb, sb = vg.sb, cl = vg.cl, sl = vg.sl, md = vg.md;
  const double p0 = xx * cb * cl + yy * sl + zz * sb * cl;        // :165
  const double p1 = -xx * cb * sl + yy * cl - zz * sb * sl;
  const double p2 = -xx * sb + zz * cb;
  SearchBox box;
  box.t0 = (float)p0, box.t1 = (float)p1, box.t2 = (float)p2;
  box.lo0 = (float)(p0 - md), box.lo1 = (float)(p1 - md), box.lo2 = (float)(p2 - md);
  box.hi0 = (float)(p0 + md), box.hi1 = (float)(p1 + md), box.hi2 = (float)(p2 + md);
  return box;
}

__device__ __forceinline__ double box_dist2(const SearchBox& box, const float4& q) {
  const double d0 = (double)q.x - (double)box.t0, d1 = (double)q.y - (double)box.t1, d2 = (double)q.z - (double)box.t2;
  return (d0 * d0 + d1 * d1) + d2 * d2;
}

// Long exact queries. A k_nn_far wave whose walk passes the far budget of node steps (targets
// deep in a hole of the data, nearly equidistant from a large part of it, take up to ~7e4) hands
// its samples to a queue instead. The tree's preorder node array is cut into kDeepSegs equal
// ranges, each with the bounding box of its leaves' points (Dataset::segbox). k_nn_deep puts
// kDeepWaves waves on every queued wave: they take its ranges one at a time from a shared counter,
// skip a range whose box cannot improve any lane's running best (started from the queued best
// so far) and walk the others, each keeping its own running best across its ranges.
// k_nn_deep_merge then takes every sample's nearest point over the waves (distance, then lowest
// original index: the order of the query, so the result is the query's). With a full queue the
// wave finishes its walk in place.
constexpr int kFarBudget = 2048;  // default; fomo_dataset_set_far_budget
constexpr int kDeepSegs = 1024;
constexpr int kDeepWaves = 32;    // waves per queued wave: kDeepWaves / 4 blocks of 256
constexpr int kDeepBlocks = 2048;
struct DeepQueue {
  unsigned* count;      // queued waves (may pass cap; the queue holds the first cap)
  int64_t* sidx;        // cap x 64 samples (-1: lane without one)
  int32_t* seed;        // cap x 64: the running best when the wave stopped (-1: none yet)
  unsigned* next;       // cap: next range of each queued wave (zeroed when it is queued)
  int32_t* part;        // cap x kDeepWaves x 64: each deep wave's nearest point
  const float4* segbox; // kDeepSegs range boxes: lo[], then hi[]
  int cap;
  int budget;           // node steps before a wave is queued (0: never)
};
// Queue capacity for a view of S samples.
int nn_deep_cap(int64_t S) { return (int)std::min<int64_t>(2048, std::max<int64_t>(16, S >> 18)); }
int64_t nn_deep_off(int64_t R) { return (R + 255) & ~(int64_t)255; }
int64_t round256(int64_t b) { return (b + 255) & ~(int64_t)255; }
// Bytes of the NN passes' scratch of a view: R ray flags, then the deep queue.
int64_t nn_scratch_bytes(int64_t R, int64_t S) {
  const int64_t cap = nn_deep_cap(S);
  return nn_deep_off(R) + 256 + round256(cap * 64 * (int64_t)sizeof(int64_t)) +
         round256(cap * 64 * (int64_t)sizeof(int32_t)) + round256(cap * (int64_t)sizeof(unsigned)) +
         cap * kDeepWaves * 64 * (int64_t)sizeof(int32_t);
}
DeepQueue deep_queue(uint8_t* scratch, int64_t R, int64_t S, int budget, const float4* segbox) {
  DeepQueue q;
  char* b = reinterpret_cast<char*>(scratch) + nn_deep_off(R);
  const int64_t cap = nn_deep_cap(S);
  q.cap = (int)cap;
  q.budget = segbox ? budget : 0;
  q.segbox = segbox;
  q.count = reinterpret_cast<unsigned*>(b);
  b += 256;
  q.sidx = reinterpret_cast<int64_t*>(b);
  b += round256(cap * 64 * (int64_t)sizeof(int64_t));
  q.seed = reinterpret_cast<int32_t*>(b);
  b += round256(cap * 64 * (int64_t)sizeof(int32_t));
  q.next = reinterpret_cast<unsigned*>(b);
  b += round256(cap * (int64_t)sizeof(unsigned));
  q.part = reinterpret_cast<int32_t*>(b);
  return q;
}

// Preorder node range g of kDeepSegs over an n-node tree.
__device__ __forceinline__ void deep_range(int n, unsigned g, int* e0, int* e1) {
  *e0 = (int)((int64_t)n * g / kDeepSegs);
  *e1 = (int)((int64_t)n * (g + 1) / kDeepSegs);
}

// Bounding box of the points of the leaves in each node range (empty range: lo = +inf, hi = -inf,
// which overlaps no search box). One block per range.
__global__ void __launch_bounds__(256) k_seg_boxes(const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                   int n, float4* __restrict__ segbox) {
  __shared__ float red[6][256];
  int e0, e1;
  deep_range(n, blockIdx.x, &e0, &e1);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int e = e0 + (int)threadIdx.x; e < e1; e += blockDim.x) {
    const float4 l = nlo[e];
    if (__float_as_int(l.w) >= 0) continue;  // internal node
    const float4 h = nhi[e];
    lo[0] = fminf(lo[0], l.x), lo[1] = fminf(lo[1], l.y), lo[2] = fminf(lo[2], l.z);
    hi[0] = fmaxf(hi[0], h.x), hi[1] = fmaxf(hi[1], h.y), hi[2] = fmaxf(hi[2], h.z);
  }
  for (int a = 0; a < 3; ++a) red[a][threadIdx.x] = lo[a], red[3 + a][threadIdx.x] = hi[a];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fminf(red[a][threadIdx.x], red[a][threadIdx.x + st]);
        red[3 + a][threadIdx.x] = fmaxf(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    segbox[blockIdx.x] = make_float4(red[0][0], red[1][0], red[2][0], 0.f);
    segbox[kDeepSegs + blockIdx.x] = make_float4(red[3][0], red[4][0], red[5][0], 0.f);
  }
}

// Second NN pass over the samples k_nn left unresolved (-2): exact octree query. Rays without
// such samples exit at once (rayflag, cleared afterwards by k_clear_u8). The starting bound is the
// previous sample's point on the ray when it is known (k_nn's, or this pass's when that sample's
// wave has finished) and lies strictly inside this sample's box; else the walk's own first descent.
template <int RS4>
__global__ void __launch_bounds__(256) k_nn_far(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                                                const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                BvhInfo bv, const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                                int32_t* __restrict__ idx, const float4* __restrict__ rec, RecInfo ri,
                                                int nlines, float* __restrict__ svlos, float2* __restrict__ sem,
                                                const uint8_t* __restrict__ rayflag, DeepQueue dq) {
  // a block takes an 8 x 8 ray tile through 4 consecutive LOS planes, a wave a 4 x 4 quarter of it
  // through the 4 planes: the most compact group of 64 samples, whose walks share the most nodes
  const int tilesx = (nx + 7) >> 3, lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
  const int j = (int)(blockIdx.x % tilesx) * 8 + (wv & 1) * 4 + (lane & 3);
  const int i = (int)(blockIdx.x / tilesx) * 8 + (wv >> 1) * 4 + ((lane >> 2) & 3);
  const int kk = (int)blockIdx.y * 4 + (lane >> 4);
  const size_t R = (size_t)nx * ny;
  const bool inside = j < nx && i < ny && kk < nz;
  const int k = inside ? kk : 0;
  const size_t ray = inside ? (size_t)i * nx + j : 0;
  const size_t sidx = (size_t)k * R + ray;
  // lanes without an unresolved sample still take part in the wave's walk (act = false)
  const bool act = inside && rayflag[ray] && idx[sidx] == -2;
  __shared__ float4 stk[4][2 * kWideStack];  // wide_walk_wave's stack of each wave
  if (__ballot(act) == 0) return;
  const ViewGeom vg = *vgp;
  const SearchBox box = sample_box(vg, nx, ny, inside ? i : 0, inside ? j : 0, k);
  double best = INFINITY;
  int bpos = -1, borig = 0x7fffffff;
  if (act && k > 0) {
    const int prev = idx[sidx - R];  // k_nn's result, or this pass's when done (any value >= 0 is a point)
    if (prev >= 0) {
      const float4 q = pts[prev];
      if (box.holds(q)) {
        best = box_dist2(box, q);
        bpos = prev;
        borig = __float_as_int(q.w);
      }
    }
  }
  if (bv.nodes > 0) {
    const int n = (int)bv.nodes;
    bool over;
    if (bv.wide) {
      // expansions and leaf scans: one per 8 steps of the preorder walk's budget
      const int wb = dq.budget > 0 ? max(1, dq.budget / 8) : 0;
      over = wide_walk_wave(pts, mpts, bv.wide, box, act, best, bpos, borig, wb, stk[threadIdx.x >> 6]);
    } else {
      if (act && bpos < 0) fomo::oct_descend(pts, mpts, nlo, nhi, box, best, bpos, borig);
      over = bvh_walk_wave(pts, mpts, nlo, nhi, n, 0, n, box, act, best, bpos, borig, dq.budget);
    }
    if (over) {
      const int lane = (int)(threadIdx.x & 63);
      unsigned job = 0;
      if (lane == 0) {
        job = atomicAdd(dq.count, 1u);
        if (job < (unsigned)dq.cap) dq.next[job] = 0;
      }
      job = (unsigned)__shfl((int)job, 0, 64);
      if (job < (unsigned)dq.cap) {
        dq.sidx[(size_t)job * 64 + lane] = act ? (int64_t)sidx : -1;
        dq.seed[(size_t)job * 64 + lane] = act ? bpos : -1;
        return;  // idx stays -2 until k_nn_deep_merge
      }
      bvh_walk_wave(pts, mpts, nlo, nhi, n, 0, n, box, act, best, bpos, borig, 0);
    }
  }
  if (!act) return;
  idx[sidx] = bpos;
  if constexpr (RS4 > 0)
    write_sample<RS4>(bpos, rec, ri, nlines, R * nz, sidx, vg.cb, vg.sb, vg.cl, vg.sl, svlos, sem);
}

// kDeepWaves waves per queued wave (kDeepWaves / 4 blocks, each block one (queued wave, block)
// pair at a time); they share the queued wave's ranges through its counter.
__global__ void __launch_bounds__(256) k_nn_deep(const float4* __restrict__ pts, const float4* __restrict__ mpts,
                                                 const float4* __restrict__ nlo, const float4* __restrict__ nhi,
                                                 BvhInfo bv, const ViewGeom* __restrict__ vgp, int nx, int ny,
                                                 DeepQueue dq) {
  constexpr unsigned kPer = kDeepWaves / 4;
  const unsigned jobs = min(*dq.count, (unsigned)dq.cap);
  const unsigned pairs = jobs * kPer;
  if (blockIdx.x >= pairs) return;
  const int lane = (int)(threadIdx.x & 63);
  const ViewGeom vg = *vgp;
  const int64_t R = (int64_t)nx * ny;
  const int n = (int)bv.nodes;
  for (unsigned p = blockIdx.x; p < pairs; p += gridDim.x) {
    const unsigned job = p / kPer, slot = (p % kPer) * 4 + (threadIdx.x >> 6);
    const int64_t sidx = dq.sidx[(size_t)job * 64 + lane];
    const bool act = sidx >= 0;
    const int64_t sa = act ? sidx : 0, ray = sa % R;
    const SearchBox box = sample_box(vg, nx, ny, (int)(ray / nx), (int)(ray % nx), (int)(sa / R));
    double best = INFINITY;
    int bpos = act ? dq.seed[(size_t)job * 64 + lane] : -1, borig = 0x7fffffff;
    if (bpos >= 0) {
      const float4 q = pts[bpos];
      best = box_dist2(box, q);
      borig = __float_as_int(q.w);
    }
    for (;;) {
      unsigned g = 0;
      if (lane == 0) g = atomicAdd(&dq.next[job], 1u);
      g = (unsigned)__shfl((int)g, 0, 64);
      if (g >= (unsigned)kDeepSegs) break;
      const float4 l = dq.segbox[g], h = dq.segbox[kDeepSegs + g];
      if (__ballot(act && box_may_improve(l, h, box, bvh_bound(best))) == 0) continue;
      int e0, e1;
      deep_range(n, g, &e0, &e1);
      bvh_walk_wave(pts, mpts, nlo, nhi, n, e0, e1, box, act, best, bpos, borig, 0);
    }
    dq.part[((size_t)job * kDeepWaves + slot) * 64 + lane] = bpos;
  }
}

// Every queued sample's nearest point over the deep waves; written like k_nn_far writes its samples.
template <int RS4>
__global__ void __launch_bounds__(256) k_nn_deep_merge(const float4* __restrict__ pts,
                                                       const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                                       int32_t* __restrict__ idx, const float4* __restrict__ rec,
                                                       RecInfo ri, int nlines, float* __restrict__ svlos,
                                                       float2* __restrict__ sem, DeepQueue dq) {
  const unsigned jobs = min(*dq.count, (unsigned)dq.cap);
  const ViewGeom vg = *vgp;
  const int64_t R = (int64_t)nx * ny;
  for (size_t u = blockIdx.x * (size_t)blockDim.x + threadIdx.x; u < (size_t)jobs * 64;
       u += (size_t)gridDim.x * blockDim.x) {
    const int64_t sidx = dq.sidx[u];
    if (sidx < 0) continue;
    const size_t job = u / 64, lane = u % 64;
    const int64_t ray = sidx % R;
    const SearchBox box = sample_box(vg, nx, ny, (int)(ray / nx), (int)(ray % nx), (int)(sidx / R));
    int bpos = -1, borig = 0x7fffffff;
    double best = INFINITY;
    for (int g = 0; g < kDeepWaves; ++g) {
      const int p = dq.part[(job * kDeepWaves + g) * 64 + lane];
      if (p < 0) continue;
      const float4 q = pts[p];
      const double d = box_dist2(box, q);
      const int o = __float_as_int(q.w);
      if (d < best || (d == best && o < borig)) {
        best = d;
        bpos = p;
        borig = o;
      }
    }
    idx[sidx] = bpos;
    if constexpr (RS4 > 0)
      write_sample<RS4>(bpos, rec, ri, nlines, (size_t)R * nz, (size_t)sidx, vg.cb, vg.sb, vg.cl, vg.sl, svlos, sem);
  }
}

__global__ void k_clear_u8(uint8_t* __restrict__ a, int64_t n) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

// The 10 meta doubles of a view (extents, half-width, path length, scale, apix) into device memory.
__global__ void k_geom_meta(const ViewGeom* __restrict__ g, double* __restrict__ meta) {
  const int t = threadIdx.x;
  if (t < 6) meta[t] = g->ext[t];
  else if (t == 6) meta[6] = g->md;
  else if (t == 7) meta[7] = g->pathlength;
  else if (t == 8) meta[8] = g->scale;
  else if (t == 9) meta[9] = g->apix;
}

// ------------------------------------------------------------------------------------
// Far groups of the lookup grid. A lookup cell left to the exact query (more than
// kLookupNearRings cells from any point, or too many candidates of its own) takes the candidate
// list of its group, kFarGroup^3 cell-list cells, when that list is short. The list is built like
// a cell's, over the group's box B: U = the smallest maxdist(B, q) among the points nearest to B's
// centre; every point with mindist(B, q) <= U that none of those 4 points dominates and no other
// such point dominates (a dominated point is never the nearest to a target in B, whichever point
// dominates it). Far from the data the Voronoi cells of the surface points fan out, so a group of
// 4^3 cell-list cells keeps a dozen or so candidates, and the queries of its samples need no
// octree walk. Inside a concave hole they converge instead: near its middle the lists grow past
// kFarCap. Points are enumerated with the octree, a wave of 64 neighbouring groups walking it
// together.
// ------------------------------------------------------------------------------------
constexpr int kFarGroup = 4;      // cell-list cells per group edge
constexpr int kFarCap = 256;      // longest group list
constexpr int kFarVisit = 65536;  // points within U a group may examine (more: exact query)
constexpr int kFarBlock = 64;     // groups per block: 4 x 4 x 4 neighbours, one wave

struct FarGrid {
  int gdim[3];   // groups per axis
  int gcell[3];  // lookup cells per group edge
  int bdim[3];   // blocks of 4 x 4 x 4 groups per axis
  int ioff[3];   // lookup cell index of group 0's low corner (negative for outer groups)
  int clip;      // boxes end at the grid's last cell (groups of the grid) or not (outer groups)
};

__device__ __forceinline__ double far_margin2(double u2, const LookupInfo& lk) {
  const double u = sqrt(u2) * (1.0 + 1e-9) + 1e-9 * lk.h[0];
  return u * u;
}

// Wave-cooperative walk over the nodes whose box comes within sqrt(r2) of the lane's box [lo, hi];
// admit(l, h) may reject a node the distance test lets in; visit(first, count) runs on the leaves
// the lane's own tests admit. A lane may lower r2 inside visit (a negative r2 ends its part of the
// walk). All lanes of the wave must call it.
template <class Admit, class Visit>
__device__ __forceinline__ void far_walk(const float4* __restrict__ nlo, const float4* __restrict__ nhi, int n,
                                         const double (&lo)[3], const double (&hi)[3], const double& r2,
                                         Admit&& admit, Visit&& visit) {
  for (int e = 0; e < n;) {
    e = __builtin_amdgcn_readfirstlane(e);
    const float4 l = nlo[e];
    const float4 h = nhi[e];
    const int w = __float_as_int(l.w);
    const double d0 = fmax(fmax((double)l.x - hi[0], lo[0] - (double)h.x), 0.0);
    const double d1 = fmax(fmax((double)l.y - hi[1], lo[1] - (double)h.y), 0.0);
    const double d2 = fmax(fmax((double)l.z - hi[2], lo[2] - (double)h.z), 0.0);
    // a point of the node is no nearer to the box than the node's box is (margin: rounding)
    const bool in = (d0 * d0 + d1 * d1) + d2 * d2 <= r2 * (1.0 + 1e-12) && admit(l, h);
    if (w < 0) {
      if (in) visit(__float_as_int(h.w), -w);
      ++e;
    } else {
      e = __ballot(in) ? e + 1 : w;
    }
  }
}

// gflag[g] = 1 for the groups holding a lookup cell left to the exact query (cnt = -1), gown[g] = 1
// for those holding a cell with a list of its own
__global__ void k_far_flags(const int32_t* __restrict__ cnt, int64_t ncl, LookupInfo lk, FarGrid fg,
                            uint8_t* __restrict__ gflag, uint8_t* __restrict__ gown) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl) return;
  const bool left = cnt[L] == -1;
  const int lx = (int)(L % lk.dims[0]);
  const int ly = (int)((L / lk.dims[0]) % lk.dims[1]);
  const int lz = (int)(L / ((int64_t)lk.dims[0] * lk.dims[1]));
  const int64_t g = ((int64_t)(lz / fg.gcell[2]) * fg.gdim[1] + ly / fg.gcell[1]) * fg.gdim[0] + lx / fg.gcell[0];
  if (left) gflag[g] = 1;
  else gown[g] = 1;
}

// Headers of the groups whose cells all take the group's list; nfull counts them.
__global__ void k_group_headers(const int2* __restrict__ ginfo, const uint8_t* __restrict__ gown, int64_t ng,
                                int64_t base, int2* __restrict__ ghdr, unsigned long long* __restrict__ nfull) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int2 gi = ginfo[g];
  const bool full = gi.y > 0 && !gown[g];
  ghdr[g] = full ? make_int2((int)(base + gi.x), gi.y) : make_int2(0, 0);
  if (full) atomicAdd(nfull, 1ull);
}

// Candidate lists of the flagged groups in blocks [b0, b0 + gridDim.x): ginfo[g] = (start in
// fpool, count), count 0 when the list is too long or fpool is full. cbuf: kFarCap entries per
// lane of the launch.
__global__ void __launch_bounds__(kFarBlock) k_far_groups(const float4* __restrict__ pts,
                                                          const float4* __restrict__ mpts,
                                                          const float4* __restrict__ nlo,
                                                          const float4* __restrict__ nhi, int nodes, LookupInfo lk,
                                                          FarGrid fg, int64_t b0, const uint8_t* __restrict__ gflag,
                                                          float4* __restrict__ cbuf, float4* __restrict__ fpool,
                                                          unsigned long long* __restrict__ fused, int64_t fcap,
                                                          int2* __restrict__ ginfo) {
  const int lane = threadIdx.x;
  const int64_t b = b0 + blockIdx.x;
  const int gi[3] = {(int)(b % fg.bdim[0]) * 4 + (lane & 3), (int)((b / fg.bdim[0]) % fg.bdim[1]) * 4 + ((lane >> 2) & 3),
                     (int)(b / ((int64_t)fg.bdim[0] * fg.bdim[1])) * 4 + (lane >> 4)};
  bool act = gi[0] < fg.gdim[0] && gi[1] < fg.gdim[1] && gi[2] < fg.gdim[2];
  const int64_t g = act ? ((int64_t)gi[2] * fg.gdim[1] + gi[1]) * fg.gdim[0] + gi[0] : 0;
  act = act && gflag[g];
  if (__ballot(act) == 0) return;
  double lo[3], hi[3], c[3], hw[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int i0 = gi[a] * fg.gcell[a] + fg.ioff[a];
    const int i1 = fg.clip ? min(i0 + fg.gcell[a], lk.dims[a]) : i0 + fg.gcell[a];
    lo[a] = lk.o[a] + (double)i0 * lk.h[a] - lk.eps;
    hi[a] = lk.o[a] + (double)i1 * lk.h[a] + lk.eps;
    c[a] = 0.5 * (lo[a] + hi[a]);
    hw[a] = 0.5 * (hi[a] - lo[a]);
  }
  // first bound: the point a greedy descent from the box centre reaches
  double U2 = INFINITY, dd = INFINITY;
  if (act) {
    fomo::SearchBox sb;
    sb.t0 = (float)c[0], sb.t1 = (float)c[1], sb.t2 = (float)c[2];
    sb.lo0 = sb.lo1 = sb.lo2 = -INFINITY;
    sb.hi0 = sb.hi1 = sb.hi2 = INFINITY;
    int bpos = -1, borig = 0x7fffffff;
    fomo::oct_descend(pts, mpts, nlo, nhi, sb, dd, bpos, borig);
    if (bpos >= 0) U2 = box_maxdist2(lo, hi, pts[bpos]);
    act = bpos >= 0;
  }
  // pass 1: the 4 points nearest to the centre, within one cell edge beyond the descent's point
  // (a walk like a 4-NN query), and U = the smallest maxdist among them. Any point's maxdist bounds
  // U and any points are sound dominators; the nearest to the centre prune the most.
  float4 dom[4];
  int dpos[4];
  double dkey[4];
  int nd = 0;
  const double ctr[3] = {c[0], c[1], c[2]};  // a box of no extent: the walk's test is the distance to c
  double r2 = -1.0;
  if (act) {
    const double r0 = sqrt(dd) + lk.h[0];
    r2 = r0 * r0;
  }
  far_walk(nlo, nhi, nodes, ctr, ctr, r2, [](const float4&, const float4&) { return true; }, [&](int first, int count) {
    for (int i = first; i < first + count; ++i) {
      const float4 q = mpts[i];
      const double d2 = box_mindist2(ctr, ctr, q);
      if (d2 > r2) continue;
      U2 = fmin(U2, box_maxdist2(lo, hi, q));
      int k = nd < 4 ? nd++ : 4;
      if (k == 4) {
        if (d2 >= dkey[3]) continue;
        k = 3;
      }
      while (k > 0 && dkey[k - 1] > d2) {
        dkey[k] = dkey[k - 1];
        dom[k] = dom[k - 1];
        dpos[k] = dpos[k - 1];
        --k;
      }
      dkey[k] = d2;
      dom[k] = q;
      dpos[k] = __float_as_int(q.w);
      if (nd == 4) r2 = fmin(r2, dkey[3]);
    }
  });
  // pass 2: the points within U (with margin) that none of the 4 dominates (whole subtrees one of
  // them dominates are skipped), pruned pairwise as they come: a point a kept one dominates is
  // skipped, and a new point drops the kept ones it dominates, so the list never holds more than
  // its final size plus the points still to lose
  const double Up2 = far_margin2(U2, lk);
  float4* const cb = cbuf + ((int64_t)blockIdx.x * kFarBlock + lane) * kFarCap;
  int n = 0, seen = 0;
  r2 = act ? Up2 : -1.0;
  const auto admit = [&](const float4& l, const float4& h) {
    for (int k = 0; k < nd; ++k)
      if (node_dominated(c, hw, l, h, dom[k])) return false;
    return true;
  };
  far_walk(nlo, nhi, nodes, lo, hi, r2, admit, [&](int first, int count) {
    for (int i = first; i < first + count && r2 >= 0.0; ++i) {
      const float4 q = mpts[i];
      if (box_mindist2(lo, hi, q) > Up2) continue;
      if (++seen > kFarVisit) {
        r2 = -1.0;  // the middle of a concave hole: every point of its rim is near
        n = kFarCap + 1;
        break;
      }
      const int p = __float_as_int(q.w);
      bool dm = false;
      for (int k = 0; k < nd && !dm; ++k)
        if (dpos[k] != p && dominated(c, hw, q, dom[k])) dm = true;
      for (int e = 0; e < n && !dm; ++e)
        if (dominated(c, hw, q, cb[e])) dm = true;
      if (dm) continue;
      for (int e = 0; e < n;) {
        if (dominated(c, hw, cb[e], q)) cb[e] = cb[--n];
        else ++e;
      }
      if (n == kFarCap) {
        r2 = -1.0;  // too many: the group stays with the exact query
        n = kFarCap + 1;
        break;
      }
      cb[n++] = q;
    }
  });
  if (!act) return;
  int2 info = make_int2(0, 0);
  if (n > 0 && n <= kFarCap) {
    const int room = (n + 7) & ~7;  // whole 128-byte lines
    const int64_t pb = (int64_t)atomicAdd(fused, (unsigned long long)room);
    if (pb + room <= fcap) {
      for (int a = 0; a < n; ++a) fpool[pb + a] = cb[a];
      info = make_int2((int)pb, n);
    }
  }
  ginfo[g] = info;
}

// Outer groups: gflag[g] = 1 for the groups reaching beyond the lookup grid.
__global__ void k_outer_flags(LookupInfo lk, FarGrid fg, int64_t ng, uint8_t* __restrict__ gflag) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int gi[3] = {(int)(g % fg.gdim[0]), (int)((g / fg.gdim[0]) % fg.gdim[1]),
                     (int)(g / ((int64_t)fg.gdim[0] * fg.gdim[1]))};
  bool inside = true;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int i0 = gi[a] * fg.gcell[a] + fg.ioff[a];
    inside = inside && i0 >= 0 && i0 + fg.gcell[a] <= lk.dims[a];
  }
  gflag[g] = inside ? 0 : 1;
}

// Outer group headers: (start in lcand, count) from the lists' places in the far pool.
__global__ void k_outer_headers(const int2* __restrict__ ginfo, int64_t ng, int64_t base, int2* __restrict__ ohdr) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int2 gi = ginfo[g];
  ohdr[g] = gi.y > 0 ? make_int2((int)(base + gi.x), gi.y) : make_int2(0, 0);
}

// Left cells of groups with a list: the record's header points at the group's list.
__global__ void k_far_headers(int32_t* __restrict__ cnt, int64_t ncl, LookupInfo lk, FarGrid fg,
                              const int2* __restrict__ ginfo, int64_t base, float4* __restrict__ slot) {
  const int64_t L = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (L >= ncl || cnt[L] != -1) return;
  const int lx = (int)(L % lk.dims[0]);
  const int ly = (int)((L / lk.dims[0]) % lk.dims[1]);
  const int lz = (int)(L / ((int64_t)lk.dims[0] * lk.dims[1]));
  const int64_t g = ((int64_t)(lz / fg.gcell[2]) * fg.gdim[1] + ly / fg.gcell[1]) * fg.gdim[0] + lx / fg.gcell[0];
  const int2 gi = ginfo[g];
  if (gi.y <= 0) return;
  slot[(int64_t)kRecSlots * L] =
      make_float4(__int_as_float((int)(base + gi.x)), __int_as_float(gi.y), 0.f, __int_as_float(-2));
  cnt[L] = gi.y;
}

// ------------------------------------------------------------------------------------
// Tetra path (fomo-CGAL.cpp:75-307): per ray, walk the Delaunay tets from sample to sample; the
// sample takes the values of the nearest vertex of its containing tet (nearest_vertex_in_cell,
// :211-222), or of none outside the convex hull (:225-228). MODE_LINEAR (an extension) takes the
// barycentric interpolation of the tet's vertex values instead.
// Orientation tests are FP64 with a conservative filter; an undecided test (a sample on a face
// plane to within 1e-12 of the determinant's magnitude) counts as "not beyond" the face, so a
// sample on a shared face stays in the tet the walk reached first.
// ------------------------------------------------------------------------------------
constexpr int kTetMaxSteps = 1 << 16;
constexpr int kTetRedoBlocks = 512;  // k_tet_redo grid (grid-stride over the redo list)
constexpr int64_t kTetThreads = 1 << 18;  // segments per launch worth splitting rays for
constexpr int kTetMaxSegment = 16;          // longest run of LOS samples one thread walks

__device__ __forceinline__ int orient_d(const float4 a, const float4 b, const float4 c, double px, double py,
                                        double pz, double* det_out = nullptr) {
  const double ax = a.x, ay = a.y, az = a.z;
  const double bx = (double)b.x - ax, by = (double)b.y - ay, bz = (double)b.z - az;
  const double cx = (double)c.x - ax, cy = (double)c.y - ay, cz = (double)c.z - az;
  const double dx = px - ax, dy = py - ay, dz = pz - az;
  const double m1 = cy * dz - cz * dy, m2 = cx * dz - cz * dx, m3 = cx * dy - cy * dx;
  const double det = bx * m1 - by * m2 + bz * m3;
  if (det_out) *det_out = det;
  const double perm = fabs(bx) * (fabs(cy * dz) + fabs(cz * dy)) + fabs(by) * (fabs(cx * dz) + fabs(cz * dx)) +
                      fabs(bz) * (fabs(cx * dy) + fabs(cy * dx));
  const double bound = 1e-12 * perm;
  return det > bound ? 1 : (det < -bound ? -1 : 0);
}

__device__ __forceinline__ int ivec(const int4 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }

__device__ __forceinline__ int tet_start(const TetGrid& tg, const int32_t* __restrict__ tstart, double px, double py,
                                         double pz) {
  int c[3];
  const double pc[3] = {px, py, pz};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double v = (pc[a] - tg.o[a]) * tg.invh;
    c[a] = v < 0 ? 0 : (v >= tg.dims[a] ? tg.dims[a] - 1 : (int)v);
  }
  return tstart[((int64_t)c[2] * tg.dims[1] + c[1]) * tg.dims[0] + c[0]];
}

// Tet records, kTetRec float4 each: the 4 vertices {x, y, z, sorted position as int bits (-1 =
// infinite)} and the 4 neighbours (int bits). The vertex coordinates sit beside the vertex ids, so
// one step of the walk is one round of loads instead of two dependent ones, and the neighbours in
// the same 80 bytes, so a step touches one 128-byte line or two (separate arrays: always two).
constexpr int kTetRec = 5;

__device__ __forceinline__ int4 tet_nbrs(const float4& f) {
  return make_int4(__float_as_int(f.x), __float_as_int(f.y), __float_as_int(f.z), __float_as_int(f.w));
}

__global__ void k_tet_records(const int4* __restrict__ tv, const int4* __restrict__ tn,
                              const float4* __restrict__ pts, int64_t n, float4* __restrict__ trec) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int4 v = tv[q];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const int id = ivec(v, m);
    float4 r = id >= 0 ? pts[id] : make_float4(0.f, 0.f, 0.f, 0.f);
    r.w = __int_as_float(id);
    trec[kTetRec * q + m] = r;
  }
  const int4 nb = tn[q];
  trec[kTetRec * q + 4] = make_float4(__int_as_float(nb.x), __int_as_float(nb.y), __int_as_float(nb.z),
                                      __int_as_float(nb.w));
}

// The outputs of one sample from its tet q (vertex records) or outside (no tet): the nearest
// vertex of the cell (first of equals in the cell's vertex order, fomo-CGAL.cpp:211-222), or with
// LINEAR the barycentric interpolation of the vertex values; vert (when given) gets the chosen
// vertex (sorted position, -1 outside).
template <int RS4, bool LINEAR>
__device__ __forceinline__ void tet_emit(const float4 (&q)[4], bool outside, double px, double py, double pz,
                                         size_t sidx, size_t S, double cb, double sb, double cl, double sl,
                                         const float4* __restrict__ rec, const RecInfo& ri, int nlines,
                                         float* __restrict__ svlos, float2* __restrict__ sem,
                                         int32_t* __restrict__ vert) {
  if (outside) {
    if (vert) vert[sidx] = -1;
    if constexpr (RS4 > 0) write_sample<RS4>(-1, rec, ri, nlines, S, sidx, cb, sb, cl, sl, svlos, sem);
    return;
  }
  if constexpr (!LINEAR) {
    // nearest vertex of the cell (first of equals in the cell's vertex order)
    int best = 0;
    double bd = INFINITY;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const double d0 = (double)q[m].x - px, d1 = (double)q[m].y - py, d2 = (double)q[m].z - pz;
      const double d = (d0 * d0 + d1 * d1) + d2 * d2;
      if (d < bd) {
        bd = d;
        best = m;
      }
    }
    const int r = __float_as_int(best == 0 ? q[0].w : (best == 1 ? q[1].w : (best == 2 ? q[2].w : q[3].w)));
    if (vert) vert[sidx] = r;
    if constexpr (RS4 > 0) write_sample<RS4>(r, rec, ri, nlines, S, sidx, cb, sb, cl, sl, svlos, sem);
  } else {
    // barycentric weights: volume of the tet with vertex m replaced by the sample, over the volume
    double w[4], tot = 0.0;
    (void)orient_d(q[2], q[1], q[3], px, py, pz, &w[0]);
    (void)orient_d(q[0], q[2], q[3], px, py, pz, &w[1]);
    (void)orient_d(q[1], q[0], q[3], px, py, pz, &w[2]);
    (void)orient_d(q[0], q[1], q[2], px, py, pz, &w[3]);
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      w[f] = fmax(w[f], 0.0);
      tot += w[f];
    }
    if (vert) vert[sidx] = __float_as_int(q[0].w);
    if constexpr (RS4 > 0) {
      constexpr int NL = 4 * RS4 - 4;  // lines in a record
      double vl = 0.0, pk[NL], fw[NL];
#pragma unroll
      for (int l = 0; l < NL; ++l) pk[l] = fw[l] = 0.0;
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        const double wm = tot > 0 ? w[m] / tot : 0.25;
        const float4* rp = rec + (size_t)__float_as_int(q[m].w) * RS4;
        float f[4 * RS4];
#pragma unroll
        for (int qq = 0; qq < RS4; ++qq) {
          const float4 a = rp[qq];
          f[4 * qq] = a.x, f[4 * qq + 1] = a.y, f[4 * qq + 2] = a.z, f[4 * qq + 3] = a.w;
        }
        vl += wm * (((0.0 + (sb * cl) * (double)f[0]) + (-sb * sl) * (double)f[1]) + cb * (double)f[2]);
        const double st = rec_st(ri.K[0], f[3]);
#pragma unroll
        for (int l = 0; l < NL; ++l) {
          pk[l] += wm * (double)f[4 + l];
          fw[l] += wm * (double)rec_fw(ri.K[l], st);
        }
      }
      svlos[sidx] = (float)vl;
#pragma unroll
      for (int l = 0; l < NL; ++l)
        if (l < nlines) sem[(size_t)l * S + sidx] = make_float2((float)pk[l], (float)fw[l]);
    }
  }
}

// One thread per segment of seglen samples of a ray: the walk is a chain of dependent loads, so
// small views split their rays to keep enough walks in flight. Each segment's first sample starts
// from the start grid, the others from the previous sample's tet. A warm start on a hull tet whose
// face does not exclude the sample (a ray entering the hull) restarts from the start grid too, so
// no walk crosses the hull surface from where the ray left it. Writes per-sample svlos / sem like
// k_nn (RS4 > 0) and, when vert != null, the chosen vertex (sorted position, -1 outside) per
// sample in k-major order. A sample whose final tet had an undecided test, or whose walk reached
// the step cap, goes to the redo list (redo[0]: entries, redo[1]: samples beyond redo_cap, then
// the sample indices), which k_tet_redo walks again with exact tests.
template <int RS4, bool LINEAR>
__global__ void __launch_bounds__(256) k_tet_walk(const float4* __restrict__ trec,
                                                  const int32_t* __restrict__ tstart, TetGrid tg,
                                                  const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                                  int seglen, const float4* __restrict__ rec, RecInfo ri, int nlines,
                                                  float* __restrict__ svlos, float2* __restrict__ sem,
                                                  int32_t* __restrict__ vert, uint32_t* __restrict__ redo,
                                                  int redo_cap) {
  const int64_t R = (int64_t)nx * ny;
  const int64_t tid = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int nseg = (nz + seglen - 1) / seglen;
  // a wave takes an 8 x 8 pixel tile of one segment: its walks stay in a compact region of the
  // mesh (shared cache lines), where a row of 64 pixels would spread them over 32 point spacings
  const int tnx = (nx + 7) >> 3, tny = (ny + 7) >> 3;
  const int64_t ntile = (int64_t)tnx * tny, w = tid >> 6;
  const int lane = (int)(tid & 63);
  const int64_t tl = w % ntile;
  const int seg = (int)(w / ntile);
  const int j = (int)(tl % tnx) * 8 + (lane & 7);
  const int i = (int)(tl / tnx) * 8 + (lane >> 3);
  if (seg >= nseg || j >= nx || i >= ny) return;
  const int64_t ray = (int64_t)i * nx + j;
  const ViewGeom vg = *vgp;
  const double minx = vg.ext[0], maxx = vg.ext[1], miny = vg.ext[2], maxy = vg.ext[3], minz = vg.ext[4],
               maxz = vg.ext[5];
  const double x = double(j) / (nx - 1) * (maxx - minx) + minx;  // fomo-CGAL.cpp:187
  const double y = double(i) / (ny - 1) * (maxy - miny) + miny;  // :188
  const double cb = vg.cb, sb = vg.sb, cl = vg.cl, sl = vg.sl;
  const size_t S = (size_t)R * nz;
  int t = -1;  // tet of the previous sample (-1: none)
  int tq = -1;  // the tet whose record q / nb hold (the previous sample's walk ended there)
  float4 q[4];
  int4 nb;
  const int k0 = seg * seglen, k1 = min(nz, k0 + seglen);
  for (int k = k0; k < k1; ++k) {
    // :189 (a single z pixel has no spacing: the sample sits at minz)
    const double z = nz > 1 ? double(k) / (nz - 1) * (maxz - minz) + minz : minz;
    const double px = x * cb * cl + y * sl + z * sb * cl;  // :192
    const double py = -x * cb * sl + y * cl - z * sb * sl;
    const double pz = -x * sb + z * cb;
    const size_t sidx = (size_t)k * R + ray;
    bool outside = px < tg.lo[0] || px > tg.hi[0] || py < tg.lo[1] || py > tg.hi[1] || pz < tg.lo[2] ||
                   pz > tg.hi[2];
    bool flag = false;  // an undecided test at the end of the walk, or the step cap
    if (!outside) {
      const bool warm = t >= 0;
      if (!warm) t = tet_start(tg, tstart, px, py, pz);
      int steps = 0;
      for (;;) {
        if (t != tq) {
          const float4* tr = trec + (size_t)t * kTetRec;
          q[0] = tr[0], q[1] = tr[1], q[2] = tr[2], q[3] = tr[3];
          nb = tet_nbrs(tr[4]);
          tq = t;
        }
        int inf = -1;
#pragma unroll
        for (int m = 0; m < 4; ++m)
          if (__float_as_int(q[m].w) < 0) inf = m;
        if (inf >= 0) {
          // hull face opposite the infinite vertex (the face order of fomo/delaunay.h)
          int o;
          switch (inf) {
            case 0: o = orient_d(q[2], q[1], q[3], px, py, pz); break;
            case 1: o = orient_d(q[0], q[2], q[3], px, py, pz); break;
            case 2: o = orient_d(q[1], q[0], q[3], px, py, pz); break;
            default: o = orient_d(q[0], q[1], q[2], px, py, pz); break;
          }
          if (o > 0) {
            outside = true;
            break;
          }
          t = (steps == 0 && warm) ? tet_start(tg, tstart, px, py, pz) : ivec(nb, inf);
        } else {
          // faces opposite v0..v3: (2,1,3), (0,2,3), (1,0,3), (0,1,2); the walk stops in a tet no
          // face of which has the sample strictly beyond it, which is certain only when no test
          // there was undecided (0). It leaves through the face whose plane the sample is furthest
          // beyond in the determinant's measure (any such face keeps the walk acyclic on a
          // Delaunay triangulation; this choice takes fewer steps than the first such face).
          double dd[4];
          const int o0 = orient_d(q[2], q[1], q[3], px, py, pz, &dd[0]);
          const int o1 = orient_d(q[0], q[2], q[3], px, py, pz, &dd[1]);
          const int o2 = orient_d(q[1], q[0], q[3], px, py, pz, &dd[2]);
          const int o3 = orient_d(q[0], q[1], q[2], px, py, pz, &dd[3]);
          const bool und = o0 == 0 || o1 == 0 || o2 == 0 || o3 == 0;
          int next = -1;
          double best = 0.0;
          if (o0 < 0 && dd[0] < best) best = dd[0], next = nb.x;
          if (o1 < 0 && dd[1] < best) best = dd[1], next = nb.y;
          if (o2 < 0 && dd[2] < best) best = dd[2], next = nb.z;
          if (o3 < 0 && dd[3] < best) best = dd[3], next = nb.w;
          if (next < 0) {
            flag = und;
            break;
          }
          t = next;
        }
        if (++steps > kTetMaxSteps) {  // never expected (the walk on a Delaunay triangulation ends)
          outside = true;
          flag = true;
          t = -1;
          break;
        }
      }
    }
    if (flag) {
      // the filtered tests could not place the sample: k_tet_redo walks it again exactly
      const unsigned e = atomicAdd(&redo[0], 1u);
      if (e < (unsigned)redo_cap) redo[2 + e] = (uint32_t)sidx;
      else atomicAdd(&redo[1], 1u);  // left as the filtered walk placed it; counted as unresolved
    }
    tet_emit<RS4, LINEAR>(q, outside, px, py, pz, sidx, S, cb, sb, cl, sl, rec, ri, nlines, svlos, sem, vert);
  }
}


// The samples of the redo list, walked again from the start grid with exact orientation tests
// (fomo::orient_exact_fp); a sample on a face plane (exact 0) stays in the first tet the walk
// reaches, as in k_tet_walk. stats[0] += samples walked here, stats[1] += samples left unresolved
// (beyond the list, or at the step cap, which an exact walk on a Delaunay triangulation does not
// reach); redo[1] holds this view's unresolved count afterwards.
__device__ __forceinline__ int orient_x(const float4& a, const float4& b, const float4& c, double px, double py,
                                        double pz) {
  // the filtered test's sign is exact where it decides; the expansion only where it cannot
  const int f = orient_d(a, b, c, px, py, pz);
  if (f != 0) return f;
  const double A[3] = {a.x, a.y, a.z}, B[3] = {b.x, b.y, b.z}, Cc[3] = {c.x, c.y, c.z}, P[3] = {px, py, pz};
  return fomo::orient_exact_fp(A, B, Cc, P);
}

template <int RS4, bool LINEAR>
__global__ void __launch_bounds__(64) k_tet_redo(const float4* __restrict__ trec,
                                                 const int32_t* __restrict__ tstart, TetGrid tg,
                                                 const ViewGeom* __restrict__ vgp, int nx, int ny, int nz,
                                                 const float4* __restrict__ rec, RecInfo ri, int nlines,
                                                 float* __restrict__ svlos, float2* __restrict__ sem,
                                                 int32_t* __restrict__ vert, uint32_t* __restrict__ redo, int redo_cap,
                                                 unsigned long long* __restrict__ stats) {
  const unsigned cnt = min(redo[0], (unsigned)redo_cap);
  const int64_t R = (int64_t)nx * ny;
  const size_t S = (size_t)R * nz;
  const ViewGeom vg = *vgp;
  const double minx = vg.ext[0], maxx = vg.ext[1], miny = vg.ext[2], maxy = vg.ext[3], minz = vg.ext[4],
               maxz = vg.ext[5];
  const double cb = vg.cb, sb = vg.sb, cl = vg.cl, sl = vg.sl;
  unsigned unresolved = 0;
  for (unsigned e = blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += gridDim.x * blockDim.x) {
    const size_t sidx = redo[2 + e];
    const int64_t ray = (int64_t)(sidx % (size_t)R);
    const int k = (int)(sidx / (size_t)R);
    const int j = (int)(ray % nx), i = (int)(ray / nx);
    const double x = double(j) / (nx - 1) * (maxx - minx) + minx;  // fomo-CGAL.cpp:187
    const double y = double(i) / (ny - 1) * (maxy - miny) + miny;  // :188
    const double z = nz > 1 ? double(k) / (nz - 1) * (maxz - minz) + minz : minz;  // :189
    const double px = x * cb * cl + y * sl + z * sb * cl;  // :192
    const double py = -x * cb * sl + y * cl - z * sb * sl;
    const double pz = -x * sb + z * cb;
    bool outside = false;
    float4 q[4];
    int t = tet_start(tg, tstart, px, py, pz);
    for (int steps = 0;; ++steps) {
      if (steps > kTetMaxSteps) {
        outside = true;
        ++unresolved;
        break;
      }
      const float4* tr = trec + (size_t)t * kTetRec;
      q[0] = tr[0], q[1] = tr[1], q[2] = tr[2], q[3] = tr[3];
      const int4 nb = tet_nbrs(tr[4]);
      int inf = -1;
#pragma unroll
      for (int m = 0; m < 4; ++m)
        if (__float_as_int(q[m].w) < 0) inf = m;
      if (inf >= 0) {
        int o;
        switch (inf) {
          case 0: o = orient_x(q[2], q[1], q[3], px, py, pz); break;
          case 1: o = orient_x(q[0], q[2], q[3], px, py, pz); break;
          case 2: o = orient_x(q[1], q[0], q[3], px, py, pz); break;
          default: o = orient_x(q[0], q[1], q[2], px, py, pz); break;
        }
        if (o > 0) {
          outside = true;
          break;
        }
        t = ivec(nb, inf);
      } else {
        int next = -1;
        if (orient_x(q[2], q[1], q[3], px, py, pz) < 0) next = nb.x;
        else if (orient_x(q[0], q[2], q[3], px, py, pz) < 0) next = nb.y;
        else if (orient_x(q[1], q[0], q[3], px, py, pz) < 0) next = nb.z;
        else if (orient_x(q[0], q[1], q[2], px, py, pz) < 0) next = nb.w;
        if (next < 0) break;
        t = next;
      }
    }
    tet_emit<RS4, LINEAR>(q, outside, px, py, pz, sidx, S, cb, sb, cl, sl, rec, ri, nlines, svlos, sem, vert);
  }
  if (unresolved) {
    atomicAdd(&redo[1], unresolved);
    atomicAdd(&stats[1], (unsigned long long)unresolved);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    atomicAdd(&stats[0], (unsigned long long)cnt);
    atomicAdd(&stats[1], (unsigned long long)(redo[0] > (unsigned)redo_cap ? redo[0] - redo_cap : 0));
  }
}

// ------------------------------------------------------------------------------------
// Projection method (fomo-projection.cpp:17-214): every point adds its line profile to the pixel
// its rotated (x', y') rounds to (:126-158). The reference scatters with OpenMP atomics, so its
// summation order is not fixed. Here the points are sorted by pixel with a stable radix sort of
// (pixel, sorted position) pairs laid out in original point order, and each (pixel, lambda) sums
// its points in that order in float, as intens.at(ind) += tempintens does (:144, :154): the sums
// of a one-thread run of the reference.
// ------------------------------------------------------------------------------------
__device__ __forceinline__ int proj_pixel(double v, double mn, double mx, int n) {
  // :129, :131 with std::round (half away from zero). A degenerate extent has no pixel formula in
  // the reference (0/0): every point goes to pixel 0. The clamp only guards memory.
  if (!(mx > mn)) return 0;
  const int q = (int)round((v - mn) * (n - 1) / (mx - mn));
  return q < 0 ? 0 : (q > n - 1 ? n - 1 : q);
}

// key[o] = pixel of the point with original index o, val[o] = its sorted position.
__global__ void k_proj_keys(const float4* __restrict__ pts, const int32_t* __restrict__ perm, int64_t n,
                            const ViewGeom* __restrict__ vgp, int nx, int ny, uint32_t* __restrict__ key,
                            int32_t* __restrict__ val) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double cb = vgp->cb, sb = vgp->sb, cl = vgp->cl, sl = vgp->sl;
  const float4 q = pts[p];
  const double g0 = q.x, g1 = q.y, g2 = q.z;
  const double xr = g0 * cb * cl - g1 * cb * sl - g2 * sb;  // :50
  const double yr = g0 * sl + g1 * cl;                      // :51
  const int j = proj_pixel(xr, vgp->ext[0], vgp->ext[1], nx);
  const int i = proj_pixel(yr, vgp->ext[2], vgp->ext[3], ny);
  const int32_t o = perm[p];
  key[o] = (uint32_t)i * (uint32_t)nx + (uint32_t)j;
  val[o] = (int32_t)p;
}

// start[q] = first position of pixel q in the sorted keys (q = 0..R; start[R] = n).
__global__ void k_proj_starts(const uint32_t* __restrict__ key, int64_t n, int64_t R, int32_t* __restrict__ start) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q > R) return;
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)key[mid] < q) lo = mid + 1;
    else hi = mid;
  }
  start[q] = (int32_t)lo;
}

// One thread per (pixel, chunk of kProjChunk lambda pixels), adding the pixel's points in point
// order; nl == 1: one thread per pixel summing the peaks (Imaging, :148-155). A point's Gaussian
// over the chunk comes from the ratio recurrence (three exps per point and chunk,
// profile_chains), and each term is added as the reference adds it: tmp = peak * g in double,
// then acc = (float)(acc + tmp) (:139-144). A point whose anchors would leave the double range
// takes the reference's exp for every lambda pixel.
constexpr int kProjChunk = 16;
template <int RS4>
__global__ void __launch_bounds__(256) k_proj_spectral(const int32_t* __restrict__ start,
                                                       const int32_t* __restrict__ order,
                                                       const float4* __restrict__ rec, RecInfo ri,
                                                       const ViewGeom* __restrict__ vgp, int64_t R, int nl,
                                                       double wla, double lambda0, float* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (nl <= 1) {
    if (t >= R) return;
    float acc = 0.f;
    for (int q = start[t]; q < start[t + 1]; ++q) acc = acc + rec[(size_t)order[q] * RS4 + 1].x;  // :154
    out[t] = (float)(vgp->scale * (double)acc);  // :200
    return;
  }
  const int nch = (nl + kProjChunk - 1) / kProjChunk;
  if (t >= R * nch) return;
  const int64_t pix = t / nch;
  const int c0 = (int)(t - pix * nch) * kProjChunk;
  const int q0 = start[pix], q1 = start[pix + 1];
  const double cb = vgp->cb, sb = vgp->sb, cl = vgp->cl, sl = vgp->sl;
  const double u0 = sb * cl, u1 = -sb * sl, u2 = cb;                          // :34
  const double dl = wla / (nl - 1);
  const double lamc = static_cast<double>(c0) / (nl - 1) * wla - wla / 2.;  // lambdaval at c0 (:138)
  const double ln2 = log(2.);
  const double K = ri.K[0];
  float acc[kProjChunk];
#pragma unroll
  for (int c = 0; c < kProjChunk; ++c) acc[c] = 0.f;
  for (int q = q0; q < q1; ++q) {
    const float4* rp = rec + (size_t)order[q] * RS4;
    const float4 a = rp[0];  // vx, vy, vz, T
    const float pk = rp[1].x;
    if (pk == 0.f) continue;  // adds zeros
    const double losvel = ((0.0 + u0 * (double)a.x) + u1 * (double)a.y) + u2 * (double)a.z;  // :53-54
    const double fw = (double)rec_fw(K, rec_st(K, a.w));
    const double mu = losvel / FOMO_C_LIGHT * lambda0;
    double xa[3];
    if (profile_args(fomo::spectral_kappa(fw), lamc - mu, dl, xa)) {
      double ex[3], g[4], RR[4], q16;
      fomo::exp3(xa, ex);
      profile_chains(ex, g, RR, q16);
#pragma unroll
      for (int c4 = 0; c4 < kProjChunk; c4 += 4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const double tmp = (double)pk * g[j];
          acc[c4 + j] = (float)((double)acc[c4 + j] + tmp);  // :144
          g[j] *= RR[j];
          RR[j] *= q16;
        }
      }
    } else {
#pragma unroll
      for (int c = 0; c < kProjChunk; ++c) {
        const double lambdaval = static_cast<double>(c0 + c) / (nl - 1) * wla - wla / 2.;  // :138
        const double d = lambdaval - losvel / FOMO_C_LIGHT * lambda0;
        const double tmp = (double)pk * exp((-(d * d)) / (fw * fw) * 4. * ln2);  // :139
        acc[c] = (float)((double)acc[c] + tmp);                                  // :144
      }
    }
  }
  float* o = out + (size_t)pix * nl + c0;
#pragma unroll
  for (int c = 0; c < kProjChunk; ++c)
    if (c0 + c < nl) o[c] = (float)(vgp->scale * (double)acc[c]);  // :200
}

// Per-point records for k_nn (sorted order): {vx, vy, vz, T, peak_0, peak_1, ...}, rs4 float4 each.
// Emission pair of every sorted point from two of its variables (a given GoftCube's).
__global__ void k_pack_em(const float* __restrict__ a, const float* __restrict__ b, int64_t n, float2* __restrict__ em) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p < n) em[p] = make_float2(a[p], b[p]);
}

__global__ void k_build_rec(const float4* __restrict__ vel, const float* __restrict__ T, const LineSet ls, int64_t n,
                            int rs4, float4* __restrict__ rec) {
  const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (p >= n) return;
  float f[32];
#pragma unroll
  for (int q = 0; q < 32; ++q) f[q] = 0.f;
  if (vel) {
    const float4 v = vel[p];
    f[0] = v.x;
    f[1] = v.y;
    f[2] = v.z;
  }
  f[3] = T[p];
#pragma unroll
  for (int l = 0; l < kMaxLines; ++l)
    if (l < ls.n) f[4 + l] = ls.em[l][p].x;
#pragma unroll
  for (int q = 0; q < 8; ++q)
    if (q < rs4) rec[(size_t)p * rs4 + q] = make_float4(f[4 * q], f[4 * q + 1], f[4 * q + 2], f[4 * q + 3]);
}

// sorted positions -> original indices, transposed to [ray][k] (parity output)
__global__ void k_idx_to_orig(const int32_t* __restrict__ idx, const int32_t* __restrict__ perm, int64_t R, int nz,
                              int32_t* __restrict__ out) {
  int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (s >= R * nz) return;
  const int64_t ray = s / nz;
  const int k = (int)(s - ray * nz);
  const int32_t p = idx[(size_t)k * R + ray];
  out[s] = p < 0 ? -1 : perm[p];
}

// ------------------------------------------------------------------------------------
// K5: LOS spectral integration (fomo-nearestneighbour.cpp:195-215, scale :235-267).
// One thread per ray, CH lambda pixels [l0, l0+CH) per pass. The Gaussian
// peak*exp(-4ln2 (lambda-mu)^2 / w^2) is evaluated by the ratio recurrence
//   g_{c+1} = g_c r_c,  r_{c+1} = r_c q,  q = exp(-2 kappa dl^2)
// in FP64 (three FP64 ops per sample x lambda). Accumulation over z is FP64 in the
// reference's order. A ray with a sample whose anchors would under/overflow is flagged and
// left to the DIRECT instantiation (one exp per sample x lambda), launched right after.
// ------------------------------------------------------------------------------------
template <int CH, bool DIRECT>
__global__ void __launch_bounds__(256) k_los_spectral(int R, int nz, const LineSet ls, const float* __restrict__ vlos,
                                                      const ViewGeom* __restrict__ vgp, int nl, int l0,
                                                      int32_t* __restrict__ flags) {
  // one lane per ray: the samples are added in LOS order (the order of k_nn_los)
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = ray < R;
  const int line = blockIdx.y;
  const float2* __restrict__ em = ls.em[line];
  const double wla = ls.wla[line], lam0_over_c = ls.l0c[line];
  int32_t* __restrict__ flag = flags + (size_t)line * R;
  if (DIRECT) {
    // whole waves leave together (no lane of this wave needs the direct pass)
    const bool need = live && flag[ray] != 0;
    if (__ballot(need) == 0) return;
  }
  const double dl = wla / (nl - 1);
  const double lamc = double(l0) / (nl - 1) * wla - wla / 2.;  // lambdaval of the first pixel of the pass (:200)
  float acc[CH];
#pragma unroll
  for (int c = 0; c < CH; ++c) acc[c] = 0.f;
  bool bad = false;
  const int rr_ = live ? ray : 0;
  // per-sample (peak, fwhm) and LOS velocity written by k_nn; one-sample lookahead
  float2 en = make_float2(0.f, 0.f);
  float vn = 0.f;
  if (nz > 0) { en = em[rr_]; vn = vlos[rr_]; }
  for (int k = 0; k < nz; ++k) {
    const float2 e = en;
    const float v = vn;
    if (k + 1 < nz) {
      en = em[(size_t)(k + 1) * R + rr_];
      vn = vlos[(size_t)(k + 1) * R + rr_];
    }
    if (!live || e.x == 0.f) continue;
    const double peak = e.x, w = e.y;
    const double mu = (double)v * lam0_over_c;
    const double kap = fomo::spectral_kappa(w);
    const double d0 = lamc - mu;
    if (DIRECT) {
#pragma unroll
      for (int c = 0; c < CH; ++c) {
        const double d = d0 + (double)c * dl;
        acc[c] = spectral_add(acc[c], (float)peak, exp(-kap * d * d));
      }
    } else {
      double xa[3];
      if (profile_args(kap, d0, dl, xa)) {
        double ex[3], g[4], RR[4], q16;
        fomo::exp3(xa, ex);
        profile_chains(ex, g, RR, q16);
#pragma unroll
        for (int c0 = 0; c0 < CH; c0 += 4) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            acc[c0 + j] = spectral_add(acc[c0 + j], (float)peak, g[j]);
            g[j] *= RR[j];
            RR[j] *= q16;
          }
        }
      } else {
        bad = true;
      }
    }
  }
  if (!DIRECT) {
    if (live) flag[ray] = bad ? 1 : 0;
    if (bad) return;
  } else {
    if (!live || flag[ray] == 0) return;
  }
  if (!live) return;
  const double sc = vgp->scale;
  float* o = ls.out[line] + (size_t)ray * nl + l0;
#pragma unroll
  for (int c = 0; c < CH; ++c)
    if (l0 + c < nl) o[c] = (float)(sc * (double)acc[c]);
}

// nl == 1 (Imaging): intens += peak (fomo-nearestneighbour.cpp:217-225)
__global__ void __launch_bounds__(256) k_los_imaging(int R, int nz, const LineSet ls,
                                                     const ViewGeom* __restrict__ vgp) {
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= R) return;
  const float2* __restrict__ em = ls.em[blockIdx.y];  // per-sample (peak, fwhm), zero when no point
  float* __restrict__ out = ls.out[blockIdx.y];
  double acc = 0.0;
  for (int k = 0; k < nz; ++k) acc += (double)em[(size_t)k * R + ray].x;
  out[ray] = (float)(vgp->scale * acc);
}

// Thomson pB: intens += max(0, pol * sin^2 chi) (fomo-thomson.cpp:325-338, scale :350-359)
// term: each point's max(0, pol * sin^2 chi) for the view (k_view_rotate_thomson)
__global__ void __launch_bounds__(256) k_los_thomson(const int32_t* __restrict__ idx, int R, int nz,
                                                     const double* __restrict__ term,
                                                     const ViewGeom* __restrict__ vgp, float* __restrict__ out) {
  const int ray = blockIdx.x * blockDim.x + threadIdx.x;
  if (ray >= R) return;
  double acc = 0.0;
  for (int k = 0; k < nz; ++k) {
    const int p = idx[(size_t)k * R + ray];
    if (p >= 0) acc += term[p];
  }
  out[ray] = (float)(vgp->scale * acc);
}

// ------------------------------------------------------------------------------------
// host-side objects
// ------------------------------------------------------------------------------------
struct Dataset {
  int64_t n = 0;
  int nvars = 0;
  GridInfo grid{};
  int64_t ncells = 0;
  int maxcount = 0;
  double cell_scale = 1.0;
  float bbox[6] = {0, 0, 0, 0, 0, 0};
  DevBuf raw[3];       // original-order coordinates (kept for re-indexing)
  std::vector<std::unique_ptr<DevBuf>> rawvars;  // original-order variables
  DevBuf pts;          // float4 sorted
  DevBuf perm;         // int32 sorted pos -> original index
  DevBuf cstart;       // int32 [ncells+1]
  std::vector<std::unique_ptr<DevBuf>> svars;  // sorted variables
  DevBuf vel;          // float4 sorted (vx,vy,vz,0) when nvars>=5
  LookupInfo lk{};     // candidate lookup grid (lk.enabled==0: ring search only)
  int lookup_subdiv = 2;         // requested subdivision
  int lookup_active_subdiv = 0;  // subdivision actually built (0: none)
  DevBuf lslot;        // kRecSlots float4 per lookup cell: inline candidates or an overflow header
  DevBuf lcand;        // float4 overflow candidates (x, y, z, sorted position as int bits)
  DevBuf ohdr;         // int2 per outer group of the lookup grid (lk.ohdr)
  DevBuf ghdr;         // int2 per far group whose cells all take its list (lk.ghdr)
  BvhInfo bvh{};       // implicit BVH over Morton-ordered points (exact NN fallback)
  DevBuf mpts;         // float4 (x, y, z, sorted position) in Morton order
  DevBuf nlo, nhi;     // node boxes, all levels
  DevBuf segbox;       // kDeepSegs bounding boxes of the node ranges of k_nn_deep (lo[], then hi[])
  DevBuf wide;         // 256-byte records of the internal nodes (BvhInfo::wide), or none
  // rectilinear datasets: the points form a product grid of distinct axis values, each
  // combination once; the sorted order is then the grid order and no cell list or candidate grid
  // is built (the NN query reads the axes: rect_geom.h)
  bool allow_rect = true;  // detect product grids when indexing (fomo_dataset_set_rectilinear)
  bool rect = false;
  // masked grids: the points are a subset (at least a quarter) of a product grid, each on its own
  // node (a Cartesian mesh cut to a region, like the corona's shell); indexed as scattered points,
  // plus the grid axes and gpos, which k_nn tries first
  bool mrect = false;
  DevBuf gpos;         // int32 per grid node: sorted position of the point on it, -1 for none
  int rect_n[3] = {0, 0, 0};
  float rect_o[3] = {0, 0, 0}, rect_inv[3] = {0, 0, 0};
  double rect_gap2[3] = {0, 0, 0};
  DevBuf rv;           // axis values, the three axes one after another
  DevBuf rbin;         // float4 bin tables, the three axes one after another
  DevBuf corners;      // float4 x 8: the grid's corner points (their rotated extents are the grid's)
  int64_t ncand = 0, nlcells = 0, nfallback = 0;
  // far and outer groups of the lookup grid: groups, groups with a candidate list, far groups
  // whose cells all take the list (group headers)
  int64_t far_groups = 0, far_lists = 0, full_groups = 0, outer_groups = 0, outer_lists = 0;
  int far_budget = kFarBudget;  // node steps of a k_nn_far walk before the deep pass (0: never)
  int tile_log2 = 3;  // ray tile of k_nn blocks: 2^tile_log2 x 256 / 2^tile_log2 (autotuned)
  // tetra path: Delaunay triangulation of the points, built on the host on first use
  int64_t ntets = 0;      // 0: not built
  DevBuf trec;            // kTetRec float4 per tet: 4 vertices {x, y, z, sorted position as int bits (-1 =
                          // infinite)}, then the neighbour opposite each vertex (int bits)
  DevBuf tstart;          // int32 per start-grid cell: a finite tet near the cell centre
  DevBuf tetstats;        // uint64 {samples walked again with exact tests, samples left unresolved}
  TetGrid tk{};           // start grid geometry
  double tet_build_s = 0;
  std::atomic<int> inuse{0};  // batches and queued/running server requests reading this dataset
};

struct Table {
  int nt = 0, nrho = 0;
  DevBuf logt, logn, g;
};

struct Goft {
  int dataset = -1;
  int kind = 0;  // 0 chianti, 1 thomson
  double lambda0 = 0;
  DevBuf em;     // float2 per sorted point
  double K = 0;  // fwhm = (float)(K * sqrt((double)T)); 0 for imaging (fwhm 1)
  DevBuf rec;    // chianti: per-point records {vx, vy, vz, T, peak} for the NN passes
  RecInfo ri{};  // K of the record's line
  int rs4 = 0;
  std::atomic<int> inuse{0};  // queued/running server requests reading this cube
};

struct ViewScratch {
  DevBuf partial;  // doubles
  DevBuf geom;     // ViewGeom
  DevBuf idx;      // int32 S
  DevBuf aux;      // float N (vlos or sin2chi)
  DevBuf out;      // float R*nl
  DevBuf flag;     // int32 R (rays needing the direct-exp pass)
  DevBuf sem;      // float2 S: per-sample (peak, fwhm)
  DevBuf rayflag;  // nn_scratch_bytes: R ray flags for the BVH pass (kept all-zero between uses), deep queue
};

struct Batch {
  int dataset = -1;
  std::vector<int> gofts;
  std::vector<double> l, b;
  int nx = 0, ny = 0, nz = 0, nl = 0, dn = 0;
  double width = 0;
  DevBuf partial, geom, aux, out, flag;
  DevBuf idx;      // per-sample sorted positions (double-buffered)
  DevBuf rayflag;  // nn_scratch_bytes per buffer: ray flags all-zero between uses, deep queue
  DevBuf rec;  // per-point records {vel, T, all lines' peaks}
  DevBuf trig;  // 5 doubles per view (view_trig), for the views' finalize in one launch
  bool fused = false;  // the views take the fused pass, all in one launch
  RecInfo ri{};
  DevBuf sem;  // per-sample (peak, fwhm), nlines x S
  int rs4 = 0;
  std::vector<hipEvent_t> ev;  // 2 per view: NN done, integration done (graph dependencies)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~Batch() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

struct ProfStat {
  int64_t count = 0;
  double ms = 0;
};

struct Server;

struct Ctx {
  bool init = false;
  int device = -1;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr;  // second stream for overlapped batch work (default priority)
  hipDeviceProp_t prop{};
  int next_id = 1;
  std::map<int, std::unique_ptr<Dataset>> datasets;
  // datasets being filled by fomo_dataset_append: the object and the point ranges written so far
  struct Ingest {
    std::unique_ptr<Dataset> ds;
    int dim = 3;
    std::vector<std::pair<int64_t, int64_t>> written;  // (offset, count)
  };
  std::map<int, Ingest> ingests;
  std::map<int, std::unique_ptr<Table>> tables;
  std::map<int, std::unique_ptr<Goft>> gofts;
  std::map<int, std::unique_ptr<Batch>> batches;
  std::map<int, std::shared_ptr<Server>> servers;
  ViewScratch vs;
  DevBuf sortkeys[2], sortvals, sorttmp, tmpf;
  bool profile = false;
  std::map<std::string, ProfStat> prof;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> evpool;
  void* noalloc = nullptr;                       // device address of `scratch` on this device
  void* stage = nullptr;                         // pinned staging of large results (copy_out)
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  std::unique_ptr<fomo::HostCopyPool> copy_pool;
};

Ctx g_ctx;

int require_init() {
  if (!g_ctx.init) return fail(FOMO_ERR_STATE, "fomo_init has not been called");
  return FOMO_OK;
}

hipEvent_t pool_event() {
  if (!g_ctx.evpool.empty()) {
    hipEvent_t e = g_ctx.evpool.back();
    g_ctx.evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Event pair around an eager launch when profiling is on.
// Events around one launch, recorded on the launch's stream. Only work on the library's main
// stream is timed (server streams run concurrently, so their durations would overlap).
struct ProfScope {
  const char* name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(const char* n, hipStream_t st) : name(n), s(st) {
    if (g_ctx.profile && s == g_ctx.stream) {
      a = pool_event();
      b = pool_event();
      (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (a) {
      (void)hipEventRecord(b, s);
      g_ctx.pending.push_back({name, {a, b}});
    }
  }
};

void collect_profile() {
  for (auto& pe : g_ctx.pending) {
    float ms = 0;
    if (hipEventSynchronize(pe.second.second) == hipSuccess &&
        hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
      auto& st = g_ctx.prof[pe.first];
      st.count++;
      st.ms += ms;
    }
    g_ctx.evpool.push_back(pe.second.first);
    g_ctx.evpool.push_back(pe.second.second);
  }
  g_ctx.pending.clear();
}

int sync_stream() {
  HIPOK(hipStreamSynchronize(g_ctx.stream));
  collect_profile();
  return FOMO_OK;
}

inline unsigned nblk(int64_t n, int bs = kBlock) { return (unsigned)((n + bs - 1) / bs); }

// Device-to-host copy of a result enqueued on s. Results under 4 MiB are left to the runtime
// (the caller synchronises). Larger ones go through two pinned 8 MiB chunks and are complete on
// return: the device fills one chunk while host threads copy the other to the destination.
constexpr int64_t kStageChunk = 8 << 20;
constexpr int64_t kStageMin = 4 << 20;
constexpr int kStageHelpers = 5;
int copy_out(void* dst, const void* src, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return FOMO_OK;
  if (!g_ctx.stage && bytes >= kStageMin) {
    // pinned staging once per context; without it the runtime's own path still works
    if (hipHostMalloc(&g_ctx.stage, 2 * kStageChunk, hipHostMallocDefault) != hipSuccess) g_ctx.stage = nullptr;
    for (auto& e : g_ctx.stage_ev)
      if (g_ctx.stage && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
        (void)hipHostFree(g_ctx.stage);
        g_ctx.stage = nullptr;
      }
  }
  if (bytes < kStageMin || !g_ctx.stage) {
    HIPOK(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, s));
    return FOMO_OK;
  }
  if (!g_ctx.copy_pool) g_ctx.copy_pool.reset(new fomo::HostCopyPool(kStageHelpers));
  char* const st = static_cast<char*>(g_ctx.stage);
  const int64_t n = (bytes + kStageChunk - 1) / kStageChunk;
  const auto len_of = [&](int64_t c) { return std::min<int64_t>(kStageChunk, bytes - c * kStageChunk); };
  const auto enq = [&](int64_t c) {
    hipError_t e = hipMemcpyAsync(st + (c & 1) * kStageChunk, static_cast<const char*>(src) + c * kStageChunk,
                                  (size_t)len_of(c), hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipEventRecord(g_ctx.stage_ev[c & 1], s) : e;
  };
  HIPOK(enq(0));
  for (int64_t c = 0; c < n; ++c) {
    // chunk c + 1 goes to the half that held chunk c - 1, copied out in the previous step
    if (c + 1 < n) HIPOK(enq(c + 1));
    HIPOK(hipEventSynchronize(g_ctx.stage_ev[c & 1]));
    g_ctx.copy_pool->copy(static_cast<char*>(dst) + c * kStageChunk, st + (c & 1) * kStageChunk,
                          (size_t)len_of(c));
  }
  return FOMO_OK;
}

// Implicit BVH over the dataset's points (Morton order of 10-bit quantised coordinates).
int build_bvh(Dataset& ds) {
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds.n;
  ds.bvh = BvhInfo{};
  if (n > (int64_t)INT32_MAX / 2) return fail(FOMO_ERR_ARG, "too many points for the octree");
  HIPOK(g_ctx.sortkeys[0].ensure(sizeof(uint32_t) * n));
  HIPOK(g_ctx.sortkeys[1].ensure(sizeof(uint32_t) * n));
  HIPOK(g_ctx.sortvals.ensure(sizeof(int32_t) * n));
  DevBuf order;
  if (order.alloc(sizeof(int32_t) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (octree)");
  k_morton_keys<<<nblk(n), kBlock, 0, s>>>(ds.pts.as<float4>(), n, ds.grid, g_ctx.sortkeys[0].as<uint32_t>(),
                                           g_ctx.sortvals.as<int32_t>());
  HIPOK(hipGetLastError());
  size_t tmpb = 0;
  HIPOK(rocprim::radix_sort_pairs(nullptr, tmpb, g_ctx.sortkeys[0].as<uint32_t>(), g_ctx.sortkeys[1].as<uint32_t>(),
                                  g_ctx.sortvals.as<int32_t>(), order.as<int32_t>(), (size_t)n, 0, 30, s));
  HIPOK(g_ctx.sorttmp.ensure(tmpb));
  HIPOK(rocprim::radix_sort_pairs(g_ctx.sorttmp.p, tmpb, g_ctx.sortkeys[0].as<uint32_t>(),
                                  g_ctx.sortkeys[1].as<uint32_t>(), g_ctx.sortvals.as<int32_t>(), order.as<int32_t>(),
                                  (size_t)n, 0, 30, s));
  if (ds.mpts.ensure(sizeof(float4) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (octree)");
  k_morton_points<<<nblk(n), kBlock, 0, s>>>(ds.pts.as<float4>(), order.as<int32_t>(), n, ds.mpts.as<float4>());
  HIPOK(hipGetLastError());
  std::vector<uint32_t> keys(n);
  HIPOK(hipMemcpyAsync(keys.data(), g_ctx.sortkeys[1].p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  fomo::OctBuilder ob;
  ob.build(keys.data(), n);
  const int64_t nodes = (int64_t)ob.w.size();
  // initial node records (boxes filled on the device): nlo.w = skip / -count, nhi.w = first
  std::vector<float4> lo(nodes), hi(nodes);
  for (int64_t i = 0; i < nodes; ++i) {
    lo[i] = make_float4(0.f, 0.f, 0.f, fomo::i2f(ob.w[i]));
    hi[i] = make_float4(0.f, 0.f, 0.f, fomo::i2f(ob.first[i]));
  }
  std::vector<int32_t> ids;
  std::vector<int64_t> doff;
  for (int d = (int)ob.bydepth.size() - 1; d >= 0; --d) {
    doff.push_back((int64_t)ids.size());
    ids.insert(ids.end(), ob.bydepth[d].begin(), ob.bydepth[d].end());
  }
  doff.push_back((int64_t)ids.size());
  DevBuf dids;
  if (ds.nlo.alloc(sizeof(float4) * std::max<int64_t>(nodes, 1)) ||
      ds.nhi.alloc(sizeof(float4) * std::max<int64_t>(nodes, 1)) ||
      dids.alloc(sizeof(int32_t) * std::max<size_t>(ids.size(), 1)))
    return fail(FOMO_ERR_NOMEM, "device allocation failed (octree)");
  HIPOK(hipMemcpyAsync(ds.nlo.p, lo.data(), sizeof(float4) * nodes, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ds.nhi.p, hi.data(), sizeof(float4) * nodes, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(dids.p, ids.data(), sizeof(int32_t) * ids.size(), hipMemcpyHostToDevice, s));
  for (size_t d = 0; d + 1 < doff.size(); ++d) {
    const int64_t cnt = doff[d + 1] - doff[d];
    if (cnt > 0)
      k_oct_boxes<<<nblk(cnt), kBlock, 0, s>>>(dids.as<int32_t>() + doff[d], cnt, ds.mpts.as<float4>(),
                                              ds.nlo.as<float4>(), ds.nhi.as<float4>());
  }
  if (ds.segbox.ensure(sizeof(float4) * 2 * kDeepSegs)) return fail(FOMO_ERR_NOMEM, "device allocation failed (octree)");
  k_seg_boxes<<<kDeepSegs, 256, 0, s>>>(ds.nlo.as<float4>(), ds.nhi.as<float4>(), (int)nodes, ds.segbox.as<float4>());
  HIPOK(hipGetLastError());
  // wide records (BvhInfo::wide) when the root is internal and every internal node lies at depth
  // <= kWideDepth (the bound of wide_walk_wave's stack); else the exact queries walk the preorder
  // records
  ds.wide.release();
  const int64_t depth = (int64_t)ob.bydepth.size() - 1;  // of the deepest node, a leaf
  std::vector<int32_t> kid, kref;
  const int64_t nint = nodes > 1 && depth - 1 <= kWideDepth ? fomo::wide_children(ob, kid, kref) : 0;
  if (nint > 0) {
    DevBuf dkid, dkref;
    if (ds.wide.alloc(sizeof(float4) * 16 * nint) || dkid.alloc(sizeof(int32_t) * kid.size()) ||
        dkref.alloc(sizeof(int32_t) * kref.size()))
      return fail(FOMO_ERR_NOMEM, "device allocation failed (octree)");
    HIPOK(hipMemcpyAsync(dkid.p, kid.data(), sizeof(int32_t) * kid.size(), hipMemcpyHostToDevice, s));
    HIPOK(hipMemcpyAsync(dkref.p, kref.data(), sizeof(int32_t) * kref.size(), hipMemcpyHostToDevice, s));
    k_wide_records<<<nblk(nint * 8), kBlock, 0, s>>>(dkid.as<int32_t>(), dkref.as<int32_t>(), nint * 8,
                                                      ds.nlo.as<float4>(), ds.nhi.as<float4>(), ds.wide.as<float4>());
    HIPOK(hipGetLastError());
    if (int rc = sync_stream()) return rc;
  }
  if (int rc = sync_stream()) return rc;
  ds.bvh.nodes = nodes;
  ds.bvh.n = n;
  ds.bvh.wide = ds.wide.p ? ds.wide.as<float4>() : nullptr;
  return FOMO_OK;
}

// Delaunay triangulation of the dataset for the tetra path (host builder, fomo/delaunay.h), then
// upload: vertex ids become sorted positions; a start grid (up to 64^3 cells over the bbox) holds
// a finite tet near each cell centre, located by a host walk, as the first tet of each ray's walk.
int build_tets(Dataset& ds) {
  if (ds.ntets > 0) return FOMO_OK;
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds.n;
  if (n < 4) return fail(FOMO_ERR_ARG, "the tetra path needs at least 4 points");
  if (n > (int64_t)INT32_MAX / 8) return fail(FOMO_ERR_ARG, "too many points for the tetra path");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<float> hx(n), hy(n), hz(n);
  std::vector<int32_t> perm(n);
  HIPOK(hipMemcpyAsync(hx.data(), ds.raw[0].p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(hy.data(), ds.raw[1].p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(hz.data(), ds.raw[2].p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(perm.data(), ds.perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  fomo::Delaunay dt;
  dt.threads = fomo_delaunay_threads();
  if (!dt.build(hx.data(), hy.data(), hz.data(), n))
    return fail(FOMO_ERR_ARG, "Delaunay triangulation failed: " + dt.error);
  const auto tdt = std::chrono::steady_clock::now();
  const fomo::TetVec& T = dt.T;
  std::vector<int32_t> sorted_of(n);
  for (int64_t p = 0; p < n; ++p) sorted_of[perm[p]] = (int32_t)p;
  // start grid: cells ~ 4 points each along an axis, at most 64 per axis
  TetGrid tk{};
  double ext = 0;
  for (int a = 0; a < 3; ++a) ext = std::max(ext, (double)ds.bbox[2 * a + 1] - ds.bbox[2 * a]);
  const double h = std::max(ext / 64.0, ext > 0 ? ext * 1e-6 : 1.0);
  for (int a = 0; a < 3; ++a) {
    tk.o[a] = tk.lo[a] = ds.bbox[2 * a];
    tk.hi[a] = ds.bbox[2 * a + 1];
    tk.dims[a] = std::max(1, std::min(64, (int)std::ceil(((double)ds.bbox[2 * a + 1] - ds.bbox[2 * a]) / h)));
  }
  tk.h = h;
  tk.invh = 1.0 / h;
  const int64_t ng = (int64_t)tk.dims[0] * tk.dims[1] * tk.dims[2];
  std::vector<int32_t> start(ng, 0);
  {
    // walks with exact predicates on the original-index tets, one z layer of cells per task (the
    // cells in order, each from the previous one's tet; each layer from the first finite tet of
    // the canonical order, so the grid does not depend on the thread count)
    int32_t t0 = 0;
    for (size_t q = 0; q < T.size(); ++q)
      if (dt.finite(T[q])) {
        t0 = (int32_t)q;
        break;
      }
    const int64_t layer = (int64_t)tk.dims[0] * tk.dims[1];
    std::atomic<int> next_z{0};
    fomo::run_threads(std::min(16, std::max(1, dt.nthreads_used)), [&](int) {
      for (int cz; (cz = next_z.fetch_add(1)) < tk.dims[2];) {
        int32_t t = t0;
        for (int64_t c = cz * layer; c < (cz + 1) * layer; ++c) {
          const int cx = (int)(c % tk.dims[0]), cy = (int)((c / tk.dims[0]) % tk.dims[1]);
          const fomo::pred::P3 q{tk.o[0] + (cx + 0.5) * h, tk.o[1] + (cy + 0.5) * h, tk.o[2] + (cz + 0.5) * h};
          for (int64_t steps = 0; steps < 10000000; ++steps) {
            const fomo::Tet& u = T[t];
            int inf = -1;
            for (int m = 0; m < 4; ++m)
              if (u.v[m] < 0) inf = m;
            if (inf >= 0) {  // outside the hull: take the finite tet across the hull face
              t = u.n[inf];
              break;
            }
            int next = -1;
            for (int f = 0; f < 4 && next < 0; ++f)
              if (fomo::pred::orient(dt.P[u.v[fomo::kFace[f][0]]], dt.P[u.v[fomo::kFace[f][1]]],
                                     dt.P[u.v[fomo::kFace[f][2]]], q) < 0)
                next = u.n[f];
            if (next < 0) break;
            t = next;
          }
          start[c] = t;
        }
      }
    });
  }
  const auto t1 = std::chrono::steady_clock::now();
  std::vector<int4, fomo::NoInitAlloc<int4>> hv(T.size()), hn(T.size());
  fomo::run_threads(std::min(16, dt.nthreads_used), [&](int th) {
    const size_t nt = std::min(16, dt.nthreads_used);
    for (size_t q = T.size() * th / nt; q < T.size() * (th + 1) / nt; ++q) {
      int32_t vv[4];
      for (int m = 0; m < 4; ++m) vv[m] = T[q].v[m] < 0 ? -1 : sorted_of[T[q].v[m]];
      hv[q] = make_int4(vv[0], vv[1], vv[2], vv[3]);
      hn[q] = make_int4(T[q].n[0], T[q].n[1], T[q].n[2], T[q].n[3]);
    }
  });
  DevBuf tv, tn;
  if (tv.alloc(sizeof(int4) * hv.size()) || tn.alloc(sizeof(int4) * hn.size()) ||
      ds.trec.alloc(kTetRec * sizeof(float4) * hv.size()) || ds.tstart.alloc(sizeof(int32_t) * ng) ||
      ds.tetstats.ensure(2 * sizeof(unsigned long long))) {
    ds.trec.release();
    ds.tstart.release();
    return fail(FOMO_ERR_NOMEM, "device allocation failed (tets)");
  }
  HIPOK(hipMemsetAsync(ds.tetstats.p, 0, 2 * sizeof(unsigned long long), s));
  HIPOK(hipMemcpyAsync(tv.p, hv.data(), sizeof(int4) * hv.size(), hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(tn.p, hn.data(), sizeof(int4) * hn.size(), hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ds.tstart.p, start.data(), sizeof(int32_t) * ng, hipMemcpyHostToDevice, s));
  k_tet_records<<<nblk((int64_t)hv.size()), kBlock, 0, s>>>(tv.as<int4>(), tn.as<int4>(), ds.pts.as<float4>(),
                                                             (int64_t)hv.size(), ds.trec.as<float4>());
  HIPOK(hipGetLastError());
  if (int rc = sync_stream()) return rc;
  ds.tk = tk;
  ds.ntets = (int64_t)T.size();
  const auto t2 = std::chrono::steady_clock::now();
  ds.tet_build_s = std::chrono::duration<double>(t2 - t0).count();
  if (const char* pe = std::getenv("FOMO_PROGRESS"); pe && pe[0] == '1')
    std::fprintf(stderr, "tets: Delaunay %.2f s, start grid %.2f s, conversion and upload %.2f s\n",
                 std::chrono::duration<double>(tdt - t0).count(), std::chrono::duration<double>(t1 - tdt).count(),
                 std::chrono::duration<double>(t2 - t1).count());
  return FOMO_OK;
}

// Build the candidate lookup grid of a dataset (after its cell list). subdiv lookup cells per
// cell-list edge; 0 disables the fast path. The grid extends kLookupMargin cell-list cells beyond
// the data bbox (so targets just outside the data, within a search box of them, take the fast
// path too) and is shifted by 0.3 lookup cell off the cell lattice.
// Far groups: candidate lists for the lookup cells left to the exact query (cnt = -1), one per
// group of kFarGroup^3 cell-list cells, appended to lcand after its first nnear entries. Optional:
// without the memory for it, or for groups with long lists, the cells keep the exact query.
int build_far_groups(Dataset& ds, LookupInfo& lk, int subdiv, int64_t ncl, int32_t* cnt, int64_t nnear) {
  if (ds.bvh.nodes <= 0 || ds.bvh.nodes >= (int64_t)INT32_MAX) return FOMO_OK;
  hipStream_t s = g_ctx.stream;
  FarGrid fg{};
  fg.gcell[0] = kFarGroup * std::max(1, subdiv / 2);  // lookup cells per cell-list cell along x
  fg.gcell[1] = fg.gcell[2] = kFarGroup * subdiv;
  fg.clip = 1;
  // outer groups: twice the edge, over a fifth of the grid's extent beyond each face
  FarGrid og{};
  int64_t ng = 1, nb = 1, ong = 1, onb = 1;
  for (int a = 0; a < 3; ++a) {
    fg.gdim[a] = (lk.dims[a] + fg.gcell[a] - 1) / fg.gcell[a];
    fg.bdim[a] = (fg.gdim[a] + 3) / 4;
    ng *= fg.gdim[a];
    nb *= fg.bdim[a];
    og.gcell[a] = 2 * fg.gcell[a];
    const int across = (lk.dims[a] + og.gcell[a] - 1) / og.gcell[a];
    const int ext = lk.dims[a] > 1 ? (across + 4) / 5 : 0;
    og.ioff[a] = -ext * og.gcell[a];
    og.gdim[a] = across + 2 * ext;
    og.bdim[a] = (og.gdim[a] + 3) / 4;
    ong *= og.gdim[a];
    onb *= og.bdim[a];
  }
  const bool outer = ong <= ng;  // the outer lattice is coarser: about a third of the groups
  const int64_t chunk = std::min<int64_t>(std::max(nb, onb), 2048);  // blocks per launch (scratch 512 MiB)
  const int64_t fcap = std::min<int64_t>(32 * (ng + (outer ? ong : 0)), (int64_t)INT32_MAX - 1 - nnear);
  if (fcap <= 0) return FOMO_OK;
  DevBuf gflag, gown, ginfo, fcnt, fpool, cbuf, oflag, oinfo;
  if (gflag.alloc(ng) || gown.alloc(ng) || ginfo.alloc(sizeof(int2) * ng) || fcnt.alloc(sizeof(unsigned long long)) ||
      fpool.alloc(sizeof(float4) * fcap) || cbuf.alloc(sizeof(float4) * chunk * kFarBlock * kFarCap)) {
    (void)hipGetLastError();
    return FOMO_OK;
  }
  const bool do_outer = outer && !oflag.alloc(ong) && !oinfo.alloc(sizeof(int2) * ong);
  if (!do_outer) (void)hipGetLastError();
  HIPOK(hipMemsetAsync(gflag.p, 0, ng, s));
  HIPOK(hipMemsetAsync(gown.p, 0, ng, s));
  HIPOK(hipMemsetAsync(ginfo.p, 0, sizeof(int2) * ng, s));
  HIPOK(hipMemsetAsync(fcnt.p, 0, sizeof(unsigned long long), s));
  k_far_flags<<<nblk(ncl), kBlock, 0, s>>>(cnt, ncl, lk, fg, gflag.as<uint8_t>(), gown.as<uint8_t>());
  HIPOK(hipGetLastError());
  const auto groups = [&](const FarGrid& G, int64_t nblocks, DevBuf& flag, DevBuf& info) {
    for (int64_t b0 = 0; b0 < nblocks; b0 += chunk)
      k_far_groups<<<(unsigned)std::min(chunk, nblocks - b0), kFarBlock, 0, s>>>(
          ds.pts.as<float4>(), ds.mpts.as<float4>(), ds.nlo.as<float4>(), ds.nhi.as<float4>(), (int)ds.bvh.nodes, lk,
          G, b0, flag.as<uint8_t>(), cbuf.as<float4>(), fpool.as<float4>(), fcnt.as<unsigned long long>(), fcap,
          info.as<int2>());
  };
  {
    ProfScope ps("k_far_groups", s);
    groups(fg, nb, gflag, ginfo);
    if (do_outer) {
      HIPOK(hipMemsetAsync(oinfo.p, 0, sizeof(int2) * ong, s));
      k_outer_flags<<<nblk(ong), kBlock, 0, s>>>(lk, og, ong, oflag.as<uint8_t>());
      groups(og, onb, oflag, oinfo);
    }
  }
  HIPOK(hipGetLastError());
  unsigned long long used = 0;
  HIPOK(hipMemcpyAsync(&used, fcnt.p, sizeof used, hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  const int64_t nfar = std::min<int64_t>((int64_t)used, fcap);  // blocks past fcap were not written
  if (nfar == 0) return FOMO_OK;
  DevBuf lc;
  if (lc.alloc(sizeof(float4) * (nnear + nfar)) || (do_outer && ds.ohdr.alloc(sizeof(int2) * ong))) {
    (void)hipGetLastError();
    ds.ohdr.release();
    return FOMO_OK;
  }
  if (nnear) HIPOK(hipMemcpyAsync(lc.p, ds.lcand.p, sizeof(float4) * nnear, hipMemcpyDeviceToDevice, s));
  HIPOK(hipMemcpyAsync(lc.as<float4>() + nnear, fpool.p, sizeof(float4) * nfar, hipMemcpyDeviceToDevice, s));
  k_far_headers<<<nblk(ncl), kBlock, 0, s>>>(cnt, ncl, lk, fg, ginfo.as<int2>(), nnear, ds.lslot.as<float4>());
  if (do_outer)
    k_outer_headers<<<nblk(ong), kBlock, 0, s>>>(oinfo.as<int2>(), ong, nnear, ds.ohdr.as<int2>());
  // the group table, kept only when some group's cells all take its list
  unsigned long long nfull = 0;
  if (!ds.ghdr.alloc(sizeof(int2) * ng)) {
    HIPOK(hipMemsetAsync(fcnt.p, 0, sizeof(unsigned long long), s));
    k_group_headers<<<nblk(ng), kBlock, 0, s>>>(ginfo.as<int2>(), gown.as<uint8_t>(), ng, nnear, ds.ghdr.as<int2>(),
                                                fcnt.as<unsigned long long>());
    HIPOK(hipMemcpyAsync(&nfull, fcnt.p, sizeof nfull, hipMemcpyDeviceToHost, s));
  } else {
    (void)hipGetLastError();
  }
  HIPOK(hipGetLastError());
  // group counts for fomo_dataset_lookup_groups
  {
    std::vector<int2> hi((size_t)ng), ho(do_outer ? (size_t)ong : 0);
    HIPOK(hipMemcpyAsync(hi.data(), ginfo.p, sizeof(int2) * ng, hipMemcpyDeviceToHost, s));
    if (do_outer) HIPOK(hipMemcpyAsync(ho.data(), oinfo.p, sizeof(int2) * ong, hipMemcpyDeviceToHost, s));
    if (int rc = sync_stream()) return rc;
    ds.far_groups = ng;
    ds.far_lists = std::count_if(hi.begin(), hi.end(), [](const int2& g) { return g.y > 0; });
    ds.outer_groups = do_outer ? ong : 0;
    ds.outer_lists = std::count_if(ho.begin(), ho.end(), [](const int2& g) { return g.y > 0; });
  }
  if (int rc = sync_stream()) return rc;
  if (nfull == 0) ds.ghdr.release();
  ds.full_groups = (int64_t)nfull;
  std::swap(ds.lcand.p, lc.p);
  std::swap(ds.lcand.bytes, lc.bytes);
  if (do_outer) {
    for (int a = 0; a < 3; ++a) {
      lk.oo[a] = lk.o[a] + (double)og.ioff[a] * lk.h[a];
      lk.oinvh[a] = 1.0 / ((double)og.gcell[a] * lk.h[a]);
      lk.odims[a] = og.gdim[a];
    }
    lk.ohdr = ds.ohdr.as<int2>();
  }
  if (ds.ghdr.p) {
    for (int a = 0; a < 3; ++a) {
      lk.gcell[a] = fg.gcell[a];
      lk.gdims[a] = fg.gdim[a];
    }
    lk.ghdr = ds.ghdr.as<int2>();
  }
  return FOMO_OK;
}

int build_lookup(Dataset& ds) {
  ds.lk = LookupInfo{};
  ds.lslot.release();
  ds.ohdr.release();
  ds.ghdr.release();
  ds.lcand.release();
  ds.ncand = ds.nlcells = ds.nfallback = 0;
  ds.far_groups = ds.far_lists = ds.full_groups = ds.outer_groups = ds.outer_lists = 0;
  if (ds.lookup_subdiv <= 0) return FOMO_OK;
  if (ds.rect || ds.mrect) {
    LookupInfo& lk = ds.lk;
    lk.rect = ds.rect ? 1 : 0;
    lk.mrect = ds.mrect ? 1 : 0;
    lk.gpos = ds.mrect ? ds.gpos.as<int32_t>() : nullptr;
    for (int a = 0; a < 3; ++a) {
      lk.rn[a] = ds.rect_n[a];
      lk.ro[a] = ds.rect_o[a];
      lk.rinv[a] = ds.rect_inv[a];
    }
    lk.rdmax = fomo::rect_dmax(ds.rect_gap2);
    lk.rv = ds.rv.as<float>();
    lk.rbin = ds.rbin.as<float4>();
    if (ds.rect) return FOMO_OK;
  }
  hipStream_t s = g_ctx.stream;
  const GridInfo& g = ds.grid;
  // subdivision: the requested one, reduced until the grid fits ~24 cells per point and 2^30.
  // Cell edges: g.h / subdiv, and along x twice that (as long as it stays within g.h), so that a
  // 128-byte record (8 slots) holds the candidates of structured data without overflow.
  const auto edges = [&](int sd, double (&h)[3]) {
    for (int a = 0; a < 3; ++a) h[a] = g.h / sd;
    h[0] = g.h / std::max(1, sd / 2);
  };
  int subdiv = ds.lookup_subdiv;
  for (; subdiv >= 1; --subdiv) {
    double h[3], ncl_est = 1;
    edges(subdiv, h);
    for (int a = 0; a < 3; ++a) {
      const double mrg = (g.dims[a] > 1 ? kLookupMargin : 0) * g.h;
      ncl_est *= std::floor(((double)g.bmax[a] - g.o[a] + 2 * mrg + 0.3 * h[a]) / h[a]) + 1;
    }
    if (ncl_est <= 24.0 * ds.n + 65536 && ncl_est <= (double)(1LL << 30)) break;
  }
  if (subdiv < 1) return FOMO_OK;
  LookupInfo lk = ds.lk;  // a masked grid's axes (set above) stay; the rest is zero
  edges(subdiv, lk.h);
  double ext = 0, mag = 0;
  int64_t ncl = 1;
  for (int a = 0; a < 3; ++a) {
    lk.invh[a] = 1.0 / lk.h[a];
    const double mrg = (g.dims[a] > 1 ? kLookupMargin : 0) * g.h;
    lk.o[a] = g.o[a] - mrg - 0.3 * lk.h[a];
    const double e = (double)g.bmax[a] - g.o[a];
    lk.dims[a] = (int)std::floor((e + 2 * mrg + 0.3 * lk.h[a]) / lk.h[a]) + 1;
    ncl *= lk.dims[a];
    ext = std::max(ext, e);
    mag = std::max(mag, std::max(std::fabs((double)g.bmin[a]), std::fabs((double)g.bmax[a])));
  }
  // keep the lookup grid within ~24 cells per point and 2^30 cells
  if (ncl > 24 * ds.n + 65536 || ncl > (1LL << 30)) return FOMO_OK;
  lk.eps = 1e-7 * lk.h[0] + 1e-12 * (mag + ext);
  lk.enabled = 1;
  // per cell of the cell list: Chebyshev distance to the nearest non-empty cell, capped
  DevBuf near0, near1;
  if (near0.alloc(ds.ncells) || near1.alloc(ds.ncells)) return FOMO_OK;
  k_occupancy<<<nblk(ds.ncells), kBlock, 0, s>>>(ds.cstart.as<int32_t>(), ds.ncells, kLookupDistCap, near0.as<uint8_t>());
  k_dist_axis<<<nblk(ds.ncells), kBlock, 0, s>>>(near0.as<uint8_t>(), near1.as<uint8_t>(), g, 0, kLookupDistCap);
  k_dist_axis<<<nblk(ds.ncells), kBlock, 0, s>>>(near1.as<uint8_t>(), near0.as<uint8_t>(), g, 1, kLookupDistCap);
  k_dist_axis<<<nblk(ds.ncells), kBlock, 0, s>>>(near0.as<uint8_t>(), near1.as<uint8_t>(), g, 2, kLookupDistCap);
  HIPOK(hipGetLastError());
  DevBuf cnt, pos, off;
  if (cnt.alloc(sizeof(int32_t) * ncl) || pos.alloc(sizeof(int32_t) * ncl) || off.alloc(sizeof(int64_t) * ncl))
    return FOMO_OK;  // not enough memory: ring search only
  if (ds.lslot.alloc(sizeof(float4) * kRecSlots * ncl)) return FOMO_OK;
  size_t tmpb = 0;
  HIPOK(rocprim::exclusive_scan(nullptr, tmpb, pos.as<int32_t>(), off.as<int64_t>(), (int64_t)0, (size_t)ncl,
                                rocprim::plus<int64_t>(), s));
  HIPOK(g_ctx.sorttmp.ensure(tmpb));
  // One pass, overflow lists bump-allocated from a pool with room for one cell in 8; if the pool
  // cannot be had or runs out, the two-pass build (count, offsets, fill).
  bool done = false;
  int64_t nnear = 0;  // entries of lcand in use
  {
    const int64_t pool_cap = 3 * ncl;  // float4 entries: ncl / 8 blocks of kLookupCap
    DevBuf pool, used;
    if (pool_cap < (int64_t)INT32_MAX && !pool.alloc(sizeof(float4) * pool_cap) &&
        !used.alloc(sizeof(unsigned long long))) {
      HIPOK(hipMemsetAsync(used.p, 0, sizeof(unsigned long long), s));
      {
        ProfScope ps("k_lookup_build", s);
        k_lookup_build<2><<<nblk(ncl), kBlock, 0, s>>>(ds.pts.as<float4>(), ds.cstart.as<int32_t>(), g, lk, ncl,
                                                        near1.as<uint8_t>(), cnt.as<int32_t>(), nullptr,
                                                        pool.as<float4>(), ds.lslot.as<float4>(),
                                                        used.as<unsigned long long>(), pool_cap);
      }
      HIPOK(hipGetLastError());
      unsigned long long pool_used = 0;
      HIPOK(hipMemcpyAsync(&pool_used, used.p, sizeof pool_used, hipMemcpyDeviceToHost, s));
      if (int rc = sync_stream()) return rc;
      if ((int64_t)pool_used <= pool_cap && !ds.lcand.alloc(sizeof(float4) * std::max<int64_t>((int64_t)pool_used, 1))) {
        if (pool_used)
          HIPOK(hipMemcpyAsync(ds.lcand.p, pool.p, sizeof(float4) * pool_used, hipMemcpyDeviceToDevice, s));
        nnear = (int64_t)pool_used;
        done = true;
      }
    }
  }
  if (!done) {
    {
      ProfScope ps("k_lookup_build", s);
      k_lookup_build<0><<<nblk(ncl), kBlock, 0, s>>>(ds.pts.as<float4>(), ds.cstart.as<int32_t>(), g, lk, ncl,
                                                      near1.as<uint8_t>(), cnt.as<int32_t>(), nullptr, nullptr,
                                                      nullptr, nullptr, 0);
    }
    // the overflow lists' offsets
    k_overflow_counts<<<nblk(ncl), kBlock, 0, s>>>(cnt.as<int32_t>(), ncl, pos.as<int32_t>());
    HIPOK(hipGetLastError());
    HIPOK(rocprim::exclusive_scan(g_ctx.sorttmp.p, tmpb, pos.as<int32_t>(), off.as<int64_t>(), (int64_t)0,
                                  (size_t)ncl, rocprim::plus<int64_t>(), s));
    int64_t last_off = 0;
    int32_t last_cnt = 0;
    HIPOK(hipMemcpyAsync(&last_off, off.as<int64_t>() + ncl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPOK(hipMemcpyAsync(&last_cnt, pos.as<int32_t>() + ncl - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (int rc = sync_stream()) return rc;
    const int64_t total = last_off + last_cnt;
    if (total >= (int64_t)INT32_MAX || ds.lcand.alloc(sizeof(float4) * std::max<int64_t>(total, 1))) {
      ds.lslot.release();
      ds.lcand.release();
      return FOMO_OK;
    }
    k_lookup_build<1><<<nblk(ncl), kBlock, 0, s>>>(ds.pts.as<float4>(), ds.cstart.as<int32_t>(), g, lk, ncl,
                                                    near1.as<uint8_t>(), cnt.as<int32_t>(), off.as<int64_t>(),
                                                    ds.lcand.as<float4>(), ds.lslot.as<float4>(), nullptr, 0);
    k_lookup_headers<<<nblk(ncl), kBlock, 0, s>>>(cnt.as<int32_t>(), off.as<int64_t>(), ncl, ds.lslot.as<float4>());
    HIPOK(hipGetLastError());
    nnear = total;
  }
  if (int rc = build_far_groups(ds, lk, subdiv, ncl, cnt.as<int32_t>(), nnear)) return rc;
  // candidates in all (statistics)
  int64_t ntotal = 0;
  {
    k_clamp_counts<<<nblk(ncl), kBlock, 0, s>>>(cnt.as<int32_t>(), ncl, pos.as<int32_t>());
    HIPOK(hipGetLastError());
    HIPOK(rocprim::exclusive_scan(g_ctx.sorttmp.p, tmpb, pos.as<int32_t>(), off.as<int64_t>(), (int64_t)0,
                                  (size_t)ncl, rocprim::plus<int64_t>(), s));
    int64_t last_off = 0;
    int32_t last_cnt = 0;
    HIPOK(hipMemcpyAsync(&last_off, off.as<int64_t>() + ncl - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPOK(hipMemcpyAsync(&last_cnt, pos.as<int32_t>() + ncl - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (int rc = sync_stream()) return rc;
    ntotal = last_off + last_cnt;
  }
  int32_t nfb = 0;
  // number of cells left to the ring search
  int* dfb = reinterpret_cast<int*>(pos.p);
  HIPOK(hipMemsetAsync(dfb, 0, sizeof(int), s));
  k_count_negative<<<nblk(ncl), kBlock, 0, s>>>(cnt.as<int32_t>(), ncl, dfb);
  HIPOK(hipMemcpyAsync(&nfb, dfb, sizeof(int), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  ds.nfallback = nfb;
  ds.lookup_active_subdiv = subdiv;
  ds.lk = lk;
  ds.ncand = ntotal;
  ds.nlcells = ncl;
  return FOMO_OK;
}

// Rectilinear detection. The distinct values of each axis (device sort + unique); when their
// counts multiply to n, every point's grid position (k_rect_map). On success *ok is set,
// ds.perm[gid] holds the original index of the point at grid position gid (the sorted order is
// the grid order), and the axis values and bin tables are on the device. Otherwise the dataset is
// indexed as scattered points.
int build_rect(Dataset& ds, bool* ok) {
  *ok = false;
  ds.rect = false;
  ds.mrect = false;
  ds.rv.release();
  ds.rbin.release();
  ds.corners.release();
  ds.gpos.release();
  const int64_t n = ds.n;
  if (!ds.allow_rect || n < 2) return FOMO_OK;
  hipStream_t s = g_ctx.stream;
  HIPOK(g_ctx.sortkeys[0].ensure(sizeof(uint32_t) * n));
  HIPOK(g_ctx.sortkeys[1].ensure(sizeof(uint32_t) * n));
  float* sorted = g_ctx.sortkeys[0].as<float>();
  float* uniq = g_ctx.sortkeys[1].as<float>();
  DevBuf dcount;
  if (dcount.alloc(sizeof(unsigned int))) return fail(FOMO_ERR_NOMEM, "device allocation failed (rectilinear test)");
  std::vector<float> vals[3];
  int64_t prod = 1;
  for (int a = 0; a < 3; ++a) {
    size_t tb = 0, tu = 0;
    HIPOK(rocprim::radix_sort_keys(nullptr, tb, ds.raw[a].as<float>(), sorted, (size_t)n, 0, 32, s));
    HIPOK(rocprim::unique(nullptr, tu, sorted, uniq, dcount.as<unsigned int>(), (size_t)n,
                          rocprim::equal_to<float>(), s));
    HIPOK(g_ctx.sorttmp.ensure(std::max(tb, tu)));
    HIPOK(rocprim::radix_sort_keys(g_ctx.sorttmp.p, tb, ds.raw[a].as<float>(), sorted, (size_t)n, 0, 32, s));
    HIPOK(rocprim::unique(g_ctx.sorttmp.p, tu, sorted, uniq, dcount.as<unsigned int>(), (size_t)n,
                          rocprim::equal_to<float>(), s));
    unsigned int na = 0;
    HIPOK(hipMemcpyAsync(&na, dcount.p, sizeof(na), hipMemcpyDeviceToHost, s));
    if (int rc = sync_stream()) return rc;
    prod *= (int64_t)na;
    // a full grid (prod == n) or a masked one, at least a quarter full (prod <= 4 n)
    if (na == 0 || prod > kMaskedGridMaxRatio * n || prod >= (int64_t)INT32_MAX) return FOMO_OK;
    vals[a].resize(na);
    HIPOK(hipMemcpyAsync(vals[a].data(), uniq, sizeof(float) * na, hipMemcpyDeviceToHost, s));
    if (int rc = sync_stream()) return rc;
    for (float f : vals[a])
      if (!std::isfinite(f)) return FOMO_OK;
  }
  const bool masked = prod != n;
  // axis tables
  int64_t nv = 0, nbt = 0;
  for (int a = 0; a < 3; ++a) {
    ds.rect_n[a] = (int)vals[a].size();
    nv += ds.rect_n[a];
    nbt += fomo::kRectBinsPerValue * ds.rect_n[a];
  }
  std::vector<float> hv;
  std::vector<float4> hb(nbt);
  hv.reserve(nv);
  int64_t boff = 0;
  for (int a = 0; a < 3; ++a) {
    hv.insert(hv.end(), vals[a].begin(), vals[a].end());
    fomo::rect_axis_table(vals[a].data(), ds.rect_n[a], hb.data() + boff, &ds.rect_o[a], &ds.rect_inv[a],
                          &ds.rect_gap2[a]);
    boff += fomo::kRectBinsPerValue * ds.rect_n[a];
  }
  // the 8 corners of the grid: a view's rotated coordinate ((x c0 - y c1) - z c2 and the like, in
  // the kernels' FP64 order) is monotone in each coordinate, since rounding is monotone, so its
  // smallest and largest values over the grid are taken at corners
  std::vector<float4> hc(8);
  for (int c = 0; c < 8; ++c)
    hc[c] = make_float4((c & 1) ? vals[0].back() : vals[0].front(), (c & 2) ? vals[1].back() : vals[1].front(),
                        (c & 4) ? vals[2].back() : vals[2].front(), 0.f);
  if (ds.rv.alloc(sizeof(float) * nv) || ds.rbin.alloc(sizeof(float4) * nbt) || ds.corners.alloc(sizeof(float4) * 8))
    return fail(FOMO_ERR_NOMEM, "device allocation failed (rectilinear axes)");
  HIPOK(hipMemcpyAsync(ds.rv.p, hv.data(), sizeof(float) * nv, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ds.rbin.p, hb.data(), sizeof(float4) * nbt, hipMemcpyHostToDevice, s));
  HIPOK(hipMemcpyAsync(ds.corners.p, hc.data(), sizeof(float4) * 8, hipMemcpyHostToDevice, s));
  // grid positions: every point on an axis value, no two on one position. A full grid's map is
  // the sorted order itself (perm); a masked grid's (original indices for now) goes to gpos
  int32_t* gmap = nullptr;
  if (masked) {
    if (ds.gpos.alloc(sizeof(int32_t) * prod)) return FOMO_OK;  // no room: index as scattered points only
    gmap = ds.gpos.as<int32_t>();
  } else {
    HIPOK(ds.perm.ensure(sizeof(int32_t) * n));
    gmap = ds.perm.as<int32_t>();
  }
  HIPOK(hipMemsetAsync(gmap, 0xff, sizeof(int32_t) * prod, s));
  int* bad = reinterpret_cast<int*>(dcount.p);
  HIPOK(hipMemsetAsync(bad, 0, sizeof(int), s));
  const float* v = ds.rv.as<float>();
  k_rect_map<<<nblk(n), kBlock, 0, s>>>(ds.raw[0].as<float>(), ds.raw[1].as<float>(), ds.raw[2].as<float>(), n, v,
                                        v + ds.rect_n[0], v + ds.rect_n[0] + ds.rect_n[1], ds.rect_n[0],
                                        ds.rect_n[1], ds.rect_n[2], gmap, bad);
  HIPOK(hipGetLastError());
  int hbad = 1;
  HIPOK(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  if (hbad) {
    ds.rv.release();
    ds.rbin.release();
    ds.corners.release();
    ds.gpos.release();
    return FOMO_OK;
  }
  if (masked) {
    ds.mrect = true;  // gpos becomes sorted positions once the cell list has sorted the points
    ds.corners.release();
    return FOMO_OK;
  }
  ds.rect = true;
  *ok = true;
  return FOMO_OK;
}

// gpos[g] = original index -> sorted position (inv[original] = sorted position), -1 kept.
__global__ void k_gpos_sorted(int32_t* __restrict__ gpos, int64_t ng, const int32_t* __restrict__ inv) {
  for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const int32_t o = gpos[g];
    if (o >= 0) gpos[g] = inv[o];
  }
}

// Cell list of a scattered dataset: points sorted by the key of their cell of ds.grid (x fastest),
// ds.perm = sorted position -> original index, ds.cstart = first sorted position of every cell.
int build_cell_list(Dataset& ds) {
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds.n;
  const GridInfo& g = ds.grid;
  // keys
  HIPOK(g_ctx.sortkeys[0].ensure(sizeof(uint32_t) * n));
  HIPOK(g_ctx.sortkeys[1].ensure(sizeof(uint32_t) * n));
  HIPOK(g_ctx.sortvals.ensure(sizeof(int32_t) * n));
  HIPOK(ds.perm.ensure(sizeof(int32_t) * n));
  {
    ProfScope ps("k_cell_keys", s);
    k_cell_keys<<<nblk(n), kBlock, 0, s>>>(ds.raw[0].as<float>(), ds.raw[1].as<float>(), ds.raw[2].as<float>(), n, g,
                                           g_ctx.sortkeys[0].as<uint32_t>(), g_ctx.sortvals.as<int32_t>());
  }
  HIPOK(hipGetLastError());
  int endbit = 1;
  while (endbit < 32 && ((int64_t)1 << endbit) < ds.ncells) endbit++;
  size_t tmpb = 0;
  HIPOK(rocprim::radix_sort_pairs(nullptr, tmpb, g_ctx.sortkeys[0].as<uint32_t>(), g_ctx.sortkeys[1].as<uint32_t>(),
                                  g_ctx.sortvals.as<int32_t>(), ds.perm.as<int32_t>(), (size_t)n, 0, endbit, s));
  HIPOK(g_ctx.sorttmp.ensure(tmpb));
  {
    ProfScope ps("radix_sort", s);
    HIPOK(rocprim::radix_sort_pairs(g_ctx.sorttmp.p, tmpb, g_ctx.sortkeys[0].as<uint32_t>(),
                                    g_ctx.sortkeys[1].as<uint32_t>(), g_ctx.sortvals.as<int32_t>(),
                                    ds.perm.as<int32_t>(), (size_t)n, 0, endbit, s));
  }
  HIPOK(ds.cstart.ensure(sizeof(int32_t) * (ds.ncells + 1)));
  {
    ProfScope ps("k_cell_start", s);
    k_cell_start<<<nblk(n + 1), kBlock, 0, s>>>(g_ctx.sortkeys[1].as<uint32_t>(), n, ds.ncells, ds.cstart.as<int32_t>());
  }
  return FOMO_OK;
}

// Build (or rebuild) the index of a dataset on the device: rectilinear axes when the points form
// a product grid, else the cell list and candidate lookup grid; then the octree.
int build_index(Dataset& ds, double cell_scale) {
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds.n;
  // the triangulation refers to sorted positions, which the new cell order changes
  ds.ntets = 0;
  ds.trec.release();
  ds.tstart.release();
  if (cell_scale <= 0) cell_scale = 1.0;
  ds.cell_scale = cell_scale;
  // bbox
  const unsigned nb = std::min<unsigned>(nblk(n), kMaxPartials);
  HIPOK(g_ctx.vs.partial.ensure(sizeof(double) * 6 * (kMaxPartials + 1)));
  double* part = g_ctx.vs.partial.as<double>();
  {
    ProfScope ps("k_bbox", s);
    k_bbox<<<nb, kBlock, 0, s>>>(ds.raw[0].as<float>(), ds.raw[1].as<float>(), ds.raw[2].as<float>(), n, part);
  }
  k_reduce_partials<<<1, kBlock, 0, s>>>(part, nb, part + 6 * kMaxPartials);
  HIPOK(hipGetLastError());
  double bb[6];
  HIPOK(hipMemcpyAsync(bb, part + 6 * kMaxPartials, sizeof(bb), hipMemcpyDeviceToHost, s));
  int rc = sync_stream();
  if (rc) return rc;
  for (int a = 0; a < 6; ++a) ds.bbox[a] = (float)bb[a];
  // bounding sphere about the bbox centre: search boxes wholly outside it hold no point
  const double sc[3] = {0.5 * (bb[0] + bb[1]), 0.5 * (bb[2] + bb[3]), 0.5 * (bb[4] + bb[5])};
  k_sphere_r2<<<nb, kBlock, 0, s>>>(ds.raw[0].as<float>(), ds.raw[1].as<float>(), ds.raw[2].as<float>(), n, sc[0],
                                    sc[1], sc[2], part);
  k_reduce_partials<<<1, kBlock, 0, s>>>(part, nb, part + 6 * kMaxPartials);
  HIPOK(hipGetLastError());
  double sp[6];
  HIPOK(hipMemcpyAsync(sp, part + 6 * kMaxPartials, sizeof(sp), hipMemcpyDeviceToHost, s));
  if ((rc = sync_stream())) return rc;
  // grid: cell edge h = scale * (V/N)^(1/3) over the non-degenerate axes
  double ext[3] = {bb[1] - bb[0], bb[3] - bb[2], bb[5] - bb[4]};
  double vol = 1.0;
  int nd = 0;
  double emax = 0;
  for (int a = 0; a < 3; ++a) {
    emax = std::max(emax, ext[a]);
    if (ext[a] > 0) { vol *= ext[a]; nd++; }
  }
  double h;
  if (nd == 0) h = 1.0;
  else h = cell_scale * std::pow(vol / (double)n, 1.0 / nd);
  if (!(h > 0) || !std::isfinite(h)) h = std::max(emax, 1.0);
  // cap the number of cells at ~4N + 64 (and 2^30)
  for (;;) {
    double nc = 1;
    for (int a = 0; a < 3; ++a) nc *= std::floor(ext[a] / h) + 1;
    if (nc <= 4.0 * (double)n + 64 && nc < 1073741824.0) break;
    h *= 1.25;
  }
  GridInfo& g = ds.grid;
  for (int a = 0; a < 3; ++a) {
    g.o[a] = bb[2 * a];
    g.dims[a] = (int)std::floor(ext[a] / h) + 1;
    g.bmin[a] = (float)bb[2 * a];
    g.bmax[a] = (float)bb[2 * a + 1];
  }
  g.h = h;
  g.invh = 1.0 / h;
  g.slack = 1e-6 * h;
  for (int a = 0; a < 3; ++a) g.sc[a] = sc[a];
  g.sr2 = sp[1] * (1.0 + 1e-9) + 1e-300;  // rounded up: the FP64 distances are within 1e-15 of exact
  ds.ncells = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
  bool rect = false;
  if (int rc2 = build_rect(ds, &rect)) return rc2;
  if (rect) {
    // no cell list: the sorted order is the grid order (ds.perm from build_rect)
    ds.ncells = 0;
    ds.cstart.release();
  } else if (int rc2 = build_cell_list(ds)) {
    return rc2;
  }
  if (ds.mrect) {
    // the masked grid's map: original indices -> sorted positions of the cell order
    DevBuf inv;
    if (inv.alloc(sizeof(int32_t) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (masked grid)");
    k_invert_perm<<<nblk(n), kBlock, 0, s>>>(ds.perm.as<int32_t>(), n, inv.as<int32_t>());
    const int64_t ng = (int64_t)ds.rect_n[0] * ds.rect_n[1] * ds.rect_n[2];
    k_gpos_sorted<<<std::min<int64_t>(nblk(ng), 65536), kBlock, 0, s>>>(ds.gpos.as<int32_t>(), ng, inv.as<int32_t>());
    HIPOK(hipGetLastError());
    if (int rc2 = sync_stream()) return rc2;
  }
  HIPOK(ds.pts.ensure(sizeof(float4) * n));
  k_permute_points<<<nblk(n), kBlock, 0, s>>>(ds.raw[0].as<float>(), ds.raw[1].as<float>(), ds.raw[2].as<float>(),
                                              ds.perm.as<int32_t>(), n, ds.pts.as<float4>());
  ds.svars.clear();
  for (int v = 0; v < ds.nvars; ++v) {
    ds.svars.emplace_back(new DevBuf());
    HIPOK(ds.svars.back()->alloc(sizeof(float) * n));
    k_permute_var<<<nblk(n), kBlock, 0, s>>>(ds.rawvars[v]->as<float>(), ds.perm.as<int32_t>(), n,
                                             ds.svars.back()->as<float>());
  }
  if (ds.nvars >= 5) {
    HIPOK(ds.vel.ensure(sizeof(float4) * n));
    k_pack_vel<<<nblk(n), kBlock, 0, s>>>(ds.rawvars[2]->as<float>(), ds.rawvars[3]->as<float>(),
                                          ds.rawvars[4]->as<float>(), ds.perm.as<int32_t>(), n, ds.vel.as<float4>());
  }
  HIPOK(hipGetLastError());
  // maximum cell occupancy (one point per grid position on rectilinear datasets)
  int hmx = 1;
  if (!rect) {
    int* mx = reinterpret_cast<int*>(part);
    HIPOK(hipMemsetAsync(mx, 0, sizeof(int), s));
    k_cell_maxcount<<<nblk(ds.ncells), kBlock, 0, s>>>(ds.cstart.as<int32_t>(), ds.ncells, mx);
    HIPOK(hipGetLastError());
    HIPOK(hipMemcpyAsync(&hmx, mx, sizeof(int), hipMemcpyDeviceToHost, s));
  }
  rc = sync_stream();
  if (rc) return rc;
  ds.maxcount = hmx;
  if (int rc2 = build_bvh(ds)) return rc2;
  return build_lookup(ds);
}

template <class T> T* find_obj(std::map<int, std::unique_ptr<T>>& m, int id) {
  auto it = m.find(id);
  return it == m.end() ? nullptr : it->second.get();
}

// Enqueue the per-view rotate + finalize kernels (mode NN or Thomson) on stream s.
// Host trig of a view (glibc): cos b, sin b, cos l, sin l and the Thomson blend of the search
// half-width, 1/(1+exp(-10(|b|-0.1745))) (fomo-thomson.cpp:249).
void view_trig(double l, double b, double t[5]) {
  t[0] = cos(b), t[1] = sin(b), t[2] = cos(l), t[3] = sin(l);
  t[4] = 1.0 / (1.0 + exp(-10 * (std::fabs(b) - 0.1745)));
}

// clear2: a fused pass's direct-pass list to clear (or null).
void enqueue_view_finalize(hipStream_t s, int mode, double l, double b, int nx, int ny, int nz, int dn,
                           const double* partial, unsigned nb, ViewGeom* geom, int32_t* clear2 = nullptr) {
  double t[5];
  view_trig(l, b, t);
  k_view_finalize<<<1, kBlock, 0, s>>>(partial, nb, mode, nx, ny, nz, dn, t[0], t[1], t[2], t[3], t[4], geom,
                                       clear2);
}

// Thomson: with em, aux receives the points' pB terms (n doubles), else sin^2 chi (n floats).
void enqueue_view_geom(hipStream_t s, const Dataset& ds, int mode, double l, double b, int nx, int ny, int nz, int dn,
                       float* aux, double* partial, ViewGeom* geom, int32_t* clear2 = nullptr,
                       const float2* em = nullptr) {
  unsigned nb = std::min<unsigned>(nblk(ds.n), kMaxPartials);
  const double cb = cos(b), sb = sin(b), cl = cos(l), sl = sin(l);
  if (mode == MODE_NN) {
    ProfScope ps("k_view_rotate_nn", s);
    (void)aux;
    // a rectilinear dataset's extents are those of its 8 corners
    const int64_t np = ds.rect ? 8 : ds.n;
    if (ds.rect) nb = 1;
    k_view_rotate_nn<<<nb, kBlock, 0, s>>>(ds.rect ? ds.corners.as<float4>() : ds.pts.as<float4>(), nullptr, np, cb,
                                           sb, cl, sl, nullptr, partial);
  } else {
    const double eps = 1e-6;  // fomo-thomson.cpp:139-152
    double cbc = cb, sbc = sb;
    if (std::abs(cbc) < eps) cbc = eps;
    if (std::abs(sbc) < eps) sbc = eps;
    const double blend = 1.0 / (1.0 + exp(-20.0 * (std::fabs(b) - 1.309)));  // :173-176
    ProfScope ps("k_view_rotate_thomson", s);
    k_view_rotate_thomson<<<nb, kBlock, 0, s>>>(ds.pts.as<float4>(), ds.n, cbc, sbc, cl, sl, blend, em, aux,
                                                reinterpret_cast<double*>(aux), partial);
  }
  enqueue_view_finalize(s, mode, l, b, nx, ny, nz, dn, partial, nb, geom, clear2);
}

// Extents of many NN views (a batch), kViewsPerPass views per pass over the points; partial of
// view v at partial + v * 6 * kMaxPartials. Returns the number of partials per view.
unsigned enqueue_views_extents(hipStream_t s, const Dataset& ds, const std::vector<double>& l,
                               const std::vector<double>& b, double* partial) {
  // a rectilinear dataset's extents are those of its 8 corners
  const float4* P = ds.rect ? ds.corners.as<float4>() : ds.pts.as<float4>();
  const int64_t np = ds.rect ? 8 : ds.n;
  const unsigned nb = std::min<unsigned>(nblk(np), kMaxPartials);
  const int nv = (int)l.size();
  for (int v0 = 0; v0 < nv; v0 += kViewsPerPass) {
    ViewTrigs t{};
    const int m = std::min(kViewsPerPass, nv - v0);
    for (int w = 0; w < m; ++w) {
      t.cb[w] = cos(b[v0 + w]), t.sb[w] = sin(b[v0 + w]);
      t.cl[w] = cos(l[v0 + w]), t.sl[w] = sin(l[v0 + w]);
    }
    bool sameb = true;
    for (int w = 1; w < m; ++w) sameb = sameb && t.cb[w] == t.cb[0] && t.sb[w] == t.sb[0];
    ProfScope ps("k_view_extents_multi", s);
    double* const pv = partial + (int64_t)v0 * 6 * kMaxPartials;
    if (sameb) k_view_extents_multi<true><<<nb, kBlock, 0, s>>>(P, np, t, m, pv, 6 * kMaxPartials);
    else k_view_extents_multi<false><<<nb, kBlock, 0, s>>>(P, np, t, m, pv, 6 * kMaxPartials);
  }
  return nb;
}

// Per-point records shared by the lines of a render: RS4 float4 per point.
int rec_width(int nlines) {
  const int f4 = (4 + nlines + 3) / 4;
  return f4 <= 2 ? 2 : (f4 <= 4 ? 4 : 8);
}

// NN query for every sample of a view (idx: S ints, always written). With rec: per-sample vlos
// and emissivities (nlines) too. Two passes: k_nn (lookup grid) and k_nn_far (exact octree query
// for the rest).
// rayflag: nn_scratch_bytes(R, S) bytes: R ray flags, zero on entry and on exit (set by k_nn,
// cleared by k_clear_u8), then k_nn_far's deep queue (its count zeroed by k_nn).
void enqueue_nn(hipStream_t s, const Dataset& ds, const ViewGeom* geom, int nx, int ny, int nz, int32_t* idx,
                uint8_t* rayflag, const float4* rec = nullptr, int rs4 = 0, int nlines = 0, float* svlos = nullptr,
                float2* sem = nullptr, const RecInfo& ri = RecInfo{}) {
  const int tlog = ds.tile_log2, tw = 1 << tlog, th = 256 >> tlog;
  const int tiles = ((nx + tw - 1) / tw) * ((ny + th - 1) / th);
  dim3 grid(tiles, nz);
  const int64_t R = (int64_t)nx * ny;
  const DeepQueue dq = deep_queue(rayflag, R, R * nz, ds.far_budget, ds.segbox.as<float4>());
  {
    ProfScope ps("k_nn", s);
#define FOMO_NN_ARGS                                                                                         \
  ds.pts.as<float4>(), ds.grid, ds.lslot.as<float4>(), ds.lcand.as<float4>(), ds.lk, geom, nx, ny, nz, idx, rec, \
      ri, nlines, svlos, sem, rayflag, tlog, dq.count
    if (!rec) k_nn<0><<<grid, 256, 0, s>>>(FOMO_NN_ARGS);
    else if (rs4 == 2) k_nn<2><<<grid, 256, 0, s>>>(FOMO_NN_ARGS);
    else if (rs4 == 4) k_nn<4><<<grid, 256, 0, s>>>(FOMO_NN_ARGS);
    else k_nn<8><<<grid, 256, 0, s>>>(FOMO_NN_ARGS);
#undef FOMO_NN_ARGS
  }
  {
    ProfScope ps("k_nn_far", s);
    const dim3 fgrid((unsigned)(((nx + 7) / 8) * ((ny + 7) / 8)), (unsigned)((nz + 3) / 4));
#define FOMO_NNF_ARGS                                                                                         \
  ds.pts.as<float4>(), ds.mpts.as<float4>(), ds.nlo.as<float4>(), ds.nhi.as<float4>(), ds.bvh, geom, nx, ny, \
      nz, idx, rec, ri, nlines, svlos, sem, rayflag, dq
    if (!rec) k_nn_far<0><<<fgrid, 256, 0, s>>>(FOMO_NNF_ARGS);
    else if (rs4 == 2) k_nn_far<2><<<fgrid, 256, 0, s>>>(FOMO_NNF_ARGS);
    else if (rs4 == 4) k_nn_far<4><<<fgrid, 256, 0, s>>>(FOMO_NNF_ARGS);
    else k_nn_far<8><<<fgrid, 256, 0, s>>>(FOMO_NNF_ARGS);
#undef FOMO_NNF_ARGS
  }
  {
    ProfScope ps("k_nn_deep", s);
    k_nn_deep<<<kDeepBlocks, 256, 0, s>>>(ds.pts.as<float4>(), ds.mpts.as<float4>(), ds.nlo.as<float4>(),
                                           ds.nhi.as<float4>(), ds.bvh, geom, nx, ny, dq);
    const unsigned mblocks = (unsigned)((dq.cap * 64 + 255) / 256);
#define FOMO_NND_ARGS ds.pts.as<float4>(), geom, nx, ny, nz, idx, rec, ri, nlines, svlos, sem, dq
    if (!rec) k_nn_deep_merge<0><<<mblocks, 256, 0, s>>>(FOMO_NND_ARGS);
    else if (rs4 == 2) k_nn_deep_merge<2><<<mblocks, 256, 0, s>>>(FOMO_NND_ARGS);
    else if (rs4 == 4) k_nn_deep_merge<4><<<mblocks, 256, 0, s>>>(FOMO_NND_ARGS);
    else k_nn_deep_merge<8><<<mblocks, 256, 0, s>>>(FOMO_NND_ARGS);
#undef FOMO_NND_ARGS
    k_clear_u8<<<nblk(R), kBlock, 0, s>>>(rayflag, R);
  }
}

// Tetra path for every sample of a view: per-sample vlos / emissivities like enqueue_nn, and the
// chosen vertex (sorted position, -1 outside the hull) in vert when given.
// Entries of a view's redo list (k_tet_walk) held in tet_redo_bytes.
constexpr int64_t kTetRedoMax = 1 << 16;
int tet_redo_cap(int64_t S) {
  return S > (int64_t)UINT32_MAX ? 0 : (int)std::max<int64_t>(0, std::min<int64_t>(S - 2, kTetRedoMax));
}
int64_t tet_redo_bytes(int64_t S) { return (int64_t)sizeof(uint32_t) * (2 + tet_redo_cap(S)); }

// redo: tet_redo_bytes(R * nz) of scratch for the samples the filtered walk cannot place.
int enqueue_tet(hipStream_t s, const Dataset& ds, const ViewGeom* geom, int nx, int ny, int nz, bool linear,
                const float4* rec, int rs4, const RecInfo& ri, int nlines, float* svlos, float2* sem, int32_t* vert,
                uint32_t* redo) {
  if (ds.ntets <= 0) return fail(FOMO_ERR_STATE, "dataset has no triangulation (fomo_dataset_triangulate)");
  const int64_t R = (int64_t)nx * ny;
  const int redo_cap = tet_redo_cap(R * nz);
  k_clear_u8<<<1, 64, 0, s>>>(reinterpret_cast<uint8_t*>(redo), 8);
  // split rays into segments only as far as needed to keep kTetThreads walks in flight
  const int64_t want = (R * nz + kTetThreads - 1) / kTetThreads;
  const int seglen = (int)std::max<int64_t>(1, std::min<int64_t>({want, (int64_t)kTetMaxSegment, (int64_t)nz}));
  const int64_t nsegs = (nz + seglen - 1) / seglen;
  const int64_t nthreads = (int64_t)((nx + 7) / 8) * ((ny + 7) / 8) * 64 * nsegs;  // 8 x 8 pixel tiles
  const unsigned nb = (unsigned)((nthreads + 255) / 256);
  ProfScope ps(linear ? "k_tet_walk_linear" : "k_tet_walk", s);
#define FOMO_TET_ARGS                                                                                        \
  ds.trec.as<float4>(), ds.tstart.as<int32_t>(), ds.tk, geom, nx, ny, nz, seglen, rec, ri, \
      nlines, svlos, sem, vert, redo, redo_cap
#define FOMO_REDO_ARGS                                                                                    \
  ds.trec.as<float4>(), ds.tstart.as<int32_t>(), ds.tk, geom, nx, ny, nz, rec, ri, nlines, \
      svlos, sem, vert, redo, redo_cap, ds.tetstats.as<unsigned long long>()
#define FOMO_TET_LAUNCH(RS, LIN)                                       \
  do {                                                                 \
    k_tet_walk<RS, LIN><<<nb, 256, 0, s>>>(FOMO_TET_ARGS);             \
    k_tet_redo<RS, LIN><<<kTetRedoBlocks, 64, 0, s>>>(FOMO_REDO_ARGS); \
  } while (0)
  if (!rec) {
    if (linear) return fail(FOMO_ERR_ARG, "linear interpolation needs per-point records");
    FOMO_TET_LAUNCH(0, false);
  } else if (rs4 == 2) {
    if (linear) FOMO_TET_LAUNCH(2, true);
    else FOMO_TET_LAUNCH(2, false);
  } else if (rs4 == 4) {
    if (linear) FOMO_TET_LAUNCH(4, true);
    else FOMO_TET_LAUNCH(4, false);
  } else {
    if (linear) FOMO_TET_LAUNCH(8, true);
    else FOMO_TET_LAUNCH(8, false);
  }
#undef FOMO_TET_LAUNCH
#undef FOMO_REDO_ARGS
#undef FOMO_TET_ARGS
  HIPOK(hipGetLastError());
  return FOMO_OK;
}

// Build records {vel, T, peaks} for a set of emission cubes of one dataset; ri gets their K.
int build_records(hipStream_t s, const Dataset& ds, const std::vector<const Goft*>& gcs, DevBuf& rec, int* rs4,
                  RecInfo* ri) {
  if (ds.nvars < 2) return fail(FOMO_ERR_STATE, "records need the temperature variable");
  for (const Goft* g : gcs)
    if ((g->K < 0) != (gcs[0]->K < 0))
      return fail(FOMO_ERR_ARG, "emission cubes made from a given GoftCube cannot share a render with computed ones");
  *rs4 = rec_width((int)gcs.size());
  if (rec.ensure(sizeof(float4) * (*rs4) * ds.n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (records)");
  LineSet ls{};
  ls.n = (int)gcs.size();
  for (int i = 0; i < ls.n; ++i) ls.em[i] = gcs[i]->em.as<float2>();
  if (ri) {
    *ri = RecInfo{};
    for (int i = 0; i < ls.n; ++i) ri->K[i] = gcs[i]->K;
  }
  k_build_rec<<<nblk(ds.n), kBlock, 0, s>>>(ds.nvars >= 5 ? ds.vel.as<float4>() : nullptr, ds.svars[1]->as<float>(), ls,
                                            ds.n, *rs4, rec.as<float4>());
  HIPOK(hipGetLastError());
  return FOMO_OK;
}

// Pick the lambda chunk: <=64 pixels per pass, chunk a multiple of 4.
int pick_chunk(int nl, int* npass) {
  int np = (nl + 63) / 64;
  int ch = (nl + np - 1) / np;
  ch = ((ch + 3) / 4) * 4;
  *npass = (nl + ch - 1) / ch;
  return ch;
}

template <int CH>
void launch_spectral_t(hipStream_t s, int R, int nz, const LineSet& ls, const float* vlos, const ViewGeom* geom,
                       int nl, int l0, int32_t* flag) {
  // 256 rays per block, one lane per ray
  dim3 grid((unsigned)((R + 255) / 256), ls.n);
  k_los_spectral<CH, false><<<grid, 256, 0, s>>>(R, nz, ls, vlos, geom, nl, l0, flag);
  k_los_spectral<CH, true><<<grid, 256, 0, s>>>(R, nz, ls, vlos, geom, nl, l0, flag);
}

// All lines of ls (per-sample emissivities written by k_nn). flag: R*ls.n ints of scratch.
void enqueue_los(hipStream_t s, int R, int nz, const LineSet& ls, const float* vlos, const ViewGeom* geom, int nl,
                 int32_t* flag) {
  if (nl <= 1) {
    ProfScope ps("k_los_imaging", s);
    k_los_imaging<<<dim3(nblk(R), ls.n), 256, 0, s>>>(R, nz, ls, geom);
    return;
  }
  int npass = 1;
  const int ch = pick_chunk(nl, &npass);
  ProfScope ps("k_los_spectral", s);
  for (int pass = 0; pass < npass; ++pass) {
    const int l0 = pass * ch;
    switch (ch) {
#define FOMO_CH(C) \
  case C: launch_spectral_t<C>(s, R, nz, ls, vlos, geom, nl, l0, flag); break;
      FOMO_CH(4) FOMO_CH(8) FOMO_CH(12) FOMO_CH(16) FOMO_CH(20) FOMO_CH(24) FOMO_CH(28) FOMO_CH(32)
      FOMO_CH(36) FOMO_CH(40) FOMO_CH(44) FOMO_CH(48) FOMO_CH(52) FOMO_CH(56) FOMO_CH(60) FOMO_CH(64)
#undef FOMO_CH
      default: break;
    }
  }
}

// Fused NN + integration pass (k_nn_los) for spectral views with one lambda pass (2 <= nl <= 64)
// and per-point records (two passes for one line with 65 <= nl <= 128). The instantiations are
// compiled in fomo_nnlos_part.hip; this picks the one for lambda chunk ch.
void launch_nn_los(int variant, int ch, unsigned blocks, hipStream_t s, const Dataset& ds, const ViewGeom* geom,
                   int nx, int ny, int nz, const float4* rec, const RecInfo& ri, const LineSet& ls, int nl,
                   int32_t* flags, int nviews, int64_t out_vs, int64_t flag_vs, int32_t* dlist) {
  const NnLosArgs a{ds.pts.as<float4>(), ds.grid, ds.lslot.as<float4>(), ds.lcand.as<float4>(), ds.lk,
                    ds.mpts.as<float4>(), ds.nlo.as<float4>(), ds.nhi.as<float4>(), ds.bvh, geom, nx, ny, nz, rec,
                    ri, ls, nl, flags, nviews, out_vs, flag_vs, dlist};
  switch (ch) {
#define FOMO_NNL_CH(C)                              \
  case C:                                           \
    if (ds.lk.rect)                                 \
      nn_los_launch<C, true>(variant, blocks, s, a);  \
    else                                            \
      nn_los_launch<C, false>(variant, blocks, s, a); \
    break;
    FOMO_NNL_CH(4) FOMO_NNL_CH(8) FOMO_NNL_CH(12) FOMO_NNL_CH(16) FOMO_NNL_CH(20) FOMO_NNL_CH(24) FOMO_NNL_CH(28)
    FOMO_NNL_CH(32) FOMO_NNL_CH(36) FOMO_NNL_CH(40) FOMO_NNL_CH(44) FOMO_NNL_CH(48) FOMO_NNL_CH(52)
    FOMO_NNL_CH(56) FOMO_NNL_CH(60) FOMO_NNL_CH(64)
#undef FOMO_NNL_CH
    default: break;
  }
}

// True when the lines of ls have one line profile: the same lambda window and rest wavelength
// (so the same Doppler shift) and the same thermal-width factor K (fwhm = K sqrt T), bit for bit.
// Their Gaussians then differ only in the peak.
bool shared_profile(const RecInfo& ri, const LineSet& ls) {
  for (int l = 1; l < ls.n; ++l)
    if (ri.K[l] != ri.K[0] || ls.wla[l] != ls.wla[0] || ls.l0c[l] != ls.l0c[0]) return false;
  return true;
}

// Rays below which a single-line view runs k_nn + k_los_spectral rather than the fused pass.
constexpr int64_t kFusedMinRays = 65536;

// True when a single-line spectral view of R rays runs the fused pass (k_nn_los): the 4 lanes of
// a ray run the profile's 4 chains, of up to two lambda passes. Small images have too few rays to
// fill the GPU this way; they run k_nn (one thread per sample) + k_los_spectral, which add the
// same values in the same order.
bool fused_single_line(const float4* rec, int rs4, int64_t R, int nl) {
  if (!rec || nl < 2) return false;
  int npass = 1;
  (void)pick_chunk(nl, &npass);
  return R >= kFusedMinRays && rs4 == 2 && npass <= 2;
}

// flags: R * ls.n ints of scratch. False when the view is left to k_nn + the integration kernels.
// The fused pass of nviews views of one resolution (geom[v]; outputs ls.out[l] + v * out_vs; ray
// flags flags + v * flag_vs, R * lines each) in one launch, when it applies to them (false:
// nothing enqueued). dlist: nn_los_list_ints(R, nviews) ints for the direct pass's list, cleared by
// the views' finalize before. launch = false only answers whether it applies.
bool enqueue_nn_los(hipStream_t s, const Dataset& ds, const ViewGeom* geom, int nx, int ny, int nz,
                    const float4* rec, int rs4, const RecInfo& ri, const LineSet& ls, int nl, int32_t* flags,
                    int32_t* dlist, bool launch = true, int nviews = 1, int64_t out_vs = 0, int64_t flag_vs = 0) {
  if (!rec || nl < 2 || ls.n < 1 || ls.n > kMaxLines) return false;
  const int64_t R = (int64_t)nx * ny;
  // k_los_spectral's lambda passes and their anchors
  int npass = 1;
  const int ch = pick_chunk(nl, &npass);
  if (ls.n == 1) {
    if (!fused_single_line(rec, rs4, R, nl)) return false;
    if (!launch) return true;
    ProfScope ps("k_nn_los", s);
    const unsigned blocks = (unsigned)((R + 63) / 64);
    launch_nn_los(npass == 1 ? kNnLosSingle : kNnLosSingle2, ch, blocks, s, ds, geom, nx, ny, nz, rec, ri, ls, nl,
                  flags, nviews, out_vs, flag_vs, dlist);
    return true;
  }
  if (npass != 1) return false;
  // several lines share each query
  const bool shared = rs4 == 2 && ls.n <= 4 && shared_profile(ri, ls);
  const int G = shared ? 4 : (ls.n <= 2 ? 2 : (ls.n <= 4 ? 4 : 8));
  if ((rs4 == 2 && G > 4) || (rs4 != 2 && rs4 != 4)) return false;
  if (!launch) return true;
  const int64_t rays_per_block = 4 * (64 / G);
  const unsigned blocks = (unsigned)((R + rays_per_block - 1) / rays_per_block);
  ProfScope ps("k_nn_los", s);
  const int variant = shared ? kNnLosShared : (rs4 == 2 && G == 2) ? kNnLosLines2 : rs4 == 2 ? kNnLosLines4 : kNnLosLines8;
  launch_nn_los(variant, ch, blocks, s, ds, geom, nx, ny, nz, rec, ri, ls, nl, flags, nviews, out_vs, flag_vs, dlist);
  return true;
}

LineSet make_lineset(const std::vector<const float2*>& em, const std::vector<float*>& out,
                     const std::vector<double>& lambda0, double lambda_width) {
  LineSet ls{};
  ls.n = (int)em.size();
  for (int i = 0; i < ls.n; ++i) {
    ls.em[i] = em[i];
    ls.out[i] = out[i];
    ls.wla[i] = lambda_width * lambda0[i] / FOMO_C_LIGHT;  // fomo-nearestneighbour.cpp:116
    ls.l0c[i] = lambda0[i] / FOMO_C_LIGHT;
  }
  return ls;
}

int check_res(int nx, int ny, int nz, int nl) {
  if (nx < 2 || ny < 2) return fail(FOMO_ERR_ARG, "x_pixel and y_pixel must be >= 2 (the reference divides by x_pixel-1)");
  if (nz < 1 || nl < 1) return fail(FOMO_ERR_ARG, "z_pixel and lambda_pixel must be >= 1");
  if ((double)nx * ny * nz > 2147483647.0) return fail(FOMO_ERR_ARG, "too many LOS samples in one view (>2^31)");
  if ((double)nx * ny * nl > 2147483647.0) return fail(FOMO_ERR_ARG, "output cube too large (>2^31 values)");
  return FOMO_OK;
}

// Device buffers of one view render (plain render or server request), in render_sizes order.
struct RenderBufs {
  double* partial;
  ViewGeom* geom;
  int32_t* idx;      // S; projection: the ProjLayout arrays
  float* aux;        // NN: per-sample LOS velocity (S); Thomson: per-point sin^2 chi (N); projection: sort scratch
  float* out;        // R * max(nl,1)
  int32_t* flag;     // NN: R rays needing the direct-exp pass
  float2* sem;       // NN: per-sample (peak, fwhm), S
  uint8_t* rayflag;  // nn_scratch_bytes: R ray flags, zero on entry (left zero), and the deep queue
};
constexpr int kRenderBufs = 8;

// Projection buffers inside RenderBufs::idx (int32 offsets, 256-byte aligned): pixel keys and
// sorted positions in original point order, the same sorted by pixel, and the pixel starts.
struct ProjLayout {
  int64_t key_in, val_in, key_out, val_out, start, total;
  int bits;  // key bits to sort
};
ProjLayout proj_layout(int64_t n, int64_t R) {
  const auto up = [](int64_t a) { return (a + 63) & ~(int64_t)63; };
  ProjLayout L;
  L.key_in = 0;
  L.val_in = up(n);
  L.key_out = 2 * up(n);
  L.val_out = 3 * up(n);
  L.start = 4 * up(n);
  L.total = L.start + up(R + 1);
  L.bits = 1;
  while (L.bits < 32 && ((uint64_t)(R - 1) >> L.bits) != 0) ++L.bits;
  return L;
}
size_t proj_sort_bytes(int64_t n, int bits) {
  size_t b = 0;
  (void)rocprim::radix_sort_pairs(nullptr, b, (uint32_t*)nullptr, (uint32_t*)nullptr, (int32_t*)nullptr,
                                  (int32_t*)nullptr, (size_t)n, 0, bits, (hipStream_t)0);
  return b;
}

void render_sizes(int mode, int64_t n, int64_t R, int64_t S, int nl, int64_t sz[kRenderBufs]) {
  sz[0] = (int64_t)sizeof(double) * 6 * (kMaxPartials + 1);
  sz[1] = (int64_t)sizeof(ViewGeom);
  if (mode == MODE_PROJ) {
    const ProjLayout L = proj_layout(n, R);
    sz[2] = (int64_t)sizeof(int32_t) * L.total;
    sz[3] = (int64_t)proj_sort_bytes(n, L.bits);
    sz[4] = (int64_t)sizeof(float) * R * std::max(nl, 1);
    sz[5] = sz[6] = sz[7] = 0;
    return;
  }
  sz[2] = (int64_t)sizeof(int32_t) * S;
  sz[3] = mode != MODE_THOMSON ? (int64_t)sizeof(float) * S : (int64_t)sizeof(double) * n;  // (pB terms)
  sz[4] = (int64_t)sizeof(float) * R * std::max(nl, 1);
  // ray flags, then (NN) the fused pass's direct-pass list
  sz[5] = mode != MODE_THOMSON ? (int64_t)sizeof(int32_t) * (R + nn_los_list_ints(R, 1)) : 0;
  sz[6] = mode != MODE_THOMSON ? (int64_t)sizeof(float2) * S : 0;
  sz[7] = nn_scratch_bytes(R, S);
}

// Argument and object checks shared by the synchronous renders and the server.
int resolve_render(int goft, int mode, int nx, int ny, int nz, int nl, Goft** gcp, Dataset** dsp) {
  if (int rc = check_res(nx, ny, nz, nl)) return rc;
  Goft* gc = find_obj(g_ctx.gofts, goft);
  if (!gc) return fail(FOMO_ERR_NOTFOUND, "unknown emission id");
  if ((mode == MODE_THOMSON) != (gc->kind == 1))
    return fail(FOMO_ERR_STATE, "emission cube kind does not match the render method");
  Dataset* ds = find_obj(g_ctx.datasets, gc->dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "dataset of this emission cube was destroyed");
  if (mode != MODE_THOMSON && nl > 1 && ds->nvars < 5)
    return fail(FOMO_ERR_ARG, "spectroscopic render needs vx,vy,vz");
  *gcp = gc;
  *dsp = ds;
  return FOMO_OK;
}

// Projection of one view after its geometry (rb.geom): pixel keys, stable sort by pixel, pixel
// starts, per-(pixel, lambda) sums into rb.out.
int enqueue_projection(hipStream_t s, const Dataset& ds, const Goft& gc, const RenderBufs& rb, int nx, int ny,
                       int nl, double width) {
  const int64_t n = ds.n, R = (int64_t)nx * ny;
  const ProjLayout L = proj_layout(n, R);
  uint32_t* key_in = reinterpret_cast<uint32_t*>(rb.idx + L.key_in);
  uint32_t* key_out = reinterpret_cast<uint32_t*>(rb.idx + L.key_out);
  int32_t* val_in = rb.idx + L.val_in;
  int32_t* val_out = rb.idx + L.val_out;
  int32_t* start = rb.idx + L.start;
  {
    ProfScope ps("k_proj_keys", s);
    k_proj_keys<<<nblk(n), kBlock, 0, s>>>(ds.pts.as<float4>(), ds.perm.as<int32_t>(), n, rb.geom, nx, ny, key_in,
                                           val_in);
    HIPOK(hipGetLastError());
  }
  {
    ProfScope ps("proj_sort", s);
    size_t tmpb = proj_sort_bytes(n, L.bits);
    HIPOK(rocprim::radix_sort_pairs(rb.aux, tmpb, key_in, key_out, val_in, val_out, (size_t)n, 0, L.bits, s));
  }
  k_proj_starts<<<nblk(R + 1), kBlock, 0, s>>>(key_out, n, R, start);
  HIPOK(hipGetLastError());
  const int64_t nt = nl > 1 ? R * ((nl + kProjChunk - 1) / kProjChunk) : R;
  const double wla = width * gc.lambda0 / FOMO_C_LIGHT;  // fomo-projection.cpp:69
  ProfScope ps("k_proj_spectral", s);
#define FOMO_PROJ_ARGS start, val_out, gc.rec.as<float4>(), gc.ri, rb.geom, R, nl, wla, gc.lambda0, rb.out
  if (gc.rs4 == 2) k_proj_spectral<2><<<nblk(nt), 256, 0, s>>>(FOMO_PROJ_ARGS);
  else if (gc.rs4 == 4) k_proj_spectral<4><<<nblk(nt), 256, 0, s>>>(FOMO_PROJ_ARGS);
  else k_proj_spectral<8><<<nblk(nt), 256, 0, s>>>(FOMO_PROJ_ARGS);
#undef FOMO_PROJ_ARGS
  HIPOK(hipGetLastError());
  return FOMO_OK;
}

// All kernels of one view on stream s; the image lands in rb.out, the geometry in rb.geom.
int enqueue_render(hipStream_t s, const Dataset& ds, const Goft& gc, int mode, double l, double b, int nx, int ny,
                   int nz, int nl, double width, int dn, const RenderBufs& rb) {
  const int64_t R = (int64_t)nx * ny;
  // NN renders: the fused pass's direct-pass list follows the R ray flags in rb.flag
  int32_t* const dlist = mode == MODE_NN ? rb.flag + R : nullptr;
  enqueue_view_geom(s, ds, mode == MODE_THOMSON ? MODE_THOMSON : MODE_NN, l, b, nx, ny, nz, dn, rb.aux, rb.partial,
                    rb.geom, dlist, mode == MODE_THOMSON ? gc.em.as<float2>() : nullptr);
  if (mode == MODE_TET || mode == MODE_LINEAR) {
    // the tetra path leaves rb.idx (S ints) unused: it holds the view's redo list
    if (int rc = enqueue_tet(s, ds, rb.geom, nx, ny, nz, mode == MODE_LINEAR, gc.rec.as<float4>(), gc.rs4, gc.ri, 1,
                             rb.aux, rb.sem, nullptr, reinterpret_cast<uint32_t*>(rb.idx)))
      return rc;
    LineSet ls = make_lineset({rb.sem}, {rb.out}, {gc.lambda0}, width);
    enqueue_los(s, (int)R, nz, ls, rb.aux, rb.geom, nl, rb.flag);
  } else if (mode == MODE_PROJ) {
    if (int rc = enqueue_projection(s, ds, gc, rb, nx, ny, nl, width)) return rc;
  } else if (mode == MODE_NN) {
    LineSet ls = make_lineset({rb.sem}, {rb.out}, {gc.lambda0}, width);
    if (!enqueue_nn_los(s, ds, rb.geom, nx, ny, nz, gc.rec.as<float4>(), gc.rs4, gc.ri, ls, nl, rb.flag, dlist)) {
      enqueue_nn(s, ds, rb.geom, nx, ny, nz, rb.idx, rb.rayflag, gc.rec.as<float4>(), gc.rs4, 1, rb.aux, rb.sem,
                 gc.ri);
      enqueue_los(s, (int)R, nz, ls, rb.aux, rb.geom, nl, rb.flag);
    }
  } else {
    enqueue_nn(s, ds, rb.geom, nx, ny, nz, rb.idx, rb.rayflag);
    ProfScope ps("k_los_thomson", s);
    k_los_thomson<<<nblk(R), 256, 0, s>>>(rb.idx, (int)R, nz, reinterpret_cast<const double*>(rb.aux), rb.geom,
                                          rb.out);
  }
  HIPOK(hipGetLastError());
  return FOMO_OK;
}

void geom_to_meta(const ViewGeom& hg, double* meta) {
  for (int a = 0; a < 6; ++a) meta[a] = hg.ext[a];
  meta[6] = hg.md;
  meta[7] = hg.pathlength;
  meta[8] = hg.scale;
  meta[9] = hg.apix;
}


// ---------------------------------------------------------------------------------------
// Render server: continuous batching of independent view requests over resident data.
//
// Each request is one view (NN or Thomson) of a resident emission cube. A worker thread admits
// queued requests in arrival order whenever an in-flight slot and a block of the server's
// device arena (the request's footprint: per-sample NN buffers, image, scratch) and of its
// pinned host arena (image + geometry staging) are free; smaller requests may pass one that
// does not fit, at most kMaxBypass times. An admitted request is enqueued on the least loaded of
// the server's streams, so requests overlap on the device. Completion is seen through an
// event; the worker then copies the image to the caller's buffer and frees the blocks.
// Requests run the same kernels as the synchronous renders on private buffers, so a result does
// not depend on what else is in flight. The worker takes no global lock: datasets and emission
// cubes referenced by queued or running requests cannot be destroyed or re-indexed (in-use
// counts), and every buffer it writes belongs to the request.
// ---------------------------------------------------------------------------------------
constexpr int kMaxBypass = 16;
constexpr int kCopyThreads = 6;  // host threads copying finished images out of pinned staging
enum { kTicketQueued = 0, kTicketRunning = 1, kTicketDone = 2 };

struct Ticket {
  Dataset* ds = nullptr;
  Goft* gc = nullptr;
  int mode = MODE_NN;
  double l = 0, b = 0, width = 0;
  int nx = 0, ny = 0, nz = 0, nl = 1, dn = 0;
  float* out = nullptr;
  double* meta = nullptr;
  int state = kTicketQueued;  // or a negative FOMO_ERR_* code
  std::string err;
  double t_submit = 0, t_admit = 0, t_done = 0;
  int64_t sz[kRenderBufs] = {};
  int64_t dev_bytes = 0, host_bytes = 0;
  int64_t dev_off = -1, host_off = -1;
  int stream = -1;
  int bypassed = 0;
  hipEvent_t done = nullptr;
  std::atomic<int> chunks_left{0};  // copy-out chunks not yet copied
  Ticket() = default;
  Ticket(Ticket&& o) noexcept
      : ds(o.ds), gc(o.gc), mode(o.mode), l(o.l), b(o.b), width(o.width), nx(o.nx), ny(o.ny), nz(o.nz),
        nl(o.nl), dn(o.dn), out(o.out), meta(o.meta), state(o.state), err(std::move(o.err)),
        t_submit(o.t_submit), t_admit(o.t_admit), t_done(o.t_done), dev_bytes(o.dev_bytes),
        host_bytes(o.host_bytes), dev_off(o.dev_off), host_off(o.host_off), stream(o.stream),
        bypassed(o.bypassed), done(o.done) {
    std::memcpy(sz, o.sz, sizeof(sz));
  }
};

// One piece of a finished request's image copy (pinned staging -> caller's buffer).
struct CopyJob {
  Ticket* t;
  int64_t off, len;
};
constexpr int64_t kCopyChunk = 4 << 20;

int64_t footprint_device(const int64_t sz[kRenderBufs]) {
  int64_t t = 0;
  for (int i = 0; i < kRenderBufs; ++i) t += fomo::Arena::round_up(sz[i]);
  return t;
}
int64_t footprint_host(const int64_t sz[kRenderBufs]) {
  return fomo::Arena::round_up(sz[4]) + fomo::Arena::round_up(sizeof(ViewGeom));
}

struct Server {
  int device = 0;
  int max_inflight = 1;
  std::vector<hipStream_t> streams;
  std::vector<int> load;  // requests in flight per stream
  void* dbase = nullptr;
  void* hbase = nullptr;
  fomo::Arena dev, host;
  std::vector<hipEvent_t> evpool;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<int64_t> pending;
  std::vector<int64_t> inflight;
  std::deque<Ticket> tickets;  // ticket id = position (references stay valid on push_back)
  std::deque<CopyJob> copyq;   // copy-out work for the copier threads
  std::condition_variable cv_copy;
  std::vector<std::thread> copiers;
  int64_t copying = 0;         // requests whose image is being copied out
  bool stop = false;
  int64_t completed = 0, failed = 0, peak_inflight = 0;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  std::thread worker;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  bool idle() const { return pending.empty() && inflight.empty() && copying == 0; }
  ~Server() {
    for (auto e : evpool) (void)hipEventDestroy(e);
    for (auto& t : tickets)
      if (t.done) (void)hipEventDestroy(t.done);
    for (auto st : streams) (void)hipStreamDestroy(st);
    if (dbase) (void)hipFree(dbase);
    if (hbase) (void)hipHostFree(hbase);
  }
};

void ticket_release_objects(Ticket& t) {
  if (t.ds) t.ds->inuse.fetch_sub(1);
  if (t.gc) t.gc->inuse.fetch_sub(1);
  t.ds = nullptr;
  t.gc = nullptr;
}

// Enqueue one admitted request (server lock held). Returns FOMO_OK or an error for this ticket.
int server_enqueue(Server& sv, Ticket& t) {
  hipStream_t s = sv.streams[t.stream];
  char* d = static_cast<char*>(sv.dbase) + t.dev_off;
  char* p[kRenderBufs];
  for (int i = 0; i < kRenderBufs; ++i) {
    p[i] = d;
    d += fomo::Arena::round_up(t.sz[i]);
  }
  const RenderBufs rb{reinterpret_cast<double*>(p[0]),  reinterpret_cast<ViewGeom*>(p[1]),
                      reinterpret_cast<int32_t*>(p[2]), reinterpret_cast<float*>(p[3]),
                      reinterpret_cast<float*>(p[4]),   reinterpret_cast<int32_t*>(p[5]),
                      reinterpret_cast<float2*>(p[6]),  reinterpret_cast<uint8_t*>(p[7])};
  HIPOK(hipMemsetAsync(rb.rayflag, 0, t.sz[7], s));
  if (int rc = enqueue_render(s, *t.ds, *t.gc, t.mode, t.l, t.b, t.nx, t.ny, t.nz, t.nl, t.width, t.dn, rb))
    return rc;
  char* h = static_cast<char*>(sv.hbase) + t.host_off;
  HIPOK(hipMemcpyAsync(h, rb.out, t.sz[4], hipMemcpyDeviceToHost, s));
  HIPOK(hipMemcpyAsync(h + fomo::Arena::round_up(t.sz[4]), rb.geom, sizeof(ViewGeom), hipMemcpyDeviceToHost, s));
  if (sv.evpool.empty()) {
    hipEvent_t e = nullptr;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    sv.evpool.push_back(e);
  }
  t.done = sv.evpool.back();
  sv.evpool.pop_back();
  HIPOK(hipEventRecord(t.done, s));
  return FOMO_OK;
}

void server_fail_ticket(Server& sv, Ticket& t, int code, const std::string& msg) {
  t.state = code;
  t.err = msg;
  t.t_done = sv.now();
  ticket_release_objects(t);
  ++sv.failed;
}

// Admit what fits (server lock held). Returns true when anything was admitted or failed.
bool server_admit(Server& sv) {
  bool any = false;
  Ticket* blocked = nullptr;  // oldest request that did not fit
  for (auto it = sv.pending.begin(); it != sv.pending.end();) {
    if ((int64_t)sv.inflight.size() >= sv.max_inflight) break;
    Ticket& t = sv.tickets[*it];
    const int64_t doff = sv.dev.alloc(t.dev_bytes);
    const int64_t hoff = doff >= 0 ? sv.host.alloc(t.host_bytes) : -1;
    if (hoff < 0) {
      if (doff >= 0) sv.dev.release(doff);
      if (!blocked) {
        blocked = &t;
        if (t.bypassed >= kMaxBypass) break;  // nothing may pass it any more
      }
      ++it;
      continue;
    }
    t.dev_off = doff;
    t.host_off = hoff;
    int best = 0;
    for (int i = 1; i < (int)sv.streams.size(); ++i)
      if (sv.load[i] < sv.load[best]) best = i;
    t.stream = best;
    const int64_t id = *it;
    it = sv.pending.erase(it);
    any = true;
    t.t_admit = sv.now();
    if (int rc = server_enqueue(sv, t)) {
      // launches already made stay on the stream; keep the blocks until it drains
      (void)hipStreamSynchronize(sv.streams[t.stream]);
      sv.dev.release(t.dev_off);
      sv.host.release(t.host_off);
      server_fail_ticket(sv, t, rc, g_err);
      sv.cv_done.notify_all();
      continue;
    }
    t.state = kTicketRunning;
    sv.load[best]++;
    sv.inflight.push_back(id);
    sv.peak_inflight = std::max<int64_t>(sv.peak_inflight, (int64_t)sv.inflight.size());
    if (blocked) blocked->bypassed++;
  }
  return any;
}

// Last step of a request (server lock held): geometry to meta, free the staging block, done.
void server_finish(Server& sv, Ticket& t) {
  if (t.meta) {
    ViewGeom hg;
    std::memcpy(&hg, static_cast<const char*>(sv.hbase) + t.host_off + fomo::Arena::round_up(t.sz[4]),
                sizeof(ViewGeom));
    geom_to_meta(hg, t.meta);
  }
  sv.host.release(t.host_off);
  t.state = kTicketDone;
  t.t_done = sv.now();
  ticket_release_objects(t);
  ++sv.completed;
  --sv.copying;
  sv.cv_done.notify_all();
  sv.cv_work.notify_one();  // the staging block may let a queued request in
}

// One pass of the worker: retire requests whose device work is done (their device block is freed
// at once; the image copy goes to the copier threads), then admit. Returns whether anything changed.
bool server_step(Server& sv) {
  std::lock_guard<std::mutex> lk(sv.mu);
  bool changed = false;
  for (size_t i = 0; i < sv.inflight.size();) {
    Ticket& t = sv.tickets[sv.inflight[i]];
    const hipError_t q = hipEventQuery(t.done);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    changed = true;
    sv.inflight[i] = sv.inflight.back();
    sv.inflight.pop_back();
    sv.dev.release(t.dev_off);
    sv.load[t.stream]--;
    sv.evpool.push_back(t.done);
    t.done = nullptr;
    if (q != hipSuccess) {
      sv.host.release(t.host_off);
      server_fail_ticket(sv, t, FOMO_ERR_HIP, std::string("request failed on the device: ") + hipGetErrorString(q));
      sv.cv_done.notify_all();
      continue;
    }
    ++sv.copying;
    const int64_t bytes = t.out ? t.sz[4] : 0;
    const int n = (int)std::max<int64_t>(1, (bytes + kCopyChunk - 1) / kCopyChunk);
    t.chunks_left.store(n);
    for (int c = 0; c < n; ++c)
      sv.copyq.push_back({&t, c * kCopyChunk, std::min<int64_t>(kCopyChunk, bytes - c * kCopyChunk)});
    sv.cv_copy.notify_all();
  }
  if (server_admit(sv)) changed = true;
  return changed;
}

void server_copier(Server* sv) {
  for (;;) {
    CopyJob j;
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_copy.wait(lk, [&] { return !sv->copyq.empty() || (sv->stop && sv->idle()); });
      if (sv->copyq.empty()) return;
      j = sv->copyq.front();
      sv->copyq.pop_front();
    }
    if (j.len > 0)
      std::memcpy(reinterpret_cast<char*>(j.t->out) + j.off,
                  static_cast<const char*>(sv->hbase) + j.t->host_off + j.off, (size_t)j.len);
    if (j.t->chunks_left.fetch_sub(1) == 1) {
      std::lock_guard<std::mutex> lk(sv->mu);
      server_finish(*sv, *j.t);
      if (sv->stop && sv->idle()) sv->cv_copy.notify_all();
    }
  }
}

void server_loop(Server* sv) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (sv->pending.empty() && sv->inflight.empty()) {
        if (sv->stop) return;
        // copies in progress free staging blocks and wake this thread through cv_work
        sv->cv_work.wait(lk, [&] { return sv->stop || !sv->pending.empty(); });
        continue;
      }
    }
    if (!server_step(*sv)) {
      std::unique_lock<std::mutex> lk(sv->mu);
      sv->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
  }
}

// Stop a server: queued requests fail with FOMO_ERR_STATE, running ones complete.
void server_shutdown(Server& sv) {
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    sv.stop = true;
    for (int64_t id : sv.pending) server_fail_ticket(sv, sv.tickets[id], FOMO_ERR_STATE, "server destroyed");
    sv.pending.clear();
    sv.cv_work.notify_all();
    sv.cv_done.notify_all();
  }
  if (sv.worker.joinable()) sv.worker.join();
  {
    std::lock_guard<std::mutex> lk(sv.mu);
    sv.cv_copy.notify_all();
  }
  for (auto& c : sv.copiers)
    if (c.joinable()) c.join();
}

std::shared_ptr<Server> find_server(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  auto it = g_ctx.servers.find(id);
  return it == g_ctx.servers.end() ? nullptr : it->second;
}

// Re-index a dataset (new cell scale or lookup subdivision) and carry its emission cubes and their
// per-point records over to the new sorted order, so cubes made before stay valid.
int reindex_dataset(int id, Dataset& ds, double cell_scale) {
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds.n;
  DevBuf old_perm;
  if (old_perm.alloc(sizeof(int32_t) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (reindex)");
  HIPOK(hipMemcpyAsync(old_perm.p, ds.perm.p, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  if (int rc = build_index(ds, cell_scale)) return rc;
  DevBuf inv;
  if (inv.alloc(sizeof(int32_t) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (reindex)");
  k_invert_perm<<<nblk(n), kBlock, 0, s>>>(ds.perm.as<int32_t>(), n, inv.as<int32_t>());
  for (auto& kv : g_ctx.gofts) {
    Goft& gc = *kv.second;
    if (gc.dataset != id) continue;
    DevBuf em;
    if (em.alloc(sizeof(float2) * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (reindex)");
    k_reorder_f2<<<nblk(n), kBlock, 0, s>>>(gc.em.as<float2>(), old_perm.as<int32_t>(), inv.as<int32_t>(), n,
                                           em.as<float2>());
    std::swap(gc.em.p, em.p);
    std::swap(gc.em.bytes, em.bytes);
    if (gc.rec.p && gc.rs4 > 0) {
      DevBuf rec;
      if (rec.alloc(sizeof(float4) * gc.rs4 * n)) return fail(FOMO_ERR_NOMEM, "device allocation failed (reindex)");
      k_reorder_f4<<<nblk(n), kBlock, 0, s>>>(gc.rec.as<float4>(), old_perm.as<int32_t>(), inv.as<int32_t>(), n,
                                             gc.rs4, rec.as<float4>());
      std::swap(gc.rec.p, rec.p);
      std::swap(gc.rec.bytes, rec.bytes);
    }
    HIPOK(hipGetLastError());
  }
  return sync_stream();
}

int method_mode(int method, int* mode) {
  switch (method) {
    case FOMO_METHOD_NN: *mode = MODE_NN; return FOMO_OK;
    case FOMO_METHOD_THOMSON: *mode = MODE_THOMSON; return FOMO_OK;
    case FOMO_METHOD_TET: *mode = MODE_TET; return FOMO_OK;
    case FOMO_METHOD_LINEAR: *mode = MODE_LINEAR; return FOMO_OK;
    case FOMO_METHOD_PROJECTION: *mode = MODE_PROJ; return FOMO_OK;
    default: return fail(FOMO_ERR_ARG, "unknown method");
  }
}

// Allocation-free NN render: the kernels of fomo_render_nn with their per-view buffers carved
// from the `scratch` array (at Ctx::noalloc) instead of the grow-only context scratch, the image
// written to the caller's device buffer, everything enqueued on s without waiting. Views whose
// buffers do not fit the array are rejected. Nothing is allocated: no hipMalloc/hipFree, no
// events (kernel profiling must be off), no library routine that allocates.
int render_nn_noalloc(int goft, double l, double b, int nx, int ny, int nz, int nl, double width, int dn,
                      float* out, double* meta, hipStream_t s) {
  if (int rc = require_init()) return rc;
  if (!g_ctx.noalloc) return fail(FOMO_ERR_STATE, "the scratch array has no device address");
  if (g_ctx.profile)
    return fail(FOMO_ERR_STATE, "kernel profiling records events; turn it off for the allocation-free render");
  if (!out) return fail(FOMO_ERR_ARG, "null output");
  Goft* gc = nullptr;
  Dataset* ds = nullptr;
  if (int rc = resolve_render(goft, MODE_NN, nx, ny, nz, nl, &gc, &ds)) return rc;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  int64_t sz[kRenderBufs];
  render_sizes(MODE_NN, ds->n, R, S, nl, sz);
  sz[4] = 0;  // the image is the caller's
  if (fused_single_line(gc->rec.as<float4>(), gc->rs4, R, nl)) sz[2] = sz[3] = sz[6] = sz[7] = 0;  // no per-sample buffers
  int64_t off[kRenderBufs], total = 0;
  for (int k = 0; k < kRenderBufs; ++k) {
    off[k] = total;
    total += (sz[k] + 255) & ~(int64_t)255;
  }
  if (total > (int64_t)FOMO_NOALLOC_SCRATCH_BYTES) {
    char msg[160];
    std::snprintf(msg, sizeof msg, "view needs %lld bytes of scratch, more than the %lld of the allocation-free render",
                  (long long)total, (long long)FOMO_NOALLOC_SCRATCH_BYTES);
    return fail(FOMO_ERR_ARG, msg);
  }
  char* const base = static_cast<char*>(g_ctx.noalloc);
  const auto at = [&](int k) { return sz[k] ? static_cast<void*>(base + off[k]) : nullptr; };
  const RenderBufs rb{static_cast<double*>(at(0)),  static_cast<ViewGeom*>(at(1)), static_cast<int32_t*>(at(2)),
                      static_cast<float*>(at(3)),   out,                           static_cast<int32_t*>(at(5)),
                      static_cast<float2*>(at(6)), static_cast<uint8_t*>(at(7))};
  // the ray flags must read zero on entry; the array keeps whatever the last call left
  if (rb.rayflag) k_clear_u8<<<nblk(R), kBlock, 0, s>>>(rb.rayflag, R);
  HIPOK(hipGetLastError());
  if (int rc = enqueue_render(s, *ds, *gc, MODE_NN, l, b, nx, ny, nz, nl, width, dn, rb)) return rc;
  if (meta) k_geom_meta<<<1, 64, 0, s>>>(rb.geom, meta);
  HIPOK(hipGetLastError());
  return FOMO_OK;
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

const char* fomo_last_error(void) { return g_err.c_str(); }
const char* fomo_version(void) { return "fomo-mi355x 0.1"; }

int fomo_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (g_ctx.init) {
    if (device != g_ctx.device) return fail(FOMO_ERR_STATE, "already initialised on another device");
    return FOMO_OK;
  }
  int cnt = 0;
  HIPOK(hipGetDeviceCount(&cnt));
  if (device < 0 || device >= cnt) return fail(FOMO_ERR_ARG, "no such device");
  HIPOK(hipSetDevice(device));
  HIPOK(hipGetDeviceProperties(&g_ctx.prop, device));
  HIPOK(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  HIPOK(hipStreamCreateWithFlags(&g_ctx.stream2, hipStreamNonBlocking));
  // device address of the allocation-free render's scratch array on this device
  void* sp = nullptr;
  HIPOK(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
  g_ctx.noalloc = sp;
  g_ctx.device = device;
  g_ctx.init = true;
  return FOMO_OK;
}

int fomo_shutdown(void) {
  {
    // servers first, outside the global lock (their workers finish in-flight requests)
    std::vector<std::shared_ptr<Server>> svs;
    {
      std::lock_guard<std::recursive_mutex> lk(g_mu);
      for (auto& kv : g_ctx.servers) svs.push_back(kv.second);
      g_ctx.servers.clear();
    }
    for (auto& sv : svs) server_shutdown(*sv);
  }
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!g_ctx.init) return FOMO_OK;
  (void)hipStreamSynchronize(g_ctx.stream);
  collect_profile();
  g_ctx.batches.clear();
  g_ctx.gofts.clear();
  g_ctx.tables.clear();
  g_ctx.datasets.clear();
  g_ctx.vs.partial.release();
  g_ctx.vs.geom.release();
  g_ctx.vs.idx.release();
  g_ctx.vs.aux.release();
  g_ctx.vs.out.release();
  g_ctx.vs.flag.release();
  g_ctx.vs.sem.release();
  g_ctx.vs.rayflag.release();
  for (auto& b : g_ctx.sortkeys) b.release();
  g_ctx.sortvals.release();
  g_ctx.sorttmp.release();
  g_ctx.tmpf.release();
  for (auto e : g_ctx.evpool) (void)hipEventDestroy(e);
  g_ctx.evpool.clear();
  g_ctx.copy_pool.reset();
  if (g_ctx.stage) (void)hipHostFree(g_ctx.stage);
  g_ctx.stage = nullptr;
  for (auto& e : g_ctx.stage_ev) {
    if (e) (void)hipEventDestroy(e);
    e = nullptr;
  }
  (void)hipStreamDestroy(g_ctx.stream);
  (void)hipStreamDestroy(g_ctx.stream2);
  g_ctx.stream2 = nullptr;
  g_ctx.stream = nullptr;
  g_ctx.init = false;
  g_ctx.device = -1;
  return FOMO_OK;
}

int fomo_device_info(char* name, int name_len, double* hbm_bytes, int* compute_units) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (name && name_len > 0) std::snprintf(name, name_len, "%s (%s)", g_ctx.prop.name, g_ctx.prop.gcnArchName);
  if (hbm_bytes) *hbm_bytes = (double)g_ctx.prop.totalGlobalMem;
  if (compute_units) *compute_units = g_ctx.prop.multiProcessorCount;
  return FOMO_OK;
}

int fomo_dataset_create(int64_t n, int dim, const float* x, const float* y, const float* z, int nvars,
                        const float* const* vars, double cell_scale, int* out_id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (n < 1 || n > 2147483647LL) return fail(FOMO_ERR_ARG, "point count must be in [1, 2^31)");
  if (dim != 2 && dim != 3) return fail(FOMO_ERR_ARG, "dim must be 2 or 3");
  if (!x || !y || (dim == 3 && !z) || !out_id) return fail(FOMO_ERR_ARG, "null coordinate pointer");
  if (nvars < 0 || nvars > 16 || (nvars > 0 && !vars)) return fail(FOMO_ERR_ARG, "bad nvars");
  for (int v = 0; v < nvars; ++v)
    if (!vars[v]) return fail(FOMO_ERR_ARG, "null variable pointer");
  std::unique_ptr<Dataset> ds(new Dataset());
  ds->n = n;
  ds->nvars = nvars;
  hipStream_t s = g_ctx.stream;
  const float* src[3] = {x, y, z};
  for (int a = 0; a < 3; ++a) {
    if (ds->raw[a].alloc(sizeof(float) * n) != hipSuccess) return fail(FOMO_ERR_NOMEM, "device allocation failed");
    if (a == 2 && dim == 2) {
      HIPOK(hipMemsetAsync(ds->raw[a].p, 0, sizeof(float) * n, s));
    } else {
      HIPOK(hipMemcpyAsync(ds->raw[a].p, src[a], sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
  }
  for (int v = 0; v < nvars; ++v) {
    ds->rawvars.emplace_back(new DevBuf());
    if (ds->rawvars.back()->alloc(sizeof(float) * n) != hipSuccess) return fail(FOMO_ERR_NOMEM, "device allocation failed");
    HIPOK(hipMemcpyAsync(ds->rawvars.back()->p, vars[v], sizeof(float) * n, hipMemcpyHostToDevice, s));
  }
  int rc = build_index(*ds, cell_scale);
  if (rc) return rc;
  int id = g_ctx.next_id++;
  g_ctx.datasets[id] = std::move(ds);
  *out_id = id;
  return FOMO_OK;
}

int fomo_dataset_begin(int64_t n, int dim, int nvars, int* out_id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (n < 1 || n > 2147483647LL) return fail(FOMO_ERR_ARG, "point count must be in [1, 2^31)");
  if (dim != 2 && dim != 3) return fail(FOMO_ERR_ARG, "dim must be 2 or 3");
  if (nvars < 0 || nvars > 16) return fail(FOMO_ERR_ARG, "bad nvars");
  if (!out_id) return fail(FOMO_ERR_ARG, "null output id");
  Ctx::Ingest in;
  in.ds.reset(new Dataset());
  in.ds->n = n;
  in.ds->nvars = nvars;
  in.dim = dim;
  for (int a = 0; a < 3; ++a)
    if (in.ds->raw[a].alloc(sizeof(float) * n) != hipSuccess) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  if (dim == 2) HIPOK(hipMemsetAsync(in.ds->raw[2].p, 0, sizeof(float) * n, g_ctx.stream));
  for (int v = 0; v < nvars; ++v) {
    in.ds->rawvars.emplace_back(new DevBuf());
    if (in.ds->rawvars.back()->alloc(sizeof(float) * n) != hipSuccess)
      return fail(FOMO_ERR_NOMEM, "device allocation failed");
  }
  if (int rc = sync_stream()) return rc;
  const int id = g_ctx.next_id++;
  g_ctx.ingests[id] = std::move(in);
  *out_id = id;
  return FOMO_OK;
}

int fomo_dataset_append(int id, int64_t offset, int64_t count, const float* x, const float* y, const float* z,
                        const float* const* vars) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  auto it = g_ctx.ingests.find(id);
  if (it == g_ctx.ingests.end()) return fail(FOMO_ERR_NOTFOUND, "no dataset of that id is being filled");
  Ctx::Ingest& in = it->second;
  Dataset& ds = *in.ds;
  if (count < 0 || offset < 0 || offset + count > ds.n) return fail(FOMO_ERR_ARG, "chunk outside [0, n)");
  if (count == 0) return FOMO_OK;
  if (!x || !y || (in.dim == 3 && !z) || (ds.nvars > 0 && !vars)) return fail(FOMO_ERR_ARG, "null pointer");
  for (int v = 0; v < ds.nvars; ++v)
    if (!vars[v]) return fail(FOMO_ERR_ARG, "null variable pointer");
  hipStream_t s = g_ctx.stream;
  const size_t off = sizeof(float) * (size_t)offset, len = sizeof(float) * (size_t)count;
  const float* src[3] = {x, y, z};
  for (int a = 0; a < in.dim; ++a)
    HIPOK(hipMemcpyAsync(static_cast<char*>(ds.raw[a].p) + off, src[a], len, hipMemcpyHostToDevice, s));
  for (int v = 0; v < ds.nvars; ++v)
    HIPOK(hipMemcpyAsync(static_cast<char*>(ds.rawvars[v]->p) + off, vars[v], len, hipMemcpyHostToDevice, s));
  if (int rc = sync_stream()) return rc;  // the caller's chunk may be reused at once
  in.written.emplace_back(offset, count);
  return FOMO_OK;
}

int fomo_dataset_commit(int id, double cell_scale) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  auto it = g_ctx.ingests.find(id);
  if (it == g_ctx.ingests.end()) return fail(FOMO_ERR_NOTFOUND, "no dataset of that id is being filled");
  Ctx::Ingest& in = it->second;
  // every point written exactly once
  std::sort(in.written.begin(), in.written.end());
  int64_t next = 0;
  for (const auto& w : in.written) {
    if (w.first != next) return fail(FOMO_ERR_STATE, w.first > next ? "points missing before commit" : "overlapping chunks");
    next = w.first + w.second;
  }
  if (next != in.ds->n) return fail(FOMO_ERR_STATE, "points missing before commit");
  if (int rc = build_index(*in.ds, cell_scale)) return rc;
  g_ctx.datasets[id] = std::move(in.ds);
  g_ctx.ingests.erase(it);
  return FOMO_OK;
}

int fomo_dataset_destroy(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  (void)hipStreamSynchronize(g_ctx.stream);
  if (g_ctx.ingests.erase(id)) return FOMO_OK;  // a dataset still being filled
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  g_ctx.datasets.erase(id);
  return FOMO_OK;
}

int fomo_dataset_info(int id, int64_t* n, int* dims3, int* maxcount) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (n) *n = ds->n;
  if (dims3) for (int a = 0; a < 3; ++a) dims3[a] = ds->grid.dims[a];
  if (maxcount) *maxcount = ds->maxcount;
  return FOMO_OK;
}

int fomo_dataset_set_ray_tile(int id, int width) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  int tl = 3;
  while (tl <= 6 && (1 << tl) != width) ++tl;
  if (tl > 6) return fail(FOMO_ERR_ARG, "ray tile width must be 8, 16, 32 or 64");
  ds->tile_log2 = tl;
  return FOMO_OK;
}

int fomo_dataset_tet_stats(int id, int64_t* out2) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (!out2) return fail(FOMO_ERR_ARG, "null output");
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  out2[0] = out2[1] = 0;
  if (!ds->tetstats.p) return FOMO_OK;
  unsigned long long v[2];
  HIPOK(hipMemcpyAsync(v, ds->tetstats.p, sizeof(v), hipMemcpyDeviceToHost, g_ctx.stream));
  if (int rc = sync_stream()) return rc;
  out2[0] = (int64_t)v[0];
  out2[1] = (int64_t)v[1];
  return FOMO_OK;
}

int fomo_dataset_set_far_budget(int id, int steps) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  if (steps < 0) return fail(FOMO_ERR_ARG, "far budget must be >= 0");
  ds->far_budget = steps;
  return FOMO_OK;
}

int fomo_dataset_set_lookup(int id, int subdiv) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (subdiv < 0 || subdiv > 4) return fail(FOMO_ERR_ARG, "lookup subdivision must be in [0,4]");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  ds->lookup_subdiv = subdiv;
  return build_lookup(*ds);
}

int fomo_dataset_lookup_info(int id, int* subdiv, int64_t* cells, int64_t* candidates, int64_t* fallback_cells) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (subdiv) *subdiv = ds->lk.enabled ? ds->lookup_active_subdiv : 0;
  if (cells) *cells = ds->nlcells;
  if (candidates) *candidates = ds->ncand;
  if (fallback_cells) *fallback_cells = ds->nfallback;
  return FOMO_OK;
}

int fomo_dataset_set_rectilinear(int id, int on) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  ds->allow_rect = on != 0;
  return reindex_dataset(id, *ds, ds->cell_scale);
}

int fomo_dataset_rect_info(int id, int* axes3) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (!axes3) return fail(FOMO_ERR_ARG, "null output");
  for (int a = 0; a < 3; ++a) axes3[a] = ds->rect ? ds->rect_n[a] : 0;
  return FOMO_OK;
}

int fomo_dataset_grid_info(int id, int* kind, int* axes3) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (!kind || !axes3) return fail(FOMO_ERR_ARG, "null output");
  *kind = ds->rect ? 1 : (ds->mrect ? 2 : 0);
  for (int a = 0; a < 3; ++a) axes3[a] = (ds->rect || ds->mrect) ? ds->rect_n[a] : 0;
  return FOMO_OK;
}

int fomo_dataset_lookup_groups(int id, int64_t* counts5) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (!counts5) return fail(FOMO_ERR_ARG, "null output");
  counts5[0] = ds->far_groups;
  counts5[1] = ds->far_lists;
  counts5[2] = ds->full_groups;
  counts5[3] = ds->outer_groups;
  counts5[4] = ds->outer_lists;
  return FOMO_OK;
}

int fomo_dataset_reindex(int id, double cell_scale) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, id);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  return reindex_dataset(id, *ds, cell_scale);
}

int fomo_table_create(int nt, int nrho, const float* logt, const float* logn, const float* g, int* out_id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (nt < 2 || nrho < 2 || !logt || !logn || !g || !out_id)
    return fail(FOMO_ERR_ARG, "table needs nt>=2, nrho>=2 and non-null arrays");
  std::unique_ptr<Table> t(new Table());
  t->nt = nt;
  t->nrho = nrho;
  if (t->logt.alloc(sizeof(float) * nt) || t->logn.alloc(sizeof(float) * nrho) ||
      t->g.alloc(sizeof(float) * (size_t)nt * nrho))
    return fail(FOMO_ERR_NOMEM, "device allocation failed");
  HIPOK(hipMemcpyAsync(t->logt.p, logt, sizeof(float) * nt, hipMemcpyHostToDevice, g_ctx.stream));
  HIPOK(hipMemcpyAsync(t->logn.p, logn, sizeof(float) * nrho, hipMemcpyHostToDevice, g_ctx.stream));
  HIPOK(hipMemcpyAsync(t->g.p, g, sizeof(float) * (size_t)nt * nrho, hipMemcpyHostToDevice, g_ctx.stream));
  if (int rc = sync_stream()) return rc;
  int id = g_ctx.next_id++;
  g_ctx.tables[id] = std::move(t);
  *out_id = id;
  return FOMO_OK;
}

int fomo_table_destroy(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (!g_ctx.tables.erase(id)) return fail(FOMO_ERR_NOTFOUND, "unknown table id");
  return FOMO_OK;
}

int fomo_emission_chianti(int dataset, int table, double lambda0, double atweight, int imaging, double abundratio,
                          int* out_goft) {
  return fomo_emission_chianti_interp(dataset, table, lambda0, atweight, imaging, abundratio, 0, out_goft);
}

int fomo_emission_chianti_interp(int dataset, int table, double lambda0, double atweight, int imaging,
                                 double abundratio, int interp, int* out_goft) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  Table* tb = find_obj(g_ctx.tables, table);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (!tb) return fail(FOMO_ERR_NOTFOUND, "unknown table id");
  if (ds->nvars < 2) return fail(FOMO_ERR_ARG, "emission needs variables n and T");
  if (!out_goft) return fail(FOMO_ERR_ARG, "null output id");
  if (!imaging && !(atweight > 0)) return fail(FOMO_ERR_ARG, "atomic weight must be > 0");
  if (interp != 0 && interp != 1) return fail(FOMO_ERR_ARG, "interp must be 0 (reference) or 1 (bilinear)");
  if (interp == 1 && tb->nrho < 2) return fail(FOMO_ERR_ARG, "bilinear G(T, n) needs at least 2 density rows");
  std::unique_ptr<Goft> gc(new Goft());
  gc->dataset = dataset;
  gc->kind = 0;
  gc->lambda0 = lambda0;
  if (gc->em.alloc(sizeof(float2) * ds->n)) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  // fomo-CHIANTI.cpp:224 width constant; :257-274 normalisation
  const double K = 2 * sqrt(2 * log(2)) * sqrt(FOMO_K_BOLTZ / FOMO_M_PROTON / (imaging ? 1.0 : atweight)) * lambda0 /
                   FOMO_C_LIGHT;
  const double normc = imaging ? 1.0 : (1. / FOMO_ALPHA) * abundratio;
  {
    ProfScope ps("k_emission_chianti", g_ctx.stream);
    const unsigned nb = std::min<unsigned>(nblk(ds->n), 4096);
    k_emission_chianti<<<nb, kBlock, 0, g_ctx.stream>>>(ds->svars[0]->as<float>(), ds->svars[1]->as<float>(), ds->n,
                                                         tb->logt.as<float>(), tb->logn.as<float>(), tb->g.as<float>(),
                                                         tb->nt, tb->nrho, K, normc, imaging, interp,
                                                         gc->em.as<float2>());
  }
  HIPOK(hipGetLastError());
  gc->K = imaging ? 0.0 : K;
  if (int rc = build_records(g_ctx.stream, *ds, {gc.get()}, gc->rec, &gc->rs4, &gc->ri)) return rc;
  if (int rc = sync_stream()) return rc;
  int id = g_ctx.next_id++;
  g_ctx.gofts[id] = std::move(gc);
  *out_goft = id;
  return FOMO_OK;
}

int fomo_emission_given(int dataset, int kind, double lambda0, int* out_goft) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (kind != 0 && kind != 1) return fail(FOMO_ERR_ARG, "kind must be 0 (line) or 1 (Thomson)");
  if (ds->nvars < 2) return fail(FOMO_ERR_ARG, "a GoftCube holds at least two variables");
  if (!out_goft) return fail(FOMO_ERR_ARG, "null output id");
  std::unique_ptr<Goft> gc(new Goft());
  gc->dataset = dataset;
  gc->kind = kind;
  gc->lambda0 = kind == 0 ? lambda0 : 0.0;
  if (gc->em.alloc(sizeof(float2) * ds->n)) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  k_pack_em<<<nblk(ds->n), kBlock, 0, g_ctx.stream>>>(ds->svars[0]->as<float>(), ds->svars[1]->as<float>(), ds->n,
                                                       gc->em.as<float2>());
  HIPOK(hipGetLastError());
  if (kind == 0) {
    gc->K = -1.0;  // the records' T slot (variable 1) holds the fwhm
    if (int rc = build_records(g_ctx.stream, *ds, {gc.get()}, gc->rec, &gc->rs4, &gc->ri)) return rc;
  }
  if (int rc = sync_stream()) return rc;
  int id = g_ctx.next_id++;
  g_ctx.gofts[id] = std::move(gc);
  *out_goft = id;
  return FOMO_OK;
}

int fomo_emission_thomson(int dataset, int* out_goft) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->nvars < 1) return fail(FOMO_ERR_ARG, "Thomson emission needs the density variable");
  if (!out_goft) return fail(FOMO_ERR_ARG, "null output id");
  std::unique_ptr<Goft> gc(new Goft());
  gc->dataset = dataset;
  gc->kind = 1;
  gc->lambda0 = 0;
  if (gc->em.alloc(sizeof(float2) * ds->n)) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  {
    ProfScope ps("k_emission_thomson", g_ctx.stream);
    const unsigned nb = std::min<unsigned>(nblk(ds->n), 4096);
    k_emission_thomson<<<nb, kBlock, 0, g_ctx.stream>>>(ds->pts.as<float4>(), ds->svars[0]->as<float>(), ds->n,
                                                         gc->em.as<float2>());
  }
  HIPOK(hipGetLastError());
  if (int rc = sync_stream()) return rc;
  int id = g_ctx.next_id++;
  g_ctx.gofts[id] = std::move(gc);
  *out_goft = id;
  return FOMO_OK;
}

int fomo_goft_destroy(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Goft* gc = find_obj(g_ctx.gofts, id);
  if (!gc) return fail(FOMO_ERR_NOTFOUND, "unknown emission id");
  if (gc->inuse.load() > 0) return fail(FOMO_ERR_STATE, "emission cube is in use by a server request");
  g_ctx.gofts.erase(id);
  return FOMO_OK;
}

int fomo_goft_read(int id, float* var0, float* var1) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Goft* gc = find_obj(g_ctx.gofts, id);
  if (!gc) return fail(FOMO_ERR_NOTFOUND, "unknown emission id");
  Dataset* ds = find_obj(g_ctx.datasets, gc->dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "dataset of this emission cube was destroyed");
  std::vector<float> em(2 * ds->n);
  std::vector<int32_t> perm(ds->n);
  HIPOK(hipMemcpyAsync(em.data(), gc->em.p, sizeof(float2) * ds->n, hipMemcpyDeviceToHost, g_ctx.stream));
  HIPOK(hipMemcpyAsync(perm.data(), ds->perm.p, sizeof(int32_t) * ds->n, hipMemcpyDeviceToHost, g_ctx.stream));
  if (int rc = sync_stream()) return rc;
  for (int64_t p = 0; p < ds->n; ++p) {
    if (var0) var0[perm[p]] = em[2 * p];
    if (var1) var1[perm[p]] = em[2 * p + 1];
  }
  return FOMO_OK;
}

static int render_one(int goft, int mode, double l, double b, int nx, int ny, int nz, int nl, double width, int dn,
                      float* out, double* meta) {
  if (int rc = require_init()) return rc;
  Goft* gc = nullptr;
  Dataset* ds = nullptr;
  if (int rc = resolve_render(goft, mode, nx, ny, nz, nl, &gc, &ds)) return rc;
  if (mode == MODE_TET || mode == MODE_LINEAR) {
    if (ds->inuse.load() > 0 && ds->ntets == 0)
      return fail(FOMO_ERR_STATE, "dataset is in use: triangulate it before serving tetra requests");
    if (int rc = build_tets(*ds)) return rc;
  }
  hipStream_t s = g_ctx.stream;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  int64_t sz[kRenderBufs];
  render_sizes(mode, ds->n, R, S, nl, sz);
  ViewScratch& vs = g_ctx.vs;
  HIPOK(vs.partial.ensure(sz[0]));
  HIPOK(vs.geom.ensure(sz[1]));
  HIPOK(vs.idx.ensure(sz[2]));
  HIPOK(vs.aux.ensure(sz[3]));
  HIPOK(vs.out.ensure(sz[4]));
  HIPOK(vs.flag.ensure(sz[5]));
  HIPOK(vs.sem.ensure(sz[6]));
  HIPOK(vs.rayflag.ensure_zeroed(sz[7], s));
  const RenderBufs rb{vs.partial.as<double>(), vs.geom.as<ViewGeom>(), vs.idx.as<int32_t>(), vs.aux.as<float>(),
                      vs.out.as<float>(),      vs.flag.as<int32_t>(),  vs.sem.as<float2>(),  vs.rayflag.as<uint8_t>()};
  if (int rc = enqueue_render(s, *ds, *gc, mode, l, b, nx, ny, nz, nl, width, dn, rb)) return rc;
  if (out)
    if (int rc = copy_out(out, rb.out, sz[4], s)) return rc;
  ViewGeom hg;
  HIPOK(hipMemcpyAsync(&hg, rb.geom, sizeof(ViewGeom), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  if (meta) geom_to_meta(hg, meta);
  return FOMO_OK;
}

int fomo_render_nn(int goft, double l, double b, int nx, int ny, int nz, int nl, double width, int dn, float* out,
                   double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return render_one(goft, MODE_NN, l, b, nx, ny, nz, nl, width, dn, out, meta);
}

int fomo_render_thomson(int goft, double l, double b, int nx, int ny, int nz, float* out, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return render_one(goft, MODE_THOMSON, l, b, nx, ny, nz, 1, 0.0, 0, out, meta);
}

int fomo_render_tet(int goft, double l, double b, int nx, int ny, int nz, int nl, double width, int dn, int linear,
                    float* out, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return render_one(goft, linear ? MODE_LINEAR : MODE_TET, l, b, nx, ny, nz, nl, width, dn, out, meta);
}

int fomo_render_projection(int goft, double l, double b, int nx, int ny, int nz, int nl, double width, int dn,
                           float* out, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  return render_one(goft, MODE_PROJ, l, b, nx, ny, nz, nl, width, dn, out, meta);
}

int fomo_render_nn_noalloc(int goft, double l, double b, int nx, int ny, int nz, int nl, double width, int dn,
                           float* out_dev, double* meta_dev, void* stream) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : g_ctx.stream;
  return render_nn_noalloc(goft, l, b, nx, ny, nz, nl, width, dn, out_dev, meta_dev, s);
}

int fomo_projection_pixels(int dataset, double l, double b, int nx, int ny, int nz, int32_t* out_pixel,
                           double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (int rc = check_res(nx, ny, nz, 1)) return rc;
  if (!out_pixel) return fail(FOMO_ERR_ARG, "null output");
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  hipStream_t s = g_ctx.stream;
  const int64_t n = ds->n;
  ViewScratch& vs = g_ctx.vs;
  HIPOK(vs.partial.ensure(sizeof(double) * 6 * (kMaxPartials + 1)));
  HIPOK(vs.geom.ensure(sizeof(ViewGeom)));
  HIPOK(vs.idx.ensure(sizeof(int32_t) * 2 * n));
  ViewGeom* geom = vs.geom.as<ViewGeom>();
  enqueue_view_geom(s, *ds, MODE_NN, l, b, nx, ny, nz, 0, nullptr, vs.partial.as<double>(), geom);
  uint32_t* key = reinterpret_cast<uint32_t*>(vs.idx.as<int32_t>());
  k_proj_keys<<<nblk(n), kBlock, 0, s>>>(ds->pts.as<float4>(), ds->perm.as<int32_t>(), n, geom, nx, ny, key,
                                         vs.idx.as<int32_t>() + n);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(out_pixel, key, sizeof(int32_t) * n, hipMemcpyDeviceToHost, s));
  ViewGeom hg;
  HIPOK(hipMemcpyAsync(&hg, geom, sizeof(ViewGeom), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  if (meta) geom_to_meta(hg, meta);
  return FOMO_OK;
}

// Autotuner: times the NN passes of one representative view (k_nn + k_nn_far, gathering the same
// per-point records a batch of these emission cubes gathers) for each candidate index
// configuration, and keeps the fastest if it beats the configuration on entry by more than 3 %.
// Candidates: cell-list scale x lookup subdivision.
int fomo_autotune(int ngoft, const int* gofts, double l, double b, int nx, int ny, int nz, int reps,
                  double* result) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (int rc = check_res(nx, ny, nz, 1)) return rc;
  if (ngoft < 1 || ngoft > kMaxLines || !gofts) return fail(FOMO_ERR_ARG, "1..8 emission cubes");
  Goft* gc0 = find_obj(g_ctx.gofts, gofts[0]);
  if (!gc0) return fail(FOMO_ERR_NOTFOUND, "unknown emission id");
  const int dsid = gc0->dataset;
  Dataset* ds = find_obj(g_ctx.datasets, dsid);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "dataset of this emission cube was destroyed");
  std::vector<Goft*> gcs;
  for (int i = 0; i < ngoft; ++i) {
    Goft* g = find_obj(g_ctx.gofts, gofts[i]);
    if (!g || g->dataset != dsid || g->kind != gc0->kind)
      return fail(FOMO_ERR_ARG, "emission cubes must exist, share one dataset and one kind");
    gcs.push_back(g);
  }
  if (ds->inuse.load() > 0) return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  if (reps < 1) reps = 1;
  const int mode = gc0->kind == 1 ? MODE_THOMSON : MODE_NN;
  hipStream_t s = g_ctx.stream;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  ViewScratch& vs = g_ctx.vs;
  HIPOK(vs.partial.ensure(sizeof(double) * 6 * (kMaxPartials + 1)));
  HIPOK(vs.geom.ensure(sizeof(ViewGeom)));
  HIPOK(vs.idx.ensure(sizeof(int32_t) * S));
  HIPOK(vs.aux.ensure(sizeof(float) * std::max<int64_t>(ds->n, S)));
  HIPOK(vs.sem.ensure(sizeof(float2) * S * ngoft));
  HIPOK(vs.rayflag.ensure_zeroed(nn_scratch_bytes(R, S), s));
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  auto time_nn = [&](double* ms) -> int {
    // records in the current sorted order, as a batch of these cubes builds them
    DevBuf rec;
    int rs4 = 0;
    RecInfo ri{};
    if (mode == MODE_NN) {
      std::vector<const Goft*> cg(gcs.begin(), gcs.end());
      if (int rc = build_records(s, *ds, cg, rec, &rs4, &ri)) return rc;
    }
    ViewGeom* geom = vs.geom.as<ViewGeom>();
    enqueue_view_geom(s, *ds, mode, l, b, nx, ny, nz, 0, vs.aux.as<float>(), vs.partial.as<double>(), geom);
    auto once = [&] {
      enqueue_nn(s, *ds, geom, nx, ny, nz, vs.idx.as<int32_t>(), vs.rayflag.as<uint8_t>(),
                 rs4 ? rec.as<float4>() : nullptr, rs4, rs4 ? ngoft : 0, rs4 ? vs.aux.as<float>() : nullptr,
                 rs4 ? vs.sem.as<float2>() : nullptr, ri);
    };
    once();  // warm-up
    HIPOK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) once();
    HIPOK(hipEventRecord(e1, s));
    HIPOK(hipEventSynchronize(e1));
    float t = 0;
    HIPOK(hipEventElapsedTime(&t, e0, e1));
    *ms = t / reps;
    return FOMO_OK;
  };
  // the cell scale and lookup subdivision do not apply to a rectilinear dataset's index
  const std::vector<double> scales = ds->rect ? std::vector<double>{ds->cell_scale} : std::vector<double>{0.7, 1.0, 1.4};
  const std::vector<int> subdivs = ds->rect ? std::vector<int>{ds->lookup_subdiv} : std::vector<int>{1, 2, 3};
  const int tlogs[] = {3, 4, 5, 6};  // ray tiles 8x32, 16x16, 32x8, 64x4
  const double cs0 = ds->cell_scale;
  const int sd0 = ds->lookup_subdiv, tl0 = ds->tile_log2;
  double base = 0, best = 1e300, bscale = cs0;
  int bsub = sd0, btl = tl0, rc = FOMO_OK;
  if ((rc = time_nn(&base))) goto done;
  for (double sc : scales)
    for (int sd : subdivs) {
      ds->lookup_subdiv = sd;
      if ((rc = reindex_dataset(dsid, *ds, sc))) goto done;
      if (!ds->rect && ds->lookup_active_subdiv != sd && sd > 1) continue;  // did not fit: same as a smaller one
      for (int tl : tlogs) {
        ds->tile_log2 = tl;
        double ms = 0;
        if ((rc = time_nn(&ms))) goto done;
        if (ms < best) best = ms, bscale = sc, bsub = sd, btl = tl;
      }
    }
  if (!(best < 0.97 * base)) best = base, bscale = cs0, bsub = sd0, btl = tl0;
  ds->lookup_subdiv = bsub;
  ds->tile_log2 = btl;
  rc = reindex_dataset(dsid, *ds, bscale);
done:
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    ds->lookup_subdiv = sd0;
    ds->tile_log2 = tl0;
    (void)reindex_dataset(dsid, *ds, cs0);
    return rc;
  }
  if (result) {
    result[0] = bscale;
    result[1] = bsub;
    result[2] = best;
    result[3] = base;
    result[4] = (double)(1 << btl);
  }
  return FOMO_OK;
}

int fomo_dataset_triangulate(int dataset, int64_t* ntets, double* seconds) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->ntets == 0 && ds->inuse.load() > 0)
    return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  if (int rc = build_tets(*ds)) return rc;
  if (ntets) *ntets = ds->ntets;
  if (seconds) *seconds = ds->tet_build_s;
  return FOMO_OK;
}

int fomo_tet_vertices(int dataset, double l, double b, int nx, int ny, int nz, int32_t* out_vertex, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (int rc = check_res(nx, ny, nz, 1)) return rc;
  if (!out_vertex) return fail(FOMO_ERR_ARG, "null output");
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  if (ds->ntets == 0 && ds->inuse.load() > 0)
    return fail(FOMO_ERR_STATE, "dataset is in use by a batch or a server request");
  if (int rc = build_tets(*ds)) return rc;
  hipStream_t s = g_ctx.stream;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  ViewScratch& vs = g_ctx.vs;
  HIPOK(vs.partial.ensure(sizeof(double) * 6 * (kMaxPartials + 1)));
  HIPOK(vs.geom.ensure(sizeof(ViewGeom)));
  HIPOK(vs.idx.ensure(sizeof(int32_t) * S));
  HIPOK(vs.aux.ensure(sizeof(float) * std::max<int64_t>(ds->n, S)));
  ViewGeom* geom = vs.geom.as<ViewGeom>();
  enqueue_view_geom(s, *ds, MODE_NN, l, b, nx, ny, nz, 0, vs.aux.as<float>(), vs.partial.as<double>(), geom);
  HIPOK(vs.sem.ensure(tet_redo_bytes(S)));
  uint32_t* redo = vs.sem.as<uint32_t>();
  if (int rc = enqueue_tet(s, *ds, geom, nx, ny, nz, false, nullptr, 0, RecInfo{}, 0, nullptr, nullptr,
                           vs.idx.as<int32_t>(), redo))
    return rc;
  int32_t* tmp = reinterpret_cast<int32_t*>(vs.aux.p);
  k_idx_to_orig<<<nblk(S), kBlock, 0, s>>>(vs.idx.as<int32_t>(), ds->perm.as<int32_t>(), R, nz, tmp);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(out_vertex, tmp, sizeof(int32_t) * S, hipMemcpyDeviceToHost, s));
  ViewGeom hg;
  HIPOK(hipMemcpyAsync(&hg, geom, sizeof(ViewGeom), hipMemcpyDeviceToHost, s));
  uint32_t rc2[2] = {0, 0};
  HIPOK(hipMemcpyAsync(rc2, redo, sizeof(rc2), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  if (meta) {
    geom_to_meta(hg, meta);
    meta[10] = (double)std::min<uint32_t>(rc2[0], (uint32_t)tet_redo_cap(S));  // walked again exactly
    meta[11] = (double)rc2[1];                                                 // left unresolved
  }
  return FOMO_OK;
}

int fomo_nn_indices(int dataset, double l, double b, int nx, int ny, int nz, int thomson, int32_t* out_idx,
                    double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (int rc = check_res(nx, ny, nz, 1)) return rc;
  if (!out_idx) return fail(FOMO_ERR_ARG, "null output");
  Dataset* ds = find_obj(g_ctx.datasets, dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "unknown dataset id");
  const int mode = thomson ? MODE_THOMSON : MODE_NN;
  hipStream_t s = g_ctx.stream;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  ViewScratch& vs = g_ctx.vs;
  HIPOK(vs.partial.ensure(sizeof(double) * 6 * (kMaxPartials + 1)));
  HIPOK(vs.geom.ensure(sizeof(ViewGeom)));
  HIPOK(vs.idx.ensure(sizeof(int32_t) * S));
  HIPOK(vs.aux.ensure(sizeof(float) * std::max<int64_t>(ds->n, S)));
  HIPOK(vs.rayflag.ensure_zeroed(nn_scratch_bytes(R, S), s));
  ViewGeom* geom = vs.geom.as<ViewGeom>();
  enqueue_view_geom(s, *ds, mode, l, b, nx, ny, nz, 0, vs.aux.as<float>(), vs.partial.as<double>(), geom);
  enqueue_nn(s, *ds, geom, nx, ny, nz, vs.idx.as<int32_t>(), vs.rayflag.as<uint8_t>());
  int32_t* tmp = reinterpret_cast<int32_t*>(vs.aux.p);
  k_idx_to_orig<<<nblk(S), kBlock, 0, s>>>(vs.idx.as<int32_t>(), ds->perm.as<int32_t>(), R, nz, tmp);
  HIPOK(hipGetLastError());
  HIPOK(hipMemcpyAsync(out_idx, tmp, sizeof(int32_t) * S, hipMemcpyDeviceToHost, s));
  ViewGeom hg;
  HIPOK(hipMemcpyAsync(&hg, geom, sizeof(ViewGeom), hipMemcpyDeviceToHost, s));
  const DeepQueue dq = deep_queue(vs.rayflag.as<uint8_t>(), R, S, ds->far_budget, ds->segbox.as<float4>());
  unsigned queued = 0;
  HIPOK(hipMemcpyAsync(&queued, dq.count, sizeof(unsigned), hipMemcpyDeviceToHost, s));
  if (int rc = sync_stream()) return rc;
  if (meta) {
    for (int a = 0; a < 6; ++a) meta[a] = hg.ext[a];
    meta[6] = hg.md;
    meta[7] = hg.pathlength;
    meta[8] = hg.scale;
    meta[9] = hg.apix;
    meta[10] = (double)queued;  // waves that reached the far budget (the first cap went to k_nn_deep)
    meta[11] = (double)dq.cap;
  }
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------
// batches: all views x lines in one captured graph
// ---------------------------------------------------------------------------------------
// Enqueue every view x line of a batch on stream s (captured into the graph, or eager when profiling).
// Enqueue every view x line of a batch. With s2 (graph capture): the NN pass of view v runs on s
// and the line integration of view v on s2, so the (memory-bound) NN search of view v+1 overlaps
// the (VALU-bound) integration of view v; per-sample buffers are double-buffered and the NN pass
// of view v+2 waits for the integration of view v. Without s2 everything runs in order on s.
static int enqueue_batch(Batch& bt, hipStream_t s, hipStream_t s2) {
  Dataset* ds = find_obj(g_ctx.datasets, bt.dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "dataset of this batch was destroyed");
  std::vector<Goft*> gs;
  for (int id : bt.gofts) {
    Goft* gc = find_obj(g_ctx.gofts, id);
    if (!gc) return fail(FOMO_ERR_NOTFOUND, "emission cube of this batch was destroyed");
    gs.push_back(gc);
  }
  const int nv = (int)bt.l.size();
  const int64_t R = (int64_t)bt.nx * bt.ny;
  const size_t S = (size_t)R * bt.nz;
  const size_t outv = (size_t)R * std::max(bt.nl, 1);
  const int L = (int)gs.size();
  if (s2 && (int)bt.ev.size() < 2 * nv) return fail(FOMO_ERR_STATE, "batch events missing");
  // extents of all views: one pass over the points per kViewsPerPass views
  const unsigned nbp = enqueue_views_extents(s, *ds, bt.l, bt.b, bt.partial.as<double>());
  if (bt.fused) {
    // every view's geometry in one launch, then every view's fused pass in one launch (blockIdx.y =
    // view): no launch gaps or per-view tails between the views
    int32_t* const dlist = bt.flag.as<int32_t>() + (size_t)nv * R * kMaxLines;
    k_view_finalize_views<<<nv, kBlock, 0, s>>>(bt.partial.as<double>(), 6 * kMaxPartials, (int)nbp, MODE_NN, bt.nx,
                                                bt.ny, bt.nz, bt.dn, bt.trig.as<double>(), bt.geom.as<ViewGeom>(),
                                                dlist);
    HIPOK(hipGetLastError());
    constexpr int kMaxViewsPerLaunch = 65535;  // grid y extent
    for (int v0 = 0; v0 < nv; v0 += kMaxViewsPerLaunch) {
      const int nvl = std::min(kMaxViewsPerLaunch, nv - v0);
      std::vector<const float2*> ems(L, nullptr);
      std::vector<float*> outs;
      std::vector<double> l0s;
      for (int gi = 0; gi < L; ++gi) {
        outs.push_back(bt.out.as<float>() + ((size_t)v0 * L + gi) * outv);
        l0s.push_back(gs[gi]->lambda0);
      }
      const LineSet ls = make_lineset(ems, outs, l0s, bt.width);
      if (!enqueue_nn_los(s, *ds, bt.geom.as<ViewGeom>() + v0, bt.nx, bt.ny, bt.nz, bt.rec.as<float4>(), bt.rs4, bt.ri,
                          ls, bt.nl, bt.flag.as<int32_t>() + (size_t)v0 * R * kMaxLines, dlist, true, nvl,
                          (int64_t)L * (int64_t)outv, R * kMaxLines))
        return fail(FOMO_ERR_STATE, "batch: the fused pass no longer applies");
    }
    return FOMO_OK;
  }
  for (int v = 0; v < nv; ++v) {
    const int buf = s2 ? (v & 1) : 0;
    float* svl = bt.aux.as<float>() + buf * S;
    float2* sem = bt.sem.as<float2>() + (size_t)buf * S * L;
    int32_t* flag = bt.flag.as<int32_t>() + (size_t)buf * R * kMaxLines;
    double* part = bt.partial.as<double>() + (size_t)v * 6 * kMaxPartials;
    ViewGeom* geom = bt.geom.as<ViewGeom>() + v;
    // the per-sample buffers of view v - 2 are free once its integration is done
    if (s2 && v >= 2) HIPOK(hipStreamWaitEvent(s, bt.ev[nv + v - 2], 0));
    enqueue_view_finalize(s, MODE_NN, bt.l[v], bt.b[v], bt.nx, bt.ny, bt.nz, bt.dn, part, nbp, geom);
    std::vector<const float2*> ems;
    std::vector<float*> outs;
    std::vector<double> l0s;
    for (int gi = 0; gi < L; ++gi) {
      ems.push_back(sem + (size_t)gi * S);
      outs.push_back(bt.out.as<float>() + ((size_t)v * L + gi) * outv);
      l0s.push_back(gs[gi]->lambda0);
    }
    LineSet ls = make_lineset(ems, outs, l0s, bt.width);
    enqueue_nn(s, *ds, geom, bt.nx, bt.ny, bt.nz, bt.idx.as<int32_t>() + buf * S, bt.rayflag.as<uint8_t>() + buf * nn_scratch_bytes(R, S),
               bt.rec.as<float4>(), bt.rs4, L, svl, sem, bt.ri);
    hipStream_t si = s;
    if (s2) {
      HIPOK(hipEventRecord(bt.ev[v], s));
      HIPOK(hipStreamWaitEvent(s2, bt.ev[v], 0));
      si = s2;
    }
    enqueue_los(si, (int)R, bt.nz, ls, svl, geom, bt.nl, flag);
    if (s2) HIPOK(hipEventRecord(bt.ev[nv + v], s2));
  }
  if (s2) HIPOK(hipStreamWaitEvent(s, bt.ev[2 * nv - 1], 0));  // join
  return FOMO_OK;
}

static int capture_batch(Batch& bt) {
  if (bt.exec) { (void)hipGraphExecDestroy(bt.exec); bt.exec = nullptr; }
  if (bt.graph) { (void)hipGraphDestroy(bt.graph); bt.graph = nullptr; }
  hipStream_t s = g_ctx.stream;
  const bool prof = g_ctx.profile;
  g_ctx.profile = false;  // no event records inside a capture
  HIPOK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_batch(bt, s, g_ctx.stream2);
  hipError_t ce = hipStreamEndCapture(s, &bt.graph);
  g_ctx.profile = prof;
  if (rc) return rc;
  if (ce != hipSuccess) return fail(FOMO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
  HIPOK(hipGraphInstantiate(&bt.exec, bt.graph, nullptr, nullptr, 0));
  return FOMO_OK;
}

int fomo_batch_create(int ngoft, const int* gofts, int nviews, const double* l, const double* b, int nx, int ny,
                      int nz, int nl, double width, int dn, int* out_batch) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (ngoft < 1 || !gofts || nviews < 1 || !l || !b || !out_batch) return fail(FOMO_ERR_ARG, "bad batch arguments");
  if (ngoft > kMaxLines) return fail(FOMO_ERR_ARG, "a batch takes at most 8 spectral lines");
  if (int rc = check_res(nx, ny, nz, nl)) return rc;
  std::unique_ptr<Batch> bt(new Batch());
  for (int i = 0; i < ngoft; ++i) {
    Goft* gc = find_obj(g_ctx.gofts, gofts[i]);
    if (!gc) return fail(FOMO_ERR_NOTFOUND, "unknown emission id in batch");
    if (gc->kind != 0) return fail(FOMO_ERR_STATE, "batches take spectral-line emission cubes");
    if (i == 0) bt->dataset = gc->dataset;
    else if (gc->dataset != bt->dataset) return fail(FOMO_ERR_ARG, "all emission cubes of a batch must share a dataset");
    bt->gofts.push_back(gofts[i]);
  }
  Dataset* ds = find_obj(g_ctx.datasets, bt->dataset);
  if (!ds) return fail(FOMO_ERR_NOTFOUND, "dataset of emission cube was destroyed");
  if (nl > 1 && ds->nvars < 5) return fail(FOMO_ERR_ARG, "spectroscopic render needs vx,vy,vz");
  bt->l.assign(l, l + nviews);
  bt->b.assign(b, b + nviews);
  bt->nx = nx; bt->ny = ny; bt->nz = nz; bt->nl = nl; bt->dn = dn; bt->width = width;
  const int64_t R = (int64_t)nx * ny, S = R * nz;
  const int nlo = std::max(nl, 1);
  {
    std::vector<const Goft*> cg;
    for (int gid : bt->gofts) cg.push_back(find_obj(g_ctx.gofts, gid));
    if (int rc = build_records(g_ctx.stream, *ds, cg, bt->rec, &bt->rs4, &bt->ri)) return rc;
    if (int rc = sync_stream()) return rc;
  }
  // whether the views take the fused pass (the same for all: one resolution and line set)
  {
    std::vector<const float2*> ems0(ngoft, nullptr);
    std::vector<float*> outs0(ngoft, nullptr);
    std::vector<double> l0s;
    for (int gid : bt->gofts) l0s.push_back(find_obj(g_ctx.gofts, gid)->lambda0);
    const LineSet ls0 = make_lineset(ems0, outs0, l0s, width);
    bt->fused = enqueue_nn_los(g_ctx.stream, *ds, nullptr, nx, ny, nz, bt->rec.as<float4>(), bt->rs4, bt->ri, ls0, nl,
                               nullptr, nullptr, false);
  }
  // the fused pass keeps ray flags per view and no per-sample buffers; the per-view passes
  // double-buffer both
  const bool fz = bt->fused;
  std::vector<double> trig(5 * (size_t)nviews);
  for (int v = 0; v < nviews; ++v) view_trig(l[v], b[v], trig.data() + 5 * v);
  if (bt->partial.alloc(sizeof(double) * 6 * kMaxPartials * nviews) || bt->geom.alloc(sizeof(ViewGeom) * nviews) ||
      bt->trig.alloc(sizeof(double) * trig.size()) ||
      (!fz && (bt->sem.alloc(sizeof(float2) * S * ngoft * 2) || bt->aux.alloc(sizeof(float) * S * 2) ||
               bt->idx.alloc(sizeof(int32_t) * S * 2) || bt->rayflag.alloc(nn_scratch_bytes(R, S) * 2))) ||
      bt->flag.alloc(sizeof(int32_t) * (fz ? R * kMaxLines * nviews + nn_los_list_ints(R, nviews)
                                           : R * kMaxLines * 2)) ||
      bt->out.alloc(sizeof(float) * (size_t)R * nlo * ngoft * nviews))
    return fail(FOMO_ERR_NOMEM, "device allocation failed for batch");
  HIPOK(hipMemcpyAsync(bt->trig.p, trig.data(), sizeof(double) * trig.size(), hipMemcpyHostToDevice, g_ctx.stream));
  if (!fz) HIPOK(hipMemsetAsync(bt->rayflag.p, 0, nn_scratch_bytes(R, S) * 2, g_ctx.stream));
  for (int i = 0; i < 2 * nviews; ++i) {
    hipEvent_t e;
    HIPOK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    bt->ev.push_back(e);
  }
  if (int rc = capture_batch(*bt)) return rc;
  int id = g_ctx.next_id++;
  ds->inuse.fetch_add(1);
  g_ctx.batches[id] = std::move(bt);
  *out_batch = id;
  return FOMO_OK;
}

int fomo_batch_launch(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  HIPOK(hipGraphLaunch(bt->exec, g_ctx.stream));
  return FOMO_OK;
}

int fomo_batch_sync(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  HIPOK(hipStreamSynchronize(g_ctx.stream));
  return FOMO_OK;
}

int fomo_batch_profile(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  const bool prof = g_ctx.profile;
  g_ctx.profile = true;
  int rc = enqueue_batch(*bt, g_ctx.stream, nullptr);
  HIPOK(hipGetLastError());
  int rc2 = sync_stream();
  g_ctx.profile = prof;
  return rc ? rc : rc2;
}

int fomo_batch_run(int id) {
  int rc = fomo_batch_launch(id);
  if (rc) return rc;
  return fomo_batch_sync(id);
}

int fomo_batch_read(int id, float* out, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  const int nv = (int)bt->l.size();
  const size_t total = (size_t)bt->nx * bt->ny * std::max(bt->nl, 1) * bt->gofts.size() * nv;
  if (out)
    if (int rc = copy_out(out, bt->out.p, (int64_t)sizeof(float) * total, g_ctx.stream)) return rc;
  std::vector<ViewGeom> hg(nv);
  HIPOK(hipMemcpyAsync(hg.data(), bt->geom.p, sizeof(ViewGeom) * nv, hipMemcpyDeviceToHost, g_ctx.stream));
  HIPOK(hipStreamSynchronize(g_ctx.stream));
  if (meta)
    for (int v = 0; v < nv; ++v) {
      for (int a = 0; a < 6; ++a) meta[10 * v + a] = hg[v].ext[a];
      meta[10 * v + 6] = hg[v].md;
      meta[10 * v + 7] = hg[v].pathlength;
      meta[10 * v + 8] = hg[v].scale;
      meta[10 * v + 9] = hg[v].apix;
    }
  return FOMO_OK;
}

int fomo_batch_read_view(int id, int view, int line, float* out, double* meta) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  const int nv = (int)bt->l.size(), ng = (int)bt->gofts.size();
  if (view < 0 || view >= nv || line < 0 || line >= ng) return fail(FOMO_ERR_ARG, "view or line out of range");
  const size_t img = (size_t)bt->nx * bt->ny * std::max(bt->nl, 1);
  const float* src = bt->out.as<float>() + ((size_t)view * ng + line) * img;
  if (out)
    if (int rc = copy_out(out, src, (int64_t)sizeof(float) * img, g_ctx.stream)) return rc;
  ViewGeom hg;
  HIPOK(hipMemcpyAsync(&hg, bt->geom.as<ViewGeom>() + view, sizeof(ViewGeom), hipMemcpyDeviceToHost, g_ctx.stream));
  HIPOK(hipStreamSynchronize(g_ctx.stream));
  if (meta) geom_to_meta(hg, meta);
  return FOMO_OK;
}

int fomo_batch_destroy(int id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  (void)hipStreamSynchronize(g_ctx.stream);
  Batch* bt = find_obj(g_ctx.batches, id);
  if (!bt) return fail(FOMO_ERR_NOTFOUND, "unknown batch id");
  if (Dataset* ds = find_obj(g_ctx.datasets, bt->dataset)) ds->inuse.fetch_sub(1);
  g_ctx.batches.erase(id);
  return FOMO_OK;
}

// Bandwidth probes. k_copy_f4: a streaming copy in one pass over the buffer, each thread moving
// 4 float4 (stride n/4) with nontemporal loads and stores, so 4 loads per thread are in flight
// (a grid-stride loop with one float4 per thread in flight reaches about 80 % of this rate).
// k_gather_lines: random aligned 128-byte lines, 8 lanes per line and 2 lines per lane group in
// flight, the access pattern of lookup-record and point-record gathers. Both are in bounds by
// construction: every index is below n (copy) or the buffer's line count (gather).
typedef float fomo_v4f __attribute__((ext_vector_type(4)));
__global__ void k_copy_f4(const float4* __restrict__ a, float4* __restrict__ b, int64_t n) {
  const int64_t stride = (n + 3) / 4;
  const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i0 >= stride) return;
  const fomo_v4f* src = reinterpret_cast<const fomo_v4f*>(a);
  fomo_v4f* dst = reinterpret_cast<fomo_v4f*>(b);
  fomo_v4f v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (i0 + u * stride < n) v[u] = __builtin_nontemporal_load(src + i0 + u * stride);
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (i0 + u * stride < n) __builtin_nontemporal_store(v[u], dst + i0 + u * stride);
}

__device__ __forceinline__ uint64_t probe_mix(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__global__ void k_gather_lines(const float4* __restrict__ a, int64_t nlines, int64_t nitems, uint64_t seed,
                               float* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t item = t >> 3;
  const int sub = (int)(t & 7);
  float4 v[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int64_t it = 2 * item + r;
    const int64_t line = it < nitems ? (int64_t)(probe_mix((uint64_t)it + seed) % (uint64_t)nlines) : 0;
    v[r] = a[line * 8 + sub];
  }
  const float s = v[0].x + v[1].w;
  if (s == 1.2345e30f) out[0] = s;  // keeps the loads; never true for the zeroed buffer
}

int fomo_copy_probe(int64_t bytes, int reps, double* ms_per_copy) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (bytes < 64 || reps < 1) return fail(FOMO_ERR_ARG, "bytes >= 64 and reps >= 1 required");
  const int64_t n = bytes / 16;
  DevBuf a, b;
  if (a.alloc(n * 16) || b.alloc(n * 16)) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  HIPOK(hipMemsetAsync(a.p, 0, n * 16, g_ctx.stream));
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  const unsigned nb = (unsigned)(((n + 3) / 4 + 255) / 256);
  k_copy_f4<<<nb, 256, 0, g_ctx.stream>>>(a.as<float4>(), b.as<float4>(), n);  // warm
  HIPOK(hipEventRecord(e0, g_ctx.stream));
  for (int r = 0; r < reps; ++r) k_copy_f4<<<nb, 256, 0, g_ctx.stream>>>(a.as<float4>(), b.as<float4>(), n);
  HIPOK(hipEventRecord(e1, g_ctx.stream));
  HIPOK(hipEventSynchronize(e1));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_copy) *ms_per_copy = ms / reps;
  return FOMO_OK;
}

int fomo_gather_probe(int64_t bytes, int64_t lines, int reps, double* ms_per_gather) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (bytes < 128 || lines < 2 || reps < 1) return fail(FOMO_ERR_ARG, "bytes >= 128, lines >= 2, reps >= 1 required");
  const int64_t nlines = bytes / 128;
  const int64_t items = lines & ~(int64_t)1;
  DevBuf a, o;
  if (a.alloc(nlines * 128) || o.alloc(64)) return fail(FOMO_ERR_NOMEM, "device allocation failed");
  HIPOK(hipMemsetAsync(a.p, 0, nlines * 128, g_ctx.stream));
  hipEvent_t e0, e1;
  HIPOK(hipEventCreate(&e0));
  HIPOK(hipEventCreate(&e1));
  const unsigned nb = (unsigned)((items / 2 * 8 + 255) / 256);
  k_gather_lines<<<nb, 256, 0, g_ctx.stream>>>(a.as<float4>(), nlines, items, 0, o.as<float>());  // warm
  HIPOK(hipEventRecord(e0, g_ctx.stream));
  for (int r = 0; r < reps; ++r)
    k_gather_lines<<<nb, 256, 0, g_ctx.stream>>>(a.as<float4>(), nlines, items, (uint64_t)(r + 1) << 40, o.as<float>());
  HIPOK(hipEventRecord(e1, g_ctx.stream));
  HIPOK(hipEventSynchronize(e1));
  float ms = 0;
  HIPOK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms_per_gather) *ms_per_gather = ms / reps;
  return FOMO_OK;
}

int fomo_profile_enable(int on) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  g_ctx.profile = on != 0;
  return FOMO_OK;
}

int fomo_profile_reset(void) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  g_ctx.prof.clear();
  return FOMO_OK;
}

int fomo_profile_stats(const char* kernel, int64_t* count, double* total_ms) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!kernel) return fail(FOMO_ERR_ARG, "null kernel name");
  auto it = g_ctx.prof.find(kernel);
  if (count) *count = it == g_ctx.prof.end() ? 0 : it->second.count;
  if (total_ms) *total_ms = it == g_ctx.prof.end() ? 0.0 : it->second.ms;
  return FOMO_OK;
}

// ---------------------------------------------------------------------------------------
// render server
// ---------------------------------------------------------------------------------------
int fomo_server_footprint(int goft, int method, int nx, int ny, int nz, int nl, int64_t* device_bytes,
                          int64_t* host_bytes) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  int mode = 0;
  if (int rc = method_mode(method, &mode)) return rc;
  if (mode == MODE_THOMSON) nl = 1;
  Goft* gc = nullptr;
  Dataset* ds = nullptr;
  if (int rc = resolve_render(goft, mode, nx, ny, nz, nl, &gc, &ds)) return rc;
  int64_t sz[kRenderBufs];
  render_sizes(mode, ds->n, (int64_t)nx * ny, (int64_t)nx * ny * nz, nl, sz);
  if (device_bytes) *device_bytes = footprint_device(sz);
  if (host_bytes) *host_bytes = footprint_host(sz);
  return FOMO_OK;
}

int fomo_server_create(int64_t device_bytes, int64_t host_bytes, int max_inflight, int nstreams, int* out_id) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  if (!out_id || device_bytes < (1 << 20) || host_bytes < (1 << 16) || max_inflight < 1 || nstreams < 1 ||
      nstreams > 16)
    return fail(FOMO_ERR_ARG, "server needs >= 1 MiB device, >= 64 KiB host, max_inflight >= 1, 1..16 streams");
  auto sv = std::make_shared<Server>();
  sv->device = g_ctx.device;
  sv->max_inflight = max_inflight;
  if (hipMalloc(&sv->dbase, (size_t)device_bytes) != hipSuccess) {
    sv->dbase = nullptr;
    (void)hipGetLastError();
    return fail(FOMO_ERR_NOMEM, "cannot allocate the server's device arena");
  }
  if (hipHostMalloc(&sv->hbase, (size_t)host_bytes, hipHostMallocDefault) != hipSuccess) {
    sv->hbase = nullptr;
    (void)hipGetLastError();
    return fail(FOMO_ERR_NOMEM, "cannot allocate the server's pinned host arena");
  }
  sv->dev.reset(device_bytes);
  sv->host.reset(host_bytes);
  for (int i = 0; i < nstreams; ++i) {
    hipStream_t st = nullptr;
    HIPOK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    sv->streams.push_back(st);
    sv->load.push_back(0);
  }
  sv->worker = std::thread(server_loop, sv.get());
  for (int i = 0; i < kCopyThreads; ++i) sv->copiers.emplace_back(server_copier, sv.get());
  const int id = g_ctx.next_id++;
  g_ctx.servers[id] = sv;
  *out_id = id;
  return FOMO_OK;
}

int fomo_server_submit(int server, int goft, int method, double l, double b, int nx, int ny, int nz, int nl,
                       double lambda_width, int dn, float* out, double* meta, int64_t* out_ticket) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (int rc = require_init()) return rc;
  auto it = g_ctx.servers.find(server);
  if (it == g_ctx.servers.end()) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  Server& sv = *it->second;
  if (!out || !out_ticket) return fail(FOMO_ERR_ARG, "null output");
  Ticket t;
  if (int rc = method_mode(method, &t.mode)) return rc;
  if (t.mode == MODE_THOMSON) nl = 1, lambda_width = 0.0, dn = 0;
  if (int rc = resolve_render(goft, t.mode, nx, ny, nz, nl, &t.gc, &t.ds)) return rc;
  if ((t.mode == MODE_TET || t.mode == MODE_LINEAR) && t.ds->ntets == 0)
    return fail(FOMO_ERR_STATE, "tetra requests need a triangulated dataset (fomo_dataset_triangulate)");
  t.l = l, t.b = b, t.width = lambda_width;
  t.nx = nx, t.ny = ny, t.nz = nz, t.nl = nl, t.dn = dn;
  t.out = out;
  t.meta = meta;
  render_sizes(t.mode, t.ds->n, (int64_t)nx * ny, (int64_t)nx * ny * nz, nl, t.sz);
  t.dev_bytes = footprint_device(t.sz);
  t.host_bytes = footprint_host(t.sz);
  if (t.dev_bytes > sv.dev.capacity() || t.host_bytes > sv.host.capacity())
    return fail(FOMO_ERR_NOMEM, "request footprint exceeds the server's memory budget");
  t.ds->inuse.fetch_add(1);
  t.gc->inuse.fetch_add(1);
  std::lock_guard<std::mutex> slk(sv.mu);
  if (sv.stop) {
    ticket_release_objects(t);
    return fail(FOMO_ERR_STATE, "server is shutting down");
  }
  t.t_submit = sv.now();
  const int64_t id = (int64_t)sv.tickets.size();
  sv.tickets.push_back(std::move(t));
  sv.pending.push_back(id);
  *out_ticket = id;
  sv.cv_work.notify_one();
  return FOMO_OK;
}

int fomo_server_poll(int server, int64_t ticket, int* state) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (ticket < 0 || ticket >= (int64_t)sv->tickets.size()) return fail(FOMO_ERR_NOTFOUND, "unknown ticket");
  const Ticket& t = sv->tickets[ticket];
  if (state) *state = t.state;
  if (t.state < 0) return fail(t.state, t.err);
  return FOMO_OK;
}

int fomo_server_wait(int server, int64_t ticket, double timeout_s, int* state) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  std::unique_lock<std::mutex> lk(sv->mu);
  if (ticket < 0 || ticket >= (int64_t)sv->tickets.size()) return fail(FOMO_ERR_NOTFOUND, "unknown ticket");
  const Ticket& t = sv->tickets[ticket];
  auto finished = [&] { return t.state == kTicketDone || t.state < 0; };
  if (timeout_s < 0) sv->cv_done.wait(lk, finished);
  else sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
  if (state) *state = t.state;
  if (t.state < 0) return fail(t.state, t.err);
  if (t.state != kTicketDone) return fail(FOMO_ERR_STATE, "timed out");
  return FOMO_OK;
}

int fomo_server_cancel(int server, int64_t ticket) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (ticket < 0 || ticket >= (int64_t)sv->tickets.size()) return fail(FOMO_ERR_NOTFOUND, "unknown ticket");
  Ticket& t = sv->tickets[ticket];
  if (t.state != kTicketQueued) return fail(FOMO_ERR_STATE, "the request is no longer queued");
  sv->pending.erase(std::find(sv->pending.begin(), sv->pending.end(), ticket));
  server_fail_ticket(*sv, t, FOMO_ERR_CANCELLED, "cancelled before it was admitted");
  sv->cv_done.notify_all();
  sv->cv_work.notify_one();  // a request it held back may fit now
  return FOMO_OK;
}

int fomo_server_wait_all(int server, double timeout_s) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto idle = [&] { return sv->idle(); };
  if (timeout_s < 0) sv->cv_done.wait(lk, idle);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle))
    return fail(FOMO_ERR_STATE, "timed out");
  return FOMO_OK;
}

int fomo_server_times(int server, int64_t ticket, double* times3) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  if (!times3) return fail(FOMO_ERR_ARG, "null output");
  std::lock_guard<std::mutex> lk(sv->mu);
  if (ticket < 0 || ticket >= (int64_t)sv->tickets.size()) return fail(FOMO_ERR_NOTFOUND, "unknown ticket");
  const Ticket& t = sv->tickets[ticket];
  times3[0] = t.t_submit;
  times3[1] = t.state == kTicketQueued ? -1.0 : t.t_admit;
  times3[2] = (t.state == kTicketDone || t.state < 0) ? t.t_done : -1.0;
  return FOMO_OK;
}

int fomo_server_stats(int server, double* st) {
  auto sv = find_server(server);
  if (!sv) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
  if (!st) return fail(FOMO_ERR_ARG, "null output");
  std::lock_guard<std::mutex> lk(sv->mu);
  st[0] = (double)sv->tickets.size();
  st[1] = (double)sv->completed;
  st[2] = (double)sv->failed;
  st[3] = (double)sv->pending.size();
  st[4] = (double)(sv->inflight.size() + sv->copying);
  st[5] = (double)sv->peak_inflight;
  st[6] = (double)sv->dev.in_use();
  st[7] = (double)sv->dev.peak();
  st[8] = (double)sv->dev.capacity();
  st[9] = (double)sv->host.in_use();
  st[10] = (double)sv->host.peak();
  st[11] = (double)sv->host.capacity();
  return FOMO_OK;
}

int fomo_server_destroy(int server) {
  std::shared_ptr<Server> sv;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    if (int rc = require_init()) return rc;
    auto it = g_ctx.servers.find(server);
    if (it == g_ctx.servers.end()) return fail(FOMO_ERR_NOTFOUND, "unknown server id");
    sv = it->second;
    g_ctx.servers.erase(it);
  }
  server_shutdown(*sv);
  return FOMO_OK;
}

}  // extern "C"
