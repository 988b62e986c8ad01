// hipng: w-stacking NUFFT imaging core for MI355X (gfx950).
//
// Replaces the arithmetic RASCIL hands to ducc0.wgridder (ms2dirty / dirty2ms) in
// rascil/processing_components/imaging/ng.py:91-120 and :232-275. The algorithm is
// this library's own design (exponential-of-semicircle kernel, improved w-stacking,
// bucket sort by (w plane, uv tile), per-plane FFT), laid out for CDNA4:
//
//   k_extent       uvw extent reduction (one host sync to choose the plan)
//   k_digits / k_rekey  FP64 uvw -> grid coordinates: in-tile origin digit, then the
//                  (plane, tile) bucket key of each visibility in digit order
//   k_rs_*         stable LSD radix sort (8-bit digits, wave-ballot ranking): origin digit
//                  passes, then bucket key passes
//   k_bounds       bucket [start, end) table from the sorted keys
//   k_records      sorted, compact per-visibility records: packed first cell, FP32/FP64
//                  kernel offsets and (for invert) the weighted, phase-shifted values
//   k_tasks        per-tile task counts (chunks of a tile's candidate visibilities)
//   k_grid8        sub-tile gridder: persistent waves pop (tile, chunk, 8-column strip)
//                  tasks; candidates queued in LDS per 8 x 8 sub-tile, staged 64 at a time
//                  and accumulated on the matrix cores (FP32 or FP64), the pass's planes and
//                  all polarisations in registers; one store (or atomic, for split tiles) per cell
//   rocFFT         pruned plane transform: batched 1-D c2c along gy over the band of grid
//                  rows footprints can touch, k_tr_in transpose of the image's ny columns,
//                  batched 1-D c2c along gx (forward: the same in reverse order)
//   k_screen_acc_t fused w-screen exp(-2 pi i w_p (n-1)) * checkerboard * plane -> real
//                  image accumulator (FP64 except in FP32 mode), transposed image layout
//   k_finalize_t   grid correction 1/(Phi_x Phi_y Phi_w n) in FP64 with the transpose back
//   k_prescale_t / k_screen_fill_t / k_tr_out / k_degrid / k_unsort   the adjoint chain for
//                  predict; k_degrid is a gather (one lane per visibility, LDS-resident
//                  tile with halo), no atomics.
//   k_dft          FP64 point-source DFT (dft.py:181-258 / :261-281 semantics)

#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/hipng.h"

namespace {

constexpr double kSpeedOfLight = 299792458.0;  // rascil/phyconst.py:6
constexpr double kPi = 3.14159265358979323846;
constexpr int kBlock = 256;
constexpr int kXcd = 8;                      // XCDs per device: work queues of the sub-tile gridder
constexpr int kMaxBatch = 4;
constexpr size_t kLdsBudget = 65536;  // dynamic LDS per workgroup the tile kernels may use
constexpr size_t kDegridLds = 40960;  // LDS of a multi-plane degridder workgroup (4 per CU)

thread_local std::string g_err;

struct HipngError {
  int code;
  std::string msg;
};

#define HIPNG_THROW(code, ...)                           \
  do {                                                   \
    char _b[512];                                        \
    snprintf(_b, sizeof(_b), __VA_ARGS__);               \
    throw HipngError{code, std::string(_b)};             \
  } while (0)

#define HIP_TRY(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      HIPNG_THROW(HIPNG_ERUNTIME, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                  __FILE__, __LINE__);                                                  \
  } while (0)

#define LAUNCH_CHECK() HIP_TRY(hipGetLastError())

#define FFT_TRY(expr)                                                                    \
  do {                                                                                   \
    rocfft_status _s = (expr);                                                           \
    if (_s != rocfft_status_success)                                                     \
      HIPNG_THROW(HIPNG_ERUNTIME, "%s failed: rocfft status %d", #expr, (int)_s);       \
  } while (0)

// ---------------------------------------------------------------------------------
// Exponential-of-semicircle kernel table: psi(z) = exp(beta (sqrt(1 - z^2) - 1)), |z| < 1,
// support W cells, grid oversampling sigma. beta/W and the measured 1-D L2 error of a
// type-1 NUFFT (random points, exact sum as reference) per (sigma, W); generated by
// rascil_amd/csrc/gen_es_table.py.
// ---------------------------------------------------------------------------------
constexpr int kNSig = 5;
constexpr int kWMin = 4, kWMax = 16;
const double kSigmas[kNSig] = {1.25, 1.5, 1.75, 2.0, 2.5};
const double kBetaW[kNSig][13] = {
    {1.750, 1.775, 1.800, 1.825, 1.825, 1.850, 1.800, 1.800, 1.800, 1.825, 1.825, 1.850, 1.875},
    {1.950, 2.025, 2.025, 2.025, 2.050, 1.975, 2.050, 2.075, 2.025, 2.025, 2.075, 2.050, 2.050},
    {2.125, 2.150, 2.175, 2.175, 2.200, 2.125, 2.150, 2.175, 2.200, 2.225, 2.200, 2.200, 2.200},
    {2.225, 2.250, 2.300, 2.300, 2.325, 2.250, 2.250, 2.275, 2.325, 2.300, 2.325, 2.300, 2.325},
    {2.400, 2.425, 2.450, 2.475, 2.475, 2.425, 2.450, 2.500, 2.500, 2.450, 2.475, 2.375, 2.500}};
const double kErr1D[kNSig][13] = {
    {5.4e-3, 1.2e-3, 3.1e-4, 8.3e-5, 2.1e-5, 5.5e-6, 1.5e-6, 4.1e-7, 1.1e-7, 3.3e-8, 7.3e-9, 1.5e-9, 3.2e-10},
    {1.4e-3, 2.9e-4, 4.8e-5, 8.4e-6, 1.4e-6, 2.6e-7, 4.5e-8, 5.9e-9, 1.3e-9, 2.1e-10, 3.1e-11, 5.0e-12, 9.5e-13},
    {7.9e-4, 1.0e-4, 1.5e-5, 2.0e-6, 2.3e-7, 4.0e-8, 5.9e-9, 5.9e-10, 7.7e-11, 9.4e-12, 1.2e-12, 1.7e-13, 4.5e-14},
    {5.1e-4, 6.1e-5, 6.5e-6, 7.5e-7, 7.8e-8, 1.0e-8, 1.2e-9, 1.3e-10, 1.3e-11, 1.5e-12, 2.1e-13, 2.6e-14, 1.7e-14},
    {2.6e-4, 2.7e-5, 2.5e-6, 2.4e-7, 2.1e-8, 2.1e-9, 2.0e-10, 1.7e-11, 1.5e-12, 1.4e-13, 2.3e-14, 2.2e-14, 2.2e-14}};

// Gauss-Legendre nodes/weights on [-1, 1] (Newton on P_n), positive half returned.
void gauss_legendre_half(int n, std::vector<double>& x, std::vector<double>& w) {
  x.clear();
  w.clear();
  for (int i = 1; i <= n / 2; ++i) {
    double z = std::cos(kPi * (i - 0.25) / (n + 0.5));
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      double dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    double dp = n * (z * p1 - p0) / (z * z - 1.0);
    x.push_back(z);
    w.push_back(2.0 / ((1.0 - z * z) * dp * dp));
  }
}

constexpr int kQuadHalf = 64;  // 128-node Gauss-Legendre rule for the kernel's Fourier transform

// Phi(xi) = int phi(s) exp(2 pi i s xi) ds, phi(s) = psi(2 s / W), as sum_q c_q cos(pi W z_q xi).
struct KernelFT {
  std::vector<double> z, c;
  void init(int W, double beta) {
    // the rule depends only on its size: computed once (the Newton iterations cost ~0.2 ms,
    // and planning evaluates the transform for every candidate support and oversampling)
    static const std::pair<std::vector<double>, std::vector<double>> rule = [] {
      std::pair<std::vector<double>, std::vector<double>> r;
      gauss_legendre_half(2 * kQuadHalf, r.first, r.second);
      return r;
    }();
    const std::vector<double>& x = rule.first;
    const std::vector<double>& w = rule.second;
    z = x;
    c.resize(x.size());
    for (size_t q = 0; q < x.size(); ++q)
      c[q] = W * w[q] * std::exp(beta * (std::sqrt(1.0 - x[q] * x[q]) - 1.0));
  }
  double operator()(double xi, int W) const {
    double s = 0.0;
    for (size_t q = 0; q < z.size(); ++q) s += c[q] * std::cos(kPi * W * z[q] * xi);
    return s;
  }
};

bool good_factor(int64_t m) {
  for (int f : {2, 3, 5}) while (m % f == 0) m /= f;
  return m == 1;
}

// 2^a, 3 * 2^a or 5 * 2^a: the large transform lengths rocFFT runs at full speed on MI355X
// (measured at 4 x G^2 complex64: 12288, 16384, 20480 at 19-22 ps per point; 11520, 12800,
// 13824, 14400, 15360, 17280, 18432 at 42-49 ps per point)
bool fast_large(int64_t g) {
  while (g % 2 == 0) g /= 2;
  return g == 1 || g == 3 || g == 5;
}

// Smallest multiple of `tile` >= n whose quotient has only factors 2, 3, 5; from 4096 on
// also a fast rocFFT length.
int good_grid(double n, int tile) {
  int64_t m = (int64_t)std::ceil(n / tile);
  if (m < 2) m = 2;
  while (!good_factor(m) || (m * tile >= 4096 && !fast_large(m * tile))) ++m;
  return (int)(m * tile);
}

int bits_for(uint64_t v) {
  int b = 0;
  while (v) {
    ++b;
    v >>= 1;
  }
  return b;
}

// Mixed and FP32 modes (FP32 kernel, accumulation and FFTs) refuse epsilon below 1e-5, like
// ducc0 for single precision: the measured floor at 8192^2 is 1.7e-6 to 3.2e-6.
constexpr double kMixedEpsMin = 1e-5;

// FP32 mode: a grid coordinate fmaf((float)u, (float)s, G/2) carries up to 2^-24 |u s| of
// error from each rounded operand plus half an ulp of the result (cells); a point source at
// the image edge, npix/2 pixels out, sees 2 pi delta (npix/2) / G radians of phase error.
// The w coordinate (planes) works the same way with plane spacing dw = 1 / (2 sigma h),
// h = max |(n-1) - nshift|: 2 pi delta_w dw h = pi delta_w / sigma.
double fp32_coord_error(double cells, double top) {
  const double ulp = top > 0.0 ? std::ldexp(1.0, std::ilogb(top) - 23) : 0.0;
  return 2.0 * cells * std::ldexp(1.0, -24) + 0.5 * ulp;
}

// Roundoff of the plane transforms, amplified by the grid correction. The transforms' rounding
// (relative eps_r of a plane) lands evenly on the image, which is then divided by
// Phi_x Phi_y: at the image corner that is D = Phi(0)^2 / (Phi(nx/2/gx) Phi(ny/2/gy)) times
// the centre's gain, large for low oversampling (D ~ 1190 at sigma 1.25, W 8; 64 at 1.5, W 8;
// 4.7 at 1.75, W 6). Measured with FP32 planes: 1.0-1.2 eps_r D of extra relative L2 error on
// noise-like images (60 snapshot plans, sigma 1.25-2.5) and ~3e-6 (D = 64) at the
// 4.29e9-visibility 8192^2 plan; 1.25 eps_r D is used.
// With w-stacking the correction also divides by Phi_w(dw ((n-1) - nshift)), whose argument
// reaches 1/(2 sigma) at the image centre and corners whatever the grid rounding (the x / y
// arguments shrink when the grid is rounded up). The L2 error then follows the RMS over the
// image of A = Phi(0)^3 / (Phi_x Phi_y Phi_w): measured 0.3-1.3 eps_r rms(A) on 100 snapshot
// jobs (32^2-128^2, every oversampling), 1.3 where the w amplification is largest (sigma 1.25,
// W 9: Phi_w(0) / Phi_w(edge) = 53); 1.5 eps_r rms(A) is used when it exceeds the term above.
// rms(A) over a 24 x 24 midpoint grid of one image quadrant (A is even in x and y), Phi_w
// interpolated linearly in l^2 + m^2 from 65 nodes.
double transform_floor(int W, double beta, int nx, int ny, int gx, int gy, double eps_r, bool wstack = false,
                       double px = 0.0, double py = 0.0, double dw = 0.0, double nshift = 0.0) {
  KernelFT ft;
  ft.init(W, beta);
  const double p0 = ft(0.0, W);
  const double dx = p0 / ft(0.5 * nx / gx, W), dy = p0 / ft(0.5 * ny / gy, W);
  double fl = 1.25 * eps_r * dx * dy;
  if (wstack) {
    constexpr int kS = 24, kR = 64;
    const double xm = 0.5 * nx * px, ym = 0.5 * ny * py, r2max = xm * xm + ym * ym;
    double aw[kR + 1];
    for (int i = 0; i <= kR; ++i) {
      const double r2 = r2max * i / kR;
      const double nm1 = -r2 / (std::sqrt(1.0 - r2) + 1.0);
      aw[i] = p0 / ft(dw * (nm1 - nshift), W);
    }
    double ax[kS], ay[kS];
    for (int i = 0; i < kS; ++i) {
      ax[i] = p0 / ft((i + 0.5) / kS * 0.5 * nx / gx, W);
      ay[i] = p0 / ft((i + 0.5) / kS * 0.5 * ny / gy, W);
    }
    double s2 = 0.0;
    for (int i = 0; i < kS; ++i)
      for (int j = 0; j < kS; ++j) {
        const double x = (i + 0.5) / kS * xm, y = (j + 0.5) / kS * ym;
        const double t = (x * x + y * y) / r2max * kR;
        const int k0 = std::min(kR - 1, (int)t);
        const double a = ax[i] * ay[j] * (aw[k0] + (t - k0) * (aw[k0 + 1] - aw[k0]));
        s2 += a * a;
      }
    fl = std::max(fl, 1.5 * eps_r * std::sqrt(s2 / (kS * kS)));
  }
  return fl;
}

void make_plan_uncached(double umax, double vmax, double wmin, double wmax, int nx, int ny, double px,
                        double py, double eps, int wstack, int prec, int nbatch, int64_t nitems,
                        const hipng_plan_opts* opts, hipng_plan* P) {
  if (nx <= 0 || ny <= 0 || (nx & 1) || (ny & 1))
    HIPNG_THROW(HIPNG_EINVAL, "npix_x and npix_y must be positive and even (got %d, %d)", nx, ny);
  if (!(px > 0) || !(py > 0)) HIPNG_THROW(HIPNG_EINVAL, "pixel sizes must be positive");
  if (prec < 0 || prec > 2) HIPNG_THROW(HIPNG_EINVAL, "unknown precision mode %d", prec);
  if (nbatch < 1 || nbatch > kMaxBatch)
    HIPNG_THROW(HIPNG_EINVAL, "nbatch must be in [1, %d] (got %d)", kMaxBatch, nbatch);
  hipng_plan_opts o;
  std::memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  if (o.sigma != 0.0) {
    bool known = false;
    for (double sg : kSigmas) known = known || std::fabs(sg - o.sigma) < 1e-9;
    if (!known) HIPNG_THROW(HIPNG_EINVAL, "opts.sigma %g is not one of 1.25, 1.5, 1.75, 2, 2.5", o.sigma);
  }
  if (o.W != 0 && (o.W < kWMin || o.W > kWMax)) HIPNG_THROW(HIPNG_EINVAL, "opts.W %d outside [%d, %d]", o.W, kWMin, kWMax);
  if (o.tile != 0 && o.tile != 16 && o.tile != 32) HIPNG_THROW(HIPNG_EINVAL, "opts.tile must be 16 or 32 (got %d)", o.tile);
  if (o.use_extent) {
    if (!(o.umax >= 0.0) || !(o.vmax >= 0.0) || !(o.wmin >= 0.0) || !(o.wmax >= o.wmin))
      HIPNG_THROW(HIPNG_EINVAL, "opts extent must satisfy umax, vmax >= 0 and 0 <= wmin <= wmax");
    umax = std::max(umax, o.umax);
    vmax = std::max(vmax, o.vmax);
    wmin = nitems > 0 ? std::min(wmin, o.wmin) : o.wmin;
    wmax = std::max(wmax, o.wmax);
  }
  const double epsmin = prec == HIPNG_FP64 ? 1e-13 : kMixedEpsMin;
  if (!(eps >= epsmin) || !(eps < 1.0))
    HIPNG_THROW(HIPNG_EINVAL, "epsilon %g outside [%g, 1) for %s precision", eps, epsmin,
                prec == HIPNG_FP64 ? "fp64" : prec == HIPNG_MIXED ? "mixed" : "fp32");
  std::memset(P, 0, sizeof(*P));
  P->precision = prec;
  P->wstacking = wstack ? 1 : 0;
  P->nx = nx;
  P->ny = ny;
  P->pixsize_x = px;
  P->pixsize_y = py;
  P->umax = umax;
  P->vmax = vmax;
  P->wmin = wmin;
  P->wmax = wmax;
  P->epsilon = eps;
  // corner of the image: largest direction cosine radius
  const double xc = 0.5 * nx * px, yc = 0.5 * ny * py;
  const double r2 = xc * xc + yc * yc;
  if (wstack && !(r2 < 1.0))
    HIPNG_THROW(HIPNG_EINVAL, "image extends beyond the celestial sphere (l^2+m^2 = %g)", r2);
  const double nm1min = wstack ? -r2 / (std::sqrt(1.0 - r2) + 1.0) : 0.0;
  const double nm1half = -0.5 * nm1min;  // max |(n-1) - nshift|
  const int ndim = wstack ? 3 : 2;
  const double safety = 2.0 * std::sqrt((double)ndim);
  const size_t lds_budget = kLdsBudget;
  const size_t csz = prec == HIPNG_FP64 ? 16 : 8;
  // bytes per cell of the degridder's LDS tile: FP32 polarisations in pairs (16-byte units)
  const size_t cellb = csz * (size_t)(prec != HIPNG_FP64 && nbatch == 3 ? 4 : nbatch);
  double best_cost = 1e300;
  bool lds_limited = false;
  double fp32_floor_best = 1e300;  // smallest FP32 coordinate floor over the candidates
  for (int is = 0; is < kNSig; ++is) {
    if (o.sigma != 0.0 && std::fabs(kSigmas[is] - o.sigma) > 1e-9) continue;
    const double sigma = kSigmas[is];
    // smallest support whose kernel error (with the safety factor) plus the roundoff floor of
    // this mode reaches epsilon (or the pinned support, if it does)
    int W = -1, T = 16, gx = 0, gy = 0, nplanes = 1;
    double dw = 1.0, w0 = 0.0, floor_err = 0.0;
    for (int w = (o.W ? o.W : kWMin); w <= (o.W ? o.W : kWMax); ++w) {
      // tile: 16 x 16 by default (the sub-tile gridder scans a tile's candidates once per
      // 8 x 8 sub-tile, and a smaller tile has fewer candidates per sub-tile); 32 on request
      // when the degridder's LDS tile with halo (T+W-1)^2 * batch fits
      int t = o.tile == 32 ? 32 : 16;
      if ((size_t)(t + w - 1) * (t + w - 1) * cellb > (size_t)lds_budget) {
        if (o.tile == 32) continue;
        t = 16;
      }
      if ((size_t)(t + w - 1) * (t + w - 1) * cellb > (size_t)lds_budget) {
        lds_limited = true;
        continue;
      }
      if (t < w) continue;
      const int cgx = good_grid(std::max(sigma * nx, 2.0 * t), t);
      const int cgy = good_grid(std::max(sigma * ny, 2.0 * t), t);
      const double beta = kBetaW[is][w - kWMin] * w;
      const double eps_r = prec == HIPNG_FP64 ? std::ldexp(1.0, -53) : std::ldexp(1.0, -24);
      int np = 1;
      double cdw = 1.0, cw0 = 0.0;
      if (wstack) {
        // planes: a visibility at wh = (w - w0) / dw starts at s = floor(wh - W/2) + 1. With
        // w0 = wmin - (W/2 - 1/2) dw, s runs from floor(-1/2) + 1 = 0 at wmin to at most
        // floor(span) + 1 at wmax (half a plane of margin for rounding at both ends), so
        // span + W + 1 planes hold every footprint and none stays empty
        cdw = 0.5 / (sigma * nm1half);
        cw0 = wmin - (0.5 * w - 0.5) * cdw;
        const double span = std::floor((wmax - wmin) / cdw);
        if (!(span < 1e6)) continue;
        np = (int)span + w + 1;
      }
      double fl = transform_floor(w, beta, nx, ny, cgx, cgy, eps_r);
      // the w-stacking term costs ~100 kernel transforms: only for a support that passes without it
      if (wstack && kErr1D[is][w - kWMin] * safety + fl <= eps)
        fl = transform_floor(w, beta, nx, ny, cgx, cgy, eps_r, true, px, py, cdw, 0.5 * nm1min);
      if (prec == HIPNG_FP32) {
        const double ux = umax * px * cgx, vy = vmax * py * cgy;
        const double fx = 2.0 * kPi * fp32_coord_error(ux, 0.5 * cgx + ux + 0.5 * w) * (0.5 * nx) / cgx;
        const double fy = 2.0 * kPi * fp32_coord_error(vy, 0.5 * cgy + vy + 0.5 * w) * (0.5 * ny) / cgy;
        const double fw = wstack ? kPi * fp32_coord_error(wmax / cdw, (double)np) / sigma : 0.0;
        fl += std::max(fx, fy) + fw;
        fp32_floor_best = std::min(fp32_floor_best, fl);
      }
      if (kErr1D[is][w - kWMin] * safety + fl > eps) continue;
      W = w;
      T = t;
      gx = cgx;
      gy = cgy;
      nplanes = np;
      dw = cdw;
      w0 = cw0;
      floor_err = fl;
      break;
    }
    if (W < 0) continue;
    const double ntiles = (double)(gx / T) * (gy / T);
    if ((double)nplanes * ntiles >= 4294967295.0) continue;
    // cost model, constants measured on MI355X at the production shape (mixed precision,
    // 4 polarisations, 16 x 16 tiles): sub-tile gridder per visibility-plane-pol x 8x8
    // sub-tiles touched, degridder per cell-pol, pruned plane FFT with its transposes and
    // fill / screen per grid cell-pol (both directions)
    const double ww = wstack ? W : 1;
    const double sv = (1.0 + (double)(W - 1) / 8) * (1.0 + (double)(W - 1) / 8);
    const double g2 = (double)gx * gy;
    const double pf = prec == HIPNG_FP64 ? 2.0 : 1.0;
    const double cost = pf * (6.2e-12 * (double)nitems * nbatch * ww * sv +
                              2.8e-13 * (double)nitems * nbatch * ww * W * W +
                              2.0 * 7.4e-12 * nplanes * nbatch * g2 +
                              2.0 * 1.5e-12 * nplanes * nbatch * (g2 + (double)nx * ny));
    if (cost < best_cost) {
      best_cost = cost;
      P->W = W;
      P->tile = T;
      P->gx = gx;
      P->gy = gy;
      P->ntx = gx / T;
      P->nty = gy / T;
      P->nplanes = nplanes;
      P->beta = kBetaW[is][W - kWMin] * W;
      P->sigma = sigma;
      P->dw = dw;
      P->w0 = w0;
      P->nshift = wstack ? 0.5 * nm1min : 0.0;
      P->est_error = kErr1D[is][W - kWMin] * std::sqrt((double)ndim) + floor_err;
    }
  }
  if (P->W == 0 && lds_limited && nbatch > 1)
    HIPNG_THROW(HIPNG_EINVAL, "nbatch %d too large for the LDS tile at epsilon %g: split the batch", nbatch, eps);
  if (P->W == 0 && prec == HIPNG_FP32 && fp32_floor_best < 1e300)
    HIPNG_THROW(HIPNG_EINVAL,
                "epsilon %g is below the %.2e that FP32 grid coordinates and transforms reach for "
                "this image and uv extent: use mixed or fp64 precision",
                eps, fp32_floor_best);
  if (P->W == 0)
    HIPNG_THROW(HIPNG_EINVAL, "no kernel/grid reaches epsilon %g for this problem%s", eps,
                opts ? " with these plan options" : "");
  const uint64_t nkeys = (uint64_t)P->nplanes * P->ntx * P->nty;
  P->keybits = bits_for(nkeys - 1);
  if (P->keybits == 0) P->keybits = 1;
}

// The plan is a pure function of its arguments, and choosing it evaluates the kernel's Fourier
// transform for every candidate support and oversampling (0.5-2 ms on the host while the GPU
// waits, per one-shot call). The last few choices are kept; rejected arguments are not.
struct PlanArgs {
  double d[9];  // umax, vmax, wmin, wmax, px, py, eps, nitems, has_opts
  int32_t i[6];  // nx, ny, wstack, prec, nbatch, 0
  hipng_plan_opts o;
};
void make_plan(double umax, double vmax, double wmin, double wmax, int nx, int ny, double px,
               double py, double eps, int wstack, int prec, int nbatch, int64_t nitems,
               const hipng_plan_opts* opts, hipng_plan* P) {
  constexpr size_t kKeep = 16;
  static std::mutex mu;
  static std::vector<std::pair<PlanArgs, hipng_plan>> kept;
  PlanArgs a;
  std::memset(&a, 0, sizeof(a));
  const double dv[9] = {umax, vmax, wmin, wmax, px, py, eps, (double)nitems, opts ? 1.0 : 0.0};
  std::memcpy(a.d, dv, sizeof(dv));
  const int32_t iv[6] = {nx, ny, wstack, prec, nbatch, 0};
  std::memcpy(a.i, iv, sizeof(iv));
  if (opts) std::memcpy(&a.o, opts, sizeof(a.o));
  {
    std::lock_guard<std::mutex> lk(mu);
    for (const auto& e : kept)
      if (std::memcmp(&e.first, &a, sizeof(a)) == 0) {
        *P = e.second;
        return;
      }
  }
  make_plan_uncached(umax, vmax, wmin, wmax, nx, ny, px, py, eps, wstack, prec, nbatch, nitems, opts, P);
  std::lock_guard<std::mutex> lk(mu);
  if (kept.size() >= kKeep) kept.erase(kept.begin());
  kept.emplace_back(a, *P);
}

// ---------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------
template <typename R>
struct cplx {
  R re, im;
};

template <typename R>
struct rec3 {
  R x, y, w;
};

// In FP32, sqrt(1 - z^2) - 1 is evaluated as -z^2 / (1 + sqrt(1 - z^2)): the direct difference
// cancels near z = 0 and leaves an absolute error of ~beta * 6e-8 in the exponent of every tap
// (a kernel-shape error of ~1e-6 that the FP64 grid correction does not undo).
__device__ __forceinline__ float esk(float z, float beta) {
  const float z2 = z * z;
  const float t = fmaf(-z, z, 1.0f);
  return t > 0.0f ? expf(-beta * z2 / (1.0f + sqrtf(t))) : 0.0f;
}
__device__ __forceinline__ double esk(double z, double beta) {
  const double t = 1.0 - z * z;
  return t > 0.0 ? exp(beta * (sqrt(t) - 1.0)) : 0.0;
}

// Plan constants the kernels need, passed by value.
struct KP {
  int gx, gy, hgx, hgy;  // grid and half grid
  int T, ntx, nty;
  uint32_t ntiles;
  int W, Ww, nplanes, wstack;
  int nx, ny, nb, nchan;
  int tshift;                  // log2 T (T is 16 or 32)
  double rnchan;               // 1 / nchan, rounded (udiv32)
  double hgxd, hgyd, hw, mw0;  // G/2, W/2, -w0/dw
  double beta, px, py, nshift, dw, w0;
  // packed first cell ax | ay << 16 of a record: amask clears the dense-run flag dflag
  // (bit 31, set on records the dense-run gridder owns; both 0 / all ones when unused)
  uint32_t amask, dflag;
};

// q = i / d and *r = i - q d for i < 2^32 and 1 <= d < 2^31 without an integer divide:
// i * rd (rd = 1/d rounded) lies within 2^-20 of i / d, so its truncation is off by at most
// one and a single correction step either way gives the exact quotient
__device__ __forceinline__ uint32_t udiv32(uint32_t i, uint32_t d, double rd, uint32_t* r) {
  uint32_t q = (uint32_t)((double)i * rd);
  int rem = (int)(i - q * d);
  if (rem < 0) {
    --q;
    rem += (int)d;
  } else if (rem >= (int)d) {
    ++q;
    rem -= (int)d;
  }
  *r = (uint32_t)rem;
  return q;
}

// a mod g in [0, g): one add or subtract when |a| < 2 g (footprints wrap at most once unless
// the image is undersampled several times over), the remainder otherwise
__device__ __forceinline__ long long wrap_cell(long long a, int g) {
  if (a >= 0 && a < g) return a;
  if (a < 0 && a >= -(long long)g) return a + g;
  if (a >= g && a < 2LL * g) return a - g;
  const long long m = a % g;
  return m < 0 ? m + g : m;
}

struct ItemPos {
  long long ax, ay;
  int s;
  double fx, fy, fw;
  bool flip;
};

// The coordinate contract shared with the CPU reference sort: with F32C false every
// operation is an IEEE FP64 fma/sub/floor; with F32C true the same in FP32.
template <bool F32C>
__device__ __forceinline__ ItemPos item_pos(double u, double v, double w, double sxc, double syc,
                                            double swc, const KP& k) {
  ItemPos r;
  r.flip = w < 0.0;
  if (r.flip) {
    u = -u;
    v = -v;
    w = -w;
  }
  if (!F32C) {
    const double ux = fma(u, sxc, k.hgxd);
    const double a0 = floor(ux - k.hw) + 1.0;
    r.fx = a0 - ux;
    r.ax = (long long)a0;
    const double vy = fma(v, syc, k.hgyd);
    const double b0 = floor(vy - k.hw) + 1.0;
    r.fy = b0 - vy;
    r.ay = (long long)b0;
    if (k.wstack) {
      const double wh = fma(w, swc, k.mw0);
      const double s0 = floor(wh - k.hw) + 1.0;
      r.fw = s0 - wh;
      r.s = (int)s0;
    } else {
      r.fw = 0.0;
      r.s = 0;
    }
  } else {
    const float hw = (float)k.hw;
    const float ux = fmaf((float)u, (float)sxc, (float)k.hgxd);
    const float a0 = floorf(ux - hw) + 1.0f;
    r.fx = a0 - ux;
    r.ax = (long long)a0;
    const float vy = fmaf((float)v, (float)syc, (float)k.hgyd);
    const float b0 = floorf(vy - hw) + 1.0f;
    r.fy = b0 - vy;
    r.ay = (long long)b0;
    if (k.wstack) {
      const float wh = fmaf((float)w, (float)swc, (float)k.mw0);
      const float s0 = floorf(wh - hw) + 1.0f;
      r.fw = s0 - wh;
      r.s = (int)s0;
    } else {
      r.fw = 0.0;
      r.s = 0;
    }
  }
  r.ax = wrap_cell(r.ax, k.gx);
  r.ay = wrap_cell(r.ay, k.gy);
  if (r.s < 0) r.s = 0;
  if (r.s > k.nplanes - k.Ww) r.s = k.nplanes - k.Ww;
  return r;
}

__device__ __forceinline__ uint32_t item_key(const ItemPos& p, const KP& k) {
  return (uint32_t)p.s * k.ntiles + ((uint32_t)p.ax >> k.tshift) * (uint32_t)k.nty +
         ((uint32_t)p.ay >> k.tshift);
}

// ---------------------------------------------------------------------------------
// Kernels: extent, keys
// ---------------------------------------------------------------------------------
__global__ void k_extent(const double* __restrict__ uvw, int64_t nrow,
                         unsigned long long* __restrict__ out) {
  __shared__ double s[4][kBlock];
  double mu = 0.0, mv = 0.0, wlo = INFINITY, whi = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrow;
       r += (int64_t)gridDim.x * kBlock) {
    const double u = fabs(uvw[3 * r]), v = fabs(uvw[3 * r + 1]), w = fabs(uvw[3 * r + 2]);
    mu = fmax(mu, u);
    mv = fmax(mv, v);
    wlo = fmin(wlo, w);
    whi = fmax(whi, w);
  }
  s[0][threadIdx.x] = mu;
  s[1][threadIdx.x] = mv;
  s[2][threadIdx.x] = wlo;
  s[3][threadIdx.x] = whi;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[0][threadIdx.x] = fmax(s[0][threadIdx.x], s[0][threadIdx.x + o]);
      s[1][threadIdx.x] = fmax(s[1][threadIdx.x], s[1][threadIdx.x + o]);
      s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + o]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // non-negative doubles order like their bit patterns
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(s[0][0]));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(s[1][0]));
    atomicMin(&out[2], (unsigned long long)__double_as_longlong(s[2][0]));
    atomicMax(&out[3], (unsigned long long)__double_as_longlong(s[3][0]));
  }
}

// in-tile origin digit (first cell mod tile along u, then v): the secondary sort key
__device__ __forceinline__ uint32_t item_digit(const ItemPos& p, const KP& k) {
  return (((uint32_t)p.ax & (uint32_t)(k.T - 1)) << k.tshift) + ((uint32_t)p.ay & (uint32_t)(k.T - 1));
}

// Sort words of item i. Unpacked (db = 0): digit and index. Packed (db > 0, the digit's bit
// count): the bucket key rides along, its low 32 - db bits above the digit in the key word
// and the rest above the vb index bits of the value word, so that no later pass has to
// recompute it (see unpack_key).
template <bool F32C>
__global__ void k_digits(const double* __restrict__ uvw, int64_t nitems, const double* __restrict__ sx,
                         const double* __restrict__ sy, const double* __restrict__ sw, KP k,
                         uint32_t* __restrict__ digits, uint32_t* __restrict__ vals, int db, int vb) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nitems) return;
  uint32_t ch;
  const int64_t row = udiv32((uint32_t)i, (uint32_t)k.nchan, k.rnchan, &ch);
  const ItemPos p = item_pos<F32C>(uvw[3 * row], uvw[3 * row + 1], uvw[3 * row + 2], sx[ch],
                                   sy[ch], sw[ch], k);
  const uint32_t dg = item_digit(p, k);
  if (db == 0) {
    digits[i] = dg;
    vals[i] = (uint32_t)i;
  } else {
    const uint32_t key = item_key(p, k);
    digits[i] = (key << db) | dg;
    vals[i] = (uint32_t)i | ((key >> (32 - db)) << vb);
  }
}

// bucket key and item index from a packed pair of sort words (k_digits with db > 0)
__device__ __forceinline__ uint32_t unpack_key(uint32_t kw, uint32_t vw, int db, int vb) {
  return (kw >> db) | ((vw >> vb) << (32 - db));
}
__device__ __forceinline__ uint32_t unpack_index(uint32_t vw, int vb) {
  return vb >= 32 ? vw : vw & ((1u << vb) - 1u);
}

// bucket keys of the items in the order perm gives them
template <bool F32C>
__global__ void k_rekey(const double* __restrict__ uvw, const uint32_t* __restrict__ perm, int64_t nitems,
                        const double* __restrict__ sx, const double* __restrict__ sy,
                        const double* __restrict__ sw, KP k, uint32_t* __restrict__ keys) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= nitems) return;
  const uint32_t i = perm[j];
  uint32_t ch;
  const int64_t row = udiv32(i, (uint32_t)k.nchan, k.rnchan, &ch);
  const ItemPos p = item_pos<F32C>(uvw[3 * row], uvw[3 * row + 1], uvw[3 * row + 2], sx[ch],
                                   sy[ch], sw[ch], k);
  keys[j] = item_key(p, k);
}

// ---------------------------------------------------------------------------------
// Stable LSD radix sort, 8-bit digits. A block owns kRsItems consecutive items and
// walks them in rounds of 256 (thread order = item order); ranks inside a round come
// from wave ballots (lanes with equal digit) plus per-wave digit counts in LDS.
// ---------------------------------------------------------------------------------
constexpr int kRsIpt = 16;
constexpr int kRsItems = kBlock * kRsIpt;

// orders the LDS accesses of one wave (no workgroup barrier)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Per-block digit histogram of one pass: each wave counts its contiguous quarter of the
// block's items 64 at a time; lanes with equal digits are found by ballots and their leader
// adds the group size to the wave's own LDS histogram (no atomics).
// BITS-bit digits (8 or 9) from bit `shift`; db > 0: the input is packed (k_digits) and
// the pass sorts by the unpacked bucket key
// Logical block of a radix-sort pass: consecutive logical blocks run on one XCD (workgroups
// are placed on XCDs round-robin by index), so that the adjacent digit runs neighbouring
// blocks write go through one L2 and leave it as whole lines. The grid has kXcd * per
// workgroups; the few past nblocks return at once.
__device__ __forceinline__ int64_t rs_block(int64_t nblocks) {
  const int64_t per = (nblocks + kXcd - 1) / kXcd;
  return (int64_t)(blockIdx.x % kXcd) * per + (int64_t)(blockIdx.x / kXcd);
}
inline unsigned rs_grid(int64_t nblocks) { return (unsigned)(kXcd * ((nblocks + kXcd - 1) / kXcd)); }

template <int BITS>
__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals, int db, int vb, int64_t n,
                                                    int shift, uint32_t* __restrict__ hist, int64_t nblocks) {
  constexpr int NW = kBlock / 64;
  constexpr int R = kRsItems / kBlock;
  constexpr int ND = 1 << BITS, DPT = ND / kBlock;  // digits, digits per thread
  static_assert(DPT >= 1 && ND % kBlock == 0, "digits per thread");
  __shared__ uint32_t h[NW][ND];
  const int64_t lb = rs_block(nblocks);
  if (lb >= nblocks) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int w = 0; w < NW; ++w)
#pragma unroll
    for (int e = 0; e < DPT; ++e) h[w][threadIdx.x + e * kBlock] = 0;
  __syncthreads();
  const int64_t wbase = lb * kRsItems + (int64_t)wv * (kRsItems / NW);
  uint32_t key[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = wbase + r * 64 + lane;
    key[r] = i < n ? keys[i] : 0u;
  }
  if (db > 0) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int64_t i = wbase + r * 64 + lane;
      key[r] = unpack_key(key[r], i < n ? vals[i] : 0u, db, vb);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool ok = wbase + r * 64 + lane < n;
    const uint32_t d = (key[r] >> shift) & (uint32_t)(ND - 1);
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const unsigned long long m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    if (ok && lane == __ffsll((long long)peers) - 1) h[wv][d] += (uint32_t)__popcll(peers);
    wave_sync();
  }
  __syncthreads();
#pragma unroll
  for (int e = 0; e < DPT; ++e) {
    const int dd = threadIdx.x + e * kBlock;
    uint32_t tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) tot += h[w][dd];
    hist[(int64_t)dd * nblocks + lb] = tot;
  }
}


// exclusive scan helpers (uint32), 4096 elements per block
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// returns the exclusive prefix of v over the block and the block total
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan(v);
  if (lane == 63) wsum[wv] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int w = 0; w < kBlock / 64; ++w) {
    if (w < wv) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// A thread's kRsIpt consecutive elements are read (written) as 16-byte vectors when they are
// all in range (arrays are 16-byte aligned), element by element at the end of the array.
static_assert(kRsIpt % 4 == 0, "scan segments are whole 16-byte vectors");
__global__ void k_scan_partial(const uint32_t* __restrict__ a, int64_t m, uint32_t* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kRsItems + (int64_t)threadIdx.x * kRsIpt;
  uint32_t s = 0;
  if (base + kRsIpt <= m) {
    const uint4* p = reinterpret_cast<const uint4*>(a + base);
#pragma unroll
    for (int k = 0; k < kRsIpt / 4; ++k) {
      const uint4 v = p[k];
      s += v.x + v.y + v.z + v.w;
    }
  } else {
    for (int k = 0; k < kRsIpt; ++k)
      if (base + k < m) s += a[base + k];
  }
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// one block, kRsItems partials per step (kRsIpt consecutive ones per thread): a scan over the
// 2.6e5 partials of a 1.07e9-entry array takes 64 steps instead of 1024
__global__ void k_scan_top(uint32_t* __restrict__ part, int64_t np) {
  const bool al = ((uintptr_t)part & 15u) == 0;
  uint32_t carry = 0;
  for (int64_t c = 0; c < np; c += kRsItems) {
    const int64_t base = c + (int64_t)threadIdx.x * kRsIpt;
    const bool whole = al && base + kRsIpt <= np;
    uint32_t loc[kRsIpt];
    if (whole) {
      const uint4* q = reinterpret_cast<const uint4*>(part + base);
#pragma unroll
      for (int k = 0; k < kRsIpt / 4; ++k) {
        const uint4 v = q[k];
        loc[4 * k] = v.x;
        loc[4 * k + 1] = v.y;
        loc[4 * k + 2] = v.z;
        loc[4 * k + 3] = v.w;
      }
    } else {
#pragma unroll
      for (int k = 0; k < kRsIpt; ++k) loc[k] = base + k < np ? part[base + k] : 0u;
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kRsIpt; ++k) s += loc[k];
    uint32_t tot;
    uint32_t run = carry + block_excl_scan(s, &tot);
#pragma unroll
    for (int k = 0; k < kRsIpt; ++k) {
      const uint32_t v = loc[k];
      loc[k] = run;
      run += v;
    }
    if (whole) {
      uint4* q = reinterpret_cast<uint4*>(part + base);
#pragma unroll
      for (int k = 0; k < kRsIpt / 4; ++k) q[k] = make_uint4(loc[4 * k], loc[4 * k + 1], loc[4 * k + 2], loc[4 * k + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < kRsIpt; ++k)
        if (base + k < np) part[base + k] = loc[k];
    }
    carry += tot;
  }
}

__global__ void k_scan_apply(uint32_t* __restrict__ a, int64_t m, const uint32_t* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kRsItems + (int64_t)threadIdx.x * kRsIpt;
  const bool whole = base + kRsIpt <= m;
  uint32_t loc[kRsIpt];
  if (whole) {
    const uint4* p = reinterpret_cast<const uint4*>(a + base);
#pragma unroll
    for (int k = 0; k < kRsIpt / 4; ++k) {
      const uint4 v = p[k];
      loc[4 * k] = v.x;
      loc[4 * k + 1] = v.y;
      loc[4 * k + 2] = v.z;
      loc[4 * k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < kRsIpt; ++k) loc[k] = base + k < m ? a[base + k] : 0u;
  }
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kRsIpt; ++k) s += loc[k];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot) + part[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kRsIpt; ++k) {
    const uint32_t x = loc[k];
    loc[k] = run;
    run += x;
  }
  if (whole) {
    uint4* p = reinterpret_cast<uint4*>(a + base);
#pragma unroll
    for (int k = 0; k < kRsIpt / 4; ++k) p[k] = make_uint4(loc[4 * k], loc[4 * k + 1], loc[4 * k + 2], loc[4 * k + 3]);
  } else {
    for (int k = 0; k < kRsIpt; ++k)
      if (base + k < m) a[base + k] = loc[k];
  }
}

// Scatter of one pass. Wave w of block b ranks the contiguous items
// [b kRsItems + w kRsItems/4, ...) 64 at a time (ballots find lanes with equal digit), so
// wave-major order is item order and the result is stable. The block then places its items
// in LDS in digit order and writes each digit's run to consecutive output positions
// (coalesced) starting at the block's offset from the scanned histogram.
// BITS-bit digits (8 or 9). db > 0: the input is packed (k_digits); the pass sorts by the
// bucket key and writes the unpacked key and index
template <int BITS>
__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t* __restrict__ kin,
                                                       const uint32_t* __restrict__ vin,
                                                       uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       int64_t n, int shift, const uint32_t* __restrict__ hist,
                                                       int64_t nblocks, int db, int vb) {
  constexpr int NW = kBlock / 64;
  constexpr int R = kRsItems / kBlock;  // rounds of 64 items per wave
  constexpr int ND = 1 << BITS, DPT = ND / kBlock;  // digits, digits per thread
  static_assert(DPT >= 1 && ND % kBlock == 0, "digits per thread");
  __shared__ uint32_t s_k[kRsItems];
  __shared__ uint32_t s_v[kRsItems];
  __shared__ uint32_t wcnt[NW][ND];
  __shared__ uint32_t lbase[ND];
  __shared__ uint32_t gbase[ND];
  const int64_t lb = rs_block(nblocks);
  if (lb >= nblocks) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const int64_t base = lb * kRsItems;
  const int64_t wbase = base + (int64_t)wv * (kRsItems / NW);
  for (int w = 0; w < NW; ++w)
#pragma unroll
    for (int e = 0; e < DPT; ++e) wcnt[w][threadIdx.x + e * kBlock] = 0;
  __syncthreads();
  uint32_t key[R], val[R], rk[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int64_t i = wbase + r * 64 + lane;
    const bool ok = i < n;
    key[r] = ok ? kin[i] : 0u;
    val[r] = ok ? vin[i] : 0u;
    if (db > 0) {
      key[r] = unpack_key(key[r], val[r], db, vb);
      val[r] = unpack_index(val[r], vb);
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const bool ok = wbase + r * 64 + lane < n;
    const uint32_t d = (key[r] >> shift) & (uint32_t)(ND - 1);
    unsigned long long peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
      const unsigned long long m = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? m : ~m;
    }
    const int leader = ok ? __ffsll((long long)peers) - 1 : 0;
    uint32_t before = 0;
    if (ok && lane == leader) {
      before = wcnt[wv][d];
      wcnt[wv][d] = before + (uint32_t)__popcll(peers);
    }
    before = __shfl(before, leader, 64);
    rk[r] = before + (uint32_t)__popcll(peers & lt);
    wave_sync();
  }
  __syncthreads();
  {
    // digits DPT t .. DPT t + DPT - 1 = thread t: block totals, exclusive scan over digits,
    // per-wave offsets
    uint32_t tot[DPT], sum = 0;
#pragma unroll
    for (int e = 0; e < DPT; ++e) {
      const int d = DPT * threadIdx.x + e;
      tot[e] = 0;
      for (int w = 0; w < NW; ++w) tot[e] += wcnt[w][d];
      sum += tot[e];
    }
    uint32_t all;
    uint32_t ex = block_excl_scan(sum, &all);
#pragma unroll
    for (int e = 0; e < DPT; ++e) {
      const int d = DPT * threadIdx.x + e;
      lbase[d] = ex;
      gbase[d] = hist[(int64_t)d * nblocks + lb];
      uint32_t run = ex;
      for (int w = 0; w < NW; ++w) {
        const uint32_t c = wcnt[w][d];
        wcnt[w][d] = run;
        run += c;
      }
      ex += tot[e];
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (wbase + r * 64 + lane < n) {
      const uint32_t pos = wcnt[wv][(key[r] >> shift) & (uint32_t)(ND - 1)] + rk[r];
      s_k[pos] = key[r];
      s_v[pos] = val[r];
    }
  }
  __syncthreads();
  const int cnt = (int)min((int64_t)kRsItems, n - base);
  for (int idx = threadIdx.x; idx < cnt; idx += kBlock) {
    const uint32_t kk = s_k[idx];
    const uint32_t d = (kk >> shift) & (uint32_t)(ND - 1);
    const uint32_t g = gbase[d] + (uint32_t)idx - lbase[d];
    kout[g] = kk;
    vout[g] = s_v[idx];
  }
}

// FFT input check: a finite pseudo-random pattern in n reals, and the wrapping sum of the
// 32-bit words of n reals (integer sums: independent of the order of the additions)
template <typename R>
__global__ void k_pattern(R* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint32_t h = (uint32_t)(i * 2654435761ull) ^ (uint32_t)(i >> 17);
    a[i] = (R)((double)(h & 0xfffffu) * (1.0 / 1048576.0) - 0.5);
  }
}
__global__ void k_word_sum(const uint32_t* __restrict__ a, int64_t nw, unsigned long long* __restrict__ out) {
  __shared__ unsigned long long s[kBlock];
  unsigned long long acc = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nw; i += (int64_t)gridDim.x * kBlock)
    acc += (unsigned long long)a[i] * (unsigned long long)((i & 1023) + 1);
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, s[0]);
}

__global__ void k_fill_u32(uint32_t* __restrict__ a, int64_t n, uint32_t v) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    a[i] = v;
}

// Invert with the band transforms out of place: the gridded planes then hold only what the
// gridders wrote, and lastw[tile] records the last pass that wrote each tile. At pass pj only
// the tiles written at pass pj - 1 can be nonzero (every earlier one was cleared when it went
// stale); those are cleared unless the sub-tile gatherer stores them in full this pass.
template <typename R>
__global__ void k_zero_stale(KP k, int xlo, int xhi, int nplb, const uint32_t* __restrict__ task_off,
                             const uint32_t* __restrict__ counters, uint32_t solo,
                             const uint32_t* __restrict__ lastw, uint32_t prev, cplx<R>* __restrict__ plane) {
  // one block per tile of the band (tile rows from xlo / T); the whole block leaves when the
  // tile is not stale or the gatherer stores it in full this pass
  const int tx = xlo / k.T + (int)blockIdx.y, ty = (int)blockIdx.x;
  const uint32_t tile = (uint32_t)tx * (uint32_t)k.nty + (uint32_t)ty;
  if (lastw[tile] != prev) return;
  const uint32_t total = counters[1];
  const uint32_t nt = (tile + 1 < k.ntiles ? task_off[tile + 1] : total) - task_off[tile];
  if (nt == solo) return;
  for (int c = threadIdx.x; c < k.T * k.T; c += kBlock) {
    const int x = tx * k.T + c / k.T, y = ty * k.T + c % k.T;
    if (x < xlo || x >= xhi) continue;
    for (int i = 0; i < nplb; ++i) plane[((int64_t)i * k.gx + x) * k.gy + y] = cplx<R>{(R)0, (R)0};
  }
}

// lastw[tile] = pass for the tiles with sub-tile gatherer tasks this pass
__global__ void k_mark_tiles(KP k, const uint32_t* __restrict__ task_off, const uint32_t* __restrict__ counters,
                             uint32_t* __restrict__ lastw, uint32_t pass) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= k.ntiles) return;
  const uint32_t total = counters[1];
  const uint32_t nt = (t + 1 < k.ntiles ? task_off[t + 1] : total) - task_off[t];
  if (nt) lastw[t] = pass;
}

// clears the band rows [xlo, xhi) of the nplb planes except the cells of tiles with exactly
// `solo` tasks (those are stored in full by the sub-tile gatherer)
template <typename R>
__global__ void k_zero_shared(KP k, int xlo, int xhi, int nplb, const uint32_t* __restrict__ task_off,
                              const uint32_t* __restrict__ counters, uint32_t solo, cplx<R>* __restrict__ plane) {
  const int64_t nb = (int64_t)(xhi - xlo) * k.gy;
  const uint32_t total = counters[1];
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < nb; e += (int64_t)gridDim.x * kBlock) {
    const int x = xlo + (int)(e / k.gy), y = (int)(e - (e / k.gy) * k.gy);
    const uint32_t tile = (uint32_t)(x / k.T) * (uint32_t)k.nty + (uint32_t)(y / k.T);
    const uint32_t nt = (tile + 1 < k.ntiles ? task_off[tile + 1] : total) - task_off[tile];
    if (nt == solo) continue;
    for (int i = 0; i < nplb; ++i) plane[((int64_t)i * k.gx + x) * k.gy + y] = cplx<R>{(R)0, (R)0};
  }
}

template <typename T>
__global__ void k_zero(T* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
    a[i] = T{};
}

// zero columns [c0, c1) of nrows consecutive rows of length rowlen
template <typename T>
__global__ void k_zero_cols(T* __restrict__ a, int64_t nrows, int rowlen, int c0, int c1) {
  const int w = c1 - c0;
  const int64_t n = nrows * w;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const int64_t row = i / w;
    a[row * rowlen + c0 + (int)(i - row * w)] = T{};
  }
}


__global__ void k_bounds(const uint32_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ bstart,
                         uint32_t* __restrict__ bend) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t k = keys[j];
  if (j == 0 || keys[j - 1] != k) bstart[k] = (uint32_t)j;
  if (j == n - 1 || keys[j + 1] != k) bend[k] = (uint32_t)(j + 1);
}

// ---------------------------------------------------------------------------------
// Sorted records
// ---------------------------------------------------------------------------------
template <typename R>
struct VisT;
template <>
struct VisT<double> {
  using C = cplx<double>;
  using W = double;
};
template <>
struct VisT<float> {
  using C = cplx<float>;
  using W = float;
};

// turns -> (cos, sin)(2 pi turns); FP64 range reduction unless F32C
template <typename R, bool F32C>
__device__ __forceinline__ void phase_cs(double turns, R* c, R* s);
template <>
__device__ __forceinline__ void phase_cs<double, false>(double turns, double* c, double* s) {
  const double t = turns - rint(turns);
  sincospi(2.0 * t, s, c);
}
template <>
__device__ __forceinline__ void phase_cs<float, false>(double turns, float* c, float* s) {
  const double t = turns - rint(turns);
  sincospif((float)(2.0 * t), s, c);
}
template <>
__device__ __forceinline__ void phase_cs<float, true>(double turns, float* c, float* s) {
  float t = (float)turns;
  t = t - rintf(t);
  sincospif(2.0f * t, s, c);
}

// Per sorted item, from the geometry alone: packed first cell, kernel offsets, the phase
// exp(+2 pi i |w| nshift f/c) of the predict epilogue (the invert prologue applies its
// conjugate) and the hermitian flip bit (w < 0), packed 32 items per word. The phase and flip
// are kept so that the per-call prologue / epilogue read them in sorted order instead of
// gathering uvw rows again.
template <typename R, bool F32C>
__global__ void k_records(const double* __restrict__ uvw, const uint32_t* __restrict__ perm,
                          int64_t n, const double* __restrict__ sx, const double* __restrict__ sy,
                          const double* __restrict__ sw, const double* __restrict__ pn, KP k,
                          uint32_t* __restrict__ rec_a, rec3<R>* __restrict__ rec_f,
                          cplx<R>* __restrict__ rec_p, uint32_t* __restrict__ flipb) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const int lane = threadIdx.x & 63;
  bool flip = false;
  if (j < n) {
    const uint32_t i = perm[j];
    uint32_t ch;
    const int64_t row = udiv32(i, (uint32_t)k.nchan, k.rnchan, &ch);
    const double w = uvw[3 * row + 2];
    const ItemPos p = item_pos<F32C>(uvw[3 * row], uvw[3 * row + 1], w, sx[ch], sy[ch], sw[ch], k);
    rec_a[j] = (uint32_t)p.ax | ((uint32_t)p.ay << 16);
    rec3<R> f;
    f.x = (R)p.fx;
    f.y = (R)p.fy;
    f.w = (R)p.fw;
    rec_f[j] = f;
    flip = p.flip;
    cplx<R> ph{(R)1, (R)0};
    if (k.wstack) phase_cs<R, F32C>(fabs(w) * pn[ch], &ph.re, &ph.im);
    rec_p[j] = ph;
  }
  const unsigned long long m = __ballot(flip);
  const int64_t w0 = (j - lane) >> 5;  // waves start at multiples of 64 items
  if (lane == 0 && j < n) flipb[w0] = (uint32_t)m;
  if (lane == 32 && j < n) flipb[w0 + 1] = (uint32_t)(m >> 32);
}

// NB complex values of one item: 16-byte accesses when VEC (the caller checked alignment;
// FP32 with an even batch, or FP64)
template <typename R, int NB, bool VEC>
__device__ __forceinline__ void load_vals(const cplx<R>* __restrict__ p, cplx<R> (&v)[NB]) {
  if constexpr (VEC && std::is_same<R, float>::value && NB % 2 == 0) {
    const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
    for (int h = 0; h < NB / 2; ++h) {
      const float4 t = q[h];
      v[2 * h] = cplx<R>{t.x, t.y};
      v[2 * h + 1] = cplx<R>{t.z, t.w};
    }
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) v[b] = p[b];
  }
}
template <typename R, int NB, bool VEC>
__device__ __forceinline__ void store_vals(cplx<R>* __restrict__ p, const cplx<R> (&v)[NB]) {
  if constexpr (VEC && std::is_same<R, float>::value && NB % 2 == 0) {
    float4* q = reinterpret_cast<float4*>(p);
#pragma unroll
    for (int h = 0; h < NB / 2; ++h) q[h] = make_float4(v[2 * h].re, v[2 * h].im, v[2 * h + 1].re, v[2 * h + 1].im);
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) p[b] = v[b];
  }
}

// NB weights of one item: one 16-byte (NB 4) or 8-byte (NB 2) access when VEC (the caller
// checked the alignment). The item's weights sit at a random place, and one access per item
// instead of NB is what the gathers are bound by (k_values at C1: 42.5 -> 32.0 ms).
template <typename R, int NB, bool VEC>
__device__ __forceinline__ void load_wgts(const R* __restrict__ p, R (&g)[NB]) {
  if constexpr (VEC && std::is_same<R, float>::value && NB == 4) {
    const float4 t = *reinterpret_cast<const float4*>(p);
    g[0] = t.x;
    g[1] = t.y;
    g[2] = t.z;
    g[3] = t.w;
  } else if constexpr (VEC && std::is_same<R, float>::value && NB == 2) {
    const float2 t = *reinterpret_cast<const float2*>(p);
    g[0] = t.x;
    g[1] = t.y;
  } else {
#pragma unroll
    for (int b = 0; b < NB; ++b) g[b] = p[b];
  }
}

// invert prologue: rec_v[j] = conj?(wgt ms[perm[j]]) exp(-2 pi i |w| nshift f/c)
template <typename R, int NB, bool VEC>
__global__ void k_values(const uint32_t* __restrict__ perm, int64_t n, const cplx<R>* __restrict__ rec_p,
                         const uint32_t* __restrict__ flipb, const cplx<R>* __restrict__ ms,
                         const R* __restrict__ wgt, cplx<R>* __restrict__ rec_v) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  const bool flip = (flipb[j >> 5] >> (j & 31)) & 1u;
  const cplx<R> ph = rec_p[j];
  const R c = ph.re, s = -ph.im;
  cplx<R> v[NB], o[NB];
  load_vals<R, NB, VEC>(ms + (int64_t)i * NB, v);
  R gw[NB];
  if (wgt) load_wgts<R, NB, VEC>(wgt + (int64_t)i * NB, gw);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (flip) v[b].im = -v[b].im;
    if (wgt) {
      v[b].re *= gw[b];
      v[b].im *= gw[b];
    }
    o[b].re = v[b].re * c - v[b].im * s;
    o[b].im = v[b].re * s + v[b].im * c;
  }
  store_vals<R, NB, VEC>(rec_v + j * NB, o);
}

// ---------------------------------------------------------------------------------
// Work queue for the tile kernels. For plane p the candidate visibilities of a tile are
// the concatenation of its buckets (start plane s in [p - Ww + 1, p]; with the gridder
// also the three buckets of the left/lower neighbour tiles whose footprints reach in). The
// list is cut into chunks of `chunk` candidates; every chunk is one task. Persistent
// workgroups pop tasks from a device counter until the plane's tasks are exhausted.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ void plane_range(const KP& k, int p, int* slo, int* shi) {
  *slo = max(0, p - (k.Ww - 1));
  *shi = min(p, k.nplanes - k.Ww);
}

__device__ __forceinline__ uint32_t seg_key(const KP& k, int s, int tx, int ty, int q) {
  int stx = tx - (q >> 1), sty = ty - (q & 1);
  if (stx < 0) stx += k.ntx;
  if (sty < 0) sty += k.nty;
  return (uint32_t)s * k.ntiles + (uint32_t)stx * k.nty + (uint32_t)sty;
}

// Strip ranges of the sub-tile gatherer. A bucket's records are sorted by in-tile origin with
// the x origin ax % T first, so the records whose footprints can reach the 8 columns of strip
// s (x cells 8 s .. 8 s + 7 of the tile) form one contiguous range per bucket: own-tile and
// y-neighbour buckets need ax % T in [8 s - W + 1, 8 s + 7], x-neighbour and diagonal buckets
// (the tile to the left, offsets ax % T - T) need ax % T >= T + 8 s - W + 1. split[v][key] is
// the first record of bucket `key` with ax % T >= vals[v] (the bucket end if none); a bound
// code is a split index, or kBStart / kBEnd for the bucket's start / end, or kEmpty.
constexpr int kBStart = -1, kBEnd = -2, kEmpty = -3;
constexpr int kMaxSplit = 8;
struct Strips {
  int nv;
  int vals[kMaxSplit];
  int lo[4][2], hi[4][2];  // [strip][x-neighbour] bound codes
};

Strips make_strips(int T, int W) {
  Strips sr;
  std::memset(&sr, 0, sizeof(sr));
  std::vector<int> vals;
  auto code = [&](int v) -> int {
    if (v <= 0) return kBStart;
    if (v >= T) return kBEnd;
    for (size_t i = 0; i < vals.size(); ++i)
      if (vals[i] == v) return (int)i;
    vals.push_back(v);
    return (int)vals.size() - 1;
  };
  for (int st = 0; st < T / 8; ++st) {
    sr.lo[st][0] = code(8 * st - W + 1);
    sr.hi[st][0] = code(8 * st + 8);
    const int nlo = T + 8 * st - W + 1;
    if (nlo >= T) {
      sr.lo[st][1] = kEmpty;
      sr.hi[st][1] = kEmpty;
    } else {
      sr.lo[st][1] = code(nlo);
      sr.hi[st][1] = kBEnd;
    }
  }
  if ((int)vals.size() > kMaxSplit) HIPNG_THROW(HIPNG_ERUNTIME, "strip split table too large");
  sr.nv = (int)vals.size();
  for (int i = 0; i < sr.nv; ++i) sr.vals[i] = vals[i];
  return sr;
}

__device__ __forceinline__ uint32_t strip_bound(int code, uint32_t key, const uint32_t* __restrict__ bstart,
                                                const uint32_t* __restrict__ bend, const uint32_t* __restrict__ split,
                                                uint32_t nkeys) {
  if (code == kBStart) return bstart[key];
  if (code == kBEnd) return bend[key];
  return split[(size_t)code * nkeys + key];
}

// [start, end) of the records of bucket `key` that strip `st` (x-neighbour bucket: xn) needs
__device__ __forceinline__ void strip_range(const Strips& sr, int st, int xn, uint32_t key,
                                            const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
                                            const uint32_t* __restrict__ split, uint32_t nkeys, uint32_t* a,
                                            uint32_t* b) {
  const int lo = sr.lo[st][xn], hi = sr.hi[st][xn];
  if (lo == kEmpty) {
    *a = *b = 0u;
    return;
  }
  *a = strip_bound(lo, key, bstart, bend, split, nkeys);
  *b = strip_bound(hi, key, bstart, bend, split, nkeys);
  if (*b < *a) *b = *a;
}

__global__ void k_split_init(const uint32_t* __restrict__ bend, uint32_t nkeys, int nv, uint32_t* __restrict__ split) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (int64_t)nkeys * nv;
       i += (int64_t)gridDim.x * kBlock)
    split[i] = bend[i % nkeys];
}

// record i of the gatherer's list (srec entries, or all sorted records): its bucket from the
// sorted keys and its x origin; where the origin first reaches vals[v] inside the bucket,
// split[v][bucket] = i
__global__ void k_split_points(const uint32_t* __restrict__ keys, const uint2* __restrict__ srec,
                               const uint32_t* __restrict__ rec_a, int64_t nrec, int T, uint32_t amask, Strips sr,
                               uint32_t nkeys, uint32_t* __restrict__ split) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nrec) return;
  auto get = [&](int64_t r, uint32_t* key, int* ax) {
    uint32_t j = (uint32_t)r, A;
    if (srec) {
      const uint2 e = srec[r];
      j = e.x;
      A = e.y;
    } else {
      A = rec_a[r];
    }
    *key = keys[j];
    *ax = (int)((A & amask) & 0xffffu) & (T - 1);  // T is a power of two
  };
  uint32_t key, kp = 0xffffffffu;
  int a, ap = -1;
  get(i, &key, &a);
  if (i > 0) get(i - 1, &kp, &ap);
  if (kp != key) ap = -1;
  for (int v = 0; v < sr.nv; ++v)
    if (ap < sr.vals[v] && a >= sr.vals[v]) split[(size_t)v * nkeys + key] = (uint32_t)i;
}

// gatherer tasks: per tile the longest strip candidate list over the start planes [slo, shi]
// and the four buckets, in chunks, times the strips; counters[1] += ntask[t]
__global__ void k_tasks_grid(KP k, int slo, int shi, const uint32_t* __restrict__ bstart,
                             const uint32_t* __restrict__ bend, const uint32_t* __restrict__ split, Strips sr,
                             uint32_t chunk, uint32_t* __restrict__ ntask, uint32_t* __restrict__ counters) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)k.ntiles) return;
  const int tx = (int)(t / k.nty), ty = (int)(t - (int64_t)tx * k.nty);
  const uint32_t nkeys = (uint32_t)k.nplanes * k.ntiles;
  const int ns = k.T / 8;
  uint32_t best = 0;
  for (int st = 0; st < ns; ++st) {
    uint32_t cnt = 0;
    for (int s = slo; s <= shi; ++s)
      for (int q = 0; q < 4; ++q) {
        uint32_t a, b;
        strip_range(sr, st, q >> 1, seg_key(k, s, tx, ty, q), bstart, bend, split, nkeys, &a, &b);
        cnt += b - a;
      }
    best = max(best, cnt);
  }
  const uint32_t nt = (best + chunk - 1) / chunk * (uint32_t)ns;
  ntask[t] = nt;
  if (nt) atomicAdd(&counters[1], nt);
}

// ntask[t] = ceil(candidates / chunk) for start planes [slo, shi]; counters[1] += ntask[t]
__global__ void k_tasks(KP k, int slo, int shi, int nq, const uint32_t* __restrict__ bstart,
                        const uint32_t* __restrict__ bend, uint32_t chunk, uint32_t* __restrict__ ntask,
                        uint32_t* __restrict__ counters, uint32_t mult) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= (int64_t)k.ntiles) return;
  const int tx = (int)(t / k.nty), ty = (int)(t - (int64_t)tx * k.nty);
  uint32_t cnt = 0;
  for (int s = slo; s <= shi; ++s)
    for (int q = 0; q < nq; ++q) {
      const uint32_t key = seg_key(k, s, tx, ty, q);
      cnt += bend[key] - bstart[key];
    }
  const uint32_t nt = (cnt + chunk - 1) / chunk * mult;
  ntask[t] = nt;
  if (nt) atomicAdd(&counters[1], nt);
}

// tab[i] = tile of task i, for every task of the pass (task_off = exclusive scan of the per-tile
// task counts, counters[1] = their total): one load per task instead of a binary search
__global__ void k_task_tiles(const uint32_t* __restrict__ task_off, int64_t ntiles,
                             const uint32_t* __restrict__ counters, uint32_t* __restrict__ tab) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= ntiles) return;
  const uint32_t end = t + 1 < ntiles ? task_off[t + 1] : counters[1];
  for (uint32_t i = task_off[t]; i < end; ++i) tab[i] = (uint32_t)t;
}

// ---------------------------------------------------------------------------------
// Gridding constants and the plane window of a pass
// ---------------------------------------------------------------------------------
constexpr int kMaxPC = 2;                    // w planes gridded per pass
// FP32 modes: passes whose image-side transforms T2 holds at once, so that one update of the
// FP64 image accumulator (invert, kScreenGroup) or one read of the image (predict,
// kFillGroup) serves all of them. Measured per production-shape step: invert screens 46.3 /
// 35.7 / 30.5 ms for groups of 1 / 2 / 4; predict fills 33.0 / 27.9 / 30.1 ms.
constexpr int kScreenGroup = 4;
constexpr int kFillGroup = 2;
// Work-queue counters, each on a 128-byte line of its own (see queue_pop):
// counters[kQueueBase + x kQueueStride]: pops from XCD queue x
constexpr int kQueueStride = 32;
constexpr int kQueueBase = 32;
constexpr int kMaxSeg = 4 * (16 + kMaxPC - 1);  // 4 neighbour buckets x start planes
constexpr int kDenseCnt = kQueueBase + kXcd * kQueueStride;  // counters[kDenseCnt]: dense-run chunk pops
constexpr int kDenseTotal = kDenseCnt + kQueueStride;        // counters[kDenseTotal]: dense-run chunks
constexpr int kRuns = kDenseTotal + 1;                       // counters[kRuns]: runs (geometry build)
constexpr int kNCounters = kRuns + 1;

// Pop from a work-queue counter: the next index, or `lim` once it is exhausted. Atomics on
// one 128-byte line are served one after another, so each counter has a line of its own: with
// the eight XCD queues and the dense-run counter on one line, the pops of a launch (and the
// eight failed pops every wave makes when the queues run dry) cost ~0.4 ms per launch
// whatever the work, and the production-shape gatherer 16 % of its time. A relaxed load
// before the atomic to skip exhausted queues was slower still (loads of a line under atomic
// traffic wait for it).
__device__ __forceinline__ uint32_t queue_pop(uint32_t* c, uint32_t lim) { return min(atomicAdd(c, 1u), lim); }
// Shortest run the dense path takes: a dense task pays a flush of 64 x PC x NB x 2 global
// atomic adds (gridder) or a window load (degridder) per pass, which for short runs costs
// more than the sub-tile gatherer and the tile degridder spend on them (production shape,
// invert + predict: 3130 / 3182 / 3200 / 3204 / 3224 / 3181 / 3152 Mvis/s for minimum runs
// of 64 / 96 / 128 / 192 / 256 / 384 / 512).
constexpr uint32_t kDenseMin = 256;
// visibilities per dense-run task: a task pays one flush (gridder) or window load (degridder)
// per pass, so long tasks pay off until the longest runs unbalance the waves (dense-run
// gridder + degridder per production-shape step: 205 / 175 / 175 / 191 ms for tasks of at
// most 512 / 2048 / 8192 / 65536)
constexpr uint32_t kDenseChunk = 2048;

// start planes whose W_w-plane window meets planes [p0, p0 + npl)
__device__ __host__ __forceinline__ void chunk_range(int Ww, int nplanes, int p0, int npl, int* slo,
                                                     int* shi) {
  *slo = max(0, p0 - (Ww - 1));
  *shi = min(p0 + npl - 1, nplanes - Ww);
}

// ---------------------------------------------------------------------------------
// Gridder, sub-tile variant. A task is (tile, chunk, 8-column strip of the tile) and belongs
// to one wave; the wave grids the strip in groups of RG rows (NQ = RG / 8 sub-tiles of 8 x 8
// cells), all PC planes and nb polarisations of every cell in registers. Per row group the
// wave filters the chunk's candidates 64 at a time once and queues each overlapping one in
// LDS for every sub-tile of the group it meets (one queue per sub-tile); each full batch of
// 64 of a queue is staged with one visibility per lane (kernel taps of the sub-tile's 8
// columns and 8 rows, zero outside the support; w taps; values), then the wave accumulates
// all staged visibilities into that sub-tile without branches. Only wave-level ordering is
// used, so the waves of a workgroup never wait for each other.
// ---------------------------------------------------------------------------------

// per-wave LDS bytes of k_grid8: values [NB][64], taps [64][(16 + PC) | 1], nq queues of
// [2][128] words (record; packed offsets and start plane), segments. The tap row stride is
// odd so that the 64 staging lanes write to distinct banks.
__host__ __device__ constexpr int grid8_taps(int pc) { return (16 + pc) | 1; }
__host__ __device__ constexpr size_t grid8_region(size_t rsz, int pc, int nb, int nq) {
  return ((size_t)nb * 64 * 2 * rsz + (size_t)64 * grid8_taps(pc) * rsz + (size_t)nq * 2 * 128 * 4 +
          (2 * kMaxSeg + 1) * 4 + 15) / 16 * 16;
}

// exp(beta (sqrt(1 - z^2) - 1)) from the hardware base-2 exponential and square root
// (FP32 kernel modes; beta2 = beta log2(e))
__device__ __forceinline__ float esk_fast(float z, float beta2) {
  const float z2 = z * z;
  const float t = fmaf(-z, z, 1.0f);
  return t > 0.0f ? __builtin_amdgcn_exp2f(-beta2 * z2 * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_sqrtf(t)))
                  : 0.0f;
}
__device__ __forceinline__ double esk_fast(double z, double beta) { return esk(z, beta); }

// Piecewise-polynomial kernel taps (FP32 kernel modes). Tap i of an axis is psi((f + i) 2/W)
// for the record's offset f in (-W/2, -W/2 + 1]. For the interior taps 0 < i < W - 1 that is
// a smooth function of s = 2 f + W - 1 in (-1, 1], replaced by a degree-kTapDeg polynomial
// fitted at geometry build by Chebyshev interpolation (the piecewise Horner evaluation of
// Barnett, Magland & af Klinteberg, SIAM J. Sci. Comput. 41 (2019) C479, for their ES
// kernel): kTapDeg FMAs instead of a square root, a reciprocal and an exponential. The two
// edge taps hold the kernel's square-root end point, where a polynomial converges slowly; they
// keep the direct formula unless `on_all` (their fit error is also within the bound). `on` is 0
// where the fit does not reach the FP32 evaluation level (small supports): direct formula.
constexpr int kTapDeg = 7;
struct TapPoly {
  int on;       // interior taps from the polynomials
  int on_all;   // edge taps too (their fit also meets the bound below)
  float c[kWMax][kTapDeg + 1];  // c[i][d]: coefficient of s^d of tap i
};
// Horner's rule over one coefficient row held as two float4 (c0..c3, c4..c7)
static_assert(kTapDeg == 7, "poly_row holds 8 coefficients");
__device__ __forceinline__ float poly_row(const float4* r, float s) {
  const float4 a = r[0], b = r[1];
  float p = fmaf(b.w, s, b.z);
  p = fmaf(p, s, b.y);
  p = fmaf(p, s, b.x);
  p = fmaf(p, s, a.w);
  p = fmaf(p, s, a.z);
  p = fmaf(p, s, a.y);
  return fmaf(p, s, a.x);
}
__device__ __forceinline__ float tap_poly(const TapPoly& tp, int i, float s) {
  float p = tp.c[i][kTapDeg];
#pragma unroll
  for (int d = kTapDeg - 1; d >= 0; --d) p = fmaf(p, s, tp.c[i][d]);
  return p;
}
// all W taps of one axis (W a compile-time support)
template <typename R, int W>
__device__ __forceinline__ void taps_axis(const TapPoly& tp, R f, R scale, R beta, R (&kt)[W]) {
  if constexpr (std::is_same<R, float>::value) {
    if (tp.on_all) {
      const float s = fmaf(2.0f, f, (float)(W - 1));
#pragma unroll
      for (int i = 0; i < W; ++i) kt[i] = tap_poly(tp, i, s);
      return;
    }
    if (tp.on) {
      const float s = fmaf(2.0f, f, (float)(W - 1));
      kt[0] = esk_fast(f * scale, beta);
#pragma unroll
      for (int i = 1; i < W - 1; ++i) kt[i] = tap_poly(tp, i, s);
      kt[W - 1] = esk_fast((f + (float)(W - 1)) * scale, beta);
      return;
    }
  }
#pragma unroll
  for (int i = 0; i < W; ++i) kt[i] = esk_fast((f + (R)i) * scale, beta);
}
// taps i < 8 of one axis for a run-time support W <= 8 (zero for i >= W or a dead lane)
__device__ __forceinline__ void taps8(const TapPoly& tp, int W, float f, float scale, float beta, bool live,
                                      float (&kt)[8]) {
  if (tp.on_all) {
    const float s = fmaf(2.0f, f, (float)(W - 1));
#pragma unroll
    for (int i = 0; i < 8; ++i) kt[i] = i < W ? tap_poly(tp, i, s) : 0.f;
  } else if (tp.on) {
    const float s = fmaf(2.0f, f, (float)(W - 1));
    const float e0 = esk_fast(f * scale, beta), e1 = esk_fast((f + (float)(W - 1)) * scale, beta);
    kt[0] = e0;
#pragma unroll
    for (int i = 1; i < 8; ++i) kt[i] = i == W - 1 ? e1 : (i < W - 1 ? tap_poly(tp, i, s) : 0.f);
  } else {
#pragma unroll
    for (int i = 0; i < 8; ++i) kt[i] = i < W ? esk_fast((f + (float)i) * scale, beta) : 0.f;
  }
  if (!live) {
#pragma unroll
    for (int i = 0; i < 8; ++i) kt[i] = 0.f;
  }
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

// Staged entry e: taps s_t[e * NT + i] (i < 8: x taps of the sub-tile's 8 columns; 8 + i:
// y taps of its 8 rows; 16 + q: w taps of the pass's planes), values s_v[b * 64 + e].
// Per-wave accumulator of one row group of a strip (sub-tiles h < NH of 8 x 8 cells): PC
// planes x NB polarisations of every cell; accumulate(h, ...) adds a staged batch of
// sub-tile h.
// Matrix-core form (PC <= 2): per sub-tile h and polarisation b the 16 x 16 block
//   G_hb[(a, q)][(c, p)] = sum_e kx_e[a] kw_e[q] * ky_e[c] (p ? Im : Re)(val_e[b])
// for cells (sx0 + a, sy0 + 8 h + c), taken four entries at a time by
// v_mfma_f32_16x16x4_f32 (FP32) or v_mfma_f64_16x16x4_f64 (FP64): lane l supplies row l % 16
// of A and column l % 16 of B for entry l / 16 of the group. Register r of lane l receives
// column l % 16 and row 4 (l / 16) + r (FP32) or row 4 r + l / 16 (FP64; both checked on the
// device against a host product). Entries past the count up to the end of the batch of 64 must
// be staged as zeros.
typedef double f64x4 __attribute__((ext_vector_type(4)));
template <typename R>
struct MfmaT;
template <>
struct MfmaT<float> {
  using V = f32x4;
  static __device__ __forceinline__ V mac(float a, float b, V c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int gk, int r) { return 4 * gk + r; }
};
template <>
struct MfmaT<double> {
  using V = f64x4;
  static __device__ __forceinline__ V mac(double a, double b, V c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int gk, int r) { return 4 * r + gk; }
};
template <typename R, int PC, int NB, int NT, int NH>
struct Grid8Mx {
  static_assert(PC <= 2, "the 16 rows of a block are 8 cells x 2 planes");
  static constexpr bool kZeroPad = true;
  using M = MfmaT<R>;
  typename M::V acc[NH][NB];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int h = 0; h < NH; ++h)
#pragma unroll
      for (int b = 0; b < NB; ++b) acc[h][b] = typename M::V{0, 0, 0, 0};
  }
  // rounds of 16 entries (4 MFMA groups, unrolled); the staging pads the batch of 64 with
  // zeros, so a round may run past the count
  template <int H>
  __device__ __forceinline__ void accumulate(const R* s_t, const cplx<R>* s_v, int cnt, int lane) {
    const int gk = lane >> 4, c8 = lane & 7, h2 = (lane >> 3) & 1;
    const int n16 = __builtin_amdgcn_readfirstlane(cnt);
    for (int e0 = 0; e0 < n16; e0 += 16) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + 4 * u + gk;
        const R* t = s_t + e * NT;
        const R av = t[c8] * (h2 < PC ? t[16 + h2] : (R)0);
        const R ty = t[8 + c8];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          const R vb = reinterpret_cast<const R*>(s_v + b * 64 + e)[h2];
          acc[H][b] = M::mac(av, ty * vb, acc[H][b]);
        }
      }
    }
  }
  __device__ __forceinline__ void store(cplx<R>* plane, const KP& k, int npl, int x0, int y0, bool solo,
                                        int lane) const {
    const int gk = lane >> 4, c8 = lane & 7, h2 = (lane >> 3) & 1;
#pragma unroll
    for (int h = 0; h < NH; ++h)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = M::row(gk, r);
        const int a = row & 7, q = row >> 3;
        if (q < PC && q < npl) {
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const R val = acc[h][b][r];
            if (b < k.nb && (solo || val != (R)0)) {
              R* o = reinterpret_cast<R*>(plane + ((int64_t)(q * k.nb + b) * k.gx + x0 + a) * k.gy + y0 + 8 * h + c8) +
                     h2;
              if (solo) *o = val;
              else atomicAdd(o, val);
            }
          }
        }
      }
  }
};

template <typename R, int PC, int NB, int NT, int NH>
using Grid8Acc = Grid8Mx<R, PC, NB, NT, NH>;

// rows per candidate scan of the sub-tile gatherer (RG of k_grid8)
template <typename R>
__host__ __device__ constexpr int grid8_rows() {
  return std::is_same<R, float>::value ? 16 : 8;
}

// FP32: at most 128 VGPRs, so that 4 waves per SIMD stay resident (as many as the LDS allows);
// FP64: at most 168, 3 waves per SIMD (C1 FP64 gatherer 353 -> 305 ms per 2-polarisation invert)
template <typename R, int T, int PC, int NB, int RG>
__global__ __launch_bounds__(kBlock, (std::is_same<R, float>::value ? 4 : 3)) void k_grid8(KP k, int p0, int npl, const uint32_t* __restrict__ bstart,
                                                  const uint32_t* __restrict__ bend,
                                                  const uint2* __restrict__ srec,
                                                  const uint32_t* __restrict__ rec_a,
                                                  const rec3<R>* __restrict__ rec_f,
                                                  const cplx<R>* __restrict__ rec_v,
                                                  const uint32_t* __restrict__ task_off,
                                                  const uint32_t* __restrict__ task_tab, uint32_t chunk,
                                                  uint32_t* __restrict__ counters,
                                                  cplx<R>* __restrict__ plane,
                                                  const uint32_t* __restrict__ split, Strips sr, TapPoly tp) {
  constexpr int NS = T / 8;               // strips per tile
  static_assert((RG == 8 || RG == 16) && T % RG == 0, "row groups of 8 or 16 tile the strip");
  constexpr int NT = grid8_taps(PC);      // per staged visibility: 8 x taps, 8 y taps, PC w taps
  constexpr int NQ = RG / 8;              // sub-tiles (queues) of a row group
  using Acc = Grid8Acc<R, PC, NB, NT, NQ>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int W = k.W;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  unsigned char* base = smem_raw + wv * grid8_region(sizeof(R), PC, NB, NQ);
  cplx<R>* s_v = reinterpret_cast<cplx<R>*>(base);  // [NB][64]
  R* s_t = reinterpret_cast<R*>(s_v + NB * 64);       // [64][NT]
  uint32_t* q_j = reinterpret_cast<uint32_t*>(s_t + 64 * NT);  // [NQ][128] records
  uint32_t* q_p = q_j + NQ * 128;  // [NQ][128] offsets dx + 128, dy + 128 (8 bits each), start plane - slo
  uint32_t* s_segv = q_p + NQ * 128;        // [kMaxSeg + 1]
  uint32_t* s_segj = s_segv + kMaxSeg + 1;  // [kMaxSeg]
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const R scale = (R)2 / (R)W;
  // the FP32 tap evaluation works in base 2
  const R beta = std::is_same<R, float>::value ? (R)(k.beta * 1.4426950408889634) : (R)k.beta;
  // FP32 with every tap fitted (on_all): x / y taps from the polynomials, whose coefficient
  // rows -8 .. W + 7 sit in LDS as two float4 each, zero outside the support, so a staging lane
  // reads the row of its own tap index (always in -7 .. W + 6 for an accepted candidate)
  constexpr bool kF32 = std::is_same<R, float>::value;
  constexpr int kTapRows = kWMax + 16;
  __shared__ float4 s_tap[kF32 ? 2 * kTapRows : 1];
  const bool polytaps = kF32 && tp.on_all;
  if (polytaps) {
    for (int r = threadIdx.x; r < 2 * kTapRows; r += kBlock) {
      const int row = r / 2 - 8, c0 = 4 * (r % 2);
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (row >= 0 && row < W) v = make_float4(tp.c[row][c0], tp.c[row][c0 + 1], tp.c[row][c0 + 2], tp.c[row][c0 + 3]);
      s_tap[kF32 ? r : 0] = v;
    }
    __syncthreads();
  }
  const uint32_t total = counters[1];
  int slo, shi;
  chunk_range(k.Ww, k.nplanes, p0, npl, &slo, &shi);
  const int nseg = (shi - slo + 1) * 4;
  // XCD-local work queues: the NS strip tasks of a (tile, chunk) read the same records, and
  // so do the tasks of the next tile in y (its neighbour buckets), so queue x holds the
  // contiguous x-th eighth of the (tile, chunk) groups and a wave pops from the queue of its
  // own XCD (workgroups are placed on XCDs round-robin by index) before helping the others
  const uint32_t ngroups = total / NS;
  const uint32_t gper = (ngroups + kXcd - 1) / kXcd;
  const int xcd0 = (int)(blockIdx.x % kXcd);
  int qi = 0;
  for (;;) {
    uint32_t task = 0xffffffffu;
    while (qi < kXcd) {
      const int x = (xcd0 + qi) % kXcd;
      const uint32_t g0 = min(ngroups, (uint32_t)x * gper), qsize = min(ngroups - g0, gper);
      uint32_t l = 0;
      if (lane == 0) l = queue_pop(&counters[kQueueBase + x * kQueueStride], qsize * NS);
      l = __shfl(l, 0, 64);
      if (l < qsize * NS) {
        task = (g0 + l / NS) * NS + l % NS;
        break;
      }
      ++qi;
    }
    if (task == 0xffffffffu) break;
    const int tile = (int)task_tab[task];
    const uint32_t local = task - task_off[tile];
    const uint32_t ci = local / NS;
    const int sx0 = (int)(local - ci * NS) * 8;
    const uint32_t tnext = tile + 1 < (int)k.ntiles ? task_off[tile + 1] : total;
    const bool solo = tnext - task_off[tile] == (uint32_t)NS;
    const int tx = tile / k.nty, ty = tile - tx * k.nty;
    // segment table: one bucket per lane (the strip's range of it), inclusive scan by shuffles
    const uint32_t nkeys = (uint32_t)k.nplanes * k.ntiles;
    for (int sg0 = 0; sg0 < nseg; sg0 += 64) {
      const int sg = sg0 + lane;
      uint32_t len = 0;
      if (sg < nseg) {
        uint32_t a, b;
        strip_range(sr, sx0 / 8, (sg & 3) >> 1, seg_key(k, slo + sg / 4, tx, ty, sg & 3), bstart, bend, split, nkeys,
                    &a, &b);
        len = b - a;
        s_segj[sg] = a;
      }
      uint32_t inc = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t tt = __shfl_up(inc, o, 64);
        if (lane >= o) inc += tt;
      }
      const uint32_t carry = sg0 ? s_segv[sg0] : 0u;
      if (sg < nseg) s_segv[sg + 1] = carry + inc;
      if (sg0 == 0 && lane == 0) s_segv[0] = 0;
      wave_sync();
    }
    const uint32_t vlo = ci * chunk;
    const uint32_t vhi = min(vlo + chunk, s_segv[nseg]);
    for (int sy0 = 0; sy0 < T; sy0 += RG) {
      Acc acc;
      acc.zero();
      // stage entries [0, cnt) of queue H (sub-tile rows sy0 + 8 H + [0, 8)) and accumulate them
      auto stage_gather = [&](auto hq, int cnt) {
        constexpr int H = decltype(hq)::value;
        const int ry0 = sy0 + 8 * H;
        // staging: one queued visibility per lane; its values, the taps of the sub-tile's 8
        // columns and 8 rows (zero outside the support) and its w taps for the pass's planes
        if (Acc::kZeroPad && lane >= cnt) {
#pragma unroll
          for (int b = 0; b < NB; ++b) s_v[b * 64 + lane] = cplx<R>{0, 0};
          R* t = s_t + lane * NT;
#pragma unroll
          for (int i = 0; i < NT; ++i) t[i] = (R)0;
        }
        if (lane < cnt) {
          const uint32_t j = q_j[H * 128 + lane], pk = q_p[H * 128 + lane];
          const int dx = (int)(pk & 255u) - 128, dy = (int)((pk >> 8) & 255u) - 128;
          const int sp = slo + (int)(pk >> 16);
          const rec3<R> f = rec_f[j];
#pragma unroll
          for (int b = 0; b < NB; ++b) s_v[b * 64 + lane] = rec_v[(int64_t)j * NB + b];
          R* t = s_t + lane * NT;
#pragma unroll
          for (int q = 0; q < PC; ++q) {
            const int iw = p0 + q - sp;
            t[16 + q] = (q < npl && iw >= 0 && iw < k.Ww)
                            ? (k.wstack ? esk_fast((f.w + (R)iw) * scale, beta) : (R)1)
                            : (R)0;
          }
          if (polytaps) {
            const float sxv = fmaf(2.0f, (float)f.x, (float)(W - 1)), syv = fmaf(2.0f, (float)f.y, (float)(W - 1));
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              const int rx = sx0 + i - dx, ry = ry0 + i - dy;
              t[i] = (R)poly_row(s_tap + 2 * (rx + 8), sxv);
              t[8 + i] = (R)poly_row(s_tap + 2 * (ry + 8), syv);
            }
          } else {
#pragma unroll
            for (int i = 0; i < 8; ++i) {
              const int rx = sx0 + i - dx, ry = ry0 + i - dy;
              t[i] = (unsigned)rx < (unsigned)W ? esk_fast((f.x + (R)rx) * scale, beta) : (R)0;
              t[8 + i] = (unsigned)ry < (unsigned)W ? esk_fast((f.y + (R)ry) * scale, beta) : (R)0;
            }
          }
        }
        wave_sync();
        acc.template accumulate<H>(s_t, s_v, cnt, lane);
        wave_sync();
      };
      int qn[NQ];
#pragma unroll
      for (int h = 0; h < NQ; ++h) qn[h] = 0;
      // queue H with 64 or more entries: stage and accumulate 64, move the rest to the front
      auto drain = [&](auto hq) {
        constexpr int H = decltype(hq)::value;
        if (qn[H] >= 64) {
          stage_gather(hq, 64);
          qn[H] -= 64;
          uint32_t a0 = 0, a1 = 0;
          if (lane < qn[H]) {
            a0 = q_j[H * 128 + 64 + lane];
            a1 = q_p[H * 128 + 64 + lane];
          }
          wave_sync();
          if (lane < qn[H]) {
            q_j[H * 128 + lane] = a0;
            q_p[H * 128 + lane] = a1;
          }
          wave_sync();
        }
      };
      // segment of the chunk's first candidate by bisection, then walked forward block by block
      // (a block of 64 consecutive candidates rarely crosses more than one segment boundary)
      int cur = 0;
      {
        int hi = nseg;
        while (hi - cur > 1) {
          const int mid = (cur + hi) >> 1;
          if (s_segv[mid] <= vlo) cur = mid;
          else hi = mid;
        }
      }
      // candidate v: its segment (walked forward from `from`), record index and origin word;
      // the next block's are loaded while the current block is filtered
      auto fetch = [&](uint32_t v, int from, uint32_t& j, uint32_t& Ar, int& lo) {
        lo = from;
        j = 0;
        Ar = 0;
        if (v < vhi) {
          while (lo + 1 < nseg && s_segv[lo + 1] <= v) ++lo;
          const uint32_t sj = s_segj[lo] + (v - s_segv[lo]);
          if (srec) {
            const uint2 e = srec[sj];
            j = e.x;
            Ar = e.y;
          } else {
            j = sj;
            Ar = rec_a[sj];
          }
        }
      };
      uint32_t j_n, A_n;
      int lo_n;
      fetch(vlo + (uint32_t)lane, cur, j_n, A_n, lo_n);
      for (uint32_t vb = vlo; vb < vhi; vb += 64) {
        const uint32_t v = vb + (uint32_t)lane;
        const uint32_t j = j_n, Ar = A_n;
        const int lo = lo_n;  // last segment starting at or before v
        fetch(v + 64u, __shfl(lo, 63, 64), j_n, A_n, lo_n);
        bool okq[NQ];
#pragma unroll
        for (int h = 0; h < NQ; ++h) okq[h] = false;
        int dx = 0, dy = 0, sp = 0;
        if (v < vhi) {
          sp = slo + lo / 4;
          const uint32_t A = Ar & k.amask;
          dx = (int)(A & 0xffffu) - tx * T;
          dy = (int)(A >> 16) - ty * T;
          if (dx >= k.hgx) dx -= k.gx;
          if (dx < -k.hgx) dx += k.gx;
          if (dy >= k.hgy) dy -= k.gy;
          if (dy < -k.hgy) dy += k.gy;
          const bool okx = (Ar & k.dflag) == 0u && (dx + W > sx0) && (dx < sx0 + 8);
#pragma unroll
          for (int h = 0; h < NQ; ++h) okq[h] = okx && (dy + W > sy0 + 8 * h) && (dy < sy0 + 8 * h + 8);
        }
        // accepted offsets lie in (-16, 32): 8 bits each with a bias of 128; the start plane
        // relative to the pass's first candidate plane is at most W + PC - 2
        const uint32_t pk = (uint32_t)(dx + 128) | ((uint32_t)(dy + 128) << 8) | ((uint32_t)(sp - slo) << 16);
#pragma unroll
        for (int h = 0; h < NQ; ++h) {
          const unsigned long long m = __ballot(okq[h]);
          if (okq[h]) {
            const int pos = qn[h] + (int)__popcll(m & lt);
            q_j[h * 128 + pos] = j;
            q_p[h * 128 + pos] = pk;
          }
          qn[h] += (int)__popcll(m);
        }
        wave_sync();
        drain(std::integral_constant<int, 0>{});
        if (NQ > 1) drain(std::integral_constant<int, NQ - 1>{});
      }
      if (qn[0] > 0) stage_gather(std::integral_constant<int, 0>{}, qn[0]);
      if (NQ > 1 && qn[NQ - 1] > 0) stage_gather(std::integral_constant<int, NQ - 1>{}, qn[NQ - 1]);
      // a tile with a single task is owned by it: every cell is stored (zeros included; the
      // plane is not cleared there beforehand); shared tiles add their non-zero cells
      acc.store(plane, k, npl, tx * T + sx0, ty * T + sy0, solo, lane);
    }
  }
}

// ---------------------------------------------------------------------------------
// Dense runs. The sort orders a bucket by in-tile origin, so visibilities with the same
// start plane and the same first cell are consecutive: a run. Runs of at least kDenseMin
// visibilities (the dense uv core: short baselines at many channels and times) are
// gridded and degridded on the matrix cores with the run's footprint as the block
// (k_grid_dense_mx, k_degrid_dense_mx); the sub-tile gatherer and k_degrid skip their
// records (flag bit in rec_a).
// ---------------------------------------------------------------------------------
// head[j] = 1 where a run starts (bucket key or first cell differ from the previous record)
// Run heads of the kRsIpt consecutive records from `base`: a run starts where the bucket key
// or the first cell changes (h = 0 past the end). Vector loads when all are in range.
__device__ __forceinline__ void run_heads(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rec_a,
                                          int64_t n, int64_t base, uint32_t (&h)[kRsIpt]) {
  uint32_t kk[kRsIpt], aa[kRsIpt];
  if (base + kRsIpt <= n) {
    const uint4* pk = reinterpret_cast<const uint4*>(keys + base);
    const uint4* pa = reinterpret_cast<const uint4*>(rec_a + base);
#pragma unroll
    for (int q = 0; q < kRsIpt / 4; ++q) {
      const uint4 a = pk[q], b = pa[q];
      kk[4 * q] = a.x, kk[4 * q + 1] = a.y, kk[4 * q + 2] = a.z, kk[4 * q + 3] = a.w;
      aa[4 * q] = b.x, aa[4 * q + 1] = b.y, aa[4 * q + 2] = b.z, aa[4 * q + 3] = b.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kRsIpt; ++i) {
      kk[i] = base + i < n ? keys[base + i] : 0u;
      aa[i] = base + i < n ? rec_a[base + i] : 0u;
    }
  }
  uint32_t kp = 0, ap = 0;
  if (base > 0 && base <= n) {
    kp = keys[base - 1];
    ap = rec_a[base - 1];
  }
#pragma unroll
  for (int i = 0; i < kRsIpt; ++i) {
    const int64_t j = base + i;
    h[i] = (j < n && (j == 0 || kk[i] != kp || aa[i] != ap)) ? 1u : 0u;
    kp = kk[i];
    ap = aa[i];
  }
}

// Exclusive scan of the run heads, computed on the fly (no heads array), in kRsItems blocks:
// partial sums per block, then A[j] = number of run heads before record j, with each head's
// position written to rstart[A[j]] and, at the last record, rstart[nruns] = n and
// counters[kRuns] = nruns.
__global__ void k_runscan_partial(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rec_a, int64_t n,
                                  uint32_t* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kRsItems + (int64_t)threadIdx.x * kRsIpt;
  uint32_t h[kRsIpt];
  run_heads(keys, rec_a, n, base, h);
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kRsIpt; ++i) s += h[i];
  uint32_t tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void k_runscan_apply(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rec_a, int64_t n,
                                const uint32_t* __restrict__ part, uint32_t* __restrict__ A,
                                uint32_t* __restrict__ rstart, uint32_t* __restrict__ counters) {
  const int64_t base = (int64_t)blockIdx.x * kRsItems + (int64_t)threadIdx.x * kRsIpt;
  uint32_t h[kRsIpt];
  run_heads(keys, rec_a, n, base, h);
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < kRsIpt; ++i) s += h[i];
  uint32_t tot;
  uint32_t run = block_excl_scan(s, &tot) + part[blockIdx.x];
  uint32_t loc[kRsIpt];
#pragma unroll
  for (int i = 0; i < kRsIpt; ++i) {
    loc[i] = run;
    if (h[i]) rstart[run] = (uint32_t)(base + i);
    run += h[i];
  }
  if (base + kRsIpt <= n) {
    uint4* p = reinterpret_cast<uint4*>(A + base);
#pragma unroll
    for (int q = 0; q < kRsIpt / 4; ++q) p[q] = make_uint4(loc[4 * q], loc[4 * q + 1], loc[4 * q + 2], loc[4 * q + 3]);
  } else {
    for (int i = 0; i < kRsIpt; ++i)
      if (base + i < n) A[base + i] = loc[i];
  }
  if (base < n && base + kRsIpt >= n) {  // the thread holding the last record
    rstart[run] = (uint32_t)n;
    counters[kRuns] = run;
  }
}

// per run r < nruns: tcnt[r] = tasks of a dense run (len >= minlen), dlen[r] = its length
// (0 for short runs); entry nruns is 0 so that the exclusive scans end with the totals
__global__ void k_run_info(const uint32_t* __restrict__ rstart, int64_t nruns, uint32_t minlen, uint32_t chunk,
                           uint32_t* __restrict__ tcnt, uint32_t* __restrict__ dlen) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r > nruns) return;
  uint32_t t = 0, d = 0;
  if (r < nruns) {
    const uint32_t len = rstart[r + 1] - rstart[r];
    if (len >= minlen) {
      t = (len + chunk - 1) / chunk;
      d = len;
    }
  }
  tcnt[r] = t;
  dlen[r] = d;
}

// with toff the exclusive scan of the task counts: the task list (first record, count,
// packed first cell, start plane) and each task's start plane alone; tasks are in record
// order, so their start planes ascend
__global__ void k_run_tasks(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ rec_a,
                            const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ toff,
                            int64_t nruns, uint32_t minlen, uint32_t chunk, uint32_t ntiles, uint32_t amask,
                            uint4* __restrict__ dtask, uint32_t* __restrict__ dplane,
                            uint32_t* __restrict__ counters) {
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (r == 0) counters[kDenseTotal] = toff[nruns];
  if (r >= nruns) return;
  const uint32_t st = rstart[r], len = rstart[r + 1] - st;
  if (len < minlen) return;
  const uint32_t nt = (len + chunk - 1) / chunk;
  const uint32_t sp = keys[st] / ntiles, A = rec_a[st] & amask;
  for (uint32_t c = 0; c < nt; ++c) {
    dtask[toff[r] + c] = make_uint4(st + c * chunk, min(chunk, len - c * chunk), A, sp);
    dplane[toff[r] + c] = sp;
  }
}

// per record: its run (A = exclusive scan of the run heads), dense runs flagged in rec_a,
// the other records listed in srec (record index, origin word) at their rank among non-dense
// records (dpre = exclusive
// scan of the dense run lengths)
__global__ void k_run_mark(uint32_t* __restrict__ rec_a, int64_t n, const uint32_t* __restrict__ A,
                           const uint32_t* __restrict__ rstart, const uint32_t* __restrict__ dpre,
                           const uint32_t* __restrict__ counters, uint32_t minlen, uint32_t flag,
                           uint2* __restrict__ srec) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t nruns = counters[kRuns];
  const uint32_t a = A[j];
  const uint32_t r = (a < nruns && rstart[a] == (uint32_t)j) ? a : a - 1u;
  if (rstart[r + 1] - rstart[r] >= minlen) rec_a[j] |= flag;
  else srec[(uint32_t)j - dpre[r]] = make_uint2((uint32_t)j, rec_a[j]);
}

// bucket table of the non-dense records: a bucket's first record and end are run starts
// (or n), so the non-dense records before them are the position minus the dense ones
__global__ void k_sparse_bounds(const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
                                int64_t nkeys, int64_t n, const uint32_t* __restrict__ A,
                                const uint32_t* __restrict__ dpre, const uint32_t* __restrict__ counters,
                                uint32_t* __restrict__ sbstart, uint32_t* __restrict__ sbend) {
  const int64_t key = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (key >= nkeys) return;
  const uint32_t nruns = counters[kRuns];
  const uint32_t x0 = bstart[key], x1 = bend[key];
  const uint32_t d0 = (int64_t)x0 < n ? dpre[A[x0]] : dpre[nruns];
  const uint32_t d1 = (int64_t)x1 < n ? dpre[A[x1]] : dpre[nruns];
  sbstart[key] = x0 - d0;
  sbend[key] = x1 - d1;
}

// diagnostics: out[0] += dense-run tasks, out[1] += visibilities they hold
__global__ void k_dense_stats(const uint4* __restrict__ dtask, const uint32_t* __restrict__ counters,
                              unsigned long long* __restrict__ out) {
  const uint32_t total = counters[kDenseTotal];
  unsigned long long nv = 0, nt = 0;
  for (uint32_t t = blockIdx.x * kBlock + threadIdx.x; t < total; t += gridDim.x * kBlock) {
    nv += dtask[t].y;
    ++nt;
  }
  if (nt) {
    atomicAdd(&out[0], nt);
    atomicAdd(&out[1], nv);
  }
}

// (plane, tile) pairs that some footprint touches: plane p is met by the records starting at
// planes p - Ww + 1 .. p, and tile (tx, ty) by those whose first cell lies in tiles
// (tx or tx - 1, ty or ty - 1) (periodic), since W <= T. One thread per (plane, tile).
__global__ void k_occupied(KP k, const uint32_t* __restrict__ bstart, const uint32_t* __restrict__ bend,
                           unsigned long long* __restrict__ out) {
  const int64_t total = (int64_t)k.nplanes * k.ntiles;
  unsigned long long cnt = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total; i += (int64_t)gridDim.x * kBlock) {
    const int p = (int)(i / k.ntiles);
    const uint32_t t = (uint32_t)(i - (int64_t)p * k.ntiles);
    const int tx = (int)(t / (uint32_t)k.nty), ty = (int)(t % (uint32_t)k.nty);
    const int txm = tx == 0 ? k.ntx - 1 : tx - 1, tym = ty == 0 ? k.nty - 1 : ty - 1;
    const uint32_t nb4[4] = {(uint32_t)tx * k.nty + ty, (uint32_t)txm * k.nty + ty, (uint32_t)tx * k.nty + tym,
                             (uint32_t)txm * k.nty + tym};
    bool occ = false;
    const int s0 = max(0, p - k.Ww + 1), s1 = min(p, k.nplanes - k.Ww);
    for (int sp = s0; sp <= s1 && !occ; ++sp)
      for (int q = 0; q < 4 && !occ; ++q) {
        const int64_t key = (int64_t)sp * k.ntiles + nb4[q];
        occ = bend[key] > bstart[key];
      }
    cnt += occ ? 1ull : 0ull;
  }
  if (cnt) atomicAdd(out, cnt);
}

// first index in [0, n) with a[i] >= v (a ascending)
__device__ __forceinline__ uint32_t lower_bound_u32(const uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// lastw[tile] = pass for the tiles the footprints of this pass's dense-run tasks touch (a W x W
// footprint from the packed first cell ax | ay << 16 meets at most 2 x 2 tiles, W <= tile)
__global__ void k_mark_dense(KP k, int p0, int npl, const uint4* __restrict__ dtask,
                             const uint32_t* __restrict__ dplane, const uint32_t* __restrict__ counters,
                             uint32_t* __restrict__ lastw, uint32_t pass) {
  int slo, shi;
  chunk_range(k.Ww, k.nplanes, p0, npl, &slo, &shi);
  const uint32_t total = counters[kDenseTotal];
  const uint32_t lo = lower_bound_u32(dplane, total, (uint32_t)slo);
  const uint32_t hi = lower_bound_u32(dplane, total, (uint32_t)shi + 1u);
  for (uint32_t i = lo + blockIdx.x * kBlock + threadIdx.x; i < hi; i += gridDim.x * kBlock) {
    const uint32_t A = dtask[i].z;
    const int ax = (int)(A & 0xffffu), ay = (int)(A >> 16);
#pragma unroll
    for (int cx = 0; cx < 2; ++cx)
#pragma unroll
      for (int cy = 0; cy < 2; ++cy) {
        const int x = (ax + cx * (k.W - 1)) % k.gx, y = (ay + cy * (k.W - 1)) % k.gy;
        lastw[(uint32_t)(x / k.T) * (uint32_t)k.nty + (uint32_t)(y / k.T)] = pass;
      }
  }
}

// Dense-run gridder on the matrix cores (W <= 8, 2 planes per pass; FP32 or FP64).
// For polarisation b the wave accumulates the 16 x 16 block
//   G_b[(a, q)][(c, h)] = sum_v kx_v[a] kw_v[q] * ky_v[c] (h ? Im : Re)(val_v[b])
// over the task's visibilities v, for footprint cells (ax + a, ay + c), a, c < 8, and the
// pass's planes q < 2: a sum of rank-1 terms that v_mfma_f32_16x16x4_f32 (FP64:
// v_mfma_f64_16x16x4_f64) takes four visibilities at a time, with the lane layouts of
// Grid8Mx. Tasks (up to kDenseChunk visibilities of one run: same start plane, same first
// cell) are popped one ahead and staged 64 at a time, the next records loaded while the
// current ones are accumulated; each task ends with one global atomic add per cell, plane
// and polarisation. Staging lanes past the end of a batch write zeros, so a partial group of
// four adds nothing. Runs after the sub-tile gatherer of the same pass (whose solo tiles are
// written with plain stores).

// taps i < 8 of one axis for a run-time support W <= 8, FP64 (direct formula)
__device__ __forceinline__ void taps8(const TapPoly&, int W, double f, double scale, double beta, bool live,
                                      double (&kt)[8]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) kt[i] = (live && i < W) ? esk((f + (double)i) * scale, beta) : 0.0;
}

template <typename R, int NB>
__global__ __launch_bounds__(kBlock) void k_grid_dense_mx(KP k, int p0, int npl, const uint4* __restrict__ dtask,
                                                          const uint32_t* __restrict__ dplane,
                                                          const rec3<R>* __restrict__ rec_f,
                                                          const cplx<R>* __restrict__ rec_v,
                                                          uint32_t* __restrict__ counters,
                                                          cplx<R>* __restrict__ plane, TapPoly tp) {
  static_assert(kMaxPC == 2, "the 16 rows of a block are 8 cells x 2 planes");
  constexpr int PC = kMaxPC;
  constexpr int NW = kBlock / 64;
  using M = MfmaT<R>;
  __shared__ R s_tx[NW][64][9];
  __shared__ R s_ty[NW][64][9];
  __shared__ R s_tw[NW][64][PC];
  __shared__ R s_v[NW][NB][64][2];
  const int W = k.W;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gk = lane >> 4, gr = lane & 15;  // visibility of the group, row / column of the block
  const int c8 = gr & 7, h2 = gr >> 3;       // cell, and plane (A side) or Re / Im (B side)
  const R scale = (R)2 / (R)W;
  // the FP32 tap evaluation works in base 2
  const R beta = std::is_same<R, float>::value ? (R)(k.beta * 1.4426950408889634) : (R)k.beta;
  int slo, shi;
  chunk_range(k.Ww, k.nplanes, p0, npl, &slo, &shi);
  const uint32_t total = counters[kDenseTotal];
  const uint32_t lo = lower_bound_u32(dplane, total, (uint32_t)slo);
  const uint32_t hi = lower_bound_u32(dplane, total, (uint32_t)shi + 1u);
  auto pop = [&]() {
    uint32_t t = 0;
    if (lane == 0) t = queue_pop(&counters[kDenseCnt], hi - lo);
    return lo + __shfl(t, 0, 64);
  };
  auto task_at = [&](uint32_t i) {  // wave-uniform task fields
    const uint4 d = dtask[i];
    return make_uint4(__builtin_amdgcn_readfirstlane(d.x), __builtin_amdgcn_readfirstlane(d.y),
                      __builtin_amdgcn_readfirstlane(d.z), __builtin_amdgcn_readfirstlane(d.w));
  };
  uint32_t t = pop();
  if (t >= hi) return;
  uint4 cur = task_at(t);  // first record, count, packed first cell, start plane
  uint32_t tn = pop();
  rec3<R> f{};
  cplx<R> v[NB];
  auto load = [&](const uint4& tk, uint32_t b0) {
    if (b0 + (uint32_t)lane < tk.y) {
      const uint32_t j = tk.x + b0 + (uint32_t)lane;
      f = rec_f[j];
#pragma unroll
      for (int b = 0; b < NB; ++b) v[b] = rec_v[(int64_t)j * NB + b];
    }
  };
  load(cur, 0);
  typename M::V acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = typename M::V{0, 0, 0, 0};
  uint32_t b0 = 0;
  for (;;) {
    const int cnt = (int)min(64u, cur.y - b0);
    const int sp = (int)cur.w;
    {
      const bool live = lane < cnt;
      R tx8[8], ty8[8];
      taps8(tp, W, f.x, scale, beta, live, tx8);
      taps8(tp, W, f.y, scale, beta, live, ty8);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        s_tx[wv][lane][i] = tx8[i];
        s_ty[wv][lane][i] = ty8[i];
      }
#pragma unroll
      for (int q = 0; q < PC; ++q) {
        const int iw = p0 + q - sp;
        s_tw[wv][lane][q] = (live && q < npl && iw >= 0 && iw < k.Ww)
                                ? (k.wstack ? esk_fast((f.w + (R)iw) * scale, beta) : (R)1)
                                : (R)0;
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        s_v[wv][b][lane][0] = live ? v[b].re : (R)0;
        s_v[wv][b][lane][1] = live ? v[b].im : (R)0;
      }
    }
    wave_sync();
    const bool last = b0 + 64u >= cur.y;
    uint4 nxt = cur;
    uint32_t nb0 = b0 + 64u;
    if (last) {
      nb0 = 0;
      nxt = tn < hi ? task_at(tn) : make_uint4(0u, 0u, 0u, 0u);
    }
    load(nxt, nb0);
    // rounds of 16 visibilities (4 MFMA groups, unrolled); the batch is zero-padded to 64
    for (int e0 = 0; e0 < cnt; e0 += 16) {
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + 4 * u + gk;
        const R av = s_tx[wv][e][c8] * s_tw[wv][e][h2];
        const R ty = s_ty[wv][e][c8];
#pragma unroll
        for (int b = 0; b < NB; ++b) acc[b] = M::mac(av, ty * s_v[wv][b][e][h2], acc[b]);
      }
    }
    wave_sync();
    if (!last) {
      b0 = nb0;
      continue;
    }
    {
      const uint32_t A = cur.z;
      int gyi = (int)(A >> 16) + c8;
      if (gyi >= k.gy) gyi -= k.gy;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = M::row(gk, r);
        const int ca = row & 7, q = row >> 3;
        if (q < npl && ca < W && c8 < W) {
          int gxi = (int)(A & 0xffffu) + ca;
          if (gxi >= k.gx) gxi -= k.gx;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const R val = acc[b][r];
            if (val != (R)0)
              atomicAdd(reinterpret_cast<R*>(plane + ((int64_t)(q * NB + b) * k.gx + gxi) * k.gy + gyi) + h2, val);
          }
        }
      }
    }
    if (tn >= hi) break;
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[b] = typename M::V{0, 0, 0, 0};
    cur = nxt;
    b0 = 0;
    tn = pop();
  }
}

// ---------------------------------------------------------------------------------
// Image-domain kernels
// ---------------------------------------------------------------------------------
// (n-1) - nshift at pixel (tx, ty) relative to the image centre, FP64 or FP32
template <bool F32C>
__device__ __forceinline__ double nm1s_at(int tx, int ty, const KP& k, double* n_out) {
  if (!F32C) {
    const double x = tx * k.px, y = ty * k.py;
    const double r2 = x * x + y * y;
    const double rt = sqrt(1.0 - r2);
    const double nm1 = -r2 / (rt + 1.0);
    if (n_out) *n_out = rt;
    return nm1 - k.nshift;
  } else {
    const float x = tx * (float)k.px, y = ty * (float)k.py;
    const float r2 = x * x + y * y;
    const float rt = sqrtf(1.0f - r2);
    const float nm1 = -r2 / (rt + 1.0f);
    if (n_out) *n_out = rt;
    return (double)(nm1 - (float)k.nshift);
  }
}

// Fourier transform of the W-point kernel at xi (cycles per W cells): sum_q c_q cos(pi W z_q xi)
// over the quadrature nodes z_q in [0, 1) with weights c_q that fold in the kernel's values
__device__ __forceinline__ double kernel_ft(double xi, int W, const double* __restrict__ qz,
                                            const double* __restrict__ qc, int nq) {
  double s = 0.0;
  for (int q = 0; q < nq; ++q) s += qc[q] * cospi((double)W * qz[q] * xi);
  return s;
}

// sum_m c[m] T_m(t) for t in [-1, 1] (Clenshaw)
__host__ __device__ inline double cheb_eval(const double* c, int n, double t) {
  double b1 = 0.0, b2 = 0.0;
  for (int m = n - 1; m >= 1; --m) {
    const double b0 = 2.0 * t * b1 - b2 + c[m];
    b2 = b1;
    b1 = b0;
  }
  return t * b1 - b2 + c[0];
}

// invert: acc[b][ix][iy] += Re(exp(-2 pi i w_p nm1s) (-1)^(tx+ty) plane[b][tx mod G][ty mod G])
// 1 / (Phi_x Phi_y [Phi_w(dw nm1s) n]) at pixel (ix, iy). Phi_w comes from the quadrature sum
// (nq > 0: nodes qz, weights qc) or from a Chebyshev series of log Phi_w over the image's
// range of dw nm1s (nq < 0: -nq coefficients in qc, qz = {centre, 1 / half-width})
template <typename A>
__device__ __forceinline__ A corr_at(int ix, int iy, const KP& k, const double* __restrict__ phix,
                                     const double* __restrict__ phiy, const double* __restrict__ qz,
                                     const double* __restrict__ qc, int nq) {
  double den = phix[ix] * phiy[iy];
  if (k.wstack) {
    double n;
    const double xi = k.dw * nm1s_at<false>(ix - k.nx / 2, iy - k.ny / 2, k, &n);
    den *= (nq > 0 ? kernel_ft(xi, k.W, qz, qc, nq) : exp(cheb_eval(qc, -nq, (xi - qz[0]) * qz[1]))) * n;
  }
  return (A)(1.0 / den);
}

// phix[i] = Phi((i - nx/2) / gx) for i < nx, then phiy likewise: one thread per pixel column
// or row (evaluating the nx + ny quadrature sums on the host cost ~8 ms per call at 8192^2)
__global__ void k_phi_tables(double* __restrict__ phix, int nx, int gx, double* __restrict__ phiy, int ny, int gy,
                             int W, const double* __restrict__ qz, const double* __restrict__ qc, int nq) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < nx)
    phix[i] = kernel_ft((double)(i - nx / 2) / gx, W, qz, qc, nq);
  else if (i < nx + ny)
    phiy[i - nx] = kernel_ft((double)(i - nx - ny / 2) / gy, W, qz, qc, nq);
}

// ---------------------------------------------------------------------------------
// Pruned plane transforms. A plane's 2-D FFT runs as two batched 1-D passes with a
// transpose between them, skipping what is known to be zero or unused: grid rows outside
// the band [xlo, xhi) that visibility footprints can reach are zero on the gridding side and
// unused on the degridding side, and only the image's ny columns of the other axis are kept.
// The image-side buffer T holds, per plane, ny rows of gx complex values: T[i][iy][gxi].
// ---------------------------------------------------------------------------------
__device__ __forceinline__ int img_col_to_grid(int iy, int ny, int gy) {  // image column -> grid column
  const int h = ny / 2;
  return iy < h ? gy - h + iy : iy - h;
}
__device__ __forceinline__ int grid_col_to_img(int c, int ny, int gy) {  // -1 outside the image
  const int h = ny / 2;
  if (c < h) return c + h;
  if (c >= gy - h) return c - gy + h;
  return -1;
}

// T[i][iy][x] = plane[i][x][col(iy)] for x inside the band [xlo, xhi) (T stays zero outside
// it); 32 x 32 tiles, the blocks covering the band from column xbase
template <typename R>
__global__ __launch_bounds__(kBlock) void k_tr_in(KP k, int xlo, int xhi, int xbase,
                                                  const cplx<R>* __restrict__ plane, cplx<R>* __restrict__ tb) {
  __shared__ cplx<R> t[32][33];
  const int i = blockIdx.z;
  const int x0 = xbase + blockIdx.x * 32, y0 = blockIdx.y * 32;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  for (int r = ly; r < 32; r += 8) {
    const int x = x0 + r, iy = y0 + lx;
    cplx<R> v{0, 0};
    if (x >= xlo && x < xhi && iy < k.ny) v = plane[((int64_t)i * k.gx + x) * k.gy + img_col_to_grid(iy, k.ny, k.gy)];
    t[r][lx] = v;
  }
  __syncthreads();
  for (int r = ly; r < 32; r += 8) {
    const int iy = y0 + r, x = x0 + lx;
    if (iy < k.ny && x >= xlo && x < xhi) tb[((int64_t)i * k.ny + iy) * k.gx + x] = t[lx][r];
  }
}

// band rows: plane[i][x][c] = T[i][img(c)][x] for image columns c, and 0 for the others when
// ZERO (without it those columns are left as they are)
template <typename R, bool ZERO>
__global__ __launch_bounds__(kBlock) void k_tr_out(KP k, int xlo, int xhi, const cplx<R>* __restrict__ tb,
                                                   cplx<R>* __restrict__ plane) {
  __shared__ cplx<R> t[32][33];
  const int i = blockIdx.z;
  const int x0 = xlo + blockIdx.x * 32, c0 = blockIdx.y * 32;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  for (int r = ly; r < 32; r += 8) {
    const int c = c0 + r, x = x0 + lx;
    cplx<R> v{0, 0};
    if (c < k.gy && x < xhi) {
      const int iy = grid_col_to_img(c, k.ny, k.gy);
      if (iy >= 0) v = tb[((int64_t)i * k.ny + iy) * k.gx + x];
    }
    t[r][lx] = v;
  }
  __syncthreads();
  const bool img = ZERO || grid_col_to_img(c0 + lx, k.ny, k.gy) >= 0;
  for (int r = ly; r < 32; r += 8) {
    const int x = x0 + r, c = c0 + lx;
    if (x < xhi && c < k.gy && img) plane[((int64_t)i * k.gx + x) * k.gy + c] = t[lx][r];
  }
}

// invert: accT[b][iy][ix] += (-1)^(tx+ty) sum_q Re(T_q[b][iy][gxi] exp(-2 pi i w_q nm1s))
template <typename R, typename A, bool F32C>
__global__ void k_screen_acc_t(KP k, int p0, int npl, const cplx<R>* __restrict__ tb, A* __restrict__ acc,
                               int first) {
  const int ix = blockIdx.x * kBlock + threadIdx.x;
  const int iy = blockIdx.y;
  if (ix >= k.nx) return;
  const int tx = ix - k.nx / 2, ty = iy - k.ny / 2;
  const int gxi = tx < 0 ? tx + k.gx : tx;
  const double nm1s = k.wstack ? nm1s_at<F32C>(tx, ty, k, nullptr) : 0.0;
  const R sign = ((tx + ty) & 1) ? (R)-1 : (R)1;
  A sum[kMaxBatch] = {0, 0, 0, 0};
  for (int q = 0; q < npl; ++q) {
    R c = 1, s = 0;
    if (k.wstack) phase_cs<R, F32C>(-(k.w0 + (p0 + q) * k.dw) * nm1s, &c, &s);
#pragma unroll
    for (int b = 0; b < kMaxBatch; ++b) {
      if (b < k.nb) {
        const cplx<R> F = tb[((int64_t)(q * k.nb + b) * k.ny + iy) * k.gx + gxi];
        sum[b] += (A)(F.re * c - F.im * s);
      }
    }
  }
  // the first update stores (the accumulator is not cleared beforehand)
  for (int b = 0; b < k.nb; ++b) {
    A* a = &acc[((int64_t)b * k.ny + iy) * k.nx + ix];
    *a = first ? (A)sign * sum[b] : *a + (A)sign * sum[b];
  }
}

// predict: T_q[b][iy][gxi] = (-1)^(tx+ty) dcT[b][iy][ix] exp(+2 pi i w_q nm1s) on image rows (T stays
// zero on the others)
template <typename R, typename A, bool F32C>
__global__ void k_screen_fill_t(KP k, int p0, int npl, const A* __restrict__ dc, cplx<R>* __restrict__ tb) {
  const int gxi = blockIdx.x * kBlock + threadIdx.x;
  const int iy = blockIdx.y;
  if (gxi >= k.gx) return;
  const int hx = k.nx / 2;
  int tx = 0;
  bool in = true;
  if (gxi < hx) tx = gxi;
  else if (gxi >= k.gx - hx) tx = gxi - k.gx;
  else in = false;
  if (!in) return;
  const int ty = iy - k.ny / 2;
  const double nm1s = k.wstack ? nm1s_at<F32C>(tx, ty, k, nullptr) : 0.0;
  const R sign = ((tx + ty) & 1) ? (R)-1 : (R)1;
  const int ix = tx + hx;
  R d[kMaxBatch];
#pragma unroll
  for (int b = 0; b < kMaxBatch; ++b) d[b] = b < k.nb ? (R)dc[((int64_t)b * k.ny + iy) * k.nx + ix] * sign : (R)0;
  for (int q = 0; q < npl; ++q) {
    R c = 1, s = 0;
    if (k.wstack) phase_cs<R, F32C>((k.w0 + (p0 + q) * k.dw) * nm1s, &c, &s);
    for (int b = 0; b < k.nb; ++b)
      tb[((int64_t)(q * k.nb + b) * k.ny + iy) * k.gx + gxi] = cplx<R>{d[b] * c, d[b] * s};
  }
}

// ---------------------------------------------------------------------------------
// Four-step image-side transform (FP64 planes of G = N1 x kFsN2 points along x, N1 = 8 M,
// M in {10, 12, 16, 20, 24}: G = 10240 ... 24576). rocFFT runs FP64 transforms of these
// lengths as two kernels (a strided pass and a row pass), so transpose + transform + screen
// move the image-side data five times. Split x = kFsN2 n1 + n2 and the output index
// g = k1 + N1 k2:
//   X[k1 + N1 k2] = sum_n2 w128^(n2 k2) [ wG^(n2 k1) sum_n1 wN1^(n1 k1) T[kFsN2 n1 + n2] ]
// (w_N = exp(+2 pi i / N), the inverse transform). Stage 1 reads the band plane in 256-byte
// strips of 16 image columns (no transpose), transforms along n1 (radix 8, then radix M),
// applies wG^(n2 k1) and writes U[i][iy][n2][k1]; stage 2 transforms along n2 and adds the
// screened real values of one plane to the image accumulator. 8192^2 image, G = 12288: 0.57 +
// 0.52 ms per plane and polarisation against 2.48 ms for transpose, rocFFT and screen.
// ---------------------------------------------------------------------------------
constexpr int kFsN2 = 128;
// M of a grid length, or 0 when the four-step does not cover it
inline int fs_m_for(int g) {
  if (g % (8 * kFsN2) != 0) return 0;
  const int m = g / (8 * kFsN2);
  return (m == 10 || m == 12 || m == 16 || m == 20 || m == 24) ? m : 0;
}
// twiddle table: wG^j (j < G), then wN1^j, then w128^j
inline int fs_tw_len(int m) { return 8 * m * kFsN2 + 8 * m + kFsN2; }
// compile-time sine / cosine (Taylor series, |x| <= pi after reduction) for the constant
// twiddles of the in-register transforms
constexpr double kFsPi = 3.14159265358979323846;
constexpr double fs_sin_red(double x) {
  double t = x, s = x;
  for (int n = 1; n < 30; ++n) {
    t *= -x * x / ((2 * n) * (2 * n + 1));
    s += t;
  }
  return s;
}
constexpr double fs_red(double x) {
  while (x > kFsPi) x -= 2 * kFsPi;
  while (x < -kFsPi) x += 2 * kFsPi;
  return x;
}
constexpr double fs_sin(double x) { return fs_sin_red(fs_red(x)); }
constexpr double fs_cos(double x) { return fs_sin(x + kFsPi / 2); }
template <typename R>
__device__ __forceinline__ cplx<R> fs_mul(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename R>
__device__ __forceinline__ cplx<R> fs_add(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.re + b.re, a.im + b.im};
}
template <typename R>
__device__ __forceinline__ cplx<R> fs_sub(cplx<R> a, cplx<R> b) {
  return cplx<R>{a.re - b.re, a.im - b.im};
}
// exp(S 2 pi i j / N), j < N, as compile-time tables
template <int N, int S>
struct FsTw {
  double c[N], s[N];
  constexpr FsTw() : c(), s() {
    for (int j = 0; j < N; ++j) {
      c[j] = fs_cos(S * 2.0 * kFsPi * j / N);
      s[j] = fs_sin(S * 2.0 * kFsPi * j / N);
    }
  }
};
// in-register DFTs of size 2, 3, 4 and Cooley-Tukey compositions P x Q (n = Q p + q,
// k = k1 + P k2); S = +1 inverse, -1 forward
template <typename R, int N, int S>
struct FsDft;
template <typename R, int S>
struct FsDft<R, 2, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[2]) {
    const cplx<R> t = x[0];
    x[0] = fs_add(t, x[1]);
    x[1] = fs_sub(t, x[1]);
  }
};
template <typename R, int S>
struct FsDft<R, 3, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[3]) {
    const R c = (R)-0.5, sn = (R)(S * 0.86602540378443864676);
    const cplx<R> t1 = fs_add(x[1], x[2]), t2 = fs_sub(x[1], x[2]);
    const cplx<R> m{x[0].re + c * t1.re, x[0].im + c * t1.im};
    const cplx<R> r{-sn * t2.im, sn * t2.re};
    x[0] = fs_add(x[0], t1);
    x[1] = fs_add(m, r);
    x[2] = fs_sub(m, r);
  }
};
template <typename R, int S>
struct FsDft<R, 4, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[4]) {
    const cplx<R> a0 = fs_add(x[0], x[2]), a1 = fs_sub(x[0], x[2]), a2 = fs_add(x[1], x[3]), a3 = fs_sub(x[1], x[3]);
    const cplx<R> b3 = S > 0 ? cplx<R>{-a3.im, a3.re} : cplx<R>{a3.im, -a3.re};
    x[0] = fs_add(a0, a2);
    x[2] = fs_sub(a0, a2);
    x[1] = fs_add(a1, b3);
    x[3] = fs_sub(a1, b3);
  }
};
template <typename R, int S>
struct FsDft<R, 5, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[5]) {
    // X_k = sum_n x_n exp(S 2 pi i n k / 5): cos / sin of 2 pi / 5 and 4 pi / 5
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
    const R s1 = (R)(S * 0.95105651629515357212), s2 = (R)(S * 0.58778525229247312917);
    const cplx<R> t1 = fs_add(x[1], x[4]), t2 = fs_add(x[2], x[3]);
    const cplx<R> t3 = fs_sub(x[1], x[4]), t4 = fs_sub(x[2], x[3]);
    const cplx<R> a1{x[0].re + c1 * t1.re + c2 * t2.re, x[0].im + c1 * t1.im + c2 * t2.im};
    const cplx<R> a2{x[0].re + c2 * t1.re + c1 * t2.re, x[0].im + c2 * t1.im + c1 * t2.im};
    const cplx<R> b1{s1 * t3.re + s2 * t4.re, s1 * t3.im + s2 * t4.im};
    const cplx<R> b2{s2 * t3.re - s1 * t4.re, s2 * t3.im - s1 * t4.im};
    x[0] = fs_add(x[0], fs_add(t1, t2));
    // a +- i b
    x[1] = cplx<R>{a1.re - b1.im, a1.im + b1.re};
    x[4] = cplx<R>{a1.re + b1.im, a1.im - b1.re};
    x[2] = cplx<R>{a2.re - b2.im, a2.im + b2.re};
    x[3] = cplx<R>{a2.re + b2.im, a2.im - b2.re};
  }
};
template <typename R, int P, int Q, int S>
__device__ __forceinline__ void fs_ct(cplx<R> (&x)[P * Q]) {
  constexpr FsTw<P * Q, S> tw{};
  cplx<R> y[Q][P];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    cplx<R> t[P];
#pragma unroll
    for (int p = 0; p < P; ++p) t[p] = x[Q * p + q];
    FsDft<R, P, S>::run(t);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) y[q][k1] = t[k1];
  }
#pragma unroll
  for (int q = 1; q < Q; ++q)
#pragma unroll
    for (int k1 = 1; k1 < P; ++k1)
      y[q][k1] = fs_mul(y[q][k1], cplx<R>{(R)tw.c[(q * k1) % (P * Q)], (R)tw.s[(q * k1) % (P * Q)]});
#pragma unroll
  for (int k1 = 0; k1 < P; ++k1) {
    cplx<R> t[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) t[q] = y[q][k1];
    FsDft<R, Q, S>::run(t);
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) x[k1 + P * k2] = t[k2];
  }
}
template <typename R, int S>
struct FsDft<R, 8, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[8]) { fs_ct<R, 2, 4, S>(x); }
};
template <typename R, int S>
struct FsDft<R, 12, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[12]) { fs_ct<R, 3, 4, S>(x); }
};
template <typename R, int S>
struct FsDft<R, 10, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[10]) { fs_ct<R, 2, 5, S>(x); }
};
template <typename R, int S>
struct FsDft<R, 20, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[20]) { fs_ct<R, 4, 5, S>(x); }
};
template <typename R, int S>
struct FsDft<R, 24, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[24]) { fs_ct<R, 8, 3, S>(x); }
};
template <typename R, int S>
struct FsDft<R, 16, S> {
  static __device__ __forceinline__ void run(cplx<R> (&x)[16]) { fs_ct<R, 4, 4, S>(x); }
};

// twiddle table of the inverse transform (fs_tw_len(m) entries)
__global__ void k_fs_twiddles(cplx<double>* __restrict__ tw, int m) {
  const int j = blockIdx.x * kBlock + threadIdx.x;
  const int n1 = 8 * m, G = n1 * kFsN2;
  if (j >= G + n1 + kFsN2) return;
  int n, e;
  if (j < G) { n = G; e = j; }
  else if (j < G + n1) { n = n1; e = j - G; }
  else { n = kFsN2; e = j - G - n1; }
  double sn, cs;
  sincospi(2.0 * (double)e / (double)n, &sn, &cs);
  tw[j] = cplx<double>{cs, sn};
}

// stage 1 (invert): block = (n2, 16 image columns, plane i); 16 M threads
template <typename R, int M>
__global__ __launch_bounds__(16 * M) void k_fs_inv1(KP k, int xlo, int xhi, const cplx<R>* __restrict__ plane,
                                                    cplx<R>* __restrict__ U, const cplx<R>* __restrict__ tw) {
  constexpr int N1 = 8 * M, NT = 16 * M, G = N1 * kFsN2;
  __shared__ cplx<R> s[16][N1 + 1];
  const int n2 = blockIdx.x, iy0 = blockIdx.y * 16, i = blockIdx.z;
  const int t = threadIdx.x;
  {
    const int c16 = t & 15, iy = iy0 + c16;
    const bool ok = iy < k.ny;
    const int col = ok ? img_col_to_grid(iy, k.ny, k.gy) : 0;
    // 8 loads in flight per thread (rows n1 = t / 16 + M a), then the LDS writes
    const cplx<R>* pc = plane + (int64_t)i * k.gx * k.gy + col;
    cplx<R> v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int x = kFsN2 * ((t >> 4) + M * a) + n2;
      const bool in = ok && x >= xlo && x < xhi;
      v[a] = pc[in ? (int64_t)x * k.gy : 0];
      if (!in) v[a] = cplx<R>{0, 0};
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) s[c16][(t >> 4) + M * a] = v[a];
  }
  __syncthreads();
  {
    // radix 8 over a of n1 = M a + q, then the twiddle wN1^(q d)
    const int c = t / M, q = t % M;
    cplx<R> x[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) x[a] = s[c][M * a + q];
    FsDft<R, 8, 1>::run(x);
#pragma unroll
    for (int d = 1; d < 8; ++d) x[d] = fs_mul(x[d], tw[G + (q * d) % N1]);
    __syncthreads();
#pragma unroll
    for (int d = 0; d < 8; ++d) s[c][M * d + q] = x[d];
  }
  __syncthreads();
  {
    // radix M over q for each d: k1 = d + 8 e; then wG^(n2 k1)
    const int c = (t >> 3) & 15, d = t & 7;
    cplx<R> x[M];
    if (t < 128) {
#pragma unroll
      for (int q = 0; q < M; ++q) x[q] = s[c][M * d + q];
      FsDft<R, M, 1>::run(x);
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
      for (int e = 0; e < M; ++e) {
        const int k1 = d + 8 * e;
        s[c][k1] = fs_mul(x[e], tw[(n2 * k1) % G]);
      }
    }
  }
  __syncthreads();
  for (int e = t; e < 16 * N1; e += NT) {
    const int c = e / N1, k1 = e - c * N1;
    if (iy0 + c < k.ny) U[(((int64_t)i * k.ny + iy0 + c) * kFsN2 + n2) * N1 + k1] = s[c][k1];
  }
}

// stage 2 (invert): block = (16 values of k1, image row iy, polarisation b); 256 threads; one
// launch per plane (a loop over the planes of a pass makes the compiler hold the transforms'
// constants and both unswitched copies in registers: 214 VGPRs against 129). Transform along
// n2 (128 = 16 x 8: n2 = 8 p + r, k2 = j + 16 e) and add the screened real part
// (-1)^(tx+ty) Re(X exp(-2 pi i w nm1s)) of the image pixels to the accumulator (first != 0
// stores instead)
template <typename R, typename A, bool F32C>
__global__ __launch_bounds__(256) void k_fs_inv2(KP k, int N1, double w, const cplx<R>* __restrict__ U,
                                                 const cplx<R>* __restrict__ tw, A* __restrict__ acc, int first) {
  __shared__ cplx<R> s[16][kFsN2 + 1];
  const int kb = blockIdx.x * 16, iy = blockIdx.y, b = blockIdx.z;
  const int t = threadIdx.x;
  const int G = N1 * kFsN2;
  const cplx<R>* u = U + ((int64_t)b * k.ny + iy) * G;
  {
    // 8 loads in flight per thread, then the LDS writes
    cplx<R> v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = u[((t >> 4) + 16 * i) * N1 + kb + (t & 15)];
#pragma unroll
    for (int i = 0; i < 8; ++i) s[t & 15][(t >> 4) + 16 * i] = v[i];
  }
  __syncthreads();
  {
    // radix 16 over p of n2 = 8 p + r (threads t < 128), then the twiddle w128^(r j)
    const int c = (t >> 3) & 15, r = t & 7;
    cplx<R> x[16];
    if (t < 128) {
#pragma unroll
      for (int pp = 0; pp < 16; ++pp) x[pp] = s[c][8 * pp + r];
      FsDft<R, 16, 1>::run(x);
      // w128^(r j) = (w128^r)^j: one table read, then products
      const cplx<R> step = tw[G + N1 + r];
      cplx<R> cur = step;
#pragma unroll
      for (int j = 1; j < 16; ++j) {
        x[j] = fs_mul(x[j], cur);
        cur = fs_mul(cur, step);
      }
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
      for (int j = 0; j < 16; ++j) s[c][8 * j + r] = x[j];
    }
  }
  __syncthreads();
  // radix 8 over r for each j: k2 = j + 16 e; lanes run along k1, so the accumulator rows are
  // updated in 128-byte pieces
  const int c2 = t & 15, j2 = t >> 4;
  const int ty = iy - k.ny / 2, hx = k.nx / 2;
  cplx<R> x[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) x[r] = s[c2][8 * j2 + r];
  FsDft<R, 8, 1>::run(x);
  A* arow = acc + ((int64_t)b * k.ny + iy) * k.nx + hx;
  // the accumulator reads go out together (pixels off the image read pixel 0 and are dropped)
  A av[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int gxi = kb + c2 + N1 * (j2 + 16 * e);
    const bool in = gxi < hx || gxi >= k.gx - hx;
    av[e] = first ? (A)0 : arow[in ? (gxi < hx ? gxi : gxi - k.gx) : 0];
  }
  // all sums first, then the stores (a store ahead of a use would hold the use's wait, as
  // the memory counter covers both)
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int gxi = kb + c2 + N1 * (j2 + 16 * e);
    const int tx = gxi < hx ? gxi : gxi - k.gx;
    R c = 1, sn = 0;
    if (k.wstack && (gxi < hx || gxi >= k.gx - hx)) phase_cs<R, F32C>(-w * nm1s_at<F32C>(tx, ty, k, nullptr), &c, &sn);
    const R re = x[e].re * c - x[e].im * sn;
    av[e] += ((tx + ty) & 1) ? -(A)re : (A)re;
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int gxi = kb + c2 + N1 * (j2 + 16 * e);
    if (gxi < hx || gxi >= k.gx - hx) arow[gxi < hx ? gxi : gxi - k.gx] = av[e];
  }
}

// predict, the adjoint split (w_N^-1 = conj(w_N)): input g = k1 + N1 k2 (image side),
// output x = kFsN2 n1 + n2 (grid side):
//   X[kFsN2 n1 + n2] = sum_k1 wN1^-(k1 n1) [ wG^-(k1 n2) sum_k2 w128^-(k2 n2) T[k1 + N1 k2] ]
template <typename R>
__device__ __forceinline__ cplx<R> fs_conj(cplx<R> a) {
  return cplx<R>{a.re, -a.im};
}

// stage 2' (predict): block = (16 values of k1, image row iy, polarisation b); 256 threads; one
// launch per plane. Forms T = (-1)^(tx+ty) dc exp(+2 pi i w nm1s) on the image pixels (zero
// on the others) as it loads, transforms along k2 (128 = 8 x 16: k2 = j + 16 e,
// n2 = a + 8 b'), applies wG^-(k1 n2) and writes V[iy][n2][k1]
template <typename R, typename A, bool F32C>
__global__ __launch_bounds__(256) void k_fs_fwd2(KP k, int N1, double w, const A* __restrict__ dc, cplx<R>* __restrict__ V,
                                                 const cplx<R>* __restrict__ tw) {
  __shared__ cplx<R> s[16][kFsN2 + 1];
  const int kb = blockIdx.x * 16, iy = blockIdx.y, b = blockIdx.z;
  const int G = N1 * kFsN2;
  const int t = threadIdx.x;
  const int c2 = t & 15, j2 = t >> 4;  // lanes run along k1: 128-byte pieces of the image rows
  const int ty = iy - k.ny / 2, hx = k.nx / 2;
  const A* drow = dc + ((int64_t)b * k.ny + iy) * k.nx + hx;
  // the 8 image loads go out together (pixels off the image read pixel 0 and are dropped)
  A dv[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int gxi = kb + c2 + N1 * (j2 + 16 * e);
    const bool in = gxi < hx || gxi >= k.gx - hx;
    dv[e] = drow[in ? (gxi < hx ? gxi : gxi - k.gx) : 0];
  }
  cplx<R> x[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int gxi = kb + c2 + N1 * (j2 + 16 * e);
    x[e] = cplx<R>{0, 0};
    if (gxi < hx || gxi >= k.gx - hx) {
      const int tx = gxi < hx ? gxi : gxi - k.gx;
      R d = (R)dv[e];
      if ((tx + ty) & 1) d = -d;
      R c = 1, sn = 0;
      if (k.wstack) phase_cs<R, F32C>(w * nm1s_at<F32C>(tx, ty, k, nullptr), &c, &sn);
      x[e] = cplx<R>{d * c, d * sn};
    }
  }
  // radix 8 over e, then the twiddle w128^-(j a); Y[j][a] goes to 16 a + j
  FsDft<R, 8, -1>::run(x);
#pragma unroll
  for (int a2 = 1; a2 < 8; ++a2) x[a2] = fs_mul(x[a2], fs_conj(tw[G + N1 + (j2 * a2) % kFsN2]));
#pragma unroll
  for (int a2 = 0; a2 < 8; ++a2) s[c2][16 * a2 + j2] = x[a2];
  __syncthreads();
  {
    // radix 16 over j for each a (threads t < 128): n2 = a + 8 b'; then wG^-(k1 n2)
    const int c = (t >> 3) & 15, a2 = t & 7;
    cplx<R> y[16];
    if (t < 128) {
#pragma unroll
      for (int j = 0; j < 16; ++j) y[j] = s[c][16 * a2 + j];
      FsDft<R, 16, -1>::run(y);
    }
    __syncthreads();
    if (t < 128) {
      // wG^-(k1 (a + 8 b')) = wG^-(k1 a) (wG^-(8 k1))^b': two table reads, then products
      const cplx<R> step = fs_conj(tw[(kb + c) * 8]);
      cplx<R> cur = fs_conj(tw[(kb + c) * a2]);
#pragma unroll
      for (int bb = 0; bb < 16; ++bb) {
        s[c][a2 + 8 * bb] = fs_mul(y[bb], cur);
        cur = fs_mul(cur, step);
      }
    }
  }
  __syncthreads();
  cplx<R>* v = V + ((int64_t)b * k.ny + iy) * G;
  for (int e = t; e < 16 * kFsN2; e += 256) {
    const int n2 = e >> 4, c = e & 15;
    v[n2 * N1 + kb + c] = s[c][n2];
  }
}

// stage 1' (predict): block = (n2, 16 image columns, plane i); 16 M threads. Transforms
// V[i][iy][n2][.] along k1 (N1 = 8 M: k1 = d + 8 e, n1 = q + M a) and writes the band rows
// x = kFsN2 n1 + n2 in [xlo, xhi) of the image columns of plane i (the other columns of the
// band rows are left as they are: zero)
template <typename R, int M>
__global__ __launch_bounds__(16 * M) void k_fs_fwd1(KP k, int xlo, int xhi, const cplx<R>* __restrict__ V,
                                                    cplx<R>* __restrict__ plane, const cplx<R>* __restrict__ tw) {
  constexpr int N1 = 8 * M, NT = 16 * M, G = N1 * kFsN2;
  __shared__ cplx<R> s[16][N1 + 1];
  const int n2 = blockIdx.x, iy0 = blockIdx.y * 16, i = blockIdx.z;
  const int t = threadIdx.x;
  {
    // 8 loads in flight per thread (16 N1 = 8 NT elements), then the LDS writes; columns past
    // the image read row iy0 and are dropped
    cplx<R> v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int e = t + NT * a, c = e / N1, k1 = e - c * N1;
      const bool ok = iy0 + c < k.ny;
      v[a] = V[(((int64_t)i * k.ny + iy0 + (ok ? c : 0)) * kFsN2 + n2) * N1 + k1];
      if (!ok) v[a] = cplx<R>{0, 0};
    }
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int e = t + NT * a, c = e / N1;
      s[c][e - c * N1] = v[a];
    }
  }
  __syncthreads();
  {
    // radix M over e for each d (threads t < 128), then the twiddle wN1^-(d q); Y[d][q] goes
    // to M d + q
    const int c = (t >> 3) & 15, d = t & 7;
    cplx<R> x[M];
    if (t < 128) {
#pragma unroll
      for (int e = 0; e < M; ++e) x[e] = s[c][d + 8 * e];
      FsDft<R, M, -1>::run(x);
#pragma unroll
      for (int q = 1; q < M; ++q) x[q] = fs_mul(x[q], fs_conj(tw[G + (d * q) % N1]));
    }
    __syncthreads();
    if (t < 128) {
#pragma unroll
      for (int q = 0; q < M; ++q) s[c][M * d + q] = x[q];
    }
  }
  __syncthreads();
  {
    // radix 8 over d for each q: n1 = q + M a
    const int c = t / M, q = t % M;
    cplx<R> x[8];
#pragma unroll
    for (int d = 0; d < 8; ++d) x[d] = s[c][M * d + q];
    FsDft<R, 8, -1>::run(x);
    __syncthreads();
#pragma unroll
    for (int a2 = 0; a2 < 8; ++a2) s[c][q + M * a2] = x[a2];
  }
  __syncthreads();
  const int c16 = t & 15, iy = iy0 + c16;
  if (iy >= k.ny) return;
  cplx<R>* pc = plane + (int64_t)i * k.gx * k.gy + img_col_to_grid(iy, k.ny, k.gy);
#pragma unroll
  for (int a2 = 0; a2 < 8; ++a2) {
    const int n1 = (t >> 4) + M * a2, x = kFsN2 * n1 + n2;
    if (x >= xlo && x < xhi) pc[(int64_t)x * k.gy] = s[c16][n1];
  }
}

// launches of the stages templated on M
template <typename R, int M>
void fs_stage1_m(bool fwd, const KP& k, int xlo, int xhi, int nplanes, const cplx<R>* in, cplx<R>* out,
                 const cplx<R>* tw, hipStream_t st) {
  const dim3 grid(kFsN2, (k.ny + 15) / 16, nplanes);
  if (fwd) k_fs_fwd1<R, M><<<grid, 16 * M, 0, st>>>(k, xlo, xhi, in, out, tw);
  else k_fs_inv1<R, M><<<grid, 16 * M, 0, st>>>(k, xlo, xhi, in, out, tw);
}
// stage 1 (fwd false: band plane -> U) or stage 1' (fwd true: V -> band plane) over nplanes
template <typename R>
void fs_stage1(int m, bool fwd, const KP& k, int xlo, int xhi, int nplanes, const cplx<R>* in, cplx<R>* out,
               const cplx<R>* tw, hipStream_t st) {
  switch (m) {
    case 10: fs_stage1_m<R, 10>(fwd, k, xlo, xhi, nplanes, in, out, tw, st); break;
    case 12: fs_stage1_m<R, 12>(fwd, k, xlo, xhi, nplanes, in, out, tw, st); break;
    case 16: fs_stage1_m<R, 16>(fwd, k, xlo, xhi, nplanes, in, out, tw, st); break;
    case 20: fs_stage1_m<R, 20>(fwd, k, xlo, xhi, nplanes, in, out, tw, st); break;
    case 24: fs_stage1_m<R, 24>(fwd, k, xlo, xhi, nplanes, in, out, tw, st); break;
    default: HIPNG_THROW(HIPNG_ERUNTIME, "four-step transform: no stage for M = %d", m);
  }
}

// dirty[b][ix][iy] = accT[b][iy][ix] * correction(ix, iy); 32 x 32 tiles through LDS
template <typename A, typename O>
__global__ __launch_bounds__(kBlock) void k_finalize_t(KP k, const A* __restrict__ acc, O* __restrict__ dirty,
                                                       const double* __restrict__ phix,
                                                       const double* __restrict__ phiy,
                                                       const double* __restrict__ qz,
                                                       const double* __restrict__ qc, int nq) {
  __shared__ A t[32][33];
  const int ix0 = blockIdx.x * 32, iy0 = blockIdx.y * 32;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  A f[4];
#pragma unroll
  for (int r4 = 0; r4 < 4; ++r4) {
    const int ix = ix0 + ly + 8 * r4, iy = iy0 + lx;
    f[r4] = (ix < k.nx && iy < k.ny) ? corr_at<A>(ix, iy, k, phix, phiy, qz, qc, nq) : (A)0;
  }
  for (int b = 0; b < k.nb; ++b) {
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const int iy = iy0 + ly + 8 * r4, ix = ix0 + lx;
      t[ly + 8 * r4][lx] = (ix < k.nx && iy < k.ny) ? acc[((int64_t)b * k.ny + iy) * k.nx + ix] : (A)0;
    }
    __syncthreads();
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const int ix = ix0 + ly + 8 * r4, iy = iy0 + lx;
      if (ix < k.nx && iy < k.ny) dirty[((int64_t)b * k.nx + ix) * k.ny + iy] = (O)(t[lx][ly + 8 * r4] * f[r4]);
    }
    __syncthreads();
  }
}

// The same two passes for images in [b][iy][ix] order (the accumulator's own): no transpose,
// and invert scales polarisation b by scale[b] (the normalisation by the sum of weights).
struct Scale4 {
  double s[4];
};
template <typename A, typename O>
__global__ __launch_bounds__(kBlock) void k_finalize_yx(KP k, const A* __restrict__ acc, O* __restrict__ dirty,
                                                        const double* __restrict__ phix,
                                                        const double* __restrict__ phiy,
                                                        const double* __restrict__ qz,
                                                        const double* __restrict__ qc, int nq, Scale4 sc) {
  const int ix = blockIdx.x * kBlock + threadIdx.x, iy = blockIdx.y;
  if (ix >= k.nx) return;
  const A f = corr_at<A>(ix, iy, k, phix, phiy, qz, qc, nq);
  for (int b = 0; b < k.nb; ++b) {
    const int64_t i = ((int64_t)b * k.ny + iy) * k.nx + ix;
    dirty[i] = (O)((double)(acc[i] * f) * sc.s[b]);
  }
}
template <typename I, typename A>
__global__ __launch_bounds__(kBlock) void k_prescale_yx(KP k, const I* __restrict__ dirty, A* __restrict__ dc,
                                                        const double* __restrict__ phix,
                                                        const double* __restrict__ phiy,
                                                        const double* __restrict__ qz,
                                                        const double* __restrict__ qc, int nq) {
  const int ix = blockIdx.x * kBlock + threadIdx.x, iy = blockIdx.y;
  if (ix >= k.nx) return;
  const A f = corr_at<A>(ix, iy, k, phix, phiy, qz, qc, nq);
  for (int b = 0; b < k.nb; ++b) {
    const int64_t i = ((int64_t)b * k.ny + iy) * k.nx + ix;
    dc[i] = (A)dirty[i] * f;
  }
}

// dcT[b][iy][ix] = dirty[b][ix][iy] * correction(ix, iy)
template <typename I, typename A>
__global__ __launch_bounds__(kBlock) void k_prescale_t(KP k, const I* __restrict__ dirty, A* __restrict__ dc,
                                                       const double* __restrict__ phix,
                                                       const double* __restrict__ phiy,
                                                       const double* __restrict__ qz,
                                                       const double* __restrict__ qc, int nq) {
  __shared__ A t[32][33];
  const int ix0 = blockIdx.x * 32, iy0 = blockIdx.y * 32;
  const int lx = threadIdx.x & 31, ly = threadIdx.x >> 5;
  A f[4];
#pragma unroll
  for (int r4 = 0; r4 < 4; ++r4) {
    const int ix = ix0 + ly + 8 * r4, iy = iy0 + lx;
    f[r4] = (ix < k.nx && iy < k.ny) ? corr_at<A>(ix, iy, k, phix, phiy, qz, qc, nq) : (A)0;
  }
  for (int b = 0; b < k.nb; ++b) {
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const int ix = ix0 + ly + 8 * r4, iy = iy0 + lx;
      t[ly + 8 * r4][lx] = (ix < k.nx && iy < k.ny) ? (A)dirty[((int64_t)b * k.nx + ix) * k.ny + iy] * f[r4] : (A)0;
    }
    __syncthreads();
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
      const int iy = iy0 + ly + 8 * r4, ix = ix0 + lx;
      if (ix < k.nx && iy < k.ny) dc[((int64_t)b * k.ny + iy) * k.nx + ix] = t[lx][ly + 8 * r4];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------
// Degridder: persistent workgroups over (tile, chunk) tasks of one pass of npl w planes.
// A task loads the tile of all npl planes with its (W-1) halo into LDS, cell-major with the
// NB polarisations of a cell contiguous, and degrids one chunk of the tile's visibilities:
// one lane per visibility, u/v taps evaluated once per pass, separable contraction in
// registers for every plane of the pass the visibility's w window meets, one
// read-modify-write of the sorted accumulator per pass; no atomics. The candidates of a
// start-plane segment share their plane range, so the plane loop does not diverge.
// ---------------------------------------------------------------------------------
// polarisations per 16-byte unit of the degridder's LDS tile
template <typename R, int NB>
__host__ __device__ constexpr int dg_per() {
  return NB < (int)(16 / sizeof(cplx<R>)) ? NB : (int)(16 / sizeof(cplx<R>));
}
template <typename R, int NB>
__host__ __device__ constexpr int dg_slots() {
  return (NB + dg_per<R, NB>() - 1) / dg_per<R, NB>() * dg_per<R, NB>();
}

template <typename R, int W, int NB>
__global__ __launch_bounds__(kBlock) void k_degrid(KP k, int p0, int npl, const uint32_t* __restrict__ bstart,
                                                   const uint32_t* __restrict__ bend,
                                                   const uint2* __restrict__ srec,
                                                   const uint32_t* __restrict__ rec_a,
                                                   const rec3<R>* __restrict__ rec_f,
                                                   const cplx<R>* __restrict__ plane, int64_t plane_stride,
                                                   cplx<R>* __restrict__ accs,
                                                   const uint32_t* __restrict__ task_off,
                                                   const uint32_t* __restrict__ task_tab, uint32_t chunk,
                                                   uint32_t* __restrict__ counters, TapPoly tp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // [U][npl][L][L][P]: the NB polarisations of a cell in U units of P (16 bytes: two FP32
  // or one FP64 polarisation), each unit a plane of its own, so that lanes reading nearby
  // cells of a unit touch consecutive 16-byte slots (a cell-major [L][L][NB] tile puts the
  // same half of consecutive cells 32 bytes apart, so two of every eight lanes share banks)
  constexpr int P = dg_per<R, NB>(), U = (NB + P - 1) / P;
  cplx<R>* g = reinterpret_cast<cplx<R>*>(smem_raw);
  __shared__ uint32_t s_task;
  __shared__ uint32_t s_segj[64], s_segv[65];  // per start plane: first record, exclusive prefix
  const int T = k.T, L = k.T + W - 1, LL = L * L;
  const int ustride = npl * LL * P;  // elements between units
  const R scale = (R)2 / (R)W;
  // the FP32 tap evaluation works in base 2
  const R beta = std::is_same<R, float>::value ? (R)(k.beta * 1.4426950408889634) : (R)k.beta;
  const uint32_t total = counters[1];
  int slo, shi;
  chunk_range(k.Ww, k.nplanes, p0, npl, &slo, &shi);
  const int nseg = shi - slo + 1;  // <= W_w + npl - 1 <= 64
  int loaded = -1;
  // XCD-local work queues: tasks of neighbouring tiles in y share halo rows of the planes, so
  // queue x holds the contiguous x-th eighth of the tasks and a workgroup pops from the queue
  // of its own XCD (workgroups are placed on XCDs round-robin by index) before the others
  const uint32_t tper = (total + kXcd - 1) / kXcd;
  const int xcd0 = (int)(blockIdx.x % kXcd);
  int qi = 0;  // used by thread 0 only
  for (;;) {
    if (threadIdx.x == 0) {
      uint32_t t = 0xffffffffu;
      while (qi < kXcd) {
        const int x = (xcd0 + qi) % kXcd;
        const uint32_t t0 = min(total, (uint32_t)x * tper), qsize = min(total - t0, tper);
        const uint32_t l = queue_pop(&counters[kQueueBase + x * kQueueStride], qsize);
        if (l < qsize) {
          t = t0 + l;
          break;
        }
        ++qi;
      }
      s_task = t;
    }
    __syncthreads();
    const uint32_t task = s_task;
    if (task == 0xffffffffu) break;
    const int tile = (int)task_tab[task];
    const uint32_t vlo = (task - task_off[tile]) * chunk, vhi = vlo + chunk;
    const int tx = tile / k.nty, ty = tile - tx * k.nty;
    if (tile != loaded) {
      // coalesced along grid columns; LDS written cell-major (polarisation fastest). Thread t
      // takes column t % CW (CW >= L) of every RS-th row, row = (q * NB + b) * L + rr advanced
      // without divisions, and has kTileLoads loads in flight before it writes them to LDS
      constexpr int kTileLoads = sizeof(R) == 4 ? 8 : 4;
      const int nrows = npl * NB * L;
      const int CW = L > 32 ? 64 : 32, RS = kBlock / CW;  // RS < L: a row step wraps rr at most once
      const int cc = threadIdx.x & (CW - 1);
      if (cc < L) {
        int gyi = ty * T + cc;
        if (gyi >= k.gy) gyi -= k.gy;
        int row = threadIdx.x / CW, rr = row, qb = 0;
        while (row < nrows) {
          cplx<R> v[kTileLoads];
          int dst[kTileLoads];
#pragma unroll
          for (int u = 0; u < kTileLoads; ++u) {
            const bool ok = row < nrows;
            const int r1 = ok ? rr : 0, qb1 = ok ? qb : 0;  // past the end: reload row 0
            const int b = qb1 % NB, q = qb1 / NB;
            int gxi = tx * T + r1;
            if (gxi >= k.gx) gxi -= k.gx;
            v[u] = plane[q * plane_stride + ((int64_t)b * k.gx + gxi) * k.gy + gyi];
            dst[u] = ok ? ((b / P) * npl * LL + (q * L + r1) * L + cc) * P + b % P : -1;
            row += RS;
            rr += RS;
            if (rr >= L) {
              rr -= L;
              ++qb;
            }
          }
#pragma unroll
          for (int u = 0; u < kTileLoads; ++u)
            if (dst[u] >= 0) g[dst[u]] = v[u];
        }
      }
      loaded = tile;
      __syncthreads();
    }
    // the task's candidates as one list over the buckets of start planes [slo, shi]
    // (segment table in LDS), so that every lane takes an item whatever the bucket sizes
    if (threadIdx.x < 64) {
      const int sg = threadIdx.x;
      uint32_t len = 0;
      if (sg < nseg) {
        const uint32_t key = (uint32_t)(slo + sg) * k.ntiles + (uint32_t)tile;
        const uint32_t st0 = bstart[key];
        len = bend[key] - st0;
        s_segj[sg] = st0;
      }
      uint32_t inc = len;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t tt = __shfl_up(inc, o, 64);
        if (sg >= o) inc += tt;
      }
      if (sg < nseg) s_segv[sg + 1] = inc;
      if (sg == 0) s_segv[0] = 0;
    }
    __syncthreads();
    const uint32_t vend = min(vhi, s_segv[nseg]);
    for (uint32_t vi = vlo + threadIdx.x; vi < vend; vi += kBlock) {
      int sg = 0;
      {
        int hi = nseg;  // last segment starting at or before vi
        while (hi - sg > 1) {
          const int mid = (sg + hi) >> 1;
          if (s_segv[mid] <= vi) sg = mid;
          else hi = mid;
        }
      }
      // candidate: a record, or (srec given) an entry of the non-dense record list
      const uint32_t sj = s_segj[sg] + (vi - s_segv[sg]);
      uint32_t j = sj, Ar;
      if (srec) {
        const uint2 e = srec[sj];
        j = e.x;
        Ar = e.y;
      } else {
        Ar = rec_a[sj];
      }
      const int s = slo + sg;
      // planes of the pass inside the W_w-plane window of start plane s
      const int qlo = max(0, s - p0), qhi = min(npl, s + k.Ww - p0);
      {
        const uint32_t A = Ar & k.amask;
        const int rx0 = (int)(A & 0xffffu) - tx * T, ry0 = (int)(A >> 16) - ty * T;
        const rec3<R> f = rec_f[j];
        R kx[W], ky[W];
        taps_axis<R, W>(tp, f.x, scale, beta, kx);
        taps_axis<R, W>(tp, f.y, scale, beta, ky);
        R sr[NB], si[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) sr[b] = si[b] = 0;
        for (int q = qlo; q < qhi; ++q) {
          const R kw = k.wstack ? esk_fast((f.w + (R)(p0 + q - s)) * scale, beta) : (R)1;
          const cplx<R>* gq = g + ((q * L + rx0) * L + ry0) * P;
          R tr[NB], ti[NB];
#pragma unroll
          for (int b = 0; b < NB; ++b) tr[b] = ti[b] = 0;
#pragma unroll
          for (int ix = 0; ix < W; ++ix) {
            R rr[NB], ri[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) rr[b] = ri[b] = 0;
#pragma unroll
            for (int iy = 0; iy < W; ++iy) {
#pragma unroll
              for (int b = 0; b < NB; ++b) {
                const cplx<R> v = gq[(b / P) * ustride + (ix * L + iy) * P + b % P];
                rr[b] += ky[iy] * v.re;
                ri[b] += ky[iy] * v.im;
              }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              tr[b] += kx[ix] * rr[b];
              ti[b] += kx[ix] * ri[b];
            }
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            sr[b] += kw * tr[b];
            si[b] += kw * ti[b];
          }
        }
        if (qlo < qhi) {
          // the launch holding the start plane is the first to reach the record: it stores
          const bool first = s >= p0;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            cplx<R> o = first ? cplx<R>{(R)0, (R)0} : accs[(int64_t)j * NB + b];
            o.re += sr[b];
            o.im += si[b];
            accs[(int64_t)j * NB + b] = o;
          }
        }
      }
    }
    __syncthreads();
  }
}

// Dense-run degridder on the matrix cores (W <= 8, up to 2 planes per launch; FP32 or FP64).
// The visibilities of a run share their footprint window, so per polarisation b
//   U_b[(c, h)][v] = sum_(a, q) (h ? Im : Re) G_q[ax + a][ay + c][b] * kx_v[a] kw_v[q]
// is the product of the window's 16 x 16 matrix and a 16-row column per visibility, taken
// 16 visibilities x 4 terms at a time by v_mfma_f32_16x16x4_f32 / v_mfma_f64_16x16x4_f64
// (lane l supplies A row l % 16, column 4 s + l / 16 and B row 4 s + l / 16, column l % 16 of
// step s). In FP32 register r of lane l receives row 4 (l / 16) + r, so a lane holds 4 of the
// 8 y cells of one visibility and part h: the y contraction sum_c ky_v[c] U_b[(c, h)][v] is 4
// products in registers plus one exchange with lane l ^ 16. In FP64 it receives row
// 4 r + l / 16 (y cells l / 16 and 4 + l / 16 of both parts): 2 products per part and a sum
// over the four lanes of the visibility. The results are staged in LDS and added to the
// visibilities' accumulators one visibility per lane. Tasks are the dense-run gridder's; the
// other records go through k_degrid.
template <typename R, int NB>
__global__ __launch_bounds__(kBlock) void k_degrid_dense_mx(KP k, int p0, int npl, const uint4* __restrict__ dtask,
                                                            const uint32_t* __restrict__ dplane,
                                                            const rec3<R>* __restrict__ rec_f,
                                                            const cplx<R>* __restrict__ plane,
                                                            int64_t plane_stride, cplx<R>* __restrict__ accs,
                                                            uint32_t* __restrict__ counters, TapPoly tp) {
  static_assert(kMaxPC == 2, "the 16 columns of A are 8 cells x 2 planes");
  constexpr int NW = kBlock / 64;
  // taps tap-major, row strides chosen so that the reads of a half wave (16 visibilities x
  // 2 rows) fall in distinct banks: x rows 4 s + {0, 1} are 16 banks apart (80 = 16 mod 32),
  // y rows r and 4 + r likewise (4 x 68 = 16 mod 32); results with an odd stride per visibility
  using M = MfmaT<R>;
  __shared__ R s_tx[NW][8][80];
  __shared__ R s_ty[NW][8][68];
  __shared__ R s_tw[NW][64][2];
  __shared__ R s_out[NW][64][2 * NB + 1];
  const int W = k.W;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int gk = lane >> 4, gm = lane & 15;
  const int c8 = gm & 7, h2 = gm >> 3;
  const R scale = (R)2 / (R)W;
  // the FP32 tap evaluation works in base 2
  const R beta = std::is_same<R, float>::value ? (R)(k.beta * 1.4426950408889634) : (R)k.beta;
  int slo, shi;
  chunk_range(k.Ww, k.nplanes, p0, npl, &slo, &shi);
  const uint32_t total = counters[kDenseTotal];
  const uint32_t lo = lower_bound_u32(dplane, total, (uint32_t)slo);
  const uint32_t hi = lower_bound_u32(dplane, total, (uint32_t)shi + 1u);
  for (;;) {
    uint32_t t = 0;
    if (lane == 0) t = queue_pop(&counters[kDenseCnt], hi - lo);
    t = lo + __shfl(t, 0, 64);
    if (t >= hi) break;
    const uint4 d = dtask[t];  // first record, count, packed first cell, start plane
    const uint4 cur = make_uint4(__builtin_amdgcn_readfirstlane(d.x), __builtin_amdgcn_readfirstlane(d.y),
                                 __builtin_amdgcn_readfirstlane(d.z), __builtin_amdgcn_readfirstlane(d.w));
    const int sp = (int)cur.w;
    // A: the run's window, row (c, h) = l % 16, column (a, q) = 4 s + l / 16 of step s
    R bw[NB][4];
    {
      int gyi = (int)(cur.z >> 16) + c8;
      if (gyi >= k.gy) gyi -= k.gy;
#pragma unroll
      for (int st = 0; st < 4; ++st) {
        const int kk = 4 * st + gk;
        const int a = kk & 7, q = kk >> 3;
        int gxi = (int)(cur.z & 0xffffu) + a;
        if (gxi >= k.gx) gxi -= k.gx;
        const bool in = q < npl && a < W && c8 < W;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          bw[b][st] = in ? reinterpret_cast<const R*>(
                               plane + q * plane_stride + ((int64_t)b * k.gx + gxi) * k.gy + gyi)[h2]
                         : (R)0;
      }
    }
    for (uint32_t b0 = 0; b0 < cur.y; b0 += 64) {
      const int cnt = (int)min(64u, cur.y - b0);
      {
        const bool live = lane < cnt;
        rec3<R> f{};
        if (live) f = rec_f[cur.x + b0 + (uint32_t)lane];
        R tx8[8], ty8[8];
        taps8(tp, W, f.x, scale, beta, live, tx8);
        taps8(tp, W, f.y, scale, beta, live, ty8);
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          s_tx[wv][i][lane] = tx8[i];
          s_ty[wv][i][lane] = ty8[i];
        }
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int iw = p0 + q - sp;
          s_tw[wv][lane][q] = (live && q < npl && iw >= 0 && iw < k.Ww)
                                  ? (k.wstack ? esk_fast((f.w + (R)iw) * scale, beta) : (R)1)
                                  : (R)0;
        }
      }
      wave_sync();
      for (int g0 = 0; g0 < cnt; g0 += 16) {
        const int m = g0 + gm;  // this lane's visibility (B column, D column)
        R av[4], ky[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const int kk = 4 * st + gk;
          av[st] = s_tx[wv][kk & 7][m] * s_tw[wv][m][kk >> 3];
        }
        if constexpr (std::is_same<R, float>::value) {
          // D rows 4 gk + r = (c, h): c = 4 (gk & 1) + r, h = gk >> 1
#pragma unroll
          for (int r = 0; r < 4; ++r) ky[r] = s_ty[wv][4 * (gk & 1) + r][m];
        } else {
          // D rows 4 r + gk = (c, h): c = gk (r even) or 4 + gk (r odd), h = r >> 1
          ky[0] = ky[2] = s_ty[wv][gk][m];
          ky[1] = ky[3] = s_ty[wv][4 + gk][m];
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          typename M::V u = {0, 0, 0, 0};
#pragma unroll
          for (int st = 0; st < 4; ++st) u = M::mac(bw[b][st], av[st], u);
          if constexpr (std::is_same<R, float>::value) {
            R x = u[0] * ky[0] + u[1] * ky[1] + u[2] * ky[2] + u[3] * ky[3];
            x += __shfl_xor(x, 16, 64);
            if ((gk & 1) == 0) s_out[wv][m][2 * b + (gk >> 1)] = x;
          } else {
            // this lane holds c in {gk, 4 + gk} of both h; the four lanes of visibility m add up
            R x0 = u[0] * ky[0] + u[1] * ky[1], x1 = u[2] * ky[2] + u[3] * ky[3];
            x0 += __shfl_xor(x0, 16, 64);
            x1 += __shfl_xor(x1, 16, 64);
            x0 += __shfl_xor(x0, 32, 64);
            x1 += __shfl_xor(x1, 32, 64);
            if (gk == 0) {
              s_out[wv][m][2 * b] = x0;
              s_out[wv][m][2 * b + 1] = x1;
            }
          }
        }
      }
      wave_sync();
      if (lane < cnt) {
        cplx<R>* o = accs + (int64_t)(cur.x + b0 + (uint32_t)lane) * NB;
        const bool first = (int)cur.w >= p0;  // the launch holding the run's start plane stores
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          cplx<R> a = first ? cplx<R>{0, 0} : o[b];
          a.re += s_out[wv][lane][2 * b];
          a.im += s_out[wv][lane][2 * b + 1];
          o[b] = a;
        }
      }
      wave_sync();
    }
  }
}

// predict epilogue: ms[perm[j]] = wgt conj?(acc_sorted[j] exp(+2 pi i |w| nshift f/c))
template <typename R, int NB, bool VEC>
__global__ void k_unsort(const uint32_t* __restrict__ perm, int64_t n, const cplx<R>* __restrict__ rec_p,
                         const uint32_t* __restrict__ flipb, const cplx<R>* __restrict__ accs,
                         const R* __restrict__ wgt, cplx<R>* __restrict__ ms) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  const bool flip = (flipb[j >> 5] >> (j & 31)) & 1u;
  const cplx<R> ph = rec_p[j];
  const R c = ph.re, s = ph.im;
  cplx<R> a[NB], o[NB];
  load_vals<R, NB, true>(accs + j * NB, a);
  R gw[NB];
  if (wgt) load_wgts<R, NB, VEC>(wgt + (int64_t)i * NB, gw);
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    o[b].re = a[b].re * c - a[b].im * s;
    o[b].im = a[b].re * s + a[b].im * c;
    if (flip) o[b].im = -o[b].im;
    if (wgt) {
      o[b].re *= gw[b];
      o[b].im *= gw[b];
    }
  }
  store_vals<R, NB, VEC>(ms + (int64_t)i * NB, o);
}

// ---------------------------------------------------------------------------------
// Imaging weights (uniform / robust): nearest-cell weight density on the image-sized uv
// grid and the per-visibility reweighting of RASCIL's weight_visibility. Grid layout
// [nimchan][ny][nx][npol]: the polarisations of a cell share one 32-byte segment.
// ---------------------------------------------------------------------------------
constexpr int kWgtMaxPol = 4;

struct WgtArgs {
  int64_t ntimes, nbl;  // rows are [time][baseline]
  int nchan, npol, nimchan, nx, ny;
  double du, dv;        // grid cell sizes in wavelengths (du negative: RA increases to the left)
  int ch_lanes;         // channels per wave (min(nchan, 64))
  int bl_per_wave;      // baselines per wave (64 / ch_lanes)
  int tseg;             // times per work unit of the density pass
};

__device__ __forceinline__ bool weight_cell(const double* __restrict__ uvw, int64_t row, double fc,
                                            const WgtArgs& a, int* pu, int* pv) {
  const double u = uvw[3 * row] * fc, v = uvw[3 * row + 1] * fc;
  *pu = (int)rint(u / a.du + (double)(a.nx / 2 + 1) - 1.0);
  *pv = (int)rint(v / a.dv + (double)(a.ny / 2 + 1) - 1.0);
  return *pu >= 0 && *pu < a.nx && *pv >= 0 && *pv < a.ny;
}

// Density grid[imchan][pv][pu][pol] += imaging_weight, and red[1] += sum(imaging_weight).
// A lane holds one polarisation of one (baseline, channel): the npol lanes of a visibility
// are adjacent, so their adds to a cell are one 32-byte segment of one instruction. A work
// unit is (segment of tseg times, bl_per_wave baselines, ch_lanes channels); each lane
// follows its (baseline, channel, pol) along the segment and sums the weights of
// consecutive times that fall on the same cell in registers (a short baseline moves a
// fraction of a cell per time step), adding them to the grid when the cell changes. The hot
// cells of the short baselines thus see a fraction of the atomics of one add per
// visibility. Persistent waves stride over the units; out-of-grid visibilities raise flag[0].
__global__ __launch_bounds__(kBlock) void k_weight_grid(const double* __restrict__ uvw,
                                                        const double* __restrict__ fc,
                                                        const int* __restrict__ vis_to_im, WgtArgs a,
                                                        const double* __restrict__ wt,
                                                        double* __restrict__ grid, double* __restrict__ red,
                                                        unsigned* __restrict__ flag, int64_t nunits) {
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = ((int64_t)gridDim.x * kBlock) >> 6;
  const int np = a.npol;
  const int isub = lane / np, p = lane - isub * np;
  const int bsub = isub / a.ch_lanes, csub = isub - bsub * a.ch_lanes;
  const int ncb = (a.nchan + a.ch_lanes - 1) / a.ch_lanes;
  const int64_t nbg = (a.nbl + a.bl_per_wave - 1) / a.bl_per_wave;
  double tot = 0.0;
  bool bad = false;
  for (int64_t u = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6; u < nunits; u += nwaves) {
    const int cb = (int)(u % ncb);
    const int64_t r = u / ncb;
    const int64_t bg = r % nbg, ts = r / nbg;
    const int64_t b = bg * a.bl_per_wave + bsub;
    const int c = cb * a.ch_lanes + csub;
    if (bsub >= a.bl_per_wave || b >= a.nbl || c >= a.nchan) continue;
    const double f = fc[c];
    const int64_t icell = (int64_t)vis_to_im[c] * a.ny;
    const int64_t t0 = ts * a.tseg, t1 = min(a.ntimes, t0 + a.tseg);
    int64_t cur = -1;
    double sum = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
      const int64_t row = t * a.nbl + b;
      int pu, pv;
      if (!weight_cell(uvw, row, f, a, &pu, &pv)) {
        bad = true;
        continue;
      }
      const int64_t cell = (icell + pv) * a.nx + pu;
      const double w = wt[(row * a.nchan + c) * np + p];
      if (cell != cur) {
        if (cur >= 0 && sum != 0.0) atomicAdd(&grid[cur * np + p], sum);
        cur = cell;
        sum = 0.0;
      }
      sum += w;
      tot += w;
    }
    if (cur >= 0 && sum != 0.0) atomicAdd(&grid[cur * np + p], sum);
  }
  // one add of the weight total per wave
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) tot += __shfl_xor(tot, o, 64);
  if (lane == 0 && tot != 0.0) atomicAdd(&red[1], tot);
  if (bad) atomicOr(flag, 1u);
}

// Threads stride over the weights (row, chan, pol) in memory order (a version that walks the
// density pass's time tracks for cell locality measured slower at the production shape:
// fewer independent loads in flight).
// mode 1: w <- w * w / g ; mode 2: w <- w * (1 + f2 w) / (1 + f2 g) ; g <= 0 -> 0
__global__ void k_weight_apply(const double* __restrict__ uvw, const double* __restrict__ fc,
                               const int* __restrict__ vis_to_im, WgtArgs a,
                               const double* __restrict__ grid, int mode, double f2,
                               double* __restrict__ wt) {
  const int per_row = a.nchan * a.npol;
  const int64_t total = a.ntimes * a.nbl * per_row;
  for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < total; j += (int64_t)gridDim.x * kBlock) {
    const int64_t row = j / per_row;
    const int rem = (int)(j - row * per_row);
    const int ch = rem / a.npol, p = rem - ch * a.npol;
    int pu, pv;
    if (!weight_cell(uvw, row, fc[ch], a, &pu, &pv)) continue;
    const double g = grid[(((int64_t)vis_to_im[ch] * a.ny + pv) * a.nx + pu) * a.npol + p];
    const double w = wt[j];
    wt[j] = g > 0.0 ? (mode == 1 ? w * (w / g) : w * ((1.0 + f2 * w) / (1.0 + f2 * g))) : 0.0;
  }
}

// out[0] += sum(x^2), out[1] += sum(y)
__global__ void k_sum2(const double* __restrict__ x, int64_t nx, const double* __restrict__ y,
                       int64_t ny, double* __restrict__ out) {
  __shared__ double s[2][kBlock];
  double a = 0.0, b = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nx; i += (int64_t)gridDim.x * kBlock)
    a += x[i] * x[i];
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < ny; i += (int64_t)gridDim.x * kBlock)
    b += y[i];
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s[0][threadIdx.x] += s[0][threadIdx.x + o];
      s[1][threadIdx.x] += s[1][threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicAdd(&out[0], s[0][0]);
    atomicAdd(&out[1], s[1][0]);
  }
}

// ---------------------------------------------------------------------------------
// FP64 point-source DFT: one thread per (time*baseline, channel); components looped,
// polarisations in registers.
// ---------------------------------------------------------------------------------
__global__ void k_dft(const double* __restrict__ dc, const cplx<double>* __restrict__ flux,
                      const double* __restrict__ uvwl, int ncomp, int64_t ntb, int nchan, int npol,
                      cplx<double>* __restrict__ vis) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= ntb * nchan) return;
  const int ch = (int)(t % nchan);
  const double u = uvwl[3 * t], v = uvwl[3 * t + 1], w = uvwl[3 * t + 2];
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0};
  for (int c = 0; c < ncomp; ++c) {
    const double ph = u * dc[3 * c] + v * dc[3 * c + 1] + w * dc[3 * c + 2];
    double s, co;
    sincospi(-2.0 * ph, &s, &co);
    const cplx<double>* fl = flux + ((int64_t)c * nchan + ch) * npol;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p < npol) {
        const cplx<double> f = fl[p];
        ar[p] += f.re * co - f.im * s;
        ai[p] += f.re * s + f.im * co;
      }
    }
  }
#pragma unroll
  for (int p = 0; p < 4; ++p)
    if (p < npol) vis[t * npol + p] = cplx<double>{ar[p], ai[p]};
}

// Inverse DFT of components (dft.py:335-380): per component k, channel c and polarisation p,
// partial sums over a block's share of the (time, baseline) rows of
//   w (1 - f) v exp(+2 pi i (u l_k + v m_k + w (n_k - 1)))   and   w (1 - f),
// reduced in LDS in a fixed order; k_idft_final adds the blocks' partials in block order and
// divides (0 where the weight sum is not positive). Deterministic: no atomics.
constexpr int kIdftBlocks = 256;  // blocks per (component, channel) over the rows
__global__ __launch_bounds__(kBlock) void k_idft_partial(const double* __restrict__ dc,
                                                         const cplx<double>* __restrict__ vis,
                                                         const double* __restrict__ wt,
                                                         const int* __restrict__ flags,
                                                         const double* __restrict__ uvwl, int64_t ntb, int nchan,
                                                         int npol, cplx<double>* __restrict__ part,
                                                         double* __restrict__ wpart) {
  const int ch = (int)blockIdx.y, k = (int)blockIdx.z;
  const double l = dc[3 * k], m = dc[3 * k + 1], nm1 = dc[3 * k + 2];
  double ar[4] = {0, 0, 0, 0}, ai[4] = {0, 0, 0, 0}, ws[4] = {0, 0, 0, 0};
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < ntb; r += (int64_t)gridDim.x * kBlock) {
    const int64_t t = r * nchan + ch;
    const double ph = uvwl[3 * t] * l + uvwl[3 * t + 1] * m + uvwl[3 * t + 2] * nm1;
    double sn, co;
    sincospi(2.0 * ph, &sn, &co);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p < npol) {
        const int64_t i = t * npol + p;
        const double w = flags ? wt[i] * (double)(1 - flags[i]) : wt[i];
        const cplx<double> x = vis[i];
        ar[p] += w * (x.re * co - x.im * sn);
        ai[p] += w * (x.re * sn + x.im * co);
        ws[p] += w;
      }
    }
  }
  __shared__ double red[3 * 4][kBlock];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    red[3 * p][threadIdx.x] = ar[p];
    red[3 * p + 1][threadIdx.x] = ai[p];
    red[3 * p + 2][threadIdx.x] = ws[p];
  }
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int q = 0; q < 12; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < npol) {
    const int p = threadIdx.x;
    const int64_t o = (((int64_t)k * nchan + ch) * npol + p) * gridDim.x + blockIdx.x;
    part[o] = cplx<double>{red[3 * p][0], red[3 * p + 1][0]};
    if (k == 0) wpart[((int64_t)ch * npol + p) * gridDim.x + blockIdx.x] = red[3 * p + 2][0];
  }
}

__global__ void k_idft_final(const cplx<double>* __restrict__ part, const double* __restrict__ wpart, int ncomp,
                             int nchan, int npol, int nblk, cplx<double>* __restrict__ flux,
                             double* __restrict__ weight) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // (k, ch, p)
  if (e >= (int64_t)ncomp * nchan * npol) return;
  const int64_t cp = e % ((int64_t)nchan * npol);
  double w = 0.0, re = 0.0, im = 0.0;
  for (int b = 0; b < nblk; ++b) {
    w += wpart[cp * nblk + b];
    re += part[e * nblk + b].re;
    im += part[e * nblk + b].im;
  }
  flux[e] = w > 0.0 ? cplx<double>{re / w, im / w} : cplx<double>{0.0, 0.0};
  if (e < (int64_t)nchan * npol) weight[e] = w;
}

// ---------------------------------------------------------------------------------
// Gain application: V' = M1 V conj(M2) per (time, baseline, channel), M the gain of the
// baseline's antennas (or its inverse) for the gain-table row of the visibility's time.
// ---------------------------------------------------------------------------------
__device__ __forceinline__ cplx<double> cmul(cplx<double> a, cplx<double> b) {
  return cplx<double>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cplx<double> cmulc(cplx<double> a, cplx<double> b) {  // a * conj(b)
  return cplx<double>{a.re * b.re + a.im * b.im, a.im * b.re - a.re * b.im};
}
__device__ __forceinline__ cplx<double> cadd(cplx<double> a, cplx<double> b) {
  return cplx<double>{a.re + b.re, a.im + b.im};
}
__device__ __forceinline__ cplx<double> cdiv(cplx<double> a, cplx<double> b) {
  // scaled division (Smith's method): no overflow for large |b|
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return cplx<double>{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return cplx<double>{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

// Per (gain row, antenna, channel): the matrix applied (nrec x nrec, row-major) and whether
// it exists. Forward: the gain. Inverse, nrec = 1: 1/g, or 0 where g = 0. Inverse, nrec = 2:
// the 2x2 inverse, or "no inverse" where the determinant is exactly zero.
__global__ void k_gain_prep(const cplx<double>* __restrict__ gain, int64_t nent, int nrec, int inverse,
                            cplx<double>* __restrict__ m, int* __restrict__ valid) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= nent) return;
  const int n2 = nrec * nrec;
  const cplx<double>* g = gain + e * n2;
  cplx<double>* o = m + e * n2;
  int ok = 1;
  if (!inverse) {
    for (int k = 0; k < n2; ++k) o[k] = g[k];
  } else if (nrec == 1) {
    const cplx<double> x = g[0];
    o[0] = (x.re != 0.0 || x.im != 0.0) ? cdiv(cplx<double>{1.0, 0.0}, x) : cplx<double>{0.0, 0.0};
  } else {
    const cplx<double> a = g[0], b = g[1], c = g[2], d = g[3];
    const cplx<double> ad = cmul(a, d), bc = cmul(b, c);
    const cplx<double> det{ad.re - bc.re, ad.im - bc.im};
    if (det.re == 0.0 && det.im == 0.0) {
      ok = 0;
      for (int k = 0; k < 4; ++k) o[k] = cplx<double>{0.0, 0.0};
    } else {
      o[0] = cdiv(d, det);
      o[1] = cdiv(cplx<double>{-b.re, -b.im}, det);
      o[2] = cdiv(cplx<double>{-c.re, -c.im}, det);
      o[3] = cdiv(a, det);
    }
  }
  valid[e] = ok;
}

struct GainArgs {
  int64_t ntimes;
  int nbl, nchan, npol, gnchan, nant, nrec, ngain;
};

// One thread per (time, baseline, channel) with a gain-table row; channels at or beyond the
// table's channel count get only the flag mask below. npol 1: v *= sum_k m1[k] conj(m2[k]) and the
// weight becomes the product of the antennas' gain weights (both zero where that sum is
// zero). npol 2: the diagonal of M1 diag(v) conj(M2). npol 4: M1 [[v0 v1] [v2 v3]] conj(M2).
// Without an inverse (inverse mode, singular 2x2 gain): npol 4 zeroes the visibility and its
// weights, npol 2 its first polarisation. time_flagged[t] != 0: the visibility and weight are
// multiplied by (1 - flags) first.
__global__ void k_apply_gains(cplx<double>* __restrict__ vis, double* __restrict__ wt,
                              const double* __restrict__ flags, const int8_t* __restrict__ time_flagged,
                              const int* __restrict__ ant1, const int* __restrict__ ant2,
                              const cplx<double>* __restrict__ m, const int* __restrict__ valid,
                              const double* __restrict__ gw, const int* __restrict__ time_to_gain,
                              GainArgs a) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.ntimes * a.nbl * a.nchan) return;
  const int64_t tb = i / a.nchan;
  const int ch = (int)(i - tb * a.nchan);
  const int64_t t = tb / a.nbl;
  const int b = (int)(tb - t * a.nbl);
  const int r = time_to_gain[t];
  const int a1 = ant1[b], a2 = ant2[b];
  // the host validates these; out-of-range entries are left unchanged rather than read past
  if (r < 0 || r >= a.ngain || a1 < 0 || a1 >= a.nant || a2 < 0 || a2 >= a.nant) return;
  const bool fl = time_flagged && time_flagged[t];
  if (ch >= a.gnchan && !fl) return;
  const int np = a.npol, n2 = a.nrec * a.nrec;
  const int64_t e1 = ((int64_t)r * a.nant + a1) * a.gnchan + ch;
  const int64_t e2 = ((int64_t)r * a.nant + a2) * a.gnchan + ch;
  const cplx<double>* m1 = m + e1 * n2;
  const cplx<double>* m2 = m + e2 * n2;
  // weights are read only where they can change (flag mask, one polarisation, no inverse)
  cplx<double> v[4];
  double w[4] = {0.0, 0.0, 0.0, 0.0};
  bool wdirty = false;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (p >= np) break;
    v[p] = vis[i * np + p];
    if (fl) {
      const double k = 1.0 - flags[i * np + p];
      v[p] = cplx<double>{v[p].re * k, v[p].im * k};
      w[p] = wt[i * np + p] * k;
    }
  }
  wdirty = fl;
  if (ch >= a.gnchan) {
    // beyond the table's channels: only the flag mask
  } else if (np == 1) {
    cplx<double> s{0.0, 0.0};
    for (int k = 0; k < n2; ++k) s = cadd(s, cmulc(m1[k], m2[k]));
    if (s.re != 0.0 || s.im != 0.0) {
      v[0] = cmul(v[0], s);
      w[0] = gw[e1 * n2] * gw[e2 * n2];
    } else {
      v[0] = cplx<double>{0.0, 0.0};
      w[0] = 0.0;
    }
    wdirty = true;
  } else if (!(valid[e1] && valid[e2])) {
    if (!fl) {
#pragma unroll
      for (int p = 0; p < 4; ++p)
        if (p < np) w[p] = wt[i * np + p];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p >= (np == 4 ? 4 : 1)) break;
      v[p] = cplx<double>{0.0, 0.0};
      w[p] = 0.0;
    }
    wdirty = true;
  } else if (np == 2) {
    const cplx<double> x0 = v[0], x1 = v[1];
    v[0] = cadd(cmulc(cmul(m1[0], x0), m2[0]), cmulc(cmul(m1[1], x1), m2[2]));
    v[1] = cadd(cmulc(cmul(m1[2], x0), m2[1]), cmulc(cmul(m1[3], x1), m2[3]));
  } else {
    // T = M1 V, then T conj(M2)
    cplx<double> tm[4];
    for (int ii = 0; ii < 2; ++ii)
      for (int l = 0; l < 2; ++l)
        tm[2 * ii + l] = cadd(cmul(m1[2 * ii], v[l]), cmul(m1[2 * ii + 1], v[2 + l]));
    for (int ii = 0; ii < 2; ++ii)
      for (int k = 0; k < 2; ++k)
        v[2 * ii + k] = cadd(cmulc(tm[2 * ii], m2[k]), cmulc(tm[2 * ii + 1], m2[2 + k]));
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (p >= np) break;
    vis[i * np + p] = v[p];
    if (wdirty) wt[i * np + p] = w[p];
  }
}

// ---------------------------------------------------------------------------------
// Visibility prologue / epilogue of invert_ng / predict_ng: flag mask, phase shift to the
// image centre, polarisation conversion and the gridder's layout and precision, in one pass.
// ---------------------------------------------------------------------------------
// ---------------------------------------------------------------------------------
// Visibility columns in the BlockVisibility layout [row][column channel][npv], read and written
// by the value gather (invert) and the unsort (predict) directly, with the passes around the
// gridder fused in: flag mask, imaging weights, phase shift, polarisation conversion,
// per-antenna scalar gains and the cast to / from the gridder's precision.
// ---------------------------------------------------------------------------------
struct ColArgs {
  int nchan_g, nchan_t, chan0;  // geometry channels = column channels chan0 .. chan0 + nchan_g - 1
  double rnchan_g;              // 1 / nchan_g, rounded (udiv32)
  int vis_f32, wgt_f32, flag_i8;
  int shift;                    // phase exp(+-2 pi i f/c (u L + v M + w N)) on the geometry's uvw
  double L, M, N;
  int nant;                     // gains [nant][nchan_t] complex128 with ant1 / ant2 [row], or none
  int psf;                      // invert: image polarisation 0 gets 1, the others 0 (weights kept)
  int vec;                      // bits: vis (1), weights (2), flags (4) aligned for one access per item
  double cre[16], cim[16];      // conversion [nout][nin]
};

__device__ __forceinline__ cplx<double> col_load(const void* p, int64_t i, int f32) {
  if (f32) {
    const float2 v = reinterpret_cast<const float2*>(p)[i];
    return cplx<double>{v.x, v.y};
  }
  const double2 v = reinterpret_cast<const double2*>(p)[i];
  return cplx<double>{v.x, v.y};
}
__device__ __forceinline__ void col_store(void* p, int64_t i, int f32, cplx<double> v) {
  if (f32) reinterpret_cast<float2*>(p)[i] = make_float2((float)v.re, (float)v.im);
  else reinterpret_cast<double2*>(p)[i] = make_double2(v.re, v.im);
}
__device__ __forceinline__ double col_real(const void* p, int64_t i, int f32) {
  return f32 ? (double)reinterpret_cast<const float*>(p)[i] : reinterpret_cast<const double*>(p)[i];
}
__device__ __forceinline__ double col_keep(const void* f, int64_t i, int i8) {
  if (!f) return 1.0;
  return i8 ? 1.0 - (double)reinterpret_cast<const int8_t*>(f)[i] : 1.0 - (double)reinterpret_cast<const int*>(f)[i];
}
// The NP polarisations of one item of a column (vis, weights, flags): one or two 16-byte
// accesses (or one 4-byte access for four int8 flags) when `vec` says the column is aligned for
// it. Items sit at random places in the columns, and the gathers are bound by accesses.
template <int NP>
__device__ __forceinline__ void col_load_item(const void* p, int64_t i, int f32, bool vec, cplx<double> (&v)[NP]) {
  if (vec && f32 && NP % 2 == 0) {
    const float4* q = reinterpret_cast<const float4*>(p) + i * (NP / 2);
#pragma unroll
    for (int h = 0; h < NP / 2; ++h) {
      const float4 t = q[h];
      v[2 * h] = cplx<double>{t.x, t.y};
      v[2 * h + 1] = cplx<double>{t.z, t.w};
    }
  } else {
#pragma unroll
    for (int k = 0; k < NP; ++k) v[k] = col_load(p, i * NP + k, f32);
  }
}
template <int NP>
__device__ __forceinline__ void col_store_item(void* p, int64_t i, int f32, bool vec, const cplx<double> (&v)[NP]) {
  if (vec && f32 && NP % 2 == 0) {
    float4* q = reinterpret_cast<float4*>(p) + i * (NP / 2);
#pragma unroll
    for (int h = 0; h < NP / 2; ++h)
      q[h] = make_float4((float)v[2 * h].re, (float)v[2 * h].im, (float)v[2 * h + 1].re, (float)v[2 * h + 1].im);
  } else {
#pragma unroll
    for (int k = 0; k < NP; ++k) col_store(p, i * NP + k, f32, v[k]);
  }
}
template <int NP>
__device__ __forceinline__ void col_real_item(const void* p, int64_t i, int f32, bool vec, double (&w)[NP]) {
  if (vec && f32 && NP == 4) {
    const float4 t = reinterpret_cast<const float4*>(p)[i];
    w[0] = t.x;
    w[1] = t.y;
    w[2] = t.z;
    w[3] = t.w;
  } else if (vec && f32 && NP == 2) {
    const float2 t = reinterpret_cast<const float2*>(p)[i];
    w[0] = t.x;
    w[1] = t.y;
  } else if (vec && !f32 && NP % 2 == 0) {
    const double2* q = reinterpret_cast<const double2*>(p) + i * (NP / 2);
#pragma unroll
    for (int h = 0; h < NP / 2; ++h) {
      const double2 t = q[h];
      w[2 * h] = t.x;
      w[2 * h + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int k = 0; k < NP; ++k) w[k] = col_real(p, i * NP + k, f32);
  }
}
template <int NP>
__device__ __forceinline__ void col_keep_item(const void* f, int64_t i, int i8, bool vec, double (&k)[NP]) {
  if (!f) {
#pragma unroll
    for (int p = 0; p < NP; ++p) k[p] = 1.0;
  } else if (vec && i8 && NP == 4) {
    const uint32_t t = reinterpret_cast<const uint32_t*>(f)[i];
#pragma unroll
    for (int p = 0; p < 4; ++p) k[p] = 1.0 - (double)(int8_t)(t >> (8 * p));
  } else if (vec && !i8 && NP == 4) {
    const int4 t = reinterpret_cast<const int4*>(f)[i];
    k[0] = 1.0 - (double)t.x;
    k[1] = 1.0 - (double)t.y;
    k[2] = 1.0 - (double)t.z;
    k[3] = 1.0 - (double)t.w;
  } else {
#pragma unroll
    for (int p = 0; p < NP; ++p) k[p] = col_keep(f, i * NP + p, i8);
  }
}

// the product of the row's antenna gains g1 conj(g2) (or its conjugate)
__device__ __forceinline__ cplx<double> row_gain(const cplx<double>* __restrict__ gains, const int* __restrict__ ant1,
                                                 const int* __restrict__ ant2, int64_t row, int ch, int nchan_t,
                                                 bool adjoint) {
  const cplx<double> g1 = gains[(int64_t)ant1[row] * nchan_t + ch], g2 = gains[(int64_t)ant2[row] * nchan_t + ch];
  const cplx<double> c{g1.re * g2.re + g1.im * g2.im, g1.im * g2.re - g1.re * g2.im};  // g1 conj(g2)
  return adjoint ? cplx<double>{c.re, -c.im} : c;
}

// invert: per sorted item, x_p = vis_p (1 - f_p) [* exp(+2 pi i ph)] [* conj(g1 conj(g2))],
// y_q = sum_p C[q][p] x_p, w_q = weight_q (1 - f_q); then the value gather of k_values:
// rec_v[j][q] = conj?(w_q y_q) exp(-2 pi i |w| nshift f/c). FP64 until the final store.
template <typename R, int NI, int NO>
__global__ void k_values_cols(const uint32_t* __restrict__ perm, int64_t n, const cplx<R>* __restrict__ rec_p,
                              const uint32_t* __restrict__ flipb, const double* __restrict__ uvw,
                              const double* __restrict__ fc, ColArgs a, const void* __restrict__ vis,
                              const void* __restrict__ wgt, const void* __restrict__ flags,
                              const cplx<double>* __restrict__ gains, const int* __restrict__ ant1,
                              const int* __restrict__ ant2, cplx<R>* __restrict__ rec_v) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  uint32_t chu;
  const int64_t row = udiv32(i, (uint32_t)a.nchan_g, a.rnchan_g, &chu);
  const int chl = (int)chu, ch = a.chan0 + chl;
  const int64_t ci = row * a.nchan_t + ch;
  double keep[NI], wv[NI];
  cplx<double> x[NI];
  col_keep_item<NI>(flags, ci, a.flag_i8, (a.vec & 4) != 0, keep);
  col_load_item<NI>(vis, ci, a.vis_f32, (a.vec & 1) != 0, x);
  if (wgt) col_real_item<NI>(wgt, ci, a.wgt_f32, (a.vec & 2) != 0, wv);
#pragma unroll
  for (int p = 0; p < NI; ++p) x[p] = cplx<double>{x[p].re * keep[p], x[p].im * keep[p]};
  cplx<double> m{1.0, 0.0};
  if (a.shift) {
    double sn, cs;
    const double ph = fc[chl] * (uvw[3 * row] * a.L + uvw[3 * row + 1] * a.M + uvw[3 * row + 2] * a.N);
    sincospi(2.0 * ph, &sn, &cs);
    m = cplx<double>{cs, sn};
  }
  if (gains) m = cmul(m, row_gain(gains, ant1, ant2, row, ch, a.nchan_t, true));
  const bool flip = (flipb[j >> 5] >> (j & 31)) & 1u;
  const cplx<R> pr = rec_p[j];
  const double c = (double)pr.re, s = -(double)pr.im;
  cplx<R> o[NO];
#pragma unroll
  for (int q = 0; q < NO; ++q) {
    cplx<double> y{0.0, 0.0};
#pragma unroll
    for (int p = 0; p < NI; ++p) y = cadd(y, cmul(cplx<double>{a.cre[q * NI + p], a.cim[q * NI + p]}, x[p]));
    y = cmul(y, m);
    if (a.psf) y = cplx<double>{q == 0 ? 1.0 : 0.0, 0.0};
    if (flip) y.im = -y.im;
    const double w = (wgt ? wv[q] : 1.0) * keep[q];
    y.re *= w;
    y.im *= w;
    o[q] = cplx<R>{(R)(y.re * c - y.im * s), (R)(y.re * s + y.im * c)};
  }
  store_vals<R, NO, false>(rec_v + j * NO, o);
}

// predict: the unsort of k_unsort into the columns: o_q = conj?(acc_q exp(2 pi i |w| nshift f/c)),
// y_p = sum_q C[p][q] o_q [* exp(-2 pi i ph)]; model (nullable) = y; vis = y [* g1 conj(g2)]
template <typename R, int NI, int NO>
__global__ void k_unsort_cols(const uint32_t* __restrict__ perm, int64_t n, const cplx<R>* __restrict__ rec_p,
                              const uint32_t* __restrict__ flipb, const cplx<R>* __restrict__ accs,
                              const double* __restrict__ uvw, const double* __restrict__ fc, ColArgs a,
                              const cplx<double>* __restrict__ gains, const int* __restrict__ ant1,
                              const int* __restrict__ ant2, void* __restrict__ vis, void* __restrict__ model) {
  const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (j >= n) return;
  const uint32_t i = perm[j];
  uint32_t chu;
  const int64_t row = udiv32(i, (uint32_t)a.nchan_g, a.rnchan_g, &chu);
  const int chl = (int)chu, ch = a.chan0 + chl;
  const int64_t ci = row * a.nchan_t + ch;
  const bool flip = (flipb[j >> 5] >> (j & 31)) & 1u;
  const cplx<R> pr = rec_p[j];
  const double c = (double)pr.re, s = (double)pr.im;
  cplx<R> av[NI];
  load_vals<R, NI, false>(accs + j * NI, av);
  cplx<double> o[4];
#pragma unroll
  for (int q = 0; q < NI; ++q) {
    const double re = (double)av[q].re, im = (double)av[q].im;
    o[q] = cplx<double>{re * c - im * s, flip ? -(re * s + im * c) : re * s + im * c};
  }
  cplx<double> m{1.0, 0.0};
  if (a.shift) {
    double sn, cs;
    const double ph = fc[chl] * (uvw[3 * row] * a.L + uvw[3 * row + 1] * a.M + uvw[3 * row + 2] * a.N);
    sincospi(-2.0 * ph, &sn, &cs);
    m = cplx<double>{cs, sn};
  }
  const cplx<double> g = gains ? row_gain(gains, ant1, ant2, row, ch, a.nchan_t, false) : cplx<double>{1.0, 0.0};
  cplx<double> y[NO], yg[NO];
#pragma unroll
  for (int p = 0; p < NO; ++p) {
    y[p] = cplx<double>{0.0, 0.0};
#pragma unroll
    for (int q = 0; q < NI; ++q) y[p] = cadd(y[p], cmul(cplx<double>{a.cre[p * NI + q], a.cim[p * NI + q]}, o[q]));
    y[p] = cmul(y[p], m);
    yg[p] = gains ? cmul(y[p], g) : y[p];
  }
  if (model) col_store_item<NO>(model, ci, a.vis_f32, (a.vec & 8) != 0, y);
  col_store_item<NO>(vis, ci, a.vis_f32, (a.vec & 1) != 0, yg);
}

// Sums over rows of weight (1 - flags) per (channel, polarisation) of columns [nrow][nchan][npol]:
// per-block partials in a fixed order, then the blocks' partials added in block order.
constexpr int kSumBlocks = 512;
__global__ __launch_bounds__(kBlock) void k_wsum_partial(const void* __restrict__ wgt, const void* __restrict__ flags,
                                                         int64_t nrow, int nchan, int npol, int wgt_f32, int flag_i8,
                                                         double* __restrict__ part) {
  const int ch = (int)blockIdx.y;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x; r < nrow; r += (int64_t)gridDim.x * kBlock) {
    const int64_t b = (r * nchan + ch) * npol;
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < npol) acc[p] += (wgt ? col_real(wgt, b + p, wgt_f32) : 1.0) * col_keep(flags, b + p, flag_i8);
  }
  __shared__ double red[4][kBlock];
#pragma unroll
  for (int p = 0; p < 4; ++p) red[p][threadIdx.x] = acc[p];
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o)
      for (int p = 0; p < 4; ++p) red[p][threadIdx.x] += red[p][threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x < npol) part[((int64_t)ch * npol + threadIdx.x) * gridDim.x + blockIdx.x] = red[threadIdx.x][0];
}
__global__ void k_wsum_final(const double* __restrict__ part, int64_t ncp, int nblk, double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= ncp) return;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += part[e * nblk + b];
  out[e] = s;
}

// Gradient of a real loss through V = g1 conj(g2) M (per-antenna scalar gains, all
// polarisations): with G the loss gradient with respect to V (torch's convention),
// dL/dg[a1] += sum_p G_p conj(M_p) g[a2] and dL/dg[a2] += conj(sum_p G_p conj(M_p)) g[a1].
// One thread per (row, channel); FP64 atomics into grad [nant][nchan].
__global__ void k_gain_grad(const void* __restrict__ model, const void* __restrict__ grad, int vis_f32,
                            const cplx<double>* __restrict__ gains, const int* __restrict__ ant1,
                            const int* __restrict__ ant2, int64_t nrow, int nchan, int npol,
                            double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= nrow * nchan) return;
  const int64_t row = i / nchan;
  const int ch = (int)(i - row * nchan);
  cplx<double> s{0.0, 0.0};
  for (int p = 0; p < npol; ++p) {
    const cplx<double> gv = col_load(grad, i * npol + p, vis_f32), mv = col_load(model, i * npol + p, vis_f32);
    s = cadd(s, cmulc(gv, mv));
  }
  const int a1 = ant1[row], a2 = ant2[row];
  const cplx<double> g1 = gains[(int64_t)a1 * nchan + ch], g2 = gains[(int64_t)a2 * nchan + ch];
  const cplx<double> d1 = cmul(s, g2), d2 = cmul(cplx<double>{s.re, -s.im}, g1);
  double* o1 = out + 2 * ((int64_t)a1 * nchan + ch);
  double* o2 = out + 2 * ((int64_t)a2 * nchan + ch);
  atomicAdd(o1, d1.re);
  atomicAdd(o1 + 1, d1.im);
  atomicAdd(o2, d2.re);
  atomicAdd(o2 + 1, d2.im);
}

struct PolArgs {
  int64_t nrow;
  int nchan, nin, nout, shift;
  double l, m, nm1;             // direction cosines of the image centre seen from the vis centre
  double cre[16], cim[16];      // conversion [nout][nin], row-major
};

__device__ __forceinline__ void to_c(cplx<double> v, cplx<float>* o) { *o = cplx<float>{(float)v.re, (float)v.im}; }
__device__ __forceinline__ void to_c(cplx<double> v, cplx<double>* o) { *o = v; }
__device__ __forceinline__ cplx<double> from_c(cplx<float> v) { return cplx<double>{v.re, v.im}; }
__device__ __forceinline__ cplx<double> from_c(cplx<double> v) { return v; }

// ph = f/c (u l + v m + w (n-1)) of item (row, channel)
__device__ __forceinline__ double shift_phase(const double* __restrict__ uvw, int64_t row, double fc,
                                              const PolArgs& a) {
  return (uvw[3 * row] * fc) * a.l + (uvw[3 * row + 1] * fc) * a.m + (uvw[3 * row + 2] * fc) * a.nm1;
}

// invert: x_p = vis_p (1 - flags_p) e^{+2 pi i ph}; ms_q = sum_p C[q][p] x_p; wgt_q =
// imaging_weight_q (1 - flags_q); nonzero[chan][q] = 1 where some ms_q of the channel is not 0
template <typename CT, typename RT, int NI, int NO>
__global__ void k_vis_prologue(const cplx<double>* __restrict__ vis, const int* __restrict__ flags,
                               const double* __restrict__ iwt, const double* __restrict__ uvw,
                               const double* __restrict__ fc, PolArgs a, CT* __restrict__ ms,
                               RT* __restrict__ wgt, int* __restrict__ nonzero) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nrow * a.nchan) return;
  const int64_t row = i / a.nchan;
  const int ch = (int)(i - row * a.nchan);
  constexpr int ni = NI, no = NO;
  double k[4] = {1.0, 1.0, 1.0, 1.0};
  cplx<double> x[4];
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (p >= ni) break;
    if (flags) k[p] = 1.0 - (double)flags[i * ni + p];
    const cplx<double> v = vis[i * ni + p];
    x[p] = cplx<double>{v.re * k[p], v.im * k[p]};
  }
  if (a.shift) {
    double sn, cs;
    sincospi(2.0 * shift_phase(uvw, row, fc[ch], a), &sn, &cs);
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ni) x[p] = cmul(x[p], cplx<double>{cs, sn});
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (q >= no) break;
    cplx<double> y{0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 4; ++p)
      if (p < ni) y = cadd(y, cmul(cplx<double>{a.cre[q * ni + p], a.cim[q * ni + p]}, x[p]));
    to_c(y, &ms[i * no + q]);
    int* nz = &nonzero[ch * no + q];
    if ((y.re != 0.0 || y.im != 0.0) && *nz == 0) *nz = 1;
    wgt[i * no + q] = (RT)(iwt[i * ni + q] * k[q]);
  }
}

// predict: y_p = sum_q C[p][q] vist_q, times e^{-2 pi i ph} when shifting
template <typename CT, int NI, int NO>
__global__ void k_vis_epilogue(const CT* __restrict__ vist, const double* __restrict__ uvw,
                               const double* __restrict__ fc, PolArgs a, cplx<double>* __restrict__ vis) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= a.nrow * a.nchan) return;
  const int64_t row = i / a.nchan;
  const int ch = (int)(i - row * a.nchan);
  constexpr int ni = NI, no = NO;
  cplx<double> x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    if (q < ni) x[q] = from_c(vist[i * ni + q]);
  cplx<double> ph{1.0, 0.0};
  if (a.shift) {
    double sn, cs;
    sincospi(-2.0 * shift_phase(uvw, row, fc[ch], a), &sn, &cs);
    ph = cplx<double>{cs, sn};
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    if (p >= no) break;
    cplx<double> y{0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (q < ni) y = cadd(y, cmul(cplx<double>{a.cre[p * ni + q], a.cim[p * ni + q]}, x[q]));
    vis[i * no + p] = a.shift ? cmul(y, ph) : y;
  }
}

// ---------------------------------------------------------------------------------
// Optional per-kernel timing with HIP events on the launch stream
// ---------------------------------------------------------------------------------
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
};
bool g_prof = false;
std::vector<ProfRec> g_pending;
std::vector<hipEvent_t> g_evpool;
std::vector<std::pair<std::string, std::pair<double, int64_t>>> g_profacc;

hipEvent_t ev_get() {
  if (!g_evpool.empty()) {
    hipEvent_t e = g_evpool.back();
    g_evpool.pop_back();
    return e;
  }
  hipEvent_t e;
  HIP_TRY(hipEventCreate(&e));
  return e;
}

struct ProfScope {
  const char* name;
  hipStream_t st;
  hipEvent_t a = nullptr;
  ProfScope(const char* n, hipStream_t s) : name(n), st(s) {
    if (g_prof) {
      a = ev_get();
      HIP_TRY(hipEventRecord(a, st));
    }
  }
  ~ProfScope() {
    if (a) {
      hipEvent_t b = ev_get();
      if (hipEventRecord(b, st) == hipSuccess) g_pending.push_back({name, a, b});
    }
  }
};

void prof_flush() {
  for (auto& r : g_pending) {
    float ms = 0.f;
    HIP_TRY(hipEventSynchronize(r.b));
    HIP_TRY(hipEventElapsedTime(&ms, r.a, r.b));
    bool found = false;
    for (auto& e : g_profacc)
      if (e.first == r.name) {
        e.second.first += ms;
        e.second.second += 1;
        found = true;
      }
    if (!found) g_profacc.push_back({r.name, {(double)ms, 1}});
    g_evpool.push_back(r.a);
    g_evpool.push_back(r.b);
  }
  g_pending.clear();
}

#define PROF(name) ProfScope _prof_scope_##__LINE__(name, st)

// ---------------------------------------------------------------------------------
// Host side: workspace cache, FFT plans, orchestration
// ---------------------------------------------------------------------------------
enum Slot {
  S_KEYS0, S_VALS0, S_KEYS1, S_VALS1, S_HIST, S_PART, S_BSTART, S_BEND, S_RECA, S_RECF, S_RECV,
  S_PLANE, S_ACCIMG, S_ACCS, S_SMALL, S_EXT, S_FFTWORK, S_TASK, S_TASKPART, S_TASKTAB, S_CNT, S_TBUF, S_TBUF2, S_DTASK,
  S_DPLANE, S_RECP, S_FLIPB, S_SREC, S_SBSTART, S_SBEND, S_RUNT, S_RUND, S_GAIN, S_PLANE2, S_LASTW, S_IDFT, S_SPLIT, S_FSTW, S_NSLOTS
};

struct FftKey {
  int gx, gy, nb, prec, fwd;
  bool operator<(const FftKey& o) const {
    return std::tie(gx, gy, nb, prec, fwd) < std::tie(o.gx, o.gy, o.nb, o.prec, o.fwd);
  }
};

struct DevCtx {
  void* buf[S_NSLOTS] = {};
  size_t cap[S_NSLOTS] = {};
  std::map<FftKey, rocfft_plan> plans;
  // out-of-place plans checked for leaving their input untouched (value: it does)
  std::map<rocfft_plan, bool> keeps_input;
  // the last stream that enqueued work on the shared workspace, and an event after that work
  hipEvent_t ws_ev = nullptr;
  hipStream_t ws_st = nullptr;
  bool ws_set = false;
};

std::mutex g_mu;
std::map<int, DevCtx> g_ctx;
bool g_fft_ready = false;

DevCtx& ctx() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  return g_ctx[dev];
}

void* ws(DevCtx& c, Slot s, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (c.cap[s] < bytes) {
    if (c.buf[s]) HIP_TRY(hipFree(c.buf[s]));
    c.buf[s] = nullptr;
    c.cap[s] = 0;
    size_t nb = bytes + bytes / 8;  // headroom against repeated regrowth
    if (hipMalloc(&c.buf[s], nb) != hipSuccess) {
      (void)hipGetLastError();
      HIPNG_THROW(HIPNG_ERUNTIME, "device allocation of %zu bytes failed", nb);
    }
    c.cap[s] = nb;
  }
  return c.buf[s];
}

// batched in-place 1-D transforms of `len` contiguous points, batches `len` apart
// (out of place: from one contiguous batch to another)
rocfft_plan fft_plan1(DevCtx& c, int len, int batch, int prec, int fwd, bool inplace = true) {
  if (!g_fft_ready) {
    FFT_TRY(rocfft_setup());
    g_fft_ready = true;
  }
  FftKey key{len, batch, inplace ? 1 : 2, prec == HIPNG_FP64 ? 0 : 1, fwd};
  auto it = c.plans.find(key);
  if (it != c.plans.end()) return it->second;
  rocfft_plan plan = nullptr;
  const size_t lengths[1] = {(size_t)len};
  FFT_TRY(rocfft_plan_create(&plan, inplace ? rocfft_placement_inplace : rocfft_placement_notinplace,
                             fwd ? rocfft_transform_type_complex_forward
                                 : rocfft_transform_type_complex_inverse,
                             prec == HIPNG_FP64 ? rocfft_precision_double : rocfft_precision_single,
                             1, lengths, (size_t)batch, nullptr));
  c.plans[key] = plan;
  return plan;
}

// One-shot calls borrow the per-device workspace and return before their kernels finish, so a
// call on another stream first waits for the work the previous call enqueued on it; the
// scope then records where its own work ends (also when it throws).
struct WsUse {
  DevCtx& c;
  hipStream_t st;
  WsUse(DevCtx& c_, hipStream_t s) : c(c_), st(s) {
    if (c.ws_set && c.ws_st != st) HIP_TRY(hipStreamWaitEvent(st, c.ws_ev, 0));
  }
  ~WsUse() {
    if (!c.ws_ev && hipEventCreateWithFlags(&c.ws_ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      c.ws_ev = nullptr;
      return;
    }
    if (hipEventRecord(c.ws_ev, st) == hipSuccess) {
      c.ws_set = true;
      c.ws_st = st;
    } else {
      (void)hipGetLastError();
    }
  }
};

unsigned grid1(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
unsigned gridcap(int64_t n) { return (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 1 << 20); }
// zero columns [c0, c1) of nrows rows of length rowlen
template <typename T>
void zero_cols(T* a, int64_t nrows, int rowlen, int c0, int c1, hipStream_t st) {
  if (c1 <= c0 || nrows <= 0) return;
  k_zero_cols<T><<<gridcap(nrows * (c1 - c0)), kBlock, 0, st>>>(a, nrows, rowlen, c0, c1);
}

void extent(DevCtx& c, const double* uvw, int64_t nrow, const double* freq, int nchan, double out[4],
            hipStream_t st) {
  double fmin = INFINITY, fmax = 0.0;
  for (int i = 0; i < nchan; ++i) {
    if (!(freq[i] > 0)) HIPNG_THROW(HIPNG_EINVAL, "frequencies must be positive");
    fmin = std::min(fmin, freq[i]);
    fmax = std::max(fmax, freq[i]);
  }
  if (nrow == 0) {
    out[0] = out[1] = out[2] = out[3] = 0.0;
    return;
  }
  unsigned long long h[4] = {0ull, 0ull, (unsigned long long)0x7ff0000000000000ull, 0ull};
  auto* d = (unsigned long long*)ws(c, S_EXT, sizeof(h));
  HIP_TRY(hipMemcpyAsync(d, h, sizeof(h), hipMemcpyHostToDevice, st));
  k_extent<<<std::min<int64_t>(grid1(nrow), 2048), kBlock, 0, st>>>(uvw, nrow, d);
  LAUNCH_CHECK();
  HIP_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  double e[4];
  std::memcpy(e, h, sizeof(e));
  if (!std::isfinite(e[0]) || !std::isfinite(e[1]) || !std::isfinite(e[3]))
    HIPNG_THROW(HIPNG_EINVAL, "uvw contains non-finite values");
  out[0] = e[0] * fmax / kSpeedOfLight;
  out[1] = e[1] * fmax / kSpeedOfLight;
  out[2] = e[2] * fmin / kSpeedOfLight;
  out[3] = e[3] * fmax / kSpeedOfLight;
}

void channel_scales(const hipng_plan& P, const double* freq, int nchan, double* sx, double* sy,
                    double* sw, double* pn) {
  for (int c = 0; c < nchan; ++c) {
    const double fc = freq[c] / kSpeedOfLight;
    sx[c] = fc * P.pixsize_x * (double)P.gx;
    sy[c] = fc * P.pixsize_y * (double)P.gy;
    sw[c] = P.wstacking ? fc / P.dw : 0.0;
    if (pn) pn[c] = fc * P.nshift;
  }
}

// Degree-kTapDeg polynomials of the interior kernel taps (TapPoly): Chebyshev interpolation
// of tap i at kTapDeg + 1 Chebyshev points of s in [-1, 1] in FP64, converted to monomials in
// s and rounded to FP32, for every tap. on = 1 when the fit's largest deviation from the
// kernel, sampled over every interior tap's interval, is at the FP32 evaluation level
// (1.2e-7) or below 5 % of the kernel's own 1-D error at this oversampling and support; on_all
// when the two edge taps' deviation is also below 2e-7 or 10 % of that error.
TapPoly make_tap_poly(int W, double beta, double sigma) {
  TapPoly tp{};
  if (W < 4 || W > kWMax) return tp;
  double err1d = 0.0;
  for (int is = 0; is < kNSig; ++is)
    if (std::fabs(kSigmas[is] - sigma) < 1e-9) err1d = kErr1D[is][W - kWMin];
  auto psi = [&](double z) {
    const double t = 1.0 - z * z;
    return t > 0.0 ? std::exp(beta * (std::sqrt(t) - 1.0)) : 0.0;
  };
  auto zof = [&](int i, double sv) { return ((sv + 1.0) * 0.5 - 0.5 * W + i) * 2.0 / W; };
  constexpr int N = kTapDeg + 1;
  double worst = 0.0, worst_edge = 0.0;
  for (int i = 0; i < W; ++i) {
    const bool edge = i == 0 || i == W - 1;
    double y[N], a[N], mono[N] = {0}, tm[N] = {0}, t0[N] = {0}, t1[N] = {0};
    for (int q = 0; q < N; ++q) y[q] = psi(zof(i, std::cos(kPi * (q + 0.5) / N)));
    for (int j = 0; j < N; ++j) {
      double acc = 0.0;
      for (int q = 0; q < N; ++q) acc += y[q] * std::cos(kPi * j * (q + 0.5) / N);
      a[j] = (j ? 2.0 : 1.0) * acc / N;
    }
    // sum_j a_j T_j(s) in monomials: T_0 = 1, T_1 = s, T_(j+1) = 2 s T_j - T_(j-1)
    t0[0] = 1.0;
    t1[1] = 1.0;
    mono[0] = a[0];
    for (int d = 0; d < N; ++d) mono[d] += a[1] * t1[d];
    for (int j = 2; j < N; ++j) {
      for (int d = 0; d < N; ++d) tm[d] = (d ? 2.0 * t1[d - 1] : 0.0) - t0[d];
      for (int d = 0; d < N; ++d) {
        mono[d] += a[j] * tm[d];
        t0[d] = t1[d];
        t1[d] = tm[d];
      }
    }
    for (int d = 0; d < N; ++d) tp.c[i][d] = (float)mono[d];
    for (int m = 0; m <= 512; ++m) {
      const double sv = -1.0 + m / 256.0;
      double pv = tp.c[i][kTapDeg];
      for (int d = kTapDeg - 1; d >= 0; --d) pv = pv * sv + tp.c[i][d];
      const double e = std::fabs(pv - psi(zof(i, sv)));
      if (edge) worst_edge = std::max(worst_edge, e);
      else worst = std::max(worst, e);
    }
  }
  tp.on = worst <= std::max(1.2e-7, 0.05 * err1d) ? 1 : 0;
  // the edge taps: an absolute error of 2e-7 (the FP32 level of the kernel's peak) or 10 % of
  // the kernel's own 1-D error
  tp.on_all = tp.on && worst_edge <= std::max(2e-7, 0.1 * err1d) ? 1 : 0;
  return tp;
}

KP make_kp(const hipng_plan& P, int nb, int nchan) {
  KP k;
  k.gx = P.gx;
  k.gy = P.gy;
  k.hgx = P.gx / 2;
  k.hgy = P.gy / 2;
  k.T = P.tile;
  k.ntx = P.ntx;
  k.nty = P.nty;
  k.ntiles = (uint32_t)P.ntx * (uint32_t)P.nty;
  k.W = P.W;
  k.Ww = P.wstacking ? P.W : 1;
  k.nplanes = P.nplanes;
  k.wstack = P.wstacking;
  k.nx = P.nx;
  k.ny = P.ny;
  k.nb = nb;
  k.nchan = nchan;
  k.rnchan = 1.0 / (double)std::max(nchan, 1);
  k.tshift = 0;
  while ((1 << k.tshift) < P.tile) ++k.tshift;
  k.hgxd = 0.5 * P.gx;
  k.hgyd = 0.5 * P.gy;
  k.hw = 0.5 * P.W;
  k.mw0 = P.wstacking ? -(P.w0 / P.dw) : 0.0;
  k.beta = P.beta;
  k.px = P.pixsize_x;
  k.py = P.pixsize_y;
  k.nshift = P.nshift;
  k.dw = P.dw;
  k.w0 = P.w0;
  k.amask = 0xffffffffu;
  k.dflag = 0u;
  return k;
}

// Buffer provider: one-shot calls borrow the per-device workspace; a persistent geometry owns
// its buffers (allocated once at creation, so its executions allocate nothing).
struct Buffers {
  DevCtx* borrow = nullptr;
  std::vector<void*> owned;
  void* get(Slot s, size_t bytes) {
    if (borrow) return ws(*borrow, s, bytes);
    if (bytes == 0) bytes = 256;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      HIPNG_THROW(HIPNG_ERUNTIME, "device allocation of %zu bytes failed", bytes);
    }
    owned.push_back(p);
    return p;
  }
  void release() {
    for (void* p : owned) (void)hipFree(p);
    owned.clear();
  }
};

// Small device tables: sx, sy, sw, pn [nchan]; phix [nx], phiy [ny]; quadrature z, c.
struct Tables {
  double *sx, *sy, *sw, *pn, *phix, *phiy, *qz, *qc, *fc;  // fc[c] = f_c / c0
  int nq;
  // Phi_w of the grid correction as corr_at takes it: a Chebyshev series (cn < 0) or the
  // quadrature (cz = qz, cc = qc, cn = nq)
  double *cz, *cc;
  int cn;
};

// Chebyshev coefficients of log Phi_w over the range of dw nm1s an image covers: degree 32 at
// the first-kind nodes, kept when exp of the series matches the quadrature sum to 1e-13 of its
// value at each of 97 points of the range (the correction divides by Phi_w, so the relative
// error is what counts; one evaluation then costs 32 FMAs and an exp instead of nq cosines)
constexpr int kWCheb = 33;
bool w_cheb(const hipng_plan& P, const KernelFT& ft, double coef[kWCheb], double* centre, double* inv_half) {
  if (!P.wstacking || !(P.dw > 0.0)) return false;
  const double x = -(double)(P.nx / 2) * P.pixsize_x, y = -(double)(P.ny / 2) * P.pixsize_y;
  const double r2 = x * x + y * y;
  if (!(r2 < 1.0)) return false;
  const double nm1c = -r2 / (std::sqrt(1.0 - r2) + 1.0);
  double a = P.dw * (nm1c - P.nshift), b = P.dw * (-P.nshift);
  if (a > b) std::swap(a, b);
  const double pad = 1e-6 * (b - a);
  a -= pad;
  b += pad;
  if (!(b > a)) return false;
  auto f = [&](double xi) {
    double sum = 0.0;
    for (size_t q = 0; q < ft.z.size(); ++q) sum += ft.c[q] * std::cos(M_PI * P.W * ft.z[q] * xi);
    return sum;
  };
  double fv[kWCheb];
  for (int j = 0; j < kWCheb; ++j) {
    const double v = f(0.5 * (a + b) + 0.5 * (b - a) * std::cos(M_PI * (j + 0.5) / kWCheb));
    if (!(v > 0.0)) return false;
    fv[j] = std::log(v);
  }
  for (int m = 0; m < kWCheb; ++m) {
    double sum = 0.0;
    for (int j = 0; j < kWCheb; ++j) sum += fv[j] * std::cos(M_PI * m * (j + 0.5) / kWCheb);
    coef[m] = (m == 0 ? 1.0 : 2.0) * sum / kWCheb;
  }
  *centre = 0.5 * (a + b);
  *inv_half = 2.0 / (b - a);
  for (int i = 0; i < 97; ++i) {
    const double xi = a + (b - a) * (i + 0.37) / 97.0, ex = f(xi);
    if (!(ex > 0.0) || !(std::fabs(std::exp(cheb_eval(coef, kWCheb, (xi - *centre) * *inv_half)) - ex) <= 1e-13 * ex))
      return false;
  }
  return true;
}

Tables upload_tables(Buffers& B, const hipng_plan& P, const double* freq, int nchan, hipStream_t st) {
  KernelFT ft;
  ft.init(P.W, P.beta);
  const int nq = (int)ft.z.size();
  const size_t n = 5 * (size_t)nchan + P.nx + P.ny + 2 * (size_t)nq + kWCheb + 2;
  std::vector<double> host(n, 0.0);
  double* h = host.data();
  channel_scales(P, freq, nchan, h, h + nchan, h + 2 * nchan, h + 3 * nchan);
  // phix, phiy are filled on the device below
  double* hz = h + 4 * nchan + P.nx + P.ny;
  double* hc = hz + nq;
  for (int q = 0; q < nq; ++q) {
    hz[q] = ft.z[q];
    hc[q] = ft.c[q];
  }
  double* hf = hc + nq;
  for (int c = 0; c < nchan; ++c) hf[c] = freq[c] / kSpeedOfLight;
  double* hw = hf + nchan;  // Chebyshev coefficients of Phi_w, then centre and 1 / half-width
  const bool cheb = w_cheb(P, ft, hw, hw + kWCheb, hw + kWCheb + 1);
  auto* d = (double*)B.get(S_SMALL, n * sizeof(double));
  HIP_TRY(hipMemcpyAsync(d, h, n * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));  // host buffer is released on return
  Tables t;
  t.sx = d;
  t.sy = d + nchan;
  t.sw = d + 2 * nchan;
  t.pn = d + 3 * nchan;
  t.phix = d + 4 * nchan;
  t.phiy = t.phix + P.nx;
  t.qz = t.phiy + P.ny;
  t.qc = t.qz + nq;
  t.fc = t.qc + nq;
  t.nq = nq;
  t.cz = cheb ? t.fc + nchan + kWCheb : t.qz;
  t.cc = cheb ? t.fc + nchan : t.qc;
  t.cn = cheb ? -kWCheb : nq;
  const int nphi = P.nx + P.ny;
  k_phi_tables<<<(nphi + kBlock - 1) / kBlock, kBlock, 0, st>>>(t.phix, P.nx, P.gx, t.phiy, P.ny, P.gy, P.W, t.qz,
                                                               t.qc, nq);
  LAUNCH_CHECK();
  return t;
}

// in-place exclusive scan of a[0, m); part holds ceil(m / kRsItems) entries. Arrays of up to
// kScanOneBlock entries are scanned by one block in one launch (k_scan_top scans any array in
// place): small jobs run dozens of these scans per call, and three launches cost more there.
constexpr int64_t kScanOneBlock = 4 * (int64_t)kRsItems;
void excl_scan(uint32_t* a, int64_t m, uint32_t* part, hipStream_t st) {
  if (m <= kScanOneBlock) {
    k_scan_top<<<1, kBlock, 0, st>>>(a, m);
    LAUNCH_CHECK();
    return;
  }
  const int64_t np = (m + kRsItems - 1) / kRsItems;
  k_scan_partial<<<(unsigned)np, kBlock, 0, st>>>(a, m, part);
  LAUNCH_CHECK();
  k_scan_top<<<1, kBlock, 0, st>>>(part, np);
  LAUNCH_CHECK();
  k_scan_apply<<<(unsigned)np, kBlock, 0, st>>>(a, m, part);
  LAUNCH_CHECK();
}

// keys + stable radix sort in the shared workspace; returns sorted keys and permutation
// The bucket keys are recomputed from uvw in digit order (k_rekey): uvw rows are shared by all
// channels of a row and stay cache resident, which makes this cheaper than keeping the keys
// of every item and gathering them (measured at the production shape).
void sort_items(DevCtx& c, const double* uvw, int64_t n, const Tables& tb, const KP& k,
                const hipng_plan& P, bool allow_pack, hipStream_t st, uint32_t** keys_out,
                uint32_t** perm_out, uint32_t** free0 = nullptr, uint32_t** free1 = nullptr) {
  // one spare entry each: the ping-pong pair left free after the sort is reused as scratch
  // of n + 1 entries (run starts of the dense-run detection)
  auto* k0 = (uint32_t*)ws(c, S_KEYS0, (n + 1) * 4);
  auto* v0 = (uint32_t*)ws(c, S_VALS0, (n + 1) * 4);
  auto* k1 = (uint32_t*)ws(c, S_KEYS1, (n + 1) * 4);
  auto* v1 = (uint32_t*)ws(c, S_VALS1, (n + 1) * 4);
  const int64_t nblocks = (n + kRsItems - 1) / kRsItems;
  const int64_t mmax = 512 * nblocks;  // histogram entries of a 9-bit pass
  auto* hist = (uint32_t*)ws(c, S_HIST, mmax * 4);
  auto* part = (uint32_t*)ws(c, S_PART, ((mmax + kRsItems - 1) / kRsItems) * 4);
  // the digit passes carry the bucket key packed into the sort words when key, digit and
  // index bits fit in 64; the first key pass unpacks it (no second pass over uvw)
  const int dbits = 2 * bits_for((uint64_t)P.tile - 1);
  const int ibits = std::max(1, bits_for((uint64_t)std::max<int64_t>(n, 1) - 1));
  const bool pack = allow_pack && dbits < 32 && ibits < 32 && P.keybits + dbits + ibits <= 64;
  // `bits` bits from bit 0 in 8-bit passes, or in 9-bit passes where that saves a pass (a
  // 9-bit pass while the remaining bits need more than 8 per remaining pass)
  auto passes = [&](int bits, int udb) {
    PROF("radix_sort");
    const int np = std::min((bits + 7) / 8, (bits + 8) / 9);
    int shift = 0;
    for (int ps = 0; ps < np; ++ps) {
      const int wbits = bits - shift > 8 * (np - ps) ? 9 : 8;
      const int db = ps == 0 ? udb : 0;
      const int64_t m = ((int64_t)1 << wbits) * nblocks;
      if (wbits == 9)
        k_rs_hist<9><<<rs_grid(nblocks), kBlock, 0, st>>>(k0, v0, db, ibits, n, shift, hist, nblocks);
      else
        k_rs_hist<8><<<rs_grid(nblocks), kBlock, 0, st>>>(k0, v0, db, ibits, n, shift, hist, nblocks);
      LAUNCH_CHECK();
      excl_scan(hist, m, part, st);
      if (wbits == 9)
        k_rs_scatter<9><<<rs_grid(nblocks), kBlock, 0, st>>>(k0, v0, k1, v1, n, shift, hist, nblocks, db, ibits);
      else
        k_rs_scatter<8><<<rs_grid(nblocks), kBlock, 0, st>>>(k0, v0, k1, v1, n, shift, hist, nblocks, db, ibits);
      LAUNCH_CHECK();
      std::swap(k0, k1);
      std::swap(v0, v1);
      shift += wbits;
    }
  };
  // LSD: the in-tile origin digit first, then the bucket key of each item in that order
  {
    PROF("keys");
    const int db = pack ? dbits : 0;
    if (P.precision == HIPNG_FP32)
      k_digits<true><<<grid1(n), kBlock, 0, st>>>(uvw, n, tb.sx, tb.sy, tb.sw, k, k0, v0, db, ibits);
    else
      k_digits<false><<<grid1(n), kBlock, 0, st>>>(uvw, n, tb.sx, tb.sy, tb.sw, k, k0, v0, db, ibits);
    LAUNCH_CHECK();
  }
  passes(dbits, 0);
  if (!pack) {
    PROF("keys");
    if (P.precision == HIPNG_FP32)
      k_rekey<true><<<grid1(n), kBlock, 0, st>>>(uvw, v0, n, tb.sx, tb.sy, tb.sw, k, k0);
    else
      k_rekey<false><<<grid1(n), kBlock, 0, st>>>(uvw, v0, n, tb.sx, tb.sy, tb.sw, k, k0);
    LAUNCH_CHECK();
  }
  passes(std::max(P.keybits, 1), pack ? dbits : 0);
  *keys_out = k0;
  *perm_out = v0;
  if (free0) *free0 = k1;
  if (free1) *free1 = v1;
}

void set_lds_limits() {
  static bool done = false;
  if (done) return;
  done = true;
  // dynamic LDS up to 64 KiB for the tile kernels (the plan keeps tiles within it)
#define HIPNG_LDS1(RR, TT, QQ, BB)                                                                          \
  (void)hipFuncSetAttribute((const void*)k_grid8<RR, TT, QQ, BB, grid8_rows<RR>()>,                               \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsBudget)
#define HIPNG_LDS(RR, TT, QQ) \
  HIPNG_LDS1(RR, TT, QQ, 1); HIPNG_LDS1(RR, TT, QQ, 2); HIPNG_LDS1(RR, TT, QQ, 3); HIPNG_LDS1(RR, TT, QQ, 4)
  HIPNG_LDS(float, 32, 1); HIPNG_LDS(float, 32, 2); HIPNG_LDS(float, 16, 1); HIPNG_LDS(float, 16, 2);
  HIPNG_LDS(double, 32, 1); HIPNG_LDS(double, 32, 2); HIPNG_LDS(double, 16, 1); HIPNG_LDS(double, 16, 2);
#undef HIPNG_LDS
#undef HIPNG_LDS1
  (void)hipGetLastError();
}

// Everything that depends only on (uvw, freq, image geometry, epsilon, precision, batch):
// plan, device tables, sorted permutation, bucket table, sorted coordinate records and the
// execution buffers. Built once; executed any number of times.
struct Geom {
  hipng_plan P;
  KP k;
  Tables tb;
  int64_t nrow = 0, n = 0;
  int nchan = 0, nb = 0, prec = 0, device = 0;
  const double* uvw = nullptr;
  uint32_t *perm = nullptr, *bstart = nullptr, *bend = nullptr, *rec_a = nullptr, *flipb = nullptr;
  void* rec_p = nullptr;  // per sorted item phase (cplx<R>) of the unsort epilogue
  void *rec_f = nullptr, *plane = nullptr, *accimg = nullptr, *recv = nullptr, *fftwork = nullptr;
  size_t fftwork_size = 0;
  // pruned plane transforms: grid rows [xlo, xhi) (band, along gy), image side T (along gx)
  rocfft_plan fft_b_inv = nullptr, fft_b_fwd = nullptr, fft_t_inv = nullptr, fft_t_fwd = nullptr;
  // FP32 planes: predict's band transforms go out of place into plane2, so the non-image
  // columns of the band rows in plane stay zero through a call
  void* plane2 = nullptr;
  cplx<double>* fstw = nullptr;  // four-step twiddles (FP64 planes of 8 fsm x 128 points along x)
  int fsm = 0;
  rocfft_plan fft_b_fwd_oop = nullptr;
  // ... and invert's, with lastw[tile] = the last pass that gridded into the tile
  rocfft_plan fft_b_inv_oop = nullptr;
  uint32_t* lastw = nullptr;
  // every out-of-place transform of this geometry leaves its input untouched (checked once per
  // rocFFT plan on a pattern): the buffers cleared once per call stay clear. Otherwise the
  // band transforms run in place and the transposed buffers are cleared before every pass.
  bool oop_ok = true;
  int xlo = 0, xhi = 0;
  void* tbuf = nullptr;   // transposed band / image rows, zero where no kernel writes (cleared per call)
  void* tbuf2 = nullptr;  // its transforms along the grid rows
  uint32_t *task = nullptr, *taskpart = nullptr, *counters = nullptr;
  uint32_t* tasktab = nullptr;  // tile of every task of a pass
  uint32_t chunk_grid = 1024, chunk_degrid = 1024;
  int npersist = 1024;
  int pc = 1;  // w planes per gridding pass / FFT batch
  // dense runs: tasks (first record, count) and their start planes, ascending
  bool dense = false;
  uint4* dtask = nullptr;
  uint32_t* dplane = nullptr;
  // records outside dense runs: sorted record indices and their bucket table
  uint2* srec = nullptr;  // non-dense records: (record index, origin word)
  uint32_t *sbstart = nullptr, *sbend = nullptr;
  // strip ranges of the sub-tile gatherer over its buckets (gb0 / gb1 below)
  Strips strips;
  uint32_t* split = nullptr;
  TapPoly taps{};  // interior kernel taps as polynomials (FP32 kernel modes)
  Buffers buf;
};

// shortest run the dense-run gridder takes: kDenseMin unless the plan options pin it
// (dense_min -1 turns the path off, every record then goes through the sub-tile gatherer)
uint32_t dense_min_run(const hipng_plan_opts& o) { return o.dense_min >= 2 ? (uint32_t)o.dense_min : kDenseMin; }

int device_cus() {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cus <= 0) cus = 256;
  }
  return cus;
}

uint32_t pick_chunk(double visits_per_plane, int npersist) {
  double c = visits_per_plane / (2.0 * npersist);
  c = std::min(16384.0, std::max(256.0, c));
  return (uint32_t)((uint64_t)(c + 255) / 256 * 256);
}


void check_fft_inputs(Geom& g, DevCtx& c, hipStream_t st);

void build_geom(Geom& g, DevCtx& c, const double* uvw, const double* freq, int64_t nrow, int nchan,
                int nb, int nx, int ny, double px, double py, double eps, int wstack, int prec,
                const hipng_plan_opts* opts, hipStream_t st, bool own) {
  if (!freq) HIPNG_THROW(HIPNG_EINVAL, "freq must not be NULL");
  if (nrow < 0 || nchan <= 0) HIPNG_THROW(HIPNG_EINVAL, "bad nrow/nchan (%lld, %d)", (long long)nrow, nchan);
  if (nrow > 0 && !uvw) HIPNG_THROW(HIPNG_EINVAL, "uvw must not be NULL");
  const int64_t n = nrow * (int64_t)nchan;
  if (n >= 4294967295LL) HIPNG_THROW(HIPNG_EINVAL, "too many visibilities for 32-bit indices (%lld)", (long long)n);
  hipng_plan_opts o;
  std::memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  if (o.dense_min < -1 || o.dense_min == 1) HIPNG_THROW(HIPNG_EINVAL, "opts.dense_min must be -1, 0 or >= 2");
  if (o.chunk_grid < 0 || o.chunk_degrid < 0) HIPNG_THROW(HIPNG_EINVAL, "opts chunk sizes must be >= 0");
  set_lds_limits();
  HIP_TRY(hipGetDevice(&g.device));
  g.buf.borrow = own ? nullptr : &c;
  double e[4];
  extent(c, uvw, nrow, freq, nchan, e, st);
  make_plan(e[0], e[1], e[2], e[3], nx, ny, px, py, eps, wstack, prec, nb, n, opts, &g.P);
  g.k = make_kp(g.P, nb, nchan);
  if (prec != HIPNG_FP64) g.taps = make_tap_poly(g.P.W, g.P.beta, g.P.sigma);
  g.tb = upload_tables(g.buf, g.P, freq, nchan, st);
  g.nrow = nrow;
  g.n = n;
  g.nchan = nchan;
  g.nb = nb;
  g.prec = prec;
  const hipng_plan& P = g.P;
  const size_t csz = prec == HIPNG_FP64 ? 16 : 8;
  const size_t asz = prec == HIPNG_FP32 ? 4 : 8;
  g.accimg = g.buf.get(S_ACCIMG, (size_t)nb * nx * ny * asz);
  if (n == 0) return;
  if (own) {
    auto* u = (double*)g.buf.get(S_EXT, (size_t)nrow * 3 * sizeof(double));
    HIP_TRY(hipMemcpyAsync(u, uvw, (size_t)nrow * 3 * sizeof(double), hipMemcpyDeviceToDevice, st));
    g.uvw = u;
  } else {
    g.uvw = uvw;
  }
  uint32_t *keys, *perm, *scr0, *scr1;
  g.rec_a = (uint32_t*)g.buf.get(S_RECA, n * 4);
  sort_items(c, g.uvw, n, g.tb, g.k, P, o.sort_pack != -1, st, &keys, &perm, &scr0, &scr1);
  // dense-run gridding needs W <= 8 (one lane per footprint cell) and a free flag bit
  // in the packed first cell (grids up to 32768 cells per axis)
  g.dense = o.dense_min != -1 && P.W <= 8 && P.gx <= 32768 && P.gy <= 32768;
  if (g.dense) {
    g.k.amask = 0x7fff7fffu;
    g.k.dflag = 0x80000000u;
    g.P.dense_min = (int32_t)dense_min_run(o);
  }
  if (own) {
    g.perm = (uint32_t*)g.buf.get(S_VALS0, n * 4);
    HIP_TRY(hipMemcpyAsync(g.perm, perm, n * 4, hipMemcpyDeviceToDevice, st));
  } else {
    g.perm = perm;
  }
  const int64_t nkeys = (int64_t)P.nplanes * P.ntx * P.nty;
  g.bstart = (uint32_t*)g.buf.get(S_BSTART, nkeys * 4);
  g.bend = (uint32_t*)g.buf.get(S_BEND, nkeys * 4);
  k_fill_u32<<<gridcap(nkeys), kBlock, 0, st>>>(g.bstart, nkeys, 0u);
  k_fill_u32<<<gridcap(nkeys), kBlock, 0, st>>>(g.bend, nkeys, 0u);
  LAUNCH_CHECK();
  {
    PROF("bounds");
    k_bounds<<<grid1(n), kBlock, 0, st>>>(keys, n, g.bstart, g.bend);
    LAUNCH_CHECK();
  }
  g.rec_f = g.buf.get(S_RECF, n * (prec == HIPNG_FP64 ? sizeof(rec3<double>) : sizeof(rec3<float>)));
  g.rec_p = g.buf.get(S_RECP, (size_t)n * csz);
  g.flipb = (uint32_t*)g.buf.get(S_FLIPB, (size_t)((n + 31) / 32 + 2) * 4);
  {
    PROF("records");
    if (prec == HIPNG_FP64)
      k_records<double, false><<<grid1(n), kBlock, 0, st>>>(g.uvw, g.perm, n, g.tb.sx, g.tb.sy, g.tb.sw, g.tb.pn,
                                                            g.k, g.rec_a, (rec3<double>*)g.rec_f,
                                                            (cplx<double>*)g.rec_p, g.flipb);
    else if (prec == HIPNG_MIXED)
      k_records<float, false><<<grid1(n), kBlock, 0, st>>>(g.uvw, g.perm, n, g.tb.sx, g.tb.sy, g.tb.sw, g.tb.pn,
                                                           g.k, g.rec_a, (rec3<float>*)g.rec_f,
                                                           (cplx<float>*)g.rec_p, g.flipb);
    else
      k_records<float, true><<<grid1(n), kBlock, 0, st>>>(g.uvw, g.perm, n, g.tb.sx, g.tb.sy, g.tb.sw, g.tb.pn,
                                                          g.k, g.rec_a, (rec3<float>*)g.rec_f,
                                                          (cplx<float>*)g.rec_p, g.flipb);
    LAUNCH_CHECK();
  }
  g.recv = g.buf.get(S_RECV, (size_t)n * nb * csz);
  // planes per pass: up to kMaxPC while the plane batch stays within 40 GB (the gridders'
  // per-lane accumulators, one cell x planes x batch x complex, stay small)
  const size_t plane_bytes = (size_t)nb * P.gx * P.gy * csz;
  g.pc = 1;
  for (int q = kMaxPC; q > 1; q /= 2) {
    if ((double)q * plane_bytes <= 40e9 && q <= P.nplanes) {
      g.pc = q;
      break;
    }
  }
  g.plane = g.buf.get(S_PLANE, g.pc * plane_bytes);
  // band of grid rows visibility footprints can touch: |u| sx <= umax px gx around gx/2, plus
  // the kernel half-width and two rows of rounding margin; the whole grid when it would wrap
  {
    const double U = P.umax * P.pixsize_x * P.gx;
    const double lo = std::floor(0.5 * P.gx - U - 0.5 * P.W - 2.0);
    const double hi = std::ceil(0.5 * P.gx + U + 0.5 * P.W + 2.0) + 1.0;
    g.xlo = lo < 0.0 ? 0 : (int)lo;
    g.xhi = hi > P.gx ? P.gx : (int)hi;
    if (lo < 0.0 || hi > P.gx) {
      g.xlo = 0;
      g.xhi = P.gx;
    }
  }
  g.tbuf = g.buf.get(S_TBUF, (size_t)g.pc * nb * P.ny * P.gx * csz);
  // FP32 planes: room for two passes, so that invert screens both into the FP64 image
  // accumulator with one read-modify-write
  g.tbuf2 = g.buf.get(S_TBUF2, (size_t)(csz == 8 ? std::max(kScreenGroup, kFillGroup) : 1) * g.pc * nb * P.ny *
                                   P.gx * csz);
  g.fft_b_inv = fft_plan1(c, P.gy, g.xhi - g.xlo, prec, 0);
  g.fft_b_fwd = fft_plan1(c, P.gy, g.xhi - g.xlo, prec, 1);
  g.fft_t_inv = fft_plan1(c, P.gx, g.pc * nb * P.ny, prec, 0, false);
  g.fft_t_fwd = fft_plan1(c, P.gx, g.pc * nb * P.ny, prec, 1, false);
  // the band transforms go out of place into plane2, so that a pass clears only the tiles the
  // previous pass gridded into: always for FP32 planes, for FP64 ones when the second set of
  // planes takes at most 1/16 of the device memory (C1 FP64 with 2 polarisations: 9.7 GB; the
  // per-pass clear 31 -> 9 ms per invert, 29 ms less per invert + predict job)
  size_t mem_free = 0, mem_total = 0;
  if (csz != 8 && hipMemGetInfo(&mem_free, &mem_total) != hipSuccess) {
    (void)hipGetLastError();
    mem_total = 0;
  }
  if (csz == 8 || (size_t)g.pc * plane_bytes <= mem_total / 16) {
    g.plane2 = g.buf.get(S_PLANE2, g.pc * plane_bytes);
    g.fft_b_fwd_oop = fft_plan1(c, P.gy, g.xhi - g.xlo, prec, 1, false);
    g.fft_b_inv_oop = fft_plan1(c, P.gy, g.xhi - g.xlo, prec, 0, false);
    g.lastw = (uint32_t*)g.buf.get(S_LASTW, (size_t)P.ntx * P.nty * 4);
  }
  g.fsm = csz == 16 ? fs_m_for(P.gx) : 0;
  g.fstw = nullptr;
  if (g.fsm) {
    const int ntw = fs_tw_len(g.fsm);
    g.fstw = (cplx<double>*)g.buf.get(S_FSTW, (size_t)ntw * sizeof(cplx<double>));
    k_fs_twiddles<<<(ntw + kBlock - 1) / kBlock, kBlock, 0, st>>>(g.fstw, g.fsm);
    LAUNCH_CHECK();
  }
  g.fftwork_size = 0;
  for (rocfft_plan pl : {g.fft_b_inv, g.fft_b_fwd, g.fft_t_inv, g.fft_t_fwd, g.fft_b_fwd_oop, g.fft_b_inv_oop}) {
    if (!pl) continue;
    size_t w = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(pl, &w));
    g.fftwork_size = std::max(g.fftwork_size, w);
  }
  if (g.fftwork_size) g.fftwork = g.buf.get(S_FFTWORK, g.fftwork_size);
  check_fft_inputs(g, c, st);
  const int64_t ntiles = (int64_t)P.ntx * P.nty;
  g.task = (uint32_t*)g.buf.get(S_TASK, ntiles * 4);
  g.taskpart = (uint32_t*)g.buf.get(S_TASKPART, ((ntiles + kRsItems - 1) / kRsItems) * 4);
  g.counters = (uint32_t*)g.buf.get(S_CNT, kNCounters * sizeof(uint32_t));
  g.npersist = device_cus() * 4;
  const double ww = P.wstacking ? P.W : 1;
  g.chunk_grid = pick_chunk((double)n * ww / P.nplanes * 4.0, g.npersist);
  g.chunk_degrid = pick_chunk((double)n * ww / P.nplanes, g.npersist);
  // the plan options may pin the candidates per task (calibration)
  if (o.chunk_grid > 0) g.chunk_grid = (uint32_t)std::max(256, o.chunk_grid) / 256 * 256;
  if (o.chunk_degrid > 0) g.chunk_degrid = (uint32_t)std::max(256, o.chunk_degrid) / 256 * 256;
  {
    // tasks of a pass: sum over tiles of ceil(candidates / chunk) x strips, where a record is
    // a candidate of at most 4 tiles (gridder) or 1 (degridder)
    const int64_t tg = (4 * n / g.chunk_grid + ntiles) * (P.tile / 8);
    const int64_t td = n / g.chunk_degrid + ntiles;
    g.tasktab = (uint32_t*)g.buf.get(S_TASKTAB, (size_t)(std::max(tg, td) + 1) * 4);
  }
  int64_t nsparse = n;  // records on the gatherer's list
  if (g.dense) {
    PROF("runs");
    // scr0 = exclusive scan of the run heads (run of each record); scr1 = run starts
    // (n + 1 entries); per-run task counts and dense lengths in workspace, scanned over the
    // runs only. Two synchronisations: the run count and the dense total size the lists.
    uint32_t* part = (uint32_t*)ws(c, S_HIST, (size_t)((n + kRsItems) / kRsItems + 1) * 4);
    const uint32_t dmin = dense_min_run(o);
    {
      const int64_t np = (n + kRsItems - 1) / kRsItems;
      k_runscan_partial<<<(unsigned)np, kBlock, 0, st>>>(keys, g.rec_a, n, part);
      LAUNCH_CHECK();
      k_scan_top<<<1, kBlock, 0, st>>>(part, np);
      LAUNCH_CHECK();
      k_runscan_apply<<<(unsigned)np, kBlock, 0, st>>>(keys, g.rec_a, n, part, scr0, scr1, g.counters);
      LAUNCH_CHECK();
    }
    uint32_t nruns = 0;
    HIP_TRY(hipMemcpyAsync(&nruns, g.counters + kRuns, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    auto* tcnt = (uint32_t*)ws(c, S_RUNT, ((size_t)nruns + 1) * 4);
    auto* dpre = (uint32_t*)ws(c, S_RUND, ((size_t)nruns + 1) * 4);
    k_run_info<<<grid1((int64_t)nruns + 1), kBlock, 0, st>>>(scr1, nruns, dmin, kDenseChunk, tcnt, dpre);
    LAUNCH_CHECK();
    excl_scan(tcnt, (int64_t)nruns + 1, part, st);
    excl_scan(dpre, (int64_t)nruns + 1, part, st);
    uint32_t ntask = 0, ndense = 0;
    HIP_TRY(hipMemcpyAsync(&ntask, tcnt + nruns, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&ndense, dpre + nruns, 4, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    g.dtask = (uint4*)g.buf.get(S_DTASK, (size_t)std::max<uint32_t>(ntask, 1) * sizeof(uint4));
    g.dplane = (uint32_t*)g.buf.get(S_DPLANE, (size_t)std::max<uint32_t>(ntask, 1) * 4);
    k_run_tasks<<<grid1(std::max<int64_t>(nruns, 1)), kBlock, 0, st>>>(keys, g.rec_a, scr1, tcnt, nruns, dmin,
                                                                       kDenseChunk, g.k.ntiles, g.k.amask,
                                                                       g.dtask, g.dplane, g.counters);
    LAUNCH_CHECK();
    const int64_t ns = n - (int64_t)ndense;
    nsparse = ns;
    g.srec = (uint2*)g.buf.get(S_SREC, (size_t)std::max<int64_t>(ns, 1) * sizeof(uint2));
    k_run_mark<<<grid1(n), kBlock, 0, st>>>(g.rec_a, n, scr0, scr1, dpre, g.counters, dmin, g.k.dflag, g.srec);
    LAUNCH_CHECK();
    g.sbstart = (uint32_t*)g.buf.get(S_SBSTART, nkeys * 4);
    g.sbend = (uint32_t*)g.buf.get(S_SBEND, nkeys * 4);
    k_sparse_bounds<<<grid1(nkeys), kBlock, 0, st>>>(g.bstart, g.bend, nkeys, n, scr0, dpre, g.counters,
                                                     g.sbstart, g.sbend);
    LAUNCH_CHECK();
  }
  {
    // strip split points over the gatherer's list (the compacted non-dense records, or all)
    PROF("runs");
    g.strips = make_strips(P.tile, P.W);
    const uint32_t* gb1 = g.srec ? g.sbend : g.bend;
    const int64_t nrec = g.srec ? nsparse : n;
    g.split = (uint32_t*)g.buf.get(S_SPLIT, (size_t)std::max(1, g.strips.nv) * nkeys * 4);
    if (g.strips.nv > 0) {
      k_split_init<<<gridcap((int64_t)nkeys * g.strips.nv), kBlock, 0, st>>>(gb1, (uint32_t)nkeys, g.strips.nv,
                                                                               g.split);
      LAUNCH_CHECK();
      if (nrec > 0) {
        k_split_points<<<grid1(nrec), kBlock, 0, st>>>(keys, g.srec, g.rec_a, nrec, P.tile, g.k.amask, g.strips,
                                                       (uint32_t)nkeys, g.split);
        LAUNCH_CHECK();
      }
    }
  }
  if (own) HIP_TRY(hipStreamSynchronize(st));
}

// per-plane task list: counters reset, per-tile task counts, exclusive scan -> task offsets
// strips: the gatherer's tasks (the longest strip range list of a tile, times its strips)
void build_tasks(const Geom& g, int slo, int shi, int nq, uint32_t chunk, hipStream_t st, uint32_t mult = 1,
                 const uint32_t* bs = nullptr, const uint32_t* be = nullptr, bool strips = false) {
  const int64_t ntiles = g.k.ntiles;
  k_fill_u32<<<1, kBlock, 0, st>>>(g.counters, kDenseCnt + 1, 0u);
  LAUNCH_CHECK();
  if (strips)
    k_tasks_grid<<<grid1(ntiles), kBlock, 0, st>>>(g.k, slo, shi, bs ? bs : g.bstart, be ? be : g.bend, g.split,
                                                   g.strips, chunk, g.task, g.counters);
  else
    k_tasks<<<grid1(ntiles), kBlock, 0, st>>>(g.k, slo, shi, nq, bs ? bs : g.bstart, be ? be : g.bend, chunk, g.task,
                                              g.counters, mult);
  LAUNCH_CHECK();
  excl_scan(g.task, ntiles, g.taskpart, st);
  k_task_tiles<<<grid1(ntiles), kBlock, 0, st>>>(g.task, ntiles, g.counters, g.tasktab);
  LAUNCH_CHECK();
}

// in place on `data`, or from `data` to `out` with an out-of-place plan
void geom_fft(const Geom& g, rocfft_plan plan, void* data, hipStream_t st, void* out = nullptr) {
  rocfft_execution_info info = nullptr;
  FFT_TRY(rocfft_execution_info_create(&info));
  rocfft_status e = rocfft_execution_info_set_stream(info, st);
  if (e == rocfft_status_success && g.fftwork_size)
    e = rocfft_execution_info_set_work_buffer(info, g.fftwork, g.fftwork_size);
  void* bufs[1] = {data};
  void* obufs[1] = {out};
  if (e == rocfft_status_success) e = rocfft_execute(plan, bufs, out ? obufs : nullptr, info);
  rocfft_execution_info_destroy(info);
  if (e != rocfft_status_success) HIPNG_THROW(HIPNG_ERUNTIME, "rocfft failed: %d", (int)e);
}

// Does an out-of-place plan leave its input untouched? Checked once per rocFFT plan: a
// pattern in `in` (n complex values of the geometry's precision), its word sum before and
// after one execution into `out`. Every buffer involved is re-initialised at the start of
// each invert / predict, so the check may write them.
bool fft_keeps_input(const Geom& g, DevCtx& c, rocfft_plan plan, void* in, int64_t n, void* out, hipStream_t st) {
  auto it = c.keeps_input.find(plan);
  if (it != c.keeps_input.end()) return it->second;
  const size_t rsz = g.prec == HIPNG_FP64 ? 8 : 4;
  const int64_t nr = 2 * n, nw = nr * (int64_t)rsz / 4;
  if (g.prec == HIPNG_FP64) k_pattern<double><<<gridcap(nr), kBlock, 0, st>>>((double*)in, nr);
  else k_pattern<float><<<gridcap(nr), kBlock, 0, st>>>((float*)in, nr);
  LAUNCH_CHECK();
  auto* d = (unsigned long long*)ws(c, S_EXT, 2 * sizeof(unsigned long long));
  k_zero<unsigned long long><<<1, kBlock, 0, st>>>(d, 2);
  k_word_sum<<<std::min<unsigned>(gridcap(nw), 4096), kBlock, 0, st>>>((const uint32_t*)in, nw, d);
  LAUNCH_CHECK();
  geom_fft(g, plan, in, st, out);
  k_word_sum<<<std::min<unsigned>(gridcap(nw), 4096), kBlock, 0, st>>>((const uint32_t*)in, nw, d + 1);
  LAUNCH_CHECK();
  unsigned long long h[2] = {0ull, 0ull};
  HIP_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const bool keeps = h[0] == h[1];
  c.keeps_input[plan] = keeps;
  return keeps;
}

// The out-of-place transforms of a geometry: band rows plane -> plane2 (FP32 planes, both
// directions) and the transposed rows T -> T2 (invert) / S -> D (predict).
void check_fft_inputs(Geom& g, DevCtx& c, hipStream_t st) {
  const hipng_plan& P = g.P;
  const int64_t nband = (int64_t)(g.xhi - g.xlo) * P.gy;
  const int64_t ntr = (int64_t)g.pc * g.nb * P.ny * P.gx;
  const size_t csz = g.prec == HIPNG_FP64 ? 16 : 8;
  char* plane = (char*)g.plane + (size_t)g.xlo * P.gy * csz;
  char* plane2 = g.plane2 ? (char*)g.plane2 + (size_t)g.xlo * P.gy * csz : nullptr;
  const bool grouped = csz == 8;  // FP32 predict fills S = T2 and transforms into D = T
  bool ok = fft_keeps_input(g, c, g.fft_t_inv, g.tbuf, ntr, g.tbuf2, st);
  ok = fft_keeps_input(g, c, g.fft_t_fwd, grouped ? g.tbuf2 : g.tbuf, ntr, grouped ? g.tbuf : g.tbuf2, st) && ok;
  if (g.fft_b_inv_oop) ok = fft_keeps_input(g, c, g.fft_b_inv_oop, plane, nband, plane2, st) && ok;
  if (g.fft_b_fwd_oop) ok = fft_keeps_input(g, c, g.fft_b_fwd_oop, plane, nband, plane2, st) && ok;
  g.oop_ok = ok;
}

// polarisation slots per cell of the degridder's LDS tile (whole 16-byte units)
int degrid_slots(const Geom& g) {
  if (g.prec == HIPNG_FP64) return g.nb;
  return g.nb == 1 ? dg_slots<float, 1>() : g.nb == 2 ? dg_slots<float, 2>() : g.nb == 3 ? dg_slots<float, 3>()
                                                                                          : dg_slots<float, 4>();
}

// planes whose tiles the degridder holds in LDS at once: as many of the pass's planes as fit
// in kDegridLds (four workgroups per CU), at least one
int degrid_planes(const Geom& g, int npl) {
  const size_t L = (size_t)g.k.T + g.k.W - 1;
  const size_t per = L * L * degrid_slots(g) * (g.prec == HIPNG_FP64 ? 16 : 8);
  int q = npl;
  while (q > 1 && q * per > kDegridLds) --q;
  return q;
}

// dense-run kernels: 5 workgroups per CU fit their LDS (at most 30 KiB each, FP32)
constexpr int kDenseBlocksPerCu = 5;

// with dense runs: the dense runs on the matrix cores, the other records in k_degrid;
// otherwise every record in k_degrid
template <typename R>
bool degrid_split(const Geom& g) {
  return g.srec != nullptr;
}

template <typename R>
void launch_degrid_dense(const Geom& g, int p0, int npl, const cplx<R>* plane, cplx<R>* accs, hipStream_t st) {
  const KP& k = g.k;
  const int64_t pstride = (int64_t)g.nb * k.gx * k.gy;
  const auto* rf = (const rec3<R>*)g.rec_f;
  const int nblk = device_cus() * kDenseBlocksPerCu;
  if (g.nb == 1)
    k_degrid_dense_mx<R, 1><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, plane, pstride, accs, g.counters, g.taps);
  else if (g.nb == 2)
    k_degrid_dense_mx<R, 2><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, plane, pstride, accs, g.counters, g.taps);
  else if (g.nb == 3)
    k_degrid_dense_mx<R, 3><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, plane, pstride, accs, g.counters, g.taps);
  else
    k_degrid_dense_mx<R, 4><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, plane, pstride, accs, g.counters, g.taps);
  LAUNCH_CHECK();
}

template <typename R>
void launch_degrid(const Geom& g, int p0, int npl, const cplx<R>* plane, cplx<R>* accs, hipStream_t st) {
  const KP& k = g.k;
  const int L = k.T + k.W - 1;
  const size_t sm = (size_t)npl * degrid_slots(g) * L * L * sizeof(cplx<R>);
  const int64_t pstride = (int64_t)g.nb * k.gx * k.gy;
  const auto* rf = (const rec3<R>*)g.rec_f;
  const bool split = degrid_split<R>(g);
  const uint32_t* b0 = split ? g.sbstart : g.bstart;
  const uint32_t* b1 = split ? g.sbend : g.bend;
  const uint2* sr = split ? g.srec : nullptr;
#define HIPNG_DG1(WW, BB)                                                                          \
  k_degrid<R, WW, BB><<<g.npersist, kBlock, sm, st>>>(k, p0, npl, b0, b1, sr, g.rec_a, rf, plane, \
                                                      pstride, accs, g.task, g.tasktab, g.chunk_degrid,  \
                                                      g.counters, g.taps)
#define HIPNG_DG(WW)                        \
  case WW:                                  \
    if (g.nb == 1) HIPNG_DG1(WW, 1);        \
    else if (g.nb == 2) HIPNG_DG1(WW, 2);   \
    else if (g.nb == 3) HIPNG_DG1(WW, 3);   \
    else HIPNG_DG1(WW, 4);                  \
    break;
  switch (k.W) {
    HIPNG_DG(4) HIPNG_DG(5) HIPNG_DG(6) HIPNG_DG(7) HIPNG_DG(8) HIPNG_DG(9) HIPNG_DG(10)
    HIPNG_DG(11) HIPNG_DG(12) HIPNG_DG(13) HIPNG_DG(14) HIPNG_DG(15) HIPNG_DG(16)
    default: HIPNG_THROW(HIPNG_EINVAL, "unsupported kernel support %d", k.W);
  }
#undef HIPNG_DG
#undef HIPNG_DG1
  LAUNCH_CHECK();
}

// 16-byte value accesses need the caller's arrays 16-byte aligned (FP32 with an even batch),
// and the weights (if any) aligned to one item's NB weights
template <typename R>
bool vec_ok(const Geom& g, const void* p, const void* wgt) {
  return std::is_same<R, float>::value && g.nb % 2 == 0 && ((uintptr_t)p & 15u) == 0 &&
         (!wgt || ((uintptr_t)wgt & (uintptr_t)(g.nb * sizeof(R) - 1)) == 0);
}

template <typename R>
void launch_values(const Geom& g, const cplx<R>* ms, const R* wgt, cplx<R>* recv, hipStream_t st) {
  const auto* rp = (const cplx<R>*)g.rec_p;
  const bool vec = vec_ok<R>(g, ms, wgt);
#define HIPNG_KV(BB)                                                                                    \
  if (vec) k_values<R, BB, true><<<grid1(g.n), kBlock, 0, st>>>(g.perm, g.n, rp, g.flipb, ms, wgt, recv); \
  else k_values<R, BB, false><<<grid1(g.n), kBlock, 0, st>>>(g.perm, g.n, rp, g.flipb, ms, wgt, recv)
  if (g.nb == 1) { HIPNG_KV(1); } else if (g.nb == 2) { HIPNG_KV(2); } else if (g.nb == 3) { HIPNG_KV(3); } else { HIPNG_KV(4); }
#undef HIPNG_KV
  LAUNCH_CHECK();
}

template <typename R>
void launch_unsort(const Geom& g, const cplx<R>* accs, const R* wgt, cplx<R>* ms, hipStream_t st) {
  const auto* rp = (const cplx<R>*)g.rec_p;
  const bool vec = vec_ok<R>(g, ms, wgt);
#define HIPNG_KU(BB)                                                                                    \
  if (vec) k_unsort<R, BB, true><<<grid1(g.n), kBlock, 0, st>>>(g.perm, g.n, rp, g.flipb, accs, wgt, ms); \
  else k_unsort<R, BB, false><<<grid1(g.n), kBlock, 0, st>>>(g.perm, g.n, rp, g.flipb, accs, wgt, ms)
  if (g.nb == 1) { HIPNG_KU(1); } else if (g.nb == 2) { HIPNG_KU(2); } else if (g.nb == 3) { HIPNG_KU(3); } else { HIPNG_KU(4); }
#undef HIPNG_KU
  LAUNCH_CHECK();
}

// dense-run gridder launches (matrix cores, FP32 or FP64 records)
void launch_grid_dense(const Geom& g, const KP& k, int p0, int npl, const rec3<float>* rf,
                       const cplx<float>* recv, cplx<float>* plane, hipStream_t st) {
  const int nblk = device_cus() * kDenseBlocksPerCu;
  if (g.nb == 1)
    k_grid_dense_mx<float, 1><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else if (g.nb == 2)
    k_grid_dense_mx<float, 2><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else if (g.nb == 3)
    k_grid_dense_mx<float, 3><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else
    k_grid_dense_mx<float, 4><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
}
void launch_grid_dense(const Geom& g, const KP& k, int p0, int npl, const rec3<double>* rf,
                       const cplx<double>* recv, cplx<double>* plane, hipStream_t st) {
  const int nblk = device_cus() * kDenseBlocksPerCu;
  if (g.nb == 1)
    k_grid_dense_mx<double, 1><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else if (g.nb == 2)
    k_grid_dense_mx<double, 2><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else if (g.nb == 3)
    k_grid_dense_mx<double, 3><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
  else
    k_grid_dense_mx<double, 4><<<nblk, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, rf, recv, g.counters, plane, g.taps);
}

// Columns of one call (hipng_geometry_invert_columns / _predict_columns)
struct ColCall {
  ColArgs a;
  int npv = 1;
  bool image_yx = false, image_f32 = false;  // image arrays [nb][ny][nx] (else [nb][nx][ny]), element type
  Scale4 scale{{1.0, 1.0, 1.0, 1.0}};
  const void *vis = nullptr, *wgt = nullptr, *flags = nullptr;
  const cplx<double>* gains = nullptr;
  const int *ant1 = nullptr, *ant2 = nullptr;
  void* model = nullptr;
};

// ColArgs.vec: which columns allow one access per item (npol items of the element type,
// aligned to 16 bytes for complex64 pairs, to npol elements for weights and flags)
inline bool aligned_to(const void* p, size_t a) { return p && ((uintptr_t)p % a) == 0; }
int col_vec_bits(const ColCall& cc, int npol, const void* vis, const void* model) {
  int v = 0;
  const bool even = npol % 2 == 0;
  if (cc.a.vis_f32 && even && aligned_to(vis, 16)) v |= 1;
  const size_t ws = cc.a.wgt_f32 ? 4 : 8;
  if ((npol == 4 || npol == 2) && aligned_to(cc.wgt, cc.a.wgt_f32 ? ws * npol : 16)) v |= 2;
  const size_t fs = cc.a.flag_i8 ? 1 : 4;
  if (npol == 4 && aligned_to(cc.flags, fs * 4)) v |= 4;
  if (cc.a.vis_f32 && even && aligned_to(model, 16)) v |= 8;
  return v;
}

// calls f(integral_constant<NI>, integral_constant<NO>) for 1 <= NI, NO <= 4
template <typename F>
void pol_dispatch(int ni, int no, F&& f) {
  using I1 = std::integral_constant<int, 1>;
  using I2 = std::integral_constant<int, 2>;
  using I3 = std::integral_constant<int, 3>;
  using I4 = std::integral_constant<int, 4>;
  auto on_no = [&](auto nic) {
    if (no == 1) f(nic, I1{});
    else if (no == 2) f(nic, I2{});
    else if (no == 3) f(nic, I3{});
    else f(nic, I4{});
  };
  if (ni == 1) on_no(I1{});
  else if (ni == 2) on_no(I2{});
  else if (ni == 3) on_no(I3{});
  else on_no(I4{});
}

template <typename R>
void launch_values_cols(const Geom& g, const ColCall& cc, cplx<R>* recv, hipStream_t st) {
  const auto* rp = (const cplx<R>*)g.rec_p;
  ColArgs a = cc.a;
  a.vec = col_vec_bits(cc, cc.npv, cc.vis, nullptr);
  pol_dispatch(cc.npv, g.nb, [&](auto ni, auto no) {
    k_values_cols<R, decltype(ni)::value, decltype(no)::value><<<grid1(g.n), kBlock, 0, st>>>(
        g.perm, g.n, rp, g.flipb, g.uvw, g.tb.fc, a, cc.vis, cc.wgt, cc.flags, cc.gains, cc.ant1, cc.ant2, recv);
  });
  LAUNCH_CHECK();
}

template <typename R>
void launch_unsort_cols(const Geom& g, const ColCall& cc, const cplx<R>* accs, void* vis, hipStream_t st) {
  const auto* rp = (const cplx<R>*)g.rec_p;
  ColArgs a = cc.a;
  a.vec = col_vec_bits(cc, cc.npv, vis, cc.model);
  pol_dispatch(g.nb, cc.npv, [&](auto ni, auto no) {
    k_unsort_cols<R, decltype(ni)::value, decltype(no)::value><<<grid1(g.n), kBlock, 0, st>>>(
        g.perm, g.n, rp, g.flipb, accs, g.uvw, g.tb.fc, a, cc.gains, cc.ant1, cc.ant2, vis, cc.model);
  });
  LAUNCH_CHECK();
}

// R: kernel/grid real type; A: image accumulator type; F32C: coordinates/phase in FP32.
// Kernel launches and rocFFT only: no allocation, no host synchronisation. `cc` (nullable): the
// visibilities come from BlockVisibility-layout columns through the fused value gather.
template <typename R, typename A, bool F32C>
void exec_ms2dirty(const Geom& g, const void* ms, const void* wgt, void* dirty, hipStream_t st,
                   const ColCall* cc = nullptr) {
  using O = typename std::conditional<std::is_same<R, double>::value, double, float>::type;
  const KP& k = g.k;
  const hipng_plan& P = g.P;
  const int64_t nimg = (int64_t)g.nb * P.nx * P.ny;
  if (g.n == 0) {
    k_zero<O><<<gridcap(nimg), kBlock, 0, st>>>((O*)dirty, nimg);
    LAUNCH_CHECK();
    return;
  }
  if (!ms && !(cc && cc->vis)) HIPNG_THROW(HIPNG_EINVAL, "ms must not be NULL");
  auto* recv = (cplx<R>*)g.recv;
  {
    PROF("values");
    if (cc) launch_values_cols<R>(g, *cc, recv, st);
    else launch_values<R>(g, (const cplx<R>*)ms, (const R*)wgt, recv, st);
    LAUNCH_CHECK();
  }
  auto* acc = (A*)g.accimg;  // the first screen update stores into it
  auto* plane = (cplx<R>*)g.plane;
  auto* tbuf = (cplx<R>*)g.tbuf;
  auto* tbuf2 = (cplx<R>*)g.tbuf2;
  // FP64 planes of a length the four-step image-side transform covers (stage 1 writes all of
  // its buffer, so T needs no clearing)
  const bool fs = sizeof(R) == 8 && g.fstw != nullptr;
  if (!fs) {
    // the per-pass kernels write only the band (invert) or the image (predict) part of T
    PROF("transpose");
    // k_tr_in rewrites the band columns [xlo, xhi) of T every pass; the others stay zero
    const int64_t rows = (int64_t)g.pc * g.nb * P.ny;
    zero_cols(tbuf, rows, P.gx, 0, g.xlo, st);
    zero_cols(tbuf, rows, P.gx, g.xhi, P.gx, st);
    LAUNCH_CHECK();
  }
  const int PCn = g.pc;
  const int64_t np1 = (int64_t)g.nb * P.gx * P.gy;
  const uint32_t* gb0 = g.srec ? g.sbstart : g.bstart;
  const uint32_t* gb1 = g.srec ? g.sbend : g.bend;
  // per-wave LDS of the sub-tile gatherer (at most 62.6 KiB: FP64, 4 polarisations, 2 planes)
  const size_t smw = (kBlock / 64) * grid8_region(sizeof(R), PCn, g.nb, grid8_rows<R>() / 8);
  // sub-tile gatherer row groups: two sub-tiles per candidate scan in the FP32 modes, one in
  // FP64 (whose vector accumulators would not leave room for a second sub-tile)
  constexpr int RG = grid8_rows<R>();
  if (smw > kLdsBudget) HIPNG_THROW(HIPNG_ERUNTIME, "gridder LDS %zu exceeds %zu", smw, kLdsBudget);
  const int ngrp = sizeof(R) == 4 ? kScreenGroup : 1;  // passes of FP32 planes T2 holds
  // FP32 planes: band transforms out of place into plane2, so a pass clears only the tiles the
  // previous pass gridded into (lastw); the band is cleared once per call
  const bool oop = g.plane2 != nullptr && g.oop_ok;
  cplx<R>* bplane = oop ? (cplx<R>*)g.plane2 : plane;
  const int64_t nband = (int64_t)(g.xhi - g.xlo) * P.gy;
  if (oop) {
    PROF("grid_setup");
    for (int i = 0; i < PCn * g.nb; ++i)
      k_zero<cplx<R>><<<gridcap(nband), kBlock, 0, st>>>(plane + ((int64_t)i * P.gx + g.xlo) * P.gy, nband);
    k_fill_u32<<<gridcap(k.ntiles), kBlock, 0, st>>>(g.lastw, k.ntiles, 0xfffffff0u);
    LAUNCH_CHECK();
  }
  for (int p0 = 0; p0 < P.nplanes; p0 += PCn) {
    const int npl = std::min(PCn, P.nplanes - p0);
    const uint32_t pass = (uint32_t)(p0 / PCn);
    {
      PROF("grid_setup");
      int slo, shi;
      chunk_range(k.Ww, P.nplanes, p0, npl, &slo, &shi);
      // the sub-tile gatherer walks the compacted table of records outside dense runs
      build_tasks(g, slo, shi, 4, g.chunk_grid, st, P.tile / 8, gb0, gb1, true);
      // footprints stay inside the band, so only its rows need clearing; tiles the gatherer
      // owns alone are stored in full by it and are skipped here
      if (oop) {
        const dim3 zg((unsigned)P.nty, (unsigned)((g.xhi - 1) / P.tile - g.xlo / P.tile + 1));
        k_zero_stale<R><<<zg, kBlock, 0, st>>>(k, g.xlo, g.xhi, npl * g.nb, g.task, g.counters,
                                               (uint32_t)(P.tile / 8), g.lastw, pass - 1u, plane);
        k_mark_tiles<<<grid1(k.ntiles), kBlock, 0, st>>>(k, g.task, g.counters, g.lastw, pass);
        if (g.dense)
          k_mark_dense<<<1024, kBlock, 0, st>>>(k, p0, npl, g.dtask, g.dplane, g.counters, g.lastw, pass);
      } else {
        k_zero_shared<R><<<gridcap(nband), kBlock, 0, st>>>(k, g.xlo, g.xhi, npl * g.nb, g.task, g.counters,
                                                           (uint32_t)(P.tile / 8), plane);
      }
      LAUNCH_CHECK();
    }
    {
      PROF("grid");
      {
#define HIPNG_GW(TT, QQ, BB)                                                                      \
  k_grid8<R, TT, QQ, BB, RG><<<g.npersist, kBlock, smw, st>>>(k, p0, npl, gb0, gb1, g.srec, g.rec_a,  \
                                                              (const rec3<R>*)g.rec_f, recv, g.task, \
                                                              g.tasktab, g.chunk_grid, g.counters, plane, \
                                                              g.split, g.strips, g.taps)
#define HIPNG_GWB(TT, QQ)                 \
  if (g.nb == 1) HIPNG_GW(TT, QQ, 1);     \
  else if (g.nb == 2) HIPNG_GW(TT, QQ, 2); \
  else if (g.nb == 3) HIPNG_GW(TT, QQ, 3); \
  else HIPNG_GW(TT, QQ, 4)
        if (P.tile == 32) {
          if (PCn == 2) { HIPNG_GWB(32, 2); } else { HIPNG_GWB(32, 1); }
        } else {
          if (PCn == 2) { HIPNG_GWB(16, 2); } else { HIPNG_GWB(16, 1); }
        }
#undef HIPNG_GWB
#undef HIPNG_GW
        LAUNCH_CHECK();
      }
    }
    if (g.dense) {
      PROF("grid_dense");
      launch_grid_dense(g, k, p0, npl, (const rec3<R>*)g.rec_f, recv, plane, st);
      LAUNCH_CHECK();
    }
    {
      PROF("fft_c2c_inverse");
      for (int i = 0; i < npl * g.nb; ++i) {
        const int64_t off = ((int64_t)i * P.gx + g.xlo) * P.gy;
        if (oop) geom_fft(g, g.fft_b_inv_oop, plane + off, st, bplane + off);
        else geom_fft(g, g.fft_b_inv, plane + off, st);
      }
    }
    if constexpr (sizeof(R) == 8) {
      if (fs) {
        const auto* tw = reinterpret_cast<const cplx<R>*>(g.fstw);
        const int n1 = 8 * g.fsm;
        {
          PROF("fourstep_1");
          fs_stage1<R>(g.fsm, false, k, g.xlo, g.xhi, npl * g.nb, bplane, tbuf, tw, st);
          LAUNCH_CHECK();
        }
        {
          PROF("fourstep_2");
          for (int q = 0; q < npl; ++q) {
            k_fs_inv2<R, A, F32C><<<dim3(n1 / 16, P.ny, g.nb), 256, 0, st>>>(
                k, n1, k.w0 + (p0 + q) * k.dw, tbuf + (int64_t)q * g.nb * P.ny * P.gx, tw, acc, p0 + q == 0 ? 1 : 0);
            LAUNCH_CHECK();
          }
        }
        continue;
      }
    }
    {
      PROF("transpose");
      if (!g.oop_ok) {
        // the transform along the grid rows may have overwritten T: clear it again
        const int64_t rows = (int64_t)g.pc * g.nb * P.ny;
        zero_cols(tbuf, rows, P.gx, 0, g.xlo, st);
        zero_cols(tbuf, rows, P.gx, g.xhi, P.gx, st);
      }
      const int xb = g.xlo & ~31;
      k_tr_in<R><<<dim3((g.xhi - xb + 31) / 32, (P.ny + 31) / 32, npl * g.nb), kBlock, 0, st>>>(
          k, g.xlo, g.xhi, xb, bplane, tbuf);
      LAUNCH_CHECK();
    }
    // FP32 planes of a group of kScreenGroup passes wait in their slots of T2, and the screen
    // adds them to the accumulator at once (the last pass flushes what is held)
    const int slot = (p0 / PCn) % ngrp;
    {
      PROF("fft_c2c_inverse");
      geom_fft(g, g.fft_t_inv, tbuf, st, tbuf2 + (int64_t)slot * PCn * g.nb * P.ny * P.gx);
    }
    if (slot == ngrp - 1 || p0 + npl >= P.nplanes) {
      PROF("screen_acc");
      const int s0 = p0 - slot * PCn;
      k_screen_acc_t<R, A, F32C><<<dim3(grid1(P.nx), P.ny), kBlock, 0, st>>>(k, s0, p0 + npl - s0, tbuf2, acc,
                                                                               s0 == 0 ? 1 : 0);
      LAUNCH_CHECK();
    }
  }
  PROF("finalize");
  if (cc && cc->image_yx) {
    const dim3 fg((P.nx + kBlock - 1) / kBlock, P.ny);
    if (cc->image_f32)
      k_finalize_yx<A, float><<<fg, kBlock, 0, st>>>(k, acc, (float*)dirty, g.tb.phix, g.tb.phiy, g.tb.cz, g.tb.cc,
                                                      g.tb.cn, cc->scale);
    else
      k_finalize_yx<A, double><<<fg, kBlock, 0, st>>>(k, acc, (double*)dirty, g.tb.phix, g.tb.phiy, g.tb.cz,
                                                       g.tb.cc, g.tb.cn, cc->scale);
  } else {
    k_finalize_t<A, O><<<dim3((P.nx + 31) / 32, (P.ny + 31) / 32), kBlock, 0, st>>>(
        k, acc, (O*)dirty, g.tb.phix, g.tb.phiy, g.tb.cz, g.tb.cc, g.tb.cn);
  }
  LAUNCH_CHECK();
}

template <typename R, typename A, bool F32C>
void exec_dirty2ms(const Geom& g, const void* dirty, const void* wgt, void* ms, hipStream_t st,
                   const ColCall* cc = nullptr) {
  using I = typename std::conditional<std::is_same<R, double>::value, double, float>::type;
  if (g.n == 0) return;
  if (!dirty || !ms) HIPNG_THROW(HIPNG_EINVAL, "dirty and ms must not be NULL");
  const KP& k = g.k;
  const hipng_plan& P = g.P;
  auto* dc = (A*)g.accimg;
  {
    PROF("prescale");
    if (cc && cc->image_yx) {
      const dim3 fg((P.nx + kBlock - 1) / kBlock, P.ny);
      if (cc->image_f32)
        k_prescale_yx<float, A><<<fg, kBlock, 0, st>>>(k, (const float*)dirty, dc, g.tb.phix, g.tb.phiy, g.tb.cz,
                                                        g.tb.cc, g.tb.cn);
      else
        k_prescale_yx<double, A><<<fg, kBlock, 0, st>>>(k, (const double*)dirty, dc, g.tb.phix, g.tb.phiy,
                                                         g.tb.cz, g.tb.cc, g.tb.cn);
    } else {
      k_prescale_t<I, A><<<dim3((P.nx + 31) / 32, (P.ny + 31) / 32), kBlock, 0, st>>>(
          k, (const I*)dirty, dc, g.tb.phix, g.tb.phiy, g.tb.cz, g.tb.cc, g.tb.cn);
    }
    LAUNCH_CHECK();
  }
  // the sorted accumulators need no clearing: the degridders store at each record's first
  // pass (the one holding its start plane) and add at the later ones
  auto* accs = (cplx<R>*)g.recv;
  auto* plane = (cplx<R>*)g.plane;
  auto* tbuf = (cplx<R>*)g.tbuf;
  auto* tbuf2 = (cplx<R>*)g.tbuf2;
  // The screens fill the image part of S (cleared once per call) and the transforms along
  // the grid rows go out of place into D. FP32 planes: S = T2 (kFillGroup passes of it), so
  // one read of the image fills the planes of a group of passes; FP64: S = T.
  const int ngrp = sizeof(R) == 4 ? kFillGroup : 1;
  cplx<R>* sbuf = ngrp > 1 ? tbuf2 : tbuf;
  cplx<R>* dbuf = ngrp > 1 ? tbuf : tbuf2;
  const int64_t np1 = (int64_t)g.nb * P.gx * P.gy;
  // band transforms out of place: the band rows of plane are cleared once and only their
  // image columns are rewritten per pass
  const bool oop = g.plane2 != nullptr && g.oop_ok;
  cplx<R>* bplane = oop ? (cplx<R>*)g.plane2 : plane;
  // FP64 planes of a length the four-step image-side transform covers: it writes the image
  // columns of the band rows itself (no S)
  const bool fs = sizeof(R) == 8 && g.fstw != nullptr;
  if (!fs) {
    PROF("transpose");
    // the fills rewrite the image columns of S; the others stay zero
    zero_cols(sbuf, (int64_t)ngrp * g.pc * g.nb * P.ny, P.gx, P.nx / 2, P.gx - P.nx / 2, st);
    LAUNCH_CHECK();
  }
  if (oop) {
    PROF("transpose");
    // k_tr_out rewrites the image columns of the band rows; the others stay zero
    for (int i = 0; i < g.pc * g.nb; ++i)
      zero_cols(plane + ((int64_t)i * P.gx + g.xlo) * P.gy, (int64_t)(g.xhi - g.xlo), P.gy, P.ny / 2,
                P.gy - P.ny / 2, st);
    LAUNCH_CHECK();
  }
  for (int p0 = 0; p0 < P.nplanes; p0 += g.pc) {
    const int npl = std::min(g.pc, P.nplanes - p0);
    const int slot = (p0 / g.pc) % ngrp;
    bool fs_done = false;
    if constexpr (sizeof(R) == 8) {
      if (fs) {
        const auto* tw = reinterpret_cast<const cplx<R>*>(g.fstw);
        const int n1 = 8 * g.fsm;
        {
          PROF("fourstep_2");
          for (int q = 0; q < npl; ++q) {
            k_fs_fwd2<R, A, F32C><<<dim3(n1 / 16, P.ny, g.nb), 256, 0, st>>>(
                k, n1, k.w0 + (p0 + q) * k.dw, dc, tbuf + (int64_t)q * g.nb * P.ny * P.gx, tw);
            LAUNCH_CHECK();
          }
        }
        {
          PROF("fourstep_1");
          if (!oop) {
            // the in-place band transforms of the previous pass wrote every column of the band
            // rows: clear the others again
            for (int i = 0; i < npl * g.nb; ++i)
              zero_cols(plane + ((int64_t)i * P.gx + g.xlo) * P.gy, (int64_t)(g.xhi - g.xlo), P.gy, P.ny / 2,
                        P.gy - P.ny / 2, st);
          }
          fs_stage1<R>(g.fsm, true, k, g.xlo, g.xhi, npl * g.nb, tbuf, plane, tw, st);
          LAUNCH_CHECK();
        }
        fs_done = true;
      }
    }
    if (!fs_done && slot == 0 && !g.oop_ok) {
      PROF("transpose");
      // the transforms along the grid rows may have overwritten S: clear it again
      zero_cols(sbuf, (int64_t)ngrp * g.pc * g.nb * P.ny, P.gx, P.nx / 2, P.gx - P.nx / 2, st);
      LAUNCH_CHECK();
    }
    if (!fs_done && slot == 0) {
      PROF("screen_fill");
      const int nfill = std::min(ngrp * g.pc, P.nplanes - p0);
      k_screen_fill_t<R, A, F32C><<<dim3(grid1(P.gx), P.ny), kBlock, 0, st>>>(k, p0, nfill, dc, sbuf);
      LAUNCH_CHECK();
    }
    if (!fs_done) {
      PROF("fft_c2c_forward");
      geom_fft(g, g.fft_t_fwd, sbuf + (int64_t)slot * g.pc * g.nb * P.ny * P.gx, st, dbuf);
    }
    if (!fs_done) {
      PROF("transpose");
      const dim3 trg((g.xhi - g.xlo + 31) / 32, (P.gy + 31) / 32, npl * g.nb);
      if (oop) k_tr_out<R, false><<<trg, kBlock, 0, st>>>(k, g.xlo, g.xhi, dbuf, plane);
      else k_tr_out<R, true><<<trg, kBlock, 0, st>>>(k, g.xlo, g.xhi, dbuf, plane);
      LAUNCH_CHECK();
    }
    {
      PROF("fft_c2c_forward");
      for (int i = 0; i < npl * g.nb; ++i) {
        const int64_t off = ((int64_t)i * P.gx + g.xlo) * P.gy;
        if (oop) geom_fft(g, g.fft_b_fwd_oop, plane + off, st, bplane + off);
        else geom_fft(g, g.fft_b_fwd, plane + off, st);
      }
    }
    const int dq = degrid_planes(g, npl);
    for (int q = 0; q < npl; q += dq) {
      const int nq = std::min(dq, npl - q);
      {
        PROF("degrid_setup");
        int slo, shi;
        chunk_range(k.Ww, P.nplanes, p0 + q, nq, &slo, &shi);
        if (degrid_split<R>(g)) build_tasks(g, slo, shi, 1, g.chunk_degrid, st, 1, g.sbstart, g.sbend);
        else build_tasks(g, slo, shi, 1, g.chunk_degrid, st);
      }
      {
        PROF("degrid");
        launch_degrid<R>(g, p0 + q, nq, bplane + q * np1, accs, st);
      }
      if (degrid_split<R>(g)) {
        PROF("degrid_dense");
        launch_degrid_dense<R>(g, p0 + q, nq, bplane + q * np1, accs, st);
      }
    }
  }
  PROF("unsort");
  if (cc) launch_unsort_cols<R>(g, *cc, accs, ms, st);
  else launch_unsort<R>(g, accs, (const R*)wgt, (cplx<R>*)ms, st);
  LAUNCH_CHECK();
}

void exec_dispatch(const Geom& g, bool invert, const void* in, const void* wgt, void* out, hipStream_t st,
                   const ColCall* cc = nullptr) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != g.device) HIPNG_THROW(HIPNG_EINVAL, "geometry was built on device %d, current device is %d", g.device, dev);
  if (g.prec == HIPNG_FP64) {
    if (invert) exec_ms2dirty<double, double, false>(g, in, wgt, out, st, cc);
    else exec_dirty2ms<double, double, false>(g, in, wgt, out, st, cc);
  } else if (g.prec == HIPNG_MIXED) {
    if (invert) exec_ms2dirty<float, double, false>(g, in, wgt, out, st, cc);
    else exec_dirty2ms<float, double, false>(g, in, wgt, out, st, cc);
  } else {
    if (invert) exec_ms2dirty<float, float, true>(g, in, wgt, out, st, cc);
    else exec_dirty2ms<float, float, true>(g, in, wgt, out, st, cc);
  }
}

// a validated ColCall for geometry g from the caller's column description
ColCall col_call(const Geom& g, const hipng_columns* c, bool invert) {
  if (!c) HIPNG_THROW(HIPNG_EINVAL, "columns must not be NULL");
  if (c->npol < 1 || c->npol > 4) HIPNG_THROW(HIPNG_EINVAL, "columns npol must be 1..4 (got %d)", c->npol);
  if (g.nb > c->npol)
    HIPNG_THROW(HIPNG_EINVAL, "the geometry's %d polarisations do not fit %d column polarisations", g.nb, c->npol);
  if (c->nchan < 1 || c->chan0 < 0 || c->chan0 + g.nchan > c->nchan)
    HIPNG_THROW(HIPNG_EINVAL, "geometry channels %d.. +%d outside the columns' %d", c->chan0, g.nchan, c->nchan);
  if ((c->vis_type != HIPNG_COMPLEX128 && c->vis_type != HIPNG_COMPLEX64) ||
      (c->weight_type != HIPNG_FLOAT64 && c->weight_type != HIPNG_FLOAT32) ||
      (c->flag_type != HIPNG_INT32 && c->flag_type != HIPNG_INT8))
    HIPNG_THROW(HIPNG_EINVAL, "unknown column element type");
  if (c->nant < 0 || (c->nant > 0 && (!c->gains || !c->ant1 || !c->ant2)))
    HIPNG_THROW(HIPNG_EINVAL, "gains need the gain array and both antenna index arrays");
  ColCall cc;
  std::memset(&cc.a, 0, sizeof(cc.a));
  cc.a.nchan_g = g.nchan;
  cc.a.rnchan_g = 1.0 / (double)std::max(g.nchan, 1);
  cc.a.nchan_t = c->nchan;
  cc.a.chan0 = c->chan0;
  cc.a.vis_f32 = c->vis_type == HIPNG_COMPLEX64;
  cc.a.wgt_f32 = c->weight_type == HIPNG_FLOAT32;
  cc.a.flag_i8 = c->flag_type == HIPNG_INT8;
  cc.a.shift = c->shift ? 1 : 0;
  cc.a.L = c->lmn[0];
  cc.a.M = c->lmn[1];
  cc.a.N = c->lmn[2];
  cc.a.nant = c->nant;
  const int nin = invert ? c->npol : g.nb, nout = invert ? g.nb : c->npol;
  for (int k = 0; k < nin * nout; ++k) {
    cc.a.cre[k] = c->conv[2 * k];
    cc.a.cim[k] = c->conv[2 * k + 1];
  }
  cc.npv = c->npol;
  cc.a.psf = c->psf ? 1 : 0;
  if (c->image_order != 0 && c->image_order != 1) HIPNG_THROW(HIPNG_EINVAL, "image_order must be 0 or 1");
  if (c->image_type != HIPNG_FLOAT64 && c->image_type != HIPNG_FLOAT32) HIPNG_THROW(HIPNG_EINVAL, "unknown image type");
  cc.image_yx = c->image_order == 1;
  cc.image_f32 = c->image_type == HIPNG_FLOAT32;
  if (!cc.image_yx && cc.image_f32 != (g.prec != HIPNG_FP64))
    HIPNG_THROW(HIPNG_EINVAL, "images in [npix_x][npix_y] order have the geometry's precision");
  for (int b = 0; b < 4; ++b) cc.scale.s[b] = (invert && cc.image_yx && b < g.nb) ? c->scale[b] : 1.0;
  if (c->nant > 0) {
    cc.gains = (const cplx<double>*)c->gains;
    cc.ant1 = c->ant1;
    cc.ant2 = c->ant2;
  }
  return cc;
}

// dense-run tasks and the visibilities they hold (device reduction, one synchronisation)
void dense_counts(const Geom& g, hipStream_t st, int64_t* ntask, int64_t* nvis) {
  *ntask = 0;
  *nvis = 0;
  if (!g.dense || g.n == 0) return;
  auto* d = (unsigned long long*)ws(ctx(), S_EXT, 2 * sizeof(unsigned long long));
  k_zero<unsigned long long><<<1, kBlock, 0, st>>>(d, 2);
  LAUNCH_CHECK();
  k_dense_stats<<<256, kBlock, 0, st>>>(g.dtask, g.counters, d);
  LAUNCH_CHECK();
  unsigned long long h[2] = {0ull, 0ull};
  HIP_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  *ntask = (int64_t)h[0];
  *nvis = (int64_t)h[1];
}

// (plane, tile) pairs that footprints touch (device reduction, one synchronisation)
int64_t occupied_tiles(const Geom& g, hipStream_t st) {
  if (g.n == 0) return 0;
  auto* d = (unsigned long long*)ws(ctx(), S_EXT, sizeof(unsigned long long));
  k_zero<unsigned long long><<<1, kBlock, 0, st>>>(d, 1);
  LAUNCH_CHECK();
  k_occupied<<<2048, kBlock, 0, st>>>(g.k, g.bstart, g.bend, d);
  LAUNCH_CHECK();
  unsigned long long h = 0ull;
  HIP_TRY(hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return (int64_t)h;
}

void fill_plan_out(const Geom& g, hipStream_t st, hipng_plan* out) {
  if (!out) return;
  *out = g.P;
  int64_t nt = 0;
  dense_counts(g, st, &nt, &out->dense_items);
  out->occupied_tiles = occupied_tiles(g, st);
}

template <typename F>
int32_t guarded(F&& f) {
  try {
    std::lock_guard<std::mutex> lk(g_mu);
    g_err.clear();
    f();
    if (g_prof) prof_flush();
    return HIPNG_OK;
  } catch (const HipngError& e) {
    g_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_err = std::string("internal error: ") + e.what();
    return HIPNG_ERUNTIME;
  }
}

}  // namespace

// =================================================================================
// C-ABI
// =================================================================================
namespace {
PolArgs pol_args(int64_t nrow, int32_t nchan, int32_t nin, int32_t nout, const double* conv, const double* lmn) {
  if (nrow < 0 || nchan <= 0 || nin < 1 || nin > 4 || nout < 1 || nout > 4) HIPNG_THROW(HIPNG_EINVAL, "bad sizes");
  if (!conv) HIPNG_THROW(HIPNG_EINVAL, "NULL conversion matrix");
  PolArgs a{};
  a.nrow = nrow;
  a.nchan = nchan;
  a.nin = nin;
  a.nout = nout;
  for (int k = 0; k < nin * nout; ++k) {
    a.cre[k] = conv[2 * k];
    a.cim[k] = conv[2 * k + 1];
  }
  a.shift = lmn ? 1 : 0;
  if (lmn) {
    a.l = lmn[0];
    a.m = lmn[1];
    a.nm1 = lmn[2];
  }
  return a;
}

GainArgs gain_args(int64_t ntimes, int32_t nbaselines, int32_t nchan, int32_t npol, int32_t ngain_times,
                   int32_t nant, int32_t gain_nchan, int32_t nrec) {
  if (ntimes < 0 || nbaselines < 0 || nchan <= 0 || ngain_times < 0 || nant <= 0 || gain_nchan <= 0)
    HIPNG_THROW(HIPNG_EINVAL, "bad sizes");
  if (npol != 1 && npol != 2 && npol != 4) HIPNG_THROW(HIPNG_EINVAL, "npol must be 1, 2 or 4");
  if (nrec != 1 && nrec != 2) HIPNG_THROW(HIPNG_EINVAL, "nrec must be 1 or 2");
  if (npol > 1 && nrec != 2) HIPNG_THROW(HIPNG_EINVAL, "npol %d needs 2x2 gains", npol);
  return GainArgs{ntimes, nbaselines, nchan, npol, gain_nchan, nant, nrec, ngain_times};
}

// prepared matrices (nrec x nrec complex128) and validity flags of every gain-table entry
size_t gain_scratch_bytes(const GainArgs& a) {
  const size_t nent = (size_t)a.ngain * a.nant * a.gnchan;
  return nent * ((size_t)a.nrec * a.nrec * sizeof(cplx<double>) + sizeof(int));
}

// k_gain_prep into `buf` (gain_scratch_bytes(a) bytes), then k_apply_gains; kernels only
void launch_gains(const GainArgs& a, void* vis, double* weight, const double* flags, const int8_t* time_flagged,
                  const int32_t* ant1, const int32_t* ant2, const void* gain, const double* gain_weight,
                  const int32_t* time_to_gain, int32_t inverse, void* buf, hipStream_t st) {
  if (!vis || !weight || !ant1 || !ant2 || !gain || !gain_weight || !time_to_gain)
    HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
  if (time_flagged && !flags) HIPNG_THROW(HIPNG_EINVAL, "time_flagged needs flags");
  const int64_t n = a.ntimes * a.nbl * (int64_t)a.nchan;
  const int64_t nent = (int64_t)a.ngain * a.nant * a.gnchan;
  const int n2 = a.nrec * a.nrec;
  auto* m = (cplx<double>*)buf;
  auto* valid = (int*)((char*)buf + nent * n2 * sizeof(cplx<double>));
  k_gain_prep<<<grid1(nent), kBlock, 0, st>>>((const cplx<double>*)gain, nent, a.nrec, inverse ? 1 : 0, m, valid);
  LAUNCH_CHECK();
  // (a version specialised on npol measured slower at the production shape: 39.9 against 35.9 ms)
  k_apply_gains<<<grid1(n), kBlock, 0, st>>>((cplx<double>*)vis, weight, flags, time_flagged, ant1, ant2, m,
                                             valid, gain_weight, time_to_gain, a);
  LAUNCH_CHECK();
}

// Scratch of the allocation-free gain entry point: the prepared matrices live in this array,
// one copy per GPU, held until the process ends. Host code uses the address
// hipGetSymbolAddress gives for the current device, never the array's name (a host-side
// stand-in, not a device address).
constexpr size_t kGainScratchBytes = (size_t)8 << 20;
__device__ __align__(256) unsigned char scratch[kGainScratchBytes];
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

// the device address of `scratch` on the current device (looked up on its first use there)
void* gain_scratch() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) HIPNG_THROW(HIPNG_ERUNTIME, "device index %d out of range", dev);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) HIPNG_THROW(HIPNG_ERUNTIME, "hipGetSymbolAddress returned NULL for the gain scratch");
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

double* upload_fc(DevCtx& c, const double* freq, int32_t nchan, hipStream_t st) {
  if (!freq) HIPNG_THROW(HIPNG_EINVAL, "NULL freq");
  std::vector<double> h(nchan);
  for (int i = 0; i < nchan; ++i) h[i] = freq[i] / kSpeedOfLight;
  auto* d = (double*)ws(c, S_SMALL, nchan * sizeof(double) + 64);
  HIP_TRY(hipMemcpyAsync(d, h.data(), nchan * sizeof(double), hipMemcpyHostToDevice, st));
  return d;
}
}  // namespace

extern "C" {

int32_t hipng_version(void) { return 100; }

const char* hipng_last_error(void) { return g_err.c_str(); }

int32_t hipng_plan_params(double umax, double vmax, double wmin, double wmax, int32_t npix_x,
                          int32_t npix_y, double pixsize_x, double pixsize_y, double epsilon,
                          int32_t do_wstacking, int32_t precision, int32_t nbatch, int64_t nvis,
                          const hipng_plan_opts* opts, hipng_plan* out) {
  return guarded([&] {
    if (!out) HIPNG_THROW(HIPNG_EINVAL, "out must not be NULL");
    make_plan(umax, vmax, wmin, wmax, npix_x, npix_y, pixsize_x, pixsize_y, epsilon, do_wstacking,
              precision, nbatch, nvis, opts, out);
    // the dense-run threshold the compute calls would use (0: that path is off)
    if (!(opts && opts->dense_min == -1) && out->W <= 8 && out->gx <= 32768 && out->gy <= 32768)
      out->dense_min = opts && opts->dense_min >= 2 ? opts->dense_min : (int32_t)kDenseMin;
  });
}

int32_t hipng_channel_scales(const hipng_plan* plan, const double* freq, int32_t nchan, double* sx,
                             double* sy, double* sw) {
  return guarded([&] {
    if (!plan || !freq || !sx || !sy || !sw) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    channel_scales(*plan, freq, nchan, sx, sy, sw, nullptr);
  });
}

int32_t hipng_uvw_extent(const double* uvw, int64_t nrow, const double* freq, int32_t nchan,
                         double* out, void* stream) {
  return guarded([&] {
    if (!uvw || !freq || !out || nchan <= 0 || nrow < 0) HIPNG_THROW(HIPNG_EINVAL, "bad argument");
    DevCtx& c = ctx();
    WsUse use(c, (hipStream_t)stream);
    extent(c, uvw, nrow, freq, nchan, out, (hipStream_t)stream);
  });
}

int32_t hipng_bucket_sort(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                          const hipng_plan* plan, const hipng_plan_opts* opts, uint32_t* keys_out,
                          uint32_t* perm_out, void* stream) {
  return guarded([&] {
    if (!freq || !plan || nchan <= 0 || nrow < 0) HIPNG_THROW(HIPNG_EINVAL, "bad argument");
    const int64_t n = nrow * (int64_t)nchan;
    if (n == 0) return;
    if (!uvw || !keys_out || !perm_out) HIPNG_THROW(HIPNG_EINVAL, "NULL device pointer");
    if (n >= 4294967295LL) HIPNG_THROW(HIPNG_EINVAL, "too many visibilities");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    Buffers B;
    B.borrow = &c;
    const KP k = make_kp(*plan, 1, nchan);
    Tables tb = upload_tables(B, *plan, freq, nchan, st);
    uint32_t *keys, *perm;
    sort_items(c, uvw, n, tb, k, *plan, !(opts && opts->sort_pack == -1), st, &keys, &perm);
    HIP_TRY(hipMemcpyAsync(keys_out, keys, n * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(perm_out, perm, n * 4, hipMemcpyDeviceToDevice, st));
  });
}

int32_t hipng_ms2dirty(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                       const void* ms, const void* wgt, int32_t nbatch, int32_t npix_x,
                       int32_t npix_y, double pixsize_x, double pixsize_y, double epsilon,
                       int32_t do_wstacking, int32_t precision, void* dirty, void* stream,
                       const hipng_plan_opts* opts, hipng_plan* plan_out) {
  return guarded([&] {
    if (!dirty) HIPNG_THROW(HIPNG_EINVAL, "dirty must not be NULL");
    hipStream_t st = (hipStream_t)stream;
    Geom g;
    DevCtx& c = ctx();
    WsUse use(c, st);
    build_geom(g, c, uvw, freq, nrow, nchan, nbatch, npix_x, npix_y, pixsize_x, pixsize_y,
               epsilon, do_wstacking, precision, opts, st, false);
    fill_plan_out(g, st, plan_out);
    exec_dispatch(g, true, ms, wgt, dirty, st);
  });
}

int32_t hipng_dirty2ms(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                       const void* dirty, const void* wgt, int32_t nbatch, int32_t npix_x,
                       int32_t npix_y, double pixsize_x, double pixsize_y, double epsilon,
                       int32_t do_wstacking, int32_t precision, void* ms, void* stream,
                       const hipng_plan_opts* opts, hipng_plan* plan_out) {
  return guarded([&] {
    hipStream_t st = (hipStream_t)stream;
    Geom g;
    DevCtx& c = ctx();
    WsUse use(c, st);
    build_geom(g, c, uvw, freq, nrow, nchan, nbatch, npix_x, npix_y, pixsize_x, pixsize_y,
               epsilon, do_wstacking, precision, opts, st, false);
    fill_plan_out(g, st, plan_out);
    exec_dispatch(g, false, dirty, wgt, ms, st);
  });
}

struct hipng_geometry_s {
  Geom g;
};

int32_t hipng_geometry_create(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                              int32_t nbatch, int32_t npix_x, int32_t npix_y, double pixsize_x,
                              double pixsize_y, double epsilon, int32_t do_wstacking,
                              int32_t precision, void* stream, const hipng_plan_opts* opts,
                              hipng_geometry* out, hipng_plan* plan_out) {
  return guarded([&] {
    if (!out) HIPNG_THROW(HIPNG_EINVAL, "out must not be NULL");
    *out = nullptr;
    auto* h = new hipng_geometry_s();
    DevCtx& c = ctx();
    WsUse use(c, (hipStream_t)stream);
    try {
      build_geom(h->g, c, uvw, freq, nrow, nchan, nbatch, npix_x, npix_y, pixsize_x, pixsize_y,
                 epsilon, do_wstacking, precision, opts, (hipStream_t)stream, true);
    } catch (...) {
      h->g.buf.release();
      delete h;
      throw;
    }
    fill_plan_out(h->g, (hipStream_t)stream, plan_out);
    *out = h;
  });
}

int32_t hipng_geometry_ms2dirty(hipng_geometry geom, const void* ms, const void* wgt, void* dirty,
                                void* stream) {
  return guarded([&] {
    if (!geom || !dirty) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    exec_dispatch(geom->g, true, ms, wgt, dirty, (hipStream_t)stream);
  });
}

int32_t hipng_geometry_dirty2ms(hipng_geometry geom, const void* dirty, const void* wgt, void* ms,
                                void* stream) {
  return guarded([&] {
    if (!geom) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    exec_dispatch(geom->g, false, dirty, wgt, ms, (hipStream_t)stream);
  });
}

int32_t hipng_geometry_invert_columns(hipng_geometry geom, const hipng_columns* cols, const void* vis,
                                      const void* weight, const void* flags, void* dirty, void* stream) {
  return guarded([&] {
    if (!geom || !dirty) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    ColCall cc = col_call(geom->g, cols, true);
    if (geom->g.n > 0 && !vis) HIPNG_THROW(HIPNG_EINVAL, "vis must not be NULL");
    cc.vis = vis;
    cc.wgt = weight;
    cc.flags = flags;
    exec_dispatch(geom->g, true, vis, nullptr, dirty, (hipStream_t)stream, &cc);
  });
}

int32_t hipng_geometry_predict_columns(hipng_geometry geom, const hipng_columns* cols, const void* dirty, void* vis,
                                       void* model_vis, void* stream) {
  return guarded([&] {
    if (!geom || !dirty || !vis) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    ColCall cc = col_call(geom->g, cols, false);
    cc.model = model_vis;
    exec_dispatch(geom->g, false, dirty, nullptr, vis, (hipStream_t)stream, &cc);
  });
}

int32_t hipng_weight_sums(const void* weight, const void* flags, int64_t nrow, int32_t nchan, int32_t npol,
                          int32_t weight_type, int32_t flag_type, double* out, void* stream) {
  return guarded([&] {
    if (!out || nrow < 0 || nchan < 1 || nchan > 65535 || npol < 1 || npol > 4)
      HIPNG_THROW(HIPNG_EINVAL, "bad arguments");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kSumBlocks, (nrow + kBlock - 1) / kBlock));
    const int64_t ncp = (int64_t)nchan * npol;
    auto* part = (double*)ws(c, S_IDFT, (size_t)ncp * nblk * sizeof(double) + 256);
    PROF("weight_sums");
    k_wsum_partial<<<dim3(nblk, nchan), kBlock, 0, st>>>(weight, flags, nrow, nchan, npol,
                                                         weight_type == HIPNG_FLOAT32, flag_type == HIPNG_INT8, part);
    LAUNCH_CHECK();
    k_wsum_final<<<grid1(ncp), kBlock, 0, st>>>(part, ncp, nblk, out);
    LAUNCH_CHECK();
  });
}

int32_t hipng_gain_gradient(const void* model_vis, const void* grad_vis, int32_t vis_type, const void* gains,
                            const int32_t* ant1, const int32_t* ant2, int64_t nrow, int32_t nchan, int32_t npol,
                            int32_t nant, void* grad_gains, void* stream) {
  return guarded([&] {
    if (nrow < 0 || nchan < 1 || npol < 1 || npol > 4 || nant < 1) HIPNG_THROW(HIPNG_EINVAL, "bad sizes");
    if (vis_type != HIPNG_COMPLEX128 && vis_type != HIPNG_COMPLEX64) HIPNG_THROW(HIPNG_EINVAL, "unknown vis type");
    if (!grad_gains || !gains || !ant1 || !ant2 || (nrow > 0 && (!model_vis || !grad_vis)))
      HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    PROF("gain_gradient");
    k_zero<double><<<gridcap(2 * (int64_t)nant * nchan), kBlock, 0, st>>>((double*)grad_gains, 2 * (int64_t)nant * nchan);
    LAUNCH_CHECK();
    if (nrow == 0) return;
    k_gain_grad<<<grid1(nrow * nchan), kBlock, 0, st>>>(model_vis, grad_vis, vis_type == HIPNG_COMPLEX64,
                                                        (const cplx<double>*)gains, ant1, ant2, nrow, nchan, npol,
                                                        (double*)grad_gains);
    LAUNCH_CHECK();
  });
}

int32_t hipng_geometry_stats(hipng_geometry geom, int64_t* out, int32_t nout) {
  return guarded([&] {
    if (!geom || !out || nout < 1) HIPNG_THROW(HIPNG_EINVAL, "geometry and output array required");
    const Geom& g = geom->g;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != g.device) HIPNG_THROW(HIPNG_EINVAL, "geometry was built on device %d, current device is %d", g.device, dev);
    int64_t v[5] = {g.n, 0, 0, g.dense ? 1 : 0, g.oop_ok ? 1 : 0};
    WsUse use(ctx(), nullptr);
    dense_counts(g, nullptr, &v[1], &v[2]);
    for (int i = 0; i < nout && i < 5; ++i) out[i] = v[i];
  });
}

int32_t hipng_geometry_destroy(hipng_geometry geom) {
  return guarded([&] {
    if (!geom) return;
    HIP_TRY(hipDeviceSynchronize());
    geom->g.buf.release();
    delete geom;
  });
}

int32_t hipng_dft(const double* direction_cosines, const void* flux, const double* uvw_lambda,
                  int32_t ncomp, int64_t ntimes_baselines, int32_t nchan, int32_t npol, void* vis,
                  void* stream) {
  return guarded([&] {
    if (!direction_cosines || !flux || !uvw_lambda || !vis) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    if (npol < 1 || npol > 4 || nchan < 1 || ncomp < 0 || ntimes_baselines < 0)
      HIPNG_THROW(HIPNG_EINVAL, "bad dft sizes");
    const int64_t nt = ntimes_baselines * nchan;
    if (nt == 0) return;
    hipStream_t st = (hipStream_t)stream;
    k_dft<<<grid1(nt), kBlock, 0, st>>>(direction_cosines, (const cplx<double>*)flux, uvw_lambda,
                                        ncomp, ntimes_baselines, nchan, npol, (cplx<double>*)vis);
    LAUNCH_CHECK();
  });
}

int32_t hipng_idft(const double* direction_cosines, const void* vis, const double* weight, const int32_t* flags,
                   const double* uvw_lambda, int32_t ncomp, int64_t ntimes_baselines, int32_t nchan, int32_t npol,
                   void* flux, double* weight_sum, void* stream) {
  return guarded([&] {
    if (npol < 1 || npol > 4 || nchan < 1 || ncomp < 0 || ntimes_baselines < 0)
      HIPNG_THROW(HIPNG_EINVAL, "bad idft sizes");
    if (ncomp == 0) return;
    if (!direction_cosines || !vis || !weight || !uvw_lambda || !flux || !weight_sum)
      HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    if ((int64_t)ncomp * nchan > 65535LL * 65535LL || ncomp > 65535 || nchan > 65535)
      HIPNG_THROW(HIPNG_EINVAL, "too many components or channels for one call");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    const int nblk = (int)std::max<int64_t>(1, std::min<int64_t>(kIdftBlocks, (ntimes_baselines + kBlock - 1) / kBlock));
    const int64_t np = (int64_t)ncomp * nchan * npol * nblk;
    auto* buf = (char*)ws(c, S_IDFT, np * sizeof(cplx<double>) + (size_t)nchan * npol * nblk * sizeof(double) + 256);
    auto* part = (cplx<double>*)buf;
    auto* wpart = (double*)(buf + np * sizeof(cplx<double>));
    PROF("idft");
    k_idft_partial<<<dim3(nblk, nchan, ncomp), kBlock, 0, st>>>(direction_cosines, (const cplx<double>*)vis, weight,
                                                                flags, uvw_lambda, ntimes_baselines, nchan, npol, part,
                                                                wpart);
    LAUNCH_CHECK();
    k_idft_final<<<grid1((int64_t)ncomp * nchan * npol), kBlock, 0, st>>>(part, wpart, ncomp, nchan, npol, nblk,
                                                                          (cplx<double>*)flux, weight_sum);
    LAUNCH_CHECK();
  });
}

int32_t hipng_apply_gaintable(void* vis, double* weight, const double* flags, const int8_t* time_flagged,
                              int64_t ntimes, int32_t nbaselines, int32_t nchan, int32_t npol,
                              const int32_t* ant1, const int32_t* ant2, const void* gain,
                              const double* gain_weight, int32_t ngain_times, int32_t nant,
                              int32_t gain_nchan, int32_t nrec, const int32_t* time_to_gain,
                              int32_t inverse, void* stream) {
  return guarded([&] {
    const GainArgs a = gain_args(ntimes, nbaselines, nchan, npol, ngain_times, nant, gain_nchan, nrec);
    if (a.ntimes * a.nbl * (int64_t)a.nchan == 0 || ngain_times == 0) return;
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    void* buf = ws(c, S_GAIN, gain_scratch_bytes(a) + 256);
    PROF("apply_gains");
    launch_gains(a, vis, weight, flags, time_flagged, ant1, ant2, gain, gain_weight, time_to_gain, inverse, buf, st);
  });
}

int32_t hipng_apply_gaintable_noalloc(void* vis, double* weight, const double* flags, const int8_t* time_flagged,
                                      int64_t ntimes, int32_t nbaselines, int32_t nchan, int32_t npol,
                                      const int32_t* ant1, const int32_t* ant2, const void* gain,
                                      const double* gain_weight, int32_t ngain_times, int32_t nant,
                                      int32_t gain_nchan, int32_t nrec, const int32_t* time_to_gain,
                                      int32_t inverse, void* stream) {
  return guarded([&] {
    const GainArgs a = gain_args(ntimes, nbaselines, nchan, npol, ngain_times, nant, gain_nchan, nrec);
    if (gain_scratch_bytes(a) > kGainScratchBytes)
      HIPNG_THROW(HIPNG_EINVAL,
                  "gain table of %d rows x %d antennas x %d channels needs %zu bytes of prepared matrices; "
                  "the allocation-free entry point holds at most %zu (use hipng_apply_gaintable)",
                  ngain_times, nant, gain_nchan, gain_scratch_bytes(a), kGainScratchBytes);
    if (a.ntimes * a.nbl * (int64_t)a.nchan == 0 || ngain_times == 0) return;
    launch_gains(a, vis, weight, flags, time_flagged, ant1, ant2, gain, gain_weight, time_to_gain, inverse,
                 gain_scratch(), (hipStream_t)stream);
  });
}

int32_t hipng_vis_prologue(const void* vis, const int32_t* flags, const double* imaging_weight, const double* uvw,
                           const double* freq, int64_t nrow, int32_t nchan, int32_t npol_vis, int32_t npol_image,
                           const double* conv, const double* lmn, int32_t precision, void* ms, void* wgt,
                           int32_t* nonzero, void* stream) {
  return guarded([&] {
    const PolArgs a = pol_args(nrow, nchan, npol_vis, npol_image, conv, lmn);
    if (npol_image > npol_vis) HIPNG_THROW(HIPNG_EINVAL, "more image than visibility polarisations");
    if (!vis || !imaging_weight || !ms || !wgt || !nonzero || (lmn && !uvw)) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    const double* fc = upload_fc(c, freq, nchan, st);
    k_zero<int><<<1, kBlock, 0, st>>>(nonzero, (int64_t)nchan * npol_image);
    LAUNCH_CHECK();
    const int64_t n = nrow * (int64_t)nchan;
    if (n == 0) return;
    PROF("vis_prologue");
    auto launch = [&](auto ni, auto no) {
      constexpr int NI = decltype(ni)::value, NO = decltype(no)::value;
      if (precision == HIPNG_FP64)
        k_vis_prologue<cplx<double>, double, NI, NO><<<grid1(n), kBlock, 0, st>>>(
            (const cplx<double>*)vis, flags, imaging_weight, uvw, fc, a, (cplx<double>*)ms, (double*)wgt, nonzero);
      else
        k_vis_prologue<cplx<float>, float, NI, NO><<<grid1(n), kBlock, 0, st>>>(
            (const cplx<double>*)vis, flags, imaging_weight, uvw, fc, a, (cplx<float>*)ms, (float*)wgt, nonzero);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    const int key = npol_vis * 8 + npol_image;
    if (key == 8 + 1) launch(I1{}, I1{});
    else if (key == 16 + 1) launch(I2{}, I1{});
    else if (key == 16 + 2) launch(I2{}, I2{});
    else if (key == 32 + 1) launch(I4{}, I1{});
    else if (key == 32 + 2) launch(I4{}, I2{});
    else if (key == 32 + 4) launch(I4{}, I4{});
    else HIPNG_THROW(HIPNG_EINVAL, "polarisations %d -> %d not supported", npol_vis, npol_image);
    LAUNCH_CHECK();
  });
}

int32_t hipng_vis_epilogue(const void* vist, const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                           int32_t npol_image, int32_t npol_vis, const double* conv, const double* lmn,
                           int32_t precision, void* vis, void* stream) {
  return guarded([&] {
    const PolArgs a = pol_args(nrow, nchan, npol_image, npol_vis, conv, lmn);
    if (!vist || !vis || (lmn && !uvw)) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    const double* fc = upload_fc(c, freq, nchan, st);
    const int64_t n = nrow * (int64_t)nchan;
    if (n == 0) return;
    PROF("vis_epilogue");
    auto launch = [&](auto ni, auto no) {
      constexpr int NI = decltype(ni)::value, NO = decltype(no)::value;
      if (precision == HIPNG_FP64)
        k_vis_epilogue<cplx<double>, NI, NO><<<grid1(n), kBlock, 0, st>>>((const cplx<double>*)vist, uvw, fc, a,
                                                                          (cplx<double>*)vis);
      else
        k_vis_epilogue<cplx<float>, NI, NO><<<grid1(n), kBlock, 0, st>>>((const cplx<float>*)vist, uvw, fc, a,
                                                                         (cplx<double>*)vis);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I4 = std::integral_constant<int, 4>;
    const int key = npol_image * 8 + npol_vis;
    if (key == 8 + 1) launch(I1{}, I1{});
    else if (key == 8 + 2) launch(I1{}, I2{});
    else if (key == 8 + 4) launch(I1{}, I4{});
    else if (key == 16 + 2) launch(I2{}, I2{});
    else if (key == 16 + 4) launch(I2{}, I4{});
    else if (key == 32 + 4) launch(I4{}, I4{});
    else if (key == 32 + 1) launch(I4{}, I1{});
    else if (key == 16 + 1) launch(I2{}, I1{});
    else if (key == 32 + 2) launch(I4{}, I2{});
    else HIPNG_THROW(HIPNG_EINVAL, "polarisations %d -> %d not supported", npol_image, npol_vis);
    LAUNCH_CHECK();
  });
}

int32_t hipng_profile_enable(int32_t on) {
  return guarded([&] {
    prof_flush();
    g_profacc.clear();
    g_prof = on != 0;
  });
}

int32_t hipng_profile_query(int32_t idx, char* name, int32_t name_len, double* total_ms,
                            int64_t* launches) {
  return guarded([&] {
    if (idx < 0 || idx >= (int32_t)g_profacc.size()) HIPNG_THROW(HIPNG_EINVAL, "no profile entry %d", idx);
    const auto& e = g_profacc[idx];
    if (name && name_len > 0) {
      std::strncpy(name, e.first.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = e.second.first;
    if (launches) *launches = e.second.second;
  });
}

int32_t hipng_weight_visibility(const double* uvw, const double* freq, int64_t ntimes, int64_t nbaselines,
                                int32_t nchan, int32_t npol, const int32_t* vis_to_im, int32_t nimchan,
                                int32_t npix_x, int32_t npix_y, double cell_u, double cell_v, int32_t weighting,
                                double robustness, double* imaging_weight, double* grid, void* stream) {
  return guarded([&] {
    if (!freq || !vis_to_im || !imaging_weight || !grid) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    if (ntimes < 0 || nbaselines < 0 || nchan <= 0 || npol <= 0 || nimchan <= 0 || npix_x <= 0 || npix_y <= 0)
      HIPNG_THROW(HIPNG_EINVAL, "bad sizes");
    if (npol > kWgtMaxPol) HIPNG_THROW(HIPNG_EINVAL, "npol %d > %d", npol, kWgtMaxPol);
    if (weighting != 1 && weighting != 2) HIPNG_THROW(HIPNG_EINVAL, "weighting must be 1 (uniform) or 2 (robust)");
    hipStream_t st = (hipStream_t)stream;
    DevCtx& c = ctx();
    WsUse use(c, st);
    const int64_t n = ntimes * nbaselines * (int64_t)nchan;
    const int64_t ng = (int64_t)nimchan * npol * npix_x * npix_y;
    if (n && !uvw) HIPNG_THROW(HIPNG_EINVAL, "NULL argument");
    std::vector<double> hfc(nchan);
    std::vector<int> hmap(nchan);
    for (int i = 0; i < nchan; ++i) {
      hfc[i] = freq[i] / kSpeedOfLight;
      hmap[i] = vis_to_im[i];
      if (hmap[i] < 0 || hmap[i] >= nimchan) HIPNG_THROW(HIPNG_EINVAL, "vis_to_im out of range");
    }
    auto* small = (char*)ws(c, S_SMALL, nchan * (sizeof(double) + sizeof(int)) + 64);
    auto* dfc = (double*)small;
    auto* dmap = (int*)(small + nchan * sizeof(double));
    auto* red = (double*)ws(c, S_EXT, 4 * sizeof(double));
    HIP_TRY(hipMemcpyAsync(dfc, hfc.data(), nchan * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(dmap, hmap.data(), nchan * sizeof(int), hipMemcpyHostToDevice, st));
    k_zero<double><<<gridcap(ng), kBlock, 0, st>>>(grid, ng);
    k_zero<double><<<1, kBlock, 0, st>>>(red, 4);
    LAUNCH_CHECK();
    WgtArgs a{ntimes, nbaselines, nchan, npol, nimchan, npix_x, npix_y, cell_u, cell_v, 0, 0, 0};
    a.ch_lanes = std::min(nchan, 64 / npol);  // visibilities per wave: 64 / npol
    a.bl_per_wave = 64 / npol / a.ch_lanes;
    const int64_t ncb = (nchan + a.ch_lanes - 1) / a.ch_lanes;
    const int64_t nbg = (nbaselines + a.bl_per_wave - 1) / a.bl_per_wave;
    // whole time tracks per work unit unless that leaves fewer than 64k units
    int64_t tseg = std::max<int64_t>(ntimes, 1);
    while (tseg > 1 && ncb * nbg * ((ntimes + tseg - 1) / tseg) < 65536) tseg = (tseg + 1) / 2;
    a.tseg = (int)std::min<int64_t>(tseg, 1 << 30);
    const int64_t nunits = ncb * nbg * ((ntimes + a.tseg - 1) / a.tseg);
    if (n) {
      PROF("weight_grid");
      const int64_t blocks = std::min<int64_t>((nunits + 3) / 4, 2048);  // 4 waves per block
      k_weight_grid<<<(unsigned)blocks, kBlock, 0, st>>>(uvw, dfc, dmap, a, imaging_weight, grid, red,
                                                         (unsigned*)(red + 3), nunits);
      LAUNCH_CHECK();
    }
    if (weighting == 2) {
      PROF("weight_sums");
      k_sum2<<<std::min<int64_t>(1024, gridcap(ng)), kBlock, 0, st>>>(grid, ng, nullptr, 0, red);
      LAUNCH_CHECK();
    }
    double h[4];
    HIP_TRY(hipMemcpyAsync(h, red, sizeof(h), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    unsigned fl;
    std::memcpy(&fl, &h[3], sizeof(fl));
    if (fl) HIPNG_THROW(HIPNG_EINVAL, "visibilities fall outside the uv grid: cellsize is too large");
    // Briggs' robust parameter with RASCIL's normalisation (equation 3.15 of the thesis)
    const double f2 = weighting == 2 ? std::pow(5.0 * std::pow(10.0, -robustness), 2.0) * h[1] / h[0] : 0.0;
    if (n) {
      PROF("weight_apply");
      k_weight_apply<<<gridcap(n * npol), kBlock, 0, st>>>(uvw, dfc, dmap, a, grid, weighting, f2, imaging_weight);
      LAUNCH_CHECK();
    }
  });
}

int32_t hipng_release(void) {
  return guarded([&] {
    for (auto& kv : g_ctx) {
      HIP_TRY(hipSetDevice(kv.first));
      HIP_TRY(hipDeviceSynchronize());
      for (int s = 0; s < S_NSLOTS; ++s)
        if (kv.second.buf[s]) (void)hipFree(kv.second.buf[s]);
      for (auto& pl : kv.second.plans) rocfft_plan_destroy(pl.second);
      if (kv.second.ws_ev) (void)hipEventDestroy(kv.second.ws_ev);
    }
    g_ctx.clear();
  });
}

}  // extern "C"
