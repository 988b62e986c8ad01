/*
 * hipng — MI355X (gfx950) w-stacking imaging core behind RASCIL's invert_ng / predict_ng.
 *
 * C-ABI boundary. Plain pointers and sizes only. Large arrays are DEVICE pointers
 * (HBM-resident, e.g. torch tensors' data_ptr()); small per-channel arrays are host
 * pointers. Every entry point returns 0 on success and a negative code on failure;
 * hipng_last_error() then describes the failure (thread-local).
 *
 * The two compute entry points replace the ducc0.wgridder calls RASCIL makes:
 *   hipng_ms2dirty  <- ducc0.wgridder.ms2dirty  (rascil/processing_components/imaging/ng.py:232-248, 259-275)
 *   hipng_dirty2ms  <- ducc0.wgridder.dirty2ms  (rascil/processing_components/imaging/ng.py:91-103, 108-120)
 * with the ducc0 0.24 argument meaning: uvw [nrow][3] in metres, freq [nchan] in Hz,
 * pixel sizes in radians, image axes [npix_x][npix_y] (x = RA-like first axis),
 * x = (ix - npix_x/2) * pixsize_x, y = (iy - npix_y/2) * pixsize_y and
 *   dirty(x,y) = sum_{row,ch} wgt * Re[ ms * exp(+2 pi i f/c (u x + v y - w (n-1))) ] / n
 *   ms(row,ch) = wgt * sum_{x,y} dirty(x,y)/n * exp(-2 pi i f/c (u x + v y - w (n-1)))
 * where n = sqrt(1 - x^2 - y^2) with w-stacking and n = 1 (no w term) without it.
 * Extension over ducc0: a batch of `nbatch` independent images (e.g. polarisations)
 * that share uvw/freq is processed in one call; ms and wgt are then laid out
 * [nrow][nchan][nbatch] (nbatch = 1 is exactly the ducc0 layout) and dirty
 * [nbatch][npix_x][npix_y].
 *
 * Precision modes (`precision`):
 *   HIPNG_FP64  : complex128 ms, float64 wgt/dirty; everything in FP64.
 *   HIPNG_MIXED : complex64 ms, float32 wgt/dirty; uvw->grid scaling, the w-plane
 *                 phase (n-1)w and the grid-correction divide in FP64; kernel
 *                 evaluation, tile accumulation and the FFT in FP32.
 *   HIPNG_FP32  : complex64 ms, float32 wgt/dirty; everything in FP32.
 */
#ifndef HIPNG_H
#define HIPNG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { HIPNG_FP64 = 0, HIPNG_MIXED = 1, HIPNG_FP32 = 2 };

enum {
  HIPNG_OK = 0,
  HIPNG_EINVAL = -1,  /* invalid argument (maps to ValueError) */
  HIPNG_ERUNTIME = -2 /* HIP / rocFFT / allocation failure (maps to RuntimeError) */
};

/* Parameters of one gridding job: chosen by hipng_plan_params, reported back by the
 * compute calls. Units: u/v/w extents in wavelengths, pixel sizes in radians. */
typedef struct hipng_plan {
  int32_t precision;  /* HIPNG_FP64 / HIPNG_MIXED / HIPNG_FP32 */
  int32_t wstacking;  /* 1: improved w-stacking with 1/n, 0: plain 2-D */
  int32_t W;          /* kernel support in cells (u, v and w) */
  int32_t tile;       /* uv tile side in cells */
  int32_t nx, ny;     /* image size */
  int32_t gx, gy;     /* padded grid size (multiples of tile) */
  int32_t ntx, nty;   /* tiles per grid side */
  int32_t nplanes;    /* number of w planes (1 without w-stacking) */
  int32_t keybits;    /* significant bits of the (plane, tile) bucket key */
  double beta;        /* ES kernel exp(beta (sqrt(1-z^2) - 1)) */
  double sigma;       /* nominal oversampling factor */
  double pixsize_x, pixsize_y;
  double w0, dw;      /* plane p sits at w = w0 + p dw (wavelengths) */
  double nshift;      /* (n-1) centre: phase exp(-2 pi i w nshift) is applied per visibility */
  double umax, vmax, wmin, wmax; /* |u|,|v| maxima and |w| range, wavelengths */
  double epsilon;     /* requested accuracy */
  double est_error;   /* modelled kernel error */
  int32_t dense_min;  /* shortest run for the dense-run gridder, 0 if that path is off */
  int32_t reserved;
  int64_t dense_items; /* visibilities (rows x channels) on the dense-run path; filled in
                          plan_out by the compute calls (one extra device synchronisation) */
  int64_t occupied_tiles; /* (w plane, uv tile) pairs some visibility footprint touches, i.e.
                             the grid cells (x tile^2) gridding must write; filled like
                             dense_items */
} hipng_plan;

/* Plan options: every field 0 means "automatic" (the library's own choice). They pin a plan
 * for tests, studies and autotuning without touching the process environment; a pinned plan
 * still has to reach epsilon in its precision mode (HIPNG_EINVAL otherwise). Pass NULL for
 * all-automatic. No counterpart in ducc0 0.24 (which has no plan overrides). */
typedef struct hipng_plan_opts {
  double sigma;         /* oversampling factor: 1.25, 1.5, 1.75, 2 or 2.5 */
  int32_t W;            /* kernel support 4..16 (at least what epsilon needs) */
  int32_t tile;         /* uv tile side: 16 or 32 */
  int32_t dense_min;    /* shortest run the dense-run path takes (>= 2); -1: path off */
  int32_t chunk_grid;   /* candidates per gridding task (rounded down to a multiple of 256) */
  int32_t chunk_degrid; /* visibilities per degridding task (likewise) */
  int32_t sort_pack;    /* -1: recompute the bucket keys after the digit passes (k_rekey) */
  int32_t use_extent;   /* 1: plan for at least the extent below (the union with the data's) */
  int32_t reserved;
  double umax, vmax, wmin, wmax; /* wavelengths, after the w >= 0 hermitian flip */
} hipng_plan_opts;

/* Library version (major*10000 + minor*100 + patch). */
int32_t hipng_version(void);

/* Description of the last error on this thread ("" if none). */
const char* hipng_last_error(void);

/* Host-only: choose W, beta, oversampling, grid, tile and w planes for the given uvw
 * extents (wavelengths, after the w >= 0 hermitian flip), image and number of
 * visibilities nvis = nrow * nchan (the cost model weighs gridding against FFTs).
 * The compute calls make the same choice internally. No GPU use.
 * Accuracy contract (like ducc0, which rejects an epsilon its data type cannot reach):
 * HIPNG_EINVAL unless epsilon >= 1e-13 (HIPNG_FP64), >= 1e-5 (HIPNG_MIXED: FP32 kernel,
 * grid and FFT) or, for HIPNG_FP32, >= max(1e-5, the phase error of FP32 grid coordinates
 * at the image edge, 2 pi (2 |u| + ulp(u_max)/2) (npix/2) / grid, in cells), so that an
 * accepted epsilon is met. est_error in the plan is the modelled total error. */
int32_t hipng_plan_params(double umax, double vmax, double wmin, double wmax,
                          int32_t npix_x, int32_t npix_y, double pixsize_x, double pixsize_y,
                          double epsilon, int32_t do_wstacking, int32_t precision,
                          int32_t nbatch, int64_t nvis, const hipng_plan_opts* opts, hipng_plan* out);

/* Host-only: per-channel scale factors the coordinate kernel uses (bit-exact contract
 * with the CPU reference sort): sx[c] = f_c/c0 * pixsize_x * gx, sy[c] likewise,
 * sw[c] = f_c/c0 / dw (0 without w-stacking). Arrays of nchan doubles. */
int32_t hipng_channel_scales(const hipng_plan* plan, const double* freq, int32_t nchan,
                             double* sx, double* sy, double* sw);

/* Device: extent of the uvw set in wavelengths: out[0]=max|u|, out[1]=max|v|,
 * out[2]=min|w|, out[3]=max|w| over all rows and channels. out is a host array. */
int32_t hipng_uvw_extent(const double* uvw, int64_t nrow, const double* freq, int32_t nchan,
                         double* out, void* stream);

/* Device: bucket sort of the nrow*nchan visibilities by (w plane, uv tile) for `plan`,
 * and inside a bucket by the in-tile origin digit (first cell mod tile along u) * tile +
 * (first cell mod tile along v). keys_out[j] is the bucket key of the j-th visibility in
 * sorted order, perm_out[j] its original index row*nchan + chan. Stable: equal
 * (key, digit) pairs keep increasing original index. */
int32_t hipng_bucket_sort(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                          const hipng_plan* plan, const hipng_plan_opts* opts, uint32_t* keys_out,
                          uint32_t* perm_out, void* stream);

/* Device: adjoint NUFFT with w-stacking (ducc0.wgridder.ms2dirty semantics, see above).
 * uvw [nrow][3] f64 (device); freq [nchan] f64 (host); ms [nrow][nchan][nbatch] complex
 * (c16 for HIPNG_FP64, c8 otherwise); wgt same layout, real, or NULL; dirty out
 * [nbatch][npix_x][npix_y] real. opts (host, nullable) pins plan choices; plan_out (host,
 * nullable) receives the parameters used. */
int32_t hipng_ms2dirty(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                       const void* ms, const void* wgt, int32_t nbatch,
                       int32_t npix_x, int32_t npix_y, double pixsize_x, double pixsize_y,
                       double epsilon, int32_t do_wstacking, int32_t precision,
                       void* dirty, void* stream, const hipng_plan_opts* opts, hipng_plan* plan_out);

/* Device: forward NUFFT with w-stacking (ducc0.wgridder.dirty2ms semantics).
 * dirty [nbatch][npix_x][npix_y] real in; ms out [nrow][nchan][nbatch] complex; wgt as for
 * hipng_ms2dirty (multiplies the output) or NULL. */
int32_t hipng_dirty2ms(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                       const void* dirty, const void* wgt, int32_t nbatch,
                       int32_t npix_x, int32_t npix_y, double pixsize_x, double pixsize_y,
                       double epsilon, int32_t do_wstacking, int32_t precision,
                       void* ms, void* stream, const hipng_plan_opts* opts, hipng_plan* plan_out);

/* Reusable geometry ("plan once, execute many"): everything that depends only on uvw, freq,
 * the image, epsilon, precision and nbatch (plan, bucket sort, coordinate records, FFT plans,
 * all device buffers) is built once by hipng_geometry_create (which allocates and
 * synchronises); hipng_geometry_ms2dirty / _dirty2ms then only launch kernels on `stream`
 * (no allocation, no host synchronisation, so they can be captured in a hipGraph) and give the
 * same results as hipng_ms2dirty / hipng_dirty2ms with the same arguments. The geometry keeps
 * its own copy of uvw. No counterpart in ducc0 0.24; it serves callers that repeat
 * invert/predict on one visibility set (major cycles, gradient descent, graph replay). */
typedef struct hipng_geometry_s* hipng_geometry;
int32_t hipng_geometry_create(const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                              int32_t nbatch, int32_t npix_x, int32_t npix_y, double pixsize_x,
                              double pixsize_y, double epsilon, int32_t do_wstacking,
                              int32_t precision, void* stream, const hipng_plan_opts* opts,
                              hipng_geometry* out, hipng_plan* plan_out);
int32_t hipng_geometry_ms2dirty(hipng_geometry geom, const void* ms, const void* wgt, void* dirty,
                                void* stream);
int32_t hipng_geometry_dirty2ms(hipng_geometry geom, const void* dirty, const void* wgt, void* ms,
                                void* stream);
int32_t hipng_geometry_destroy(hipng_geometry geom);

/* Visibility columns in the BlockVisibility layout, read or written by the gridder itself:
 * vis / weight / flags [nrow][nchan][npol] (row = time x baseline), the geometry covering
 * column channels chan0 .. chan0 + (its nchan) - 1. The passes RASCIL's invert_ng / predict_ng
 * run on the host around ducc0 (ng.py:56, 122-133, 169-190: copy, phase shift to the image
 * centre, flag masking, polarisation conversion, casts) and the gain product of
 * apply_gaintable (calibration/operations.py:135-137, V = g1 conj(g2) M, per-antenna scalar
 * gains) are fused into the gridder's value gather (invert) and its unsort (predict):
 *   invert:  x_p = vis_p (1 - f_p) [e^{+2 pi i f/c (u L + v M + w N)}] [conj(g1 conj(g2))],
 *            ms_q = sum_p conv[q][p] x_p, wgt_q = weight_q (1 - f_q) (weight NULL: 1), q < image pols;
 *   predict: y_p = sum_q conv[p][q] vis_q [e^{-2 pi i f/c (u L + v M + w N)}],
 *            model_vis_p = y_p, vis_p = y_p [g1 conj(g2)],
 * with (u, v, w) the geometry's own uvw (so a caller that negates u and w for the gridder folds
 * that into L, M, N), g [nant][nchan] complex128 and ant1 / ant2 [nrow] (device). */
enum { HIPNG_COMPLEX128 = 0, HIPNG_COMPLEX64 = 1 };
enum { HIPNG_FLOAT64 = 0, HIPNG_FLOAT32 = 1 };
enum { HIPNG_INT32 = 0, HIPNG_INT8 = 1 };
typedef struct hipng_columns {
  int32_t nchan;        /* channels of the columns */
  int32_t chan0;        /* column channel of the geometry's first channel */
  int32_t npol;         /* polarisations of the columns (>= the geometry's nbatch) */
  int32_t vis_type;     /* HIPNG_COMPLEX128 / HIPNG_COMPLEX64 */
  int32_t weight_type;  /* HIPNG_FLOAT64 / HIPNG_FLOAT32 */
  int32_t flag_type;    /* HIPNG_INT32 / HIPNG_INT8 */
  int32_t shift;        /* 1: apply the phase with the coefficients lmn */
  int32_t nant;         /* antennas of gains (0: no gains) */
  double lmn[3];        /* L, M, N */
  double conv[32];      /* [npol_out][npol_in] complex as (re, im) pairs: columns -> image
                           (invert, out = the geometry's nbatch), image -> columns (predict) */
  int32_t psf;          /* invert: image polarisation 0 gets ms = 1, the others 0 (the weights
                           still apply): the point spread function of invert_ng (ng.py:221-223) */
  int32_t image_order;  /* 0: image arrays [nbatch][npix_x][npix_y] (ducc0), 1: [nbatch][npix_y][npix_x]
                           (RASCIL's pixel order, no transpose) */
  int32_t image_type;   /* HIPNG_FLOAT64 / HIPNG_FLOAT32 (order 0: the geometry's precision) */
  int32_t reserved;
  double scale[4];      /* invert, order 1: factor per image polarisation (1: none; 1 / sumwt
                           normalises, 0 zeroes); order 0 and predict ignore it */
  const void* gains;    /* device [nant][nchan] complex128, or NULL */
  const int32_t* ant1;  /* device [nrow] */
  const int32_t* ant2;  /* device [nrow] */
} hipng_columns;

/* Geometry invert from columns (no counterpart in ducc0: it replaces ng.py:169-190 + ms2dirty
 * ng.py:232-248 on the device): dirty [nbatch][npix_x][npix_y] as hipng_geometry_ms2dirty of
 * the ms / wgt above. Kernels only (graph capturable). */
int32_t hipng_geometry_invert_columns(hipng_geometry geom, const hipng_columns* cols, const void* vis,
                                      const void* weight, const void* flags, void* dirty, void* stream);

/* Geometry predict into columns (dirty2ms ng.py:91-103 + ng.py:122-133 on the device): writes
 * vis (and model_vis, nullable: the visibilities before the gains) for the geometry's
 * channels. Kernels only (graph capturable). */
int32_t hipng_geometry_predict_columns(hipng_geometry geom, const hipng_columns* cols, const void* dirty, void* vis,
                                       void* model_vis, void* stream);

/* Device: out[c][p] = sum over rows of weight (1 - flags) for columns [nrow][nchan][npol]
 * (weight NULL: 1, flags NULL: 0), FP64, fixed summation order: the sumwt of invert_ng
 * (ng.py:178-190 sums flagged imaging weights per channel and polarisation). */
int32_t hipng_weight_sums(const void* weight, const void* flags, int64_t nrow, int32_t nchan, int32_t npol,
                          int32_t weight_type, int32_t flag_type, double* out, void* stream);

/* Device: gradient of a real loss with respect to per-antenna scalar gains through
 * V = g[ant1] conj(g[ant2]) M (calibration/operations.py:135-137): grad_vis = dL/dV in torch's
 * complex convention, model_vis = M, both [nrow][nchan][npol] of vis_type; grad_gains
 * [nant][nchan] complex128 = sum over the rows of each antenna of G conj(M) g[other] (as ant1)
 * and conj(G conj(M)) g[other] (as ant2), summed over polarisations. No counterpart in the
 * reference (it has no gradients); FP64 atomics, so the summation order varies. */
int32_t hipng_gain_gradient(const void* model_vis, const void* grad_vis, int32_t vis_type, const void* gains,
                            const int32_t* ant1, const int32_t* ant2, int64_t nrow, int32_t nchan, int32_t npol,
                            int32_t nant, void* grad_gains, void* stream);

/* Diagnostics (no counterpart in the reference): out[0] = visibilities (rows x channels),
 * out[1] = dense-run tasks, out[2] = visibilities gridded by the dense-run path, out[3] = 1
 * if that path is enabled for this geometry (W <= 8, grid <= 32768 per axis), out[4] = 1 if
 * every out-of-place rocFFT plan of the geometry was found to leave its input untouched
 * (checked once per plan on a pattern; 0: the buffers they read are cleared before every
 * pass instead of once per call). Writes min(nout, 5) entries; synchronises the device. */
int32_t hipng_geometry_stats(hipng_geometry geom, int64_t* out, int32_t nout);

/* Device: FP64 point-source DFT, replacing dft_kernel / dft_cpu_looped
 * (rascil/processing_components/imaging/dft.py:123-178, 261-281; CUDA precedent
 * dft.py:181-258): vis[t][b][c][p] = sum_k flux[k][c][p] exp(-2 pi i (u l_k + v m_k + w (n_k-1)))
 * with uvw_lambda [t][b][c][3], direction_cosines [k][3] = (l, m, n-1), flux [k][c][p]
 * complex128, vis complex128 out. All device pointers. */
int32_t hipng_dft(const double* direction_cosines, const void* flux, const double* uvw_lambda,
                  int32_t ncomp, int64_t ntimes_baselines, int32_t nchan, int32_t npol,
                  void* vis, void* stream);

/* Device: FP64 inverse DFT of point components, replacing idft_visibility_skycomponent
 * (rascil/processing_components/imaging/dft.py:335-380): per component k, channel c and
 * polarisation p, flux[k][c][p] = sum_{t,b} w (1 - f) vis exp(+2 pi i (u l_k + v m_k + w (n_k-1)))
 * / sum_{t,b} w (1 - f), 0 where that weight sum is not positive; weight_sum[c][p] receives the
 * sum. vis complex128, weight f64 and flags int32 (or NULL) [t][b][c][p], uvw_lambda
 * [t][b][c][3], direction_cosines [k][3] = (l, m, n-1); all device pointers. Deterministic
 * (fixed-order reductions). */
int32_t hipng_idft(const double* direction_cosines, const void* vis, const double* weight, const int32_t* flags,
                   const double* uvw_lambda, int32_t ncomp, int64_t ntimes_baselines, int32_t nchan, int32_t npol,
                   void* flux, double* weight_sum, void* stream);

/* Device: imaging weights, replacing weight_visibility
 * (rascil/processing_components/imaging/weighting.py:28-52) ->
 * grid_blockvisibility_weight_to_griddata + griddata_blockvisibility_reweight
 * (griddata/gridding.py:260-308, 336-424). Rows are [ntimes][nbaselines] as in a
 * BlockVisibility. Every (row, chan) falls on the nearest cell
 * pu = rint(u/cell_u + npix_x/2), pv = rint(v/cell_v + npix_y/2) of an image-sized grid
 * (cell_u = 1/(npix_x * cdelt_x), negative for RA); the grid sums imaging_weight
 * [nrow][nchan][npol] (flagged entries already zero), npol <= 4. weighting 1 (uniform):
 * w <- w*w/g; 2 (robust): f2 = (5*10^-R)^2 sum(w)/sum(g^2), w <- w*(1+f2 w)/(1+f2 g);
 * g <= 0 -> 0. imaging_weight is updated in place; grid [nimchan][npix_y][npix_x][npol]
 * f64 (polarisation innermost) is workspace and holds the density on return.
 * HIPNG_EINVAL if any visibility falls outside the grid (the reference asserts). */
int32_t hipng_weight_visibility(const double* uvw, const double* freq, int64_t ntimes, int64_t nbaselines,
                                int32_t nchan, int32_t npol, const int32_t* vis_to_im, int32_t nimchan,
                                int32_t npix_x, int32_t npix_y, double cell_u, double cell_v,
                                int32_t weighting, double robustness, double* imaging_weight, double* grid,
                                void* stream);

/* Device: gain application, replacing apply_gaintable
 * (rascil/processing_components/calibration/operations.py:37-274). vis complex128 and
 * weight f64 [ntimes][nbaselines][nchan][npol] are updated in place; ant1/ant2 [nbaselines]
 * index the antennas (0 <= a < nant); gain complex128 and gain_weight f64
 * [ngain_times][nant][gain_nchan][nrec][nrec]; time_to_gain [ntimes] names the gain-table row
 * applied to each visibility time (-1: none; the caller matches times within half the row's
 * interval and calls once per match when a time matches several rows, in row order).
 * Channels >= gain_nchan are left unchanged, as in the reference (a T or G table has one
 * channel). With M the gain (inverse = 0) or its inverse: npol 1: V *= sum M1 conj(M2) and
 * the weight becomes gain_weight[a1] * gain_weight[a2] (both 0 where that sum is 0; 1/g is
 * 0 for g = 0); npol 2: V = diag(M1 diag(V) conj(M2)); npol 4: V = M1 V conj(M2) (conj
 * elementwise, no transpose, as the reference). A singular 2x2 gain in inverse mode zeroes
 * the visibility and weight (npol 4) or their first polarisation (npol 2). time_flagged
 * [ntimes] (or NULL) with flags [like vis] (or NULL): where set, V and the weight are first
 * multiplied by (1 - flags) (the reference's use_flags). All arrays are device memory. */
int32_t hipng_apply_gaintable(void* vis, double* weight, const double* flags, const int8_t* time_flagged,
                              int64_t ntimes, int32_t nbaselines, int32_t nchan, int32_t npol,
                              const int32_t* ant1, const int32_t* ant2, const void* gain,
                              const double* gain_weight, int32_t ngain_times, int32_t nant,
                              int32_t gain_nchan, int32_t nrec, const int32_t* time_to_gain,
                              int32_t inverse, void* stream);

/* Device: hipng_apply_gaintable without any device allocation (no counterpart in the reference;
 * for callers that cannot allocate: code under HIP graph capture, tight calibration loops,
 * harnesses that forbid hipMalloc). Same arguments, kernels and results; the prepared gain
 * matrices go to a static per-GPU device array of 8 MiB (nrec^2 x 16 + 4 bytes per gain-table
 * entry: 123,361 entries of 2x2 gains) instead of the library's workspace, and a larger table
 * is refused with HIPNG_EINVAL. All calls on one GPU share that array: calls that overlap on
 * one GPU (several streams or threads) race. */
int32_t hipng_apply_gaintable_noalloc(void* vis, double* weight, const double* flags, const int8_t* time_flagged,
                                      int64_t ntimes, int32_t nbaselines, int32_t nchan, int32_t npol,
                                      const int32_t* ant1, const int32_t* ant2, const void* gain,
                                      const double* gain_weight, int32_t ngain_times, int32_t nant,
                                      int32_t gain_nchan, int32_t nrec, const int32_t* time_to_gain,
                                      int32_t inverse, void* stream);

/* Device: the visibility passes of invert_ng before the gridder, fused
 * (rascil/processing_components/imaging/ng.py:169-181 -> shift_vis_to_image
 * imaging/base.py:67-106 with phaserotate_visibility visibility/base.py:296-357,
 * flagged_vis / flagged_imaging_weight data_models/memory_data_models.py:1502-1524,
 * convert_pol_frame data_models/polarisation.py:164-419). Per (row, chan):
 * x = vis (1 - flags) e^{+2 pi i f/c (u l + v m + w nm1)} (no phase term when lmn is NULL),
 * ms = conv x with conv [npol_image][npol_vis] complex (interleaved re, im; host),
 * wgt[q] = imaging_weight[q] (1 - flags[q]) for q < npol_image; ms / wgt in the gridder's
 * precision (HIPNG_FP64: complex128 / f64, else complex64 / f32), layout [nrow][nchan][npol_image];
 * nonzero [nchan][npol_image] int32 = 1 where some ms of the channel is not 0 (the
 * reference skips all-zero polarisations). vis complex128, flags int32 (or NULL) and
 * imaging_weight f64 [nrow][nchan][npol_vis], uvw f64 [nrow][3]: device; freq, conv, lmn: host. */
int32_t hipng_vis_prologue(const void* vis, const int32_t* flags, const double* imaging_weight, const double* uvw,
                           const double* freq, int64_t nrow, int32_t nchan, int32_t npol_vis, int32_t npol_image,
                           const double* conv, const double* lmn, int32_t precision, void* ms, void* wgt,
                           int32_t* nonzero, void* stream);

/* Device: the visibility passes of predict_ng after the degridder, fused (ng.py:122-133 ->
 * convert_pol_frame, then shift_vis_to_image(inverse=True)): vis = (conv vist)
 * e^{-2 pi i f/c (u l + v m + w nm1)} (no phase term when lmn is NULL), conv
 * [npol_vis][npol_image] complex (host); vist [nrow][nchan][npol_image] in the gridder's
 * precision, vis complex128 [nrow][nchan][npol_vis]. */
int32_t hipng_vis_epilogue(const void* vist, const double* uvw, const double* freq, int64_t nrow, int32_t nchan,
                           int32_t npol_image, int32_t npol_vis, const double* conv, const double* lmn,
                           int32_t precision, void* vis, void* stream);

/* Diagnostics (no counterpart in the reference): per-kernel-class timing with HIP
 * events recorded on the launch stream around every kernel (and rocFFT execution).
 * hipng_profile_enable(1) resets and starts, (0) stops. hipng_profile_query returns
 * entry `idx` (name, total milliseconds, number of launches) or HIPNG_EINVAL past the end. */
int32_t hipng_profile_enable(int32_t on);
int32_t hipng_profile_query(int32_t idx, char* name, int32_t name_len, double* total_ms,
                            int64_t* launches);

/* Release cached device workspace and FFT plans (optional; also released at exit). */
int32_t hipng_release(void);

#ifdef __cplusplus
}
#endif

#endif /* HIPNG_H */
