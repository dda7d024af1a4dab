/*
 * fastbarnes.h -- C ABI of libfastbarnes.so, the MI355X (gfx950) implementation of
 * fast (convolutional) Barnes interpolation.
 *
 * Every entry point returns an int status (FB_OK == 0) and never throws; the
 * text of the last error of the calling thread is available from fb_last_error().
 * Pointers are plain host or device pointers; sizes are int64_t.  Arrays follow
 * the reference's conventions: `size` is given in (x, y, z) order and grids are
 * stored row-major with index [y][x] / [z][y][x]; pts is N x dim, row-major.
 *
 * Reference interfaces replaced (MeteoSwiss fast-barnes-py 2.0.0,
 * fastbarnes/interpolation.py) are cited beside each declaration.
 */
#ifndef FASTBARNES_H
#define FASTBARNES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define FB_OK 0
#define FB_ERR_INVALID_ARGUMENT 1   /* maps to the reference's RuntimeError */
#define FB_ERR_HIP 2                /* a HIP runtime call failed */
#define FB_ERR_UNSUPPORTED 3        /* arguments valid for the reference, outside this library's limits */
#define FB_ERR_OUT_OF_MEMORY 4

/* ---- methods (the `method` string of barnes(), interpolation.py:65-76,169-199) */
#define FB_METHOD_OPTIMIZED_CONVOLUTION 0
#define FB_METHOD_CONVOLUTION 1
#define FB_METHOD_NAIVE 2
#define FB_METHOD_RADIUS 3

/* ---- arithmetic of the convolution methods ----------------------------- */
/* FB_PREC_FP64: the reference's own FP64 operation order (injection in point order,
 *   sequential running sums per line, FP64 quotient); float32 results equal the
 *   reference's bit for bit.
 * FP32C: FP32 storage between passes with every window sum exact in FP64 (outputs and the
 *   quotient formed in float32, the value offset added in FP64, the NaN mask decided as in
 *   FP64); deterministic, RMSE <= 1e-5 against FP64. */
#define FB_PREC_FP64 0
#define FB_PREC_FP32C 1

/* ---- flags ------------------------------------------------------------- */
#define FB_FLAG_DEVICE_PTRS 1u      /* pts/val/out (and line buffers) are device pointers */

/* Library version (major*10000 + minor*100 + patch). */
int fb_version(void);

/* Text of the last error raised on the calling thread ("" if none). */
const char *fb_last_error(void);

/* Number of visible HIP devices (0 when none).  Host-only. */
int fb_device_count(int *count);

/* Architecture name of the current device (hipDeviceProp_t::gcnArchName, e.g.
 * "gfx950:sramecc+:xnack-", truncated to cap-1 characters) and its compute-unit count (may be
 * NULL); a key for cached tuning results.  No reference counterpart. */
int fb_device_arch(char *arch, int cap, int *compute_units);

/* ---- kernel parameters (host-only, no GPU needed) ---------------------- */

/* Half kernel size T per axis.
 *   method FB_METHOD_OPTIMIZED_CONVOLUTION: _get_half_kernel_size_opt  interpolation.py:548-552
 *   method FB_METHOD_CONVOLUTION:           _get_half_kernel_size      interpolation.py:782-785 */
int fb_half_kernel_size(int method, int dim, const double *sigma, const double *step, int num_iter,
                        int32_t *half_kernel_size);

/* Tail value alpha per axis: _get_tail_value  interpolation.py:560-569 */
int fb_tail_value(int dim, const double *sigma, const double *step, int num_iter, double *tail_value);

/* Effective sigma of the plain rect kernel: _get_sigma_effective  interpolation.py:796-803 */
int fb_sigma_effective(int dim, const double *sigma, const double *step, int num_iter, double *sigma_eff);

/* NaN-mask threshold conv_scale_factor of the convolution methods:
 *   tailed interpolation.py:424-425 (1D :389-390, 3D :472-473); plain :669-670 (:634-635, :717-718).
 * kernel_size and tail_value are per axis (tail_value ignored for FB_METHOD_CONVOLUTION). */
int fb_threshold(int method, int dim, const double *sigma, const double *step, const int32_t *kernel_size,
                 int num_iter, const double *tail_value, double max_dist_weight, double *threshold);

/* ---- the drop-in ------------------------------------------------------- */

/* barnes(pts, val, sigma, x0, step, size, method, num_iter, max_dist, min_weight)
 *   interpolation.py:31-199 (dispatch :169-199) -> _interpolate_opt_convol :328-367,
 *   _interpolate_convol :575-612, _interpolate_naive :861-938, _interpolate_radius :809-855.
 * max_dist_weight = exp(-max_dist^2/2) as computed at interpolation.py:167.
 * out: float32 of W*H*D elements for the convolution methods, float64 for naive/radius.
 * flags: FB_FLAG_DEVICE_PTRS when pts/val/out live in device memory; stream is a
 * hipStream_t (NULL = the library's stream).  The call returns after `out` is written. */
int fb_barnes(int method, int precision, int dim, int64_t n, const double *pts, const double *val,
              const double *sigma, const double *x0, const double *step, const int64_t *size,
              int num_iter, double max_dist_weight, double min_weight, void *out, unsigned flags,
              void *stream);

/* fb_barnes for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.
 * The convolution methods only (FB_METHOD_OPTIMIZED_CONVOLUTION, FB_METHOD_CONVOLUTION);
 * pts, val and out (float32) are device pointers, the other arguments as for fb_barnes.
 * The workspace is a fixed array of FB_NOALLOC_SCRATCH_BYTES per device, reserved with the
 * library's code object and held until the process ends: enough for one 2400x1200 field in
 * either precision; a call whose workspace exceeds it fails with FB_ERR_UNSUPPORTED before
 * any kernel runs.  Calls no hipMalloc/hipFree.  All calls on one device share that array,
 * which keeps the last call's contents: calls that overlap in time on one device (other host
 * threads or streams) race.  stream is a hipStream_t (NULL = the null stream); the call
 * returns after `out` is written. */
#define FB_NOALLOC_SCRATCH_BYTES (128ull << 20)
int fb_barnes_noalloc(int method, int precision, int dim, int64_t n, const double *pts, const double *val,
                      const double *sigma, const double *x0, const double *step, const int64_t *size,
                      int num_iter, double max_dist_weight, float *out, void *stream);

/* barnes_S2(pts, val, sigma, x0, step, size, method, num_iter, max_dist, resample)
 *   interpolationS2.py:31-154 (dispatch :146-154):
 *   FB_METHOD_OPTIMIZED_CONVOLUTION = 'optimized_convolution_S2' :157-219 -- Lambert conformal
 *     projection (util/lambert_conformal.py:115-125, projection create_proj(11.5, 34.5,
 *     42.5, 65.5) :50-97) onto the fixed grid x0 = (-32, -2), size (int(64/step_x),
 *     int(44/step_y)), the optimized convolution there, then _resample :221-267 to the
 *     lon/lat grid.  out: float32, size[1] x size[0] if resample, else the Lambert field
 *     (float32, int(44/step_y) x int(64/step_x)).
 *   FB_METHOD_NAIVE = 'naive_S2' :272-320 -- spherical distance; out float64 size[1] x size[0].
 * pts rows are (lon, lat) in degrees; sigma, x0, step, size have 2 entries.  flags and
 * stream as for fb_barnes. */
int fb_barnes_s2(int method, int precision, int64_t n, const double *pts, const double *val,
                 const double *sigma, const double *x0, const double *step, const int64_t *size,
                 int num_iter, double max_dist_weight, int resample, void *out, unsigned flags, void *stream);

/* ---- line and grid primitives behind the reference's private helpers ---- */

/* _inject_data_1d/2d/3d  interpolation.py:218-322: adds the 2^dim corner weights w of every
 * sample point (bilinear / trilinear, algorithm B.1) as w*val[k] to vg and w to wg, FP64 grids
 * of shape size[::-1] (x fastest) whose existing contents are kept.  Each cell receives its
 * contributions in increasing point index, the reference's order, so the sums are bitwise the
 * reference's.  val is used as given (the caller has normalised it).  Points outside
 * [x0, x0 + (size-1)*step) along any axis are skipped.  flags, stream as for fb_barnes; the
 * call returns after vg and wg are written. */
int fb_inject(int dim, int64_t n, const double *pts, const double *val, const double *x0, const double *step,
              const int64_t *size, double *vg, double *wg, unsigned flags, void *stream);

/* num_iter-fold box filter of `nlines` lines of length len (row-major nlines x len, FP64):
 *   alpha-tailed  _accumulate_tail_array  interpolation.py:484-533  (tail != 0)
 *   plain         _accumulate_array       interpolation.py:729-772  (tail == 0)
 * precision FB_PREC_FP64 reproduces the reference bit for bit; FB_PREC_FP32C runs the
 * fast kernel (FP32 storage, FP64 accumulators).  in and out may alias. */
int fb_accumulate_lines(int tail, int precision, int64_t nlines, int64_t len, int rect_len, int num_iter,
                        double alpha, const double *in, double *out, unsigned flags, void *stream);

/* In-place num_iter-fold convolution of the FP64 fields vg, wg (shape size[::-1]) along every
 * axis followed by wg[wg < threshold] = NaN:
 *   _convolve_tail_1d/2d/3d interpolation.py:372-479 (tail != 0),
 *   _convolve_1d/2d/3d      interpolation.py:617-724 (tail == 0). */
int fb_convolve_grid(int tail, int precision, int dim, const int64_t *size, const int32_t *kernel_size,
                     int num_iter, const double *tail_value, double threshold, double *vg, double *wg,
                     unsigned flags, void *stream);

/* ---- batched fields (no reference counterpart: the reference interpolates one field per call) */

/* One field of a batch.  Its points are rows [pt_offset, pt_offset+n) of the batch's
 * pts array (row-major, dim columns) and val array; its result occupies
 * size[0]*size[1]*size[2] float32 elements at out_offset of the batch output. */
typedef struct fb_field {
    int32_t dim;            /* 1, 2 or 3 (all fields of a plan share it) */
    int32_t method;         /* FB_METHOD_OPTIMIZED_CONVOLUTION or FB_METHOD_CONVOLUTION */
    int32_t num_iter;
    int32_t reserved;
    int64_t n;
    int64_t pt_offset;
    int64_t out_offset;
    int64_t size[3];        /* unused axes = 1 */
    double sigma[3];
    double x0[3];
    double step[3];
    double max_dist_weight;
} fb_field;

typedef struct fb_plan fb_plan;

/* Validates the fields, derives per-field kernel parameters on the host, uploads the
 * descriptors and allocates the plan's device workspace on the current device. */
int fb_plan_create(fb_plan **plan, const fb_field *fields, int nfields, int precision, unsigned flags);

/* Runs the batch on `stream`: pts/val/out are device pointers laid out as described
 * for fb_field.  Asynchronous with respect to the host. */
int fb_plan_execute(fb_plan *plan, const double *d_pts, const double *d_val, float *d_out, void *stream);

/* Captures fb_plan_execute for fixed buffers into a hipGraph and replays it; the first
 * call for a (pts, val, out, stream) tuple captures, later calls replay. */
int fb_plan_execute_graph(fb_plan *plan, const double *d_pts, const double *d_val, float *d_out, void *stream);

/* Training forward: fb_plan_execute that also writes the normalised weight field of every
 * output node (float32, laid out like d_out) to d_wsave, for fb_plan_backward.
 * FP32C plans only.  No counterpart in the reference (it has no gradient); it serves
 * callers that differentiate through the interpolation. */
int fb_plan_forward(fb_plan *plan, const double *d_pts, const double *d_val, float *d_out, float *d_wsave,
                    void *stream);

/* Adjoint of fb_plan_forward.  Given d_gout = dL/d(out) (float32, laid out like d_out;
 * nodes where out is NaN are ignored), writes dL/d(val) to d_gval (one double per point)
 * and, when d_gpts is not NULL, dL/d(pts) to d_gpts (n x dim doubles).  Points the
 * injection drops (outside the grid) get zero gradients.  pts, val, out and wsave must be
 * those of the forward call.  Uses the plan's workspace: calls on one plan are
 * stream-ordered, not concurrent.  No counterpart in the reference. */
int fb_plan_backward(fb_plan *plan, const double *d_pts, const double *d_val, const float *d_out,
                     const float *d_wsave, const float *d_gout, double *d_gval, double *d_gpts, void *stream);

/* Kinds of FP32C line pass (fb_plan_pass_groups): the sliding-window pass (float32 lines in
 * LDS, FP64 window sums) and the prefix pass (FP64 prefix sums of the lines in LDS). */
#define FB_PASS_SLIDE 0
#define FB_PASS_PREFIX 1

/* Autotuner of the FP32C line passes (no reference counterpart: the reference has no
 * GPU tiling).  For every axis and line length of the plan it times each pass kind, segment
 * length (elements per lane; it sets the lanes per line and the lines per workgroup) and
 * workgroup size (128, 256 or 512 threads) on the plan's own buffers and keeps the fastest;
 * untuned plans use a cost model.  Leaves d_out undefined: execute again afterwards.  A
 * no-op for FP64 plans. */
int fb_plan_autotune(fb_plan *plan, const double *d_pts, const double *d_val, float *d_out, void *stream);

/* Forces the kind (FB_PASS_SLIDE or FB_PASS_PREFIX), segment length and workgroup size
 * (128, 256 or 512 threads) of the fast pass along `axis` for the fields whose lines along it
 * have length `len`, and rebuilds the plan's passes; for manual tuning and tests (no
 * reference counterpart).  FP32C plans only. */
int fb_plan_choose_pass(fb_plan *plan, int axis, int64_t len, int kind, int seg, int threads);

/* The line-length classes of the fast pass along `axis` of an FP32C plan, ascending by the
 * length of the lines the pass runs (a chunk's length for lines that run in chunks), with the
 * choice in force for each: kind (FB_PASS_SLIDE or FB_PASS_PREFIX), segment length and
 * workgroup threads, autotuned or forced or else the cost model's.  *nclasses receives the
 * number of classes; at most max_classes entries are written.  field_class (nfields entries,
 * may be NULL) receives each field's class index.  Any output array may be NULL.  With
 * fb_plan_choose_pass this saves and restores an autotuned plan's choices.  No reference
 * counterpart. */
int fb_plan_pass_choices(fb_plan *plan, int axis, int64_t *len, int32_t *kind, int32_t *seg, int32_t *threads,
                         int max_classes, int *nclasses, int32_t *field_class);

/* The launches of the fast pass along `axis`: segment length, workgroup size (threads), work
 * items and kind (FB_PASS_SLIDE or FB_PASS_PREFIX) of each; any output array may be NULL. */
int fb_plan_pass_groups(fb_plan *plan, int axis, int32_t *seg, int32_t *threads, int32_t *items, int32_t *kind,
                        int max_groups, int *ngroups);

/* Layout of field `field`'s intermediate v, w fields between the x and y passes of an FP32C
 * plan: *blocked = 1 when they are kept in blocks of 8 rows x 4 columns (one 128-byte line
 * each, so that a y-pass tile of 4 columns reads whole lines), 0 when row-major.  2D fields
 * with W a multiple of 4 and H of 8 whose passes allow it are blocked; the results are the
 * same bits either way.  No reference counterpart (the reference has no GPU layout). */
int fb_plan_field_layout(fb_plan *plan, int field, int *blocked);

/* Allows (enable = 1, the default) or forbids (0) the blocked intermediate for the plan's
 * fields and rebuilds its passes; for tests and measurements (the results are the same bits
 * either way).  No reference counterpart. */
int fb_plan_set_blocked(fb_plan *plan, int enable);

/* LDS layout of the FP32C sliding pass over lines of length `len` with half kernel T, `seg`
 * elements per lane (an instantiated segment length) and workgroups of `threads` threads
 * (128, 256 or 512); column_tiles = 1 for a y or z pass, 0 for an x pass.  The workgroup's
 * LDS holds T+1 zeros, then `lines` v lines and `lines` w lines at a stride of len + gap
 * floats (each line's cells, then a gap whose last T+1 zeros are the next line's left
 * padding), then a reserve of 64 floats; *lanes_per_line lanes take `seg` consecutive cells
 * each.  *lds_bytes is the workgroup's LDS size.  Host code only (no device call): for tests
 * of the layout's bounds.  No reference counterpart. */
int fb_pass_layout(int64_t len, int T, int seg, int threads, int column_tiles, int *lanes_per_line, int *lines,
                   int *gap, int64_t *lds_bytes);

/* Per-stage device time (ms) of the last fb_plan_execute run with timing enabled,
 * measured with HIP events recorded on the execution stream.  names receives a
 * comma-separated stage list.  enable != 0 turns event recording on. */
int fb_plan_set_timing(fb_plan *plan, int enable);
int fb_plan_stage_times(fb_plan *plan, float *ms, int max_stages, int *nstages, char *names, int names_len);

/* Total points of the batch's output grids, and device workspace bytes held. */
int fb_plan_info(fb_plan *plan, int64_t *total_points, int64_t *workspace_bytes);

int fb_plan_destroy(fb_plan *plan);

#ifdef __cplusplus
}
#endif

#endif /* FASTBARNES_H */
