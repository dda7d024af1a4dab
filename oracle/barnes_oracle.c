/*
 * barnes_oracle.c -- CPU oracle for fast Barnes interpolation (TEST INFRASTRUCTURE ONLY).
 *
 * This file is a plain-C, FP64 restatement of the reference algorithm
 * (MeteoSwiss fast-barnes-py 2.0.0, fastbarnes/interpolation.py).  It keeps the
 * reference's loop order and operation order so that its results are the
 * reference's results bit for bit, and it is built with -ffp-contract=off so that
 * no multiply-add is fused.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it, and only as the checker: the product path
 * (fast-barnes-py_amd/) never calls it.
 *
 * Function                       reference (fastbarnes/interpolation.py)
 *   orc_normalize_values         _normalize_values            :204-212
 *   orc_inject                   _inject_data_1d/2d/3d        :218-322
 *   orc_accumulate_tail_array    _accumulate_tail_array       :484-533
 *   orc_accumulate_array         _accumulate_array            :729-772
 *   orc_convolve                 _convolve_tail_1d/2d/3d      :372-479
 *                                _convolve_1d/2d/3d           :617-724
 *   orc_threshold                conv_scale_factor            :389-390,424-425,472-473,634-635,669-670,717-718
 *   orc_half_kernel_size_opt     _get_half_kernel_size_opt    :548-552
 *   orc_tail_value               _get_tail_value              :560-569
 *   orc_half_kernel_size         _get_half_kernel_size        :782-785
 *   orc_sigma_effective          _get_sigma_effective         :796-803
 *   orc_interpolate_convol       _interpolate_opt_convol      :328-367  (tail=1)
 *                                _interpolate_convol          :575-612  (tail=0)
 *   orc_interpolate_naive        _interpolate_naive           :861-938
 *   orc_interpolate_radius_bf    _interpolate_radius          :809-855  (brute force instead of kd-tree)
 *
 * Powers.  The tailed threshold (kernel_size+2*tail_value)**num_iter (:424) is a
 * float64 array raised to an integer: NumPy's power ufunc runs its float64 loop
 * there (the exponent cast to float64), which calls the C library's pow();
 * orc_threshold() calls pow() the same way.  The plain threshold
 * kernel_size**num_iter (:669) is an integer array to an integer power, which
 * wraps around in 64 bits (orc_ipow_int).  `x**2` with a literal exponent is a
 * single product (sq).
 */
#define _GNU_SOURCE
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* int ** int with 64-bit wrap-around, as the integer power of the reference. */
static int64_t orc_ipow_int(int64_t a, int64_t e)
{
    uint64_t r = 1, b = (uint64_t)a;
    while (e > 0) {
        if (e & 1) r *= b;
        e >>= 1;
        b *= b;
    }
    return (int64_t)r;
}

static double sq(double x) { return x * x; }

/* ---- kernel parameters ------------------------------------------------- */

/* _get_half_kernel_size_opt :548-552 */
void orc_half_kernel_size_opt(int dim, const double *sigma, const double *step, int64_t num_iter, int32_t *T)
{
    for (int m = 0; m < dim; m++) {
        double s = sigma[m] / step[m];
        double r = (sqrt(1.0 + 12 * s * s / (double)num_iter) - 1.0) / 2.0;
        T[m] = (int32_t)r;
    }
}

/* _get_tail_value :560-569 */
void orc_tail_value(int dim, const double *sigma, const double *step, int64_t num_iter, double *alpha)
{
    int32_t T[3];
    orc_half_kernel_size_opt(dim, sigma, step, num_iter, T);
    for (int m = 0; m < dim; m++) {
        int64_t hks = T[m];
        int64_t kernel_size = 2 * hks + 1;
        double sigma_rect_sqr = (double)((hks + 1) * hks) / 3.0 * sq(step[m]);
        double num = 0.5 * (double)kernel_size * (sq(sigma[m]) / (double)num_iter - sigma_rect_sqr);
        double den = sq((double)(hks + 1) * step[m]) - sq(sigma[m]) / (double)num_iter;
        alpha[m] = num / den;
    }
}

/* _get_half_kernel_size :782-785 */
void orc_half_kernel_size(int dim, const double *sigma, const double *step, int64_t num_iter, int32_t *T)
{
    double c = sqrt(3.0 / (double)num_iter);
    for (int m = 0; m < dim; m++)
        T[m] = (int32_t)(c * sigma[m] / step[m] + 0.5);
}

/* _get_sigma_effective :796-803 */
void orc_sigma_effective(int dim, const double *sigma, const double *step, int64_t num_iter, double *out)
{
    int32_t T[3];
    orc_half_kernel_size(dim, sigma, step, num_iter, T);
    for (int m = 0; m < dim; m++)
        out[m] = sqrt((double)num_iter / 3.0 * (double)T[m] * (double)(T[m] + 1)) * step[m];
}

/* conv_scale_factor: tailed :424-425 (and :389-390, :472-473); plain :669-670 (and :634-635, :717-718) */
double orc_threshold(int tail, int dim, const double *sigma, const double *step, const int32_t *kernel_size,
                     int64_t num_iter, const double *tail_value, double max_dist_weight)
{
    double prod = 1.0;
    for (int m = 0; m < dim; m++) {
        double p;
        if (tail)
            p = pow((double)kernel_size[m] + 2 * tail_value[m], (double)num_iter);
        else
            p = (double)orc_ipow_int(kernel_size[m], num_iter);
        double f = p / sqrt(2 * 3.141592653589793) / (sigma[m] / step[m]);
        prod *= f;
    }
    return prod * max_dist_weight;
}

/* ---- values ------------------------------------------------------------ */

/* _normalize_values :204-212; val is modified in place, offset returned.  np.amin and
 * np.amax propagate NaN: any NaN value gives a NaN offset (and so an all-NaN field). */
double orc_normalize_values(double *val, int64_t n)
{
    double mn = val[0], mx = val[0];
    int has_nan = isnan(val[0]);
    for (int64_t k = 1; k < n; k++) {
        if (isnan(val[k])) has_nan = 1;
        if (val[k] < mn) mn = val[k];
        if (val[k] > mx) mx = val[k];
    }
    double offset = has_nan ? NAN : (mn + mx) / 2.0;
    for (int64_t k = 0; k < n; k++) val[k] -= offset;
    return offset;
}

/* ---- injection --------------------------------------------------------- */

/* _inject_data_1d/2d/3d :218-322.  vg, wg are zeroed arrays of shape size[::-1]. */
void orc_inject(int dim, double *vg, double *wg, const double *pts, const double *val, int64_t n,
                const double *x0, const double *step, const int64_t *size)
{
    if (dim == 1) {
        for (int64_t k = 0; k < n; k++) {
            double xc = (pts[k] - x0[0]) / step[0];
            if (xc < 0.0 || xc >= (double)(size[0] - 1)) continue;
            int64_t xi = (int64_t)xc;
            double xw = xc - (double)xi;
            double w = (1.0 - xw);
            vg[xi] += w * val[k]; wg[xi] += w;
            w = xw;
            vg[xi + 1] += w * val[k]; wg[xi + 1] += w;
        }
    } else if (dim == 2) {
        const int64_t W = size[0];
        for (int64_t k = 0; k < n; k++) {
            double xc = (pts[2 * k] - x0[0]) / step[0];
            double yc = (pts[2 * k + 1] - x0[1]) / step[1];
            if (xc < 0.0 || yc < 0.0 || xc >= (double)(size[0] - 1) || yc >= (double)(size[1] - 1)) continue;
            int64_t xi = (int64_t)xc, yi = (int64_t)yc;
            double xw = xc - (double)xi, yw = yc - (double)yi;
            double w;
            w = (1.0 - xw) * (1.0 - yw);
            vg[yi * W + xi] += w * val[k];             wg[yi * W + xi] += w;
            w = xw * (1.0 - yw);
            vg[yi * W + xi + 1] += w * val[k];         wg[yi * W + xi + 1] += w;
            w = xw * yw;
            vg[(yi + 1) * W + xi + 1] += w * val[k];   wg[(yi + 1) * W + xi + 1] += w;
            w = (1.0 - xw) * yw;
            vg[(yi + 1) * W + xi] += w * val[k];       wg[(yi + 1) * W + xi] += w;
        }
    } else {
        const int64_t W = size[0], H = size[1];
        for (int64_t k = 0; k < n; k++) {
            double xc = (pts[3 * k] - x0[0]) / step[0];
            double yc = (pts[3 * k + 1] - x0[1]) / step[1];
            double zc = (pts[3 * k + 2] - x0[2]) / step[2];
            if (xc < 0.0 || yc < 0.0 || zc < 0.0 ||
                xc >= (double)(size[0] - 1) || yc >= (double)(size[1] - 1) || zc >= (double)(size[2] - 1))
                continue;
            int64_t xi = (int64_t)xc, yi = (int64_t)yc, zi = (int64_t)zc;
            double xw = xc - (double)xi, yw = yc - (double)yi, zw = zc - (double)zi;
#define IDX3(z, y, x) (((z) * H + (y)) * W + (x))
            double w;
            w = (1.0 - xw) * (1.0 - yw) * (1.0 - zw);
            vg[IDX3(zi, yi, xi)] += w * val[k];             wg[IDX3(zi, yi, xi)] += w;
            w = xw * (1.0 - yw) * (1.0 - zw);
            vg[IDX3(zi, yi, xi + 1)] += w * val[k];         wg[IDX3(zi, yi, xi + 1)] += w;
            w = xw * yw * (1.0 - zw);
            vg[IDX3(zi, yi + 1, xi + 1)] += w * val[k];     wg[IDX3(zi, yi + 1, xi + 1)] += w;
            w = (1.0 - xw) * yw * (1.0 - zw);
            vg[IDX3(zi, yi + 1, xi)] += w * val[k];         wg[IDX3(zi, yi + 1, xi)] += w;
            w = (1.0 - xw) * (1.0 - yw) * zw;
            vg[IDX3(zi + 1, yi, xi)] += w * val[k];         wg[IDX3(zi + 1, yi, xi)] += w;
            w = xw * (1.0 - yw) * zw;
            vg[IDX3(zi + 1, yi, xi + 1)] += w * val[k];     wg[IDX3(zi + 1, yi, xi + 1)] += w;
            w = xw * yw * zw;
            vg[IDX3(zi + 1, yi + 1, xi + 1)] += w * val[k]; wg[IDX3(zi + 1, yi + 1, xi + 1)] += w;
            w = (1.0 - xw) * yw * zw;
            vg[IDX3(zi + 1, yi + 1, xi)] += w * val[k];     wg[IDX3(zi + 1, yi + 1, xi)] += w;
#undef IDX3
        }
    }
}

/* ---- line passes ------------------------------------------------------- */

/* _accumulate_tail_array :484-533.  in_arr and h_arr are both scratch of length
 * arr_len; the function returns whichever of the two holds the result, exactly
 * as the reference returns its ping-pong buffer. */
double *orc_accumulate_tail_array(double *in_arr, double *h_arr, int64_t arr_len, int64_t rect_len,
                                  int64_t num_iter, double alpha)
{
    const int64_t h0 = (rect_len - 1) / 2;   /* rect_len >= 1, so floor division == C division */
    const int64_t h0_1 = h0 + 1;
    const int64_t h1 = rect_len - h0;
    for (int64_t it = 0; it < num_iter; it++) {
        double accu = 0.0;
        int64_t k;
        for (k = -h0; k < 0; k++)                       /* phase a */
            accu += in_arr[k + h0];
        for (k = 0; k < h1; k++) {                      /* phase b */
            accu += in_arr[k + h0];
            h_arr[k] = accu + alpha * in_arr[k + h0_1];
        }
        for (k = h1; k < arr_len - h0_1; k++) {         /* phase c */
            accu += (in_arr[k + h0] - in_arr[k - h1]);
            h_arr[k] = accu + alpha * (in_arr[k - h1] + in_arr[k + h0_1]);
        }
        k = arr_len - h0_1;                             /* phase c' */
        accu += (in_arr[k + h0] - in_arr[k - h1]);
        h_arr[k] = accu + alpha * in_arr[k - h1];
        for (k = arr_len - h0; k < arr_len; k++) {      /* phase d */
            accu -= in_arr[k - h1];
            h_arr[k] = accu + alpha * in_arr[k - h1];
        }
        double *h = in_arr; in_arr = h_arr; h_arr = h;
    }
    return in_arr;
}

/* _accumulate_array :729-772 */
double *orc_accumulate_array(double *in_arr, double *h_arr, int64_t arr_len, int64_t rect_len, int64_t num_iter)
{
    const int64_t h0 = (rect_len - 1) / 2;
    const int64_t h1 = rect_len - h0;
    for (int64_t it = 0; it < num_iter; it++) {
        double accu = 0.0;
        int64_t k;
        for (k = -h0; k < 0; k++)                       /* phase a */
            accu += in_arr[k + h0];
        for (k = 0; k < h1; k++) {                      /* phase b */
            accu += in_arr[k + h0];
            h_arr[k] = accu;
        }
        for (k = h1; k < arr_len - h0; k++) {           /* phase c */
            accu += (in_arr[k + h0] - in_arr[k - h1]);
            h_arr[k] = accu;
        }
        for (k = arr_len - h0; k < arr_len; k++) {      /* phase d */
            accu -= in_arr[k - h1];
            h_arr[k] = accu;
        }
        double *h = in_arr; in_arr = h_arr; h_arr = h;
    }
    return in_arr;
}

/* Convolve every line of `grid` along one axis.  Lines have length len and element
 * stride estride; the line starting at base is walked as grid[base + i*estride].
 * This is the `vg[j,:] = _accumulate_*(vg[j,:].copy(), h_arr, ...)` idiom of
 * :404-418, :440-466, :649-663, :685-711. */
static void convolve_axis(double *grid, int64_t nlines, const int64_t *bases, int64_t len, int64_t estride,
                          int64_t rect_len, int64_t num_iter, int tail, double alpha)
{
    /* Lines are independent, so an OpenMP build (liboracle_omp.so, the timed CPU baseline)
     * shares them out over threads, each with its own line buffers; the values do not change. */
#pragma omp parallel
    {
        double *a = (double *)malloc(sizeof(double) * (len > 0 ? len : 1));
        double *b = (double *)malloc(sizeof(double) * (len > 0 ? len : 1));
#pragma omp for schedule(static)
        for (int64_t l = 0; l < nlines; l++) {
            const int64_t base = bases[l];
            for (int64_t i = 0; i < len; i++) a[i] = grid[base + i * estride];
            double *r = tail ? orc_accumulate_tail_array(a, b, len, rect_len, num_iter, alpha)
                             : orc_accumulate_array(a, b, len, rect_len, num_iter);
            for (int64_t i = 0; i < len; i++) grid[base + i * estride] = r[i];
        }
        free(a);
        free(b);
    }
}

/* Threads of the OpenMP build's line loops (no effect on the single-threaded build); returns
 * the number in use. */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* _convolve_tail_{1,2,3}d (tail=1) :372-479 and _convolve_{1,2,3}d (tail=0) :617-724.
 * vg and wg are modified in place: all passes x, y(, z) for vg then wg per line,
 * then wg < thr is replaced by NaN. */
void orc_convolve(int tail, int dim, double *vg, double *wg, const double *sigma, const double *step,
                  const int64_t *size, const int32_t *kernel_size, int64_t num_iter, const double *tail_value,
                  double max_dist_weight)
{
    int64_t W = size[0], H = dim > 1 ? size[1] : 1, D = dim > 2 ? size[2] : 1;
    int64_t nmax = W * H > W * D ? W * H : W * D; if (H * D > nmax) nmax = H * D;
    int64_t *bases = (int64_t *)malloc(sizeof(int64_t) * (nmax > 0 ? nmax : 1));
    double alpha[3] = {0, 0, 0};
    if (tail) for (int m = 0; m < dim; m++) alpha[m] = tail_value[m];

    /* the reference interleaves vg and wg line by line; the lines are independent, so
     * the per-field order below produces identical values */
    double *fields[2] = {vg, wg};
    /* x pass: lines (z, y) */
    {
        int64_t nl = 0;
        for (int64_t z = 0; z < D; z++) for (int64_t y = 0; y < H; y++) bases[nl++] = (z * H + y) * W;
        for (int f = 0; f < 2; f++)
            convolve_axis(fields[f], nl, bases, W, 1, kernel_size[0], num_iter, tail, alpha[0]);
    }
    if (dim >= 2) {   /* y pass: lines (z, x) */
        int64_t nl = 0;
        for (int64_t z = 0; z < D; z++) for (int64_t x = 0; x < W; x++) bases[nl++] = z * H * W + x;
        for (int f = 0; f < 2; f++)
            convolve_axis(fields[f], nl, bases, H, W, kernel_size[1], num_iter, tail, alpha[1]);
    }
    if (dim >= 3) {   /* z pass: lines (y, x) */
        int64_t nl = 0;
        for (int64_t y = 0; y < H; y++) for (int64_t x = 0; x < W; x++) bases[nl++] = y * W + x;
        for (int f = 0; f < 2; f++)
            convolve_axis(fields[f], nl, bases, D, W * H, kernel_size[2], num_iter, tail, alpha[2]);
    }
    double thr = orc_threshold(tail, dim, sigma, step, kernel_size, num_iter, alpha, max_dist_weight);
    const int64_t total = W * H * D;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total; i++)
        if (wg[i] < thr) wg[i] = NAN;
    free(bases);
}

/* _interpolate_opt_convol :328-367 (tail=1) and _interpolate_convol :575-612 (tail=0).
 * out is float32 of shape size[::-1].  vg_out/wg_out, when non-NULL, receive the
 * masked FP64 fields before the quotient (for diagnostics). */
void orc_interpolate_convol(int tail, int dim, const double *pts, const double *val_in, int64_t n,
                            const double *sigma, const double *x0, const double *step, const int64_t *size,
                            int64_t num_iter, double max_dist_weight, float *out, double *vg_out, double *wg_out)
{
    double *val = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(val, val_in, sizeof(double) * n);
    double offset = orc_normalize_values(val, n);
    int64_t total = 1;
    for (int m = 0; m < dim; m++) total *= size[m];
    double *vg = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    double *wg = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total; i++) vg[i] = wg[i] = 0.0;
    orc_inject(dim, vg, wg, pts, val, n, x0, step, size);
    int32_t T[3], K[3];
    double alpha[3] = {0, 0, 0};
    if (tail) {
        orc_half_kernel_size_opt(dim, sigma, step, num_iter, T);
        orc_tail_value(dim, sigma, step, num_iter, alpha);
    } else {
        orc_half_kernel_size(dim, sigma, step, num_iter, T);
    }
    for (int m = 0; m < dim; m++) K[m] = 2 * T[m] + 1;
    orc_convolve(tail, dim, vg, wg, sigma, step, size, K, num_iter, alpha, max_dist_weight);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < total; i++) out[i] = (float)(vg[i] / wg[i] + offset);
    if (vg_out) memcpy(vg_out, vg, sizeof(double) * total);
    if (wg_out) memcpy(wg_out, wg, sizeof(double) * total);
    free(vg); free(wg); free(val);
}

/* _interpolate_naive :861-938.  Sums run sequentially in point order; the
 * reference's np.dot may use a different association (parity to a tolerance). */
void orc_interpolate_naive(int dim, const double *pts, const double *val_in, int64_t n, const double *sigma,
                           const double *x0, const double *step, const int64_t *size, double *out)
{
    double *val = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(val, val_in, sizeof(double) * n);
    double offset = orc_normalize_values(val, n);
    double scale[3];
    for (int m = 0; m < dim; m++) scale[m] = 2 * sq(sigma[m]);
    int64_t W = size[0], H = dim > 1 ? size[1] : 1, D = dim > 2 ? size[2] : 1;
    for (int64_t z = 0; z < D; z++) {
        double zc = dim > 2 ? x0[2] + (double)z * step[2] : 0.0;
        for (int64_t y = 0; y < H; y++) {
            double yc = dim > 1 ? x0[1] + (double)y * step[1] : 0.0;
            for (int64_t x = 0; x < W; x++) {
                double xc = x0[0] + (double)x * step[0];
                double ws = 0.0, wt = 0.0;
                for (int64_t k = 0; k < n; k++) {
                    double d = sq(pts[k * dim] - xc) / scale[0];
                    if (dim > 1) d = d + sq(pts[k * dim + 1] - yc) / scale[1];
                    if (dim > 2) d = d + sq(pts[k * dim + 2] - zc) / scale[2];
                    double w = exp(-d);
                    ws += w * val[k];
                    wt += w;
                }
                out[(z * H + y) * W + x] = wt > 0.0 ? ws / wt + offset : NAN;
            }
        }
    }
    free(val);
}

/* _interpolate_radius :809-855 with the kd-tree replaced by a brute-force scan that
 * keeps the tree's inclusion test sqr_dist <= radius^2 (util/kdtree.py:305-308).
 * 2D only, scalar sigma. */
void orc_interpolate_radius_bf(const double *pts, const double *val_in, int64_t n, double sigma,
                               const double *x0, const double *step, const int64_t *size,
                               double max_dist_weight, double min_weight, double *out)
{
    double *val = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(val, val_in, sizeof(double) * n);
    double offset = orc_normalize_values(val, n);
    double search_radius = sqrt(-2.0 * log(min_weight)) * sigma;
    double radius_sqr = search_radius * search_radius;
    double scale = 2 * sq(sigma);
    for (int64_t j = 0; j < size[1]; j++) {
        double cy = x0[1] + (double)j * step[1];
        for (int64_t i = 0; i < size[0]; i++) {
            double cx = x0[0] + (double)i * step[0];
            double ws = 0.0, wt = 0.0;
            for (int64_t k = 0; k < n; k++) {
                double dx = cx - pts[2 * k], dy = cy - pts[2 * k + 1];
                double d2 = dx * dx + dy * dy;
                if (d2 <= radius_sqr) {
                    double w = exp(-d2 / scale);
                    ws += w * val[k];
                    wt += w;
                }
            }
            out[j * size[0] + i] = wt >= max_dist_weight ? ws / wt + offset : NAN;
        }
    }
    free(val);
}

/* ---- adjoint (no reference counterpart: the reference has no gradient) ----
 *
 * Derived from the forward path above (:328-367).  On unmasked nodes the forward value
 * is f = V/W + offset with V = B S (val - offset) and W = B S 1: S the bilinear
 * injection (orc_inject), B the per-axis n-fold line filters of orc_convolve.  Each
 * line filter is a banded symmetric matrix (the same weights left and right of the
 * centre, zero padding), so B is self-adjoint.  For g = dL/df:
 *   a = g/W, b = -g V/W^2 (zero on masked nodes), A = B a, Bb = B b
 *   dL/dval_k  = sum_c w_c A[c]
 *   dL/dpts_km = sum_c (dw_c/dp_m) ((val_k - offset) A[c] + Bb[c])
 * where dw_c/dp_m differentiates the corner weight product in axis m and divides by
 * step[m].  Points dropped by the injection get zero gradients. */
static void filter_grid(int tail, int dim, double *grid, const int64_t *size, const int32_t *kernel_size,
                        int64_t num_iter, const double *alpha)
{
    int64_t W = size[0], H = dim > 1 ? size[1] : 1, D = dim > 2 ? size[2] : 1;
    int64_t nmax = W * H > W * D ? W * H : W * D; if (H * D > nmax) nmax = H * D;
    int64_t *bases = (int64_t *)calloc(nmax > 0 ? nmax : 1, sizeof(int64_t));
    int64_t nl = 0;
    for (int64_t z = 0; z < D; z++) for (int64_t y = 0; y < H; y++) bases[nl++] = (z * H + y) * W;
    convolve_axis(grid, nl, bases, W, 1, kernel_size[0], num_iter, tail, alpha[0]);
    if (dim >= 2) {
        nl = 0;
        for (int64_t z = 0; z < D; z++) for (int64_t x = 0; x < W; x++) bases[nl++] = z * H * W + x;
        convolve_axis(grid, nl, bases, H, W, kernel_size[1], num_iter, tail, alpha[1]);
    }
    if (dim >= 3) {
        nl = 0;
        for (int64_t y = 0; y < H; y++) for (int64_t x = 0; x < W; x++) bases[nl++] = y * W + x;
        convolve_axis(grid, nl, bases, D, W * H, kernel_size[2], num_iter, tail, alpha[2]);
    }
    free(bases);
}

/* out: the filtered copy of grid_in (B x); exposes the self-adjoint operator for tests */
void orc_filter_grid(int tail, int dim, const double *grid_in, double *out, const double *sigma, const double *step,
                     const int64_t *size, int64_t num_iter)
{
    int32_t T[3], K[3];
    double alpha[3] = {0, 0, 0};
    if (tail) {
        orc_half_kernel_size_opt(dim, sigma, step, num_iter, T);
        orc_tail_value(dim, sigma, step, num_iter, alpha);
    } else {
        orc_half_kernel_size(dim, sigma, step, num_iter, T);
    }
    for (int m = 0; m < dim; m++) K[m] = 2 * T[m] + 1;
    int64_t total = 1;
    for (int m = 0; m < dim; m++) total *= size[m];
    memcpy(out, grid_in, sizeof(double) * total);
    filter_grid(tail, dim, out, size, K, num_iter, alpha);
}

void orc_adjoint(int tail, int dim, const double *pts, const double *val_in, int64_t n, const double *sigma,
                 const double *x0, const double *step, const int64_t *size, int64_t num_iter,
                 double max_dist_weight, const double *g, double *gval, double *gpts)
{
    int64_t total = 1;
    for (int m = 0; m < dim; m++) total *= size[m];
    float *fout = (float *)malloc(sizeof(float) * total);
    double *vg = (double *)malloc(sizeof(double) * total);
    double *wg = (double *)malloc(sizeof(double) * total);
    orc_interpolate_convol(tail, dim, pts, val_in, n, sigma, x0, step, size, num_iter, max_dist_weight, fout, vg, wg);
    double *val = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(val, val_in, sizeof(double) * n);
    orc_normalize_values(val, n);
    /* wg is NaN on masked nodes (orc_convolve), so the comparison drops them */
    for (int64_t i = 0; i < total; i++) {
        if (wg[i] == wg[i] && g[i] == g[i]) {
            double a = g[i] / wg[i];
            vg[i] = -a * (vg[i] / wg[i]);
            wg[i] = a;
        } else {
            vg[i] = 0.0;
            wg[i] = 0.0;
        }
    }
    double *A = wg, *Bb = vg;
    double sg[3], st[3];
    for (int m = 0; m < dim; m++) { sg[m] = sigma[m]; st[m] = step[m]; }
    double *tmp = (double *)malloc(sizeof(double) * total);
    orc_filter_grid(tail, dim, A, tmp, sg, st, size, num_iter);
    memcpy(A, tmp, sizeof(double) * total);
    orc_filter_grid(tail, dim, Bb, tmp, sg, st, size, num_iter);
    memcpy(Bb, tmp, sizeof(double) * total);
    const int64_t W = size[0], H = dim > 1 ? size[1] : 1;
    for (int64_t k = 0; k < n; k++) {
        double c[3], w[3];
        int64_t ci[3];
        int inside = 1;
        for (int m = 0; m < dim; m++) {
            c[m] = (pts[k * dim + m] - x0[m]) / step[m];
            if (!(c[m] >= 0.0) || c[m] >= (double)(size[m] - 1)) inside = 0;
        }
        double gv = 0.0, gp[3] = {0.0, 0.0, 0.0};
        if (inside) {
            for (int m = 0; m < dim; m++) { ci[m] = (int64_t)c[m]; w[m] = c[m] - (double)ci[m]; }
            for (int cb = 0; cb < (1 << dim); cb++) {
                int64_t idx = 0, mul = 1;
                double f[3], d[3];
                for (int m = 0; m < dim; m++) {
                    int bit = (cb >> m) & 1;
                    idx += (ci[m] + bit) * mul;
                    mul *= size[m];
                    f[m] = bit ? w[m] : 1.0 - w[m];
                    d[m] = bit ? 1.0 : -1.0;
                }
                double wc = 1.0;
                for (int m = 0; m < dim; m++) wc *= f[m];
                gv += wc * A[idx];
                double t = val[k] * A[idx] + Bb[idx];
                for (int m = 0; m < dim; m++) {
                    double dm = d[m];
                    for (int q = 0; q < dim; q++) if (q != m) dm *= f[q];
                    gp[m] += dm * t;
                }
            }
        }
        gval[k] = gv;
        if (gpts) for (int m = 0; m < dim; m++) gpts[k * dim + m] = gp[m] / step[m];
    }
    (void)W; (void)H;
    free(tmp); free(val); free(vg); free(wg); free(fout);
}

/* ---- spherical variant ---------------------------------------------------
 * fastbarnes/util/lambert_conformal.py and fastbarnes/interpolationS2.py. */

static const double ORC_RAD = 3.141592653589793 / 180.0;

/* create_proj, lambert_conformal.py:50-97: proj = (center_lon, n, n_inv, F, rho0) */
void orc_lambert_create(double center_lon, double center_lat, double lat1, double lat2, double *proj)
{
    const double h = ORC_RAD / 2.0;
    double n;
    if (lat1 != lat2)
        n = log(cos(lat1 * ORC_RAD) / cos(lat2 * ORC_RAD)) / log(tan((90.0 + lat2) * h) / tan((90.0 + lat1) * h));
    else
        n = sin(lat1 * ORC_RAD);
    const double F = cos(lat1 * ORC_RAD) * pow(tan((90.0 + lat1) * h), n) / n;
    const double rho0 = F / pow(tan((90.0 + center_lat) * h), n);
    proj[0] = center_lon; proj[1] = n; proj[2] = 1.0 / n; proj[3] = F; proj[4] = rho0;
}

/* get_scale, lambert_conformal.py:102-110 */
double orc_lambert_scale(double lat, const double *proj)
{
    return proj[3] * proj[1] / (cos(lat * ORC_RAD) * pow(tan((90.0 + lat) * (ORC_RAD / 2.0)), proj[1]));
}

/* to_map, lambert_conformal.py:115-125 (geo rows are (lon, lat)) */
void orc_lambert_to_map(const double *geo, double *map, int64_t n, const double *proj)
{
    for (int64_t k = 0; k < n; k++) {
        const double rho = proj[3] / pow(tan((90.0 + geo[2 * k + 1]) * (ORC_RAD / 2.0)), proj[1]);
        const double arg = proj[1] * (geo[2 * k] - proj[0]) * ORC_RAD;
        map[2 * k] = rho * sin(arg) / ORC_RAD;
        map[2 * k + 1] = (proj[4] - rho * cos(arg)) / ORC_RAD;
    }
}

/* to_geo, lambert_conformal.py:143-160 */
void orc_lambert_to_geo(const double *map, double *geo, int64_t n, const double *proj)
{
    for (int64_t k = 0; k < n; k++) {
        const double x = map[2 * k] * ORC_RAD, y = map[2 * k + 1] * ORC_RAD;
        const double arg = proj[4] - y;
        double rho = sqrt(x * x + arg * arg);
        if (proj[1] < 0.0) rho = -rho;
        const double theta = atan2(x, arg);
        geo[2 * k + 1] = atan(pow(proj[3] / rho, proj[2])) / (ORC_RAD / 2.0) - 90.0;
        geo[2 * k] = proj[0] + theta / proj[1] / ORC_RAD;
    }
}

/* _resample, interpolationS2.py:221-267.  Nodes whose 2x2 Lambert neighbourhood leaves the
 * Lambert grid are NaN here (the reference does not check the bounds). */
void orc_resample(const float *lam, const double *lam_x0, const int64_t *lam_size, const double *x0,
                  const double *step, const int64_t *size, const double *proj, float *out)
{
    for (int64_t j = 0; j < size[1]; j++) {
        const double geoy = (double)j * step[1] + x0[1];
        const double rho = proj[3] / pow(tan((90.0 + geoy) * (ORC_RAD / 2.0)), proj[1]);
        for (int64_t i = 0; i < size[0]; i++) {
            const double geox = x0[0] + (double)i * step[0];
            const double arg = proj[1] * (geox - proj[0]) * ORC_RAD;
            double mx = rho * sin(arg) / ORC_RAD;
            double my = (proj[4] - rho * cos(arg)) / ORC_RAD;
            mx = (mx - lam_x0[0]) / step[0];
            my = (my - lam_x0[1]) / step[1];
            const int32_t ix = (int32_t)mx, iy = (int32_t)my;
            mx -= (double)ix;
            my -= (double)iy;
            if (!(ix >= 0 && iy >= 0 && ix + 1 < lam_size[0] && iy + 1 < lam_size[1])) {
                out[j * size[0] + i] = NAN;
                continue;
            }
            const int64_t W = lam_size[0];
            out[j * size[0] + i] = (float)((1.0 - my) * (1.0 - mx) * (double)lam[iy * W + ix] +
                                           my * (1.0 - mx) * (double)lam[(iy + 1) * W + ix] +
                                           my * mx * (double)lam[(iy + 1) * W + ix + 1] +
                                           (1.0 - my) * mx * (double)lam[iy * W + ix + 1]);
        }
    }
}

/* _interpolate_opt_convol_S2 with its two parts, interpolationS2.py:157-206: project the
 * points onto the fixed Lambert grid (lam_x0 = (-32, -2), size (64/step, 44/step)), run the
 * optimized convolution there, resample to the lon/lat grid (resample != 0) or return the
 * Lambert field (out sized lam_size[1] x lam_size[0]). */
void orc_barnes_s2_opt(const double *pts, const double *val, int64_t n, const double *sigma, const double *x0,
                       const double *step, const int64_t *size, int64_t num_iter, double max_dist_weight,
                       int resample, float *out)
{
    double proj[5];
    orc_lambert_create(11.5, 34.5, 42.5, 65.5, proj);
    const double lam_x0[2] = {-32.0, -2.0};
    const int64_t lam_size[2] = {(int64_t)(64.0 / step[0]), (int64_t)(44.0 / step[1])};
    double *lam_pts = (double *)malloc(sizeof(double) * 2 * (n > 0 ? n : 1));
    orc_lambert_to_map(pts, lam_pts, n, proj);
    float *lam = resample ? (float *)malloc(sizeof(float) * lam_size[0] * lam_size[1]) : out;
    orc_interpolate_convol(1, 2, lam_pts, val, n, sigma, lam_x0, step, lam_size, num_iter, max_dist_weight, lam,
                           NULL, NULL);
    if (resample) {
        orc_resample(lam, lam_x0, lam_size, x0, step, size, proj, out);
        free(lam);
    }
    free(lam_pts);
}

/* _interpolate_naive_S2 and _dist_S2, interpolationS2.py:272-320: spherical distance in
 * degrees, sequential sums (the reference's np.dot may associate differently). */
void orc_naive_s2(const double *pts, const double *val_in, int64_t n, const double *sigma, const double *x0,
                  const double *step, const int64_t *size, double *out)
{
    double *val = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
    memcpy(val, val_in, sizeof(double) * n);
    const double offset = orc_normalize_values(val, n);
    const double scale = 2 * sq(sigma[0]);
    for (int64_t j = 0; j < size[1]; j++) {
        const double yc = x0[1] + (double)j * step[1];
        const double lat0 = yc * ORC_RAD;
        for (int64_t i = 0; i < size[0]; i++) {
            const double xc = x0[0] + (double)i * step[0];
            double ws = 0.0, wt = 0.0;
            for (int64_t k = 0; k < n; k++) {
                const double lat1 = pts[2 * k + 1] * ORC_RAD;
                double arg = sin(lat0) * sin(lat1) + cos(lat0) * cos(lat1) * cos((pts[2 * k] - xc) * ORC_RAD);
                if (arg > 1.0) arg = 1.0;
                const double d = acos(arg) / ORC_RAD;
                const double w = exp(-d * d / scale);
                ws += w * val[k];
                wt += w;
            }
            out[j * size[0] + i] = wt > 0.0 ? ws / wt + offset : NAN;
        }
    }
    free(val);
}
