// fb_launch.h -- host-side launchers shared between translation units.
#pragma once

#include "fastbarnes.h"
#include "fb_internal.h"

namespace fb {

// fb_sort.hip
int64_t scan_temp_elems(int64_t n);
int scan_exclusive(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t stream);
// per-field sort in LDS (fields of at most kLocalSortMax points): the sorted in-grid points
// of field f from out_vals[pt_prefix[f]] on, in the same order as the row buckets; from the
// per-row point counts d_row_count it writes the rows' starts d_row_start (the range of field
// f begins at pt_prefix[f], so no scan over all rows is needed)
int field_sort_init();
int launch_field_sort(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int max_n,
                      const uint32_t *d_keys, const uint32_t *d_row_count, uint32_t *d_row_start,
                      uint32_t *d_out_vals, hipStream_t stream);

// fb_kernels.hip -- injection front end
// (min+max)/2 per field; with d_part (offsets_part_elems doubles) large fields are reduced
// by many workgroups
int64_t offsets_part_elems(int nfields, int64_t max_n);
int launch_offsets(const FieldDev *d_fields, int nfields, int64_t max_n, const double *d_val, double *d_part,
                   double *d_offsets, hipStream_t stream);
// one (base cell, point) pair per point and, unless d_row_count is NULL, per-base-row point
// counts (row = x-line index) by integer atomics; with d_slot/d_rowg (row buckets) each point's
// slot in its row and its row instead of the point indices.  max_rows: the most x-lines of any
// field.  d_rec (may be NULL): each in-grid point's PointRec, for the FP32C injection
int launch_point_keys(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                      const double *d_pts, uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_row_count,
                      uint32_t *d_slot, uint32_t *d_rowg, int64_t max_rows, PointRec *d_rec, hipStream_t stream);
// launch_point_keys and, when the offsets take one workgroup per field (offsets_single_pass),
// launch_offsets in one launch
bool offsets_single_pass(int64_t max_n, const double *d_part);
int launch_point_keys_offsets(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                              const double *d_pts, uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_row_count,
                              uint32_t *d_slot, uint32_t *d_rowg, const double *d_val, double *d_offsets,
                              PointRec *d_rec, hipStream_t stream);
// fb_sort.hip -- row buckets for fields of more than kLocalSortMax points: scatter the in-grid
// points into their rows (d_row_start from the scan of the counts), then sort each row by
// (key, point index); d_svals/d_skeys receive the sorted list, d_tkeys/d_tvals are scratch for
// rows of more than kLocalSortMax points; d_big_list holds 2*nrows entries and d_big_count 2
// counters, which must be zero
int row_bucket_sort(const uint32_t *d_keys, const uint32_t *d_slot, const uint32_t *d_rowg, int64_t npoints,
                    const uint32_t *d_row_start, const uint32_t *d_row_count, int64_t nrows, uint32_t *d_skeys,
                    uint32_t *d_svals, uint32_t *d_big_list, uint32_t *d_big_count, uint32_t *d_tkeys,
                    uint32_t *d_tvals, hipStream_t stream);

// fb_kernels.hip -- FP64 path with the reference's operation order
// Each sorted entry's FP64 injection record (local: the per-field sort's layout).
int launch_inject_records(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                          bool local, int64_t nxlines, const uint32_t *d_vals, const uint32_t *d_row_start,
                          const uint32_t *d_row_count, const double *d_pts, const double *d_val,
                          const double *d_offsets, InjectRec *d_irec, hipStream_t stream);
int launch_exact_inject(const FieldDev *d_fields, const int64_t *d_xline_prefix, int nfields, int64_t nxlines,
                        const InjectRec *d_irec, const uint32_t *d_row_start, const uint32_t *d_row_count,
                        double *d_vg, double *d_wg, int dim, bool zeroed, hipStream_t stream);
// One thread per (line, field component); lines of axis `axis` of every field.
// line_prefix[f] = first (2*line) slot of field f; total = line_prefix[nfields].
// rect_len < 0 selects 2*T[axis]+1 from the field.
int launch_exact_lines(const FieldDev *d_fields, const int64_t *d_line_prefix, int nfields, int64_t nslots,
                       int axis, int rect_len, double *d_vg, double *d_wg, double *d_vs, double *d_ws,
                       hipStream_t stream);
// FP64 x lines through LDS (k_exact_rows), the same bits as launch_exact_lines along x:
// d_wg_prefix[f] = first workgroup of field f (ceil(2*H*D/slots_per_wg) each, nwg in all;
// slots_per_wg = exact_rows_slots of the plan's slots: 32 or 64), d_slot_prefix
// the line slots as for launch_exact_lines; max_rect = the widest kernel (2T+1 or
// rect_override) of the fields.  FB_ERR_UNSUPPORTED when its ring would not fit the LDS
// (exact_rows_ring(max_rect) == 0): take launch_exact_lines then.
int launch_exact_rows(const FieldDev *d_fields, const int64_t *d_wg_prefix, const int64_t *d_slot_prefix, int nfields,
                      int64_t nwg, int slots_per_wg, int rect_override, int64_t max_rect, double *d_vg, double *d_wg,
                      double *d_vs, double *d_ws, hipStream_t stream);
int launch_exact_finalize(const FieldDev *d_fields, const int64_t *d_cell_prefix, int nfields, int64_t ncells,
                          const double *d_vg, const double *d_wg, const double *d_offsets, float *d_out,
                          hipStream_t stream);
int launch_mask(const FieldDev *d_fields, const int64_t *d_cell_prefix, int nfields, int64_t ncells, double *d_wg,
                int fast, hipStream_t stream);

// fb_kernels.hip -- FP32C path
int fast_pass_max_lds();
int fast_pass_init();
bool fast_pass_seg_supported(int kind, int seg);
// kind: kPassSlide (k_fast_pass) or kPassPrefix (k_prefix_pass).  Injecting passes: d_vals is the
// sorted point list (rows' ranges d_line_start/d_line_count), d_rec and d_val each point's
// injection record and value by point index.  d_adj_g (first pass of the adjoint): the tile
// load forms a = g/W, b = -a(f - offset) from d_out = f, d_wsave = W, d_adj_g = g.
int launch_fast_pass(int kind, bool inject, int seg, const FieldDev *d_fields, const double *d_offsets, const PassWork *d_work,
                     int nwork, const PassField *d_pass, int threads, size_t lds_bytes, const PointRec *d_rec,
                     const uint32_t *d_vals, const uint32_t *d_line_start, const uint32_t *d_line_count,
                     const double *d_val, float *d_v, float *d_w, float *d_out,
                     float *d_wsave, int final_pass, const float *d_adj_g, hipStream_t stream);

// fb_kernels.hip -- adjoint of the FP32C batch
int launch_adj_gather(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                      const double *d_pts, const double *d_val, const double *d_offsets, const float *d_A,
                      const float *d_Bb, const PassField *d_pass_y, double *d_gval, double *d_gpts,
                      hipStream_t stream);

// fb_s2.hip -- spherical variant (proj = center_lon, n, n_inv, F, rho0)
int launch_lambert_points(const double *d_pts, double *d_out, int64_t n, const double *proj, hipStream_t stream);
int launch_resample(const float *d_lam, const int64_t *lam_size, const double *lam_x0, const double *x0,
                    const double *step, const int64_t *size, const double *proj, float *d_out, hipStream_t stream);
int launch_naive_s2(const double *d_pts, const double *d_val, const double *d_offset, int64_t n, double sigma,
                    const double *x0, const double *step, const int64_t *size, double *d_out, hipStream_t stream);

// fb_kernels.hip -- elementwise helpers
int launch_zero(void *p, int64_t bytes, hipStream_t stream);   // bytes: multiple of 4
int launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t stream);
int launch_f32_to_f64(const float *in, double *out, int64_t n, double scale, hipStream_t stream);

// fb_kernels.hip -- naive / radius (O(N*W*H*D) brute force, FP64)
int launch_naive(const FieldDev *d_field, int radius, double radius_sqr, double max_dist_weight,
                 const double *d_pts, const double *d_val, const double *d_offsets, double *d_out, int64_t ncells,
                 hipStream_t stream);

}  // namespace fb
