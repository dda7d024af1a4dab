// fb_kernels.hip -- device kernels of fast Barnes interpolation on gfx950.
//
// Pipeline of one batch (every stage covers all fields of the batch):
//   k_offsets       value offset (min+max)/2 per field           interpolation.py:204-212
//   k_point_keys    per point: base cell (lower corner) -> (cell key, point) pair,
//                   per-row point counts                          interpolation.py:218-322
//   (fb_sort.hip)   scan of the row counts; per-field LDS sort or row buckets ordering the
//                   pairs by (base cell, point index)
//   FP64 path:      k_exact_inject, k_exact_lines (per axis), k_exact_finalize
//                   -- the reference's operation order, one thread per line
//   FP32C path:     k_fast_pass<inject> along x, k_fast_pass along y (and z);
//                   quotient, offset and NaN mask fused into the last pass
#include <climits>
#include <cmath>
#include <mutex>
#include <string>

#include "fb_launch.h"

namespace fb {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int find_seg(const int64_t *__restrict__ prefix, int n, int64_t i)
{
    // largest f in [0, n) with prefix[f] <= i
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (prefix[mid] <= i) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// value offsets: (min + max)/2 over all values of a field, NaN if any value is NaN
// (numpy's amin/amax propagate NaN), interpolation.py:209

// min, max and a NaN flag of v[k] for k = k0, k0 + kBlock, ... < k1, the loads of eight steps
// issued before their first use (one thread's share of a workgroup's values)
__device__ __forceinline__ void minmax_range(const double *__restrict__ v, int64_t k0, int64_t k1, double &mn,
                                             double &mx, int &isn)
{
    constexpr int kUnroll = 8;
    int64_t k = k0;
    for (; k + (kUnroll - 1) * kBlock < k1; k += kUnroll * kBlock) {
        double x[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) x[u] = v[k + u * kBlock];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (x[u] != x[u]) isn = 1;
            mn = fmin(mn, x[u]);
            mx = fmax(mx, x[u]);
        }
    }
    for (; k < k1; k += kBlock) {
        const double x = v[k];
        if (x != x) isn = 1;
        mn = fmin(mn, x);
        mx = fmax(mx, x);
    }
}

__device__ __forceinline__ void offsets_block(const FieldDev *__restrict__ fields, const double *__restrict__ val,
                                              double *__restrict__ offsets, int f)
{
    __shared__ double smin[kBlock], smax[kBlock];
    __shared__ int snan;
    const FieldDev &F = fields[f];
    const int tid = threadIdx.x;
    if (tid == 0) snan = 0;
    double mn = INFINITY, mx = -INFINITY;
    int isn = 0;
    minmax_range(val + F.pt_off, tid, F.n, mn, mx, isn);
    smin[tid] = mn;
    smax[tid] = mx;
    __syncthreads();
    if (isn) snan = 1;
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if (tid < s) {
            smin[tid] = fmin(smin[tid], smin[tid + s]);
            smax[tid] = fmax(smax[tid], smax[tid + s]);
        }
        __syncthreads();
    }
    if (tid == 0) offsets[f] = snan ? __builtin_nan("") : (smin[0] + smax[0]) / 2.0;
}

__global__ __launch_bounds__(kBlock) void k_offsets(const FieldDev *__restrict__ fields,
                                                    const double *__restrict__ val, double *__restrict__ offsets)
{
    offsets_block(fields, val, offsets, blockIdx.x);
}

// The same over many workgroups per field (large fields): chunk c of field f (blockIdx.x,
// blockIdx.y) reduces kOffChunk values to (min, max, NaN flag); k_offsets_final combines a
// field's chunks.  min/max do not depend on the order, so the result equals k_offsets'.
constexpr int kOffChunk = 8192;

__global__ __launch_bounds__(kBlock) void k_offsets_part(const FieldDev *__restrict__ fields,
                                                         const double *__restrict__ val, double *__restrict__ part)
{
    __shared__ double smin[kBlock], smax[kBlock];
    __shared__ int snan;
    const FieldDev &F = fields[blockIdx.y];
    const int tid = threadIdx.x;
    if (tid == 0) snan = 0;
    const int64_t b0 = (int64_t)blockIdx.x * kOffChunk;
    const int64_t b1 = min(F.n, b0 + kOffChunk);
    double mn = INFINITY, mx = -INFINITY;
    int isn = 0;
    minmax_range(val + F.pt_off, b0 + tid, b1, mn, mx, isn);
    smin[tid] = mn;
    smax[tid] = mx;
    __syncthreads();
    if (isn) snan = 1;
    for (int st = kBlock / 2; st > 0; st >>= 1) {
        if (tid < st) {
            smin[tid] = fmin(smin[tid], smin[tid + st]);
            smax[tid] = fmax(smax[tid], smax[tid + st]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        double *o = part + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 3;
        o[0] = smin[0];
        o[1] = smax[0];
        o[2] = snan ? 1.0 : 0.0;
    }
}

__global__ __launch_bounds__(kBlock) void k_offsets_final(const double *__restrict__ part, int nfields, int nchunk,
                                                          double *__restrict__ offsets)
{
    const int f = blockIdx.x * kBlock + threadIdx.x;
    if (f >= nfields) return;
    double mn = INFINITY, mx = -INFINITY;
    bool isn = false;
    for (int c = 0; c < nchunk; ++c) {
        const double *o = part + ((int64_t)f * nchunk + c) * 3;
        mn = fmin(mn, o[0]);
        mx = fmax(mx, o[1]);
        isn = isn || o[2] != 0.0;
    }
    offsets[f] = isn ? __builtin_nan("") : (mn + mx) / 2.0;
}

// ---------------------------------------------------------------------------
// grid position of a point: xc = (p - x0)/step per axis, dropped when outside
// [0, size-1) on any axis (interpolation.py:247-250).  NaN coordinates are
// dropped as well (the reference's int(NaN) is undefined).

struct PointPos {
    int64_t i[3];
    double w[3];
};

__device__ __forceinline__ bool point_pos(const FieldDev &F, const double *__restrict__ p, PointPos &pp)
{
    bool inside = true;
    for (int m = 0; m < 3; ++m) {
        pp.i[m] = 0;
        pp.w[m] = 0.0;
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        if (m < F.dim) {
            const double c = (p[m] - F.x0[m]) / F.step[m];
            if (!(c >= 0.0) || c >= (double)(F.size[m] - 1)) inside = false;
            const int64_t ci = inside ? (int64_t)c : 0;
            pp.i[m] = ci;
            pp.w[m] = c - (double)ci;
        }
    }
    return inside;
}

// bilinear / trilinear corner weight of corner (bx, by, bz), formed left to right as
// in the reference, e.g. (1.0-xw)*(1.0-yw) at interpolation.py:256
__device__ __forceinline__ double corner_weight(int dim, const PointPos &pp, int64_t bx, int64_t by, int64_t bz)
{
    double w = bx ? pp.w[0] : (1.0 - pp.w[0]);
    if (dim >= 2) w = w * (by ? pp.w[1] : (1.0 - pp.w[1]));
    if (dim >= 3) w = w * (bz ? pp.w[2] : (1.0 - pp.w[2]));
    return w;
}

// One (key, point) pair per point: the key is the point's base cell (the grid cell of its
// lower corner), kInvalidKey for points the injection drops.  The corners of a point lie in
// the base row (x-line) and the rows above it; counts per base row give, after the scan,
// each row's range in the sorted pairs.
__device__ __forceinline__ void point_keys_block(const FieldDev *__restrict__ fields,
                                                 const int64_t *__restrict__ pt_prefix, int nfields, int64_t npoints,
                                                 const double *__restrict__ pts, uint32_t *__restrict__ keys,
                                                 uint32_t *__restrict__ vals, uint32_t *__restrict__ row_count,
                                                 uint32_t *__restrict__ slot, uint32_t *__restrict__ rowg,
                                                 PointRec *__restrict__ rec, int64_t blk)
{
    // row_count (may be null): per-row point counts (integer atomics).  slot/rowg (row buckets): the point's
    // position among its row's points as the atomics handed them out, and its global row
    // (kInvalidKey outside the grid); vals (point indices) is then not written.
    const int64_t i = blk * kBlock + threadIdx.x;
    if (i >= npoints) return;
    const int f = find_seg(pt_prefix, nfields, i);
    const FieldDev &F = fields[f];
    PointPos pp;
    const bool inside = point_pos(F, pts + i * F.dim, pp);
    if (!slot) vals[i] = (uint32_t)i;
    if (!inside) {
        keys[i] = kInvalidKey;
        if (rowg) rowg[i] = kInvalidKey;
        return;
    }
    const int64_t row = pp.i[2] * F.size[1] + pp.i[1];
    keys[i] = (uint32_t)(F.cell_off + row * F.size[0] + pp.i[0]);
    if (rec) rec[i] = PointRec{{pp.w[0], pp.w[1], pp.w[2]}, (int32_t)pp.i[0], 0};
    if (row_count) {
        const uint32_t sl = atomicAdd(&row_count[F.xline_off + row], 1u);
        if (slot) {
            slot[i] = sl;
            rowg[i] = (uint32_t)(F.xline_off + row);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_point_keys(const FieldDev *__restrict__ fields,
                                                       const int64_t *__restrict__ pt_prefix, int nfields,
                                                       int64_t npoints, const double *__restrict__ pts,
                                                       uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                       uint32_t *__restrict__ row_count,
                                                       uint32_t *__restrict__ slot, uint32_t *__restrict__ rowg,
                                                       PointRec *__restrict__ rec)
{
    point_keys_block(fields, pt_prefix, nfields, npoints, pts, keys, vals, row_count, slot, rowg, rec, blockIdx.x);
}

// k_point_keys and k_offsets in one launch (they are independent): the first nkb workgroups
// take the points, the next nfields the fields' offsets
__global__ __launch_bounds__(kBlock) void k_point_keys_offsets(const FieldDev *__restrict__ fields,
                                                               const int64_t *__restrict__ pt_prefix, int nfields,
                                                               int64_t npoints, const double *__restrict__ pts,
                                                               uint32_t *__restrict__ keys,
                                                               uint32_t *__restrict__ vals,
                                                               uint32_t *__restrict__ row_count,
                                                               uint32_t *__restrict__ slot,
                                                               uint32_t *__restrict__ rowg,
                                                               PointRec *__restrict__ rec, int64_t nkb,
                                                               const double *__restrict__ val,
                                                               double *__restrict__ offsets)
{
    if ((int64_t)blockIdx.x < nkb)
        point_keys_block(fields, pt_prefix, nfields, npoints, pts, keys, vals, row_count, slot, rowg, rec,
                         blockIdx.x);
    else
        offsets_block(fields, val, offsets, (int)(blockIdx.x - nkb));
}

// k_point_keys for large point sets: each workgroup takes ppt*kBlock consecutive points and,
// when the rows of the fields they belong to number at most nbins, counts them per row in
// LDS first; one global atomic per row the workgroup touches then reserves the row's slots for
// all of them (instead of one contended atomic per point), and each point's slot is that base
// plus its rank among the workgroup's points of the row.  The slots come in another order than
// with one atomic per point; the row sort orders every row by (base cell, point index), so the
// sorted list is the same.  Workgroups whose rows do not fit count with one atomic per point.
constexpr int kKeysMaxPpt = 16;     // points per thread, at most
constexpr int kKeysBins = 8192;     // LDS row counters of a workgroup (32 KB) ...
constexpr int kKeysBinsMax = 16384; // ... or, for plans with fields of up to this many rows, 64 KB

__global__ __launch_bounds__(kBlock) void k_point_keys_hist(const FieldDev *__restrict__ fields,
                                                            const int64_t *__restrict__ pt_prefix, int nfields,
                                                            int64_t npoints, int ppt, int nbins,
                                                            const double *__restrict__ pts,
                                                            uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                            uint32_t *__restrict__ row_count,
                                                            uint32_t *__restrict__ slot, uint32_t *__restrict__ rowg,
                                                            PointRec *__restrict__ rec)
{
    extern __shared__ uint32_t hist[];   // nbins row counters (dynamic LDS)
    __shared__ int64_t s_rbase;
    __shared__ int s_span;
    const int64_t p0 = (int64_t)blockIdx.x * ppt * kBlock;
    const int64_t p1 = min(p0 + (int64_t)ppt * kBlock, npoints);
    if (threadIdx.x == 0) {
        const FieldDev &A = fields[find_seg(pt_prefix, nfields, p0)];
        const FieldDev &B = fields[find_seg(pt_prefix, nfields, p1 - 1)];
        const int64_t r0 = A.xline_off;
        const int64_t span = B.xline_off + B.size[1] * B.size[2] - r0;
        s_rbase = r0;
        s_span = span <= nbins ? (int)span : -1;
    }
    __syncthreads();
    const int64_t rbase = s_rbase;
    const int span = s_span;
    for (int b = threadIdx.x; b < span; b += kBlock) hist[b] = 0u;
    __syncthreads();
    int rr[kKeysMaxPpt];
    uint32_t lr[kKeysMaxPpt];
    // a thread's points ascend: its field is found once and then advanced, and the field's
    // geometry stays in registers until its points pass into the next field
    int f = p0 + threadIdx.x < p1 ? find_seg(pt_prefix, nfields, p0 + threadIdx.x) : 0;
    int64_t fnext = f + 1 < nfields ? pt_prefix[f + 1] : INT64_MAX;
    const int dim = fields[0].dim;   // the same for every field of a plan
    double gx0[3], gst[3];
    int64_t gsz[3], gcell = 0, gxl = 0;
    auto geometry = [&](int ff) {
        const FieldDev &F = fields[ff];
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            gx0[m] = F.x0[m];
            gst[m] = F.step[m];
            gsz[m] = F.size[m];
        }
        gcell = F.cell_off;
        gxl = F.xline_off;
    };
    geometry(f);
    // kKeysGrp points at a time: their coordinates are loaded first, all in flight together
    constexpr int kKeysGrp = 4;
#pragma unroll
    for (int k0 = 0; k0 < kKeysMaxPpt; k0 += kKeysGrp) {
        double c[kKeysGrp][3];
#pragma unroll
        for (int u = 0; u < kKeysGrp; ++u) {
            const int64_t i = p0 + (int64_t)(k0 + u) * kBlock + threadIdx.x;
            const bool ok = k0 + u < ppt && i < p1;
#pragma unroll
            for (int m = 0; m < 3; ++m) c[u][m] = ok && m < dim ? pts[i * dim + m] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kKeysGrp; ++u) {
            const int k = k0 + u;
            rr[k] = -1;
            lr[k] = 0u;
            const int64_t i = p0 + (int64_t)k * kBlock + threadIdx.x;
            if (k >= ppt || i >= p1) continue;
            if (i >= fnext) {
                while (i >= fnext) {
                    ++f;
                    fnext = f + 1 < nfields ? pt_prefix[f + 1] : INT64_MAX;
                }
                geometry(f);
            }
            // point_pos on the cached geometry
            bool inside = true;
            int64_t ci[3] = {0, 0, 0};
            double w[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int m = 0; m < 3; ++m) {
                if (m < dim) {
                    const double cc = (c[u][m] - gx0[m]) / gst[m];
                    if (!(cc >= 0.0) || cc >= (double)(gsz[m] - 1)) inside = false;
                    const int64_t q = inside ? (int64_t)cc : 0;
                    ci[m] = q;
                    w[m] = cc - (double)q;
                }
            }
            if (!slot) vals[i] = (uint32_t)i;
            if (!inside) {
                keys[i] = kInvalidKey;
                if (rowg) rowg[i] = kInvalidKey;
                continue;
            }
            const int64_t row = ci[2] * gsz[1] + ci[1];
            keys[i] = (uint32_t)(gcell + row * gsz[0] + ci[0]);
            if (rec) rec[i] = PointRec{{w[0], w[1], w[2]}, (int32_t)ci[0], 0};
            const int64_t grow = gxl + row;
            if (span >= 0) {
                rr[k] = (int)(grow - rbase);
                lr[k] = atomicAdd(&hist[rr[k]], 1u);
            } else {
                const uint32_t sl = atomicAdd(&row_count[grow], 1u);
                if (slot) {
                    slot[i] = sl;
                    rowg[i] = (uint32_t)grow;
                }
            }
        }
    }
    if (span < 0) return;   // uniform for the workgroup
    __syncthreads();
    // one global atomic per touched row; the LDS counter becomes the row's first slot
    for (int b = threadIdx.x; b < span; b += kBlock) {
        const uint32_t c = hist[b];
        if (c) hist[b] = atomicAdd(&row_count[rbase + b], c);
    }
    if (!slot) return;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kKeysMaxPpt; ++k) {
        if (rr[k] < 0) continue;
        const int64_t i = p0 + (int64_t)k * kBlock + threadIdx.x;
        slot[i] = hist[rr[k]] + lr[k];
        rowg[i] = (uint32_t)(rbase + rr[k]);
    }
}

// The sub-streams of x-line (y, z): for q < 2*nrows, the points of base row (y - dy, z - dz)
// contribute their corner (bx, dy, dz) at x = base x + bx, with bx = q & 1, dy = (q >> 1) & 1,
// dz = q >> 2.  Within a base row the sorted points ascend in base x (point order for equal
// x), so every sub-stream ascends in target cell.
__device__ __forceinline__ int source_rows(int dim) { return dim >= 3 ? 4 : (dim == 2 ? 2 : 1); }

__device__ __forceinline__ void line_streams(const FieldDev &F, int64_t line, bool enable,
                                             const uint32_t *__restrict__ row_start,
                                             const uint32_t *__restrict__ row_count, uint32_t (&rs)[8],
                                             uint32_t (&rc)[8])
{
    // sub-stream q reads base row (y - dy, z - dz) with dy = (q >> 1) & 1, dz = q >> 2 for its
    // corner column bx = q & 1: both corner columns of a row share the row's range, loaded
    // once; (y, z) are formed once per line (a 2D line index is y itself)
    const int nsub = 2 * source_rows(F.dim);
    const int64_t H = F.size[1];
    const int64_t y = F.dim >= 3 ? line % H : line;
    const int64_t z = F.dim >= 3 ? line / H : 0;
#pragma unroll
    for (int q = 0; q < 8; q += 2) {
        const int dy = (q >> 1) & 1, dz = q >> 2;
        uint32_t st = 0, ct = 0;
        if (enable && q < nsub && y - dy >= 0 && z - dz >= 0) {
            const int64_t row = line - dy - dz * H;
            st = row_start[F.xline_off + row];
            ct = row_count[F.xline_off + row];
        }
        rs[q] = rs[q + 1] = st;
        rc[q] = rc[q + 1] = ct;
    }
}

// ---------------------------------------------------------------------------
// FP64 path, reference operation order

// Each sorted entry's injection record, in list order (the merge then reads one record per
// step): base column, point, position within the base cell and value minus the offset, as
// point_pos forms them.  The per-field sort (`local`) writes a field's in-grid points from its
// first point index on and leaves the slots after them unused; the row-bucket sort packs all
// fields' entries by row.  Slots past the list's last entry are skipped.
__global__ __launch_bounds__(kBlock) void k_inject_records(
    const FieldDev *__restrict__ fields, const int64_t *__restrict__ pt_prefix, int nfields, int64_t npoints,
    int local, int64_t nxlines, const uint32_t *__restrict__ vals, const uint32_t *__restrict__ row_start,
    const uint32_t *__restrict__ row_count, const double *__restrict__ pts, const double *__restrict__ val,
    const double *__restrict__ offsets, InjectRec *__restrict__ irec)
{
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= npoints) return;
    int64_t last = nxlines - 1;   // the list's last row
    if (local) {
        const FieldDev &Fe = fields[find_seg(pt_prefix, nfields, e)];
        last = Fe.xline_off + Fe.size[1] * Fe.size[2] - 1;   // the last row of the slot's field
    }
    if (e >= (int64_t)row_start[last] + row_count[last]) return;
    const uint32_t p = vals[e];
    const int f = find_seg(pt_prefix, nfields, p);
    const FieldDev &F = fields[f];
    PointPos pp;
    point_pos(F, pts + (int64_t)p * F.dim, pp);
    irec[e] = InjectRec{(int)pp.i[0], p, pp.w[0], pp.w[1], pp.w[2], val[p] - offsets[f]};
}

// One thread per x-line: merges the line's sub-streams by (x, point index) and adds
// w*val_c and w to the FP64 grids in that order, which is the per-cell order of the
// reference's point loop (each cell gets its contributions in increasing point index).
// The contributions of a cell are consecutive in the merged order, so a cell's two sums are
// held in registers and stored once; they start at 0.0 when the grids were cleared (zeroed),
// else at the grids' values.  One wave per workgroup, as k_exact_lines: a field has only as
// many threads as x lines.
//
// The merge is a chain of dependent steps, one entry record read per step.  When a
// workgroup's lines lie in one 1D or 2D field and their rows (each line's own and, in 2D,
// the row below) hold at most stage_cap points, the wave first copies those records to LDS
// (dynamic, stage_cap records: kInjectStage for plans of 1D and 2D fields, none for 3D) and
// the merge reads them there.  A field's rows are contiguous in the sorted list (each
// row starts where the previous one ends), so the workgroup's records are one range of it.
constexpr int kExactInjectThreads = 64;
constexpr int kInjectStage = 1024;

// head of sub-stream q at entry e: target x, point, corner weight, value (x = INT_MAX when
// the stream is exhausted); STAGED reads the record from LDS.  q may differ between the
// lanes: its corner (q & 1, (q >> 1) & 1, q >> 2) selects the weight's factors.
template <bool STAGED>
__device__ __forceinline__ void inject_head(const FieldDev &F, int q, uint32_t e, uint32_t end,
                                            const InjectRec *rec, uint32_t e0,
                                            const InjectRec *__restrict__ irec, int &x, uint32_t &p,
                                            double &w, double &v)
{
    if (e >= end) {
        x = INT_MAX;
        p = 0;
        w = v = 0.0;
        return;
    }
    const InjectRec &r = STAGED ? rec[e - e0] : irec[e];
    PointPos pp;
    pp.i[0] = r.x;
    pp.i[1] = pp.i[2] = 0;
    pp.w[0] = r.wx;
    pp.w[1] = r.wy;
    pp.w[2] = r.wz;
    p = r.p;
    v = r.vc;
    x = r.x + (q & 1);
    w = corner_weight(F.dim, pp, q & 1, (q >> 1) & 1, q >> 2);
}

// the merge of one line's NS sub-streams (4 in 1D and 2D, 8 in 3D).  Each step takes the
// smallest head and reads the next entry of that stream only: one read per step whichever
// stream each lane advances.  A cell's sums start at 0.0 (ZEROED) or at the grids' values,
// and are stored when the merge moves past the cell.
template <int NS, bool STAGED, bool ZEROED>
__device__ __forceinline__ void inject_merge(const FieldDev &F, uint32_t (&pos)[8], const uint32_t (&end)[8],
                                             const InjectRec *rec, uint32_t e0,
                                             const InjectRec *__restrict__ irec, double *__restrict__ vg,
                                             double *__restrict__ wg, int64_t base)
{
    // hr: entries left in the stream after its head (loop-carried, like the heads, so that
    // taking the advanced stream's values stays a chain of selects, not an indexed array)
    uint32_t hp[NS], hr[NS];
    int hx[NS];
    double hw[NS], hv[NS];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
        inject_head<STAGED>(F, q, pos[q], end[q], rec, e0, irec, hx[q], hp[q], hw[q], hv[q]);
        hr[q] = pos[q] < end[q] ? end[q] - pos[q] - 1 : 0;
    }
    int cx = INT_MAX;
    double av = 0.0, aw = 0.0;
    while (true) {
        int best = 0;
#pragma unroll
        for (int q = 1; q < NS; ++q)
            if (hx[q] < hx[best] || (hx[q] == hx[best] && hp[q] < hp[best])) best = q;
        int bx = hx[0];
        double bw = hw[0], bv = hv[0];
        uint32_t e = pos[0] + 1, r = hr[0];
#pragma unroll
        for (int q = 1; q < NS; ++q)
            if (q == best) {
                bx = hx[q];
                bw = hw[q];
                bv = hv[q];
                e = pos[q] + 1;
                r = hr[q];
            }
        // every stream exhausted (at once for a line without points): store the last cell
        if (bx == INT_MAX) {
            if (cx != INT_MAX) {
                vg[base + cx] = av;
                wg[base + cx] = aw;
            }
            break;
        }
        int nx;
        uint32_t np;
        double nw, nv;
        inject_head<STAGED>(F, best, e, r ? e + 1 : e, rec, e0, irec, nx, np, nw, nv);
#pragma unroll
        for (int q = 0; q < NS; ++q)
            if (q == best) {
                pos[q] = e;
                hr[q] = r ? r - 1 : 0;
                hx[q] = nx;
                hp[q] = np;
                hw[q] = nw;
                hv[q] = nv;
            }
        if (bx != cx) {
            if (cx != INT_MAX) {
                vg[base + cx] = av;
                wg[base + cx] = aw;
            }
            cx = bx;
            av = ZEROED ? 0.0 : vg[base + cx];
            aw = ZEROED ? 0.0 : wg[base + cx];
        }
        av += bw * bv;
        aw += bw;
    }
}

__global__ __launch_bounds__(kExactInjectThreads) void k_exact_inject(
    const FieldDev *__restrict__ fields, const int64_t *__restrict__ xline_prefix, int nfields, int64_t nxlines,
    const InjectRec *__restrict__ irec, const uint32_t *__restrict__ row_start,
    const uint32_t *__restrict__ row_count, double *__restrict__ vg, double *__restrict__ wg, int zeroed,
    int stage_cap)
{
    extern __shared__ InjectRec rec[];
    const int64_t g0 = (int64_t)blockIdx.x * kExactInjectThreads;
    const int64_t glast = min(g0 + kExactInjectThreads - 1, nxlines - 1);
    const int64_t gl = g0 + threadIdx.x;
    const bool live = gl < nxlines;
    const int f = find_seg(xline_prefix, nfields, live ? gl : glast);
    const FieldDev &F = fields[f];
    const int64_t line = (live ? gl : glast) - F.xline_off;
    const int64_t base = F.cell_off + line * F.size[0];
    uint32_t pos[8], end[8];
    {
        uint32_t rs[8], rc[8];
        line_streams(F, line, live, row_start, row_count, rs, rc);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            pos[q] = rs[q];
            end[q] = rs[q] + rc[q];
        }
    }
    // stage the workgroup's records when its lines share one 1D/2D field (then F is the same
    // for every lane and both conditions are uniform) and every stream lies in their range
    bool staged = false;
    uint32_t e0 = 0;
    if (find_seg(xline_prefix, nfields, g0) == find_seg(xline_prefix, nfields, glast) && F.dim <= 2) {
        const int64_t lo = (F.dim == 2 && g0 > F.xline_off) ? g0 - 1 : g0;
        e0 = row_start[lo];
        const uint32_t e1 = row_start[glast] + row_count[glast];
        bool ok = e0 <= e1 && e1 - e0 <= (uint32_t)stage_cap;
#pragma unroll
        for (int q = 0; q < 8; ++q) ok = ok && (pos[q] == end[q] || (pos[q] >= e0 && end[q] <= e1));
        staged = __all(ok);
        if (staged) {
            for (uint32_t e = e0 + threadIdx.x; e < e1; e += kExactInjectThreads) rec[e - e0] = irec[e];
            __syncthreads();
        }
    }
    if (!live) return;
    // staged lines are 1D or 2D (4 streams)
    if (staged) {
        if (zeroed)
            inject_merge<4, true, true>(F, pos, end, rec, e0, irec, vg, wg, base);
        else
            inject_merge<4, true, false>(F, pos, end, rec, e0, irec, vg, wg, base);
    } else if (F.dim <= 2) {
        if (zeroed)
            inject_merge<4, false, true>(F, pos, end, rec, e0, irec, vg, wg, base);
        else
            inject_merge<4, false, false>(F, pos, end, rec, e0, irec, vg, wg, base);
    } else {
        if (zeroed)
            inject_merge<8, false, true>(F, pos, end, rec, e0, irec, vg, wg, base);
        else
            inject_merge<8, false, false>(F, pos, end, rec, e0, irec, vg, wg, base);
    }
}

// _accumulate_tail_array (interpolation.py:484-533) and _accumulate_array (:729-772)
// on one strided line, ping-ponging between the grid line G and a scratch line S.
constexpr int kExactBlk = 32;  // interior steps of exact_line whose loads are issued together

// UNIT: element stride 1 (rows), known at compile time so that adjacent loads merge
template <bool UNIT>
__device__ void exact_line(double *__restrict__ G, double *__restrict__ S, int64_t L, int64_t st_, int64_t rect_len,
                           int num_iter, int tail, double alpha)
{
    const int64_t st = UNIT ? 1 : st_;
    const int64_t h0 = (rect_len - 1) / 2;
    const int64_t h0_1 = h0 + 1;
    const int64_t h1 = rect_len - h0;
    double *in = G;
    double *h = S;
#define IN(i) in[(i) * st]
#define H_(i) h[(i) * st]
    for (int it = 0; it < num_iter; ++it) {
        double accu = 0.0;
        int64_t k;
        if (tail) {
            for (k = -h0; k < 0; ++k) accu += IN(k + h0);
            for (k = 0; k < h1; ++k) {
                accu += IN(k + h0);
                H_(k) = accu + alpha * IN(k + h0_1);
            }
            // the interior in blocks of kExactBlk steps: the block's input terms are loaded
            // first (they do not depend on accu; `in` and `h` are different arrays), then the
            // steps run in the reference's order, so the memory latency is paid once per block
            for (k = h1; k + kExactBlk <= L - h0_1; k += kExactBlk) {
                double a[kExactBlk], b[kExactBlk + 1];
#pragma unroll
                for (int j = 0; j < kExactBlk; ++j) {
                    a[j] = IN(k + j + h0);
                    b[j] = IN(k + j - h1);
                }
                b[kExactBlk] = IN(k + kExactBlk + h0);   // IN(k + j + h0_1) is a[j + 1]
#pragma unroll
                for (int j = 0; j < kExactBlk; ++j) {
                    const double nxt = j + 1 < kExactBlk ? a[j + 1] : b[kExactBlk];
                    accu += (a[j] - b[j]);
                    H_(k + j) = accu + alpha * (b[j] + nxt);
                }
            }
            for (; k < L - h0_1; ++k) {
                accu += (IN(k + h0) - IN(k - h1));
                H_(k) = accu + alpha * (IN(k - h1) + IN(k + h0_1));
            }
            k = L - h0_1;
            accu += (IN(k + h0) - IN(k - h1));
            H_(k) = accu + alpha * IN(k - h1);
            for (k = L - h0; k < L; ++k) {
                accu -= IN(k - h1);
                H_(k) = accu + alpha * IN(k - h1);
            }
        } else {
            for (k = -h0; k < 0; ++k) accu += IN(k + h0);
            for (k = 0; k < h1; ++k) {
                accu += IN(k + h0);
                H_(k) = accu;
            }
            for (k = h1; k + kExactBlk <= L - h0; k += kExactBlk) {
                double a[kExactBlk], b[kExactBlk];
#pragma unroll
                for (int j = 0; j < kExactBlk; ++j) {
                    a[j] = IN(k + j + h0);
                    b[j] = IN(k + j - h1);
                }
#pragma unroll
                for (int j = 0; j < kExactBlk; ++j) {
                    accu += (a[j] - b[j]);
                    H_(k + j) = accu;
                }
            }
            for (; k < L - h0; ++k) {
                accu += (IN(k + h0) - IN(k - h1));
                H_(k) = accu;
            }
            for (k = L - h0; k < L; ++k) {
                accu -= IN(k - h1);
                H_(k) = accu;
            }
        }
        double *t = in; in = h; h = t;
    }
    if (in != G)
        for (int64_t i = 0; i < L; ++i) G[i * st] = in[i * st];
#undef IN
#undef H_
}

// One wave per workgroup: a field's few thousand line chains then spread over as many CUs as
// they have waves instead of sharing a quarter as many CUs' memory pipelines.
constexpr int kExactLineThreads = 64;

__global__ __launch_bounds__(kExactLineThreads) void k_exact_lines(const FieldDev *__restrict__ fields,
                                                        const int64_t *__restrict__ line_prefix, int nfields,
                                                        int64_t nslots, int axis, int rect_override,
                                                        double *__restrict__ vg, double *__restrict__ wg,
                                                        double *__restrict__ vs, double *__restrict__ ws)
{
    const int64_t g = (int64_t)blockIdx.x * kExactLineThreads + threadIdx.x;
    if (g >= nslots) return;
    const int f = find_seg(line_prefix, nfields, g);
    const FieldDev &F = fields[f];
    const int64_t r = g - line_prefix[f];
    const int which = (int)(r & 1);
    const int64_t l = r >> 1;
    const int64_t W = F.size[0], H = F.size[1], D = F.size[2];
    int64_t L, st, base;
    if (axis == 0) {
        L = W; st = 1; base = l * W;
    } else if (axis == 1) {
        L = H; st = W; base = (l / W) * W * H + (l % W);
    } else {
        L = D; st = W * H; base = l;
    }
    const int64_t rect = rect_override > 0 ? rect_override : 2 * (int64_t)F.T[axis] + 1;
    double *G = (which ? wg : vg) + F.cell_off + base;
    double *S = (which ? ws : vs) + F.cell_off + base;
    if (st == 1) exact_line<true>(G, S, L, st, rect, F.num_iter, F.tail, F.alpha[axis]);
    else exact_line<false>(G, S, L, st, rect, F.num_iter, F.tail, F.alpha[axis]);
}

// FP64 x lines through LDS.  One wave takes 32 consecutive x lines of one field, v and w each
// (64 (line, field) slots, lane = slot).  Each iteration streams the slots' rows in tiles of
// kRowTile columns: coalesced loads (two rows of 32 doubles per instruction) into a ring of
// `ring` tiles per slot in LDS, one tile ahead of use, the line's steps from the ring in the
// reference's order (exact_line's phases, interpolation.py:503-523), the outputs through an
// LDS tile and coalesced stores into the other buffer (ping-pong, as exact_line).  One lane
// runs one line's running sum exactly as one thread of k_exact_lines does, so the results are
// the same bits; the rows are read 512 bytes per instruction instead of 64 lines of 8 bytes.
// ring: a power of two >= h0/C + ceil(h1/C) + 2 tiles (C = kRowTile), which holds every term
// the chunk of outputs [i*C, (i+1)*C) reads (tiles (i*C - h1)/C .. ((i+1)*C + h0)/C) while
// the next tile loads.  Slot stride in LDS: ring*C + 1 doubles (odd: the lanes' reads of one
// column fall on different banks).

template <bool TAIL, int S, bool MIR>
__device__ __forceinline__ void exact_rows(const FieldDev &F, int64_t s0, int nsl, int rect_override, int ring,
                                           double *__restrict__ rl, double *__restrict__ vg,
                                           double *__restrict__ wg, double *__restrict__ vs,
                                           double *__restrict__ ws)
{
    constexpr int C = kRowTile;
    constexpr int NP = S / 2;   // slot pairs: a load or store instruction takes two rows
    const int lane = threadIdx.x;
    const int64_t L = F.size[0];
    const int64_t rect = rect_override > 0 ? rect_override : 2 * (int64_t)F.T[0] + 1;
    const int64_t h0 = (rect - 1) / 2, h0_1 = h0 + 1, h1 = rect - h0;
    const double alpha = F.alpha[0];
    const int mask = ring * C - 1;             // the ring holds ring*C columns: a power of two
    const int64_t RS = exact_rows_slot_stride(ring, MIR);
    const int rlen = ring * C;                 // the ring's positions; with MIR, [rlen, rlen + C] repeat [0, C]
    double *ot = rl + S * RS;                  // output tile, slot stride C + 1
    const int64_t NT = (L + C - 1) / C;
    // loader / storer mapping: instruction j takes slot 2j + (lane >> 5), column lane & 31; the
    // loads clamp row and column into the field (values past them are never used)
    const int wl = lane >> 5, cl = lane & 31;
    const int64_t line0 = s0 >> 1;
    const int jmax = max(0, (nsl - 1 - wl) >> 1);
    double *const bufG = (wl ? wg : vg) + F.cell_off + line0 * L;
    double *const bufS = (wl ? ws : vs) + F.cell_off + line0 * L;
    const double *my = rl + lane * RS;
    const bool act = lane < nsl;

    // the tile in flight: slot 2j + wl, column t*C + cl
    double pa_[NP];
    auto fetch = [&](double (&pre)[NP], const double *__restrict__ g, int64_t t) {
        const int64_t c = min(t * C + cl, L - 1);
#pragma unroll
        for (int j = 0; j < NP; ++j) pre[j] = g[(int64_t)min(j, jmax) * L + c];
    };
    auto put = [&](const double (&pre)[NP], int64_t t) {
        const int p = (int)(t * C + cl) & mask;
#pragma unroll
        for (int j = 0; j < NP; ++j) rl[(2 * j + wl) * RS + p] = pre[j];
        if (MIR && p <= C) {
#pragma unroll
            for (int j = 0; j < NP; ++j) rl[(2 * j + wl) * RS + rlen + p] = pre[j];
        }
    };
    auto in_at = [&](int64_t p) { return my[(int)p & mask]; };
    for (int it = 0; it < F.num_iter; ++it) {
        const double *__restrict__ src = (it & 1) ? bufS : bufG;
        double *__restrict__ dst = (it & 1) ? bufG : bufS;
        // tiles the chunk of outputs [i*C, (i+1)*C) reads: up to ((i+1)*C + h0) / C; the ring
        // receives one more tile after each chunk, loaded while the chunk runs
        const int64_t need0 = min(NT - 1, (C + h0) / C);
        for (int64_t t = 0; t <= need0; ++t) {
            fetch(pa_, src, t);
            put(pa_, t);
        }
        int64_t next = need0 + 1;   // the next tile to put
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        double accu = 0.0;
        if (act)
            for (int64_t k = -h0; k < 0; ++k) accu += in_at(k + h0);
        auto chunk = [&](int64_t i) {
            const int64_t k0 = i * C, k1 = min(k0 + C, L);
            if (act) {
                double *o = ot + lane * (C + 1);
                if (k0 >= h1 && k0 + C <= L - (TAIL ? h0_1 : h0)) {
                    // interior chunk: its terms into registers first (one LDS wait), then the
                    // steps in the reference's order
                    double a[C + 1], b[C];
                    const int pa = (int)(k0 + h0), pb = (int)(k0 - h1);
                    if (MIR) {   // positions (pa & mask) .. + C, read past the ring's end from the mirror
                        const double *ma = my + (pa & mask), *mb = my + (pb & mask);
#pragma unroll
                        for (int j = 0; j <= C; ++j) a[j] = ma[j];
#pragma unroll
                        for (int j = 0; j < C; ++j) b[j] = mb[j];
                    } else {
#pragma unroll
                        for (int j = 0; j <= C; ++j) a[j] = my[(pa + j) & mask];
#pragma unroll
                        for (int j = 0; j < C; ++j) b[j] = my[(pb + j) & mask];
                    }
#pragma unroll
                    for (int j = 0; j < C; ++j) {
                        accu += (a[j] - b[j]);
                        o[j] = TAIL ? accu + alpha * (b[j] + a[j + 1]) : accu;
                    }
                } else {
                    for (int64_t k = k0; k < k1; ++k) {
                        double r;
                        if (TAIL) {
                            if (k < h1) {
                                accu += in_at(k + h0);
                                r = accu + alpha * in_at(k + h0_1);
                            } else if (k < L - h0_1) {
                                const double bk = in_at(k - h1);
                                accu += (in_at(k + h0) - bk);
                                r = accu + alpha * (bk + in_at(k + h0_1));
                            } else if (k == L - h0_1) {
                                const double bk = in_at(k - h1);
                                accu += (in_at(k + h0) - bk);
                                r = accu + alpha * bk;
                            } else {
                                const double bk = in_at(k - h1);
                                accu -= bk;
                                r = accu + alpha * bk;
                            }
                        } else {
                            if (k < h1) accu += in_at(k + h0);
                            else if (k < L - h0) accu += (in_at(k + h0) - in_at(k - h1));
                            else accu -= in_at(k - h1);
                            r = accu;
                        }
                        o[k - k0] = r;
                    }
                }
            }
        };
        // the chunk's outputs to the other buffer (the storer mapping reads other lanes' rows)
        auto store = [&](int64_t i) {
            const int64_t c = i * C + cl;
            double v[NP];   // all LDS reads first (one wait), then the stores
#pragma unroll
            for (int j = 0; j < NP; ++j) v[j] = ot[(2 * j + wl) * (C + 1) + cl];
            if (c < L) {
                if (nsl == S) {   // every slot of the workgroup: stores without a branch each
#pragma unroll
                    for (int j = 0; j < NP; ++j) dst[(int64_t)j * L + c] = v[j];
                } else {
#pragma unroll
                    for (int j = 0; j < NP; ++j)
                        if (j <= jmax && 2 * j + wl < nsl) dst[(int64_t)j * L + c] = v[j];
                }
            }
        };
        // per chunk: the next tile's loads issued, the steps, the tile into the ring (its wait
        // comes after the steps), the chunk's stores
        for (int64_t i = 0; i < NT; ++i) {
            const bool more = next < NT;
            if (more) fetch(pa_, src, next);
            chunk(i);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            if (more) put(pa_, next++);
            store(i);
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    }
    if (F.num_iter & 1) {
        // the result is in S: copy it to G a tile at a time (each lane its own stores' addresses)
        for (int64_t t = 0; t < NT; ++t) {
            fetch(pa_, bufS, t);
            const int64_t c = t * C + cl;
#pragma unroll
            for (int j = 0; j < NP; ++j)
                if (2 * j + wl < nsl && c < L) bufG[(int64_t)j * L + c] = pa_[j];
        }
    }
}

// One wave per workgroup: workgroup b takes the S slots from S*(b - wg_prefix[f]) on of field
// f (slot 2l + c: x line l, field c = v, w of the field's 2*H*D slots).  S = 32 or 16 (lanes
// past S load and store, but run no line) spreads the few lines of a small plan over more CUs,
// each chunk then moving a half or a quarter of the bytes.
template <int S, bool MIR>
__global__ __launch_bounds__(64) void k_exact_rows(const FieldDev *__restrict__ fields,
                                                   const int64_t *__restrict__ wg_prefix,
                                                   const int64_t *__restrict__ slot_prefix, int nfields,
                                                   int rect_override, int ring, double *__restrict__ vg,
                                                   double *__restrict__ wg, double *__restrict__ vs,
                                                   double *__restrict__ ws)
{
    extern __shared__ double rl[];
    const int f = find_seg(wg_prefix, nfields, blockIdx.x);
    const FieldDev &F = fields[f];
    const int64_t s0 = (blockIdx.x - wg_prefix[f]) * S;
    const int nsl = (int)min((int64_t)S, slot_prefix[f + 1] - slot_prefix[f] - s0);
    if (F.tail) exact_rows<true, S, MIR>(F, s0, nsl, rect_override, ring, rl, vg, wg, vs, ws);
    else exact_rows<false, S, MIR>(F, s0, nsl, rect_override, ring, rl, vg, wg, vs, ws);
}

// wg[wg < thr] = NaN, then (vg/wg + offset).astype(float32): interpolation.py:428-430, :367
__global__ __launch_bounds__(kBlock) void k_exact_finalize(const FieldDev *__restrict__ fields,
                                                           const int64_t *__restrict__ cell_prefix, int nfields,
                                                           int64_t ncells, const double *__restrict__ vg,
                                                           const double *__restrict__ wg,
                                                           const double *__restrict__ offsets, float *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= ncells) return;
    const int f = find_seg(cell_prefix, nfields, g);
    const FieldDev &F = fields[f];
    const int64_t i = g - F.cell_off;
    if (i >= F.size[0] * F.size[1] * F.size[2]) return;   // alignment gap between fields
    const double w = wg[g];
    const double v = vg[g];
    out[F.out_off + i] = (w < F.thr) ? __builtin_nanf("") : (float)(v / w + offsets[f]);
}

__global__ __launch_bounds__(kBlock) void k_mask(const FieldDev *__restrict__ fields,
                                                 const int64_t *__restrict__ cell_prefix, int nfields, int64_t ncells,
                                                 double *__restrict__ wg, int fast)
{
    const int64_t g = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= ncells) return;
    const FieldDev &F = fields[find_seg(cell_prefix, nfields, g)];
    if (g - F.cell_off >= F.size[0] * F.size[1] * F.size[2]) return;   // alignment gap between fields
    const double thr = fast ? F.thr_fast : F.thr;
    if (wg[g] < thr) wg[g] = __builtin_nan("");
}

// ---------------------------------------------------------------------------
// FP32C path
//
// One workgroup processes `nlines` lines of one field held in LDS as float32,
// with T+1 zeros of padding on both sides of every line (the reference's zero
// padding, interpolation.py:503-523).  A line is split into lpl contiguous
// segments, one per lane; each lane re-anchors its window sum at its segment
// start with a direct (2T+1)-term sum in FP64 and slides it across the segment
// in FP64, so no running sum ever spans more than one segment and no FP32 sum
// is ever formed.  Each iteration is computed into registers (one field at a
// time) and written back in place after a barrier.  The last iteration scales
// by (1/(K+2*alpha))^n and, on the last pass, forms the masked FP64 quotient
// directly from the FP64 window sums of both fields.

// Segment lengths (elements per lane) the fast pass is instantiated for; odd, so the LDS
// lane stride is conflict-free.  The host picks one per pass and line length.
#define FB_FAST_SEGS(X) X(19) X(21) X(23) X(25) X(29) X(33) X(37) X(39) X(41) X(45) X(49)
constexpr int kFastBlock = 512;   // threads of a fast-pass workgroup: 256 per field
constexpr int kBlk = 4;           // add-stream block (LDS loads in flight)

// One iteration of the tailed box filter over the lane's segment starting at s of one
// field's line `lin`.  The subtract stream lin[s-T+j] (j < kSegArr) is read in one burst
// into registers and doubles as the first part of the initial window sum lin[s-T .. s+T];
// the add stream lin[s+T+1+j] is read in blocks of kBlk with one block of lookahead, so
// many LDS loads are in flight.  All kSegArr steps run unconditionally; steps past the
// lane's segment read in-bounds LDS (next line or the tail reserve) and their outputs are
// discarded by stage_store.  Results replace sb[j] (dead after step j: its value lives on
// in p), so the register array doubles as the staging buffer of the in-place update.
//   SCALE: multiply the results by `scale` (last iteration of a pass).
// Masked FP32C quotient v/w + offset of one node (interpolation.py:367, :428-430): a
// float32 reciprocal refined by one Newton step on the quotient (~1 ulp of v/w; w is
// normalised, well inside the float32 range), the offset added in FP64 before the final
// rounding to float32, and w < thr decided exactly as in FP64 through the float32 threshold.
__device__ __forceinline__ float fast_quotient(float vf, float wf, double off, float thr_f)
{
    const float r = __builtin_amdgcn_rcpf(wf);
    float q = vf * r;
    q = fmaf(fmaf(-q, wf, vf), r, q);
    return wf < thr_f ? __builtin_nanf("") : (float)((double)q + off);
}

template <int SEG, bool SCALE>
__device__ __forceinline__ void slide1(const float *__restrict__ lin, int s, int T, double alpha, double scale,
                                       float (&sb)[SEG])
{
    constexpr int kSegArr = SEG;
    constexpr int kNBlk = (SEG + kBlk - 1) / kBlk;   // blocks covering the SEG steps
    const int K = 2 * T + 1;
    const float *subp = lin + s - T;
#pragma unroll
    for (int j = 0; j < kSegArr; ++j) sb[j] = subp[j];
    const float p0 = lin[s - T - 1];
    // window sum: the first min(K, kSegArr) terms are sb[], the rest (K > kSegArr) from LDS,
    // four at a time without per-term conditions (K is the same for the whole workgroup)
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (K >= kSegArr) {
        int i = kSegArr;
        for (; i + 4 <= K; i += 4) {
            float x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = subp[i + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] += (double)x[u];
        }
        for (; i < K; ++i) v[0] += (double)subp[i];
#pragma unroll
        for (int j = 0; j < kSegArr; ++j) v[j & 3] += (double)sb[j];
    } else {
        // short kernels (2T+1 < kSegArr): direct sum from LDS
        int i = 0;
        for (; i + 4 <= K; i += 4) {
            float x[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) x[u] = subp[i + u];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] += (double)x[u];
        }
        for (; i < K; ++i) v[0] += (double)subp[i];
    }
    double a = (v[0] + v[1]) + (v[2] + v[3]);
    // Each step: output = a + alpha*(left + right tail), then a slides by the entering minus
    // the leaving term, exactly in FP64.  The output rounds a to float32 and adds the tails
    // in float32 (one float32 FMA; a float32 operation costs half an FP64 one): it is stored
    // as float32 anyway, and the tails are a small part of it.  The scaling iteration
    // multiplies by the float32 scale, whose rounding is the same for the v and w fields of
    // an axis and cancels in their quotient.
    float pf = p0;
    const float alpha_f = (float)alpha;
    const float scale_f = (float)scale;
    const float *addp = lin + s + T + 1;
    float cur[kBlk], nxt[kBlk];
#pragma unroll
    for (int u = 0; u < kBlk; ++u) cur[u] = addp[u];
#pragma unroll
    for (int b = 0; b < kNBlk; ++b) {
        if (b + 1 < kNBlk) {
#pragma unroll
            for (int u = 0; u < kBlk; ++u)
                if ((b + 1) * kBlk + u < kSegArr) nxt[u] = addp[(b + 1) * kBlk + u];
        }
#pragma unroll
        for (int u = 0; u < kBlk; ++u) {
            const int j = b * kBlk + u;
            if (j < kSegArr) {
                const float svf = sb[j];
                float r = fmaf(alpha_f, pf + cur[u], (float)a);
                if (SCALE) r *= scale_f;
                pf = svf;
                a += (double)cur[u] - (double)svf;
                // the empty asm pins each float32 result at its step; otherwise the compiler
                // sinks all conversions to the store and keeps kSegArr doubles live
                asm volatile("" : "+v"(r));
                sb[j] = r;
            }
        }
        if (b + 1 < kNBlk) {
#pragma unroll
            for (int u = 0; u < kBlk; ++u) cur[u] = nxt[u];
        }
    }
}

// n: elements of the segment inside the line; full: the line's right padding holds at
// least kSegArr-1 floats, so the whole segment may be stored (the caller re-zeroes the padding)
template <int SEG>
__device__ __forceinline__ void stage_store(float *__restrict__ lin, int s, int n, bool full,
                                            const float (&sb)[SEG])
{
    constexpr int kSegArr = SEG;
    if (full || n == kSegArr) {
#pragma unroll
        for (int j = 0; j < kSegArr; ++j) lin[s + j] = sb[j];
    } else {
#pragma unroll
        for (int j = 0; j < kSegArr; ++j)
            if (j < n) lin[s + j] = sb[j];
    }
}

__device__ __forceinline__ bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

// global tile <-> LDS.  Tile: nl lines of length L; element (l, j) at g[l*ls + j*es] maps
// to lds[l*lstride + j].  Either ls == 1 (columns: rows of nl contiguous floats) or es == 1
// (rows of L contiguous floats); float4 global accesses when the layout allows.  All
// threads of the workgroup take part.  Pointers advance by increments (no 64-bit index
// products per element).  A store takes the value of each LDS element through `src`: the
// element itself, or the masked quotient of the v element and the w element `wdiff` floats
// after it (the last pass's result, formed on the way out).
struct LdsValue {
    __device__ __forceinline__ float operator()(const float *d) const { return *d; }
};
struct LdsQuotient {
    int wdiff;
    double off;
    float thr_f;
    __device__ __forceinline__ float operator()(const float *d) const { return fast_quotient(d[0], d[wdiff], off, thr_f); }
};

template <bool LOAD, class SRC = LdsValue>
__device__ __forceinline__ void tile_copy(float *__restrict__ g, int64_t ls, int64_t es, int nl, int L,
                                          float *__restrict__ lds, int lstride, SRC src = SRC())
{
    const int tid = threadIdx.x, nthr = blockDim.x;
    if (ls == 1) {
        if ((nl & 3) == 0 && (es & 3) == 0 && aligned16(g)) {
            const int nq = nl >> 2;
            const int rstep = nthr / nq;
            const int q = tid % nq;
            const int r0 = tid / nq;
            if (r0 < rstep) {
                float *gp = g + (int64_t)r0 * es + 4 * q;
                const int64_t gstep = (int64_t)rstep * es;
                float *d = lds + (4 * q) * lstride + r0;
#pragma unroll 4
                for (int r = r0; r < L; r += rstep, gp += gstep, d += rstep) {
                    if (LOAD) {
                        const float4 x = *reinterpret_cast<const float4 *>(gp);
                        d[0] = x.x;
                        d[lstride] = x.y;
                        d[2 * lstride] = x.z;
                        d[3 * lstride] = x.w;
                    } else {
                        *reinterpret_cast<float4 *>(gp) =
                            make_float4(src(d), src(d + lstride), src(d + 2 * lstride), src(d + 3 * lstride));
                    }
                }
            }
        } else {
            const int rstep = nthr / nl;
            const int l = tid % nl;
            const int r0 = tid / nl;
            if (r0 < rstep) {
                float *gp = g + (int64_t)r0 * es + l;
                const int64_t gstep = (int64_t)rstep * es;
                float *d = lds + l * lstride + r0;
#pragma unroll 4
                for (int r = r0; r < L; r += rstep, gp += gstep, d += rstep) {
                    if (LOAD) *d = *gp;
                    else *gp = src(d);
                }
            }
        }
    } else {
        if ((L & 3) == 0 && (ls & 3) == 0 && aligned16(g)) {
            const int Lq = L >> 2;
            for (int l = 0; l < nl; ++l) {
                float4 *gp = reinterpret_cast<float4 *>(g + l * ls) + tid;
                float4 *d = reinterpret_cast<float4 *>(lds + l * lstride) + tid;
#pragma unroll 4
                for (int q = tid; q < Lq; q += nthr, gp += nthr, d += nthr) {
                    // LDS rows start at pad offsets: float4 LDS access only when 16-byte aligned
                    float *dd = reinterpret_cast<float *>(d);
                    if (LOAD) {
                        const float4 x = *gp;
                        dd[0] = x.x;
                        dd[1] = x.y;
                        dd[2] = x.z;
                        dd[3] = x.w;
                    } else {
                        *gp = make_float4(src(dd), src(dd + 1), src(dd + 2), src(dd + 3));
                    }
                }
            }
        } else {
            for (int l = 0; l < nl; ++l) {
                float *gp = g + l * ls + (int64_t)tid * es;
                const int64_t gstep = (int64_t)nthr * es;
                float *d = lds + l * lstride + tid;
#pragma unroll 4
                for (int j = tid; j < L; j += nthr, gp += gstep, d += nthr) {
                    if (LOAD) *d = *gp;
                    else *gp = src(d);
                }
            }
        }
    }
}

// x-pass store of nl rows (y0 ..) of a 2D field into the blocked intermediate: consecutive
// threads take the rows of one 4-column group, so a 2-row tile writes 32-byte pieces.
__device__ __forceinline__ void blocked_store_rows(float *__restrict__ g, int64_t W, int64_t y0, int nl, int L,
                                                   const float *__restrict__ lds, int lstride)
{
    const int nq = L >> 2;
    const int nthr = blockDim.x;
    if (nthr % nl == 0) {
        // thread t keeps row t % nl and steps through the 4-column groups t / nl + k*(nthr/nl):
        // consecutive groups of a row are 32 floats apart in the blocked layout
        const int l = threadIdx.x % nl;
        const int qstep = nthr / nl;
        const float *d = lds + l * lstride;
        float *gp = g + blocked_index(y0 + l, 0, W);
        for (int q = threadIdx.x / nl; q < nq; q += qstep)
            *reinterpret_cast<float4 *>(gp + 32 * (int64_t)q) =
                make_float4(d[4 * q], d[4 * q + 1], d[4 * q + 2], d[4 * q + 3]);
    } else {
        for (int i = threadIdx.x; i < nl * nq; i += nthr) {
            const int l = i % nl, q = i / nl;
            const float *d = lds + l * lstride + 4 * q;
            *reinterpret_cast<float4 *>(g + blocked_index(y0 + l, 4 * q, W)) = make_float4(d[0], d[1], d[2], d[3]);
        }
    }
}

// y-pass load of the columns x0 .. x0+nl-1 (nl a multiple of 4) of both intermediate fields
// from the blocked layout into LDS lines (line l, element j at lds[l*lstride + j]).  Thread t
// takes column group t % (nl/4) and rows t / (nl/4) + k*rstep, k < kBlkRows (the host checks
// that they cover the column); all loads are issued before the first LDS write.
__device__ __forceinline__ void blocked_load_cols(const float *__restrict__ gv, const float *__restrict__ gw,
                                                  int64_t W, int64_t x0, int nl, int L, float *__restrict__ dv,
                                                  float *__restrict__ dw, int lstride)
{
    const int nq = nl >> 2;
    const int rstep = blockDim.x / nq;
    const int q = threadIdx.x % nq, r0 = threadIdx.x / nq;
    if (r0 >= rstep) return;
    float4 rv[kBlkRows], rw[kBlkRows];
#pragma unroll
    for (int k = 0; k < kBlkRows; ++k) {
        const int r = r0 + k * rstep;
        if (r < L) {
            const int64_t o = blocked_index(r, x0 + 4 * q, W);
            rv[k] = *reinterpret_cast<const float4 *>(gv + o);
            rw[k] = *reinterpret_cast<const float4 *>(gw + o);
        }
    }
    float *a = dv + (4 * q) * lstride + r0, *b = dw + (4 * q) * lstride + r0;
#pragma unroll
    for (int k = 0; k < kBlkRows; ++k)
        if (r0 + k * rstep < L) {
            const int e = k * rstep;
            a[e] = rv[k].x;
            a[lstride + e] = rv[k].y;
            a[2 * lstride + e] = rv[k].z;
            a[3 * lstride + e] = rv[k].w;
            b[e] = rw[k].x;
            b[lstride + e] = rw[k].y;
            b[2 * lstride + e] = rw[k].z;
            b[3 * lstride + e] = rw[k].w;
        }
}

// Store of a y pass that is not the last one (the adjoint's, whose v and w lines hold A and
// B): in place, as (A, B) pairs (adj_pair_index), within the 8-row blocks of the columns
// x0 .. x0+nl-1 it read, so no other workgroup's nodes are touched.  Thread (q, r) writes the
// pairs of row r, columns x0+4q .. x0+4q+3: 32 contiguous bytes.
__device__ __forceinline__ void blocked_store_cols(float *__restrict__ gv, float *__restrict__ gw, int64_t W,
                                                   int64_t x0, int nl, int L, const float *__restrict__ dv,
                                                   const float *__restrict__ dw, int lstride)
{
    const int nq = nl >> 2;
    const int rstep = blockDim.x / nq;
    const int q = threadIdx.x % nq, r0 = threadIdx.x / nq;
    if (r0 >= rstep) return;
    const float *a = dv + (4 * q) * lstride + r0, *b = dw + (4 * q) * lstride + r0;
#pragma unroll
    for (int k = 0; k < kBlkRows; ++k) {
        const int r = r0 + k * rstep;
        if (r < L) {
            const int e = k * rstep;
            float4 *p = reinterpret_cast<float4 *>(((r & 4) ? gw : gv) + adj_pair_index(r, x0 + 4 * q, W));
            p[0] = make_float4(a[e], b[e], a[lstride + e], b[lstride + e]);
            p[1] = make_float4(a[2 * lstride + e], b[2 * lstride + e], a[3 * lstride + e], b[3 * lstride + e]);
        }
    }
}

// Load of the adjoint's first pass (rows, element stride 1): from the forward output f, the
// saved normalised weights W and dL/df = g of a row tile it forms a = g / W and
// b = -a (f - offset) (the two fields the adjoint filters; zero where f or g is NaN) straight
// into the v and w lines, instead of a separate pass writing them to HBM.  float32
// arithmetic: the quotient with one Newton step, and f - offset as (f - hi) - lo with the
// offset split into float32 hi + lo (f - hi is exact while f is within a factor 2 of hi).
// float4 global loads when the rows allow.
__device__ __forceinline__ void adjoint_ab(float fo, float go, float wv, float hi, float lo, float &a, float &b)
{
    a = 0.f;
    b = 0.f;
    if (fo == fo && go == go) {
        const float r = __builtin_amdgcn_rcpf(wv);
        float q = go * r;
        q = fmaf(fmaf(-q, wv, go), r, q);
        a = q;
        b = -q * ((fo - hi) - lo);
    }
}

__device__ __forceinline__ void adjoint_load(const float *__restrict__ f, const float *__restrict__ w,
                                             const float *__restrict__ g, double off, int64_t ls, int nl, int L,
                                             float *__restrict__ sa, float *__restrict__ sb, int lstride)
{
    const float hi = (float)off;
    const float lo = (float)(off - (double)hi);
    const int tid = threadIdx.x, nthr = blockDim.x;
    if ((L & 3) == 0 && (ls & 3) == 0 && aligned16(f) && aligned16(w) && aligned16(g)) {
        const int Lq = L >> 2;
        for (int l = 0; l < nl; ++l) {
            const float4 *fl = reinterpret_cast<const float4 *>(f + l * ls);
            const float4 *wl = reinterpret_cast<const float4 *>(w + l * ls);
            const float4 *gl = reinterpret_cast<const float4 *>(g + l * ls);
            float *da = sa + l * lstride, *db = sb + l * lstride;
            for (int q = tid; q < Lq; q += nthr) {
                const float4 fo = fl[q], go = gl[q], wv = wl[q];
                float a, b;
                adjoint_ab(fo.x, go.x, wv.x, hi, lo, a, b);
                da[4 * q] = a;
                db[4 * q] = b;
                adjoint_ab(fo.y, go.y, wv.y, hi, lo, a, b);
                da[4 * q + 1] = a;
                db[4 * q + 1] = b;
                adjoint_ab(fo.z, go.z, wv.z, hi, lo, a, b);
                da[4 * q + 2] = a;
                db[4 * q + 2] = b;
                adjoint_ab(fo.w, go.w, wv.w, hi, lo, a, b);
                da[4 * q + 3] = a;
                db[4 * q + 3] = b;
            }
        }
    } else {
        for (int l = 0; l < nl; ++l) {
            const float *fl = f + l * ls, *wl = w + l * ls, *gl = g + l * ls;
            for (int j = tid; j < L; j += nthr) {
                float a, b;
                adjoint_ab(fl[j], gl[j], wl[j], hi, lo, a, b);
                sa[l * lstride + j] = a;
                sb[l * lstride + j] = b;
            }
        }
    }
}

// Lane i receives lane i+1's value (from_next) or lane i-1's (from_prev) by a DPP wave shift;
// the lane without a source keeps its own value.
__device__ __forceinline__ int dpp_from_next(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }
__device__ __forceinline__ int dpp_from_prev(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ __forceinline__ double dpp_from_next(double v)
{
    return __hiloint2double(dpp_from_next(__double2hiint(v)), dpp_from_next(__double2loint(v)));
}

// Deterministic fused injection of one line's points into its float32 row (element x at
// row[x]).  The line's sub-streams (line_streams: one base row, one corner column),
// concatenated, are taken in chunks of up to 64 entries by the line's first `iw` lanes
// (`injector`).  A run sum over adjacent lanes of equal target cell (in lane order) gives
// each run's total to its first lane.  Equal cells adjacent in one
// sub-stream form one run; a cell can head two runs of a chunk only from different
// sub-streams, so the heads add their totals in one round per sub-stream of the head
// (program order), without conflicts.  The result does not depend on timing.  Called by
// every thread of the workgroup (wave-wide shuffles).
__device__ __forceinline__ void inject_line(const FieldDev &F, int64_t line_in_field, bool injector, int iw, int sid,
                                            int fsel, double off, const uint32_t *__restrict__ vals,
                                            const uint32_t *__restrict__ line_start,
                                            const uint32_t *__restrict__ line_count, const PointRec *__restrict__ rec,
                                            const double *__restrict__ val, float *row, int lo, int len)
{
    // row[x - lo] receives cell x of the line for x in [lo, lo + len) (a chunk of a long line);
    // cells outside are skipped
    const int nsub = 2 * source_rows(F.dim);
    uint32_t rsq[8], rcq[8], pre[9];
    line_streams(F, line_in_field, injector, line_start, line_count, rsq, rcq);
    pre[0] = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) pre[q + 1] = pre[q] + rcq[q];
    int nch = (int)((pre[8] + iw - 1) / iw);
    for (int d = 32; d >= 1; d >>= 1) nch = max(nch, __shfl_xor(nch, d));
    // This lane's entry of chunk ch: its sub-stream and its slot in the sorted point list.
    auto entry = [&](int ch, int &hq, uint32_t &src) {
        const uint32_t e = (uint32_t)ch * iw + sid;
        hq = 0;
        src = 0;
        if (!injector || e >= pre[8]) return false;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if (e >= pre[q] && e < pre[q + 1]) {
                hq = q;
                src = rsq[q] + (e - pre[q]);
            }
        return true;
    };
    struct PointData {
        PointRec r;
        double v;
    };
    auto load_point = [&](bool ok, uint32_t p, PointData &o) {
        o.r = PointRec{{0.0, 0.0, 0.0}, 0, 0};
        o.v = 0.0;
        if (!ok) return;
        o.r = rec[p];
        o.v = val[p];
    };
    // Software pipeline over the chunks: while chunk ch is summed, the point data of chunk ch+1
    // and the point index of chunk ch+2 are in flight.
    auto fetch_index = [&](int ch, bool &ok, int &hq, uint32_t &p) {
        uint32_t src;
        ok = entry(ch, hq, src);
        p = ok ? vals[src] : 0u;
    };
    bool ok_a, ok_b;
    int hq_a, hq_b;
    uint32_t p_a, p_b;
    PointData pd_a;
    fetch_index(0, ok_a, hq_a, p_a);
    load_point(ok_a, p_a, pd_a);
    fetch_index(1, ok_b, hq_b, p_b);
    for (int ch = 0; ch < nch; ++ch) {
        const bool valid = ok_a;
        const int hq = hq_a;
        const PointData pd = pd_a;
        ok_a = ok_b;
        hq_a = hq_b;
        load_point(ok_b, p_b, pd_a);
        fetch_index(ch + 2, ok_b, hq_b, p_b);
        {
            int x = -1;
            double c = 0.0;
            if (valid) {
                PointPos pp;
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    pp.i[m] = 0;
                    pp.w[m] = pd.r.w[m];
                }
                x = pd.r.xi + (hq & 1);
                const double w = corner_weight(F.dim, pp, hq & 1, (hq >> 1) & 1, hq >> 2);
                c = fsel ? w : w * (pd.v - off);
            }
            // runs: maximal groups of adjacent lanes with equal cell; run id = inclusive count
            // of run starts (the concatenated stream is not sorted, so equal cells further
            // apart belong to different runs)
            const int px = dpp_from_prev(x);
            const bool start = sid == 0 || px != x;
            const int lane = threadIdx.x & 63;
            const int rid = __popcll(__ballot(start) & ((2ull << lane) - 1ull));
            // The run sum only where some run of this wave's chunk has more than one entry:
            // otherwise every lane is its own run and keeps c.  Each run's first lane adds the
            // contributions of the run's next lanes one at a time (DPP wave shifts, no LDS
            // round trips), in lane order; runs are short, so this stops after a few steps.
            if (__ballot(valid && !start)) {
                double o = c;
                int orid = rid;
                for (int t = 1; t < iw; ++t) {
                    o = dpp_from_next(o);
                    orid = dpp_from_next(orid);
                    const bool take = sid + t < iw && orid == rid;
                    if (take) c += o;
                    if (!__ballot(take)) break;
                }
            }
            const bool head = valid && start;
            const bool here = head && x >= lo && x < lo + len;
            float *cell = row + (here ? x - lo : 0);
            for (int q = 0; q < nsub; ++q)
                if (here && hq == q) *cell = (float)((double)*cell + c);
        }
    }
}

// Orders the LDS accesses of one phase of a line pass before the next: a workgroup
// barrier when a line spans several waves; a line within one wave needs only program order
// (a wave's LDS instructions execute in order), kept by a compiler fence.
__device__ __forceinline__ void line_sync(int lpl)
{
    if (lpl > 64) __syncthreads();
    else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// One workgroup = 128, 256 or 512 threads: the first half works on the v field, the second
// half on the w field of the same nlines lines (each half: up to lines_cap lines of lpl
// lanes).  LDS holds the v lines, then the w lines, each padded by zeros on both sides.
template <bool INJECT, int SEG>
__global__ __launch_bounds__(kFastBlock, 1) void k_fast_pass(
    const FieldDev *__restrict__ fields, const double *__restrict__ offsets, const PassWork *__restrict__ work,
    const PassField *__restrict__ passes, const PointRec *__restrict__ rec, const uint32_t *__restrict__ vals,
    const uint32_t *__restrict__ line_start, const uint32_t *__restrict__ line_count,
    const double *__restrict__ val, float *__restrict__ gv, float *__restrict__ gw, float *__restrict__ out,
    float *__restrict__ wsave, int final_pass, const float *__restrict__ adj_g)
{
    constexpr int kSegArr = SEG;
    extern __shared__ __align__(16) unsigned char smem[];
    const PassWork wk = work[blockIdx.x];
    const FieldDev &F = fields[wk.field];
    const PassField P = passes[wk.field];
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int half = nthr >> 1;
    const int fsel = tid >= half;            // 0: v worker, 1: w worker
    const int ltid = tid - fsel * half;
    // line range of this work item: the whole line, or for a chunked field the chunk's nodes
    // [lo + so, lo + so + sc) with their halos, [lo, lo + L); the LDS is laid out for P.len
    int L = (int)P.len;
    int lo = 0, so = 0, sc = L;
    if (P.chunk) {
        const int64_t c0 = (int64_t)wk.chunk * P.chunk;
        const int64_t c1 = min(c0 + (int64_t)P.chunk, P.full_len);
        const int64_t a0 = max(c0 - (int64_t)P.halo, (int64_t)0);
        lo = (int)a0;
        L = (int)(min(c1 + (int64_t)P.halo, P.full_len) - a0);
        so = (int)(c0 - a0);
        sc = (int)(c1 - c0);
    }
    const int T = P.T;
    const int lpl = P.lpl;
    const int lines_cap = P.lines;
    // LDS: T+1 zeros, then per line its P.len cells and a gap of P.rpad floats: line l's
    // overflow area (full stores, below) followed by zeros, whose last T+1 are line l+1's
    // left padding.  The w lines follow the v lines, their first left padding the end of the
    // last v line's gap.
    const int pad = T + 1;
    const int rpad = P.rpad;
    const int lstride = (int)P.len + rpad;
    const int nl = wk.nlines;
    const int64_t es = P.elem_stride;
    float *sv = reinterpret_cast<float *>(smem);
    float *sw = sv + (size_t)lines_cap * lstride;
    float *mine = fsel ? sw : sv;
    const int my_line = ltid / lpl;
    const int sid = ltid - my_line * lpl;
    const bool line_ok = my_line < nl;
    const double off = offsets[wk.field];

    // injection: the first (up to 64) lanes of each line
    const int iw = lpl < 64 ? lpl : 64;
    const bool injector = INJECT && line_ok && sid < iw;
    const int64_t line_in_field = INJECT ? wk.base / F.size[0] + my_line : 0;

    if (INJECT) {
        // each line's lanes zero the line's left padding, cells and gap (the left padding is
        // the previous line's gap end, which that line also zeroes and never writes again:
        // zeros over zeros); lines past nl and the reserve after the w lines feed only
        // discarded steps
        if (line_ok) {
            float *z = mine + (size_t)my_line * lstride;
            const int zlen = pad + lstride;
            const int head = (int)((16 - ((uintptr_t)z & 15u)) & 15u) >> 2;   // floats before a 16-byte boundary
            const int nq = (zlen - head) >> 2;
            float4 *z4 = reinterpret_cast<float4 *>(z + head);
            for (int i = sid; i < head; i += lpl) z[i] = 0.f;
            for (int i = sid; i < nq; i += lpl) z4[i] = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int i = head + 4 * nq + sid; i < zlen; i += lpl) z[i] = 0.f;
        }
        line_sync(lpl);
    } else {
        // zero what follows each line's cells (the rest of a chunk's cells and the gap) and the
        // left padding of the v lines
        const int nz = lstride - L;
        if (nz <= 64) {
            // wave w zeroes the gaps of lines w, w + waves, ... (lane = position in the gap)
            const int lane = tid & 63;
            if (lane < nz) {
                for (int l = tid >> 6; l < lines_cap; l += nthr >> 6) {
                    sv[pad + l * lstride + L + lane] = 0.f;
                    sw[pad + l * lstride + L + lane] = 0.f;
                }
            }
        } else {
            for (int i = tid; i < lines_cap * nz; i += nthr) {
                const int l = i / nz, r = i - l * nz;
                const int o = pad + l * lstride + L + r;
                sv[o] = 0.f;
                sw[o] = 0.f;
            }
        }
        for (int i = tid; i < pad; i += nthr) sv[i] = 0.f;
        if (adj_g)
            adjoint_load(out + F.out_off + wk.base + lo, wsave + F.out_off + wk.base + lo,
                         adj_g + F.out_off + wk.base + lo, off, P.line_stride, nl, L, sv + pad, sw + pad, lstride);
        else if (P.blocked == kBlockedIn)
            blocked_load_cols(gv + F.cell_off, gw + F.cell_off, F.size[0], wk.base, nl, L, sv + pad, sw + pad,
                              lstride);
        else {
            tile_copy<true>(gv + F.cell_off + wk.base + lo * es, P.line_stride, es, nl, L, sv + pad, lstride);
            tile_copy<true>(gw + F.cell_off + wk.base + lo * es, P.line_stride, es, nl, L, sw + pad, lstride);
        }
        __syncthreads();
    }

    if (INJECT) {
        // a line's injection writes cells 0..L-1 of that line only
        inject_line(F, line_in_field, injector, iw, sid, fsel, off, vals, line_start, line_count, rec, val,
                    mine + my_line * lstride + pad, lo, L);
        line_sync(lpl);
    }

    // n-fold tailed box filter, in place.  Every lane owns kSegArr consecutive elements;
    // the host picks lpl with lpl*kSegArr >= L and reserves 64 floats past the w lines so
    // that the branch-free slides never read out of bounds.
    const int s0 = sid * kSegArr;
    int myseg = L - s0;
    myseg = myseg < 0 ? 0 : (myseg > kSegArr ? kSegArr : myseg);
    const bool act = line_ok && myseg > 0;
    float *ln = mine + (size_t)(line_ok ? my_line : 0) * lstride + pad;
    const double alpha = P.alpha;
    const int niter = F.num_iter;
    // whole-segment stores (when the host widened the gap by the overflow of the last
    // segment): the segment crossing the line end writes `over` floats into the gap's
    // overflow area, short of the next line's left padding; the line's lanes then re-zero
    // those of them the next iteration reads (within the first T+1), after the stores: by the
    // wave's program order when the line lies in one wave, else after a workgroup barrier
    const int over = (kSegArr - L % kSegArr) % kSegArr;
    const bool full = L > kSegArr && rpad - pad >= over;   // the overflow stays clear of the next left padding
    const int nclr = full ? min(pad, over) : 0;
    // An iteration reads and writes only its own line and its left padding, which no lane
    // writes after the zeroing (reads past the line feed discarded steps),
    // so the lines synchronise on their own: by wave order when a line's lanes lie in one
    // wave (lpl <= 64, a power of two; each field's half of the workgroup is whole waves),
    // else by workgroup barriers.  The tile IO after the loop spans all lines.
    float sb[kSegArr];
    for (int it = 0; it < niter; ++it) {
        const bool last = it + 1 == niter;
        const bool grp = (it + 1) % P.sgroup == 0;
        if (act) {
            if (last || grp) slide1<SEG, true>(ln, s0, T, alpha, last ? P.slast : P.sgscale, sb);
            else slide1<SEG, false>(ln, s0, T, alpha, 1.0, sb);
        }
        line_sync(lpl);
        // an opaque copy of the lane's element count keeps the compiler from hoisting the
        // kSegArr store predicates of the partial case out of the loop (into spilled SGPRs)
        int nst = myseg;
        asm volatile("" : "+v"(nst));
        if (act) stage_store(ln, s0, nst, full, sb);
        // the lane whose segment crosses the line end stores into the padding; in a line of
        // several waves it is not in the wave of the lanes that re-zero the padding, so a
        // barrier orders its stores before theirs (nclr and lpl are the same for the whole
        // workgroup)
        if (lpl > 64 && nclr > 0) __syncthreads();
        if (line_ok)
            for (int e = sid; e < nclr; e += lpl) ln[L + e] = 0.f;
        line_sync(lpl);
    }
    __syncthreads();

    if (final_pass) {
        // masked quotient + offset (interpolation.py:367, :428-430), formed as the v lines are stored
        tile_copy<false>(out + F.out_off + wk.base + (lo + so) * es, P.line_stride, es, nl, sc, sv + pad + so, lstride,
                         LdsQuotient{(int)(sw - sv), off, F.thr_fast_f});
        // training forward: keep the normalised weight field for the adjoint
        if (wsave)
            tile_copy<false>(wsave + F.out_off + wk.base + (lo + so) * es, P.line_stride, es, nl, sc, sw + pad + so,
                             lstride);
    } else if (P.blocked == kBlockedOut) {
        const int64_t y0 = wk.base / F.size[0];
        blocked_store_rows(gv + F.cell_off, F.size[0], y0, nl, L, sv + pad, lstride);
        blocked_store_rows(gw + F.cell_off, F.size[0], y0, nl, L, sw + pad, lstride);
    } else if (P.blocked == kBlockedIn) {
        blocked_store_cols(gv + F.cell_off, gw + F.cell_off, F.size[0], wk.base, nl, L, sv + pad, sw + pad, lstride);
    } else {
        tile_copy<false>(gv + F.cell_off + wk.base + (lo + so) * es, P.line_stride, es, nl, sc, sv + pad + so,
                         lstride);
        tile_copy<false>(gw + F.cell_off + wk.base + (lo + so) * es, P.line_stride, es, nl, sc, sw + pad + so,
                         lstride);
    }
}

// ---------------------------------------------------------------------------
// FP32C prefix pass.  The same n-fold tailed box filter (interpolation.py:484-533), computed
// on FP64 prefix sums of the line held in LDS.  With P_m = sum of the line's elements
// before position m (zero padding outside the line), one iteration is
//   o_k = (1 - alpha) (P_{k+T+1} - P_{k-T}) + alpha (P_{k+T+2} - P_{k-T-1})
// (the window of 2T+1 terms plus the two tail terms).  Each lane owns SEG consecutive
// positions: it chains the FP64 prefix of its outputs, the line's lanes scan their segment
// totals, and the lanes store the next iteration's P = base + chain.  Nothing is rounded to
// float32 between iterations; the last iteration writes float32 outputs into the line's
// tile.  5 FP64 operations per element and iteration, no conversions.
//
// LDS per (field, line): prefix_line_stride(lpl, SEG, T) doubles holding positions
// -T-1 .. lpl*SEG+T+1 (position m at index m+T+1): zero for m <= 0, P_L for m >= L.  The
// float32 tile of the line (tile IO, injection, results) aliases the same storage.  After
// all lines: one double per wave for the segment-total scan of lines wider than a wave.

// Segment lengths the prefix pass is instantiated for; odd, so the 8-byte LDS lane stride is
// conflict-free.
#define FB_PFX_SEGS(X) X(15) X(17) X(19) X(21) X(23)
constexpr int kPfxBlk = 4;   // prefix reads in flight per stream

// x + (x of the lane CTRL selects), as a DPP move of both halves; lanes whose source is
// outside the row, and rows outside ROWMASK, add 0
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_add(double x)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(x), CTRL, ROWMASK, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(x), CTRL, ROWMASK, 0xf, false);
    return x + __hiloint2double(hi, lo);
}

// Exclusive scan of the lanes' segment totals t along each line (lpl lanes of consecutive
// threads) and the line total.  Lines of 16 or more lanes scan with DPP row shifts and row
// broadcasts (a row is 16 lanes), shorter ones with shuffles.  For lines of several waves it
// contains the workgroup barrier that ends the read phase; every thread of the workgroup
// calls it.
__device__ __forceinline__ void line_scan(double t, int sid, int lpl, double *wtot, double &base, double &total)
{
    const int w = lpl < 64 ? lpl : 64;
    const int li = sid & (w - 1);
    double inc = t;
    if (w >= 16) {
        inc = dpp_add<0x111, 0xf>(inc);   // row_shr:1
        inc = dpp_add<0x112, 0xf>(inc);   // row_shr:2
        inc = dpp_add<0x114, 0xf>(inc);   // row_shr:4
        inc = dpp_add<0x118, 0xf>(inc);   // row_shr:8
        if (w >= 32) inc = dpp_add<0x142, 0xa>(inc);   // row_bcast:15 into rows 1, 3
        if (w >= 64) inc = dpp_add<0x143, 0xc>(inc);   // row_bcast:31 into rows 2, 3
    } else {
        for (int d = 1; d < w; d <<= 1) {
            const double o = __shfl_up(inc, d, w);
            if (li >= d) inc += o;
        }
    }
    const double ex = inc - t;
    if (lpl > 64) {
        // lines of several waves: wave totals through LDS
        if ((threadIdx.x & 63) == 63) wtot[threadIdx.x >> 6] = inc;
        __syncthreads();
        const int wv = sid >> 6;
        const int w0 = (int)(threadIdx.x >> 6) - wv;   // first wave of the line
        double pre = 0.0, tot = 0.0;
        for (int k = 0; k < (lpl >> 6); ++k) {
            const double x = wtot[w0 + k];
            pre += k < wv ? x : 0.0;
            tot += x;
        }
        base = pre + ex;
        total = tot;
    } else {
        base = ex;
        total = __shfl(inc, w - 1, w);
    }
}

// One iteration over the lane's segment: pl[j] = P_{s+j-T-1}, pr[j] = P_{s+j+T+1} for
// j = 0..SEG.  OUT: o_j = c1 D_j + c2 E_j into res[]; else the chain
// q[j+1] = q[j] + o_j, and q[nvp] is returned (the partial lane's total).
template <int SEG, bool OUT>
__device__ __forceinline__ double prefix_step(const double *pl, const double *pr, double c1, double c2, int nvp,
                                              double (&q)[SEG + 1], float (&res)[SEG])
{
    constexpr int NB = (SEG + kPfxBlk - 1) / kPfxBlk;
    double la = pl[0], ra = pr[0];
    double lc[kPfxBlk], rc[kPfxBlk], ln[kPfxBlk], rn[kPfxBlk];
#pragma unroll
    for (int u = 0; u < kPfxBlk; ++u) {
        lc[u] = pl[1 + u];
        rc[u] = pr[1 + u];
    }
    double snap = 0.0;
    q[0] = 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        if (b + 1 < NB) {
#pragma unroll
            for (int u = 0; u < kPfxBlk; ++u) {
                const int j = (b + 1) * kPfxBlk + u;
                if (j < SEG) {
                    ln[u] = pl[1 + j];
                    rn[u] = pr[1 + j];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kPfxBlk; ++u) {
            const int j = b * kPfxBlk + u;
            if (j < SEG) {
                const double D = ra - lc[u];   // P_{k+T+1} - P_{k-T}
                const double E = rc[u] - la;   // P_{k+T+2} - P_{k-T-1}
                if (OUT) {
                    res[j] = (float)fma(c1, D, c2 * E);
                } else {
                    q[j + 1] = fma(c1, D, fma(c2, E, q[j]));
                }
                la = lc[u];
                ra = rc[u];
            }
        }
        if (b + 1 < NB) {
#pragma unroll
            for (int u = 0; u < kPfxBlk; ++u) {
                lc[u] = ln[u];
                rc[u] = rn[u];
            }
        }
    }
    if (!OUT) {
        // q[nvp] with the wave-uniform nvp: scalar branches, no per-element selects
        switch (nvp) {
#define FB_SNAP(J) case J: if (J <= SEG) snap = q[J <= SEG ? J : 0]; break;
            FB_SNAP(1) FB_SNAP(2) FB_SNAP(3) FB_SNAP(4) FB_SNAP(5) FB_SNAP(6) FB_SNAP(7) FB_SNAP(8)
            FB_SNAP(9) FB_SNAP(10) FB_SNAP(11) FB_SNAP(12) FB_SNAP(13) FB_SNAP(14) FB_SNAP(15) FB_SNAP(16)
            FB_SNAP(17) FB_SNAP(18) FB_SNAP(19) FB_SNAP(20) FB_SNAP(21) FB_SNAP(22) FB_SNAP(23) FB_SNAP(24)
#undef FB_SNAP
        default: break;
        }
    }
    return snap;
}

template <bool INJECT, int SEG>
__global__ __launch_bounds__(kFastBlock, 1) void k_prefix_pass(
    const FieldDev *__restrict__ fields, const double *__restrict__ offsets, const PassWork *__restrict__ work,
    const PassField *__restrict__ passes, const PointRec *__restrict__ rec, const uint32_t *__restrict__ vals,
    const uint32_t *__restrict__ line_start, const uint32_t *__restrict__ line_count,
    const double *__restrict__ val, float *__restrict__ gv, float *__restrict__ gw, float *__restrict__ out,
    float *__restrict__ wsave, int final_pass, const float *__restrict__ adj_g)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const PassWork wk = work[blockIdx.x];
    const FieldDev &F = fields[wk.field];
    const PassField P = passes[wk.field];
    const int tid = threadIdx.x;
    const int nthr = blockDim.x;
    const int half = nthr >> 1;
    const int fsel = tid >= half;            // 0: v worker, 1: w worker
    const int ltid = tid - fsel * half;
    const int L = (int)P.len;
    const int T = P.T;
    const int lpl = P.lpl;
    const int lines_cap = P.lines;
    const int lstride = prefix_line_stride(lpl, SEG, T);
    const int fstride = 2 * lstride;         // floats per line of the float32 tiles
    const int pad = T + 1;                   // index of position 0
    const int nl = wk.nlines;
    double *pall = reinterpret_cast<double *>(smem);
    double *wtot = pall + (size_t)2 * lines_cap * lstride;
    float *tv = reinterpret_cast<float *>(smem);
    float *tw = tv + (size_t)lines_cap * fstride;
    const int my_line = ltid / lpl;
    const int sid = ltid - my_line * lpl;
    const bool line_ok = my_line < nl;
    double *myP = pall + (size_t)(fsel * lines_cap + (line_ok ? my_line : 0)) * lstride;
    float *myX = reinterpret_cast<float *>(myP);
    const double off = offsets[wk.field];

    if (INJECT) {
        float4 *z = reinterpret_cast<float4 *>(smem);
        const int nz = (2 * lines_cap * fstride) >> 2;
        for (int i = tid; i < nz; i += nthr) z[i] = make_float4(0.f, 0.f, 0.f, 0.f);
        __syncthreads();
        const int iw = lpl < 64 ? lpl : 64;
        const int64_t line_in_field = wk.base / F.size[0] + my_line;
        inject_line(F, line_in_field, line_ok && sid < iw, iw, sid, fsel, off, vals, line_start, line_count, rec, val,
                    myX, 0, L);
    } else {
        if (adj_g)
            adjoint_load(out + F.out_off + wk.base, wsave + F.out_off + wk.base, adj_g + F.out_off + wk.base, off,
                         P.line_stride, nl, L, tv, tw, fstride);
        else {
            tile_copy<true>(gv + F.cell_off + wk.base, P.line_stride, P.elem_stride, nl, L, tv, fstride);
            tile_copy<true>(gw + F.cell_off + wk.base, P.line_stride, P.elem_stride, nl, L, tw, fstride);
        }
    }
    __syncthreads();

    // lane classes: nfull lanes of SEG positions, then (when L mod SEG != 0) one partial lane
    // of nvp positions, then lanes past the line end
    const int s0 = sid * SEG;
    const int nfull = L / SEG;
    const int nvp = L - nfull * SEG;
    const int nv = !line_ok ? 0 : (sid < nfull ? SEG : (sid == nfull ? nvp : 0));
    const bool partial = line_ok && nvp > 0 && sid == nfull;
    double q[SEG + 1];
    float res[SEG];
    double base, total;

    // prefix of the input line
    q[0] = 0.0;
#pragma unroll
    for (int j = 0; j < SEG; ++j) {
        const float xv = j < nv ? myX[s0 + j] : 0.f;
        q[j + 1] = q[j] + (double)xv;
    }
    line_scan(q[SEG], sid, lpl, wtot, base, total);
    line_sync(lpl);   // the float32 reads above end before the prefix stores alias them
    if (line_ok) {
#pragma unroll
        for (int j = 0; j < SEG; ++j) myP[pad + s0 + j] = base + q[j];
        for (int e = sid; e < pad; e += lpl) myP[e] = 0.0;
        for (int e = sid; e < T + 2; e += lpl) myP[pad + lpl * SEG + e] = total;
    }
    line_sync(lpl);

    const double alpha = P.alpha;
    const int niter = F.num_iter;
    const double *pl = myP + s0;                 // P_{s0-T-1}
    const double *pr = myP + s0 + 2 * T + 2;     // P_{s0+T+1}
    for (int it = 0; it < niter; ++it) {
        const bool last = it + 1 == niter;
        const bool grp = (it + 1) % P.sgroup == 0;
        const double sc = last ? P.slast : (grp ? P.sgscale : 1.0);
        const double c1 = nv > 0 ? sc * (1.0 - alpha) : 0.0;
        const double c2 = nv > 0 ? sc * alpha : 0.0;
        if (!last) {
            const double snap = prefix_step<SEG, false>(pl, pr, c1, c2, nvp, q, res);
            line_scan(partial ? snap : q[SEG], sid, lpl, wtot, base, total);
            line_sync(lpl);
            if (line_ok) {
#pragma unroll
                for (int j = 0; j < SEG; ++j) myP[pad + s0 + j] = base + q[j];
                // the partial lane's chain ran on past the line end: its positions after L
                // hold P_L (program order after its own stores above)
                if (partial)
                    for (int e = nvp + 1; e < SEG; ++e) myP[pad + s0 + e] = total;
                for (int e = sid; e < T + 2; e += lpl) myP[pad + lpl * SEG + e] = total;
            }
            line_sync(lpl);
        } else {
            prefix_step<SEG, true>(pl, pr, c1, c2, nvp, q, res);
            line_sync(lpl);
            if (line_ok) {
#pragma unroll
                for (int j = 0; j < SEG; ++j) myX[s0 + j] = res[j];
            }
            __syncthreads();   // the tile IO below spans all lines
        }
    }

    if (final_pass) {
        // masked quotient + offset (interpolation.py:367, :428-430), formed as the v tiles are stored
        tile_copy<false>(out + F.out_off + wk.base, P.line_stride, P.elem_stride, nl, L, tv, fstride,
                         LdsQuotient{(int)(tw - tv), off, F.thr_fast_f});
        if (wsave) tile_copy<false>(wsave + F.out_off + wk.base, P.line_stride, P.elem_stride, nl, L, tw, fstride);
    } else {
        tile_copy<false>(gv + F.cell_off + wk.base, P.line_stride, P.elem_stride, nl, L, tv, fstride);
        tile_copy<false>(gw + F.cell_off + wk.base, P.line_stride, P.elem_stride, nl, L, tw, fstride);
    }
}

// ---------------------------------------------------------------------------
// elementwise helpers

// zero fill by kernel (a fill kernel instead of hipMemsetAsync keeps the batch graph free
// of memset nodes)
__global__ __launch_bounds__(kBlock) void k_zero_u32(uint32_t *__restrict__ p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// the same for a 16-byte aligned buffer of n words: four words per thread (one 16-byte store),
// the last thread's remainder word by word
__global__ __launch_bounds__(kBlock) void k_zero_u32x4(uint32_t *__restrict__ p, int64_t n)
{
    const int64_t i = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 4;
    if (i + 4 <= n) {
        *reinterpret_cast<uint4 *>(p + i) = make_uint4(0u, 0u, 0u, 0u);
    } else {
        for (int64_t k = i; k < n; ++k) p[k] = 0u;
    }
}

__global__ __launch_bounds__(kBlock) void k_f64_to_f32(const double *__restrict__ in, float *__restrict__ out,
                                                       int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

__global__ __launch_bounds__(kBlock) void k_f32_to_f64(const float *__restrict__ in, double *__restrict__ out,
                                                       int64_t n, double scale)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = (double)in[i] * scale;
}

// ---------------------------------------------------------------------------
// naive (interpolation.py:861-938) and radius (:809-855, brute force in place of the
// kd-tree, same inclusion test sqr_dist <= radius^2) in FP64, one thread per node.

__global__ __launch_bounds__(kBlock) void k_naive(const FieldDev *__restrict__ fp, int radius, double r2, double mdw,
                                                  const double *__restrict__ pts, const double *__restrict__ val,
                                                  const double *__restrict__ offsets, double *__restrict__ out,
                                                  int64_t ncells)
{
    __shared__ double sp[3][kBlock];
    __shared__ double sval[kBlock];
    const FieldDev F = *fp;
    const int tid = threadIdx.x;
    const int64_t g = (int64_t)blockIdx.x * kBlock + tid;
    const int dim = F.dim;
    const int64_t W = F.size[0], H = F.size[1];
    const int64_t x = g % W, y = (g / W) % H, z = g / (W * H);
    const double c0 = F.x0[0] + (double)x * F.step[0];
    const double c1 = dim > 1 ? F.x0[1] + (double)y * F.step[1] : 0.0;
    const double c2 = dim > 2 ? F.x0[2] + (double)z * F.step[2] : 0.0;
    // scale = 2*sigma**2 per axis (interpolation.py:872)
    const double s0 = 2 * (F.sigma[0] * F.sigma[0]);
    const double s1 = 2 * (F.sigma[1] * F.sigma[1]);
    const double s2 = 2 * (F.sigma[2] * F.sigma[2]);
    const double off = offsets[0];
    double ws = 0.0, wt = 0.0;
    for (int64_t b = 0; b < F.n; b += kBlock) {
        __syncthreads();
        const int64_t k = b + tid;
        if (k < F.n) {
            for (int m = 0; m < dim; ++m) sp[m][tid] = pts[k * dim + m];
            sval[tid] = val[k] - off;
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)kBlock, F.n - b);
        if (g < ncells) {
            if (!radius) {
                for (int t = 0; t < cnt; ++t) {
                    const double d0 = sp[0][t] - c0;
                    double d = d0 * d0 / s0;
                    if (dim > 1) {
                        const double d1 = sp[1][t] - c1;
                        d = d + d1 * d1 / s1;
                    }
                    if (dim > 2) {
                        const double d2 = sp[2][t] - c2;
                        d = d + d2 * d2 / s2;
                    }
                    const double w = exp(-d);
                    ws += w * sval[t];
                    wt += w;
                }
            } else {
                for (int t = 0; t < cnt; ++t) {
                    const double dx = c0 - sp[0][t];
                    const double dy = c1 - sp[1][t];
                    const double d2 = dx * dx + dy * dy;
                    if (d2 <= r2) {
                        const double w = exp(-d2 / s0);
                        ws += w * sval[t];
                        wt += w;
                    }
                }
            }
        }
    }
    if (g < ncells) {
        if (!radius) out[g] = wt > 0.0 ? ws / wt + off : __builtin_nan("");
        else out[g] = wt >= mdw ? ws / wt + off : __builtin_nan("");
    }
}

// ---------------------------------------------------------------------------
// Adjoint of the FP32C batch (no counterpart in the reference, which has no gradient).
//
// On unmasked nodes f = V/W + offset with V = B S (val - offset), W = B S 1, where S is
// the bilinear injection and B the product of the per-axis n-fold tailed box filters
// (zero-padded banded symmetric matrices, so B is self-adjoint).  For g = dL/df:
//   a = g/W,  b = -g V/W^2 = -a (f - offset),  A = B a,  Bb = B b
//   dL/dval_k  = sum_c w_c(p_k) A[c]
//   dL/dp_k,m  = sum_c (dw_c/dp_m) ((val_k - offset) A[c] + Bb[c])
// The offset drops out (V/W does not depend on it).  The FP32C passes carry the
// normalisation (1/(K+2 alpha))^n per axis in both V and W; applying the same scaled
// operator to a = g/W_scaled gives the same A, so the gradient needs no rescaling.

// a and b at every node of the batch in the cell layout; zero where the forward result
// is NaN (masked) or the incoming gradient is NaN, and in the alignment gaps.
// Transposed injection: one thread per point gathers A and Bb at its 2^dim corners.
// Points dropped by the injection (outside the grid, NaN) get zero gradients.  (Taking the
// points in the forward's sorted order makes the A, Bb reads local but the point reads and
// gradient writes scattered; measured slower.)
__global__ __launch_bounds__(kBlock) void k_adj_gather(const FieldDev *__restrict__ fields,
                                                       const int64_t *__restrict__ pt_prefix, int nfields,
                                                       int64_t npoints, const double *__restrict__ pts,
                                                       const double *__restrict__ val,
                                                       const double *__restrict__ offsets,
                                                       const float *__restrict__ A, const float *__restrict__ Bb,
                                                       const PassField *__restrict__ pass_y,
                                                       double *__restrict__ gval, double *__restrict__ gpts)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npoints) return;
    const int f = find_seg(pt_prefix, nfields, i);
    const FieldDev &F = fields[f];
    const int dim = F.dim;
    // A and B of a 2D field whose y pass ran on the blocked intermediate are (A, B) pairs
    // (adj_pair_index), the pairs of rows y % 8 < 4 in A's field and the others in Bb's
    const bool blocked = pass_y && pass_y[f].blocked == kBlockedIn;
    PointPos pp;
    const bool inside = point_pos(F, pts + i * dim, pp);
    double gv = 0.0, gp[3] = {0.0, 0.0, 0.0};
    if (inside) {
        const double vk = val[i] - offsets[f];
        const int64_t W = F.size[0], H = F.size[1];
        const int corners = 1 << dim;
        for (int cb = 0; cb < corners; ++cb) {
            const int bx = cb & 1, by = (cb >> 1) & 1, bz = (cb >> 2) & 1;
            double a, b;
            if (blocked) {
                const int64_t y = pp.i[1] + by;
                const float2 ab = *reinterpret_cast<const float2 *>((y & 4 ? Bb : A) + F.cell_off +
                                                                    adj_pair_index(y, pp.i[0] + bx, W));
                a = (double)ab.x;
                b = (double)ab.y;
            } else {
                const int64_t cell = F.cell_off + ((pp.i[2] + bz) * H + (pp.i[1] + by)) * W + (pp.i[0] + bx);
                a = (double)A[cell];
                b = (double)Bb[cell];
            }
            const double fx = bx ? pp.w[0] : 1.0 - pp.w[0];
            const double fy = dim >= 2 ? (by ? pp.w[1] : 1.0 - pp.w[1]) : 1.0;
            const double fz = dim >= 3 ? (bz ? pp.w[2] : 1.0 - pp.w[2]) : 1.0;
            const double sx = bx ? 1.0 : -1.0, sy = by ? 1.0 : -1.0, sz = bz ? 1.0 : -1.0;
            gv += corner_weight(dim, pp, bx, by, bz) * a;
            const double t = vk * a + b;
            gp[0] += sx * fy * fz * t;
            gp[1] += fx * sy * fz * t;
            gp[2] += fx * fy * sz * t;
        }
    }
    gval[i] = gv;
    if (gpts)
        for (int m = 0; m < dim; ++m) gpts[i * dim + m] = gp[m] / F.step[m];
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ---------------------------------------------------------------------------
// launchers

int64_t offsets_part_elems(int nfields, int64_t max_n)
{
    return (int64_t)nfields * ((max_n + kOffChunk - 1) / kOffChunk) * 3;
}

int launch_offsets(const FieldDev *d_fields, int nfields, int64_t max_n, const double *d_val, double *d_part,
                   double *d_offsets, hipStream_t s)
{
    if (offsets_single_pass(max_n, d_part)) {
        hipLaunchKernelGGL(k_offsets, dim3(nfields), dim3(kBlock), 0, s, d_fields, d_val, d_offsets);
    } else {
        const int nchunk = (int)((max_n + kOffChunk - 1) / kOffChunk);
        hipLaunchKernelGGL(k_offsets_part, dim3(nchunk, nfields), dim3(kBlock), 0, s, d_fields, d_val, d_part);
        FB_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_offsets_final, dim3((nfields + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                           (const double *)d_part, nfields, nchunk, d_offsets);
    }
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

bool offsets_single_pass(int64_t max_n, const double *d_part) { return !d_part || max_n <= kOffChunk; }

int launch_point_keys_offsets(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                              const double *d_pts, uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_row_count,
                              uint32_t *d_slot, uint32_t *d_rowg, const double *d_val, double *d_offsets,
                              PointRec *d_rec, hipStream_t s)
{
    if (nfields <= 0) return FB_OK;
    const int64_t nkb = npoints > 0 ? nblocks(npoints) : 0;
    hipLaunchKernelGGL(k_point_keys_offsets, dim3((unsigned)(nkb + nfields)), dim3(kBlock), 0, s, d_fields,
                       d_pt_prefix, nfields, npoints, d_pts, d_keys, d_vals, d_row_count, d_slot, d_rowg, d_rec, nkb,
                       d_val, d_offsets);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_point_keys(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                      const double *d_pts, uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_row_count,
                      uint32_t *d_slot, uint32_t *d_rowg, int64_t max_rows, PointRec *d_rec, hipStream_t s)
{
    if (npoints <= 0) return FB_OK;
    // points per thread: enough for each workgroup to see many points per row, while keeping
    // at least ~512 workgroups; small point sets, and fields with more rows than the LDS
    // counters (a workgroup of k_point_keys_hist over them would take one atomic per point, for
    // several points per thread), take one atomic per point and one point per thread
    const int64_t q = npoints / (512 * (int64_t)kBlock);
    const int64_t ppt = q < kKeysMaxPpt ? q : kKeysMaxPpt;
    if (ppt >= 4 && d_row_count && max_rows <= kKeysBinsMax) {
        // 64 KB of counters only where the fields need them (they halve the workgroups per CU;
        // fast_pass_init raised the kernel's dynamic-LDS limit)
        const int nbins = max_rows <= kKeysBins ? kKeysBins : kKeysBinsMax;
        const int64_t per = ppt * kBlock;
        hipLaunchKernelGGL(k_point_keys_hist, dim3((unsigned)((npoints + per - 1) / per)), dim3(kBlock),
                           (size_t)nbins * 4, s, d_fields, d_pt_prefix, nfields, npoints, (int)ppt, nbins, d_pts,
                           d_keys, d_vals, d_row_count, d_slot, d_rowg, d_rec);
    } else {
        hipLaunchKernelGGL(k_point_keys, dim3(nblocks(npoints)), dim3(kBlock), 0, s, d_fields, d_pt_prefix, nfields,
                           npoints, d_pts, d_keys, d_vals, d_row_count, d_slot, d_rowg, d_rec);
    }
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_inject_records(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                          bool local, int64_t nxlines, const uint32_t *d_vals, const uint32_t *d_row_start,
                          const uint32_t *d_row_count, const double *d_pts, const double *d_val,
                          const double *d_offsets, InjectRec *d_irec, hipStream_t s)
{
    if (npoints <= 0 || nxlines <= 0) return FB_OK;
    if (!d_irec) {
        set_error("FP64 injection: the plan has no record array");
        return FB_ERR_INVALID_ARGUMENT;
    }
    hipLaunchKernelGGL(k_inject_records, dim3((unsigned)((npoints + kBlock - 1) / kBlock)), dim3(kBlock), 0, s,
                       d_fields, d_pt_prefix, nfields, npoints, local ? 1 : 0, nxlines, d_vals, d_row_start,
                       d_row_count, d_pts, d_val, d_offsets, d_irec);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_exact_inject(const FieldDev *d_fields, const int64_t *d_xline_prefix, int nfields, int64_t nxlines,
                        const InjectRec *d_irec, const uint32_t *d_row_start, const uint32_t *d_row_count,
                        double *d_vg, double *d_wg, int dim, bool zeroed, hipStream_t s)
{
    if (nxlines <= 0) return FB_OK;
    if (!d_irec) {
        set_error("FP64 injection: the plan has no record array");
        return FB_ERR_INVALID_ARGUMENT;
    }
    const int cap = dim <= 2 ? kInjectStage : 0;   // 3D lines never stage
    hipLaunchKernelGGL(k_exact_inject, dim3((unsigned)((nxlines + kExactInjectThreads - 1) / kExactInjectThreads)),
                       dim3(kExactInjectThreads), (size_t)cap * sizeof(InjectRec), s, d_fields, d_xline_prefix,
                       nfields, nxlines, d_irec, d_row_start, d_row_count, d_vg, d_wg, zeroed ? 1 : 0, cap);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_exact_lines(const FieldDev *d_fields, const int64_t *d_line_prefix, int nfields, int64_t nslots,
                       int axis, int rect_len, double *d_vg, double *d_wg, double *d_vs, double *d_ws, hipStream_t s)
{
    if (nslots <= 0) return FB_OK;
    hipLaunchKernelGGL(k_exact_lines, dim3((unsigned)((nslots + kExactLineThreads - 1) / kExactLineThreads)),
                       dim3(kExactLineThreads), 0, s, d_fields, d_line_prefix, nfields, nslots, axis, rect_len, d_vg,
                       d_wg, d_vs, d_ws);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_exact_rows(const FieldDev *d_fields, const int64_t *d_wg_prefix, const int64_t *d_slot_prefix, int nfields,
                      int64_t nwg, int slots_per_wg, int rect_override, int64_t max_rect, double *d_vg, double *d_wg,
                      double *d_vs, double *d_ws, hipStream_t s)
{
    if (nwg <= 0) return FB_OK;
    const int ring = exact_rows_ring(max_rect);
    if (!ring) {
        set_error("FP64 x lines: kernel of " + std::to_string(max_rect) + " terms too wide for the LDS ring");
        return FB_ERR_UNSUPPORTED;
    }
    // the mirrored ring where its LDS fits (the plain ring's always does: exact_rows_ring)
    const bool mir = exact_rows_lds(ring, slots_per_wg, true) <= 160 * 1024;
    const size_t lds = (size_t)exact_rows_lds(ring, slots_per_wg, mir);
#define FB_EXACT_ROWS(S_, M_)                                                                                       \
    hipLaunchKernelGGL((k_exact_rows<S_, M_>), dim3((unsigned)nwg), dim3(64), lds, s, d_fields, d_wg_prefix,         \
                       d_slot_prefix, nfields, rect_override, ring, d_vg, d_wg, d_vs, d_ws)
    if (slots_per_wg == 16) {
        if (mir) FB_EXACT_ROWS(16, true);
        else FB_EXACT_ROWS(16, false);
    } else if (slots_per_wg == 32) {
        if (mir) FB_EXACT_ROWS(32, true);
        else FB_EXACT_ROWS(32, false);
    } else {
        if (mir) FB_EXACT_ROWS(64, true);
        else FB_EXACT_ROWS(64, false);
    }
#undef FB_EXACT_ROWS
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_exact_finalize(const FieldDev *d_fields, const int64_t *d_cell_prefix, int nfields, int64_t ncells,
                          const double *d_vg, const double *d_wg, const double *d_offsets, float *d_out,
                          hipStream_t s)
{
    if (ncells <= 0) return FB_OK;
    hipLaunchKernelGGL(k_exact_finalize, dim3(nblocks(ncells)), dim3(kBlock), 0, s, d_fields, d_cell_prefix,
                       nfields, ncells, d_vg, d_wg, d_offsets, d_out);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_mask(const FieldDev *d_fields, const int64_t *d_cell_prefix, int nfields, int64_t ncells, double *d_wg,
                int fast, hipStream_t s)
{
    if (ncells <= 0) return FB_OK;
    hipLaunchKernelGGL(k_mask, dim3(nblocks(ncells)), dim3(kBlock), 0, s, d_fields, d_cell_prefix, nfields, ncells,
                       d_wg, fast);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int fast_pass_max_lds() { return 160 * 1024; }

// Raises the dynamic-LDS limit of the fast passes and of k_point_keys_hist once per process
// (before any stream capture, where attribute calls are not allowed).
int fast_pass_init()
{
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        const void *fns[] = {
#define FB_FN(S) (const void *)k_fast_pass<true, S>, (const void *)k_fast_pass<false, S>,
            FB_FAST_SEGS(FB_FN)
#undef FB_FN
#define FB_FN(S) (const void *)k_prefix_pass<true, S>, (const void *)k_prefix_pass<false, S>,
            FB_PFX_SEGS(FB_FN)
#undef FB_FN
        };
        for (const void *fn : fns) {
            err = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, fast_pass_max_lds());
            if (err != hipSuccess) break;
        }
        if (err == hipSuccess)   // the point keys' row counters
            err = hipFuncSetAttribute((const void *)k_point_keys_hist, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      kKeysBinsMax * (int)sizeof(uint32_t));
        // the FP64 x lines' rings
        const void *rows_fns[6] = {(const void *)k_exact_rows<64, false>, (const void *)k_exact_rows<64, true>,
                                   (const void *)k_exact_rows<32, false>, (const void *)k_exact_rows<32, true>,
                                   (const void *)k_exact_rows<16, false>, (const void *)k_exact_rows<16, true>};
        for (int i = 0; i < 6 && err == hipSuccess; ++i)
            err = hipFuncSetAttribute(rows_fns[i], hipFuncAttributeMaxDynamicSharedMemorySize, fast_pass_max_lds());
    });
    if (err != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string("hipFuncSetAttribute failed: ") + hipGetErrorString(err));
        return FB_ERR_HIP;
    }
    return FB_OK;
}

bool fast_pass_seg_supported(int kind, int seg)
{
    if (kind == kPassPrefix) {
        switch (seg) {
#define FB_CASE(S) case S:
            FB_PFX_SEGS(FB_CASE)
#undef FB_CASE
            return true;
        default:
            return false;
        }
    }
    switch (seg) {
#define FB_CASE(S) case S:
        FB_FAST_SEGS(FB_CASE)
#undef FB_CASE
        return true;
    default:
        return false;
    }
}

int launch_fast_pass(int kind, bool inject, int seg, const FieldDev *d_fields, const double *d_offsets,
                     const PassWork *d_work, int nwork, const PassField *d_pass, int threads, size_t lds_bytes,
                     const PointRec *d_rec, const uint32_t *d_vals, const uint32_t *d_line_start,
                     const uint32_t *d_line_count, const double *d_val, float *d_v, float *d_w,
                     float *d_out, float *d_wsave, int final_pass, const float *d_adj_g, hipStream_t s)
{
    if (nwork <= 0) return FB_OK;
    if (threads != 128 && threads != 256 && threads != kFastBlock) {
        set_error("fast pass: unsupported workgroup size");
        return FB_ERR_UNSUPPORTED;
    }
    if (!fast_pass_seg_supported(kind, seg)) {
        set_error("fast pass: unsupported segment length " + std::to_string(seg));
        return FB_ERR_UNSUPPORTED;
    }
#define FB_LAUNCH(K, S)                                                                                            \
    case S:                                                                                                        \
        if (inject)                                                                                                \
            hipLaunchKernelGGL((K<true, S>), dim3(nwork), dim3(threads), lds_bytes, s, d_fields, d_offsets, d_work, \
                               d_pass, d_rec, d_vals, d_line_start, d_line_count, d_val, d_v, d_w, d_out,           \
                               d_wsave, final_pass, d_adj_g);                                                      \
        else                                                                                                       \
            hipLaunchKernelGGL((K<false, S>), dim3(nwork), dim3(threads), lds_bytes, s, d_fields, d_offsets,        \
                               d_work, d_pass, d_rec, d_vals, d_line_start, d_line_count, d_val, d_v, d_w,          \
                               d_out, d_wsave, final_pass, d_adj_g);                                               \
        break;
#define FB_LAUNCH_SLIDE(S) FB_LAUNCH(k_fast_pass, S)
#define FB_LAUNCH_PFX(S) FB_LAUNCH(k_prefix_pass, S)
    if (kind == kPassPrefix) {
        switch (seg) { FB_PFX_SEGS(FB_LAUNCH_PFX) default: break; }
    } else {
        switch (seg) { FB_FAST_SEGS(FB_LAUNCH_SLIDE) default: break; }
    }
#undef FB_LAUNCH_PFX
#undef FB_LAUNCH_SLIDE
#undef FB_LAUNCH
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_adj_gather(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int64_t npoints,
                      const double *d_pts, const double *d_val, const double *d_offsets, const float *d_A,
                      const float *d_Bb, const PassField *d_pass_y, double *d_gval, double *d_gpts, hipStream_t s)
{
    if (npoints <= 0) return FB_OK;
    hipLaunchKernelGGL(k_adj_gather, dim3(nblocks(npoints)), dim3(kBlock), 0, s, d_fields, d_pt_prefix, nfields,
                       npoints, d_pts, d_val, d_offsets, d_A, d_Bb, d_pass_y, d_gval, d_gpts);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_zero(void *p, int64_t bytes, hipStream_t s)
{
    const int64_t n = bytes / 4;
    if (n <= 0) return FB_OK;
    if (((uintptr_t)p & 15) == 0 && n >= 4 * kBlock)
        hipLaunchKernelGGL(k_zero_u32x4, dim3(nblocks((n + 3) / 4)), dim3(kBlock), 0, s, (uint32_t *)p, n);
    else
        hipLaunchKernelGGL(k_zero_u32, dim3(nblocks(n)), dim3(kBlock), 0, s, (uint32_t *)p, n);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_f64_to_f32(const double *in, float *out, int64_t n, hipStream_t s)
{
    if (n <= 0) return FB_OK;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nblocks(n)), dim3(kBlock), 0, s, in, out, n);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_f32_to_f64(const float *in, double *out, int64_t n, double scale, hipStream_t s)
{
    if (n <= 0) return FB_OK;
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nblocks(n)), dim3(kBlock), 0, s, in, out, n, scale);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int launch_naive(const FieldDev *d_field, int radius, double radius_sqr, double max_dist_weight, const double *d_pts,
                 const double *d_val, const double *d_offsets, double *d_out, int64_t ncells, hipStream_t s)
{
    if (ncells <= 0) return FB_OK;
    hipLaunchKernelGGL(k_naive, dim3(nblocks(ncells)), dim3(kBlock), 0, s, d_field, radius, radius_sqr,
                       max_dist_weight, d_pts, d_val, d_offsets, d_out, ncells);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

}  // namespace fb
