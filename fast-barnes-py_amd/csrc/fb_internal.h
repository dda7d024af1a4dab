// fb_internal.h -- shared definitions of the HIP implementation (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace fb {

// Sentinel sort key of a contribution that falls outside the grid
// (reference: the `continue` of interpolation.py:226,249,283).
constexpr uint32_t kInvalidKey = 0xFFFFFFFFu;
// largest field (points) the per-field LDS sort takes (64-bit entries: 64 KiB of LDS)
constexpr int kLocalSortMax = 8192;

// Per-field descriptor as the kernels see it.  Everything is derived on the host in
// FP64 from the reference's own formulas, except the value offset, which the
// offset kernel computes on the device into a separate array.
struct FieldDev {
    int32_t dim;
    int32_t tail;          // 1: optimized_convolution (alpha tails), 0: convolution
    int32_t num_iter;
    int32_t pad0;
    int64_t n;             // points of this field
    int64_t pt_off;        // first point (row of pts / element of val)
    int64_t out_off;       // first output element
    int64_t size[3];       // W, H, D (unused axes = 1)
    int64_t cell_off;      // first cell of the field in the batch's cell space
    int64_t xline_off;     // first x-line of the field in the batch's x-line space
    double x0[3];
    double step[3];
    double sigma[3];
    int32_t T[3];          // half kernel size per axis
    float thr_fast_f;      // smallest float32 >= thr_fast: (double)w < thr_fast iff w < thr_fast_f
    double alpha[3];       // tail value per axis (0 for plain convolution)
    double pscale[3];      // (1/(K+2*alpha))^num_iter per axis, FP32C normalisation
    double thr;            // reference NaN-mask threshold (FP64 path)
    double thr_fast;       // thr * prod(pscale) (FP32C path)
};

// One workgroup of a fast line pass: `nlines` lines of one field whose first
// element offsets (within the field) are base + l*line_stride; for a field whose lines run in
// chunks (PassField::chunk), chunk `chunk` of each.
struct PassWork {
    int32_t field;
    int32_t nlines;
    int64_t base;
    int32_t chunk = 0;
    int32_t pad_ = 0;
};

// Doubles per line in the LDS of the prefix pass: positions -T-1 .. lpl*seg+T+1, rounded up
// to an even count so that every line's float32 view starts 16-byte aligned.
__host__ __device__ inline int prefix_line_stride(int lpl, int seg, int T)
{
    return (lpl * seg + 2 * T + 3 + 1) & ~1;
}

// Kinds of FP32C line pass: the sliding-window pass (float32 tile, FP64 window sums) and the
// prefix pass (FP64 prefix sums in LDS).
enum PassKind { kPassSlide = 0, kPassPrefix = 1 };

// Per-field geometry of one fast line pass along one axis.
struct PassField {
    int64_t len;           // line length L the workgroup's LDS is laid out for (a chunk's, when chunked)
    int64_t line_stride;   // offset between consecutive lines of a workgroup
    int64_t elem_stride;   // offset between consecutive elements of a line
    int32_t T;
    int32_t lpl;           // lanes per line (power of two)
    int32_t rpad;          // gap after each line's cells in LDS (>= T+1; its last T+1 zeros are the next line's left padding)
    int32_t lines;         // line capacity of a workgroup (lines * lpl <= 256)
    double alpha;
    double pscale;         // (1/(K+2*alpha))^num_iter: the pass's total normalisation
    // The normalisation is applied in groups so that float32 storage between iterations
    // cannot overflow: after every sgroup-th iteration multiply by sgscale, after the last
    // one by slast (the product of all applied factors is pscale).
    int32_t sgroup;
    int32_t seg;           // elements per lane (the kernel instantiation the pass runs)
    double sgscale;
    double slast;
    // Layout of the intermediate v, w fields between the x and y passes of a 2D field: row-major
    // (element (y, x) at y*W + x) unless the x pass writes them blocked (blocked_index,
    // kBlockedOut) and the y pass reads them so (kBlockedIn).  Every other pass input and
    // output is row-major.
    int32_t blocked;
    // Lines longer than the pass takes run in chunks: chunk c stores nodes [c*chunk,
    // (c+1)*chunk) of the line (clipped to full_len) and loads `halo` = num_iter*(T+1) more on
    // each side within the line, since an output after num_iter iterations depends only on
    // inputs that near.  chunk = 0: whole lines (len = full_len).
    int32_t chunk;
    int64_t full_len;
    int32_t halo;
    int32_t pad_;
};

enum BlockedRole { kRowMajor = 0, kBlockedOut = 1, kBlockedIn = 2 };

// float4 rows per thread of a blocked y-pass tile load (all in flight at once)
constexpr int kBlkRows = 5;

// Blocked intermediate of a 2D field (W a multiple of 4; the field's cell space holds H
// rounded up to 8 rows): blocks of 8 rows x 4 columns, one 128-byte line each, block
// (y/8, x/4) at line (y/8)*(W/4) + x/4, element (y%8, x%4) at float (y%8)*4 + x%4 of it.  A
// y-pass tile of 4 columns reads whole lines.
__host__ __device__ inline int64_t blocked_index(int64_t y, int64_t x, int64_t W)
{
    return (((y >> 3) * (W >> 2) + (x >> 2)) << 5) + ((y & 7) << 2) + (x & 3);
}

// The adjoint's A and B of such a field, as its y pass leaves them for the gather: (A, B)
// pairs in blocks of 4 rows x 4 columns, one 128-byte line each, within the footprint of the
// blocked v and w fields.  The pair of node (y, x) starts at float adj_pair_index(y, x, W) of
// the v field when y % 8 < 4, else of the w field.  A point's 2x2 corners then touch 1.56
// lines on average, against 2.81 for A and B at the same blocked index of two fields.
__host__ __device__ inline int64_t adj_pair_index(int64_t y, int64_t x, int64_t W)
{
    return (((y >> 3) * (W >> 2) + (x >> 2)) << 5) + ((y & 3) << 3) + ((x & 3) << 1);
}

// A point's grid position as the FP32C injection uses it, formed once per point by the
// point-key kernels (point_pos): its base cell along x and its fractional offsets per axis.
struct PointRec {
    double w[3];
    int32_t xi;
    int32_t pad_;
};

// FP64 x lines through LDS (k_exact_rows): tiles of kRowTile columns, a ring of `ring` tiles
// per (line, field) slot, 32 or 64 slots per workgroup.  exact_rows_ring: the ring's tiles for
// a kernel of rect terms (0 when a 64-slot workgroup's ring would not fit the LDS);
// exact_rows_lds: the bytes of a workgroup of `slots` slots.
constexpr int kRowTile = 32;
// mir: the ring's first kRowTile + 1 positions repeated past its end, so that a chunk's reads
// never wrap (slot stride odd either way)
__host__ __device__ inline int64_t exact_rows_slot_stride(int ring, bool mir = false)
{
    return (int64_t)ring * kRowTile + (mir ? kRowTile + 1 : 1);
}
__host__ __device__ inline int64_t exact_rows_lds(int ring, int slots, bool mir = false)
{
    return (slots * exact_rows_slot_stride(ring, mir) + slots * (kRowTile + 1)) * (int64_t)sizeof(double);
}
inline int exact_rows_ring(int64_t rect)
{
    const int64_t h0 = (rect - 1) / 2, h1 = rect - h0;
    const int64_t need = h0 / kRowTile + (h1 + kRowTile - 1) / kRowTile + 2;
    int ring = 1;
    while (ring < need) ring <<= 1;
    return exact_rows_lds(ring, 64) <= 160 * 1024 ? ring : 0;
}
// slots per workgroup of k_exact_rows for a plan of `slots` x-line slots: 64, or 32 or 16
// while fewer than 256 workgroups (one per CU of an MI355X) would take the lines
inline int exact_rows_slots(int64_t slots)
{
    int s = 64;
    while (s > 16 && (slots + s - 1) / s < 256) s >>= 1;
    return s;
}

// FP64 injection: one sorted entry's record (k_inject_records), in list order.
struct InjectRec {
    int x;              // base column
    uint32_t p;         // point index
    double wx, wy, wz;  // position within the base cell
    double vc;          // value minus the field's offset
};

// Contribution record of the sorted injection list.
struct Contrib {
    uint32_t key;          // global cell index (kInvalidKey when outside)
    uint32_t pidx;         // global point index
};

}  // namespace fb

// ---- host-side error plumbing (shared by the .hip translation units) ------
namespace fb {
void set_error(const std::string &msg);
const char *last_error();
}  // namespace fb

#define FB_HIP_CHECK(expr)                                                          \
    do {                                                                            \
        hipError_t _e = (expr);                                                     \
        if (_e != hipSuccess) {                                                     \
            fb::set_error(std::string("HIP error '") + hipGetErrorString(_e) +      \
                          "' at " __FILE__ ":" + std::to_string(__LINE__) + " in " #expr); \
            return FB_ERR_HIP;                                                      \
        }                                                                           \
    } while (0)
