// fb_api.hip -- host side of libfastbarnes.so: the C ABI of include/fastbarnes.h.
//
// Kernel parameters (T, alpha, threshold) are derived here on the host in FP64
// with the reference's formulas (fastbarnes/interpolation.py:548-569, :782-803,
// :424-425, :669-670); the device kernels only do the data-parallel arithmetic.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <limits>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "fb_launch.h"

// Workspace of fb_barnes_noalloc, reserved with the code object on every device the library
// runs on.  Host code uses the per-device address noalloc_scratch() returns, never this name.
constexpr size_t kNoAllocBytes = FB_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace fb {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }

namespace {

constexpr int kBlock = 256;             // largest number of lanes per field of a fast-pass
                                         // workgroup (it holds as many v workers as w workers)
constexpr size_t kLdsBudget = 80 * 1024; // LDS of a 2*kBlock-thread workgroup: two per CU
constexpr int kLpfs[] = {64, 128, 256};  // lanes per field the passes may run with
constexpr size_t kLdsMax = 160 * 1024;

int fail(int code, const std::string &msg)
{
    set_error(msg);
    return code;
}

// ---- reference parameter formulas ------------------------------------------

// int ** int with 64-bit wrap-around: the plain convolution's kernel_size**num_iter, an
// integer array to an integer power (interpolation.py:669)
int64_t ipow_int(int64_t a, int64_t e)
{
    uint64_t r = 1, b = (uint64_t)a;
    while (e > 0) {
        if (e & 1) r *= b;
        e >>= 1;
        b *= b;
    }
    return (int64_t)r;
}

double sq(double x) { return x * x; }

int32_t to_i32(double r)
{
    if (!(r == r)) return 0;
    if (r >= 2147483647.0) return INT32_MAX;
    if (r <= -2147483648.0) return INT32_MIN;
    return (int32_t)r;
}

// _get_half_kernel_size_opt, interpolation.py:548-552
int32_t half_kernel_opt(double sigma, double step, int64_t n)
{
    const double s = sigma / step;
    return to_i32((std::sqrt(1.0 + 12 * s * s / (double)n) - 1.0) / 2.0);
}

// _get_half_kernel_size, interpolation.py:782-785
int32_t half_kernel_conv(double sigma, double step, int64_t n)
{
    return to_i32(std::sqrt(3.0 / (double)n) * sigma / step + 0.5);
}

// _get_tail_value, interpolation.py:560-569
double tail_value(double sigma, double step, int64_t n)
{
    const int64_t hks = half_kernel_opt(sigma, step, n);
    const int64_t K = 2 * hks + 1;
    const double sigma_rect_sqr = (double)((hks + 1) * hks) / 3.0 * sq(step);
    return 0.5 * (double)K * (sq(sigma) / (double)n - sigma_rect_sqr) /
           (sq((double)(hks + 1) * step) - sq(sigma) / (double)n);
}

// _get_sigma_effective, interpolation.py:796-803
double sigma_effective(double sigma, double step, int64_t n)
{
    const int64_t T = half_kernel_conv(sigma, step, n);
    return std::sqrt((double)n / 3.0 * (double)T * (double)(T + 1)) * step;
}

// conv_scale_factor, interpolation.py:424-425 (tailed) and :669-670 (plain).  The tailed
// (kernel_size+2*tail_value)**num_iter is a float64 array to an integer power: NumPy's power
// ufunc runs its float64 loop on it (the exponent cast to float64), which calls the C
// library's pow, as here.  The plain kernel_size**num_iter stays an integer power.
double threshold(int tail, int dim, const double *sigma, const double *step, const int32_t *K, int64_t n,
                 const double *alpha, double mdw)
{
    double prod = 1.0;
    for (int m = 0; m < dim; ++m) {
        const double p = tail ? std::pow((double)K[m] + 2 * alpha[m], (double)n) : (double)ipow_int(K[m], n);
        prod *= p / std::sqrt(2 * 3.141592653589793) / (sigma[m] / step[m]);
    }
    return prod * mdw;
}

// ---- device memory pool ----------------------------------------------------

struct Pool {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void *> free_blocks;
};
Pool &pool()
{
    static Pool *p = new Pool();   // never destroyed: blocks live until process exit
    return *p;
}

size_t round_bytes(size_t b) { return (b + 255) & ~(size_t)255; }

int pool_alloc(int device, size_t bytes, void **out)
{
    bytes = round_bytes(bytes ? bytes : 256);
    {
        std::lock_guard<std::mutex> g(pool().mu);
        auto it = pool().free_blocks.find({device, bytes});
        if (it != pool().free_blocks.end()) {
            *out = it->second;
            pool().free_blocks.erase(it);
            return FB_OK;
        }
    }
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        // release cached blocks of this device and retry once
        std::vector<void *> rel;
        {
            std::lock_guard<std::mutex> g(pool().mu);
            for (auto it = pool().free_blocks.begin(); it != pool().free_blocks.end();) {
                if (it->first.first == device) {
                    rel.push_back(it->second);
                    it = pool().free_blocks.erase(it);
                } else {
                    ++it;
                }
            }
        }
        for (void *p : rel) (void)hipFree(p);
        e = hipMalloc(out, bytes);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return fail(FB_ERR_OUT_OF_MEMORY, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
        }
    }
    return FB_OK;
}

void pool_free(int device, void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> g(pool().mu);
    pool().free_blocks.insert({{device, round_bytes(bytes ? bytes : 256)}, p});
}

// per-device library stream (default priority)
hipStream_t lib_stream(int device)
{
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    std::lock_guard<std::mutex> g(mu);
    auto it = streams.find(device);
    if (it != streams.end()) return it->second;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s = nullptr;
    }
    streams[device] = s;
    return s;
}

// Grouped FP32C normalisation of a pass with per-iteration gain g = K + 2*alpha: groups of
// m iterations with g^m <= 2^60, so values stored in float32 between iterations stay far
// below its range (2^128) for any num_iter; a single group (scaling only after the last
// iteration) whenever g^num_iter <= 2^60.
void scale_groups(double g, int n, double total, PassField &p)
{
    int m = n;
    if (g > 1.0) m = std::max(1, std::min(n, (int)std::floor(60.0 * std::log(2.0) / std::log(g))));
    const int ngroups = (n + m - 1) / m;
    const int last = n - m * (ngroups - 1);
    p.pscale = total;
    p.sgroup = m;
    p.sgscale = ngroups > 1 ? std::pow(1.0 / g, (double)m) : 1.0;
    p.slast = ngroups > 1 ? std::pow(1.0 / g, (double)last) : total;
}

// LDS bank multiplicity of a line layout: the most lanes of one half-wave (32 lanes, served
// together by the 32 four-byte banks; 1 is conflict-free) that meet on one bank.
//  * slide_bank_conflicts: the slides and their stores, lane sid of line l of a field's half
//    of the workgroup at l*lstride + sid*seg + j.  Within a line the odd segment length keeps
//    the lanes apart; lines shorter than 32 lanes share a half-wave, and their lanes fall on
//    disjoint banks when lstride = seg*lpl*odd (mod 32).
//  * column_bank_conflicts: the column tiles of a y or z pass (tile_copy, blocked_load_cols),
//    element (l, r) at l*lstride + r: thread t takes lines 4q .. 4q+3, q = t % (nl/4), at rows
//    t / (nl/4) + k*nthr/(nl/4) when nl is a multiple of 4 (float4 columns), else line t % nl.
int slide_bank_conflicts(int lstride, int lpl, int seg, int nl)
{
    int worst = 0;
    for (int t0 = 0; t0 < nl * lpl; t0 += 32) {
        int bank[32] = {0};
        for (int t = t0; t < std::min(t0 + 32, nl * lpl); ++t)
            worst = std::max(worst, ++bank[((int64_t)(t / lpl) * lstride + (int64_t)(t % lpl) * seg) % 32]);
    }
    return worst;
}

int column_bank_conflicts(int lstride, int nl, int nthr)
{
    const bool quad = nl % 4 == 0;
    const int groups = quad ? nl / 4 : nl;
    int worst = 0;
    for (int t0 = 0; t0 < nthr; t0 += 32)
        for (int c = 0; c < (quad ? 4 : 1); ++c) {
            int bank[32] = {0};
            for (int t = t0; t < t0 + 32; ++t) {
                const int64_t a = (int64_t)(quad ? 4 * (t % groups) + c : t % groups) * lstride + t / groups;
                worst = std::max(worst, ++bank[a % 32]);
            }
        }
    return worst;
}

// lanes per line / lines per workgroup of a fast line pass with `seg` elements per lane and
// workgroups of 2*lpf threads (lpf lanes per field); the LDS budget scales with the
// workgroup so that every size keeps 16 waves per CU
// What pass_geometry's choice of the gap width looks at besides the lines per workgroup: nothing
// (the cost model), the slides only (x passes: row tiles), the slides then the column tiles
enum BankSearch { kBanksNone = 0, kBanksRows = 1, kBanksColumns = 2 };

int pass_geometry(int64_t L, int T, int seg, int lpf, int &lpl, int &lines, int &rpad, size_t &lds, int banks)
{
    const int64_t need = (L + seg - 1) / seg;
    lpl = 1;
    while (lpl < need) lpl <<= 1;
    if (lpl > lpf)
        return fail(FB_ERR_UNSUPPORTED, "line length " + std::to_string(L) + " exceeds the FP32C limit of " +
                                            std::to_string((int64_t)lpf * seg) + " for this workgroup size");
    // branch-free slides read up to seg-1 floats past the last w line: reserve 64 floats
    const size_t reserve = 64 * sizeof(float);
    const int lane_lines = lpf / lpl;
    const size_t budget = kLdsBudget / (kBlock / lpf);
    // LDS: T+1 zeros, then lines of L cells each followed by a gap of rpad floats, v lines
    // then w lines.  The gap's last T+1 zeros are the next line's left padding (the
    // reference's zero padding on both sides of a line, shared by neighbouring lines).
    const size_t lead = (size_t)(T + 1) * sizeof(float);
    auto fit = [&](int rp) {
        const size_t per_line = (size_t)(L + rp) * 2 * sizeof(float);
        int64_t n = (int64_t)((budget - reserve - lead) / per_line);
        n = std::min<int64_t>(n, lane_lines);
        if (n >= 16) n &= ~(int64_t)3;   // float4 column tiles, at a loss below 1/4
        return (int)std::max<int64_t>(n, 1);
    };
    // Gap: the reference's T+1 zeros.  A line of several segments may instead widen it by the
    // overflow of its last segment (seg - L mod seg), so that the lane whose segment crosses
    // the line end stores all of it (into the widening, clear of the next line's padding) and
    // the pass re-zeroes what the line reads, instead of predicating every store.  The wider
    // gap is taken when it does not cost lines per workgroup.
    const int over = (int)((seg - L % seg) % seg);
    rpad = T + 1;
    lines = fit(rpad);
    if (need > 1 && over > 0) {
        const int wide = fit(T + 1 + over);
        if (wide >= lines) {
            rpad = T + 1 + over;
            lines = wide;
        }
    }
    // A wider gap changes nothing but the line stride: up to 31 more floats, where they cost no
    // lines, for the fewest LDS bank conflicts of the slides (about 14 LDS accesses per element
    // and pass in four iterations), then, in passes with column tiles, of those (2)
    if (banks != kBanksNone) {
        const int base = rpad;
        // nor workgroups per CU (the LDS's, up to 32 waves)
        auto per_cu = [&](int rp) {
            const size_t b = round_bytes(lead + lines * (size_t)(L + rp) * 2 * sizeof(float) + reserve);
            return std::min<size_t>(kLdsMax / b, (size_t)(32 / std::max(1, 2 * lpf / 64)));
        };
        const size_t wgs = per_cu(base);
        int sbest = 1 << 30, cbest = 1 << 30;
        int sc[32];
        for (int extra = 0; extra < 32; ++extra) {
            sc[extra] = extra == 0 || (fit(base + extra) >= lines && per_cu(base + extra) >= wgs)
                            ? slide_bank_conflicts((int)L + base + extra, lpl, seg, lines) : 1 << 30;
            sbest = std::min(sbest, sc[extra]);
        }
        for (int extra = 0; extra < 32 && cbest > 1; ++extra) {
            if (sc[extra] != sbest) continue;
            const int c = banks == kBanksColumns ? column_bank_conflicts((int)L + base + extra, lines, 2 * lpf) : 1;
            if (c < cbest) {
                cbest = c;
                rpad = base + extra;
            }
        }
    }
    const size_t per_line = (size_t)(L + rpad) * 2 * sizeof(float);
    lds = round_bytes(lead + lines * per_line + reserve);
    if (lds > kLdsMax)
        return fail(FB_ERR_UNSUPPORTED, "line of length " + std::to_string(L) + " with half kernel " +
                                            std::to_string(T) + " does not fit the local data share");
    return FB_OK;
}

// Segment lengths of the fast-pass instantiations (FB_FAST_SEGS of fb_kernels.hip).
constexpr int kSegs[] = {19, 21, 23, 25, 29, 33, 37, 39, 41, 45, 49};

// Cost model of a fast pass per line: every lane of a workgroup issues the same
// instructions, about 7 per element of its segment (slide), 2 per window term of the
// re-anchoring (2T+1 terms) and a fixed ~30 per iteration, plus 2 per element when the
// last segment's stores are predicated; a workgroup holds `lines` lines.
double pass_cost(int64_t L, int T, int seg, int lpf = kBlock, bool *ok = nullptr)
{
    int lpl, lines, rpad;
    size_t lds;
    const std::string keep = last_error();
    const int rc = pass_geometry(L, T, seg, lpf, lpl, lines, rpad, lds, kBanksNone);   // the bank search keeps the lines
    set_error(keep);
    if (ok) *ok = rc == FB_OK;
    if (rc) return 1e300;
    const int over = (int)((seg - L % seg) % seg);
    const bool full = L > seg && rpad - (T + 1) >= over;
    return (7.0 * seg + 2.0 * (2.0 * T + 1.0) + 30.0 + (full ? 0.0 : 2.0 * seg)) / lines * (kBlock / lpf);
}

int model_seg(int64_t L, int T, int lpf = kBlock)
{
    int best = 41;
    double bc = 1e300;
    for (int seg : kSegs) {
        const double c = pass_cost(L, T, seg, lpf);
        if (c < bc) {
            bc = c;
            best = seg;
        }
    }
    return best;
}

// Segment lengths of the prefix-pass instantiations (FB_PFX_SEGS of fb_kernels.hip).
constexpr int kPfxSegs[] = {15, 17, 19, 21, 23};

// Geometry of a prefix pass: lpl lanes per line, `lines` lines per field in a workgroup of
// 2*lpf threads (all lanes busy unless the LDS limit cuts lines), LDS bytes, and the
// workgroups per CU the LDS allows.  A line takes 8 bytes per position (FP64 prefix sums),
// so the LDS, not the thread count, bounds how many lines a CU holds.
int prefix_geometry(int64_t L, int T, int seg, int lpf, int &lpl, int &lines, size_t &lds, int &wg_per_cu)
{
    const int64_t need = (L + seg - 1) / seg;
    lpl = 1;
    while (lpl < need) lpl <<= 1;
    if (lpl > lpf)
        return fail(FB_ERR_UNSUPPORTED, "line length " + std::to_string(L) + " exceeds the prefix-pass limit of " +
                                            std::to_string((int64_t)lpf * seg) + " for this workgroup size");
    const size_t per_line = (size_t)prefix_line_stride(lpl, seg, T) * 2 * sizeof(double);   // v and w
    const size_t reserve = 16 * sizeof(double);   // wave totals of the segment scan
    int n = lpf / lpl;
    while (n > 1 && n * per_line + reserve > kLdsMax) --n;
    if (n >= 16) n &= ~3;   // float4 column tiles
    lines = n;
    lds = round_bytes(n * per_line + reserve);
    if (lds > kLdsMax)
        return fail(FB_ERR_UNSUPPORTED, "line of length " + std::to_string(L) + " with half kernel " +
                                            std::to_string(T) + " does not fit the local data share");
    const int waves = 2 * lpf / 64;
    wg_per_cu = (int)std::min<size_t>(kLdsMax / lds, (size_t)(32 / std::max(waves, 1)));
    return FB_OK;
}

// Cost model of a prefix pass per line, in the units of pass_cost: about 5.5 instructions
// per element (two differences, two FMAs of the chain, the store of base + chain) plus ~60
// per iteration (segment-total scan, barriers, addressing), divided by the lines a
// workgroup holds; below 12 waves per CU the FP64 latency is no longer hidden.
double prefix_cost(int64_t L, int T, int seg, int lpf = kBlock, bool *ok = nullptr)
{
    int lpl, lines, wgs;
    size_t lds;
    const std::string keep = last_error();
    const int rc = prefix_geometry(L, T, seg, lpf, lpl, lines, lds, wgs);
    set_error(keep);
    if (ok) *ok = rc == FB_OK;
    if (rc) return 1e300;
    const double waves = (double)wgs * (2 * lpf / 64);
    const double occ = std::min(1.0, waves / 12.0);
    return (5.5 * seg + 60.0) / lines * (kBlock / lpf) / occ;
}

// segment length, lanes per field and kind of a pass (autotuned or from the cost model)
struct PassChoice {
    int seg;
    int lpf;
    int kind;
};

double choice_cost(int64_t L, int T, const PassChoice &c, bool *ok = nullptr)
{
    return c.kind == kPassPrefix ? prefix_cost(L, T, c.seg, c.lpf, ok) : pass_cost(L, T, c.seg, c.lpf, ok);
}

// Untuned default: the sliding-window pass with 256-thread workgroups (faster than 512 on
// every measured workload) when a segment length fits the line in 128 lanes, else
// 512-thread workgroups.  The prefix pass is an autotuner candidate only: it halves the
// elements a CU holds in LDS (FP64) and was slower on the 2400x1200 and 512^2x128 fields.
PassChoice model_choice(int64_t L, int T)
{
    for (int lpf : {kBlock / 2, kBlock}) {
        const int seg = model_seg(L, T, lpf);
        bool ok = false;
        pass_cost(L, T, seg, lpf, &ok);
        if (ok) return PassChoice{seg, lpf, kPassSlide};
    }
    return PassChoice{41, kBlock, kPassSlide};   // pass_geometry reports the limit
}

// order work items so that blockIdx b runs tile t with consecutive tiles on one XCD
// (workgroups are dispatched round-robin over the 8 XCDs, each with its own L2)
template <class V>
V xcd_order(const V &items)
{
    const size_t n = items.size();
    const size_t nx = 8;
    const size_t per = (n + nx - 1) / nx;
    V out;
    out.reserve(n);
    for (size_t b = 0; b < per * nx; ++b) {
        const size_t t = (b % nx) * per + b / nx;
        if (t < n) out.push_back(items[t]);
    }
    return out;
}

struct DevBuf {
    void *p;
    size_t bytes;
};

}  // namespace
}  // namespace fb

using namespace fb;

// ---- plan --------------------------------------------------------------------

namespace fb {
namespace {
// one launch of a fast pass: work items [first, first + count) of the axis's list
struct PassGroup {
    int seg;
    int lpf;      // lanes per field: the workgroup has 2*lpf threads
    int kind;     // kPassSlide or kPassPrefix
    int first;
    int count;
    size_t lds;
};

}  // namespace
}  // namespace fb

struct fb_plan {
    int device = 0;
    int precision = FB_PREC_FP64;
    int nfields = 0, dim = 0, corners = 0;
    int64_t npoints = 0, ncells = 0, nxlines = 0, nentries = 0, total_out = 0, total_cells = 0;
    std::vector<FieldDev> fh;
    std::vector<DevBuf> bufs;
    size_t ws_bytes = 0;

    FieldDev *d_fields = nullptr;
    int64_t *d_pt_prefix = nullptr, *d_xline_prefix = nullptr, *d_cell_prefix = nullptr;
    int64_t *d_lslot_prefix[3] = {nullptr, nullptr, nullptr};
    int64_t lslots[3] = {0, 0, 0};
    // FP64 x lines through LDS (k_exact_rows): first workgroup of each field, workgroups in
    // all, slots per workgroup, and the widest x kernel (2T+1) of the fields
    int64_t *d_rowwg_prefix = nullptr;
    int64_t rowwgs = 0, row_rect = 0;
    int row_slots = 64;
    double *d_offsets = nullptr, *d_off_part = nullptr;
    int64_t max_n = 0;
    int64_t max_rows = 0;   // most x-lines of any field
    uint32_t *d_keys = nullptr, *d_vals = nullptr, *d_keys2 = nullptr, *d_vals2 = nullptr;
    uint32_t *d_line_count = nullptr, *d_line_start = nullptr, *d_scan_tmp = nullptr;
    uint32_t *d_skeys = nullptr, *d_big_list = nullptr, *d_big_count = nullptr;   // row buckets
    PointRec *d_rec = nullptr;                  // FP32C: each point's position for the injection
    InjectRec *d_irec = nullptr;                // FP64: each sorted entry's injection record
    double *d_vg = nullptr, *d_wg = nullptr, *d_vs = nullptr, *d_ws = nullptr;
    float *d_v = nullptr, *d_w = nullptr;
    PassWork *d_work[3] = {nullptr, nullptr, nullptr};
    PassField *d_pass[3] = {nullptr, nullptr, nullptr};
    std::vector<PassGroup> groups[3];          // launches of each axis pass, one per segment length
    std::map<std::pair<int, int64_t>, PassChoice> seg_choice;   // (axis, line length) -> autotuned choice
    const uint32_t *sorted_vals = nullptr;      // sorted point indices of the last run
    int allow_blocked = 1;                      // fb_plan_set_blocked
    bool inplace_passes = false;                // every pass filters its planes in place (no chunks)

    int timing = 0;
    std::vector<hipEvent_t> events;
    std::vector<std::string> stage_names;
    std::vector<float> stage_ms;

    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    const void *g_pts = nullptr, *g_val = nullptr, *g_out = nullptr;
    hipStream_t g_stream = nullptr;

    // The last use of the plan's buffers on each stream it ran on: the destructor waits for
    // these before the buffers go back to the pool, where another plan may take them at once.
    std::vector<std::pair<hipStream_t, hipEvent_t>> uses;

    // Records the plan's work queued so far on s.  Work queued while the caller captures s into
    // a graph of its own is not recorded: that graph holds the plan's buffers, so the plan must
    // outlive it.
    int mark_use(hipStream_t s)
    {
        hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
        if (hipStreamIsCapturing(s, &st) != hipSuccess) {
            (void)hipGetLastError();
            st = hipStreamCaptureStatusNone;
        }
        if (st != hipStreamCaptureStatusNone) return FB_OK;
        for (auto &u : uses)
            if (u.first == s) {
                FB_HIP_CHECK(hipEventRecord(u.second, s));
                return FB_OK;
            }
        hipEvent_t e = nullptr;
        FB_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        if (hipEventRecord(e, s) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipEventDestroy(e);
            FB_HIP_CHECK(hipStreamSynchronize(s));   // no event: wait here instead
            return FB_OK;
        }
        uses.push_back({s, e});
        return FB_OK;
    }

    // With `arena` set, the buffers are carved from [arena, arena + arena_cap) in order
    // (256-byte aligned) instead of the pool, and nothing is allocated or returned.
    char *arena = nullptr;
    size_t arena_cap = 0;

    template <class T>
    int alloc(T **p, size_t count)
    {
        void *q = nullptr;
        const size_t bytes = count * sizeof(T);
        if (arena) {
            const size_t need = round_bytes(bytes ? bytes : 256);
            if (ws_bytes + need > arena_cap)
                return fail(FB_ERR_UNSUPPORTED, "the workspace of this call exceeds the fixed scratch array of " +
                                                    std::to_string(arena_cap) + " bytes of fb_barnes_noalloc; " +
                                                    "use fb_barnes");
            *p = (T *)(arena + ws_bytes);
            ws_bytes += need;
            return FB_OK;
        }
        int rc = pool_alloc(device, bytes, &q);
        if (rc) return rc;
        bufs.push_back({q, bytes});
        ws_bytes += round_bytes(bytes ? bytes : 256);
        *p = (T *)q;
        return FB_OK;
    }
    // several host arrays into one device block with one copy, each 256-byte aligned (a plan's
    // descriptors and prefixes: one synchronous copy instead of one per array)
    struct Upload {
        void **dst;
        const void *src;
        size_t bytes;
    };
    int upload_all(const std::vector<Upload> &ups)
    {
        size_t total = 0;
        for (const Upload &u : ups) total += round_bytes(u.bytes ? u.bytes : 1);
        char *d = nullptr;
        int rc = alloc(&d, total);
        if (rc) return rc;
        std::vector<char> h(total);
        size_t off = 0;
        for (const Upload &u : ups) {
            if (u.bytes) std::memcpy(h.data() + off, u.src, u.bytes);
            *u.dst = d + off;
            off += round_bytes(u.bytes ? u.bytes : 1);
        }
        FB_HIP_CHECK(hipMemcpy(d, h.data(), total, hipMemcpyHostToDevice));
        return FB_OK;
    }
    template <class T>
    int upload(T **p, const std::vector<T> &h)
    {
        int rc = alloc(p, h.size() ? h.size() : 1);
        if (rc) return rc;
        if (!h.empty()) FB_HIP_CHECK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
        return FB_OK;
    }
    ~fb_plan()
    {
        // kernels queued on a caller's stream may still read or write the buffers
        for (auto &u : uses) {
            (void)hipEventSynchronize(u.second);
            (void)hipEventDestroy(u.second);
        }
        for (hipEvent_t e : events) {
            (void)hipEventSynchronize(e);
            (void)hipEventDestroy(e);
        }
        (void)hipGetLastError();
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &b : bufs) pool_free(device, b.p, b.bytes);
    }
};

namespace fb {
namespace {
// Chunks of an FP32C line pass whose lines are longer than the pass takes (no segment length
// and workgroup size fits them): lines of length L run as chunks that each store C nodes and
// load H = num_iter*(T+1) more on either side within the line; C is balanced over the fewest
// chunks of at most Lc = C + 2H nodes that the pass takes.  Lc = 0: the whole line fits.
struct ChunkGeom {
    int64_t Lc = 0, C = 0, H = 0;
};

int chunk_geom(int64_t L, int T, int num_iter, ChunkGeom &g)
{
    g = ChunkGeom{};
    bool ok = false;
    choice_cost(L, T, model_choice(L, T), &ok);
    if (ok) return FB_OK;
    const int64_t H = (int64_t)num_iter * (T + 1);
    for (int64_t Lc : {8192, 6144, 4096, 3072, 2048}) {
        if (Lc >= L || Lc <= 2 * H + 256) continue;
        choice_cost(Lc, T, model_choice(Lc, T), &ok);
        if (!ok) continue;
        const int64_t nc = (L + (Lc - 2 * H) - 1) / (Lc - 2 * H);
        g.C = (L + nc - 1) / nc;
        g.H = H;
        g.Lc = g.C + 2 * H;
        return FB_OK;
    }
    return fail(FB_ERR_UNSUPPORTED, "lines of length " + std::to_string(L) + " with half kernel " + std::to_string(T) +
                                        " and " + std::to_string(num_iter) +
                                        " iterations fit no FP32C line pass, whole or in chunks; use precision fp64");
}

// Length the fast pass along `axis` lays its LDS out for: the field's line length, or its
// chunks' when the lines run in chunks (0 and an error when they cannot).
int64_t pass_len(const FieldDev &F, int axis)
{
    ChunkGeom g;
    const std::string keep = last_error();
    const int rc = chunk_geom(F.size[axis], F.T[axis], F.num_iter, g);
    set_error(keep);
    if (rc) return 0;
    return g.Lc ? g.Lc : F.size[axis];
}

// the pass choice of an FP32C plan for lines of length L along `axis` (autotuned, forced or the
// cost model's)
PassChoice pass_choice(const fb_plan *P, int axis, int64_t L, int T)
{
    auto it = P->seg_choice.find({axis, L});
    return it != P->seg_choice.end() ? it->second : model_choice(L, T);
}

// Whether field f of a 2D FP32C plan keeps its intermediate between the x and y passes blocked
// (blocked_index): W a multiple of 4 (plan_create then gives the field room for H rounded up
// to 8 rows), a float4-aligned cell offset, the sliding pass on both axes, y tiles a multiple
// of 4 columns that kBlkRows float4 rows per thread cover.
bool field_blocked(const fb_plan *P, int f)
{
    const FieldDev &F = P->fh[f];
    if (!P->allow_blocked || F.dim != 2 || F.size[0] % 4 || F.cell_off % 4) return false;
    if (pass_len(F, 0) != F.size[0] || pass_len(F, 1) != F.size[1]) return false;   // chunked lines
    const PassChoice cx = pass_choice(P, 0, F.size[0], F.T[0]);
    const PassChoice cy = pass_choice(P, 1, F.size[1], F.T[1]);
    if (cx.kind != kPassSlide || cy.kind != kPassSlide) return false;
    int lpl, lines, rpad;
    size_t lds;
    const std::string keep = last_error();
    const int rc = pass_geometry(F.size[1], F.T[1], cy.seg, cy.lpf, lpl, lines, rpad, lds, kBanksNone);
    set_error(keep);
    if (rc || lines % 4) return false;
    const int64_t rstep = 2 * cy.lpf / (lines / 4);
    return rstep >= 1 && rstep * kBlkRows >= F.size[1];
}
}  // namespace
}  // namespace fb

namespace fb {
namespace {

bool finite_pos(double x) { return std::isfinite(x) && x > 0.0; }

// The smallest float32 t with (double)t >= thr, so that for every float32 w the FP64
// comparison (double)w < thr equals the float32 comparison w < t.
float float_threshold(double thr)
{
    if (std::isnan(thr)) return thr;
    float t = (float)thr;
    if ((double)t < thr) t = std::nextafter(t, std::numeric_limits<float>::infinity());
    return t;
}

// validates one field and fills its device descriptor (kernel parameters included)
int make_field(const fb_field &f, FieldDev &d, int fast)
{
    if (f.dim < 1 || f.dim > 3) return fail(FB_ERR_INVALID_ARGUMENT,
                                           "Barnes interpolation supports only sample points in dimensions 1, 2 or 3");
    if (f.method != FB_METHOD_OPTIMIZED_CONVOLUTION && f.method != FB_METHOD_CONVOLUTION)
        return fail(FB_ERR_INVALID_ARGUMENT, "batched plans support the convolution methods only");
    if (f.num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
    if (f.n < 1) return fail(FB_ERR_INVALID_ARGUMENT, "a field needs at least one sample point");
    std::memset(&d, 0, sizeof(d));
    d.dim = f.dim;
    d.tail = f.method == FB_METHOD_OPTIMIZED_CONVOLUTION;
    d.num_iter = f.num_iter;
    d.n = f.n;
    d.pt_off = f.pt_offset;
    d.out_off = f.out_offset;
    int32_t K[3] = {1, 1, 1};
    double prod_scale = 1.0;
    for (int m = 0; m < 3; ++m) d.size[m] = 1;
    for (int m = 0; m < f.dim; ++m) {
        if (f.size[m] < 2) return fail(FB_ERR_INVALID_ARGUMENT, "grid size must be >= 2 along every axis");
        if (!finite_pos(f.step[m]) || !finite_pos(f.sigma[m]) || !std::isfinite(f.x0[m]))
            return fail(FB_ERR_INVALID_ARGUMENT, "sigma and step must be finite and positive, x0 finite");
        d.size[m] = f.size[m];
        d.x0[m] = f.x0[m];
        d.step[m] = f.step[m];
        d.sigma[m] = f.sigma[m];
        d.T[m] = d.tail ? half_kernel_opt(f.sigma[m], f.step[m], f.num_iter)
                        : half_kernel_conv(f.sigma[m], f.step[m], f.num_iter);
        if (d.T[m] < 0) return fail(FB_ERR_INVALID_ARGUMENT, "negative half kernel size");
        K[m] = 2 * d.T[m] + 1;
        if ((int64_t)K[m] >= f.size[m] || d.T[m] == INT32_MAX)
            return fail(FB_ERR_INVALID_ARGUMENT,
                        "resulting rectangular kernel size should be smaller w.r.t. specified grid");
        if (fast && !d.tail && std::pow((double)K[m], (double)f.num_iter) >= 9.2233720368547758e18)
            return fail(FB_ERR_UNSUPPORTED,
                        "FP32C does not support method convolution with kernel_size**num_iter >= 2**63: the "
                        "reference's integer threshold wraps around there and leaves nodes with weights below "
                        "the float32 range unmasked; use precision fp64");
        d.alpha[m] = d.tail ? tail_value(f.sigma[m], f.step[m], f.num_iter) : 0.0;
        d.pscale[m] = fast ? std::pow(1.0 / ((double)K[m] + 2 * d.alpha[m]), (double)f.num_iter) : 1.0;
        prod_scale *= d.pscale[m];
    }
    d.thr = threshold(d.tail, f.dim, d.sigma, d.step, K, f.num_iter, d.alpha, f.max_dist_weight);
    d.thr_fast = d.thr * prod_scale;
    d.thr_fast_f = float_threshold(d.thr_fast);
    return FB_OK;
}

// Work list and per-field geometry of the fast pass along `axis`: each field runs with the
// kind and segment length chosen for its line length (autotuned choice, else the cost
// model); the work items are grouped by (kind, segment length, workgroup size), one launch
// per group.
int build_pass(fb_plan *P, int axis, std::vector<PassWork> &work, std::vector<PassField> &pf,
               std::vector<PassGroup> &groups, const std::vector<char> *sel = nullptr)
{
    pf.assign(P->nfields, PassField{});
    std::map<int, std::vector<PassWork>> by_seg;
    std::map<int, size_t> lds_of;
    for (int f = 0; f < P->nfields; ++f) {
        if (sel && !(*sel)[f]) continue;   // a pass over the selected fields only (autotuning)
        const FieldDev &F = P->fh[f];
        const int64_t W = F.size[0], H = F.size[1], D = F.size[2];
        ChunkGeom cg;
        int rc = chunk_geom(F.size[axis], F.T[axis], F.num_iter, cg);
        if (rc) return rc;
        // a chunk's halo is read from the pass's input while other chunks store into it: only
        // passes whose input and output differ (the injecting first pass, the last pass) chunk
        if (cg.Lc && P->inplace_passes)
            return fail(FB_ERR_UNSUPPORTED, "lines of length " + std::to_string(F.size[axis]) +
                                                " exceed the FP32C line pass here; use precision fp64");
        if (cg.Lc && axis != 0 && axis != P->dim - 1)
            return fail(FB_ERR_UNSUPPORTED, "FP32C takes y lines longer than " + std::to_string(cg.Lc - 2 * cg.H) +
                                                " in 2D only; use precision fp64");
        const int64_t L = cg.Lc ? cg.Lc : F.size[axis];
        PassChoice ch = pass_choice(P, axis, L, F.T[axis]);
        if (cg.Lc && ch.kind != kPassSlide) ch = model_choice(L, F.T[axis]);   // chunks: the sliding pass
        const int seg = ch.seg;
        int lpl = 1, lines = 1, rpad = 0, wgs = 0;
        size_t l = 0;
        rc = ch.kind == kPassPrefix ? prefix_geometry(L, F.T[axis], seg, ch.lpf, lpl, lines, l, wgs)
                                    : pass_geometry(L, F.T[axis], seg, ch.lpf, lpl, lines, rpad, l,
                                                    axis == 0 ? kBanksRows : kBanksColumns);
        if (rc) return rc;
        const int nch = cg.Lc ? (int)((F.size[axis] + cg.C - 1) / cg.C) : 1;
        const int gkey = (ch.kind * 1024 + seg) * 1024 + ch.lpf;
        lds_of[gkey] = std::max(lds_of[gkey], l);
        PassField &p = pf[f];
        p.len = L;
        p.full_len = F.size[axis];
        p.chunk = (int32_t)cg.C;
        p.halo = (int32_t)cg.H;
        p.T = F.T[axis];
        p.lpl = lpl;
        p.lines = lines;
        p.rpad = rpad;
        p.seg = seg;
        p.alpha = F.alpha[axis];
        p.blocked = axis < 2 && field_blocked(P, f) ? (axis == 0 ? kBlockedOut : kBlockedIn) : kRowMajor;
        if (F.pscale[axis] != 1.0)
            scale_groups((double)(2 * F.T[axis] + 1) + 2 * F.alpha[axis], F.num_iter, F.pscale[axis], p);
        else
            scale_groups(1.0, F.num_iter, 1.0, p);
        std::vector<PassWork> &items = by_seg[gkey];
        if (axis == 0) {
            p.line_stride = W;
            p.elem_stride = 1;
            const int64_t nlines = H * D;
            for (int64_t l0 = 0; l0 < nlines; l0 += lines)
                for (int c = 0; c < nch; ++c)
                    items.push_back({f, (int32_t)std::min<int64_t>(lines, nlines - l0), l0 * W, c});
        } else if (axis == 1) {
            p.line_stride = 1;
            p.elem_stride = W;
            for (int64_t z = 0; z < D; ++z)
                for (int64_t x0 = 0; x0 < W; x0 += lines)
                    for (int c = 0; c < nch; ++c)
                        items.push_back({f, (int32_t)std::min<int64_t>(lines, W - x0), z * W * H + x0, c});
        } else {
            p.line_stride = 1;
            p.elem_stride = W * H;
            for (int64_t c0 = 0; c0 < W * H; c0 += lines)
                for (int c = 0; c < nch; ++c)
                    items.push_back({f, (int32_t)std::min<int64_t>(lines, W * H - c0), c0, c});
        }
    }
    work.clear();
    groups.clear();
    for (auto &kv : by_seg) {
        // x passes too when they write a blocked intermediate: the 4 workgroups whose rows share
        // its lines then run on one XCD, so their partial lines merge in its L2 (measured
        // slower for row-major x passes of 3D fields)
        bool any_blocked = false;
        for (const PassWork &w : kv.second)
            if (pf[w.field].blocked != kRowMajor) { any_blocked = true; break; }
        std::vector<PassWork> items = axis == 0 && !any_blocked ? kv.second : xcd_order(kv.second);
        groups.push_back({(kv.first / 1024) % 1024, kv.first % 1024, kv.first / (1024 * 1024), (int)work.size(),
                          (int)items.size(), lds_of[kv.first]});
        work.insert(work.end(), items.begin(), items.end());
    }
    if (work.size() > (size_t)INT32_MAX) return fail(FB_ERR_UNSUPPORTED, "too many work items");
    return FB_OK;
}

// (Re)builds and uploads the fast passes of all axes of an FP32C plan.
int build_passes(fb_plan *P)
{
    std::vector<PassWork> work[3];
    std::vector<PassField> pf[3];
    std::vector<fb_plan::Upload> ups;
    for (int a = 0; a < P->dim; ++a) {
        const int rc = build_pass(P, a, work[a], pf[a], P->groups[a]);
        if (rc) return rc;
        ups.push_back({reinterpret_cast<void **>(&P->d_work[a]), work[a].data(), work[a].size() * sizeof(PassWork)});
        ups.push_back({reinterpret_cast<void **>(&P->d_pass[a]), pf[a].data(), pf[a].size() * sizeof(PassField)});
    }
    return P->upload_all(ups);   // one copy for all axes
}

// One fast pass along `axis` over the whole batch: a launch per segment-length group.
int run_pass(fb_plan *P, int axis, bool inject, const uint32_t *vals, const double *val,
             float *out, float *wsave, int final_pass, hipStream_t s, const float *adj_g = nullptr)
{
    for (const PassGroup &g : P->groups[axis]) {
        int rc = launch_fast_pass(g.kind, inject, g.seg, P->d_fields, P->d_offsets, P->d_work[axis] + g.first, g.count,
                                  P->d_pass[axis], 2 * g.lpf, g.lds, P->d_rec, vals, P->d_line_start,
                                  P->d_line_count, val, P->d_v, P->d_w, out, wsave, final_pass, adj_g, s);
        if (rc) return rc;
    }
    return FB_OK;
}

// arena: the plan's device buffers are carved from [arena, arena + arena_cap) instead of the
// pool (fb_barnes_noalloc)
int plan_create(fb_plan **out, const fb_field *fields, int nfields, int precision, void *arena = nullptr,
                size_t arena_cap = 0, bool grids = true)
{
    *out = nullptr;
    if (!fields || nfields < 1) return fail(FB_ERR_INVALID_ARGUMENT, "no fields");
    if (precision != FB_PREC_FP64 && precision != FB_PREC_FP32C)
        return fail(FB_ERR_INVALID_ARGUMENT, "unknown precision");
    int rc0 = fast_pass_init();
    if (!rc0) rc0 = field_sort_init();
    if (rc0) return rc0;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    fb_plan *P = new fb_plan();
    P->device = device;
    P->arena = (char *)arena;
    P->arena_cap = arena_cap;
    P->precision = precision;
    P->nfields = nfields;
    P->dim = fields[0].dim;
    P->corners = 1 << P->dim;
    const int fast = precision == FB_PREC_FP32C;
    P->fh.resize(nfields);
    std::vector<int64_t> pt_prefix, xline_prefix, cell_prefix, lslot[3], rowwg;
    for (int f = 0; f < nfields; ++f) {
        int rc = make_field(fields[f], P->fh[f], fast);
        if (rc) { delete P; return rc; }
        FieldDev &d = P->fh[f];
        if (d.dim != P->dim) { delete P; return fail(FB_ERR_INVALID_ARGUMENT, "all fields of a plan share dim"); }
        if (d.pt_off != P->npoints) {
            delete P;
            return fail(FB_ERR_INVALID_ARGUMENT, "fields must hold consecutive point ranges (pt_offset)");
        }
        pt_prefix.push_back(P->npoints);
        xline_prefix.push_back(P->nxlines);
        cell_prefix.push_back(P->ncells);
        d.cell_off = P->ncells;
        d.xline_off = P->nxlines;
        const int64_t W = d.size[0], H = d.size[1], D = d.size[2];
        const int64_t nl[3] = {H * D, W * D, W * H};
        for (int a = 0; a < 3; ++a) {
            lslot[a].push_back(P->lslots[a]);
            if (a < P->dim) P->lslots[a] += 2 * nl[a];
        }
        P->row_rect = std::max(P->row_rect, 2 * (int64_t)d.T[0] + 1);
        P->npoints += d.n;
        P->nxlines += H * D;
        // 256-byte aligned field starts (float4 tiles); an FP32C 2D field that may keep its
        // intermediate blocked (8-row blocks) gets room for H rounded up to 8 rows
        const int64_t Hc = fast && P->dim == 2 && W % 4 == 0 ? (H + 7) / 8 * 8 : H;
        P->ncells += (W * Hc * D + 63) / 64 * 64;
        P->total_cells += W * H * D;
        P->total_out = std::max(P->total_out, d.out_off + W * H * D);
    }
    pt_prefix.push_back(P->npoints);
    xline_prefix.push_back(P->nxlines);
    cell_prefix.push_back(P->ncells);
    for (int a = 0; a < 3; ++a) lslot[a].push_back(P->lslots[a]);
    P->row_slots = exact_rows_slots(P->lslots[0]);
    for (const FieldDev &d : P->fh) {
        rowwg.push_back(P->rowwgs);
        P->rowwgs += (2 * d.size[1] * d.size[2] + P->row_slots - 1) / P->row_slots;
    }
    rowwg.push_back(P->rowwgs);
    P->nentries = P->npoints;   // one (base cell, point) pair per point
    if (P->ncells >= (int64_t)kInvalidKey || P->npoints >= (int64_t)UINT32_MAX || P->nentries >= (int64_t)UINT32_MAX) {
        delete P;
        return fail(FB_ERR_UNSUPPORTED, "a plan holds at most 2^32-2 grid cells and 2^32-1 contributions");
    }

    int rc = FB_OK;
#define TRY(x) do { rc = (x); if (rc) { delete P; return rc; } } while (0)
    {
        auto item = [](auto **dst, const auto &v) {
            return fb_plan::Upload{reinterpret_cast<void **>(dst), v.data(), v.size() * sizeof(v[0])};
        };
        TRY(P->upload_all({item(&P->d_fields, P->fh), item(&P->d_pt_prefix, pt_prefix),
                           item(&P->d_xline_prefix, xline_prefix), item(&P->d_cell_prefix, cell_prefix),
                           item(&P->d_lslot_prefix[0], lslot[0]), item(&P->d_lslot_prefix[1], lslot[1]),
                           item(&P->d_lslot_prefix[2], lslot[2]), item(&P->d_rowwg_prefix, rowwg)}));
    }
    TRY(P->alloc(&P->d_offsets, nfields));
    for (const FieldDev &d : P->fh) {
        P->max_n = std::max(P->max_n, d.n);
        P->max_rows = std::max(P->max_rows, d.size[1] * d.size[2]);
    }
    TRY(P->alloc(&P->d_off_part, offsets_part_elems(nfields, P->max_n)));
    TRY(P->alloc(&P->d_keys, P->nentries));
    TRY(P->alloc(&P->d_vals, P->nentries));
    TRY(P->alloc(&P->d_keys2, P->nentries));
    TRY(P->alloc(&P->d_vals2, P->nentries));
    if (P->max_n > kLocalSortMax) {
        TRY(P->alloc(&P->d_skeys, P->nentries));
        TRY(P->alloc(&P->d_big_list, 2 * P->nxlines));   // rows for the wave kernel, then larger rows
        TRY(P->alloc(&P->d_big_count, 2));
    }
    TRY(P->alloc(&P->d_line_count, P->nxlines));
    TRY(P->alloc(&P->d_line_start, P->nxlines));
    TRY(P->alloc(&P->d_scan_tmp, scan_temp_elems(P->nxlines)));
    if (!grids) {
        // the sort and injection stages only (fb_inject): the caller's grids take the sums
        TRY(P->alloc(&P->d_irec, P->npoints));
    } else if (!fast) {
        TRY(P->alloc(&P->d_vg, P->ncells));
        TRY(P->alloc(&P->d_wg, P->ncells));
        TRY(P->alloc(&P->d_vs, P->ncells));
        TRY(P->alloc(&P->d_ws, P->ncells));
        TRY(P->alloc(&P->d_irec, P->npoints));
    } else {
        TRY(P->alloc(&P->d_v, P->ncells));
        TRY(P->alloc(&P->d_w, P->ncells));
        TRY(P->alloc(&P->d_rec, P->npoints));
        TRY(build_passes(P));
    }
#undef TRY
    *out = P;
    return FB_OK;
}

void stage_mark(fb_plan *P, hipStream_t s, const char *name)
{
    if (!P->timing) return;
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) { (void)hipGetLastError(); return; }
    (void)hipEventRecord(e, s);
    P->events.push_back(e);
    P->stage_names.push_back(name);
}

// The sort stage of a run: the value offsets and the points ordered by base row, (base cell,
// point index); *sorted receives the sorted point indices.
int plan_sort(fb_plan *P, const double *pts, const double *val, hipStream_t s, const uint32_t **sorted)
{
    int rc;
    // Per-row point counts by integer atomics and their exclusive scan give the rows' ranges of
    // the sorted list.  Fields of at most kLocalSortMax points: one workgroup sorts each field
    // in LDS.  Larger fields: every point goes to its atomic slot of its row's range, then each
    // row is sorted by (key, point index).  Both give the order of a stable sort of the (key,
    // point index) pairs.
    const bool local = P->max_n <= kLocalSortMax;
    if ((rc = launch_zero(P->d_line_count, P->nxlines * 4, s))) return rc;
    if (!local && (rc = launch_zero(P->d_big_count, 8, s))) return rc;
    if (offsets_single_pass(P->max_n, P->d_off_part)) {
        if ((rc = launch_point_keys_offsets(P->d_fields, P->d_pt_prefix, P->nfields, P->npoints, pts, P->d_keys,
                                            P->d_vals, P->d_line_count, local ? nullptr : P->d_vals,
                                            local ? nullptr : P->d_keys2, val, P->d_offsets, P->d_rec, s)))
            return rc;
    } else {
        if ((rc = launch_offsets(P->d_fields, P->nfields, P->max_n, val, P->d_off_part, P->d_offsets, s))) return rc;
        if ((rc = launch_point_keys(P->d_fields, P->d_pt_prefix, P->nfields, P->npoints, pts, P->d_keys, P->d_vals,
                                    P->d_line_count, local ? nullptr : P->d_vals, local ? nullptr : P->d_keys2, P->max_rows, P->d_rec, s)))
            return rc;
    }
    uint32_t *vs = P->d_vals2;
    if (local) {
        // the field sort scans each field's row counts itself (a field's list starts at its
        // first point index)
        if ((rc = launch_field_sort(P->d_fields, P->d_pt_prefix, P->nfields, (int)P->max_n, P->d_keys,
                                    P->d_line_count, P->d_line_start, P->d_vals2, s)))
            return rc;
    } else {
        if ((rc = scan_exclusive(P->d_line_count, P->d_line_start, P->nxlines, P->d_scan_tmp, s))) return rc;
        // d_vals holds the slots and d_keys2 the rows; both, with d_keys, are free after the
        // scatter and serve as scratch for rows of more than kLocalSortMax points
        if ((rc = row_bucket_sort(P->d_keys, P->d_vals, P->d_keys2, P->npoints, P->d_line_start, P->d_line_count,
                                  P->nxlines, P->d_skeys, P->d_vals2, P->d_big_list, P->d_big_count, P->d_keys,
                                  P->d_vals, s)))
            return rc;
    }
    *sorted = vs;
    return FB_OK;
}

int plan_run(fb_plan *P, const double *pts, const double *val, float *out, float *wsave, hipStream_t s)
{
    int rc;
    for (hipEvent_t e : P->events) (void)hipEventDestroy(e);
    P->events.clear();
    P->stage_names.clear();
    stage_mark(P, s, "begin");
    const uint32_t *vs = nullptr;
    if ((rc = plan_sort(P, pts, val, s, &vs))) return rc;
    P->sorted_vals = vs;
    stage_mark(P, s, "sort");
    if (P->precision == FB_PREC_FP64) {
        if ((rc = launch_zero(P->d_vg, P->ncells * 8, s))) return rc;
        if ((rc = launch_zero(P->d_wg, P->ncells * 8, s))) return rc;
        if ((rc = launch_inject_records(P->d_fields, P->d_pt_prefix, P->nfields, P->npoints,
                                        P->max_n <= kLocalSortMax, P->nxlines, vs, P->d_line_start,
                                        P->d_line_count, pts, val, P->d_offsets, P->d_irec, s)) ||
            (rc = launch_exact_inject(P->d_fields, P->d_xline_prefix, P->nfields, P->nxlines, P->d_irec,
                                      P->d_line_start, P->d_line_count, P->d_vg, P->d_wg, P->dim, true, s)))
            return rc;
        stage_mark(P, s, "inject");
        static const char *ln[3] = {"lines_x", "lines_y", "lines_z"};
        for (int a = 0; a < P->dim; ++a) {
            if (a == 0 && exact_rows_ring(P->row_rect))
                rc = launch_exact_rows(P->d_fields, P->d_rowwg_prefix, P->d_lslot_prefix[0], P->nfields, P->rowwgs,
                                       P->row_slots, -1, P->row_rect, P->d_vg, P->d_wg, P->d_vs, P->d_ws, s);
            else
                rc = launch_exact_lines(P->d_fields, P->d_lslot_prefix[a], P->nfields, P->lslots[a], a, -1, P->d_vg,
                                        P->d_wg, P->d_vs, P->d_ws, s);
            if (rc) return rc;
            stage_mark(P, s, ln[a]);
        }
        if ((rc = launch_exact_finalize(P->d_fields, P->d_cell_prefix, P->nfields, P->ncells, P->d_vg, P->d_wg,
                                        P->d_offsets, out, s)))
            return rc;
        stage_mark(P, s, "finalize");
    } else {
        static const char *pn[3] = {"pass_x", "pass_y", "pass_z"};
        for (int a = 0; a < P->dim; ++a) {
            const int fin = a == P->dim - 1;
            if ((rc = run_pass(P, a, a == 0, vs, val, out, fin ? wsave : nullptr, fin, s))) return rc;
            stage_mark(P, s, pn[a]);
        }
    }
    return FB_OK;
}

int check_device_ptr(const void *p, const char *what)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return fail(FB_ERR_INVALID_ARGUMENT, std::string(what) + " is not a device pointer");
    }
    if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)
        return fail(FB_ERR_INVALID_ARGUMENT, std::string(what) + " is not device memory");
    return FB_OK;
}

// the batch descriptor of the one field of a drop-in call
fb_field single_field(int method, int dim, int64_t n, const double *sigma, const double *x0, const double *step,
                      const int64_t *size, int num_iter, double max_dist_weight)
{
    fb_field f;
    std::memset(&f, 0, sizeof(f));
    f.dim = dim;
    f.method = method;
    f.num_iter = num_iter;
    f.n = n;
    for (int m = 0; m < 3; ++m) f.size[m] = 1;
    for (int m = 0; m < dim; ++m) {
        f.size[m] = size[m];
        f.sigma[m] = sigma[m];
        f.x0[m] = x0[m];
        f.step[m] = step[m];
    }
    f.max_dist_weight = max_dist_weight;
    return f;
}

// Device address of this device's copy of `scratch`, looked up once per device (with that
// device current) by its first caller.
int noalloc_scratch(int device, void **ptr)
{
    static std::mutex mu;
    static std::map<int, void *> addr;
    std::lock_guard<std::mutex> g(mu);
    auto it = addr.find(device);
    if (it != addr.end()) {
        *ptr = it->second;
        return FB_OK;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(FB_ERR_HIP, std::string("address of the no-allocation scratch array: ") + hipGetErrorString(e));
    }
    addr[device] = p;
    *ptr = p;
    return FB_OK;
}

// scoped device buffers for the single-call entry points; with `stream` set, the destructor
// waits for the work queued there (an error return may leave kernels in flight) before the
// buffers go back to the pool
struct Scratch {
    int device = 0;
    hipStream_t stream = nullptr;
    bool has_stream = false;
    std::vector<DevBuf> b;
    void on(hipStream_t s)
    {
        stream = s;
        has_stream = true;
    }
    template <class T>
    int get(T **p, size_t count)
    {
        void *q = nullptr;
        int rc = pool_alloc(device, count * sizeof(T), &q);
        if (rc) return rc;
        b.push_back({q, count * sizeof(T)});
        *p = (T *)q;
        return FB_OK;
    }
    ~Scratch()
    {
        if (has_stream && !b.empty() && hipStreamSynchronize(stream) != hipSuccess) (void)hipGetLastError();
        for (auto &x : b) pool_free(device, x.p, x.bytes);
    }
};

}  // namespace
}  // namespace fb

// ============================================================================
// C ABI
// ============================================================================

extern "C" {

int fb_version(void) { return 10000; }

const char *fb_last_error(void) { return fb::last_error(); }

int fb_device_arch(char *arch, int cap, int *compute_units)
{
    if (!arch || cap <= 0) return fail(FB_ERR_INVALID_ARGUMENT, "arch buffer is NULL or empty");
    int dev = 0;
    FB_HIP_CHECK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    FB_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    std::snprintf(arch, (size_t)cap, "%s", prop.gcnArchName);
    if (compute_units) *compute_units = prop.multiProcessorCount;
    return FB_OK;
}

int fb_device_count(int *count)
{
    if (!count) return fail(FB_ERR_INVALID_ARGUMENT, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        n = 0;
    }
    *count = n;
    return FB_OK;
}

int fb_half_kernel_size(int method, int dim, const double *sigma, const double *step, int num_iter, int32_t *out)
{
    if (dim < 1 || dim > 3 || !sigma || !step || !out) return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    if (num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
    for (int m = 0; m < dim; ++m)
        out[m] = method == FB_METHOD_CONVOLUTION ? half_kernel_conv(sigma[m], step[m], num_iter)
                                                 : half_kernel_opt(sigma[m], step[m], num_iter);
    return FB_OK;
}

int fb_tail_value(int dim, const double *sigma, const double *step, int num_iter, double *out)
{
    if (dim < 1 || dim > 3 || !sigma || !step || !out) return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    if (num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
    for (int m = 0; m < dim; ++m) out[m] = tail_value(sigma[m], step[m], num_iter);
    return FB_OK;
}

int fb_sigma_effective(int dim, const double *sigma, const double *step, int num_iter, double *out)
{
    if (dim < 1 || dim > 3 || !sigma || !step || !out) return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    if (num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
    for (int m = 0; m < dim; ++m) out[m] = sigma_effective(sigma[m], step[m], num_iter);
    return FB_OK;
}

int fb_threshold(int method, int dim, const double *sigma, const double *step, const int32_t *kernel_size,
                 int num_iter, const double *tail_value_, double max_dist_weight, double *thr)
{
    if (dim < 1 || dim > 3 || !sigma || !step || !kernel_size || !thr)
        return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    const int tail = method == FB_METHOD_OPTIMIZED_CONVOLUTION;
    if (tail && !tail_value_) return fail(FB_ERR_INVALID_ARGUMENT, "tail_value is NULL");
    double zero[3] = {0, 0, 0};
    *thr = threshold(tail, dim, sigma, step, kernel_size, num_iter, tail ? tail_value_ : zero, max_dist_weight);
    return FB_OK;
}

int fb_plan_create(fb_plan **plan, const fb_field *fields, int nfields, int precision, unsigned flags)
{
    (void)flags;
    if (!plan) return fail(FB_ERR_INVALID_ARGUMENT, "plan is NULL");
    return plan_create(plan, fields, nfields, precision);
}

int fb_plan_execute(fb_plan *P, const double *d_pts, const double *d_val, float *d_out, void *stream)
{
    if (!P || !d_pts || !d_val || !d_out) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    const int rc = plan_run(P, d_pts, d_val, d_out, nullptr, (hipStream_t)stream);
    const int rc2 = P->mark_use((hipStream_t)stream);
    return rc ? rc : rc2;
}

int fb_plan_execute_graph(fb_plan *P, const double *d_pts, const double *d_val, float *d_out, void *stream)
{
    if (!P || !d_pts || !d_val || !d_out) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    if (!s) return fail(FB_ERR_INVALID_ARGUMENT, "graph execution needs a non-default stream");
    if (!P->gexec || P->g_pts != d_pts || P->g_val != d_val || P->g_out != d_out || P->g_stream != s) {
        if (P->gexec) { (void)hipGraphExecDestroy(P->gexec); P->gexec = nullptr; }
        if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
        const int timing = P->timing;
        P->timing = 0;
        FB_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = plan_run(P, d_pts, d_val, d_out, nullptr, s);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s, &g);
        P->timing = timing;
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(FB_ERR_HIP, std::string("graph capture failed: ") + hipGetErrorString(e));
        P->graph = g;
        FB_HIP_CHECK(hipGraphInstantiate(&P->gexec, g, nullptr, nullptr, 0));
        P->g_pts = d_pts; P->g_val = d_val; P->g_out = d_out; P->g_stream = s;
    }
    FB_HIP_CHECK(hipGraphLaunch(P->gexec, s));
    return P->mark_use(s);
}

int fb_plan_forward(fb_plan *P, const double *d_pts, const double *d_val, float *d_out, float *d_wsave, void *stream)
{
    if (!P || !d_pts || !d_val || !d_out || !d_wsave) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P->precision != FB_PREC_FP32C)
        return fail(FB_ERR_UNSUPPORTED, "the training forward and the adjoint need a plan of precision FP32C");
    const int rc = plan_run(P, d_pts, d_val, d_out, d_wsave, (hipStream_t)stream);
    const int rc2 = P->mark_use((hipStream_t)stream);
    return rc ? rc : rc2;
}

int fb_plan_backward(fb_plan *P, const double *d_pts, const double *d_val, const float *d_out, const float *d_wsave,
                     const float *d_gout, double *d_gval, double *d_gpts, void *stream)
{
    if (!P || !d_pts || !d_val || !d_out || !d_wsave || !d_gout || !d_gval)
        return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P->precision != FB_PREC_FP32C)
        return fail(FB_ERR_UNSUPPORTED, "the training forward and the adjoint need a plan of precision FP32C");
    // the adjoint's later passes filter their planes in place, which chunked lines cannot
    for (const FieldDev &F : P->fh)
        for (int a = 1; a < P->dim; ++a)
            if (pass_len(F, a) != F.size[a])
                return fail(FB_ERR_UNSUPPORTED, "the adjoint takes lines of at most the FP32C line pass's length "
                                                "along y and z");
    hipStream_t s = (hipStream_t)stream;
    int rc = launch_offsets(P->d_fields, P->nfields, P->max_n, d_val, P->d_off_part, P->d_offsets, s);
    // the same scaled line passes; the first one forms a = g/W and b = -a(f - offset) in its tile
    // load, so a and b never round-trip through HBM
    for (int a = 0; a < P->dim && !rc; ++a)
        rc = run_pass(P, a, false, nullptr, nullptr, a == 0 ? const_cast<float *>(d_out) : nullptr,
                      a == 0 ? const_cast<float *>(d_wsave) : nullptr, 0, s, a == 0 ? d_gout : nullptr);
    if (!rc)
        rc = launch_adj_gather(P->d_fields, P->d_pt_prefix, P->nfields, P->npoints, d_pts, d_val, P->d_offsets,
                               P->d_v, P->d_w, P->dim >= 2 ? P->d_pass[1] : nullptr, d_gval, d_gpts, s);
    const int rc2 = P->mark_use(s);
    return rc ? rc : rc2;
}

// Device time (ms, mean of 5 after a warm-up) of the pass `fw`/`pf`/`fg` along axis a (built
// by build_pass, for all fields or a selection of them), on the plan's own buffers.
static int time_pass(fb_plan *P, int a, const std::vector<PassWork> &fw, const std::vector<PassField> &pf,
                     const std::vector<PassGroup> &fg, const double *val, float *out,
                     hipStream_t s, float *ms)
{
    *ms = 0.f;
    if (fw.empty()) return FB_OK;
    Scratch sc;
    sc.device = P->device;
    sc.on(s);
    PassWork *dw;
    PassField *dp;
    int rc;
    if ((rc = sc.get(&dw, fw.size())) || (rc = sc.get(&dp, pf.size()))) return rc;
    FB_HIP_CHECK(hipMemcpy(dw, fw.data(), fw.size() * sizeof(PassWork), hipMemcpyHostToDevice));
    FB_HIP_CHECK(hipMemcpy(dp, pf.data(), pf.size() * sizeof(PassField), hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    FB_HIP_CHECK(hipEventCreate(&e0));
    FB_HIP_CHECK(hipEventCreate(&e1));
    const int fin = a == P->dim - 1;
    for (int rep = 0; rep < 6 && !rc; ++rep) {
        if (rep == 1) (void)hipEventRecord(e0, s);
        for (const PassGroup &g : fg) {
            rc = launch_fast_pass(g.kind, a == 0, g.seg, P->d_fields, P->d_offsets, dw + g.first, g.count, dp, 2 * g.lpf,
                                  g.lds, P->d_rec, P->sorted_vals, P->d_line_start, P->d_line_count, val, P->d_v,
                                  P->d_w, out, nullptr, fin, nullptr, s);
            if (rc) break;
        }
    }
    (void)hipEventRecord(e1, s);
    hipError_t e = hipEventSynchronize(e1);
    if (!rc && e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (e != hipSuccess) return fail(FB_ERR_HIP, std::string("autotune timing failed: ") + hipGetErrorString(e));
    *ms /= 5.f;
    return FB_OK;
}

int fb_plan_autotune(fb_plan *P, const double *d_pts, const double *d_val, float *d_out, void *stream)
{
    if (!P || !d_pts || !d_val || !d_out) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P->precision != FB_PREC_FP32C) return FB_OK;   // the FP64 path has nothing to tune
    hipStream_t s = (hipStream_t)stream;
    int rc = plan_run(P, d_pts, d_val, d_out, nullptr, s);   // sorted injection lists for the x pass
    const int rc2 = P->mark_use(s);
    if (rc || rc2) return rc ? rc : rc2;
    FB_HIP_CHECK(hipStreamSynchronize(s));
    for (int a = 0; a < P->dim; ++a) {
        std::map<int64_t, std::vector<char>> classes;   // pass line length -> fields with it
        std::map<int64_t, bool> chunked;                 // ... some of them in chunks
        for (int f = 0; f < P->nfields; ++f) {
            const int64_t len = pass_len(P->fh[f], a);
            if (len != P->fh[f].size[a]) chunked[len] = true;
            std::vector<char> &sel = classes[len];
            sel.resize(P->nfields, 0);
            sel[f] = 1;
        }
        for (auto &cl : classes) {
            // the cost model's choice (for the class's first field) is the baseline; a
            // candidate replaces it only when at least 2 % faster (timing noise is ~1 %)
            int f0 = 0;
            while (!cl.second[f0]) ++f0;
            const PassChoice base = model_choice(cl.first, P->fh[f0].T[a]);
            PassChoice best{-1, kBlock, kPassSlide};
            std::vector<PassChoice> cands{base};
            auto same = [](const PassChoice &x, const PassChoice &y) {
                return x.seg == y.seg && x.lpf == y.lpf && x.kind == y.kind;
            };
            for (int lpf : kLpfs) {
                for (int seg : kPfxSegs)   // chunked lines run the sliding pass only
                    if (!chunked[cl.first] && !same({seg, lpf, kPassPrefix}, base))
                        cands.push_back({seg, lpf, kPassPrefix});
                for (int seg : kSegs)
                    if (!same({seg, lpf, kPassSlide}, base)) cands.push_back({seg, lpf, kPassSlide});
            }
            std::vector<PassChoice> adm;
            for (const PassChoice &cand : cands) {
                bool ok = true;
                for (int f = 0; f < P->nfields && ok; ++f)
                    if (cl.second[f]) choice_cost(cl.first, P->fh[f].T[a], cand, &ok);
                if (ok) adm.push_back(cand);
            }
            // Each candidate is timed on a pass over the class's fields alone (built, grouped and
            // XCD-ordered as a pass of its own: a subset of a pass shared with other classes
            // would lose the XCD order of its work items), in two interleaved rounds keeping
            // each candidate's faster time, so that clock drift during the sweep favours none.
            auto time_choice = [&](const PassChoice &c, const std::vector<char> *sel, float *ms) {
                P->seg_choice[{a, cl.first}] = c;
                std::vector<PassWork> work;
                std::vector<PassField> pf;
                std::vector<PassGroup> groups;
                int r = build_pass(P, a, work, pf, groups, sel);
                return r ? r : time_pass(P, a, work, pf, groups, d_val, d_out, s, ms);
            };
            std::vector<float> tms(adm.size(), 0.f);
            for (int round = 0; round < 2; ++round)
                for (size_t c = 0; c < adm.size(); ++c) {
                    float ms = 0.f;
                    if ((rc = time_choice(adm[c], &cl.second, &ms))) return rc;
                    tms[c] = round == 0 ? ms : std::min(tms[c], ms);
                }
            if (!adm.empty()) {
                size_t ib = 0, im = 0;   // the base (when admissible) and the fastest candidate
                for (size_t c = 0; c < adm.size(); ++c) {
                    if (same(adm[c], base)) ib = c;
                    if (tms[c] < tms[im]) im = c;
                }
                best = same(adm[ib], base) && !(tms[im] < 0.98f * tms[ib]) ? adm[ib] : adm[im];
                // A winner other than the base is confirmed on the whole pass (all classes, two
                // interleaved rounds): it is kept only if the pass is not slower with it.
                if (!same(best, adm[ib]) && same(adm[ib], base) && classes.size() > 1) {
                    float tb = 0.f, tw = 0.f;
                    for (int round = 0; round < 2; ++round) {
                        float m0 = 0.f, m1 = 0.f;
                        if ((rc = time_choice(base, nullptr, &m0)) || (rc = time_choice(best, nullptr, &m1)))
                            return rc;
                        tb = round == 0 ? m0 : std::min(tb, m0);
                        tw = round == 0 ? m1 : std::min(tw, m1);
                    }
                    if (!(tw <= tb)) best = base;
                }
            }
            if (best.seg < 0) P->seg_choice.erase({a, cl.first});
            else P->seg_choice[{a, cl.first}] = best;
        }
    }
    if ((rc = build_passes(P))) return rc;
    if (P->gexec) { (void)hipGraphExecDestroy(P->gexec); P->gexec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    return FB_OK;
}

int fb_plan_choose_pass(fb_plan *P, int axis, int64_t len, int kind, int seg, int threads)
{
    if (!P) return fail(FB_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (P->precision != FB_PREC_FP32C) return fail(FB_ERR_INVALID_ARGUMENT, "pass choices apply to FP32C plans");
    if (axis < 0 || axis >= P->dim) return fail(FB_ERR_INVALID_ARGUMENT, "axis out of range");
    if (kind != kPassSlide && kind != kPassPrefix) return fail(FB_ERR_INVALID_ARGUMENT, "unknown pass kind");
    if (threads != 128 && threads != 256 && threads != 2 * kBlock)
        return fail(FB_ERR_INVALID_ARGUMENT, "threads must be 128, 256 or 512");
    if (!fast_pass_seg_supported(kind, seg))
        return fail(FB_ERR_INVALID_ARGUMENT, "no pass instantiation for segment length " + std::to_string(seg));
    const PassChoice ch{seg, threads / 2, kind};
    bool any = false;
    for (int f = 0; f < P->nfields; ++f) {
        if (pass_len(P->fh[f], axis) != len) continue;
        if (kind != kPassSlide && len != P->fh[f].size[axis])
            return fail(FB_ERR_UNSUPPORTED, "lines that run in chunks take the sliding pass only");
        any = true;
        bool ok = false;
        choice_cost(len, P->fh[f].T[axis], ch, &ok);
        if (!ok) return fail(FB_ERR_UNSUPPORTED, "the pass choice does not fit lines of length " + std::to_string(len));
    }
    if (!any) return fail(FB_ERR_INVALID_ARGUMENT, "no field has lines of length " + std::to_string(len));
    const auto key = std::make_pair(axis, len);
    const bool had = P->seg_choice.count(key) != 0;
    const PassChoice old = had ? P->seg_choice[key] : ch;
    P->seg_choice[key] = ch;
    int rc = build_passes(P);
    if (rc) {
        if (had) P->seg_choice[key] = old;
        else P->seg_choice.erase(key);
        const std::string keep = last_error();
        (void)build_passes(P);
        set_error(keep);
        return rc;
    }
    if (P->gexec) { (void)hipGraphExecDestroy(P->gexec); P->gexec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    return FB_OK;
}

int fb_plan_pass_choices(fb_plan *P, int axis, int64_t *len, int32_t *kind, int32_t *seg, int32_t *threads,
                         int max_classes, int *nclasses, int32_t *field_class)
{
    if (!P || !nclasses) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (P->precision != FB_PREC_FP32C) return fail(FB_ERR_INVALID_ARGUMENT, "pass choices apply to FP32C plans");
    if (axis < 0 || axis >= P->dim) return fail(FB_ERR_INVALID_ARGUMENT, "axis out of range");
    std::map<int64_t, int> cls;   // pass line length -> a field with it (the first)
    for (int f = 0; f < P->nfields; ++f) cls.emplace(pass_len(P->fh[f], axis), f);
    std::map<int64_t, int> index;
    int i = 0;
    for (const auto &c : cls) {
        index[c.first] = i;
        if (i < max_classes) {
            const PassChoice ch = pass_choice(P, axis, c.first, P->fh[c.second].T[axis]);
            if (len) len[i] = c.first;
            if (kind) kind[i] = ch.kind;
            if (seg) seg[i] = ch.seg;
            if (threads) threads[i] = 2 * ch.lpf;
        }
        ++i;
    }
    *nclasses = i;
    if (field_class)
        for (int f = 0; f < P->nfields; ++f) field_class[f] = index[pass_len(P->fh[f], axis)];
    return FB_OK;
}

int fb_plan_pass_groups(fb_plan *P, int axis, int32_t *seg, int32_t *threads, int32_t *items, int32_t *kind,
                        int max_groups, int *ngroups)
{
    if (!P || !ngroups) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (axis < 0 || axis > 2) return fail(FB_ERR_INVALID_ARGUMENT, "axis must be 0, 1 or 2");
    const std::vector<PassGroup> &g = P->groups[axis];
    *ngroups = (int)g.size();
    for (int i = 0; i < (int)g.size() && i < max_groups; ++i) {
        if (seg) seg[i] = g[i].seg;
        if (threads) threads[i] = 2 * g[i].lpf;
        if (items) items[i] = g[i].count;
        if (kind) kind[i] = g[i].kind;
    }
    return FB_OK;
}

int fb_plan_set_blocked(fb_plan *P, int enable)
{
    if (!P) return fail(FB_ERR_INVALID_ARGUMENT, "plan is NULL");
    P->allow_blocked = enable ? 1 : 0;
    if (P->precision != FB_PREC_FP32C) return FB_OK;
    int rc = build_passes(P);
    if (rc) return rc;
    if (P->gexec) { (void)hipGraphExecDestroy(P->gexec); P->gexec = nullptr; }
    if (P->graph) { (void)hipGraphDestroy(P->graph); P->graph = nullptr; }
    return FB_OK;
}

int fb_plan_field_layout(fb_plan *P, int field, int *blocked)
{
    if (!P || !blocked) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (field < 0 || field >= P->nfields) return fail(FB_ERR_INVALID_ARGUMENT, "field out of range");
    *blocked = P->precision == FB_PREC_FP32C && field_blocked(P, field) ? 1 : 0;
    return FB_OK;
}

int fb_pass_layout(int64_t len, int T, int seg, int threads, int column_tiles, int *lanes_per_line, int *lines,
                   int *gap, int64_t *lds_bytes)
{
    if (!lanes_per_line || !lines || !gap || !lds_bytes) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (len < 1 || T < 0) return fail(FB_ERR_INVALID_ARGUMENT, "len must be >= 1 and T >= 0");
    if (threads != 128 && threads != 256 && threads != 2 * kBlock)
        return fail(FB_ERR_INVALID_ARGUMENT, "threads must be 128, 256 or 512");
    if (!fast_pass_seg_supported(kPassSlide, seg))
        return fail(FB_ERR_INVALID_ARGUMENT, "no pass instantiation for segment length " + std::to_string(seg));
    size_t lds = 0;
    const int rc = pass_geometry(len, T, seg, threads / 2, *lanes_per_line, *lines, *gap, lds,
                                 column_tiles ? kBanksColumns : kBanksRows);
    *lds_bytes = (int64_t)lds;
    return rc;
}

int fb_plan_set_timing(fb_plan *P, int enable)
{
    if (!P) return fail(FB_ERR_INVALID_ARGUMENT, "plan is NULL");
    P->timing = enable ? 1 : 0;
    return FB_OK;
}

int fb_plan_stage_times(fb_plan *P, float *ms, int max_stages, int *nstages, char *names, int names_len)
{
    if (!P || !nstages) return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    const int n = P->events.size() > 1 ? (int)P->events.size() - 1 : 0;
    *nstages = n;
    std::string all;
    for (int i = 0; i < n; ++i) {
        FB_HIP_CHECK(hipEventSynchronize(P->events[i + 1]));
        float t = 0.f;
        FB_HIP_CHECK(hipEventElapsedTime(&t, P->events[i], P->events[i + 1]));
        if (ms && i < max_stages) ms[i] = t;
        if (i) all += ",";
        all += P->stage_names[i + 1];
    }
    if (names && names_len > 0) {
        std::strncpy(names, all.c_str(), names_len - 1);
        names[names_len - 1] = 0;
    }
    return FB_OK;
}

int fb_plan_info(fb_plan *P, int64_t *total_points, int64_t *workspace_bytes)
{
    if (!P) return fail(FB_ERR_INVALID_ARGUMENT, "plan is NULL");
    if (total_points) *total_points = P->total_cells;
    if (workspace_bytes) *workspace_bytes = (int64_t)P->ws_bytes;
    return FB_OK;
}

int fb_plan_destroy(fb_plan *P)
{
    delete P;
    return FB_OK;
}

int fb_barnes(int method, int precision, int dim, int64_t n, const double *pts, const double *val,
              const double *sigma, const double *x0, const double *step, const int64_t *size, int num_iter,
              double max_dist_weight, double min_weight, void *out, unsigned flags, void *stream)
{
    if (dim < 1 || dim > 3)
        return fail(FB_ERR_INVALID_ARGUMENT, "Barnes interpolation supports only sample points in dimensions 1, 2 or 3");
    if (n < 1 || !pts || !val || !sigma || !x0 || !step || !size || !out)
        return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument or no sample points");
    const bool on_dev = flags & FB_FLAG_DEVICE_PTRS;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    hipStream_t s = stream ? (hipStream_t)stream : lib_stream(device);
    int64_t cells = 1;
    for (int m = 0; m < dim; ++m) {
        if (size[m] < 1) return fail(FB_ERR_INVALID_ARGUMENT, "grid size must be positive");
        cells *= size[m];
    }
    Scratch sc;
    sc.device = device;
    sc.on(s);
    int rc;
    const double *d_pts = pts, *d_val = val;
    if (on_dev) {
        if ((rc = check_device_ptr(pts, "pts")) || (rc = check_device_ptr(val, "val")) ||
            (rc = check_device_ptr(out, "out")))
            return rc;
    } else {
        double *a, *b;
        if ((rc = sc.get(&a, n * dim)) || (rc = sc.get(&b, n))) return rc;
        FB_HIP_CHECK(hipMemcpyAsync(a, pts, n * dim * sizeof(double), hipMemcpyHostToDevice, s));
        FB_HIP_CHECK(hipMemcpyAsync(b, val, n * sizeof(double), hipMemcpyHostToDevice, s));
        d_pts = a;
        d_val = b;
    }

    if (method == FB_METHOD_NAIVE || method == FB_METHOD_RADIUS) {
        if (method == FB_METHOD_RADIUS && dim != 2)
            return fail(FB_ERR_INVALID_ARGUMENT, "radius algorithm works only in 2D");
        FieldDev F;
        std::memset(&F, 0, sizeof(F));
        F.dim = dim;
        F.n = n;
        for (int m = 0; m < 3; ++m) F.size[m] = 1;
        for (int m = 0; m < dim; ++m) {
            F.size[m] = size[m];
            F.x0[m] = x0[m];
            F.step[m] = step[m];
            F.sigma[m] = sigma[m];
        }
        FieldDev *d_f;
        double *d_off, *d_out = (double *)out;
        if ((rc = sc.get(&d_f, 1)) || (rc = sc.get(&d_off, 1))) return rc;
        if (!on_dev && (rc = sc.get(&d_out, cells))) return rc;
        FB_HIP_CHECK(hipMemcpy(d_f, &F, sizeof(F), hipMemcpyHostToDevice));
        double *d_part;
        if ((rc = sc.get(&d_part, offsets_part_elems(1, n)))) return rc;
        if ((rc = launch_offsets(d_f, 1, n, d_val, d_part, d_off, s))) return rc;
        double r2 = 0.0;
        if (method == FB_METHOD_RADIUS) {
            // kdtree.prepare_search: radius**2 with radius = sqrt(-2 log(min_weight)) * sigma
            const double radius = std::sqrt(-2.0 * std::log(min_weight)) * sigma[0];
            r2 = radius * radius;
        }
        if ((rc = launch_naive(d_f, method == FB_METHOD_RADIUS, r2, max_dist_weight, d_pts, d_val, d_off, d_out,
                               cells, s)))
            return rc;
        if (!on_dev) FB_HIP_CHECK(hipMemcpyAsync(out, d_out, cells * sizeof(double), hipMemcpyDeviceToHost, s));
        FB_HIP_CHECK(hipStreamSynchronize(s));
        return FB_OK;
    }
    if (method != FB_METHOD_OPTIMIZED_CONVOLUTION && method != FB_METHOD_CONVOLUTION)
        return fail(FB_ERR_INVALID_ARGUMENT, "encountered invalid Barnes interpolation method");

    const fb_field f = single_field(method, dim, n, sigma, x0, step, size, num_iter, max_dist_weight);
    fb_plan *P = nullptr;
    if ((rc = plan_create(&P, &f, 1, precision))) return rc;
    float *d_out = (float *)out;
    if (!on_dev && (rc = sc.get(&d_out, cells))) { delete P; return rc; }
    rc = plan_run(P, d_pts, d_val, d_out, nullptr, s);
    if (const int rc2 = P->mark_use(s)) rc = rc ? rc : rc2;
    if (!rc && !on_dev) {
        hipError_t e = hipMemcpyAsync(out, d_out, cells * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = fail(FB_ERR_HIP, std::string("copy of the result failed: ") + hipGetErrorString(e));
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(FB_ERR_HIP, std::string("execution failed: ") + hipGetErrorString(e));
    }
    delete P;
    return rc;
}

int fb_barnes_noalloc(int method, int precision, int dim, int64_t n, const double *d_pts, const double *d_val,
                      const double *sigma, const double *x0, const double *step, const int64_t *size, int num_iter,
                      double max_dist_weight, float *d_out, void *stream)
{
    if (dim < 1 || dim > 3)
        return fail(FB_ERR_INVALID_ARGUMENT, "Barnes interpolation supports only sample points in dimensions 1, 2 or 3");
    if (n < 1 || !d_pts || !d_val || !sigma || !x0 || !step || !size || !d_out)
        return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument or no sample points");
    if (method != FB_METHOD_OPTIMIZED_CONVOLUTION && method != FB_METHOD_CONVOLUTION)
        return fail(FB_ERR_UNSUPPORTED, "fb_barnes_noalloc runs the convolution methods only");
    for (int m = 0; m < dim; ++m)
        if (size[m] < 1) return fail(FB_ERR_INVALID_ARGUMENT, "grid size must be positive");
    int rc;
    if ((rc = check_device_ptr(d_pts, "pts")) || (rc = check_device_ptr(d_val, "val")) ||
        (rc = check_device_ptr(d_out, "out")))
        return rc;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    void *ws = nullptr;
    if ((rc = noalloc_scratch(device, &ws))) return rc;
    const fb_field f = single_field(method, dim, n, sigma, x0, step, size, num_iter, max_dist_weight);
    fb_plan *P = nullptr;
    if ((rc = plan_create(&P, &f, 1, precision, ws, kNoAllocBytes))) return rc;
    const hipStream_t s = (hipStream_t)stream;
    rc = plan_run(P, d_pts, d_val, d_out, nullptr, s);
    // wait in any case: kernels queued before an error may still use the scratch array
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (!rc) rc = fail(FB_ERR_HIP, std::string("execution failed: ") + hipGetErrorString(e));
    }
    delete P;
    return rc;
}

int fb_inject(int dim, int64_t n, const double *pts, const double *val, const double *x0, const double *step,
              const int64_t *size, double *vg, double *wg, unsigned flags, void *stream)
{
    if (dim < 1 || dim > 3)
        return fail(FB_ERR_INVALID_ARGUMENT, "Barnes interpolation supports only sample points in dimensions 1, 2 or 3");
    if (n < 0 || (n > 0 && (!pts || !val)) || !x0 || !step || !size || !vg || !wg)
        return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument");
    int64_t cells = 1;
    bool empty = n == 0;
    for (int m = 0; m < dim; ++m) {
        if (size[m] < 1) return fail(FB_ERR_INVALID_ARGUMENT, "grid size must be positive");
        if (!finite_pos(step[m]) || !std::isfinite(x0[m]))
            return fail(FB_ERR_INVALID_ARGUMENT, "step must be finite and positive, x0 finite");
        cells *= size[m];
        empty = empty || size[m] < 2;   // no point lies below size-1 along that axis
    }
    if (empty) return FB_OK;
    const bool on_dev = flags & FB_FLAG_DEVICE_PTRS;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    hipStream_t s = stream ? (hipStream_t)stream : lib_stream(device);
    Scratch sc;
    sc.device = device;
    sc.on(s);
    int rc;
    const double *d_pts = pts, *d_val = val;
    double *d_vg = vg, *d_wg = wg;
    if (on_dev) {
        if ((rc = check_device_ptr(pts, "pts")) || (rc = check_device_ptr(val, "val")) ||
            (rc = check_device_ptr(vg, "vg")) || (rc = check_device_ptr(wg, "wg")))
            return rc;
    } else {
        double *a, *b;
        if ((rc = sc.get(&a, n * dim)) || (rc = sc.get(&b, n)) || (rc = sc.get(&d_vg, cells)) ||
            (rc = sc.get(&d_wg, cells)))
            return rc;
        FB_HIP_CHECK(hipMemcpyAsync(a, pts, n * dim * sizeof(double), hipMemcpyHostToDevice, s));
        FB_HIP_CHECK(hipMemcpyAsync(b, val, n * sizeof(double), hipMemcpyHostToDevice, s));
        FB_HIP_CHECK(hipMemcpyAsync(d_vg, vg, cells * sizeof(double), hipMemcpyHostToDevice, s));
        FB_HIP_CHECK(hipMemcpyAsync(d_wg, wg, cells * sizeof(double), hipMemcpyHostToDevice, s));
        d_pts = a;
        d_val = b;
    }
    // A one-field FP64 plan for its sort stage; its kernel parameters are unused (sigma =
    // step/10 and one iteration keep them valid for any grid of at least 2 nodes per axis).
    double sig[3];
    for (int m = 0; m < dim; ++m) sig[m] = 0.1 * step[m];
    const fb_field f = single_field(FB_METHOD_OPTIMIZED_CONVOLUTION, dim, n, sig, x0, step, size, 1, 0.0);
    fb_plan *P = nullptr;
    if ((rc = plan_create(&P, &f, 1, FB_PREC_FP64, nullptr, 0, false))) return rc;
    const uint32_t *vs = nullptr;
    rc = plan_sort(P, d_pts, d_val, s, &vs);
    // the values are injected as given: offset 0 (one field's grid starts at cell 0, x-fastest,
    // exactly the caller's layout)
    if (!rc && hipMemsetAsync(P->d_offsets, 0, sizeof(double), s) != hipSuccess) {
        (void)hipGetLastError();
        rc = fail(FB_ERR_HIP, "clearing the offset failed");
    }
    if (!rc)
        rc = launch_inject_records(P->d_fields, P->d_pt_prefix, 1, P->npoints, P->max_n <= kLocalSortMax,
                                   P->nxlines, vs, P->d_line_start, P->d_line_count, d_pts, d_val, P->d_offsets,
                                   P->d_irec, s);
    if (!rc)
        rc = launch_exact_inject(P->d_fields, P->d_xline_prefix, 1, P->nxlines, P->d_irec, P->d_line_start,
                                 P->d_line_count, d_vg, d_wg, P->dim, false, s);
    if (!rc && !on_dev) {
        if (hipMemcpyAsync(vg, d_vg, cells * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipMemcpyAsync(wg, d_wg, cells * sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess) {
            (void)hipGetLastError();
            rc = fail(FB_ERR_HIP, "copy of the grids failed");
        }
    }
    const hipError_t e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (!rc) rc = fail(FB_ERR_HIP, std::string("execution failed: ") + hipGetErrorString(e));
    }
    delete P;
    return rc;
}

int fb_barnes_s2(int method, int precision, int64_t n, const double *pts, const double *val, const double *sigma,
                 const double *x0, const double *step, const int64_t *size, int num_iter, double max_dist_weight,
                 int resample, void *out, unsigned flags, void *stream)
{
    if (n < 1 || !pts || !val || !sigma || !x0 || !step || !size || !out)
        return fail(FB_ERR_INVALID_ARGUMENT, "NULL argument or no sample points");
    if (method != FB_METHOD_OPTIMIZED_CONVOLUTION && method != FB_METHOD_NAIVE)
        return fail(FB_ERR_INVALID_ARGUMENT, "encountered invalid Barnes interpolation method");
    for (int m = 0; m < 2; ++m)
        if (size[m] < 1 || !(step[m] > 0.0) || !(sigma[m] > 0.0) || !std::isfinite(x0[m]))
            return fail(FB_ERR_INVALID_ARGUMENT, "grid size, step and sigma must be positive, x0 finite");
    const bool on_dev = flags & FB_FLAG_DEVICE_PTRS;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    hipStream_t s = stream ? (hipStream_t)stream : lib_stream(device);
    Scratch sc;
    sc.device = device;
    sc.on(s);
    int rc;
    const double *d_pts = pts, *d_val = val;
    if (on_dev) {
        if ((rc = check_device_ptr(pts, "pts")) || (rc = check_device_ptr(val, "val")) ||
            (rc = check_device_ptr(out, "out")))
            return rc;
    } else {
        double *a, *b;
        if ((rc = sc.get(&a, n * 2)) || (rc = sc.get(&b, n))) return rc;
        FB_HIP_CHECK(hipMemcpyAsync(a, pts, n * 2 * sizeof(double), hipMemcpyHostToDevice, s));
        FB_HIP_CHECK(hipMemcpyAsync(b, val, n * sizeof(double), hipMemcpyHostToDevice, s));
        d_pts = a;
        d_val = b;
    }
    // get_lambert_proj: create_proj(11.5, 34.5, 42.5, 65.5), lambert_conformal.py:50-97
    double proj[5];
    {
        const double rad = 3.141592653589793 / 180.0, h = rad / 2.0;
        const double lat1 = 42.5, lat2 = 65.5, clat = 34.5;
        const double nn = std::log(std::cos(lat1 * rad) / std::cos(lat2 * rad)) /
                          std::log(std::tan((90.0 + lat2) * h) / std::tan((90.0 + lat1) * h));
        const double F = std::cos(lat1 * rad) * std::pow(std::tan((90.0 + lat1) * h), nn) / nn;
        proj[0] = 11.5;
        proj[1] = nn;
        proj[2] = 1.0 / nn;
        proj[3] = F;
        proj[4] = F / std::pow(std::tan((90.0 + clat) * h), nn);
    }
    const int64_t cells = size[0] * size[1];
    if (method == FB_METHOD_NAIVE) {
        FieldDev F;
        std::memset(&F, 0, sizeof(F));
        F.dim = 2;
        F.n = n;
        FieldDev *d_f;
        double *d_off, *d_out = (double *)out;
        if ((rc = sc.get(&d_f, 1)) || (rc = sc.get(&d_off, 1))) return rc;
        if (!on_dev && (rc = sc.get(&d_out, cells))) return rc;
        FB_HIP_CHECK(hipMemcpy(d_f, &F, sizeof(F), hipMemcpyHostToDevice));
        double *d_part;
        if ((rc = sc.get(&d_part, offsets_part_elems(1, n)))) return rc;
        if ((rc = launch_offsets(d_f, 1, n, d_val, d_part, d_off, s))) return rc;
        if ((rc = launch_naive_s2(d_pts, d_val, d_off, n, sigma[0], x0, step, size, d_out, s))) return rc;
        if (!on_dev) FB_HIP_CHECK(hipMemcpyAsync(out, d_out, cells * sizeof(double), hipMemcpyDeviceToHost, s));
        FB_HIP_CHECK(hipStreamSynchronize(s));
        return FB_OK;
    }
    // the fixed Lambert grid, interpolationS2.py:186-188
    const double lam_x0[2] = {-32.0, -2.0};
    const int64_t lam_size[2] = {(int64_t)(64.0 / step[0]), (int64_t)(44.0 / step[1])};
    const int64_t lam_cells = lam_size[0] * lam_size[1];
    double *d_lpts;
    if ((rc = sc.get(&d_lpts, n * 2))) return rc;
    if ((rc = launch_lambert_points(d_pts, d_lpts, n, proj, s))) return rc;
    fb_field f;
    std::memset(&f, 0, sizeof(f));
    f.dim = 2;
    f.method = FB_METHOD_OPTIMIZED_CONVOLUTION;
    f.num_iter = num_iter;
    f.n = n;
    f.size[2] = 1;
    for (int m = 0; m < 2; ++m) {
        f.size[m] = lam_size[m];
        f.sigma[m] = sigma[m];
        f.x0[m] = lam_x0[m];
        f.step[m] = step[m];
    }
    f.max_dist_weight = max_dist_weight;
    fb_plan *P = nullptr;
    if ((rc = plan_create(&P, &f, 1, precision))) return rc;
    const int64_t out_cells = resample ? cells : lam_cells;
    float *d_out = (float *)out, *d_lam = nullptr;
    if (!on_dev) rc = sc.get(&d_out, out_cells);
    if (!rc) {
        if (resample) rc = sc.get(&d_lam, lam_cells);
        else d_lam = d_out;
    }
    if (!rc) {
        rc = plan_run(P, d_lpts, d_val, d_lam, nullptr, s);
        if (const int rc2 = P->mark_use(s)) rc = rc ? rc : rc2;
    }
    if (!rc && resample) rc = launch_resample(d_lam, lam_size, lam_x0, x0, step, size, proj, d_out, s);
    if (!rc && !on_dev) {
        hipError_t e = hipMemcpyAsync(out, d_out, out_cells * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) rc = fail(FB_ERR_HIP, std::string("copy of the result failed: ") + hipGetErrorString(e));
    }
    if (!rc) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) rc = fail(FB_ERR_HIP, std::string("execution failed: ") + hipGetErrorString(e));
    }
    delete P;
    return rc;
}

int fb_accumulate_lines(int tail, int precision, int64_t nlines, int64_t len, int rect_len, int num_iter, double alpha,
                        const double *in, double *out, unsigned flags, void *stream)
{
    if (nlines < 1 || len < 1 || rect_len < 1 || num_iter < 0 || !in || !out)
        return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    if (rect_len >= len) return fail(FB_ERR_INVALID_ARGUMENT, "rect_len must be smaller than the line length");
    const bool on_dev = flags & FB_FLAG_DEVICE_PTRS;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    hipStream_t s = stream ? (hipStream_t)stream : lib_stream(device);
    Scratch sc;
    sc.device = device;
    sc.on(s);
    int rc;
    const int64_t cells = nlines * len;
    double *d_io;
    if ((rc = sc.get(&d_io, cells))) return rc;
    FB_HIP_CHECK(hipMemcpyAsync(d_io, in, cells * sizeof(double),
                                on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    FieldDev F;
    std::memset(&F, 0, sizeof(F));
    F.dim = 2;
    F.tail = tail ? 1 : 0;
    F.num_iter = num_iter;
    F.size[0] = len;
    F.size[1] = nlines;
    F.size[2] = 1;
    F.T[0] = (rect_len - 1) / 2;
    F.alpha[0] = tail ? alpha : 0.0;
    F.pscale[0] = 1.0;
    FieldDev *d_f;
    if ((rc = sc.get(&d_f, 1))) return rc;
    FB_HIP_CHECK(hipMemcpy(d_f, &F, sizeof(F), hipMemcpyHostToDevice));
    if (precision == FB_PREC_FP64) {
        double *d_w, *d_vs, *d_ws;
        int64_t *d_pre;
        if ((rc = sc.get(&d_w, cells)) || (rc = sc.get(&d_vs, cells)) || (rc = sc.get(&d_ws, cells)) ||
            (rc = sc.get(&d_pre, 4)))
            return rc;
        if ((rc = launch_zero(d_w, cells * 8, s))) return rc;
        if ((rc = fast_pass_init())) return rc;   // k_exact_rows' LDS limit
        // line slots, then the x-line workgroups of k_exact_rows
        const int sl = exact_rows_slots(2 * nlines);
        const int64_t nwg = (2 * nlines + sl - 1) / sl;
        const int64_t pre[4] = {0, 2 * nlines, 0, nwg};
        FB_HIP_CHECK(hipMemcpy(d_pre, pre, sizeof(pre), hipMemcpyHostToDevice));
        rc = exact_rows_ring(rect_len)
                 ? launch_exact_rows(d_f, d_pre + 2, d_pre, 1, nwg, sl, rect_len, rect_len, d_io, d_w, d_vs, d_ws, s)
                 : launch_exact_lines(d_f, d_pre, 1, 2 * nlines, 0, rect_len, d_io, d_w, d_vs, d_ws, s);
        if (rc) return rc;
    } else {
        if (rect_len % 2 == 0) return fail(FB_ERR_UNSUPPORTED, "FP32C line passes need an odd rect_len");
        if ((rc = fast_pass_init())) return rc;
        if (num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
        float *d_v, *d_w;
        double *d_off;
        PassWork *d_work;
        PassField *d_pf;
        if ((rc = sc.get(&d_v, cells)) || (rc = sc.get(&d_w, cells)) || (rc = sc.get(&d_off, 1))) return rc;
        int lpl, lines, rpad;
        size_t lds;
        const int seg = model_seg(len, F.T[0]);
        if ((rc = pass_geometry(len, F.T[0], seg, kBlock, lpl, lines, rpad, lds, kBanksRows))) return rc;
        PassField pf{};
        pf.seg = seg;
        pf.len = len;
        pf.line_stride = len;
        pf.elem_stride = 1;
        pf.T = F.T[0];
        pf.lpl = lpl;
        pf.lines = lines;
        pf.rpad = rpad;
        pf.alpha = F.alpha[0];
        scale_groups(1.0, num_iter, 1.0, pf);
        std::vector<PassWork> work;
        for (int64_t l0 = 0; l0 < nlines; l0 += lines)
            work.push_back({0, (int32_t)std::min<int64_t>(lines, nlines - l0), l0 * len});
        if ((rc = sc.get(&d_work, work.size())) || (rc = sc.get(&d_pf, 1))) return rc;
        FB_HIP_CHECK(hipMemcpy(d_work, work.data(), work.size() * sizeof(PassWork), hipMemcpyHostToDevice));
        FB_HIP_CHECK(hipMemcpy(d_pf, &pf, sizeof(pf), hipMemcpyHostToDevice));
        if ((rc = launch_zero(d_off, 8, s))) return rc;
        if ((rc = launch_f64_to_f32(d_io, d_v, cells, s))) return rc;
        if ((rc = launch_zero(d_w, cells * 4, s))) return rc;
        if ((rc = launch_fast_pass(kPassSlide, false, seg, d_f, d_off, d_work, (int)work.size(), d_pf, 2 * kBlock, lds,
                                   nullptr, nullptr,
                                   nullptr, nullptr, nullptr, d_v, d_w, nullptr, nullptr, 0, nullptr, s)))
            return rc;
        if ((rc = launch_f32_to_f64(d_v, d_io, cells, 1.0, s))) return rc;
    }
    FB_HIP_CHECK(hipMemcpyAsync(out, d_io, cells * sizeof(double),
                                on_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    FB_HIP_CHECK(hipStreamSynchronize(s));
    return FB_OK;
}

int fb_convolve_grid(int tail, int precision, int dim, const int64_t *size, const int32_t *kernel_size, int num_iter,
                     const double *tail_value_, double thr, double *vg, double *wg, unsigned flags, void *stream)
{
    if (dim < 1 || dim > 3 || !size || !kernel_size || !vg || !wg || num_iter < 0)
        return fail(FB_ERR_INVALID_ARGUMENT, "bad arguments");
    const bool on_dev = flags & FB_FLAG_DEVICE_PTRS;
    int device = 0;
    FB_HIP_CHECK(hipGetDevice(&device));
    hipStream_t s = stream ? (hipStream_t)stream : lib_stream(device);
    FieldDev F;
    std::memset(&F, 0, sizeof(F));
    F.dim = dim;
    F.tail = tail ? 1 : 0;
    F.num_iter = num_iter;
    int64_t cells = 1;
    for (int m = 0; m < 3; ++m) F.size[m] = 1;
    for (int m = 0; m < dim; ++m) {
        if (size[m] < 1) return fail(FB_ERR_INVALID_ARGUMENT, "grid size must be positive");
        if (kernel_size[m] < 1 || kernel_size[m] % 2 == 0 || kernel_size[m] >= size[m])
            return fail(FB_ERR_INVALID_ARGUMENT, "kernel sizes must be odd and smaller than the grid");
        F.size[m] = size[m];
        F.T[m] = (kernel_size[m] - 1) / 2;
        F.alpha[m] = (tail && tail_value_) ? tail_value_[m] : 0.0;
        F.pscale[m] = 1.0;
        cells *= size[m];
    }
    F.thr = thr;
    F.thr_fast = thr;
    F.thr_fast_f = float_threshold(thr);
    Scratch sc;
    sc.device = device;
    sc.on(s);
    int rc;
    FieldDev *d_f;
    double *d_vg, *d_wg;
    int64_t *d_cpre;
    if ((rc = sc.get(&d_f, 1)) || (rc = sc.get(&d_vg, cells)) || (rc = sc.get(&d_wg, cells)) ||
        (rc = sc.get(&d_cpre, 2)))
        return rc;
    const hipMemcpyKind kin = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    FB_HIP_CHECK(hipMemcpy(d_f, &F, sizeof(F), hipMemcpyHostToDevice));
    FB_HIP_CHECK(hipMemcpyAsync(d_vg, vg, cells * 8, kin, s));
    FB_HIP_CHECK(hipMemcpyAsync(d_wg, wg, cells * 8, kin, s));
    const int64_t cpre[2] = {0, cells};
    FB_HIP_CHECK(hipMemcpy(d_cpre, cpre, sizeof(cpre), hipMemcpyHostToDevice));
    const int64_t W = F.size[0], H = F.size[1], D = F.size[2];
    const int64_t nl[3] = {H * D, W * D, W * H};
    if (precision == FB_PREC_FP64) {
        double *d_vs, *d_ws;
        int64_t *d_pre;
        if ((rc = sc.get(&d_vs, cells)) || (rc = sc.get(&d_ws, cells)) || (rc = sc.get(&d_pre, 4))) return rc;
        if ((rc = fast_pass_init())) return rc;   // k_exact_rows' LDS limit
        for (int a = 0; a < dim; ++a) {
            // line slots, then (x lines) the workgroups of k_exact_rows
            const int sl = exact_rows_slots(2 * nl[a]);
            const int64_t nwg = (2 * nl[a] + sl - 1) / sl;
            const int64_t pre[4] = {0, 2 * nl[a], 0, nwg};
            const int64_t rect = 2 * (int64_t)F.T[0] + 1;
            FB_HIP_CHECK(hipMemcpy(d_pre, pre, sizeof(pre), hipMemcpyHostToDevice));
            rc = a == 0 && exact_rows_ring(rect)
                     ? launch_exact_rows(d_f, d_pre + 2, d_pre, 1, nwg, sl, -1, rect, d_vg, d_wg, d_vs, d_ws, s)
                     : launch_exact_lines(d_f, d_pre, 1, 2 * nl[a], a, -1, d_vg, d_wg, d_vs, d_ws, s);
            if (rc) return rc;
            FB_HIP_CHECK(hipStreamSynchronize(s));   // d_pre is reused by the next axis
        }
        if ((rc = launch_mask(d_f, d_cpre, 1, cells, d_wg, 0, s))) return rc;
    } else {
        if (num_iter < 1) return fail(FB_ERR_INVALID_ARGUMENT, "num_iter must be >= 1");
        if ((rc = fast_pass_init())) return rc;
        float *d_v, *d_w;
        double *d_off;
        if ((rc = sc.get(&d_v, cells)) || (rc = sc.get(&d_w, cells)) || (rc = sc.get(&d_off, 1))) return rc;
        if ((rc = launch_zero(d_off, 8, s))) return rc;
        if ((rc = launch_f64_to_f32(d_vg, d_v, cells, s))) return rc;
        if ((rc = launch_f64_to_f32(d_wg, d_w, cells, s))) return rc;
        fb_plan tmp;
        tmp.device = device;
        tmp.nfields = 1;
        tmp.dim = dim;
        tmp.inplace_passes = true;
        tmp.fh.assign(1, F);
        for (int a = 0; a < dim; ++a) {
            std::vector<PassWork> work;
            std::vector<PassField> pf;
            std::vector<PassGroup> groups;
            if ((rc = build_pass(&tmp, a, work, pf, groups))) return rc;
            PassWork *d_work;
            PassField *d_pf;
            if ((rc = sc.get(&d_work, work.size())) || (rc = sc.get(&d_pf, 1))) return rc;
            FB_HIP_CHECK(hipMemcpy(d_work, work.data(), work.size() * sizeof(PassWork), hipMemcpyHostToDevice));
            FB_HIP_CHECK(hipMemcpy(d_pf, pf.data(), sizeof(PassField), hipMemcpyHostToDevice));
            for (const PassGroup &g : groups)
                if ((rc = launch_fast_pass(g.kind, false, g.seg, d_f, d_off, d_work + g.first, g.count, d_pf, 2 * g.lpf,
                                           g.lds, nullptr, nullptr, nullptr, nullptr, nullptr, d_v, d_w,
                                           nullptr, nullptr, 0, nullptr, s)))
                    return rc;
            FB_HIP_CHECK(hipStreamSynchronize(s));   // host vectors above go out of scope
        }
        if ((rc = launch_f32_to_f64(d_v, d_vg, cells, 1.0, s))) return rc;
        if ((rc = launch_f32_to_f64(d_w, d_wg, cells, 1.0, s))) return rc;
        if ((rc = launch_mask(d_f, d_cpre, 1, cells, d_wg, 1, s))) return rc;
    }
    const hipMemcpyKind kout = on_dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
    FB_HIP_CHECK(hipMemcpyAsync(vg, d_vg, cells * 8, kout, s));
    FB_HIP_CHECK(hipMemcpyAsync(wg, d_wg, cells * 8, kout, s));
    FB_HIP_CHECK(hipStreamSynchronize(s));
    return FB_OK;
}

}  // extern "C"
