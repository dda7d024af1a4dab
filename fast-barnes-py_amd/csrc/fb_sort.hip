// fb_sort.hip -- exclusive scan, the per-field sort in LDS (fields of at most kLocalSortMax
// points) and the row-bucket sort (larger fields).
//
// Used to order the injection's points by base cell while keeping, inside a cell, the point
// order of the input (the order in which the reference's _inject_data_*d loop adds them,
// interpolation.py:246-270): both sorts order the composite (cell key, point index), so the
// result does not depend on timing.
#include <mutex>

#include "fb_internal.h"
#include "fb_launch.h"

namespace fb {

namespace {

constexpr int kThreads = 256;
constexpr int kItems = 16;
constexpr int kTile = kThreads * kItems;   // 4096 elements per workgroup of the multi-level scan
constexpr int kLocalSortThreads = 1024;

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- scan ----------------------------------------------------------------

__global__ __launch_bounds__(kThreads) void k_scan_tiles(const uint32_t *__restrict__ in,
                                                         uint32_t *__restrict__ out,
                                                         uint32_t *__restrict__ tile_sums, int64_t n)
{
    __shared__ uint32_t s_tot[kThreads];
    const int tid = threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x * kTile + (int64_t)tid * kItems;
    uint32_t v[kItems];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        sum += v[i];
    }
    s_tot[tid] = sum;
    __syncthreads();
    for (int off = 1; off < kThreads; off <<= 1) {
        const uint32_t t = tid >= off ? s_tot[tid - off] : 0u;
        __syncthreads();
        s_tot[tid] += t;
        __syncthreads();
    }
    uint32_t run = tid ? s_tot[tid - 1] : 0u;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
        const int64_t idx = base + i;
        if (idx < n) out[idx] = run;
        run += v[i];
    }
    if (tid == kThreads - 1 && tile_sums) tile_sums[blockIdx.x] = s_tot[kThreads - 1];
}

__global__ __launch_bounds__(kThreads) void k_scan_add(uint32_t *__restrict__ out,
                                                       const uint32_t *__restrict__ tile_offs, int64_t n)
{
    const int64_t idx = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (idx < n) out[idx] += tile_offs[idx / kTile];
}

// Bitonic network over sk[0, np2) (np2 a power of two) in LDS by the whole workgroup, ascending.
// Pair t of a stage is (i, i + j) with i = t with a zero bit inserted at bit log2(j).  A thread
// takes pairs t = tid + m*blockDim.x, so for j <= 64 wave w touches only the 128-element blocks
// w + (waves per workgroup)*m: consecutive stages that both have j <= 64 depend on the wave's own
// program order only; a barrier follows every stage where this stage or the next one has j > 64.
// Ends with a workgroup barrier.
__device__ __forceinline__ void block_bitonic(unsigned long long *sk, int np2)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 2; k <= np2; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = tid; t < (np2 >> 1); t += nt) {
                const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                const int l = i + j;
                const unsigned long long a = sk[i], b = sk[l];
                const bool up = (i & k) == 0;
                if ((a > b) == up) {
                    sk[i] = b;
                    sk[l] = a;
                }
            }
            const int jn = j > 1 ? (j >> 1) : k;   // the next stage's distance
            if (j > 64 || jn > 64) __syncthreads();
            else __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        }
    }
    __syncthreads();
}

// One workgroup per field: the field's points inside the grid written in (cell key, point
// index) order (equal keys stay in point order, as in the row-bucket sort) into the field's
// range of the list, which begins at the field's first point index (pt_prefix[f]; the range
// holds n slots, of which the in-grid points fill the first).  Keys order by row first (key =
// cell_off + row*W + x), and k_point_keys has counted each row's points: the workgroup scans the
// field's row counts into the rows' starts (written to row_start, absolute).  When the field
// has at most kFieldRows rows and none holds more than kRankMax points, every point goes to
// its row's start plus its rank among the row's points: the points are dropped into their
// rows' ranges in LDS (atomic cursors, any order), then each counts the row's entries below
// its own composite.  Otherwise a bitonic network sorts the 64-bit composites in LDS.  Both
// give the same order.  Fields of at most kLocalSortMax points.
constexpr int kFieldRows = 4096;                                    // LDS row starts and cursors
constexpr int kRankMax = 64;                                        // largest row placed by rank
constexpr int kSortPerThread = kLocalSortMax / kLocalSortThreads;   // points per thread
constexpr int kSortWaves = kLocalSortThreads / 64;

__global__ __launch_bounds__(kLocalSortThreads) void k_field_sort(const FieldDev *__restrict__ fields,
                                                                  const int64_t *__restrict__ pt_prefix,
                                                                  const uint32_t *__restrict__ keys,
                                                                  const uint32_t *__restrict__ row_count,
                                                                  uint32_t *__restrict__ row_start,
                                                                  uint32_t *__restrict__ out_vals)
{
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ uint32_t cur[kFieldRows];
    __shared__ uint32_t rst[kFieldRows];
    __shared__ uint32_t wsum[kSortWaves];
    unsigned long long *sk = reinterpret_cast<unsigned long long *>(smem);
    const int f = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const FieldDev &F = fields[f];
    const int64_t p0 = pt_prefix[f];
    const int n = (int)(pt_prefix[f + 1] - p0);
    const int64_t r0 = F.xline_off;
    const int64_t nr = F.size[1] * F.size[2];   // x-lines (rows) of the field
    const uint32_t dst = (uint32_t)p0;

    // exclusive scan of the row counts (chunks of one row per thread) and the largest row
    uint32_t carry = 0, m = 0;
    for (int64_t c0 = 0; c0 < nr; c0 += kLocalSortThreads) {
        const int64_t r = c0 + tid;
        const uint32_t v = r < nr ? row_count[r0 + r] : 0u;
        m = max(m, v);
        uint32_t inc = v;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d);
            if (lane >= d) inc += o;
        }
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kSortWaves; ++w) {
            const uint32_t x = wsum[w];
            before += w < wv ? x : 0u;
            total += x;
        }
        const uint32_t ex = carry + before + inc - v;
        if (r < nr) {
            row_start[r0 + r] = dst + ex;
            if (r < kFieldRows) rst[r] = ex;
        }
        carry += total;
        __syncthreads();   // wsum is rewritten by the next chunk
    }
    for (int d = 32; d >= 1; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d));
    if (lane == 0) wsum[wv] = m;
    __syncthreads();
    uint32_t mx = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) mx = max(mx, wsum[w]);

    if (nr <= kFieldRows && mx <= (uint32_t)kRankMax) {
        for (int r = tid; r < nr; r += kLocalSortThreads) cur[r] = 0u;
        __syncthreads();
        const uint32_t cbase = (uint32_t)F.cell_off, W = (uint32_t)F.size[0];
        unsigned long long ev[kSortPerThread];
        uint32_t rs[kSortPerThread], rc[kSortPerThread];
#pragma unroll
        for (int k = 0; k < kSortPerThread; ++k) {
            const int i = tid + k * kLocalSortThreads;
            rc[k] = 0u;
            rs[k] = 0u;
            ev[k] = 0ull;
            if (i < n) {
                const uint32_t key = keys[p0 + i];
                if (key != kInvalidKey) {
                    const uint32_t r = (key - cbase) / W;
                    rs[k] = rst[r];   // the row's range in the field's list
                    rc[k] = row_count[r0 + r];
                    ev[k] = ((unsigned long long)key << 32) | (unsigned long long)(uint32_t)(p0 + i);
                    sk[rs[k] + atomicAdd(&cur[r], 1u)] = ev[k];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kSortPerThread; ++k) {
            if (rc[k] == 0u) continue;
            uint32_t rank = 0;
            for (uint32_t j = 0; j < rc[k]; ++j) rank += sk[rs[k] + j] < ev[k] ? 1u : 0u;
            out_vals[dst + rs[k] + rank] = (uint32_t)ev[k];
        }
        return;
    }

    int np2 = 1;
    while (np2 < n) np2 <<= 1;
    for (int i = tid; i < np2; i += kLocalSortThreads)
        sk[i] = i < n ? ((unsigned long long)keys[p0 + i] << 32) | (unsigned long long)(uint32_t)(p0 + i) : ~0ull;
    __syncthreads();
    block_bitonic(sk, np2);
    for (int i = tid; i < n; i += kLocalSortThreads) {
        const unsigned long long e = sk[i];
        if ((uint32_t)(e >> 32) != kInvalidKey) out_vals[dst + i] = (uint32_t)e;
    }
}

// ---- row buckets (fields of more than kLocalSortMax points) ------------------
// The order of a stable sort of the (cell key, point index) pairs: every in-grid point
// has a slot in its base row (k_point_keys' atomic count), the rows start at the exclusive scan
// of the counts, and each row's entries are then sorted by the composite (key, point index).

constexpr int kRowSmall = 16;           // rows of at most this many entries: one thread, in registers
constexpr int kRowWave = 1024;          // rows of at most this many entries: one wave, in LDS ...
constexpr int kRankWave = 256;          // ... by rank up to this many, by a bitonic network above
constexpr int kRowSortThreads = 256;    // workgroup of the wave kernel (4 rows at a time) and of larger rows

__device__ __forceinline__ unsigned long long row_entry(const uint32_t *keys, const uint32_t *vals, uint32_t i)
{
    return ((unsigned long long)keys[i] << 32) | (unsigned long long)vals[i];
}

__global__ __launch_bounds__(kThreads) void k_row_scatter(int64_t n, const uint32_t *__restrict__ keys,
                                                          const uint32_t *__restrict__ slot,
                                                          const uint32_t *__restrict__ rowg,
                                                          const uint32_t *__restrict__ row_start,
                                                          uint32_t *__restrict__ skeys, uint32_t *__restrict__ svals)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t r = rowg[i];
    if (r == kInvalidKey) return;   // outside its grid
    const uint32_t dst = row_start[r] + slot[i];
    skeys[dst] = keys[i];
    svals[dst] = (uint32_t)i;
}

// One thread per row: rows of 2..kRowSmall entries sorted by a bitonic network in registers
// (padded with the largest composite); rows of up to kRowWave entries are appended to
// lists[0] (count counts[0]), larger ones to lists[1] (counts[1]).
__global__ __launch_bounds__(kThreads) void k_row_sort_small(int64_t nrows, const uint32_t *__restrict__ row_start,
                                                             const uint32_t *__restrict__ row_count,
                                                             uint32_t *__restrict__ skeys,
                                                             uint32_t *__restrict__ svals,
                                                             uint32_t *__restrict__ lists,
                                                             uint32_t *__restrict__ counts)
{
    const int64_t r = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const uint32_t n = r < nrows ? row_count[r] : 0;
    // list appends aggregated per wave: one atomic per wave and list (a single counter taking
    // one atomic per row serialises when most rows are large)
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const bool mine = c == 0 ? (n > (uint32_t)kRowSmall && n <= (uint32_t)kRowWave) : n > (uint32_t)kRowWave;
        const unsigned long long m = __ballot(mine);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(&counts[c], (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (mine) lists[c * nrows + base + (uint32_t)__popcll(m & below)] = (uint32_t)r;
        }
    }
    if (n <= 1 || n > (uint32_t)kRowSmall) return;
    const uint32_t s0 = row_start[r];
    unsigned long long e[kRowSmall];
#pragma unroll
    for (int q = 0; q < kRowSmall; ++q) e[q] = (uint32_t)q < n ? row_entry(skeys, svals, s0 + q) : ~0ull;
#pragma unroll
    for (int k = 2; k <= kRowSmall; k <<= 1)
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
            for (int i = 0; i < kRowSmall; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    const bool up = (i & k) == 0;
                    const unsigned long long a = e[i], b = e[l];
                    if ((a > b) == up) {
                        e[i] = b;
                        e[l] = a;
                    }
                }
            }
#pragma unroll
    for (int q = 0; q < kRowSmall; ++q)
        if ((uint32_t)q < n) {
            skeys[s0 + q] = (uint32_t)(e[q] >> 32);
            svals[s0 + q] = (uint32_t)e[q];
        }
}

// One wave sorts the row of n <= 64*NU entries at s0 by rank: each entry's slot is the number of
// the row's composites below it (all differ), counted against the row's copy in LDS `sk`, two
// composites per read (every lane reads the same pair at a time: broadcast reads; an odd row
// is padded with ~0, which no composite exceeds).  NU = ceil(n/64) keeps the compares to the
// registers that hold entries.
template <int NU>
__device__ __forceinline__ void rank_row(uint32_t *__restrict__ skeys, uint32_t *__restrict__ svals, uint32_t s0,
                                         uint32_t n, unsigned long long *__restrict__ sk, int lane)
{
    unsigned long long e[NU];
    uint32_t rk[NU];
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        const uint32_t i = lane + 64 * u;
        e[u] = i < n ? row_entry(skeys, svals, s0 + i) : ~0ull;
        if (i < n) sk[i] = e[u];
        rk[u] = 0u;
    }
    if (lane == 0 && (n & 1u)) sk[n] = ~0ull;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    const ulonglong2 *sk2 = reinterpret_cast<const ulonglong2 *>(sk);
    for (uint32_t j = 0; j < (n + 1) >> 1; ++j) {
        const ulonglong2 x = sk2[j];
#pragma unroll
        for (int u = 0; u < NU; ++u) rk[u] += (x.x < e[u] ? 1u : 0u) + (x.y < e[u] ? 1u : 0u);
    }
#pragma unroll
    for (int u = 0; u < NU; ++u) {
        if (lane + 64 * u < (int)n) {
            skeys[s0 + rk[u]] = (uint32_t)(e[u] >> 32);
            svals[s0 + rk[u]] = (uint32_t)e[u];
        }
    }
}

// Waves loop over the rows in list (at most kRowWave entries each): each wave sorts one row at a
// time in its own kRowWave-entry slice of LDS, by rank (rows of at most kRankWave entries) or by
// the bitonic network.  Within one wave every step only needs its LDS operations complete (a
// wavefront fence), never a workgroup barrier.
__global__ __launch_bounds__(kRowSortThreads) void k_row_sort_wave(const uint32_t *__restrict__ row_start,
                                                                   const uint32_t *__restrict__ row_count,
                                                                   uint32_t *__restrict__ skeys,
                                                                   uint32_t *__restrict__ svals,
                                                                   const uint32_t *__restrict__ list,
                                                                   const uint32_t *__restrict__ count)
{
    __shared__ __align__(16) unsigned long long lds[(kRowSortThreads / 64) * kRowWave];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    unsigned long long *sk = lds + wv * kRowWave;
    const uint32_t nb = *count;
    const uint32_t nwaves = gridDim.x * (kRowSortThreads / 64);
    for (uint32_t b = blockIdx.x * (kRowSortThreads / 64) + wv; b < nb; b += nwaves) {
        const uint32_t r = list[b];
        const uint32_t n = row_count[r];
        const uint32_t s0 = row_start[r];
        if (n <= (uint32_t)kRankWave) {
            // rows of at most kRankWave entries (more than kRowSmall): by rank
            switch ((n + 63) >> 6) {
            case 1: rank_row<1>(skeys, svals, s0, n, sk, lane); break;
            case 2: rank_row<2>(skeys, svals, s0, n, sk, lane); break;
            case 3: rank_row<3>(skeys, svals, s0, n, sk, lane); break;
            default: rank_row<kRankWave / 64>(skeys, svals, s0, n, sk, lane); break;
            }
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // before the next row reuses sk
            continue;
        }
        int np2 = 1;
        while ((uint32_t)np2 < n) np2 <<= 1;
        for (int i = lane; i < np2; i += 64) sk[i] = (uint32_t)i < n ? row_entry(skeys, svals, s0 + i) : ~0ull;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (int k = 2; k <= np2; k <<= 1)
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int t = lane; t < (np2 >> 1); t += 64) {
                    const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                    const int l = i + j;
                    const unsigned long long a = sk[i], c = sk[l];
                    const bool up = (i & k) == 0;
                    if ((a > c) == up) {
                        sk[i] = c;
                        sk[l] = a;
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
            }
        for (int i = lane; (uint32_t)i < n; i += 64) {
            const unsigned long long e = sk[i];
            skeys[s0 + i] = (uint32_t)(e >> 32);
            svals[s0 + i] = (uint32_t)e;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // before the next row reuses sk
    }
}

// Workgroups loop over the rows in big_list (more than kRowWave entries): up to kLocalSortMax
// entries by the bitonic network in LDS, more by rank (each entry counts the row's smaller composites; all composites differ)
// through the scratch arrays tkeys/tvals.
__global__ __launch_bounds__(kRowSortThreads) void k_row_sort_big(const uint32_t *__restrict__ row_start,
                                                                  const uint32_t *__restrict__ row_count,
                                                                  uint32_t *__restrict__ skeys,
                                                                  uint32_t *__restrict__ svals,
                                                                  const uint32_t *__restrict__ big_list,
                                                                  const uint32_t *__restrict__ big_count,
                                                                  uint32_t *__restrict__ tkeys,
                                                                  uint32_t *__restrict__ tvals)
{
    extern __shared__ __align__(16) unsigned char smem[];
    unsigned long long *sk = reinterpret_cast<unsigned long long *>(smem);
    const int tid = threadIdx.x, nt = blockDim.x;
    const uint32_t nb = *big_count;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint32_t r = big_list[b];
        const uint32_t n = row_count[r];
        const uint32_t s0 = row_start[r];
        if (n <= (uint32_t)kLocalSortMax) {
            int np2 = 1;
            while ((uint32_t)np2 < n) np2 <<= 1;
            for (int i = tid; i < np2; i += nt) sk[i] = (uint32_t)i < n ? row_entry(skeys, svals, s0 + i) : ~0ull;
            __syncthreads();
            block_bitonic(sk, np2);
            for (int i = tid; (uint32_t)i < n; i += nt) {
                const unsigned long long e = sk[i];
                skeys[s0 + i] = (uint32_t)(e >> 32);
                svals[s0 + i] = (uint32_t)e;
            }
        } else {
            for (uint32_t i = tid; i < n; i += nt) {
                const unsigned long long ei = row_entry(skeys, svals, s0 + i);
                uint32_t rank = 0;
                for (uint32_t j = 0; j < n; ++j) rank += row_entry(skeys, svals, s0 + j) < ei;
                tkeys[s0 + rank] = (uint32_t)(ei >> 32);
                tvals[s0 + rank] = (uint32_t)ei;
            }
            __syncthreads();
            for (uint32_t i = tid; i < n; i += nt) {
                skeys[s0 + i] = tkeys[s0 + i];
                svals[s0 + i] = tvals[s0 + i];
            }
        }
        __syncthreads();   // the next row reuses sk / reads what this one wrote
    }
}

}  // namespace

int64_t scan_temp_elems(int64_t n)
{
    int64_t total = 0;
    while (n > kTile) {
        n = cdiv(n, kTile);
        total += n;
    }
    return total + 1;
}

int scan_exclusive(const uint32_t *in, uint32_t *out, int64_t n, uint32_t *tmp, hipStream_t stream)
{
    if (n <= 0) return FB_OK;
    const int64_t ntiles = cdiv(n, kTile);
    if (ntiles == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, stream, in, out, (uint32_t *)nullptr, n);
        FB_HIP_CHECK(hipGetLastError());
        return FB_OK;
    }
    uint32_t *sums = tmp;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(kThreads), 0, stream, in, out, sums, n);
    FB_HIP_CHECK(hipGetLastError());
    int rc = scan_exclusive(sums, sums, ntiles, tmp + ntiles, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)cdiv(n, kThreads)), dim3(kThreads), 0, stream, out,
                       (const uint32_t *)sums, n);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int local_sort_max() { return kLocalSortMax; }

int row_bucket_sort(const uint32_t *d_keys, const uint32_t *d_slot, const uint32_t *d_rowg, int64_t npoints,
                    const uint32_t *d_row_start, const uint32_t *d_row_count, int64_t nrows, uint32_t *d_skeys,
                    uint32_t *d_svals, uint32_t *d_big_list, uint32_t *d_big_count, uint32_t *d_tkeys,
                    uint32_t *d_tvals, hipStream_t stream)
{
    if (npoints <= 0 || nrows <= 0) return FB_OK;
    hipLaunchKernelGGL(k_row_scatter, dim3((unsigned)cdiv(npoints, kThreads)), dim3(kThreads), 0, stream, npoints,
                       d_keys, d_slot, d_rowg, d_row_start, d_skeys, d_svals);
    FB_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_row_sort_small, dim3((unsigned)cdiv(nrows, kThreads)), dim3(kThreads), 0, stream, nrows,
                       d_row_start, d_row_count, d_skeys, d_svals, d_big_list, d_big_count);
    FB_HIP_CHECK(hipGetLastError());
    const int64_t wgrid = std::min<int64_t>(cdiv(nrows, kRowSortThreads / 64), 2048);
    hipLaunchKernelGGL(k_row_sort_wave, dim3((unsigned)wgrid), dim3(kRowSortThreads), 0, stream, d_row_start,
                       d_row_count, d_skeys, d_svals, (const uint32_t *)d_big_list, (const uint32_t *)d_big_count);
    FB_HIP_CHECK(hipGetLastError());
    const int64_t grid = std::min<int64_t>(nrows, 1024);
    hipLaunchKernelGGL(k_row_sort_big, dim3((unsigned)grid), dim3(kRowSortThreads),
                       (size_t)kLocalSortMax * sizeof(unsigned long long), stream, d_row_start, d_row_count, d_skeys,
                       d_svals, (const uint32_t *)(d_big_list + nrows), (const uint32_t *)(d_big_count + 1), d_tkeys,
                       d_tvals);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

int field_sort_init()
{
    static std::once_flag once;
    static hipError_t err = hipSuccess;
    std::call_once(once, [] {
        err = hipFuncSetAttribute((const void *)k_field_sort, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  kLocalSortMax * (int)sizeof(unsigned long long));
        if (err == hipSuccess)
            err = hipFuncSetAttribute((const void *)k_row_sort_big, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      kLocalSortMax * (int)sizeof(unsigned long long));
    });
    if (err != hipSuccess) {
        (void)hipGetLastError();
        set_error(std::string("hipFuncSetAttribute failed: ") + hipGetErrorString(err));
        return FB_ERR_HIP;
    }
    return FB_OK;
}

int launch_field_sort(const FieldDev *d_fields, const int64_t *d_pt_prefix, int nfields, int max_n,
                      const uint32_t *d_keys, const uint32_t *d_row_count, uint32_t *d_row_start,
                      uint32_t *d_out_vals, hipStream_t stream)
{
    if (nfields <= 0) return FB_OK;
    if (max_n > kLocalSortMax) {
        set_error("field sort: a field has more than " + std::to_string(kLocalSortMax) + " points");
        return FB_ERR_UNSUPPORTED;
    }
    int np2 = 1;
    while (np2 < max_n) np2 <<= 1;
    hipLaunchKernelGGL(k_field_sort, dim3(nfields), dim3(kLocalSortThreads), (size_t)np2 * 8, stream, d_fields,
                       d_pt_prefix, d_keys, d_row_count, d_row_start, d_out_vals);
    FB_HIP_CHECK(hipGetLastError());
    return FB_OK;
}

}  // namespace fb
