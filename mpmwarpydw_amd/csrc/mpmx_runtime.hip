// libmpmx runtime pieces: error text, debug synchronisation, fills and copies,
// a three-pass exclusive scan and the counting sort by key that both the P2G
// binning and the XPBD hash grid are built on.
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <vector>

#include "mpmx_device.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

static thread_local char g_err[512] = "";
// tiles dealt to the XCDs in runs of 16, split constitutive update on, G2P per tile
static int g_tuning[MPMX_TUNE_COUNT] = {2, 1, 1, 0};

int tuning(int knob) { return (knob >= 0 && knob < MPMX_TUNE_COUNT) ? g_tuning[knob] : 0; }

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

static int debug_sync_enabled() {
    static int v = -1;
    if (v < 0) {
        const char* e = getenv("MPMX_DEBUG");
        v = (e && e[0] == '1') ? 1 : 0;
    }
    return v;
}

// ---------------------------------------------------------------------------
// launch profiler: when on, every kernel launched through MPMX_LAUNCH carries
// a start/stop event pair stamped by its own dispatch (hipExtLaunchKernelGGL);
// check_launch names the pair.  Pairs sit at (2k, 2k + 1) in the pool.
// ---------------------------------------------------------------------------
struct Profiler {
    bool on = false;
    std::vector<hipEvent_t> pool;
    std::vector<const char*> names;  // name of the kernel whose stop event this is, else nullptr
    size_t used = 0;
    long pending = -1;  // stop-event index of the last pair not yet named
};
static Profiler g_prof;

bool prof_pair(hipStream_t s, hipEvent_t* start, hipEvent_t* stop) {
    if (!g_prof.on || g_prof.used + 2 > g_prof.pool.size()) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
    *start = g_prof.pool[g_prof.used];
    *stop = g_prof.pool[g_prof.used + 1];
    g_prof.names[g_prof.used] = nullptr;
    g_prof.names[g_prof.used + 1] = nullptr;
    g_prof.pending = (long)g_prof.used + 1;
    g_prof.used += 2;
    return true;
}

static void prof_record(const char* what, hipStream_t) {
    if (!g_prof.on || g_prof.pending < 0) return;
    g_prof.names[g_prof.pending] = what;
    g_prof.pending = -1;
}

int check_launch(const char* what, hipStream_t s) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return (int)e;
    }
    prof_record(what, s);
    if (debug_sync_enabled()) {
        e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            set_error("%s failed on the device: %s", what, hipGetErrorString(e));
            return (int)e;
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// fills / copies
// ---------------------------------------------------------------------------
template <typename T>
__global__ void fill_kernel(T* __restrict__ dst, long n, T value) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = value;
}

__global__ void copy_kernel(float* __restrict__ dst, const float* __restrict__ src, long n) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

static unsigned fill_grid(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

// ---------------------------------------------------------------------------
// exclusive scan: tiles of 2048 ints per 256-thread block
// ---------------------------------------------------------------------------
constexpr int kScanPer = 8;
constexpr int kScanTile = kBlock * kScanPer;
static_assert(kScanTile == kScanChunk, "scan chunk size");

__device__ int block_exclusive_scan(int v, int* sh, int* total) {
    // Hillis-Steele inclusive scan over the 256 thread values in LDS
    int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        int add = (t >= off) ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    int incl = sh[t];
    *total = sh[kBlock - 1];
    __syncthreads();
    return incl - v;
}

__global__ void scan_reduce_kernel(const int32_t* __restrict__ in, long n, int32_t* __restrict__ partials) {
    __shared__ int sh[kBlock];
    long base = (long)blockIdx.x * kScanTile;
    int s = 0;
    for (int k = 0; k < kScanPer; k++) {
        long i = base + (long)k * kBlock + threadIdx.x;
        if (i < n) s += in[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// one block: exclusive scan of partials[0..nb) in place, total in partials[nb]
__global__ void scan_partials_kernel(int32_t* __restrict__ partials, long nb) {
    __shared__ int sh[kBlock];
    int carry = 0;
    for (long base = 0; base < nb; base += kBlock) {
        long i = base + threadIdx.x;
        int v = (i < nb) ? partials[i] : 0;
        int tot;
        int ex = block_exclusive_scan(v, sh, &tot);
        if (i < nb) partials[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) partials[nb] = carry;
}

__global__ void scan_apply_kernel(const int32_t* __restrict__ in, long n, const int32_t* __restrict__ partials,
                                  int32_t* __restrict__ out, long nb) {
    __shared__ int sh[kBlock];
    __shared__ int vals[kScanTile];
    long base = (long)blockIdx.x * kScanTile;
    for (int k = 0; k < kScanPer; k++) {
        long i = base + (long)k * kBlock + threadIdx.x;
        vals[k * kBlock + threadIdx.x] = (i < n) ? in[i] : 0;
    }
    __syncthreads();
    int local = 0;
    for (int k = 0; k < kScanPer; k++) local += vals[threadIdx.x * kScanPer + k];
    int tot;
    int ex = block_exclusive_scan(local, sh, &tot) + partials[blockIdx.x];
    for (int k = 0; k < kScanPer; k++) {
        long i = base + (long)threadIdx.x * kScanPer + k;
        if (i < n) out[i] = ex;
        ex += vals[threadIdx.x * kScanPer + k];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = partials[nb];
}

// scan_apply when partials[] holds the plain total of every chunk: each block
// first sums the totals of the chunks before it (in a fixed order), and block 0
// also writes the grand total to out[n]
__device__ __forceinline__ void scan_apply_chunks_tile(const int32_t* __restrict__ in, long n,
                                                       const int32_t* __restrict__ totals, int32_t* __restrict__ out,
                                                       long nb) {
    __shared__ int sh[kBlock];
    __shared__ int vals[kScanTile];
    const long lim = blockIdx.x == 0 ? nb : (long)blockIdx.x;
    int part = 0;
    for (long c = threadIdx.x; c < lim; c += kBlock) part += totals[c];
    sh[threadIdx.x] = part;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
        __syncthreads();
    }
    const int before = blockIdx.x == 0 ? 0 : sh[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = sh[0];
    __syncthreads();
    const long base = (long)blockIdx.x * kScanTile;
    for (int k = 0; k < kScanPer; k++) {
        long i = base + (long)k * kBlock + threadIdx.x;
        vals[k * kBlock + threadIdx.x] = (i < n) ? in[i] : 0;
    }
    __syncthreads();
    int local = 0;
    for (int k = 0; k < kScanPer; k++) local += vals[threadIdx.x * kScanPer + k];
    int tot;
    int ex = block_exclusive_scan(local, sh, &tot) + before;
    for (int k = 0; k < kScanPer; k++) {
        long i = base + (long)threadIdx.x * kScanPer + k;
        if (i < n) out[i] = ex;
        ex += vals[threadIdx.x * kScanPer + k];
    }
}

__global__ void scan_apply_chunks_kernel(const int32_t* __restrict__ in, long n, const int32_t* __restrict__ totals,
                                         int32_t* __restrict__ out, long nb) {
    scan_apply_chunks_tile(in, n, totals, out, nb);
}

static_assert(kMaxSortSegs == 2, "the table kernels select seg[0] or seg[1]");

// the same for each table of a multi-table sort (blockIdx.y = table)
__global__ void scan_apply_chunks_seg_kernel(SortSegs t, long n, long nb) {
    const SortSeg g = blockIdx.y == 0 ? t.seg[0] : t.seg[1];
    if (seg_skipped(g)) return;  // the whole block
    const long gn = g.nkeys ? g.nkeys : n, gnb = (gn + kScanTile - 1) / kScanTile;
    if ((long)blockIdx.x < gnb) scan_apply_chunks_tile(g.count, gn, g.totals, g.cell_start, gnb);
}

size_t scan_partials_count(long n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

int exclusive_scan_i32(const int32_t* counts, int32_t* out, long n, int32_t* partials, hipStream_t s) {
    long nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) {
        MPMX_LAUNCH(fill_kernel<int32_t>, dim3(1), dim3(kBlock), 0, s, out, 1L, 0);
        return check_launch("scan(empty)", s);
    }
    MPMX_LAUNCH(scan_reduce_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, counts, n, partials);
    int rc = check_launch("scan_reduce", s);
    if (rc) return rc;
    MPMX_LAUNCH(scan_partials_kernel, dim3(1), dim3(kBlock), 0, s, partials, nb);
    rc = check_launch("scan_partials", s);
    if (rc) return rc;
    MPMX_LAUNCH(scan_apply_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, counts, n, partials, out, nb);
    return check_launch("scan_apply", s);
}

// ---------------------------------------------------------------------------
// counting sort of ids by key
// ---------------------------------------------------------------------------
// Runs of equal keys in consecutive lanes (particles stored in spatial order)
// take one atomic per run: the run's first lane adds the run length and every
// lane's rank is the returned base plus its offset in the run.  Ranks only
// place ids inside their bin, which is sorted by id afterwards.
__global__ void __launch_bounds__(kBlock) bin_count_kernel(const int32_t* __restrict__ keys, long n,
                                                           int32_t* __restrict__ count, int32_t* __restrict__ rank) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (i < n) ? keys[i] : -1;  // every lane stays for the cross-lane operations
    const int lane = (int)(threadIdx.x & 63);
    const int prev = __shfl_up(k, 1);
    const bool head = lane == 0 || prev != k;
    const unsigned long long heads = __ballot(head);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    const int my_head = 63 - __clzll((long long)(heads & upto));
    const unsigned long long above = heads & ~upto;
    const int next_head = above ? __ffsll((long long)above) - 1 : 64;
    int base = 0;
    if (head && k >= 0) base = atomicAdd(&count[k], next_head - lane);
    base = __shfl(base, my_head);
    if (i < n && k >= 0) rank[i] = base + (lane - my_head);
}

__global__ void bin_scatter_kernel(const int32_t* __restrict__ keys, long n, const int32_t* __restrict__ cell_start,
                                   const int32_t* __restrict__ rank, int32_t* __restrict__ sorted) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = keys[i];
    if (k >= 0) sorted[cell_start[k] + rank[i]] = (int32_t)i;
}

// Ids inside each key's run in ascending order of order[id] (of id itself when
// order is null): the order a stable sort by key gives when ids are numbered
// by order.  Runs are short (the particles of one cell): up to kRunReg they are
// sorted in registers by a bitonic network (static indices, one load and one
// store per entry, nothing written when the run is already in order); longer
// runs are heap-sorted in place.
constexpr int kRunReg = 16;

__device__ __forceinline__ void cswap(int& oa, int& va, int& ob, int& vb, bool up) {
    if ((oa > ob) == up) {
        int t = oa; oa = ob; ob = t;
        t = va; va = vb; vb = t;
    }
}

// POS: also pos[id] = its new position for the ids of a run it reorders
template <bool POS = false>
__device__ __forceinline__ void sort_run(long c, const int32_t* __restrict__ cell_start, int32_t* __restrict__ sorted,
                                         const int32_t* __restrict__ order, int32_t* __restrict__ pos = nullptr) {
    const int b = cell_start[c], e = cell_start[c + 1];
    const int len = e - b;
    if (len <= 1) return;
    if (len <= kRunReg) {
        int v[kRunReg], o[kRunReg];
        bool in_order = true;
#pragma unroll
        for (int k = 0; k < kRunReg; k++) {
            v[k] = k < len ? sorted[b + k] : 0x7fffffff;
            o[k] = k < len ? (order ? order[v[k]] : v[k]) : 0x7fffffff;
            if (k > 0 && k < len && o[k - 1] > o[k]) in_order = false;
        }
        if (in_order) return;
#pragma unroll
        for (int kk = 2; kk <= kRunReg; kk <<= 1)
#pragma unroll
            for (int j = kk >> 1; j > 0; j >>= 1)
#pragma unroll
                for (int i = 0; i < kRunReg; i++) {
                    const int l = i ^ j;
                    if (l > i) cswap(o[i], v[i], o[l], v[l], (i & kk) == 0);
                }
#pragma unroll
        for (int k = 0; k < kRunReg; k++)
            if (k < len) {
                sorted[b + k] = v[k];
                if constexpr (POS) pos[v[k]] = b + k;
            }
        return;
    }
    // heapsort in place: O(len log len) for the rare long run (an orphan bin, a
    // crowded hash cell); the keys are distinct, so the result is that of any sort
    int32_t* r = sorted + b;
    auto key = [&](int v) { return order ? order[v] : v; };
    auto sift = [&](int root, int m) {
        const int v = r[root], kv = key(v);
        while (2 * root + 1 < m) {
            int c = 2 * root + 1;
            int kc = key(r[c]);
            if (c + 1 < m) {
                const int k2 = key(r[c + 1]);
                if (k2 > kc) {
                    c++;
                    kc = k2;
                }
            }
            if (kc <= kv) break;
            r[root] = r[c];
            root = c;
        }
        r[root] = v;
    };
    for (int i = len / 2 - 1; i >= 0; i--) sift(i, len);
    for (int m = len - 1; m > 0; m--) {
        const int top = r[0];
        r[0] = r[m];
        r[m] = top;
        sift(0, m);
    }
    if constexpr (POS)
        for (int k = 0; k < len; k++) pos[r[k]] = b + k;
}

__global__ void __launch_bounds__(kBlock) bin_sort_runs_kernel(long nkeys, const int32_t* __restrict__ cell_start,
                                                               int32_t* __restrict__ sorted,
                                                               const int32_t* __restrict__ order) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nkeys) sort_run(c, cell_start, sorted, order);
}

__global__ void __launch_bounds__(kBlock) bin_sort_runs_seg_kernel(SortSegs t, long nkeys,
                                                                   const int32_t* __restrict__ order) {
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const SortSeg g = blockIdx.y == 0 ? t.seg[0] : t.seg[1];
    if (c >= (g.nkeys ? g.nkeys : nkeys) || seg_skipped(g)) return;
    int32_t* pos = seg_pos(g);
    if (pos)
        sort_run<true>(c, g.cell_start, g.sorted, order, pos);
    else
        sort_run<false>(c, g.cell_start, g.sorted, order);
}

__global__ void bin_scatter_seg_kernel(SortSegs t, int nseg, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
#pragma unroll
    for (int g = 0; g < kMaxSortSegs; g++) {
        if (g >= nseg) break;
        if (seg_skipped(t.seg[g])) continue;
        const int k = t.seg[g].keys[i];
        const int p = k >= 0 ? t.seg[g].cell_start[k] + t.seg[g].rank[i] : -1;
        if (k >= 0) t.seg[g].sorted[p] = (int32_t)i;
        int32_t* pos = seg_pos(t.seg[g]);
        if (pos) pos[i] = p;  // the in-run sort moves the points it reorders
    }
}

__global__ void invert_permutation_kernel(const int32_t* __restrict__ perm, int32_t* __restrict__ inv, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[perm[i]] = (int32_t)i;
}

// dst[(p*n + i)*w + k] = src[(p*n + perm[i])*w + k] for planes p, items i, words k
__global__ void gather_words_kernel(uint32_t* __restrict__ dst, const uint32_t* __restrict__ src,
                                    const int32_t* __restrict__ perm, long n, int planes, int words) {
    const long total = n * planes;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const long p = t / n, i = t - p * n;
        const long from = (p * n + perm[i]) * words, to = (p * n + i) * words;
        for (int k = 0; k < words; k++) dst[to + k] = src[from + k];
    }
}

size_t bin_workspace_bytes(long n, long nkeys) {
    return Carver::bytes_for(nkeys, 4) + Carver::bytes_for(nkeys + 1, 4) + Carver::bytes_for(n > 0 ? n : 1, 4) +
           Carver::bytes_for(scan_partials_count(nkeys), 4);
}

bool carve_bin(Carver& cv, long n, long nkeys, BinWork& w) {
    w.count = cv.take<int32_t>(nkeys);
    w.cell_start = cv.take<int32_t>(nkeys + 1);
    w.rank = cv.take<int32_t>(n > 0 ? n : 1);
    w.partials = cv.take<int32_t>(scan_partials_count(nkeys));
    return !cv.overflow;
}

int counting_sort_ids(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w, hipStream_t s,
                      const int32_t* order) {
    MPMX_LAUNCH(fill_kernel<int32_t>, dim3(fill_grid(nkeys)), dim3(kBlock), 0, s, w.count, nkeys, 0);
    int rc = check_launch("bin_zero", s);
    if (rc) return rc;
    if (n > 0) {
        MPMX_LAUNCH(bin_count_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, keys, n, w.count, w.rank);
        if ((rc = check_launch("bin_count", s))) return rc;
    }
    if ((rc = exclusive_scan_i32(w.count, w.cell_start, nkeys, w.partials, s))) return rc;
    if (n > 0) {
        MPMX_LAUNCH(bin_scatter_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, keys, n, w.cell_start,
                           w.rank, sorted);
        if ((rc = check_launch("bin_scatter", s))) return rc;
    }
    MPMX_LAUNCH(bin_sort_runs_kernel, dim3(grid_for(nkeys, kBlock)), dim3(kBlock), 0, s, nkeys, w.cell_start,
                       sorted, order);
    return check_launch("bin_sort_runs", s);
}

int counting_sort_from_chunk_counts(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w,
                                    hipStream_t s, const int32_t* order) {
    const long nb = (nkeys + kScanTile - 1) / kScanTile;
    int rc;
    if (nb == 0) {
        MPMX_LAUNCH(fill_kernel<int32_t>, dim3(1), dim3(kBlock), 0, s, w.cell_start, 1L, 0);
        if ((rc = check_launch("scan(empty)", s))) return rc;
    } else {
        MPMX_LAUNCH(scan_apply_chunks_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, w.count, nkeys, w.partials,
                    w.cell_start, nb);
        if ((rc = check_launch("scan_apply", s))) return rc;
    }
    if (n > 0) {
        MPMX_LAUNCH(bin_scatter_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, keys, n, w.cell_start,
                    w.rank, sorted);
        if ((rc = check_launch("bin_scatter", s))) return rc;
    }
    MPMX_LAUNCH(bin_sort_runs_kernel, dim3(grid_for(nkeys, kBlock)), dim3(kBlock), 0, s, nkeys, w.cell_start,
                sorted, order);
    return check_launch("bin_sort_runs", s);
}

int fill_i32(int32_t* dst, long n, int32_t value, hipStream_t s) {
    if (n <= 0) return 0;
    MPMX_LAUNCH(fill_kernel<int32_t>, dim3(fill_grid(n)), dim3(kBlock), 0, s, dst, n, value);
    return check_launch("fill_i32", s);
}

int counting_sort_segments(const SortSegs& t, int nseg, long n, long nkeys, hipStream_t s, const int32_t* order) {
    MPMX_REQUIRE(nseg >= 1 && nseg <= kMaxSortSegs, "counting_sort_segments: %d tables", nseg);
    const long nb = (nkeys + kScanTile - 1) / kScanTile;
    int rc;
    if (nb == 0) {
        for (int g = 0; g < nseg; g++) {
            MPMX_LAUNCH(fill_kernel<int32_t>, dim3(1), dim3(kBlock), 0, s, t.seg[g].cell_start, 1L, 0);
            if ((rc = check_launch("scan(empty)", s))) return rc;
        }
        return 0;
    }
    MPMX_LAUNCH(scan_apply_chunks_seg_kernel, dim3((unsigned)nb, nseg), dim3(kBlock), 0, s, t, nkeys, nb);
    if ((rc = check_launch("scan_apply(tables)", s))) return rc;
    if (n > 0) {
        MPMX_LAUNCH(bin_scatter_seg_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, t, nseg, n);
        if ((rc = check_launch("bin_scatter(tables)", s))) return rc;
    }
    MPMX_LAUNCH(bin_sort_runs_seg_kernel, dim3(grid_for(nkeys, kBlock), nseg), dim3(kBlock), 0, s, t, nkeys, order);
    return check_launch("bin_sort_runs(tables)", s);
}

// The scan / scatter / in-run sort half of counting_sort_ids, for callers that
// produced keys, ranks and counts themselves.
int counting_sort_from_counts(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w, hipStream_t s,
                              const int32_t* order) {
    int rc = exclusive_scan_i32(w.count, w.cell_start, nkeys, w.partials, s);
    if (rc) return rc;
    if (n > 0) {
        MPMX_LAUNCH(bin_scatter_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, keys, n, w.cell_start,
                           w.rank, sorted);
        if ((rc = check_launch("bin_scatter", s))) return rc;
    }
    MPMX_LAUNCH(bin_sort_runs_kernel, dim3(grid_for(nkeys, kBlock)), dim3(kBlock), 0, s, nkeys, w.cell_start,
                       sorted, order);
    return check_launch("bin_sort_runs", s);
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

const char* mpmx_last_error(void) { return g_err; }

int mpmx_invert_permutation(const int32_t* perm, int32_t* inv, long n, void* stream) {
    MPMX_REQUIRE(n >= 0, "mpmx_invert_permutation: n < 0");
    if (n == 0) return 0;
    MPMX_REQUIRE(perm && inv && perm != inv, "mpmx_invert_permutation: null or aliased pointer");
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(invert_permutation_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, perm, inv, n);
    return check_launch("invert_permutation", s);
}

int mpmx_gather_words(void* dst, const void* src, const int32_t* perm, long n, int planes, int item_words,
                      void* stream) {
    MPMX_REQUIRE(n >= 0 && planes >= 1 && item_words >= 1, "mpmx_gather_words: bad sizes");
    if (n == 0) return 0;
    MPMX_REQUIRE(dst && src && perm && dst != src, "mpmx_gather_words: null or aliased pointer");
    const long total = n * planes;
    long nb = (total + kBlock - 1) / kBlock;
    if (nb > 65536) nb = 65536;
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(gather_words_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, (uint32_t*)dst, (const uint32_t*)src, perm,
                n, planes, item_words);
    return check_launch("gather_words", s);
}

int mpmx_profile_begin(int capacity) {
    MPMX_REQUIRE(capacity > 0 && capacity <= (1 << 22), "mpmx_profile_begin: capacity out of range");
    if ((int)g_prof.pool.size() < capacity) {
        size_t have = g_prof.pool.size();
        g_prof.pool.resize(capacity);
        for (size_t i = have; i < (size_t)capacity; i++) MPMX_HIP(hipEventCreate(&g_prof.pool[i]));
        g_prof.names.resize(capacity);
    }
    g_prof.used = 0;
    g_prof.pending = -1;
    g_prof.on = true;
    return 0;
}

int mpmx_profile_mark(void* stream) {
    // an unnamed pair: a point in the stream that no kernel's time is taken from
    if (g_prof.on && g_prof.used + 2 <= g_prof.pool.size()) {
        MPMX_HIP(hipEventRecord(g_prof.pool[g_prof.used], as_stream(stream)));
        MPMX_HIP(hipEventRecord(g_prof.pool[g_prof.used + 1], as_stream(stream)));
        g_prof.names[g_prof.used] = g_prof.names[g_prof.used + 1] = nullptr;
        g_prof.used += 2;
    }
    return 0;
}

// one thread waits until the wall clock has advanced by `ticks` (always ends)
__global__ void hold_kernel(unsigned long long ticks) {
    const unsigned long long t0 = wall_clock64();
    while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(127);
}

int mpmx_profile_hold(void* stream, int microseconds) {
    MPMX_REQUIRE(microseconds >= 0 && microseconds <= 1000000, "mpmx_profile_hold: 0..1e6 microseconds");
    int dev = 0, khz = 0;
    MPMX_HIP(hipGetDevice(&dev));
    MPMX_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
    MPMX_REQUIRE(khz > 0, "mpmx_profile_hold: no wall clock rate");
    const unsigned long long ticks = (unsigned long long)khz * (unsigned long long)microseconds / 1000ull;
    hipLaunchKernelGGL(hold_kernel, dim3(1), dim3(1), 0, as_stream(stream), ticks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of profile hold failed: %s", hipGetErrorString(e));
        return (int)e;
    }
    return 0;
}

int mpmx_profile_end(void) {
    g_prof.on = false;
    return 0;
}

int mpmx_profile_report(char* buf, size_t buflen) {
    MPMX_REQUIRE(buf && buflen > 0, "mpmx_profile_report: bad buffer");
    std::map<std::string, std::pair<long, double>> acc;
    for (size_t i = 1; i < g_prof.used; i += 2) {
        if (!g_prof.names[i]) continue;
        float ms = 0.0f;
        hipError_t he = hipEventSynchronize(g_prof.pool[i]);
        if (he == hipSuccess) he = hipEventElapsedTime(&ms, g_prof.pool[i - 1], g_prof.pool[i]);
        if (he != hipSuccess) {
            (void)hipGetLastError();  // do not leave the error for the next launch check
            set_error("mpmx_profile_report: event timing failed: %s", hipGetErrorString(he));
            return (int)he;
        }
        auto& e = acc[g_prof.names[i]];
        e.first += 1;
        e.second += ms;
    }
    size_t off = 0;
    buf[0] = 0;
    for (auto& kv : acc) {
        int w = snprintf(buf + off, buflen - off, "%s %ld %.6f\n", kv.first.c_str(), kv.second.first,
                         kv.second.second);
        if (w < 0 || (size_t)w >= buflen - off) break;
        off += (size_t)w;
    }
    return 0;
}
int mpmx_version(void) { return 1; }

int mpmx_set_tuning(int knob, int value) {
    MPMX_REQUIRE(knob >= 0 && knob < MPMX_TUNE_COUNT, "mpmx_set_tuning: unknown knob %d", knob);
    const int vmax =
        (knob == MPMX_TUNE_G2P_TILES || knob == MPMX_TUNE_P2G_CHUNK || knob == MPMX_TUNE_XCD_TILES) ? 2 : 1;
    MPMX_REQUIRE(value >= 0 && value <= vmax, "mpmx_set_tuning: knob %d takes 0 to %d", knob, vmax);
    g_tuning[knob] = value;
    return 0;
}

int mpmx_get_tuning(int knob) { return tuning(knob); }

int mpmx_struct_layout(long* out, int cap) {
    const long v[10] = {(long)sizeof(mpmx_particles), (long)offsetof(mpmx_particles, order),
                        (long)sizeof(mpmx_grid), (long)sizeof(mpmx_step_params), (long)offsetof(mpmx_step_params, bc),
                        (long)offsetof(mpmx_particles, scene), (long)sizeof(mpmx_scene),
                        (long)sizeof(mpmx_xpbd_bounds), (long)offsetof(mpmx_xpbd_bounds, delta),
                        (long)offsetof(mpmx_xpbd_bounds, xpbd_count)};
    int k = 0;
    for (; out && k < cap && k < 10; k++) out[k] = v[k];
    return k;
}

// ---- contexts, buffers and captured substep sequences ----------------------

struct mpmx_ctx {
    int device;
    hipStream_t stream;
    bool capturing;
};
struct mpmx_graph {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

int mpmx_ctx_create(int device, mpmx_ctx** out) {
    MPMX_REQUIRE(out, "mpmx_ctx_create: null out");
    int count = 0;
    MPMX_HIP(hipGetDeviceCount(&count));
    MPMX_REQUIRE(device >= 0 && device < count, "mpmx_ctx_create: device %d of %d", device, count);
    MPMX_HIP(hipSetDevice(device));
    hipStream_t s;
    MPMX_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));  // default priority
    mpmx_ctx* c = new mpmx_ctx{device, s, false};
    *out = c;
    return 0;
}

int mpmx_ctx_destroy(mpmx_ctx* ctx) {
    if (!ctx) return 0;
    MPMX_HIP(hipStreamSynchronize(ctx->stream));
    MPMX_HIP(hipStreamDestroy(ctx->stream));
    delete ctx;
    return 0;
}

void* mpmx_ctx_stream(mpmx_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mpmx_buf_alloc(mpmx_ctx* ctx, size_t bytes, void** out) {
    MPMX_REQUIRE(ctx && out, "mpmx_buf_alloc: null argument");
    MPMX_REQUIRE(!ctx->capturing, "mpmx_buf_alloc: not allowed while a substep sequence is being captured");
    void* p = nullptr;
    MPMX_HIP(hipMalloc(&p, bytes > 0 ? bytes : 1));
    hipError_t e = hipMemsetAsync(p, 0, bytes > 0 ? bytes : 1, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(p);
        set_error("mpmx_buf_alloc: clearing failed: %s", hipGetErrorString(e));
        return (int)e;
    }
    *out = p;
    return 0;
}

int mpmx_buf_free(mpmx_ctx* ctx, void* p) {
    MPMX_REQUIRE(ctx, "mpmx_buf_free: null context");
    if (!p) return 0;
    MPMX_HIP(hipStreamSynchronize(ctx->stream));
    MPMX_HIP(hipFree(p));
    return 0;
}

int mpmx_buf_upload(mpmx_ctx* ctx, void* dst, const void* host, size_t bytes) {
    MPMX_REQUIRE(ctx && (bytes == 0 || (dst && host)), "mpmx_buf_upload: null argument");
    if (bytes == 0) return 0;
    MPMX_HIP(hipMemcpyAsync(dst, host, bytes, hipMemcpyHostToDevice, ctx->stream));
    MPMX_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int mpmx_buf_download(mpmx_ctx* ctx, void* host, const void* src, size_t bytes) {
    MPMX_REQUIRE(ctx && (bytes == 0 || (src && host)), "mpmx_buf_download: null argument");
    if (bytes == 0) return 0;
    MPMX_HIP(hipMemcpyAsync(host, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    MPMX_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
}

int mpmx_graph_begin(mpmx_ctx* ctx) {
    MPMX_REQUIRE(ctx && !ctx->capturing, "mpmx_graph_begin: no context or already capturing");
    MPMX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    ctx->capturing = true;
    return 0;
}

int mpmx_graph_end(mpmx_ctx* ctx, mpmx_graph** out) {
    MPMX_REQUIRE(ctx && ctx->capturing && out, "mpmx_graph_end: not capturing");
    ctx->capturing = false;
    hipGraph_t g;
    MPMX_HIP(hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ex;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        set_error("mpmx_graph_end: instantiate failed: %s", hipGetErrorString(e));
        return (int)e;
    }
    *out = new mpmx_graph{g, ex};
    return 0;
}

int mpmx_graph_launch(mpmx_graph* g, void* stream) {
    MPMX_REQUIRE(g, "mpmx_graph_launch: null graph");
    MPMX_HIP(hipGraphLaunch(g->exec, as_stream(stream)));
    return 0;
}

int mpmx_graph_destroy(mpmx_graph* g) {
    if (!g) return 0;
    MPMX_HIP(hipGraphExecDestroy(g->exec));
    MPMX_HIP(hipGraphDestroy(g->graph));
    delete g;
    return 0;
}

int mpmx_stream_sync(void* stream) {
    MPMX_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int mpmx_fill_i32(int32_t* dst, long n, int32_t value, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || dst), "mpmx_fill_i32: bad arguments");
    if (n == 0) return 0;
    MPMX_LAUNCH(fill_kernel<int32_t>, dim3(fill_grid(n)), dim3(kBlock), 0, as_stream(stream), dst, n, value);
    return check_launch("fill_i32", as_stream(stream));
}

int mpmx_fill_f32(float* dst, long n, float value, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || dst), "mpmx_fill_f32: bad arguments");
    if (n == 0) return 0;
    MPMX_LAUNCH(fill_kernel<float>, dim3(fill_grid(n)), dim3(kBlock), 0, as_stream(stream), dst, n, value);
    return check_launch("fill_f32", as_stream(stream));
}

int mpmx_copy_f32(float* dst, const float* src, long n, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (dst && src)), "mpmx_copy_f32: bad arguments");
    if (n == 0 || dst == src) return 0;
    MPMX_LAUNCH(copy_kernel, dim3(fill_grid(n)), dim3(kBlock), 0, as_stream(stream), dst, src, n);
    return check_launch("copy_f32", as_stream(stream));
}

}  // extern "C"
