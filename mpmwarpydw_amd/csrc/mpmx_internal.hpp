// Host-side helpers shared by the libmpmx translation units: error reporting,
// launch checking and the workspace carving used by the entry points that need
// scratch memory.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/mpmx.h"

namespace mpmx {

void set_error(const char* fmt, ...);
int check_launch(const char* what, hipStream_t s);
// launch profiler hook: records a start event before a kernel when profiling is on
// when the profiler is on, the events that time the next launch (true)
bool prof_pair(hipStream_t s, hipEvent_t* start, hipEvent_t* stop);
inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }
// launch-shape knobs (mpmx_set_tuning); read on the host when a launch is issued
int tuning(int knob);
inline unsigned grid_for(long n, int block) { return (unsigned)((n + block - 1) / block); }

// Bump allocator over a caller-provided workspace; every slice is 256-byte aligned.
struct Carver {
    char* base;
    size_t cap;
    size_t off = 0;
    bool overflow = false;
    Carver(void* b, size_t c) : base(static_cast<char*>(b)), cap(c) {}
    template <typename T>
    T* take(size_t count) {
        size_t bytes = (count * sizeof(T) + 255) & ~size_t(255);
        if (off + bytes > cap) {
            overflow = true;
            return nullptr;
        }
        T* p = reinterpret_cast<T*>(base + off);
        off += bytes;
        return p;
    }
    static size_t bytes_for(size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); }
};

// Exclusive scan of n int32 counts into out[0..n] (out[n] = total), on stream s.
// partials must hold scan_partials_count(n) ints.
// exclusive scans cut their input into chunks of kScanChunk values
constexpr int kScanChunk = 2048;
size_t scan_partials_count(long n);
int exclusive_scan_i32(const int32_t* counts, int32_t* out, long n, int32_t* partials, hipStream_t s);

// Counting sort of particle ids by int32 key (key < 0 = not binned):
//   count[key]++ (rank kept per id), cell_start = exclusive scan(count),
//   sorted[cell_start[key] + rank] = id, then each key's id run is sorted by
// order[id] (by id when order is null), so the result equals a stable sort by
// key of the ids numbered by order.
struct BinWork {
    int32_t* count;      // [nkeys]
    int32_t* cell_start; // [nkeys + 1]
    int32_t* rank;       // [n]
    int32_t* partials;   // scan partials
};
size_t bin_workspace_bytes(long n, long nkeys);
bool carve_bin(Carver& cv, long n, long nkeys, BinWork& w);
int counting_sort_ids(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w, hipStream_t s,
                      const int32_t* order = nullptr);
// scan + scatter + in-run sort only (w.count and w.rank already filled)
int counting_sort_from_counts(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w, hipStream_t s,
                              const int32_t* order = nullptr);
// as counting_sort_from_counts, for producers that also accumulated the total
// of every chunk of kScanChunk keys into w.partials (one scan pass fewer)
int counting_sort_from_chunk_counts(const int32_t* keys, long n, long nkeys, int32_t* sorted, BinWork& w,
                                    hipStream_t s, const int32_t* order = nullptr);
// Several tables sorted over the same n items in one set of launches (nkeys:
// the most keys of any table): table g has the key of every item (-1: not in the table), its rank
// inside its key, the counts in key order and their totals per chunk of
// kScanChunk keys; it receives its cell ranges ([nkeys + 1]) and sorted ids.
constexpr int kMaxSortSegs = 2;
struct SortSeg {
    const int32_t* keys;
    const int32_t* rank;
    const int32_t* count;
    const int32_t* totals;
    int32_t* cell_start;
    int32_t* sorted;
    long nkeys;  // keys of this table (nkeys of counting_sort_segments when 0)
    int32_t* pos;  // may be null: pos[i] = i's position in sorted, -1 where i has no key
    const int32_t* pos_gate;  // may be null: pos is written only while *pos_gate != 0
    const int32_t* skip_gate; // may be null: nothing of this table is scanned, scattered or sorted
                              // while *skip_gate == 0
};
__device__ __forceinline__ bool seg_skipped(const SortSeg& g) { return g.skip_gate && *g.skip_gate == 0; }
// the position array of g that the scatter and the in-run sort write (null: none)
__device__ __forceinline__ int32_t* seg_pos(const SortSeg& g) {
    return (g.pos && (!g.pos_gate || *g.pos_gate != 0)) ? g.pos : nullptr;
}
struct SortSegs {
    SortSeg seg[kMaxSortSegs];
};
int counting_sort_segments(const SortSegs& t, int nseg, long n, long nkeys, hipStream_t s, const int32_t* order);
// dst[0..n) = value on stream s (a kernel: safe inside a captured graph)
int fill_i32(int32_t* dst, long n, int32_t value, hipStream_t s);

// the constitutive update of every active MPM particle with the learned
// projection judged by the exact map (params NULL: exact projection for all);
// counters[3] (may be NULL) += {plastic states, accepted predictions, updated};
// flag: n bytes of scratch (needed with params)
int launch_stress_learned(const mpmx_particles& P, float eff, const float* params, float tol, int64_t* counters,
                          uint8_t* flag, hipStream_t s);

}  // namespace mpmx

// every kernel launch of the library goes through this macro: with the
// profiler on, the launch carries a start and a stop event that the dispatch
// itself timestamps (the kernel's own begin and end, as a profiler sees them);
// check_launch then names the pair
#define MPMX_LAUNCH(kernel, grid, block, shmem, stream, ...)                                               \
    do {                                                                                                   \
        hipEvent_t mpmx_ev0_, mpmx_ev1_;                                                                   \
        if (::mpmx::prof_pair((stream), &mpmx_ev0_, &mpmx_ev1_))                                           \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, mpmx_ev0_, mpmx_ev1_, 0, ##__VA_ARGS__); \
        else                                                                                               \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ##__VA_ARGS__);                         \
    } while (0)

#define MPMX_REQUIRE(cond, ...)                 \
    do {                                        \
        if (!(cond)) {                          \
            ::mpmx::set_error(__VA_ARGS__);     \
            return MPMX_EINVAL;                 \
        }                                       \
    } while (0)

#define MPMX_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            ::mpmx::set_error("%s failed: %s", #call, hipGetErrorString(e_));       \
            return (int)e_;                                                         \
        }                                                                           \
    } while (0)
