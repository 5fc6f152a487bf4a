// Per-substep reductions of the driver loop (simulationRoutines.py:538-675),
// accumulated in FP64 on the device.  The reference adds float atomics into
// one address; here a fixed grid of blocks each reduces a fixed grid-stride
// subset in a fixed tree order and one block sums the block partials in order,
// so the result is identical from run to run.
#include <atomic>

#include "mpmx_device.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

constexpr int kRedBlocks = 512;

enum RedMode { RED_VCONV = 0, RED_SUM_ACTIVE = 1, RED_BY_TYPE = 2 };

struct RedAcc {
    double s;
    int c[4];
};

template <int MODE>
__global__ void __launch_bounds__(kBlock) reduce_pass1(int n, const int32_t* __restrict__ active,
                                                       const int32_t* __restrict__ mat, const float* __restrict__ a,
                                                       const float* __restrict__ radius, double* __restrict__ psum,
                                                       int32_t* __restrict__ pcnt) {
    __shared__ double ss[kBlock];
    __shared__ int sc[4][kBlock];
    double s = 0.0;
    int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        int act = active[p] == 1;
        if (MODE == RED_VCONV) {
            if (act) {
                float vx = a[p], vy = a[n + p], vz = a[2L * n + p];
                float l = sqrtf(vx * vx + vy * vy + vz * vz);
                s += (double)(l / radius[p]);
                c0++;
            }
        } else if (MODE == RED_SUM_ACTIVE) {
            if (act) {
                s += (double)a[p];
                c0++;
            }
        } else {
            int xp = mat[p] == 2;
            if (xp) {
                if (act) c1++; else c3++;
            } else {
                if (act) c0++; else c2++;
            }
        }
    }
    int t = threadIdx.x;
    ss[t] = s;
    sc[0][t] = c0; sc[1][t] = c1; sc[2][t] = c2; sc[3][t] = c3;
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
            ss[t] += ss[t + off];
#pragma unroll
            for (int k = 0; k < 4; k++) sc[k][t] += sc[k][t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        psum[blockIdx.x] = ss[0];
#pragma unroll
        for (int k = 0; k < 4; k++) pcnt[blockIdx.x * 4 + k] = sc[k][0];
    }
}

__global__ void __launch_bounds__(kBlock) reduce_pass2(int nb, const double* __restrict__ psum,
                                                       const int32_t* __restrict__ pcnt, double* __restrict__ out_sum,
                                                       int32_t* __restrict__ out_cnt, int ncnt) {
    __shared__ double ss[kBlock];
    __shared__ int sc[4][kBlock];
    int t = threadIdx.x;
    double s = 0.0;
    int c[4] = {0, 0, 0, 0};
    for (int b = t; b < nb; b += kBlock) {
        s += psum[b];
#pragma unroll
        for (int k = 0; k < 4; k++) c[k] += pcnt[b * 4 + k];
    }
    ss[t] = s;
#pragma unroll
    for (int k = 0; k < 4; k++) sc[k][t] = c[k];
    __syncthreads();
    for (int off = kBlock / 2; off > 0; off >>= 1) {
        if (t < off) {
            ss[t] += ss[t + off];
#pragma unroll
            for (int k = 0; k < 4; k++) sc[k][t] += sc[k][t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        if (out_sum) *out_sum = ss[0];
        for (int k = 0; k < ncnt; k++) out_cnt[k] = sc[k][0];
    }
}

static int nblocks(int n) {
    int b = (n + kBlock - 1) / kBlock;
    if (b > kRedBlocks) b = kRedBlocks;
    if (b < 1) b = 1;
    return b;
}

template <int MODE>
static int run_reduce(int n, const int32_t* active, const int32_t* mat, const float* a, const float* radius,
                      double* out_sum, int32_t* out_cnt, int ncnt, void* ws, size_t ws_bytes, hipStream_t s) {
    MPMX_REQUIRE(n >= 0 && out_cnt, "reduce: bad arguments");
    MPMX_REQUIRE(ws_bytes >= mpmx_reduce_workspace_bytes(n) && ws, "reduce: workspace too small");
    int nb = nblocks(n);
    Carver cv(ws, ws_bytes);
    double* psum = cv.take<double>(kRedBlocks);
    int32_t* pcnt = cv.take<int32_t>(kRedBlocks * 4);
    MPMX_LAUNCH(reduce_pass1<MODE>, dim3(nb), dim3(kBlock), 0, s, n, active, mat, a, radius, psum, pcnt);
    int rc = check_launch("reduce_pass1", s);
    if (rc) return rc;
    MPMX_LAUNCH(reduce_pass2, dim3(1), dim3(kBlock), 0, s, nb, psum, pcnt, out_sum, out_cnt, ncnt);
    return check_launch("reduce_pass2", s);
}

// ---- the allocation-free residual (mpmx_velocity_convergence_noalloc) ------
// Its block partials live in a file-scope device array instead of a caller's
// workspace: kRedBlocks doubles and kRedBlocks x 4 counts, 12 KB per GPU, held
// until the process ends.  Host code never passes the array's name to a kernel
// (on the host it is a stand-in, not a device address): the address comes from
// hipGetSymbolAddress, taken once per device with that device current, and only
// that pointer reaches the kernels.  Every call on one GPU shares the array;
// calls that overlap on one GPU race.
constexpr size_t kNoAllocPsumBytes = (size_t)kRedBlocks * sizeof(double);
constexpr size_t kNoAllocScratchBytes = kNoAllocPsumBytes + (size_t)kRedBlocks * 4 * sizeof(int32_t);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::atomic<void*> g_scratch_ptr[kMaxDevices];

static int scratch_for_current_device(unsigned char** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("mpmx_velocity_convergence_noalloc: hipGetDevice failed: %s", hipGetErrorString(e));
        return (int)e;
    }
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("mpmx_velocity_convergence_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
        return MPMX_ERANGE;
    }
    void* p = g_scratch_ptr[dev].load(std::memory_order_acquire);
    if (!p) {
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("mpmx_velocity_convergence_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return e != hipSuccess ? (int)e : MPMX_EINVAL;
        }
        g_scratch_ptr[dev].store(p, std::memory_order_release);  // the same address whichever thread wins
    }
    *out = static_cast<unsigned char*>(p);
    return 0;
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

size_t mpmx_reduce_workspace_bytes(int n) {
    (void)n;
    return Carver::bytes_for(kRedBlocks, sizeof(double)) + Carver::bytes_for(kRedBlocks * 4, sizeof(int32_t));
}

int mpmx_velocity_convergence(int n, const int32_t* activeLabel, const float* v, const float* radius,
                              double* out_sum, int32_t* out_count, void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n == 0 || (activeLabel && v && radius && out_sum), "mpmx_velocity_convergence: null pointer");
    return run_reduce<RED_VCONV>(n, activeLabel, nullptr, v, radius, out_sum, out_count, 1, workspace, ws_bytes,
                                 as_stream(stream));
}

int mpmx_velocity_convergence_noalloc(int n, const int32_t* activeLabel, const float* v, const float* radius,
                                      double* out_sum, int32_t* out_count, void* stream) {
    MPMX_REQUIRE(n >= 0, "mpmx_velocity_convergence_noalloc: n < 0");
    MPMX_REQUIRE(n == 0 || (activeLabel && v && radius), "mpmx_velocity_convergence_noalloc: null input");
    MPMX_REQUIRE(out_sum && out_count, "mpmx_velocity_convergence_noalloc: null output");
    // the partials of nblocks(n) <= kRedBlocks blocks: the scratch fits every int n
    MPMX_REQUIRE(nblocks(n) <= kRedBlocks, "mpmx_velocity_convergence_noalloc: n needs more than the scratch");
    unsigned char* base = nullptr;
    const int rc = scratch_for_current_device(&base);
    if (rc) return rc;
    double* psum = reinterpret_cast<double*>(base);
    int32_t* pcnt = reinterpret_cast<int32_t*>(base + kNoAllocPsumBytes);
    const hipStream_t s = as_stream(stream);
    const int nb = nblocks(n);
    MPMX_LAUNCH(reduce_pass1<RED_VCONV>, dim3(nb), dim3(kBlock), 0, s, n, activeLabel, nullptr, v, radius, psum, pcnt);
    int r = check_launch("reduce_pass1", s);
    if (r) return r;
    MPMX_LAUNCH(reduce_pass2, dim3(1), dim3(kBlock), 0, s, nb, psum, pcnt, out_sum, out_count, 1);
    return check_launch("reduce_pass2", s);
}

int mpmx_sum_float_array_active(int n, const float* arr, const int32_t* activeLabel, double* out_sum,
                                int32_t* out_count, void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n == 0 || (activeLabel && arr && out_sum), "mpmx_sum_float_array_active: null pointer");
    return run_reduce<RED_SUM_ACTIVE>(n, activeLabel, nullptr, arr, nullptr, out_sum, out_count, 1, workspace,
                                      ws_bytes, as_stream(stream));
}

int mpmx_count_particles_by_type(int n, const int32_t* activeLabel, const int32_t* materialLabel, int32_t* out4,
                                 void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n == 0 || (activeLabel && materialLabel), "mpmx_count_particles_by_type: null pointer");
    return run_reduce<RED_BY_TYPE>(n, activeLabel, materialLabel, nullptr, nullptr, nullptr, out4, 4, workspace,
                                   ws_bytes, as_stream(stream));
}

}  // extern "C"
