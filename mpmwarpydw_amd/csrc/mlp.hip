// Forward, backward and Adam kernels of the learned return map's network
// (mlp.hpp).  Samples are rows of X[B][10] (features) and T[B][3] (targets),
// predictions rows of Y[B][3].  The loss is the mean squared error over all
// B*3 outputs.
//
// Training kernel: every wavefront runs tiles of 16 samples through the
// forward pass, back-propagates (bfloat16 MFMA, FP32 accumulation) and keeps
// its share of the weight and bias gradients in registers.  At the end the
// four wavefronts of a workgroup add their gradients in LDS in a fixed order
// and the workgroup writes one partial gradient; a second kernel sums the
// partials of all workgroups in a fixed order (and optionally applies Adam).
// The tile -> workgroup assignment depends only on B, so results are
// reproducible bit for bit.
#include "mlp.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

constexpr int kMlpBlocks = 256;  // workgroups of the training kernel (at most)

// B operands of the inputs, high and low bfloat16 halves: lane l holds features
// 4(l/16)..+3 of sample tile*16 + l%16
__device__ __forceinline__ void x_frag(const float* __restrict__ X, int B, int tile, int lane, bf16x4& hi,
                                       bf16x4& lo) {
    const int b = tile * 16 + (lane & 15);
    const int f0 = 4 * (lane >> 4);
    float v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) v[j] = (b < B && f0 + j < kNI) ? X[(long)b * kNI + f0 + j] : 0.0f;
    hi = pack_bf16(v[0], v[1], v[2], v[3]);
    lo = pack_bf16_lo(v[0], v[1], v[2], v[3]);
}

__global__ void __launch_bounds__(kBlock) mlp_forward_kernel(int B, const float* __restrict__ prm,
                                                             const float* __restrict__ X, float* __restrict__ Y) {
    __shared__ MlpWeights w;
    load_weights(prm, w);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int ntiles = (B + 15) / 16;
    for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {  // no barrier inside
        MlpTile t;
        bf16x4 xh, xl;
        x_frag(X, B, tile, lane, xh, xl);
        mlp_tile_forward(w, xh, xl, lane, t);
        const int b = tile * 16 + lane;
        if (lane < 16 && b < B) {
            Y[(long)b * kNO + 0] = t.y[0];
            Y[(long)b * kNO + 1] = t.y[1];
            Y[(long)b * kNO + 2] = t.y[2];
        }
    }
}

// per-wavefront LDS of the training kernel (bfloat16 bits)
struct TrainWave {
    unsigned short H1[kNH][16];  // [feature][sample], high halves
    unsigned short H1L[kNH][16];  // low halves
    unsigned short H2[kNH][16];
    unsigned short H2L[kNH][16];
    unsigned short G1[kNH][16];
    unsigned short G2[kNH][16];
    unsigned short G3[16][16];
};
union TrainScratch {
    TrainWave wv[4];
    float red[kNParam];
};

// sum over the 16 lanes that share lane/16 (a fixed butterfly)
__device__ __forceinline__ float sum16(float v) {
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

__global__ void __launch_bounds__(kBlock) mlp_train_kernel(int B, const float* __restrict__ prm,
                                                           const float* __restrict__ X, const float* __restrict__ T,
                                                           float scale, float* __restrict__ gpart,
                                                           double* __restrict__ lpart) {
    __shared__ MlpWeights w;
    __shared__ unsigned short W2T[kNH][kNH];  // [i][o]
    __shared__ unsigned short W3T[kNH][16];   // [i][o], o >= 3 zero
    __shared__ TrainScratch sc;
    __shared__ double lred[4];
    load_weights(prm, w);
    for (int k = threadIdx.x; k < kNH * kNH; k += blockDim.x) {
        const int o = k / kNH, i = k % kNH;
        W2T[i][o] = bf16_bits(prm[kOW2 + o * kNH + i]);
    }
    for (int k = threadIdx.x; k < kNH * 16; k += blockDim.x) {
        const int i = k / 16, o = k % 16;
        W3T[i][o] = o < kNO ? bf16_bits(prm[kOW3 + o * kNH + i]) : 0;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int r0 = 4 * (lane >> 4);
    const int col = lane & 15;
    TrainWave& tw = sc.wv[wave];

    f32x4 gW1[4], gW2[4][4], gW3[4];
    float gb1[4][4], gb2[4][4], gb3[3];
    const f32x4 zero4 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < 4; a++) {
        gW1[a] = zero4;
        gW3[a] = zero4;
#pragma unroll
        for (int b = 0; b < 4; b++) {
            gW2[a][b] = zero4;
            gb1[a][b] = 0.0f;
            gb2[a][b] = 0.0f;
        }
    }
    gb3[0] = gb3[1] = gb3[2] = 0.0f;
    double loss = 0.0;

    const int ntiles = (B + 15) / 16;
    const int per_round = gridDim.x * 4;
    const int rounds = (ntiles + per_round - 1) / per_round;  // the same for every wavefront: barriers below
    for (int rd = 0; rd < rounds; rd++) {
        const int tile = (rd * gridDim.x + blockIdx.x) * 4 + wave;
        const int nvalid = max(0, min(16, B - tile * 16));  // 0 for a tile past the end: all terms vanish
        MlpTile t;
        bf16x4 xh, xl;
        x_frag(X, B, tile, lane, xh, xl);
        mlp_tile_forward(w, xh, xl, lane, t);
        // X^T fragments (high and low halves) for the W1 gradient (k = sample 4(l/16)+j, column = feature l%16)
        float xt[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int b = r0 + j;
            xt[j] = (b < nvalid && col < kNI) ? X[(long)(tile * 16 + b) * kNI + col] : 0.0f;
        }
        const bf16x4 xtb = pack_bf16(xt[0], xt[1], xt[2], xt[3]);
        const bf16x4 xtl = pack_bf16_lo(xt[0], xt[1], xt[2], xt[3]);
#pragma unroll
        for (int mb = 0; mb < 4; mb++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                tw.H1[mb * 16 + r0 + j][col] = (unsigned short)t.h1b[mb][j];
                tw.H1L[mb * 16 + r0 + j][col] = (unsigned short)t.h1l[mb][j];
                tw.H2[mb * 16 + r0 + j][col] = (unsigned short)t.h2b[mb][j];
                tw.H2L[mb * 16 + r0 + j][col] = (unsigned short)t.h2l[mb][j];
            }
        // output error: lanes 0..15 hold outputs 0..2 of sample l
        float g3[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        if (lane < 16 && lane < nvalid) {
            const long b = (long)tile * 16 + lane;
#pragma unroll
            for (int j = 0; j < kNO; j++) {
                const float d = t.y[j] - T[b * kNO + j];
                loss += (double)d * (double)d;
                g3[j] = d * scale;
                gb3[j] += g3[j];
            }
        }
        const bf16x4 g3b = pack_bf16(g3[0], g3[1], g3[2], g3[3]);
#pragma unroll
        for (int j = 0; j < 4; j++) tw.G3[r0 + j][col] = (unsigned short)g3b[j];
        // back to the second hidden layer
        bf16x4 g2b[4];
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const f32x4 gh = mfma16(a_frag<16>(W3T, mb, 0, lane), g3b, zero4);
            float g2[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                g2[j] = gh[j] * gelu_grad_f(t.z2[mb][j]);
                gb2[mb][j] += g2[j];
            }
            g2b[mb] = pack_bf16(g2[0], g2[1], g2[2], g2[3]);
#pragma unroll
            for (int j = 0; j < 4; j++) tw.G2[mb * 16 + r0 + j][col] = (unsigned short)g2b[mb][j];
        }
        // back to the first hidden layer
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            f32x4 c = zero4;
#pragma unroll
            for (int kb = 0; kb < 4; kb++) c = mfma16(a_frag<kNH>(W2T, mb, kb, lane), g2b[kb], c);
            float g1[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                g1[j] = c[j] * gelu_grad_f(t.z1[mb][j]);
                gb1[mb][j] += g1[j];
                tw.G1[mb * 16 + r0 + j][col] = bf16_bits(g1[j]);
            }
        }
        __syncthreads();  // the tile's transposes are in LDS
        // weight gradients: sum over the tile's 16 samples (the K dimension)
#pragma unroll
        for (int mb = 0; mb < 4; mb++) {
            const bf16x4 ag1 = a_frag<16>(tw.G1, mb, 0, lane);
            gW1[mb] = mfma16(ag1, xtl, mfma16(ag1, xtb, gW1[mb]));
            const bf16x4 ag2 = a_frag<16>(tw.G2, mb, 0, lane);
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
                gW2[mb][nb] = mfma16(ag2, a_frag<16>(tw.H1L, nb, 0, lane),
                                     mfma16(ag2, a_frag<16>(tw.H1, nb, 0, lane), gW2[mb][nb]));
        }
        const bf16x4 ag3 = a_frag<16>(tw.G3, 0, 0, lane);
#pragma unroll
        for (int nb = 0; nb < 4; nb++)
            gW3[nb] = mfma16(ag3, a_frag<16>(tw.H2L, nb, 0, lane), mfma16(ag3, a_frag<16>(tw.H2, nb, 0, lane), gW3[nb]));
        __syncthreads();  // before the next tile overwrites the transposes
    }

    // biases: sum over the samples held by the 16 lanes of each row group
#pragma unroll
    for (int mb = 0; mb < 4; mb++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            gb1[mb][j] = sum16(gb1[mb][j]);
            gb2[mb][j] = sum16(gb2[mb][j]);
        }
#pragma unroll
    for (int j = 0; j < 3; j++) gb3[j] = sum16(gb3[j]);
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) loss += __shfl_xor(loss, o);
    if (lane == 0) lred[wave] = loss;

    // wavefront 0 stores, wavefronts 1..3 add in turn (fixed order)
    for (int wv = 0; wv < 4; wv++) {
        if (wave == wv) {
            const bool first = wv == 0;
#pragma unroll
            for (int mb = 0; mb < 4; mb++) {
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int o = mb * 16 + r0 + j;
                    if (col < kNI) {
                        float& d = sc.red[kOW1 + o * kNI + col];
                        d = first ? gW1[mb][j] : d + gW1[mb][j];
                    }
#pragma unroll
                    for (int nb = 0; nb < 4; nb++) {
                        float& d = sc.red[kOW2 + o * kNH + nb * 16 + col];
                        d = first ? gW2[mb][nb][j] : d + gW2[mb][nb][j];
                    }
                    if (col == 0) {
                        float& d1 = sc.red[kOB1 + o];
                        d1 = first ? gb1[mb][j] : d1 + gb1[mb][j];
                        float& d2 = sc.red[kOB2 + o];
                        d2 = first ? gb2[mb][j] : d2 + gb2[mb][j];
                    }
                }
            }
#pragma unroll
            for (int nb = 0; nb < 4; nb++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const int o = r0 + j;
                    if (o < kNO) {
                        float& d = sc.red[kOW3 + o * kNH + nb * 16 + col];
                        d = first ? gW3[nb][j] : d + gW3[nb][j];
                    }
                }
            if (lane == 0) {
#pragma unroll
                for (int j = 0; j < kNO; j++) {
                    float& d = sc.red[kOB3 + j];
                    d = first ? gb3[j] : d + gb3[j];
                }
            }
        }
        __syncthreads();
    }
    float* out = gpart + (long)blockIdx.x * kNParam;
    for (int k = threadIdx.x; k < kNParam; k += blockDim.x) out[k] = sc.red[k];
    if (threadIdx.x == 0) lpart[blockIdx.x] = ((lred[0] + lred[1]) + lred[2]) + lred[3];
}

// grad = sum of the workgroup partials in order; optionally one Adam step
__global__ void mlp_reduce_kernel(int nb, const float* __restrict__ gpart, const double* __restrict__ lpart,
                                  double inv_count, float* __restrict__ grad, double* __restrict__ loss) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < kNParam) {
        float g = 0.0f;
        for (int b = 0; b < nb; b++) g += gpart[(long)b * kNParam + k];
        grad[k] = g;
    }
    if (k == 0 && loss) {
        double s = 0.0;
        for (int b = 0; b < nb; b++) s += lpart[b];
        *loss = s * inv_count;
    }
}

__global__ void adam_kernel(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

static int train_blocks(int B) {
    const int ntiles = (B + 15) / 16;
    int nb = (ntiles + 3) / 4;
    return nb < 1 ? 1 : (nb > kMlpBlocks ? kMlpBlocks : nb);
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

int mpmx_mlp_nparams(void) { return kNParam; }

int mpmx_mlp_forward(int B, const float* params, const float* X, float* Y, void* stream) {
    MPMX_REQUIRE(B >= 0, "mpmx_mlp_forward: B < 0");
    if (B == 0) return 0;
    MPMX_REQUIRE(params && X && Y, "mpmx_mlp_forward: null pointer");
    hipStream_t s = as_stream(stream);
    const int ntiles = (B + 15) / 16;
    int nb = (ntiles + 3) / 4;
    if (nb > 4096) nb = 4096;
    MPMX_LAUNCH(mlp_forward_kernel, dim3(nb), dim3(kBlock), 0, s, B, params, X, Y);
    return check_launch("mlp_forward", s);
}

size_t mpmx_mlp_workspace_bytes(int B) {
    if (B < 0) return 0;
    const size_t nb = (size_t)train_blocks(B);
    return Carver::bytes_for(nb * kNParam, sizeof(float)) + Carver::bytes_for(nb, sizeof(double));
}

int mpmx_mlp_loss_grad(int B, const float* params, const float* X, const float* T, float* grad, double* loss,
                       void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(B > 0, "mpmx_mlp_loss_grad: B must be positive");
    MPMX_REQUIRE(params && X && T && grad, "mpmx_mlp_loss_grad: null pointer");
    MPMX_REQUIRE(workspace && ws_bytes >= mpmx_mlp_workspace_bytes(B), "mpmx_mlp_loss_grad: workspace too small");
    hipStream_t s = as_stream(stream);
    const int nb = train_blocks(B);
    Carver cv(workspace, ws_bytes);
    float* gpart = cv.take<float>((size_t)nb * kNParam);
    double* lpart = cv.take<double>(nb);
    MPMX_REQUIRE(!cv.overflow, "mpmx_mlp_loss_grad: workspace carve failed");
    const float scale = (float)(2.0 / (3.0 * (double)B));
    MPMX_LAUNCH(mlp_train_kernel, dim3(nb), dim3(kBlock), 0, s, B, params, X, T, scale, gpart, lpart);
    int rc = check_launch("mlp_train", s);
    if (rc) return rc;
    MPMX_LAUNCH(mlp_reduce_kernel, dim3(grid_for(kNParam, kBlock)), dim3(kBlock), 0, s, nb, gpart, lpart,
                1.0 / (3.0 * (double)B), grad, loss);
    return check_launch("mlp_reduce", s);
}

int mpmx_mlp_adam_step(int n, float* params, const float* grad, float* m, float* v, float lr, float beta1, float beta2,
                       float eps, int t, void* stream) {
    MPMX_REQUIRE(n >= 0 && t >= 1, "mpmx_mlp_adam_step: bad n or t");
    if (n == 0) return 0;
    MPMX_REQUIRE(params && grad && m && v, "mpmx_mlp_adam_step: null pointer");
    const float c1 = (float)(1.0 - pow((double)beta1, t)), c2 = (float)(1.0 - pow((double)beta2, t));
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(adam_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, params, grad, m, v, lr, beta1, beta2,
                eps, c1, c2);
    return check_launch("adam", s);
}

}  // extern "C"
