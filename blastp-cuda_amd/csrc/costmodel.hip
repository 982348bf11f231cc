// costmodel.hip -- the learned per-subject cost model that orders and bins
// scan work.  It never filters: a scan's results do not depend on the order
// it processes subjects in.
//
//   k_features   one wavefront per subject: log(1 + length), the 20
//                standard-residue composition fractions and log(1 + hits of
//                the subject's words against a sampled word-count table);
//                32 floats per subject (the rest zero).
//   MLP          32 -> 64 -> 64 -> 1, ReLU, bf16 weights and activations,
//                fp32 accumulation on MFMA (v_mfma_f32_16x16x32_bf16):
//                k_mlp_forward, k_mlp_backward (weight gradients reduce over
//                the batch rows with the same MFMA), k_adam (fp32 master
//                weights, bf16 copies in both layouts the kernels read).
//
// MFMA fragment layout (16x16x32, wave64): lane l holds A[l%16][8*(l/16)+v]
// and B[8*(l/16)+v][l%16] (v = 0..7) and receives D[4*(l/16)+v][l%16]
// (v = 0..3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "bpx.h"

namespace {

constexpr int kF = BPX_CM_FEATURES;  // 32 input features
constexpr int kH = 64;               // hidden width
constexpr int kO = 16;               // output tile width (column 0 is the prediction)
constexpr int kRows = 32;            // rows per wave tile
constexpr int kWaves = 4;            // waves per workgroup

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Parameter block (fp32 master copy and gradients share it):
//   W1 [kF][kH], b1 [kH], W2 [kH][kH], b2 [kH], W3 [kH][kO], b3 [kO]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kF * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kO, kParams = kOffB3 + kO;
static_assert(kParams == BPX_CM_PARAMS, "parameter count");

// bf16 weight copies: forward layout (transposed, [out][in]) and backward
// layout ([in][out]) of each matrix
struct BfWeights {
    uint16_t w1t[kH * kF], w2t[kH * kH], w3t[kO * kH];  // forward: B[k][n] read as Wt[n][k..k+7]
    uint16_t w2[kH * kH], w3[kH * kO];                  // backward: B[k=out][n=in] read as W[in][out..out+7]
};

__host__ __device__ inline uint16_t to_bf16(float f) {  // round to nearest even (finite inputs)
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ inline bf16x8 frag(const uint16_t* p) {  // 8 consecutive bf16 (16-byte aligned)
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u16x8*>(p));
}
__device__ inline bf16x8 frag_strided(const uint16_t* p, int stride) {
    u16x8 v;
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = p[i * stride];
    return __builtin_bit_cast(bf16x8, v);
}
__device__ inline f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// ---- features ------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_features(const uint8_t* residues, const uint64_t* soff,
                                                  const uint32_t* slen, uint32_t n, const uint16_t* word_counts,
                                                  int word, float* out) {
    __shared__ uint32_t hist[kWaves * 32];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t* h = hist + wave * 32;
    for (uint64_t sid = (uint64_t)blockIdx.x * kWaves + wave; sid < n; sid += (uint64_t)gridDim.x * kWaves) {
        if (lane < 32) h[lane] = 0;
        __builtin_amdgcn_wave_barrier();
        const uint8_t* S = residues + soff[sid];
        const int32_t L = (int32_t)slen[sid];
        float hits = 0.f;
        for (int32_t p = (int32_t)lane; p < L; p += 64) {
            const uint32_t c = S[p];
            atomicAdd(&h[c & 31u], 1u);
            if (p + word <= L) {
                uint32_t w = 0;
                for (int i = 0; i < word; ++i) w = (w << 5) | S[p + i];
                hits += (float)word_counts[w];
            }
        }
        for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off, 64);
        __builtin_amdgcn_wave_barrier();
        // ncbistdaa codes of the 20 standard residues, in ARNDCQEGHILKMFPSTWYV order
        const uint8_t std20[20] = {1, 16, 13, 4, 3, 15, 5, 7, 8, 9, 11, 10, 12, 6, 14, 17, 18, 20, 22, 19};
        float* o = out + sid * kF;
        if (lane < kF) {
            float v = 0.f;
            if (lane == 0) v = logf(1.f + (float)L);
            else if (lane <= 20) v = L > 0 ? (float)h[std20[lane - 1]] / (float)L : 0.f;
            else if (lane == 21) v = logf(1.f + hits);
            o[lane] = v;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- MLP -----------------------------------------------------------------------

// One 32-row tile of a wave: 2 row-subtiles of 16.  X, H1, H2 tiles are bf16
// in LDS ([row][col]).
struct alignas(16) Tile {
    uint16_t x[kRows * kF];
    uint16_t h1[kRows * kH];
    uint16_t h2[kRows * kH];
    uint16_t dz[kRows * kH];   // backward: bf16 dZ of the current layer
    float y[kRows];
};

// Forward through a dense layer: out[16 x 64] += A[16 x K] * W (Wt [64][K]),
// for row-subtile rs of the tile; returns D fragments of the 4 column tiles.
template <int K>
__device__ inline void dense16(const uint16_t* a, int lda, const uint16_t* wt, uint32_t lane, f32x4 acc[4]) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kt = 0; kt < K / 32; ++kt) {
        const bf16x8 A = frag(a + (lane & 15u) * lda + kt * 32 + 8 * (lane >> 4));
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 B = frag(wt + (nt * 16 + (lane & 15u)) * K + kt * 32 + 8 * (lane >> 4));
            acc[nt] = mfma(A, B, acc[nt]);
        }
    }
}

// bias + ReLU of D fragments into a bf16 [16 x 64] LDS tile (row-subtile base)
__device__ inline void relu_store(const f32x4 acc[4], const float* bias, uint16_t* out, uint32_t lane) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = 4 * (lane >> 4) + v, col = nt * 16 + (lane & 15u);
            out[row * kH + col] = to_bf16(fmaxf(acc[nt][v] + bias[col], 0.f));
        }
}

// forward of one tile (rows r0 .. r0+31, rows >= n are zero inputs)
__device__ inline void forward_tile(Tile& t, const float* X, uint32_t n, uint32_t r0, const BfWeights& w,
                                    const float* prm, uint32_t lane) {
    for (uint32_t i = lane; i < kRows * kF; i += 64) {
        const uint32_t r = r0 + i / kF;
        t.x[i] = r < n ? to_bf16(X[(size_t)r * kF + i % kF]) : (uint16_t)0;
    }
    __builtin_amdgcn_wave_barrier();
    f32x4 acc[4];
    for (int rs = 0; rs < 2; ++rs) {
        dense16<kF>(t.x + rs * 16 * kF, kF, w.w1t, lane, acc);
        relu_store(acc, prm + kOffB1, t.h1 + rs * 16 * kH, lane);
    }
    __builtin_amdgcn_wave_barrier();
    for (int rs = 0; rs < 2; ++rs) {
        dense16<kH>(t.h1 + rs * 16 * kH, kH, w.w2t, lane, acc);
        relu_store(acc, prm + kOffB2, t.h2 + rs * 16 * kH, lane);
    }
    __builtin_amdgcn_wave_barrier();
    for (int rs = 0; rs < 2; ++rs) {
        f32x4 o = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 2; ++kt) {
            const bf16x8 A = frag(t.h2 + (rs * 16 + (lane & 15u)) * kH + kt * 32 + 8 * (lane >> 4));
            const bf16x8 B = frag(w.w3t + (lane & 15u) * kH + kt * 32 + 8 * (lane >> 4));
            o = mfma(A, B, o);
        }
        if ((lane & 15u) == 0)
#pragma unroll
            for (int v = 0; v < 4; ++v) t.y[rs * 16 + 4 * (lane >> 4) + v] = o[v] + prm[kOffB3];
    }
    __builtin_amdgcn_wave_barrier();
}

__global__ __launch_bounds__(64 * kWaves) void k_mlp_forward(const float* X, uint32_t n, const BfWeights* w,
                                                             const float* prm, float* y) {
    __shared__ Tile tiles[kWaves];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    Tile& t = tiles[wave];
    for (uint32_t r0 = (blockIdx.x * kWaves + wave) * kRows; r0 < n; r0 += gridDim.x * kWaves * kRows) {
        forward_tile(t, X, n, r0, *w, prm, lane);
        if (lane < kRows && r0 + lane < n) y[r0 + lane] = t.y[lane];
        __builtin_amdgcn_wave_barrier();
    }
}

// Weight gradient of one layer over the 32 rows of a tile:
// G[m][n] += sum_r A[r][m] * D[r][n]  (A: [32][M] bf16, D: [32][N] bf16),
// M = 16 * MT, N = 16 * NT; accumulators g[mt][nt].
template <int MT, int NT>
__device__ inline void wgrad(const uint16_t* a, int lda, const uint16_t* d, int ldd, uint32_t lane,
                             f32x4 (&g)[MT][NT]) {
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
        // A operand: rows m of G = columns of a; K = the 32 tile rows
        const bf16x8 A = frag_strided(a + (8 * (lane >> 4)) * lda + mt * 16 + (lane & 15u), lda);
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const bf16x8 B = frag_strided(d + (8 * (lane >> 4)) * ldd + nt * 16 + (lane & 15u), ldd);
            g[mt][nt] = mfma(A, B, g[mt][nt]);
        }
    }
}

// Backward of a mini-batch: every row's loss gradient, the hidden gradients
// and the weight/bias gradients, added into grad (kParams floats) and
// loss[0] (sum of squared errors).  scale = 2 / batch (mean squared error).
__global__ __launch_bounds__(64 * kWaves) void k_mlp_backward(const float* X, const float* target, uint32_t n,
                                                              float scale, const BfWeights* wp, const float* prm,
                                                              float* grad, float* loss) {
    __shared__ Tile tiles[kWaves];
    __shared__ float gsh[kParams];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < kParams; i += blockDim.x) gsh[i] = 0.f;
    __syncthreads();
    Tile& t = tiles[wave];
    const BfWeights& w = *wp;
    f32x4 g1[2][4], g2[4][4], g3[4][1];
    for (auto& a : g1)
        for (auto& b : a) b = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto& a : g2)
        for (auto& b : a) b = f32x4{0.f, 0.f, 0.f, 0.f};
    for (auto& a : g3)
        for (auto& b : a) b = f32x4{0.f, 0.f, 0.f, 0.f};
    float sq = 0.f;
    for (uint32_t r0 = (blockIdx.x * kWaves + wave) * kRows; r0 < n; r0 += gridDim.x * kWaves * kRows) {
        forward_tile(t, X, n, r0, w, prm, lane);
        // output layer: dZ3[r][0] = scale * (y - target); other columns zero
        float dy = 0.f;
        if (lane < kRows && r0 + lane < n) {
            const float e = t.y[lane] - target[r0 + lane];
            sq += e * e;
            dy = scale * e;
        }
        float db3 = dy;
        for (int off = 32; off > 0; off >>= 1) db3 += __shfl_xor(db3, off, 64);
        if (lane == 0) atomicAdd(&gsh[kOffB3], db3);
        for (uint32_t i = lane; i < kRows * kO; i += 64) t.dz[i] = 0;
        __builtin_amdgcn_wave_barrier();
        if (lane < kRows) t.dz[lane * kO] = to_bf16(dy);
        __builtin_amdgcn_wave_barrier();
        wgrad<4, 1>(t.h2, kH, t.dz, kO, lane, g3);
        // dH2 = dZ3 * W3^T (K = 16 outputs, zero-padded to 32), masked by ReLU'
        float dz2[2][4][4];
        for (int rs = 0; rs < 2; ++rs) {
            const uint32_t kb = 8 * (lane >> 4);
            u16x8 av;
#pragma unroll
            for (int v = 0; v < 8; ++v) av[v] = kb + v < kO ? t.dz[(rs * 16 + (lane & 15u)) * kO + kb + v] : 0;
            const bf16x8 A = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                u16x8 bv;
                const uint32_t unit = nt * 16 + (lane & 15u);
#pragma unroll
                for (int v = 0; v < 8; ++v) bv[v] = kb + v < kO ? w.w3[unit * kO + kb + v] : 0;
                const f32x4 d = mfma(A, __builtin_bit_cast(bf16x8, bv), f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = rs * 16 + 4 * (lane >> 4) + v;
                    dz2[rs][nt][v] = t.h2[row * kH + unit] != 0 ? d[v] : 0.f;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // bias 2 and bf16 dZ2 tile
        for (int nt = 0; nt < 4; ++nt) {
            float s = 0.f;
            for (int rs = 0; rs < 2; ++rs)
                for (int v = 0; v < 4; ++v) {
                    const int row = rs * 16 + 4 * (lane >> 4) + v;
                    const float val = r0 + row < n ? dz2[rs][nt][v] : 0.f;
                    s += val;
                    t.dz[row * kH + nt * 16 + (lane & 15u)] = to_bf16(val);
                }
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (lane < 16) atomicAdd(&gsh[kOffB2 + nt * 16 + lane], s);
        }
        __builtin_amdgcn_wave_barrier();
        wgrad<4, 4>(t.h1, kH, t.dz, kH, lane, g2);
        // dH1 = dZ2 * W2^T, masked by ReLU'
        float dz1[2][4][4];
        for (int rs = 0; rs < 2; ++rs)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 d = f32x4{0.f, 0.f, 0.f, 0.f};
                const uint32_t unit = nt * 16 + (lane & 15u);
#pragma unroll
                for (int kt = 0; kt < 2; ++kt) {
                    const bf16x8 A = frag(t.dz + (rs * 16 + (lane & 15u)) * kH + kt * 32 + 8 * (lane >> 4));
                    const bf16x8 B = frag(w.w2 + unit * kH + kt * 32 + 8 * (lane >> 4));
                    d = mfma(A, B, d);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int row = rs * 16 + 4 * (lane >> 4) + v;
                    dz1[rs][nt][v] = t.h1[row * kH + unit] != 0 ? d[v] : 0.f;
                }
            }
        __builtin_amdgcn_wave_barrier();
        for (int nt = 0; nt < 4; ++nt) {
            float s = 0.f;
            for (int rs = 0; rs < 2; ++rs)
                for (int v = 0; v < 4; ++v) {
                    const int row = rs * 16 + 4 * (lane >> 4) + v;
                    const float val = r0 + row < n ? dz1[rs][nt][v] : 0.f;
                    s += val;
                    t.dz[row * kH + nt * 16 + (lane & 15u)] = to_bf16(val);
                }
            s += __shfl_xor(s, 16, 64);
            s += __shfl_xor(s, 32, 64);
            if (lane < 16) atomicAdd(&gsh[kOffB1 + nt * 16 + lane], s);
        }
        __builtin_amdgcn_wave_barrier();
        wgrad<2, 4>(t.x, kF, t.dz, kH, lane, g1);
        __builtin_amdgcn_wave_barrier();
    }
    // weight gradients: D fragments -> workgroup sums -> global
    auto put = [&](int off, int ld, int mt, int nt, const f32x4& g) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = mt * 16 + 4 * (lane >> 4) + v, nn = nt * 16 + (lane & 15u);
            atomicAdd(&gsh[off + m * ld + nn], g[v]);
        }
    };
    for (int mt = 0; mt < 2; ++mt)
        for (int nt = 0; nt < 4; ++nt) put(kOffW1, kH, mt, nt, g1[mt][nt]);
    for (int mt = 0; mt < 4; ++mt)
        for (int nt = 0; nt < 4; ++nt) put(kOffW2, kH, mt, nt, g2[mt][nt]);
    for (int mt = 0; mt < 4; ++mt) put(kOffW3, kO, mt, 0, g3[mt][0]);
    for (int off = 32; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
    if (lane == 0) atomicAdd(loss, sq);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kParams; i += blockDim.x)
        if (gsh[i] != 0.f) atomicAdd(&grad[i], gsh[i]);
}

// bf16 copies of the fp32 parameters in both layouts
__device__ inline void convert_one(const float* prm, BfWeights* w, uint32_t i) {
    if (i < kF * kH) {  // W1[k][n]
        const uint32_t k = i / kH, nn = i % kH;
        w->w1t[nn * kF + k] = to_bf16(prm[kOffW1 + i]);
    } else if (i < kF * kH + kH * kH) {  // W2[k][n]
        const uint32_t j = i - kF * kH, k = j / kH, nn = j % kH;
        const uint16_t b = to_bf16(prm[kOffW2 + j]);
        w->w2t[nn * kH + k] = b;
        w->w2[k * kH + nn] = b;  // backward reads W2[in][out..]: row = in = k
    } else if (i < kF * kH + kH * kH + kH * kO) {  // W3[k][n]
        const uint32_t j = i - kF * kH - kH * kH, k = j / kO, nn = j % kO;
        const uint16_t b = to_bf16(prm[kOffW3 + j]);
        w->w3t[nn * kH + k] = b;
        w->w3[k * kO + nn] = b;
    }
}

__global__ void k_convert(const float* prm, BfWeights* w) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kF * kH + kH * kH + kH * kO;
         i += gridDim.x * blockDim.x)
        convert_one(prm, w, i);
}

// Adam on the fp32 parameters (bias-corrected step size lr_t); clears the
// gradient for the next step.  k_convert then refreshes the bf16 copies.
__global__ void k_adam(float* prm, float* grad, float* m, float* v, float lr_t, float b1, float b2, float eps) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < kParams; i += gridDim.x * blockDim.x) {
        const float g = grad[i];
        const float mi = b1 * m[i] + (1.f - b1) * g;
        const float vi = b2 * v[i] + (1.f - b2) * g * g;
        m[i] = mi;
        v[i] = vi;
        prm[i] -= lr_t * mi / (sqrtf(vi) + eps);
        grad[i] = 0.f;
    }
}

__global__ void k_zero(float* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0.f;
}

thread_local std::string g_cm_err;

}  // namespace

// ---- host side -----------------------------------------------------------------

struct bpx_costmodel {
    int device = 0;
    hipStream_t stream = nullptr;
    float* d_prm = nullptr;   // kParams
    float* d_grad = nullptr;  // kParams
    float* d_m = nullptr;
    float* d_v = nullptr;
    float* d_loss = nullptr;
    BfWeights* d_w = nullptr;
    int64_t steps = 0;
    int num_cus = 1;
};

namespace {

#define CM_TRY(expr)                                       \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) {                            \
            g_cm_err = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return BPX_E_HIP;                              \
        }                                                  \
    } while (0)

int grid_for(const bpx_costmodel* cm, uint32_t n) {
    const uint32_t tiles = (n + kRows * kWaves - 1) / (kRows * kWaves);
    return (int)std::max<uint32_t>(1, std::min<uint32_t>(tiles, (uint32_t)cm->num_cus * 4));
}

// device buffers of X (n x kF fp32) and optional targets
struct DevData {
    float* X = nullptr;
    float* y = nullptr;
    ~DevData() {
        (void)hipFree(X);
        (void)hipFree(y);
    }
};

int upload(bpx_costmodel* cm, const float* X, const float* y, uint32_t n, DevData* d) {
    CM_TRY(hipMalloc(&d->X, std::max<size_t>(n, 1) * kF * 4));
    CM_TRY(hipMemcpyAsync(d->X, X, (size_t)n * kF * 4, hipMemcpyHostToDevice, cm->stream));
    if (y) {
        CM_TRY(hipMalloc(&d->y, std::max<size_t>(n, 1) * 4));
        CM_TRY(hipMemcpyAsync(d->y, y, (size_t)n * 4, hipMemcpyHostToDevice, cm->stream));
    }
    return BPX_OK;
}

}  // namespace

extern "C" {

const char* bpx_costmodel_last_error(void) { return g_cm_err.c_str(); }

int bpx_costmodel_create(int device, uint64_t seed, bpx_costmodel** out) {
    if (!out) return BPX_E_ARG;
    *out = nullptr;
    CM_TRY(hipSetDevice(device));
    auto* cm = new (std::nothrow) bpx_costmodel;
    if (!cm) return BPX_E_NOMEM;
    cm->device = device;
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0)
        cm->num_cus = v;
    if (hipStreamCreateWithFlags(&cm->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&cm->d_prm, kParams * 4) != hipSuccess || hipMalloc(&cm->d_grad, kParams * 4) != hipSuccess ||
        hipMalloc(&cm->d_m, kParams * 4) != hipSuccess || hipMalloc(&cm->d_v, kParams * 4) != hipSuccess ||
        hipMalloc(&cm->d_loss, 16) != hipSuccess || hipMalloc(&cm->d_w, sizeof(BfWeights)) != hipSuccess) {
        bpx_costmodel_destroy(cm);
        g_cm_err = "allocation failed";
        return BPX_E_NOMEM;
    }
    // He-uniform weights, zero biases (seeded, host side)
    std::vector<float> p(kParams, 0.f);
    std::mt19937_64 rng(seed);
    auto init = [&](int off, int fan_in, int count) {
        std::uniform_real_distribution<float> u(-std::sqrt(6.f / fan_in), std::sqrt(6.f / fan_in));
        for (int i = 0; i < count; ++i) p[off + i] = u(rng);
    };
    init(kOffW1, kF, kF * kH);
    init(kOffW2, kH, kH * kH);
    init(kOffW3, kH, kH * kO);
    const int rc = bpx_costmodel_params(cm, p.data(), 1);
    if (rc) {
        bpx_costmodel_destroy(cm);
        return rc;
    }
    *out = cm;
    return BPX_OK;
}

int bpx_costmodel_params(bpx_costmodel* cm, float* params, int set) {
    if (!cm || !params) return BPX_E_ARG;
    CM_TRY(hipSetDevice(cm->device));
    if (set) {
        CM_TRY(hipMemcpyAsync(cm->d_prm, params, kParams * 4, hipMemcpyHostToDevice, cm->stream));
        hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, cm->stream, cm->d_grad, (uint32_t)kParams);
        hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, cm->stream, cm->d_m, (uint32_t)kParams);
        hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, cm->stream, cm->d_v, (uint32_t)kParams);
        hipLaunchKernelGGL(k_convert, dim3(16), dim3(256), 0, cm->stream, cm->d_prm, cm->d_w);
        CM_TRY(hipGetLastError());
        cm->steps = 0;
    } else {
        CM_TRY(hipMemcpyAsync(params, cm->d_prm, kParams * 4, hipMemcpyDeviceToHost, cm->stream));
    }
    CM_TRY(hipStreamSynchronize(cm->stream));
    return BPX_OK;
}

int bpx_costmodel_predict(bpx_costmodel* cm, const float* X, uint32_t n, float* y) {
    if (!cm || (n && (!X || !y))) return BPX_E_ARG;
    if (n == 0) return BPX_OK;
    CM_TRY(hipSetDevice(cm->device));
    DevData d;
    int rc = upload(cm, X, nullptr, n, &d);
    if (rc) return rc;
    CM_TRY(hipMalloc(&d.y, (size_t)n * 4));
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(cm, n)), dim3(64 * kWaves), 0, cm->stream, d.X, n, cm->d_w,
                       cm->d_prm, d.y);
    CM_TRY(hipGetLastError());
    CM_TRY(hipMemcpyAsync(y, d.y, (size_t)n * 4, hipMemcpyDeviceToHost, cm->stream));
    CM_TRY(hipStreamSynchronize(cm->stream));
    return BPX_OK;
}

int bpx_costmodel_grad(bpx_costmodel* cm, const float* X, const float* target, uint32_t n, float* loss,
                       float* grad) {
    if (!cm || !X || !target || !grad || n == 0) return BPX_E_ARG;
    CM_TRY(hipSetDevice(cm->device));
    DevData d;
    int rc = upload(cm, X, target, n, &d);
    if (rc) return rc;
    hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, cm->stream, cm->d_grad, (uint32_t)kParams);
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, cm->stream, cm->d_loss, 4u);
    hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(cm, n)), dim3(64 * kWaves), 0, cm->stream, d.X, d.y, n,
                       2.f / (float)n, cm->d_w, cm->d_prm, cm->d_grad, cm->d_loss);
    CM_TRY(hipGetLastError());
    float l = 0.f;
    CM_TRY(hipMemcpyAsync(grad, cm->d_grad, kParams * 4, hipMemcpyDeviceToHost, cm->stream));
    CM_TRY(hipMemcpyAsync(&l, cm->d_loss, 4, hipMemcpyDeviceToHost, cm->stream));
    CM_TRY(hipStreamSynchronize(cm->stream));
    hipLaunchKernelGGL(k_zero, dim3(8), dim3(256), 0, cm->stream, cm->d_grad, (uint32_t)kParams);
    CM_TRY(hipStreamSynchronize(cm->stream));
    if (loss) *loss = l / (float)n;
    return BPX_OK;
}

int bpx_costmodel_train(bpx_costmodel* cm, const float* X, const float* target, uint32_t n, int32_t epochs,
                        int32_t batch, float lr, uint64_t seed, float* epoch_loss) {
    if (!cm || !X || !target || n == 0 || epochs <= 0 || batch <= 0 || !(lr > 0.f)) return BPX_E_ARG;
    CM_TRY(hipSetDevice(cm->device));
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937_64 rng(seed);
    std::vector<float> Xp((size_t)n * kF), yp(n);
    DevData d;
    CM_TRY(hipMalloc(&d.X, (size_t)n * kF * 4));
    CM_TRY(hipMalloc(&d.y, (size_t)n * 4));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int32_t ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        for (uint32_t i = 0; i < n; ++i) {
            std::memcpy(&Xp[(size_t)i * kF], &X[(size_t)perm[i] * kF], kF * 4);
            yp[i] = target[perm[i]];
        }
        CM_TRY(hipMemcpyAsync(d.X, Xp.data(), (size_t)n * kF * 4, hipMemcpyHostToDevice, cm->stream));
        CM_TRY(hipMemcpyAsync(d.y, yp.data(), (size_t)n * 4, hipMemcpyHostToDevice, cm->stream));
        hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, cm->stream, cm->d_loss, 4u);
        for (uint32_t b0 = 0; b0 < n; b0 += (uint32_t)batch) {
            const uint32_t nb = std::min<uint32_t>((uint32_t)batch, n - b0);
            hipLaunchKernelGGL(k_mlp_backward, dim3(grid_for(cm, nb)), dim3(64 * kWaves), 0, cm->stream,
                               d.X + (size_t)b0 * kF, d.y + b0, nb, 2.f / (float)nb, cm->d_w, cm->d_prm,
                               cm->d_grad, cm->d_loss);
            ++cm->steps;
            const float lr_t = lr * std::sqrt(1.f - std::pow(b2, (float)cm->steps)) /
                               (1.f - std::pow(b1, (float)cm->steps));
            hipLaunchKernelGGL(k_adam, dim3(8), dim3(256), 0, cm->stream, cm->d_prm, cm->d_grad, cm->d_m, cm->d_v,
                               lr_t, b1, b2, eps);
            hipLaunchKernelGGL(k_convert, dim3(16), dim3(256), 0, cm->stream, cm->d_prm, cm->d_w);
        }
        CM_TRY(hipGetLastError());
        float l = 0.f;
        CM_TRY(hipMemcpyAsync(&l, cm->d_loss, 4, hipMemcpyDeviceToHost, cm->stream));
        CM_TRY(hipStreamSynchronize(cm->stream));
        if (epoch_loss) epoch_loss[ep] = l / (float)n;
    }
    return BPX_OK;
}

int bpx_costmodel_destroy(bpx_costmodel* cm) {
    if (!cm) return BPX_OK;
    (void)hipSetDevice(cm->device);
    if (cm->stream) (void)hipStreamSynchronize(cm->stream);
    (void)hipFree(cm->d_prm);
    (void)hipFree(cm->d_grad);
    (void)hipFree(cm->d_m);
    (void)hipFree(cm->d_v);
    (void)hipFree(cm->d_loss);
    (void)hipFree(cm->d_w);
    if (cm->stream) (void)hipStreamDestroy(cm->stream);
    delete cm;
    return BPX_OK;
}

}  // extern "C"

// features of a database (called from runtime.hip, which owns the arena)
namespace bpx {
hipError_t launch_features(const uint8_t* residues, const uint64_t* soff, const uint32_t* slen, uint32_t n,
                           const uint16_t* word_counts, int word, float* out, int num_cus, hipStream_t stream) {
    const uint32_t blocks = std::max<uint32_t>(1, std::min<uint32_t>((n + kWaves - 1) / kWaves, num_cus * 8u));
    hipLaunchKernelGGL(k_features, dim3(blocks), dim3(64 * kWaves), 0, stream, residues, soff, slen, n, word_counts,
                       word, out);
    return hipGetLastError();
}
}  // namespace bpx
