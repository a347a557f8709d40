// k_mlp.hip -- BF16 MLP surrogate of the light curve: training and inference
// on the matrix cores of gfx950.
//
// Layers: y = W_L silu(... silu(W_1 x + b_1) ...) + b_L, weights row-major
// [fan_in x fan_out].  Master weights and Adam moments are FP32; the GEMMs
// read BF16 copies and accumulate in FP32 with v_mfma_f32_32x32x16_bf16.
//
// One GEMM kernel serves the forward pass (X W), the weight gradients
// (X^T G) and the input gradients (G W^T): the operands can be read
// transposed while they are staged in LDS, and the epilogue adds the bias,
// keeps the pre-activation and applies SiLU.  A training step is
// gather -> forward -> MSE gradient -> backward -> Adam, all on one stream.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lcurve_amd.h"
#include "lc_host.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

struct MlpError : std::runtime_error {
    int code;
    MlpError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define MLPCHK(call)                                                                                  \
    do {                                                                                              \
        hipError_t e_ = (call);                                                                       \
        if (e_ != hipSuccess) throw MlpError(LC_E_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------ numerics
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (NaN kept quiet)
__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float sigmoid(float z) { return 1.f / (1.f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
    const float s = sigmoid(z);
    return s * (1.f + z * (1.f - s));
}

// ------------------------------------------------------------ GEMM
// C[M,N] = op(A)[M,K] op(B)[K,N].  A is stored [M,K] (TA false) or [K,M]
// (TA true) with leading dimension lda; B is stored [K,N] (TB false) or
// [N,K] (TB true).  Workgroup tile 64x64, four waves of 32x32, K staged in
// steps of 32 through LDS as A[m][k] and B[n][k] so that every MFMA operand
// is one 16-byte LDS read.  Each thread moves 8 consecutive elements of A
// and of B per step: one 16-byte load when they are in range and aligned
// (vec), element loads with zero fill at the edges.  The next step's
// elements are loaded into registers while the current step's MFMAs run.
// blockIdx.z splits K (kchunk per split): split s writes its partial
// product to Cf + s * M * ldc, summed afterwards by k_sum_splits.
constexpr int TM = 64, TN = 64, TK = 32, LDK = TK + 8;

enum { EPI_BIAS = 1, EPI_SILU = 2 };

struct Seg {
    uint4 v;  // 8 bf16
};

// 8 consecutive elements p[0..7] (valid: how many are in range)
__device__ __forceinline__ Seg load8(const uint16_t *p, int valid, bool vec) {
    Seg s;
    if (vec && valid >= 8) {
        s.v = *reinterpret_cast<const uint4 *>(p);
    } else {
        uint16_t e[8];
#pragma unroll
        for (int i = 0; i < 8; i++) e[i] = i < valid ? p[i] : (uint16_t)0;
        s.v.x = e[0] | ((uint32_t)e[1] << 16);
        s.v.y = e[2] | ((uint32_t)e[3] << 16);
        s.v.z = e[4] | ((uint32_t)e[5] << 16);
        s.v.w = e[6] | ((uint32_t)e[7] << 16);
    }
    return s;
}

__device__ __forceinline__ uint16_t seg_at(const Seg &s, int i) {
    const uint32_t w = i < 2 ? s.v.x : i < 4 ? s.v.y : i < 6 ? s.v.z : s.v.w;
    return (uint16_t)(w >> (16 * (i & 1)));
}

template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, int kchunk, const uint16_t *__restrict__ A, int lda,
                                              bool avec, const uint16_t *__restrict__ B, int ldb, bool bvec,
                                              const float *__restrict__ bias, float *__restrict__ Z,
                                              float *__restrict__ Cf, uint16_t *__restrict__ Cb, int ldc) {
    __shared__ __attribute__((aligned(16))) __bf16 As[TM * LDK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[TN * LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    f32x16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.f;

    // this thread's segments: (row, first column) in the stored layouts
    const int ar = TA ? (t >> 3) : (t >> 2), ac = TA ? (t & 7) * 8 : (t & 3) * 8;
    const int br = TB ? (t >> 2) : (t >> 3), bc = TB ? (t & 3) * 8 : (t & 7) * 8;
    auto fetch = [&](int k0, Seg &sa, Seg &sb) {
        if (!TA) {  // row m0 + ar, columns k0 + ac .. + 8
            const int m = m0 + ar, k = k0 + ac;
            sa = load8(A + (size_t)m * lda + k, m < M ? min(8, ke - k) : 0, avec);
        } else {    // row k0 + ar of [K][M], columns m0 + ac .. + 8
            const int k = k0 + ar, m = m0 + ac;
            sa = load8(A + (size_t)k * lda + m, k < ke ? min(8, M - m) : 0, avec);
        }
        if (!TB) {  // row k0 + br of [K][N], columns n0 + bc .. + 8
            const int k = k0 + br, n = n0 + bc;
            sb = load8(B + (size_t)k * ldb + n, k < ke ? min(8, N - n) : 0, bvec);
        } else {    // row n0 + br of [N][K], columns k0 + bc .. + 8
            const int n = n0 + br, k = k0 + bc;
            sb = load8(B + (size_t)n * ldb + k, n < N ? min(8, ke - k) : 0, bvec);
        }
    };
    auto stage = [&](const Seg &sa, const Seg &sb) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            if (!TA)
                As[ar * LDK + ac + i] = __builtin_bit_cast(__bf16, seg_at(sa, i));
            else
                As[(ac + i) * LDK + ar] = __builtin_bit_cast(__bf16, seg_at(sa, i));
            if (!TB)
                Bs[(bc + i) * LDK + br] = __builtin_bit_cast(__bf16, seg_at(sb, i));
            else
                Bs[br * LDK + bc + i] = __builtin_bit_cast(__bf16, seg_at(sb, i));
        }
    };

    Seg sa, sb;
    if (kb < ke) fetch(kb, sa, sb);
    for (int k0 = kb; k0 < ke; k0 += TK) {
        stage(sa, sb);
        __syncthreads();
        if (k0 + TK < ke) fetch(k0 + TK, sa, sb);  // in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < TK; kk += 16) {
            // lane l: row (l % 32), k block 8 * (l / 32) of this 16-wide step
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&As[(wm + (lane & 31)) * LDK + kk + 8 * (lane >> 5)]);
            const bf16x8 b = *reinterpret_cast<const bf16x8 *>(&Bs[(wn + (lane & 31)) * LDK + kk + 8 * (lane >> 5)]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // accumulator r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
    const int j = n0 + wn + (lane & 31);
    if (j >= N) return;
    const float bj = (EPI & EPI_BIAS) ? bias[j] : 0.f;
    float *Cs = Cf ? Cf + (size_t)blockIdx.z * M * ldc : nullptr;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int i = m0 + wm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
        if (i >= M) continue;
        const float z = acc[r] + bj;
        const size_t o = (size_t)i * ldc + j;
        if (Z) Z[o] = z;
        const float a = (EPI & EPI_SILU) ? silu(z) : z;
        if (Cs) Cs[o] = a;
        if (Cb) Cb[o] = f2bf(a);
    }
}

// ------------------------------------------------------------ GEMM, large
// C[M,N] = A[M,K] B[N,K]^T with both operands K-contiguous (row strides lda,
// ldb), the layout every training GEMM is given: the forward pass reads the
// transposed weight copy, the weight gradients read transposed activations
// and gradients that their producers write, the input gradients read W as
// it is.  Workgroup tile BM x BN x BK (BM, BN = 64 or 128; BK = 64, or 128
// for the 64 x 64 tile, whose K stages are then half as many), four waves of
// BM/2 x BN/2, each a BM/64 x BN/64 block of 32x32x16 MFMAs.  LDS holds two K
// stages of A[m][k] and B[n][k] (rows padded by 16 bytes, so the 16-byte
// operand reads of 8 consecutive rows fall in distinct banks); a thread
// stages 16-byte chunks (8 consecutive k of one row): one global load and
// one LDS write each, zero-filled past M, N or the K range, element loads
// where a chunk is cut by the K range or unaligned.  The next stage's
// chunks are loaded into registers while the current stage's MFMAs run and
// written to the other LDS buffer after them: one barrier per K stage.
// blockIdx.z splits K as in k_gemm.  Epilogue: bias, SiLU (EPI_SILU), or
// the SiLU gradient of the pre-activation Zin (EPI_DSILU); outputs Z
// (pre-activation), Cf (FP32), Cb (BF16) and CbT (BF16, transposed:
// CbT[j * ldt + i], four consecutive rows per 8-byte store).
constexpr int NBK = 64;  // K per split is a multiple of this
enum { EPI_DSILU = 4 };

__device__ __forceinline__ uint4 load_chunk(const uint16_t *p, bool row_ok, int valid, bool vec) {
    if (row_ok && vec && valid >= 8) return *reinterpret_cast<const uint4 *>(p);
    uint16_t e[8];
#pragma unroll
    for (int i = 0; i < 8; i++) e[i] = (row_ok && i < valid) ? p[i] : (uint16_t)0;
    return make_uint4(e[0] | ((uint32_t)e[1] << 16), e[2] | ((uint32_t)e[3] << 16), e[4] | ((uint32_t)e[5] << 16),
                      e[6] | ((uint32_t)e[7] << 16));
}

template <int BM, int BN, int BK, int EPI>
__global__ __launch_bounds__(256, 2) void k_gemm_nt(int M, int N, int K, int kchunk, const uint16_t *__restrict__ A,
                                                    int lda, const uint16_t *__restrict__ B, int ldb, bool vec,
                                                    const float *__restrict__ bias, const float *__restrict__ Zin,
                                                    float *__restrict__ Z, float *__restrict__ Cf,
                                                    uint16_t *__restrict__ Cb, int ldc, uint16_t *__restrict__ CbT,
                                                    int ldt) {
    constexpr int NI = BM / 64, NJ = BN / 64;           // 32x32 tiles per wave
    constexpr int KC = BK / 8, NLDK = BK + 8;             // 16-byte chunks per row; LDS row length
    constexpr int CA = BM * KC / 256, CB = BN * KC / 256;  // chunks per thread and stage
    __shared__ __attribute__((aligned(16))) __bf16 As[2][BM * NLDK];
    __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN * NLDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (w >> 1) * (BM / 2), wn = (w & 1) * (BN / 2);
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    f32x16 acc[NI][NJ];
#pragma unroll
    for (int i = 0; i < NI; i++)
#pragma unroll
        for (int j = 0; j < NJ; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.f;

    // chunk c of a stage: row c / KC, k offset 8 (c % KC); thread t takes
    // chunks t, t + 256, ...
    uint4 ra[CA], rb[CB];
    auto fetch = [&](int k0) {
#pragma unroll
        for (int i = 0; i < CA; i++) {
            const int c = t + 256 * i, row = c / KC, k = k0 + 8 * (c % KC);
            ra[i] = load_chunk(A + (size_t)(m0 + row) * lda + k, m0 + row < M, ke - k, vec);
        }
#pragma unroll
        for (int i = 0; i < CB; i++) {
            const int c = t + 256 * i, row = c / KC, k = k0 + 8 * (c % KC);
            rb[i] = load_chunk(B + (size_t)(n0 + row) * ldb + k, n0 + row < N, ke - k, vec);
        }
    };
    auto stage = [&](int buf) {
#pragma unroll
        for (int i = 0; i < CA; i++) {
            const int c = t + 256 * i;
            *reinterpret_cast<uint4 *>(&As[buf][(c / KC) * NLDK + 8 * (c % KC)]) = ra[i];
        }
#pragma unroll
        for (int i = 0; i < CB; i++) {
            const int c = t + 256 * i;
            *reinterpret_cast<uint4 *>(&Bs[buf][(c / KC) * NLDK + 8 * (c % KC)]) = rb[i];
        }
    };

    if (kb < ke) {
        fetch(kb);
        stage(0);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += BK) {
        const bool more = k0 + BK < ke;
        if (more) fetch(k0 + BK);  // in flight during the MFMAs
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            // lane l: row (l % 32), k block 8 * (l / 32) of this 16-wide step
            const int ko = kk + 8 * (lane >> 5);
            bf16x8 a[NI], b[NJ];
#pragma unroll
            for (int i = 0; i < NI; i++)
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[buf][(wm + 32 * i + (lane & 31)) * NLDK + ko]);
#pragma unroll
            for (int j = 0; j < NJ; j++)
                b[j] = *reinterpret_cast<const bf16x8 *>(&Bs[buf][(wn + 32 * j + (lane & 31)) * NLDK + ko]);
#pragma unroll
            for (int i = 0; i < NI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) stage(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // accumulator r of lane l: row 8 (r / 4) + 4 (l / 32) + r % 4, column l % 32
    float *Cs = Cf ? Cf + (size_t)blockIdx.z * M * ldc : nullptr;
#pragma unroll
    for (int j = 0; j < NJ; j++) {
        const int col = n0 + wn + 32 * j + (lane & 31);
        if (col >= N) continue;
        const float bj = (EPI & EPI_BIAS) ? bias[col] : 0.f;
#pragma unroll
        for (int i = 0; i < NI; i++) {
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int row0 = m0 + wm + 32 * i + 8 * g + 4 * (lane >> 5);
                uint16_t tb[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int row = row0 + q;
                    tb[q] = 0;
                    if (row >= M) continue;
                    const size_t o = (size_t)row * ldc + col;
                    float z = acc[i][j][4 * g + q] + bj;
                    if (EPI & EPI_DSILU) z *= silu_grad(Zin[o]);
                    if (Z) Z[o] = z;
                    const float v = (EPI & EPI_SILU) ? silu(z) : z;
                    if (Cs) Cs[o] = v;
                    const uint16_t h = f2bf(v);
                    if (Cb) Cb[o] = h;
                    tb[q] = h;
                }
                if (CbT) {
                    uint16_t *dst = CbT + (size_t)col * ldt + row0;
                    if (row0 + 4 <= M && (ldt & 3) == 0) {
                        *reinterpret_cast<uint2 *>(dst) =
                            make_uint2(tb[0] | ((uint32_t)tb[1] << 16), tb[2] | ((uint32_t)tb[3] << 16));
                    } else {
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            if (row0 + q < M) dst[q] = tb[q];
                    }
                }
            }
        }
    }
}

// dst[j * ldd + i] = src[i * lds + j] for a rows x cols BF16 matrix, in
// 64 x 64 tiles through LDS (both sides coalesced); blockIdx.z selects one
// of up to 8 matrices of a TransposeSet
struct TransposeSet {
    const uint16_t *src[8];
    uint16_t *dst[8];
    int rows[8], cols[8], ldd[8];  // source rows are cols long; destination rows ldd apart
};
__global__ __launch_bounds__(256) void k_transpose_bf16(TransposeSet ts) {
    __shared__ uint16_t tile[64][65];
    const int z = blockIdx.z;
    const int rows = ts.rows[z], cols = ts.cols[z];
    const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
    if (r0 >= rows || c0 >= cols) return;
    const uint16_t *src = ts.src[z];
    uint16_t *dst = ts.dst[z];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int r = ty; r < 64; r += 4)
        if (r0 + r < rows && c0 + tx < cols) tile[r][tx] = src[(size_t)(r0 + r) * cols + c0 + tx];
    __syncthreads();
    for (int c = ty; c < 64; c += 4)
        if (c0 + c < cols && r0 + tx < rows) dst[(size_t)(c0 + c) * ts.ldd[z] + r0 + tx] = tile[tx][c];
}

// out[o] = sum over splits of part[s * n + o] (in split order)
__global__ void k_sum_splits(size_t n, int splits, const float *__restrict__ part, float *__restrict__ out) {
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        float a = 0.f;
        for (int s = 0; s < splits; s++) a += part[(size_t)s * n + o];
        out[o] = a;
    }
}

// ------------------------------------------------------------ elementwise
__global__ void k_cast_bf16(size_t n, const float *__restrict__ x, uint16_t *__restrict__ y) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        y[i] = f2bf(x[i]);
}

// dZ = dA * silu'(Z)
__global__ void k_silu_bwd(size_t n, const float *__restrict__ dA, const float *__restrict__ Z, float *__restrict__ dZ,
                           uint16_t *__restrict__ dZb) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float g = dA[i] * silu_grad(Z[i]);
        dZ[i] = g;
        dZb[i] = f2bf(g);
    }
}

// db[j] = sum_i G[i][j] in two deterministic stages: CS_PART row chunks
// (blockIdx.y) of partial column sums, then their sum in chunk order
constexpr int CS_PART = 64;
__global__ void k_colsum_part(int M, int N, const float *__restrict__ G, float *__restrict__ part) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const int rows = (M + CS_PART - 1) / CS_PART, i0 = blockIdx.y * rows, i1 = min(M, i0 + rows);
    float s = 0.f;
    for (int i = i0; i < i1; i++) s += G[(size_t)i * N + j];
    part[(size_t)blockIdx.y * N + j] = s;
}

// MSE over M x N: dY = 2 (Y - T) / (M N); the sum of squares of each
// workgroup into loss_part[blockIdx.x] (k_sum_loss adds them in order)
__global__ __launch_bounds__(256) void k_mse(size_t n, float scale, const float *__restrict__ Y,
                                             const float *__restrict__ T, float *__restrict__ dY,
                                             uint16_t *__restrict__ dYb, double *__restrict__ loss) {
    __shared__ double part[256];
    double s = 0.;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        s += (double)d * d;
        if (dY) {
            const float g = scale * d;
            dY[i] = g;
            dYb[i] = f2bf(g);
        }
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss[blockIdx.x] = part[0];
}

// *out = sum of part[0 .. n) in a fixed order (one workgroup)
__global__ __launch_bounds__(256) void k_sum_loss(int n, const double *__restrict__ part, double *__restrict__ out) {
    __shared__ double red[256];
    double s = 0.;
    for (int i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Adam (bias-corrected) on FP32 masters, refreshing the BF16 copies
__global__ void k_adam(size_t n, float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, uint16_t *__restrict__ wb, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[i] = wi;
        if (wb) wb[i] = f2bf(wi);
    }
}

// Minibatch rows of the device data set: idx from a stateless hash of
// (seed, step, row), so no host round trip per step.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

__global__ void k_gather(int B, int nin, int nout, size_t nrows, uint64_t seed, uint64_t step,
                         const float *__restrict__ X, const float *__restrict__ T, uint16_t *__restrict__ Xb,
                         float *__restrict__ Tb) {
    const int r = blockIdx.x;
    if (r >= B) return;
    const size_t src = (size_t)(mix64(seed ^ mix64(step * 0x100000001b3ull + (uint64_t)r)) % nrows);
    for (int c = threadIdx.x; c < nin; c += blockDim.x) Xb[(size_t)r * nin + c] = f2bf(X[src * nin + c]);
    for (int c = threadIdx.x; c < nout; c += blockDim.x) Tb[(size_t)r * nout + c] = T[src * nout + c];
}

// chi**2 of each row's predicted curve against the data, with the optimal
// scale factor s = sum w f m / sum w m^2 (re_scale.cpp:16-43):
// chi**2 = sum w f^2 - (sum w f m)^2 / sum w m^2.  The network's output row
// is the standardised curve at nbin phases k / nbin: v = y * scale + mu is
// the log flux (log_output) or the relative flux itself; the model flux at
// a data time is the linear interpolation of the flux at the phase
// (t - t0) / period (t0 per row: an exact shift).  One workgroup per row.
__global__ __launch_bounds__(256) void k_sur_chisq(int nbin, int log_output, const float *__restrict__ y,
                                                   const float *__restrict__ mu,
                                                   float scale, const double *__restrict__ t0, double period, int n,
                                                   const double *__restrict__ time, const double *__restrict__ flux,
                                                   const double *__restrict__ w, double *__restrict__ chisq) {
    __shared__ double red[3][256];
    const int row = blockIdx.x;
    const float *yr = y + (size_t)row * nbin;
    const double tz = t0[row];
    double sfm = 0., smm = 0., sff = 0.;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double ph = (time[k] - tz) / period;
        ph -= floor(ph);
        const double x = ph * nbin;
        int i0 = (int)floor(x);
        const double f = x - i0;
        i0 = i0 >= nbin ? i0 - nbin : i0;
        const int i1 = i0 + 1 == nbin ? 0 : i0 + 1;
        const double v0 = (double)yr[i0] * scale + mu[i0], v1 = (double)yr[i1] * scale + mu[i1];
        const double m0 = log_output ? exp(v0) : v0, m1 = log_output ? exp(v1) : v1;
        const double mk = m0 * (1.0 - f) + m1 * f;
        const double wk = w[k], fk = flux[k];
        sfm += wk * fk * mk;
        smm += wk * mk * mk;
        sff += wk * fk * fk;
    }
    red[0][threadIdx.x] = sfm;
    red[1][threadIdx.x] = smm;
    red[2][threadIdx.x] = sff;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) chisq[row] = red[1][0] > 0. ? red[2][0] - red[0][0] * red[0][0] / red[1][0] : INFINITY;
}

// y[i][j] += sum_k x[i][k] S[k][j]: the FP32 linear skip of inference
__global__ void k_skip_add(int B, int nin, int nout, const float *__restrict__ x, const float *__restrict__ S,
                           float *__restrict__ y) {
    const size_t n = (size_t)B * nout;
    for (size_t o = blockIdx.x * (size_t)blockDim.x + threadIdx.x; o < n; o += (size_t)gridDim.x * blockDim.x) {
        const size_t i = o / nout, j = o % nout;
        float a = y[o];
        for (int k = 0; k < nin; k++) a = fmaf(x[i * nin + k], S[(size_t)k * nout + j], a);
        y[o] = a;
    }
}

inline int nblocks(size_t n, int bs = 256, int cap = 4096) {
    return (int)std::max<size_t>(1, std::min<size_t>((n + bs - 1) / bs, (size_t)cap));
}

// K per split (a multiple of the K step) and the number of splits it gives
inline int split_chunk(int K, int splits) {
    const int c = (K + std::max(splits, 1) - 1) / std::max(splits, 1);
    return std::max(TK, (c + TK - 1) / TK * TK);
}
inline int split_actual(int K, int splits) { return (K + split_chunk(K, splits) - 1) / split_chunk(K, splits); }

// Launches the GEMM; with splits > 1, Cf must hold splits x M x ldc floats
// (the partial products).  Returns the number of splits used.
template <bool TA, bool TB, int EPI>
int gemm(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
         const float *bias, float *Z, float *Cf, uint16_t *Cb, int ldc, int splits = 1) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    const bool avec = (lda % 8 == 0) && ((uintptr_t)A % 16 == 0);
    const bool bvec = (ldb % 8 == 0) && ((uintptr_t)B % 16 == 0);
    const int kchunk = split_chunk(K, splits);
    const int ns = split_actual(K, splits);
    dim3 grid((N + TN - 1) / TN, (M + TM - 1) / TM, ns);
    hipLaunchKernelGGL((k_gemm<TA, TB, EPI>), grid, dim3(256), 0, s, M, N, K, kchunk, A, lda, avec, B, ldb, bvec,
                       bias, Z, Cf, Cb, ldc);
    return ns;
}

// splits of K that give the grid about 512 workgroups (at least 256
// elements of K per split)
inline int split_count(int M, int N, int K) {
    const int tiles = ((M + TM - 1) / TM) * ((N + TN - 1) / TN);
    return std::max(1, std::min(512 / std::max(tiles, 1), K / 256));
}

// Large-tile GEMM (k_gemm_nt): the largest tile (128 x 128, 128 x 64,
// 64 x 64) that still gives two workgroups per CU (512).  K per split is a
// multiple of the 64-wide K stage.  Returns the number of splits used.
template <int EPI>
int gemm_nt(hipStream_t s, int M, int N, int K, const uint16_t *A, int lda, const uint16_t *B, int ldb,
            const float *bias, const float *Zin, float *Z, float *Cf, uint16_t *Cb, int ldc, uint16_t *CbT, int ldt,
            int splits = 1) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    const bool vec = (lda % 8 == 0) && (ldb % 8 == 0) && ((uintptr_t)A % 16 == 0) && ((uintptr_t)B % 16 == 0);
    const int c = (K + std::max(splits, 1) - 1) / std::max(splits, 1);
    const int kchunk = std::max(NBK, (c + NBK - 1) / NBK * NBK);
    const int ns = (K + kchunk - 1) / kchunk;
    auto tiles = [&](int bm, int bn) { return ((M + bm - 1) / bm) * ((N + bn - 1) / bn) * ns; };
#define LC_GEMM_NT(BM_, BN_, BK_)                                                                                     \
    hipLaunchKernelGGL((k_gemm_nt<BM_, BN_, BK_, EPI>), dim3((N + BN_ - 1) / BN_, (M + BM_ - 1) / BM_, ns), dim3(256), 0, \
                       s, M, N, K, kchunk, A, lda, B, ldb, vec, bias, Zin, Z, Cf, Cb, ldc, CbT, ldt)
    if (tiles(128, 128) >= 512)
        LC_GEMM_NT(128, 128, 64);
    else if (tiles(128, 64) >= 512)
        LC_GEMM_NT(128, 64, 64);
    else
        LC_GEMM_NT(64, 64, 128);
#undef LC_GEMM_NT
    return ns;
}

// splits of K for k_gemm_nt: about 512 workgroups of 64 x 64, at least
// 256 elements of K per split
inline int split_count_nt(int M, int N, int K) {
    const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
    return std::max(1, std::min(512 / std::max(tiles, 1), K / 256));
}

// Batches of BF16 matrix transposes (at most 8 per launch)
struct Transposer {
    TransposeSet ts{};
    int n = 0, rmax = 0, cmax = 0;
    void add(hipStream_t s, const uint16_t *src, uint16_t *dst, int rows, int cols, int ldd = 0) {
        if (rows <= 0 || cols <= 0) return;
        ts.src[n] = src;
        ts.dst[n] = dst;
        ts.rows[n] = rows;
        ts.cols[n] = cols;
        ts.ldd[n] = ldd > 0 ? ldd : rows;
        rmax = std::max(rmax, rows);
        cmax = std::max(cmax, cols);
        if (++n == 8) flush(s);
    }
    void flush(hipStream_t s) {
        if (!n) return;
        dim3 grid((cmax + 63) / 64, (rmax + 63) / 64, n);
        hipLaunchKernelGGL(k_transpose_bf16, grid, dim3(256), 0, s, ts);
        n = rmax = cmax = 0;
    }
};

}  // namespace

// ================================================================== MLP
struct lc_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::vector<int> width;     // width[0] inputs ... width[L] outputs
    int L = 0, max_batch = 0;
    size_t nparam = 0;
    std::vector<size_t> woff, boff;  // offsets of W_l and b_l in the parameter vector
    // device: parameters and optimiser state (one block each, same layout)
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *pb = nullptr;     // BF16 copy of the parameters
    uint16_t *pbT = nullptr;    // the same with each W_l transposed ([fan_out x fan_in] at woff[l])
    // activations for max_batch rows
    uint16_t *xb = nullptr;                  // inputs, BF16
    uint16_t *xbT = nullptr;                 // inputs, BF16, transposed [inputs x rows]
    float *x32 = nullptr;                    // inputs, FP32 (for the skip)
    float *skip = nullptr;                   // optional FP32 linear skip [inputs x outputs]
    std::vector<float *> z;                  // pre-activations per layer (1..L)
    std::vector<uint16_t *> ab;              // BF16 activations per layer (1..L-1)
    std::vector<uint16_t *> abT;             // the same, transposed [width x rows] (training)
    float *y = nullptr, *t = nullptr;        // output and target
    float *ga = nullptr, *gz = nullptr;      // gradient scratch (fp32)
    float *part = nullptr;                   // split-K partial products of the weight gradients
    float *cpart = nullptr;                  // partial column sums of the bias gradients
    uint16_t *gzb = nullptr;                 // gradient, BF16
    uint16_t *gzb2 = nullptr;                // second BF16 gradient buffer (large-tile backward)
    uint16_t *gzbT = nullptr;                // BF16 gradient, transposed [width x rows]
    double *loss = nullptr;                  // the last loss (sum of squares)
    double *loss_part = nullptr;             // per-workgroup sums of k_mse
    // data set on the device
    float *dx = nullptr, *dt = nullptr;
    // likelihood for lc_mlp_chisq: data and output de-standardisation
    double *lt = nullptr, *lf = nullptr, *lw = nullptr, *lt0 = nullptr, *lchi = nullptr;
    float *lmu = nullptr;
    float lscale = 1.f;
    double lperiod = 1.;
    int llog = 1;
    int ln = 0;
    std::vector<void *> lallocs;
    size_t nrows = 0;
    long long adam_t = 0;
    double gemm_flop = 0., train_ms = 0.;
    std::vector<void *> allocs;

    template <class T> T *alloc(size_t n) {
        void *q = nullptr;
        MLPCHK(hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
        allocs.push_back(q);
        return static_cast<T *>(q);
    }
    ~lc_mlp() {
        for (void *q : allocs) (void)hipFree(q);
        for (void *q : lallocs) (void)hipFree(q);
        if (skip) (void)hipFree(skip);
        if (dx) (void)hipFree(dx);
        if (dt) (void)hipFree(dt);
        if (stream) (void)hipStreamDestroy(stream);
    }
    float *W(int l) { return p + woff[l]; }
    uint16_t *Wb(int l) { return pb + woff[l]; }
    uint16_t *WbT(int l) { return pbT + woff[l]; }
    float *bvec(int l) { return p + boff[l]; }

    // Batches of at least this many rows run on the large-tile GEMM
    // (k_gemm_nt); smaller ones on k_gemm, whose 64-row tiles waste less.
    static constexpr int NT_ROWS = 128;

    // the transposed BF16 weights, after every change of pb
    void refresh_wt() {
        Transposer tr;
        for (int l = 1; l <= L; l++) tr.add(stream, Wb(l), WbT(l), width[l - 1], width[l]);
        tr.flush(stream);
    }
    // the transposed BF16 inputs of a training batch
    void transpose_inputs(int B) {
        if (B < NT_ROWS) return;
        Transposer tr;
        tr.add(stream, xb, xbT, B, width[0], max_batch);
        tr.flush(stream);
    }

    // forward of B rows whose BF16 inputs are in xb; output in y.  with_skip:
    // add the linear skip of the FP32 inputs in x32 (inference only; the
    // network is trained on what the skip leaves).  train: keep what the
    // backward pass reads (transposed activations on the large-tile path)
    void forward(int B, bool with_skip = false, bool train = false) {
        hipStream_t s = stream;
        const uint16_t *in = xb;
        const bool nt = B >= NT_ROWS;
        for (int l = 1; l <= L; l++) {
            const int K = width[l - 1], N = width[l];
            if (nt) {
                if (l < L)
                    gemm_nt<EPI_BIAS | EPI_SILU>(s, B, N, K, in, K, WbT(l), K, bvec(l), nullptr, z[l], nullptr, ab[l],
                                                 N, train ? abT[l] : nullptr, max_batch);
                else
                    gemm_nt<EPI_BIAS>(s, B, N, K, in, K, WbT(l), K, bvec(l), nullptr, nullptr, y, nullptr, N, nullptr,
                                      0);
            } else if (l < L) {
                gemm<false, false, EPI_BIAS | EPI_SILU>(s, B, N, K, in, K, Wb(l), N, bvec(l), z[l], nullptr, ab[l], N);
            } else {
                gemm<false, false, EPI_BIAS>(s, B, N, K, in, K, Wb(l), N, bvec(l), nullptr, y, nullptr, N);
            }
            gemm_flop += 2.0 * B * N * K;
            in = ab[l];
        }
        if (with_skip && skip) {
            const size_t n = (size_t)B * width[L];
            hipLaunchKernelGGL(k_skip_add, dim3(nblocks(n)), dim3(256), 0, s, B, width[0], width[L], x32, skip, y);
        }
    }

    // gradients of mean squared error over B rows (targets in t) into g
    void backward(int B) {
        hipStream_t s = stream;
        const int NL = width[L];
        const size_t n = (size_t)B * NL;
        const int nb = nblocks(n);
        hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, n, 2.f / (float)n, y, t, gz, gzb, loss_part);
        hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, s, nb, loss_part, loss);
        if (B >= NT_ROWS) {
            backward_nt(B);
            return;
        }
        for (int l = L; l >= 1; l--) {
            const int K = width[l - 1], N = width[l];
            const uint16_t *in = l == 1 ? xb : ab[l - 1];
            // dW_l = A_{l-1}^T G_l   (M = K, N = N, K = B), K split over the batch
            const int sp = split_count(K, N, B);
            if (sp > 1) {
                const int ns = gemm<true, false, 0>(s, K, N, B, in, K, gzb, N, nullptr, nullptr, part, nullptr, N, sp);
                const size_t nw_ = (size_t)K * N;
                hipLaunchKernelGGL(k_sum_splits, dim3(nblocks(nw_)), dim3(256), 0, s, nw_, ns, part, g + woff[l]);
            } else {
                gemm<true, false, 0>(s, K, N, B, in, K, gzb, N, nullptr, nullptr, g + woff[l], nullptr, N);
            }
            hipLaunchKernelGGL(k_colsum_part, dim3((N + 255) / 256, CS_PART), dim3(256), 0, s, B, N, gz, cpart);
            hipLaunchKernelGGL(k_sum_splits, dim3(nblocks(N)), dim3(256), 0, s, (size_t)N, CS_PART, cpart,
                               g + boff[l]);
            gemm_flop += 2.0 * B * N * K;
            if (l > 1) {
                // dA_{l-1} = G_l W_l^T   (M = B, N = K, K = N), then through SiLU
                gemm<false, true, 0>(s, B, K, N, gzb, N, Wb(l), N, nullptr, nullptr, ga, nullptr, K);
                gemm_flop += 2.0 * B * N * K;
                const size_t nk = (size_t)B * K;
                hipLaunchKernelGGL(k_silu_bwd, dim3(nblocks(nk)), dim3(256), 0, s, nk, ga, z[l - 1], gz, gzb);
            }
        }
    }

    // backward on the large-tile GEMM: dW_l = A_{l-1}^T G_l from the
    // transposed activations and gradients (rows max_batch apart), whose
    // extra last row of ones makes the GEMM's last output row the bias
    // gradient sum_i G_l[i] (b_l follows W_l in the parameter vector);
    // dA_{l-1} = G_l W_l^T with the SiLU gradient applied in its epilogue,
    // which also writes the next layer's transposed gradient (the BF16
    // gradient alternates between two buffers: the GEMM reads one and
    // writes the other)
    void backward_nt(int B) {
        hipStream_t s = stream;
        {
            Transposer tr;
            tr.add(s, gzb, gzbT, B, width[L], max_batch);
            tr.flush(s);
        }
        uint16_t *cur = gzb, *nxt = gzb2;
        for (int l = L; l >= 1; l--) {
            const int K = width[l - 1], N = width[l];
            const uint16_t *inT = l == 1 ? xbT : abT[l - 1];
            const int sp = split_count_nt(K + 1, N, B);
            if (sp > 1) {
                const int ns = gemm_nt<0>(s, K + 1, N, B, inT, max_batch, gzbT, max_batch, nullptr, nullptr, nullptr,
                                          part, nullptr, N, nullptr, 0, sp);
                const size_t nw_ = (size_t)(K + 1) * N;
                hipLaunchKernelGGL(k_sum_splits, dim3(nblocks(nw_)), dim3(256), 0, s, nw_, ns, part, g + woff[l]);
            } else {
                gemm_nt<0>(s, K + 1, N, B, inT, max_batch, gzbT, max_batch, nullptr, nullptr, nullptr, g + woff[l],
                           nullptr, N, nullptr, 0);
            }
            gemm_flop += 2.0 * B * N * K;
            if (l > 1) {
                gemm_nt<EPI_DSILU>(s, B, K, N, cur, N, Wb(l), N, nullptr, z[l - 1], nullptr, gz, nxt, K, gzbT,
                                   max_batch);
                gemm_flop += 2.0 * B * N * K;
                std::swap(cur, nxt);
            }
        }
    }

    void adam(float lr) {
        adam_t++;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float c1 = 1.f - (float)std::pow(0.9, (double)adam_t), c2 = 1.f - (float)std::pow(0.999, (double)adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nblocks(nparam)), dim3(256), 0, stream, nparam, p, g, m, v, pb, lr, b1, b2,
                           eps, c1, c2);
        refresh_wt();
    }

    void upload_inputs(const float *x, int B) {
        const size_t n = (size_t)B * width[0];
        MLPCHK(hipMemcpyAsync(x32, x, n * sizeof(float), hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(k_cast_bf16, dim3(nblocks(n)), dim3(256), 0, stream, n, x32, xb);
    }
};

namespace {

template <class F> lc_status mlp_guarded(lc_mlp *mm, F f) {
    try {
        return (lc_status)f();
    } catch (const MlpError &e) {
        if (mm) mm->err = e.what();
        lch::set_thread_error(e.what());
        return (lc_status)e.code;
    } catch (const std::exception &e) {
        if (mm) mm->err = e.what();
        lch::set_thread_error(e.what());
        return LC_E_DEVICE;
    }
}

}  // namespace

extern "C" {

lc_status lc_mlp_create(int device, const int32_t *widths, int nwidth, int max_batch, lc_mlp **out) {
    if (!out || !widths || nwidth < 2 || max_batch <= 0) return LC_E_ARG;
    *out = nullptr;
    return mlp_guarded(nullptr, [&]() -> int {
        for (int i = 0; i < nwidth; i++)
            if (widths[i] <= 0 || widths[i] > 65536) throw MlpError(LC_E_ARG, "lc_mlp_create: bad layer width");
        int ndev = 0;
        MLPCHK(hipGetDeviceCount(&ndev));
        if (ndev <= 0) throw MlpError(LC_E_DEVICE, "no HIP device available");
        if (device < 0 || device >= ndev) device = 0;
        MLPCHK(hipSetDevice(device));
        std::unique_ptr<lc_mlp> mm(new lc_mlp());
        mm->device = device;
        MLPCHK(hipStreamCreateWithFlags(&mm->stream, hipStreamNonBlocking));
        mm->width.assign(widths, widths + nwidth);
        mm->L = nwidth - 1;
        mm->max_batch = max_batch;
        mm->woff.assign(nwidth, 0);
        mm->boff.assign(nwidth, 0);
        size_t off = 0;
        int wmax = 0;
        for (int l = 1; l <= mm->L; l++) {
            mm->woff[l] = off;
            off += (size_t)widths[l - 1] * widths[l];
            mm->boff[l] = off;
            off += widths[l];
        }
        for (int i = 0; i < nwidth; i++) wmax = std::max(wmax, widths[i]);
        mm->nparam = off;
        mm->p = mm->alloc<float>(off);
        mm->g = mm->alloc<float>(off);
        mm->m = mm->alloc<float>(off);
        mm->v = mm->alloc<float>(off);
        mm->pb = mm->alloc<uint16_t>(off);
        mm->pbT = mm->alloc<uint16_t>(off);
        MLPCHK(hipMemset(mm->p, 0, off * sizeof(float)));
        MLPCHK(hipMemset(mm->g, 0, off * sizeof(float)));
        MLPCHK(hipMemset(mm->m, 0, off * sizeof(float)));
        MLPCHK(hipMemset(mm->v, 0, off * sizeof(float)));
        MLPCHK(hipMemset(mm->pb, 0, off * sizeof(uint16_t)));
        MLPCHK(hipMemset(mm->pbT, 0, off * sizeof(uint16_t)));
        const size_t B = (size_t)max_batch;
        mm->xb = mm->alloc<uint16_t>(B * std::max(widths[0], 1));
        mm->x32 = mm->alloc<float>(B * std::max(widths[0], 1));
        // transposed inputs and activations: width + 1 rows of max_batch,
        // the last one all ones (the bias column of the weight gradients)
        auto ones_row = [&](uint16_t *q, int rows) {
            std::vector<uint16_t> one(B, (uint16_t)0x3f80);
            MLPCHK(hipMemset(q, 0, (size_t)rows * B * sizeof(uint16_t)));
            MLPCHK(hipMemcpy(q + (size_t)rows * B, one.data(), B * sizeof(uint16_t), hipMemcpyHostToDevice));
        };
        mm->xbT = mm->alloc<uint16_t>(B * (std::max(widths[0], 1) + 1));
        ones_row(mm->xbT, widths[0]);
        mm->z.assign(nwidth, nullptr);
        mm->ab.assign(nwidth, nullptr);
        mm->abT.assign(nwidth, nullptr);
        for (int l = 1; l <= mm->L; l++) {
            mm->z[l] = mm->alloc<float>(B * widths[l]);
            if (l < mm->L) {
                mm->ab[l] = mm->alloc<uint16_t>(B * widths[l]);
                mm->abT[l] = mm->alloc<uint16_t>(B * (widths[l] + 1));
                ones_row(mm->abT[l], widths[l]);
            }
        }
        mm->y = mm->alloc<float>(B * widths[mm->L]);
        mm->t = mm->alloc<float>(B * widths[mm->L]);
        mm->ga = mm->alloc<float>(B * wmax);
        size_t part_n = 1;
        for (int l = 1; l <= mm->L; l++) {
            const int sp = std::max(split_count(widths[l - 1], widths[l], max_batch),
                                    split_count_nt(widths[l - 1] + 1, widths[l], max_batch));
            if (sp > 1) part_n = std::max(part_n, (size_t)sp * (widths[l - 1] + 1) * widths[l]);  // splits <= sp
        }
        mm->part = mm->alloc<float>(part_n);
        mm->cpart = mm->alloc<float>((size_t)CS_PART * wmax);
        mm->gz = mm->alloc<float>(B * wmax);
        mm->gzb = mm->alloc<uint16_t>(B * wmax);
        mm->gzb2 = mm->alloc<uint16_t>(B * wmax);
        mm->gzbT = mm->alloc<uint16_t>(B * wmax);
        mm->loss = mm->alloc<double>(1);
        mm->loss_part = mm->alloc<double>(4096);   // nblocks() caps grids at 4096
        *out = mm.release();
        return LC_OK;
    });
}

void lc_mlp_destroy(lc_mlp *mm) {
    if (!mm) return;
    (void)hipSetDevice(mm->device);
    if (mm->stream) (void)hipStreamSynchronize(mm->stream);
    delete mm;
}

const char *lc_mlp_last_error(const lc_mlp *mm) { return mm ? mm->err.c_str() : "null mlp"; }

size_t lc_mlp_num_params(const lc_mlp *mm) { return mm ? mm->nparam : 0; }

lc_status lc_mlp_set_params(lc_mlp *mm, const float *params, size_t n) {
    if (!mm || !params) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (n != mm->nparam) throw MlpError(LC_E_ARG, "lc_mlp_set_params: wrong parameter count");
        MLPCHK(hipSetDevice(mm->device));
        MLPCHK(hipMemcpyAsync(mm->p, params, n * sizeof(float), hipMemcpyHostToDevice, mm->stream));
        MLPCHK(hipMemsetAsync(mm->m, 0, n * sizeof(float), mm->stream));
        MLPCHK(hipMemsetAsync(mm->v, 0, n * sizeof(float), mm->stream));
        hipLaunchKernelGGL(k_cast_bf16, dim3(nblocks(n)), dim3(256), 0, mm->stream, n, mm->p, mm->pb);
        mm->refresh_wt();
        MLPCHK(hipStreamSynchronize(mm->stream));
        mm->adam_t = 0;
        return LC_OK;
    });
}

lc_status lc_mlp_get_params(lc_mlp *mm, float *params, size_t n) {
    if (!mm || !params) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (n != mm->nparam) throw MlpError(LC_E_ARG, "lc_mlp_get_params: wrong parameter count");
        MLPCHK(hipSetDevice(mm->device));
        MLPCHK(hipMemcpyAsync(params, mm->p, n * sizeof(float), hipMemcpyDeviceToHost, mm->stream));
        MLPCHK(hipStreamSynchronize(mm->stream));
        return LC_OK;
    });
}

lc_status lc_mlp_forward(lc_mlp *mm, const float *x, size_t batch, float *y) {
    if (!mm || !x || !y) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        MLPCHK(hipSetDevice(mm->device));
        const int nin = mm->width[0], nout = mm->width[mm->L];
        for (size_t r0 = 0; r0 < batch; r0 += (size_t)mm->max_batch) {
            const int B = (int)std::min<size_t>(mm->max_batch, batch - r0);
            mm->upload_inputs(x + r0 * nin, B);
            mm->forward(B, true);
            MLPCHK(hipMemcpyAsync(y + r0 * nout, mm->y, (size_t)B * nout * sizeof(float), hipMemcpyDeviceToHost,
                                  mm->stream));
        }
        MLPCHK(hipGetLastError());
        MLPCHK(hipStreamSynchronize(mm->stream));
        return LC_OK;
    });
}

lc_status lc_mlp_grad(lc_mlp *mm, const float *x, const float *target, size_t batch, float *grad, double *loss) {
    if (!mm || !x || !target || !grad || !loss) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (batch == 0 || batch > (size_t)mm->max_batch) throw MlpError(LC_E_ARG, "lc_mlp_grad: batch out of range");
        MLPCHK(hipSetDevice(mm->device));
        const int B = (int)batch, nout = mm->width[mm->L];
        mm->upload_inputs(x, B);
        mm->transpose_inputs(B);
        MLPCHK(hipMemcpyAsync(mm->t, target, (size_t)B * nout * sizeof(float), hipMemcpyHostToDevice, mm->stream));
        mm->forward(B, false, true);
        mm->backward(B);
        MLPCHK(hipGetLastError());
        MLPCHK(hipMemcpyAsync(grad, mm->g, mm->nparam * sizeof(float), hipMemcpyDeviceToHost, mm->stream));
        MLPCHK(hipMemcpyAsync(loss, mm->loss, sizeof(double), hipMemcpyDeviceToHost, mm->stream));
        MLPCHK(hipStreamSynchronize(mm->stream));
        *loss /= (double)B * nout;
        return LC_OK;
    });
}

lc_status lc_mlp_set_dataset(lc_mlp *mm, size_t nrows) {
    if (!mm) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        MLPCHK(hipSetDevice(mm->device));
        MLPCHK(hipStreamSynchronize(mm->stream));
        if (mm->dx) MLPCHK(hipFree(mm->dx));
        if (mm->dt) MLPCHK(hipFree(mm->dt));
        mm->dx = mm->dt = nullptr;
        mm->nrows = 0;
        if (nrows) {
            MLPCHK(hipMalloc(&mm->dx, nrows * mm->width[0] * sizeof(float)));
            MLPCHK(hipMalloc(&mm->dt, nrows * mm->width[mm->L] * sizeof(float)));
        }
        mm->nrows = nrows;
        return LC_OK;
    });
}

lc_status lc_mlp_load_dataset(lc_mlp *mm, size_t row0, const float *x, const float *target, size_t count) {
    if (!mm || !x || !target) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (row0 + count > mm->nrows) throw MlpError(LC_E_ARG, "lc_mlp_load_dataset: rows beyond the data set");
        MLPCHK(hipSetDevice(mm->device));
        const size_t nin = mm->width[0], nout = mm->width[mm->L];
        MLPCHK(hipMemcpyAsync(mm->dx + row0 * nin, x, count * nin * sizeof(float), hipMemcpyHostToDevice, mm->stream));
        MLPCHK(hipMemcpyAsync(mm->dt + row0 * nout, target, count * nout * sizeof(float), hipMemcpyHostToDevice,
                              mm->stream));
        MLPCHK(hipStreamSynchronize(mm->stream));
        return LC_OK;
    });
}

lc_status lc_mlp_train(lc_mlp *mm, int steps, int batch, float lr, uint64_t seed, double *last_loss) {
    if (!mm || !last_loss) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (!mm->nrows) throw MlpError(LC_E_ARG, "lc_mlp_train: no data set");
        if (batch <= 0 || batch > mm->max_batch) throw MlpError(LC_E_ARG, "lc_mlp_train: batch out of range");
        MLPCHK(hipSetDevice(mm->device));
        const auto t0 = std::chrono::steady_clock::now();
        const int nin = mm->width[0], nout = mm->width[mm->L];
        for (int s = 0; s < steps; s++) {
            hipLaunchKernelGGL(k_gather, dim3(batch), dim3(256), 0, mm->stream, batch, nin, nout, mm->nrows, seed,
                               (uint64_t)(mm->adam_t + 1), mm->dx, mm->dt, mm->xb, mm->t);
            mm->transpose_inputs(batch);
            mm->forward(batch, false, true);
            mm->backward(batch);
            mm->adam(lr);
        }
        MLPCHK(hipGetLastError());
        MLPCHK(hipMemcpyAsync(last_loss, mm->loss, sizeof(double), hipMemcpyDeviceToHost, mm->stream));
        MLPCHK(hipStreamSynchronize(mm->stream));
        *last_loss /= (double)batch * nout;
        mm->train_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return LC_OK;
    });
}

lc_status lc_mlp_eval_dataset(lc_mlp *mm, size_t row0, size_t count, double *mse) {
    if (!mm || !mse) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (row0 + count > mm->nrows) throw MlpError(LC_E_ARG, "lc_mlp_eval_dataset: rows beyond the data set");
        MLPCHK(hipSetDevice(mm->device));
        const int nin = mm->width[0], nout = mm->width[mm->L];
        double total = 0.;
        for (size_t r0 = row0; r0 < row0 + count; r0 += (size_t)mm->max_batch) {
            const int B = (int)std::min<size_t>(mm->max_batch, row0 + count - r0);
            const size_t n = (size_t)B * nin;
            hipLaunchKernelGGL(k_cast_bf16, dim3(nblocks(n)), dim3(256), 0, mm->stream, n, mm->dx + r0 * nin, mm->xb);
            mm->forward(B);
            const size_t no = (size_t)B * nout;
            const int nb = nblocks(no);
            hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, mm->stream, no, 0.f, mm->y, mm->dt + r0 * nout, nullptr,
                               nullptr, mm->loss_part);
            hipLaunchKernelGGL(k_sum_loss, dim3(1), dim3(256), 0, mm->stream, nb, mm->loss_part, mm->loss);
            double l = 0.;
            MLPCHK(hipMemcpyAsync(&l, mm->loss, sizeof(double), hipMemcpyDeviceToHost, mm->stream));
            MLPCHK(hipStreamSynchronize(mm->stream));
            total += l;
        }
        *mse = total / ((double)count * nout);
        return LC_OK;
    });
}

lc_status lc_mlp_set_likelihood(lc_mlp *mm, const double *time, const double *flux, const double *weight, size_t n,
                                const float *mu, float scale, double period, int log_output) {
    if (!mm || !time || !flux || !weight || !mu || n == 0) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        MLPCHK(hipSetDevice(mm->device));
        MLPCHK(hipStreamSynchronize(mm->stream));
        for (void *q : mm->lallocs) MLPCHK(hipFree(q));
        mm->lallocs.clear();
        auto take = [&](size_t bytes) {
            void *q = nullptr;
            MLPCHK(hipMalloc(&q, bytes));
            mm->lallocs.push_back(q);
            return q;
        };
        const int nbin = mm->width[mm->L];
        mm->lt = static_cast<double *>(take(n * 8));
        mm->lf = static_cast<double *>(take(n * 8));
        mm->lw = static_cast<double *>(take(n * 8));
        mm->lmu = static_cast<float *>(take((size_t)nbin * 4));
        mm->lt0 = static_cast<double *>(take((size_t)mm->max_batch * 8));
        mm->lchi = static_cast<double *>(take((size_t)mm->max_batch * 8));
        MLPCHK(hipMemcpy(mm->lt, time, n * 8, hipMemcpyHostToDevice));
        MLPCHK(hipMemcpy(mm->lf, flux, n * 8, hipMemcpyHostToDevice));
        MLPCHK(hipMemcpy(mm->lw, weight, n * 8, hipMemcpyHostToDevice));
        MLPCHK(hipMemcpy(mm->lmu, mu, (size_t)nbin * 4, hipMemcpyHostToDevice));
        mm->ln = (int)n;
        mm->lscale = scale;
        mm->lperiod = period;
        mm->llog = log_output ? 1 : 0;
        return LC_OK;
    });
}

lc_status lc_mlp_chisq(lc_mlp *mm, const float *x, const double *t0, size_t m, double *chisq) {
    if (!mm || !x || !t0 || !chisq) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        if (!mm->ln) throw MlpError(LC_E_ARG, "lc_mlp_chisq: no likelihood set");
        MLPCHK(hipSetDevice(mm->device));
        const int nin = mm->width[0], nbin = mm->width[mm->L];
        for (size_t r0 = 0; r0 < m; r0 += (size_t)mm->max_batch) {
            const int B = (int)std::min<size_t>(mm->max_batch, m - r0);
            mm->upload_inputs(x + r0 * nin, B);
            MLPCHK(hipMemcpyAsync(mm->lt0, t0 + r0, (size_t)B * 8, hipMemcpyHostToDevice, mm->stream));
            mm->forward(B, true);
            hipLaunchKernelGGL(k_sur_chisq, dim3(B), dim3(256), 0, mm->stream, nbin, mm->llog, mm->y, mm->lmu, mm->lscale,
                               mm->lt0, mm->lperiod, mm->ln, mm->lt, mm->lf, mm->lw, mm->lchi);
            MLPCHK(hipMemcpyAsync(chisq + r0, mm->lchi, (size_t)B * 8, hipMemcpyDeviceToHost, mm->stream));
        }
        MLPCHK(hipGetLastError());
        MLPCHK(hipStreamSynchronize(mm->stream));
        return LC_OK;
    });
}

lc_status lc_mlp_set_skip(lc_mlp *mm, const float *S, size_t n) {
    if (!mm) return LC_E_ARG;
    return mlp_guarded(mm, [&]() -> int {
        MLPCHK(hipSetDevice(mm->device));
        MLPCHK(hipStreamSynchronize(mm->stream));
        if (!S) {
            if (mm->skip) MLPCHK(hipFree(mm->skip));
            mm->skip = nullptr;
            return LC_OK;
        }
        const size_t want = (size_t)mm->width[0] * mm->width[mm->L];
        if (n != want) throw MlpError(LC_E_ARG, "lc_mlp_set_skip: size must be inputs x outputs");
        if (!mm->skip) MLPCHK(hipMalloc(&mm->skip, want * sizeof(float)));
        MLPCHK(hipMemcpy(mm->skip, S, want * sizeof(float), hipMemcpyHostToDevice));
        return LC_OK;
    });
}

lc_status lc_mlp_stats(const lc_mlp *mm, double *gemm_flop, double *train_ms) {
    if (!mm || !gemm_flop || !train_ms) return LC_E_ARG;
    *gemm_flop = mm->gemm_flop;
    *train_ms = mm->train_ms;
    return LC_OK;
}

}  // extern "C"
