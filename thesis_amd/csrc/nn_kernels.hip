// Kernels of the learned warm start: a 3x3-convolution encoder-decoder
// trained and run on the GPU.  Convolutions are im2col + a bf16 MFMA GEMM
// with FP32 accumulation (weights and im2col operands in bf16, activations
// and gradients in FP32, NHWC layout).  The reference has no surrogate; this
// is new work for the warm start of S/Iterate.cu:211-275 (initial fields).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "lbm_nn.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t u16;

__device__ inline u16 f2bf(float f) {   // round to nearest even
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (u16)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    return (u16)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ inline float bf2f(u16 b) { return __uint_as_float((uint32_t)b << 16); }

// ---------------------------------------------------------------------------
// GEMM: C[m][n] (+)= sum_r A(m, r) B(n, r) (+ bias[n]) (ReLU)
//   A(m, r) = A[m*lda + r] (TA = false) or A[r*lda + m] (TA = true)
//   B(n, r) = B[n*ldb + r] (TB = false) or B[r*ldb + n] (TB = true)
// Workgroup: 4 waves in a WM x WN grid; each wave owns TM x TN tiles of
// 32 x 32 computed with v_mfma_f32_32x32x16_bf16 (lane l holds row l%32,
// k = 8*(l/32)..+7 of A and B; accumulator register r of lane l is row
// 8*(r/4) + 4*(l/32) + r%4, column l%32).  Operand tiles of 32 reduction
// steps go through LDS, transposed there when the operand is stored with the
// reduction index strided.  blockIdx.z splits the reduction (rsplit each);
// with NN_ATOMIC the partial sums are added into C; NN_TRANSC stores C[n][m].

constexpr int BK = 32;
constexpr int LDK = BK + 8;   // LDS row pitch in bf16 (80 B: 16-byte aligned rows)

template <int BX, bool T>
__device__ inline void load_tile(const u16* __restrict__ P, int ld, int X, int x0, int r0, int r_end, u16* s) {
    constexpr int NV = BX * BK / 8;   // 8-element vectors per tile
    for (int v = threadIdx.x; v < NV; v += 256) {
        if (!T) {
            const int x = v >> 2, rr = (v & 3) * 8;
            const int gx = x0 + x, gr = r0 + rr;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (gx < X) {
                const size_t base = (size_t)gx * ld + gr;
                if (gr + 8 <= r_end && (ld & 7) == 0) {
                    val = *(const uint4*)(P + base);
                } else {
                    u16 e[8];
                    for (int i = 0; i < 8; ++i) e[i] = gr + i < r_end ? P[base + i] : (u16)0;
                    val = make_uint4(e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16,
                                     e[4] | (uint32_t)e[5] << 16, e[6] | (uint32_t)e[7] << 16);
                }
            }
            *(uint4*)(s + x * LDK + rr) = val;
        } else {
            constexpr int VPR = BX / 8;   // vectors per reduction row
            const int r = v / VPR, xx = (v % VPR) * 8;
            const int gr = r0 + r, gx = x0 + xx;
            u16 e[8];
            if (gr < r_end && gx + 8 <= X && (ld & 7) == 0) {
                const uint4 val = *(const uint4*)(P + (size_t)gr * ld + gx);
                e[0] = (u16)val.x; e[1] = (u16)(val.x >> 16); e[2] = (u16)val.y; e[3] = (u16)(val.y >> 16);
                e[4] = (u16)val.z; e[5] = (u16)(val.z >> 16); e[6] = (u16)val.w; e[7] = (u16)(val.w >> 16);
            } else {
                for (int i = 0; i < 8; ++i)
                    e[i] = (gr < r_end && gx + i < X) ? P[(size_t)gr * ld + gx + i] : (u16)0;
            }
            for (int i = 0; i < 8; ++i) s[(xx + i) * LDK + r] = e[i];
        }
    }
}

template <int WM, int WN, int TM, int TN, bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int R, const u16* __restrict__ A, int lda,
                                              const u16* __restrict__ B, int ldb, float* __restrict__ C, int ldc,
                                              const float* __restrict__ bias, int flags, int rsplit) {
    static_assert(WM * WN == 4, "four waves");
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    __shared__ __align__(16) u16 sA[BM * LDK];
    __shared__ __align__(16) u16 sB[BN * LDK];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WN, wn = wave % WN;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int r_begin = blockIdx.z * rsplit;
    const int r_end = min(R, r_begin + rsplit);
    f32x16 acc[TM][TN];
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
    for (int r0 = r_begin; r0 < r_end; r0 += BK) {
        load_tile<BM, TA>(A, lda, M, m0, r0, r_end, sA);
        load_tile<BN, TB>(B, ldb, N, n0, r0, r_end, sB);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            bf16x8 a[TM], b[TN];
            const int ko = kk + 8 * (lane >> 5);
#pragma unroll
            for (int i = 0; i < TM; ++i) a[i] = *(const bf16x8*)(sA + ((wm * TM + i) * 32 + (lane & 31)) * LDK + ko);
#pragma unroll
            for (int j = 0; j < TN; ++j) b[j] = *(const bf16x8*)(sB + ((wn * TN + j) * 32 + (lane & 31)) * LDK + ko);
#pragma unroll
            for (int i = 0; i < TM; ++i)
#pragma unroll
                for (int j = 0; j < TN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    const bool relu = flags & NN_RELU, atomic = flags & NN_ATOMIC, transc = flags & NN_TRANSC;
    const bool add_bias = bias && blockIdx.z == 0;
#pragma unroll
    for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
            const int n = n0 + (wn * TN + j) * 32 + (lane & 31);
            if (n >= N) continue;
            const float bn = add_bias ? bias[n] : 0.f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int m = m0 + (wm * TM + i) * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                if (m >= M) continue;
                float v = acc[i][j][r] + bn;
                if (relu) v = fmaxf(v, 0.f);
                float* c = transc ? C + (size_t)n * ldc + m : C + (size_t)m * ldc + n;
                if (atomic) atomicAdd(c, v);
                else *c = v;
            }
        }
}

template <int WM, int WN, int TM, int TN, bool TA, bool TB>
hipError_t gemm_t(int M, int N, int R, const u16* A, int lda, const u16* B, int ldb, float* C, int ldc,
                  const float* bias, int flags, int splits, hipStream_t st) {
    constexpr int BM = WM * TM * 32, BN = WN * TN * 32;
    int rsplit = (R + splits - 1) / splits;
    rsplit = rsplit > 0 ? (rsplit + BK - 1) / BK * BK : BK;
    const int gz = (R + rsplit - 1) / rsplit;
    dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, gz > 0 ? gz : 1);
    hipLaunchKernelGGL((k_gemm<WM, WN, TM, TN, TA, TB>), grid, dim3(256), 0, st, M, N, R, A, lda, B, ldb, C, ldc, bias,
                       flags, rsplit);
    return hipGetLastError();
}

template <bool TA, bool TB>
hipError_t gemm_shape(int M, int N, int R, const u16* A, int lda, const u16* B, int ldb, float* C, int ldc,
                      const float* bias, int flags, int splits, hipStream_t st) {
    if (N <= 32) return gemm_t<4, 1, 2, 1, TA, TB>(M, N, R, A, lda, B, ldb, C, ldc, bias, flags, splits, st);
    if (N <= 64) return gemm_t<2, 2, 2, 1, TA, TB>(M, N, R, A, lda, B, ldb, C, ldc, bias, flags, splits, st);
    return gemm_t<2, 2, 2, 2, TA, TB>(M, N, R, A, lda, B, ldb, C, ldc, bias, flags, splits, st);
}

// ---------------------------------------------------------------------------
// elementwise and data-movement kernels (NHWC, FP32 unless noted)

// Quotient by an invariant divisor d (1 <= d < 2^31) for n < 2^32:
// ((mulhi(m, n) + n) >> s) with s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1.
struct Div {
    uint32_t m, s;
};
inline Div make_div(uint32_t d) {
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    return Div{(uint32_t)(((1ull << 32) * ((1ull << s) - d)) / d + 1), s};
}
__device__ inline uint32_t qdiv(uint32_t n, Div d) {
    return (uint32_t)(((uint64_t)__umulhi(d.m, n) + n) >> d.s);
}

// col[p][k], k = (ky*3 + kx)*C + c for k < 9C, zero for 9C <= k < Kp; zero padding
// outside the image.  One thread per (pixel, k), consecutive threads along k
// (coalesced stores); 32-bit indices with invariant-divisor quotients (the
// launcher checks that every index fits).
__global__ void k_im2col(const float* __restrict__ x, int H, int W, int C, int Kp, Div dKp, Div dC, Div dW, Div dH,
                         uint32_t total, u16* __restrict__ col) {
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const uint32_t p = qdiv(e, dKp);
        const int k = (int)(e - p * (uint32_t)Kp);
        float v = 0.f;
        if (k < 9 * C) {
            const uint32_t tap = qdiv((uint32_t)k, dC);
            const int c = k - (int)tap * C;
            const int ky = (int)tap / 3, kx = (int)tap - ky * 3;
            const uint32_t row = qdiv(p, dW);
            const int xw = (int)(p - row * (uint32_t)W);
            const uint32_t b = qdiv(row, dH);
            const int yh = (int)(row - b * (uint32_t)H);
            const int ys = yh + ky - 1, xs = xw + kx - 1;
            if (ys >= 0 && ys < H && xs >= 0 && xs < W) v = x[((size_t)(b * H + ys) * W + xs) * C + c];
        }
        col[e] = f2bf(v);
    }
}

// dx[b][y][x][c] = sum over taps of dcol at the output pixel that read it.
__global__ void k_col2im(const float* __restrict__ dcol, int Bn, int H, int W, int C, int Kp, float* __restrict__ dx) {
    const uint32_t total = (uint32_t)Bn * H * W * C;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < total; e += gridDim.x * blockDim.x) {
        const uint32_t p = e / (uint32_t)C;
        const int c = (int)(e - p * (uint32_t)C);
        const uint32_t row = p / (uint32_t)W;
        const int xw = (int)(p - row * (uint32_t)W);
        const uint32_t b = row / (uint32_t)H;
        const int yh = (int)(row - b * (uint32_t)H);
        float s = 0.f;
        for (int ky = 0; ky < 3; ++ky) {
            const int yo = yh - ky + 1;
            if (yo < 0 || yo >= H) continue;
            for (int kx = 0; kx < 3; ++kx) {
                const int xo = xw - kx + 1;
                if (xo < 0 || xo >= W) continue;
                s += dcol[(size_t)((b * H + yo) * W + xo) * Kp + (ky * 3 + kx) * C + c];
            }
        }
        dx[e] = s;
    }
}

// dz = dy * (y > 0) (relu) or dy, as bf16
__global__ void k_grad_bf16(const float* __restrict__ dy, const float* __restrict__ y, size_t n, int relu,
                            u16* __restrict__ dz) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dz[e] = f2bf(relu && !(y[e] > 0.f) ? 0.f : dy[e]);
}

__global__ void k_to_bf16(const float* __restrict__ x, size_t n, u16* __restrict__ y) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        y[e] = f2bf(x[e]);
}

// db[n] += sum_m dz[m][n] (dz FP32 with the relu mask applied through y).
// Each workgroup strides over rows with its 256 threads spread over the N
// columns (consecutive threads read consecutive columns of a row), sums per
// column in LDS and adds one FP32 atomic per column.
__global__ void k_colsum(const float* __restrict__ dy, const float* __restrict__ y, int relu, int M, int N,
                         float* __restrict__ db) {
    const int rows = blockDim.x / N;          // rows per pass (N <= 256)
    const int n = threadIdx.x % N, r = threadIdx.x / N;
    float s = 0.f;
    if (r < rows) {
        for (int m = blockIdx.x * rows + r; m < M; m += gridDim.x * rows) {
            const size_t i = (size_t)m * N + n;
            const float g = dy[i];
            s += (relu && !(y[i] > 0.f)) ? 0.f : g;
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    if ((int)threadIdx.x < N) {
        float t = 0.f;
        for (int k = 0; k < rows; ++k) t += red[k * N + threadIdx.x];
        atomicAdd(db + threadIdx.x, t);
    }
}

__global__ void k_avgpool2(const float* __restrict__ x, int Bn, int H, int W, int C, float* __restrict__ y) {
    const int Ho = H / 2, Wo = W / 2;
    const size_t total = (size_t)Bn * Ho * Wo * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t p = e / C;
        const int xo = (int)(p % Wo), yo = (int)((p / Wo) % Ho);
        const size_t b = p / ((size_t)Wo * Ho);
        const size_t i0 = ((b * H + 2 * yo) * W + 2 * xo) * C + c;
        y[e] = 0.25f * (x[i0] + x[i0 + C] + x[i0 + (size_t)W * C] + x[i0 + (size_t)W * C + C]);
    }
}

__global__ void k_avgpool2_bwd(const float* __restrict__ dy, int Bn, int H, int W, int C, float* __restrict__ dx) {
    const size_t total = (size_t)Bn * H * W * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t p = e / C;
        const int xw = (int)(p % W), yh = (int)((p / W) % H);
        const size_t b = p / ((size_t)W * H);
        dx[e] = 0.25f * dy[((b * (H / 2) + yh / 2) * (W / 2) + xw / 2) * C + c];
    }
}

// nearest-neighbour x2 upsampling; accumulate = 1 adds into y
__global__ void k_up2(const float* __restrict__ x, int Bn, int H, int W, int C, float* __restrict__ y) {
    const size_t total = (size_t)Bn * 2 * H * 2 * W * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t p = e / C;
        const int xw = (int)(p % (2 * W)), yh = (int)((p / (2 * W)) % (2 * H));
        const size_t b = p / ((size_t)4 * W * H);
        y[e] = x[((b * H + yh / 2) * W + xw / 2) * C + c];
    }
}

__global__ void k_up2_bwd(const float* __restrict__ dy, int Bn, int H, int W, int C, float* __restrict__ dx) {
    const size_t total = (size_t)Bn * H * W * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t p = e / C;
        const int xw = (int)(p % W), yh = (int)((p / W) % H);
        const size_t b = p / ((size_t)W * H);
        const size_t i0 = ((b * 2 * H + 2 * yh) * 2 * W + 2 * xw) * C + c;
        const size_t row = (size_t)2 * W * C;
        dx[e] = dy[i0] + dy[i0 + C] + dy[i0 + row] + dy[i0 + row + C];
    }
}

// y = x + (1/W) sum_{x' >= x} x (reverse running sum along the width);
// backward dx = dy + (1/W) sum_{x' <= x} dy.  One thread per (b, y, c) row.
__global__ void k_xint(const float* __restrict__ x, int Bn, int H, int W, int C, int bwd, float* __restrict__ y) {
    const size_t rows = (size_t)Bn * H * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < rows; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t bh = e / C;
        const float inv = 1.f / (float)W;
        float run = 0.f;
        if (!bwd) {
            for (int xw = W - 1; xw >= 0; --xw) {
                const size_t i = (bh * W + xw) * C + c;
                run += x[i];
                y[i] = x[i] + inv * run;
            }
        } else {
            for (int xw = 0; xw < W; ++xw) {
                const size_t i = (bh * W + xw) * C + c;
                run += x[i];
                y[i] = x[i] + inv * run;
            }
        }
    }
}

// y = x + (1/H) sum_{y'} x (mean over the height broadcast back); the operator
// is symmetric, so it is its own adjoint.  One thread per (b, x, c) column.
__global__ void k_ymix(const float* __restrict__ x, int Bn, int H, int W, int C, float* __restrict__ y) {
    const size_t cols = (size_t)Bn * W * C;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < cols; e += (size_t)gridDim.x * blockDim.x) {
        const int c = (int)(e % C);
        const size_t bw = e / C;
        const int xw = (int)(bw % W);
        const size_t b = bw / W;
        float s = 0.f;
        for (int yh = 0; yh < H; ++yh) s += x[((b * H + yh) * W + xw) * C + c];
        s /= (float)H;
        for (int yh = 0; yh < H; ++yh) {
            const size_t i = ((b * H + yh) * W + xw) * C + c;
            y[i] = x[i] + s;
        }
    }
}

__global__ void k_axpy(float a, const float* __restrict__ x, size_t n, float* __restrict__ y) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        y[e] += a * x[e];
}

__global__ void k_add(const float* __restrict__ x, const float* __restrict__ y, size_t n, float* __restrict__ z) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        z[e] = x[e] + y[e];
}

__global__ void k_fill(float v, size_t n, float* __restrict__ y) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        y[e] = v;
}

// masked mean-squared error over the C channels of pixels with mask > 0:
// grad = 2 w_c (pred - target) / count; loss partials per block
__global__ void k_mse(const float* __restrict__ pred, const float* __restrict__ target,
                      const float* __restrict__ mask, const float* __restrict__ wch, size_t npix, int C,
                      float inv_count, float* __restrict__ grad, float* __restrict__ loss) {
    float s = 0.f;
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < npix; p += (size_t)gridDim.x * blockDim.x) {
        const float m = mask[p];
        for (int c = 0; c < C; ++c) {
            const size_t i = p * C + c;
            const float d = m > 0.f ? pred[i] - target[i] : 0.f;
            s += wch[c] * d * d;
            grad[i] = 2.f * wch[c] * d * inv_count;
        }
    }
    __shared__ float red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_count);
}

// Adam with bias correction on FP32 master weights; refreshes the bf16 copy
// (the first n_bf entries, the GEMM weights) of w.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, size_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       u16* __restrict__ wbf, size_t n_bf) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
        w[e] = wi;
        if (e < n_bf) wbf[e] = f2bf(wi);
    }
}

inline dim3 grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return dim3((unsigned)(b ? b : 1));
}

}  // namespace

extern "C" {

int nn_gemm_bf16(int M, int N, int R, const void* A, int lda, int ta, const void* B, int ldb, int tb, float* C,
                 int ldc, const float* bias, int flags, int splits, void* stream) {
    if (M < 0 || N < 0 || R < 0 || splits < 1) return 3;
    if (splits > 1 && !(flags & NN_ATOMIC)) return 3;   // partial sums must be added
    if (M == 0 || N == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const u16* a = (const u16*)A;
    const u16* b = (const u16*)B;
    hipError_t e;
    if (!ta && !tb) e = gemm_shape<false, false>(M, N, R, a, lda, b, ldb, C, ldc, bias, flags, splits, st);
    else if (!ta && tb) e = gemm_shape<false, true>(M, N, R, a, lda, b, ldb, C, ldc, bias, flags, splits, st);
    else if (ta && !tb) e = gemm_shape<true, false>(M, N, R, a, lda, b, ldb, C, ldc, bias, flags, splits, st);
    else e = gemm_shape<true, true>(M, N, R, a, lda, b, ldb, C, ldc, bias, flags, splits, st);
    return e == hipSuccess ? 0 : 4;
}

#define NN_LAUNCH(kernel, n, ...)                                                                          \
    do {                                                                                                   \
        hipLaunchKernelGGL(kernel, grid_for(n), dim3(256), 0, (hipStream_t)stream, __VA_ARGS__);           \
        return hipGetLastError() == hipSuccess ? 0 : 4;                                                    \
    } while (0)

int nn_im2col3x3(const float* x, int Bn, int H, int W, int C, int Kp, void* col, void* stream) {
    if (Kp < 9 * C || C < 1 || H < 1 || W < 1 || Bn < 0) return 3;
    const uint64_t total = (uint64_t)Bn * H * W * Kp;
    if (total >= (1ull << 31)) return 3;
    if (total == 0) return 0;
    NN_LAUNCH(k_im2col, (size_t)total, x, H, W, C, Kp, make_div((uint32_t)Kp), make_div((uint32_t)C),
              make_div((uint32_t)W), make_div((uint32_t)H), (uint32_t)total, (u16*)col);
}
int nn_col2im3x3(const float* dcol, int Bn, int H, int W, int C, int Kp, float* dx, void* stream) {
    if (Kp < 9 * C) return 3;
    if ((uint64_t)Bn * H * W * C >= (1ull << 32)) return 3;
    NN_LAUNCH(k_col2im, (size_t)Bn * H * W * C, dcol, Bn, H, W, C, Kp, dx);
}
int nn_grad_bf16(const float* dy, const float* y, int64_t n, int relu, void* dz, void* stream) {
    NN_LAUNCH(k_grad_bf16, (size_t)n, dy, y, (size_t)n, relu, (u16*)dz);
}
int nn_to_bf16(const float* x, int64_t n, void* y, void* stream) {
    NN_LAUNCH(k_to_bf16, (size_t)n, x, (size_t)n, (u16*)y);
}
int nn_colsum(const float* dy, const float* y, int relu, int M, int N, float* db, void* stream) {
    if (N <= 0 || M <= 0) return 0;
    if (N > 256) return 3;
    const int rows = 256 / N;
    const int blocks = (int)std::min<int64_t>(1024, ((int64_t)M + rows - 1) / rows);
    hipLaunchKernelGGL(k_colsum, dim3(blocks), dim3(256), 0, (hipStream_t)stream, dy, y, relu, M, N, db);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}
int nn_avgpool2(const float* x, int Bn, int H, int W, int C, float* y, void* stream) {
    if ((H | W) & 1) return 3;
    NN_LAUNCH(k_avgpool2, (size_t)Bn * (H / 2) * (W / 2) * C, x, Bn, H, W, C, y);
}
int nn_avgpool2_bwd(const float* dy, int Bn, int H, int W, int C, float* dx, void* stream) {
    if ((H | W) & 1) return 3;
    NN_LAUNCH(k_avgpool2_bwd, (size_t)Bn * H * W * C, dy, Bn, H, W, C, dx);
}
int nn_up2(const float* x, int Bn, int H, int W, int C, float* y, void* stream) {
    NN_LAUNCH(k_up2, (size_t)Bn * 4 * H * W * C, x, Bn, H, W, C, y);
}
int nn_up2_bwd(const float* dy, int Bn, int H, int W, int C, float* dx, void* stream) {
    NN_LAUNCH(k_up2_bwd, (size_t)Bn * H * W * C, dy, Bn, H, W, C, dx);
}
int nn_xint(const float* x, int Bn, int H, int W, int C, int bwd, float* y, void* stream) {
    NN_LAUNCH(k_xint, (size_t)Bn * H * C, x, Bn, H, W, C, bwd, y);
}
int nn_ymix(const float* x, int Bn, int H, int W, int C, float* y, void* stream) {
    NN_LAUNCH(k_ymix, (size_t)Bn * W * C, x, Bn, H, W, C, y);
}
int nn_axpy(float a, const float* x, int64_t n, float* y, void* stream) {
    NN_LAUNCH(k_axpy, (size_t)n, a, x, (size_t)n, y);
}
int nn_add(const float* x, const float* y, int64_t n, float* z, void* stream) {
    NN_LAUNCH(k_add, (size_t)n, x, y, (size_t)n, z);
}
int nn_fill(float v, int64_t n, float* y, void* stream) {
    NN_LAUNCH(k_fill, (size_t)n, v, (size_t)n, y);
}
int nn_mse(const float* pred, const float* target, const float* mask, const float* wch, int64_t npix, int C,
           float inv_count, float* grad, float* loss, void* stream) {
    hipLaunchKernelGGL(k_mse, dim3(512), dim3(256), 0, (hipStream_t)stream, pred, target, mask, wch, (size_t)npix, C,
                       inv_count, grad, loss);
    return hipGetLastError() == hipSuccess ? 0 : 4;
}
int nn_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float b1, float b2, float eps,
            int64_t t, void* wbf, int64_t n_bf, void* stream) {
    if (t < 1) return 3;
    const float c1 = 1.f / (1.f - powf(b1, (float)t)), c2 = 1.f / (1.f - powf(b2, (float)t));
    NN_LAUNCH(k_adam, (size_t)n, w, g, m, v, (size_t)n, lr, b1, b2, eps, c1, c2, (u16*)wbf, (size_t)n_bf);
}

}  // extern "C"
