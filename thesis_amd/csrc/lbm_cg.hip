// Colour-gradient two-phase step for D2Q9 (periodic box), fused with LDS halo.
//
// Follows the reference's GPU formulation: gradient of the phase field
// (rho_r - rho_b)/rho with the 8- or 24-neighbour isotropic stencil
// (calculateColorGradient / calculateHOColorGradient case 0,
// S/GpuCollision.cu:112-281), BGK collision towards the two-phase equilibrium
// with effective viscosity and rest fraction, perturbation and Latva-Kokko
// recolouring (gpuCollBgkwGC2D, S/GpuCollision.cu:552-622), pull streaming of
// both colours (gpuStreaming2DCG, S/GpuStream.cu:30-158) and moments
// (gpuUpdateMacro2DCG, S/GpuUpdateMacro.cu:43-135).  Rotated like the
// single-phase step: the stored state is the post-collision colours.
//
// Per link the storage holds h = f - w (total population deviation) and the
// red population r (b = f - r); both colours are pulled, the phase field of
// the 32x16 node tile and a 1- or 2-node halo goes to LDS, then every node
// collides and recolours in registers.
#include "lbm_cg.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {
namespace cg {

constexpr int TX = 32, TY = 16, THREADS = 256;   // tile 32 x 16 nodes, 2 per thread
__device__ constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__host__ __device__ constexpr double W9(int i) { return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
// phi_k, theta_k (S/GpuInit.cu:150-158): rest equilibrium phi + theta*alpha
__host__ __device__ constexpr double PHI(int i) { return i == 0 ? 0.0 : (i <= 4 ? 0.2 : 0.05); }
__host__ __device__ constexpr double TETA(int i) { return i == 0 ? 1.0 : (i <= 4 ? -0.2 : -0.05); }
// perturbation constants B_k (S/GpuInit.cu:170-171)
__host__ __device__ constexpr double BPERT(int i) { return i == 0 ? -4.0 / 27.0 : (i <= 4 ? 2.0 / 27.0 : 5.0 / 108.0); }
__host__ __device__ constexpr double CNORM(int i) { return i == 0 ? 1.0 : (i <= 4 ? 1.0 : 1.4142135623730950488); }
// enhanced-model chi_k, psi_k (S/GpuInit.cu:158-165)
__host__ __device__ constexpr double CHI(int i) { return i == 0 ? -8.0 / 3.0 : (i <= 4 ? -1.0 / 6.0 : 1.0 / 12.0); }
__host__ __device__ constexpr double PSI(int i) { return i == 0 ? 0.0 : (i <= 4 ? 0.5 : 0.125); }
// 24-neighbour stencil weight by |c|^2 (S/GpuInit.cu:186-199)
__host__ __device__ constexpr double HOW(int c2) {
    return c2 == 1 ? 960.0 / 5040 : c2 == 2 ? 448.0 / 5040 : c2 == 4 ? 84.0 / 5040 : c2 == 5 ? 32.0 / 5040 : 1.0 / 5040;
}

// storage: FP64 / FP32 raw; FP16 scaled by 1024/w (deviation) and 1/w (red)
template <int S> struct CIo;
template <> struct CIo<0> {
    using raw = double; using T = double;
    __device__ static T ld(const raw* p, uint32_t k, int, bool) { return p[k]; }
    __device__ static void st(raw* p, uint32_t k, T v, int, bool) { __builtin_nontemporal_store(v, p + k); }
};
template <> struct CIo<1> {
    using raw = float; using T = float;
    __device__ static T ld(const raw* p, uint32_t k, int, bool) { return p[k]; }
    __device__ static void st(raw* p, uint32_t k, T v, int, bool) { __builtin_nontemporal_store(v, p + k); }
};
template <> struct CIo<2> {
    using raw = uint16_t; using T = float;
    __device__ static T ld(const raw* p, uint32_t k, int i, bool red) {
        const float s = red ? (float)W9(i) : (float)(W9(i) / 1024.0);
        return (float)__builtin_bit_cast(_Float16, p[k]) * s;
    }
    __device__ static void st(raw* p, uint32_t k, T v, int i, bool red) {
        const float s = red ? (float)(1.0 / W9(i)) : (float)(1024.0 / W9(i));
        __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, (_Float16)(v * s)), p + k);
    }
};

template <class T> __device__ inline T crecip(T x) { return (T)1 / x; }

// Gather node (x, y) from the post-collision buffer: pull with periodic wrap.
template <int S>
__device__ inline void pull(const typename CIo<S>::raw* __restrict__ src, const CgParams& cp, int x, int y,
                            typename CIo<S>::T* h, typename CIo<S>::T* r) {
    const int nx = cp.nx, ny = cp.ny;
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        int sx = x - CX[i], sy = y - CY[i];
        sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
        sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
        const uint32_t k = (uint32_t)sy * (uint32_t)nx + (uint32_t)sx;
        h[i] = CIo<S>::ld(src, (uint32_t)i * cp.plane + k, i, false);
        r[i] = CIo<S>::ld(src, (uint32_t)(9 + i) * cp.plane + k, i, true);
    }
}

template <class T>
__device__ inline void densities(const T* h, const T* r, T& rho, T& rr) {
    T d = 0, s = 0;
#pragma unroll
    for (int i = 0; i < 9; ++i) { d += h[i]; s += r[i]; }
    rho = (T)1 + d;
    rr = s;
}

// Collision, perturbation and recolouring of one node (gpuCollBgkwGC2D;
// with E the density-gradient terms of gpuCollEnhancedBgkwGC2D,
// S/GpuCollision.cu:624-707, from the density gradient (dx, dy)).
// Writes the new deviation h' = f* - w and red r'.
template <class T, bool E>
__device__ inline void cg_collide(const CgParams& cp, const T* h, const T* r, T gx, T gy, T dx, T dy, T* ho, T* ro,
                                  T* mac) {
    T rho, rr;
    densities<T>(h, r, rho, rr);
    const T rb = rho - rr;
    T jx = 0, jy = 0;
#pragma unroll
    for (int i = 1; i < 9; ++i) {   // (zero components left out: 0 * x is not folded)
        if (CX[i]) jx += CX[i] > 0 ? h[i] : -h[i];
        if (CY[i]) jy += CY[i] > 0 ? h[i] : -h[i];
    }
    const T ir = crecip(rho);
    const T u = jx * ir, v = jy * ir;
    const T mean_nu = (T)1 / (rr / (rho * (T)cp.r_nu) + rb / (rho * (T)cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T alpha = (T)cp.r_alpha * rr * ir + (T)cp.b_alpha * rb * ir;
    const T gn = sqrt(gx * gx + gy * gy);
    const bool on = gn > (T)cp.g_limit;
    const T ign = on ? (T)1 / gn : (T)0;
    const T kr = rr * ir, kk = (T)cp.beta * rr * rb * ir;
    const T uu = u * u + v * v;
    const T ugr = u * dx + v * dy;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const T cg = cdot<T>(CX[k], CY[k], 0, gx, gy, (T)0);
        const T cs = (on && k) ? cg * ign * (T)(1.0 / CNORM(k)) : (T)0;
        const T pert = on ? (T)cp.A * gn * ((T)W9(k) * cg * cg * ign * ign - (T)BPERT(k)) : (T)0;
        const T rest = (T)PHI(k) + (T)TETA(k) * alpha;
        const T cu = cdot<T>(CX[k], CY[k], 0, u, v, (T)0);
        // f_eq - w = (rho - 1) rest + (rest - w) + rho w (3cu + 4.5cu^2 - 1.5uu)
        T feqd = (rho - (T)1) * rest + (rest - (T)W9(k)) +
                 rho * (T)W9(k) * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu);
        if (E) {   // nu (chi u.grad rho + psi G:cc), G:cc = 2 (c.u)(c.grad rho)
            const T cgr = cdot<T>(CX[k], CY[k], 0, dx, dy, (T)0);
            feqd += mean_nu * ((T)CHI(k) * ugr + (T)PSI(k) * (T)2 * cu * cgr);
        }
        const T fsd = om * feqd + ((T)1 - om) * h[k] + pert;       // f* - w
        const T fs = fsd + (T)W9(k);
        ho[k] = fsd;
        ro[k] = kr * fs + kk * cs * rest;
    }
    mac[0] = rr; mac[1] = rb; mac[2] = rho; mac[3] = u; mac[4] = v;
}

template <int H>
__device__ inline void gradient(const float* __restrict__ psi, int lx, int ly, float& gx, float& gy) {
    // psi tile row length TX + 2H, node (lx, ly) of the interior at (lx+H, ly+H)
    constexpr int RW = TX + 2 * H;
    float ax = 0, ay = 0;
#pragma unroll
    for (int dy = -H; dy <= H; ++dy)
#pragma unroll
        for (int dx = -H; dx <= H; ++dx) {
            const int c2 = dx * dx + dy * dy;
            if (c2 == 0) continue;
            const float wk = H == 1 ? (c2 == 1 ? 4.0f / 12.0f : 1.0f / 12.0f) : (float)HOW(c2);
            const float p = psi[(ly + H + dy) * RW + (lx + H + dx)];
            if (dx) ax += wk * (float)dx * p;
            if (dy) ay += wk * (float)dy * p;
        }
    gx = ax; gy = ay;
}
template <int H>
__device__ inline void gradient(const double* __restrict__ psi, int lx, int ly, double& gx, double& gy) {
    constexpr int RW = TX + 2 * H;
    double ax = 0, ay = 0;
#pragma unroll
    for (int dy = -H; dy <= H; ++dy)
#pragma unroll
        for (int dx = -H; dx <= H; ++dx) {
            const int c2 = dx * dx + dy * dy;
            if (c2 == 0) continue;
            const double wk = H == 1 ? (c2 == 1 ? 4.0 / 12.0 : 1.0 / 12.0) : HOW(c2);
            const double p = psi[(ly + H + dy) * RW + (lx + H + dx)];
            if (dx) ax += wk * (double)dx * p;
            if (dy) ay += wk * (double)dy * p;
        }
    gx = ax; gy = ay;
}

// One fused step.  Block (32, 8): thread (tx, ty) owns interior nodes
// (tx, ty) and (tx, ty + 8) of the 32 x 16 tile.  With E the density goes to
// a second LDS tile for the density gradient of the enhanced model.
template <int S, int H, bool E>
__global__ __launch_bounds__(THREADS) void k_cg_step(const typename CIo<S>::raw* __restrict__ src,
                                                     typename CIo<S>::raw* __restrict__ dst, CgParams cp,
                                                     double* __restrict__ mac, int write_macros) {
    using T = typename CIo<S>::T;
    constexpr int RW = TX + 2 * H, RH = TY + 2 * H;
    __shared__ T psi[RW * RH];
    __shared__ T rt[E ? RW * RH : 1];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    // tiles renumbered for the XCDs (xcd_block), so that the tiles an XCD
    // runs together are neighbours in x and share their halo columns' lines
    const uint32_t b = xcd_block(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y, cp.xcd_shift,
                                 cp.xchunk_shift);
    const uint32_t by = b / gridDim.x;
    const int x0 = (int)(b - by * gridDim.x) * TX, y0 = (int)by * TY;
    const int nx = cp.nx, ny = cp.ny;
    // periodic index; the division only where a coordinate leaves the box
    auto pwrap = [](int a, int n) { return a >= 0 && a < n ? a : ((a % n) + n) % n; };
    T h[2][9], r[2][9];
    bool inside[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int lx = tx, ly = ty + 8 * k, x = x0 + lx, y = y0 + ly;
        inside[k] = x < nx && y < ny;
        const int gxp = pwrap(x, nx), gyp = pwrap(y, ny);   // partial tiles wrap onto periodic neighbours
        pull<S>(src, cp, gxp, gyp, h[k], r[k]);
        T rho, rr;
        densities<T>(h[k], r[k], rho, rr);
        psi[(ly + H) * RW + (lx + H)] = ((T)2 * rr - rho) / rho;
        if (E) rt[(ly + H) * RW + (lx + H)] = rho;
    }
    // halo ring (positions of the padded tile outside the interior)
    constexpr int NHALO = RW * RH - TX * TY;
    for (int e = tid; e < NHALO; e += THREADS) {
        int px, py;
        // enumerate: top H rows, bottom H rows (full width), then the left/right H columns
        if (e < H * RW) { py = e / RW; px = e % RW; }
        else if (e < 2 * H * RW) { const int q = e - H * RW; py = TY + H + q / RW; px = q % RW; }
        else { const int q = e - 2 * H * RW; py = H + q / (2 * H); const int c = q % (2 * H); px = c < H ? c : TX + c; }
        const int x = pwrap(x0 + px - H, nx), y = pwrap(y0 + py - H, ny);
        T hh[9], rr9[9];
        pull<S>(src, cp, x, y, hh, rr9);
        T rho, rr;
        densities<T>(hh, rr9, rho, rr);
        psi[py * RW + px] = ((T)2 * rr - rho) / rho;
        if (E) rt[py * RW + px] = rho;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (!inside[k]) continue;
        const int lx = tx, ly = ty + 8 * k, x = x0 + lx, y = y0 + ly;
        T gx, gy, dx = 0, dy = 0;
        gradient<H>(psi, lx, ly, gx, gy);
        if (E) gradient<H>(rt, lx, ly, dx, dy);
        T ho[9], ro[9], m[5];
        cg_collide<T, E>(cp, h[k], r[k], gx, gy, dx, dy, ho, ro, m);
        const uint32_t idx = (uint32_t)y * (uint32_t)nx + (uint32_t)x;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            CIo<S>::st(dst, (uint32_t)i * cp.plane + idx, ho[i], i, false);
            CIo<S>::st(dst, (uint32_t)(9 + i) * cp.plane + idx, ro[i], i, true);
        }
        if (write_macros) {
            const size_t n = (size_t)nx * ny;
            for (int a = 0; a < 5; ++a) mac[a * n + idx] = (double)m[a];
        }
    }
}

// First post-collision state from colour densities (rest equilibria,
// S/GpuInit.cu:323-352), or from given colour populations rf, bf (FP64,
// 9 planes of N) when rf is not null: gradient straight from the density
// arrays rr0, rb0 (the densities of rf, bf in the second case).
template <int S, int H, bool E>
__global__ __launch_bounds__(256) void k_cg_init(typename CIo<S>::raw* __restrict__ dst, CgParams cp,
                                                 const double* __restrict__ rr0, const double* __restrict__ rb0,
                                                 const double* __restrict__ rf, const double* __restrict__ bf,
                                                 double* __restrict__ mac) {
    using T = typename CIo<S>::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int nx = cp.nx, ny = cp.ny;
    if (x >= nx) return;
    T gx = 0, gy = 0, ex = 0, ey = 0;
    for (int dy = -H; dy <= H; ++dy)
        for (int dx = -H; dx <= H; ++dx) {
            const int c2 = dx * dx + dy * dy;
            if (c2 == 0) continue;
            const double wk = H == 1 ? (c2 == 1 ? 4.0 / 12.0 : 1.0 / 12.0) : HOW(c2);
            const int xx = ((x + dx) % nx + nx) % nx, yy = ((y + dy) % ny + ny) % ny;
            const size_t j = (size_t)yy * nx + xx;
            const double p = (rr0[j] - rb0[j]) / (rr0[j] + rb0[j]);
            gx += (T)(wk * dx * p);
            gy += (T)(wk * dy * p);
            if (E) {
                ex += (T)(wk * dx * (rr0[j] + rb0[j]));
                ey += (T)(wk * dy * (rr0[j] + rb0[j]));
            }
        }
    const size_t idx = (size_t)y * nx + x;
    const size_t n = (size_t)nx * ny;
    const T rr = (T)rr0[idx], rb = (T)rb0[idx];
    T h[9], r[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        if (rf) {
            r[i] = (T)rf[i * n + idx];
            h[i] = (T)(rf[i * n + idx] + bf[i * n + idx] - W9(i));
        } else {
            const T rr_i = rr * (T)(PHI(i) + TETA(i) * cp.r_alpha), rb_i = rb * (T)(PHI(i) + TETA(i) * cp.b_alpha);
            r[i] = rr_i;
            h[i] = rr_i + rb_i - (T)W9(i);
        }
    }
    T ho[9], ro[9], m[5];
    cg_collide<T, E>(cp, h, r, gx, gy, ex, ey, ho, ro, m);
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        CIo<S>::st(dst, (uint32_t)i * cp.plane + (uint32_t)idx, ho[i], i, false);
        CIo<S>::st(dst, (uint32_t)(9 + i) * cp.plane + (uint32_t)idx, ro[i], i, true);
    }
    mac[idx] = (double)rr; mac[n + idx] = (double)rb; mac[2 * n + idx] = (double)(rr + rb);
    mac[3 * n + idx] = rf ? (double)m[3] : 0.0;
    mac[4 * n + idx] = rf ? (double)m[4] : 0.0;
}

// Colour densities of FP64 populations (q planes of n).
__global__ __launch_bounds__(256) void k_cg_dens(int q, const double* __restrict__ rf, const double* __restrict__ bf,
                                                 size_t n, double* __restrict__ rr, double* __restrict__ rb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0, b = 0;
    for (int k = 0; k < q; ++k) { a += rf[k * n + i]; b += bf[k * n + i]; }
    rr[i] = a;
    rb[i] = b;
}

// Post-stream colour populations of the state in src (FP64, full).
template <int S>
__global__ __launch_bounds__(256) void k_cg_recon(const typename CIo<S>::raw* __restrict__ src, CgParams cp,
                                                  double* __restrict__ rf, double* __restrict__ bf) {
    using T = typename CIo<S>::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= cp.nx) return;
    T h[9], r[9];
    pull<S>(src, cp, x, y, h, r);
    const size_t n = (size_t)cp.nx * cp.ny, idx = (size_t)y * cp.nx + x;
    for (int i = 0; i < 9; ++i) {
        rf[i * n + idx] = (double)r[i];
        bf[i * n + idx] = (double)h[i] + W9(i) - (double)r[i];
    }
}

// FP64 reductions over the macroscopic fields: 0 sum rho_r, 1 sum rho_b,
// 2/3 reference Laplace sums over chi >= kappa (rho_r, count), 4/5 over
// chi <= -kappa (rho_b, count) (S/GpuUpdateMacro.cu:124-133, :347-357),
// 6/7 bulk rho over chi > 0.99 (sum, count), 8/9 over chi < -0.99,
// 10 max |u|, 11 non-finite count.  w is null for D2Q9.
__global__ __launch_bounds__(256) void k_cg_reduce(const double* __restrict__ mac, const double* __restrict__ wz,
                                                   size_t n, double kappa, double* __restrict__ partials) {
    double a[NRED];
    for (int k = 0; k < NRED; ++k) a[k] = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double rr = mac[i], rb = mac[n + i], rho = mac[2 * n + i], u = mac[3 * n + i], v = mac[4 * n + i];
        const double chi = (rr - rb) / rho;
        a[0] += rr; a[1] += rb;
        if (chi >= kappa) { a[2] += rr; a[3] += 1; }
        else if (chi <= -kappa) { a[4] += rb; a[5] += 1; }
        if (chi > 0.99) { a[6] += rho; a[7] += 1; }
        if (chi < -0.99) { a[8] += rho; a[9] += 1; }
        const double w = wz ? wz[i] : 0.0;
        const double um = sqrt(u * u + v * v + w * w);
        a[10] = fmax(a[10], um);
        if (!(um == um) || !(rho == rho)) a[11] += 1;
    }
    __shared__ double sh[NRED][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int k = 0; k < NRED; ++k) {
        double v = a[k];
        for (int off = 32; off > 0; off >>= 1) {
            const double o = __shfl_xor(v, off, 64);
            v = k == 10 ? fmax(v, o) : v + o;
        }
        if (lane == 0) sh[k][wv] = v;
    }
    __syncthreads();
    if (threadIdx.x < NRED) {
        const int k = threadIdx.x;
        double s = sh[k][0];
        for (int w = 1; w < 4; ++w) s = k == 10 ? fmax(s, sh[k][w]) : s + sh[k][w];
        partials[(size_t)blockIdx.x * NRED + k] = s;
    }
}

}  // namespace cg

// ---------------------------------------------------------------------------
// launchers (lbm_cg.h)
hipError_t cg_launch_step(int storage, int ho, const void* src, void* dst, const cg::CgParams& cp, double* mac,
                          bool wm, hipStream_t st) {
    dim3 grid((cp.nx + cg::TX - 1) / cg::TX, (cp.ny + cg::TY - 1) / cg::TY), block(cg::TX, 8);
    const int w = wm ? 1 : 0;
#define CG_STEP(S_, H_, E_)                                                                                     \
    hipLaunchKernelGGL((cg::k_cg_step<S_, H_, E_>), grid, block, 0, st, (const typename cg::CIo<S_>::raw*)src, \
                       (typename cg::CIo<S_>::raw*)dst, cp, mac, w)
#define CG_STEP_E(S_, H_) do { if (cp.enhanced) CG_STEP(S_, H_, true); else CG_STEP(S_, H_, false); } while (0)
    if (storage == 0) { if (ho) CG_STEP_E(0, 2); else CG_STEP_E(0, 1); }
    else if (storage == 1) { if (ho) CG_STEP_E(1, 2); else CG_STEP_E(1, 1); }
    else { if (ho) CG_STEP_E(2, 2); else CG_STEP_E(2, 1); }
#undef CG_STEP_E
#undef CG_STEP
    return hipGetLastError();
}

hipError_t cg_launch_init(int storage, int ho, void* dst, const cg::CgParams& cp, const double* rr, const double* rb,
                          const double* rf, const double* bf, double* mac, hipStream_t st) {
    dim3 grid((cp.nx + 255) / 256, cp.ny), block(256);
#define CG_INIT(S_, H_, E_)                                                                                       \
    hipLaunchKernelGGL((cg::k_cg_init<S_, H_, E_>), grid, block, 0, st, (typename cg::CIo<S_>::raw*)dst, cp, rr, \
                       rb, rf, bf, mac)
#define CG_INIT_E(S_, H_) do { if (cp.enhanced) CG_INIT(S_, H_, true); else CG_INIT(S_, H_, false); } while (0)
    if (storage == 0) { if (ho) CG_INIT_E(0, 2); else CG_INIT_E(0, 1); }
    else if (storage == 1) { if (ho) CG_INIT_E(1, 2); else CG_INIT_E(1, 1); }
    else { if (ho) CG_INIT_E(2, 2); else CG_INIT_E(2, 1); }
#undef CG_INIT_E
#undef CG_INIT
    return hipGetLastError();
}

hipError_t cg_launch_recon(int storage, const void* src, const cg::CgParams& cp, double* rf, double* bf,
                           hipStream_t st) {
    dim3 grid((cp.nx + 255) / 256, cp.ny), block(256);
    if (storage == 0) hipLaunchKernelGGL((cg::k_cg_recon<0>), grid, block, 0, st, (const double*)src, cp, rf, bf);
    else if (storage == 1) hipLaunchKernelGGL((cg::k_cg_recon<1>), grid, block, 0, st, (const float*)src, cp, rf, bf);
    else hipLaunchKernelGGL((cg::k_cg_recon<2>), grid, block, 0, st, (const uint16_t*)src, cp, rf, bf);
    return hipGetLastError();
}

hipError_t cg_launch_dens(int q, const double* rf, const double* bf, size_t n, double* rr, double* rb,
                          hipStream_t st) {
    hipLaunchKernelGGL(cg::k_cg_dens, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, q, rf, bf, n, rr, rb);
    return hipGetLastError();
}

hipError_t cg_launch_reduce(const double* mac, size_t n, bool three_d, double kappa, double* partials, int nblocks,
                            hipStream_t st) {
    hipLaunchKernelGGL(cg::k_cg_reduce, dim3(nblocks), dim3(256), 0, st, mac, three_d ? mac + 5 * n : nullptr, n,
                       kappa, partials);
    return hipGetLastError();
}

}  // namespace lbm
