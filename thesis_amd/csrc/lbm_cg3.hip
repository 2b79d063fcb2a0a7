// Colour-gradient two-phase step for D3Q19 (periodic box).
//
// Follows the reference's GPU formulation in 3D: gradient of the phase field
// (rho_r - rho_b)/rho with the 18-neighbour stencil or the 104-neighbour one
// (calculateColorGradient3D / calculateHOColorGradient3D case 0,
// S/GpuCollision.cu:283-420), BGK collision towards the two-phase
// equilibrium with effective viscosity and rest fraction, perturbation and
// recolouring (gpuCollBgkwGC3D, S/GpuCollision.cu:709-779), pull streaming of
// both colours (gpuStreaming3DCG, S/GpuStream.cu:196-285) and moments
// (gpuUpdateMacro3DCG, S/GpuUpdateMacro.cu:215-361).  The stored state is the
// post-collision colours, as in the single-phase step.
//
// Per link the storage holds h = f - w and the red population r (b = f - r),
// 38 planes.  With the 104-point gradient (and on boxes the single pass does
// not tile) a step is two launches: k_cg3_psi pulls each node's
// post-streaming colours and stores the phase field psi (and the density,
// for the enhanced model's density gradient); k_cg3_step pulls
// them again, takes the gradient of psi over its neighbours (served from
// L1/L2: each psi value is read by 18 or 104 nodes), collides and recolours
// in registers and stores the new post-collision state.  With the 18-point
// gradient k_cg3_fz does both in one pass over z-columns of row tiles.
#include "lbm_cg.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {
namespace cg3 {

using cg::CgParams;

constexpr int BX = 64, BY = 4;   // one wavefront per x row segment, 4 rows per block
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
        const float s = red ? (float)W19(i) : (float)(W19(i) / 1024.0);
        return (float)__builtin_bit_cast(_Float16, p[k]) * s;
    }
    __device__ static void st(raw* p, uint32_t k, T v, int i, bool red) {
        const float s = red ? (float)(1.0 / W19(i)) : (float)(1024.0 / W19(i));
        __builtin_nontemporal_store(__builtin_bit_cast(uint16_t, (_Float16)(v * s)), p + k);
    }
};

__device__ inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// Post-streaming colours of node (x, y, z): pull from the post-collision
// buffer with periodic wrap.
template <int S>
__device__ inline void pull(const typename CIo<S>::raw* __restrict__ src, const CgParams& cp, int x, int y, int z,
                            typename CIo<S>::T* h, typename CIo<S>::T* r) {
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int sx = wrap(x - CX[i], cp.nx), sy = wrap(y - CY[i], cp.ny), sz = wrap(z - CZ[i], cp.nz);
        const uint32_t k = ((uint32_t)sz * (uint32_t)cp.ny + (uint32_t)sy) * (uint32_t)cp.nx + (uint32_t)sx;
        h[i] = CIo<S>::ld(src, (uint32_t)i * cp.plane + k, i, false);
        r[i] = CIo<S>::ld(src, (uint32_t)(19 + i) * cp.plane + k, i, true);
    }
}

template <class T>
__device__ inline void densities(const T* h, const T* r, T& rho, T& rr) {
    T d = 0, s = 0;
#pragma unroll
    for (int i = 0; i < 19; ++i) { d += h[i]; s += r[i]; }
    rho = (T)1 + d;
    rr = s;
}

// Collision, perturbation and recolouring of one node (gpuCollBgkwGC3D;
// with E the density-gradient terms of gpuCollEnhancedBgkwGC3D,
// S/GpuCollision.cu:781-880, from the density gradient d): new deviation
// h' = f* - w and red r'; mac = rho_r, rho_b, rho, u, v, w of the incoming
// state.
template <class T, bool E, class Out>
__device__ inline void collide_to(const CgParams& cp, const T* h, const T* r, T gx, T gy, T gz, const T* d, Out out,
                                  T* mac) {
    T rho, rr;
    densities<T>(h, r, rho, rr);
    const T rb = rho - rr;
    T jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 1; i < 19; ++i) {   // (zero components left out: 0 * x is not folded)
        if (CX[i]) jx += CX[i] > 0 ? h[i] : -h[i];
        if (CY[i]) jy += CY[i] > 0 ? h[i] : -h[i];
        if (CZ[i]) jz += CZ[i] > 0 ? h[i] : -h[i];
    }
    const T ir = (T)1 / rho;
    const T u = jx * ir, v = jy * ir, w = jz * ir;
    const T mean_nu = (T)1 / (rr / (rho * (T)cp.r_nu) + rb / (rho * (T)cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T alpha = (T)cp.r_alpha * rr * ir + (T)cp.b_alpha * rb * ir;
    const T gn = sqrt(gx * gx + gy * gy + gz * gz);
    const bool on = gn > (T)cp.g_limit;
    const T ign = on ? (T)1 / gn : (T)0;
    const T agn = on ? (T)cp.A * gn : (T)0;
    const T kr = rr * ir, kk = (T)cp.beta * rr * rb * ir;
    const T uu = u * u + v * v + w * w;
    const T ugr = E ? u * d[0] + v * d[1] + w * d[2] : (T)0;
    // The node's factors of each speed shell (rest, axis, diagonal), so that a
    // link adds only what depends on its own c:
    //   rest    = phi + theta alpha
    //   base    = f_eq - w at u = 0 = (rho - 1) rest + (rest - w)
    //   pa, pb  : perturbation A |g| (w (c.g)^2 / |g|^2 - B) = pa (c.g)^2 - pb
    //   ic      : cos = (c.g) ic, ic = 1 / (|g| |c|)
    //   kw, kq  : red = kr (f* - w) + kr w + kk cos rest = kr (f* - w) + kw + kq cos
    T rest[3], base[3], rw[3], pa[3], pb[3], ic[3], kw[3], kq[3], ec[3], ep[3];
#pragma unroll
    for (int sh = 0; sh < 3; ++sh) {
        const int k = sh == 0 ? 0 : (sh == 1 ? 1 : 7);
        rest[sh] = (T)PHI(k) + (T)TETA(k) * alpha;
        base[sh] = (rho - (T)1) * rest[sh] + (rest[sh] - (T)W19(k));
        rw[sh] = rho * (T)W19(k);
        pa[sh] = agn * ign * ign * (T)W19(k);
        pb[sh] = agn * (T)BPERT(k);
        ic[sh] = ign * (T)(1.0 / CNORM(k));
        kw[sh] = kr * (T)W19(k);
        kq[sh] = kk * rest[sh];
        // enhanced model: nu (chi u.grad rho + psi G:cc), G:cc = 2 (c.u)(c.grad rho)
        ec[sh] = E ? mean_nu * (T)CHI(k) * ugr : (T)0;
        ep[sh] = E ? mean_nu * (T)PSI(k) * (T)2 : (T)0;
    }
    const T omc = (T)1 - om, m15uu = (T)-1.5 * uu;
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        const int sh = k == 0 ? 0 : (k <= 6 ? 1 : 2);
        const T cg = cg::cdot<T>(CX[k], CY[k], CZ[k], gx, gy, gz);
        const T cs = k ? cg * ic[sh] : (T)0;
        const T pert = pa[sh] * (cg * cg) - pb[sh];
        const T cu = cg::cdot<T>(CX[k], CY[k], CZ[k], u, v, w);
        // f_eq - w = base + rho w (3cu + 4.5cu^2 - 1.5uu)
        T feqd = rw[sh] * (cu * ((T)3 + (T)4.5 * cu) + m15uu) + base[sh];
        if (E) {
            const T cgr = cg::cdot<T>(CX[k], CY[k], CZ[k], d[0], d[1], d[2]);
            feqd += ec[sh] + ep[sh] * cu * cgr;
        }
        const T fsd = om * feqd + (omc * h[k] + pert);               // f* - w
        out(k, fsd, kr * fsd + (kw[sh] + kq[sh] * cs));
    }
    mac[0] = rr; mac[1] = rb; mac[2] = rho; mac[3] = u; mac[4] = v; mac[5] = w;
}

// the same into arrays: ho[k] = f*_k - w_k, ro[k] = red
template <class T, bool E>
__device__ inline void collide(const CgParams& cp, const T* h, const T* r, T gx, T gy, T gz, const T* d, T* ho, T* ro,
                               T* mac) {
    collide_to<T, E>(cp, h, r, gx, gy, gz, d, [&](int k, T hv, T rv) { ho[k] = hv; ro[k] = rv; }, mac);
}

// Gradient of the phase field at (x, y, z): psi(i) returns the phase field of
// linear node index i.
// SKIP leaves out the products with a zero offset component (the compiler does
// not fold 0 * x; same result for finite values).  Measured at 256^3 with the
// 104-point stencil: 7 % faster with 16-bit storage, 5 % slower with FP32.
template <int H, class T, bool SKIP = false, class Psi>
__device__ inline void gradient(const CgParams& cp, int x, int y, int z, Psi psi, T& gx, T& gy, T& gz) {
    int xs[2 * H + 1], ys[2 * H + 1], zs[2 * H + 1];
#pragma unroll
    for (int d = -H; d <= H; ++d) {
        xs[d + H] = wrap(x + d, cp.nx);
        ys[d + H] = wrap(y + d, cp.ny);
        zs[d + H] = wrap(z + d, cp.nz);
    }
    T ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int dz = -H; dz <= H; ++dz)
#pragma unroll
        for (int dy = -H; dy <= H; ++dy)
#pragma unroll
            for (int dx = -H; dx <= H; ++dx) {
                const int c2 = dx * dx + dy * dy + dz * dz;
                const double wk = H == 1 ? LOW(c2) : HOW(c2);
                if (wk == 0.0) continue;
                const uint32_t j = ((uint32_t)zs[dz + H] * (uint32_t)cp.ny + (uint32_t)ys[dy + H]) * (uint32_t)cp.nx +
                                   (uint32_t)xs[dx + H];
                const T p = (T)wk * psi(j);
                if (!SKIP || dx) ax += (T)dx * p;
                if (!SKIP || dy) ay += (T)dy * p;
                if (!SKIP || dz) az += (T)dz * p;
            }
    gx = ax; gy = ay; gz = az;
}

// The workgroup's (x, y, z) block from its dispatch order renumbered by
// xcd_block, so that the row blocks an XCD runs together are neighbours in y
// and share their neighbour rows' lines in that XCD's L2.
__device__ inline bool node_of(const CgParams& cp, int& x, int& y, int& z) {
    const uint32_t gx = gridDim.x, gxy = gx * gridDim.y;
    const uint32_t b = xcd_block(blockIdx.x + gx * blockIdx.y + gxy * blockIdx.z, gxy * gridDim.z, cp.xcd_shift,
                                 cp.xchunk_shift);
    const uint32_t bz = b / gxy, rxy = b - bz * gxy, by = rxy / gx;
    x = (int)((rxy - by * gx) * BX + threadIdx.x);
    y = (int)(by * BY + threadIdx.y);
    z = (int)bz;
    return x < cp.nx && y < cp.ny;
}

// Phase field of the post-streaming state (and with E its density in the
// second half of psi).
template <int S, bool E>
__global__ __launch_bounds__(BX * BY) void k_cg3_psi(const typename CIo<S>::raw* __restrict__ src, CgParams cp,
                                                     typename CIo<S>::T* __restrict__ psi) {
    using T = typename CIo<S>::T;
    int x, y, z;
    if (!node_of(cp, x, y, z)) return;
    T h[19], r[19], rho, rr;
    pull<S>(src, cp, x, y, z, h, r);
    densities<T>(h, r, rho, rr);
    const uint32_t j = ((uint32_t)z * (uint32_t)cp.ny + (uint32_t)y) * (uint32_t)cp.nx + (uint32_t)x;
    psi[j] = ((T)2 * rr - rho) / rho;
    if (E) psi[(size_t)cp.nx * cp.ny * cp.nz + j] = rho;
}

template <int S, int H, bool E>
__global__ __launch_bounds__(BX * BY) void k_cg3_step(const typename CIo<S>::raw* __restrict__ src,
                                                      typename CIo<S>::raw* __restrict__ dst,
                                                      const typename CIo<S>::T* __restrict__ psi, CgParams cp,
                                                      double* __restrict__ mac, int write_macros) {
    using T = typename CIo<S>::T;
    int x, y, z;
    if (!node_of(cp, x, y, z)) return;
    T h[19], r[19];
    pull<S>(src, cp, x, y, z, h, r);
    T gx, gy, gz, d[3] = {0, 0, 0};
    gradient<H, T, (S >= 2)>(cp, x, y, z, [psi](uint32_t j) { return psi[j]; }, gx, gy, gz);
    if (E) {
        const T* rh = psi + (size_t)cp.nx * cp.ny * cp.nz;
        gradient<H, T, (S >= 2)>(cp, x, y, z, [rh](uint32_t j) { return rh[j]; }, d[0], d[1], d[2]);
    }
    T ho[19], ro[19], m[6];
    collide<T, E>(cp, h, r, gx, gy, gz, d, ho, ro, m);
    const uint32_t idx = ((uint32_t)z * (uint32_t)cp.ny + (uint32_t)y) * (uint32_t)cp.nx + (uint32_t)x;
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        CIo<S>::st(dst, (uint32_t)i * cp.plane + idx, ho[i], i, false);
        CIo<S>::st(dst, (uint32_t)(19 + i) * cp.plane + idx, ro[i], i, true);
    }
    if (write_macros) {
        const size_t n = (size_t)cp.nx * cp.ny * cp.nz;
        for (int a = 0; a < 6; ++a) mac[a * n + idx] = (double)m[a];
    }
}

// Single-pass step for the 18-point gradient.  A workgroup owns a 64 x TY
// column of rows and walks up zchunk planes of it.  For each plane it pulls
// the post-streaming colours of its nodes and of the one-node ring around
// them, puts their phase field (and with E their density) into a ring of four
// LDS planes, and keeps its own nodes' populations in registers; once the
// plane above is in, it finishes the plane below it: gradient from LDS,
// collision, store.  Every population is read once, plus the ring nodes'
// (2 (64 + TY) + 4 per plane) and two extra planes per chunk, instead of
// twice.  The gradient sums in the order of gradient<1>.  Needs nx % 64 == 0
// and ny % TY == 0.
constexpr int FX = 64;

template <int TY> struct Frame {
    static constexpr int W = FX + 2, H = TY + 2, N = W * H, RING = N - FX * TY;
    // LDS cell of ring node e: the first and last rows of the frame, then the
    // first and last columns of the rows between
    __device__ static int cell(int e, int& lx, int& ly) {
        if (e < W) { lx = e; ly = 0; }
        else if (e < 2 * W) { lx = e - W; ly = H - 1; }
        else { const int c = e - 2 * W; ly = 1 + (c >> 1); lx = (c & 1) ? W - 1 : 0; }
        return ly * W + lx;
    }
};

// gradient<1, T, SKIP> over three LDS planes (below, at, above the node) of
// row width FW
template <int FW, class T, bool SKIP>
__device__ inline void gradient_lds(const T* p0, const T* p1, const T* p2, int l, T& gx, T& gy, T& gz) {
    const T* pl[3] = {p0, p1, p2};
    T ax = 0, ay = 0, az = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                const int c2 = dx * dx + dy * dy + dz * dz;
                const double wk = LOW(c2);
                if (wk == 0.0) continue;
                const T p = (T)wk * pl[dz + 1][l + dy * FW + dx];
                if (!SKIP || dx) ax += (T)dx * p;
                if (!SKIP || dy) ay += (T)dy * p;
                if (!SKIP || dz) az += (T)dz * p;
            }
    gx = ax; gy = ay; gz = az;
}

template <int S, bool E, int TY>
__global__ __launch_bounds__(FX * TY) void k_cg3_fz(const typename CIo<S>::raw* __restrict__ src,
                                                    typename CIo<S>::raw* __restrict__ dst, CgParams cp,
                                                    double* __restrict__ mac, int write_macros) {
    using T = typename CIo<S>::T;
    using F = Frame<TY>;
    __shared__ T lpsi[4][F::N];
    __shared__ T lrho[E ? 4 : 1][E ? F::N : 1];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * FX + tx;
    const uint32_t tiles_x = (uint32_t)cp.nx / FX, txy = tiles_x * ((uint32_t)cp.ny / TY);
    const uint32_t b = xcd_block(blockIdx.x, gridDim.x, cp.xcd_shift, cp.xchunk_shift);
    const uint32_t chunk = b / txy, rxy = b - chunk * txy, tyi = rxy / tiles_x, txi = rxy - tyi * tiles_x;
    const int x0 = (int)txi * FX, y0 = (int)tyi * TY, z0 = (int)chunk * cp.zchunk;
    const int z1 = min(z0 + cp.zchunk, cp.nz);
    const int x = x0 + tx, y = y0 + ty, own = (ty + 1) * F::W + tx + 1;
    const bool ring = tid < F::RING;
    int rx = 0, ry = 0, rl = 0;
    if (ring) {
        int lx, ly;
        rl = F::cell(tid, lx, ly);
        rx = wrap(x0 - 1 + lx, cp.nx);
        ry = wrap(y0 - 1 + ly, cp.ny);
    }
    // plane z goes to LDS slot (z - z0 + 1) & 3
    auto put = [&](int slot, int l, const T* h, const T* r) {
        T rho, rr;
        densities<T>(h, r, rho, rr);
        lpsi[slot][l] = ((T)2 * rr - rho) / rho;
        if (E) lrho[slot][l] = rho;
    };
    auto put_ring = [&](int slot, int z) {
        if (!ring) return;
        T h[19], r[19];
        pull<S>(src, cp, rx, ry, z, h, r);
        put(slot, rl, h, r);
    };
    T ch[19], cr[19];
    {   // the plane below the chunk: phase field only
        const int zb = wrap(z0 - 1, cp.nz);
        put_ring(0, zb);
        pull<S>(src, cp, x, y, zb, ch, cr);
        put(0, own, ch, cr);
    }
    put_ring(1, z0);
    pull<S>(src, cp, x, y, z0, ch, cr);
    put(1, own, ch, cr);
    const uint32_t xy = (uint32_t)y * (uint32_t)cp.nx + (uint32_t)x, nxy = (uint32_t)cp.nx * (uint32_t)cp.ny;
    for (int z = z0; z < z1; ++z) {
        const int za = z + 1 == cp.nz ? 0 : z + 1, s0 = (z - z0) & 3, s1 = (z - z0 + 1) & 3, s2 = (z - z0 + 2) & 3;
        put_ring(s2, za);
        T nh[19], nr[19];
        pull<S>(src, cp, x, y, za, nh, nr);
        put(s2, own, nh, nr);
        // slot s2 was last read two planes ago, before the previous barrier
        __syncthreads();
        T gx, gy, gz, d[3] = {0, 0, 0};
        gradient_lds<F::W, T, (S >= 2)>(lpsi[s0], lpsi[s1], lpsi[s2], own, gx, gy, gz);
        if (E) gradient_lds<F::W, T, (S >= 2)>(lrho[s0], lrho[s1], lrho[s2], own, d[0], d[1], d[2]);
        T m[6];
        const uint32_t idx = (uint32_t)z * nxy + xy;
        collide_to<T, E>(cp, ch, cr, gx, gy, gz, d, [&](int i, T hv, T rv) {
            CIo<S>::st(dst, (uint32_t)i * cp.plane + idx, hv, i, false);
            CIo<S>::st(dst, (uint32_t)(19 + i) * cp.plane + idx, rv, i, true);
        }, m);
        if (write_macros) {
            const size_t n = (size_t)nxy * cp.nz;
            for (int a = 0; a < 6; ++a) mac[a * n + idx] = (double)m[a];
        }
#pragma unroll
        for (int i = 0; i < 19; ++i) { ch[i] = nh[i]; cr[i] = nr[i]; }
    }
}

// Phase field and density of colour densities.
template <class T>
__global__ __launch_bounds__(BX * BY) void k_cg3_psi0(const double* __restrict__ rr0, const double* __restrict__ rb0,
                                                      CgParams cp, T* __restrict__ psi) {
    int x, y, z;
    if (!node_of(cp, x, y, z)) return;
    const uint32_t j = ((uint32_t)z * (uint32_t)cp.ny + (uint32_t)y) * (uint32_t)cp.nx + (uint32_t)x;
    psi[j] = (T)((rr0[j] - rb0[j]) / (rr0[j] + rb0[j]));
    psi[(size_t)cp.nx * cp.ny * cp.nz + j] = (T)(rr0[j] + rb0[j]);
}

// First post-collision state from colour densities rr0, rb0 (rest
// equilibria phi + theta alpha) or, when rf is not null, from the colour
// populations rf, bf (FP64, 19 planes of N; rr0, rb0 their densities);
// psi holds the phase field of rr0, rb0.
template <int S, int H, bool E>
__global__ __launch_bounds__(BX * BY) void k_cg3_init(typename CIo<S>::raw* __restrict__ dst, CgParams cp,
                                                      const double* __restrict__ rr0, const double* __restrict__ rb0,
                                                      const double* __restrict__ rf, const double* __restrict__ bf,
                                                      const typename CIo<S>::T* __restrict__ psi,
                                                      double* __restrict__ mac) {
    using T = typename CIo<S>::T;
    int x, y, z;
    if (!node_of(cp, x, y, z)) return;
    const size_t n = (size_t)cp.nx * cp.ny * cp.nz;
    T gx, gy, gz, d[3] = {0, 0, 0};
    gradient<H, T>(cp, x, y, z, [psi](uint32_t j) { return psi[j]; }, gx, gy, gz);
    if (E) {
        const T* rh = psi + n;
        gradient<H, T>(cp, x, y, z, [rh](uint32_t j) { return rh[j]; }, d[0], d[1], d[2]);
    }
    const uint32_t idx = ((uint32_t)z * (uint32_t)cp.ny + (uint32_t)y) * (uint32_t)cp.nx + (uint32_t)x;
    const T rr = (T)rr0[idx], rb = (T)rb0[idx];
    T h[19], r[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        if (rf) {
            r[i] = (T)rf[i * n + idx];
            h[i] = (T)(rf[i * n + idx] + bf[i * n + idx] - W19(i));
        } else {
            const T rr_i = rr * (T)(PHI(i) + TETA(i) * cp.r_alpha), rb_i = rb * (T)(PHI(i) + TETA(i) * cp.b_alpha);
            r[i] = rr_i;
            h[i] = rr_i + rb_i - (T)W19(i);
        }
    }
    T ho[19], ro[19], m[6];
    collide<T, E>(cp, h, r, gx, gy, gz, d, ho, ro, m);
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        CIo<S>::st(dst, (uint32_t)i * cp.plane + idx, ho[i], i, false);
        CIo<S>::st(dst, (uint32_t)(19 + i) * cp.plane + idx, ro[i], i, true);
    }
    mac[idx] = (double)rr; mac[n + idx] = (double)rb; mac[2 * n + idx] = (double)(rr + rb);
    for (int a = 3; a < 6; ++a) mac[a * n + idx] = rf ? (double)m[a] : 0.0;
}

// Post-streaming colour populations of the state in src (FP64).
template <int S>
__global__ __launch_bounds__(BX * BY) void k_cg3_recon(const typename CIo<S>::raw* __restrict__ src, CgParams cp,
                                                       double* __restrict__ rf, double* __restrict__ bf) {
    using T = typename CIo<S>::T;
    int x, y, z;
    if (!node_of(cp, x, y, z)) return;
    T h[19], r[19];
    pull<S>(src, cp, x, y, z, h, r);
    const size_t n = (size_t)cp.nx * cp.ny * cp.nz;
    const size_t idx = ((size_t)z * cp.ny + y) * cp.nx + x;
    for (int i = 0; i < 19; ++i) {
        rf[i * n + idx] = (double)r[i];
        bf[i * n + idx] = (double)h[i] + W19(i) - (double)r[i];
    }
}

inline dim3 grid_of(const CgParams& cp) { return dim3((cp.nx + BX - 1) / BX, (cp.ny + BY - 1) / BY, cp.nz); }

}  // namespace cg3

// ---------------------------------------------------------------------------
// launchers (lbm_cg.h)
hipError_t cg3_launch_step(int storage, int ho, const void* src, void* dst, void* psi, const cg::CgParams& cp,
                           double* mac, bool wm, hipStream_t st) {
    const dim3 grid = cg3::grid_of(cp), block(cg3::BX, cg3::BY);
    const int w = wm ? 1 : 0;
    if (!ho && cp.zchunk > 0) {
        const int ty = cp.zrows;
        if ((ty != 4 && ty != 8) || cp.nx % cg3::FX || cp.ny % ty) return hipErrorInvalidValue;
        const uint32_t nb = (uint32_t)(cp.nx / cg3::FX) * (uint32_t)(cp.ny / ty) *
                            (uint32_t)((cp.nz + cp.zchunk - 1) / cp.zchunk);
        const dim3 fgrid(nb), fblock(cg3::FX, ty);
#define CG3_FZT_(S_, E_, TY_)                                                                                       \
    hipLaunchKernelGGL((cg3::k_cg3_fz<S_, E_, TY_>), fgrid, fblock, 0, st,                                        \
                       (const typename cg3::CIo<S_>::raw*)src, (typename cg3::CIo<S_>::raw*)dst, cp, mac, w)
#define CG3_FZ_(S_, E_) do { if (ty == 8) CG3_FZT_(S_, E_, 8); else CG3_FZT_(S_, E_, 4); } while (0)
#define CG3_FZ(S_) do { if (cp.enhanced) CG3_FZ_(S_, true); else CG3_FZ_(S_, false); } while (0)
        if (storage == 0) CG3_FZ(0);
        else if (storage == 1) CG3_FZ(1);
        else CG3_FZ(2);
#undef CG3_FZ
#undef CG3_FZ_
#undef CG3_FZT_
        return hipGetLastError();
    }
#define CG3_STEP_(S_, H_, E_)                                                                                  \
    do {                                                                                                       \
        using R_ = typename cg3::CIo<S_>::raw;                                                                 \
        using T_ = typename cg3::CIo<S_>::T;                                                                   \
        hipLaunchKernelGGL((cg3::k_cg3_psi<S_, E_>), grid, block, 0, st, (const R_*)src, cp, (T_*)psi);        \
        hipLaunchKernelGGL((cg3::k_cg3_step<S_, H_, E_>), grid, block, 0, st, (const R_*)src, (R_*)dst,        \
                           (const T_*)psi, cp, mac, w);                                                        \
    } while (0)
#define CG3_STEP(S_, H_) do { if (cp.enhanced) CG3_STEP_(S_, H_, true); else CG3_STEP_(S_, H_, false); } while (0)
    if (storage == 0) { if (ho) CG3_STEP(0, 2); else CG3_STEP(0, 1); }
    else if (storage == 1) { if (ho) CG3_STEP(1, 2); else CG3_STEP(1, 1); }
    else { if (ho) CG3_STEP(2, 2); else CG3_STEP(2, 1); }
#undef CG3_STEP
#undef CG3_STEP_
    return hipGetLastError();
}

hipError_t cg3_launch_init(int storage, int ho, void* dst, void* psi, const cg::CgParams& cp, const double* rr,
                           const double* rb, const double* rf, const double* bf, double* mac, hipStream_t st) {
    const dim3 grid = cg3::grid_of(cp), block(cg3::BX, cg3::BY);
#define CG3_INIT_(S_, H_, E_)                                                                                      \
    do {                                                                                                           \
        using T_ = typename cg3::CIo<S_>::T;                                                                       \
        hipLaunchKernelGGL((cg3::k_cg3_psi0<T_>), grid, block, 0, st, rr, rb, cp, (T_*)psi);                       \
        hipLaunchKernelGGL((cg3::k_cg3_init<S_, H_, E_>), grid, block, 0, st, (typename cg3::CIo<S_>::raw*)dst, cp, \
                           rr, rb, rf, bf, (const T_*)psi, mac);                                                   \
    } while (0)
#define CG3_INIT(S_, H_) do { if (cp.enhanced) CG3_INIT_(S_, H_, true); else CG3_INIT_(S_, H_, false); } while (0)
    if (storage == 0) { if (ho) CG3_INIT(0, 2); else CG3_INIT(0, 1); }
    else if (storage == 1) { if (ho) CG3_INIT(1, 2); else CG3_INIT(1, 1); }
    else { if (ho) CG3_INIT(2, 2); else CG3_INIT(2, 1); }
#undef CG3_INIT
#undef CG3_INIT_
    return hipGetLastError();
}

hipError_t cg3_launch_recon(int storage, const void* src, const cg::CgParams& cp, double* rf, double* bf,
                            hipStream_t st) {
    const dim3 grid = cg3::grid_of(cp), block(cg3::BX, cg3::BY);
    if (storage == 0) hipLaunchKernelGGL((cg3::k_cg3_recon<0>), grid, block, 0, st, (const double*)src, cp, rf, bf);
    else if (storage == 1) hipLaunchKernelGGL((cg3::k_cg3_recon<1>), grid, block, 0, st, (const float*)src, cp, rf, bf);
    else hipLaunchKernelGGL((cg3::k_cg3_recon<2>), grid, block, 0, st, (const uint16_t*)src, cp, rf, bf);
    return hipGetLastError();
}

}  // namespace lbm
