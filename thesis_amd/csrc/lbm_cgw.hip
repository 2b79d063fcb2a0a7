// Colour-gradient D2Q9 / D3Q19 step on a box with walls on the y rows
// (planes; the other axes periodic) and a body force, in the reference's own
// order: collide (reading the stored macros), stream, wall condition,
// macroscopic update.
//
// The fused periodic step (lbm_cg.hip) recomputes every macro from the
// streamed populations; the wall cases need what that rotation drops: the
// wall condition reads each wall node's velocity from the previous step's
// macroscopic update, the body force shifts the stored velocity with the
// previous step's effective viscosity, and a wall row keeps its links from
// beyond the wall until the condition sets them.  So this path stores the
// colour populations after the wall condition (r, b: 2 x 9 planes) and after
// collision (rc, bc) and the macros, one launch per half step:
//   k_cgw_collide: gradient (mirrored on wall rows, S/GpuCollision.cu:135-156;
//     8-point next to them with the 24-point stencil, :262-275), BGK towards
//     the two-phase equilibrium, perturbation, recolouring (gpuCollBgkwGC2D /
//     gpuCollEnhancedBgkwGC2D, :552-707) -> rc, bc;
//   k_cgw_stream: pull with x wrap (gpuStreaming2DCG, S/GpuStream.cu:30-158),
//     per-colour Zou-He or bounce-back walls (gpuBcPeriodic2D,
//     S/GpuBoundaries.cu:1876-1968), moments with the force shift
//     (gpuUpdateMacro2DCG, S/GpuUpdateMacro.cu:43-135) -> r, b, macros.
// D3Q19 (k_cgw3_*) follows the same order with the 3D kernels: gradient
// mirrored on the wall faces (calculateColorGradient3D cases 1 and 2,
// S/GpuCollision.cu:306-333; 18-point next to them with the 104-point
// stencil, case 11, :505-518), gpuCollBgkwGC3D / gpuCollEnhancedBgkwGC3D
// (:709-851), gpuStreaming3DCG (S/GpuStream.cu:196-285), the single-phase
// faces applied per colour (S/Iterate3D.cu:576-598: the 3D inlet condition
// with the face velocity, or on-node bounce-back), gpuUpdateMacro3DCG
// (S/GpuUpdateMacro.cu:215-345).
// Each thread reads its neighbours only from arrays the launch does not
// write, and writes only its own node.
#include "lbm_cg.h"
#include "lbm_device.h"

#include <hip/hip_runtime.h>
#include <cstdint>

namespace lbm {
namespace cgw {

using cg::CgParams;
using cg::WallParams;

__device__ constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
__device__ constexpr int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
__host__ __device__ constexpr double W9(int i) { return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
__host__ __device__ constexpr double PHI(int i) { return i == 0 ? 0.0 : (i <= 4 ? 0.2 : 0.05); }
__host__ __device__ constexpr double TETA(int i) { return i == 0 ? 1.0 : (i <= 4 ? -0.2 : -0.05); }
__host__ __device__ constexpr double BPERT(int i) { return i == 0 ? -4.0 / 27.0 : (i <= 4 ? 2.0 / 27.0 : 5.0 / 108.0); }
__host__ __device__ constexpr double CNORM(int i) { return i == 0 ? 1.0 : (i <= 4 ? 1.0 : 1.4142135623730950488); }
__host__ __device__ constexpr double CHI(int i) { return i == 0 ? -8.0 / 3.0 : (i <= 4 ? -1.0 / 6.0 : 1.0 / 12.0); }
__host__ __device__ constexpr double PSI(int i) { return i == 0 ? 0.0 : (i <= 4 ? 0.5 : 0.125); }
__host__ __device__ constexpr double HOW(int c2) {
    return c2 == 1 ? 960.0 / 5040 : c2 == 2 ? 448.0 / 5040 : c2 == 4 ? 84.0 / 5040 : c2 == 5 ? 32.0 / 5040 : 1.0 / 5040;
}

__device__ inline int wrapx(int x, int nx) { return x < 0 ? x + nx : (x >= nx ? x - nx : x); }

// Phase-field and density gradients of node (x, y) from the stored macros
// (mac: rho_r, rho_b, rho blocks of n).
template <class T>
__device__ inline void gradient(const double* __restrict__ mac, const WallParams& wp, int x, int y, size_t n, T* g) {
    const int nx = wp.cp.nx, ny = wp.cp.ny;
    T ax = 0, ay = 0, bx = 0, by = 0;
    auto take = [&](int xx, int yy, T w, int kx, int ky) {
        const size_t j = (size_t)yy * nx + (size_t)xx;
        const T rho = (T)mac[2 * n + j];
        const T psi = w * ((T)mac[j] - (T)mac[n + j]) / rho;
        const T rw = w * rho;
        if (kx) { ax += kx > 0 ? psi : -psi; bx += kx > 0 ? rw : -rw; }
        if (ky) { ay += ky > 0 ? psi : -psi; by += ky > 0 ? rw : -rw; }
    };
    const bool wall_row = wp.walls && (y == 0 || y == ny - 1);
    const bool near = wp.walls && wp.ho && (y == 1 || y == ny - 2);
    if (wall_row || near || !wp.ho) {
        // 8-point stencil; on a wall row the rows beyond the wall are mirrored
        // onto the row inside and only the x components are taken
        const int into = y == 0 ? 1 : -1;
#pragma unroll
        for (int k = 1; k < 9; ++k) {
            const T w = (T)(k < 5 ? 4.0 / 12.0 : 1.0 / 12.0);
            int yy = y + (wall_row ? into * (CY[k] != 0) : CY[k]);
            if (!wp.walls) yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
            take(wrapx(x + CX[k], nx), yy, w, CX[k], wall_row ? 0 : CY[k]);
        }
    } else {
        // 24-point stencil (rows at least two away from any wall)
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                const int c2 = dx * dx + dy * dy;
                if (c2 == 0) continue;
                int yy = y + dy;
                if (!wp.walls) yy = yy < 0 ? yy + ny : (yy >= ny ? yy - ny : yy);
                const int xx = x + dx < 0 ? x + dx + nx : (x + dx >= nx ? x + dx - nx : x + dx);
                const size_t j = (size_t)yy * nx + (size_t)xx;
                const T rho = (T)mac[2 * n + j];
                const T w = (T)HOW(c2);
                const T psi = w * ((T)mac[j] - (T)mac[n + j]) / rho;
                const T rw = w * rho;
                ax += psi * (T)dx; ay += psi * (T)dy;
                bx += rw * (T)dx; by += rw * (T)dy;
            }
    }
    g[0] = ax; g[1] = ay; g[2] = bx; g[3] = by;
}

template <class T, bool E>
__global__ __launch_bounds__(256) void k_cgw_collide(const T* __restrict__ P, T* __restrict__ Pc,
                                                     const double* __restrict__ mac, WallParams wp) {
    const int nx = wp.cp.nx, ny = wp.cp.ny;
    const size_t n = (size_t)nx * ny;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    const int x = (int)(ind % (size_t)nx), y = (int)(ind / (size_t)nx);
    const CgParams& cp = wp.cp;
    T g[4];
    gradient<T>(mac, wp, x, y, n, g);
    const T rr = (T)mac[ind], br = (T)mac[n + ind], r = (T)mac[2 * n + ind];
    const T u = (T)mac[3 * n + ind], v = (T)mac[4 * n + ind];
    const T mean_nu = (T)1 / (rr / (r * (T)cp.r_nu) + br / (r * (T)cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T alpha = (T)cp.r_alpha * rr / r + (T)cp.b_alpha * br / r;
    const T gn = sqrt(g[0] * g[0] + g[1] * g[1]);
    const T kr = rr / r, kb = br / r, kk = (T)cp.beta * rr * br / r;
    const T uu = u * u + v * v;
    const T ugr = u * g[2] + v * g[3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        T cs = 0, pert = 0;
        if (gn > (T)cp.g_limit) {
            const T cg = cg::cdot<T>(CX[k], CY[k], 0, g[0], g[1], (T)0);
            if (k) cs = cg / (gn * (T)CNORM(k));
            pert = (T)cp.A * gn * ((T)W9(k) * cg * cg / (gn * gn) - (T)BPERT(k));
        }
        const T cu = cg::cdot<T>(CX[k], CY[k], 0, u, v, (T)0);
        const T rest = (T)PHI(k) + (T)TETA(k) * alpha;
        T feq = r * (rest + (T)W9(k) * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu));
        if (E) {
            const T cgr = cg::cdot<T>(CX[k], CY[k], 0, g[2], g[3], (T)0);
            feq += mean_nu * ((T)CHI(k) * ugr + (T)PSI(k) * (T)2 * cu * cgr);
        }
        const T f = P[(size_t)k * n + ind] + P[(size_t)(9 + k) * n + ind];
        const T fs = om * feq + ((T)1 - om) * f + pert;
        Pc[(size_t)k * n + ind] = kr * fs + kk * cs * rest;
        Pc[(size_t)(9 + k) * n + ind] = kb * fs - kk * cs * rest;
    }
}

// Zou-He velocity condition of one colour (S/GpuBoundaries.cu:1888-1899, :1929-1942)
template <class T>
__device__ inline T zouhe(T* f, T u, T v, bool north) {
    if (north) {
        const T r = ((T)1 / ((T)1 + v)) * (f[0] + f[1] + f[3] + (T)2 * (f[2] + f[6] + f[5]));
        f[4] = f[2] - (T)(2.0 / 3.0) * r * v;
        f[7] = f[5] + (T)0.5 * (f[1] - f[3]) - (T)(1.0 / 6.0) * (r * v) - (T)0.5 * (r * u);
        f[8] = f[6] + (T)0.5 * (f[3] - f[1]) - (T)(1.0 / 6.0) * (r * v) + (T)0.5 * (r * u);
        return r;
    }
    const T r = ((T)1 / ((T)1 - v)) * (f[0] + f[1] + f[3] + (T)2 * (f[4] + f[7] + f[8]));
    f[2] = f[4] + (T)(2.0 / 3.0) * r * v;
    f[5] = f[7] - (T)0.5 * (f[1] - f[3]) + (T)(1.0 / 6.0) * (r * v) + (T)0.5 * (r * u);
    f[6] = f[8] + (T)0.5 * (f[1] - f[3]) + (T)(1.0 / 6.0) * (r * v) - (T)0.5 * (r * u);
    return r;
}

template <class T>
__device__ inline void bounce(T* f, bool north) {
    if (north) { f[4] = f[2]; f[7] = f[5]; f[8] = f[6]; }
    else { f[2] = f[4]; f[5] = f[7]; f[6] = f[8]; }
}

template <class T>
__global__ __launch_bounds__(256) void k_cgw_stream(const T* __restrict__ Pc, T* __restrict__ P, double* __restrict__ mac,
                                                    WallParams wp) {
    const int nx = wp.cp.nx, ny = wp.cp.ny;
    const size_t n = (size_t)nx * ny;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    const int x = (int)(ind % (size_t)nx), y = (int)(ind / (size_t)nx);
    T fr[9], fb[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        int sy = y - CY[k];
        const bool beyond = sy < 0 || sy >= ny;
        if (beyond && wp.walls) {
            // a wall row keeps its links from beyond the wall; the condition sets them
            fr[k] = P[(size_t)k * n + ind];
            fb[k] = P[(size_t)(9 + k) * n + ind];
            continue;
        }
        sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
        const size_t src = (size_t)sy * nx + (size_t)wrapx(x - CX[k], nx);
        fr[k] = Pc[(size_t)k * n + src];
        fb[k] = Pc[(size_t)(9 + k) * n + src];
    }
    T rr0 = (T)mac[ind], br0 = (T)mac[n + ind], r0 = (T)mac[2 * n + ind];
    if (wp.walls && (y == 0 || y == ny - 1)) {
        const bool north = y == ny - 1;
        if (wp.walls == 1) {
            const T u = (T)mac[3 * n + ind], v = (T)mac[4 * n + ind];
            rr0 = zouhe<T>(fr, u, v, north);
            br0 = zouhe<T>(fb, u, v, north);
            r0 = rr0 + br0;
        } else {
            bounce<T>(fr, north);
            bounce<T>(fb, north);
        }
    }
    T R = 0, B = 0, jx = 0, jy = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        P[(size_t)k * n + ind] = fr[k];
        P[(size_t)(9 + k) * n + ind] = fb[k];
        R += fr[k];
        B += fb[k];
        const T f = fr[k] + fb[k];
        if (CX[k]) jx += CX[k] > 0 ? f : -f;
        if (CY[k]) jy += CY[k] > 0 ? f : -f;
    }
    // omega_eff from the densities stored before this update (S/GpuUpdateMacro.cu:62-64)
    const T mean_nu = (T)1 / (rr0 / (r0 * (T)wp.cp.r_nu) + br0 / (r0 * (T)wp.cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T r = R + B;
    const T g = (T)wp.g;
    mac[ind] = (double)R;
    mac[n + ind] = (double)B;
    mac[2 * n + ind] = (double)r;
    mac[3 * n + ind] = (double)(jx / r + (wp.external_force ? g / (r * om) : (T)0));
    mac[4 * n + ind] = (double)(jy / r + (wp.external_force ? (T)0 : g / om));
}

// Rest-state colours from colour densities (S/GpuInit.cu:323-352), or given
// colour populations; macros from them; the moving north wall's velocity
// (initInletVelocity, S/Multiphase.cu:1199-1209).
template <class T>
__global__ __launch_bounds__(256) void k_cgw_init(T* __restrict__ P, double* __restrict__ mac, const double* __restrict__ rr,
                                                  const double* __restrict__ rb, const double* __restrict__ rf,
                                                  const double* __restrict__ bf, WallParams wp) {
    const int nx = wp.cp.nx, ny = wp.cp.ny;
    const size_t n = (size_t)nx * ny;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    double R = 0, B = 0, jx = 0, jy = 0;
    for (int k = 0; k < 9; ++k) {
        const double a = rf ? rf[(size_t)k * n + ind] : rr[ind] * (PHI(k) + TETA(k) * wp.cp.r_alpha);
        const double b = bf ? bf[(size_t)k * n + ind] : rb[ind] * (PHI(k) + TETA(k) * wp.cp.b_alpha);
        P[(size_t)k * n + ind] = (T)a;
        P[(size_t)(9 + k) * n + ind] = (T)b;
        R += a;
        B += b;
        jx += CX[k] * (a + b);
        jy += CY[k] * (a + b);
    }
    mac[ind] = R;
    mac[n + ind] = B;
    mac[2 * n + ind] = R + B;
    double u = rf ? jx / (R + B) : 0.0, v = rf ? jy / (R + B) : 0.0;
    if (!rf && wp.walls == 1 && ind / (size_t)nx == (size_t)ny - 1) u = wp.u_wall;
    mac[3 * n + ind] = u;
    mac[4 * n + ind] = v;
    mac[5 * n + ind] = 0.0;
}

template <class T>
__global__ __launch_bounds__(256) void k_cgw_out(const T* __restrict__ P, size_t qn, double* __restrict__ rf,
                                                 double* __restrict__ bf) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= qn) return;
    rf[i] = (double)P[i];
    bf[i] = (double)P[qn + i];
}

// ---- D3Q19 ------------------------------------------------------------------
__device__ inline int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }
// opposite link: every component negated
__device__ constexpr int opp3(int i) {
    for (int j = 0; j < 19; ++j)
        if (cg3::CX[j] == -cg3::CX[i] && cg3::CY[j] == -cg3::CY[i] && cg3::CZ[j] == -cg3::CZ[i]) return j;
    return 0;
}

// Phase-field and density gradients of node (x, y, z) from the stored macros
// (g[0..2] colour, g[3..5] density).
template <class T>
__device__ inline void gradient3(const double* __restrict__ mac, const WallParams& wp, int x, int y, int z, size_t n,
                                 T* g) {
    const int nx = wp.cp.nx, ny = wp.cp.ny, nz = wp.cp.nz;
    T ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    auto take = [&](int xx, int yy, int zz, T w, int kx, int ky, int kz) {
        const size_t j = ((size_t)zz * ny + (size_t)yy) * nx + (size_t)xx;
        const T rho = (T)mac[2 * n + j];
        const T psi = w * ((T)mac[j] - (T)mac[n + j]) / rho;
        const T rw = w * rho;
        if (kx) { ax += kx > 0 ? psi : -psi; bx += kx > 0 ? rw : -rw; }
        if (ky) { ay += ky > 0 ? psi : -psi; by += ky > 0 ? rw : -rw; }
        if (kz) { az += kz > 0 ? psi : -psi; bz += kz > 0 ? rw : -rw; }
    };
    const bool wall_face = wp.walls && (y == 0 || y == ny - 1);
    const bool near = wp.walls && wp.ho && (y == 1 || y == ny - 2);
    if (wall_face || near || !wp.ho) {
        // 18-point stencil; on a wall face the plane beyond the wall is
        // mirrored onto the plane inside and only the x and z components are taken
        const int into = y == 0 ? 1 : -1;
#pragma unroll
        for (int k = 1; k < 19; ++k) {
            const T w = (T)(k <= 6 ? 1.0 / 6.0 : 1.0 / 12.0);
            int yy = y + (wall_face ? into * (cg3::CY[k] != 0) : cg3::CY[k]);
            if (!wp.walls) yy = wrap(yy, ny);
            take(wrap(x + cg3::CX[k], nx), yy, wrap(z + cg3::CZ[k], nz), w, cg3::CX[k], wall_face ? 0 : cg3::CY[k],
                 cg3::CZ[k]);
        }
    } else {
        // 104-point stencil (planes at least two away from any wall)
        for (int dz = -2; dz <= 2; ++dz)
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    const double wk = cg3::HOW(dx * dx + dy * dy + dz * dz);
                    if (wk == 0.0) continue;
                    int yy = y + dy;
                    if (!wp.walls) yy = wrap(yy, ny);
                    const size_t j = ((size_t)wrap(z + dz, nz) * ny + (size_t)yy) * nx + (size_t)wrap(x + dx, nx);
                    const T rho = (T)mac[2 * n + j];
                    const T psi = (T)wk * ((T)mac[j] - (T)mac[n + j]) / rho;
                    const T rw = (T)wk * rho;
                    ax += psi * (T)dx; ay += psi * (T)dy; az += psi * (T)dz;
                    bx += rw * (T)dx; by += rw * (T)dy; bz += rw * (T)dz;
                }
    }
    g[0] = ax; g[1] = ay; g[2] = az;
    g[3] = bx; g[4] = by; g[5] = bz;
}

template <class T, bool E>
__global__ __launch_bounds__(256) void k_cgw3_collide(const T* __restrict__ P, T* __restrict__ Pc,
                                                      const double* __restrict__ mac, WallParams wp) {
    const int nx = wp.cp.nx, ny = wp.cp.ny;
    const size_t n = (size_t)nx * ny * wp.cp.nz;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    const int x = (int)(ind % (size_t)nx), y = (int)(ind / (size_t)nx % (size_t)ny);
    const int z = (int)(ind / ((size_t)nx * ny));
    const CgParams& cp = wp.cp;
    T g[6];
    gradient3<T>(mac, wp, x, y, z, n, g);
    const T rr = (T)mac[ind], br = (T)mac[n + ind], r = (T)mac[2 * n + ind];
    const T u = (T)mac[3 * n + ind], v = (T)mac[4 * n + ind], w = (T)mac[5 * n + ind];
    const T mean_nu = (T)1 / (rr / (r * (T)cp.r_nu) + br / (r * (T)cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T alpha = (T)cp.r_alpha * rr / r + (T)cp.b_alpha * br / r;
    const T gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    const T kr = rr / r, kb = br / r, kk = (T)cp.beta * rr * br / r;
    const T uu = u * u + v * v + w * w;
    const T ugr = u * g[3] + v * g[4] + w * g[5];
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        T cs = 0, pert = 0;
        if (gn > (T)cp.g_limit) {
            const T cg = cg::cdot<T>(cg3::CX[k], cg3::CY[k], cg3::CZ[k], g[0], g[1], g[2]);
            if (k) cs = cg / (gn * (T)cg3::CNORM(k));
            pert = (T)cp.A * gn * ((T)cg3::W19(k) * cg * cg / (gn * gn) - (T)cg3::BPERT(k));
        }
        const T cu = cg::cdot<T>(cg3::CX[k], cg3::CY[k], cg3::CZ[k], u, v, w);
        const T rest = (T)cg3::PHI(k) + (T)cg3::TETA(k) * alpha;
        T feq = r * (rest + (T)cg3::W19(k) * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu));
        if (E) {
            const T cgr = cg::cdot<T>(cg3::CX[k], cg3::CY[k], cg3::CZ[k], g[3], g[4], g[5]);
            feq += mean_nu * ((T)cg3::CHI(k) * ugr + (T)cg3::PSI(k) * (T)2 * cu * cgr);
        }
        const T f = P[(size_t)k * n + ind] + P[(size_t)(19 + k) * n + ind];
        const T fs = om * feq + ((T)1 - om) * f + pert;
        Pc[(size_t)k * n + ind] = kr * fs + kk * cs * rest;
        Pc[(size_t)(19 + k) * n + ind] = kb * fs - kk * cs * rest;
    }
}

// One colour's populations on a y wall face: the 3D inlet condition with the
// face velocity, or on-node bounce-back (each unknown link takes its
// opposite's streamed value, the STRAIGHT case of gpuBcSimpleWall3D,
// S/GpuBoundaries.cu:482-484).
template <class T>
__device__ inline void wall3(T* f, int walls, bool north, T u_wall) {
    if (walls == 1) {
        const T U[3] = {north ? u_wall : (T)0, (T)0, (T)0};
        if (north) vface<D3Q19, T, 1, -1>(f, U, false);
        else vface<D3Q19, T, 1, 1>(f, U, false);
        return;
    }
#pragma unroll
    for (int k = 1; k < 19; ++k)
        if (cg3::CY[k] == (north ? -1 : 1)) f[k] = f[opp3(k)];
}

template <class T>
__global__ __launch_bounds__(256) void k_cgw3_stream(const T* __restrict__ Pc, T* __restrict__ P,
                                                     double* __restrict__ mac, WallParams wp) {
    const int nx = wp.cp.nx, ny = wp.cp.ny, nz = wp.cp.nz;
    const size_t n = (size_t)nx * ny * nz;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    const int x = (int)(ind % (size_t)nx), y = (int)(ind / (size_t)nx % (size_t)ny);
    const int z = (int)(ind / ((size_t)nx * ny));
    T fr[19], fb[19];
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        const int sy = y - cg3::CY[k];
        if ((sy < 0 || sy >= ny) && wp.walls) {
            // a wall face keeps its links from beyond the wall; the condition sets them
            fr[k] = P[(size_t)k * n + ind];
            fb[k] = P[(size_t)(19 + k) * n + ind];
            continue;
        }
        const size_t src = ((size_t)wrap(z - cg3::CZ[k], nz) * ny + (size_t)wrap(sy, ny)) * nx +
                           (size_t)wrap(x - cg3::CX[k], nx);
        fr[k] = Pc[(size_t)k * n + src];
        fb[k] = Pc[(size_t)(19 + k) * n + src];
    }
    if (wp.walls && (y == 0 || y == ny - 1)) {
        wall3<T>(fr, wp.walls, y == ny - 1, (T)wp.u_wall);
        wall3<T>(fb, wp.walls, y == ny - 1, (T)wp.u_wall);
    }
    const T rr0 = (T)mac[ind], br0 = (T)mac[n + ind], r0 = (T)mac[2 * n + ind];
    T R = 0, B = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int k = 0; k < 19; ++k) {
        P[(size_t)k * n + ind] = fr[k];
        P[(size_t)(19 + k) * n + ind] = fb[k];
        R += fr[k];
        B += fb[k];
        const T f = fr[k] + fb[k];
        if (cg3::CX[k]) jx += cg3::CX[k] > 0 ? f : -f;
        if (cg3::CY[k]) jy += cg3::CY[k] > 0 ? f : -f;
        if (cg3::CZ[k]) jz += cg3::CZ[k] > 0 ? f : -f;
    }
    // omega_eff from the densities stored before this update (S/GpuUpdateMacro.cu:233-235)
    const T mean_nu = (T)1 / (rr0 / (r0 * (T)wp.cp.r_nu) + br0 / (r0 * (T)wp.cp.b_nu));
    const T om = (T)1 / ((T)3 * mean_nu + (T)0.5);
    const T r = R + B;
    const T g = (T)wp.g;
    mac[ind] = (double)R;
    mac[n + ind] = (double)B;
    mac[2 * n + ind] = (double)r;
    mac[3 * n + ind] = (double)(jx / r + (wp.external_force ? g / (r * om) : (T)0));
    mac[4 * n + ind] = (double)(jy / r + (wp.external_force ? (T)0 : g / om));
    mac[5 * n + ind] = (double)(jz / r);
}

// Rest-state colours from colour densities (initCGBubble3D,
// S/GpuInit.cu:536-700, with the rest fractions), or given colour
// populations; macros from them.
template <class T>
__global__ __launch_bounds__(256) void k_cgw3_init(T* __restrict__ P, double* __restrict__ mac,
                                                   const double* __restrict__ rr, const double* __restrict__ rb,
                                                   const double* __restrict__ rf, const double* __restrict__ bf,
                                                   WallParams wp) {
    const size_t n = (size_t)wp.cp.nx * wp.cp.ny * wp.cp.nz;
    const size_t ind = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ind >= n) return;
    double R = 0, B = 0, jx = 0, jy = 0, jz = 0;
    for (int k = 0; k < 19; ++k) {
        const double a = rf ? rf[(size_t)k * n + ind] : rr[ind] * (cg3::PHI(k) + cg3::TETA(k) * wp.cp.r_alpha);
        const double b = bf ? bf[(size_t)k * n + ind] : rb[ind] * (cg3::PHI(k) + cg3::TETA(k) * wp.cp.b_alpha);
        P[(size_t)k * n + ind] = (T)a;
        P[(size_t)(19 + k) * n + ind] = (T)b;
        R += a;
        B += b;
        jx += cg3::CX[k] * (a + b);
        jy += cg3::CY[k] * (a + b);
        jz += cg3::CZ[k] * (a + b);
    }
    mac[ind] = R;
    mac[n + ind] = B;
    mac[2 * n + ind] = R + B;
    mac[3 * n + ind] = rf ? jx / (R + B) : 0.0;
    mac[4 * n + ind] = rf ? jy / (R + B) : 0.0;
    mac[5 * n + ind] = rf ? jz / (R + B) : 0.0;
}

}  // namespace cgw

hipError_t cgw_launch_step(int storage, const void* P, void* Pc, double* mac, const cg::WallParams& wp, hipStream_t st) {
    const size_t n = (size_t)wp.cp.nx * wp.cp.ny * (wp.cp.nz > 1 ? wp.cp.nz : 1);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (wp.cp.nz > 1) {
#define CGW3_(T_)                                                                                              \
    do {                                                                                                       \
        if (wp.cp.enhanced)                                                                                    \
            hipLaunchKernelGGL((cgw::k_cgw3_collide<T_, true>), dim3(blocks), dim3(256), 0, st, (const T_*)P,  \
                               (T_*)Pc, mac, wp);                                                              \
        else                                                                                                   \
            hipLaunchKernelGGL((cgw::k_cgw3_collide<T_, false>), dim3(blocks), dim3(256), 0, st, (const T_*)P, \
                               (T_*)Pc, mac, wp);                                                              \
        hipLaunchKernelGGL(cgw::k_cgw3_stream<T_>, dim3(blocks), dim3(256), 0, st, (const T_*)Pc, (T_*)P, mac, \
                           wp);                                                                                \
    } while (0)
        if (storage == 0) CGW3_(double);
        else CGW3_(float);
#undef CGW3_
        return hipGetLastError();
    }
    if (storage == 0) {
        if (wp.cp.enhanced)
            hipLaunchKernelGGL((cgw::k_cgw_collide<double, true>), dim3(blocks), dim3(256), 0, st, (const double*)P,
                               (double*)Pc, mac, wp);
        else
            hipLaunchKernelGGL((cgw::k_cgw_collide<double, false>), dim3(blocks), dim3(256), 0, st, (const double*)P,
                               (double*)Pc, mac, wp);
        hipLaunchKernelGGL(cgw::k_cgw_stream<double>, dim3(blocks), dim3(256), 0, st, (const double*)Pc, (double*)P,
                           mac, wp);
    } else {
        if (wp.cp.enhanced)
            hipLaunchKernelGGL((cgw::k_cgw_collide<float, true>), dim3(blocks), dim3(256), 0, st, (const float*)P,
                               (float*)Pc, mac, wp);
        else
            hipLaunchKernelGGL((cgw::k_cgw_collide<float, false>), dim3(blocks), dim3(256), 0, st, (const float*)P,
                               (float*)Pc, mac, wp);
        hipLaunchKernelGGL(cgw::k_cgw_stream<float>, dim3(blocks), dim3(256), 0, st, (const float*)Pc, (float*)P, mac,
                           wp);
    }
    return hipGetLastError();
}

hipError_t cgw_launch_init(int storage, void* P, double* mac, const double* rr, const double* rb, const double* rf,
                           const double* bf, const cg::WallParams& wp, hipStream_t st) {
    const size_t n = (size_t)wp.cp.nx * wp.cp.ny * (wp.cp.nz > 1 ? wp.cp.nz : 1);
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (wp.cp.nz > 1) {
        if (storage == 0)
            hipLaunchKernelGGL(cgw::k_cgw3_init<double>, dim3(blocks), dim3(256), 0, st, (double*)P, mac, rr, rb, rf,
                               bf, wp);
        else
            hipLaunchKernelGGL(cgw::k_cgw3_init<float>, dim3(blocks), dim3(256), 0, st, (float*)P, mac, rr, rb, rf,
                               bf, wp);
        return hipGetLastError();
    }
    if (storage == 0)
        hipLaunchKernelGGL(cgw::k_cgw_init<double>, dim3(blocks), dim3(256), 0, st, (double*)P, mac, rr, rb, rf, bf, wp);
    else
        hipLaunchKernelGGL(cgw::k_cgw_init<float>, dim3(blocks), dim3(256), 0, st, (float*)P, mac, rr, rb, rf, bf, wp);
    return hipGetLastError();
}

hipError_t cgw_launch_out(int storage, int q, const void* P, size_t n, double* rf, double* bf, hipStream_t st) {
    const size_t qn = (size_t)q * n;
    const unsigned blocks = (unsigned)((qn + 255) / 256);
    if (storage == 0)
        hipLaunchKernelGGL(cgw::k_cgw_out<double>, dim3(blocks), dim3(256), 0, st, (const double*)P, qn, rf, bf);
    else
        hipLaunchKernelGGL(cgw::k_cgw_out<float>, dim3(blocks), dim3(256), 0, st, (const float*)P, qn, rf, bf);
    return hipGetLastError();
}

}  // namespace lbm
