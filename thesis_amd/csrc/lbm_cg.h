// Shared declarations of the colour-gradient kernels (D2Q9 and D3Q19,
// periodic box) and their C ABI.
#pragma once

#include "lbm_internal.h"

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace lbm {
namespace cg {

struct CgParams {
    int nx, ny;
    double r_alpha, b_alpha, r_nu, b_nu, beta, A, g_limit, kappa;
    uint32_t plane;          // elements per plane (2Q planes: Q deviations h = f - w, Q red)
    int nz;                  // 1 for D2Q9
    int enhanced;            // 1: density-gradient terms in f_eq (enhanced model)
    uint32_t xcd_shift;      // workgroup renumbering of the D3Q19 kernels (xcd_block)
    uint32_t xchunk_shift;
    int zchunk;              // D3Q19, 18-point gradient: planes per workgroup of the
                             // single-pass step (k_cg3_fz); 0: phase-field pass + step
    int zrows;               // rows of its 64-wide tiles (4 or 8)
};

constexpr int NRED = 12;     // reductions of k_cg_reduce

// D2Q9 / D3Q19 box with walls on the y rows (planes) and a body force (lbm_cgw.hip)
struct WallParams {
    CgParams cp;
    int walls;               // 0 none (periodic), 1 Zou-He velocity walls, 2 bounce-back walls
    int ho;                  // 24- / 104-point stencil away from the walls
    double u_wall;           // north wall velocity along x (walls == 1): D2Q9 its initial stored
                             // velocity, D3Q19 the face velocity of the inlet condition
    double g;                // body force, by the velocity shift of the macroscopic update
    int external_force;      // 1: along x, u += g / (rho omega_eff); 0: along y, v += g / omega_eff
};

}  // namespace cg

// D3Q19 colour-gradient constants, shared by the periodic step (lbm_cg3.hip)
// and the reference-order wall path (lbm_cgw.hip)
namespace cg3 {
__device__ constexpr int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
__device__ constexpr int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
__device__ constexpr int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
__host__ __device__ constexpr double W19(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }
// phi_k, theta_k, B_k by shell (S/GpuInit.cu:900-915)
__host__ __device__ constexpr double PHI(int i) { return i == 0 ? 0.0 : (i <= 6 ? 1.0 / 12.0 : 1.0 / 24.0); }
__host__ __device__ constexpr double TETA(int i) { return i == 0 ? 1.0 : (i <= 6 ? -1.0 / 12.0 : -1.0 / 24.0); }
__host__ __device__ constexpr double BPERT(int i) { return i == 0 ? -2.0 / 9.0 : (i <= 6 ? 1.0 / 54.0 : 1.0 / 27.0); }
__host__ __device__ constexpr double CNORM(int i) { return i <= 6 ? 1.0 : 1.4142135623730950488; }
// enhanced-model chi_k, psi_k (S/GpuInit.cu:916-926)
__host__ __device__ constexpr double CHI(int i) { return i == 0 ? -2.5 : (i <= 6 ? -1.0 / 6.0 : 1.0 / 24.0); }
__host__ __device__ constexpr double PSI(int i) { return i == 0 ? 0.0 : (i <= 6 ? 0.25 : 0.125); }
// gradient weights by |c|^2: 18-neighbour stencil 1/6, 1/12 (S/GpuInit.cu:928);
// 104-neighbour stencil (S/GpuInit.cu:950-990): shells 1, 2, 3, 4, 5, 6 and 9
// (the (2,2,1) points), 0 for the points of [-2,2]^3 it leaves out
__host__ __device__ constexpr double LOW(int c2) { return c2 == 1 ? 1.0 / 6.0 : c2 == 2 ? 1.0 / 12.0 : 0.0; }
__host__ __device__ constexpr double HOW(int c2) {
    return c2 == 1 ? 4.0 / 45 : c2 == 2 ? 1.0 / 21 : c2 == 3 ? 2.0 / 105 : c2 == 4 ? 5.0 / 504
         : c2 == 5 ? 1.0 / 315 : c2 == 6 ? 1.0 / 630 : c2 == 9 ? 1.0 / 5040 : 0.0;
}
}  // namespace cg3

namespace cg {

// (kx, ky, kz) . (a, b, c) for components in {-1, 0, 1} (constants once the
// direction loop is unrolled) with the zero components left out: the compiler
// does not fold 0 * x, and the result is the same for finite operands.
template <class T>
__device__ inline T cdot(int kx, int ky, int kz, T a, T b, T c) {
    T r = 0;
    bool any = false;
    const int k[3] = {kx, ky, kz};
    const T v[3] = {a, b, c};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (k[d] == 0) continue;
        const T t = k[d] > 0 ? v[d] : -v[d];
        r = any ? r + t : t;
        any = true;
    }
    return r;
}

}  // namespace cg

// D2Q9 (lbm_cg.hip)
hipError_t cg_launch_step(int storage, int ho, const void* src, void* dst, const cg::CgParams& cp, double* mac,
                          bool wm, hipStream_t st);
hipError_t cg_launch_init(int storage, int ho, void* dst, const cg::CgParams& cp, const double* rr,
                          const double* rb, const double* rf, const double* bf, double* mac, hipStream_t st);
hipError_t cg_launch_recon(int storage, const void* src, const cg::CgParams& cp, double* rf, double* bf,
                           hipStream_t st);
// colour densities rr, rb of FP64 populations rf, bf (q planes of n)
hipError_t cg_launch_dens(int q, const double* rf, const double* bf, size_t n, double* rr, double* rb,
                          hipStream_t st);
// mac: rho_r, rho_b, rho, u, v, w (n each; w ignored when nz == 1)
hipError_t cg_launch_reduce(const double* mac, size_t n, bool three_d, double kappa, double* partials,
                            int nblocks, hipStream_t st);

// D2Q9 / D3Q19 with walls / force (lbm_cgw.hip; D3Q19 when wp.cp.nz > 1): P
// holds the colour populations after the wall condition (2 x Q planes of n),
// Pc after collision
hipError_t cgw_launch_step(int storage, const void* P, void* Pc, double* mac, const cg::WallParams& wp,
                           hipStream_t st);
hipError_t cgw_launch_init(int storage, void* P, double* mac, const double* rr, const double* rb, const double* rf,
                           const double* bf, const cg::WallParams& wp, hipStream_t st);
hipError_t cgw_launch_out(int storage, int q, const void* P, size_t n, double* rf, double* bf, hipStream_t st);

// D3Q19 (lbm_cg3.hip): phase field of the post-streaming state, then the step
hipError_t cg3_launch_step(int storage, int ho, const void* src, void* dst, void* psi, const cg::CgParams& cp,
                           double* mac, bool wm, hipStream_t st);
hipError_t cg3_launch_init(int storage, int ho, void* dst, void* psi, const cg::CgParams& cp, const double* rr,
                           const double* rb, const double* rf, const double* bf, double* mac, hipStream_t st);
hipError_t cg3_launch_recon(int storage, const void* src, const cg::CgParams& cp, double* rf, double* bf,
                            hipStream_t st);

}  // namespace lbm
