// Batched D2Q9 channels in one persistent launch.
//
// Many independent lattices (instances) of different sizes advance together in
// one cooperative kernel.  Every step the workgroups stride over a table of
// 256-node tiles of all instances, skip the tiles of retired instances and run
// the single-lattice node update (pull with half-way walls, Zou-He inlet and
// outlet, moments, BGK/TRT) with the instance's own parameters; a grid barrier
// separates steps.  Every check_every steps each tile adds its sum |u - u_prev|^2
// and sum |u|^2 to its instance (FP64 atomics); after the barrier workgroup 0
// turns them into the instance's residual sqrt(sum du^2 / sum u^2), retires the
// instances below the tolerance (and non-finite ones), and publishes how many
// are left.  All workgroups leave the loop at the same barrier, when no
// instance is left or at max_steps.  The reference runs one lattice per
// process, host-synchronised per phase (S/Iterate.cu:385-625, residual
// S/ComputeResiduals.cu:13-63).
#include "lbm_device.h"

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include "lbm_batch_internal.h"

namespace lbm {

namespace {

// Single-lattice parameters of one instance (open x faces, walls in y).
__device__ inline KParams inst_params(const BInst& I) {
    KParams kp{};
    kp.nx = I.nx; kp.ny = I.ny; kp.nz = 1;
    kp.per[2] = 1;
    kp.omega = I.omega; kp.omega_m = I.omega_m; kp.rho_out = I.rho_out;
    kp.tile = 1;
    return kp;
}

// Post-boundary deviations h of compact node `own` of instance I; idx, cls out.
template <int S>
__device__ inline void batch_gather(const BInst& I, const KParams& kp, const typename Io<D2Q9, S>::raw* isrc,
                                    const uint32_t* __restrict__ didx, const uint32_t* __restrict__ nodew,
                                    const uint32_t* __restrict__ cidx, const Fields& fl, uint32_t own,
                                    uint32_t& idx, uint32_t& cls, typename Io<D2Q9, S>::T* h) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    idx = didx[I.cbase + own];
    const uint32_t nd = nodew[I.cbase + own];
    const uint32_t yq = fdiv(idx, I.mx, I.sx);
    const int x = (int)(idx - yq * (uint32_t)I.nx), y = (int)yq;
    uint32_t cn[L::Q];
    source_index<L, true>(cidx + I.dbase, kp, x, y, 0, cn);
    cls = nd >> kClsShift;
    gather<L, S, true>(isrc, kp, own, nd, cn, h);
    Fields f = fl;
    f.ubc = fl.ubc + I.dbase;
    f.vbc = fl.vbc + I.dbase;
    if (cls == VEL_W) zou_inlet<L, T>(kp, f, idx, nd, h);
    if (cls == PRES_E) zou_outlet<L, T>(kp, h);
}

}  // namespace

template <int S, int C>
__global__ __launch_bounds__(256) void k_batch(const BInst* __restrict__ inst, int ninst,
                                               const uint2* __restrict__ tiles, int ntiles,
                                               typename Io<D2Q9, S>::raw* P0, typename Io<D2Q9, S>::raw* P1,
                                               const uint32_t* __restrict__ didx, const uint32_t* __restrict__ nodew,
                                               const uint32_t* __restrict__ cidx, Fields fl, double* uprev,
                                               double* vprev, BState bs, double tol, int check_every,
                                               int max_steps) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    namespace cgr = cooperative_groups;
    cgr::grid_group grid = cgr::this_grid();
    __shared__ double red[2][4];
    __shared__ int left;
    for (int step = 0; step < max_steps; ++step) {
        // written by workgroup 0 before the last barrier: the same in every workgroup
        if (*(volatile int*)bs.remaining == 0) break;
        const bool check = (step + 1) % check_every == 0 || step + 1 == max_steps;
        for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
            const uint2 tl = tiles[t];
            const int i = (int)tl.x;
            if (!*(volatile int*)(bs.active + i)) continue;   // uniform over the workgroup
            const BInst I = inst[i];
            const KParams kp = inst_params(I);
            const int sb = (bs.par[i] + step) & 1;             // buffer holding the instance's newest state
            const size_t off = (size_t)I.cbase * L::Q;
            const typename Io<L, S>::raw* isrc = (sb ? P1 : P0) + off;
            typename Io<L, S>::raw* idst = (sb ? P0 : P1) + off;
            const uint32_t own = tl.y + threadIdx.x;
            double du2 = 0, uu2 = 0;
            if (own < I.nf) {
                uint32_t idx, cls;
                T h[L::Q];
                batch_gather<S>(I, kp, isrc, didx, nodew, cidx, fl, own, idx, cls, h);
                T drho, rho, u[3];
                moments<L, T>(kp, cls, h, drho, rho, u);
                if (check) {
                    const size_t d = (size_t)I.dbase + idx;
                    const double ud = (double)u[0], vd = (double)u[1];
                    const double dx = ud - uprev[d], dy = vd - vprev[d];
                    du2 = dx * dx + dy * dy;
                    uu2 = ud * ud + vd * vd;
                    if (!(uu2 <= 1e300)) du2 = uu2 = (double)INFINITY;   // NaN or inf
                    uprev[d] = ud; vprev[d] = vd;
                    fl.rho[d] = (double)rho; fl.u[d] = ud; fl.v[d] = vd;
                }
                T out[L::Q];
                collide<L, T, C>(kp, h, drho, rho, u, out);
#pragma unroll
                for (int q = 0; q < L::Q; ++q) Io<L, S>::st(idst, paddr<L>(kp, q, own), out[q], q);
            }
            if (check) {
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) {
                    du2 += __shfl_xor(du2, o, 64);
                    uu2 += __shfl_xor(uu2, o, 64);
                }
                const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
                if (lane == 0) { red[0][wv] = du2; red[1][wv] = uu2; }
                __syncthreads();
                if (threadIdx.x == 0) {
                    atomicAdd(bs.acc + 2 * i, (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]));
                    atomicAdd(bs.acc + 2 * i + 1, (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
                }
                __syncthreads();
            }
        }
        grid.sync();
        if (check) {
            if (blockIdx.x == 0) {
                if (threadIdx.x == 0) left = 0;
                __syncthreads();
                for (int i = threadIdx.x; i < ninst; i += blockDim.x) {
                    if (!bs.active[i]) continue;
                    const double a = bs.acc[2 * i], b = bs.acc[2 * i + 1];
                    const double r = b > 0 ? sqrt(a / b) : sqrt(a);
                    bs.res[i] = r;
                    bs.acc[2 * i] = 0;
                    bs.acc[2 * i + 1] = 0;
                    const bool bad = !(r <= 1e300);
                    if (r < tol || bad || step + 1 == max_steps) {
                        bs.active[i] = 0;
                        bs.done[i] = step + 1;
                        bs.par[i] = ((bs.par[i] + step) & 1) ^ 1;   // the buffer just written
                    } else {
                        atomicAdd(&left, 1);
                    }
                }
                __syncthreads();
                if (threadIdx.x == 0) *bs.remaining = left;
            }
            grid.sync();
        }
    }
}

// First post-collision state: equilibrium at the instance's rho, u (dense
// fields), then one collision, as the single-lattice initialisation does.
template <int S, int C>
__global__ __launch_bounds__(256) void k_batch_init(const BInst* __restrict__ inst, const uint2* __restrict__ tiles,
                                                    typename Io<D2Q9, S>::raw* P, const uint32_t* __restrict__ didx,
                                                    Fields fl, int only) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    const uint2 tl = tiles[blockIdx.x];
    if (only >= 0 && (int)tl.x != only) return;
    const BInst I = inst[tl.x];
    const uint32_t own = tl.y + threadIdx.x;
    if (own >= I.nf) return;
    const KParams kp = inst_params(I);
    const size_t d = (size_t)I.dbase + didx[I.cbase + own];
    const double r = fl.rho[d];
    const T drho = (T)(r - 1.0), rho = (T)r;
    const T u[3] = {(T)fl.u[d], (T)fl.v[d], (T)0};
    const T uu = u[0] * u[0] + u[1] * u[1];
    T h[L::Q];
#pragma unroll
    for (int q = 0; q < L::Q; ++q) {
        const T cu = (T)L::cx(q) * u[0] + (T)L::cy(q) * u[1];
        h[q] = (T)L::w(q) * (drho + rho * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu));
    }
    T out[L::Q];
    collide<L, T, C>(kp, h, drho, rho, u, out);
    typename Io<L, S>::raw* base = P + (size_t)I.cbase * L::Q;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) Io<L, S>::st(base, paddr<L>(kp, q, own), out[q], q);
}

// Full post-boundary populations of instance i (dense, FP64) from buffer P.
template <int S>
__global__ __launch_bounds__(256) void k_batch_recon(const BInst* __restrict__ inst, int i,
                                                     const typename Io<D2Q9, S>::raw* __restrict__ P,
                                                     const uint32_t* __restrict__ didx,
                                                     const uint32_t* __restrict__ nodew,
                                                     const uint32_t* __restrict__ cidx, Fields fl, double* fout) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    const BInst I = inst[i];
    const uint32_t own = blockIdx.x * blockDim.x + threadIdx.x;
    if (own >= I.nf) return;
    const KParams kp = inst_params(I);
    uint32_t idx, cls;
    T h[L::Q];
    batch_gather<S>(I, kp, P + (size_t)I.cbase * L::Q, didx, nodew, cidx, fl, own, idx, cls, h);
    const size_t n = (size_t)I.nx * I.ny;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) fout[q * n + idx] = (double)h[q] + L::w(q);
}

// ---------------------------------------------------------------------------
// launchers

template <int S, int C> static const void* batch_kernel() { return (const void*)k_batch<S, C>; }

static const void* batch_kernel_for(int storage, int collision) {
    if (storage == 0) return collision == 0 ? batch_kernel<0, 0>() : batch_kernel<0, 1>();
    return collision == 0 ? batch_kernel<1, 0>() : batch_kernel<1, 1>();
}

int batch_max_blocks(int storage, int collision) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_kernel_for(storage, collision), 256, 0) !=
        hipSuccess)
        return 0;
    return ncu * per;
}

hipError_t batch_run(int storage, int collision, const BInst* inst, int ninst, const uint2* tiles, int ntiles,
                     void* P0, void* P1, const uint32_t* didx, const uint32_t* nodew, const uint32_t* cidx,
                     Fields fl, double* uprev, double* vprev, BState bs, double tol, int check_every, int max_steps,
                     int grid_blocks, hipStream_t st) {
    void* args[] = {(void*)&inst, (void*)&ninst, (void*)&tiles, (void*)&ntiles, (void*)&P0, (void*)&P1,
                    (void*)&didx, (void*)&nodew, (void*)&cidx, (void*)&fl, (void*)&uprev, (void*)&vprev,
                    (void*)&bs, (void*)&tol, (void*)&check_every, (void*)&max_steps};
    return hipLaunchCooperativeKernel(batch_kernel_for(storage, collision), dim3(grid_blocks), dim3(256), args, 0,
                                      st);
}

hipError_t batch_init(int storage, int collision, const BInst* inst, const uint2* tiles, int ntiles, void* P,
                      const uint32_t* didx, Fields fl, int only, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    if (storage == 0) {
        if (collision == 0)
            hipLaunchKernelGGL((k_batch_init<0, 0>), dim3(ntiles), dim3(256), 0, st, inst, tiles, (double*)P, didx,
                               fl, only);
        else
            hipLaunchKernelGGL((k_batch_init<0, 1>), dim3(ntiles), dim3(256), 0, st, inst, tiles, (double*)P, didx,
                               fl, only);
    } else {
        if (collision == 0)
            hipLaunchKernelGGL((k_batch_init<1, 0>), dim3(ntiles), dim3(256), 0, st, inst, tiles, (float*)P, didx,
                               fl, only);
        else
            hipLaunchKernelGGL((k_batch_init<1, 1>), dim3(ntiles), dim3(256), 0, st, inst, tiles, (float*)P, didx,
                               fl, only);
    }
    return hipGetLastError();
}

hipError_t batch_recon(int storage, const BInst* inst, int i, uint32_t nf, const void* P, const uint32_t* didx,
                       const uint32_t* nodew, const uint32_t* cidx, Fields fl, double* fout, hipStream_t st) {
    const unsigned blocks = (nf + 255) / 256;
    if (blocks == 0) return hipSuccess;
    if (storage == 0)
        hipLaunchKernelGGL(k_batch_recon<0>, dim3(blocks), dim3(256), 0, st, inst, i, (const double*)P, didx, nodew,
                           cidx, fl, fout);
    else
        hipLaunchKernelGGL(k_batch_recon<1>, dim3(blocks), dim3(256), 0, st, inst, i, (const float*)P, didx, nodew,
                           cidx, fl, fout);
    return hipGetLastError();
}

}  // namespace lbm
