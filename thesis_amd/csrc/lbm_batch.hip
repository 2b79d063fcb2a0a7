// Batched D2Q9 channels in one persistent launch.
//
// Many independent lattices (instances) of different sizes advance together in
// one cooperative kernel.  Every step the workgroups stride over the 1024-node
// tiles of all instances, skip the tiles of retired instances and run the
// single-lattice node update (pull with half-way walls, Zou-He inlet and
// outlet, moments, BGK/TRT) with the instance's own parameters; a grid barrier
// separates steps.  Nodes are stored densely per instance, so the source of
// every pulled population is the node index minus (cx + cy nx): no index maps
// are loaded, and the instance table, activity flags and buffer parities are
// staged in LDS, so a tile's only global loads are its node words and
// populations.  The grid barrier costs about 0.1 us per workgroup on MI355X
// (each one writes back and invalidates its caches), so the grid is one
// 1024-thread workgroup per CU rather than many small ones.  Every check_every steps each tile adds its sum |u - u_prev|^2
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
    set_float_params(kp);
    return kp;
}

// Population slot of plane q of dense node j: tiles of 64 nodes x 9 planes.
__device__ inline uint32_t bslot(int q, uint32_t j) { return (j >> 6) * (64u * 9u) + (uint32_t)q * 64u + (j & 63u); }

// Post-boundary deviations h of node `own` (local dense index, not solid) of
// instance I, pulled from the instance's population block src.  Walls (bit
// opp(i) of the node word) return the node's own opposite population
// (half-way bounce-back); sources beyond the open x faces keep the node's own
// P_i, replaced by the Zou-He conditions on inlet and outlet nodes.
template <int S>
__device__ inline void dense_gather(const BInst& I, const KParams& kp, const typename Io<D2Q9, S>::raw* src,
                                    const Fields& fl, uint32_t own, uint32_t nd, typename Io<D2Q9, S>::T* h) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    const uint32_t yq = fdiv(own, I.mx, I.sx);
    const int x = (int)(own - yq * (uint32_t)I.nx);
    h[0] = Io<L, S>::ld(src, bslot(0, own), 0);
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int o = L::opp(i), cx = L::cx(i), cy = L::cy(i);
        uint32_t j = own - (uint32_t)(cx + cy * I.nx);
        int pl = i;
        const bool wall = (nd >> o) & 1u;
        const bool open = (cx > 0 && x == 0) || (cx < 0 && x == I.nx - 1);
        if (wall || open) {
            j = own;
            pl = wall ? o : i;
        }
        h[i] = Io<L, S>::ld(src, bslot(pl, j), i);
    }
    const uint32_t cls = nd >> kClsShift;
    Fields f = fl;
    f.ubc = fl.ubc + I.dbase;
    f.vbc = fl.vbc + I.dbase;
    if (cls == VEL_W) zou_inlet<L, T>(kp, f, own, nd, h);
    if (cls == PRES_E) zou_outlet<L, T>(kp, f, own, h);   // kp.outlet = 0: Zou-He pressure
}

__device__ inline int find_inst(const BInst* I, int ninst, uint32_t t) {
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (I[mid].tbase <= t) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace

template <int S, int C>
__global__ __launch_bounds__(kBatchTile) void k_batch(const BInst* __restrict__ inst_g, int ninst, int ntiles,
                                               typename Io<D2Q9, S>::raw* P0, typename Io<D2Q9, S>::raw* P1,
                                               const uint32_t* __restrict__ nodew, Fields fl, double* uprev,
                                               double* vprev, BState bs, double tol, int check_every,
                                               int max_steps) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    using raw = typename Io<L, S>::raw;
    namespace cgr = cooperative_groups;
    cgr::grid_group grid = cgr::this_grid();
    extern __shared__ __align__(16) unsigned char lds_raw[];
    BInst* inst = (BInst*)lds_raw;
    int* act = (int*)(inst + ninst);
    int* par = act + ninst;
    constexpr int W = kBatchTile / 64;
    __shared__ double red[2][W];
    __shared__ int left;
    for (int k = threadIdx.x; k < ninst; k += blockDim.x) inst[k] = inst_g[k];
    for (int step = 0; step < max_steps; ++step) {
        // activity and parity were last written by workgroup 0 before the
        // previous barrier: every workgroup sees the same values
        for (int k = threadIdx.x; k < ninst; k += blockDim.x) {
            act[k] = *(volatile int*)(bs.active + k);
            par[k] = *(volatile int*)(bs.par + k);
        }
        const int remaining = *(volatile int*)bs.remaining;
        __syncthreads();
        if (remaining == 0) break;
        const bool check = (step + 1) % check_every == 0 || step + 1 == max_steps;
        int ci = 0;
        // the workgroups an XCD runs together take consecutive tiles (xcd_block,
        // 8 XCDs, chunks of 8), so tiles that share boundary rows meet in one L2
        for (int t = (int)xcd_block(blockIdx.x, gridDim.x, 3, 3); t < ntiles; t += gridDim.x) {
            while (ci + 1 < ninst && inst[ci + 1].tbase <= (uint32_t)t) ++ci;   // tiles ascend with t
            const int i = ci;
            if (!act[i]) continue;                          // uniform over the workgroup
            const BInst I = inst[i];
            const KParams kp = inst_params(I);
            const int sb = (par[i] + step) & 1;             // buffer holding the instance's newest state
            const size_t off = (size_t)I.pbase * L::Q;
            const raw* isrc = (sb ? P1 : P0) + off;
            raw* idst = (sb ? P0 : P1) + off;
            const uint32_t own = ((uint32_t)t - I.tbase) * (uint32_t)kBatchTile + threadIdx.x;
            double du2 = 0, uu2 = 0;
            const uint32_t nd = own < I.n ? nodew[I.dbase + own] : 0u;
            if ((nd >> kClsShift) != SOLID) {
                T h[L::Q];
                dense_gather<S>(I, kp, isrc, fl, own, nd, h);
                T drho, rho, u[3];
                moments<L, T>(kp, nd >> kClsShift, h, drho, rho, u);
                if (check) {
                    const size_t d = (size_t)I.dbase + own;
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
                for (int q = 0; q < L::Q; ++q) Io<L, S>::st(idst, bslot(q, own), out[q], q);
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
                    double a = 0, b = 0;
                    for (int k = 0; k < W; ++k) { a += red[0][k]; b += red[1][k]; }
                    atomicAdd(bs.acc + 2 * i, a);
                    atomicAdd(bs.acc + 2 * i + 1, b);
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
                    if (!act[i]) continue;
                    const double a = bs.acc[2 * i], b = bs.acc[2 * i + 1];
                    const double r = b > 0 ? sqrt(a / b) : sqrt(a);
                    bs.res[i] = r;
                    bs.acc[2 * i] = 0;
                    bs.acc[2 * i + 1] = 0;
                    const bool bad = !(r <= 1e300);
                    if (r < tol || bad || step + 1 == max_steps) {
                        bs.active[i] = 0;
                        bs.done[i] = step + 1;
                        bs.par[i] = ((par[i] + step) & 1) ^ 1;   // the buffer just written
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
__global__ __launch_bounds__(kBatchTile) void k_batch_init(const BInst* __restrict__ inst, int ninst,
                                                    typename Io<D2Q9, S>::raw* P, const uint32_t* __restrict__ nodew,
                                                    Fields fl, int only) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    const int i = find_inst(inst, ninst, blockIdx.x);
    if (only >= 0 && i != only) return;
    const BInst I = inst[i];
    const uint32_t own = (blockIdx.x - I.tbase) * (uint32_t)kBatchTile + threadIdx.x;
    if (own >= I.n) return;
    if ((nodew[I.dbase + own] >> kClsShift) == SOLID) return;
    const KParams kp = inst_params(I);
    const size_t d = (size_t)I.dbase + own;
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
    typename Io<L, S>::raw* base = P + (size_t)I.pbase * L::Q;
#pragma unroll
    for (int q = 0; q < L::Q; ++q) Io<L, S>::st(base, bslot(q, own), out[q], q);
}

// Full post-boundary populations of instance i (dense, FP64) from buffer P.
template <int S>
__global__ __launch_bounds__(256) void k_batch_recon(const BInst* __restrict__ inst, int i,
                                                     const typename Io<D2Q9, S>::raw* __restrict__ P,
                                                     const uint32_t* __restrict__ nodew, Fields fl, double* fout) {
    using L = D2Q9;
    using T = typename Io<L, S>::T;
    const BInst I = inst[i];
    const uint32_t own = blockIdx.x * blockDim.x + threadIdx.x;
    if (own >= I.n) return;
    const uint32_t nd = nodew[I.dbase + own];
    if ((nd >> kClsShift) == SOLID) return;
    const KParams kp = inst_params(I);
    T h[L::Q];
    dense_gather<S>(I, kp, P + (size_t)I.pbase * L::Q, fl, own, nd, h);
#pragma unroll
    for (int q = 0; q < L::Q; ++q) fout[(size_t)q * I.n + own] = (double)h[q] + L::w(q);
}

// ---------------------------------------------------------------------------
// launchers

size_t batch_lds_bytes(int ninst) { return (size_t)ninst * (sizeof(BInst) + 2 * sizeof(int)); }

template <int S, int C> static const void* batch_kernel() { return (const void*)k_batch<S, C>; }

static const void* batch_kernel_for(int storage, int collision) {
    if (storage == 0) return collision == 0 ? batch_kernel<0, 0>() : batch_kernel<0, 1>();
    return collision == 0 ? batch_kernel<1, 0>() : batch_kernel<1, 1>();
}

int batch_max_blocks(int storage, int collision, int ninst) {
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, batch_kernel_for(storage, collision), kBatchTile,
                                                     batch_lds_bytes(ninst)) != hipSuccess)
        return 0;
    return ncu * per;
}

hipError_t batch_run(int storage, int collision, const BInst* inst, int ninst, int ntiles, void* P0, void* P1,
                     const uint32_t* nodew, Fields fl, double* uprev, double* vprev, BState bs, double tol,
                     int check_every, int max_steps, int grid_blocks, hipStream_t st) {
    void* args[] = {(void*)&inst, (void*)&ninst, (void*)&ntiles, (void*)&P0, (void*)&P1, (void*)&nodew,
                    (void*)&fl, (void*)&uprev, (void*)&vprev, (void*)&bs, (void*)&tol, (void*)&check_every,
                    (void*)&max_steps};
    return hipLaunchCooperativeKernel(batch_kernel_for(storage, collision), dim3(grid_blocks), dim3(kBatchTile), args,
                                      (unsigned)batch_lds_bytes(ninst), st);
}

hipError_t batch_init(int storage, int collision, const BInst* inst, int ninst, int ntiles, void* P,
                      const uint32_t* nodew, Fields fl, int only, hipStream_t st) {
    if (ntiles == 0) return hipSuccess;
    if (storage == 0) {
        if (collision == 0)
            hipLaunchKernelGGL((k_batch_init<0, 0>), dim3(ntiles), dim3(kBatchTile), 0, st, inst, ninst, (double*)P, nodew,
                               fl, only);
        else
            hipLaunchKernelGGL((k_batch_init<0, 1>), dim3(ntiles), dim3(kBatchTile), 0, st, inst, ninst, (double*)P, nodew,
                               fl, only);
    } else {
        if (collision == 0)
            hipLaunchKernelGGL((k_batch_init<1, 0>), dim3(ntiles), dim3(kBatchTile), 0, st, inst, ninst, (float*)P, nodew,
                               fl, only);
        else
            hipLaunchKernelGGL((k_batch_init<1, 1>), dim3(ntiles), dim3(kBatchTile), 0, st, inst, ninst, (float*)P, nodew,
                               fl, only);
    }
    return hipGetLastError();
}

hipError_t batch_recon(int storage, const BInst* inst, int i, uint32_t n, const void* P, const uint32_t* nodew,
                       Fields fl, double* fout, hipStream_t st) {
    const unsigned blocks = (n + 255) / 256;
    if (blocks == 0) return hipSuccess;
    if (storage == 0)
        hipLaunchKernelGGL(k_batch_recon<0>, dim3(blocks), dim3(256), 0, st, inst, i, (const double*)P, nodew, fl,
                           fout);
    else
        hipLaunchKernelGGL(k_batch_recon<1>, dim3(blocks), dim3(256), 0, st, inst, i, (const float*)P, nodew, fl,
                           fout);
    return hipGetLastError();
}

}  // namespace lbm
