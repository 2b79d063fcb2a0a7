// Internal interface between the batched C ABI (lbm_batch_api.cpp) and its kernels.
#pragma once
#include "lbm_internal.h"

namespace lbm {

// One instance of the batch: its place in the concatenated arrays and its
// relaxation and boundary parameters.
struct BInst {
    int nx, ny;
    uint32_t nf;          // non-solid nodes
    uint32_t cbase;       // first compact (population) index, a multiple of 64
    uint32_t dbase;       // first dense index in the concatenated per-node arrays
    uint32_t sx;          // invariant-divisor shift and multiplier for nx
    uint64_t mx;
    double omega, omega_m, rho_out;
};

// Device state of a run, one entry per instance unless noted.
struct BState {
    double* acc;          // 2 per instance: sum |du|^2, sum |u|^2 since the last check
    int* active;          // 1 while the instance runs
    int* done;            // steps taken in this run when retired
    int* par;             // buffer (0/1) holding the instance's newest state
    double* res;          // residual at the last check
    int* remaining;       // one int: instances still running
};

int batch_max_blocks(int storage, int collision);
hipError_t batch_run(int storage, int collision, const BInst* inst, int ninst, const uint2* tiles, int ntiles,
                     void* P0, void* P1, const uint32_t* didx, const uint32_t* nodew, const uint32_t* cidx,
                     Fields fl, double* uprev, double* vprev, BState bs, double tol, int check_every, int max_steps,
                     int grid_blocks, hipStream_t st);
hipError_t batch_init(int storage, int collision, const BInst* inst, const uint2* tiles, int ntiles, void* P,
                      const uint32_t* didx, Fields fl, int only, hipStream_t st);
hipError_t batch_recon(int storage, const BInst* inst, int i, uint32_t nf, const void* P, const uint32_t* didx,
                       const uint32_t* nodew, const uint32_t* cidx, Fields fl, double* fout, hipStream_t st);

}  // namespace lbm
