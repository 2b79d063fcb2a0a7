// Internal interface between the batched C ABI (lbm_batch_api.cpp) and its kernels.
#pragma once
#include "lbm_internal.h"

namespace lbm {

// One instance of the batch: its place in the concatenated arrays and its
// relaxation and boundary parameters.  Nodes are stored densely (solid ones
// included, never updated) in the instance's own row-major order.
struct BInst {
    int nx, ny;
    uint32_t n;           // nodes, nx * ny
    uint32_t pbase;       // first population slot of the instance (a multiple of 64)
    uint32_t dbase;       // first index in the concatenated per-node arrays
    uint32_t tbase;       // first kBatchTile-node tile of the instance in the tile numbering
    uint32_t sx, pad;     // invariant-divisor shift and multiplier for nx
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

constexpr int kBatchTile = 1024;          // nodes per tile = threads per workgroup

constexpr int kBatchMaxInstances = 512;   // instance table, activity and parity live in LDS

size_t batch_lds_bytes(int ninst);
int batch_max_blocks(int storage, int collision, int ninst);
hipError_t batch_run(int storage, int collision, const BInst* inst, int ninst, int ntiles, void* P0, void* P1,
                     const uint32_t* nodew, Fields fl, double* uprev, double* vprev, BState bs, double tol,
                     int check_every, int max_steps, int grid_blocks, hipStream_t st);
hipError_t batch_init(int storage, int collision, const BInst* inst, int ninst, int ntiles, void* P,
                      const uint32_t* nodew, Fields fl, int only, hipStream_t st);
hipError_t batch_recon(int storage, const BInst* inst, int i, uint32_t n, const void* P, const uint32_t* nodew,
                       Fields fl, double* fout, hipStream_t st);

}  // namespace lbm
