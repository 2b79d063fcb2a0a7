// Internal interface between the C ABI (lbm_api.cpp) and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <cstdlib>

namespace lbm {

enum Cls : uint32_t {
    SOLID = 0, FLUID = 1, VEL_W = 2, PRES_E = 3, VEL_N = 4,
    VEL_E = 5, VEL_S = 6, VEL_B = 7, VEL_T = 8,
    OUT_W = 9, OUT_S = 10, OUT_N = 11, OUT_B = 12, OUT_T = 13,
    SYM_W = 14, SYM_E = 15, SYM_S = 16, SYM_N = 17, SYM_B = 18, SYM_T = 19
};
enum Mode : int { INIT_EQ = 0, INIT_ZERO = 1, INIT_FROM_F = 2 };

// Per-node word: bits 0..18 bounce-back links, bits 19..23 the wet-node wall
// closure id (3D), bits 24..31 node class.
constexpr uint32_t kBBMask = 0x7FFFFu;
constexpr int kHHShift = 19;
constexpr int kClsShift = 24;

// The 16-bit rank step on periodic lattices addresses population slots from a
// base kRankShift compact nodes (whole 64-node tiles) before the buffer, so
// that every per-lane offset it selects is non-negative; the buffers keep that
// much room in front (lbm_create).
constexpr uint32_t kRankShift = 256;

// Everything a kernel needs, passed by value as a kernel argument.
struct KParams {
    int nx, ny, nz;
    int per[3];
    size_t qstride;          // elements between population planes
    uint32_t qstride32;
    double omega, omega_m;   // BGK/TRT rates
    double s[19];            // MRT diag(S)
    double g[3];
    int force;               // 0 none, 1 shift, 2 guo
    int wall;                // 0 halfway, 1 on-node, 2 wet-node (3D), 3 curved (Bouzidi)
    int inlet;               // 0 zou-he, 1 reference 2D (west inlet and north lid)
    int outlet;              // 0 pressure, 1 reference velocity (2D), 2/3 extrapolation, 4 f_eq (3D)
    uint32_t parity;
    int mref;                // D3Q19 MRT: populations stored as f - (w - r), r = 3 M^-1 e_eps
    double rho_out;
    uint32_t nf;             // non-solid nodes (compact population count)
    uint64_t mx, my;         // invariant-divisor multipliers for nx, ny
    uint32_t sx, sy;
    uint32_t xcd_shift;      // log2 of the XCDs the workgroups are dealt over (0: no renumbering)
    uint32_t xchunk_shift;   // log2 of the consecutive workgroups an XCD takes per turn
    int tile;                // population layout: 0 SoA planes, 1 tiles of 64 nodes x Q
    const uint32_t* didx;    // compact -> dense node index
    const uint32_t* nodew;   // compact node words
    int rank;                // use the row-bitmask step (no per-node index loads)
    int sr;                  // stochastic rounding of 16-bit stores
    uint32_t seed;           // per-launch random seed for stochastic rounding
    int simple;              // walls are exactly the links into solid nodes; all fluid plain
    int words;               // 32-node words per row
    int lgw;                 // log2 of the words per row rounded up to a power of two (LDS pitch)
    int rblock;              // nodes per workgroup of the row-bitmask step
    int npt;                 // nodes per thread of its 16-bit periodic variant (0 or 1: one, 2, 4)
    int rmax;                // rows staged per workgroup (LDS sizing)
    const uint2* rowmeta;    // per row and word: (fluid mask, compact index of first fluid node)
    const uint2* blk;        // per rblock-node workgroup: (first row, rows spanned)
    const uint32_t* qslot;   // curved walls: per compact node, its row of qv (0xFFFFFFFF: none)
    const double* qv;        // curved walls: 19 wall fractions per row
    // FP32 copies of omega, omega_m, g and g / 2 (set_float_params): float
    // kernels read these rather than converting the doubles in every wave
    float omegaf, omega_mf, gf[3], hgf[3];
};

__host__ __device__ inline void set_float_params(KParams& k) {
    k.omegaf = (float)k.omega;
    k.omega_mf = (float)k.omega_m;
    for (int a = 0; a < 3; ++a) {
        k.gf[a] = (float)k.g[a];
        k.hgf[a] = (float)(0.5 * k.g[a]);
    }
}

struct Fields {              // FP64 side arrays (N each)
    double *rho, *u, *v, *w; // macroscopic output / input
    const double *ubc, *vbc, *wbc;
    double* fout;            // Q*N full populations (reconstruct) or input (INIT_FROM_F)
    // f_eq outlet: the macros of the previous step (rho, u, v, w blocks of N)
    // read by this step, and the block this step writes (null: none)
    const double* om_in;
    double* om_out;
    // frozen macros of velocity-face nodes (rho, u, v, w blocks of N; null:
    // none): their collision and written macros use these, not the moments
    const double* frz;
};

// Launchers; q in {9,19}; storage 0..3; collision 0..2.
hipError_t launch_step(int q, int storage, int collision, bool open, const KParams& kp,
                       const void* src, void* dst, const uint32_t* node, const uint32_t* cidx, Fields fl,
                       bool write_macros, hipStream_t st);
hipError_t launch_recon(int q, int storage, bool open, const KParams& kp, const void* src,
                        const uint32_t* node, const uint32_t* cidx, Fields fl, hipStream_t st);
hipError_t launch_init(int q, int storage, int collision, int mode, const KParams& kp, void* dst,
                       const uint32_t* node, const uint32_t* cidx, Fields fl, hipStream_t st);
// FP64 reductions over non-solid nodes: out8 (device) = {sum du^2, sum u^2,
// sum rho, nonfinite, max |du|, max |dv|, max |dw|, max |drho|}; partials
// holds 8 doubles per block.
hipError_t launch_reduce(const KParams& kp, const uint32_t* node, Fields fl, double* uprev, double* vprev,
                         double* wprev, double* rprev, double* partials, int nblocks, double* out8,
                         hipStream_t st);
// out (device) = sum over non-solid nodes of (f - fColl)^2, f dense FP64 in fout.
hipError_t launch_fdiff2(int q, int storage, const KParams& kp, const void* pc, const uint32_t* cidx,
                         const double* fout, double* partials, int nblocks, double* out, hipStream_t st);
// out (device) = max over non-solid links of |f - fprev| / |f1| (dense FP64).
hipError_t launch_fdrel(int q, const double* f, const double* f1, const double* fprev, const uint32_t* cidx,
                        size_t n, double* partials, int nblocks, double* out, hipStream_t st);
int reduce_blocks();
// Allocation-free L2 residual: the sum of k_fdiff2 with the post-boundary
// populations rebuilt in registers from pc; partials (nblocks doubles) and out
// may be the device-resident scratch of l2_scratch.
hipError_t launch_l2_fused(int q, int storage, bool open, const KParams& kp, const void* pc, const uint32_t* node,
                           const uint32_t* cidx, Fields fl, double* partials, int nblocks, double* out,
                           hipStream_t st);
// Device address of the calling thread's current device's copy of the static
// scratch (kL2ScratchDoubles doubles); looked up once per device.
constexpr int kL2ScratchDoubles = 1024 + 8;
hipError_t l2_scratch(double** out);

// log2 of a power-of-two setting read from the environment (rounded down, at
// most 256), for the workgroup renumbering below; out is kept when unset.
inline void log2_env(const char* name, uint32_t& out) {
    if (const char* e = std::getenv(name)) {
        const unsigned long v = std::strtoul(e, nullptr, 10);
        uint32_t sh = 0;
        while ((2ul << sh) <= v && sh < 8) ++sh;
        out = v ? sh : 0;
    }
}

// Workgroups are dealt round-robin to the 2^xs XCDs; renumber them so that each
// XCD's turn takes the next 2^cs consecutive workgroups of the lattice, and
// neighbouring rows (and their cache lines of cidx and P) meet in the same L2.
// The tail that fills no whole round keeps its numbers; xs = cs = 0 is the
// identity.  Shifts only: the XCD count and the chunk are powers of two, so
// no integer division runs in the kernel.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t xs, uint32_t cs) {
    const uint32_t rs = xs + cs;
    if (b >= (nb >> rs) << rs) return b;
    const uint32_t k = b >> xs;
    return ((k >> cs) << rs) + ((b & ((1u << xs) - 1u)) << cs) + (k & ((1u << cs) - 1u));
}

}  // namespace lbm
