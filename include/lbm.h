/*
 * lbm.h -- C ABI of the MI355X lattice-Boltzmann time step.
 *
 * Replaces the kernel-level surface of ZPaulo/Thesis `lbm-solver_final`:
 * the __global__ kernels over raw device pointers with parameters in
 * __constant__ symbols (S/include/GpuFunctions.h:63-472,
 * S/include/GpuConstants.h:43-118) and the driver entry points
 * Iterate2D/Iterate3D (S/include/Iterate.h:57-65).  Prefix S/ stands for
 * `Source code/lbm-solver_final/` of the reference tree.
 *
 * Plain C types only.  A context owns one HIP stream and every device
 * buffer; host pointers passed in are borrowed for the duration of the call.
 * Errors are status codes (the library never exits); lbm_last_error()
 * returns the message.  A context is not re-entrant; distinct contexts may
 * be used from distinct host threads.
 */
#ifndef LBM_H
#define LBM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_ABI_VERSION 3

typedef struct lbm_ctx lbm_ctx;

/* Status codes.  LBM_DIVERGED / LBM_EIO keep the reference's driver return
 * values 1 and 2 (S/include/Iterate.h:62). */
typedef enum {
    LBM_OK = 0,
    LBM_DIVERGED = 1,
    LBM_EIO = 2,
    LBM_EINVAL = 3,
    LBM_EDEVICE = 4,
    LBM_ENOMEM = 5,
    LBM_EUNSUPPORTED = 6
} lbm_status;

/* node classes: one byte per node (latticegen2 fluid flag + boundary kind).
 * Faces of the box: W x = 0, E x = nx-1, S y = 0, N y = ny-1, B z = 0,
 * T z = nz-1.  LBM_VEL_*: velocity face (u_bc, v_bc, w_bc); in 2D only W (the
 * inlet) and N (the moving lid, the reference's 2D north "inlet",
 * S/GpuBoundaries.cu:21-41).  LBM_PRES_E and LBM_OUT_*: outlets
 * (lbm_params.outlet_model); 2D has the east one only.  LBM_SYM_*: symmetry
 * planes (3D, S/GpuBoundaries.cu:2086-2167). */
enum {
    LBM_SOLID = 0, LBM_FLUID = 1, LBM_VEL_W = 2, LBM_PRES_E = 3, LBM_VEL_N = 4,
    LBM_VEL_E = 5, LBM_VEL_S = 6, LBM_VEL_B = 7, LBM_VEL_T = 8,
    LBM_OUT_W = 9, LBM_OUT_S = 10, LBM_OUT_N = 11, LBM_OUT_B = 12, LBM_OUT_T = 13,
    LBM_SYM_W = 14, LBM_SYM_E = 15, LBM_SYM_S = 16, LBM_SYM_N = 17, LBM_SYM_B = 18, LBM_SYM_T = 19
};
/* collision models (S/include/Arguments.h: BGKW, TRT, MRT) */
enum { LBM_BGK = 0, LBM_TRT = 1, LBM_MRT = 2 };
/* population storage; compute is FP64 for F64 storage and FP32 otherwise */
enum { LBM_STORE_F64 = 0, LBM_STORE_F32 = 1, LBM_STORE_F16 = 2, LBM_STORE_BF16 = 3 };
/* body force: NONE, the reference's stored-velocity shift u += g/omega
 * (S/GpuUpdateMacro.cu:207), or Guo forcing */
enum { LBM_FORCE_NONE = 0, LBM_FORCE_SHIFT = 1, LBM_FORCE_GUO = 2 };
/* walls: half-way bounce-back, or the reference's on-node
 * f[opp] = f[dir] (S/GpuBoundaries.cu:435-437, :482-484) */
/* LBM_WALL_HH: the reference's wet-node walls ("HHmodel",
 * S/GpuBoundaries.cu:492-1560), D3Q19: nodes with 5, 9 or 12 wall links are
 * closed as a face, an edge or a corner of a box; other wall links keep the
 * node's own population, as in the reference */
enum { LBM_WALL_HALFWAY = 0, LBM_WALL_ONNODE = 1, LBM_WALL_HH = 2, LBM_WALL_CURVED = 3 };
/* LBM_WALL_CURVED: Bouzidi linear interpolation (Bouzidi, Firdaouss &
 * Lallemand 2001) with the wall fraction q of each wall link
 * (lbm_lattice.wall_q; 1/2 is half-way bounce-back), the textbook form of the
 * reference's CURVED walls (S/GpuBoundaries.cu:414-431, :457-471). */
/* velocity faces: textbook Zou-He, or the reference's 2D forms without
 * transverse terms, corners skipped (west S/GpuBoundaries.cu:45-64; north
 * lid :21-41); the reference forms impose u0 exactly on equilibrium states */
enum { LBM_INLET_ZOUHE = 0, LBM_INLET_REF2D = 1 };
/* east outlet (2D): Zou-He pressure (rho_out; T/LBM/Tom/Lecture Script/LBMchannel.m:254-270),
 * the reference's velocity outlet with its (1 - u0) density denominator and
 * u0 = u_bc at the node (S/GpuBoundaries.cu:1571-1605), or its second- and
 * first-order extrapolations, which overwrite the streamed-in f1, f5, f8 from
 * the two upstream nodes (:1609-1640) */
enum { LBM_OUTLET_PRESSURE = 0, LBM_OUTLET_REF_VELOCITY = 1, LBM_OUTLET_EXTRAP2 = 2, LBM_OUTLET_EXTRAP1 = 3,
       LBM_OUTLET_FEQ = 4 };
/* In 3D (S/GpuBoundaries.cu:1655-1874) the outlets act on any LBM_OUT_* face
 * and LBM_PRES_E: LBM_OUTLET_PRESSURE is a textbook pressure face (rho_out,
 * zero tangential velocity; the reference has none), the extrapolations set
 * the face's unknown populations from the two upstream nodes, and
 * LBM_OUTLET_FEQ sets them to f_eq of the node's macros of the previous step
 * (the reference's HEmodel outlet); LBM_OUTLET_REF_VELOCITY is 2D only. */
/* reference quirks reproduced on request */
#define LBM_PARITY_MRT_RHO2   0x1u /* 2D MRT m_eq with rho^2 (S/GpuCollision.cu:987-988) */
#define LBM_PARITY_MRT_NO_F18 0x2u /* 3D MRT leaves f18 unwritten (S/GpuCollision.cu:1161) */
#define LBM_PARITY_INLET_FROZEN 0x4u /* macros of velocity-face nodes keep the values lbm_init set:
                                       the reference's UpdateInltOutl = no (S/GpuUpdateMacro.cu:150-151) */
#define LBM_PARITY_OUTLET_V0  0x8u /* 2D v := 0 on outlet nodes (S/GpuUpdateMacro.cu:30) */
#define LBM_PARITY_INLET3D_FLIP 0x10u /* 3D W/E velocity faces divide by 1 + G u (S/GpuBoundaries.cu:102,148) */

/* Solver parameters; mirrors the numerical part of the reference's
 * Arguments struct (S/include/Arguments.h:86-135) plus the new knobs. */
typedef struct {
    uint32_t abi_version;   /* LBM_ABI_VERSION */
    int32_t q;              /* 9 (D2Q9) or 19 (D3Q19) */
    int32_t nx, ny, nz;     /* n, m, h of the reference; nz = 1 for D2Q9 */
    int32_t periodic[3];    /* per axis, modular wrap */
    int32_t collision;
    double omega;           /* 1/tau = 1/(3 nu + 1/2) (S/GpuInit.cu:102) */
    double trt_magic;       /* Lambda; 0 selects 3/16, the reference's omegaA (S/GpuInit.cu:103) */
    double mrt_s[19];       /* diag(S); all zero selects S/CellFunctions.cu:44-52 / :122-140 */
    int32_t force_model;
    double force[3];        /* acceleration g */
    int32_t storage;
    int32_t wall_model;
    int32_t inlet_model;
    uint32_t parity_flags;
    double rho_out;         /* Zou-He pressure outlet density */
    int32_t outlet_model;   /* LBM_OUTLET_*; applies to LBM_PRES_E nodes */
} lbm_params;

/* Lattice description; node_class is required, the rest may be NULL.
 * bb_links[i] bit d (1 <= d < q): the link from node i in direction d
 * crosses a wall (the _BC.dat line "i j k d WALL").
 * u_bc/v_bc/w_bc: prescribed velocity on velocity-face nodes, and u0 of the
 * reference velocity outlet on LBM_PRES_E nodes.
 * wall_q[i*19 + d]: fraction of the link from node i in direction d at which
 * it meets the wall (LBM_WALL_CURVED; entries <= 0, or NULL, mean 1/2). */
typedef struct {
    const uint8_t* node_class;
    const uint32_t* bb_links;
    const double* u_bc;
    const double* v_bc;
    const double* w_bc;
    const double* wall_q;
} lbm_lattice;

int lbm_abi_version(void);
/* Number of visible HIP devices (0 when none); never fails. */
int lbm_device_count(void);

/* Allocate a solver; replaces the allocation/initialisation part of
 * Iterate2D/Iterate3D (S/Iterate.cu:23-330, S/Iterate3D.cu:30-490). */
lbm_status lbm_create(const lbm_params* p, const lbm_lattice* lat, lbm_ctx** out);
void lbm_destroy(lbm_ctx* ctx);
const char* lbm_last_error(const lbm_ctx* ctx);

/* Initial state.  mode 0: f = f_eq(rho, u); mode 1: f = 0 with the given
 * macroscopic fields, the reference's start (S/Iterate.cu:244-245,
 * S/Iterate3D.cu:301-302).  NULL fields mean rho = 1, u = 0. */
lbm_status lbm_init(lbm_ctx* ctx, int mode, const double* rho, const double* u,
                    const double* v, const double* w);

/* Advance nsteps fused steps (collide, stream, boundaries, moments: the
 * loop body of S/Iterate.cu:385-520 / S/Iterate3D.cu:493-653).  Enqueues
 * everything on the context stream and waits once at the end. */
lbm_status lbm_step(lbm_ctx* ctx, int64_t nsteps);

/* Steps in blocks of check_every until the relative L2 change of the
 * velocity over one block falls below tol (FP64 reduction), or max_steps.
 * Returns LBM_DIVERGED on a non-finite field (S/Iterate.cu:552-564). */
lbm_status lbm_run_to_residual(lbm_ctx* ctx, double tol, int64_t check_every, int64_t max_steps,
                               int64_t* steps_done, double* final_residual);

/* Reductions over non-solid nodes, FP64: out[0] = sum |u - u_prev|^2,
 * out[1] = sum |u|^2, out[2] = sum rho (mass), out[3] = non-finite count.
 * u_prev is the field at the previous call (zero before the first). */
lbm_status lbm_reduce(lbm_ctx* ctx, double out[4]);

/* lbm_reduce plus the reference's MacroDiff residual (S/Iterate3D.cu:696-756,
 * S/GpuSum.cu:10-23): out[4..7] = max |u - u_prev|, max |v - v_prev|,
 * max |w - w_prev|, max |rho - rho_prev| over non-solid nodes, the fields of
 * the previous reduction (either call) as reference. */
lbm_status lbm_reduce_all(lbm_ctx* ctx, double out[8]);

/* The reference's L2 residual (S/ComputeResiduals.cu:13-27, S/Iterate.cu:551):
 * sqrt of the sum over non-solid nodes and populations of (f - fColl)^2, with
 * f the post-boundary populations of the last step and fColl the
 * post-collision populations that step streamed.  Needs at least one step. */
lbm_status lbm_residual_l2(lbm_ctx* ctx, double* out);
/* lbm_residual_l2 without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); it has no counterpart in the reference.  lbm_residual_l2
 * stages the post-boundary populations in a Q*N FP64 array the context
 * allocates on first use (2.55 GB at 256^3 D3Q19); this variant rebuilds them
 * in registers and keeps its 1024 partial sums and the result in a static
 * device array of 8.3 KB per device (held until the process ends).  The
 * scratch does not grow with the lattice, so no lattice size is refused; it
 * must be called with the context's device current.  Calls that overlap on
 * one device race on the shared scratch. */
lbm_status lbm_residual_l2_noalloc(lbm_ctx* ctx, double* out);
/* The reference's FdRelDiff residual (computeNewResidual3D + gpu_NewResidual,
 * S/ComputeResiduals.cu:28-36, :55-63; driven by S/Iterate3D.cu:677-693):
 * max over non-solid links of |f - f_prev| / |f1|, f the populations after the
 * last step, f_prev those of the previous call, f1 those of the first call
 * since lbm_init / lbm_set_f.  The first call returns 1, the reference's
 * comparison with its zero initial populations.  Keeps two FP64 copies of f. */
lbm_status lbm_residual_fdrel(lbm_ctx* ctx, double* out);

/* Macroscopic fields of the current state (N doubles each; any may be NULL).
 * The velocity is the stored velocity of the force model in use. */
lbm_status lbm_get_macros(lbm_ctx* ctx, double* rho, double* u, double* v, double* w);
/* Post-boundary populations f[q*N + i] of the current state (FP64). */
lbm_status lbm_get_f(lbm_ctx* ctx, double* f);
/* Replace the state by f (q*N doubles); moments are recomputed from f. */
lbm_status lbm_set_f(lbm_ctx* ctx, const double* f);

/* Timing for benchmarks (no counterpart in the reference, whose loop times
 * phases with host clocks, S/Iterate3D.cu:493-653): runs nsteps steps
 * bracketed by HIP events on the context stream; *ms is the elapsed device
 * time of those steps.  The timed launches write no macroscopic fields; one
 * untimed moments pass after the window brings rho, u, v, w up to date. */
lbm_status lbm_time_steps(lbm_ctx* ctx, int64_t nsteps, double* ms);
/* As lbm_time_steps with one event pair per launch: ms_each[k] is the device
 * time of step k (1 <= nsteps <= 100000). */
lbm_status lbm_time_launches(lbm_ctx* ctx, int64_t nsteps, double* ms_each);

/* Bytes of device memory held by the context. */
int64_t lbm_device_bytes(const lbm_ctx* ctx);

/* ---- lattice files (host only; no device needed) ------------------------ */
/* latticegen2 reader: <name>_Node.dat ("i j [k] x y [z] fluid") and
 * <name>_BC.dat ("i j [k] dir type x y [z] id"), replacing readNodeFile /
 * readConnFile / getGridSpacing (S/FilesReading.cu:270-347, :473-490); format
 * of L/main.cpp:59-66 and L/FluidTracking.cpp:229-236.  First call with
 * node_class == NULL returns dims (nx, ny, nz) and the link count; the second
 * fills node_class (0 solid / 1 fluid from the fluid column; in 2D, nodes with
 * INLE links on x = 0 become LBM_VEL_W and nodes with OUTL links on
 * x = nx-1 become LBM_PRES_E), bb_links (WALL links), type_counts[6] (links
 * per type 0..5), dir_counts[19] (links per direction), *delta (grid
 * spacing) and cyclic[3] (an axis whose faces carry CYCL links).  Any output
 * pointer may be NULL.  Returns LBM_EIO on a missing or malformed file. */
lbm_status lbm_read_lattice(const char* node_path, const char* bc_path, int32_t dims[3],
                            int64_t* nlinks, uint8_t* node_class, uint32_t* bb_links,
                            int64_t type_counts[6], int64_t dir_counts[19], double* delta,
                            int32_t cyclic[3]);

/* Wall fractions of the WALL links of a latticegen2 lattice for
 * LBM_WALL_CURVED: wall_q[i*19 + d] = |x_bc - x_i| / (delta |c_d|) from the BC
 * line's intersection point and the node's coordinates (S/GpuInit.cu:1088,
 * :1205); 0 for links without a WALL line.  wall_q holds nx*ny*nz*19 doubles. */
lbm_status lbm_read_wall_q(const char* node_path, const char* bc_path, double* wall_q);

/* Binary copy of a latticegen2 pair: node rows (indices, coordinates, fluid
 * flag) and BC rows (indices, direction, type, point, id) as arrays after an
 * "LBMLAT01" header.  lbm_read_lattice and lbm_read_wall_q take such a file as
 * node_path (bc_path is then ignored) and return what the text pair gives. */
lbm_status lbm_write_lattice_binary(const char* node_path, const char* bc_path, const char* out_path);

/* Lattice generator compatible with the reference's latticegen2
 * (L/main.cpp:17-70, L/GeomRead.cpp:14-189, L/MeshGen.cpp:19-80,
 * L/FluidTracking.cpp:24-236, L/Intersect.cpp:43-528): reads a STAR-CD
 * vertex file (lines "id x y z") and boundary file (lines "n p1 p2 p3 p4 id
 * n type", type WALL / INLE / OUTL / CYCL / SYMP), meshes the vertices'
 * bounding box with nx nodes along x, fills the fluid from the node nearest
 * seed[3] and writes the node and BC files (byte-identical to the reference
 * generator's, complexity 1: links written during the fill, 2: after it) and,
 * when mesh_out is not NULL, the padded mesh listing (its xtecplot.dat).
 * dims receives n, m, h; nlinks the number of BC lines. */
lbm_status lbm_latticegen(const char* vrt_path, const char* bnd_path, int32_t nx, const double seed[3],
                          int32_t complexity, const char* node_out, const char* bc_out, const char* mesh_out,
                          int32_t dims[3], int64_t* nlinks);
/* Message of the calling thread's last lbm_latticegen failure. */
const char* lbm_latticegen_last_error(void);

/* ---- colour-gradient two-phase model (D2Q9 or D3Q19, periodic box) ------- */
/* Replaces the reference's multiphase path.  D2Q9: gpuCollBgkwGC2D +
 * gradient (S/GpuCollision.cu:112-281, :552-622), gpuStreaming2DCG
 * (S/GpuStream.cu:30-158), gpuUpdateMacro2DCG (S/GpuUpdateMacro.cu:43-135) and
 * the Laplace sums of S/Iterate.cu:495-500, driven by S/Iterate.cu:392-500.
 * D3Q19 (nz >= 3): gpuCollBgkwGC3D + gradient (S/GpuCollision.cu:283-420,
 * :709-779), gpuStreaming3DCG (S/GpuStream.cu:196-285), gpuUpdateMacro3DCG
 * (S/GpuUpdateMacro.cu:215-361) and the sums of S/Iterate3D.cu:637-645,
 * driven by S/Iterate3D.cu:495-800.  Colour parameters as in S/main.cu:52-72.
 * A box with walls or a body force runs the reference's order instead
 * (collide, stream, wall condition, macros; two launches per step). */
typedef struct lbm_cg_ctx lbm_cg_ctx;
typedef struct {
    uint32_t abi_version;
    int32_t nx, ny;
    double r_alpha, b_alpha;  /* rest fractions alpha_k */
    double r_nu, b_nu;        /* kinematic viscosities */
    double beta;              /* recolouring parameter */
    double A;                 /* surface-tension parameter: sigma = 4 A / (9 omega) */
    double g_limit;           /* gradient threshold */
    int32_t high_order;       /* 0: 8 / 18-neighbour gradient, 1: 24 / 104-neighbour */
    double kappa;             /* Laplace sampling threshold (control_param) */
    int32_t storage;          /* LBM_STORE_F64, _F32 or _F16 */
    int32_t nz;               /* 0 or 1: D2Q9 box nx x ny; >= 3: D3Q19 box nx x ny x nz */
    int32_t enhanced;         /* 1: f_eq += nu (chi u.grad rho + psi G:cc), the reference's enhanced_distrib
                                 (gpuCollEnhancedBgkwGC2D/3D, S/GpuCollision.cu:624-707, :781-880) */
    /* F64 or F32 storage: walls on the rows (planes) y = 0 and y = ny-1, the
     * other axes periodic; 0 none.  D2Q9 (gpuBcPeriodic2D,
     * S/GpuBoundaries.cu:1876-1968): 1 per-colour Zou-He velocity walls with
     * each wall node's stored velocity, the north row starting at (u_wall, 0)
     * and the south row at rest (the reference's Couette test_case 2); 2
     * on-node bounce-back walls (its Rayleigh-Taylor test_case 6).  Wall rows
     * take the mirrored gradient of S/GpuCollision.cu:135-156 and, with
     * high_order, the next rows the 8-point one.  D3Q19 (the single-phase
     * faces per colour, S/Iterate3D.cu:576-598): 1 the 3D inlet condition
     * (gpuBcInlet3D) with (u_wall, 0, 0) on the north face and rest on the
     * south face (3D Couette test_case 4); 2 on-node bounce-back (gpuBcSimpleWall3D).
     * Wall faces take the mirrored 18-point gradient (S/GpuCollision.cu:306-333)
     * and, with high_order, the next planes the 18-point one (:505-518). */
    int32_t walls;
    double u_wall;
    /* body force by the velocity shift of gpuUpdateMacro2DCG / 3DCG
     * (S/GpuUpdateMacro.cu:118-119, :340-341): external_force 1 along x
     * (u += g / (rho omega_eff)), 0 along y (v += g / omega_eff), omega_eff from
     * the densities of the previous step */
    double g;
    int32_t external_force;
} lbm_cg_params;

lbm_status lbm_cg_create(const lbm_cg_params* p, lbm_cg_ctx** out);
void lbm_cg_destroy(lbm_cg_ctx* ctx);
const char* lbm_cg_last_error(const lbm_cg_ctx* ctx);
/* rest-state colours from colour densities (S/GpuInit.cu:323-352) */
lbm_status lbm_cg_init(lbm_cg_ctx* ctx, const double* r_rho, const double* b_rho);
/* start from colour populations r_f[q*N + i], b_f[q*N + i] (q = 9 or 19),
 * e.g. the reference's steady bubble of initCGBubble3D (S/GpuInit.cu:543-592) */
lbm_status lbm_cg_init_f(lbm_cg_ctx* ctx, const double* r_f, const double* b_f);
lbm_status lbm_cg_step(lbm_cg_ctx* ctx, int64_t nsteps);
lbm_status lbm_cg_get_macros(lbm_cg_ctx* ctx, double* r_rho, double* b_rho, double* rho, double* u, double* v);
/* the same with the z velocity (zero for D2Q9); any pointer may be NULL */
lbm_status lbm_cg_get_macros3d(lbm_cg_ctx* ctx, double* r_rho, double* b_rho, double* rho, double* u, double* v,
                               double* w);
/* post-streaming colour populations r_f[q*N + i], b_f[q*N + i] (needs one step) */
lbm_status lbm_cg_get_f(lbm_cg_ctx* ctx, double* r_f, double* b_f);
/* FP64 sums: 0 sum rho_r, 1 sum rho_b, 2-5 the reference's Laplace sums
 * (rho_r over chi >= kappa and count, rho_b over chi <= -kappa and count;
 * S/GpuUpdateMacro.cu:124-133, :347-357), 6-9 bulk rho over chi > 0.99 and
 * < -0.99 with counts, 10 max |u| (spurious currents), 11 non-finite count. */
lbm_status lbm_cg_reduce(lbm_cg_ctx* ctx, double out[12]);
lbm_status lbm_cg_time_steps(lbm_cg_ctx* ctx, int64_t nsteps, double* ms);

/* ---- batched channels: many D2Q9 lattices in one persistent launch ------- */
/* The reference runs one lattice per process: Iterate2D's loop of
 * collide/stream/boundary/macro launches with a residual check every
 * ShowMacroDiff steps (S/Iterate.cu:385-625, S/ComputeResiduals.cu:13-63).
 * A batch advances N independent D2Q9 channels (open x faces, walls along
 * y given by bb_links and by solid nodes, Zou-He velocity inlet on LBM_VEL_W
 * and pressure outlet on LBM_PRES_E nodes, half-way bounce-back) in one
 * cooperative kernel with a grid barrier per step; each instance stops at
 * its own residual. */
typedef struct lbm_batch lbm_batch;
typedef struct {
    int32_t nx, ny;
    const uint8_t* node_class;  /* nx*ny, x fastest; LBM_SOLID .. LBM_PRES_E */
    const uint32_t* bb_links;   /* nx*ny wall links as in lbm_lattice; NULL: only links into solid nodes */
    const double* u_bc;         /* nx*ny inlet velocity on LBM_VEL_W nodes (NULL: 0) */
    const double* v_bc;
    double omega;               /* 1/tau */
    double trt_magic;           /* Lambda; 0 selects 3/16 */
    double rho_out;             /* outlet density; 0 selects 1 */
} lbm_batch_instance;

/* collision LBM_BGK or LBM_TRT; storage LBM_STORE_F64 (LBM_STORE_F32 returns
 * LBM_EUNSUPPORTED for now).  Instances start at rest (rho = 1, u = 0). */
lbm_status lbm_batch_create(const lbm_batch_instance* inst, int32_t n, int32_t collision, int32_t storage,
                            lbm_batch** out);
void lbm_batch_destroy(lbm_batch* b);
const char* lbm_batch_last_error(const lbm_batch* b);
/* Equilibrium start of instance i (i = -1: every instance) at rho, u, v
 * (nx*ny doubles each; NULL: 1, 0, 0).  Also the residual's reference field. */
lbm_status lbm_batch_init(lbm_batch* b, int32_t i, const double* rho, const double* u, const double* v);
/* One persistent launch: every instance steps until its residual
 * sqrt(sum |u - u_prev|^2 / sum |u|^2) over check_every steps (FP64, over
 * non-solid nodes, as lbm_run_to_residual) falls below tol, or max_steps.
 * steps[i] = steps instance i took, residual[i] = its last residual (NaN:
 * non-finite field), *ms = device time of the launch; any may be NULL.
 * Returns LBM_DIVERGED when an instance went non-finite (outputs filled). */
lbm_status lbm_batch_run(lbm_batch* b, double tol, int32_t check_every, int64_t max_steps, int64_t* steps,
                         double* residual, double* ms);
/* Fields of instance i at its last check step (nx*ny doubles; NULL skips). */
lbm_status lbm_batch_get_macros(lbm_batch* b, int32_t i, double* rho, double* u, double* v);
/* Post-boundary populations f[q*nx*ny + node] of instance i's newest state (FP64). */
lbm_status lbm_batch_get_f(lbm_batch* b, int32_t i, double* f);
/* Total nodes of all instances; *fluid (may be NULL) = non-solid nodes. */
int64_t lbm_batch_nodes(const lbm_batch* b, int64_t* fluid);

#ifdef __cplusplus
}
#endif
#endif
