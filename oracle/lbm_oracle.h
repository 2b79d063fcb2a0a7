/*
 * FP64 CPU oracle for the lattice-Boltzmann time step of ZPaulo/Thesis
 * (`Source code/lbm-solver_final`).
 *
 * TEST INFRASTRUCTURE ONLY.  This restatement is the checker that the HIP
 * product path is compared against; it is linked by the tests, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg, and by nothing
 * that ships.  The product never calls into it.
 *
 * It keeps the reference's data flow: populations f and post-collision
 * populations fColl are separate SoA arrays f[q*N + ind] (S/include/Iterate.h:10-35),
 * macroscopic fields rho,u,v,w are stored and the collision reads them
 * (S/GpuCollision.cu:528-550), and one step is
 *   collide -> stream -> boundaries -> macroscopic update
 * (S/Iterate.cu:385-520, S/Iterate3D.cu:493-653).
 *
 * Every reference behaviour that the north-star configurations replace by a
 * textbook one has a switch here (wall model, inlet model, forcing, MRT
 * quirks), so the same oracle pins both the reference-parity and the
 * textbook modes of the product.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* node classes (shared with the product's lattice description) */
enum { ORC_SOLID = 0, ORC_FLUID = 1, ORC_VEL_W = 2, ORC_PRES_E = 3, ORC_VEL_N = 4,
       /* 3D faces: velocity (W and N above), outlets (east = ORC_PRES_E), symmetry planes */
       ORC_VEL_E = 5, ORC_VEL_S = 6, ORC_VEL_B = 7, ORC_VEL_T = 8,
       ORC_OUT_W = 9, ORC_OUT_S = 10, ORC_OUT_N = 11, ORC_OUT_B = 12, ORC_OUT_T = 13,
       ORC_SYM_W = 14, ORC_SYM_E = 15, ORC_SYM_S = 16, ORC_SYM_N = 17, ORC_SYM_B = 18, ORC_SYM_T = 19 };
enum { ORC_BGK = 0, ORC_TRT = 1, ORC_MRT = 2 };
enum { ORC_FORCE_NONE = 0, ORC_FORCE_SHIFT = 1, ORC_FORCE_GUO = 2 };
/* ORC_WALL_HH: the reference's wet-node walls (HHmodel, S/GpuBoundaries.cu:492-1560),
 * 3D, on nodes with 5 (face), 9 (edge) or 12 (corner) wall links */
/* ORC_WALL_CURVED: Bouzidi linear interpolation with the wall fraction q of
 * each wall link (orc_set_wall_q; 1/2 is half-way bounce-back) */
enum { ORC_WALL_HALFWAY = 0, ORC_WALL_ONNODE = 1, ORC_WALL_HH = 2, ORC_WALL_CURVED = 3 };
enum { ORC_INLET_ZOUHE = 0, ORC_INLET_REF2D = 1 };
/* east outlet: Zou-He pressure, the reference's velocity outlet, its second-
 * and first-order extrapolations (S/GpuBoundaries.cu:1562-1653) */
enum { ORC_OUTLET_PRESSURE = 0, ORC_OUTLET_REF_VELOCITY = 1, ORC_OUTLET_EXTRAP2 = 2, ORC_OUTLET_EXTRAP1 = 3,
       ORC_OUTLET_FEQ = 4 /* 3D: f_eq of the stored macros (S/GpuBoundaries.cu:1840-1852) */ };
/* parity flags: reference quirks reproduced on request */
#define ORC_PARITY_MRT_RHO2      0x1u /* 2D MRT m_eq uses rho^2 (S/GpuCollision.cu:987-988) */
#define ORC_PARITY_MRT_NO_F18    0x2u /* 3D MRT never writes f18 (S/GpuCollision.cu:1161) */
#define ORC_PARITY_INLET_FROZEN  0x4u /* 3D macros skip inlet nodes (S/GpuUpdateMacro.cu:150-151) */
#define ORC_PARITY_OUTLET_V0     0x8u /* 2D v := 0 on east-outlet nodes (S/GpuUpdateMacro.cu:30) */
#define ORC_PARITY_INLET3D_FLIP 0x10u /* 3D W/E velocity faces use 1/(1 +- u) (S/GpuBoundaries.cu:102,148) */

typedef struct {
    int q;              /* 9 (D2Q9) or 19 (D3Q19) */
    int nx, ny, nz;     /* n, m, h in the reference's naming; nz = 1 for 2D */
    int periodic[3];
    int collision;
    double omega;       /* 1/tau = 1/(3 nu + 1/2) (S/GpuInit.cu:102) */
    double trt_magic;   /* Lambda; 3/16 reproduces omegaA = 8(2-w)/(8-w) (S/GpuInit.cu:103) */
    double mrt_s[19];   /* diag(S); all zero selects the reference's tables */
    int force_model;
    double g[3];
    int wall_model;
    int inlet_model;
    unsigned parity;
    double rho_out;     /* Zou-He pressure outlet density */
    int nthreads;       /* OpenMP threads; <= 0 uses the runtime default */
    int outlet_model;   /* ORC_OUTLET_*, on ORC_PRES_E nodes */
} orc_params;

typedef struct orc_sim orc_sim;

/* Lattice tables (reference numbering, S/GpuInit.cu:94-99 and :822-842). */
int orc_lattice(int q, int* cx, int* cy, int* cz, int* opp, double* w);

/* MRT matrices built from the moment polynomials; M and Minv are Q*Q row-major.
 * s receives the reference's diag(S) for the given omega
 * (S/CellFunctions.cu:44-52 for D2Q9, :122-140 for D3Q19). */
int orc_mrt_matrices(int q, double omega, double* M, double* Minv, double* s);

/* Create a simulation.  node_class[N], bb[N] (bit d: the link from the node
 * in direction d crosses a wall), ubc/vbc/wbc: prescribed velocity on
 * velocity nodes (west inlet, north lid) and u0 of the reference's velocity
 * outlet (may be NULL). */
orc_sim* orc_create(const orc_params* p, const uint8_t* node_class, const uint32_t* bb,
                    const double* ubc, const double* vbc, const double* wbc);
void orc_destroy(orc_sim* s);
/* q[ind*19 + d]: fraction of the link from node ind in direction d at which it
 * meets the wall (values <= 0 mean 1/2) */
void orc_set_wall_q(orc_sim* s, const double* q);

/* mode 0: f = f_eq(rho,u); mode 1: f = 0 (the reference's start, S/Iterate3D.cu:301-302).
 * rho/u/v/w may be NULL (1 and 0). Stored macros are set to the given fields. */
void orc_init(orc_sim* s, int mode, const double* rho, const double* u, const double* v, const double* w);
void orc_step(orc_sim* s, int nsteps);

void orc_get_f(const orc_sim* s, double* f);         /* post-boundary f (Q*N) */
void orc_set_f(orc_sim* s, const double* f);         /* also recomputes the stored macros */
void orc_get_macros(const orc_sim* s, double* rho, double* u, double* v, double* w);
double orc_mass(const orc_sim* s);                   /* sum of f over non-solid nodes */
double orc_residual_l2(const orc_sim* s);            /* sqrt(sum (f - fColl)^2) of the last step */

/* Kernels exposed for closed-form tests. */
void orc_collide_node(const orc_sim* s, const double* f, double rho, const double* uvw, double* out);
void orc_equilibrium(int q, double rho, const double* uvw, double* feq);
/* The node-local open-boundary condition of node ind applied to its streamed
 * populations f[Q]: 2D west inlet, north lid, east outlet models 0 and 1;
 * 3D velocity faces and symmetry planes. */
void orc_open_boundary(const orc_sim* s, long ind, double* f);
/* The wall condition of node ind (on-node bounce-back or wet-node walls). */
void orc_wall_node(const orc_sim* s, long ind, double* f);

/* ---- colour-gradient two-phase model (D2Q9, or D3Q19 when nz > 1) ------ */
typedef struct {
    int nx, ny;
    double r_alpha, b_alpha;     /* rest fractions alpha_k */
    double r_nu, b_nu;           /* kinematic viscosities */
    double beta;                 /* recolouring parameter */
    double A;                    /* surface-tension parameter */
    double g_limit;              /* gradient threshold */
    int high_order;              /* 0: 8-neighbour stencil, 1: 24-neighbour stencil */
    double g; int external_force;/* force (S/GpuUpdateMacro.cu:118-119) */
    double kappa;                /* Laplace sampling threshold (control_param) */
    int nthreads;
    int nz;                      /* > 1: D3Q19 box nx x ny x nz */
    int enhanced;                /* 1: density-gradient correction of f_eq (gpuCollEnhancedBgkwGC2D/3D) */
    /* Walls on the rows (planes) y = 0 (south) and y = ny-1 (north), the
     * other axes periodic.  D2Q9: 1: per-colour Zou-He velocity walls with
     * the node's stored velocity, north row initialised to (u_wall, 0), south
     * to rest (test case 2, Couette); 2: on-node bounce-back of the incoming
     * links (test case 6, Rayleigh-Taylor).  gpuBcPeriodic2D,
     * S/GpuBoundaries.cu:1876-1968; wall rows take the mirrored gradient of
     * S/GpuCollision.cu:135-156 and, with the 24-point stencil, the next rows
     * the 8-point one (:262-275).  D3Q19: the single-phase faces applied per
     * colour (S/Iterate3D.cu:576-598): 1: the 3D inlet condition with
     * (u_wall, 0, 0) on the north face and rest on the south one; 2: on-node
     * bounce-back; wall faces take the mirrored 18-point gradient
     * (S/GpuCollision.cu:306-333), the planes next to them the 18-point one
     * with the 104-point stencil (:505-518). */
    int walls;
    double u_wall;
} orc_cg_params;

typedef struct orc_cg orc_cg;
orc_cg* orc_cg_create(const orc_cg_params* p);
void orc_cg_destroy(orc_cg* s);
/* rest-state colours from the given density fields (S/GpuInit.cu:323-352) */
void orc_cg_init(orc_cg* s, const double* r_rho, const double* b_rho);
/* colour populations r_f[q*N + i], b_f[q*N + i] as the state before the
 * first collision, e.g. the reference's steady bubble of initCGBubble3D (S/GpuInit.cu:543-592) */
void orc_cg_init_f(orc_cg* s, const double* r_f, const double* b_f);
void orc_cg_step(orc_cg* s, int nsteps);
void orc_cg_get(const orc_cg* s, double* r_f, double* b_f, double* r_rho, double* b_rho,
                double* rho, double* u, double* v);
void orc_cg_get_w(const orc_cg* s, double* w);
/* out[0]=sum p_in, out[1]=n_in, out[2]=sum p_out, out[3]=n_out, out[4]=max|u| */
void orc_cg_laplace(const orc_cg* s, double* out);
/* gradient field (cg_x, cg_y, cg_z) for tests of the stencil; cgz may be NULL */
void orc_cg_gradient(const orc_cg* s, double* cgx, double* cgy, double* cgz);

/* ---- the reference's CPU colour-gradient path (D2Q9 and D3Q19; test
 * infrastructure, a second opinion beside the GPU formulation) ---- */
typedef struct {
    int nx, ny;                  /* the full lattice, edges included */
    double r_alpha, b_alpha, r_nu, b_nu, beta, A, g_limit;
    int test_case;               /* 2: Zou-He walls on the first and last rows; else edge copies (D2Q9) */
    double u_wall;               /* north-row velocity of test case 2 */
    int nz;                      /* > 1: the D3Q19 path on an nx x ny x nz box */
} orc_cgcpu_params;
typedef struct orc_cgcpu orc_cgcpu;
orc_cgcpu* orc_cgcpu_create(const orc_cgcpu_params* p);
void orc_cgcpu_destroy(orc_cgcpu* s);
void orc_cgcpu_init(orc_cgcpu* s, const double* r_rho, const double* b_rho);
void orc_cgcpu_step(orc_cgcpu* s, int nsteps);
void orc_cgcpu_get(const orc_cgcpu* s, double* r_f, double* b_f, double* r_rho, double* b_rho, double* rho,
                   double* u, double* v);
void orc_cgcpu_get_coll(const orc_cgcpu* s, double* r_c, double* b_c, double* gx, double* gy);
/* colour populations r_f[q*N + i], b_f[q*N + i] as the state before the
 * first collision; the macros of every node follow from them */
void orc_cgcpu_init_f(orc_cgcpu* s, const double* r_f, const double* b_f);
/* D3Q19: the z velocity and the z component of the last step's gradient */
void orc_cgcpu_get_w(const orc_cgcpu* s, double* w, double* gz);

#ifdef __cplusplus
}
#endif
#endif
