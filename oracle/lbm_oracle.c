/*
 * FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (see lbm_oracle.h).
 *
 * Restates the reference's per-node arithmetic; each function names the
 * reference lines it follows.  Written for clarity, parallelised with OpenMP
 * over nodes so that it can also serve as the timed CPU baseline.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* lattice tables: reference numbering (S/GpuInit.cu:94-99, :822-842)        */
static const int CX9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int CX19[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int CY19[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int CZ19[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};

int orc_lattice(int q, int* cx, int* cy, int* cz, int* opp, double* w)
{
    if (q != 9 && q != 19) return -1;
    for (int i = 0; i < q; ++i) {
        cx[i] = q == 9 ? CX9[i] : CX19[i];
        cy[i] = q == 9 ? CY9[i] : CY19[i];
        cz[i] = q == 9 ? 0 : CZ19[i];
    }
    for (int i = 0; i < q; ++i) {
        int c2 = cx[i] * cx[i] + cy[i] * cy[i] + cz[i] * cz[i];
        if (q == 9) w[i] = c2 == 0 ? 4.0 / 9.0 : (c2 == 1 ? 1.0 / 9.0 : 1.0 / 36.0);
        else w[i] = c2 == 0 ? 1.0 / 3.0 : (c2 == 1 ? 1.0 / 18.0 : 1.0 / 36.0);
        for (int j = 0; j < q; ++j)
            if (cx[j] == -cx[i] && cy[j] == -cy[i] && cz[j] == -cz[i]) opp[i] = j;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MRT: rows of M as moment polynomials of c.  The reference tabulates the
 * same matrices (S/CellFunctions.cu:15-36 for D2Q9, :73-114 for D3Q19);
 * rows are orthogonal, so Minv = M^T diag(1/|row|^2).                       */
static double mrt_row(int q, int r, int cx, int cy, int cz)
{
    double c2 = cx * cx + cy * cy + cz * cz;
    if (q == 9) {
        switch (r) {
        case 0: return 1;
        case 1: return 3 * c2 - 4;
        case 2: return (9 * c2 * c2 - 21 * c2 + 8) / 2;
        case 3: return cx;
        case 4: return (3 * c2 - 5) * cx;
        case 5: return cy;
        case 6: return (3 * c2 - 5) * cy;
        case 7: return cx * cx - cy * cy;
        default: return cx * cy;
        }
    }
    switch (r) {
    case 0: return 1;
    case 1: return 19 * c2 - 30;
    case 2: return (21 * c2 * c2 - 53 * c2 + 24) / 2;
    case 3: return cx;
    case 4: return (5 * c2 - 9) * cx;
    case 5: return cy;
    case 6: return (5 * c2 - 9) * cy;
    case 7: return cz;
    case 8: return (5 * c2 - 9) * cz;
    case 9: return 3 * cx * cx - c2;
    case 10: return (3 * c2 - 5) * (3 * cx * cx - c2);
    case 11: return cy * cy - cz * cz;
    case 12: return (3 * c2 - 5) * (cy * cy - cz * cz);
    case 13: return cx * cy;
    case 14: return cy * cz;
    case 15: return cx * cz;
    case 16: return (cy * cy - cz * cz) * cx;
    case 17: return (cz * cz - cx * cx) * cy;
    default: return (cx * cx - cy * cy) * cz;
    }
}

int orc_mrt_matrices(int q, double omega, double* M, double* Minv, double* s)
{
    int cx[19], cy[19], cz[19], opp[19];
    double w[19];
    if (orc_lattice(q, cx, cy, cz, opp, w)) return -1;
    for (int r = 0; r < q; ++r) {
        double n2 = 0;
        for (int j = 0; j < q; ++j) {
            M[r * q + j] = mrt_row(q, r, cx[j], cy[j], cz[j]);
            n2 += M[r * q + j] * M[r * q + j];
        }
        for (int j = 0; j < q; ++j) Minv[j * q + r] = M[r * q + j] / n2;
    }
    if (q == 9) { /* S/CellFunctions.cu:44-52 */
        const double s9[9] = {1.0, 1.4, 1.4, 1.0, 1.2, 1.0, 1.2, omega, omega};
        memcpy(s, s9, sizeof s9);
    } else {      /* S/CellFunctions.cu:122-140 */
        const double s19[19] = {0.0, 1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 1.2, omega,
                                1.4, omega, 1.4, omega, omega, omega, 1.98, 1.98, 1.98};
        memcpy(s, s19, sizeof s19);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
struct orc_sim {
    orc_params p;
    long n;
    int Q;
    int cx[19], cy[19], cz[19], opp[19];
    double w[19];
    double omega_m;                 /* TRT antisymmetric rate */
    double M[361], Minv[361], s[19];
    uint8_t* cls;
    uint32_t* bb;
    double *ubc, *vbc, *wbc;
    double *f, *fcoll, *rho, *u, *v, *w3;
    double* wq;                     /* wall fractions (N x 19) or NULL */
};

/* f_eq = rho w (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u)  (S/GpuCollision.cu:21-32) */
static void feq_all(const orc_sim* s, double rho, const double* uvw, double* feq)
{
    double uu = uvw[0] * uvw[0] + uvw[1] * uvw[1] + uvw[2] * uvw[2];
    for (int i = 0; i < s->Q; ++i) {
        double cu = s->cx[i] * uvw[0] + s->cy[i] * uvw[1] + s->cz[i] * uvw[2];
        feq[i] = rho * s->w[i] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * uu);
    }
}

void orc_equilibrium(int q, double rho, const double* uvw, double* feq)
{
    orc_sim tmp;
    memset(&tmp, 0, sizeof tmp);
    tmp.Q = q;
    orc_lattice(q, tmp.cx, tmp.cy, tmp.cz, tmp.opp, tmp.w);
    feq_all(&tmp, rho, uvw, feq);
}

/* Guo source without its relaxation prefactor:
 * S_i = w_i [3 (c_i - u) + 9 (c_i.u) c_i] . F, with F = rho g. */
static void guo_source(const orc_sim* s, double rho, const double* uvw, double* S)
{
    const double* g = s->p.g;
    double F[3] = {rho * g[0], rho * g[1], rho * g[2]};
    for (int i = 0; i < s->Q; ++i) {
        double c[3] = {s->cx[i], s->cy[i], s->cz[i]};
        double cu = c[0] * uvw[0] + c[1] * uvw[1] + c[2] * uvw[2];
        double acc = 0;
        for (int d = 0; d < 3; ++d) acc += (3.0 * (c[d] - uvw[d]) + 9.0 * cu * c[d]) * F[d];
        S[i] = s->w[i] * acc;
    }
}

/* MRT equilibrium moments.
 * D2Q9: S/GpuCollision.cu:986-994 (rho^2 in e, eps under ORC_PARITY_MRT_RHO2).
 * D3Q19: S/GpuCollision.cu:1065-1083 (j = rho u; incompressible j^2 forms). */
static void mrt_meq(const orc_sim* s, double r, const double* uvw, double* meq)
{
    if (s->Q == 9) {
        double u = uvw[0], v = uvw[1], uu = u * u + v * v;
        double k = (s->p.parity & ORC_PARITY_MRT_RHO2) ? r : 1.0;
        meq[0] = r;
        meq[1] = r * (-2.0 + 3.0 * k * uu);
        meq[2] = r * (1.0 - 3.0 * k * uu);
        meq[3] = r * u;
        meq[4] = -r * u;
        meq[5] = r * v;
        meq[6] = -r * v;
        meq[7] = r * (u * u - v * v);
        meq[8] = r * u * v;
        return;
    }
    double jx = r * uvw[0], jy = r * uvw[1], jz = r * uvw[2];
    double j2 = jx * jx + jy * jy + jz * jz;
    meq[0] = r;
    meq[1] = -11.0 * r + 19.0 * j2;
    meq[2] = -475.0 / 63.0 * j2;
    meq[3] = jx;
    meq[4] = -2.0 / 3.0 * jx;
    meq[5] = jy;
    meq[6] = -2.0 / 3.0 * jy;
    meq[7] = jz;
    meq[8] = -2.0 / 3.0 * jz;
    meq[9] = 2.0 * jx * jx - (jy * jy + jz * jz);
    meq[10] = 0.0;
    meq[11] = jy * jy - jz * jz;
    meq[12] = 0.0;
    meq[13] = jx * jy;
    meq[14] = jy * jz;
    meq[15] = jx * jz;
    meq[16] = meq[17] = meq[18] = 0.0;
}

/* Collision of one node from its populations f and stored macros (rho, u).
 * BGK: S/GpuCollision.cu:528-550, :882-916.  TRT: :917-969 with a general
 * magic parameter.  MRT: :971-1037, :1039-1170.  Forcing: the reference
 * shifts the stored velocity (S/GpuUpdateMacro.cu:207), which enters here
 * through u; GUO adds the source term with the (1 - s/2) prefactors. */
void orc_collide_node(const orc_sim* s, const double* f, double rho, const double* uvw, double* out)
{
    const int Q = s->Q;
    double feq[19], S[19];
    const int guo = s->p.force_model == ORC_FORCE_GUO;
    if (guo) guo_source(s, rho, uvw, S);
    if (s->p.collision == ORC_BGK) {
        const double om = s->p.omega;
        feq_all(s, rho, uvw, feq);
        for (int i = 0; i < Q; ++i) {
            out[i] = om * feq[i] + (1.0 - om) * f[i];
            if (guo) out[i] += (1.0 - 0.5 * om) * S[i];
        }
    } else if (s->p.collision == ORC_TRT) {
        const double op = s->p.omega, om = s->omega_m;
        feq_all(s, rho, uvw, feq);
        for (int i = 0; i < Q; ++i) {
            int o = s->opp[i];
            double fp = 0.5 * (f[i] + f[o]), fm = 0.5 * (f[i] - f[o]);
            double ep = 0.5 * (feq[i] + feq[o]), em = 0.5 * (feq[i] - feq[o]);
            out[i] = f[i] - op * (fp - ep) - om * (fm - em);
            if (guo) {
                double sp = 0.5 * (S[i] + S[o]), sm = 0.5 * (S[i] - S[o]);
                out[i] += (1.0 - 0.5 * op) * sp + (1.0 - 0.5 * om) * sm;
            }
        }
    } else {
        double m[19], meq[19], dm[19];
        mrt_meq(s, rho, uvw, meq);
        for (int r = 0; r < Q; ++r) {
            double acc = 0;
            for (int j = 0; j < Q; ++j) acc += s->M[r * Q + j] * f[j];
            m[r] = acc;
            dm[r] = s->s[r] * (m[r] - meq[r]);
        }
        double sm[19];
        if (guo) {
            for (int r = 0; r < Q; ++r) {
                double acc = 0;
                for (int j = 0; j < Q; ++j) acc += s->M[r * Q + j] * S[j];
                sm[r] = (1.0 - 0.5 * s->s[r]) * acc;
            }
        }
        for (int i = 0; i < Q; ++i) {
            double acc = 0;
            for (int r = 0; r < Q; ++r) acc += s->Minv[i * Q + r] * dm[r];
            out[i] = f[i] - acc;
            if (guo) {
                double a2 = 0;
                for (int r = 0; r < Q; ++r) a2 += s->Minv[i * Q + r] * sm[r];
                out[i] += a2;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
static void set_threads(int nt)
{
#ifdef _OPENMP
    if (nt > 0) omp_set_num_threads(nt);
#else
    (void)nt;
#endif
}

orc_sim* orc_create(const orc_params* p, const uint8_t* node_class, const uint32_t* bb,
                    const double* ubc, const double* vbc, const double* wbc)
{
    if (p->q != 9 && p->q != 19) return NULL;
    orc_sim* s = (orc_sim*)calloc(1, sizeof *s);
    s->p = *p;
    if (s->p.nz < 1) s->p.nz = 1;
    s->Q = p->q;
    s->n = (long)p->nx * p->ny * s->p.nz;
    orc_lattice(s->Q, s->cx, s->cy, s->cz, s->opp, s->w);
    {
        /* Lambda = (1/w+ - 1/2)(1/w- - 1/2) */
        double lam = p->trt_magic > 0 ? p->trt_magic : 3.0 / 16.0;
        s->omega_m = 1.0 / (0.5 + lam / (1.0 / p->omega - 0.5));
    }
    orc_mrt_matrices(s->Q, p->omega, s->M, s->Minv, s->s);
    {
        int any = 0;
        for (int i = 0; i < s->Q; ++i) any |= p->mrt_s[i] != 0.0;
        if (any) memcpy(s->s, p->mrt_s, sizeof(double) * s->Q);
    }
    long n = s->n;
    s->cls = (uint8_t*)malloc(n);
    s->bb = (uint32_t*)malloc(n * sizeof(uint32_t));
    memcpy(s->cls, node_class, n);
    if (bb) memcpy(s->bb, bb, n * sizeof(uint32_t));
    else memset(s->bb, 0, n * sizeof(uint32_t));
    s->ubc = (double*)calloc(n, sizeof(double));
    s->vbc = (double*)calloc(n, sizeof(double));
    s->wbc = (double*)calloc(n, sizeof(double));
    if (ubc) memcpy(s->ubc, ubc, n * sizeof(double));
    if (vbc) memcpy(s->vbc, vbc, n * sizeof(double));
    if (wbc) memcpy(s->wbc, wbc, n * sizeof(double));
    s->f = (double*)calloc((size_t)n * s->Q, sizeof(double));
    s->fcoll = (double*)calloc((size_t)n * s->Q, sizeof(double));
    s->rho = (double*)calloc(n, sizeof(double));
    s->u = (double*)calloc(n, sizeof(double));
    s->v = (double*)calloc(n, sizeof(double));
    s->w3 = (double*)calloc(n, sizeof(double));
    return s;
}

void orc_set_wall_q(orc_sim* s, const double* q)
{
    if (!s->wq) s->wq = (double*)malloc(sizeof(double) * 19 * (size_t)s->n);
    memcpy(s->wq, q, sizeof(double) * 19 * (size_t)s->n);
}

void orc_destroy(orc_sim* s)
{
    if (!s) return;
    free(s->wq);
    free(s->cls); free(s->bb); free(s->ubc); free(s->vbc); free(s->wbc);
    free(s->f); free(s->fcoll); free(s->rho); free(s->u); free(s->v); free(s->w3);
    free(s);
}

/* macroscopic update of one node (S/GpuUpdateMacro.cu:6-41, :137-213) */
static void macro_node(orc_sim* s, long ind)
{
    const int Q = s->Q;
    const long n = s->n;
    uint8_t c = s->cls[ind];
    if (c == ORC_SOLID) return;
    if ((s->p.parity & ORC_PARITY_INLET_FROZEN) && (c == ORC_VEL_W || c == ORC_VEL_N || (c >= ORC_VEL_E && c <= ORC_VEL_T)))
        return;
    double r = 0, jx = 0, jy = 0, jz = 0;
    for (int i = 0; i < Q; ++i) {
        double fi = s->f[(size_t)i * n + ind];
        r += fi;
        jx += s->cx[i] * fi;
        jy += s->cy[i] * fi;
        jz += s->cz[i] * fi;
    }
    double u = jx / r, v = jy / r, w = jz / r;
    if (s->p.force_model == ORC_FORCE_SHIFT) {
        u += s->p.g[0] / s->p.omega;
        v += s->p.g[1] / s->p.omega;
        w += s->p.g[2] / s->p.omega;
    } else if (s->p.force_model == ORC_FORCE_GUO) {
        u += 0.5 * s->p.g[0];
        v += 0.5 * s->p.g[1];
        w += 0.5 * s->p.g[2];
    }
    if ((s->p.parity & ORC_PARITY_OUTLET_V0) && c == ORC_PRES_E) v = 0.0;
    s->rho[ind] = r;
    s->u[ind] = u;
    s->v[ind] = v;
    s->w3[ind] = w;
}

void orc_init(orc_sim* s, int mode, const double* rho, const double* u, const double* v, const double* w)
{
    const long n = s->n;
    const int Q = s->Q;
    set_threads(s->p.nthreads);
#pragma omp parallel for
    for (long ind = 0; ind < n; ++ind) {
        double r = rho ? rho[ind] : 1.0;
        double uvw[3] = {u ? u[ind] : 0.0, v ? v[ind] : 0.0, w ? w[ind] : 0.0};
        s->rho[ind] = r;
        s->u[ind] = uvw[0];
        s->v[ind] = uvw[1];
        s->w3[ind] = uvw[2];
        double feq[19];
        feq_all(s, r, uvw, feq);
        for (int i = 0; i < Q; ++i) {
            s->f[(size_t)i * n + ind] = (mode == 0 && s->cls[ind] != ORC_SOLID) ? feq[i] : 0.0;
            s->fcoll[(size_t)i * n + ind] = 0.0;
        }
    }
}

static inline long wrap_src(const orc_sim* s, int x, int y, int z, int i, int* outside)
{
    int sx = x - s->cx[i], sy = y - s->cy[i], sz = z - s->cz[i];
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    *outside = 0;
    if (sx < 0 || sx >= nx) { if (s->p.periodic[0]) sx = (sx + nx) % nx; else *outside = 1; }
    if (sy < 0 || sy >= ny) { if (s->p.periodic[1]) sy = (sy + ny) % ny; else *outside = 1; }
    if (sz < 0 || sz >= nz) { if (s->p.periodic[2]) sz = (sz + nz) % nz; else *outside = 1; }
    return ((long)sz * ny + sy) * nx + sx;
}

/* Zou-He boundaries on D2Q9 faces.
 * Velocity inlet, west face: textbook (Zou & He 1997) or the reference's
 * simplified form without the transverse terms (S/GpuBoundaries.cu:45-64).
 * Moving lid, north face: textbook, solved from the density, the two
 * momentum components and non-equilibrium bounce-back of f4, or the
 * reference's form (S/GpuBoundaries.cu:21-41): f4 = f2 and
 * f8, f7 = f6, f5 -+ B u0/6 with B = f0 + f1 + f3 + 2(f2 + f5 + f6), corner
 * nodes skipped.
 * East outlet: Zou-He pressure, the closed form of
 * T/LBM/Tom/Lecture Script/LBMchannel.m:254-270 (absent from the solver), or
 * the reference's velocity outlet (S/GpuBoundaries.cu:1571-1605) whose
 * density R = (f0 + f2 + f4 + 2(f1 + f5 + f8)) / (1 - u0) keeps its sign.
 * The extrapolating outlets are applied afterwards (outlet_extrap). */
static void zouhe_node(const orc_sim* s, long ind, double* f)
{
    uint8_t c = s->cls[ind];
    if (s->Q != 9) return;
    if (c == ORC_VEL_N) {
        double u0 = s->ubc[ind], v0 = s->vbc[ind];
        if (s->p.inlet_model == ORC_INLET_REF2D) {
            if (s->bb[ind]) return;
            double B = f[0] + f[1] + f[3] + 2.0 * (f[2] + f[6] + f[5]);
            f[4] = f[2];
            f[8] = f[6] + B * u0 / 6.0;
            f[7] = f[5] - B * u0 / 6.0;
        } else {
            double r = (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / (1.0 + v0);
            f[4] = f[2] - 2.0 / 3.0 * r * v0;
            f[7] = f[5] + 0.5 * (f[1] - f[3]) - 0.5 * r * u0 - r * v0 / 6.0;
            f[8] = f[6] - 0.5 * (f[1] - f[3]) + 0.5 * r * u0 - r * v0 / 6.0;
        }
    } else if (c == ORC_VEL_W) {
        double u0 = s->ubc[ind], v0 = s->vbc[ind];
        if (s->p.inlet_model == ORC_INLET_REF2D) {
            if (s->bb[ind]) return;   /* corner nodes are skipped (BC_CORNER) */
            double r = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u0);
            f[1] = f[3] + 2.0 * r * u0 / 3.0;
            f[5] = f[7] + r * u0 / 6.0;
            f[8] = f[6] + r * u0 / 6.0;
        } else {
            double r = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - u0);
            f[1] = f[3] + 2.0 / 3.0 * r * u0;
            f[5] = f[7] - 0.5 * (f[2] - f[4]) + r * u0 / 6.0 + 0.5 * r * v0;
            f[8] = f[6] + 0.5 * (f[2] - f[4]) + r * u0 / 6.0 - 0.5 * r * v0;
        }
    } else if (c == ORC_PRES_E && s->p.outlet_model == ORC_OUTLET_REF_VELOCITY) {
        double u0 = s->ubc[ind];
        double R = (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / (1.0 - u0);
        f[3] = f[1] - 2.0 * R * u0 / 3.0;
        f[7] = f[5] - R * u0 / 6.0;
        f[6] = f[8] - R * u0 / 6.0;
    } else if (c == ORC_PRES_E && s->p.outlet_model == ORC_OUTLET_PRESSURE) {
        double ro = s->p.rho_out;
        double ux = -1.0 + (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / ro;
        f[3] = f[1] - 2.0 / 3.0 * ro * ux;
        f[7] = f[5] + 0.5 * (f[2] - f[4]) - ro * ux / 6.0;
        f[6] = f[8] - 0.5 * (f[2] - f[4]) - ro * ux / 6.0;
    }
}

/* ---- D3Q19 faces, walls, outlets, symmetry ---------------------------------
 * A face is (axis a, inward sign g): W (0,+1), E (0,-1), S (1,+1), N (1,-1),
 * B (2,+1), T (2,-1).  Unknowns of a face node are the populations with
 * g c_a > 0 (they would stream in from outside the lattice). */
static int dir3(int x, int y, int z)
{
    for (int i = 0; i < 19; ++i)
        if (CX19[i] == x && CY19[i] == y && CZ19[i] == z) return i;
    return -1;
}
static int comp(int i, int a) { return a == 0 ? CX19[i] : a == 1 ? CY19[i] : CZ19[i]; }
static int dir_ax(int a, int ca, int b, int cb)   /* direction with c_a = ca, c_b = cb, others 0 */
{
    int c[3] = {0, 0, 0};
    c[a] += ca;
    c[b] += cb;
    return dir3(c[0], c[1], c[2]);
}

/* face of a 3D class: returns kind (1 velocity, 2 outlet, 3 symmetry, 0 none) */
static int face_of(uint8_t c, int* a, int* g)
{
    static const int VA[6] = {0, 0, 1, 1, 2, 2}, VG[6] = {1, -1, 1, -1, 1, -1};
    int k = 0, f = -1;
    switch (c) {
    case ORC_VEL_W: k = 1; f = 0; break;  case ORC_VEL_E: k = 1; f = 1; break;
    case ORC_VEL_S: k = 1; f = 2; break;  case ORC_VEL_N: k = 1; f = 3; break;
    case ORC_VEL_B: k = 1; f = 4; break;  case ORC_VEL_T: k = 1; f = 5; break;
    case ORC_OUT_W: k = 2; f = 0; break;  case ORC_PRES_E: k = 2; f = 1; break;
    case ORC_OUT_S: k = 2; f = 2; break;  case ORC_OUT_N: k = 2; f = 3; break;
    case ORC_OUT_B: k = 2; f = 4; break;  case ORC_OUT_T: k = 2; f = 5; break;
    default:
        if (c >= ORC_SYM_W && c <= ORC_SYM_T) { k = 3; f = c - ORC_SYM_W; }
    }
    if (f >= 0) { *a = VA[f]; *g = VG[f]; }
    return k;
}

/* Velocity face after Hecht & Harting (2010), the formulation of
 * gpuBcInlet3D (S/GpuBoundaries.cu:73-399): rho from the tangential and
 * outgoing populations, the normal unknown by bounce-back of the
 * non-equilibrium part, the four diagonal unknowns with the transverse
 * momentum corrections N_ab.  flip: the reference's W/E denominator. */
static void velocity_face(double* f, int a, int g, const double* U, int flip)
{
    double K = 0;
    for (int i = 0; i < 19; ++i) {
        int ca = comp(i, a);
        if (ca == 0) K += f[i];
        else if (ca == -g) K += 2.0 * f[i];
    }
    double rho = K / (flip ? 1.0 + g * U[a] : 1.0 - g * U[a]);
    f[dir_ax(a, g, a, 0)] = f[dir_ax(a, -g, a, 0)] + g * rho * U[a] / 3.0;
    for (int b = 0; b < 3; ++b) {
        if (b == a) continue;
        double N = 0;
        for (int i = 0; i < 19; ++i)
            if (comp(i, a) == 0) N += 0.5 * comp(i, b) * f[i];
        N -= rho * U[b] / 3.0;
        for (int t = -1; t <= 1; t += 2)
            f[dir_ax(a, g, b, t)] = f[dir_ax(a, -g, b, -t)] + rho * (g * U[a] + t * U[b]) / 6.0 - t * N;
    }
}

/* Wet-node walls of gpuBcComplexWall3D: a node with 5 wall links is a face
 * (the velocity face with u = 0), 9 an edge, 12 a corner; the inward normal
 * along an axis is -1 when the +axis link is a wall, +1 for the -axis link.
 * Edge with normals (a, ga), (b, gb) and tangent t: bounce-back of the normal
 * and in-plane diagonal unknowns, N_t = -(f_+t - f_-t)/4 on the diagonals
 * with a t component, the two buried diagonals and the rest population share
 * the sum S of the other 16: S/22 each, 12 S/22.  Corner: bounce-back of the
 * three normal and three diagonal unknowns, the six buried diagonals C =
 * (sum of the 12 others)/18, rest 12 C.  Other nodes keep their values. */
static void hh_node(uint32_t bb, double* f)
{
    int cnt = 0, nrm[3] = {0, 0, 0}, nn = 0;
    for (int d = 1; d < 19; ++d) cnt += (bb >> d) & 1u;
    for (int a = 0; a < 3; ++a) {
        if ((bb >> dir_ax(a, 1, a, 0)) & 1u) nrm[a] = -1;
        else if ((bb >> dir_ax(a, -1, a, 0)) & 1u) nrm[a] = 1;
        nn += nrm[a] != 0;
    }
    if (cnt == 5 && nn == 1) {
        const double U[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            if (nrm[a]) velocity_face(f, a, nrm[a], U, 0);
    } else if (cnt == 9 && nn == 2) {
        int a = -1, b = -1, t = -1;
        for (int k = 0; k < 3; ++k) {
            if (!nrm[k]) t = k;
            else if (a < 0) a = k;
            else b = k;
        }
        int ga = nrm[a], gb = nrm[b];
        f[dir_ax(a, ga, a, 0)] = f[dir_ax(a, -ga, a, 0)];
        f[dir_ax(b, gb, b, 0)] = f[dir_ax(b, -gb, b, 0)];
        f[dir_ax(a, ga, b, gb)] = f[dir_ax(a, -ga, b, -gb)];
        double N = -0.25 * (f[dir_ax(t, 1, t, 0)] - f[dir_ax(t, -1, t, 0)]);
        f[dir_ax(a, ga, t, 1)] = f[dir_ax(a, -ga, t, -1)] + N;
        f[dir_ax(a, ga, t, -1)] = f[dir_ax(a, -ga, t, 1)] - N;
        f[dir_ax(b, gb, t, 1)] = f[dir_ax(b, -gb, t, -1)] + N;
        f[dir_ax(b, gb, t, -1)] = f[dir_ax(b, -gb, t, 1)] - N;
        int q1 = dir_ax(a, ga, b, -gb), q2 = dir_ax(a, -ga, b, gb);
        double S = 0;
        for (int i = 1; i < 19; ++i)
            if (i != q1 && i != q2) S += f[i];
        f[q1] = f[q2] = S / 22.0;
        f[0] = 12.0 * S / 22.0;
    } else if (cnt == 12 && nn == 3) {
        int bur[6], nb = 0;
        for (int a = 0; a < 3; ++a) f[dir_ax(a, nrm[a], a, 0)] = f[dir_ax(a, -nrm[a], a, 0)];
        for (int a = 0; a < 3; ++a)
            for (int b = a + 1; b < 3; ++b) {
                f[dir_ax(a, nrm[a], b, nrm[b])] = f[dir_ax(a, -nrm[a], b, -nrm[b])];
                bur[nb++] = dir_ax(a, nrm[a], b, -nrm[b]);
                bur[nb++] = dir_ax(a, -nrm[a], b, nrm[b]);
            }
        double S = 0;
        for (int i = 1; i < 19; ++i) {
            int buried = 0;
            for (int k = 0; k < 6; ++k) buried |= bur[k] == i;
            if (!buried) S += f[i];
        }
        for (int k = 0; k < 6; ++k) f[bur[k]] = S / 18.0;
        f[0] = 12.0 * S / 18.0;
    }
}

/* 3D outlet of the face (a, g).  ORC_OUTLET_PRESSURE is a textbook pressure
 * face (the reference's 3D OUTLET prints "not implemented",
 * S/GpuBoundaries.cu:1671-1706).  The others follow the intended semantics of gpuBcOutlet3D
 * (S/GpuBoundaries.cu:1655-1874): the unknowns become f(x+n) (first order) or
 * 2 f(x+n) - f(x+2n) (second order) of the upstream nodes' populations after
 * their inlet and wall conditions, or f_eq of the node's stored macros.  The
 * reference's population stride there is n*m instead of n*m*h and its
 * second-order north face reads downstream; neither is reproduced.  An
 * upstream node that is solid or outside keeps the node's streamed values. */
static void outlet3d(const orc_sim* s, const double* src, double* dst, long ind, int a, int g)
{
    const long n = s->n;
    const int nx = s->p.nx, ny = s->p.ny, nz = s->p.nz;
    int x = (int)(ind % nx), y = (int)((ind / nx) % ny), z = (int)(ind / ((long)nx * ny));
    if (s->p.outlet_model == ORC_OUTLET_PRESSURE) {
        /* textbook pressure face (Hecht & Harting 2010): rho = rho_out, zero
         * tangential velocity, the normal velocity from the known populations */
        double f[19], K = 0, U[3] = {0, 0, 0};
        for (int i = 0; i < 19; ++i) {
            f[i] = src[(size_t)i * n + ind];
            int ca = comp(i, a);
            if (ca == 0) K += f[i];
            else if (ca == -g) K += 2.0 * f[i];
        }
        U[a] = g * (1.0 - K / s->p.rho_out);
        velocity_face(f, a, g, U, 0);
        for (int i = 0; i < 19; ++i) dst[(size_t)i * n + ind] = f[i];
        return;
    }
    if (s->p.outlet_model == ORC_OUTLET_FEQ) {
        double uvw[3] = {s->u[ind], s->v[ind], s->w3[ind]}, feq[19];
        feq_all(s, s->rho[ind], uvw, feq);
        for (int i = 0; i < 19; ++i)
            if (g * comp(i, a) > 0) dst[(size_t)i * n + ind] = feq[i];
        return;
    }
    int p[3] = {x, y, z}, dims[3] = {nx, ny, nz};
    long up[2];
    for (int k = 1; k <= 2; ++k) {
        int q[3] = {p[0], p[1], p[2]};
        q[a] += g * k;
        if (q[a] < 0 || q[a] >= dims[a]) return;
        up[k - 1] = ((long)q[2] * ny + q[1]) * nx + q[0];
        if (s->cls[up[k - 1]] == ORC_SOLID) return;
    }
    for (int i = 0; i < 19; ++i) {
        if (g * comp(i, a) <= 0) continue;
        double f1 = src[(size_t)i * n + up[0]], f2 = src[(size_t)i * n + up[1]];
        dst[(size_t)i * n + ind] = s->p.outlet_model == ORC_OUTLET_EXTRAP2 ? 2.0 * f1 - f2 : f1;
    }
}

/* symmetry plane (gpuBcSymm3D, S/GpuBoundaries.cu:2086-2167): each unknown
 * takes its mirror image in the face */
static void symmetry3d(double* f, int a, int g)
{
    for (int i = 0; i < 19; ++i) {
        if (g * comp(i, a) <= 0) continue;
        int c[3] = {CX19[i], CY19[i], CZ19[i]};
        c[a] = -c[a];
        f[i] = f[dir3(c[0], c[1], c[2])];
    }
}

void orc_open_boundary(const orc_sim* s, long ind, double* f)
{
    if (s->Q == 9) { zouhe_node(s, ind, f); return; }
    int a, g;
    int k = face_of(s->cls[ind], &a, &g);
    if (k == 1) {
        double U[3] = {s->ubc[ind], s->vbc[ind], s->wbc[ind]};
        velocity_face(f, a, g, U, (s->p.parity & ORC_PARITY_INLET3D_FLIP) && a == 0);
    } else if (k == 3) {
        symmetry3d(f, a, g);
    }
}

void orc_wall_node(const orc_sim* s, long ind, double* f)
{
    uint32_t bb = s->bb[ind];
    if (s->p.wall_model == ORC_WALL_HH) hh_node(bb, f);
    else if (s->p.wall_model == ORC_WALL_ONNODE)
        for (int d = 1; d < s->Q; ++d)
            if ((bb >> d) & 1u) f[s->opp[d]] = f[d];
}

/* Extrapolating east outlet (S/GpuBoundaries.cu:1609-1640): the populations
 * that streamed in from the east face's upstream side, f1, f5, f8, become
 * 2 f(x-1) - f(x-2) (second order) or f(x-1) (first order), from the two
 * upstream nodes' populations after their inlet and wall conditions.  The
 * node keeps its streamed values when an upstream node is solid or missing. */
static void outlet_extrap(orc_sim* s, long ind)
{
    const long n = s->n;
    const int nx = s->p.nx;
    int x = (int)(ind % nx);
    if (x < 2 || s->cls[ind - 1] == ORC_SOLID || s->cls[ind - 2] == ORC_SOLID) return;
    static const int Q3[3] = {1, 5, 8};
    for (int k = 0; k < 3; ++k) {
        size_t q = (size_t)Q3[k] * n;
        double a = s->f[q + ind - 1], b = s->f[q + ind - 2];
        s->f[q + ind] = s->p.outlet_model == ORC_OUTLET_EXTRAP2 ? 2.0 * a - b : a;
    }
}

void orc_step(orc_sim* s, int nsteps)
{
    const long n = s->n;
    const int Q = s->Q;
    const int nx = s->p.nx, ny = s->p.ny;
    set_threads(s->p.nthreads);
    for (int it = 0; it < nsteps; ++it) {
        /* collide (reads stored macros) */
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            if (s->cls[ind] == ORC_SOLID) continue;
            double fin[19], out[19];
            for (int i = 0; i < Q; ++i) fin[i] = s->f[(size_t)i * n + ind];
            double uvw[3] = {s->u[ind], s->v[ind], s->w3[ind]};
            orc_collide_node(s, fin, s->rho[ind], uvw, out);
            int nw = (Q == 19 && s->p.collision == ORC_MRT && (s->p.parity & ORC_PARITY_MRT_NO_F18)) ? 18 : Q;
            for (int i = 0; i < nw; ++i) s->fcoll[(size_t)i * n + ind] = out[i];
        }
        /* stream (pull, S/GpuStream.cu:8-28, :160-194) + walls + open boundaries */
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            if (s->cls[ind] == ORC_SOLID) continue;
            int x = (int)(ind % nx), y = (int)((ind / nx) % ny), z = (int)(ind / ((long)nx * ny));
            uint32_t bb = s->bb[ind];
            double f[19];
            for (int i = 0; i < Q; ++i) {
                int o = s->opp[i];
                if (i > 0 && (bb >> o) & 1u) {
                    /* the source node lies beyond a wall */
                    f[i] = s->p.wall_model == ORC_WALL_HALFWAY ? s->fcoll[(size_t)o * n + ind]
                                                               : s->fcoll[(size_t)i * n + ind];
                    continue;
                }
                int outside;
                long src = wrap_src(s, x, y, z, i, &outside);
                f[i] = outside ? s->fcoll[(size_t)i * n + ind] : s->fcoll[(size_t)i * n + src];
            }
            if (s->p.wall_model == ORC_WALL_CURVED && bb) {
                /* Bouzidi, Firdaouss & Lallemand (2001), linear: for a wall
                 * link d at fraction q the unknown opp(d) is
                 *   q < 1/2:  2q f*_d(x) + (1 - 2q) f*_d(x - c_d)
                 *   q >= 1/2: f*_d(x) / 2q + (2q - 1)/2q f*_opp(d)(x)
                 * (the textbook form of S/GpuBoundaries.cu:414-431, whose
                 * q index and second node are off by one link); q < 1/2 with
                 * no fluid node at x - c_d falls back to bounce-back. */
                double g[19];
                for (int i = 0; i < Q; ++i) g[i] = f[i];
                for (int d = 1; d < Q; ++d) {
                    if (!((bb >> d) & 1u)) continue;
                    const int o = s->opp[d];
                    double q = s->wq ? s->wq[(size_t)ind * 19 + d] : 0.5;
                    if (q <= 0) q = 0.5;
                    const double fd = s->fcoll[(size_t)d * n + ind];
                    if (q < 0.5) {
                        int outside;
                        long src = wrap_src(s, x, y, z, d, &outside);
                        if (((bb >> o) & 1u) || outside) { g[o] = fd; continue; }
                        g[o] = 2.0 * q * fd + (1.0 - 2.0 * q) * s->fcoll[(size_t)d * n + src];
                    } else {
                        g[o] = fd / (2.0 * q) + (2.0 * q - 1.0) / (2.0 * q) * s->fcoll[(size_t)o * n + ind];
                    }
                }
                for (int i = 0; i < Q; ++i) f[i] = g[i];
            }
            /* inlet -> wall -> outlet order of S/Iterate.cu:468-473 and
             * S/Iterate3D.cu:602-621 (then periodic, here by wrap, and symmetry) */
            uint8_t c = s->cls[ind];
            int a, g, kind = Q == 19 ? face_of(c, &a, &g) : 0;
            if (Q == 9 && (c == ORC_VEL_W || c == ORC_VEL_N)) zouhe_node(s, ind, f);
            if (kind == 1) orc_open_boundary(s, ind, f);
            /* on-node bounce-back, S/GpuBoundaries.cu:400-443, :445-490; wet-node walls */
            if (bb) orc_wall_node(s, ind, f);
            if (Q == 9 && c == ORC_PRES_E) zouhe_node(s, ind, f);
            for (int i = 0; i < Q; ++i) s->f[(size_t)i * n + ind] = f[i];
        }
        if (Q == 9 && s->p.outlet_model >= ORC_OUTLET_EXTRAP2) {
            /* east-face nodes only read nodes off the face: no ordering issue */
#pragma omp parallel for
            for (long ind = 0; ind < n; ++ind)
                if (s->cls[ind] == ORC_PRES_E) outlet_extrap(s, ind);
        }
        if (Q == 19) {
            /* outlets read their upstream nodes' pre-outlet state from a snapshot */
            long nout = 0;
            for (long ind = 0; ind < n; ++ind) { int a, g; nout += face_of(s->cls[ind], &a, &g) == 2; }
            if (nout) {
                double* snap = (double*)malloc(sizeof(double) * 19 * (size_t)n);
                memcpy(snap, s->f, sizeof(double) * 19 * (size_t)n);
#pragma omp parallel for
                for (long ind = 0; ind < n; ++ind) {
                    int a, g;
                    if (face_of(s->cls[ind], &a, &g) == 2) outlet3d(s, snap, s->f, ind, a, g);
                }
                free(snap);
            }
#pragma omp parallel for
            for (long ind = 0; ind < n; ++ind) {
                int a, g;
                if (face_of(s->cls[ind], &a, &g) != 3) continue;
                double f[19];
                for (int i = 0; i < 19; ++i) f[i] = s->f[(size_t)i * n + ind];
                symmetry3d(f, a, g);
                for (int i = 0; i < 19; ++i) s->f[(size_t)i * n + ind] = f[i];
            }
        }
        /* macroscopic update */
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) macro_node(s, ind);
    }
}

void orc_get_f(const orc_sim* s, double* f)
{
    memcpy(f, s->f, sizeof(double) * (size_t)s->n * s->Q);
}

void orc_set_f(orc_sim* s, const double* f)
{
    memcpy(s->f, f, sizeof(double) * (size_t)s->n * s->Q);
    for (long ind = 0; ind < s->n; ++ind) macro_node(s, ind);
}

void orc_get_macros(const orc_sim* s, double* rho, double* u, double* v, double* w)
{
    size_t b = sizeof(double) * (size_t)s->n;
    if (rho) memcpy(rho, s->rho, b);
    if (u) memcpy(u, s->u, b);
    if (v) memcpy(v, s->v, b);
    if (w) memcpy(w, s->w3, b);
}

/* The reference's L2 residual: sqrt(sum (f - fColl)^2) over all nodes and
 * populations of the last step (S/ComputeResiduals.cu:13-27; solid nodes hold
 * zeros in both here). */
double orc_residual_l2(const orc_sim* s)
{
    long double acc = 0;
    for (size_t k = 0; k < (size_t)s->n * s->Q; ++k) {
        double d = s->f[k] - s->fcoll[k];
        acc += (long double)d * d;
    }
    return sqrt((double)acc);
}

double orc_mass(const orc_sim* s)
{
    long double acc = 0;
    for (long ind = 0; ind < s->n; ++ind) {
        if (s->cls[ind] == ORC_SOLID) continue;
        for (int i = 0; i < s->Q; ++i) acc += s->f[(size_t)i * s->n + ind];
    }
    return (double)acc;
}

/* ======================================================================== */
/* colour-gradient two-phase model, D2Q9 or D3Q19, periodic box.
 * Follows the GPU formulation: gpuCollBgkwGC2D / gpuCollBgkwGC3D
 * (S/GpuCollision.cu:552-622, :709-779), or with enhanced = 1
 * gpuCollEnhancedBgkwGC2D / 3D (:624-707, :781-880), with the gradients
 * calculateColorGradient / calculateHOColorGradient (2D, :112-281) and
 * calculateColorGradient3D / calculateHOColorGradient3D case 0 (3D, :283-420),
 * streaming gpuStreaming2DCG / gpuStreaming3DCG (S/GpuStream.cu:30-158,
 * :196-285) and moments gpuUpdateMacro2DCG / gpuUpdateMacro3DCG
 * (S/GpuUpdateMacro.cu:43-135, :215-361), with periodic wrap instead of the
 * reference's edge-orientation stencils.                                    */
struct orc_cg {
    orc_cg_params p;
    long n;
    int q, nz;
    int cx[19], cy[19], cz[19];
    double w[19], phi[19], teta[19], bpert[19], cn[19], chi[19], psi[19];
    double *r, *b, *rc, *bc, *f;          /* colour populations, post-collision, total */
    double *rr, *br, *rho, *u, *v, *w3;   /* stored macros */
    int hx[105], hy[105], hz[105];
    double hw[105];
    int nst;
};

/* 2D constants (S/GpuInit.cu:150-186) */
static const double CG_PHI[9] = {0.0, 0.2, 0.2, 0.2, 0.2, 0.05, 0.05, 0.05, 0.05};
static const double CG_TETA[9] = {1.0, -0.2, -0.2, -0.2, -0.2, -0.05, -0.05, -0.05, -0.05};
static const double CG_BPERT[9] = {-4.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
                                   5.0 / 108.0, 5.0 / 108.0, 5.0 / 108.0, 5.0 / 108.0};
static const double W9[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                             1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};

orc_cg* orc_cg_create(const orc_cg_params* p)
{
    orc_cg* s = (orc_cg*)calloc(1, sizeof *s);
    s->p = *p;
    s->nz = p->nz > 1 ? p->nz : 1;
    s->q = s->nz > 1 ? 19 : 9;
    s->n = (long)p->nx * p->ny * s->nz;
    const int q = s->q;
    for (int i = 0; i < q; ++i) {
        if (q == 9) {
            s->cx[i] = CX9[i]; s->cy[i] = CY9[i]; s->cz[i] = 0;
            s->w[i] = W9[i]; s->phi[i] = CG_PHI[i]; s->teta[i] = CG_TETA[i]; s->bpert[i] = CG_BPERT[i];
            /* enhanced-model chi, psi (S/GpuInit.cu:158-165) */
            s->chi[i] = i == 0 ? -8.0 / 3.0 : (i <= 4 ? -1.0 / 6.0 : 1.0 / 12.0);
            s->psi[i] = i == 0 ? 0.0 : (i <= 4 ? 0.5 : 0.125);
        } else {
            /* 3D constants (S/GpuInit.cu:900-935): phi, theta and B by shell */
            s->cx[i] = CX19[i]; s->cy[i] = CY19[i]; s->cz[i] = CZ19[i];
            s->w[i] = i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
            s->phi[i] = i == 0 ? 0.0 : (i <= 6 ? 1.0 / 12.0 : 1.0 / 24.0);
            s->teta[i] = i == 0 ? 1.0 : (i <= 6 ? -1.0 / 12.0 : -1.0 / 24.0);
            s->bpert[i] = i == 0 ? -2.0 / 9.0 : (i <= 6 ? 1.0 / 54.0 : 1.0 / 27.0);
            /* enhanced-model chi3D, psi3D (S/GpuInit.cu:916-926) */
            s->chi[i] = i == 0 ? -2.5 : (i <= 6 ? -1.0 / 6.0 : 1.0 / 24.0);
            s->psi[i] = i == 0 ? 0.0 : (i <= 6 ? 0.25 : 0.125);
        }
        s->cn[i] = sqrt((double)(s->cx[i] * s->cx[i] + s->cy[i] * s->cy[i] + s->cz[i] * s->cz[i]));
    }
    size_t nq = (size_t)s->n * q;
    s->r = (double*)calloc(nq, 8); s->b = (double*)calloc(nq, 8);
    s->rc = (double*)calloc(nq, 8); s->bc = (double*)calloc(nq, 8);
    s->f = (double*)calloc(nq, 8);
    s->rr = (double*)calloc(s->n, 8); s->br = (double*)calloc(s->n, 8);
    s->rho = (double*)calloc(s->n, 8); s->u = (double*)calloc(s->n, 8); s->v = (double*)calloc(s->n, 8);
    s->w3 = (double*)calloc(s->n, 8);
    int k = 0;
    if (p->high_order && q == 9) {
        /* 24-neighbour isotropic stencil; weights by |c|^2 over 5040
         * (S/GpuInit.cu:186-199): 1 -> 960, 2 -> 448, 4 -> 84, 5 -> 32, 8 -> 1 */
        for (int dy = -2; dy <= 2; ++dy)
            for (int dx = -2; dx <= 2; ++dx) {
                int c2 = dx * dx + dy * dy;
                if (c2 == 0) continue;
                double wt = c2 == 1 ? 960 : c2 == 2 ? 448 : c2 == 4 ? 84 : c2 == 5 ? 32 : 1;
                s->hx[k] = dx; s->hy[k] = dy; s->hz[k] = 0; s->hw[k] = wt / 5040.0; ++k;
            }
    } else if (p->high_order) {
        /* 104-neighbour stencil of S/GpuInit.cu:950-990: the points of
         * [-2,2]^3 with |c|^2 in {1,2,3,4,5,6,9} (the (2,2,1) shell, not
         * (2,2,0)), weights 4/45, 1/21, 2/105, 5/504, 1/315, 1/630, 1/5040 */
        for (int dz = -2; dz <= 2; ++dz)
            for (int dy = -2; dy <= 2; ++dy)
                for (int dx = -2; dx <= 2; ++dx) {
                    int c2 = dx * dx + dy * dy + dz * dz;
                    double wt = c2 == 1 ? 4.0 / 45 : c2 == 2 ? 1.0 / 21 : c2 == 3 ? 2.0 / 105
                              : c2 == 4 ? 5.0 / 504 : c2 == 5 ? 1.0 / 315 : c2 == 6 ? 1.0 / 630
                              : c2 == 9 ? 1.0 / 5040 : 0.0;
                    if (wt == 0.0) continue;
                    s->hx[k] = dx; s->hy[k] = dy; s->hz[k] = dz; s->hw[k] = wt; ++k;
                }
    } else {
        for (int i = 1; i < q; ++i) {
            s->hx[i - 1] = s->cx[i]; s->hy[i - 1] = s->cy[i]; s->hz[i - 1] = s->cz[i];
            /* cg_w: 2D 4/12, 1/12 (S/GpuInit.cu:186); 3D 1/6, 1/12 (:928) */
            s->hw[i - 1] = q == 9 ? ((i < 5) ? 4.0 / 12.0 : 1.0 / 12.0) : ((i <= 6) ? 1.0 / 6.0 : 1.0 / 12.0);
        }
        k = q - 1;
    }
    s->nst = k;
    return s;
}

void orc_cg_destroy(orc_cg* s)
{
    if (!s) return;
    free(s->r); free(s->b); free(s->rc); free(s->bc); free(s->f);
    free(s->rr); free(s->br); free(s->rho); free(s->u); free(s->v); free(s->w3);
    free(s);
}

/* macros of the populations in r, b (velocity without the force shift) */
static void cg_moments(orc_cg* s, long ind)
{
    const long n = s->n;
    double R = 0, B = 0, jx = 0, jy = 0, jz = 0;
    for (int k = 0; k < s->q; ++k) {
        double rv = s->r[k * n + ind], bv = s->b[k * n + ind];
        R += rv; B += bv;
        jx += s->cx[k] * (rv + bv); jy += s->cy[k] * (rv + bv); jz += s->cz[k] * (rv + bv);
    }
    s->rr[ind] = R; s->br[ind] = B; s->rho[ind] = R + B;
    s->u[ind] = jx / (R + B); s->v[ind] = jy / (R + B); s->w3[ind] = jz / (R + B);
}

void orc_cg_init(orc_cg* s, const double* r_rho, const double* b_rho)
{
    const long n = s->n;
    for (long ind = 0; ind < n; ++ind) {
        double rr = r_rho[ind], br = b_rho[ind];
        for (int i = 0; i < s->q; ++i) {
            s->r[i * n + ind] = rr * (s->phi[i] + s->teta[i] * s->p.r_alpha);
            s->b[i * n + ind] = br * (s->phi[i] + s->teta[i] * s->p.b_alpha);
            s->f[i * n + ind] = s->r[i * n + ind] + s->b[i * n + ind];
        }
        s->rr[ind] = rr; s->br[ind] = br; s->rho[ind] = rr + br;
        s->u[ind] = 0; s->v[ind] = 0; s->w3[ind] = 0;
        /* the moving north wall's velocity (initInletVelocity, S/Multiphase.cu:1199-1209) */
        if (s->p.walls == 1 && s->q == 9 && ind / s->p.nx == s->p.ny - 1) s->u[ind] = s->p.u_wall;
    }
}

void orc_cg_init_f(orc_cg* s, const double* r_f, const double* b_f)
{
    const long n = s->n;
    for (long j = 0; j < n * s->q; ++j) {
        s->r[j] = r_f[j]; s->b[j] = b_f[j]; s->f[j] = r_f[j] + b_f[j];
    }
    for (long ind = 0; ind < n; ++ind) cg_moments(s, ind);
}

static int cg_wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

/* colour gradient g[0..2] and density gradient g[3..5] (the stencil's
 * second accumulator, S/GpuCollision.cu:291-300) */
static void cg_grad_node(const orc_cg* s, int x, int y, int z, double* g)
{
    const int nx = s->p.nx, ny = s->p.ny, nz = s->nz;
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0;
    if (s->p.walls && s->q == 9 && (y == 0 || y == ny - 1 || (s->p.high_order && (y == 1 || y == ny - 2)))) {
        /* wall rows: the 8-point stencil with the rows beyond the wall
         * mirrored onto the row inside, x components only (calculateColorGradient
         * cases 1 and 2, S/GpuCollision.cu:135-156); next to the wall rows the
         * 24-point stencil falls back to the 8-point one (case 11, :262-275) */
        const int wall = y == 0 || y == ny - 1;
        const int into = y == 0 ? 1 : -1;
        for (int k = 1; k < 9; ++k) {
            const int dy = wall ? into * abs(s->cy[k]) : s->cy[k];
            long j = (long)(y + dy) * nx + cg_wrap(x + s->cx[k], nx);
            double wk = k < 5 ? 4.0 / 12.0 : 1.0 / 12.0;
            double psi = wk * (s->rr[j] - s->br[j]) / s->rho[j];
            double rw = wk * s->rho[j];
            ax += psi * s->cx[k]; bx += rw * s->cx[k];
            if (!wall) { ay += psi * s->cy[k]; by += rw * s->cy[k]; }
        }
        g[0] = ax; g[1] = ay; g[2] = 0.0;
        g[3] = bx; g[4] = by; g[5] = 0.0;
        return;
    }
    if (s->p.walls && s->q == 19 && (y == 0 || y == ny - 1 || (s->p.high_order && (y == 1 || y == ny - 2)))) {
        /* y wall faces: the 18-point stencil with the plane beyond the wall
         * mirrored onto the plane inside, x and z components only
         * (calculateColorGradient3D cases 1 and 2, S/GpuCollision.cu:306-333);
         * next to them the 104-point stencil falls back to the 18-point one
         * (calculateHOColorGradient3D case 11, :505-518) */
        const int wall = y == 0 || y == ny - 1;
        const int into = y == 0 ? 1 : -1;
        for (int k = 1; k < 19; ++k) {
            const int dy = wall ? into * abs(s->cy[k]) : s->cy[k];
            long j = ((long)cg_wrap(z + s->cz[k], nz) * ny + (y + dy)) * nx + cg_wrap(x + s->cx[k], nx);
            double wk = k <= 6 ? 1.0 / 6.0 : 1.0 / 12.0;
            double psi = wk * (s->rr[j] - s->br[j]) / s->rho[j];
            double rw = wk * s->rho[j];
            ax += psi * s->cx[k]; bx += rw * s->cx[k];
            az += psi * s->cz[k]; bz += rw * s->cz[k];
            if (!wall) { ay += psi * s->cy[k]; by += rw * s->cy[k]; }
        }
        g[0] = ax; g[1] = ay; g[2] = az;
        g[3] = bx; g[4] = by; g[5] = bz;
        return;
    }
    for (int k = 0; k < s->nst; ++k) {
        int xx = cg_wrap(x + s->hx[k], nx), yy = cg_wrap(y + s->hy[k], ny), zz = cg_wrap(z + s->hz[k], nz);
        long j = ((long)zz * ny + yy) * nx + xx;
        double psi = s->hw[k] * (s->rr[j] - s->br[j]) / s->rho[j];
        double rw = s->hw[k] * s->rho[j];
        ax += psi * s->hx[k]; ay += psi * s->hy[k]; az += psi * s->hz[k];
        bx += rw * s->hx[k]; by += rw * s->hy[k]; bz += rw * s->hz[k];
    }
    g[0] = ax; g[1] = ay; g[2] = az;
    g[3] = bx; g[4] = by; g[5] = bz;
}

void orc_cg_gradient(const orc_cg* s, double* cgx, double* cgy, double* cgz)
{
    const long nx = s->p.nx, ny = s->p.ny;
    for (long ind = 0; ind < s->n; ++ind) {
        double g[6];
        cg_grad_node(s, (int)(ind % nx), (int)(ind / nx % ny), (int)(ind / (nx * ny)), g);
        cgx[ind] = g[0]; cgy[ind] = g[1];
        if (cgz) cgz[ind] = g[2];
    }
}

/* Zou-He velocity condition of one colour's populations f[9] on the north
 * (unknown f4, f7, f8) or south (f2, f5, f6) wall row with the node's stored
 * velocity (u, v) (S/GpuBoundaries.cu:1888-1899, :1929-1942); returns the
 * colour density it imposes */
static double cg_zouhe(double* f, double u, double v, int north)
{
    if (north) {
        double r = (1.0 / (1.0 + v)) * (f[0] + f[1] + f[3] + 2 * (f[2] + f[6] + f[5]));
        f[4] = f[2] - (2.0 / 3.0) * r * v;
        f[7] = f[5] + 0.5 * (f[1] - f[3]) - (1.0 / 6.0) * (r * v) - 0.5 * (r * u);
        f[8] = f[6] + 0.5 * (f[3] - f[1]) - (1.0 / 6.0) * (r * v) + 0.5 * (r * u);
        return r;
    }
    double r = (1.0 / (1.0 - v)) * (f[0] + f[1] + f[3] + 2 * (f[4] + f[7] + f[8]));
    f[2] = f[4] + (2.0 / 3.0) * r * v;
    f[5] = f[7] - 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * (r * v) + 0.5 * (r * u);
    f[6] = f[8] + 0.5 * (f[1] - f[3]) + (1.0 / 6.0) * (r * v) - 0.5 * (r * u);
    return r;
}

/* on-node bounce-back of the incoming links (S/GpuBoundaries.cu:1908-1915, :1951-1958) */
static void cg_bounce(double* f, int north)
{
    if (north) { f[4] = f[2]; f[7] = f[5]; f[8] = f[6]; }
    else { f[2] = f[4]; f[5] = f[7]; f[6] = f[8]; }
}

/* D3Q19 y wall face of one colour: the reference's 3D inlet condition with
 * the face velocity (gpuBcInlet3D north / south, S/GpuBoundaries.cu:73-399:
 * (u_wall, 0, 0) on the north face, rest on the south one), or on-node
 * bounce-back, each unknown link taking its opposite's streamed value (the
 * STRAIGHT case of gpuBcSimpleWall3D, S/GpuBoundaries.cu:482-484) */
static void cg_wall3(double* f, int walls, int north, double u_wall)
{
    if (walls == 1) {
        const double U[3] = {north ? u_wall : 0.0, 0.0, 0.0};
        velocity_face(f, 1, north ? -1 : 1, U, 0);
        return;
    }
    for (int k = 1; k < 19; ++k)
        if (CY19[k] == (north ? -1 : 1))
            f[k] = f[dir3(-CX19[k], -CY19[k], -CZ19[k])];
}

void orc_cg_step(orc_cg* s, int nsteps)
{
    const long n = s->n;
    const int nx = s->p.nx, ny = s->p.ny, nz = s->nz, q = s->q;
    const orc_cg_params* P = &s->p;
    set_threads(P->nthreads);
    for (int it = 0; it < nsteps; ++it) {
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            int x = (int)(ind % nx), y = (int)(ind / nx % ny), z = (int)(ind / ((long)nx * ny));
            double g[6];
            cg_grad_node(s, x, y, z, g);
            double rr = s->rr[ind], br = s->br[ind], r = s->rho[ind];
            double u = s->u[ind], v = s->v[ind], w = s->w3[ind];
            double mean_nu = 1.0 / (rr / (r * P->r_nu) + br / (r * P->b_nu));
            double om = 1.0 / (3.0 * mean_nu + 0.5);
            double alpha = P->r_alpha * rr / r + P->b_alpha * br / r;
            double gn = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            double kr = rr / r, kb = br / r, kk = P->beta * rr * br / r;
            double uu = u * u + v * v + w * w;
            double ugr = u * g[3] + v * g[4] + w * g[5];
            for (int k = 0; k < q; ++k) {
                double cs = 0, pert = 0;
                if (gn > P->g_limit) {
                    double cg = s->cx[k] * g[0] + s->cy[k] * g[1] + s->cz[k] * g[2];
                    if (k) cs = cg / (gn * s->cn[k]);
                    pert = P->A * gn * (s->w[k] * cg * cg / (gn * gn) - s->bpert[k]);
                }
                double cu = u * s->cx[k] + v * s->cy[k] + w * s->cz[k];
                double rest = s->phi[k] + s->teta[k] * alpha;
                double feq = r * (rest + s->w[k] * (3.0 * cu + 4.5 * cu * cu - 1.5 * uu));
                if (P->enhanced) {
                    /* nu (chi_k u.grad rho + psi_k G:c c), G = u grad rho + grad rho u,
                     * so G:c c = 2 (c.u)(c.grad rho) */
                    double cgr = s->cx[k] * g[3] + s->cy[k] * g[4] + s->cz[k] * g[5];
                    feq += mean_nu * (s->chi[k] * ugr + s->psi[k] * 2.0 * cu * cgr);
                }
                double fs = om * feq + (1.0 - om) * s->f[k * n + ind] + pert;
                s->rc[k * n + ind] = kr * fs + kk * cs * rest;
                s->bc[k * n + ind] = kb * fs - kk * cs * rest;
            }
        }
        const int walls = P->walls;
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            int x = (int)(ind % nx), y = (int)(ind / nx % ny), z = (int)(ind / ((long)nx * ny));
            for (int k = 0; k < q; ++k) {
                /* a wall row keeps the links from beyond the wall (gpuStreaming2DCG,
                 * S/GpuStream.cu:62-85); the wall condition below sets them */
                if (walls && (y - s->cy[k] < 0 || y - s->cy[k] >= ny)) continue;
                int sx = cg_wrap(x - s->cx[k], nx), sy = cg_wrap(y - s->cy[k], ny), sz = cg_wrap(z - s->cz[k], nz);
                long src = ((long)sz * ny + sy) * nx + sx;
                s->r[k * n + ind] = s->rc[k * n + src];
                s->b[k * n + ind] = s->bc[k * n + src];
            }
        }
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            int y = (int)(ind / nx % ny);
            double rr0 = s->rr[ind], br0 = s->br[ind], r0 = s->rho[ind];
            if (walls && q == 19 && (y == 0 || y == ny - 1)) {
                /* per-colour wall faces (S/Iterate3D.cu:576-598); the 3D
                 * condition stores no densities */
                double fr[19], fb[19];
                for (int k = 0; k < 19; ++k) { fr[k] = s->r[k * n + ind]; fb[k] = s->b[k * n + ind]; }
                cg_wall3(fr, walls, y == ny - 1, P->u_wall);
                cg_wall3(fb, walls, y == ny - 1, P->u_wall);
                for (int k = 0; k < 19; ++k) { s->r[k * n + ind] = fr[k]; s->b[k * n + ind] = fb[k]; }
            } else if (walls && (y == 0 || y == ny - 1)) {
                /* per-colour wall condition (gpuBcPeriodic2D, S/GpuBoundaries.cu:1885-1968) */
                double fr[9], fb[9];
                for (int k = 0; k < 9; ++k) { fr[k] = s->r[k * n + ind]; fb[k] = s->b[k * n + ind]; }
                if (walls == 1) {
                    double rt = cg_zouhe(fr, s->u[ind], s->v[ind], y == ny - 1);
                    double bt = cg_zouhe(fb, s->u[ind], s->v[ind], y == ny - 1);
                    /* the condition stores these colour densities, which the
                     * macro update then reads for omega_eff */
                    rr0 = rt; br0 = bt; r0 = rt + bt;
                } else {
                    cg_bounce(fr, y == ny - 1);
                    cg_bounce(fb, y == ny - 1);
                }
                for (int k = 0; k < 9; ++k) { s->r[k * n + ind] = fr[k]; s->b[k * n + ind] = fb[k]; }
            }
            double R = 0, B = 0, jx = 0, jy = 0, jz = 0;
            for (int k = 0; k < q; ++k) {
                double rv = s->r[k * n + ind], bv = s->b[k * n + ind];
                s->f[k * n + ind] = rv + bv;
                R += rv; B += bv;
                jx += s->cx[k] * (rv + bv);
                jy += s->cy[k] * (rv + bv);
                jz += s->cz[k] * (rv + bv);
            }
            /* omega_eff from the previous densities (S/GpuUpdateMacro.cu:62-64, :233-235) */
            double mean_nu = 1.0 / (rr0 / (r0 * P->r_nu) + br0 / (r0 * P->b_nu));
            double om = 1.0 / (3.0 * mean_nu + 0.5);
            double r = R + B;
            s->u[ind] = jx / r + (P->external_force ? P->g / (r * om) : 0.0);
            s->v[ind] = jy / r + (P->external_force ? 0.0 : P->g / om);
            s->w3[ind] = jz / r;
        }
#pragma omp parallel for
        for (long ind = 0; ind < n; ++ind) {
            double R = 0, B = 0;
            for (int k = 0; k < q; ++k) { R += s->r[k * n + ind]; B += s->b[k * n + ind]; }
            s->rr[ind] = R; s->br[ind] = B; s->rho[ind] = R + B;
        }
    }
}

void orc_cg_get(const orc_cg* s, double* r_f, double* b_f, double* r_rho, double* b_rho,
                double* rho, double* u, double* v)
{
    size_t n = (size_t)s->n, bq = n * s->q * 8, b1 = n * 8;
    if (r_f) memcpy(r_f, s->r, bq);
    if (b_f) memcpy(b_f, s->b, bq);
    if (r_rho) memcpy(r_rho, s->rr, b1);
    if (b_rho) memcpy(b_rho, s->br, b1);
    if (rho) memcpy(rho, s->rho, b1);
    if (u) memcpy(u, s->u, b1);
    if (v) memcpy(v, s->v, b1);
}

void orc_cg_get_w(const orc_cg* s, double* w) { memcpy(w, s->w3, (size_t)s->n * 8); }

/* Laplace sampling masks (S/GpuUpdateMacro.cu:124-133, :347-357) and max |u|. */
void orc_cg_laplace(const orc_cg* s, double* out)
{
    double pin = 0, nin = 0, pout = 0, nout = 0, umax = 0;
    for (long ind = 0; ind < s->n; ++ind) {
        double chi = (s->rr[ind] - s->br[ind]) / s->rho[ind];
        if (chi >= s->p.kappa) { pin += s->rr[ind]; nin += 1; }
        else if (chi <= -s->p.kappa) { pout += s->br[ind]; nout += 1; }
        double um = sqrt(s->u[ind] * s->u[ind] + s->v[ind] * s->v[ind] + s->w3[ind] * s->w3[ind]);
        if (um > umax) umax = um;
    }
    out[0] = pin; out[1] = nin; out[2] = pout; out[3] = nout; out[4] = umax;
}

/* ---- the reference's CPU colour-gradient path (D2Q9, the #define CUDA 0
 * branch of S/Iterate.cu:178-208, :396-399, :429-431, :460-462, :486-489):
 * mp2DColl (S/Multiphase.cu:6-138), streamMP (:851-1010),
 * peridicBoundaries (:598-739), updateMacroMP (:451-529).  A second opinion
 * beside the GPU formulation above, not what the GPU kernels follow: the
 * gradient is of rho_r - rho_b (not divided by rho), each colour collides
 * towards its own equilibrium rho_k (phi_k + w (...)) with half the
 * perturbation, and recolouring carries beta rho_r rho_b / rho^2 times
 * (rho_r phi_r + rho_b phi_b).  The box is the reference's full lattice:
 * links from beyond an edge are not streamed; the edges copy them from the
 * opposite edge (period n - 1, m - 1) or, in test case 2, impose Zou-He walls
 * on the first and last rows; corners copy in the reference's sequential
 * order.  Test case 2 updates no macros on the wall rows. */
struct orc_cgcpu {
    orc_cgcpu_params p;
    long n;
    int nz, q;
    double *rf, *bf, *rc, *bc, *rr, *br, *rho, *u, *v, *gx, *gy;
    double *w3, *f, *gz;   /* D3Q19 only: z velocity, the summed populations, z gradient */
    double rphi[9], bphi[9];
};

orc_cgcpu* orc_cgcpu_create(const orc_cgcpu_params* p)
{
    orc_cgcpu* s = (orc_cgcpu*)calloc(1, sizeof *s);
    s->p = *p;
    s->nz = p->nz > 1 ? p->nz : 1;
    s->q = s->nz > 1 ? 19 : 9;
    s->n = (long)p->nx * p->ny * s->nz;
    const size_t n = (size_t)s->n, nq = n * s->q;
    s->rf = calloc(nq, 8); s->bf = calloc(nq, 8); s->rc = calloc(nq, 8); s->bc = calloc(nq, 8);
    s->rr = calloc(n, 8); s->br = calloc(n, 8); s->rho = calloc(n, 8);
    s->u = calloc(n, 8); s->v = calloc(n, 8); s->gx = calloc(n, 8); s->gy = calloc(n, 8);
    if (s->q == 19) { s->w3 = calloc(n, 8); s->f = calloc(nq, 8); s->gz = calloc(n, 8); }
    /* r_phi, b_phi (S/Iterate.cu:194-204) */
    for (int k = 0; k < 9; ++k) {
        s->rphi[k] = k == 0 ? p->r_alpha : (k < 5 ? (1.0 - p->r_alpha) / 5.0 : (1.0 - p->r_alpha) / 20.0);
        s->bphi[k] = k == 0 ? p->b_alpha : (k < 5 ? (1.0 - p->b_alpha) / 5.0 : (1.0 - p->b_alpha) / 20.0);
    }
    return s;
}

void orc_cgcpu_destroy(orc_cgcpu* s)
{
    if (!s) return;
    free(s->rf); free(s->bf); free(s->rc); free(s->bc); free(s->rr); free(s->br); free(s->rho);
    free(s->u); free(s->v); free(s->gx); free(s->gy);
    free(s->w3); free(s->f); free(s->gz);
    free(s);
}

/* createBubble (S/Multiphase.cu:325-413): each colour's rest populations
 * where it has density; initInletVelocity (:1199-1209) in test case 2 */
void orc_cgcpu_init(orc_cgcpu* s, const double* r_rho, const double* b_rho)
{
    const long n = s->n;
    if (s->q == 19) {
        /* createBubble3D (S/Multiphase.cu:415-449): rho_k phi_i for each
         * colour, phi = (0, 1/12 x 6, 1/24 x 12) (S/Iterate3D.cu:254-273), at rest */
        for (long i = 0; i < n; ++i) {
            s->rr[i] = r_rho[i]; s->br[i] = b_rho[i]; s->rho[i] = r_rho[i] + b_rho[i];
            s->u[i] = 0.0; s->v[i] = 0.0; s->w3[i] = 0.0;
            for (int k = 0; k < 19; ++k) {
                const double ph = k == 0 ? 0.0 : (k <= 6 ? 1.0 / 12.0 : 1.0 / 24.0);
                s->rf[k * n + i] = r_rho[i] * ph;
                s->bf[k * n + i] = b_rho[i] * ph;
                s->f[k * n + i] = s->rf[k * n + i] + s->bf[k * n + i];
            }
        }
        return;
    }
    for (long i = 0; i < n; ++i) {
        s->rr[i] = r_rho[i]; s->br[i] = b_rho[i]; s->rho[i] = r_rho[i] + b_rho[i];
        s->u[i] = 0.0; s->v[i] = 0.0;
        for (int k = 0; k < 9; ++k) {
            s->rf[k * n + i] = r_rho[i] * s->rphi[k];
            s->bf[k * n + i] = b_rho[i] * s->bphi[k];
        }
    }
    if (s->p.test_case == 2)
        for (int i = 0; i < s->p.nx; ++i) s->u[(long)(s->p.ny - 1) * s->p.nx + i] = s->p.u_wall;
}

static void cgcpu_collide(orc_cgcpu* s)
{
    const int nx = s->p.nx, ny = s->p.ny;
    const long n = s->n;
    static const double cgw[9] = {0.0, 4.0 / 12.0, 4.0 / 12.0, 4.0 / 12.0, 4.0 / 12.0,
                                  1.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0, 1.0 / 12.0};
    static const double wpert[9] = {-4.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
                                    5.0 / 108.0, 5.0 / 108.0, 5.0 / 108.0, 5.0 / 108.0};
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const long ind = (long)j * nx + i;
            double gx = 0.0, gy = 0.0;
            const int north = j == ny - 1, south = j == 0, east = i == nx - 1, west = i == 0;
            for (int k = 0; k < 9; ++k) {
                /* the 8-point gradient of rho_r - rho_b; an edge mirrors the
                 * rows or columns beyond it and keeps the tangential component;
                 * corners get none */
                long t;
                if (!north && !south && !east && !west) {
                    t = (long)(j + CY9[k]) * nx + i + CX9[k];
                    gx += (s->rr[t] - s->br[t]) * CX9[k] * cgw[k];
                    gy += (s->rr[t] - s->br[t]) * CY9[k] * cgw[k];
                } else if ((north || south) && !east && !west) {
                    t = (long)(j + (north ? -1 : 1) * abs(CY9[k])) * nx + i + CX9[k];
                    gx += (s->rr[t] - s->br[t]) * CX9[k] * cgw[k];
                } else if ((east || west) && !north && !south) {
                    t = (long)(j + CY9[k]) * nx + i + (east ? -1 : 1) * abs(CX9[k]);
                    gy += (s->rr[t] - s->br[t]) * CY9[k] * cgw[k];
                }
            }
            s->gx[ind] = gx; s->gy[ind] = gy;
            const double aux1 = s->rr[ind] / (s->rho[ind] * s->p.r_nu) + s->br[ind] / (s->rho[ind] * s->p.b_nu);
            const double om = 1.0 / (3.0 / aux1 + 0.5);
            const double u = s->u[ind], v = s->v[ind], cu1 = u * u + v * v;
            const double gn = sqrt(gx * gx + gy * gy);
            const double kr = s->rr[ind] / s->rho[ind], kb = s->br[ind] / s->rho[ind];
            const double kk = s->p.beta * s->rr[ind] * s->br[ind] / (s->rho[ind] * s->rho[ind]);
            for (int k = 0; k < 9; ++k) {
                double cosin = 0.0, rp = 0.0, bp = 0.0;
                if (gn > s->p.g_limit) {
                    const double cg = CX9[k] * gx + CY9[k] * gy;
                    if (k) cosin = cg / (gn * sqrt((double)(CX9[k] * CX9[k] + CY9[k] * CY9[k])));
                    const double t = W9[k] * (cg / gn) * (cg / gn) - wpert[k];
                    rp = 0.5 * s->p.A * gn * t;
                    bp = 0.5 * s->p.A * gn * t;
                }
                const double cu2 = u * CX9[k] + v * CY9[k];
                const double sq = W9[k] * (3 * cu2 + 4.5 * cu2 * cu2 - 1.5 * cu1);
                const double rfe = s->rr[ind] * (s->rphi[k] + sq), bfe = s->br[ind] * (s->bphi[k] + sq);
                const long q = (long)k * n + ind;
                const double fs = (om * rfe + (1 - om) * s->rf[q] + rp) + (om * bfe + (1 - om) * s->bf[q] + bp);
                const double rec = kk * cosin * (s->rr[ind] * s->rphi[k] + s->br[ind] * s->bphi[k]);
                s->rc[q] = kr * fs + rec;
                s->bc[q] = kb * fs - rec;
            }
        }
}

static void cgcpu_copy(orc_cgcpu* s, long to, long from, const int* ks, int nk)
{
    for (int a = 0; a < nk; ++a) {
        s->rf[ks[a] * s->n + to] = s->rf[ks[a] * s->n + from];
        s->bf[ks[a] * s->n + to] = s->bf[ks[a] * s->n + from];
    }
}

static void cgcpu3_step(orc_cgcpu* s);

void orc_cgcpu_step(orc_cgcpu* s, int nsteps)
{
    const int nx = s->p.nx, ny = s->p.ny;
    const long n = s->n;
    if (s->q == 19) {
        for (int it = 0; it < nsteps; ++it) cgcpu3_step(s);
        return;
    }
    for (int it = 0; it < nsteps; ++it) {
        cgcpu_collide(s);
        /* streamMP: pull the links whose source lies in the box */
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i)
                for (int k = 0; k < 9; ++k) {
                    const int si = i - CX9[k], sj = j - CY9[k];
                    if (si < 0 || si >= nx || sj < 0 || sj >= ny) continue;
                    const long d = (long)k * n + (long)j * nx + i, src = (long)k * n + (long)sj * nx + si;
                    s->rf[d] = s->rc[src];
                    s->bf[d] = s->bc[src];
                }
        /* peridicBoundaries, in its order */
        for (int i = 1; i < nx - 1; ++i) {
            const long e = (long)(ny - 1) * nx + i, st = i;
            if (s->p.test_case == 2) {
                double* fs[2] = {s->rf, s->bf};
                double dens[2][2];
                for (int c = 0; c < 2; ++c) {
                    double f[9], g[9];
                    for (int k = 0; k < 9; ++k) { f[k] = fs[c][k * n + e]; g[k] = fs[c][k * n + st]; }
                    dens[c][0] = cg_zouhe(f, s->u[e], s->v[e], 1);
                    dens[c][1] = cg_zouhe(g, s->u[st], s->v[st], 0);
                    for (int k = 0; k < 9; ++k) { fs[c][k * n + e] = f[k]; fs[c][k * n + st] = g[k]; }
                }
                s->rr[e] = dens[0][0]; s->br[e] = dens[1][0]; s->rho[e] = dens[0][0] + dens[1][0];
                s->rr[st] = dens[0][1]; s->br[st] = dens[1][1]; s->rho[st] = dens[0][1] + dens[1][1];
            } else {
                static const int up[3] = {4, 7, 8}, dn[3] = {2, 5, 6};
                cgcpu_copy(s, e, st, up, 3);
                cgcpu_copy(s, st, e, dn, 3);
            }
        }
        for (int j = 1; j < ny - 1; ++j) {
            const long e = (long)j * nx + nx - 1, w = (long)j * nx;
            static const int we[3] = {3, 7, 6}, ea[3] = {1, 5, 8};
            cgcpu_copy(s, e, w, we, 3);
            cgcpu_copy(s, w, e, ea, 3);
        }
        {
            const long ne = (long)(ny - 1) * nx + nx - 1, nw = (long)(ny - 1) * nx, se = nx - 1, sw = 0;
            static const int kne[3] = {3, 4, 7}, knw[3] = {1, 4, 8}, kse[3] = {2, 3, 6}, ksw[3] = {2, 1, 5};
            cgcpu_copy(s, ne, nw, kne, 3);
            cgcpu_copy(s, nw, ne, knw, 3);
            cgcpu_copy(s, se, sw, kse, 3);
            cgcpu_copy(s, sw, se, ksw, 3);
        }
        /* updateMacroMP: every row, or the rows between the walls */
        const int j0 = s->p.test_case == 2 ? 1 : 0, j1 = s->p.test_case == 2 ? ny - 1 : ny;
        for (int j = j0; j < j1; ++j)
            for (int i = 0; i < nx; ++i) {
                const long ind = (long)j * nx + i;
                double R = 0, B = 0, jx = 0, jy = 0;
                for (int k = 0; k < 9; ++k) { R += s->rf[k * n + ind]; B += s->bf[k * n + ind]; }
                s->rr[ind] = R; s->br[ind] = B; s->rho[ind] = R + B;
                for (int k = 0; k < 9; ++k) {
                    jx += (s->rf[k * n + ind] + s->bf[k * n + ind]) * CX9[k];
                    jy += (s->rf[k * n + ind] + s->bf[k * n + ind]) * CY9[k];
                }
                s->u[ind] = jx / s->rho[ind];
                s->v[ind] = jy / s->rho[ind];
            }
    }
}

void orc_cgcpu_get(const orc_cgcpu* s, double* r_f, double* b_f, double* r_rho, double* b_rho, double* rho,
                   double* u, double* v)
{
    const size_t n = (size_t)s->n;
    if (r_f) memcpy(r_f, s->rf, s->q * n * 8);
    if (b_f) memcpy(b_f, s->bf, s->q * n * 8);
    if (r_rho) memcpy(r_rho, s->rr, n * 8);
    if (b_rho) memcpy(b_rho, s->br, n * 8);
    if (rho) memcpy(rho, s->rho, n * 8);
    if (u) memcpy(u, s->u, n * 8);
    if (v) memcpy(v, s->v, n * 8);
}

/* the post-collision colours of the last step (r_fPert, b_fPert) and the gradient */
void orc_cgcpu_get_coll(const orc_cgcpu* s, double* r_c, double* b_c, double* gx, double* gy)
{
    const size_t n = (size_t)s->n;
    if (r_c) memcpy(r_c, s->rc, s->q * n * 8);
    if (b_c) memcpy(b_c, s->bc, s->q * n * 8);
    if (gx) memcpy(gx, s->gx, n * 8);
    if (gy) memcpy(gy, s->gy, n * 8);
}

/* ---- the reference's D3Q19 CPU colour-gradient path: mp3DColl
 * (S/Multiphase.cu:140-323), streamMP3D (:1012-1033), peridicBoundaries3D
 * (:741-849), updateMacroMP3D (:531-596), in the order of the #define CUDA 0
 * branch of S/Iterate3D.cu (:500-503, :538-539, :571-572, :633-635; the
 * shipped file defines CUDA 1, :28, so this branch is compiled out).
 *
 * Unlike its D2Q9 sibling it collides the summed populations f towards one
 * equilibrium, as the GPU formulation does, with the gradient of the phase
 * field (rho_r - rho_b) / rho, half the perturbation amplitude (0.5 A), the
 * enhanced terms always on and weighted by the mean alpha, and both colours
 * recoloured with the same sign (:317-318), so blue is pushed up the phase
 * gradient as red is.
 *
 * Only interior nodes are streamed and updated; the faces keep the
 * populations, densities and velocities of the start, and their collision
 * (on those frozen values, with a gradient from the live interior) feeds the
 * interior.  The periodic copies between opposite faces move face
 * populations that nothing reads afterwards; they are kept, with the links
 * the reference copies into the red array where the blue one was meant
 * (:767, :811-812, :845-846).
 *
 * Two points the reference leaves undefined are fixed here:
 * - the gradient weights: the host branch writes 1/6 and 1/12 into cg_w[0]
 *   inside its loops over the axis and diagonal links
 *   (S/Iterate3D.cu:259, :266, :274) and never sets cg_w[1..18]; the oracle takes
 *   1/6 for the axis links and 1/12 for the diagonal ones, the values those
 *   lines write and the GPU path's;
 * - the x and y face branches (:192-235) read the planes k - 1 and k + 1
 *   beyond the box on the front and back planes; the oracle leaves those
 *   terms out.
 * Where the gradient norm is not above g_limit the reference keeps the
 * cosine of the link before (:272-288); the oracle takes 0 there. */
static void cgcpu3_collide(orc_cgcpu* s)
{
    const int n = s->p.nx, m = s->p.ny, h = s->nz;
    const long N = s->n;
    for (int k = 0; k < h; ++k)
        for (int j = 0; j < m; ++j)
            for (int i = 0; i < n; ++i) {
                const long ind = ((long)k * m + j) * n + i;
                /* the reference's branches in its order: interior, north, south,
                 * east, west, front, back; a node on a y face and an x face
                 * matches none and gets no gradient */
                int br = -1;
                if (i && j && k && i != n - 1 && j != m - 1 && k != h - 1) br = 0;
                else if (j == m - 1 && i && i != n - 1) br = 1;
                else if (j == 0 && i && i != n - 1) br = 2;
                else if (i == n - 1 && j && j != m - 1) br = 3;
                else if (i == 0 && j && j != m - 1) br = 4;
                else if (k == h - 1 && j && j != m - 1) br = 5;
                else if (k == 0 && j && j != m - 1) br = 6;
                double cg[3] = {0, 0, 0}, gr[3] = {0, 0, 0};
                for (int d = 1; br >= 0 && d < 19; ++d) {
                    const int cx = CX19[d], cy = CY19[d], cz = CZ19[d];
                    int ti = i + cx, tj = j + cy, tk = k + cz;
                    switch (br) {
                    case 1: tj = j - abs(cy); break;
                    case 2: tj = j + abs(cy); break;
                    case 3: ti = i - abs(cx); break;
                    case 4: ti = i + abs(cx); break;
                    case 5: ti = i + abs(cx); tk = k - abs(cz); break;   /* i + |cx| as written */
                    case 6: ti = i + abs(cx); tk = k + abs(cz); break;
                    default: break;
                    }
                    if (tk < 0 || tk >= h) continue;   /* beyond the box: left out */
                    const long t = ((long)tk * m + tj) * n + ti;
                    const double wk = d <= 6 ? 1.0 / 6.0 : 1.0 / 12.0;
                    const double ph = (s->rr[t] - s->br[t]) / s->rho[t];
                    const int c[3] = {cx, cy, cz};
                    for (int a = 0; a < 3; ++a) {
                        gr[a] += s->rho[t] * c[a] * wk;
                        cg[a] += ph * c[a] * wk;
                    }
                }
                /* the face branches zero the normal components */
                const int normal = br == 1 || br == 2 ? 1 : (br == 3 || br == 4 ? 0 : (br >= 5 ? 2 : -1));
                if (normal >= 0) { cg[normal] = 0.0; gr[normal] = 0.0; }
                s->gx[ind] = cg[0]; s->gy[ind] = cg[1]; s->gz[ind] = cg[2];

                const double u = s->u[ind], v = s->v[ind], w = s->w3[ind];
                const double G[9] = {2.0 * u * gr[0], u * gr[1] + v * gr[0], u * gr[2] + w * gr[0],
                                     u * gr[1] + v * gr[0], 2.0 * v * gr[1], v * gr[2] + w * gr[1],
                                     u * gr[2] + w * gr[0], v * gr[2] + w * gr[1], 2.0 * w * gr[2]};
                const double ugr = u * gr[0] + v * gr[1] + w * gr[2];
                const double cu1 = u * u + v * v + w * w;
                const double gn = sqrt(cg[0] * cg[0] + cg[1] * cg[1] + cg[2] * cg[2]);
                const double rr = s->rr[ind], bb = s->br[ind], rho = s->rho[ind];
                const double kr = rr / rho, kb = bb / rho, kk = s->p.beta * rr * bb / rho;
                const double mean_nu = 1.0 / (rr / (rho * s->p.r_nu) + bb / (rho * s->p.b_nu));
                const double om = 1.0 / (3.0 * mean_nu + 0.5);
                const double alpha = s->p.r_alpha * rr / rho + s->p.b_alpha * bb / rho;
                for (int d = 0; d < 19; ++d) {
                    const int cx = CX19[d], cy = CY19[d], cz = CZ19[d];
                    const double wd = d == 0 ? 1.0 / 3.0 : (d <= 6 ? 1.0 / 18.0 : 1.0 / 36.0);
                    const double wpert = d == 0 ? -2.0 / 9.0 : (d <= 6 ? 1.0 / 54.0 : 1.0 / 27.0);
                    const double phi = d == 0 ? 0.0 : (d <= 6 ? 1.0 / 12.0 : 1.0 / 24.0);
                    const double teta = d == 0 ? 1.0 : (d <= 6 ? -1.0 / 12.0 : -1.0 / 24.0);
                    const double chi = d == 0 ? -2.5 : (d <= 6 ? -1.0 / 6.0 : 1.0 / 24.0);
                    const double psi = d == 0 ? 0.0 : (d <= 6 ? 0.25 : 0.125);
                    double cosin = 0.0, pert = 0.0;
                    if (gn > s->p.g_limit) {
                        const double pcg = cx * cg[0] + cy * cg[1] + cz * cg[2];
                        if (d) cosin = pcg / (gn * sqrt((double)(cx * cx + cy * cy + cz * cz)));
                        pert = 0.5 * s->p.A * gn * (wd * (pcg / gn) * (pcg / gn) - wpert);
                    }
                    const int c[3] = {cx, cy, cz};
                    double tc = 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) tc += G[3 * a + b] * c[a] * c[b];
                    const double cu2 = u * cx + v * cy + w * cz;
                    const double feq = alpha * (chi * ugr + psi * tc) +
                                       rho * (phi + teta * alpha + wd * (3 * cu2 + 4.5 * cu2 * cu2 - 1.5 * cu1));
                    const long q = (long)d * N + ind;
                    const double fs = om * feq + (1 - om) * s->f[q] + pert;
                    const double rec = kk * cosin * (phi + teta * alpha);
                    s->rc[q] = kr * fs + rec;
                    s->bc[q] = kb * fs + rec;
                }
            }
}

/* one face-to-face copy of the listed links of the red and blue arrays;
 * nb of them in the blue one, the rest into the red one (the reference's
 * slips) */
static void cgcpu3_copy(orc_cgcpu* s, long to, long from, const int* ks, int nk, int nb)
{
    for (int a = 0; a < nk; ++a) s->rf[ks[a] * s->n + to] = s->rf[ks[a] * s->n + from];
    for (int a = 0; a < nk; ++a) {
        double* dst = a < nb ? s->bf : s->rf;
        const double* src = a < nb ? s->bf : s->rf;
        dst[ks[a] * s->n + to] = src[ks[a] * s->n + from];
    }
}

static void cgcpu3_step(orc_cgcpu* s)
{
    const int n = s->p.nx, m = s->p.ny, h = s->nz;
    const long N = s->n;
    cgcpu3_collide(s);
    /* streamMP3D: interior nodes pull from x - c (every link streams in a
     * box without solids) */
    for (int k = 1; k < h - 1; ++k)
        for (int j = 1; j < m - 1; ++j)
            for (int i = 1; i < n - 1; ++i) {
                const long ind = ((long)k * m + j) * n + i;
                for (int d = 0; d < 19; ++d) {
                    const long src = ((long)(k - CZ19[d]) * m + (j - CY19[d])) * n + (i - CX19[d]);
                    s->rf[d * N + ind] = s->rc[d * N + src];
                    s->bf[d * N + ind] = s->bc[d * N + src];
                }
            }
    /* peridicBoundaries3D, in its order */
    static const int north[5] = {4, 9, 10, 16, 18}, south[5] = {3, 7, 8, 15, 17};
    static const int east[5] = {2, 8, 10, 12, 14}, west[5] = {1, 7, 9, 11, 13};
    static const int front[5] = {6, 13, 14, 17, 18}, back[5] = {5, 11, 12, 15, 16};
    for (int k = 0; k < h - 1; ++k) {
        for (int i = 0; i < n; ++i) {
            const long e = ((long)k * m + m - 1) * n + i, st = (long)k * m * n + i;
            cgcpu3_copy(s, e, st, north, 5, 4);
            cgcpu3_copy(s, st, e, south, 5, 5);
        }
        for (int j = 1; j < m - 1; ++j) {
            const long e = ((long)k * m + j) * n + n - 1, w = ((long)k * m + j) * n;
            cgcpu3_copy(s, e, w, east, 5, 5);
            cgcpu3_copy(s, w, e, west, 5, 3);
        }
    }
    for (int j = 0; j < m; ++j)
        for (int i = 0; i < n; ++i) {
            const long e = ((long)(h - 1) * m + j) * n + i, st = (long)j * n + i;
            cgcpu3_copy(s, e, st, front, 5, 5);
            cgcpu3_copy(s, st, e, back, 5, 3);
        }
    /* updateMacroMP3D: interior nodes only (the surface-tension estimate it
     * also makes is a diagnostic and is left out) */
    for (int k = 1; k < h - 1; ++k)
        for (int j = 1; j < m - 1; ++j)
            for (int i = 1; i < n - 1; ++i) {
                const long ind = ((long)k * m + j) * n + i;
                double R = 0, B = 0, jx = 0, jy = 0, jz = 0;
                for (int d = 0; d < 19; ++d) {
                    const long q = d * N + ind;
                    R += s->rf[q]; B += s->bf[q];
                    s->f[q] = s->rf[q] + s->bf[q];
                }
                s->rr[ind] = R; s->br[ind] = B; s->rho[ind] = R + B;
                for (int d = 0; d < 19; ++d) {
                    const double fq = s->rf[d * N + ind] + s->bf[d * N + ind];
                    jx += fq * CX19[d]; jy += fq * CY19[d]; jz += fq * CZ19[d];
                }
                s->u[ind] = jx / s->rho[ind];
                s->v[ind] = jy / s->rho[ind];
                s->w3[ind] = jz / s->rho[ind];
            }
}

void orc_cgcpu_init_f(orc_cgcpu* s, const double* r_f, const double* b_f)
{
    const long n = s->n;
    const int q = s->q;
    for (long ind = 0; ind < n; ++ind) {
        double R = 0, B = 0, jx = 0, jy = 0, jz = 0;
        for (int d = 0; d < q; ++d) {
            const long j = (long)d * n + ind;
            s->rf[j] = r_f[j]; s->bf[j] = b_f[j];
            if (s->f) s->f[j] = r_f[j] + b_f[j];
            R += r_f[j]; B += b_f[j];
            const double fq = r_f[j] + b_f[j];
            jx += fq * (q == 19 ? CX19[d] : CX9[d]);
            jy += fq * (q == 19 ? CY19[d] : CY9[d]);
            jz += q == 19 ? fq * CZ19[d] : 0.0;
        }
        s->rr[ind] = R; s->br[ind] = B; s->rho[ind] = R + B;
        s->u[ind] = jx / (R + B); s->v[ind] = jy / (R + B);
        if (s->w3) s->w3[ind] = jz / (R + B);
    }
}

void orc_cgcpu_get_w(const orc_cgcpu* s, double* w, double* gz)
{
    const size_t n = (size_t)s->n;
    if (w && s->w3) memcpy(w, s->w3, n * 8);
    if (gz && s->gz) memcpy(gz, s->gz, n * 8);
}
