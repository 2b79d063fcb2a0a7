// C ABI of the colour-gradient step, D2Q9 and D3Q19 (declared in include/lbm.h).
#include "lbm.h"
#include "lbm_cg.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {
constexpr int kRedBlocks = 256;
}  // namespace

struct lbm_cg_ctx {
    lbm_cg_params p{};
    lbm::cg::CgParams cp{};
    size_t n = 0;
    int q = 9;
    bool three_d = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* P[2] = {nullptr, nullptr};
    int cur = 0;
    int64_t steps = -1;
    double* mac = nullptr;       // 6 N: rho_r, rho_b, rho, u, v, w
    double* rr0 = nullptr;       // 2 N initial densities
    void* psi = nullptr;         // 2 N phase field and density (D3Q19)
    double* fout = nullptr;      // 2 q N: populations in and out
    double* partials = nullptr;
    bool walls = false;          // walls or body force: the reference-order path (lbm_cgw.hip)
    lbm::cg::WallParams wp{};
    std::string err;
};

namespace {
lbm_status cfail(lbm_cg_ctx* c, lbm_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return s;
}
#define CGCHK(c, expr)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return cfail((c), LBM_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)
size_t sbytes(int s) { return s == LBM_STORE_F64 ? 8 : s == LBM_STORE_F32 ? 4 : 2; }
}  // namespace

extern "C" {

lbm_status lbm_cg_create(const lbm_cg_params* p, lbm_cg_ctx** out) {
    if (!out || !p) return LBM_EINVAL;
    *out = nullptr;
    if (p->abi_version != LBM_ABI_VERSION || p->nx < 3 || p->ny < 3) return LBM_EINVAL;
    if (p->nz < 0 || p->nz == 2) return LBM_EINVAL;
    if (p->storage < LBM_STORE_F64 || p->storage > LBM_STORE_F16) return LBM_EUNSUPPORTED;
    if (!(p->r_nu > 0 && p->b_nu > 0)) return LBM_EINVAL;
    if (p->enhanced != 0 && p->enhanced != 1) return LBM_EINVAL;
    const bool three_d = p->nz >= 3;
    if (p->walls < 0 || p->walls > 2 || (p->external_force != 0 && p->external_force != 1) || !std::isfinite(p->g) ||
        !std::isfinite(p->u_wall))
        return LBM_EINVAL;
    const bool walls = p->walls != 0 || p->g != 0.0;
    if (walls && p->storage == LBM_STORE_F16) return LBM_EUNSUPPORTED;   // F64 / F32 only
    if (p->walls && p->ny < 5) return LBM_EINVAL;
    const int nz = three_d ? p->nz : 1, q = three_d ? 19 : 9;
    const size_t n = (size_t)p->nx * p->ny * nz;
    const size_t plane = (n + 63) / 64 * 64;
    if (2 * (size_t)q * plane >= (1ull << 32)) return LBM_EUNSUPPORTED;   // 32-bit element offsets
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return LBM_EDEVICE;
    lbm_cg_ctx* c = new lbm_cg_ctx();
    c->p = *p;
    c->n = n;
    c->q = q;
    c->three_d = three_d;
    lbm::cg::CgParams& k = c->cp;
    k.nx = p->nx; k.ny = p->ny; k.nz = nz;
    k.r_alpha = p->r_alpha; k.b_alpha = p->b_alpha; k.r_nu = p->r_nu; k.b_nu = p->b_nu;
    k.beta = p->beta; k.A = p->A; k.g_limit = p->g_limit; k.kappa = p->kappa;
    k.plane = (uint32_t)plane;
    k.enhanced = p->enhanced;
    // D3Q19 kernels: the 8 XCDs each take 8 consecutive row blocks per turn
    // (256^3 droplet: -9 % per step FP32, -8 % FP16 against the round-robin
    // order; chunks of 4-64 within 2 %)
    k.xcd_shift = 3;
    k.xchunk_shift = 3;
    lbm::log2_env("LBM_XCD", k.xcd_shift);
    lbm::log2_env("LBM_XCHUNK", k.xchunk_shift);
    // D3Q19 with the 18-point gradient on boxes of whole 64 x 8 row tiles:
    // the single-pass step over z-columns of 32 planes (64 with FP16 storage:
    // 0.756 vs 0.767 ms at 256^3; FP32 1.162 at 32, 1.220 at 64);
    // LBM_CG3_ZCHUNK sets the column height (0 selects the two-launch step),
    // LBM_CG3_ROWS the tile's rows (8 or 4)
    k.zchunk = 0;
    k.zrows = 8;
    if (const char* e = std::getenv("LBM_CG3_ROWS"))
        if (std::strtol(e, nullptr, 10) == 4) k.zrows = 4;
    if (three_d && !p->high_order && p->nx % 64 == 0 && p->ny % k.zrows == 0) {
        k.zchunk = p->storage == LBM_STORE_F16 ? 64 : 32;
        if (const char* e = std::getenv("LBM_CG3_ZCHUNK")) {
            const long v = std::strtol(e, nullptr, 10);
            if (v >= 0 && v <= 4096) k.zchunk = (int)v;
        }
    }
    c->walls = walls;
    c->wp.cp = k;
    c->wp.walls = p->walls;
    c->wp.ho = p->high_order;
    c->wp.u_wall = p->u_wall;
    c->wp.g = p->g;
    c->wp.external_force = p->external_force;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess;
    const size_t pb = 2 * (size_t)q * plane * sbytes(p->storage);
    ok = ok && hipMalloc(&c->P[0], pb) == hipSuccess && hipMalloc(&c->P[1], pb) == hipSuccess &&
         hipMalloc((void**)&c->mac, 6 * n * 8) == hipSuccess && hipMalloc((void**)&c->rr0, 2 * n * 8) == hipSuccess &&
         hipMalloc((void**)&c->partials, (size_t)kRedBlocks * lbm::cg::NRED * 8) == hipSuccess &&
         (!three_d || hipMalloc(&c->psi, 2 * n * 8) == hipSuccess) &&
         hipMemset(c->mac + 5 * n, 0, n * 8) == hipSuccess;
    if (!ok) {
        lbm_cg_destroy(c);
        return LBM_ENOMEM;
    }
    *out = c;
    return LBM_OK;
}

void lbm_cg_destroy(lbm_cg_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    void* b[] = {c->P[0], c->P[1], c->mac, c->rr0, c->psi, c->fout, c->partials};
    for (void* x : b)
        if (x) hipFree(x);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char* lbm_cg_last_error(const lbm_cg_ctx* c) { return c ? c->err.c_str() : "null context"; }

static hipError_t launch_init(lbm_cg_ctx* c, const double* rf, const double* bf) {
    const double *rr = c->rr0, *rb = c->rr0 + c->n;
    if (c->walls) return lbm::cgw_launch_init(c->p.storage, c->P[0], c->mac, rr, rb, rf, bf, c->wp, c->stream);
    if (c->three_d)
        return lbm::cg3_launch_init(c->p.storage, c->p.high_order, c->P[0], c->psi, c->cp, rr, rb, rf, bf, c->mac,
                                    c->stream);
    return lbm::cg_launch_init(c->p.storage, c->p.high_order, c->P[0], c->cp, rr, rb, rf, bf, c->mac, c->stream);
}

lbm_status lbm_cg_init(lbm_cg_ctx* c, const double* r_rho, const double* b_rho) {
    if (!c || !r_rho || !b_rho) return LBM_EINVAL;
    const size_t n = c->n;
    for (size_t i = 0; i < n; ++i)
        if (!(r_rho[i] >= 0 && b_rho[i] >= 0 && r_rho[i] + b_rho[i] > 0))
            return cfail(c, LBM_EINVAL, "colour densities must be >= 0 with a positive sum (node %zu)", i);
    CGCHK(c, hipMemcpyAsync(c->rr0, r_rho, 8 * n, hipMemcpyHostToDevice, c->stream));
    CGCHK(c, hipMemcpyAsync(c->rr0 + n, b_rho, 8 * n, hipMemcpyHostToDevice, c->stream));
    CGCHK(c, launch_init(c, nullptr, nullptr));
    CGCHK(c, hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->steps = 0;
    return LBM_OK;
}

lbm_status lbm_cg_init_f(lbm_cg_ctx* c, const double* r_f, const double* b_f) {
    if (!c || !r_f || !b_f) return LBM_EINVAL;
    const size_t n = c->n, q = (size_t)c->q;
    for (size_t i = 0; i < n; ++i) {
        double a = 0, b = 0;
        bool fin = true;
        for (size_t k = 0; k < q; ++k) {
            a += r_f[k * n + i];
            b += b_f[k * n + i];
            fin = fin && std::isfinite(r_f[k * n + i]) && std::isfinite(b_f[k * n + i]);
        }
        if (!fin || !(a + b > 0))
            return cfail(c, LBM_EINVAL, "colour populations must be finite with a positive density (node %zu)", i);
    }
    if (!c->fout && hipMalloc((void**)&c->fout, 2 * q * n * 8) != hipSuccess) return LBM_ENOMEM;
    CGCHK(c, hipMemcpyAsync(c->fout, r_f, 8 * q * n, hipMemcpyHostToDevice, c->stream));
    CGCHK(c, hipMemcpyAsync(c->fout + q * n, b_f, 8 * q * n, hipMemcpyHostToDevice, c->stream));
    CGCHK(c, lbm::cg_launch_dens(c->q, c->fout, c->fout + q * n, n, c->rr0, c->rr0 + n, c->stream));
    CGCHK(c, launch_init(c, c->fout, c->fout + q * n));
    CGCHK(c, hipStreamSynchronize(c->stream));
    c->cur = 0;
    c->steps = 0;
    return LBM_OK;
}

static lbm_status cg_run(lbm_cg_ctx* c, int64_t nsteps) {
    for (int64_t k = 0; k < nsteps; ++k) {
        const bool wm = k == nsteps - 1;
        if (c->walls) {
            // P[0]: populations after the wall condition, P[1]: after collision
            CGCHK(c, lbm::cgw_launch_step(c->p.storage, c->P[0], c->P[1], c->mac, c->wp, c->stream));
            continue;
        }
        if (c->three_d)
            CGCHK(c, lbm::cg3_launch_step(c->p.storage, c->p.high_order, c->P[c->cur], c->P[c->cur ^ 1], c->psi,
                                          c->cp, c->mac, wm, c->stream));
        else
            CGCHK(c, lbm::cg_launch_step(c->p.storage, c->p.high_order, c->P[c->cur], c->P[c->cur ^ 1], c->cp,
                                         c->mac, wm, c->stream));
        c->cur ^= 1;
    }
    c->steps += nsteps;
    return LBM_OK;
}

lbm_status lbm_cg_step(lbm_cg_ctx* c, int64_t nsteps) {
    if (!c || nsteps < 0) return LBM_EINVAL;
    if (c->steps < 0) return cfail(c, LBM_EINVAL, "state not initialised");
    lbm_status s = cg_run(c, nsteps);
    if (s != LBM_OK) return s;
    CGCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_cg_get_macros(lbm_cg_ctx* c, double* rr, double* rb, double* rho, double* u, double* v) {
    return lbm_cg_get_macros3d(c, rr, rb, rho, u, v, nullptr);
}

lbm_status lbm_cg_get_macros3d(lbm_cg_ctx* c, double* rr, double* rb, double* rho, double* u, double* v, double* w) {
    if (!c) return LBM_EINVAL;
    if (c->steps < 0) return cfail(c, LBM_EINVAL, "state not initialised");
    double* d[6] = {rr, rb, rho, u, v, w};
    for (int a = 0; a < 6; ++a)
        if (d[a]) CGCHK(c, hipMemcpyAsync(d[a], c->mac + a * c->n, 8 * c->n, hipMemcpyDeviceToHost, c->stream));
    CGCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_cg_get_f(lbm_cg_ctx* c, double* rf, double* bf) {
    if (!c || !rf || !bf) return LBM_EINVAL;
    if (c->steps < 1) return cfail(c, LBM_EINVAL, "post-streaming populations exist after the first step");
    const size_t qn = (size_t)c->q * c->n;
    if (!c->fout && hipMalloc((void**)&c->fout, 2 * qn * 8) != hipSuccess) return LBM_ENOMEM;
    if (c->walls)
        CGCHK(c, lbm::cgw_launch_out(c->p.storage, c->q, c->P[0], c->n, c->fout, c->fout + qn, c->stream));
    else if (c->three_d)
        CGCHK(c, lbm::cg3_launch_recon(c->p.storage, c->P[c->cur ^ 1], c->cp, c->fout, c->fout + qn, c->stream));
    else
        CGCHK(c, lbm::cg_launch_recon(c->p.storage, c->P[c->cur ^ 1], c->cp, c->fout, c->fout + qn, c->stream));
    CGCHK(c, hipMemcpyAsync(rf, c->fout, 8 * qn, hipMemcpyDeviceToHost, c->stream));
    CGCHK(c, hipMemcpyAsync(bf, c->fout + qn, 8 * qn, hipMemcpyDeviceToHost, c->stream));
    CGCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_cg_reduce(lbm_cg_ctx* c, double out[12]) {
    if (!c || !out) return LBM_EINVAL;
    if (c->steps < 0) return cfail(c, LBM_EINVAL, "state not initialised");
    const int nr = lbm::cg::NRED;
    CGCHK(c, lbm::cg_launch_reduce(c->mac, c->n, c->three_d, c->p.kappa, c->partials, kRedBlocks, c->stream));
    std::vector<double> h((size_t)kRedBlocks * nr);
    CGCHK(c, hipMemcpyAsync(h.data(), c->partials, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    CGCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < nr; ++k) {
        double s = 0.0;
        for (int b = 0; b < kRedBlocks; ++b) {
            const double v = h[(size_t)b * nr + k];
            s = k == 10 ? (v > s ? v : s) : s + v;
        }
        out[k] = s;
    }
    return LBM_OK;
}

lbm_status lbm_cg_time_steps(lbm_cg_ctx* c, int64_t nsteps, double* ms) {
    if (!c || !ms || nsteps < 1) return LBM_EINVAL;
    if (c->steps < 0) return cfail(c, LBM_EINVAL, "state not initialised");
    CGCHK(c, hipEventRecord(c->ev0, c->stream));
    lbm_status s = cg_run(c, nsteps);
    if (s != LBM_OK) return s;
    CGCHK(c, hipEventRecord(c->ev1, c->stream));
    CGCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    CGCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
    return LBM_OK;
}

}  // extern "C"
