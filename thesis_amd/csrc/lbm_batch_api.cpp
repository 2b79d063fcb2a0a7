// C ABI of the batched channels (declared in include/lbm.h).
//
// Each instance keeps its nodes densely in row-major order: its populations
// in one block of 64-node tiles (the block starts on a tile), its node words
// and FP64 fields in one concatenated per-node index space, and its
// 1024-node workgroup tiles numbered after those of the previous instances.  The
// reference runs one lattice per process through Iterate2D
// (S/Iterate.cu:23-625).
#include "lbm.h"
#include "lbm_batch_internal.h"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace lbm;

struct lbm_batch {
    int n = 0;
    int collision = 0, storage = 0;
    size_t raw_bytes = 0;
    std::vector<BInst> hinst;
    std::vector<uint8_t> ready;     // instance initialised
    std::vector<int64_t> steps;     // steps since its initialisation
    std::vector<int> par;           // host copy of the buffer holding each instance's newest state
    size_t ndense = 0, nslots = 0, nfluid = 0, maxn = 0;
    int ntiles = 0, grid = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<void*> allocs;
    BInst* inst = nullptr;
    void* P[2] = {nullptr, nullptr};
    uint32_t* nodew = nullptr;
    double *rho = nullptr, *u = nullptr, *v = nullptr, *up = nullptr, *vp = nullptr;
    double *ubc = nullptr, *vbc = nullptr, *fout = nullptr;
    BState bs{};
    std::string err;
};

namespace {

lbm_status bfail(lbm_batch* b, lbm_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->err = buf;
    return s;
}

#define BCHK(b, expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return bfail((b), LBM_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_));         \
    } while (0)

template <class T> lbm_status balloc(lbm_batch* b, T** p, size_t count) {
    void* q = nullptr;
    const size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 256;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return bfail(b, LBM_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    b->allocs.push_back(q);
    *p = (T*)q;
    return LBM_OK;
}

Fields bfields(lbm_batch* b) {
    Fields f{};
    f.rho = b->rho; f.u = b->u; f.v = b->v; f.w = nullptr;
    f.ubc = b->ubc; f.vbc = b->vbc; f.wbc = nullptr;
    f.fout = b->fout;
    return f;
}

// D2Q9 velocities in the library's direction order.
const int kC9[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};

void magic(uint32_t d, uint64_t& m, uint32_t& s) {
    s = 0;
    while ((1ull << s) < d) ++s;
    m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
}

thread_local std::string g_create_err;

}  // namespace

extern "C" {

lbm_status lbm_batch_create(const lbm_batch_instance* in, int32_t n, int32_t collision, int32_t storage,
                            lbm_batch** out) {
    auto reject = [](const char* fmt, ...) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        g_create_err = buf;
        return LBM_EINVAL;
    };
    if (!out) return LBM_EINVAL;
    *out = nullptr;
    if (!in || n < 1) return reject("need at least one instance");
    if (collision != LBM_BGK && collision != LBM_TRT) return reject("batch collision must be BGK or TRT");
    if (storage != LBM_STORE_F64 && storage != LBM_STORE_F32) return reject("batch storage must be F64 or F32");
    if (storage == LBM_STORE_F32) {
        // switched off: its first cooperative launch (7 workgroups per CU) did
        // not finish on MI355X while FP64 (6 per CU) runs; cause not found
        g_create_err = "F32 batch storage is switched off (its persistent launch did not complete)";
        return LBM_EUNSUPPORTED;
    }

    if (n > kBatchMaxInstances) return reject("at most %d instances per batch", kBatchMaxInstances);

    std::vector<BInst> hinst(n);
    std::vector<uint32_t> nodew;
    std::vector<double> ubc, vbc;
    size_t dbase = 0, pbase = 0, tbase = 0, maxn = 0, nfluid = 0;
    for (int k = 0; k < n; ++k) {
        const lbm_batch_instance& I = in[k];
        if (I.nx < 3 || I.ny < 1 || !I.node_class) return reject("instance %d: need nx >= 3, ny >= 1, node_class", k);
        if (!(I.omega > 0.0 && I.omega < 2.0)) return reject("instance %d: omega must lie in (0, 2)", k);
        const size_t nn = (size_t)I.nx * I.ny;
        if (nn * 9 * 8 >= (1ull << 32)) return reject("instance %d too large for 32-bit offsets", k);
        std::vector<uint32_t> word(nn);
        size_t nf = 0;
        bool need_u = false;
        for (size_t j = 0; j < nn; ++j) {
            const uint8_t c = I.node_class[j];
            const int x = (int)(j % I.nx);
            if (c > LBM_PRES_E) return reject("instance %d: node class %d out of range", k, (int)c);
            if (c == LBM_VEL_W && x != 0) return reject("instance %d: LBM_VEL_W off the west face", k);
            if (c == LBM_PRES_E && x != I.nx - 1) return reject("instance %d: LBM_PRES_E off the east face", k);
            need_u |= c == LBM_VEL_W;
            nf += c != LBM_SOLID;
            word[j] = ((uint32_t)c << kClsShift) | (I.bb_links ? (I.bb_links[j] & 0x1FEu) : 0u);
        }
        if (nf == 0) return reject("instance %d has no fluid nodes", k);
        if (need_u && !I.u_bc) return reject("instance %d: velocity inlet nodes need u_bc", k);
        // a link into a solid node, or out through the y faces, is a wall
        for (size_t j = 0; j < nn; ++j) {
            if (I.node_class[j] == LBM_SOLID) {
                word[j] = 0;
                continue;
            }
            const int x = (int)(j % I.nx), y = (int)(j / I.nx);
            for (int d = 1; d < 9; ++d) {
                const int xs = x + kC9[d][0], ys = y + kC9[d][1];
                if (ys < 0 || ys >= I.ny) {
                    word[j] |= 1u << d;
                    continue;
                }
                if (xs < 0 || xs >= I.nx) continue;
                if (I.node_class[(size_t)ys * I.nx + xs] == LBM_SOLID) word[j] |= 1u << d;
            }
        }
        BInst& B = hinst[k];
        B.nx = I.nx; B.ny = I.ny;
        B.n = (uint32_t)nn;
        B.pbase = (uint32_t)pbase;
        B.dbase = (uint32_t)dbase;
        B.tbase = (uint32_t)tbase;
        B.pad = 0;
        magic((uint32_t)I.nx, B.mx, B.sx);
        B.omega = I.omega;
        const double lam = I.trt_magic > 0 ? I.trt_magic : 3.0 / 16.0;
        B.omega_m = collision == LBM_TRT ? 1.0 / (0.5 + lam / (1.0 / I.omega - 0.5)) : I.omega;
        B.rho_out = I.rho_out > 0 ? I.rho_out : 1.0;
        nodew.insert(nodew.end(), word.begin(), word.end());
        for (size_t j = 0; j < nn; ++j) {
            ubc.push_back(I.u_bc ? I.u_bc[j] : 0.0);
            vbc.push_back(I.v_bc ? I.v_bc[j] : 0.0);
        }
        dbase += nn;
        pbase += (nn + 63) / 64 * 64;
        tbase += (nn + kBatchTile - 1) / kBatchTile;
        nfluid += nf;
        maxn = nn > maxn ? nn : maxn;
        if (dbase >= 0xFFFFFFFFull || pbase >= 0xFFFFFFFFull || tbase >= (size_t)INT_MAX)
            return reject("batch too large");
    }
    const size_t raw = storage == LBM_STORE_F64 ? 8 : 4;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return reject("no HIP device");
    int coop = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop)
        return reject("device does not support cooperative launches");

    lbm_batch* b = new lbm_batch();
    b->n = n;
    b->collision = collision;
    b->storage = storage;
    b->raw_bytes = raw;
    b->hinst = hinst;
    b->ready.assign(n, 0);
    b->steps.assign(n, 0);
    b->par.assign(n, 0);
    b->ndense = dbase;
    b->nslots = pbase;
    b->nfluid = nfluid;
    b->maxn = maxn;
    b->ntiles = (int)tbase;
    auto bad = [&](lbm_status s) {
        g_create_err = b->err;
        lbm_batch_destroy(b);
        return s;
    };
    const int maxb = batch_max_blocks(storage == LBM_STORE_F64 ? 0 : 1, collision, n);
    if (maxb <= 0) {
        b->err = "occupancy query for the batch kernel failed";
        return bad(LBM_EDEVICE);
    }
    // grid.sync() needs every workgroup resident at once: leave one
    // workgroup per XCD of headroom below the occupancy query's co-residency,
    // so that a CU the runtime keeps back cannot leave the grid waiting
    const int head = maxb >= 64 ? 8 : (maxb > 1 ? 1 : 0);
    b->grid = b->ntiles < maxb - head ? b->ntiles : maxb - head;
    if (hipStreamCreate(&b->stream) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess) {
        b->err = "stream or event creation failed";
        return bad(LBM_EDEVICE);
    }
    lbm_status s;
    const size_t D = b->ndense;
    if ((s = balloc(b, &b->inst, (size_t)n)) != LBM_OK ||
        (s = balloc(b, (char**)&b->P[0], pbase * 9 * raw)) != LBM_OK ||
        (s = balloc(b, (char**)&b->P[1], pbase * 9 * raw)) != LBM_OK || (s = balloc(b, &b->nodew, D)) != LBM_OK ||
        (s = balloc(b, &b->rho, D)) != LBM_OK || (s = balloc(b, &b->u, D)) != LBM_OK ||
        (s = balloc(b, &b->v, D)) != LBM_OK || (s = balloc(b, &b->up, D)) != LBM_OK ||
        (s = balloc(b, &b->vp, D)) != LBM_OK || (s = balloc(b, &b->ubc, D)) != LBM_OK ||
        (s = balloc(b, &b->vbc, D)) != LBM_OK || (s = balloc(b, &b->fout, 9 * maxn)) != LBM_OK ||
        (s = balloc(b, &b->bs.acc, 2 * (size_t)n)) != LBM_OK || (s = balloc(b, &b->bs.active, (size_t)n)) != LBM_OK ||
        (s = balloc(b, &b->bs.done, (size_t)n)) != LBM_OK || (s = balloc(b, &b->bs.par, (size_t)n)) != LBM_OK ||
        (s = balloc(b, &b->bs.res, (size_t)n)) != LBM_OK || (s = balloc(b, &b->bs.remaining, 1)) != LBM_OK)
        return bad(s);
    const hipMemcpyKind h2d = hipMemcpyHostToDevice;
    if (hipMemcpy(b->inst, hinst.data(), sizeof(BInst) * n, h2d) != hipSuccess ||
        hipMemcpy(b->nodew, nodew.data(), 4 * D, h2d) != hipSuccess ||
        hipMemcpy(b->ubc, ubc.data(), 8 * D, h2d) != hipSuccess ||
        hipMemcpy(b->vbc, vbc.data(), 8 * D, h2d) != hipSuccess) {
        b->err = "host-to-device copy failed";
        return bad(LBM_EDEVICE);
    }
    if ((s = lbm_batch_init(b, -1, nullptr, nullptr, nullptr)) != LBM_OK) return bad(s);
    *out = b;
    return LBM_OK;
}

void lbm_batch_destroy(lbm_batch* b) {
    if (!b) return;
    if (b->stream) hipStreamSynchronize(b->stream);
    for (void* p : b->allocs) hipFree(p);
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

const char* lbm_batch_last_error(const lbm_batch* b) { return b ? b->err.c_str() : g_create_err.c_str(); }

lbm_status lbm_batch_init(lbm_batch* b, int32_t i, const double* rho, const double* u, const double* v) {
    if (!b) return LBM_EINVAL;
    if (i < -1 || i >= b->n) return bfail(b, LBM_EINVAL, "instance %d out of range", i);
    if (i == -1 && (rho || u || v)) return bfail(b, LBM_EINVAL, "fields are per instance (i >= 0)");
    const int lo = i < 0 ? 0 : i, hi = i < 0 ? b->n : i + 1;
    for (int k = lo; k < hi; ++k) {
        const BInst& B = b->hinst[k];
        const size_t nn = (size_t)B.nx * B.ny;
        std::vector<double> one(rho ? 0 : nn, 1.0), zero((u && v) ? 0 : nn, 0.0);
        const double* r = rho ? rho : one.data();
        const double* uu = u ? u : zero.data();
        const double* vv = v ? v : zero.data();
        const size_t o = B.dbase;
        BCHK(b, hipMemcpyAsync(b->rho + o, r, 8 * nn, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->u + o, uu, 8 * nn, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->v + o, vv, 8 * nn, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->up + o, uu, 8 * nn, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipMemcpyAsync(b->vp + o, vv, 8 * nn, hipMemcpyHostToDevice, b->stream));
        BCHK(b, hipStreamSynchronize(b->stream));   // the host vectors die with this iteration
        b->par[k] = 0;
        b->ready[k] = 1;
        b->steps[k] = 0;
    }
    BCHK(b, batch_init(b->storage == LBM_STORE_F64 ? 0 : 1, b->collision, b->inst, b->n, b->ntiles, b->P[0],
                       b->nodew, bfields(b), i, b->stream));
    BCHK(b, hipMemcpyAsync(b->bs.par, b->par.data(), sizeof(int) * b->n, hipMemcpyHostToDevice, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return LBM_OK;
}

lbm_status lbm_batch_run(lbm_batch* b, double tol, int32_t check_every, int64_t max_steps, int64_t* steps,
                         double* residual, double* ms) {
    if (!b) return LBM_EINVAL;
    if (check_every < 1 || max_steps < 0 || max_steps > INT_MAX)
        return bfail(b, LBM_EINVAL, "bad check interval or step limit");
    for (int k = 0; k < b->n; ++k)
        if (!b->ready[k]) return bfail(b, LBM_EINVAL, "instance %d not initialised", k);
    const int n = b->n;
    std::vector<int> ones(n, 1), zeros(n, 0);
    std::vector<double> dz(2 * (size_t)n, 0.0), rinf(n, INFINITY);
    int remaining = max_steps > 0 ? n : 0;
    const hipMemcpyKind h2d = hipMemcpyHostToDevice, d2h = hipMemcpyDeviceToHost;
    BCHK(b, hipMemcpyAsync(b->bs.active, ones.data(), sizeof(int) * n, h2d, b->stream));
    BCHK(b, hipMemcpyAsync(b->bs.done, zeros.data(), sizeof(int) * n, h2d, b->stream));
    BCHK(b, hipMemcpyAsync(b->bs.acc, dz.data(), sizeof(double) * 2 * n, h2d, b->stream));
    BCHK(b, hipMemcpyAsync(b->bs.res, rinf.data(), sizeof(double) * n, h2d, b->stream));
    BCHK(b, hipMemcpyAsync(b->bs.remaining, &remaining, sizeof(int), h2d, b->stream));
    BCHK(b, hipEventRecord(b->ev0, b->stream));
    if (max_steps > 0)
        BCHK(b, batch_run(b->storage == LBM_STORE_F64 ? 0 : 1, b->collision, b->inst, n, b->ntiles, b->P[0],
                          b->P[1], b->nodew, bfields(b), b->up, b->vp, b->bs, tol, check_every, (int)max_steps,
                          b->grid, b->stream));
    BCHK(b, hipEventRecord(b->ev1, b->stream));
    std::vector<int> done(n);
    std::vector<double> res(n);
    BCHK(b, hipMemcpyAsync(done.data(), b->bs.done, sizeof(int) * n, d2h, b->stream));
    BCHK(b, hipMemcpyAsync(res.data(), b->bs.res, sizeof(double) * n, d2h, b->stream));
    BCHK(b, hipMemcpyAsync(b->par.data(), b->bs.par, sizeof(int) * n, d2h, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    float t = 0.f;
    BCHK(b, hipEventElapsedTime(&t, b->ev0, b->ev1));
    bool diverged = false;
    for (int k = 0; k < n; ++k) {
        b->steps[k] += done[k];
        if (steps) steps[k] = done[k];
        if (residual) residual[k] = res[k];
        diverged |= done[k] > 0 && !(res[k] <= 1e300);
    }
    if (ms) *ms = t;
    if (diverged) return bfail(b, LBM_DIVERGED, "an instance went non-finite");
    return LBM_OK;
}

lbm_status lbm_batch_get_macros(lbm_batch* b, int32_t i, double* rho, double* u, double* v) {
    if (!b) return LBM_EINVAL;
    if (i < 0 || i >= b->n) return bfail(b, LBM_EINVAL, "instance %d out of range", i);
    const BInst& B = b->hinst[i];
    const size_t nn = (size_t)B.nx * B.ny, o = B.dbase;
    const double* src[3] = {b->rho + o, b->u + o, b->v + o};
    double* dst[3] = {rho, u, v};
    for (int a = 0; a < 3; ++a)
        if (dst[a]) BCHK(b, hipMemcpyAsync(dst[a], src[a], 8 * nn, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return LBM_OK;
}

lbm_status lbm_batch_get_f(lbm_batch* b, int32_t i, double* f) {
    if (!b || !f) return LBM_EINVAL;
    if (i < 0 || i >= b->n) return bfail(b, LBM_EINVAL, "instance %d out of range", i);
    if (b->steps[i] < 1) return bfail(b, LBM_EINVAL, "post-boundary populations exist after the first step");
    const BInst& B = b->hinst[i];
    const size_t nn = (size_t)B.nx * B.ny;
    // the last step read the other buffer; its post-boundary populations are
    // the reference's f after that iteration
    const int src = b->par[i] ^ 1;
    BCHK(b, hipMemsetAsync(b->fout, 0, 8 * 9 * nn, b->stream));
    BCHK(b, batch_recon(b->storage == LBM_STORE_F64 ? 0 : 1, b->inst, i, B.n, b->P[src], b->nodew, bfields(b),
                        b->fout, b->stream));
    BCHK(b, hipMemcpyAsync(f, b->fout, 8 * 9 * nn, hipMemcpyDeviceToHost, b->stream));
    BCHK(b, hipStreamSynchronize(b->stream));
    return LBM_OK;
}

int64_t lbm_batch_nodes(const lbm_batch* b, int64_t* fluid) {
    if (!b) return 0;
    if (fluid) *fluid = (int64_t)b->nfluid;
    return (int64_t)b->ndense;
}

}  // extern "C"
