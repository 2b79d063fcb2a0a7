// C ABI of the lattice-Boltzmann step (declared in include/lbm.h).
//
// Replaces the allocation, initialisation and time loop of the reference
// driver (S/Iterate.cu:23-625, S/Iterate3D.cu:30-839): one context owns one
// HIP stream and all device buffers; the loop enqueues one fused kernel per
// step and synchronises the host once per call, not once per phase.
#include "lbm.h"
#include "lbm_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

using namespace lbm;

struct lbm_ctx {
    lbm_params p{};
    KParams kp{};
    int q = 0;
    size_t n = 0;
    size_t raw_bytes = 0;     // bytes per stored population
    bool open = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    void* P[2] = {nullptr, nullptr};      // population buffers (256 B into their allocations)
    void* Palloc[2] = {nullptr, nullptr};
    int cur = 0;              // buffer holding the newest post-collision state
    int64_t steps = -1;       // steps since the state was set; -1: no state
    uint32_t* node = nullptr;
    uint32_t* cidx = nullptr;     // dense node -> compact population index (all ones: solid)
    uint32_t* didx = nullptr;     // compact -> dense
    uint32_t* nodew = nullptr;    // compact node words
    uint32_t* rowmeta = nullptr;  // per row and word: fluid mask, first compact index
    uint32_t* blk = nullptr;      // per workgroup: first row, rows spanned
    uint32_t* qslot = nullptr;    // curved walls: per compact node, row of qv
    double* qv = nullptr;         // curved walls: 19 wall fractions per row
    size_t nf = 0;                // non-solid nodes
    double *rho = nullptr, *u = nullptr, *v = nullptr, *w = nullptr;
    double *up = nullptr, *vp = nullptr, *wp = nullptr, *rp = nullptr;
    double *ubc = nullptr, *vbc = nullptr, *wbc = nullptr;
    double* fout = nullptr;
    double *f1 = nullptr, *fprev = nullptr;   // FdRelDiff: f at the first check and at the last one
    bool fdrel_started = false;
    double* om[2] = {nullptr, nullptr};   // f_eq outlet: macros of even / odd steps (4 N each)
    double* frz = nullptr;        // LBM_PARITY_INLET_FROZEN: macros at lbm_init (4 N)
    double* partials = nullptr;
    double* out4 = nullptr;       // 8 doubles: the last reduction
    int64_t bytes = 0;
    int dev = 0;                  // device current at lbm_create
    std::string err;
};

namespace {

lbm_status fail(lbm_ctx* c, lbm_status s, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf;
    return s;
}

#define HIPCHK(c, expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail((c), LBM_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_));          \
    } while (0)

lbm_status dalloc(lbm_ctx* c, void** p, size_t bytes) {
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(c, LBM_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
    c->bytes += (int64_t)bytes;
    return LBM_OK;
}

Fields fields(lbm_ctx* c) {
    Fields f;
    f.rho = c->rho; f.u = c->u; f.v = c->v; f.w = c->w;
    f.ubc = c->ubc; f.vbc = c->vbc; f.wbc = c->wbc;
    f.fout = c->fout;
    f.om_in = nullptr;
    f.om_out = nullptr;
    f.frz = c->frz;
    return f;
}

// The f_eq outlet of step s reads the macros of step s - 1 (slot (s-1) & 1)
// and writes those of step s (slot s & 1); step 0 is the initial state.
lbm_status seed_outlet_macros(lbm_ctx* c) {
    if (!c->om[0]) return LBM_OK;
    const double* src[4] = {c->rho, c->u, c->v, c->w};
    for (int a = 0; a < 4; ++a)
        HIPCHK(c, hipMemcpyAsync(c->om[0] + a * c->n, src[a], 8 * c->n, hipMemcpyDeviceToDevice, c->stream));
    return LBM_OK;
}

size_t storage_bytes(int s) { return s == LBM_STORE_F64 ? 8 : s == LBM_STORE_F32 ? 4 : 2; }

// Population plane stride: whole 64-node tiles, rounded to 256 B, plus an offset
// of one 4 KiB page and 256 B so that the Q planes of one node start on
// different HBM channels.
size_t plane_stride(size_t n, size_t eb) {
    size_t pad_bytes = 4096 + 256;
    if (const char* e = std::getenv("LBM_QPAD_BYTES")) pad_bytes = (size_t)std::strtoull(e, nullptr, 10);
    // whole 64-node tiles (the tiled layout addresses ceil(n/64)*64*Q elements)
    // and at least 256 B alignment of every plane
    size_t per = 256 / eb > 64 ? 256 / eb : 64;
    size_t s = (n + per - 1) / per * per;
    return s + (pad_bytes + eb - 1) / eb;
}

lbm_status ensure_fout(lbm_ctx* c) {
    if (c->fout) return LBM_OK;
    return dalloc(c, (void**)&c->fout, sizeof(double) * c->n * c->q);
}

// macros_last: the last launch also writes the FP64 macroscopic fields
// (timed runs leave them to an untimed moments pass, refresh_macros).
// ev_after (optional, nsteps events): recorded after each launch.
lbm_status run_steps(lbm_ctx* c, int64_t nsteps, bool macros_last = true, hipEvent_t* ev_after = nullptr) {
    for (int64_t k = 0; k < nsteps; ++k) {
        const bool wm = macros_last && (k == nsteps - 1);
        // per-launch seed of the stochastic rounding: a hash of the step count
        uint64_t z = (uint64_t)(c->steps + k + 1) * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        c->kp.seed = (uint32_t)(z ^ (z >> 31));
        Fields fl = fields(c);
        if (c->om[0]) {
            const int64_t s = c->steps + k + 1;
            fl.om_in = c->om[(s - 1) & 1];
            fl.om_out = c->om[s & 1];
        }
        hipError_t e = launch_step(c->q, c->p.storage, c->p.collision, c->open, c->kp, c->P[c->cur],
                                   c->P[c->cur ^ 1], c->node, c->cidx, fl, wm, c->stream);
        if (e != hipSuccess) return fail(c, LBM_EDEVICE, "step launch: %s", hipGetErrorString(e));
        c->cur ^= 1;
        if (ev_after) {
            e = hipEventRecord(ev_after[k], c->stream);
            if (e != hipSuccess) return fail(c, LBM_EDEVICE, "event record: %s", hipGetErrorString(e));
        }
    }
    c->steps += nsteps;
    return LBM_OK;
}

// Moments of the current state into rho, u, v, w without a step: the
// post-boundary state of the last step rebuilt from the buffer it read.
lbm_status refresh_macros(lbm_ctx* c) {
    if (c->steps < 1) return LBM_OK;
    Fields fl = fields(c);
    fl.fout = nullptr;
    if (c->om[0]) fl.om_in = c->om[(c->steps - 1) & 1];
    HIPCHK(c, launch_recon(c->q, c->p.storage, c->open, c->kp, c->P[c->cur ^ 1], c->node, c->cidx, fl, c->stream));
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char* lbm_last_error(const lbm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t lbm_device_bytes(const lbm_ctx* ctx) { return ctx ? ctx->bytes : 0; }

static thread_local std::string g_create_err;

lbm_status lbm_create(const lbm_params* p, const lbm_lattice* lat, lbm_ctx** out) {
    if (!out) return LBM_EINVAL;
    *out = nullptr;
    auto reject = [](const char* msg) { g_create_err = msg; std::fprintf(stderr, "lbm_create: %s\n", msg); return LBM_EINVAL; };
    if (!p || !lat || !lat->node_class) return reject("null parameters or lattice");
    if (p->abi_version != LBM_ABI_VERSION) return reject("ABI version mismatch");
    if (p->q != 9 && p->q != 19) return reject("q must be 9 or 19");
    const int nz = p->q == 9 ? 1 : p->nz;
    if (p->nx < 1 || p->ny < 1 || nz < 1) return reject("empty lattice");
    if (p->q == 9 && p->nz > 1) return reject("D2Q9 needs nz == 1");
    if (!(p->omega > 0.0 && p->omega < 2.0)) return reject("omega must lie in (0, 2) (tau > 1/2)");
    if (p->collision < 0 || p->collision > 2) return reject("bad collision model");
    if (p->storage < 0 || p->storage > 3) return reject("bad storage");
    if (p->force_model < 0 || p->force_model > 2) return reject("bad force model");
    if (p->wall_model < 0 || p->wall_model > 3) return reject("bad wall model");
    if (p->wall_model == LBM_WALL_HH && p->q != 19) return reject("wet-node walls are D3Q19 only");
    if (p->inlet_model < 0 || p->inlet_model > 1) return reject("bad inlet model");
    if (p->outlet_model < 0 || p->outlet_model > 4) return reject("bad outlet model");
    const size_t n = (size_t)p->nx * p->ny * nz;
    if (n > (size_t)1 << 33) return reject("lattice too large");

    // node words and open-boundary checks
    std::vector<uint32_t> node(n);
    bool any_open_cls = false, any_outlet = false, any_face3 = false;
    const uint32_t mask = p->q == 9 ? 0x1FEu : 0x7FFFEu;
    for (size_t i = 0; i < n; ++i) {
        const uint32_t cls = lat->node_class[i];
        if (cls > LBM_SYM_T) return reject("unknown node class");
        if (cls >= LBM_VEL_W) {
            // face of the class: axis and side (0 low, 1 high)
            static const int8_t AX[20] = {0, 0, 0, 0, 1, 0, 1, 2, 2, 0, 1, 1, 2, 2, 0, 0, 1, 1, 2, 2};
            static const int8_t HI[20] = {0, 0, 0, 1, 1, 1, 0, 0, 1, 0, 0, 1, 0, 1, 0, 1, 0, 1, 0, 1};
            if (p->q == 9 && cls > LBM_VEL_N) return reject("3D face classes need D3Q19");
            const int xyz[3] = {(int)(i % p->nx), (int)((i / p->nx) % p->ny), (int)(i / ((size_t)p->nx * p->ny))};
            const int dims[3] = {p->nx, p->ny, nz};
            const int a = AX[cls];
            if (xyz[a] != (HI[cls] ? dims[a] - 1 : 0)) return reject("boundary nodes must lie on their face of the box");
            if (p->periodic[a]) return reject("a face with boundary nodes needs a non-periodic axis");
            any_open_cls = true;
            any_outlet |= cls == LBM_PRES_E || (cls >= LBM_OUT_W && cls <= LBM_OUT_T);
            any_face3 |= p->q == 19;
        }
        const uint32_t bb = lat->bb_links ? (lat->bb_links[i] & mask) : 0u;
        node[i] = (cls << kClsShift) | bb;
    }
    bool simple = p->wall_model == LBM_WALL_HALFWAY;
    // A link into a solid node is a wall: complete the bounce-back bits from the
    // node classes (periodic axes wrap; links leaving a non-periodic box are
    // left to the lattice's own WALL links or the open boundaries).
    {
        static const int C9[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
                                     {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
        static const int C19[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1},
                                       {0, 0, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},
                                       {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1},
                                       {0, -1, -1}};
        const int nx = p->nx, ny = p->ny;
        const int per[3] = {p->periodic[0] != 0, p->periodic[1] != 0, p->q == 9 ? 1 : p->periodic[2] != 0};
        for (size_t i = 0; i < n; ++i) {
            if (lat->node_class[i] == LBM_SOLID) continue;
            const uint32_t user = node[i] & kBBMask;
            uint32_t derived = 0;
            const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((size_t)nx * ny));
            for (int d = 1; d < p->q; ++d) {
                const int* c = p->q == 9 ? C9[d] : C19[d];
                int xyz[3] = {x + c[0], y + c[1], z + c[2]};
                const int dims[3] = {nx, ny, nz};
                bool out = false;
                for (int a = 0; a < 3; ++a) {
                    if (xyz[a] < 0 || xyz[a] >= dims[a]) {
                        if (per[a]) xyz[a] = (xyz[a] + dims[a]) % dims[a];
                        else out = true;
                    }
                }
                if (out) continue;
                const size_t j = ((size_t)xyz[2] * ny + xyz[1]) * nx + xyz[0];
                if (lat->node_class[j] == LBM_SOLID) derived |= 1u << d;
            }
            node[i] |= derived;
            if ((user & ~derived) != 0 || lat->node_class[i] != LBM_FLUID) simple = false;
        }
    }
    if (any_outlet && p->q == 19 && p->outlet_model == LBM_OUTLET_REF_VELOCITY)
        return reject("the reference velocity outlet is 2D; 3D outlets are pressure, extrapolation or f_eq");
    if (any_outlet && p->q == 9 && p->outlet_model == LBM_OUTLET_FEQ) return reject("the f_eq outlet is 3D");
    if (p->wall_model == LBM_WALL_HH) {
        // Wet-node closure id from the wall links: 5 links with one axis
        // normal a face, 9 with two an edge, 12 with three a corner (the
        // inward normal along an axis is -1 when the +axis link is a wall).
        static const int AXD[3][2] = {{2, 1}, {4, 3}, {6, 5}};   // (-axis, +axis) directions
        for (size_t i = 0; i < n; ++i) {
            const uint32_t bb = node[i] & kBBMask;
            if (!bb || lat->node_class[i] == LBM_SOLID) continue;
            int g[3] = {0, 0, 0}, nn = 0;
            for (int a = 0; a < 3; ++a) {
                if (bb >> AXD[a][1] & 1u) g[a] = -1;
                else if (bb >> AXD[a][0] & 1u) g[a] = 1;
                nn += g[a] != 0;
            }
            const int cnt = __builtin_popcount(bb);
            uint32_t id = 0;
            if (cnt == 5 && nn == 1) {
                for (int a = 0; a < 3; ++a)
                    if (g[a]) id = 1 + 2 * a + (g[a] < 0);
            } else if (cnt == 9 && nn == 2) {
                const int A = g[0] ? 0 : 1, B = g[2] ? 2 : 1;
                const int pair = (A == 0 && B == 1) ? 0 : (A == 0 ? 1 : 2);
                id = 7 + 4 * pair + 2 * (g[A] < 0) + (g[B] < 0);
            } else if (cnt == 12 && nn == 3) {
                id = 19 + 4 * (g[0] < 0) + 2 * (g[1] < 0) + (g[2] < 0);
            }
            node[i] |= id << kHHShift;
        }
    }
    if (n >= 0xFFFFFFFFu) return reject("lattice too large for 32-bit node indices");
    std::vector<uint32_t> cidx(n), didx, nodew;
    size_t nf = 0;
    for (size_t i = 0; i < n; ++i) {
        if (lat->node_class[i] == LBM_SOLID) {
            cidx[i] = 0xFFFFFFFFu;
        } else {
            cidx[i] = (uint32_t)nf++;
            didx.push_back((uint32_t)i);
            nodew.push_back(node[i]);
        }
    }
    // Curved walls: the wall fractions of every node with wall links, one row
    // of 19 per such node (entries <= 0 or no table: 1/2)
    std::vector<uint32_t> qslot;
    std::vector<double> qv;
    if (p->wall_model == LBM_WALL_CURVED) {
        qslot.assign(nf > 0 ? nf : 1, 0xFFFFFFFFu);
        for (size_t j = 0; j < nf; ++j) {
            const size_t i = didx[j];
            if (!(node[i] & kBBMask)) continue;
            qslot[j] = (uint32_t)(qv.size() / 19);
            for (int d = 0; d < 19; ++d) {
                const double q = (lat->wall_q && d < p->q) ? lat->wall_q[i * 19 + d] : 0.0;
                if (q > 1.0) return reject("wall fractions must lie in (0, 1]");
                qv.push_back(q);
            }
        }
        if (qv.empty()) qv.assign(19, 0.0);
    }
    // Row metadata of the rank step: per row and 32-node word, the fluid mask
    // and the compact index of the word's first fluid node; per 256-node
    // workgroup, its first row and the number of rows it spans.
    const int words = (p->nx + 31) / 32;
    const size_t nrows = (size_t)p->ny * nz;
    std::vector<uint32_t> rowmeta(nrows * words * 2);
    {
        uint32_t cnt = 0;
        for (size_t r = 0; r < nrows; ++r)
            for (int w = 0; w < words; ++w) {
                uint32_t m = 0;
                for (int b = 0; b < 32 && 32 * w + b < p->nx; ++b)
                    if (lat->node_class[r * p->nx + 32 * w + b] != LBM_SOLID) m |= 1u << b;
                rowmeta[(r * words + w) * 2] = m;
                rowmeta[(r * words + w) * 2 + 1] = cnt;
                cnt += (uint32_t)__builtin_popcount(m);
            }
    }
    // nodes per workgroup of the rank step (256, 512 or 1024): 512 measured 3 %
    // faster for 4- and 8-byte storage on the 256^3 porous lattice, 256 best for
    // 16-bit storage with one node per thread.  The 16-bit periodic BGK/TRT
    // step runs two nodes per thread (k_stepr2) on 512-node workgroups:
    // 181.7 / 187.2 vs 187.2 / 191.7 us per step at 256^3 on two boxes.
    const bool narrow = p->storage == LBM_STORE_F16 || p->storage == LBM_STORE_BF16;
    const bool periodic_all = p->periodic[0] && p->periodic[1] && (p->q == 9 || p->periodic[2]);
    const int npt = narrow && simple && periodic_all && p->collision != LBM_MRT ? 2 : 1;
    size_t rblock = narrow ? (npt == 2 ? 512 : 256) : 512;
    if (const char* e = std::getenv("LBM_RANK_BLOCK")) {
        const size_t v = std::strtoul(e, nullptr, 10);
        if (v == 256 || v == 512 || v == 1024) rblock = v;
    }
    const size_t nblk = (nf + rblock - 1) / rblock;
    std::vector<uint32_t> blk(2 * (nblk > 0 ? nblk : 1));
    int rmax = 1;
    for (size_t b = 0; b < nblk; ++b) {
        const size_t first = didx[b * rblock], last = didx[std::min(nf, (b + 1) * rblock) - 1];
        const uint32_t r0 = (uint32_t)(first / p->nx), r1 = (uint32_t)(last / p->nx);
        blk[2 * b] = r0;
        blk[2 * b + 1] = r1 - r0 + 1;
        rmax = std::max(rmax, (int)(r1 - r0 + 1));
    }
    int lgw = 0;
    while ((1 << lgw) < words) ++lgw;
    // LDS: staged row words at a power-of-two pitch plus one node position per
    // thread; positions pack x in 16 bits
    const bool rank_ok = (((size_t)rmax * 9) << lgw) * 8 + 4 * rblock + 36 * (size_t)rmax <= 48 * 1024 && p->nx < 65536;
    if (any_open_cls && !lat->u_bc) {
        bool need = false;
        for (size_t i = 0; i < n; ++i) {
            const uint8_t cl = lat->node_class[i];
            need |= cl == LBM_VEL_W || cl == LBM_VEL_N || (cl >= LBM_VEL_E && cl <= LBM_VEL_T) ||
                    (cl == LBM_PRES_E && p->outlet_model == LBM_OUTLET_REF_VELOCITY);
        }
        if (need) return reject("velocity inlet, lid and velocity outlet nodes need u_bc");
    }

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return reject("no HIP device");

    lbm_ctx* c = new lbm_ctx();
    if (hipGetDevice(&c->dev) != hipSuccess) {
        delete c;
        return reject("no current HIP device");
    }
    c->p = *p;
    c->p.nz = nz;
    c->q = p->q;
    c->n = n;
    c->nf = nf;
    c->raw_bytes = storage_bytes(p->storage);
    c->open = !(p->periodic[0] && p->periodic[1] && (p->q == 9 || p->periodic[2]));
    KParams& k = c->kp;
    k.nx = p->nx; k.ny = p->ny; k.nz = nz;
    for (int a = 0; a < 3; ++a) k.per[a] = p->periodic[a] ? 1 : 0;
    if (p->q == 9) k.per[2] = 1;
    k.qstride = plane_stride(nf > 0 ? nf : 1, c->raw_bytes);
    k.qstride32 = (uint32_t)k.qstride;
    if (k.qstride * c->q * c->raw_bytes >= (1ull << 32) || n * 4 >= (1ull << 32)) {
        delete c;
        return reject("population buffer or node map above 4 GiB (32-bit offsets)");
    }
    k.nf = (uint32_t)nf;
    auto magic = [](uint32_t d, uint64_t& m, uint32_t& s) {
        s = 0;
        while ((1ull << s) < d) ++s;
        m = ((1ull << 32) * ((1ull << s) - d)) / d + 1;
    };
    magic((uint32_t)p->nx, k.mx, k.sx);
    magic((uint32_t)p->ny, k.my, k.sy);
    // the 8 XCDs of an MI355X each take 8 consecutive workgroups per turn: at
    // 256^3 -1.6 % (FP32) and -1.8 % (FP16) per step against the round-robin
    // order; chunks of 2-32 and one contiguous range per XCD measured slower
    k.xcd_shift = 3;
    k.xchunk_shift = 3;
    k.tile = 1;
    if (const char* e = std::getenv("LBM_LAYOUT")) k.tile = std::strcmp(e, "soa") == 0 ? 0 : 1;
    log2_env("LBM_XCD", k.xcd_shift);
    log2_env("LBM_XCHUNK", k.xchunk_shift);
    k.omega = p->omega;
    const double lam = p->trt_magic > 0 ? p->trt_magic : 3.0 / 16.0;
    k.omega_m = 1.0 / (0.5 + lam / (1.0 / p->omega - 0.5));
    // reference diag(S) (S/CellFunctions.cu:44-52, :122-140) unless given
    bool any_s = false;
    for (int i = 0; i < p->q; ++i) any_s |= p->mrt_s[i] != 0.0;
    if (any_s) {
        for (int i = 0; i < p->q; ++i) k.s[i] = p->mrt_s[i];
    } else if (p->q == 9) {
        const double s9[9] = {1.0, 1.4, 1.4, 1.0, 1.2, 1.0, 1.2, p->omega, p->omega};
        for (int i = 0; i < 9; ++i) k.s[i] = s9[i];
    } else {
        const double s19[19] = {0.0, 1.19, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 1.2, p->omega,
                                1.4, p->omega, 1.4, p->omega, p->omega, p->omega, 1.98, 1.98, 1.98};
        for (int i = 0; i < 19; ++i) k.s[i] = s19[i];
    }
    for (int a = 0; a < 3; ++a) k.g[a] = p->force[a];
    set_float_params(k);
    k.force = p->force_model;
    k.wall = p->wall_model;
    k.inlet = p->inlet_model;
    k.outlet = p->outlet_model;
    k.parity = p->parity_flags;
    // D3Q19 MRT stores deviations from its own rest state (w_eps = 0 moves it
    // off w), so reduced-precision copies keep their digits for the flow
    k.mref = (p->q == 19 && p->collision == LBM_MRT) ? 1 : 0;
    k.rho_out = p->rho_out > 0 ? p->rho_out : 1.0;

    lbm_status st = LBM_OK;
    auto cleanup = [&](lbm_status s) { lbm_destroy(c); return s; };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        g_create_err = "stream/event creation failed";
        return cleanup(LBM_EDEVICE);
    }
    // 256 B guard before and after the planes: the shifted vector loads of the
    // first and last rows may touch one element outside the lattice (discarded).
    // In front of that, kRankShift nodes' worth of tiles, so that the 16-bit
    // rank step's base pointer (kRankShift nodes before the buffer) still lies
    // inside the allocation.
    const size_t front = 256 + (size_t)kRankShift * c->q * c->raw_bytes;
    const size_t pbytes = c->raw_bytes * k.qstride * c->q + front + 256;
    for (int b = 0; b < 2 && st == LBM_OK; ++b) {
        st = dalloc(c, &c->Palloc[b], pbytes);
        if (st == LBM_OK) c->P[b] = (char*)c->Palloc[b] + front;
    }
    if (st == LBM_OK) st = dalloc(c, (void**)&c->node, sizeof(uint32_t) * n);
    if (st == LBM_OK) st = dalloc(c, (void**)&c->cidx, sizeof(uint32_t) * n);
    if (st == LBM_OK) st = dalloc(c, (void**)&c->didx, sizeof(uint32_t) * (nf > 0 ? nf : 1));
    if (st == LBM_OK) st = dalloc(c, (void**)&c->nodew, sizeof(uint32_t) * (nf > 0 ? nf : 1));
    if (st == LBM_OK) st = dalloc(c, (void**)&c->rowmeta, sizeof(uint32_t) * rowmeta.size());
    if (st == LBM_OK) st = dalloc(c, (void**)&c->blk, sizeof(uint32_t) * blk.size());
    if (st == LBM_OK && !qslot.empty()) {
        st = dalloc(c, (void**)&c->qslot, sizeof(uint32_t) * qslot.size());
        if (st == LBM_OK) st = dalloc(c, (void**)&c->qv, sizeof(double) * qv.size());
    }
    double** fl[] = {&c->rho, &c->u, &c->v, &c->w, &c->up, &c->vp, &c->wp, &c->rp};
    for (double** a : fl)
        if (st == LBM_OK) st = dalloc(c, (void**)a, sizeof(double) * n);
    if (st == LBM_OK && any_open_cls) {
        st = dalloc(c, (void**)&c->ubc, sizeof(double) * n);
        if (st == LBM_OK) st = dalloc(c, (void**)&c->vbc, sizeof(double) * n);
        if (st == LBM_OK) st = dalloc(c, (void**)&c->wbc, sizeof(double) * n);
    }
    if (st == LBM_OK && any_outlet && p->q == 19 && p->outlet_model == LBM_OUTLET_FEQ)
        for (int b = 0; b < 2 && st == LBM_OK; ++b) st = dalloc(c, (void**)&c->om[b], sizeof(double) * 4 * n);
    if (st == LBM_OK && any_open_cls && (p->parity_flags & LBM_PARITY_INLET_FROZEN)) {
        // until lbm_init sets them, the frozen macros are the rest state
        st = dalloc(c, (void**)&c->frz, sizeof(double) * 4 * n);
        if (st == LBM_OK) {
            std::vector<double> rest(4 * n, 0.0);
            std::fill(rest.begin(), rest.begin() + n, 1.0);
            if (hipMemcpy(c->frz, rest.data(), sizeof(double) * 4 * n, hipMemcpyHostToDevice) != hipSuccess)
                st = fail(c, LBM_EDEVICE, "cannot seed the frozen inlet macros");
        }
    }
    if (st == LBM_OK) st = dalloc(c, (void**)&c->partials, sizeof(double) * 8 * reduce_blocks());
    if (st == LBM_OK) st = dalloc(c, (void**)&c->out4, sizeof(double) * 8);
    if (st != LBM_OK) {
        g_create_err = c->err;
        return cleanup(st);
    }
    k.didx = c->didx;
    k.nodew = c->nodew;
    k.rowmeta = (const uint2*)c->rowmeta;
    k.blk = (const uint2*)c->blk;
    k.words = words;
    k.lgw = lgw;
    k.rblock = (int)rblock;
    k.rmax = rmax;
    k.npt = npt;
    if (const char* e = std::getenv("LBM_NPT")) {
        const int v = std::atoi(e);
        k.npt = (v == 2 || v == 4) && (int)rblock / v >= 64 ? v : 1;
    }
    k.simple = simple ? 1 : 0;
    k.qslot = c->qslot;
    k.qv = c->qv;
    // the rank step addresses the tile layout (SoA planes only in its 16-bit
    // periodic variant), and has no extrapolating outlet, 3D face or wet-node wall
    const bool rank_fits = rank_ok && (k.tile || (c->raw_bytes == 2 && simple && !c->open)) && !(any_outlet && p->outlet_model >= LBM_OUTLET_EXTRAP2) &&
                           !any_face3 && p->wall_model != LBM_WALL_HH && p->wall_model != LBM_WALL_CURVED;
    k.rank = rank_fits ? 1 : 0;
    if (const char* e = std::getenv("LBM_RANK")) k.rank = rank_fits && e[0] == '1';
    k.sr = 1;
    if (const char* e = std::getenv("LBM_SR")) k.sr = e[0] == '1';
    auto h2d = [&](void* d, const void* h, size_t b) { return hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, c->stream); };
    if (h2d(c->node, node.data(), sizeof(uint32_t) * n) != hipSuccess ||
        h2d(c->cidx, cidx.data(), sizeof(uint32_t) * n) != hipSuccess ||
        (nf > 0 && h2d(c->didx, didx.data(), sizeof(uint32_t) * nf) != hipSuccess) ||
        (nf > 0 && h2d(c->nodew, nodew.data(), sizeof(uint32_t) * nf) != hipSuccess) ||
        h2d(c->rowmeta, rowmeta.data(), sizeof(uint32_t) * rowmeta.size()) != hipSuccess ||
        h2d(c->blk, blk.data(), sizeof(uint32_t) * blk.size()) != hipSuccess ||
        (c->qslot && h2d(c->qslot, qslot.data(), sizeof(uint32_t) * qslot.size()) != hipSuccess) ||
        (c->qv && h2d(c->qv, qv.data(), sizeof(double) * qv.size()) != hipSuccess))
        return cleanup(LBM_EDEVICE);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(LBM_EDEVICE);
    if (c->ubc) {
        std::vector<double> zero(n, 0.0);
        if (h2d(c->ubc, lat->u_bc ? lat->u_bc : zero.data(), 8 * n) != hipSuccess ||
            h2d(c->vbc, lat->v_bc ? lat->v_bc : zero.data(), 8 * n) != hipSuccess ||
            h2d(c->wbc, lat->w_bc ? lat->w_bc : zero.data(), 8 * n) != hipSuccess)
            return cleanup(LBM_EDEVICE);
        if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(LBM_EDEVICE);
    }
    double* zs[] = {c->up, c->vp, c->wp, c->rp};
    for (double* a : zs)
        if (hipMemsetAsync(a, 0, 8 * n, c->stream) != hipSuccess) return cleanup(LBM_EDEVICE);
    if (hipStreamSynchronize(c->stream) != hipSuccess) return cleanup(LBM_EDEVICE);
    *out = c;
    return LBM_OK;
}

void lbm_destroy(lbm_ctx* c) {
    if (!c) return;
    if (c->stream) hipStreamSynchronize(c->stream);
    void* bufs[] = {c->Palloc[0], c->Palloc[1], c->node, c->cidx, c->didx, c->nodew, c->rowmeta, c->blk, c->rho, c->u, c->v, c->w, c->up, c->vp, c->wp,
                    c->rp, c->qslot, c->qv, c->ubc, c->vbc, c->wbc, c->fout, c->f1, c->fprev, c->om[0], c->om[1], c->frz, c->partials, c->out4};
    for (void* b : bufs)
        if (b) hipFree(b);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

lbm_status lbm_init(lbm_ctx* c, int mode, const double* rho, const double* u, const double* v, const double* w) {
    if (!c) return LBM_EINVAL;
    if (mode != 0 && mode != 1) return fail(c, LBM_EINVAL, "init mode must be 0 (equilibrium) or 1 (zero f)");
    const size_t n = c->n, b = 8 * n;
    std::vector<double> ones;
    if (!rho) ones.assign(n, 1.0);
    HIPCHK(c, hipMemcpyAsync(c->rho, rho ? rho : ones.data(), b, hipMemcpyHostToDevice, c->stream));
    const double* src[3] = {u, v, w};
    double* dst[3] = {c->u, c->v, c->w};
    for (int a = 0; a < 3; ++a) {
        if (src[a]) HIPCHK(c, hipMemcpyAsync(dst[a], src[a], b, hipMemcpyHostToDevice, c->stream));
        else HIPCHK(c, hipMemsetAsync(dst[a], 0, b, c->stream));
    }
    if (c->frz) {
        const double* mac[4] = {c->rho, c->u, c->v, c->w};
        for (int a = 0; a < 4; ++a)
            HIPCHK(c, hipMemcpyAsync(c->frz + a * n, mac[a], b, hipMemcpyDeviceToDevice, c->stream));
    }
    HIPCHK(c, launch_init(c->q, c->p.storage, c->p.collision, mode == 0 ? INIT_EQ : INIT_ZERO, c->kp,
                          c->P[0], c->node, c->cidx, fields(c), c->stream));
    c->cur = 0;
    c->steps = 0;
    c->fdrel_started = false;
    lbm_status s = seed_outlet_macros(c);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_step(lbm_ctx* c, int64_t nsteps) {
    if (!c) return LBM_EINVAL;
    if (c->steps < 0) return fail(c, LBM_EINVAL, "state not initialised (call lbm_init or lbm_set_f)");
    if (nsteps < 0) return fail(c, LBM_EINVAL, "negative step count");
    lbm_status s = run_steps(c, nsteps);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_reduce_all(lbm_ctx* c, double out[8]) {
    if (!c || !out) return LBM_EINVAL;
    if (c->steps < 0) return fail(c, LBM_EINVAL, "state not initialised");
    HIPCHK(c, launch_reduce(c->kp, c->node, fields(c), c->up, c->vp, c->wp, c->rp, c->partials, reduce_blocks(),
                            c->out4, c->stream));
    HIPCHK(c, hipMemcpyAsync(out, c->out4, 8 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_reduce(lbm_ctx* c, double out[4]) {
    if (!c || !out) return LBM_EINVAL;
    double r8[8];
    lbm_status s = lbm_reduce_all(c, r8);
    if (s != LBM_OK) return s;
    for (int k = 0; k < 4; ++k) out[k] = r8[k];
    return LBM_OK;
}

lbm_status lbm_residual_l2(lbm_ctx* c, double* out) {
    if (!c || !out) return LBM_EINVAL;
    if (c->steps < 1) return fail(c, LBM_EINVAL, "the L2 residual needs one step");
    lbm_status s = ensure_fout(c);
    if (s != LBM_OK) return s;
    Fields fl = fields(c);
    if (c->om[0]) fl.om_in = c->om[(c->steps - 1) & 1];
    HIPCHK(c, launch_recon(c->q, c->p.storage, c->open, c->kp, c->P[c->cur ^ 1], c->node, c->cidx, fl, c->stream));
    // the reference's fColl of step s is the collision output that step s
    // streamed, i.e. the buffer k_recon just read
    HIPCHK(c, launch_fdiff2(c->q, c->p.storage, c->kp, c->P[c->cur ^ 1], c->cidx, c->fout, c->partials,
                            reduce_blocks(), c->out4, c->stream));
    double sum = 0;
    HIPCHK(c, hipMemcpyAsync(&sum, c->out4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = std::sqrt(sum);
    return LBM_OK;
}

lbm_status lbm_residual_l2_noalloc(lbm_ctx* c, double* out) {
    if (!c || !out) return LBM_EINVAL;
    if (c->steps < 1) return fail(c, LBM_EINVAL, "the L2 residual needs one step");
    int dev = -1;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev != c->dev)
        return fail(c, LBM_EINVAL, "lbm_residual_l2_noalloc: current device %d is not the context's device %d", dev,
                    c->dev);
    const int nb = reduce_blocks();
    if (nb + 1 > kL2ScratchDoubles)
        return fail(c, LBM_EINVAL, "lbm_residual_l2_noalloc: %d reduction blocks exceed the %d-double scratch", nb,
                    kL2ScratchDoubles);
    double* scr = nullptr;
    HIPCHK(c, l2_scratch(&scr));
    Fields fl = fields(c);
    fl.fout = nullptr;
    if (c->om[0]) fl.om_in = c->om[(c->steps - 1) & 1];
    // partials in scr[0, nb), the sum in scr[nb]; every slot read is written
    // first by this call, so nothing needs clearing
    HIPCHK(c, launch_l2_fused(c->q, c->p.storage, c->open, c->kp, c->P[c->cur ^ 1], c->node, c->cidx, fl, scr, nb,
                              scr + nb, c->stream));
    double sum = 0;
    HIPCHK(c, hipMemcpyAsync(&sum, scr + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = std::sqrt(sum);
    return LBM_OK;
}

lbm_status lbm_residual_fdrel(lbm_ctx* c, double* out) {
    if (!c || !out) return LBM_EINVAL;
    if (c->steps < 1) return fail(c, LBM_EINVAL, "the FdRelDiff residual needs one step");
    lbm_status s = ensure_fout(c);
    if (s == LBM_OK && !c->f1) s = dalloc(c, (void**)&c->f1, sizeof(double) * c->n * c->q);
    if (s == LBM_OK && !c->fprev) s = dalloc(c, (void**)&c->fprev, sizeof(double) * c->n * c->q);
    if (s != LBM_OK) return s;
    Fields fl = fields(c);
    if (c->om[0]) fl.om_in = c->om[(c->steps - 1) & 1];
    HIPCHK(c, launch_recon(c->q, c->p.storage, c->open, c->kp, c->P[c->cur ^ 1], c->node, c->cidx, fl, c->stream));
    const size_t bytes = sizeof(double) * c->n * c->q;
    double r = 1.0;
    if (!c->fdrel_started) {
        // the reference compares the first check with its zero initial
        // populations: |f - 0| / f = 1 on every link
        HIPCHK(c, hipMemcpyAsync(c->f1, c->fout, bytes, hipMemcpyDeviceToDevice, c->stream));
        c->fdrel_started = true;
    } else {
        HIPCHK(c, launch_fdrel(c->q, c->fout, c->f1, c->fprev, c->cidx, c->n, c->partials, reduce_blocks(), c->out4,
                               c->stream));
        HIPCHK(c, hipMemcpyAsync(&r, c->out4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipMemcpyAsync(c->fprev, c->fout, bytes, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *out = r;
    return LBM_OK;
}

lbm_status lbm_run_to_residual(lbm_ctx* c, double tol, int64_t every, int64_t max_steps, int64_t* steps_done,
                               double* final_res) {
    if (!c) return LBM_EINVAL;
    if (every < 1 || max_steps < 0) return fail(c, LBM_EINVAL, "bad check interval or step limit");
    double r4[8];
    lbm_status s = lbm_reduce_all(c, r4);   // seeds the previous fields with the current ones
    if (s != LBM_OK) return s;
    int64_t done = 0;
    double res = INFINITY;
    while (done < max_steps) {
        const int64_t k = (max_steps - done) < every ? (max_steps - done) : every;
        if ((s = run_steps(c, k)) != LBM_OK) return s;
        done += k;
        if ((s = lbm_reduce_all(c, r4)) != LBM_OK) return s;
        if (r4[3] > 0) {
            if (steps_done) *steps_done = done;
            if (final_res) *final_res = NAN;
            return fail(c, LBM_DIVERGED, "non-finite field after %lld steps", (long long)done);
        }
        res = r4[1] > 0 ? std::sqrt(r4[0] / r4[1]) : std::sqrt(r4[0]);
        if (res < tol) break;
    }
    if (steps_done) *steps_done = done;
    if (final_res) *final_res = res;
    return LBM_OK;
}

lbm_status lbm_get_macros(lbm_ctx* c, double* rho, double* u, double* v, double* w) {
    if (!c) return LBM_EINVAL;
    if (c->steps < 0) return fail(c, LBM_EINVAL, "state not initialised");
    const size_t b = 8 * c->n;
    const double* srcs[4] = {c->rho, c->u, c->v, c->w};
    double* dsts[4] = {rho, u, v, w};
    for (int a = 0; a < 4; ++a)
        if (dsts[a]) HIPCHK(c, hipMemcpyAsync(dsts[a], srcs[a], b, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_get_f(lbm_ctx* c, double* f) {
    if (!c || !f) return LBM_EINVAL;
    if (c->steps < 1) return fail(c, LBM_EINVAL, "post-boundary populations exist after the first step");
    lbm_status s = ensure_fout(c);
    if (s != LBM_OK) return s;
    Fields fl = fields(c);
    if (c->om[0]) fl.om_in = c->om[(c->steps - 1) & 1];
    HIPCHK(c, launch_recon(c->q, c->p.storage, c->open, c->kp, c->P[c->cur ^ 1], c->node, c->cidx, fl, c->stream));
    HIPCHK(c, hipMemcpyAsync(f, c->fout, 8 * c->n * c->q, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_set_f(lbm_ctx* c, const double* f) {
    if (!c || !f) return LBM_EINVAL;
    lbm_status s = ensure_fout(c);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipMemcpyAsync(c->fout, f, 8 * c->n * c->q, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, launch_init(c->q, c->p.storage, c->p.collision, INIT_FROM_F, c->kp, c->P[0], c->node, c->cidx, fields(c),
                          c->stream));
    c->cur = 0;
    c->steps = 0;
    c->fdrel_started = false;
    s = seed_outlet_macros(c);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_time_steps(lbm_ctx* c, int64_t nsteps, double* ms) {
    if (!c || !ms || nsteps < 1) return LBM_EINVAL;
    if (c->steps < 0) return fail(c, LBM_EINVAL, "state not initialised");
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    lbm_status s = run_steps(c, nsteps, false);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    s = refresh_macros(c);
    if (s != LBM_OK) return s;
    HIPCHK(c, hipEventSynchronize(c->ev1));
    float t = 0;
    HIPCHK(c, hipEventElapsedTime(&t, c->ev0, c->ev1));
    *ms = t;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LBM_OK;
}

lbm_status lbm_time_launches(lbm_ctx* c, int64_t nsteps, double* ms_each) {
    if (!c || !ms_each || nsteps < 1 || nsteps > 100000) return LBM_EINVAL;
    if (c->steps < 0) return fail(c, LBM_EINVAL, "state not initialised");
    std::vector<hipEvent_t> ev((size_t)nsteps + 1, nullptr);
    lbm_status s = LBM_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) { s = fail(c, LBM_EDEVICE, "event create failed"); break; }
    if (s == LBM_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = fail(c, LBM_EDEVICE, "sync failed");
    if (s == LBM_OK && hipEventRecord(ev[0], c->stream) != hipSuccess) s = fail(c, LBM_EDEVICE, "event record failed");
    if (s == LBM_OK) s = run_steps(c, nsteps, false, ev.data() + 1);
    if (s == LBM_OK) s = refresh_macros(c);
    if (s == LBM_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = fail(c, LBM_EDEVICE, "sync failed");
    for (int64_t k = 0; s == LBM_OK && k < nsteps; ++k) {
        float t = 0;
        if (hipEventElapsedTime(&t, ev[(size_t)k], ev[(size_t)k + 1]) != hipSuccess) s = fail(c, LBM_EDEVICE, "elapsed time failed");
        ms_each[k] = t;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return s;
}

}  // extern "C"
