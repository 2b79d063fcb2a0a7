// Fused lattice-Boltzmann step kernels for CDNA4 (gfx950).
//
// One thread per lattice node, x fastest; a workgroup covers a run of x at a
// fixed (y, z), so the (y, z) neighbour rows of the pull are wave-uniform.
//
// Populations of the non-solid nodes only are stored, in lattice order, as SoA
// planes P[q*qstride + j] or 64-node tiles (j = compact index of the node), in two buffers
// (A/B).  A dense map cidx[node] -> j (all-ones for solid nodes, 4 B per node,
// small enough to stay in the last-level cache) turns the pull into a gather.
// A step reads the post-collision populations of the previous step from A,
// pulls them along the links (stream), applies walls and open boundaries, takes
// the moments, collides and writes B: the reference's collide -> stream -> BC
// -> macro sequence (S/Iterate.cu:385-520, S/Iterate3D.cu:493-653) rotated so
// that one node's whole update stays in registers.  Solid nodes cost no
// population traffic.
//
// Populations are held as deviations h_i = f_i - w_i (rest density 1) both
// in storage and in the arithmetic, so FP32 and 16-bit storage keep the
// small flow signal instead of rounding it against w_i.
#include "lbm_device.h"

#include <mutex>

namespace lbm {

// Store of one node's post-collision populations.  The 16-bit storages use
// stochastic rounding (unbiased: a body force smaller than half an ulp of the
// stored deviation still accumulates), with random bits from a hash of the
// compact node index and a per-launch seed, rotated per direction; results are
// deterministic for a given seed sequence.
// LAY: the buffer's layout, 0 tested at run time (kp.tile), 1 known to be
// tiles, 2 known to be SoA planes.
// X: out holds the stored values x_i = h_i / a_i of a 16-bit storage (no
// scaling before the conversion).
template <class L, int S, int LAY = 0, bool X = false>
__device__ inline void store_node(typename Io<L, S>::raw* __restrict__ dst, const KParams& kp, uint32_t own,
                                  const typename Io<L, S>::T* out) {
    using raw = typename Io<L, S>::raw;
    // With the tile layout a node's populations are 64 elements apart: one
    // 64-bit base, and each store's offset is a constant (an immediate of the
    // store instruction rather than an add per population).  With known SoA
    // planes each plane has its own (uniform) base and every store the same
    // offset, the node's index.
    raw* const tb = LAY == 1 ? (raw*)((char*)dst + (size_t)tslot<L>(0, own) * sizeof(raw)) : dst;
    auto base = [&](int i) { return LAY == 1 ? tb : LAY == 2 ? dst + (size_t)i * kp.qstride : dst; };
    auto slot = [&](int i) { return LAY == 1 ? 64u * (uint32_t)i : LAY == 2 ? own : paddr<L>(kp, i, own); };
    if constexpr (S >= 2) {
        if (kp.sr) {
            uint32_t h = own * 0x9E3779B1u + kp.seed;
            h ^= h >> 15;
            h *= 0x2C1B3C6Du;
            h ^= h >> 12;
            // the conversions write the low half of one register in turn (no
            // register to clear per population)
            typename Io<L, S>::pair acc = {0, 0};
#pragma unroll
            for (int i = 0; i < L::Q; ++i) {
                const uint32_t rnd = (h << (i & 31)) | (h >> ((32 - i) & 31));
                if constexpr (X) acc = Io<L, S>::srx(acc, out[i], rnd);
                else acc = Io<L, S>::sr(acc, out[i], i, rnd);
                sto(base(i), slot(i), (raw)acc[0]);
            }
            return;
        }
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        if constexpr (X) sto(base(i), slot(i), Io<L, S>::px(out[i]));
        else sto(base(i), slot(i), Io<L, S>::pk(out[i], i));
    }
}

// ---------------------------------------------------------------------------
// kernels

// One thread per non-solid node, in compact order: reads of the node's own
// data and all writes are contiguous and aligned across a wave.
template <class L, int S, int C, bool OPEN>
__global__ __launch_bounds__(256) void k_step(const typename Io<L, S>::raw* __restrict__ src,
                                              typename Io<L, S>::raw* __restrict__ dst,
                                              const uint32_t* __restrict__ didx,
                                              const uint32_t* __restrict__ nodew,
                                              const uint32_t* __restrict__ cidx, KParams kp,
                                              Fields fl, int write_macros) {
    using T = typename Io<L, S>::T;
    const uint32_t own = xcd_block(blockIdx.x, gridDim.x, kp.xcd_shift, kp.xchunk_shift) * blockDim.x + threadIdx.x;
    if (own >= kp.nf) return;
    const uint32_t idx = didx[own];
    const uint32_t nd = nodew[own];
    const uint32_t q1 = fdiv(idx, kp.mx, kp.sx);
    const int x = (int)(idx - q1 * (uint32_t)kp.nx);
    const uint32_t zq = fdiv(q1, kp.my, kp.sy);
    const int y = (int)(q1 - zq * (uint32_t)kp.ny), z = (int)zq;
    const uint32_t cls = nd >> kClsShift;
    T h[L::Q];
    inflow_state<L, S, OPEN>(src, kp, fl, cidx, own, nd, x, y, z, idx, h);
    if (OPEN && L::Q == 9 && cls == PRES_E) {
        zou_outlet<L, T>(kp, fl, idx, h);
        outlet_extrap<L, S, OPEN>(src, kp, fl, cidx, nodew, true, x, y, z, idx, h);
    }
    if (OPEN && L::Q == 19) {
        if (is_outlet3(cls)) outlet3d<L, S, OPEN>(src, kp, fl, cidx, nodew, true, x, y, z, idx, cls, h);
        sym3d<L, T>(cls, h);
    }
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    if (OPEN) frozen_macros<T>(kp, fl, cls, idx, drho, rho, u);
    if (write_macros) {
        fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
        fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
    }
    if (OPEN && L::Q == 19 && fl.om_out && is_outlet3(cls)) {
        const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
        fl.om_out[idx] = (double)rho; fl.om_out[n + idx] = (double)u[0];
        fl.om_out[2 * n + idx] = (double)u[1]; fl.om_out[3 * n + idx] = (double)u[2];
    }
    T out[L::Q];
    collide<L, T, C>(kp, h, drho, rho, u, out);
    store_node<L, S>(dst, kp, own, out);
}

// Rank variant of the step: no per-node index loads.  For every lattice row
// (fixed y, z) rowmeta holds, per 32-node word, the fluid bitmask and the
// compact index of the word's first fluid node.  A workgroup (256 compact
// nodes, spanning R rows) stages the words of its rows and of their eight
// neighbour rows in LDS; a node's own position and the compact index of every
// source node are then a popcount away.  With SIMPLE (every wall is a link into
// a solid node, all non-solid nodes plain fluid) the node words are not read
// either, and a step moves only the populations.

constexpr uint32_t kSolidSrc = 0xFFFFFFFEu;

template <class L, int S, int C, bool OPEN, bool SIMPLE, bool SOA = false>
__global__ __launch_bounds__(1024) void k_stepr(const typename Io<L, S>::raw* __restrict__ src,
                                               typename Io<L, S>::raw* __restrict__ dst,
                                               const uint2* __restrict__ rowmeta,
                                               const uint2* __restrict__ blk,
                                               const uint32_t* __restrict__ nodew, KParams kp,
                                               Fields fl, int write_macros) {
    using T = typename Io<L, S>::T;
    extern __shared__ uint2 smeta[];
    const int W = kp.words, lg = kp.lgw, nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x, kp.xcd_shift, kp.xchunk_shift);
    const uint2 bi = blk[bid];
    const uint32_t r0 = bi.x;
    const int R = (int)bi.y;
    // Stage the words of rows r0..r0+R-1 and of their (dy, dz) neighbours at
    // smeta[((k << lg) + w)*9 + nb], nb = (dz+1)*3 + (dy+1), rows padded to a
    // power of two: the nine neighbour words of one word are adjacent, so a
    // node reaches all of them from three base addresses with constant
    // offsets.  Rows outside a non-periodic face read as all-solid (outside
    // sources are resolved by coordinates).
    // 16-bit periodic lattices count compact indices from kRankShift (the
    // base pointer of their loads sits that many nodes before the buffer)
    constexpr bool kShifted = sizeof(typename Io<L, S>::raw) == 2 && SIMPLE && !OPEN;
    constexpr uint32_t kSh = kShifted ? kRankShift : 0u;
    // kShifted on SoA planes: the counts run 3 planes ahead instead (every
    // opposite plane lies within 3 planes of its link's own)
    constexpr bool kSoA = kShifted && SOA;
    static_assert(!SOA || kShifted, "the SoA variant is the 16-bit periodic one");
    const uint32_t kShv = kSoA ? 3u * kp.qstride32 : kSh;
    const int tot = (R * 9) << lg;
    const uint32_t own0 = bid * blockDim.x;
    uint32_t own, r;
    int k, x;
    // kShifted: the node's dense index, loaded before the staging so that its
    // latency hides behind it
    uint32_t dj = 0;
    if constexpr (kShifted) {
        if (own0 + threadIdx.x < kp.nf) dj = kp.didx[own0 + threadIdx.x];
    }
    if constexpr (kShifted) {
        // Periodic lattice: the R * 9 neighbour-row offsets first (one thread
        // each), then a plain copy of the staged words; every node's position
        // comes from its dense index (4 B per node) rather than a scatter of
        // the staged bits, which costs more instructions than it saves bytes
        // in this instruction-bound variant.
        uint32_t* rowoff = (uint32_t*)(smeta + ((kp.rmax * 9) << lg)) + kp.rblock;
        for (int t = threadIdx.x; t < R * 9; t += blockDim.x) {
            const int kk = t / 9, nb = t - 9 * kk;
            const uint32_t rr = r0 + (uint32_t)kk;
            const uint32_t zq = fdiv(rr, kp.my, kp.sy);
            int y = (int)(rr - zq * (uint32_t)ny) + nb % 3 - 1, z = (int)zq + nb / 3 - 1;
            y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
            z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
            rowoff[t] = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)W;
        }
        __syncthreads();
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int w = e & ((1 << lg) - 1), slot = e >> lg;
            const int kk = slot / 9, nb = slot - 9 * kk;
            uint2 val = make_uint2(0u, 0u);
            if (w < W) val = rowmeta[rowoff[slot] + (uint32_t)w];
            val.y += kShv;
            smeta[((kk << lg) + w) * 9 + nb] = val;
        }
        __syncthreads();
        own = own0 + threadIdx.x;
        if (own >= kp.nf) return;
        r = fdiv(dj, kp.mx, kp.sx);
        x = (int)(dj - r * (uint32_t)nx);
        k = (int)(r - r0);
    } else {
        for (int e = threadIdx.x; e < tot; e += blockDim.x) {
            const int w = e & ((1 << lg) - 1), slot = e >> lg;
            const int k = slot / 9, nb = slot - 9 * k;
            uint2 val = make_uint2(0u, 0u);
            if (w < W) {
                const uint32_t r = r0 + (uint32_t)k;
                const uint32_t zq = fdiv(r, kp.my, kp.sy);
                int y = (int)(r - zq * (uint32_t)ny) + nb % 3 - 1, z = (int)zq + nb / 3 - 1;
                bool out = false;
                if (y < 0) { y += ny; out = !kp.per[1]; } else if (y >= ny) { y -= ny; out = !kp.per[1]; }
                if (z < 0) { z += nz; out |= !kp.per[2]; } else if (z >= nz) { z -= nz; out |= !kp.per[2]; }
                if (!out) val = rowmeta[((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)W + (uint32_t)w];
            }
            val.y += kSh;
            smeta[((k << lg) + w) * 9 + nb] = val;
        }
        __syncthreads();
        // Position of every node of this workgroup: one thread per bit of the
        // staged own-row words; a fluid bit writes (row, x) at its node's slot
        // (its rank in the word plus the word's first compact index), so no
        // thread searches for its own node.
        uint32_t* pos = (uint32_t*)(smeta + ((kp.rmax * 9) << lg));
        for (int t = threadIdx.x; t < ((R << lg) << 5); t += blockDim.x) {
            const int b = t & 31, wl = t >> 5;
            const int k = wl >> lg, w = wl & ((1 << lg) - 1);
            if (w >= W) continue;
            const uint2 me = smeta[wl * 9 + 4];
            if (!((me.x >> b) & 1u)) continue;
            const uint32_t slot = me.y + __popc(me.x & ((1u << b) - 1u)) - (own0 + kSh);
            if (slot < blockDim.x) pos[slot] = ((uint32_t)k << 16) | (uint32_t)(32 * w + b);
        }
        __syncthreads();
        own = own0 + threadIdx.x;
        if (own >= kp.nf) return;
        const uint32_t p = pos[threadIdx.x];
        k = (int)(p >> 16);
        x = (int)(p & 0xFFFFu);
        r = r0 + (uint32_t)k;
    }
    const uint32_t zq = fdiv(r, kp.my, kp.sy);
    const int y = (int)(r - zq * (uint32_t)ny), z = (int)zq;
    const uint32_t idx = r * (uint32_t)nx + (uint32_t)x;
    const uint32_t nd = SIMPLE ? ((uint32_t)FLUID << kClsShift) : nodew[own];
    const uint32_t cls = nd >> kClsShift;
    bool xm_o = false, xp_o = false, ym_o = false, yp_o = false, zm_o = false, zp_o = false;
    if (OPEN) {
        xm_o = !kp.per[0] && x == 0;  xp_o = !kp.per[0] && x == nx - 1;
        ym_o = !kp.per[1] && y == 0;  yp_o = !kp.per[1] && y == ny - 1;
        zm_o = !kp.per[2] && z == 0;  zp_o = !kp.per[2] && z == nz - 1;
    }
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    using raw = typename Io<L, S>::raw;
    constexpr bool kNarrow = sizeof(raw) == 2;
    T h[L::Q];
    raw hr[L::Q];   // 16-bit storage: the stored values as loaded
    const uint32_t oslot = kSoA ? own : tslot<L>(0, own);
    if constexpr (kNarrow) hr[0] = ldo(src, oslot);
    else h[0] = Io<L, S>::ld(src, oslot, 0);
    // per source column (x+1, x, x-1 for cx = -1, 0, 1): the LDS base of its
    // word's nine neighbour words, its bit, and the mask of the bits below it
    uint32_t cb[3], cbitm[3], clow[3], cpos[3];
    {
        const int sxs[3] = {xm, x, xp};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cb[c] = (((uint32_t)k << lg) + ((uint32_t)sxs[c] >> 5)) * 9u;
            cpos[c] = (uint32_t)sxs[c] & 31u;
            cbitm[c] = 1u << cpos[c];
            clow[c] = cbitm[c] - 1u;
        }
    }
    // a source outside a non-periodic face (resolved by the face condition)
    auto outside = [&](int cx, int cy, int cz) {
        return OPEN && ((cx > 0 && xm_o) || (cx < 0 && xp_o) || (cy > 0 && ym_o) || (cy < 0 && yp_o) ||
                        (cz > 0 && zm_o) || (cz < 0 && zp_o));
    };
    if constexpr (kShifted) {
        // Periodic lattice whose walls are exactly the links into solid nodes,
        // 16-bit storage (bound by instruction issue): all LDS words and counts
        // first, then per direction the byte offset of the source slot, the
        // fluid bit of the source as an all-ones / zero mask (one bit-field
        // extract) and a bitwise select against the node's own opposite slot
        // (half-way bounce-back).  The tile slot of compact index j is
        // (j >> 6) 64 (Q - 1) + j + 64 i: one shift and one 24-bit multiply-add.
        uint32_t cnt[L::Q], word[L::Q];
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int c = 1 - L::cx(i);
            const uint2 e = smeta[cb[c] + (uint32_t)((1 - L::cz(i)) * 3 + (1 - L::cy(i)))];
            cnt[i] = e.y + __popc(e.x & clow[c]);
            word[i] = e.x;
        }
#pragma unroll
        for (int i = 1; i < L::Q; ++i) asm volatile("" : "+v"(cnt[i]));
        // counts and the own slot are taken kRankShift nodes (kRankShift / 64
        // whole tiles) further on, against a base that many tiles before the
        // buffer: the own-slot offsets of the opposite planes, down to 3 planes
        // below plane i, then stay non-negative, and the plane term 64 i of
        // every load is an immediate of the instruction
        constexpr uint32_t rb = (uint32_t)sizeof(raw);
        if constexpr (kSoA) {
            // SoA planes: the slot of compact index j in plane i is i qstride + j,
            // so a source's byte offset is its (shifted) count times 2, and plane
            // i's base (3 planes back, uniform) is a scalar base of the load
            static_assert(max_opp_gap<L>() <= 3, "opposite planes within 3 planes");
            const size_t qsb = (size_t)kp.qstride * rb;
            const char* const srcb = (const char*)src - 3 * qsb;
            const uint32_t obyte = (own + kShv) * rb;
#pragma unroll
            for (int i = 1; i < L::Q; ++i) {
                const int c = 1 - L::cx(i), o = L::opp(i);
                const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)word[i], cpos[c], 1);
                const uint32_t there = cnt[i] * rb;
                const uint32_t self = obyte + (uint32_t)(o - i) * (uint32_t)qsb;
                const uint32_t off = (m & there) | (~m & self);
                hr[i] = *(const raw*)(srcb + (size_t)i * qsb + off);
            }
        } else {
            constexpr uint32_t kShBytes = (kSh / 64u) * 64u * L::Q * rb;
            static_assert(kSh % 64u == 0 && kSh * L::Q * rb >= 64u * 3u * rb, "shift covers the plane offsets");
            const char* const srcb = (const char*)src - kShBytes;
            const uint32_t obyte = oslot * rb + kShBytes;
#pragma unroll
            for (int i = 1; i < L::Q; ++i) {
                const int c = 1 - L::cx(i), o = L::opp(i);
                const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)word[i], cpos[c], 1);
                const uint32_t there = (__umul24(cnt[i] >> 6, 64u * (L::Q - 1)) + cnt[i]) * rb;
                const uint32_t self = obyte + (uint32_t)(64 * (int)rb * (o - i));
                const uint32_t off = (m & there) | (~m & self);
                hr[i] = ((const raw*)(srcb + off))[64 * i];
            }
        }
    } else if constexpr (kNarrow) {
        // 16-bit storage is bound by instruction issue, not by HBM: first every
        // direction's LDS word and count, with no branch around the reads so
        // they are all in flight together (the count is the source's compact
        // index, meaningful when it is fluid), then the slot of every direction
        // as a select.  Measured at 256^3: 7 % faster than the per-direction
        // form below for FP16, 2.5 % slower for FP32.
        uint32_t cnt[L::Q], word[L::Q];
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int c = 1 - L::cx(i);
            const uint2 e = smeta[cb[c] + (uint32_t)((1 - L::cz(i)) * 3 + (1 - L::cy(i)))];
            cnt[i] = e.y + __popc(e.x & clow[c]);
            word[i] = e.x;
        }
        // keeps the counts ahead of the selects below: otherwise the compiler
        // sinks each read and count into a branch of its own direction, and
        // every direction then waits for its own LDS round trip
#pragma unroll
        for (int i = 1; i < L::Q; ++i) asm volatile("" : "+v"(cnt[i]));
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int cx = L::cx(i), o = L::opp(i);
            const bool out = outside(cx, L::cy(i), L::cz(i));
            const bool solid = !out && !(word[i] & cbitm[1 - cx]);
            const bool wall = ((nd >> o) & 1u) || solid;
            // a wall reflects the node's own opposite population (half-way) or
            // keeps population i for the on-node closure (SIMPLE lattices have
            // half-way walls only); a source outside a non-periodic face reads
            // the node's own population i, replaced by the face condition later
            const uint32_t wsl = SIMPLE || kp.wall == 0 ? 64u * (uint32_t)o : 64u * (uint32_t)i;
            const uint32_t self = oslot + (wall ? wsl : 64u * (uint32_t)i);
            // both slots computed in straight-line code, so the choice is a
            // select rather than a branch per direction
            const uint32_t there = tslot<L>(i, cnt[i]);
            const bool mine = wall | out;
            hr[i] = ldo(src, mine ? self : there);
        }
    } else {
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i), o = L::opp(i);
            const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
            const bool out = outside(cx, cy, cz);
            uint32_t j = kNone;
            if (!out) {
                const uint2 e = smeta[(((uint32_t)k << lg) + ((uint32_t)sx >> 5)) * 9u +
                                      (uint32_t)((1 - cz) * 3 + (1 - cy))];
                const uint32_t bit = (uint32_t)sx & 31u;
                j = ((e.x >> bit) & 1u) ? e.y + __popc(e.x & ((1u << bit) - 1u)) : kSolidSrc;
            }
            const bool wall = ((nd >> o) & 1u) || j == kSolidSrc;
            int pl = i;
            if (wall || j == kNone) {
                j = own;
                pl = wall && kp.wall == 0 ? o : i;
            }
            h[i] = Io<L, S>::ld(src, tslot<L>(pl, j), i);
        }
    }
    if constexpr (kNarrow && SIMPLE && !OPEN && C != 2) {
        // Plain fluid, half-way walls, BGK/TRT: the update runs on the stored
        // values x_i = h_i / a_i, so neither the loads nor the stores scale:
        // the moments weigh x_i by a_i, and the pair update (a_i = a_opp(i))
        // divided by a_i is the same update with w_i / a_i in place of w_i.
        T xs[L::Q];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) xs[i] = Io<L, S>::x(hr[i]);
        T drho, rho, u[3];
        moments<L, T>(kp, cls, xs, drho, rho, u, [](int i) { return Io<L, S>::a(i); });
        if (write_macros) {
            fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
            fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
        }
        T out[L::Q];
        collide_pairs<L, T, C>(kp, xs, drho, rho, u, out, [](int) { return Io<L, S>::kw; });
        store_node<L, S, kSoA ? 2 : 1, true>(dst, kp, own, out);
        return;
    } else if constexpr (kNarrow) {
#pragma unroll
        for (int i = 0; i < L::Q; ++i) h[i] = Io<L, S>::cv(hr[i], i);
    }
    // 2D only when open: the host keeps extrapolating outlets and every 3D
    // face or wet-node wall on k_step
    if (OPEN && L::Q == 9 && cls == VEL_W) zou_inlet<L, T>(kp, fl, idx, nd, h);
    if (OPEN && L::Q == 9 && cls == VEL_N) zou_lid<L, T>(kp, fl, idx, nd, h);
    if (kp.wall == 1 && (nd & kBBMask)) onnode_walls<L, T>(nd, h);
    if (OPEN && L::Q == 9 && cls == PRES_E) zou_outlet<L, T>(kp, fl, idx, h);
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    if (OPEN) frozen_macros<T>(kp, fl, cls, idx, drho, rho, u);
    if (write_macros) {
        fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
        fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
    }
    T out[L::Q];
    collide<L, T, C>(kp, h, drho, rho, u, out);
    store_node<L, S, kSoA ? 2 : 1>(dst, kp, own, out);
}

// NPT nodes per thread (own0 + t + h blockDim, h < NPT) for the 16-bit
// periodic variant of the rank step with BGK/TRT: every node's loads are
// issued before any node's arithmetic, so a wave keeps NPT times the gathers
// in flight.  Per node the same staging, addresses, arithmetic, rounding and
// stores as k_stepr<L, S, C, false, true>.
template <class L, int S, int C, int NPT>
__global__ __launch_bounds__(512) void k_stepr2(const typename Io<L, S>::raw* __restrict__ src,
                                               typename Io<L, S>::raw* __restrict__ dst,
                                               const uint2* __restrict__ rowmeta,
                                               const uint2* __restrict__ blk, KParams kp, Fields fl,
                                               int write_macros) {
    using T = typename Io<L, S>::T;
    using raw = typename Io<L, S>::raw;
    static_assert(sizeof(raw) == 2 && C != 2, "16-bit storage, BGK or TRT");
    extern __shared__ uint2 smeta[];
    const int W = kp.words, lg = kp.lgw, nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const uint32_t bid = xcd_block(blockIdx.x, gridDim.x, kp.xcd_shift, kp.xchunk_shift);
    const uint2 bi = blk[bid];
    const uint32_t r0 = bi.x;
    const int R = (int)bi.y;
    constexpr uint32_t kSh = kRankShift;
    const uint32_t nt = blockDim.x;
    const uint32_t own0 = bid * (uint32_t)NPT * nt;
    uint32_t dj[NPT];
#pragma unroll
    for (int h = 0; h < NPT; ++h) dj[h] = 0u;
#pragma unroll
    for (int h = 0; h < NPT; ++h)
        if (own0 + (uint32_t)h * nt + threadIdx.x < kp.nf) dj[h] = kp.didx[own0 + (uint32_t)h * nt + threadIdx.x];
    const int tot = (R * 9) << lg;
    uint32_t* rowoff = (uint32_t*)(smeta + ((kp.rmax * 9) << lg)) + kp.rblock;
    for (int t = threadIdx.x; t < R * 9; t += nt) {
        const int kk = t / 9, nb = t - 9 * kk;
        const uint32_t rr = r0 + (uint32_t)kk;
        const uint32_t zq = fdiv(rr, kp.my, kp.sy);
        int y = (int)(rr - zq * (uint32_t)ny) + nb % 3 - 1, z = (int)zq + nb / 3 - 1;
        y = y < 0 ? y + ny : (y >= ny ? y - ny : y);
        z = z < 0 ? z + nz : (z >= nz ? z - nz : z);
        rowoff[t] = ((uint32_t)z * (uint32_t)ny + (uint32_t)y) * (uint32_t)W;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < tot; e += nt) {
        const int w = e & ((1 << lg) - 1), slot = e >> lg;
        const int kk = slot / 9, nb = slot - 9 * kk;
        uint2 val = make_uint2(0u, 0u);
        if (w < W) val = rowmeta[rowoff[slot] + (uint32_t)w];
        val.y += kSh;
        smeta[((kk << lg) + w) * 9 + nb] = val;
    }
    __syncthreads();
    constexpr uint32_t rb = (uint32_t)sizeof(raw);
    constexpr uint32_t kShBytes = (kSh / 64u) * 64u * L::Q * rb;
    const char* const srcb = (const char*)src - kShBytes;
    raw hr[NPT][L::Q];
#pragma unroll
    for (int h = 0; h < NPT; ++h) {
        const uint32_t own = own0 + (uint32_t)h * nt + threadIdx.x;
        if (own >= kp.nf) continue;
        const uint32_t r = fdiv(dj[h], kp.mx, kp.sx);
        const int x = (int)(dj[h] - r * (uint32_t)nx);
        const int k = (int)(r - r0);
        const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
        const uint32_t oslot = tslot<L>(0, own);
        hr[h][0] = ldo(src, oslot);
        uint32_t cb[3], clow[3], cpos[3];
        const int sxs[3] = {xm, x, xp};
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            cb[c] = (((uint32_t)k << lg) + ((uint32_t)sxs[c] >> 5)) * 9u;
            cpos[c] = (uint32_t)sxs[c] & 31u;
            clow[c] = (1u << cpos[c]) - 1u;
        }
        uint32_t cnt[L::Q], word[L::Q];
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int c = 1 - L::cx(i);
            const uint2 e = smeta[cb[c] + (uint32_t)((1 - L::cz(i)) * 3 + (1 - L::cy(i)))];
            cnt[i] = e.y + __popc(e.x & clow[c]);
            word[i] = e.x;
        }
#pragma unroll
        for (int i = 1; i < L::Q; ++i) asm volatile("" : "+v"(cnt[i]));
        const uint32_t obyte = oslot * rb + kShBytes;
#pragma unroll
        for (int i = 1; i < L::Q; ++i) {
            const int c = 1 - L::cx(i), o = L::opp(i);
            const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int)word[i], cpos[c], 1);
            const uint32_t there = (__umul24(cnt[i] >> 6, 64u * (L::Q - 1)) + cnt[i]) * rb;
            const uint32_t self = obyte + (uint32_t)(64 * (int)rb * (o - i));
            const uint32_t off = (m & there) | (~m & self);
            hr[h][i] = ((const raw*)(srcb + off))[64 * i];
        }
    }
#pragma unroll
    for (int h = 0; h < NPT; ++h) {
        const uint32_t own = own0 + (uint32_t)h * nt + threadIdx.x;
        if (own >= kp.nf) continue;
        T xs[L::Q];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) xs[i] = Io<L, S>::x(hr[h][i]);
        T drho, rho, u[3];
        moments<L, T>(kp, (uint32_t)FLUID, xs, drho, rho, u, [](int i) { return Io<L, S>::a(i); });
        if (write_macros) {
            const uint32_t idx = dj[h];
            fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
            fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
        }
        T out[L::Q];
        collide_pairs<L, T, C>(kp, xs, drho, rho, u, out, [](int) { return Io<L, S>::kw; });
        store_node<L, S, 1, true>(dst, kp, own, out);
    }
}

// Post-boundary populations h of non-solid node idx (compact slot own) from
// the post-collision state in src: the pull, walls and open faces of the step.
template <class L, int S, bool OPEN>
__device__ inline void post_boundary(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp,
                                     const Fields& fl, const uint32_t* __restrict__ node,
                                     const uint32_t* __restrict__ cidx, uint32_t own, uint32_t nd, int x, int y,
                                     int z, size_t idx, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    const uint32_t cls = nd >> kClsShift;
    inflow_state<L, S, OPEN>(src, kp, fl, cidx, own, nd, x, y, z, idx, h);
    if (OPEN && L::Q == 9 && cls == PRES_E) {
        zou_outlet<L, T>(kp, fl, idx, h);
        outlet_extrap<L, S, OPEN>(src, kp, fl, cidx, node, false, x, y, z, idx, h);
    }
    if (OPEN && L::Q == 19) {
        if (is_outlet3(cls)) outlet3d<L, S, OPEN>(src, kp, fl, cidx, node, false, x, y, z, idx, cls, h);
        sym3d<L, T>(cls, h);
    }
}

// Post-boundary populations (dense, full, FP64) and moments of the state in src;
// with fl.fout null only the moments are written.
template <class L, int S, bool OPEN>
__global__ __launch_bounds__(256) void k_recon(const typename Io<L, S>::raw* __restrict__ src,
                                               const uint32_t* __restrict__ node,
                                               const uint32_t* __restrict__ cidx, KParams kp, Fields fl) {
    using T = typename Io<L, S>::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= kp.nx) return;
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    const size_t idx = ((size_t)z * kp.ny + y) * kp.nx + x;
    const uint32_t own = cidx[idx];
    if (own == kNone) {
        if (fl.fout) {
#pragma unroll
            for (int i = 0; i < L::Q; ++i) fl.fout[i * n + idx] = 0.0;
        }
        return;
    }
    const uint32_t nd = node[idx];
    const uint32_t cls = nd >> kClsShift;
    T h[L::Q];
    post_boundary<L, S, OPEN>(src, kp, fl, node, cidx, own, nd, x, y, z, idx, h);
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    if (OPEN) frozen_macros<T>(kp, fl, cls, idx, drho, rho, u);
    fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
    fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
    if (!fl.fout) return;   // moments only
#pragma unroll
    for (int i = 0; i < L::Q; ++i) fl.fout[i * n + idx] = (double)h[i] + (L::w(i) - (kp.mref ? mref_off<L>(i) : 0.0));
}

// First post-collision state from an initial f and macroscopic fields.
template <class L, int S, int C>
__global__ __launch_bounds__(256) void k_init(typename Io<L, S>::raw* __restrict__ dst,
                                              const uint32_t* __restrict__ node,
                                              const uint32_t* __restrict__ cidx, KParams kp,
                                              Fields fl, int mode) {
    using T = typename Io<L, S>::T;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    if (x >= kp.nx) return;
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    const size_t idx = ((size_t)z * kp.ny + y) * kp.nx + x;
    const uint32_t own = cidx[idx];
    if (own == kNone) return;
    const uint32_t nd = node[idx];
    const uint32_t cls = nd >> kClsShift;
    T h[L::Q], drho, rho, u[3];
    if (mode == INIT_FROM_F) {
#pragma unroll
        for (int i = 0; i < L::Q; ++i) h[i] = (T)(fl.fout[i * n + idx] - L::w(i));
        moments<L, T>(kp, cls, h, drho, rho, u);
        frozen_macros<T>(kp, fl, cls, idx, drho, rho, u);
        fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
        fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
    } else {
        const double r = fl.rho[idx];
        drho = (T)(r - 1.0);
        rho = (T)r;
        u[0] = (T)fl.u[idx]; u[1] = (T)fl.v[idx]; u[2] = (T)fl.w[idx];
        const T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            if (mode == INIT_ZERO) {
                h[i] = -(T)L::w(i);
            } else {
                const T cu = (T)L::cx(i) * u[0] + (T)L::cy(i) * u[1] + (T)L::cz(i) * u[2];
                h[i] = (T)L::w(i) * (drho + rho * ((T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu));
            }
        }
    }
    if (kp.mref) mref_shift<L, T>(h, 1);   // D3Q19 MRT stores f - (w - r)
    T out[L::Q];
    collide<L, T, C>(kp, h, drho, rho, u, out);
#pragma unroll
    for (int i = 0; i < L::Q; ++i) Io<L, S>::st(dst, paddr<L>(kp, i, own), out[i], i);
}

// FP64 reduction: per-block partials (wave shuffles, then LDS), then one block
// sums the partials in a fixed order, so results are deterministic.
constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;
int reduce_blocks() { return kRedBlocks; }

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// FP64 reductions over non-solid nodes, per block: sums of |du|^2, |u|^2, rho
// and the count of non-finite values, then the largest |du|, |dv|, |dw|,
// |drho| against the previous reduction's fields (the reference's MacroDiff
// residual, S/Iterate3D.cu:696-756, with a true maximum where
// S/GpuSum.cu:152-187 adds the two loads); the previous fields are updated.
__global__ __launch_bounds__(kRedThreads) void k_reduce_partial(const uint32_t* __restrict__ node,
                                                                 Fields fl, double* uprev, double* vprev,
                                                                 double* wprev, double* rprev, size_t n,
                                                                 double* partials) {
    double a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if ((node[i] >> kClsShift) == SOLID) continue;
        const double u = fl.u[i], v = fl.v[i], w = fl.w[i], r = fl.rho[i];
        const double du = u - uprev[i], dv = v - vprev[i], dw = w - wprev[i], dr = r - rprev[i];
        const double uu = u * u + v * v + w * w;
        a[0] += du * du + dv * dv + dw * dw;
        a[1] += uu;
        a[2] += r;
        if (!(uu == uu) || !(r == r) || isinf(uu) || isinf(r)) a[3] += 1.0;
        a[4] = fmax(a[4], fabs(du)); a[5] = fmax(a[5], fabs(dv));
        a[6] = fmax(a[6], fabs(dw)); a[7] = fmax(a[7], fabs(dr));
        uprev[i] = u; vprev[i] = v; wprev[i] = w; rprev[i] = r;
    }
    __shared__ double sh[8][kRedThreads / 64];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = k < 4 ? wave_sum(a[k]) : wave_max(a[k]);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 8; ++k) sh[k][wv] = a[k];
    __syncthreads();
    if (threadIdx.x < 8) {
        double s = 0;
        for (int k = 0; k < kRedThreads / 64; ++k)
            s = threadIdx.x < 4 ? s + sh[threadIdx.x][k] : fmax(s, sh[threadIdx.x][k]);
        partials[(size_t)blockIdx.x * 8 + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partials, int nb, double* out8) {
    __shared__ double sh[8][256];
    for (int k = 0; k < 8; ++k) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x)
            s = k < 4 ? s + partials[(size_t)b * 8 + k] : fmax(s, partials[(size_t)b * 8 + k]);
        sh[k][threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        double s = 0;
        for (int t = 0; t < 256; ++t) s = threadIdx.x < 4 ? s + sh[threadIdx.x][t] : fmax(s, sh[threadIdx.x][t]);
        out8[threadIdx.x] = s;
    }
}

// Sum over non-solid nodes and populations of (f - fColl)^2 per block, with f
// the post-boundary populations (dense FP64, from k_recon) and fColl the
// post-collision populations they were streamed from (the reference's L2
// residual, S/ComputeResiduals.cu:13-27).
template <class L, int S>
__global__ __launch_bounds__(kRedThreads) void k_fdiff2(const typename Io<L, S>::raw* __restrict__ pc,
                                                        const uint32_t* __restrict__ cidx, KParams kp,
                                                        const double* __restrict__ fout, size_t n,
                                                        double* partials) {
    using T = typename Io<L, S>::T;
    double acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t own = cidx[i];
        if (own == kNone) continue;
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const T hc = Io<L, S>::ld(pc, paddr<L>(kp, q, own), q);
            const double d = fout[q * n + i] - ((double)hc + L::w(q));
            acc += d * d;
        }
    }
    acc = wave_sum(acc);
    __shared__ double sh[kRedThreads / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < kRedThreads / 64; ++k) s += sh[k];
        partials[blockIdx.x] = s;
    }
}

// The same sum with f rebuilt in registers (post_boundary) instead of read
// from a dense FP64 copy: no Q*N staging array.  Same blocks, threads and
// summation order as k_fdiff2, so the two agree to the last bit when the
// rebuilt populations do.
template <class L, int S, bool OPEN>
__global__ __launch_bounds__(kRedThreads) void k_l2_fused(const typename Io<L, S>::raw* __restrict__ pc,
                                                          const uint32_t* __restrict__ node,
                                                          const uint32_t* __restrict__ cidx, KParams kp,
                                                          Fields fl, size_t n, double* partials) {
    using T = typename Io<L, S>::T;
    double acc = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t own = cidx[i];
        if (own == kNone) continue;
        const size_t row = i / (size_t)kp.nx;
        const int x = (int)(i - row * (size_t)kp.nx);
        const int y = (int)(row % (size_t)kp.ny), z = (int)(row / (size_t)kp.ny);
        T h[L::Q];
        post_boundary<L, S, OPEN>(pc, kp, fl, node, cidx, own, node[i], x, y, z, i, h);
#pragma unroll
        for (int q = 0; q < L::Q; ++q) {
            const T hc = Io<L, S>::ld(pc, paddr<L>(kp, q, own), q);
            const double d = ((double)h[q] + L::w(q)) - ((double)hc + L::w(q));
            acc += d * d;
        }
    }
    acc = wave_sum(acc);
    __shared__ double sh[kRedThreads / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        for (int k = 0; k < kRedThreads / 64; ++k) s += sh[k];
        partials[blockIdx.x] = s;
    }
}

// FdRelDiff residual: max over non-solid links of |f - f_prev| / |f1|
// (gpu_NewResidual + gpu_max_h, S/ComputeResiduals.cu:28-36, :55-63), f, f1,
// f_prev dense FP64 (q planes of n); a NaN ratio wins the maximum.
__global__ __launch_bounds__(kRedThreads) void k_fdrel(const double* __restrict__ f, const double* __restrict__ f1,
                                                       const double* __restrict__ fp, const uint32_t* __restrict__ cidx,
                                                       size_t n, int q, double* partials) {
    double m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (cidx[i] == kNone) continue;
        for (int k = 0; k < q; ++k) {
            const size_t j = (size_t)k * n + i;
            const double r = fabs(f[j] - fp[j]) / fabs(f1[j]);
            m = (r > m || r != r) ? r : m;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = (o > m || o != o) ? o : m;
    }
    __shared__ double sh[kRedThreads / 64];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int k = 1; k < kRedThreads / 64; ++k) t = (sh[k] > t || sh[k] != sh[k]) ? sh[k] : t;
        partials[blockIdx.x] = t;
    }
}

__global__ __launch_bounds__(256) void k_max_final(const double* __restrict__ partials, int nb, double* out) {
    if (threadIdx.x != 0) return;
    double t = 0;
    for (int b = 0; b < nb; ++b) t = (partials[b] > t || partials[b] != partials[b]) ? partials[b] : t;
    out[0] = t;
}

__global__ __launch_bounds__(256) void k_sum_final(const double* __restrict__ partials, int nb, double* out) {
    __shared__ double sh[256];
    double s = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < 256; ++k) t += sh[k];
        out[0] = t;
    }
}

// ---------------------------------------------------------------------------
// launchers
static dim3 node_block(const KParams& kp) {
    int bx = kp.nx >= 256 ? 256 : ((kp.nx + 63) / 64) * 64;
    return dim3(bx, 1, 1);
}
static dim3 node_grid(const KParams& kp) {
    dim3 b = node_block(kp);
    return dim3((kp.nx + b.x - 1) / b.x, kp.ny, kp.nz);
}

template <class L, int S, int C, bool OPEN>
static hipError_t step_t(const KParams& kp, const void* src, void* dst, const uint32_t* node, const uint32_t* cidx,
                         Fields fl, bool wm, hipStream_t st) {
    using raw = typename Io<L, S>::raw;
    if (kp.nf == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((kp.nf + 255) / 256);
    if (kp.rank) {
        // staged row words, node positions, neighbour-row offsets
        const size_t lds = (((size_t)kp.rmax * 9) << kp.lgw) * sizeof(uint2) + (size_t)kp.rblock * sizeof(uint32_t) +
                           (size_t)kp.rmax * 9 * sizeof(uint32_t);
        const unsigned rb = (unsigned)((kp.nf + kp.rblock - 1) / kp.rblock);
        bool soa = false;
        if constexpr (sizeof(raw) == 2 && !OPEN) soa = kp.simple && !kp.tile;
        if (soa) {
            if constexpr (sizeof(raw) == 2 && !OPEN)
                hipLaunchKernelGGL((k_stepr<L, S, C, OPEN, true, true>), dim3(rb), dim3(kp.rblock), lds, st,
                                   (const raw*)src, (raw*)dst, kp.rowmeta, kp.blk, kp.nodew, kp, fl, wm ? 1 : 0);
        } else if (kp.simple) {
            if constexpr (sizeof(raw) == 2 && !OPEN && C != 2) {
                if (kp.npt == 2 && kp.tile) {
                    hipLaunchKernelGGL((k_stepr2<L, S, C, 2>), dim3(rb), dim3(kp.rblock / 2), lds, st,
                                       (const raw*)src, (raw*)dst, kp.rowmeta, kp.blk, kp, fl, wm ? 1 : 0);
                    return hipGetLastError();
                }
                if (kp.npt == 4 && kp.tile) {
                    hipLaunchKernelGGL((k_stepr2<L, S, C, 4>), dim3(rb), dim3(kp.rblock / 4), lds, st,
                                       (const raw*)src, (raw*)dst, kp.rowmeta, kp.blk, kp, fl, wm ? 1 : 0);
                    return hipGetLastError();
                }
            }
            hipLaunchKernelGGL((k_stepr<L, S, C, OPEN, true>), dim3(rb), dim3(kp.rblock), lds, st, (const raw*)src,
                               (raw*)dst, kp.rowmeta, kp.blk, kp.nodew, kp, fl, wm ? 1 : 0);
        }
        else
            hipLaunchKernelGGL((k_stepr<L, S, C, OPEN, false>), dim3(rb), dim3(kp.rblock), lds, st, (const raw*)src,
                               (raw*)dst, kp.rowmeta, kp.blk, kp.nodew, kp, fl, wm ? 1 : 0);
    } else {
        hipLaunchKernelGGL((k_step<L, S, C, OPEN>), dim3(blocks), dim3(256), 0, st, (const raw*)src, (raw*)dst,
                           kp.didx, kp.nodew, cidx, kp, fl, wm ? 1 : 0);
    }
    return hipGetLastError();
}
template <class L, int S, int C>
static hipError_t step_o(bool open, const KParams& kp, const void* src, void* dst, const uint32_t* node, const uint32_t* cidx,
                         Fields fl, bool wm, hipStream_t st) {
    return open ? step_t<L, S, C, true>(kp, src, dst, node, cidx, fl, wm, st)
                : step_t<L, S, C, false>(kp, src, dst, node, cidx, fl, wm, st);
}
template <class L, int S>
static hipError_t step_c(int c, bool open, const KParams& kp, const void* src, void* dst,
                         const uint32_t* node, const uint32_t* cidx, Fields fl, bool wm, hipStream_t st) {
    switch (c) {
    case 0: return step_o<L, S, 0>(open, kp, src, dst, node, cidx, fl, wm, st);
    case 1: return step_o<L, S, 1>(open, kp, src, dst, node, cidx, fl, wm, st);
    default: return step_o<L, S, 2>(open, kp, src, dst, node, cidx, fl, wm, st);
    }
}
template <class L>
static hipError_t step_s(int s, int c, bool open, const KParams& kp, const void* src, void* dst,
                         const uint32_t* node, const uint32_t* cidx, Fields fl, bool wm, hipStream_t st) {
    switch (s) {
    case 0: return step_c<L, 0>(c, open, kp, src, dst, node, cidx, fl, wm, st);
    case 1: return step_c<L, 1>(c, open, kp, src, dst, node, cidx, fl, wm, st);
    case 2: return step_c<L, 2>(c, open, kp, src, dst, node, cidx, fl, wm, st);
    default: return step_c<L, 3>(c, open, kp, src, dst, node, cidx, fl, wm, st);
    }
}
hipError_t launch_step(int q, int storage, int collision, bool open, const KParams& kp, const void* src,
                       void* dst, const uint32_t* node, const uint32_t* cidx, Fields fl, bool wm, hipStream_t st) {
    return q == 9 ? step_s<D2Q9>(storage, collision, open, kp, src, dst, node, cidx, fl, wm, st)
                  : step_s<D3Q19>(storage, collision, open, kp, src, dst, node, cidx, fl, wm, st);
}

template <class L, int S>
static hipError_t recon_t(bool open, const KParams& kp, const void* src, const uint32_t* node, const uint32_t* cidx, Fields fl,
                          hipStream_t st) {
    using raw = typename Io<L, S>::raw;
    if (open)
        hipLaunchKernelGGL((k_recon<L, S, true>), node_grid(kp), node_block(kp), 0, st, (const raw*)src, node, cidx, kp, fl);
    else
        hipLaunchKernelGGL((k_recon<L, S, false>), node_grid(kp), node_block(kp), 0, st, (const raw*)src, node, cidx, kp, fl);
    return hipGetLastError();
}
template <class L>
static hipError_t recon_s(int s, bool open, const KParams& kp, const void* src, const uint32_t* node,
                          const uint32_t* cidx, Fields fl, hipStream_t st) {
    switch (s) {
    case 0: return recon_t<L, 0>(open, kp, src, node, cidx, fl, st);
    case 1: return recon_t<L, 1>(open, kp, src, node, cidx, fl, st);
    case 2: return recon_t<L, 2>(open, kp, src, node, cidx, fl, st);
    default: return recon_t<L, 3>(open, kp, src, node, cidx, fl, st);
    }
}
hipError_t launch_recon(int q, int storage, bool open, const KParams& kp, const void* src, const uint32_t* node,
                        const uint32_t* cidx, Fields fl, hipStream_t st) {
    return q == 9 ? recon_s<D2Q9>(storage, open, kp, src, node, cidx, fl, st)
                  : recon_s<D3Q19>(storage, open, kp, src, node, cidx, fl, st);
}

template <class L, int S, int C>
static hipError_t init_t(int mode, const KParams& kp, void* dst, const uint32_t* node, const uint32_t* cidx, Fields fl, hipStream_t st) {
    using raw = typename Io<L, S>::raw;
    hipLaunchKernelGGL((k_init<L, S, C>), node_grid(kp), node_block(kp), 0, st, (raw*)dst, node, cidx, kp, fl, mode);
    return hipGetLastError();
}
template <class L, int S>
static hipError_t init_c(int c, int mode, const KParams& kp, void* dst, const uint32_t* node, const uint32_t* cidx, Fields fl,
                         hipStream_t st) {
    switch (c) {
    case 0: return init_t<L, S, 0>(mode, kp, dst, node, cidx, fl, st);
    case 1: return init_t<L, S, 1>(mode, kp, dst, node, cidx, fl, st);
    default: return init_t<L, S, 2>(mode, kp, dst, node, cidx, fl, st);
    }
}
template <class L>
static hipError_t init_s(int s, int c, int mode, const KParams& kp, void* dst, const uint32_t* node, const uint32_t* cidx, Fields fl,
                         hipStream_t st) {
    switch (s) {
    case 0: return init_c<L, 0>(c, mode, kp, dst, node, cidx, fl, st);
    case 1: return init_c<L, 1>(c, mode, kp, dst, node, cidx, fl, st);
    case 2: return init_c<L, 2>(c, mode, kp, dst, node, cidx, fl, st);
    default: return init_c<L, 3>(c, mode, kp, dst, node, cidx, fl, st);
    }
}
hipError_t launch_init(int q, int storage, int collision, int mode, const KParams& kp, void* dst,
                       const uint32_t* node, const uint32_t* cidx, Fields fl, hipStream_t st) {
    return q == 9 ? init_s<D2Q9>(storage, collision, mode, kp, dst, node, cidx, fl, st)
                  : init_s<D3Q19>(storage, collision, mode, kp, dst, node, cidx, fl, st);
}

hipError_t launch_reduce(const KParams& kp, const uint32_t* node, Fields fl, double* uprev, double* vprev,
                         double* wprev, double* rprev, double* partials, int nblocks, double* out8, hipStream_t st) {
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nblocks), dim3(kRedThreads), 0, st, node, fl, uprev, vprev, wprev,
                       rprev, n, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double*)partials, nblocks, out8);
    return hipGetLastError();
}

template <class L, int S>
static hipError_t fdiff2_t(const KParams& kp, const void* pc, const uint32_t* cidx, const double* fout,
                           double* partials, int nblocks, double* out, hipStream_t st) {
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    hipLaunchKernelGGL((k_fdiff2<L, S>), dim3(nblocks), dim3(kRedThreads), 0, st,
                       (const typename Io<L, S>::raw*)pc, cidx, kp, fout, n, partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, (const double*)partials, nblocks, out);
    return hipGetLastError();
}
template <class L>
static hipError_t fdiff2_s(int s, const KParams& kp, const void* pc, const uint32_t* cidx, const double* fout,
                           double* partials, int nblocks, double* out, hipStream_t st) {
    switch (s) {
    case 0: return fdiff2_t<L, 0>(kp, pc, cidx, fout, partials, nblocks, out, st);
    case 1: return fdiff2_t<L, 1>(kp, pc, cidx, fout, partials, nblocks, out, st);
    case 2: return fdiff2_t<L, 2>(kp, pc, cidx, fout, partials, nblocks, out, st);
    default: return fdiff2_t<L, 3>(kp, pc, cidx, fout, partials, nblocks, out, st);
    }
}
hipError_t launch_fdrel(int q, const double* f, const double* f1, const double* fprev, const uint32_t* cidx,
                        size_t n, double* partials, int nblocks, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_fdrel, dim3(nblocks), dim3(kRedThreads), 0, st, f, f1, fprev, cidx, n, q, partials);
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(64), 0, st, (const double*)partials, nblocks, out);
    return hipGetLastError();
}
hipError_t launch_fdiff2(int q, int storage, const KParams& kp, const void* pc, const uint32_t* cidx,
                         const double* fout, double* partials, int nblocks, double* out, hipStream_t st) {
    return q == 9 ? fdiff2_s<D2Q9>(storage, kp, pc, cidx, fout, partials, nblocks, out, st)
                  : fdiff2_s<D3Q19>(storage, kp, pc, cidx, fout, partials, nblocks, out, st);
}

// Scratch of the allocation-free L2 residual: one partial sum per reduction
// block and the final sum.  Its size does not depend on the lattice (the
// reduction strides over the nodes with a fixed grid).  Every device that runs
// the library holds its own copy until the process ends; calls that overlap
// on one device race on it.
static_assert(kRedBlocks + 1 <= kL2ScratchDoubles, "scratch holds the partials and the sum");
static __device__ __align__(256) double scratch[kL2ScratchDoubles];

hipError_t l2_scratch(double** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double* ptr[kMaxDevices] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        // the host-side name of the array is not a device address: take the
        // current device's copy once and use that pointer everywhere
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptr[dev] = (double*)p;
    }
    *out = ptr[dev];
    return hipSuccess;
}

template <class L, int S>
static hipError_t l2f_t(bool open, const KParams& kp, const void* pc, const uint32_t* node, const uint32_t* cidx,
                        Fields fl, double* partials, int nblocks, double* out, hipStream_t st) {
    using raw = typename Io<L, S>::raw;
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    if (open)
        hipLaunchKernelGGL((k_l2_fused<L, S, true>), dim3(nblocks), dim3(kRedThreads), 0, st, (const raw*)pc, node,
                           cidx, kp, fl, n, partials);
    else
        hipLaunchKernelGGL((k_l2_fused<L, S, false>), dim3(nblocks), dim3(kRedThreads), 0, st, (const raw*)pc, node,
                           cidx, kp, fl, n, partials);
    hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(256), 0, st, (const double*)partials, nblocks, out);
    return hipGetLastError();
}
template <class L>
static hipError_t l2f_s(int s, bool open, const KParams& kp, const void* pc, const uint32_t* node,
                        const uint32_t* cidx, Fields fl, double* partials, int nblocks, double* out, hipStream_t st) {
    switch (s) {
    case 0: return l2f_t<L, 0>(open, kp, pc, node, cidx, fl, partials, nblocks, out, st);
    case 1: return l2f_t<L, 1>(open, kp, pc, node, cidx, fl, partials, nblocks, out, st);
    case 2: return l2f_t<L, 2>(open, kp, pc, node, cidx, fl, partials, nblocks, out, st);
    default: return l2f_t<L, 3>(open, kp, pc, node, cidx, fl, partials, nblocks, out, st);
    }
}
hipError_t launch_l2_fused(int q, int storage, bool open, const KParams& kp, const void* pc, const uint32_t* node,
                           const uint32_t* cidx, Fields fl, double* partials, int nblocks, double* out,
                           hipStream_t st) {
    return q == 9 ? l2f_s<D2Q9>(storage, open, kp, pc, node, cidx, fl, partials, nblocks, out, st)
                  : l2f_s<D3Q19>(storage, open, kp, pc, node, cidx, fl, partials, nblocks, out, st);
}

}  // namespace lbm
