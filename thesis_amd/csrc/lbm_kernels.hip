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
#include "lbm_internal.h"

#include <hip/hip_runtime.h>
#include <cstdint>
#include <utility>

namespace lbm {

// ---------------------------------------------------------------------------
// lattices: reference numbering (S/GpuInit.cu:94-99, :822-842)
struct D2Q9 {
    static constexpr int Q = 9;
    __host__ __device__ static constexpr int cx(int i) { return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0); }
    __host__ __device__ static constexpr int cy(int i) { return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0); }
    __host__ __device__ static constexpr int cz(int) { return 0; }
    __host__ __device__ static constexpr int opp(int i) { return i == 0 ? 0 : (i <= 4 ? (i <= 2 ? i + 2 : i - 2) : (i <= 6 ? i + 2 : i - 2)); }
    __host__ __device__ static constexpr double w(int i) { return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
};

struct D3Q19 {
    static constexpr int Q = 19;
    __host__ __device__ static constexpr int cx(int i) {
        return (i == 1 || i == 7 || i == 9 || i == 11 || i == 13) ? 1 : ((i == 2 || i == 8 || i == 10 || i == 12 || i == 14) ? -1 : 0);
    }
    __host__ __device__ static constexpr int cy(int i) {
        return (i == 3 || i == 7 || i == 8 || i == 15 || i == 17) ? 1 : ((i == 4 || i == 9 || i == 10 || i == 16 || i == 18) ? -1 : 0);
    }
    __host__ __device__ static constexpr int cz(int i) {
        return (i == 5 || i == 11 || i == 12 || i == 15 || i == 16) ? 1 : ((i == 6 || i == 13 || i == 14 || i == 17 || i == 18) ? -1 : 0);
    }
    // pairs (1,2)(3,4)(5,6)(7,10)(8,9)(11,14)(12,13)(15,18)(16,17)
    __host__ __device__ static constexpr int opp(int i) {
        return i == 0 ? 0 : i <= 6 ? ((i & 1) ? i + 1 : i - 1)
             : i == 7 ? 10 : i == 10 ? 7 : i == 8 ? 9 : i == 9 ? 8
             : i == 11 ? 14 : i == 14 ? 11 : i == 12 ? 13 : i == 13 ? 12
             : i == 15 ? 18 : i == 18 ? 15 : i == 16 ? 17 : 16;
    }
    __host__ __device__ static constexpr double w(int i) { return i == 0 ? 1.0 / 3.0 : (i <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }
};

static_assert(D2Q9::opp(1) == 3 && D2Q9::opp(5) == 7 && D2Q9::opp(8) == 6, "D2Q9 opp");
static_assert(D3Q19::opp(7) == 10 && D3Q19::opp(16) == 17 && D3Q19::opp(2) == 1, "D3Q19 opp");

// MRT moment polynomials (the rows of M tabulated in S/CellFunctions.cu:15-36
// for D2Q9 and :73-114 for D3Q19).  Rows are orthogonal: Minv = M^T / |row|^2.
template <class L>
__host__ __device__ constexpr double mrow(int r, int i) {
    const double x = L::cx(i), y = L::cy(i), z = L::cz(i), c2 = x * x + y * y + z * z;
    if (L::Q == 9) {
        return r == 0 ? 1 : r == 1 ? 3 * c2 - 4 : r == 2 ? (9 * c2 * c2 - 21 * c2 + 8) / 2
             : r == 3 ? x : r == 4 ? (3 * c2 - 5) * x : r == 5 ? y : r == 6 ? (3 * c2 - 5) * y
             : r == 7 ? x * x - y * y : x * y;
    }
    return r == 0 ? 1 : r == 1 ? 19 * c2 - 30 : r == 2 ? (21 * c2 * c2 - 53 * c2 + 24) / 2
         : r == 3 ? x : r == 4 ? (5 * c2 - 9) * x : r == 5 ? y : r == 6 ? (5 * c2 - 9) * y
         : r == 7 ? z : r == 8 ? (5 * c2 - 9) * z : r == 9 ? 3 * x * x - c2
         : r == 10 ? (3 * c2 - 5) * (3 * x * x - c2) : r == 11 ? y * y - z * z
         : r == 12 ? (3 * c2 - 5) * (y * y - z * z) : r == 13 ? x * y : r == 14 ? y * z
         : r == 15 ? x * z : r == 16 ? (y * y - z * z) * x : r == 17 ? (z * z - x * x) * y
         : (x * x - y * y) * z;
}
template <class L>
__host__ __device__ constexpr double mnorm2(int r) {
    double s = 0;
    for (int i = 0; i < L::Q; ++i) s += mrow<L>(r, i) * mrow<L>(r, i);
    return s;
}
template <class L>
__host__ __device__ constexpr double mrest(int r) {   // (M w)_r: moments of the rest state
    double s = 0;
    for (int i = 0; i < L::Q; ++i) s += mrow<L>(r, i) * L::w(i);
    return s;
}
static_assert(mrest<D2Q9>(1) > -2.0000001 && mrest<D2Q9>(1) < -1.9999999, "D2Q9 rest e");
static_assert(mrest<D2Q9>(2) > 0.9999999 && mrest<D2Q9>(2) < 1.0000001, "D2Q9 rest eps");
static_assert(mrest<D3Q19>(1) > -11.0000001 && mrest<D3Q19>(1) < -10.9999999, "D3Q19 rest e");
static_assert(mrest<D3Q19>(2) > 2.9999999 && mrest<D3Q19>(2) < 3.0000001, "D3Q19 rest eps");

// Compile-time coefficients of M and Minv; the products below are expanded by
// fold expressions so every zero drops out and every +-1 becomes an add.
template <class L, int R, int I> struct Mc {
    static constexpr double m = mrow<L>(R, I);
    static constexpr double mi = mrow<L>(R, I) / mnorm2<L>(R);
};
template <class L, class T, int R, int... I>
__device__ inline T m_row_dot(const T* h, std::integer_sequence<int, I...>) {
    T acc = 0;
    ((Mc<L, R, I>::m != 0.0 ? (void)(acc += (T)Mc<L, R, I>::m * h[I]) : (void)0), ...);
    return acc;
}
template <class L, class T, int... R>
__device__ inline void m_apply(const T* h, T* m, std::integer_sequence<int, R...>) {
    ((m[R] = m_row_dot<L, T, R>(h, std::make_integer_sequence<int, L::Q>{})), ...);
}
template <class L, class T, int I, int... R>
__device__ inline T mi_col_dot(const T* m, std::integer_sequence<int, R...>) {
    T acc = 0;
    ((Mc<L, R, I>::mi != 0.0 ? (void)(acc += (T)Mc<L, R, I>::mi * m[R]) : (void)0), ...);
    return acc;
}
template <class L, class T, int... I>
__device__ inline void mi_apply(const T* m, T* out, std::integer_sequence<int, I...>) {
    ((out[I] = mi_col_dot<L, T, I>(m, std::make_integer_sequence<int, L::Q>{})), ...);
}

// ---------------------------------------------------------------------------
// storage: deviations h_i, raw in F64/F32; scaled by 1/(w_i a) in 16 bits
// FP16: stored value = h / w_i * 2^10 keeps typical deviations in the normal range.
template <class L> __device__ inline float f16_a(int i) { return (float)(L::w(i) / 1024.0); }
template <class L> __device__ inline float f16_ainv(int i) { return (float)(1024.0 / L::w(i)); }
template <class L> __device__ inline float bf16_a(int i) { return (float)L::w(i); }
template <class L> __device__ inline float bf16_ainv(int i) { return (float)(1.0 / L::w(i)); }

__device__ inline float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ inline uint16_t f32_to_bf16(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Loads and stores take 32-bit element offsets; the address is formed as
// base + 32-bit byte offset so it maps onto global_load/store with a scalar
// base register and one vector offset (population buffers stay below 4 GiB).
template <class R> __device__ inline R ldo(const R* base, uint32_t k) {
    return *(const R*)((const char*)base + k * (uint32_t)sizeof(R));
}
template <class R> __device__ inline void sto(R* base, uint32_t k, R v) {
    __builtin_nontemporal_store(v, (R*)((char*)base + k * (uint32_t)sizeof(R)));
}

template <class L, int S> struct Io;
template <class L> struct Io<L, 0> {
    using raw = double; using T = double;
    __device__ static T cv(raw r, int) { return r; }
    __device__ static raw pk(T v, int) { return v; }
    __device__ static T ld(const raw* p, uint32_t k, int) { return ldo(p, k); }
    __device__ static void st(raw* p, uint32_t k, T v, int) { sto(p, k, v); }
    __device__ static void st_sr(raw* p, uint32_t k, T v, int, uint32_t) { sto(p, k, v); }
};
template <class L> struct Io<L, 1> {
    using raw = float; using T = float;
    __device__ static T cv(raw r, int) { return r; }
    __device__ static raw pk(T v, int) { return v; }
    __device__ static T ld(const raw* p, uint32_t k, int) { return ldo(p, k); }
    __device__ static void st(raw* p, uint32_t k, T v, int) { sto(p, k, v); }
    __device__ static void st_sr(raw* p, uint32_t k, T v, int, uint32_t) { sto(p, k, v); }
};
template <class L> struct Io<L, 2> {
    using raw = uint16_t; using T = float;
    __device__ static T cv(raw r, int i) { return (float)__builtin_bit_cast(_Float16, r) * f16_a<L>(i); }
    __device__ static raw pk(T v, int i) { return __builtin_bit_cast(uint16_t, (_Float16)(v * f16_ainv<L>(i))); }
    __device__ static T ld(const raw* p, uint32_t k, int i) { return cv(ldo(p, k), i); }
    __device__ static void st(raw* p, uint32_t k, T v, int i) { sto(p, k, pk(v, i)); }
    // unbiased (stochastic) rounding with the random bits rnd
    __device__ static void st_sr(raw* p, uint32_t k, T v, int i, uint32_t rnd) {
        typedef short s2 __attribute__((ext_vector_type(2)));
        const s2 r = __builtin_amdgcn_cvt_sr_f16_f32(s2{0, 0}, v * f16_ainv<L>(i), rnd, false);
        sto(p, k, (uint16_t)r[0]);
    }
};
template <class L> struct Io<L, 3> {
    using raw = uint16_t; using T = float;
    __device__ static T cv(raw r, int i) { return bf16_to_f32(r) * bf16_a<L>(i); }
    __device__ static raw pk(T v, int i) { return f32_to_bf16(v * bf16_ainv<L>(i)); }
    __device__ static T ld(const raw* p, uint32_t k, int i) { return cv(ldo(p, k), i); }
    __device__ static void st(raw* p, uint32_t k, T v, int i) { sto(p, k, pk(v, i)); }
    __device__ static void st_sr(raw* p, uint32_t k, T v, int i, uint32_t rnd) {
        typedef short s2 __attribute__((ext_vector_type(2)));
        const s2 r = __builtin_amdgcn_cvt_sr_bf16_f32(s2{0, 0}, v * bf16_ainv<L>(i), rnd, false);
        sto(p, k, (uint16_t)r[0]);
    }
};

// ---------------------------------------------------------------------------
// node pipeline pieces (T = double for F64 storage, float otherwise)

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Stage 1 of the pull: compact indices of the source node of every direction
// (periodic wrap; kNone marks a source outside a non-periodic face).  These
// loads do not depend on the node's own word, so they issue together.
template <class L, bool OPEN>
__device__ inline void source_index(const uint32_t* __restrict__ cidx, const KParams& kp, int x, int y, int z,
                                    uint32_t* cn) {
    const int nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
    bool xm_o = false, xp_o = false, ym_o = false, yp_o = false, zm_o = false, zp_o = false;
    if (OPEN) {
        xm_o = !kp.per[0] && x == 0;  xp_o = !kp.per[0] && x == nx - 1;
        ym_o = !kp.per[1] && y == 0;  yp_o = !kp.per[1] && y == ny - 1;
        zm_o = !kp.per[2] && z == 0;  zp_o = !kp.per[2] && z == nz - 1;
    }
    // row bases of the three source rows in y and planes in z
    const uint32_t ux = (uint32_t)nx, plane = (uint32_t)nx * (uint32_t)ny;
    const uint32_t yb[3] = {(uint32_t)ym * ux, (uint32_t)y * ux, (uint32_t)yp * ux};     // index 1 - cy
    const uint32_t zb[3] = {(uint32_t)zm * plane, (uint32_t)z * plane, (uint32_t)zp * plane};
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
        bool out = false;
        if (OPEN)
            out = (cx > 0 && xm_o) || (cx < 0 && xp_o) || (cy > 0 && ym_o) || (cy < 0 && yp_o) ||
                  (cz > 0 && zm_o) || (cz < 0 && zp_o);
        cn[i] = out ? kNone : ldo(cidx, zb[1 - cz] + yb[1 - cy] + (uint32_t)sx);
    }
}

// Stage 2: the populations.  A link whose source lies beyond a wall takes the
// node's own value, P_opp(i) (half-way bounce-back) or P_i (on-node, fixed
// afterwards by onnode_walls) (S/GpuStream.cu:8-28, :160-194;
// S/GpuBoundaries.cu:400-490); a source outside a non-periodic face keeps the
// node's own P_i (the reference's "stream not allowed"), to be replaced by
// the open-boundary condition.
// Element offset of population i of compact node j: SoA planes
// P[i*qstride + j], or tiles of 64 nodes holding all Q planes
// P[(j/64)*64*Q + i*64 + j%64], so one wave's populations sit in one
// contiguous 64*Q-element block.
template <class L>
__device__ inline uint32_t paddr(const KParams& kp, int i, uint32_t j) {
    if (kp.tile) return (j >> 6) * (64u * L::Q) + (uint32_t)i * 64u + (j & 63u);
    return (uint32_t)i * kp.qstride32 + j;
}

template <class L, int S, bool OPEN>
__device__ inline void gather(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp,
                              uint32_t own, uint32_t nd, const uint32_t* cn, typename Io<L, S>::T* h) {
    h[0] = Io<L, S>::ld(src, paddr<L>(kp, 0, own), 0);
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int o = L::opp(i);
        uint32_t j = cn[i];
        int pl = i;
        if (((nd >> o) & 1u) || j == kNone) {
            j = own;
            pl = ((nd >> o) & 1u) && kp.wall == 0 ? o : i;
        }
        h[i] = Io<L, S>::ld(src, paddr<L>(kp, pl, j), i);
    }
}

// Open boundaries on D2Q9 faces, in full populations.
// Velocity inlet (west): Zou-He, or the reference's form (S/GpuBoundaries.cu:45-64).
// Pressure outlet (east): T/LBM/Tom/Lecture Script/LBMchannel.m:254-270.
template <class L, class T>
__device__ inline void zou_inlet(const KParams& kp, const Fields& fl, size_t idx, uint32_t nd, T* h) {
    if (L::Q != 9) return;
    if (kp.inlet == 1 && (nd & kBBMask)) return;       // reference skips corner nodes
    T f[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = h[i] + (T)L::w(i);
    const T u0 = (T)fl.ubc[idx], v0 = (T)fl.vbc[idx];
    const T r = (f[0] + f[2] + f[4] + (T)2 * (f[3] + f[6] + f[7])) / ((T)1 - u0);
    f[1] = f[3] + (T)(2.0 / 3.0) * r * u0;
    if (kp.inlet == 1) {
        f[5] = f[7] + r * u0 / (T)6;
        f[8] = f[6] + r * u0 / (T)6;
    } else {
        f[5] = f[7] - (T)0.5 * (f[2] - f[4]) + r * u0 / (T)6 + (T)0.5 * r * v0;
        f[8] = f[6] + (T)0.5 * (f[2] - f[4]) + r * u0 / (T)6 - (T)0.5 * r * v0;
    }
    h[1] = f[1] - (T)L::w(1); h[5] = f[5] - (T)L::w(5); h[8] = f[8] - (T)L::w(8);
}
template <class L, class T>
__device__ inline void zou_outlet(const KParams& kp, T* h) {
    if (L::Q != 9) return;
    T f[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = h[i] + (T)L::w(i);
    const T ro = (T)kp.rho_out;
    const T ux = (T)-1 + (f[0] + f[2] + f[4] + (T)2 * (f[1] + f[5] + f[8])) / ro;
    f[3] = f[1] - (T)(2.0 / 3.0) * ro * ux;
    f[7] = f[5] + (T)0.5 * (f[2] - f[4]) - ro * ux / (T)6;
    f[6] = f[8] - (T)0.5 * (f[2] - f[4]) - ro * ux / (T)6;
    h[3] = f[3] - (T)L::w(3); h[6] = f[6] - (T)L::w(6); h[7] = f[7] - (T)L::w(7);
}
template <class L, class T>
__device__ inline void onnode_walls(uint32_t nd, T* h) {
#pragma unroll
    for (int d = 1; d < L::Q; ++d)
        if ((nd >> d) & 1u) h[L::opp(d)] = h[d];
}

__device__ inline double recip(double x) { return 1.0 / x; }
__device__ inline float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Moments of h and the stored-velocity semantics of the force model
// (S/GpuUpdateMacro.cu:6-41, :137-213).
template <class L, class T>
__device__ inline void moments(const KParams& kp, uint32_t cls, const T* h, T& drho, T& rho, T* u) {
    T d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        d += h[i];
        if (L::cx(i)) jx += L::cx(i) > 0 ? h[i] : -h[i];
        if (L::cy(i)) jy += L::cy(i) > 0 ? h[i] : -h[i];
        if (L::cz(i)) jz += L::cz(i) > 0 ? h[i] : -h[i];
    }
    drho = d;
    rho = (T)1 + d;
    const T ir = recip(rho);
    u[0] = jx * ir; u[1] = jy * ir; u[2] = jz * ir;
    if (kp.force == 1) {
        const T io = (T)(1.0 / kp.omega);
        u[0] += (T)kp.g[0] * io; u[1] += (T)kp.g[1] * io; u[2] += (T)kp.g[2] * io;
    } else if (kp.force == 2) {
        u[0] += (T)(0.5 * kp.g[0]); u[1] += (T)(0.5 * kp.g[1]); u[2] += (T)(0.5 * kp.g[2]);
    }
    if ((kp.parity & 0x8u) && cls == PRES_E) u[1] = 0;
}

template <class L, class T>
__device__ inline void guo_terms(const KParams& kp, T rho, const T* u, T* S) {
    const T Fx = rho * (T)kp.g[0], Fy = rho * (T)kp.g[1], Fz = rho * (T)kp.g[2];
    const T uF = u[0] * Fx + u[1] * Fy + u[2] * Fz;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        const T cx = L::cx(i), cy = L::cy(i), cz = L::cz(i);
        const T cF = cx * Fx + cy * Fy + cz * Fz;
        const T cu = cx * u[0] + cy * u[1] + cz * u[2];
        S[i] = (T)L::w(i) * ((T)3 * (cF - uF) + (T)9 * cu * cF);
    }
}

// Deviations of the MRT equilibrium moments from the rest moments (M w).
// D2Q9: S/GpuCollision.cu:986-994; D3Q19: :1065-1083.
template <class L, class T>
__device__ inline void mrt_dmeq(const KParams& kp, T drho, T rho, const T* u, T* dm) {
    if (L::Q == 9) {
        const T uu = u[0] * u[0] + u[1] * u[1];
        const T k = (kp.parity & 0x1u) ? rho : (T)1;
        dm[0] = drho;
        dm[1] = (T)-2 * drho + (T)3 * k * rho * uu;
        dm[2] = drho - (T)3 * k * rho * uu;
        dm[3] = rho * u[0]; dm[4] = -rho * u[0];
        dm[5] = rho * u[1]; dm[6] = -rho * u[1];
        dm[7] = rho * (u[0] * u[0] - u[1] * u[1]);
        dm[8] = rho * u[0] * u[1];
    } else {
        const T jx = rho * u[0], jy = rho * u[1], jz = rho * u[2];
        const T j2 = jx * jx + jy * jy + jz * jz;
        dm[0] = drho;
        dm[1] = (T)-11 * drho + (T)19 * j2;
        dm[2] = (T)(-475.0 / 63.0) * j2 - (T)mrest<D3Q19>(2);
        dm[3] = jx; dm[4] = (T)(-2.0 / 3.0) * jx;
        dm[5] = jy; dm[6] = (T)(-2.0 / 3.0) * jy;
        dm[7] = jz; dm[8] = (T)(-2.0 / 3.0) * jz;
        dm[9] = (T)2 * jx * jx - (jy * jy + jz * jz);
        dm[10] = 0;
        dm[11] = jy * jy - jz * jz;
        dm[12] = 0;
        dm[13] = jx * jy; dm[14] = jy * jz; dm[15] = jx * jz;
        dm[16] = dm[17] = dm[18] = 0;
    }
}

// Collision in deviation form.  BGK: S/GpuCollision.cu:528-550, :882-916;
// TRT: :917-969 (general Lambda); MRT: :971-1037, :1039-1170.
template <class L, class T, int C>
__device__ inline void collide(const KParams& kp, const T* h, T drho, T rho, const T* u, T* out) {
    const bool guo = kp.force == 2;
    if (C == 2) {
        T S[L::Q];
        if (guo) guo_terms<L, T>(kp, rho, u, S);
        T m[L::Q], dm[L::Q], gs[L::Q], dd[L::Q];
        mrt_dmeq<L, T>(kp, drho, rho, u, dm);
        m_apply<L, T>(h, m, std::make_integer_sequence<int, L::Q>{});
        if (guo) m_apply<L, T>(S, gs, std::make_integer_sequence<int, L::Q>{});
#pragma unroll
        for (int r = 0; r < L::Q; ++r) {
            const T s = (T)kp.s[r];
            m[r] = s * (m[r] - dm[r]);
            if (guo) m[r] -= ((T)1 - (T)0.5 * s) * gs[r];
        }
        mi_apply<L, T>(m, dd, std::make_integer_sequence<int, L::Q>{});
#pragma unroll
        for (int i = 0; i < L::Q; ++i) out[i] = h[i] - dd[i];
        if (L::Q == 19 && (kp.parity & 0x2u))
            out[L::Q - 1] = -(T)L::w(L::Q - 1);   // fColl_18 stays at its zero initial value
        return;
    }
    // BGK and TRT in symmetric/antisymmetric pair form: for a link pair
    // (i, opp i) the parts of f_eq and of the Guo source are
    //   e+ = w (drho + rho (4.5 cu^2 - 1.5 u.u)),  e- = 3 w rho cu,
    //   S+ = w (9 cu cF - 3 u.F),                  S- = 3 w cF,
    // relaxed with omega+ and omega- (equal for BGK) and prefactors 1 - omega/2.
    const T op = (T)kp.omega;
    const T om = C == 0 ? op : (T)kp.omega_m;
    const T gp = (T)1 - (T)0.5 * op, gm = (T)1 - (T)0.5 * om;
    const T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    T Fx = 0, Fy = 0, Fz = 0, uF = 0;
    if (guo) {
        Fx = rho * (T)kp.g[0]; Fy = rho * (T)kp.g[1]; Fz = rho * (T)kp.g[2];
        uF = u[0] * Fx + u[1] * Fy + u[2] * Fz;
    }
    {
        const T w0 = (T)L::w(0);
        out[0] = h[0] - op * (h[0] - w0 * (drho - (T)1.5 * rho * uu));
        if (guo) out[0] -= gp * (T)3 * w0 * uF;
    }
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int o = L::opp(i);
        if (i < o) {
            const T w = (T)L::w(i);
            const T cu = (T)L::cx(i) * u[0] + (T)L::cy(i) * u[1] + (T)L::cz(i) * u[2];
            const T ep = w * (drho + rho * ((T)4.5 * cu * cu - (T)1.5 * uu));
            const T em = (T)3 * w * rho * cu;
            T dp = op * ((T)0.5 * (h[i] + h[o]) - ep);
            T dm = om * ((T)0.5 * (h[i] - h[o]) - em);
            if (guo) {
                const T cF = (T)L::cx(i) * Fx + (T)L::cy(i) * Fy + (T)L::cz(i) * Fz;
                dp -= gp * w * ((T)9 * cu * cF - (T)3 * uF);
                dm -= gm * (T)3 * w * cF;
            }
            out[i] = h[i] - dp - dm;
            out[o] = h[o] - dp + dm;
        }
    }
}

// Store of one node's post-collision populations.  The 16-bit storages use
// stochastic rounding (unbiased: a body force smaller than half an ulp of the
// stored deviation still accumulates), with random bits from a hash of the
// compact node index and a per-launch seed, rotated per direction; results are
// deterministic for a given seed sequence.
template <class L, int S>
__device__ inline void store_node(typename Io<L, S>::raw* __restrict__ dst, const KParams& kp, uint32_t own,
                                  const typename Io<L, S>::T* out) {
    if (S >= 2 && kp.sr) {
        uint32_t h = own * 0x9E3779B1u + kp.seed;
        h ^= h >> 15;
        h *= 0x2C1B3C6Du;
        h ^= h >> 12;
#pragma unroll
        for (int i = 0; i < L::Q; ++i)
            Io<L, S>::st_sr(dst, paddr<L>(kp, i, own), out[i], i, (h << (i & 31)) | (h >> ((32 - i) & 31)));
    } else {
#pragma unroll
        for (int i = 0; i < L::Q; ++i) Io<L, S>::st(dst, paddr<L>(kp, i, own), out[i], i);
    }
}

// ---------------------------------------------------------------------------
// kernels
// Quotient of a 32-bit n by an invariant divisor d: q = (((m n) >> 32) + n) >> s
// with m = floor(2^32 (2^s - d) / d) + 1, s = ceil(log2 d) (exact for all n < 2^32).
__device__ inline uint32_t fdiv(uint32_t n, uint64_t m, uint32_t s) {
    return (uint32_t)((((m * n) >> 32) + n) >> s);
}

// Workgroups are dealt round-robin to the XCDs; renumber them so that each XCD
// runs one contiguous range of the lattice and neighbouring rows (and their
// cache lines of cidx and P) meet in the same L2.  Bijective for any count.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb, uint32_t nxcd) {
    if (nxcd <= 1) return b;
    const uint32_t q = nb / nxcd, r = nb % nxcd, x = b % nxcd, k = b / nxcd;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

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
    const uint32_t own = xcd_block(blockIdx.x, gridDim.x, kp.nxcd) * blockDim.x + threadIdx.x;
    if (own >= kp.nf) return;
    const uint32_t idx = didx[own];
    const uint32_t nd = nodew[own];
    const uint32_t q1 = fdiv(idx, kp.mx, kp.sx);
    const int x = (int)(idx - q1 * (uint32_t)kp.nx);
    const uint32_t zq = fdiv(q1, kp.my, kp.sy);
    const int y = (int)(q1 - zq * (uint32_t)kp.ny), z = (int)zq;
    uint32_t cn[L::Q];
    source_index<L, OPEN>(cidx, kp, x, y, z, cn);
    const uint32_t cls = nd >> kClsShift;
    T h[L::Q];
    gather<L, S, OPEN>(src, kp, own, nd, cn, h);
    if (OPEN && cls == VEL_W) zou_inlet<L, T>(kp, fl, idx, nd, h);
    if (kp.wall == 1 && (nd & kBBMask)) onnode_walls<L, T>(nd, h);
    if (OPEN && cls == PRES_E) zou_outlet<L, T>(kp, h);
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    if (write_macros) {
        fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
        fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
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
__device__ inline int select_bit(uint32_t m, uint32_t t) {   // position of the t-th set bit
    int pos = 0;
    uint32_t c = __popc(m & 0xFFFFu);
    if (t >= c) { t -= c; m >>= 16; pos += 16; }
    c = __popc(m & 0xFFu);
    if (t >= c) { t -= c; m >>= 8; pos += 8; }
    c = __popc(m & 0xFu);
    if (t >= c) { t -= c; m >>= 4; pos += 4; }
    c = __popc(m & 0x3u);
    if (t >= c) { t -= c; m >>= 2; pos += 2; }
    if (t >= (m & 1u)) pos += 1;
    return pos;
}

constexpr uint32_t kSolidSrc = 0xFFFFFFFEu;

template <class L, int S, int C, bool OPEN, bool SIMPLE>
__global__ __launch_bounds__(256) void k_stepr(const typename Io<L, S>::raw* __restrict__ src,
                                               typename Io<L, S>::raw* __restrict__ dst,
                                               const uint2* __restrict__ rowmeta,
                                               const uint2* __restrict__ blk,
                                               const uint32_t* __restrict__ nodew, KParams kp,
                                               Fields fl, int write_macros) {
    using T = typename Io<L, S>::T;
    extern __shared__ uint2 smeta[];
    const int W = kp.words, nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const uint2 bi = blk[blockIdx.x];
    const uint32_t r0 = bi.x;
    const int R = (int)bi.y;
    // stage the words of rows r0..r0+R-1 and of their (dy, dz) neighbours;
    // slot (k, nb) with nb = (dz+1)*3 + (dy+1); rows outside a non-periodic
    // face read as all-solid (outside sources are resolved by coordinates)
    const int tot = R * 9 * W;
    for (int e = threadIdx.x; e < tot; e += blockDim.x) {
        const int k = e / (9 * W), rem = e - k * 9 * W, nb = rem / W, w = rem - nb * W;
        const uint32_t r = r0 + (uint32_t)k;
        const uint32_t zq = fdiv(r, kp.my, kp.sy);
        int y = (int)(r - zq * (uint32_t)ny), z = (int)zq;
        y += nb % 3 - 1;
        z += nb / 3 - 1;
        bool out = false;
        if (y < 0 || y >= ny) { if (kp.per[1]) y = (y + ny) % ny; else out = true; }
        if (z < 0 || z >= nz) { if (kp.per[2]) z = (z + nz) % nz; else out = true; }
        smeta[e] = out ? make_uint2(0u, 0u) : rowmeta[((size_t)z * ny + y) * W + w];
    }
    __syncthreads();
    const uint32_t own = blockIdx.x * blockDim.x + threadIdx.x;
    if (own >= kp.nf) return;
    // own row: the last staged own row whose first compact index is <= own
    int k = 0;
    while (k + 1 < R && smeta[((k + 1) * 9 + 4) * W].y <= own) ++k;
    const uint2* me = smeta + (k * 9 + 4) * W;
    int w = 0;
    while (w + 1 < W && me[w + 1].y <= own) ++w;
    const int x = 32 * w + select_bit(me[w].x, own - me[w].y);
    const uint32_t r = r0 + (uint32_t)k;
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
    T h[L::Q];
    h[0] = Io<L, S>::ld(src, paddr<L>(kp, 0, own), 0);
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int cx = L::cx(i), cy = L::cy(i), cz = L::cz(i), o = L::opp(i);
        const int sx = cx == 0 ? x : (cx > 0 ? xm : xp);
        bool out = false;
        if (OPEN)
            out = (cx > 0 && xm_o) || (cx < 0 && xp_o) || (cy > 0 && ym_o) || (cy < 0 && yp_o) ||
                  (cz > 0 && zm_o) || (cz < 0 && zp_o);
        uint32_t j = kNone;
        if (!out) {
            const uint2 e = smeta[(k * 9 + (1 - cz) * 3 + (1 - cy)) * W + (sx >> 5)];
            const uint32_t bit = (uint32_t)sx & 31u;
            j = ((e.x >> bit) & 1u) ? e.y + __popc(e.x & ((1u << bit) - 1u)) : kSolidSrc;
        }
        const bool wall = ((nd >> o) & 1u) || j == kSolidSrc;
        int pl = i;
        if (wall || j == kNone) {
            j = own;
            pl = wall && kp.wall == 0 ? o : i;
        }
        h[i] = Io<L, S>::ld(src, paddr<L>(kp, pl, j), i);
    }
    if (OPEN && cls == VEL_W) zou_inlet<L, T>(kp, fl, idx, nd, h);
    if (kp.wall == 1 && (nd & kBBMask)) onnode_walls<L, T>(nd, h);
    if (OPEN && cls == PRES_E) zou_outlet<L, T>(kp, h);
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    if (write_macros) {
        fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
        fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
    }
    T out[L::Q];
    collide<L, T, C>(kp, h, drho, rho, u, out);
    store_node<L, S>(dst, kp, own, out);
}

// Post-boundary populations (dense, full, FP64) and moments of the state in src.
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
#pragma unroll
        for (int i = 0; i < L::Q; ++i) fl.fout[i * n + idx] = 0.0;
        return;
    }
    const uint32_t nd = node[idx];
    const uint32_t cls = nd >> kClsShift;
    uint32_t cn[L::Q];
    source_index<L, OPEN>(cidx, kp, x, y, z, cn);
    T h[L::Q];
    gather<L, S, OPEN>(src, kp, own, nd, cn, h);
    if (OPEN && cls == VEL_W) zou_inlet<L, T>(kp, fl, idx, nd, h);
    if (kp.wall == 1 && (nd & kBBMask)) onnode_walls<L, T>(nd, h);
    if (OPEN && cls == PRES_E) zou_outlet<L, T>(kp, h);
    T drho, rho, u[3];
    moments<L, T>(kp, cls, h, drho, rho, u);
    fl.rho[idx] = (double)rho; fl.u[idx] = (double)u[0];
    fl.v[idx] = (double)u[1]; fl.w[idx] = (double)u[2];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) fl.fout[i * n + idx] = (double)h[i] + L::w(i);
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

__global__ __launch_bounds__(kRedThreads) void k_reduce_partial(const uint32_t* __restrict__ node,
                                                                 Fields fl, double* uprev, double* vprev,
                                                                 double* wprev, size_t n, double* partials) {
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if ((node[i] >> kClsShift) == SOLID) continue;
        const double u = fl.u[i], v = fl.v[i], w = fl.w[i], r = fl.rho[i];
        const double du = u - uprev[i], dv = v - vprev[i], dw = w - wprev[i];
        const double uu = u * u + v * v + w * w;
        a0 += du * du + dv * dv + dw * dw;
        a1 += uu;
        a2 += r;
        if (!(uu == uu) || !(r == r) || isinf(uu) || isinf(r)) a3 += 1.0;
        uprev[i] = u; vprev[i] = v; wprev[i] = w;
    }
    __shared__ double sh[4][kRedThreads / 64];
    a0 = wave_sum(a0); a1 = wave_sum(a1); a2 = wave_sum(a2); a3 = wave_sum(a3);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) { sh[0][wv] = a0; sh[1][wv] = a1; sh[2][wv] = a2; sh[3][wv] = a3; }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int k = 0; k < kRedThreads / 64; ++k) s += sh[threadIdx.x][k];
        partials[(size_t)blockIdx.x * 4 + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(256) void k_reduce_final(const double* __restrict__ partials, int nb, double* out4) {
    __shared__ double sh[4][256];
    for (int k = 0; k < 4; ++k) {
        double s = 0;
        for (int b = threadIdx.x; b < nb; b += blockDim.x) s += partials[(size_t)b * 4 + k];
        sh[k][threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int t = 0; t < 256; ++t) s += sh[threadIdx.x][t];
        out4[threadIdx.x] = s;
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
        const size_t lds = (size_t)kp.rmax * 9 * kp.words * sizeof(uint2);
        if (kp.simple)
            hipLaunchKernelGGL((k_stepr<L, S, C, OPEN, true>), dim3(blocks), dim3(256), lds, st, (const raw*)src,
                               (raw*)dst, kp.rowmeta, kp.blk, kp.nodew, kp, fl, wm ? 1 : 0);
        else
            hipLaunchKernelGGL((k_stepr<L, S, C, OPEN, false>), dim3(blocks), dim3(256), lds, st, (const raw*)src,
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
                         double* wprev, double* partials, int nblocks, double* out4, hipStream_t st) {
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nblocks), dim3(kRedThreads), 0, st, node, fl, uprev, vprev, wprev,
                       n, partials);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, st, (const double*)partials, nblocks, out4);
    return hipGetLastError();
}

}  // namespace lbm
