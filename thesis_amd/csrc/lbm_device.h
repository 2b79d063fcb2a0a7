// Device building blocks of the lattice-Boltzmann step shared by the step
// kernels (lbm_kernels.hip) and the batched persistent kernel (lbm_batch.hip):
// lattices, storage formats, pull gather, boundaries, moments, collision.
#pragma once
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

// The reference's D3Q19 MRT equilibrium has no density term in epsilon
// (w_eps = 0, S/GpuCollision.cu:1067), while the rest weights carry
// epsilon = 3: its rest state is f_ref = w - r with r = 3 M^-1 e_eps, 14-86 %
// away from w.  D3Q19 MRT runs therefore store h' = f - f_ref = (f - w) + r,
// which is small near rest, instead of f - w (whose reduced-precision copies
// would spend their digits on r).  r_i depends on |c_i| only, so bounce-back,
// on-node and curved walls and symmetry planes act on h' unchanged; boundary
// closures that mix speeds convert to f - w and back.  The MRT collision in h'
// drops the rest moment from epsilon_eq.
template <class L>
__host__ __device__ constexpr double mref_off(int i) {
    return L::Q == 19 ? 3.0 * mrow<L>(2, i) / mnorm2<L>(2) : 0.0;
}

// The MRT transforms in link-pair form.  Every moment polynomial is even or
// odd under c -> -c, so an even row of M weighs h_i and h_opp(i) alike and an
// odd row with opposite signs: with s = h_i + h_opp and a = h_i - h_opp per
// link pair, even moments are rows over (h_0, s) and odd ones rows over a, and
// the way back gives each pair E +- O from the even and odd rows of M^-1.
// Coefficients are compile-time constants; zero ones drop out and +-1 ones
// become adds, about half the operations of the full 19 x 19 products.
template <class L>
__host__ __device__ constexpr int pair_rep(int k) {   // k-th link i with i < opp(i)
    for (int i = 1; i < L::Q; ++i)
        if (i < L::opp(i) && k-- == 0) return i;
    return -1;
}
template <class L>
__host__ __device__ constexpr int mparity(int r) {    // +1 even row, -1 odd row, 0 neither
    bool even = true, odd = true;
    for (int i = 0; i < L::Q; ++i) {
        even = even && mrow<L>(r, L::opp(i)) == mrow<L>(r, i);
        odd = odd && mrow<L>(r, L::opp(i)) == -mrow<L>(r, i);
    }
    return even ? 1 : odd ? -1 : 0;
}
template <class L>
__host__ __device__ constexpr bool mrows_have_parity() {
    for (int r = 0; r < L::Q; ++r)
        if (mparity<L>(r) == 0) return false;
    return pair_rep<L>((L::Q - 1) / 2 - 1) > 0 && pair_rep<L>((L::Q - 1) / 2) < 0;
}
// Moments of the Guo source S_i = w_i (3 (c_i.F - u.F) + 9 (c_i.u)(c_i.F)):
// (M S)_r = 3 sum_d A_rd F_d - 3 (u.F) B_r + 9 sum_de C_rde u_d F_e.
template <class L>
__host__ __device__ constexpr double cc(int i, int d) { return d == 0 ? L::cx(i) : d == 1 ? L::cy(i) : L::cz(i); }
template <class L>
__host__ __device__ constexpr double guoA(int r, int d) {
    double t = 0;
    for (int i = 0; i < L::Q; ++i) t += mrow<L>(r, i) * L::w(i) * cc<L>(i, d);
    return t;
}
template <class L>
__host__ __device__ constexpr double guoC(int r, int d, int e) {
    double t = 0;
    for (int i = 0; i < L::Q; ++i) t += mrow<L>(r, i) * L::w(i) * cc<L>(i, d) * cc<L>(i, e);
    return t;
}

// The coefficients as template constants, so that each is evaluated at
// compile time (a constexpr call in an ordinary expression need not be).
template <class L, int K> inline constexpr int kPair = pair_rep<L>(K);
template <class L, int R> inline constexpr int kPar = mparity<L>(R);
template <class L, int R, int I> inline constexpr double kM = mrow<L>(R, I);
template <class L, int R, int I> inline constexpr double kMi = mrow<L>(R, I) / mnorm2<L>(R);
template <class L, int R, int D> inline constexpr double kGA = 3.0 * guoA<L>(R, D);
template <class L, int R> inline constexpr double kGB = -3.0 * mrest<L>(R);
template <class L, int R, int D, int E> inline constexpr double kGC = 9.0 * guoC<L>(R, D, E);
template <class L, int R>
inline constexpr bool kGuoRow = kGB<L, R> != 0.0 || kGA<L, R, 0> != 0.0 || kGA<L, R, 1> != 0.0 ||
                                kGA<L, R, 2> != 0.0 || kGC<L, R, 0, 0> != 0.0 || kGC<L, R, 0, 1> != 0.0 ||
                                kGC<L, R, 0, 2> != 0.0 || kGC<L, R, 1, 0> != 0.0 || kGC<L, R, 1, 1> != 0.0 ||
                                kGC<L, R, 1, 2> != 0.0 || kGC<L, R, 2, 0> != 0.0 || kGC<L, R, 2, 1> != 0.0 ||
                                kGC<L, R, 2, 2> != 0.0;

// acc (+)= c x for compile-time c, skipping zero coefficients; the first
// term starts the sum (no 0 + ...), so the operation count is the nonzero count
template <class T>
struct CSum {
    T acc = 0;
    bool any = false;
    __device__ __forceinline__ void add(double c, T x) {
#pragma clang fp contract(on)
        if (c == 0.0) return;
        acc = any ? acc + (T)c * x : (T)c * x;
        any = true;
    }
};

template <class L, class T, int... K>
__device__ __forceinline__ void pair_split(const T* h, T* s, T* a, std::integer_sequence<int, K...>) {
    ((s[K] = h[kPair<L, K>] + h[L::opp(kPair<L, K>)]), ...);
    ((a[K] = h[kPair<L, K>] - h[L::opp(kPair<L, K>)]), ...);
}
template <class L, class T, int R, int... K>
__device__ __forceinline__ T m_row_pairs(T h0, const T* s, const T* a, std::integer_sequence<int, K...>) {
    CSum<T> m;
    if constexpr (kPar<L, R> > 0) {
        m.add(kM<L, R, 0>, h0);
        (m.add(kM<L, R, kPair<L, K>>, s[K]), ...);
    } else {
        (m.add(kM<L, R, kPair<L, K>>, a[K]), ...);
    }
    return m.acc;
}
template <class L, class T, int R>
__device__ __forceinline__ T guo_row(const T* u, const T* F, T uF) {
    CSum<T> g;
    g.add(kGA<L, R, 0>, F[0]); g.add(kGA<L, R, 1>, F[1]); g.add(kGA<L, R, 2>, F[2]);
    g.add(kGB<L, R>, uF);
    g.add(kGC<L, R, 0, 0>, u[0] * F[0]); g.add(kGC<L, R, 0, 1>, u[0] * F[1]); g.add(kGC<L, R, 0, 2>, u[0] * F[2]);
    g.add(kGC<L, R, 1, 0>, u[1] * F[0]); g.add(kGC<L, R, 1, 1>, u[1] * F[1]); g.add(kGC<L, R, 1, 2>, u[1] * F[2]);
    g.add(kGC<L, R, 2, 0>, u[2] * F[0]); g.add(kGC<L, R, 2, 1>, u[2] * F[1]); g.add(kGC<L, R, 2, 2>, u[2] * F[2]);
    return g.acc;
}
// relaxed moment deviations y_r = s_r (m_r - m_eq,r) - (1 - s_r / 2) (M S)_r
template <class L, class T, int R>
__device__ __forceinline__ T mrt_relax_row(const KParams& kp, T h0, const T* s, const T* a, const T* dm, bool guo,
                                           const T* u, const T* F, T uF) {
#pragma clang fp contract(on)
    const T sr = (T)kp.s[R];
    T y = sr * (m_row_pairs<L, T, R>(h0, s, a, std::make_integer_sequence<int, (L::Q - 1) / 2>{}) - dm[R]);
    if constexpr (kGuoRow<L, R>) {
        if (guo) y -= ((T)1 - (T)0.5 * sr) * guo_row<L, T, R>(u, F, uF);
    }
    return y;
}
template <class L, class T, int... R>
__device__ __forceinline__ void mrt_relax(const KParams& kp, T h0, const T* s, const T* a, const T* dm, bool guo,
                                          const T* u, const T* F, T uF, T* y, std::integer_sequence<int, R...>) {
    ((y[R] = mrt_relax_row<L, T, R>(kp, h0, s, a, dm, guo, u, F, uF)), ...);
}
template <class L, class T, int I, int P, int... R>
__device__ __forceinline__ T mi_part(const T* y, std::integer_sequence<int, R...>) {
    CSum<T> d;
    (d.add(kPar<L, R> == P ? kMi<L, R, I> : 0.0, y[R]), ...);
    return d.acc;
}
template <class L, class T, int K>
__device__ __forceinline__ void mrt_back_pair(const T* h, const T* y, T* out) {
#pragma clang fp contract(on)
    constexpr int i = kPair<L, K>, o = L::opp(i);
    constexpr auto rows = std::make_integer_sequence<int, L::Q>{};
    const T E = mi_part<L, T, i, 1>(y, rows), O = mi_part<L, T, i, -1>(y, rows);
    out[i] = h[i] - (E + O);
    out[o] = h[o] - (E - O);
}
template <class L, class T, int... K>
__device__ __forceinline__ void mrt_back(const T* h, const T* y, T* out, std::integer_sequence<int, K...>) {
    out[0] = h[0] - mi_part<L, T, 0, 1>(y, std::make_integer_sequence<int, L::Q>{});
    (mrt_back_pair<L, T, K>(h, y, out), ...);
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
};
template <class L> struct Io<L, 1> {
    using raw = float; using T = float;
    __device__ static T cv(raw r, int) { return r; }
    __device__ static raw pk(T v, int) { return v; }
    __device__ static T ld(const raw* p, uint32_t k, int) { return ldo(p, k); }
    __device__ static void st(raw* p, uint32_t k, T v, int) { sto(p, k, v); }
};
template <class L> struct Io<L, 2> {
    using raw = uint16_t; using T = float;
    __device__ static T cv(raw r, int i) { return (float)__builtin_bit_cast(_Float16, r) * f16_a<L>(i); }
    __device__ static raw pk(T v, int i) { return __builtin_bit_cast(uint16_t, (_Float16)(v * f16_ainv<L>(i))); }
    __device__ static T ld(const raw* p, uint32_t k, int i) { return cv(ldo(p, k), i); }
    __device__ static void st(raw* p, uint32_t k, T v, int i) { sto(p, k, pk(v, i)); }
    // unbiased (stochastic) rounding with the random bits rnd into the low
    // half of acc (the high half is kept)
    typedef short pair __attribute__((ext_vector_type(2)));
    __device__ static pair sr(pair acc, T v, int i, uint32_t rnd) {
        return __builtin_amdgcn_cvt_sr_f16_f32(acc, v * f16_ainv<L>(i), rnd, false);
    }
    // the stored value x_i = h_i / a_i itself (a_i = f16_a, w_i / a_i = kw)
    static constexpr float kw = 1024.0f;
    __device__ static float a(int i) { return f16_a<L>(i); }
    __device__ static T x(raw r) { return (float)__builtin_bit_cast(_Float16, r); }
    __device__ static raw px(T v) { return __builtin_bit_cast(uint16_t, (_Float16)v); }
    __device__ static pair srx(pair acc, T v, uint32_t rnd) { return __builtin_amdgcn_cvt_sr_f16_f32(acc, v, rnd, false); }
};
template <class L> struct Io<L, 3> {
    using raw = uint16_t; using T = float;
    __device__ static T cv(raw r, int i) { return bf16_to_f32(r) * bf16_a<L>(i); }
    __device__ static raw pk(T v, int i) { return f32_to_bf16(v * bf16_ainv<L>(i)); }
    __device__ static T ld(const raw* p, uint32_t k, int i) { return cv(ldo(p, k), i); }
    __device__ static void st(raw* p, uint32_t k, T v, int i) { sto(p, k, pk(v, i)); }
    typedef short pair __attribute__((ext_vector_type(2)));
    __device__ static pair sr(pair acc, T v, int i, uint32_t rnd) {
        return __builtin_amdgcn_cvt_sr_bf16_f32(acc, v * bf16_ainv<L>(i), rnd, false);
    }
    static constexpr float kw = 1.0f;
    __device__ static float a(int i) { return bf16_a<L>(i); }
    __device__ static T x(raw r) { return bf16_to_f32(r); }
    __device__ static raw px(T v) { return f32_to_bf16(v); }
    __device__ static pair srx(pair acc, T v, uint32_t rnd) { return __builtin_amdgcn_cvt_sr_bf16_f32(acc, v, rnd, false); }
};

// ---------------------------------------------------------------------------
// node pipeline pieces (T = double for F64 storage, float otherwise)

constexpr uint32_t kNone = 0xFFFFFFFFu;

// Quotient of a 32-bit n by an invariant divisor d: q = (((m n) >> 32) + n) >> s
// with m = floor(2^32 (2^s - d) / d) + 1, s = ceil(log2 d) (exact for all n < 2^32).
// m < 2^32 for every d <= 2^31, so the high product is one v_mul_hi_u32.
__device__ inline uint32_t fdiv(uint32_t n, uint64_t m, uint32_t s) {
    return (uint32_t)(((uint64_t)__umulhi((uint32_t)m, n) + n) >> s);
}

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

// Largest |opp(i) - i|: how far a link's opposite plane lies from its own.
template <class L>
__host__ __device__ constexpr int max_opp_gap() {
    int g = 0;
    for (int i = 0; i < L::Q; ++i) g = L::opp(i) - i > g ? L::opp(i) - i : (i - L::opp(i) > g ? i - L::opp(i) : g);
    return g;
}

// Tile-layout slot of plane i of compact node j.  j >> 6 < 2^24 whenever the
// buffer is below 4 GiB, so the tile stride is a full-rate 24-bit multiply.
template <class L>
__device__ inline uint32_t tslot(int i, uint32_t j) {
    return __umul24(j >> 6, 64u * L::Q) + (uint32_t)i * 64u + (j & 63u);
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
            pl = ((nd >> o) & 1u) && (kp.wall == 0 || kp.wall == 3) ? o : i;
        }
        h[i] = Io<L, S>::ld(src, paddr<L>(kp, pl, j), i);
    }
}

// Open boundaries on D2Q9 faces, in full populations (deviations h = f - w
// in and out).
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
// Moving lid (north): Zou-He from the density, both momentum components and
// non-equilibrium bounce-back of f4; or the reference's form
// (S/GpuBoundaries.cu:21-41), f4 = f2, f8/f7 = f6/f5 +- B u0/6 with
// B = f0 + f1 + f3 + 2(f2 + f5 + f6), corner nodes skipped.
template <class L, class T>
__device__ inline void zou_lid(const KParams& kp, const Fields& fl, size_t idx, uint32_t nd, T* h) {
    if (L::Q != 9) return;
    if (kp.inlet == 1 && (nd & kBBMask)) return;
    T f[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = h[i] + (T)L::w(i);
    const T u0 = (T)fl.ubc[idx], v0 = (T)fl.vbc[idx];
    const T k = f[0] + f[1] + f[3] + (T)2 * (f[2] + f[5] + f[6]);
    if (kp.inlet == 1) {
        f[4] = f[2];
        f[8] = f[6] + k * u0 / (T)6;
        f[7] = f[5] - k * u0 / (T)6;
    } else {
        const T r = k / ((T)1 + v0);
        const T t = (T)0.5 * (f[1] - f[3]);
        f[4] = f[2] - (T)(2.0 / 3.0) * r * v0;
        f[7] = f[5] + t - (T)0.5 * r * u0 - r * v0 / (T)6;
        f[8] = f[6] - t + (T)0.5 * r * u0 - r * v0 / (T)6;
    }
    h[4] = f[4] - (T)L::w(4); h[7] = f[7] - (T)L::w(7); h[8] = f[8] - (T)L::w(8);
}
// East outlet, node-local models: Zou-He pressure, or the reference's velocity
// outlet (S/GpuBoundaries.cu:1571-1605) with R = (f0 + f2 + f4 + 2(f1 + f5 + f8)) / (1 - u0),
// u0 = ubc at the node.  The extrapolating models are applied by outlet_extrap.
template <class L, class T>
__device__ inline void zou_outlet(const KParams& kp, const Fields& fl, size_t idx, T* h) {
    if (L::Q != 9 || kp.outlet >= 2) return;
    T f[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) f[i] = h[i] + (T)L::w(i);
    const T a = f[0] + f[2] + f[4] + (T)2 * (f[1] + f[5] + f[8]);
    if (kp.outlet == 1) {
        const T u0 = (T)fl.ubc[idx];
        const T r = a / ((T)1 - u0);
        f[3] = f[1] - (T)(2.0 / 3.0) * r * u0;
        f[7] = f[5] - r * u0 / (T)6;
        f[6] = f[8] - r * u0 / (T)6;
    } else {
        const T ro = (T)kp.rho_out;
        const T ux = (T)-1 + a / ro;
        f[3] = f[1] - (T)(2.0 / 3.0) * ro * ux;
        f[7] = f[5] + (T)0.5 * (f[2] - f[4]) - ro * ux / (T)6;
        f[6] = f[8] - (T)0.5 * (f[2] - f[4]) - ro * ux / (T)6;
    }
    h[3] = f[3] - (T)L::w(3); h[6] = f[6] - (T)L::w(6); h[7] = f[7] - (T)L::w(7);
}
template <class L, class T>
__device__ inline void onnode_walls(uint32_t nd, T* h) {
#pragma unroll
    for (int d = 1; d < L::Q; ++d)
        if ((nd >> d) & 1u) h[L::opp(d)] = h[d];
}

// ---------------------------------------------------------------------------
// D3Q19 faces.  A face is (axis A, inward sign G): W (0,+1), E (0,-1),
// S (1,+1), N (1,-1), B (2,+1), T (2,-1); its unknowns are the populations
// with G c_A > 0.  Every closure below is instantiated per face, edge or
// corner, so all population indices are compile-time constants and the
// node's populations stay in registers.
template <class L> __host__ __device__ constexpr int cmp(int i, int a) {
    return a == 0 ? L::cx(i) : a == 1 ? L::cy(i) : L::cz(i);
}
template <class L> __host__ __device__ constexpr int dir_of(int x, int y, int z) {
    for (int i = 0; i < L::Q; ++i)
        if (L::cx(i) == x && L::cy(i) == y && L::cz(i) == z) return i;
    return -1;
}
// direction with c_a = ca and c_b = cb (a == b: the axis direction ca)
template <class L> __host__ __device__ constexpr int dir_ax(int a, int ca, int b, int cb) {
    return dir_of<L>((a == 0 ? ca : 0) + (b == 0 ? cb : 0), (a == 1 ? ca : 0) + (b == 1 ? cb : 0),
                     (a == 2 ? ca : 0) + (b == 2 ? cb : 0));
}
static_assert(dir_ax<D3Q19>(0, 1, 1, 1) == 7 && dir_ax<D3Q19>(1, -1, 2, -1) == 18 && dir_ax<D3Q19>(2, 1, 2, 0) == 5,
              "D3Q19 direction lookup");

// Transverse part of a velocity face for one tangential axis B (Hecht &
// Harting 2010): N = sum over c_A = 0 of c_B f / 2 - rho U_B / 3 and the two
// diagonal unknowns f(G e_A + t e_B) = f(-G e_A - t e_B) + rho (G U_A + t U_B)/6 - t N.
template <class L, class T, int A, int G, int B>
__device__ inline void vface_b(T* f, const T* U, T rho) {
    T N = 0;
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        if (cmp<L>(i, A) == 0 && cmp<L>(i, B) != 0) N += cmp<L>(i, B) > 0 ? f[i] : -f[i];
    N = (T)0.5 * N - rho * U[B] * (T)(1.0 / 3.0);
    constexpr int p = dir_ax<L>(A, G, B, 1), pm = dir_ax<L>(A, -G, B, -1);
    constexpr int m = dir_ax<L>(A, G, B, -1), mm = dir_ax<L>(A, -G, B, 1);
    f[p] = f[pm] + rho * ((T)G * U[A] + U[B]) * (T)(1.0 / 6.0) - N;
    f[m] = f[mm] + rho * ((T)G * U[A] - U[B]) * (T)(1.0 / 6.0) + N;
}
// Velocity face (S/GpuBoundaries.cu:73-399) on full populations f: rho from
// the tangential and outgoing populations over 1 - G U_A (1 + G U_A with the
// reference's west/east flip), the normal unknown by non-equilibrium
// bounce-back, the diagonal unknowns by vface_b.
template <class L, class T, int A, int G>
__device__ inline void vface(T* f, const T* U, bool flip) {
    T K = 0;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        if (cmp<L>(i, A) == 0) K += f[i];
        else if (cmp<L>(i, A) == -G) K += (T)2 * f[i];
    }
    const T rho = K / (flip ? (T)1 + (T)G * U[A] : (T)1 - (T)G * U[A]);
    constexpr int nin = dir_ax<L>(A, G, A, 0), nout = dir_ax<L>(A, -G, A, 0);
    f[nin] = f[nout] + (T)G * rho * U[A] * (T)(1.0 / 3.0);
    vface_b<L, T, A, G, (A + 1) % 3>(f, U, rho);
    vface_b<L, T, A, G, (A + 2) % 3>(f, U, rho);
}
// Wet-node wall edge with inward normals (A, GA), (B, GB) (gpuBcComplexWall3D,
// S/GpuBoundaries.cu:863-1365): bounce-back of the normal and in-plane
// unknowns, N = -(f(+t) - f(-t))/4 on the four unknowns with a tangential
// component, the two buried diagonals S/22 each and the rest population
// 12 S/22 with S the sum of the other 16.
template <class L, class T, int A, int GA, int B, int GB>
__device__ inline void hh_edge(T* f) {
    constexpr int C = 3 - A - B;
    constexpr int na = dir_ax<L>(A, GA, A, 0), na_ = dir_ax<L>(A, -GA, A, 0);
    constexpr int nb = dir_ax<L>(B, GB, B, 0), nb_ = dir_ax<L>(B, -GB, B, 0);
    constexpr int ab = dir_ax<L>(A, GA, B, GB), ab_ = dir_ax<L>(A, -GA, B, -GB);
    constexpr int tp = dir_ax<L>(C, 1, C, 0), tm = dir_ax<L>(C, -1, C, 0);
    constexpr int ap = dir_ax<L>(A, GA, C, 1), ap_ = dir_ax<L>(A, -GA, C, -1);
    constexpr int am = dir_ax<L>(A, GA, C, -1), am_ = dir_ax<L>(A, -GA, C, 1);
    constexpr int bp = dir_ax<L>(B, GB, C, 1), bp_ = dir_ax<L>(B, -GB, C, -1);
    constexpr int bm = dir_ax<L>(B, GB, C, -1), bm_ = dir_ax<L>(B, -GB, C, 1);
    constexpr int q1 = dir_ax<L>(A, GA, B, -GB), q2 = dir_ax<L>(A, -GA, B, GB);
    f[na] = f[na_];
    f[nb] = f[nb_];
    f[ab] = f[ab_];
    const T N = (T)-0.25 * (f[tp] - f[tm]);
    f[ap] = f[ap_] + N;
    f[am] = f[am_] - N;
    f[bp] = f[bp_] + N;
    f[bm] = f[bm_] - N;
    T S = 0;
#pragma unroll
    for (int i = 1; i < L::Q; ++i)
        if (i != q1 && i != q2) S += f[i];
    f[q1] = S * (T)(1.0 / 22.0);
    f[q2] = f[q1];
    f[0] = S * (T)(12.0 / 22.0);
}
// Wet-node wall corner (S/GpuBoundaries.cu:1366-1560): bounce-back of the
// three normal and three in-plane unknowns, the six buried diagonals
// C = (sum of the 12 others)/18, rest 12 C.
template <class L, class T, int GX, int GY, int GZ>
__device__ inline void hh_corner(T* f) {
    constexpr int x = dir_ax<L>(0, GX, 0, 0), x_ = dir_ax<L>(0, -GX, 0, 0);
    constexpr int y = dir_ax<L>(1, GY, 1, 0), y_ = dir_ax<L>(1, -GY, 1, 0);
    constexpr int z = dir_ax<L>(2, GZ, 2, 0), z_ = dir_ax<L>(2, -GZ, 2, 0);
    constexpr int xy = dir_ax<L>(0, GX, 1, GY), xy_ = dir_ax<L>(0, -GX, 1, -GY);
    constexpr int xz = dir_ax<L>(0, GX, 2, GZ), xz_ = dir_ax<L>(0, -GX, 2, -GZ);
    constexpr int yz = dir_ax<L>(1, GY, 2, GZ), yz_ = dir_ax<L>(1, -GY, 2, -GZ);
    constexpr int b0 = dir_ax<L>(0, GX, 1, -GY), b1 = dir_ax<L>(0, -GX, 1, GY);
    constexpr int b2 = dir_ax<L>(0, GX, 2, -GZ), b3 = dir_ax<L>(0, -GX, 2, GZ);
    constexpr int b4 = dir_ax<L>(1, GY, 2, -GZ), b5 = dir_ax<L>(1, -GY, 2, GZ);
    f[x] = f[x_]; f[y] = f[y_]; f[z] = f[z_];
    f[xy] = f[xy_]; f[xz] = f[xz_]; f[yz] = f[yz_];
    T S = 0;
#pragma unroll
    for (int i = 1; i < L::Q; ++i)
        if (i != b0 && i != b1 && i != b2 && i != b3 && i != b4 && i != b5) S += f[i];
    const T c = S * (T)(1.0 / 18.0);
    f[b0] = c; f[b1] = c; f[b2] = c; f[b3] = c; f[b4] = c; f[b5] = c;
    f[0] = S * (T)(12.0 / 18.0);
}
// Wet-node wall closure of a node word (bits kHHShift.. hold the id the host
// derived from the wall links: 1-6 face W,E,S,N,B,T; 7-18 edge, pair
// (x,y),(x,z),(y,z) times (GA<0)*2 + (GB<0); 19-26 corner, (GX<0)*4 +
// (GY<0)*2 + (GZ<0)).  Deviations in and out.
template <class L, class T>
__device__ inline void hh_walls(uint32_t nd, T* h) {
    if (L::Q != 19) return;
    const uint32_t id = (nd >> kHHShift) & 31u;
    if (!id) return;
    T f[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = h[i] + (T)L::w(i);
    const T Z[3] = {0, 0, 0};
    switch (id) {
    case 1: vface<L, T, 0, 1>(f, Z, false); break;
    case 2: vface<L, T, 0, -1>(f, Z, false); break;
    case 3: vface<L, T, 1, 1>(f, Z, false); break;
    case 4: vface<L, T, 1, -1>(f, Z, false); break;
    case 5: vface<L, T, 2, 1>(f, Z, false); break;
    case 6: vface<L, T, 2, -1>(f, Z, false); break;
    case 7: hh_edge<L, T, 0, 1, 1, 1>(f); break;
    case 8: hh_edge<L, T, 0, 1, 1, -1>(f); break;
    case 9: hh_edge<L, T, 0, -1, 1, 1>(f); break;
    case 10: hh_edge<L, T, 0, -1, 1, -1>(f); break;
    case 11: hh_edge<L, T, 0, 1, 2, 1>(f); break;
    case 12: hh_edge<L, T, 0, 1, 2, -1>(f); break;
    case 13: hh_edge<L, T, 0, -1, 2, 1>(f); break;
    case 14: hh_edge<L, T, 0, -1, 2, -1>(f); break;
    case 15: hh_edge<L, T, 1, 1, 2, 1>(f); break;
    case 16: hh_edge<L, T, 1, 1, 2, -1>(f); break;
    case 17: hh_edge<L, T, 1, -1, 2, 1>(f); break;
    case 18: hh_edge<L, T, 1, -1, 2, -1>(f); break;
    case 19: hh_corner<L, T, 1, 1, 1>(f); break;
    case 20: hh_corner<L, T, 1, 1, -1>(f); break;
    case 21: hh_corner<L, T, 1, -1, 1>(f); break;
    case 22: hh_corner<L, T, 1, -1, -1>(f); break;
    case 23: hh_corner<L, T, -1, 1, 1>(f); break;
    case 24: hh_corner<L, T, -1, 1, -1>(f); break;
    case 25: hh_corner<L, T, -1, -1, 1>(f); break;
    case 26: hh_corner<L, T, -1, -1, -1>(f); break;
    default: break;
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) h[i] = f[i] - (T)L::w(i);
}

// 3D velocity faces by node class; ubc/vbc/wbc hold the face velocity.
template <class L, class T>
__device__ inline void vel3d(const KParams& kp, const Fields& fl, size_t idx, uint32_t cls, T* h) {
    if (L::Q != 19) return;
    if (!(cls == VEL_W || cls == VEL_N || (cls >= VEL_E && cls <= VEL_T))) return;
    T f[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = h[i] + (T)L::w(i);
    const T U[3] = {(T)fl.ubc[idx], (T)fl.vbc[idx], (T)fl.wbc[idx]};
    const bool flip = (kp.parity & 0x10u) != 0;
    switch (cls) {
    case VEL_W: vface<L, T, 0, 1>(f, U, flip); break;
    case VEL_E: vface<L, T, 0, -1>(f, U, flip); break;
    case VEL_S: vface<L, T, 1, 1>(f, U, false); break;
    case VEL_N: vface<L, T, 1, -1>(f, U, false); break;
    case VEL_B: vface<L, T, 2, 1>(f, U, false); break;
    default: vface<L, T, 2, -1>(f, U, false); break;
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) h[i] = f[i] - (T)L::w(i);
}

// Symmetry planes (S/GpuBoundaries.cu:2086-2167): unknowns take their mirror
// images in the face (equal weights, so deviations copy directly).
template <class L, int A, int I> struct Mirror {
    static constexpr int v = dir_of<L>(A == 0 ? -L::cx(I) : L::cx(I), A == 1 ? -L::cy(I) : L::cy(I),
                                       A == 2 ? -L::cz(I) : L::cz(I));
};
template <class L, class T, int A, int G, int... I>
__device__ inline void sym_face_seq(T* h, std::integer_sequence<int, I...>) {
    ((G * cmp<L>(I, A) > 0 ? (void)(h[I] = h[Mirror<L, A, I>::v]) : (void)0), ...);
}
template <class L, class T, int A, int G>
__device__ inline void sym_face(T* h) {
    sym_face_seq<L, T, A, G>(h, std::make_integer_sequence<int, L::Q>{});
}
template <class L, class T>
__device__ inline void sym3d(uint32_t cls, T* h) {
    if (L::Q != 19 || cls < SYM_W || cls > SYM_T) return;
    T f[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = h[i];
    switch (cls) {
    case SYM_W: sym_face<L, T, 0, 1>(f); break;
    case SYM_E: sym_face<L, T, 0, -1>(f); break;
    case SYM_S: sym_face<L, T, 1, 1>(f); break;
    case SYM_N: sym_face<L, T, 1, -1>(f); break;
    case SYM_B: sym_face<L, T, 2, 1>(f); break;
    default: sym_face<L, T, 2, -1>(f); break;
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) h[i] = f[i];
}

__device__ inline bool is_outlet3(uint32_t cls) { return cls == PRES_E || (cls >= OUT_W && cls <= OUT_T); }

// Curved walls (Bouzidi, Firdaouss & Lallemand 2001, linear): after a
// half-way gather, h[opp d] = f*_d(x) for every wall link d; with the link's
// wall fraction q the unknown becomes
//   q < 1/2:  2q f*_d(x) + (1 - 2q) f*_d(x - c_d)   (f*_d(x - c_d) is the pulled h[d]),
//   q >= 1/2: f*_d(x) / 2q + (2q - 1)/2q f*_opp(d)(x),
// the textbook form of S/GpuBoundaries.cu:414-431 (whose q index and second
// node are off by one link).  Coefficients sum to one, so deviations combine
// directly.  q < 1/2 without a fluid node at x - c_d keeps bounce-back.
template <class L, int S>
__device__ inline void bouzidi(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp, uint32_t own,
                               uint32_t nd, const uint32_t* cn, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    const uint32_t slot = kp.qslot[own];
    if (slot == kNone) return;
    const double* q = kp.qv + (size_t)slot * 19;
    T g[L::Q];
#pragma unroll
    for (int i = 0; i < L::Q; ++i) g[i] = h[i];
#pragma unroll
    for (int d = 1; d < L::Q; ++d) {
        const int o = L::opp(d);
        if (!((nd >> d) & 1u)) continue;
        double qd = q[d];
        if (qd <= 0.0) qd = 0.5;
        const T fd = g[o];                       // f*_d(x), gathered for the unknown
        if (qd < 0.5) {
            if (((nd >> o) & 1u) || cn[d] == kNone) continue;
            h[o] = (T)(2.0 * qd) * fd + (T)(1.0 - 2.0 * qd) * g[d];
        } else {
            const T fo = Io<L, S>::ld(src, paddr<L>(kp, o, own), o);
            h[o] = (T)(0.5 / qd) * fd + (T)((2.0 * qd - 1.0) / (2.0 * qd)) * fo;
        }
    }
}

// Populations of a node after streaming and its inlet and wall conditions,
// i.e. everything before the outlet (the reference orders inlet, wall,
// outlet: S/Iterate.cu:468-473, S/Iterate3D.cu:602-621).
// h' <-> h = f - w of a D3Q19 MRT run (sign +1: add r, -1: subtract it).
template <class L, class T>
__device__ inline void mref_shift(T* h, int sign) {
#pragma unroll
    for (int i = 0; i < L::Q; ++i) h[i] += (T)sign * (T)mref_off<L>(i);
}

template <class L, int S, bool OPEN>
__device__ inline void inflow_state(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp,
                                    const Fields& fl, const uint32_t* __restrict__ cidx, uint32_t own,
                                    uint32_t nd, int x, int y, int z, size_t idx, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    uint32_t cn[L::Q];
    source_index<L, OPEN>(cidx, kp, x, y, z, cn);
    gather<L, S, OPEN>(src, kp, own, nd, cn, h);
    if (kp.wall == 3 && (nd & kBBMask)) bouzidi<L, S>(src, kp, own, nd, cn, h);
    const uint32_t cls = nd >> kClsShift;
    if (OPEN && L::Q == 9 && cls == VEL_W) zou_inlet<L, T>(kp, fl, idx, nd, h);
    if (OPEN && L::Q == 9 && cls == VEL_N) zou_lid<L, T>(kp, fl, idx, nd, h);
    // velocity faces and wet-node walls mix link speeds: they act on f - w
    const bool conv = L::Q == 19 && kp.mref && ((OPEN && cls != FLUID) || kp.wall == 2);
    if (conv) mref_shift<L, T>(h, -1);
    if (OPEN && L::Q == 19) vel3d<L, T>(kp, fl, idx, cls, h);
    if (kp.wall == 1 && (nd & kBBMask)) onnode_walls<L, T>(nd, h);
    if (L::Q == 19 && kp.wall == 2) hh_walls<L, T>(nd, h);
    if (conv) mref_shift<L, T>(h, 1);
}

// Extrapolating east outlet (S/GpuBoundaries.cu:1609-1640): f1, f5, f8 of an
// outlet node become 2 f(x-1) - f(x-2) (kp.outlet 2) or f(x-1) (kp.outlet 3),
// from the upstream nodes' populations after their own inlet and wall
// conditions, recomputed here (they lie off the east face, so no outlet
// condition applies to them).  The node keeps its streamed values when an
// upstream node is solid.  Node words are read from nodew at the compact
// index (compact) or at the dense index.
template <class L, int S, bool OPEN>
__device__ inline void outlet_extrap(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp,
                                     const Fields& fl, const uint32_t* __restrict__ cidx,
                                     const uint32_t* __restrict__ nodew, bool compact, int x, int y, int z,
                                     size_t idx, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    if (L::Q != 9 || kp.outlet < 2 || x < 2) return;
    const uint32_t a = cidx[idx - 1], b = cidx[idx - 2];
    if (a == kNone || b == kNone) return;
    T ha[L::Q];
    inflow_state<L, S, OPEN>(src, kp, fl, cidx, a, compact ? nodew[a] : nodew[idx - 1], x - 1, y, z, idx - 1, ha);
    if (kp.outlet == 2) {
        T hb[L::Q];
        inflow_state<L, S, OPEN>(src, kp, fl, cidx, b, compact ? nodew[b] : nodew[idx - 2], x - 2, y, z, idx - 2, hb);
        h[1] = (T)2 * ha[1] - hb[1]; h[5] = (T)2 * ha[5] - hb[5]; h[8] = (T)2 * ha[8] - hb[8];
    } else {
        h[1] = ha[1]; h[5] = ha[5]; h[8] = ha[8];
    }
}

// 3D outlet of face (A, G), model kp.outlet: 0 a textbook pressure face
// (rho_out, zero tangential velocity, the normal velocity from the known
// populations; the reference's 3D OUTLET is "not implemented",
// S/GpuBoundaries.cu:1671-1706); 2/3 the unknowns become 2 f(x+n) - f(x+2n)
// or f(x+n) of the upstream nodes' pre-outlet populations; 4 f_eq of the
// node's macros of the previous step (S/GpuBoundaries.cu:1840-1852).  The
// intended semantics of S/GpuBoundaries.cu:1655-1874: its population stride
// n*m (for n*m*h) and the downstream read of its second-order north face are
// not reproduced.  An extrapolating node keeps its streamed values when an
// upstream node is solid or outside.
template <class L, int S, bool OPEN, int A, int G>
__device__ inline void outlet3_face(const KParams& kp, const Fields& fl, size_t idx, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    if (kp.outlet == 0) {
        T f[L::Q];
        T K = 0;
#pragma unroll
        for (int i = 0; i < L::Q; ++i) {
            f[i] = h[i] + (T)L::w(i);
            if (cmp<L>(i, A) == 0) K += f[i];
            else if (cmp<L>(i, A) == -G) K += (T)2 * f[i];
        }
        T U[3] = {0, 0, 0};
        U[A] = (T)G * ((T)1 - K / (T)kp.rho_out);
        vface<L, T, A, G>(f, U, false);
#pragma unroll
        for (int i = 0; i < L::Q; ++i) h[i] = f[i] - (T)L::w(i);
        return;
    }
    if (kp.outlet == 4) {
        if (!fl.om_in) return;
        const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
        const T r = (T)fl.om_in[idx];
        const T u[3] = {(T)fl.om_in[n + idx], (T)fl.om_in[2 * n + idx], (T)fl.om_in[3 * n + idx]};
        const T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
#pragma unroll
        for (int i = 0; i < L::Q; ++i)
            if (G * cmp<L>(i, A) > 0) {
                const T cu = (T)L::cx(i) * u[0] + (T)L::cy(i) * u[1] + (T)L::cz(i) * u[2];
                h[i] = (T)L::w(i) * (r * ((T)1 + (T)3 * cu + (T)4.5 * cu * cu - (T)1.5 * uu) - (T)1);
            }
        return;
    }
}
template <class L, class T, int A, int G>
__device__ inline void extrap_mix(T* f, const T* ha, const T* hb, bool second) {
#pragma unroll
    for (int i = 0; i < L::Q; ++i)
        if (G * cmp<L>(i, A) > 0) f[i] = second ? (T)2 * ha[i] - hb[i] : ha[i];
}
template <class L, int S, bool OPEN>
__device__ inline void outlet3d(const typename Io<L, S>::raw* __restrict__ src, const KParams& kp, const Fields& fl,
                                const uint32_t* __restrict__ cidx, const uint32_t* __restrict__ nodew, bool compact,
                                int x, int y, int z, size_t idx, uint32_t cls, typename Io<L, S>::T* h) {
    using T = typename Io<L, S>::T;
    if (L::Q != 19 || !is_outlet3(cls)) return;
    T f[L::Q];   // a local copy keeps the caller's populations in registers across the switch
#pragma unroll
    for (int i = 0; i < L::Q; ++i) f[i] = h[i];
    if (kp.outlet == 2 || kp.outlet == 3) {
        // Extrapolation: the two upstream nodes along the inward normal are
        // recomputed once, outside the per-face code.
        const int A = cls == PRES_E || cls == OUT_W ? 0 : (cls == OUT_S || cls == OUT_N ? 1 : 2);
        const int G = cls == OUT_W || cls == OUT_S || cls == OUT_B ? 1 : -1;
        const bool second = kp.outlet == 2;
        const int dims[3] = {kp.nx, kp.ny, kp.nz};
        const int pa = A == 0 ? x : A == 1 ? y : z;
        const int reach = pa + G * (second ? 2 : 1);
        if (reach < 0 || reach >= dims[A < 1 ? 0 : (A < 2 ? 1 : 2)]) return;
        const size_t stride = A == 0 ? 1 : A == 1 ? (size_t)kp.nx : (size_t)kp.nx * kp.ny;
        const size_t i1 = G > 0 ? idx + stride : idx - stride;
        const size_t i2 = G > 0 ? idx + 2 * stride : idx - 2 * stride;
        const uint32_t a = cidx[i1];
        const uint32_t b = second ? cidx[i2] : 0u;
        if (a == kNone || b == kNone) return;
        const int dx = A == 0 ? G : 0, dy = A == 1 ? G : 0, dz = A == 2 ? G : 0;
        T ha[L::Q], hb[L::Q];
        inflow_state<L, S, OPEN>(src, kp, fl, cidx, a, compact ? nodew[a] : nodew[i1], x + dx, y + dy, z + dz, i1, ha);
        if (second)
            inflow_state<L, S, OPEN>(src, kp, fl, cidx, b, compact ? nodew[b] : nodew[i2], x + 2 * dx, y + 2 * dy,
                                     z + 2 * dz, i2, hb);
        switch (cls) {
        case PRES_E: extrap_mix<L, T, 0, -1>(f, ha, hb, second); break;
        case OUT_W: extrap_mix<L, T, 0, 1>(f, ha, hb, second); break;
        case OUT_S: extrap_mix<L, T, 1, 1>(f, ha, hb, second); break;
        case OUT_N: extrap_mix<L, T, 1, -1>(f, ha, hb, second); break;
        case OUT_B: extrap_mix<L, T, 2, 1>(f, ha, hb, second); break;
        default: extrap_mix<L, T, 2, -1>(f, ha, hb, second); break;
        }
    } else {
        // the pressure and f_eq closures act on f - w (extrapolations, whose
        // coefficients sum to one, do not need it)
        if (kp.mref) mref_shift<L, T>(f, -1);
        switch (cls) {
        case PRES_E: outlet3_face<L, S, OPEN, 0, -1>(kp, fl, idx, f); break;
        case OUT_W: outlet3_face<L, S, OPEN, 0, 1>(kp, fl, idx, f); break;
        case OUT_S: outlet3_face<L, S, OPEN, 1, 1>(kp, fl, idx, f); break;
        case OUT_N: outlet3_face<L, S, OPEN, 1, -1>(kp, fl, idx, f); break;
        case OUT_B: outlet3_face<L, S, OPEN, 2, 1>(kp, fl, idx, f); break;
        default: outlet3_face<L, S, OPEN, 2, -1>(kp, fl, idx, f); break;
        }
        if (kp.mref) mref_shift<L, T>(f, 1);
    }
#pragma unroll
    for (int i = 0; i < L::Q; ++i) h[i] = f[i];
}

__device__ inline double recip(double x) { return 1.0 / x; }
__device__ inline float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// Moments of h and the stored-velocity semantics of the force model
// (S/GpuUpdateMacro.cu:6-41, :137-213).
// A kernel parameter in the compute type: the FP32 copy for float kernels.
template <class T>
__device__ inline T kv(double d, float f) {
    if constexpr (sizeof(T) == sizeof(float)) return f;
    else return (T)d;
}

// Moments of the deviations h_i = sc(i) x_i (sc = 1: x holds h itself; the
// 16-bit storages pass their stored values x_i = h_i / a_i with sc = a_i).
// The moments and collisions contract multiply-adds only within one statement
// (each written as one), so every kernel that inlines them computes the same
// bits whatever code surrounds them.
template <class L, class T, class Sc>
__device__ inline void moments(const KParams& kp, uint32_t cls, const T* x, T& drho, T& rho, T* u, Sc sc) {
#pragma clang fp contract(on)
    T d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int i = 0; i < L::Q; ++i) {
        d += sc(i) * x[i];
        if (L::cx(i)) jx += (L::cx(i) > 0 ? sc(i) : -sc(i)) * x[i];
        if (L::cy(i)) jy += (L::cy(i) > 0 ? sc(i) : -sc(i)) * x[i];
        if (L::cz(i)) jz += (L::cz(i) > 0 ? sc(i) : -sc(i)) * x[i];
    }
    drho = d;
    rho = (T)1 + d;
    const T ir = recip(rho);
    T s0 = 0, s1 = 0, s2 = 0;
    if (kp.force == 1) {
        const T io = (T)(1.0 / kp.omega);
        s0 = (T)kp.g[0] * io; s1 = (T)kp.g[1] * io; s2 = (T)kp.g[2] * io;
    } else if (kp.force == 2) {
        s0 = kv<T>(0.5 * kp.g[0], kp.hgf[0]); s1 = kv<T>(0.5 * kp.g[1], kp.hgf[1]);
        s2 = kv<T>(0.5 * kp.g[2], kp.hgf[2]);
    }
    u[0] = jx * ir + s0; u[1] = jy * ir + s1; u[2] = jz * ir + s2;
    if ((kp.parity & 0x8u) && cls == PRES_E) u[1] = 0;
}
template <class L, class T>
__device__ inline void moments(const KParams& kp, uint32_t cls, const T* h, T& drho, T& rho, T* u) {
    moments<L, T>(kp, cls, h, drho, rho, u, [](int) { return (T)1; });
}

// Velocity-face nodes with frozen macros (fl.frz set): the collision and the
// macros written use the stored values in place of the node's moments, as
// the reference's 3D macro update skips inlet nodes when UpdateInltOutl = no
// (S/GpuUpdateMacro.cu:150-151).
template <class T>
__device__ inline void frozen_macros(const KParams& kp, const Fields& fl, uint32_t cls, size_t idx, T& drho, T& rho,
                                     T* u) {
    if (!fl.frz || !(cls == VEL_W || cls == VEL_N || (cls >= VEL_E && cls <= VEL_T))) return;
    const size_t n = (size_t)kp.nx * kp.ny * kp.nz;
    const double r = fl.frz[idx];
    drho = (T)(r - 1.0);
    rho = (T)r;
    u[0] = (T)fl.frz[n + idx]; u[1] = (T)fl.frz[2 * n + idx]; u[2] = (T)fl.frz[3 * n + idx];
}

// c_i . (a, b, c) with the zero components of c_i left out: a product with
// zero is not folded away by the compiler (0 * x is not 0 for every float x),
// and the result is the same for finite operands.
template <class L, class T>
__device__ inline T cdot(int i, T a, T b, T c) {
#pragma clang fp contract(on)
    T r = 0;
    bool any = false;
    const int k[3] = {L::cx(i), L::cy(i), L::cz(i)};
    const T v[3] = {a, b, c};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (k[d] == 0) continue;
        const T t = k[d] > 0 ? v[d] : -v[d];
        r = any ? r + t : t;
        any = true;
    }
    return r;
}

// Deviations of the MRT equilibrium moments from the rest moments (M w).
// D2Q9: S/GpuCollision.cu:986-994; D3Q19: :1065-1083.
template <class L, class T>
__device__ __forceinline__ void mrt_dmeq(const KParams& kp, T drho, T rho, const T* u, T* dm) {
#pragma clang fp contract(on)
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
        // in h' = h + r (kp.mref) the equilibrium deviation of epsilon loses
        // the rest moment: M r = 3 e_eps = mrest(eps) e_eps
        dm[2] = kp.mref ? (T)(-475.0 / 63.0) * j2 : (T)(-475.0 / 63.0) * j2 - (T)mrest<D3Q19>(2);
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
template <class L, class T, int C, class Wt>
__device__ inline void collide_pairs(const KParams& kp, const T* h, T drho, T rho, const T* u, T* out, Wt wt);

template <class L, class T, int C>
__device__ __forceinline__ void collide(const KParams& kp, const T* h, T drho, T rho, const T* u, T* out) {
#pragma clang fp contract(on)
    const bool guo = kp.force == 2;
    if (C == 2) {
        static_assert(mrows_have_parity<L>(), "MRT rows are even or odd, links come in pairs");
        constexpr int NP = (L::Q - 1) / 2;
        T s[NP], a[NP];
        pair_split<L, T>(h, s, a, std::make_integer_sequence<int, NP>{});
        T dm[L::Q], y[L::Q], F[3] = {0, 0, 0}, uF = 0;
        mrt_dmeq<L, T>(kp, drho, rho, u, dm);
        if (guo) {
            F[0] = rho * kv<T>(kp.g[0], kp.gf[0]); F[1] = rho * kv<T>(kp.g[1], kp.gf[1]);
            F[2] = rho * kv<T>(kp.g[2], kp.gf[2]);
            uF = u[0] * F[0] + u[1] * F[1] + u[2] * F[2];
        }
        mrt_relax<L, T>(kp, h[0], s, a, dm, guo, u, F, uF, y, std::make_integer_sequence<int, L::Q>{});
        mrt_back<L, T>(h, y, out, std::make_integer_sequence<int, NP>{});
        if (L::Q == 19 && (kp.parity & 0x2u))   // fColl_18 stays at its zero initial value
            out[L::Q - 1] = -(T)L::w(L::Q - 1) + (kp.mref ? (T)mref_off<L>(L::Q - 1) : (T)0);
        return;
    }
    collide_pairs<L, T, C>(kp, h, drho, rho, u, out, [](int i) { return (T)L::w(i); });
}

// BGK and TRT in symmetric/antisymmetric pair form: for a link pair
// (i, opp i) the parts of f_eq and of the Guo source are
//   e+ = w (drho + rho (4.5 cu^2 - 1.5 u.u)),  e- = 3 w rho cu,
//   S+ = w (9 cu cF - 3 u.F),                  S- = 3 w cF,
// relaxed with omega+ and omega- (equal for BGK) and prefactors 1 - omega/2:
//   out_i = h_i - op (h+ - e+) - om (h- - e-) + (1 - op/2) S+ + (1 - om/2) S-.
// Collected per pair this is
//   out_i   = a h_i + b h_opp + w (E + G),   out_opp = a h_opp + b h_i + w (E - G),
//   a = 1 - (op + om)/2,  b = (om - op)/2,
//   E = K0 + cu (P cu + Qf cF),  G = R cu + Sg cF,
// with the node constants K0 = op (drho - 1.5 rho u.u) - 3 (1 - op/2) u.F,
// P = 4.5 op rho, Qf = 9 (1 - op/2), R = 3 om rho, Sg = 3 (1 - om/2), scaled
// by w once per link class: about a dozen operations per pair instead of two
// dozen.  wt(i) is w_i for
// populations h; for stored values x_i = h_i / a_i with a_i = a_opp(i) the same
// update divided by a_i is the one with wt = w_i / a_i.
template <class L, class T, int C, class Wt>
__device__ inline void collide_pairs(const KParams& kp, const T* h, T drho, T rho, const T* u, T* out, Wt wt) {
#pragma clang fp contract(on)
    const bool guo = kp.force == 2;
    const T op = kv<T>(kp.omega, kp.omegaf);
    const T om = C == 0 ? op : kv<T>(kp.omega_m, kp.omega_mf);
    const T gp = (T)1 - (T)0.5 * op, gm = (T)1 - (T)0.5 * om;
    const T uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    T Fx = 0, Fy = 0, Fz = 0, uF = 0;
    if (guo) {
        Fx = rho * kv<T>(kp.g[0], kp.gf[0]); Fy = rho * kv<T>(kp.g[1], kp.gf[1]); Fz = rho * kv<T>(kp.g[2], kp.gf[2]);
        uF = u[0] * Fx + u[1] * Fy + u[2] * Fz;
    }
    const T a = (T)1 - (T)0.5 * (op + om), b = (T)0.5 * (om - op);
    const T K0 = op * (drho - (T)1.5 * rho * uu) - (T)3 * gp * uF;
    const T P = (T)4.5 * op * rho, Qf = (T)9 * gp, R = (T)3 * om * rho, Sg = (T)3 * gm;
    out[0] = ((T)1 - op) * h[0] + wt(0) * K0;
#pragma unroll
    for (int i = 1; i < L::Q; ++i) {
        const int o = L::opp(i);
        if (i < o) {
            const T w = wt(i);
            const T cu = cdot<L, T>(i, u[0], u[1], u[2]);
            // the source terms are applied unconditionally (zero without
            // forcing): a test per link would become a select per link
            const T cF = cdot<L, T>(i, Fx, Fy, Fz);
            // w is a literal per link class: the scaled constants are shared
            // by every pair of the class
            const T E = w * K0 + cu * ((w * P) * cu + (w * Qf) * cF);
            const T G = (w * R) * cu + (w * Sg) * cF;
            out[i] = a * h[i] + (b * h[o] + (E + G));
            out[o] = a * h[o] + (b * h[i] + (E - G));
        }
    }
}

}  // namespace lbm
