// latticegen2 lattice files: <name>_Node.dat and <name>_BC.dat.
//
// Formats (whitespace separated):
//   node, 2D: i j x y fluid                  (S/FilesReading.cu:290-291)
//   node, 3D: i j k x y z fluid              (L/main.cpp:59-66)
//   BC,   2D: i j dir type x y id            (S/FilesReading.cu:331-333)
//   BC,   3D: i j k dir type x y z id        (L/FluidTracking.cpp:229-236)
// The node-file column count decides 2D/3D (S/main.cu:109-140); n, m, h are
// the largest index + 1 (S/FilesReading.cu:469-471).  BC types: 1 WALL,
// 2 INLE, 3 OUTL, 4 CYCL, 5 SYMP (L/GeomRead.cpp:85-99).  Lines are split on
// '\n' only, so CRLF files count correctly (unlike S/FilesReading.cu:27-31).
//
// 2D node classes follow the reference's per-direction link masks
// (S/GpuInit.cu:1066-1116, S/include/BcMacros.h): an INLE link in direction 3
// on x = 0 makes a west velocity inlet, one in direction 2 on y = m-1 a north
// lid, an OUTL link in direction 1 on x = n-1 an east outlet.  A node whose
// links have different types is a corner: each diagonal whose two axis
// directions both carry links becomes a wall link.
//
// 3D (S/GpuInit.cu:1157-1238, S/include/BcMacros3D.h): per direction a
// 3-bit type; a node whose links have different types is a corner and all
// its links become wall links.  Other nodes take the class of the first face
// (W, E, S, N, B, T order, as S/GpuBoundaries.cu checks them) whose
// axis-direction link has exactly the INLE (velocity face), OUTL (outlet) or
// SYMP (symmetry) type, when the node lies on that face of the box.
#include "lbm.h"

#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <new>
#include <vector>

namespace {

bool slurp(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    return got == out.size();
}

// Parses up to maxc numbers of one line; returns the count, advances p.
int parse_line(const char*& p, const char* end, double* vals, int maxc) {
    int k = 0;
    while (p < end && *p != '\n') {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
        if (p >= end || *p == '\n') break;
        char* q = nullptr;
        double v = std::strtod(p, &q);
        if (q == p) {   // not a number: skip the token
            while (p < end && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') ++p;
            continue;
        }
        if (k < maxc) vals[k] = v;
        ++k;
        p = q;
    }
    if (p < end && *p == '\n') ++p;
    return k;
}

// One lattice in memory: node rows and BC rows as the files list them.
struct Lat {
    int cols = 0;                    // node columns: 5 (2D) or 7 (3D)
    std::vector<int64_t> i, j, k;    // node indices
    std::vector<double> x, y, z;     // node coordinates
    std::vector<uint8_t> fluid;
    bool have_bc = false;
    std::vector<int64_t> li, lj, lk; // BC rows
    std::vector<int8_t> dir, type;
    std::vector<double> bx, by, bz;
    std::vector<int32_t> id;
};

// Binary lattice file: magic, then little-endian fields
//   int32 cols, int32 has_bc, int64 nodes, int64 links,
//   nodes: int32 i[], j[], k[]; double x[], y[], z[]; uint8 fluid[]
//   links: int32 i[], j[], k[]; int8 dir[], type[]; double x[], y[], z[]; int32 id[]
constexpr char kMagic[8] = {'L', 'B', 'M', 'L', 'A', 'T', '0', '1'};

bool is_binary(const std::string& t) { return t.size() >= 8 && std::memcmp(t.data(), kMagic, 8) == 0; }

template <class T, class U>
bool take(const std::string& t, size_t& off, std::vector<U>& out, size_t cnt) {
    // division form: a huge count from a malformed header cannot wrap around
    if (off > t.size() || cnt > (t.size() - off) / sizeof(T)) return false;
    out.resize(cnt);
    for (size_t a = 0; a < cnt; ++a) {
        T v;
        std::memcpy(&v, t.data() + off + a * sizeof(T), sizeof(T));
        out[a] = (U)v;
    }
    off += cnt * sizeof(T);
    return true;
}

lbm_status load_binary(const std::string& t, Lat& L) {
    size_t off = 8;
    int32_t hdr[2];
    int64_t cnt[2];
    if (t.size() < 8 + 8 + 16) return LBM_EIO;
    std::memcpy(hdr, t.data() + off, 8); off += 8;
    std::memcpy(cnt, t.data() + off, 16); off += 16;
    if ((hdr[0] != 5 && hdr[0] != 7) || cnt[0] <= 0 || cnt[1] < 0) return LBM_EIO;
    L.cols = hdr[0];
    L.have_bc = hdr[1] != 0;
    const size_t nn = (size_t)cnt[0], nl = (size_t)cnt[1];
    bool ok = take<int32_t>(t, off, L.i, nn) && take<int32_t>(t, off, L.j, nn) && take<int32_t>(t, off, L.k, nn) &&
              take<double>(t, off, L.x, nn) && take<double>(t, off, L.y, nn) && take<double>(t, off, L.z, nn) &&
              take<uint8_t>(t, off, L.fluid, nn) && take<int32_t>(t, off, L.li, nl) &&
              take<int32_t>(t, off, L.lj, nl) && take<int32_t>(t, off, L.lk, nl) && take<int8_t>(t, off, L.dir, nl) &&
              take<int8_t>(t, off, L.type, nl) && take<double>(t, off, L.bx, nl) && take<double>(t, off, L.by, nl) &&
              take<double>(t, off, L.bz, nl) && take<int32_t>(t, off, L.id, nl);
    if (!ok || off != t.size()) return LBM_EIO;
    // a 2D lattice (5 node columns) has no z index
    if (L.cols == 5) {
        for (int64_t v : L.k) if (v != 0) return LBM_EIO;
        for (int64_t v : L.lk) if (v != 0) return LBM_EIO;
    }
    return LBM_OK;
}

lbm_status load_impl(const char* node_path, const char* bc_path, Lat& L);

// Node file (text or binary) and, for a text node file, the BC text file.
// Allocation failures come back as a status, never as an exception through
// the C entry points.
lbm_status load(const char* node_path, const char* bc_path, Lat& L) {
    try {
        return load_impl(node_path, bc_path, L);
    } catch (const std::bad_alloc&) {
        return LBM_ENOMEM;
    } catch (...) {
        return LBM_EIO;
    }
}

lbm_status load_impl(const char* node_path, const char* bc_path, Lat& L) {
    std::string txt;
    if (!slurp(node_path, txt)) return LBM_EIO;
    if (is_binary(txt)) return load_binary(txt, L);
    const char* p = txt.data();
    const char* end = p + txt.size();
    double v[16];
    while (p < end) {
        int c = parse_line(p, end, v, 16);
        if (c == 0) continue;
        if (L.cols == 0) L.cols = c;
        if (c != L.cols || (L.cols != 5 && L.cols != 7)) return LBM_EIO;
        const bool t3 = L.cols == 7;
        L.i.push_back((int64_t)v[0]); L.j.push_back((int64_t)v[1]); L.k.push_back(t3 ? (int64_t)v[2] : 0);
        L.x.push_back(v[t3 ? 3 : 2]); L.y.push_back(v[t3 ? 4 : 3]); L.z.push_back(t3 ? v[5] : 0.0);
        L.fluid.push_back((int)v[L.cols - 1] ? 1 : 0);
    }
    if (L.i.empty()) return LBM_EIO;
    L.have_bc = bc_path && bc_path[0];
    if (!L.have_bc) return LBM_OK;
    std::string bct;
    if (!slurp(bc_path, bct)) return LBM_EIO;
    const bool three_d = L.cols == 7;
    const int want = three_d ? 9 : 7;
    p = bct.data();
    end = p + bct.size();
    while (p < end) {
        int c = parse_line(p, end, v, 16);
        if (c == 0) continue;
        if (c != want) return LBM_EIO;
        L.li.push_back((int64_t)v[0]); L.lj.push_back((int64_t)v[1]); L.lk.push_back(three_d ? (int64_t)v[2] : 0);
        const int d = (int)v[three_d ? 3 : 2], t = (int)v[three_d ? 4 : 3];
        if (d < -128 || d > 127 || t < -128 || t > 127) return LBM_EIO;
        L.dir.push_back((int8_t)d); L.type.push_back((int8_t)t);
        L.bx.push_back(v[three_d ? 5 : 4]); L.by.push_back(v[three_d ? 6 : 5]); L.bz.push_back(three_d ? v[7] : 0.0);
        L.id.push_back((int32_t)v[three_d ? 8 : 6]);
    }
    return LBM_OK;
}

}  // namespace

extern "C" lbm_status lbm_read_lattice(const char* node_path, const char* bc_path, int32_t dims[3],
                                       int64_t* nlinks, uint8_t* node_class, uint32_t* bb_links,
                                       int64_t type_counts[6], int64_t dir_counts[19], double* delta,
                                       int32_t cyclic[3]) {
    if (!node_path || !dims) return LBM_EINVAL;
    Lat L;
    lbm_status st = load(node_path, bc_path, L);
    if (st != LBM_OK) return st;
    int64_t mx[3] = {-1, -1, -1};
    for (size_t r = 0; r < L.i.size(); ++r) {
        if (L.i[r] < 0 || L.j[r] < 0 || L.k[r] < 0) return LBM_EIO;
        if (L.i[r] > mx[0]) mx[0] = L.i[r];
        if (L.j[r] > mx[1]) mx[1] = L.j[r];
        if (L.k[r] > mx[2]) mx[2] = L.k[r];
    }
    dims[0] = (int32_t)(mx[0] + 1); dims[1] = (int32_t)(mx[1] + 1); dims[2] = (int32_t)(mx[2] + 1);
    const int64_t nx = dims[0], ny = dims[1];
    const int64_t n = nx * ny * dims[2];
    const bool three_d = L.cols == 7;
    const bool have_bc = L.have_bc;
    if (!node_class) {   // first pass: sizes only
        if (nlinks) *nlinks = have_bc ? (int64_t)L.li.size() : 0;
        return LBM_OK;
    }
    if (delta) {   // |x(1,0,0) - x(0,0,0)| (S/FilesReading.cu:473-490)
        double x0 = 0, x1 = 0;
        bool h0 = false, h1 = false;
        for (size_t r = 0; r < L.i.size(); ++r) {
            if (L.j[r] == 0 && L.k[r] == 0 && L.i[r] == 0) { x0 = L.x[r]; h0 = true; }
            if (L.j[r] == 0 && L.k[r] == 0 && L.i[r] == 1) { x1 = L.x[r]; h1 = true; }
        }
        *delta = (h0 && h1) ? std::fabs(x1 - x0) : 1.0;
    }
    if (cyclic) cyclic[0] = cyclic[1] = cyclic[2] = 0;
    std::memset(node_class, 0, (size_t)n);
    if (bb_links) std::memset(bb_links, 0, sizeof(uint32_t) * (size_t)n);
    if (type_counts) std::memset(type_counts, 0, sizeof(int64_t) * 6);
    if (dir_counts) std::memset(dir_counts, 0, sizeof(int64_t) * 19);
    for (size_t r = 0; r < L.i.size(); ++r)
        node_class[L.i[r] + L.j[r] * nx + L.k[r] * nx * ny] = L.fluid[r] ? LBM_FLUID : LBM_SOLID;
    int64_t cnt = 0;
    // 2D: first link type per node, corner flag, 2-bit type per direction
    std::vector<uint8_t> main_t, corner;
    std::vector<uint32_t> dmask;
    std::vector<uint64_t> dmask3;
    if (have_bc) { main_t.assign((size_t)n, 0); corner.assign((size_t)n, 0); }
    if (!three_d && have_bc) dmask.assign((size_t)n, 0);
    if (three_d && have_bc) dmask3.assign((size_t)n, 0);
    if (have_bc) {
        const int qmax = three_d ? 19 : 9;
        for (size_t r = 0; r < L.li.size(); ++r) {
            const int64_t i = L.li[r], j = L.lj[r], k = L.lk[r];
            const int dir = L.dir[r], type = L.type[r];
            if (i < 0 || i >= nx || j < 0 || j >= ny || k < 0 || k >= dims[2]) return LBM_EIO;
            if (dir < 1 || dir >= qmax || type < 0 || type > 5) return LBM_EIO;
            if (type_counts) type_counts[type]++;
            if (dir_counts) dir_counts[dir]++;
            const int64_t at = i + j * nx + k * nx * ny;
            if (type == 1 && bb_links) bb_links[at] |= 1u << dir;
            if (!three_d && type >= 1 && type <= 3) {
                dmask[at] |= (uint32_t)type << (2 * (dir - 1));
                if (!main_t[at]) main_t[at] = (uint8_t)type;
                else if (main_t[at] != type) corner[at] = 1;
            }
            if (three_d && type >= 1) {
                dmask3[at] |= (uint64_t)type << (3 * dir);
                if (!main_t[at]) main_t[at] = (uint8_t)type;
                else if (main_t[at] != type) corner[at] = 1;
            }
            if (type == 4 && cyclic) {
                static const int c9[9][2] = {{0, 0}, {1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
                static const int c19[19][3] = {{0, 0, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1},
                                               {0, 0, -1}, {1, 1, 0}, {-1, 1, 0}, {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},
                                               {-1, 0, 1}, {1, 0, -1}, {-1, 0, -1}, {0, 1, 1}, {0, -1, 1}, {0, 1, -1},
                                               {0, -1, -1}};
                const int cx = three_d ? c19[dir][0] : c9[dir][0], cy = three_d ? c19[dir][1] : c9[dir][1];
                const int cz = three_d ? c19[dir][2] : 0;
                if ((cx < 0 && i == 0) || (cx > 0 && i == nx - 1)) cyclic[0] = 1;
                if ((cy < 0 && j == 0) || (cy > 0 && j == ny - 1)) cyclic[1] = 1;
                if ((cz < 0 && k == 0) || (cz > 0 && k == dims[2] - 1)) cyclic[2] = 1;
            }
            ++cnt;
        }
    }
    if (!three_d && have_bc) {
        // direction d's 2-bit field: bits 2(d-1), 2(d-1)+1
        auto any = [&](int64_t at, int d) { return ((dmask[at] >> (2 * (d - 1))) & 3u) != 0; };
        auto is = [&](int64_t at, int d, uint32_t t) { return ((dmask[at] >> (2 * (d - 1))) & 3u) == t; };
        for (int64_t at = 0; at < n; ++at) {
            if (!main_t[at] || node_class[at] == LBM_SOLID) continue;
            const int64_t i = at % nx, j = at / nx;
            if (corner[at] && bb_links) {
                // diagonal d between axis directions a and b: NE(5) E,N; NW(6) W,N; SW(7) W,S; SE(8) E,S
                static const int dg[4][3] = {{5, 1, 2}, {6, 3, 2}, {7, 3, 4}, {8, 1, 4}};
                for (const auto& e : dg)
                    if (any(at, e[1]) && any(at, e[2])) bb_links[at] |= 1u << e[0];
            }
            if (i == 0 && is(at, 3, 2)) node_class[at] = LBM_VEL_W;
            else if (j == ny - 1 && is(at, 2, 2)) node_class[at] = LBM_VEL_N;
            else if (i == nx - 1 && is(at, 1, 3)) node_class[at] = LBM_PRES_E;
        }
    }
    if (three_d && have_bc) {
        const int64_t ny = dims[1], nz = dims[2];
        // faces in the reference's order: axis direction, class of a velocity
        // face, outlet, symmetry plane, and the face test
        struct Face { int dir, vel, out, sym, axis, high; };
        static const Face F[6] = {{2, LBM_VEL_W, LBM_OUT_W, LBM_SYM_W, 0, 0}, {1, LBM_VEL_E, LBM_PRES_E, LBM_SYM_E, 0, 1},
                                  {4, LBM_VEL_S, LBM_OUT_S, LBM_SYM_S, 1, 0}, {3, LBM_VEL_N, LBM_OUT_N, LBM_SYM_N, 1, 1},
                                  {6, LBM_VEL_B, LBM_OUT_B, LBM_SYM_B, 2, 0}, {5, LBM_VEL_T, LBM_OUT_T, LBM_SYM_T, 2, 1}};
        for (int64_t at = 0; at < n; ++at) {
            if (!main_t[at] || node_class[at] == LBM_SOLID) continue;
            if (corner[at]) {
                if (bb_links)
                    for (int d = 1; d < 19; ++d)
                        if ((dmask3[at] >> (3 * d)) & 7u) bb_links[at] |= 1u << d;
                continue;
            }
            const int64_t xyz[3] = {at % nx, (at / nx) % ny, at / (nx * ny)};
            const int64_t top[3] = {nx - 1, ny - 1, nz - 1};
            for (const Face& f : F) {
                const uint32_t t = (uint32_t)(dmask3[at] >> (3 * f.dir)) & 7u;
                if (xyz[f.axis] != (f.high ? top[f.axis] : 0)) continue;
                const int cls = t == 2 ? f.vel : t == 3 ? f.out : t == 5 ? f.sym : 0;
                if (cls) { node_class[at] = (uint8_t)cls; break; }
            }
        }
    }
    if (nlinks) *nlinks = cnt;
    return LBM_OK;
}

// Wall fractions of the WALL links: q = |x_bc - x_node| / (delta |c_d|), the
// BC line's point being where the link meets the wall (S/GpuInit.cu:1088, :1205).
extern "C" lbm_status lbm_read_wall_q(const char* node_path, const char* bc_path, double* wall_q) {
    if (!node_path || !wall_q) return LBM_EINVAL;
    Lat L;
    lbm_status st = load(node_path, bc_path, L);
    if (st != LBM_OK) return st;
    if (!L.have_bc) return LBM_EINVAL;
    int32_t dims[3];
    double delta = 1.0;
    {   // sizes and grid spacing as lbm_read_lattice reports them
        st = lbm_read_lattice(node_path, nullptr, dims, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (st != LBM_OK) return st;
        std::vector<uint8_t> cls((size_t)dims[0] * dims[1] * dims[2]);
        st = lbm_read_lattice(node_path, nullptr, dims, nullptr, cls.data(), nullptr, nullptr, nullptr, &delta, nullptr);
        if (st != LBM_OK) return st;
        if (!(delta > 0)) delta = 1.0;
    }
    const int64_t nx = dims[0], ny = dims[1], n = nx * ny * dims[2];
    const bool three_d = L.cols == 7;
    std::vector<double> xyz((size_t)n * 3, 0.0);
    for (size_t r = 0; r < L.i.size(); ++r) {
        const int64_t at = L.i[r] + L.j[r] * nx + L.k[r] * nx * ny;
        if (at < 0 || at >= n) return LBM_EIO;
        xyz[3 * at] = L.x[r]; xyz[3 * at + 1] = L.y[r]; xyz[3 * at + 2] = L.z[r];
    }
    std::memset(wall_q, 0, sizeof(double) * 19 * (size_t)n);
    for (size_t r = 0; r < L.li.size(); ++r) {
        const int dir = L.dir[r];
        if (L.type[r] != 1) continue;
        const int64_t at = L.li[r] + L.lj[r] * nx + L.lk[r] * nx * ny;
        if (at < 0 || at >= n || dir < 1 || dir >= (three_d ? 19 : 9)) return LBM_EIO;
        const double dx = L.bx[r] - xyz[3 * at], dy = L.by[r] - xyz[3 * at + 1], dz = L.bz[r] - xyz[3 * at + 2];
        // |c_d|: 1 on the axes, sqrt(2) on the diagonals (D2Q9 5-8, D3Q19 7-18)
        const double len = (three_d ? dir >= 7 : dir >= 5) ? std::sqrt(2.0) : 1.0;
        wall_q[at * 19 + dir] = std::sqrt(dx * dx + dy * dy + dz * dz) / (delta * len);
    }
    return LBM_OK;
}

// Binary copy of a lattice (node file + BC file) in the layout of kMagic.
extern "C" lbm_status lbm_write_lattice_binary(const char* node_path, const char* bc_path, const char* out_path) {
    if (!node_path || !out_path) return LBM_EINVAL;
    Lat L;
    lbm_status st = load(node_path, bc_path, L);
    if (st != LBM_OK) return st;
    std::string b(kMagic, 8);
    auto put = [&b](const void* p, size_t bytes) { b.append((const char*)p, bytes); };
    const int32_t hdr[2] = {L.cols, L.have_bc ? 1 : 0};
    const int64_t cnt[2] = {(int64_t)L.i.size(), (int64_t)L.li.size()};
    put(hdr, 8);
    put(cnt, 16);
    auto put32 = [&](const std::vector<int64_t>& v) {
        for (int64_t a : v) {
            if (a < INT32_MIN || a > INT32_MAX) return false;
            const int32_t w = (int32_t)a;
            put(&w, 4);
        }
        return true;
    };
    if (!put32(L.i) || !put32(L.j) || !put32(L.k)) return LBM_EUNSUPPORTED;
    put(L.x.data(), 8 * L.x.size()); put(L.y.data(), 8 * L.y.size()); put(L.z.data(), 8 * L.z.size());
    put(L.fluid.data(), L.fluid.size());
    if (!put32(L.li) || !put32(L.lj) || !put32(L.lk)) return LBM_EUNSUPPORTED;
    put(L.dir.data(), L.dir.size()); put(L.type.data(), L.type.size());
    put(L.bx.data(), 8 * L.bx.size()); put(L.by.data(), 8 * L.by.size()); put(L.bz.data(), 8 * L.bz.size());
    put(L.id.data(), 4 * L.id.size());
    FILE* f = std::fopen(out_path, "wb");
    if (!f) return LBM_EIO;
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    return std::fclose(f) == 0 && ok ? LBM_OK : LBM_EIO;
}
