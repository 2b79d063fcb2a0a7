// latticegen2-compatible lattice generator: STAR-CD boundary geometry
// (<name>.vrt vertices, <name>.bnd quadrilateral faces) to the lattice files
// <out>_Node.dat and <out>_BC.dat that lbm_read_lattice reads.
//
// Restates the reference generator (L/main.cpp:17-70, L/GeomRead.cpp:14-189,
// L/MeshGen.cpp:19-80, L/FluidTracking.cpp:24-236, L/Intersect.cpp:43-528)
// so that its output files are the same, byte for byte:
//  * mesh: delta = (x_max - x_min)/(n - 1) over the vertices' bounding box,
//    m = ceil(y_extent/delta) + 1, h = ceil(z_extent/delta) + 1 nodes, plus
//    one padding layer on every side (node (i, j, k) at x_min + delta (i-1), ...);
//  * fluid: a breadth-first fill over the D3Q19 links from the mesh node
//    nearest the seed point; a link that crosses a face is a boundary link;
//  * faces: only axis-aligned quadrilaterals cut links (the reference leaves
//    sloping faces as a to-do); the first face in file order that a link
//    crosses decides its type and intersection point.  The crossing tests
//    for diagonal links keep the reference's rules, including an
//    intersection coordinate that is left from an earlier test when neither
//    of its two update rules applies (state carried across faces and links).
//  * complexity 1 writes a boundary link while filling, when its far node is
//    not yet fluid; complexity 2 writes them in a second pass over the fluid
//    nodes once the fill is complete.
// Face lists per link class skip the faces that neither match nor update the
// carried state, so each test costs O(faces of the relevant orientations).
#include "lbm.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Face {
    int id = 0;       // BCID column
    int type = 0;     // 1 WALL, 2 INLE, 3 OUTL, 4 CYCL, 5 SYMP, 0 other
    int plane = 0;    // 3: x = const, 2: y = const, 1: z = const, 0: sloping
    double p2[3] = {0, 0, 0};                    // second listed vertex
    double xmax = 0, xmin = 0, ymax = 0, ymin = 0, zmax = 0, zmin = 0;
};

bool slurp(const char* path, std::string& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? std::fread(&out[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    out.push_back('\0');
    return got + 1 == out.size();
}

// Stream-style field readers: skip white space, then the longest number
// prefix (as formatted extraction does; "+1.0e+00+2.0e+00" is two numbers).
const char* skip_ws(const char* p) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f') ++p;
    return p;
}
bool read_int(const char*& p, long& v) {
    p = skip_ws(p);
    char* q = nullptr;
    v = std::strtol(p, &q, 10);
    if (q == p) return false;
    p = q;
    return true;
}
bool read_double(const char*& p, double& v) {
    p = skip_ws(p);
    char* q = nullptr;
    v = std::strtod(p, &q);
    if (q == p) return false;
    p = q;
    return true;
}
bool read_word(const char*& p, std::string& w) {
    p = skip_ws(p);
    const char* s = p;
    while (*p && !(*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
    w.assign(s, p);
    return p != s;
}

class Generator {
public:
    std::vector<double> vx, vy, vz;   // vertices
    std::vector<Face> faces;
    int n = 0, m = 0, h = 0;          // nodes per axis (without padding)
    double xmin = 0, ymin = 0, zmin = 0, delta = 0;
    std::vector<uint8_t> flag;        // node written as fluid
    std::vector<uint8_t> reached;     // node appended by the fill (the seed is not)
    std::vector<int> order;           // fill order
    // classes: 0 z-links, 1 x-links, 2 y-links, 3 xy, 4 xz, 5 yz diagonals
    std::vector<int> cls_faces[6];
    double I[3] = {0, 0, 0};          // carried intersection point
    int hit_id = -1, hit_type = -1;
    double hit[3] = {0, 0, 0};

    int P() const { return n + 2; }
    int Q() const { return m + 2; }
    double cx(int i) const { return xmin + delta * (i - 1); }
    double cy(int j) const { return ymin + delta * (j - 1); }
    double cz(int k) const { return zmin + delta * (k - 1); }

    bool read(const char* vrt, const char* bnd, std::string& err) {
        std::string t;
        if (!slurp(vrt, t)) { err = std::string("cannot read ") + vrt; return false; }
        const char* p = t.c_str();
        for (;;) {
            long id;
            double a, b, c;
            if (!read_int(p, id) || !read_double(p, a) || !read_double(p, b) || !read_double(p, c)) break;
            vx.push_back(a); vy.push_back(b); vz.push_back(c);
        }
        if (vx.empty()) { err = "no vertices"; return false; }
        if (!slurp(bnd, t)) { err = std::string("cannot read ") + bnd; return false; }
        p = t.c_str();
        const long nv = (long)vx.size();
        for (;;) {
            long i1, pt[4], bcid, i2;
            std::string word;
            if (!read_int(p, i1) || !read_int(p, pt[0]) || !read_int(p, pt[1]) || !read_int(p, pt[2]) ||
                !read_int(p, pt[3]) || !read_int(p, bcid) || !read_int(p, i2) || !read_word(p, word))
                break;
            Face f;
            for (long q : pt)
                if (q < 1 || q > nv) { err = "face vertex out of range"; return false; }
            f.id = (int)bcid;
            f.type = word == "WALL" ? 1 : word == "INLE" ? 2 : word == "OUTL" ? 3 : word == "CYCL" ? 4
                   : word == "SYMP" ? 5 : 0;
            const long a = pt[0] - 1, b = pt[1] - 1, c = pt[2] - 1, d = pt[3] - 1;
            if (vx[a] == vx[b] && vx[b] == vx[c] && vx[c] == vx[d]) f.plane = 3;
            else if (vy[a] == vy[b] && vy[b] == vy[c] && vy[c] == vy[d]) f.plane = 2;
            else if (vz[a] == vz[b] && vz[b] == vz[c] && vz[c] == vz[d]) f.plane = 1;
            f.p2[0] = vx[b]; f.p2[1] = vy[b]; f.p2[2] = vz[b];
            auto mx = [](double u, double v, double w, double z) { return std::max(std::max(std::max(v, u), w), z); };
            auto mn = [](double u, double v, double w, double z) { return std::min(std::min(std::min(v, u), w), z); };
            f.xmax = mx(vx[a], vx[b], vx[c], vx[d]); f.xmin = mn(vx[a], vx[b], vx[c], vx[d]);
            f.ymax = mx(vy[a], vy[b], vy[c], vy[d]); f.ymin = mn(vy[a], vy[b], vy[c], vy[d]);
            f.zmax = mx(vz[a], vz[b], vz[c], vz[d]); f.zmin = mn(vz[a], vz[b], vz[c], vz[d]);
            faces.push_back(f);
        }
        // faces that can match or update the carried point, per link class
        static const int want[6][2] = {{1, 0}, {3, 0}, {2, 0}, {2, 3}, {1, 3}, {1, 2}};
        for (int c = 0; c < 6; ++c)
            for (size_t i = 0; i < faces.size(); ++i)
                if (faces[i].plane == want[c][0] || faces[i].plane == want[c][1]) cls_faces[c].push_back((int)i);
        return true;
    }

    bool mesh(int nx, std::string& err) {
        if (nx < 2) { err = "need at least two nodes in x"; return false; }
        double xmax = vx[0], ymax = vy[0], zmax = vz[0];
        xmin = vx[0]; ymin = vy[0]; zmin = vz[0];
        for (size_t i = 0; i < vx.size(); ++i) {
            xmax = vx[i] > xmax ? vx[i] : xmax; xmin = vx[i] < xmin ? vx[i] : xmin;
            ymax = vy[i] > ymax ? vy[i] : ymax; ymin = vy[i] < ymin ? vy[i] : ymin;
            zmax = vz[i] > zmax ? vz[i] : zmax; zmin = vz[i] < zmin ? vz[i] : zmin;
        }
        n = nx;
        delta = (xmax - xmin) / (n - 1);
        if (!(delta > 0)) { err = "zero x extent"; return false; }
        const double my = std::ceil((ymax - ymin) / delta) + 1, mz = std::ceil((zmax - zmin) / delta) + 1;
        if ((double)(n + 2) * (my + 2) * (mz + 2) > 2.0e9) { err = "lattice too large"; return false; }
        m = (int)my;
        h = (int)mz;
        const size_t total = (size_t)P() * Q() * (h + 2);
        flag.assign(total, 0);
        reached.assign(total, 0);
        return true;
    }

    // Does the link from mesh node a to mesh node b cross a face?
    bool inter(int a, int b) {
        const int PQ = P() * Q();
        const double x1 = cx(a % P()), y1 = cy(a / P() % Q()), z1 = cz(a / PQ);
        const double x2 = cx(b % P()), y2 = cy(b / P() % Q()), z2 = cz(b / PQ);
        const double maxX = std::max(x1, x2), minX = std::min(x1, x2);
        const double maxY = std::max(y1, y2), minY = std::min(y1, y2);
        const double maxZ = std::max(z1, z2), minZ = std::min(z1, z2);
        const bool sx = x1 == x2, sy = y1 == y2, sz = z1 == z2;
        int c;
        if (sx && sy) c = 0;
        else if (sy && sz) c = 1;
        else if (sx && sz) c = 2;
        else if (sz) c = 3;
        else if (sy) c = 4;
        else c = 5;
        auto in = [](double v, double lo, double hi) { return v <= hi && v >= lo; };
        for (int fi : cls_faces[c]) {
            const Face& f = faces[fi];
            bool ok = false;
            switch (c) {
            case 0:   // z-links: faces z = const
                I[0] = x1; I[1] = y1;
                if (f.plane != 1) continue;
                I[2] = f.p2[2];
                ok = in(I[2], minZ, maxZ) && in(I[0], f.xmin, f.xmax) && in(I[1], f.ymin, f.ymax);
                break;
            case 1:   // x-links: faces x = const
                I[2] = z1; I[1] = y1;
                if (f.plane != 3) continue;
                I[0] = f.p2[0];
                ok = in(I[0], minX, maxX) && in(I[1], f.ymin, f.ymax) && in(I[2], f.zmin, f.zmax);
                break;
            case 2:   // y-links: faces y = const
                I[2] = z1; I[0] = x1;
                if (f.plane != 2) continue;
                I[1] = f.p2[1];
                ok = in(I[1], minY, maxY) && in(I[0], f.xmin, f.xmax) && in(I[2], f.zmin, f.zmax);
                break;
            case 3:   // diagonals in the xy plane
                I[2] = z1;
                if (f.plane == 2) {
                    const double yf = f.p2[1];
                    if (x2 >= x1 && x2 < f.xmax) I[0] = x1 + std::fabs(yf - y1);
                    else if (x2 < x1 && x2 >= f.xmin) I[0] = x1 - std::fabs(yf - y1);
                    if (y1 <= f.ymax && x1 >= f.xmin && y1 == f.ymax) I[0] = x1;
                    I[1] = yf;
                } else {
                    const double xf = f.p2[0];
                    if (y2 > y1 && y2 <= f.ymax) I[1] = y1 + std::fabs(xf - x1);
                    else if (y2 < y1 && y2 >= f.ymin) I[1] = y1 - std::fabs(xf - x1);
                    if (y1 <= f.ymax && y1 >= f.ymin && x1 == f.xmax) I[1] = y1;
                    I[0] = xf;
                }
                ok = in(I[0], minX, maxX) && in(I[1], minY, maxY) && in(I[2], f.zmin, f.zmax);
                break;
            case 4:   // diagonals in the xz plane
                I[1] = y1;
                if (f.plane == 1) {
                    const double zf = f.p2[2];
                    if (x2 > x1 && x2 <= f.xmax) I[0] = x1 + std::fabs(zf - z1);
                    else if (x2 < x1 && x2 >= f.xmin) I[0] = x1 - std::fabs(zf - z1);
                    if (x1 <= f.xmax && x1 >= f.xmin && z1 == f.zmax) I[0] = x1;
                    I[2] = zf;
                } else {
                    const double xf = f.p2[0];
                    if (z2 > z1 && z2 <= f.zmax) I[2] = z1 + std::fabs(xf - x1);
                    else if (z2 < z1 && z2 >= f.zmin) I[2] = z1 - std::fabs(xf - x1);
                    if (z1 <= f.zmax && z1 >= f.zmin && x1 == f.xmax) I[2] = z1;
                    I[0] = xf;
                }
                ok = in(I[0], minX, maxX) && in(I[2], minZ, maxZ) && in(I[1], f.ymin, f.ymax);
                break;
            default:  // diagonals in the yz plane
                I[0] = x1;
                if (f.plane == 1) {
                    const double zf = f.p2[2];
                    if (y2 > y1 && y2 <= f.ymax) I[1] = y1 + std::fabs(zf - z1);
                    else if (y2 < y1 && y2 >= f.ymin) I[1] = y1 - std::fabs(zf - z1);
                    if (y1 <= f.ymax && y1 >= f.ymin && z1 == f.zmax) I[1] = y1;
                    I[2] = zf;
                } else {
                    const double yf = f.p2[1];
                    if (z2 > z1 && z2 <= f.zmax) I[2] = z1 + std::fabs(yf - y1);
                    else if (z2 < z1 && z2 >= f.zmin) I[2] = z1 - std::fabs(yf - y1);
                    if (z1 <= f.zmax && z1 >= f.zmin && y1 == f.ymax) I[2] = z1;
                    I[1] = yf;
                }
                ok = in(I[2], minZ, maxZ) && in(I[1], minY, maxY) && in(I[0], f.xmin, f.xmax);
                break;
            }
            // a crossing exactly at the far node does not count
            ok = ok && (I[0] != x2 || sx) && (I[1] != y2 || sy) && (I[2] != z2 || sz);
            if (ok) {
                hit_id = f.id;
                hit_type = f.type;
                hit[0] = I[0]; hit[1] = I[1]; hit[2] = I[2];
                return true;
            }
        }
        return false;
    }

    static void put_link(std::string& out, int x, int y, int z, int dir, int type, double ix, double iy, double iz,
                         int id) {
        char line[256];
        // the reference passes the point as float
        const int len = std::snprintf(line, sizeof line, "%d\t%d\t%d\t%d\t%d\t%.6f\t%.6f\t%.6f\t%d\n", x - 1, y - 1,
                                      z - 1, dir, type, (double)(float)ix, (double)(float)iy, (double)(float)iz, id);
        out.append(line, (size_t)len);
    }

    bool fill(const double seed[3], int complexity, std::string& bc, int64_t& nlinks, std::string& err) {
        static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
        static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
        static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
        const int p = P(), q = Q(), pq = p * q;
        const int total = pq * (h + 2);
        // nearest mesh node: strictly closer than the seed's distance from the origin
        double best = std::sqrt(std::pow(seed[0], 2) + std::pow(seed[1], 2) + std::pow(seed[2], 2));
        int sid = -1;
        for (int id = 0; id < total; ++id) {
            const double d = std::sqrt(std::pow(cx(id % p) - seed[0], 2) + std::pow(cy(id / p % q) - seed[1], 2) +
                                       std::pow(cz(id / pq) - seed[2], 2));
            if (d < best) { best = d; sid = id; }
        }
        if (sid < 0) { err = "no mesh node is closer to the seed than the origin is"; return false; }
        order.clear();
        order.push_back(sid);
        flag[sid] = 1;
        nlinks = 0;
        auto interior = [&](int X, int Y, int Z) { return X > 0 && X < n + 1 && Y > 0 && Y < m + 1 && Z > 0 && Z < h + 1; };
        for (size_t k = 0; k < order.size(); ++k) {
            const int id = order[k];
            const int Z = id / pq, Y = (id - Z * pq) / p, X = id - Z * pq - Y * p;
            if (!interior(X, Y, Z)) continue;
            for (int dir = 1; dir < 19; ++dir) {
                const int nb = id + CX[dir] + CY[dir] * p + CZ[dir] * pq;
                if (!inter(id, nb)) {
                    if (!reached[nb]) {
                        order.push_back(nb);
                        flag[nb] = 1;
                        reached[nb] = 1;
                    }
                } else if (complexity == 1 && !reached[nb]) {
                    put_link(bc, X, Y, Z, dir, hit_type, hit[0], hit[1], hit[2], hit_id);
                    ++nlinks;
                }
            }
        }
        if (complexity == 2) {
            for (size_t k = 0; k < order.size(); ++k) {
                const int id = order[k];
                const int Z = id / pq, Y = (id - Z * pq) / p, X = id - Z * pq - Y * p;
                if (!interior(X, Y, Z)) continue;
                for (int dir = 1; dir < 19; ++dir) {
                    const int nb = id + CX[dir] + CY[dir] * p + CZ[dir] * pq;
                    if (inter(id, nb) && !reached[nb]) {
                        put_link(bc, X, Y, Z, dir, hit_type, hit[0], hit[1], hit[2], hit_id);
                        ++nlinks;
                    }
                }
            }
        }
        return true;
    }

    void node_text(std::string& out) const {
        const int p = P(), q = Q();
        out.reserve((size_t)n * m * h * 40);
        char line[256];
        for (int k = 1; k < h + 1; ++k)
            for (int j = 1; j < m + 1; ++j)
                for (int i = 1; i < n + 1; ++i) {
                    const size_t id = (size_t)i + (size_t)j * p + (size_t)k * p * q;
                    const int len = std::snprintf(line, sizeof line, "%d\t%d\t%d\t%.6f\t%.6f\t%.6f\t%d\n", i - 1, j - 1,
                                                  k - 1, cx(i), cy(j), cz(k), (int)flag[id]);
                    out.append(line, (size_t)len);
                }
    }

    // the reference's mesh listing (xtecplot.dat): all nodes with padding
    void mesh_text(std::string& out) const {
        char line[256];
        int len = std::snprintf(line, sizeof line,
                                "title = \"sample mesh\"\nvariables = \"x\", \"y\", \"z\" \nzone i=%d, j=%d, k=%d\n",
                                n + 2, m + 2, h + 2);
        out.append(line, (size_t)len);
        const int p = P(), q = Q();
        for (int id = 0; id < p * q * (h + 2); ++id) {
            len = std::snprintf(line, sizeof line, "%g    %g    %g\n", cx(id % p), cy(id / p % q), cz(id / (p * q)));
            out.append(line, (size_t)len);
        }
    }
};

bool spit(const char* path, const std::string& s) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(s.data(), 1, s.size(), f) == s.size();
    return std::fclose(f) == 0 && ok;
}

thread_local std::string g_lg_err;

}  // namespace

extern "C" lbm_status lbm_latticegen(const char* vrt_path, const char* bnd_path, int32_t nx, const double seed[3],
                                     int32_t complexity, const char* node_out, const char* bc_out,
                                     const char* mesh_out, int32_t dims[3], int64_t* nlinks) {
    g_lg_err.clear();
    if (!vrt_path || !bnd_path || !seed || !node_out || !bc_out || (complexity != 1 && complexity != 2)) {
        g_lg_err = "bad arguments";
        return LBM_EINVAL;
    }
    Generator g;
    std::string err, bc, txt;
    if (!g.read(vrt_path, bnd_path, err)) { g_lg_err = err; return LBM_EIO; }
    if (!g.mesh(nx, err)) { g_lg_err = err; return LBM_EINVAL; }
    int64_t nl = 0;
    if (!g.fill(seed, complexity, bc, nl, err)) { g_lg_err = err; return LBM_EINVAL; }
    g.node_text(txt);
    if (!spit(node_out, txt) || !spit(bc_out, bc)) { g_lg_err = "cannot write the lattice files"; return LBM_EIO; }
    if (mesh_out) {
        txt.clear();
        g.mesh_text(txt);
        if (!spit(mesh_out, txt)) { g_lg_err = "cannot write the mesh listing"; return LBM_EIO; }
    }
    if (dims) { dims[0] = g.n; dims[1] = g.m; dims[2] = g.h; }
    if (nlinks) *nlinks = nl;
    return LBM_OK;
}

extern "C" const char* lbm_latticegen_last_error(void) { return g_lg_err.c_str(); }
