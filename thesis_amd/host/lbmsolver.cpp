// lbmsolver: the process-level drop-in for the reference solver.
//
// Replaces main / Iterate2D / Iterate3D of `lbm-solver_final`
// (S/main.cu:23-145, S/Iterate.cu:23-728, S/Iterate3D.cu:30-943): reads the
// 23-line SetUpData.ini (S/FilesReading.cu:60-222) or the CLI flags
// (S/Arguments.cu:155-315), the latticegen2 Mesh/<name>_Node.dat and _BC.dat,
// runs the time loop through the C ABI (include/lbm.h) and writes
// Results/FinalData.{csv,dat,vti} (S/FilesWriting.cu:146-258) and
// Results/residuals.dat (S/LogWriter.cu:190-205).  Exit codes follow the
// reference driver: 0 ok, 1 diverged, 2 file error (S/include/Iterate.h:62).
//
// Lines after the 23 positional ones are ignored by the reference; here an
// optional "[mi355x]" section there sets the new knobs:
//   storage = f64|f32|f16|bf16    init = zero|equilibrium
//   walls = onnode|halfway        trt_magic = <Lambda>
//   force_model = shift|guo       rho_out = <density>
#include "lbm.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace {

struct Args {
    std::string lattice = "", node_file, bc_file, out_dir = "Results/";
    double u = 0, v = 0, w = 0, rho = 1.0, viscosity = 1.0 / 6.0, g = 0;
    int inlet_profile = 2;           // 1 INLET_PROF, 2 INLET_CTE, 3 INLET_PUL
    int collision = LBM_BGK;
    int curved = 0, hh_walls = 0, outlet = 1;
    long iterations = 1000, residual_every = 100;
    double stop[4] = {0, 0, 0, 0};
    int format = 3;                  // 1 CSV, 2 DAT (Tecplot), 3 VTI
    // new knobs
    int storage = LBM_STORE_F64, init_eq = 0, wall_model = LBM_WALL_ONNODE, force_model = LBM_FORCE_SHIFT;
    double trt_magic = 0, rho_out = -1;
};

std::string trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

std::string unquote(const std::string& s) {
    size_t a = s.find('"');
    if (a == std::string::npos) return trim(s);
    size_t b = s.find('"', a + 1);
    return s.substr(a + 1, b == std::string::npos ? std::string::npos : b - a - 1);
}

bool has(const std::string& s, const char* k) { return s.find(k) != std::string::npos; }

// 23 positional lines, value after the first "= " (S/FilesReading.cu:64-72).
bool read_ini(const std::string& path, Args& a) {
    std::ifstream in(path);
    if (!in) return false;
    std::vector<std::string> v;
    std::string line;
    for (int i = 0; i < 23 && std::getline(in, line); ++i) {
        size_t p = line.find("= ");
        v.push_back(p == std::string::npos ? std::string() : line.substr(p + 2));
    }
    if (v.size() < 23) return false;
    a.lattice = unquote(v[0]);
    a.u = atof(v[1].c_str()); a.v = atof(v[2].c_str()); a.w = atof(v[3].c_str());
    a.rho = atof(v[4].c_str());
    a.viscosity = atof(v[5].c_str());
    a.inlet_profile = has(v[6], "INLET_PROF") ? 1 : has(v[6], "INLET_PUL") ? 3 : 2;
    a.collision = has(v[7], "TRT") ? LBM_TRT : has(v[7], "MRT") ? LBM_MRT : LBM_BGK;
    a.curved = has(v[8], "CURVED");
    a.hh_walls = has(v[9], "HHmodel");
    a.outlet = has(v[10], "Vout") ? 1 : has(v[10], "1stE") ? 2 : has(v[10], "2ndE") ? 3 : 4;
    a.iterations = atol(v[11].c_str());
    sscanf(v[14].c_str(), "%lf %lf %lf %lf", &a.stop[0], &a.stop[1], &a.stop[2], &a.stop[3]);
    a.format = has(v[15], "CSV") ? 1 : has(v[15], "DAT") ? 2 : 3;
    a.residual_every = atol(v[19].c_str());
    a.g = atof(v[20].c_str());
    // optional section after line 23
    std::map<std::string, std::string> kv;
    bool sect = false;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line == "[mi355x]") { sect = true; continue; }
        size_t p = line.find('=');
        if (sect && p != std::string::npos) kv[trim(line.substr(0, p))] = trim(line.substr(p + 1));
    }
    if (kv.count("storage")) {
        const std::string& s = kv["storage"];
        a.storage = s == "f32" ? LBM_STORE_F32 : s == "f16" ? LBM_STORE_F16 : s == "bf16" ? LBM_STORE_BF16 : LBM_STORE_F64;
    }
    if (kv.count("init")) a.init_eq = kv["init"] == "equilibrium";
    if (kv.count("walls")) a.wall_model = kv["walls"] == "halfway" ? LBM_WALL_HALFWAY : LBM_WALL_ONNODE;
    if (kv.count("force_model")) a.force_model = kv["force_model"] == "guo" ? LBM_FORCE_GUO : LBM_FORCE_SHIFT;
    if (kv.count("trt_magic")) a.trt_magic = atof(kv["trt_magic"].c_str());
    if (kv.count("rho_out")) a.rho_out = atof(kv["rho_out"].c_str());
    return true;
}

void usage() {
    std::printf(
        "usage: lbmsolver [-f SetUpData.ini] [-n node.dat -b bc.dat] [-o outdir] [-u U] [-v V] [-r rho]\n"
        "                 [-s viscosity] [-c BGKW|TRT|MRT] [-i iterations] [--every N]\n"
        "                 [--format paraview|tecplot|csv] [--storage f64|f32|f16|bf16]\n"
        "       lbmsolver compare <file1> <file2>\n");
}

// Column-wise comparison of two result files, the reference's regression
// mode (S/FilesReading.cu:409-467): fails when any column's sum |a - b|
// exceeds 0.001.  Numeric tokens are read line by line; XML tags are skipped.
int compare_files(const char* f1, const char* f2) {
    auto load = [](const char* p, std::vector<std::vector<double>>& rows) {
        std::ifstream in(p);
        if (!in) return false;
        std::string line;
        while (std::getline(in, line)) {
            if (line.find('<') != std::string::npos) continue;
            for (char& ch : line) if (ch == ',') ch = ' ';
            std::vector<double> r;
            const char* s = line.c_str();
            char* e = nullptr;
            for (;;) {
                double x = std::strtod(s, &e);
                if (e == s) break;
                r.push_back(x);
                s = e;
            }
            if (!r.empty()) rows.push_back(r);
        }
        return true;
    };
    std::vector<std::vector<double>> a, b;
    if (!load(f1, a) || !load(f2, b)) { std::fprintf(stderr, "compare: cannot read input\n"); return 2; }
    if (a.size() != b.size()) { std::printf("compare: row counts differ (%zu, %zu)\n", a.size(), b.size()); return 1; }
    std::vector<double> sum, mx;
    for (size_t i = 0; i < a.size(); ++i) {
        size_t nc = std::min(a[i].size(), b[i].size());
        if (sum.size() < nc) { sum.resize(nc, 0); mx.resize(nc, 0); }
        for (size_t c = 0; c < nc; ++c) {
            double d = std::fabs(a[i][c] - b[i][c]);
            sum[c] += d;
            mx[c] = std::max(mx[c], d);
        }
    }
    int bad = 0;
    for (size_t c = 0; c < sum.size(); ++c) {
        std::printf("column %zu: sum|diff| %.6e  max|diff| %.6e\n", c, sum[c], mx[c]);
        bad |= sum[c] > 0.001;
    }
    std::printf(bad ? "compare: FILES DIFFER\n" : "compare: files match\n");
    return bad;
}

void write_results(const std::string& path, int fmt, int n, int m, int h, double dx, const std::vector<double>& rho,
                   const std::vector<double>& u, const std::vector<double>& v, const std::vector<double>& w) {
    FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return;
    const size_t N = (size_t)n * m * h;
    if (fmt == 1 || fmt == 2) {
        if (fmt == 1) std::fprintf(fp, "x,y,z,u,v,w,vel_mag,rho,press\n");
        else std::fprintf(fp, "Title = \"LBM results\"\nVariables = \"x\",\"y\",\"z\",\"u\",\"v\",\"w\",\"vel mag\",\"rho\",\"press\"\n"
                              "Zone i=%d, j=%d, k=%d, f=point\n", n, m, h);
        const char* f = fmt == 1 ? "%.10f, %.10f, %.10f, %.10f, %.10f, %.10f, %.10f, %.10f, %.10f\n"
                                 : "%.10f %.10f %.10f %.10f %.10f %.10f %.10f %.10f %.10f\n";
        for (size_t i = 0; i < N; ++i) {
            const int x = (int)(i % n), y = (int)((i / n) % m), z = (int)(i / ((size_t)n * m));
            std::fprintf(fp, f, x * dx, y * dx, z * dx, u[i], v[i], w[i],
                         std::sqrt(u[i] * u[i] + v[i] * v[i] + w[i] * w[i]), rho[i], rho[i] / 3.0);
        }
    } else {
        std::fprintf(fp, "<?xml version=\"1.0\"?>\n<!-- 3D LBM Solver -->\n"
                         "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                         "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
                         "  <Piece Extent=\"0 %d 0 %d 0 %d\">\n    <PointData Scalars=\"scalars\">\n"
                         "      <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                     n - 1, m - 1, h - 1, n - 1, m - 1, h - 1);
        for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f %.10f %.10f\n", u[i], v[i], w[i]);
        std::fprintf(fp, "      </DataArray>\n"
                         "      <DataArray type=\"Float32\" Name=\"Density\" NumberOfComponents=\"1\" format=\"ascii\">\n");
        for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f\n", rho[i]);
        std::fprintf(fp, "      </DataArray>\n"
                         "      <DataArray type=\"Float32\" Name=\"Pressure\" NumberOfComponents=\"1\" format=\"ascii\">\n");
        for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f\n", rho[i] / 3.0);
        std::fprintf(fp, "      </DataArray>\n    </PointData>\n    <CellData>\n    </CellData>\n  </Piece>\n"
                         "  </ImageData>\n</VTKFile>\n");
    }
    std::fclose(fp);
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    std::string ini = "SetUpData.ini";
    bool from_cli = false;
    if (argc > 1 && std::strcmp(argv[1], "compare") == 0) {
        if (argc < 4) { usage(); return 2; }
        return compare_files(argv[2], argv[3]);
    }
    for (int i = 1; i < argc; ++i) {
        std::string s = argv[i];
        auto next = [&](void) -> const char* { return i + 1 < argc ? argv[++i] : ""; };
        if (s == "-h" || s == "--help") { usage(); return 0; }
        else if (s == "-f") ini = next();
        else if (s == "-n") { a.node_file = next(); from_cli = true; }
        else if (s == "-b") { a.bc_file = next(); from_cli = true; }
        else if (s == "-o") a.out_dir = next();
        else if (s == "-u") { a.u = atof(next()); from_cli = true; }
        else if (s == "-v") { a.v = atof(next()); from_cli = true; }
        else if (s == "-r") { a.rho = atof(next()); from_cli = true; }
        else if (s == "-s") { a.viscosity = atof(next()); from_cli = true; }
        else if (s == "-i") { a.iterations = atol(next()); from_cli = true; }
        else if (s == "--every") a.residual_every = atol(next());
        else if (s == "-c") { std::string c = next(); a.collision = c == "TRT" ? LBM_TRT : c == "MRT" ? LBM_MRT : LBM_BGK; from_cli = true; }
        else if (s == "--format") { std::string f = next(); a.format = f == "tecplot" ? 2 : f == "csv" ? 1 : 3; }
        else if (s == "--storage") { std::string f = next(); a.storage = f == "f32" ? LBM_STORE_F32 : f == "f16" ? LBM_STORE_F16 : f == "bf16" ? LBM_STORE_BF16 : LBM_STORE_F64; }
        else { std::fprintf(stderr, "unknown option %s\n", s.c_str()); usage(); return 2; }
    }
    if (!from_cli || a.node_file.empty()) {
        Args keep = a;
        if (!read_ini(ini, a)) { std::fprintf(stderr, "cannot read %s\n", ini.c_str()); return 2; }
        a.storage = keep.storage != LBM_STORE_F64 ? keep.storage : a.storage;
        a.out_dir = keep.out_dir;
        if (keep.residual_every != 100) a.residual_every = keep.residual_every;
        a.node_file = "Mesh/" + a.lattice + "_Node.dat";   // S/main.cu:97-100
        a.bc_file = "Mesh/" + a.lattice + "_BC.dat";
    }
    if (a.curved) std::fprintf(stderr, "note: CURVED (Bouzidi) walls are not available; straight walls used\n");
    if (a.hh_walls) std::fprintf(stderr, "note: HHmodel wet-node walls are not available; bounce-back used\n");

    int32_t dims[3];
    int64_t nlinks = 0;
    if (lbm_read_lattice(a.node_file.c_str(), a.bc_file.c_str(), dims, &nlinks, nullptr, nullptr, nullptr, nullptr,
                         nullptr, nullptr) != LBM_OK) {
        std::fprintf(stderr, "cannot read lattice %s / %s\n", a.node_file.c_str(), a.bc_file.c_str());
        return 2;
    }
    const size_t N = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<uint8_t> cls(N);
    std::vector<uint32_t> bb(N);
    int64_t tc[6], dc[19];
    double dx = 1;
    int32_t cyc[3];
    if (lbm_read_lattice(a.node_file.c_str(), a.bc_file.c_str(), dims, &nlinks, cls.data(), bb.data(), tc, dc, &dx,
                         cyc) != LBM_OK) {
        std::fprintf(stderr, "cannot read lattice\n");
        return 2;
    }
    const bool three_d = dims[2] > 1;
    // Cyclic faces of latticegen2 lattices lie on the same physical plane
    // (period n - 1, S/GpuBoundaries.cu:2031-2081): drop the last plane of each
    // cyclic axis, wrap with period n - 1 and copy plane 0 back on output.
    int sim[3] = {dims[0], dims[1], dims[2]};
    for (int ax = 0; ax < 3; ++ax)
        if (cyc[ax] && dims[ax] > 1) sim[ax] = dims[ax] - 1;
    const size_t NS = (size_t)sim[0] * sim[1] * sim[2];
    std::vector<uint8_t> scls(NS);
    std::vector<uint32_t> sbb(NS);
    for (int z = 0; z < sim[2]; ++z)
        for (int y = 0; y < sim[1]; ++y)
            for (int x = 0; x < sim[0]; ++x) {
                const size_t s = ((size_t)z * sim[1] + y) * sim[0] + x, d = ((size_t)z * dims[1] + y) * dims[0] + x;
                scls[s] = cls[d];
                sbb[s] = bb[d];
            }
    // inlet profile on velocity-inlet nodes (S/GpuInit.cu:994-1006): parabolic
    // between half a spacing beyond the extreme inlet nodes, or constant
    std::vector<double> ubc(NS, 0.0), vbc(NS, 0.0);
    {
        int ymin = 1 << 30, ymax = -1;
        for (size_t i = 0; i < NS; ++i)
            if (scls[i] == LBM_VEL_W) {
                const int y = (int)((i / sim[0]) % sim[1]);
                ymin = std::min(ymin, y);
                ymax = std::max(ymax, y);
            }
        for (size_t i = 0; i < NS; ++i)
            if (scls[i] == LBM_VEL_W) {
                const double y = (double)((i / sim[0]) % sim[1]), lo = ymin - 0.5, hi = ymax + 0.5;
                ubc[i] = a.inlet_profile == 1 ? 6.0 * a.u * (y - lo) * (hi - y) / ((hi - lo) * (hi - lo)) : a.u;
                vbc[i] = a.v;
            }
    }
    lbm_params p;
    std::memset(&p, 0, sizeof p);
    p.abi_version = LBM_ABI_VERSION;
    p.q = three_d ? 19 : 9;
    p.nx = sim[0]; p.ny = sim[1]; p.nz = three_d ? sim[2] : 1;
    for (int ax = 0; ax < 3; ++ax) p.periodic[ax] = cyc[ax];
    p.collision = a.collision;
    p.omega = 1.0 / (3.0 * a.viscosity + 0.5);
    p.trt_magic = a.trt_magic;
    p.force_model = a.g != 0 ? a.force_model : LBM_FORCE_NONE;
    p.force[0] = a.g;
    p.storage = a.storage;
    p.wall_model = a.wall_model;
    p.inlet_model = LBM_INLET_ZOUHE;
    p.rho_out = a.rho_out > 0 ? a.rho_out : a.rho;
    lbm_lattice lat = {scls.data(), sbb.data(), ubc.data(), vbc.data(), nullptr};
    lbm_ctx* ctx = nullptr;
    if (lbm_create(&p, &lat, &ctx) != LBM_OK) { std::fprintf(stderr, "lbm_create failed\n"); return 2; }
    std::vector<double> rho(NS, a.rho), u(NS, 0.0), v(NS, 0.0), w(NS, 0.0);
    for (size_t i = 0; i < NS; ++i) { u[i] = scls[i] == LBM_VEL_W ? ubc[i] : a.u; v[i] = a.v; w[i] = a.w; }
    if (lbm_init(ctx, a.init_eq ? 0 : 1, rho.data(), u.data(), v.data(), w.data()) != LBM_OK) {
        std::fprintf(stderr, "lbm_init: %s\n", lbm_last_error(ctx));
        lbm_destroy(ctx);
        return 2;
    }
    mkdir(a.out_dir.c_str(), 0755);
    FILE* rf = std::fopen((a.out_dir + "residuals.dat").c_str(), "w");
    if (rf) std::fprintf(rf, "Iter L2_norm L2_norm_weighted Drag Lift\n");
    double r4[4];
    lbm_reduce(ctx, r4);
    const long every = a.residual_every > 0 ? a.residual_every : a.iterations;
    long done = 0;
    int code = 0;
    while (done < a.iterations) {
        const long k = std::min(every, a.iterations - done);
        lbm_status st = lbm_step(ctx, k);
        if (st != LBM_OK) { std::fprintf(stderr, "step: %s\n", lbm_last_error(ctx)); code = 2; break; }
        done += k;
        lbm_reduce(ctx, r4);
        const double res = r4[1] > 0 ? std::sqrt(r4[0] / r4[1]) : std::sqrt(r4[0]);
        if (rf) std::fprintf(rf, "%ld %5.4e %5.4e %f %f\n", done, std::sqrt(r4[0]), res, 0.0, 0.0);
        std::printf("iteration %ld residual %.4e mass %.10f\n", done, res, r4[2]);
        if (r4[3] > 0 || res != res) { std::fprintf(stderr, "\nDIVERGENCE!\n"); code = 1; break; }
        if (a.stop[0] > 0 && res < a.stop[0]) { std::printf("simulation converged!\n"); break; }
    }
    if (rf) std::fclose(rf);
    lbm_get_macros(ctx, rho.data(), u.data(), v.data(), w.data());
    lbm_destroy(ctx);
    // back to the file's node set (cyclic planes copied from plane 0)
    std::vector<double> R(N), U(N), V(N), W(N);
    for (int z = 0; z < dims[2]; ++z)
        for (int y = 0; y < dims[1]; ++y)
            for (int x = 0; x < dims[0]; ++x) {
                const size_t d = ((size_t)z * dims[1] + y) * dims[0] + x;
                const size_t s = ((size_t)(z % sim[2]) * sim[1] + (y % sim[1])) * sim[0] + (x % sim[0]);
                R[d] = rho[s]; U[d] = u[s]; V[d] = v[s]; W[d] = w[s];
            }
    const char* ext = a.format == 1 ? "csv" : a.format == 2 ? "dat" : "vti";
    write_results(a.out_dir + "FinalData." + ext, a.format, dims[0], dims[1], dims[2], dx, R, U, V, W);
    std::printf("results written to %sFinalData.%s\n", a.out_dir.c_str(), ext);
    return code;
}
