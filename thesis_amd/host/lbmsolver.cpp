// lbmsolver: the process-level drop-in for the reference solver.
//
// Replaces main / Iterate2D / Iterate3D of `lbm-solver_final`
// (S/main.cu:23-145, S/Iterate.cu:23-728, S/Iterate3D.cu:30-943): reads the
// 23-line SetUpData.ini (S/FilesReading.cu:60-222) or the CLI flags
// (S/Arguments.cu:155-315), the latticegen2 Mesh/<name>_Node.dat and _BC.dat,
// runs the time loop through the C ABI (include/lbm.h) and writes
// Results/FinalData.{csv,dat,vti} (S/FilesWriting.cu:146-258) and
// Results/residuals.dat (S/LogWriter.cu:190-205), lbmsolver.log
// (:5-174) and runtimes.dat (:176-188).  Exit codes follow the
// reference driver: 0 ok, 1 diverged, 2 file error (S/include/Iterate.h:62).
//
// Lines after the 23 positional ones are ignored by the reference; here an
// optional "[mi355x]" section there sets the new knobs:
//   storage = f64|f32|f16|bf16    init = zero|equilibrium
//   walls = onnode|halfway        trt_magic = <Lambda>
//   force_model = shift|guo       rho_out = <density>
//   inlet = zouhe|reference       outlet = pressure|velocity|extrap2|extrap1|feq
//   vti = appended                (binary-appended VTI output, also read back on restart)
//   multiphase = yes              (the reference's default colour-gradient steady bubble; its
//                                  parameters by cg_beta, cg_A, cg_b_alpha, cg_gamma, cg_r_density,
//                                  cg_r_viscosity, cg_b_viscosity, cg_g_limit, cg_high_order,
//                                  cg_enhanced, cg_bubble_radius, cg_control_param)
//   cg_test_case = 1..6           (2D: the reference's test_case, S/GpuInit.cu:323-534: 1 steady
//                                  bubble, 2 two layers between Zou-He walls, the north one at U
//                                  (Couette; Poiseuille with Force != 0), 3 square, 4 two bubbles,
//                                  5 ellipse, 6 Rayleigh-Taylor between bounce-back walls under
//                                  Force; cg_external_force = 0 gravity along y, 1 along x;
//                                  3D, S/GpuInit.cu:536-720: 1 sphere, 2 cube, 3 two spheres,
//                                  4 two layers between inlet faces on y (Couette; the lattice
//                                  cyclic along x and z), 5 Rayleigh-Taylor between bounce-back
//                                  y faces)
// BCoutlet maps as the reference reads it (S/FilesReading.cu:180-190 with
// S/include/Arguments.h:22-28): in 2D Vout -> velocity outlet, 1stE ->
// second-order extrapolation, 2ndE -> first-order extrapolation, HEmodel ->
// the Zou-He pressure outlet (the reference's 2D solver applies no outlet for
// it); in 3D Vout and 1stE -> second-order, 2ndE -> first-order, HEmodel ->
// the f_eq outlet.  BCwallModel = HHmodel gives the 3D wet-node walls.
#include "lbm.h"

#include <hip/hip_runtime_api.h>

#include <chrono>
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
    int curved = 0, hh_walls = 0;
    int outlet = 2;                  // S/include/Arguments.h:22-28 numbering; CLI default "second" (S/main.cu:34)
    long iterations = 1000, residual_every = 100, autosave_after = 0, autosave_every = 0;
    std::string residual_model = "L2";
    int boundary_id = 0, update_inlet_outlet = 1;
    int inlet_model = LBM_INLET_ZOUHE;
    int walls_set = 0;               // [mi355x] walls given: overrides BCwallModel
    double stop[4] = {0, 0, 0, 0};
    int format = 3;                  // 1 CSV, 2 DAT (Tecplot), 3 VTI, 4 VTI with binary-appended data
    // new knobs
    bool use_init_file = false;      // ini line 17: start from the macroscopic fields of a results file
    std::string init_file;           // ini line 18
    int storage = LBM_STORE_F64, init_eq = 0, wall_model = LBM_WALL_ONNODE, force_model = LBM_FORCE_SHIFT;
    double trt_magic = 0, rho_out = -1;
    // colour-gradient two-phase run ([mi355x] multiphase = yes): the steady
    // bubble with the reference's hard-coded parameters (S/main.cu:52-72)
    bool multiphase = false;
    double cg_r_density = 1.0, cg_gamma = 1.0, cg_b_alpha = 4.0 / 9.0, cg_r_nu = 0.5, cg_b_nu = 0.5;
    double cg_beta = 0.99, cg_A = 1e-4, cg_kappa = 0.9, cg_g_limit = 0.0, cg_radius = 18.0;
    int cg_high_order = 1, cg_enhanced = 0, cg_test_case = 1, cg_external_force = 0;
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
    a.use_init_file = has(v[16], "yes");
    a.init_file = unquote(v[17]);
    a.residual_every = atol(v[19].c_str());
    a.autosave_after = atol(v[12].c_str());
    a.autosave_every = atol(v[13].c_str());
    a.residual_model = trim(v[18]);
    a.boundary_id = atoi(v[21].c_str());
    // neither "yes" nor "no" keeps the default, yes (FilesReading.cu:211-219)
    a.update_inlet_outlet = has(v[22], "yes") || !has(v[22], "no");
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
    if (kv.count("walls")) {
        a.wall_model = kv["walls"] == "halfway" ? LBM_WALL_HALFWAY : LBM_WALL_ONNODE;
        a.walls_set = 1;
    }
    if (kv.count("force_model")) a.force_model = kv["force_model"] == "guo" ? LBM_FORCE_GUO : LBM_FORCE_SHIFT;
    if (kv.count("trt_magic")) a.trt_magic = atof(kv["trt_magic"].c_str());
    if (kv.count("rho_out")) a.rho_out = atof(kv["rho_out"].c_str());
    if (kv.count("vti") && kv["vti"] == "appended" && a.format == 3) a.format = 4;
    if (kv.count("inlet")) a.inlet_model = kv["inlet"] == "reference" ? LBM_INLET_REF2D : LBM_INLET_ZOUHE;
    if (kv.count("outlet")) {
        const std::string& o = kv["outlet"];
        a.outlet = o == "velocity" ? 1 : o == "extrap2" ? 2 : o == "extrap1" ? 3 : o == "feq" ? 4 : 5;
    }
    if (kv.count("multiphase")) a.multiphase = kv["multiphase"] == "yes";
    auto num = [&](const char* k, double& dst) { if (kv.count(k)) dst = atof(kv[k].c_str()); };
    num("cg_r_density", a.cg_r_density); num("cg_gamma", a.cg_gamma); num("cg_b_alpha", a.cg_b_alpha);
    num("cg_r_viscosity", a.cg_r_nu); num("cg_b_viscosity", a.cg_b_nu); num("cg_beta", a.cg_beta);
    num("cg_A", a.cg_A); num("cg_control_param", a.cg_kappa); num("cg_g_limit", a.cg_g_limit);
    num("cg_bubble_radius", a.cg_radius);
    if (kv.count("cg_high_order")) a.cg_high_order = atoi(kv["cg_high_order"].c_str());
    if (kv.count("cg_enhanced")) a.cg_enhanced = atoi(kv["cg_enhanced"].c_str());
    if (kv.count("cg_test_case")) a.cg_test_case = atoi(kv["cg_test_case"].c_str());
    if (kv.count("cg_external_force")) a.cg_external_force = atoi(kv["cg_external_force"].c_str());
    return true;
}

void usage() {
    std::printf(
        "usage: lbmsolver [-f SetUpData.ini] [-n node.dat -b bc.dat] [-o outdir] [-u U] [-v V] [-r rho]\n"
        "                 [-s viscosity] [-c BGKW|TRT|MRT] [-i iterations] [--every N]\n"
        "                 [--format paraview|tecplot|csv|vti-appended] [--storage f64|f32|f16|bf16]\n"
        "                 [-l|--outlet yes|second|first]\n"
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

// Initial macroscopic fields from a results file, the reference's
// readInitConditionsFile/readResultFile (S/FilesReading.cu:224-268,
// :348-407): CSV ("x," header, one line per node), Tecplot ("Ti" header and
// two more lines) or ASCII VTI ("<?" header, 3N + 18 lines: velocity
// vectors, then the density).  The node count must match; columns are
// x, y, z, u, v, w, |u|, rho, p.  Returns false (uniform start) otherwise.
// Binary-appended VTI (written with [mi355x] vti = appended): after the "_"
// that opens <AppendedData encoding="raw">, each array is a UInt32 byte count
// and its Float32 values: Velocity (3 N), Density (N), Pressure (N).
bool read_appended_vti(const std::string& path, size_t N, std::vector<double>& u, std::vector<double>& v,
                       std::vector<double>& w, std::vector<double>& rho) {
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::string all((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    const std::string tag = "<AppendedData encoding=\"raw\">";
    size_t p = all.find(tag);
    if (p == std::string::npos) return false;
    p = all.find('_', p + tag.size());
    if (p == std::string::npos) return false;
    ++p;
    auto block = [&](size_t count, std::vector<float>& out) {
        uint32_t nb = 0;
        if (p + 4 > all.size()) return false;
        std::memcpy(&nb, all.data() + p, 4);
        p += 4;
        if (nb != count * 4 || p + nb > all.size()) return false;
        out.resize(count);
        std::memcpy(out.data(), all.data() + p, nb);
        p += nb;
        return true;
    };
    std::vector<float> vel, den;
    if (!block(3 * N, vel) || !block(N, den)) return false;
    u.resize(N); v.resize(N); w.resize(N); rho.resize(N);
    for (size_t i = 0; i < N; ++i) {
        u[i] = vel[3 * i]; v[i] = vel[3 * i + 1]; w[i] = vel[3 * i + 2]; rho[i] = den[i];
    }
    return true;
}

bool read_initial_conditions(const std::string& path, size_t N, std::vector<double>& u, std::vector<double>& v,
                             std::vector<double>& w, std::vector<double>& rho) {
    if (read_appended_vti(path, N, u, v, w, rho)) return true;
    std::ifstream in(path);
    if (!in) return false;
    std::vector<std::string> lines;
    std::string line;
    while (std::getline(in, line)) lines.push_back(line);
    if (lines.empty()) return false;
    const std::string& h = lines[0];
    u.assign(N, 0.0); v.assign(N, 0.0); w.assign(N, 0.0); rho.assign(N, 1.0);
    auto row9 = [&](size_t first) {
        for (size_t i = 0; i < N; ++i) {
            std::string t = lines[first + i];
            for (char& ch : t) if (ch == ',') ch = ' ';
            double c[9];
            if (std::sscanf(t.c_str(), "%lf %lf %lf %lf %lf %lf %lf %lf %lf", c, c + 1, c + 2, c + 3, c + 4, c + 5,
                            c + 6, c + 7, c + 8) != 9)
                return false;
            u[i] = c[3]; v[i] = c[4]; w[i] = c[5]; rho[i] = c[7];
        }
        return true;
    };
    if (h.rfind("x,", 0) == 0 && lines.size() - 1 == N) return row9(1);
    if (h.rfind("Ti", 0) == 0 && lines.size() - 3 == N) return row9(3);
    if (h.rfind("<?", 0) == 0 && lines.size() == 18 + 3 * N) {
        for (size_t i = 0; i < N; ++i)
            if (std::sscanf(lines[7 + i].c_str(), "%lf %lf %lf", &u[i], &v[i], &w[i]) != 3) return false;
        for (size_t i = 0; i < N; ++i)
            if (std::sscanf(lines[9 + N + i].c_str(), "%lf", &rho[i]) != 1) return false;
        return true;
    }
    return false;
}

const char* ext_of(int fmt) { return fmt == 1 ? "csv" : fmt == 2 ? "dat" : "vti"; }

void write_results(const std::string& path, int fmt, int n, int m, int h, double dx, const std::vector<double>& rho,
                   const std::vector<double>& u, const std::vector<double>& v, const std::vector<double>& w,
                   const std::vector<double>* rr = nullptr, const std::vector<double>* rb = nullptr) {
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
    } else if (fmt == 4) {
        // binary-appended VTI: the same arrays as the ASCII file, Float32 raw
        std::fprintf(fp, "<?xml version=\"1.0\"?>\n"
                         "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt32\">\n"
                         "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
                         "  <Piece Extent=\"0 %d 0 %d 0 %d\">\n    <PointData Scalars=\"scalars\">\n"
                         "      <DataArray type=\"Float32\" Name=\"Velocity\" NumberOfComponents=\"3\" format=\"appended\" offset=\"0\"/>\n"
                         "      <DataArray type=\"Float32\" Name=\"Density\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%zu\"/>\n"
                         "      <DataArray type=\"Float32\" Name=\"Pressure\" NumberOfComponents=\"1\" format=\"appended\" offset=\"%zu\"/>\n"
                         "    </PointData>\n    <CellData>\n    </CellData>\n  </Piece>\n  </ImageData>\n"
                         "  <AppendedData encoding=\"raw\">\n_",
                     n - 1, m - 1, h - 1, n - 1, m - 1, h - 1, 4 + 12 * N, 8 + 16 * N);
        std::vector<float> buf(3 * N);
        auto put = [&](size_t count) {
            const uint32_t nb = (uint32_t)(count * 4);
            std::fwrite(&nb, 4, 1, fp);
            std::fwrite(buf.data(), 4, count, fp);
        };
        for (size_t i = 0; i < N; ++i) { buf[3 * i] = (float)u[i]; buf[3 * i + 1] = (float)v[i]; buf[3 * i + 2] = (float)w[i]; }
        put(3 * N);
        for (size_t i = 0; i < N; ++i) buf[i] = (float)rho[i];
        put(N);
        for (size_t i = 0; i < N; ++i) buf[i] = (float)(rho[i] / 3.0);
        put(N);
        std::fprintf(fp, "\n  </AppendedData>\n</VTKFile>\n");
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
        if (rr && rb) {   // two-phase output (S/FilesWriting.cu:91-112)
            std::fprintf(fp, "      </DataArray>\n"
                             "      <DataArray type=\"Float32\" Name=\"Red Density\" NumberOfComponents=\"1\" format=\"ascii\">\n");
            for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f\n", (*rr)[i]);
            std::fprintf(fp, "      </DataArray>\n"
                             "      <DataArray type=\"Float32\" Name=\"Blue Density\" NumberOfComponents=\"1\" format=\"ascii\">\n");
            for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f\n", (*rb)[i]);
        }
        std::fprintf(fp, "      </DataArray>\n"
                         "      <DataArray type=\"Float32\" Name=\"Pressure\" NumberOfComponents=\"1\" format=\"ascii\">\n");
        for (size_t i = 0; i < N; ++i) std::fprintf(fp, "%.10f\n", rho[i] / 3.0);
        std::fprintf(fp, "      </DataArray>\n    </PointData>\n    <CellData>\n    </CellData>\n  </Piece>\n"
                         "  </ImageData>\n</VTKFile>\n");
    }
    std::fclose(fp);
}

// The 3D runs' line and plane files (S/FilesWriting.cu:260-337, written at
// the end of every 3D run, S/Iterate3D.cu:911-912): midLines.txt holds u/U
// on the vertical line x = n/2, y = m/2 against z and w/U on the line
// y = m/2, z = h/2 against x (one row per z; the reference's arrays have n
// rows, so rows beyond n carry zeros there); channelPlane.txt holds the plane
// x = floor(3n/4).  Coordinates are i * delta as latticegen2 writes them.
void write_3d_lines(const std::string& dir, int n, int m, int h, double dx, const std::vector<double>& U,
                    const std::vector<double>& V, const std::vector<double>& W, double ulid) {
    const size_t nm = (size_t)n * m;
    if (FILE* f = std::fopen((dir + "midLines.txt").c_str(), "w")) {
        std::fprintf(f, "CASE: Ulid = %f Nx = %d\n", ulid, n);
        std::fprintf(f, "U/Ulid\tZ\tX\tW/Ulid\n");
        for (int k = 0; k < h; ++k) {
            const double uz = U[(size_t)(n / 2) + (size_t)(m / 2) * n + k * nm] / ulid;
            const double xx = k < n ? k * dx : 0.0;
            const double wx = k < n ? W[(size_t)k + (size_t)(m / 2) * n + (size_t)(h / 2) * nm] / ulid : 0.0;
            std::fprintf(f, "%.8f\t%.8f\t%.8f\t%.8f\n", uz, k * dx, xx, wx);
        }
        std::fclose(f);
    }
    if (FILE* f = std::fopen((dir + "channelPlane.txt").c_str(), "w")) {
        std::fprintf(f, "CASE: Uin = %f Ny = %d\n", ulid, m);
        std::fprintf(f, "X\tY\tZ\tU\tV\tW\n");
        const int x = (int)((3. / 4.) * n);
        for (int z = 0; z < h; ++z)
            for (int y = 0; y < m; ++y) {
                const size_t i = (size_t)x + (size_t)y * n + z * nm;
                std::fprintf(f, "%.10f %.10f %.10f %.10f %.10f %.10f\n", x * dx, y * dx, z * dx, U[i], V[i], W[i]);
            }
        std::fclose(f);
    }
}

// Results/lbmsolver.log in the layout of the reference's writeInitLog,
// writeNodeNumbers and writeEndLog (S/LogWriter.cu:5-174).
void write_init_log(const std::string& path, const Args& a, double delta, int n, int m, int h, long inlet_nodes,
                    double ymax, double ymin) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "This is the 3D lattice Boltzmann *.log file\n\n");
    std::fprintf(f, "\n:::: Imported variables from the *.ini file :::: \n");
    std::fprintf(f, ">>> Uavg              : %3.6f\n", a.u);
    std::fprintf(f, ">>> Vavg              : %3.6f\n", a.v);
    std::fprintf(f, ">>> Wavg              : %3.6f\n", a.w);
    std::fprintf(f, ">>> Force             : %3.6f\n", a.g);
    std::fprintf(f, ">>> Initial density   : %2.1f\n", a.rho);
    std::fprintf(f, ">>> Viscosity         : %3.8f\n", a.viscosity);
    std::fprintf(f, ">>> # of iterations   : %1.1ld\n", a.iterations);
    std::fprintf(f, ">>> Autosave after    : %1.1ld\n", a.autosave_after);
    std::fprintf(f, ">>> Autosave every    : %1.1ld\n", a.autosave_every);
    std::fprintf(f, ">>> CollisionModel    : %s\n", a.collision == LBM_TRT ? "TRT" : a.collision == LBM_MRT ? "MRT" : "BGKW");
    std::fprintf(f, ">>> InletProfile      : %s\n", a.inlet_profile == 1 ? "ON" : a.inlet_profile == 3 ? "PULSATILE" : "CONSTANT");
    std::fprintf(f, ">>> OutletProfile     : %s\n",
                 a.outlet == 1 ? "ON" : a.outlet == 2 ? "OPEN (2nd)" : a.outlet == 3 ? "OPEN (1st)" : "He model");
    std::fprintf(f, ">>> CurvedBoundaries  : %s\n", a.curved ? "ON" : "OFF");
    std::fprintf(f, ">>> Residuals type    : %s\n", a.residual_model.find("MacroDiff") != std::string::npos ? "MaxMacroDiffs"
                                                     : a.residual_model.find("FdRelDiff") != std::string::npos ? "MaxRelativeDiff" : "L2");
    std::fprintf(f, ">>> Boundary model    : %s\n", a.hh_walls ? "HHmodel" : "BBack");
    std::fprintf(f, ">>> Results format    : %s\n", a.format == 1 ? "excel " : a.format == 2 ? "Tecplot (*.dat)" : "Paraview (*.vti)");
    std::fprintf(f, ">>> Initial conditions from file : %s\n", a.use_init_file ? "YES" : "NO");
    std::fprintf(f, ">>> Initial conditions filename: %s\n", a.init_file.c_str());
    std::fprintf(f, ">>> Residuals and macro diffs calculated each: %ld iterations\n", a.residual_every);
    std::fprintf(f, ">>> Stop conditions: %f %f %f %f\n", a.stop[0], a.stop[1], a.stop[2], a.stop[3]);
    if (a.boundary_id > 0) std::fprintf(f, ">>> Drag, lift @ BC   : %d\n", a.boundary_id);
    else std::fprintf(f, ">>> Drag, lift was not calculated\n");
    std::fprintf(f, ">>> UpdateInltOutl\t:\t%s\n", a.update_inlet_outlet ? "yes" : "no");
    std::fprintf(f, "\n:::: Calculated variables from mesh :::: \n");
    std::fprintf(f, ">>> Grid spacing             = %f\n", delta);
    std::fprintf(f, ">>> # of nodes in x (n)      = %d\n", n);
    std::fprintf(f, ">>> # of nodes in y (m)      = %d\n", m);
    std::fprintf(f, ">>> # of nodes in z (h)      = %d\n", h);
    std::fprintf(f, ">>> NumInletNodes            = %ld\n", inlet_nodes);
    std::fprintf(f, ">>> MaxInletCoordY           = %f\n", ymax);
    std::fprintf(f, ">>> MinInletCoordY           = %f\n", ymin);
    std::fprintf(f, ">>> MaxInletCoordZ           = %f\n", 0.0);
    std::fprintf(f, ">>> MinInletCoordZ           = %f\n", 0.0);
    std::fclose(f);
}

// Phase times in seconds, keys of S/include/LogWriter.h:248-259.  The step is
// one fused kernel, so collision holds streaming, boundaries and moments too.
struct Times { double overall = 0, iter = 0, init = 0, coll = 0, strm = 0, bndc = 0, macr = 0, resi = 0, writ = 0; };

void write_end_log(const std::string& path, const Times& t) {
    FILE* f = std::fopen(path.c_str(), "a");
    if (!f) return;
    size_t fr = 0, tot = 1;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) { fr = 0; tot = 1; }
    std::fprintf(f, "^^^^ Free : %llu Mbytes \n", (unsigned long long)fr / 1024 / 1024);
    std::fprintf(f, "^^^^ Total: %llu Mbytes \n", (unsigned long long)tot / 1024 / 1024);
    std::fprintf(f, "^^^^ %f%% free, %f%% used\n", 100.0 * fr / (double)tot, 100.0 * (tot - fr) / (double)tot);
    std::fprintf(f, "\nOverall calculations took %f seconds", t.overall);
    std::fprintf(f, "\nMain while loop took %f seconds\n", t.iter);
    std::fprintf(f, "Initialization took %f seconds\n", t.init);
    std::fprintf(f, "Collision took %f seconds\n", t.coll);
    std::fprintf(f, "Streaming took %f seconds\n", t.strm);
    std::fprintf(f, "Calculating Boundaries took %f seconds\n", t.bndc);
    std::fprintf(f, "Update Macroscopic took %f seconds\n", t.macr);
    std::fprintf(f, "Calculating residuals took %f seconds\n", t.resi);
    std::fprintf(f, "Writing results took %f seconds\n", t.writ);
    std::fprintf(f, "\n:::: Iterations done! ::::\n");
    std::fclose(f);
}

// Results/runtimes.dat (S/LogWriter.cu:176-188).
void write_times(const std::string& path, const Times& t) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return;
    std::fprintf(f, "Overall %f\nIteration %f\nInitialisation %f\nCollision %f\nStreaming %f\nBoundaries %f\n"
                    "Macroscopic %f\nResiduals %f\nFile writing %f\n",
                 t.overall, t.iter, t.init, t.coll, t.strm, t.bndc, t.macr, t.resi, t.writ);
    std::fclose(f);
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Colour-gradient steady bubble on the lattice's box (the reference's
// default run: multiPhase = 1, test_case = 1, S/main.cu:52-72, driven by
// S/Iterate.cu:392-500 and S/Iterate3D.cu:495-800).  The box is periodic:
// in 2D always (gpuBcPeriodic2D, period n - 1), in 3D along the axes whose
// faces carry CYCL links (any other boundary link is refused).  The bubble
// holds the nodes within bubble_radius / n of (0.5, 0.5[, 0.5]) in the
// lattice's coordinates, taken as i * delta from the origin as latticegen2
// writes them (initCGBubble / initCGBubble3D, S/GpuInit.cu:323-592; the 3D
// start has f_q = rho phi_q and f_0 = 0).  Every step the Laplace sums give
// the surface-tension error |sigma - R (p_in - p_out)| / sigma in %, with
// p = (3/5)(1 - alpha) rho in 2D and (1/2)(1 - alpha) rho in 3D and
// sigma = 4 A / (9 omega_eff) (S/Multiphase.cu:1082-1102,
// S/Iterate3D.cu:222-230); it goes to "surface tension.txt" in 3D, one line
// per iteration (S/FilesWriting.cu:338-357), and "Oscilating_frequency.txt"
// in 2D, "i value" per line (S/FilesWriting.cu:384-394).
// End-of-run validation of the 2D test cases 2-6 (S/Iterate.cu:663-700 with
// S/Multiphase.cu:1104-1340 and S/FilesWriting.cu:359-394).  The Couette
// profile takes kappa = mu_r / mu_b of the run (the reference's args.kappa,
// fixed at 1 in S/main.cu:56).
template <class Mid, class Phase>
void multiphase_validation(const Args& a, int tcase, const int dims[3], const int sim[3], double dx,
                           const std::vector<double>& hr, const std::vector<double>& hb, lbm_cg_ctx* cg,
                           const std::vector<double>& trace, double sigma, double b_density, double rad, Mid mid,
                           Phase phase) {
    const int n = dims[0], m = dims[1];
    // sum of the positive phase field along the middle row
    auto red_chord = [&]() {
        double s2 = 0;
        for (int i = 0; i < n; ++i) {
            const double ph = phase(i, mid(m));
            if (ph > 0.0) s2 += ph;
        }
        return s2;
    };
    auto write_trace = [&](const char* name) {
        if (FILE* f = std::fopen((a.out_dir + name + ".txt").c_str(), "w")) {
            for (size_t i = 0; i < trace.size(); ++i) std::fprintf(f, "%zu %.10f\n", i, trace[i]);
            std::fclose(f);
        }
    };
    switch (tcase) {
    case 2: {   // analyticalCouette / analyticalPoiseuille, writeCouetteSolution
        const size_t NS = (size_t)sim[0] * sim[1];
        std::vector<double> u(NS);
        lbm_cg_get_macros(cg, nullptr, nullptr, nullptr, u.data(), nullptr);
        const int i = mid(n);
        const double r_mu = a.cg_r_density * a.cg_r_nu, b_mu = b_density * a.cg_b_nu;
        const bool pois = a.g != 0.0;
        const std::string name = pois ? "Profile_Poiseuille" : "Profile_Couette";
        if (FILE* f = std::fopen((a.out_dir + name + ".txt").c_str(), "w")) {
            std::fprintf(f, "y numerical analytical\n");
            for (int j = 0; j < m; ++j) {
                const double y = j * dx, g = a.g;
                double an;
                if (!pois) {
                    const double kappa = r_mu / b_mu;
                    an = j >= mid(m) ? (2.0 * y / (kappa + 1.0) + (kappa - 1.0) / (kappa + 1.0)) * a.u
                                     : (2.0 * kappa * y / (kappa + 1.0)) * a.u;
                } else if (j >= mid(m)) {
                    an = -g * y * y / (2.0 * r_mu) + g * ((3.0 * b_mu + r_mu) * y / (4.0 * r_mu * (r_mu + b_mu))) +
                         g / (2.0 * r_mu) - g * ((3.0 * b_mu + r_mu) / (4.0 * r_mu * (r_mu + b_mu)));
                } else {
                    an = -g * y * y / (2.0 * b_mu) + g * ((3.0 * b_mu + r_mu) * y) / (4.0 * b_mu * (r_mu + b_mu));
                }
                std::fprintf(f, "%.4f %.10f %.10f\n", y, u[(size_t)(j % sim[1]) * sim[0] + (size_t)(i % sim[0])], an);
            }
            std::fclose(f);
        }
        std::printf("%s profile written to %s in Results/\n", pois ? "Poiseuille" : "Couette", name.c_str());
        break;
    }
    case 3: {   // deformingBubbleValid
        const double got = red_chord() / 2.0, want = n / (2.0 * std::sqrt(M_PI));
        std::printf("Predicted radius: %lf Final radius: %lf\n", want, got);
        std::printf("Radius error %% = %lf \n", std::fabs(got - want) / want * 100.0);
        break;
    }
    case 4: {   // validateCoalescenceCase
        const double got = red_chord() / (n * 2.0), want = rad * std::sqrt(2.0);
        std::printf("Predicted radius: %lf Final radius: %lf  (physical units)\n", want, got);
        std::printf("Radius error %% = %lf \n", std::fabs(got - want) / want * 100.0);
        break;
    }
    case 5: {   // writeOscilatingSolution + validateOscilating
        write_trace("Oscilating_frequency");
        std::printf("Oscilation frequency written to Oscilating_frequency in Results/\n");
        const double radius = red_chord() / 2.0;
        std::vector<long> maxima, minima;
        bool rising = false;
        for (size_t i = 500; i + 1 < trace.size(); ++i) {
            const double d = (trace[i + 1] - trace[i - 1]) / 2.0;
            if (rising && d < 0.0) { maxima.push_back((long)i); rising = false; }
            else if (!rising && d > 0.0) { minima.push_back((long)i); rising = true; }
        }
        const double theo = std::sqrt(6.0 * sigma / ((a.cg_r_density + b_density) * radius * radius * radius));
        if (!maxima.empty() && minima.size() > 1) {
            const double num = M_PI / (double)(minima[1] - maxima[0]);
            std::printf("Theoretical %lf vs Numerical %lf\n", theo, num);
            std::printf("Error %% : %lf\n", std::fabs(theo - num) / theo * 100.0);
        } else {
            std::printf("Theoretical %lf vs Numerical: no full oscillation in the run\n", theo);
        }
        break;
    }
    case 6:
        write_trace("RT_Interface");
        break;
    default:
        break;
    }
    (void)hr; (void)hb;
}

// End-of-run validation of the 3D test cases 2-4 (S/Iterate3D.cu:882-901):
// deformingBubbleValid and validateCoalescenceCase on the middle row of the
// middle plane with the sphere's predictions, analyticalCouette and
// writeCouetteSolution on the middle column of the middle plane
// (S/Multiphase.cu:1104-1192, S/FilesWriting.cu:359-382); kappa as in 2D.
template <class Mid>
void multiphase_validation3d(const Args& a, int tcase, const int dims[3], const int sim[3], double dx,
                             const std::vector<double>& hr, const std::vector<double>& hb, lbm_cg_ctx* cg,
                             double b_density, double rad, Mid mid) {
    const int n = dims[0], m = dims[1], h = dims[2];
    auto at = [&](int i, int j, int k) {
        return ((size_t)(k % sim[2]) * sim[1] + (size_t)(j % sim[1])) * sim[0] + (size_t)(i % sim[0]);
    };
    auto red_chord = [&]() {
        const int j = mid(m), k = mid(h);
        double s2 = 0;
        for (int i = 0; i < n; ++i) {
            const size_t x = at(i, j, k);
            const double ph = (hr[x] - hb[x]) / (hr[x] + hb[x]);
            if (ph > 0.0) s2 += ph;
        }
        return s2;
    };
    switch (tcase) {
    case 2: {   // deformingBubbleValid: the cube of side n/2 becomes the sphere of its volume
        const double got = red_chord() / 2.0, want = n / 2.0 / std::cbrt(4.0 / 3.0 * M_PI);
        std::printf("Predicted radius: %lf Final radius: %lf\n", want, got);
        std::printf("Radius error %% = %lf \n", std::fabs(got - want) / want * 100.0);
        break;
    }
    case 3: {   // validateCoalescenceCase: two spheres of radius R merge into one of 2^(1/3) R
        const double got = red_chord() / (n * 2.0), want = rad * std::cbrt(2.0);
        std::printf("Predicted radius: %lf Final radius: %lf  (physical units)\n", want, got);
        std::printf("Radius error %% = %lf \n", std::fabs(got - want) / want * 100.0);
        break;
    }
    case 4: {   // analyticalCouette + writeCouetteSolution
        const size_t NS = (size_t)sim[0] * sim[1] * sim[2];
        std::vector<double> u(NS);
        lbm_cg_get_macros3d(cg, nullptr, nullptr, nullptr, u.data(), nullptr, nullptr);
        const double kappa = (a.cg_r_density * a.cg_r_nu) / (b_density * a.cg_b_nu);
        if (FILE* f = std::fopen((a.out_dir + "Profile_Couette.txt").c_str(), "w")) {
            std::fprintf(f, "y numerical analytical\n");
            for (int j = 0; j < m; ++j) {
                const double y = j * dx;
                const double an = j >= mid(m) ? (2.0 * y / (kappa + 1.0) + (kappa - 1.0) / (kappa + 1.0)) * a.u
                                              : (2.0 * kappa * y / (kappa + 1.0)) * a.u;
                std::fprintf(f, "%.4f %.10f %.10f\n", y, u[at(mid(n), j, mid(h))], an);
            }
            std::fclose(f);
        }
        std::printf("Couette profile written to Profile_Couette in Results/\n");
        break;
    }
    default:
        break;
    }
}

int run_multiphase(const Args& a, const int dims[3], const int sim[3], double dx, Times& tm, double t_start) {
    const bool three_d = dims[2] > 1;
    const double b_density = a.cg_r_density / a.cg_gamma;
    const double r_alpha = 1.0 - (1.0 - a.cg_b_alpha) / a.cg_gamma;
    lbm_cg_params p{};
    p.abi_version = LBM_ABI_VERSION;
    p.nx = sim[0]; p.ny = sim[1]; p.nz = three_d ? sim[2] : 1;
    p.r_alpha = r_alpha; p.b_alpha = a.cg_b_alpha;
    p.r_nu = a.cg_r_nu; p.b_nu = a.cg_b_nu;
    p.beta = a.cg_beta; p.A = a.cg_A; p.g_limit = a.cg_g_limit;
    p.high_order = a.cg_high_order; p.kappa = a.cg_kappa;
    p.storage = a.storage == LBM_STORE_BF16 ? LBM_STORE_F16 : a.storage;
    p.enhanced = a.cg_enhanced;
    const int tcase = a.cg_test_case;
    // walls on the first and last rows for the 2D Couette and Rayleigh-Taylor
    // cases (gpuBcPeriodic2D, S/GpuBoundaries.cu:1876-1968), on the y faces for
    // the 3D ones (the single-phase faces per colour, S/Iterate3D.cu:576-598:
    // the inlet condition for the Couette case, bounce-back for Rayleigh-Taylor);
    // the force of the macroscopic update in every case (S/GpuUpdateMacro.cu:118-119, :340-341)
    p.walls = three_d ? (tcase == 4 ? 1 : tcase == 5 ? 2 : 0) : (tcase == 2 ? 1 : tcase == 6 ? 2 : 0);
    p.u_wall = a.u;
    p.g = a.g;
    p.external_force = a.cg_external_force;
    lbm_cg_ctx* cg = nullptr;
    lbm_status st = lbm_cg_create(&p, &cg);
    if (st != LBM_OK) { std::fprintf(stderr, "colour-gradient set-up failed (%d)\n", (int)st); return 2; }
    const size_t NS = (size_t)sim[0] * sim[1] * sim[2];
    const int q = three_d ? 19 : 9;
    const double rad = a.cg_radius / dims[0];
    std::vector<double> rr(NS, 0.0), rb(NS, 0.0);
    for (size_t i = 0; i < NS; ++i) {
        const int ix = (int)(i % sim[0]), iy = (int)((i / sim[0]) % sim[1]);
        const double x = ix * dx, y = iy * dx;
        const double z = (double)(i / ((size_t)sim[0] * sim[1])) * dx;
        const double d2 = (x - 0.5) * (x - 0.5) + (y - 0.5) * (y - 0.5) + (three_d ? (z - 0.5) * (z - 0.5) : 0.0);
        // initCGBubble's shapes (S/GpuInit.cu:331-470), in the lattice's
        // coordinates and node indices of the full n x m lattice
        const double n = dims[0], m = dims[1];
        bool red = false;
        if (three_d) {   // initCGBubble3D's shapes (S/GpuInit.cu:543-720)
            const double h = dims[2];
            const int iz = (int)(i / ((size_t)sim[0] * sim[1]));
            switch (tcase) {
            case 2:
                red = ix < 0.75 * n && ix > 0.25 * n && iy < 0.75 * m && iy > 0.25 * m && iz < 0.75 * h && iz > 0.25 * h;
                break;
            case 3: {
                const double a2 = (x - 0.5) * (x - 0.5) + (z - 0.5) * (z - 0.5);
                red = std::sqrt(a2 + (y - 0.5 + rad) * (y - 0.5 + rad)) <= rad ||
                      std::sqrt(a2 + (y - 0.5 - rad) * (y - 0.5 - rad)) <= rad;
                break;
            }
            case 4: red = y > 0.5; break;
            case 5: red = y > 1.0 + 0.1 * std::cos(2.0 * M_PI * x); break;
            default: red = std::sqrt(d2) <= rad; break;
            }
            if (red) rr[i] = a.cg_r_density;
            else rb[i] = b_density;
            continue;
        }
        switch (tcase) {
        case 2: red = y > 0.5; break;
        case 3: red = ix < 0.75 * n && ix > 0.25 * n && iy < 0.75 * m && iy > 0.25 * m; break;
        case 4: red = std::hypot(x - 0.5, y - 0.5 + rad) <= rad || std::hypot(x - 0.5, y - 0.5 - rad) <= rad; break;
        case 5: {
            const double ex = (ix - 0.5 * n) / (0.125 * n), ey = (iy - 0.5 * m) / (0.1875 * m);
            red = ex * ex + ey * ey <= 1.0;
            break;
        }
        case 6: red = y > 2.0 + 0.1 * std::cos(2.0 * M_PI * x); break;
        default: red = std::sqrt(d2) <= rad; break;
        }
        if (red) rr[i] = a.cg_r_density;
        else rb[i] = b_density;
    }
    if (three_d) {
        std::vector<double> rf(q * NS), bf(q * NS);
        for (int k = 0; k < q; ++k) {
            const double phi = k == 0 ? 0.0 : (k <= 6 ? 1.0 / 12.0 : 1.0 / 24.0);
            for (size_t i = 0; i < NS; ++i) { rf[k * NS + i] = rr[i] * phi; bf[k * NS + i] = rb[i] * phi; }
        }
        st = lbm_cg_init_f(cg, rf.data(), bf.data());
    } else {
        st = lbm_cg_init(cg, rr.data(), rb.data());
    }
    if (st != LBM_OK) { std::fprintf(stderr, "init: %s\n", lbm_cg_last_error(cg)); lbm_cg_destroy(cg); return 2; }
    const double aux = a.cg_r_density / ((a.cg_r_density + b_density) * a.cg_r_nu) +
                       b_density / ((a.cg_r_density + b_density) * a.cg_b_nu);
    const double om = 1.0 / (3.0 / aux + 0.5);
    const double sigma = 4.0 * a.cg_A / 9.0 / om;
    const double pf = three_d ? 0.5 : 0.6;
    tm.init = now_s() - t_start;
    const double t_loop = now_s();
    std::vector<double> st_err((size_t)a.iterations, 0.0);
    double pin_mean = 0, pout_mean = 0;
    // phase field at node (i, j) of the full lattice (periodic copies map back)
    std::vector<double> hr(NS), hb(NS);
    auto phase = [&](int i, int j) {
        const size_t s2 = (size_t)(j % sim[1]) * sim[0] + (size_t)(i % sim[0]);
        return (hr[s2] - hb[s2]) / (hr[s2] + hb[s2]);
    };
    auto mid = [](int k) { return k % 2 == 0 ? k / 2 : (k + 1) / 2; };
    // the zero of the phase field between rows j and j2 on column i, in y
    auto crossing = [&](int i, int j, int j2) {
        const double p1 = phase(i, j), p2 = phase(i, j2);
        const double slope = (p1 - p2) / ((j - j2) * dx);
        return -(p1 - slope * j * dx) / slope;
    };
    // getMaxYOscilating / getMinYRT (S/Multiphase.cu:1211-1269) on the middle column
    auto interface_y = [&](bool up) {
        const int i = mid(dims[0]);
        int j = mid(dims[1]);
        if (up) {
            while (j < dims[1] - 1 && phase(i, j) >= 0.0) ++j;
            return crossing(i, j, j - 1);
        }
        while (j > 0 && phase(i, j) >= 0.0) --j;
        return crossing(i, j, j + 1);
    };
    int code = 0;
    for (long it = 0; it < a.iterations; ++it) {
        double t0 = now_s();
        st = lbm_cg_step(cg, 1);
        tm.coll += now_s() - t0;
        if (st != LBM_OK) { std::fprintf(stderr, "step: %s\n", lbm_cg_last_error(cg)); code = 2; break; }
        t0 = now_s();
        double red[12];
        lbm_cg_reduce(cg, red);
        tm.resi += now_s() - t0;
        if (red[11] > 0) { std::fprintf(stderr, "\nDIVERGENCE!\n"); code = 1; break; }
        if (!three_d && (tcase == 5 || tcase == 6)) {   // the interface trace (S/Iterate.cu:505-515)
            lbm_cg_get_macros(cg, hr.data(), hb.data(), nullptr, nullptr, nullptr);
            st_err[it] = interface_y(tcase == 5);
        } else {
            pin_mean = red[2] / red[3];
            pout_mean = red[4] / red[5];
            const double pin = pf * (1.0 - r_alpha) * pin_mean, pout = pf * (1.0 - a.cg_b_alpha) * pout_mean;
            st_err[it] = std::fabs(sigma - a.cg_radius * (pin - pout)) / sigma * 100.0;
        }
        if ((it + 1) % (a.residual_every > 0 ? a.residual_every : a.iterations) == 0)
            std::printf("iteration %ld surface tension error %.4f %%, max |u| %.3e\n", it + 1, st_err[it], red[10]);
    }
    tm.iter = now_s() - t_loop;
    const double t_write = now_s();
    if (a.iterations > 0 && tcase == 1) std::printf("Suface tension error: %lf\n", st_err[(size_t)a.iterations - 1]);
    if (tcase != 1) {
        lbm_cg_get_macros(cg, hr.data(), hb.data(), nullptr, nullptr, nullptr);
        if (three_d) multiphase_validation3d(a, tcase, dims, sim, dx, hr, hb, cg, b_density, rad, mid);
        else multiphase_validation(a, tcase, dims, sim, dx, hr, hb, cg, st_err, sigma, b_density, rad, mid, phase);
    } else if (three_d) {   // WriteArray (S/Iterate3D.cu:871-872)
        if (FILE* f = std::fopen((a.out_dir + "surface tension.txt").c_str(), "w")) {
            for (double e : st_err) std::fprintf(f, "%.10f\n", e);
            std::fclose(f);
        }
    } else {         // writeOscilatingSolution (S/Iterate.cu:667-673, S/FilesWriting.cu:384-394)
        std::printf("Pressure inside %lf Pressure outside %lf ST_predicted %lf\n", pin_mean, pout_mean, sigma);
        if (FILE* f = std::fopen((a.out_dir + "Oscilating_frequency.txt").c_str(), "w")) {
            for (size_t i = 0; i < st_err.size(); ++i) std::fprintf(f, "%zu %.10f\n", i, st_err[i]);
            std::fclose(f);
        }
    }
    std::vector<double> sr(NS), sb(NS), sp(NS), su(NS), sv(NS), sw(NS);
    lbm_cg_get_macros3d(cg, sr.data(), sb.data(), sp.data(), su.data(), sv.data(), sw.data());
    lbm_cg_destroy(cg);
    const size_t N = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<double> R(N), U(N), V(N), W(N), RR(N), RB(N);
    for (int z = 0; z < dims[2]; ++z)
        for (int y = 0; y < dims[1]; ++y)
            for (int x = 0; x < dims[0]; ++x) {
                const size_t d = ((size_t)z * dims[1] + y) * dims[0] + x;
                const size_t s = ((size_t)(z % sim[2]) * sim[1] + (y % sim[1])) * sim[0] + (x % sim[0]);
                R[d] = sp[s]; U[d] = su[s]; V[d] = sv[s]; W[d] = sw[s]; RR[d] = sr[s]; RB[d] = sb[s];
            }
    const char* ext = ext_of(a.format);
    write_results(a.out_dir + "FinalData." + ext, a.format, dims[0], dims[1], dims[2], dx, R, U, V, W, &RR, &RB);
    if (three_d) write_3d_lines(a.out_dir, dims[0], dims[1], dims[2], dx, U, V, W, a.u);
    std::printf("results written to %sFinalData.%s\n", a.out_dir.c_str(), ext);
    tm.writ += now_s() - t_write;
    tm.overall = now_s() - t_start;
    write_end_log(a.out_dir + "lbmsolver.log", tm);
    write_times(a.out_dir + "runtimes.dat", tm);
    return code;
}

}  // namespace

int main(int argc, char** argv) {
    const double t_start = now_s();
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
        else if (s == "-l" || s == "--outlet") {   // S/Arguments.cu:55-75
            std::string o = next();
            if (o == "yes") a.outlet = 1;
            else if (o == "second") a.outlet = 2;
            else if (o == "first") a.outlet = 3;
            else std::fprintf(stderr, "Invalid outlet profile (%s), going with default\n", o.c_str());
        }
        else if (s == "-c") { std::string c = next(); a.collision = c == "TRT" ? LBM_TRT : c == "MRT" ? LBM_MRT : LBM_BGK; from_cli = true; }
        else if (s == "--format") {
            std::string f = next();
            a.format = f == "tecplot" ? 2 : f == "csv" ? 1 : f == "vti-appended" ? 4 : 3;
        }
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
        // a binary copy of the lattice (lbm_write_lattice_binary), when there is
        // one, is read instead of the text pair
        const std::string bin = "Mesh/" + a.lattice + ".lbl";
        if (FILE* fb = std::fopen(bin.c_str(), "rb")) {
            std::fclose(fb);
            a.node_file = bin;
        }
    }

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
    // the reference empties AlertLog.txt and writes the MacroDiffs.txt header
    // at start-up (S/main.cu:25, :125-127, S/LogWriter.cu:222-226)
    mkdir(a.out_dir.c_str(), 0755);
    if (FILE* af = std::fopen((a.out_dir + "AlertLog.txt").c_str(), "w")) std::fclose(af);
    if (FILE* hf = std::fopen((a.out_dir + "MacroDiffs.txt").c_str(), "w")) {
        std::fprintf(hf, "iteration\tMax u diff\tMax v diff\tMax w diff\tMax rho diff\n");
        std::fclose(hf);
    }
    if (a.multiphase) {
        // 2D: periodic box (period n - 1 on both axes); 3D: CYCL faces only
        if (!three_d) {
            sim[0] = dims[0] - 1;
            // the Couette and Rayleigh-Taylor cases keep every row: walls on the first and last
            sim[1] = a.cg_test_case == 2 || a.cg_test_case == 6 ? dims[1] : dims[1] - 1;
        }
        else if (a.cg_test_case == 4 || a.cg_test_case == 5) {
            // walls on the y faces: every y plane is simulated
            if (!cyc[0] || !cyc[2] || cyc[1]) {
                std::fprintf(stderr, "the 3D Couette and Rayleigh-Taylor cases need a lattice cyclic along x and z "
                                     "with walls on the y faces\n");
                return 2;
            }
        } else if (tc[1] + tc[2] + tc[3] + tc[5] > 0) {
            std::fprintf(stderr, "the colour-gradient run needs a periodic lattice (CYCL links only)\n");
            return 2;
        }
        write_init_log(a.out_dir + "lbmsolver.log", a, dx, dims[0], dims[1], dims[2], 0, 0.0, 0.0);
        Times tmp;
        return run_multiphase(a, dims, sim, dx, tmp, t_start);
    }
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
    // inlet profile u0 (S/GpuInit.cu:994-1006): parabolic between half a
    // spacing beyond the extreme west-inlet nodes, or constant; the reference
    // defines it at every node, and its velocity outlet reads it on the east
    // face.  A north lid moves at the constant U (the parabola spans the west
    // inlet's y range).
    std::vector<double> ubc(NS, 0.0), vbc(NS, 0.0);
    {
        int ymin = 1 << 30, ymax = -1;
        for (size_t i = 0; i < NS; ++i)
            if (scls[i] == LBM_VEL_W) {
                const int y = (int)((i / sim[0]) % sim[1]);
                ymin = std::min(ymin, y);
                ymax = std::max(ymax, y);
            }
        const bool prof = a.inlet_profile == 1 && ymax >= 0;
        for (size_t i = 0; i < NS; ++i)
            if (scls[i] == LBM_VEL_W || scls[i] == LBM_PRES_E) {
                const double y = (double)((i / sim[0]) % sim[1]), lo = ymin - 0.5, hi = ymax + 0.5;
                ubc[i] = prof ? 6.0 * a.u * (y - lo) * (hi - y) / ((hi - lo) * (hi - lo)) : a.u;
                vbc[i] = a.v;
            } else if (scls[i] == LBM_VEL_N) {
                ubc[i] = a.u;
                vbc[i] = 0.0;
            }
    }
    std::vector<double> rho(NS, a.rho), u(NS, 0.0), v(NS, 0.0), w(NS, 0.0);
    for (size_t i = 0; i < NS; ++i) { u[i] = scls[i] == LBM_VEL_W ? ubc[i] : a.u; v[i] = a.v; w[i] = a.w; }
    bool from_file = false;
    if (a.use_init_file) {
        std::vector<double> fu, fv, fw, fr;
        if (read_initial_conditions(a.init_file, N, fu, fv, fw, fr)) {
            for (int z = 0; z < sim[2]; ++z)
                for (int y = 0; y < sim[1]; ++y)
                    for (int x = 0; x < sim[0]; ++x) {
                        const size_t s = ((size_t)z * sim[1] + y) * sim[0] + x, d = ((size_t)z * dims[1] + y) * dims[0] + x;
                        u[s] = fu[d]; v[s] = fv[d]; w[s] = fw[d]; rho[s] = fr[d];
                    }
            from_file = true;
            std::printf("Initial conditions loaded from file %s\n", a.init_file.c_str());
        } else {
            std::fprintf(stderr, "wrong or missing initial condition file %s: uniform start\n", a.init_file.c_str());
        }
    }
    // 3D velocity faces move at (U, V, W), or at the loaded velocities
    // (S/Iterate3D.cu:186-205: u1 = the start fields; the parabolic 3D profile
    // is not available in the reference either)
    std::vector<double> wbc(NS, 0.0);
    if (three_d)
        for (size_t i = 0; i < NS; ++i) {
            const uint8_t c = scls[i];
            if (c == LBM_VEL_W || c == LBM_VEL_N || (c >= LBM_VEL_E && c <= LBM_VEL_T)) {
                ubc[i] = from_file ? u[i] : a.u;
                vbc[i] = from_file ? v[i] : a.v;
                wbc[i] = from_file ? w[i] : a.w;
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
    // BCwallModel = HHmodel selects the wet-node walls in 3D (gpuBcComplexWall3D);
    // the reference's 2D walls are bounce-back whatever the setting
    // Walls = CURVED selects Bouzidi walls with the BC file's wall fractions,
    // unless the 3D wet-node walls take the wall links (S/Iterate3D.cu:604-614)
    p.wall_model = three_d && a.hh_walls && !a.walls_set ? LBM_WALL_HH
                 : a.curved && !a.walls_set ? LBM_WALL_CURVED : a.wall_model;
    std::vector<double> sq;
    if (p.wall_model == LBM_WALL_CURVED) {
        std::vector<double> fq((size_t)N * 19);
        if (lbm_read_wall_q(a.node_file.c_str(), a.bc_file.c_str(), fq.data()) != LBM_OK) {
            std::fprintf(stderr, "cannot read wall fractions from %s\n", a.bc_file.c_str());
            return 2;
        }
        sq.assign(NS * 19, 0.0);
        for (int z = 0; z < sim[2]; ++z)
            for (int y = 0; y < sim[1]; ++y)
                for (int x = 0; x < sim[0]; ++x) {
                    const size_t s2 = ((size_t)z * sim[1] + y) * sim[0] + x, d = ((size_t)z * dims[1] + y) * dims[0] + x;
                    for (int k = 0; k < 19; ++k) sq[s2 * 19 + k] = fq[d * 19 + k];
                }
    }
    p.inlet_model = a.inlet_model;
    // UpdateInltOutl = no: the 3D macro update leaves inlet nodes alone, so
    // their collision keeps the initial macros (S/GpuUpdateMacro.cu:150-151)
    if (three_d && !a.update_inlet_outlet) p.parity_flags |= LBM_PARITY_INLET_FROZEN;
    if (!three_d) {
        p.outlet_model = a.outlet == 1 ? LBM_OUTLET_REF_VELOCITY : a.outlet == 2 ? LBM_OUTLET_EXTRAP2
                       : a.outlet == 3 ? LBM_OUTLET_EXTRAP1 : LBM_OUTLET_PRESSURE;
    } else {
        // 3D (S/GpuBoundaries.cu:1666-1852): OUTLET only prints "not implemented,
        // go for open boundary 2nd order", which is what it gets here; HEmodel
        // is the f_eq outlet; [mi355x] outlet = pressure the pressure face
        p.outlet_model = a.outlet == 1 || a.outlet == 2 ? LBM_OUTLET_EXTRAP2 : a.outlet == 3 ? LBM_OUTLET_EXTRAP1
                       : a.outlet == 5 ? LBM_OUTLET_PRESSURE : LBM_OUTLET_FEQ;
    }
    p.rho_out = a.rho_out > 0 ? a.rho_out : a.rho;
    lbm_lattice lat = {scls.data(), sbb.data(), ubc.data(), vbc.data(), wbc.data(), sq.empty() ? nullptr : sq.data()};
    lbm_ctx* ctx = nullptr;
    if (lbm_create(&p, &lat, &ctx) != LBM_OK) { std::fprintf(stderr, "lbm_create failed\n"); return 2; }
    if (lbm_init(ctx, a.init_eq ? 0 : 1, rho.data(), u.data(), v.data(), w.data()) != LBM_OK) {
        std::fprintf(stderr, "lbm_init: %s\n", lbm_last_error(ctx));
        lbm_destroy(ctx);
        return 2;
    }
    mkdir(a.out_dir.c_str(), 0755);
    Times tm;
    {
        long inlet_nodes = 0;
        double ylo = 1e300, yhi = -1e300;
        for (size_t i = 0; i < NS; ++i)
            if (scls[i] == LBM_VEL_W) {
                ++inlet_nodes;
                const double y = (double)((i / sim[0]) % sim[1]) * dx;
                ylo = std::min(ylo, y);
                yhi = std::max(yhi, y);
            }
        if (!inlet_nodes) ylo = yhi = 0;
        const std::string logp = a.out_dir + "lbmsolver.log";
        write_init_log(logp, a, dx, dims[0], dims[1], dims[2], inlet_nodes, yhi, ylo);
        if (FILE* lf = std::fopen(logp.c_str(), "a")) {
            long fluid = 0;
            for (size_t i = 0; i < NS; ++i) fluid += scls[i] != LBM_SOLID;
            std::fprintf(lf, "\n:::: Initializing ::::\n");
            std::fprintf(lf, ">>> Number of nodes                      = %zu\n", N);
            std::fprintf(lf, ">>> Number of fluid nodes                = %ld\n", fluid);
            std::fprintf(lf, ">>> Device memory of the solver          = %lld Mbytes\n",
                         (long long)(lbm_device_bytes(ctx) / 1024 / 1024));
            std::fprintf(lf, "\n:::: Initialization done! ::::\n");
            std::fprintf(lf, "\n:::: Start Iterations ::::\n");
            std::fclose(lf);
        }
    }
    tm.init = now_s() - t_start;
    const double t_loop = now_s();
    FILE* rf = std::fopen((a.out_dir + "residuals.dat").c_str(), "w");
    if (rf) std::fprintf(rf, "Iter L2_norm L2_norm_weighted Drag Lift\n");
    // ResidualModel (ini line 19): MacroDiff compares the largest change of u,
    // v, w, rho since the last check with the four stop values and logs them to
    // MacroDiffs.txt (S/Iterate3D.cu:696-756, S/LogWriter.cu:215-221); L2 is
    // sqrt(sum (f - fColl)^2) against the first stop value (S/Iterate.cu:551);
    // FdRelDiff is max |f - f_prev| / f1 against the first stop value
    // (S/Iterate3D.cu:677-693, S/ComputeResiduals.cu:28-36).
    const bool macro_diff = a.residual_model.find("MacroDiff") != std::string::npos;
    const bool fd_rel = a.residual_model.find("FdRelDiff") != std::string::npos;
    FILE* mf = macro_diff ? std::fopen((a.out_dir + "MacroDiffs.txt").c_str(), "a") : nullptr;
    double r4[8];
    lbm_reduce_all(ctx, r4);
    const long every = a.residual_every > 0 ? a.residual_every : a.iterations;
    long done = 0;
    int code = 0;
    // Autosave (S/Iterate3D.cu:803-837, S/Iterate.cu:590-621): at every
    // multiple of AutosaveEach beyond AutosaveAfter, the fields go to
    // autosave_iter%05d.<ext> (3D) or over FinalData (2D; ParaView output to
    // Video/FinalData_<k>.vti)
    const long save_every = a.autosave_every > 0 ? a.autosave_every : 0;
    long save_k = 1;
    auto autosave = [&]() {
        std::vector<double> sr(NS), su(NS), sv(NS), sw(NS);
        lbm_get_macros(ctx, sr.data(), su.data(), sv.data(), sw.data());
        std::vector<double> R(N), U(N), V(N), W(N);
        for (int z = 0; z < dims[2]; ++z)
            for (int y = 0; y < dims[1]; ++y)
                for (int x = 0; x < dims[0]; ++x) {
                    const size_t d = ((size_t)z * dims[1] + y) * dims[0] + x;
                    const size_t q = ((size_t)(z % sim[2]) * sim[1] + (y % sim[1])) * sim[0] + (x % sim[0]);
                    R[d] = sr[q]; U[d] = su[q]; V[d] = sv[q]; W[d] = sw[q];
                }
        char name[768];
        if (three_d) {
            std::snprintf(name, sizeof name, "%sautosave_iter%05ld.%s", a.out_dir.c_str(), done, ext_of(a.format));
        } else if (a.format >= 3) {
            mkdir((a.out_dir + "Video").c_str(), 0755);
            std::snprintf(name, sizeof name, "%sVideo/FinalData_%ld.vti", a.out_dir.c_str(), save_k);
        } else {
            std::snprintf(name, sizeof name, "%sFinalData.%s", a.out_dir.c_str(), ext_of(a.format));
        }
        std::printf("autosave\n\n");
        write_results(name, a.format, dims[0], dims[1], dims[2], dx, R, U, V, W);
    };
    while (done < a.iterations) {
        long k = std::min(every - done % every, a.iterations - done);
        if (save_every) k = std::min(k, save_every * save_k - done);
        double t0 = now_s();
        lbm_status st = lbm_step(ctx, k);
        tm.coll += now_s() - t0;
        if (st != LBM_OK) { std::fprintf(stderr, "step: %s\n", lbm_last_error(ctx)); code = 2; break; }
        done += k;
        if (save_every && done == save_every * save_k) {
            ++save_k;
            if (done > a.autosave_after) {
                t0 = now_s();
                autosave();
                tm.writ += now_s() - t0;
            }
        }
        if (done % every != 0 && done != a.iterations) continue;
        t0 = now_s();
        lbm_reduce_all(ctx, r4);
        double l2 = 0;
        if (fd_rel) lbm_residual_fdrel(ctx, &l2);
        else if (!macro_diff) lbm_residual_l2(ctx, &l2);
        tm.resi += now_s() - t0;
        if (rf) std::fprintf(rf, "%ld %5.4e %5.4e %f %f\n", done, l2, l2 / (double)NS, 0.0, 0.0);
        if (r4[3] > 0 || l2 != l2) { std::fprintf(stderr, "\nDIVERGENCE!\n"); code = 1; break; }
        if (macro_diff) {
            if (mf) std::fprintf(mf, "%ld\t\t%.10f\t%.10f\t%.10f\t%.10f\n", done, r4[4], r4[5], r4[6], r4[7]);
            std::printf("iteration %ld max diffs u %.4e v %.4e w %.4e rho %.4e\n", done, r4[4], r4[5], r4[6], r4[7]);
            if (r4[4] < a.stop[0] && r4[5] < a.stop[1] && r4[6] < a.stop[2] && r4[7] < a.stop[3]) {
                std::printf("simulation converged!\n");
                break;
            }
        } else {
            std::printf("iteration %ld %s residual %.4e mass %.10f\n", done, fd_rel ? "FdRelDiff" : "L2", l2, r4[2]);
            if (l2 < a.stop[0]) { std::printf("simulation converged!\n"); break; }
        }
    }
    if (mf) std::fclose(mf);
    if (rf) std::fclose(rf);
    tm.iter = now_s() - t_loop;
    const double t_write = now_s();
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
    const char* ext = ext_of(a.format);
    write_results(a.out_dir + "FinalData." + ext, a.format, dims[0], dims[1], dims[2], dx, R, U, V, W);
    if (three_d) write_3d_lines(a.out_dir, dims[0], dims[1], dims[2], dx, U, V, W, a.u);
    std::printf("results written to %sFinalData.%s\n", a.out_dir.c_str(), ext);
    tm.writ += now_s() - t_write;
    tm.overall = now_s() - t_start;
    write_end_log(a.out_dir + "lbmsolver.log", tm);
    write_times(a.out_dir + "runtimes.dat", tm);
    return code;
}
