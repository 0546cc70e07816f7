// backend.cpp -- the JSON backend behind dplus_backend_*: the surface of
// PythonBackendWrapper (Backend/Backend/PythonBackendWrapper.h:15-46) for the
// hot path's model types.
//
// StartGenerate parses the state (CalculationInput args: {"state", "x", ...};
// PythonInterface/dplus/CalculationInput.py:56-68), builds the model tree
// (PDB leaves and Manual Symmetry nodes, LocalBackendParameterTree.cpp:150-272),
// and runs one job on a worker thread (JobManager.cpp:137-205): every PDB leaf
// becomes a device-resident amplitude grid (cached across jobs, the role of
// AmplitudeCache), symmetry nodes are docked either into grids (all-grids
// regime) or flattened into placements for OA-time docking (hybrid regime,
// Amplitude.cpp:577-691), and each population is orientation averaged.
// The curve is Scale * sum_pop PopulationSize * I_pop + Constant
// (CompositeModel.cpp:216-262).  Status and results use the reference's JSON
// shapes (Conversions.cpp:49-64, BackendWrapper.cpp:540-581).

#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <thread>
#include <vector>

#include "device.hpp"
#include "elements.hpp"
#include "json.hpp"

namespace dplus {

void hydration_boxes(int n, const double *xyz, const double *rad, double probe, double thickness, double vox,
                     bool fill_holes, std::vector<int> &dims, std::vector<double> &com);
void solvent_boxes(int n, const double *xyz, const double *rad, double probe, double thickness, double vox,
                   bool fill_holes, bool want_layer, bool want_excl, std::vector<int> &ldims, std::vector<double> &lcom,
                   std::vector<int> &edims, std::vector<double> &ecom);
void euler_rt(double a, double b, double c, double m[9]);
std::vector<std::array<double, 6>> lua_symmetry_rows(const std::string &script,
                                                     const std::vector<std::vector<double>> &params, int *nlp);
std::string amp_legacy_bytes(Grid &g, const std::string &header, hipStream_t s);
void amp_write(Grid &g, const std::string &path, const std::string &title, hipStream_t s);
Grid *amp_read(const std::string &path, int precision, std::string &header, hipStream_t s);

namespace {

// ------------------------------------------------------------------ model tree
struct ModelNode {
    std::string type, name, filename;
    unsigned ptr = 0;
    bool centered = true, use_grid = true;
    std::vector<double> extra;
    std::vector<std::vector<double>> rows;
    std::vector<std::vector<double>> script_params;  // a scripted symmetry's layer parameters
    double loc[6] = {0, 0, 0, 0, 0, 0};
    std::vector<ModelNode> children;
};

ModelNode parse_model(const json::Value &m, std::vector<unsigned> &ptrs) {
    ModelNode n;
    n.type = m.at("Type").text();
    if (const json::Value *v = m.find("Name")) n.name = v->text();
    if (const json::Value *v = m.find("Filename")) n.filename = v->text();
    if (const json::Value *v = m.find("ModelPtr")) n.ptr = (unsigned)v->number();
    if (const json::Value *v = m.find("Centered")) n.centered = v->truth(true);
    if (const json::Value *v = m.find("Use_Grid")) n.use_grid = v->truth(true);
    if (const json::Value *v = m.find("ExtraParameters"))
        for (auto &x : v->arr) n.extra.push_back(x.number());
    if (const json::Value *v = m.find("Parameters"))
        for (auto &r : v->arr) {
            std::vector<double> row;
            if (r.is_array())
                for (auto &x : r.arr) row.push_back(x.number());
            else
                row.push_back(r.number());
            n.rows.push_back(row);
        }
    if (const json::Value *l = m.find("Location")) {
        const char *k[6] = {"x", "y", "z", "alpha", "beta", "gamma"};
        for (int i = 0; i < 6; i++)
            if (const json::Value *v = l->find(k[i])) n.loc[i] = v->number();
    }
    ptrs.push_back(n.ptr);
    if (const json::Value *c = m.find("Children"))
        for (auto &ch : c->arr) n.children.push_back(parse_model(ch, ptrs));
    std::string lower = n.type;
    std::transform(lower.begin(), lower.end(), lower.begin(), ::tolower);
    if (lower == "scripted symmetry") {
        // LuaSymmetry (Symmetry.cpp:404-521): the script is the text of the file
        // the model names (ParameterTreeConverter::GetScript,
        // ParamTreeConversions.cpp:340-349); its Populate(p, nlayers) turns the
        // layer parameters into rows {x, y, z, alpha, beta, gamma} (degrees),
        // which then act as a manual symmetry's rows
        if (n.filename.empty()) throw Error(kErrInvalidArgs, "invalid scripted symmetry, no script (model.Filename) defined");
        std::ifstream f(n.filename, std::ios::binary);
        if (!f) throw Error(kErrFileNotFound, "cannot open symmetry script " + n.filename);
        std::stringstream text;
        text << f.rdbuf();
        n.type = "Scripted Symmetry";
        n.script_params = n.rows;
        n.rows.clear();
        for (const auto &r : lua_symmetry_rows(text.str(), n.script_params, nullptr))
            n.rows.emplace_back(r.begin(), r.end());
    }
    return n;
}

struct Population {
    double size = 1.0;
    std::vector<ModelNode> models;
};

struct Spec {
    int grid_size = 100, method = DPLUS_OA_MC, precision = DPLUS_PRECISION_MIXED;
    double qmax = 1.0, qmin = 0.0, eps = 0.0, scale = 1.0, constant = 0.0;
    long long iters = 1000000;
    bool two_d = false;  // StartGenerate2D: the (q_z, q_perp) matrix
    bool use_grid = true;
    unsigned long long seed = 1;
    std::vector<Population> pops;
    std::vector<double> q;
    std::vector<unsigned> ptrs;
};

bool is_pdb(const ModelNode &n) { return n.type == "PDB"; }
bool is_amp(const ModelNode &n) { return n.type == "AMP"; }
bool is_scripted_symmetry(const ModelNode &n) { return n.type == "Scripted Symmetry"; }
// a scripted symmetry holds its evaluated rows and is docked as a manual symmetry
bool is_manual_symmetry(const ModelNode &n) {
    return n.type == ",26" || n.type == "Manual Symmetry" || is_scripted_symmetry(n);
}
const char *symmetry_kind(const ModelNode &n);

Spec parse_state(const json::Value &doc) {
    Spec s;
    const json::Value &x = doc.at("x");
    for (auto &v : x.arr) s.q.push_back(v.number());
    if (s.q.empty()) throw Error(kErrInvalidArgs, "empty q vector");
    const json::Value &state = doc.at("state");
    const json::Value &dp = state.at("DomainPreferences");
    s.grid_size = (int)dp.at("GridSize").number();
    s.qmax = s.q.back();  // BackendWrapper.cpp:335
    s.iters = (long long)dp.at("OrientationIterations").number();
    s.eps = dp.at("Convergence").number();
    if (const json::Value *v = dp.find("qMin")) s.qmin = v->number(0.0);
    s.use_grid = dp.at("UseGrid").truth(true);
    const std::string method = dp.at("OrientationMethod").text();  // Common/Common.h:396-402
    if (method.find("VEGAS") != std::string::npos)
        s.method = DPLUS_OA_VEGAS;
    else if (method.find("Gauss Kronrod") != std::string::npos)
        s.method = DPLUS_OA_GK;
    else if (method.find("Monte Carlo") != std::string::npos)
        s.method = DPLUS_OA_MC;
    else
        throw Error(kErrUnimplemented, "orientation method \"" + method + "\" is not on this path");
    if (const json::Value *v = dp.find("Precision"))
        s.precision = v->text() == "FP64" || v->text() == "fp64" ? DPLUS_PRECISION_FP64 : DPLUS_PRECISION_MIXED;
    if (const json::Value *v = dp.find("Seed")) s.seed = (unsigned long long)v->number();
    const json::Value &dom = state.at("Domain");
    if (const json::Value *v = dom.find("Scale")) s.scale = v->number(1.0);
    if (const json::Value *v = dom.find("Constant")) s.constant = v->number(0.0);
    if (const json::Value *v = dom.find("ModelPtr")) s.ptrs.push_back((unsigned)v->number());
    for (auto &p : dom.at("Populations").arr) {
        Population pop;
        if (const json::Value *v = p.find("PopulationSize")) pop.size = v->number(1.0);
        if (const json::Value *v = p.find("ModelPtr")) s.ptrs.push_back((unsigned)v->number());
        for (auto &m : p.at("Models").arr) pop.models.push_back(parse_model(m, s.ptrs));
        s.pops.push_back(std::move(pop));
    }
    size_t nmodels = 0;
    for (auto &p : s.pops) nmodels += p.models.size();
    if (nmodels == 0) throw Error(kErrInvalidArgs, "the state has no models");  // BackendWrapper.cpp:351-354
    return s;
}

// ------------------------------------------------------------------ placements
void rot_matrix(double a, double b, double c, double R[9]) {
    double rt[9];
    euler_rt(a, b, c, rt);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R[3 * i + j] = rt[3 * j + i];
}

void matmul(const double A[9], const double B[9], double C[9]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) C[3 * i + j] = A[3 * i] * B[j] + A[3 * i + 1] * B[3 + j] + A[3 * i + 2] * B[6 + j];
}

void matvec(const double A[9], const double v[3], double o[3]) {
    for (int i = 0; i < 3; i++) o[i] = A[3 * i] * v[0] + A[3 * i + 1] * v[1] + A[3 * i + 2] * v[2];
}

bool is_space_filling(const ModelNode &n) { return n.type == ",25" || n.type == "Space-filling Symmetry"; }
bool is_symmetry(const ModelNode &n) { return is_manual_symmetry(n) || is_space_filling(n); }
const char *symmetry_kind(const ModelNode &n) {
    return is_space_filling(n) ? "Space-filling Symmetry" : is_scripted_symmetry(n) ? "Scripted Symmetry" : "Manual Symmetry";
}

struct SubLoc {
    double R[9];
    double t[3];
};

// space-filling rows: (distance nm, angle deg, repetitions) for a, b, c
LatticeDev lattice_of(const ModelNode &n) {
    if (n.rows.size() != 3) throw Error(kErrInvalidParamTree, "a space-filling symmetry has three layers (a, b, c)");
    double d[3], ang[3], reps[3];
    for (int k = 0; k < 3; k++) {
        if (n.rows[k].size() < 3) throw Error(kErrInvalidParamTree, "space-filling layers are (distance, angle, repetitions)");
        d[k] = n.rows[k][0];
        ang[k] = n.rows[k][1];
        reps[k] = n.rows[k][2];
    }
    return space_filling_lattice(d, ang, reps);
}

// ISymmetry::GetSubLocation: manual rows (x, y, z, alpha, beta, gamma in degrees,
// ManualSymmetry::PreCalculate Symmetries.cpp:670-716) or the lattice points of a
// space-filling symmetry with c fastest (GridSymmetry::GetSubLocation, Symmetries.cpp:461-477)
std::vector<SubLoc> sub_locations(const ModelNode &n) {
    std::vector<SubLoc> out;
    if (is_space_filling(n)) {
        const LatticeDev L = lattice_of(n);
        const int na = (int)L.n[0], nb = (int)L.n[1], nc = (int)L.n[2];
        for (int i = 0; i < na; i++)
            for (int j = 0; j < nb; j++)
                for (int k = 0; k < nc; k++) {
                    SubLoc s{{1, 0, 0, 0, 1, 0, 0, 0, 1}, {0, 0, 0}};
                    for (int c = 0; c < 3; c++) s.t[c] = i * L.v[c] + j * L.v[3 + c] + k * L.v[6 + c];
                    out.push_back(s);
                }
        return out;
    }
    for (auto &row : n.rows) {
        if (row.size() < 6) throw Error(kErrInvalidParamTree, "manual symmetry rows need x, y, z, alpha, beta, gamma");
        SubLoc s;
        rot_matrix(row[3] * kPi / 180.0, row[4] * kPi / 180.0, row[5] * kPi / 180.0, s.R);
        for (int c = 0; c < 3; c++) s.t[c] = row[c];
        out.push_back(s);
    }
    return out;
}

std::string file_key(const std::string &path) {
    struct stat st;
    long long mt = 0, sz = 0;
    if (stat(path.c_str(), &st) == 0) {
        mt = (long long)st.st_mtime;
        sz = (long long)st.st_size;
    }
    return path + "|" + std::to_string(mt) + "|" + std::to_string(sz);
}

// the reference's PDB reader holds coordinates and weights in float
// (PDBReaderOb<float>) and centres in float, in file order (moveCOMToOrigin,
// PDBReaderLib.cpp:271-292); the same arithmetic here keeps the voxelised
// layer on the reference's voxels
std::vector<float> centred_coords(const PdbAtoms &A, bool centered) {
    const int natoms = (int)A.species.size();
    std::vector<float> xf(3 * (size_t)natoms);
    for (size_t i = 0; i < xf.size(); i++) xf[i] = (float)A.xyz[i];
    if (centered) {
        float c[3] = {0.f, 0.f, 0.f}, w = 0.f;
        for (int a = 0; a < natoms; a++) {
            const float wt = (float)kElements[A.species[a]].weight;
            for (int k = 0; k < 3; k++) c[k] += xf[3 * a + k] * wt;
            w += wt;
        }
        for (int k = 0; k < 3; k++) c[k] /= w;
        for (int a = 0; a < natoms; a++)
            for (int k = 0; k < 3; k++) xf[3 * a + k] -= c[k];
    }
    return xf;
}

// ---- PDB text of a model (GetPDB)
// EulerD<float> (Amplitude.h:557-573): R = Rx(a) Ry(b) Rz(c) evaluated in float
void euler_float(float a, float b, float c, float R[9]) {
    const float c1 = std::cos(a), s1 = std::sin(a), c2 = std::cos(b), s2 = std::sin(b), c3 = std::cos(c),
                s3 = std::sin(c);
    const float m[9] = {c2 * c3, -c2 * s3, s2, c1 * s3 + c3 * s1 * s2, c1 * c3 - s1 * s2 * s3, -c2 * s1,
                        s1 * s3 - c1 * c3 * s2, c3 * s1 + c1 * s2 * s3, c1 * c2};
    for (int i = 0; i < 9; i++) R[i] = m[i];
}

struct PdbPoint {
    float p[3];
};

void rigid(const float R[9], const float t[3], PdbPoint &x) {
    float o[3];
    for (int i = 0; i < 3; i++) o[i] = R[3 * i] * x.p[0] + R[3 * i + 1] * x.p[1] + R[3 * i + 2] * x.p[2] + t[i];
    for (int i = 0; i < 3; i++) x.p[i] = o[i];
}

// one ATOM record in the reference's fixed 80-byte layout (PDBAmplitude.cpp:1469-1538):
// the fields cut to their column widths, occupancy and temperature factor as
// the first 6 characters of "%6f" each, coordinates in Angstrom as the first 8
// characters of "%8f"; bytes past the formatted text stay zero, as there
std::string pdb_record(const PdbAtoms &A, size_t i, const PdbPoint &x) {
    char occ[40] = {0};
    snprintf(occ, sizeof occ, "%6f", A.occ[i]);
    snprintf(occ + 6, sizeof occ - 6, "%6f", A.bfac[i]);
    const std::string occ12(occ, 12);
    const std::string place = "-" + std::string(22, 'p') + "-";
    char buf[160] = {0};
    snprintf(buf, sizeof buf, "ATOM  %5s %4s %3s %c%4s    %s%s      %s%s", A.serial[i].substr(0, 5).c_str(),
             A.name[i].substr(0, 4).c_str(), A.res_name[i].substr(0, 3).c_str(), A.chain[i],
             A.res_no[i].substr(0, 4).c_str(), place.c_str(), occ12.c_str(), A.seg[i].substr(0, 4).c_str(),
             A.elem4[i].substr(0, 4).c_str());
    std::string line(buf, 80);
    for (int k = 0; k < 3; k++) {
        char t[64] = {0};
        snprintf(t, sizeof t, "%8f", (double)x.p[k] * 10.);
        for (int c = 0; c < 8; c++) line[30 + 8 * k + c] = t[c];
    }
    return line;
}

std::string amp_header(const ModelNode &n, const Grid &g) {
    std::ostringstream o;
    o << "# Model: " << (is_pdb(n) ? "PDB" : n.type == "AMP" ? "AMP" : symmetry_kind(n))
      << (n.name.empty() ? "" : " \"" + n.name + "\"") << "\n";
    if (!n.filename.empty()) o << "# File: " << n.filename << "\n";
    o << "# Grid size: " << g.grid_size << ", qmax: " << json::num(g.qmax) << " nm^-1, step " << json::num(g.step) << "\n";
    o << "# Precision: " << (g.fp64() ? "FP64" : "mixed FP32/FP64") << "\n";
    return o.str();
}

}  // namespace

// ------------------------------------------------------------------ backend object
struct Backend {
    std::mutex mu, start_mu;
    std::thread worker;
    std::atomic<bool> stop_flag{false};
    bool running = false;
    double progress = 0.0;
    int code = 0;
    std::string message;
    bool have_results = false, have_results_2d = false;
    std::vector<double> graph;
    std::vector<double> graph2d;  // nq x nq, row = q_z index, column = q_perp index
    std::vector<unsigned> model_ptrs;
    std::vector<std::pair<unsigned, std::string>> headers;
    std::map<unsigned, Grid *> model_grid;  // grids of this job's models (get/save amplitude)
    std::string cache_dir;
    Spec job_spec;  // model tree of the last started job (GetPDB)
    bool have_job = false;
    // device-resident caches across jobs
    std::map<std::string, PdbAtoms> pdb_cache;
    std::map<std::string, std::unique_ptr<Structure>> structures;
    std::map<std::string, std::unique_ptr<Grid>> leaf_grids;
    std::map<std::string, std::pair<std::unique_ptr<Grid>, std::string>> amp_grids;
    std::vector<std::unique_ptr<Grid>> job_grids;

    ~Backend() {
        stop_flag = true;
        if (worker.joinable()) worker.join();
    }

    Grid *leaf_grid(const ModelNode &n, const Spec &s, hipStream_t st) {
        if (n.filename.empty()) throw Error(kErrFileNotFound, "PDB model without a file name");
        std::vector<double> p = n.extra;
        const double defaults[10] = {1, 0, 1, 0.05, 0.14, 0.28, 0, 0, 0, 4};
        while (p.size() < 10) p.push_back(defaults[p.size()]);
        const std::string fk = file_key(n.filename);
        std::ostringstream key;
        key << fk << "|" << n.centered;
        for (double v : p) key << "|" << json::num(v);
        const std::string skey = key.str();
        const std::string gkey = skey + "|" + std::to_string(s.grid_size) + "|" + json::num(s.qmax) + "|" +
                                 std::to_string(s.precision);
        auto git = leaf_grids.find(gkey);
        if (git != leaf_grids.end()) return git->second.get();
        auto sit = structures.find(skey);
        if (sit == structures.end()) {
            auto pit = pdb_cache.find(fk);
            if (pit == pdb_cache.end()) pit = pdb_cache.emplace(fk, read_pdb(n.filename)).first;
            const PdbAtoms &A = pit->second;
            const int natoms = (int)A.species.size();
            const std::vector<float> xf = centred_coords(A, n.centered);
            std::vector<double> xyz(xf.begin(), xf.end());
            // ion table: elements present, in table order (sorted ion indices)
            std::vector<int> row(kNumElements, -1);
            std::vector<double> coef9, rad;
            for (int e = 0; e < kNumElements; e++)
                for (int a = 0; a < natoms; a++)
                    if (A.species[a] == e) {
                        row[e] = (int)rad.size();
                        coef9.insert(coef9.end(), kElements[e].coef9, kElements[e].coef9 + 9);
                        rad.push_back(kElements[e].svg_radius);
                        break;
                    }
            // ExtraParameters (PDBAmplitude.cpp:1205-1214)
            const double scale = p[0], sol_ed = p[1], c1 = p[2], vox = p[3], probe = p[4], thick = p[5], out_ed = p[6];
            const bool solvent_only = std::fabs(p[8] - 1.0) < 1e-6;
            const int rad_type = (int)(p[9] + 0.1);
            const bool fill_holes = std::fabs(p[7] - 1.0) < 1e-6;
            // radii of the voxel marking by radius type (PDBReaderOb::SetRadiusType,
            // PDBReaderLib.cpp:1207-1236): 1 van der Waals (also the default), 2
            // empirical, 3 calculated, 4 and 5 the dummy-atom radii
            std::vector<int> ion(natoms);
            std::vector<double> arad(natoms);
            for (int a = 0; a < natoms; a++) {
                const ElementInfo &E = kElements[A.species[a]];
                ion[a] = row[A.species[a]];
                arad[a] = rad_type == 4 || rad_type == 5 ? E.svg_radius
                          : rad_type == 2                ? E.emp_radius
                          : rad_type == 3                ? E.calc_radius
                                                         : E.vdw_radius;
            }
            // the voxelised solvent of PDBAmplitude::PreCalculate (PDBAmplitude.cpp:1250-1263):
            // the excluded volume for radius types 1, 2, 3, 5 with a solvent density,
            // the outer-solvent layer for types 1..5 with an outer-solvent density
            const bool voxel_types = rad_type == 1 || rad_type == 2 || rad_type == 3 || rad_type == 5;
            const bool want_excl = sol_ed != 0.0 && voxel_types && vox > 0.0;
            const bool want_layer = out_ed != 0.0 && (voxel_types || rad_type == 4) && vox > 0.0;
            dplus_pdb_desc d{};
            d.natoms = natoms;
            d.xyz = xyz.data();
            d.ion = ion.data();
            d.nion = (int)rad.size();
            d.coef9 = coef9.data();
            d.rad = rad.data();
            d.sol_ed = sol_ed;
            d.c1 = c1;
            d.flags = (solvent_only ? 0 : DPLUS_CALC_ATOMIC_FORMFACTORS) |
                      ((rad_type == 4 && sol_ed != 0.0) ? DPLUS_CALC_DUMMY_SOLVENT : 0);
            d.scale = scale;
            std::vector<int> bdims, edims;
            std::vector<double> bcom, ecom;
            if (want_excl || want_layer) {
                solvent_boxes(natoms, xyz.data(), arad.data(), probe, thick, vox, fill_holes, want_layer, want_excl,
                              bdims, bcom, edims, ecom);
                d.vox_step = vox;
                d.nbox = (int)(bdims.size() / 3);
                d.box_dims = bdims.data();
                d.box_com = bcom.data();
                d.box_contrast = out_ed - sol_ed;
                d.nexcl = (int)(edims.size() / 3);
                d.excl_dims = edims.data();
                d.excl_com = ecom.data();
                d.excl_contrast = -sol_ed;
            }
            sit = structures.emplace(skey, std::unique_ptr<Structure>(new Structure(&d, st))).first;
        }
        std::unique_ptr<Grid> g(new Grid(s.grid_size, s.qmax, s.precision));
        sit->second->build(*g, false, st);
        return leaf_grids.emplace(gkey, std::move(g)).first->second.get();
    }

    // AmpGridAmplitude (Amplitude.cpp:3680-3796): a grid read from a .amp/.ampj file
    Grid *amp_leaf(const ModelNode &n, const Spec &s, std::string &header, hipStream_t st) {
        if (n.filename.empty()) throw Error(kErrFileNotFound, "AMP model without a file name");
        const std::string key = file_key(n.filename) + "|" + std::to_string(s.precision);
        auto it = amp_grids.find(key);
        if (it == amp_grids.end()) {
            std::string h;
            std::unique_ptr<Grid> g(amp_read(n.filename, s.precision, h, st));
            if (g->qmax < s.qmax * (1.0 - 1e-12))
                throw Error(kErrInvalidArgs, "amplitude file " + n.filename + " reaches qmax " + json::num(g->qmax) +
                                                 ", below the requested " + json::num(s.qmax));
            it = amp_grids.emplace(key, std::make_pair(std::move(g), h)).first;
        }
        header = it->second.second;
        return it->second.first.get();
    }

    void add_header(unsigned ptr, const std::string &h) {
        for (auto &x : headers)
            if (x.first == ptr) return;
        headers.emplace_back(ptr, h);
    }

    bool gridded(const ModelNode &n, const Spec &s) const {
        if (!s.use_grid || !n.use_grid) return false;
        for (auto &c : n.children)
            if (!gridded(c, s)) return false;
        return true;
    }

    // placements of node n under (R, t, scale): e^{iq.t} A_leaf(R^T q) * scale.
    // Composition follows Flatten (Amplitude.cpp:577-671): M = parent * T(loc) * R(loc),
    // then M * T(sub) * R(sub) for each sub-location of a symmetry.
    void collect(const ModelNode &n, const double R[9], const double t[3], double scale, const Spec &s,
                 std::vector<Placement> &out, hipStream_t st) {
        if (stop_flag) throw Error(kErrStopped, "stopped");
        double Rl[9], Rn[9], tn[3], tl[3] = {n.loc[0], n.loc[1], n.loc[2]};
        rot_matrix(n.loc[3], n.loc[4], n.loc[5], Rl);  // Location angles in radians (Common.h:1416-1418)
        matmul(R, Rl, Rn);
        matvec(R, tl, tn);
        for (int k = 0; k < 3; k++) tn[k] += t[k];
        if (is_pdb(n)) {
            Grid *g = leaf_grid(n, s, st);
            model_grid[n.ptr] = g;
            add_header(n.ptr, amp_header(n, *g));
            out.push_back(make_placement(g, Rn, tn, scale));
            return;
        }
        if (is_amp(n)) {
            std::string fh;
            Grid *g = amp_leaf(n, s, fh, st);
            model_grid[n.ptr] = g;
            add_header(n.ptr, fh.empty() ? amp_header(n, *g) : fh);
            out.push_back(make_placement(g, Rn, tn, scale * (n.extra.empty() ? 1.0 : n.extra[0])));
            return;
        }
        if (!is_symmetry(n))
            throw Error(kErrInvalidModelType, "model type \"" + n.type + "\" is outside the scattering hot path");
        if (gridded(n, s)) {
            Grid *g = docked_grid(n, s, st);
            model_grid[n.ptr] = g;
            add_header(n.ptr, amp_header(n, *g));
            out.push_back(make_placement(g, Rn, tn, scale));
            return;
        }
        // hybrid: the symmetry is expanded into its sub-locations (Amplitude.cpp:594-631)
        add_header(n.ptr, "# Model: " + std::string(symmetry_kind(n)) + " (docked at orientation-average time)\n");
        const double sym_scale = n.extra.empty() ? 1.0 : n.extra[0];
        for (const SubLoc &L : sub_locations(n)) {
            double Rc[9], tc[3];
            matmul(Rn, L.R, Rc);
            matvec(Rn, L.t, tc);
            for (int k = 0; k < 3; k++) tc[k] += tn[k];
            for (auto &c : n.children) collect(c, Rc, tc, scale * sym_scale, s, out, st);
        }
    }

    // ManualSymmetry::calculateGrid (Symmetries.cpp:879-962) and
    // GridSymmetry::calculateGrid (Symmetries.cpp:348-431): the children docked
    // into a grid of the domain's size in the symmetry's own frame
    Grid *docked_grid(const ModelNode &n, const Spec &s, hipStream_t st) {
        std::vector<Placement> inner;
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, z[3] = {0, 0, 0};
        const double sym_scale = n.extra.empty() ? 1.0 : n.extra[0];
        LatticeDev lat{};
        if (is_space_filling(n)) {
            lat = lattice_of(n);
            for (auto &c : n.children) collect(c, I, z, sym_scale, s, inner, st);
        } else {
            for (const SubLoc &L : sub_locations(n))
                for (auto &c : n.children) collect(c, L.R, L.t, sym_scale, s, inner, st);
        }
        if (inner.empty()) throw Error(kErrInvalidParamTree, "a symmetry without children or sub-locations");
        DockSet D(inner, st);
        std::unique_ptr<Grid> g(new Grid(s.grid_size, s.qmax, s.precision));
        dock_grid(D, *g, 1.0, st, lat.on ? &lat : nullptr);
        job_grids.push_back(std::move(g));
        return job_grids.back().get();
    }

    static Placement make_placement(Grid *g, const double R[9], const double t[3], double scale) {
        Placement p;
        p.grid = g;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) p.rt[3 * i + j] = R[3 * j + i];
        for (int k = 0; k < 3; k++) p.t[k] = t[k];
        p.scale = scale;
        return p;
    }

    // DomainModel::CalculateIntensity2DMatrix through CompositeModel::CalculateMatrix
    // (Amplitude.cpp:843-993, CompositeModel.cpp:265-317): for every pair (q_z =
    // x[i], q_perp = x[j]) with q = |(q_z, q_perp)| in [qMin, qMax], the azimuthal
    // average of |A(q, theta = |atan2(q_perp, q_z)|, phi)|^2 (Grid.cpp:609-657);
    // the matrix is scale * sum_pop size * I_pop + constant (pairs outside the q
    // range are 0 before the constant).  Every orientation method takes the
    // azimuthal average (the reference's CPU Gauss-Kronrod branch ignores theta,
    // Amplitude.cpp:1484-1486); nphi = OrientationIterations within [64, 16384].
    std::vector<double> run_2d(const Spec &s, hipStream_t st) {
        const size_t nq = s.q.size();
        std::vector<double> q, th;
        std::vector<size_t> at;
        for (size_t i = 0; i < nq; i++)
            for (size_t j = 0; j < nq; j++) {
                const double qz = s.q[i], qp = s.q[j];
                const double qq = std::sqrt(qz * qz + qp * qp), t = std::fabs(std::atan2(qp, qz));
                if (qq > s.qmax || qq < s.qmin) continue;
                q.push_back(qq);
                th.push_back(std::min(t, kPi));
                at.push_back(i * nq + j);
            }
        const int nphi = (int)std::min<long long>(16384, std::max<long long>(64, s.iters));
        std::vector<double> total(nq * nq, 0.0);
        const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, z[3] = {0, 0, 0};
        for (size_t pi = 0; pi < s.pops.size(); pi++) {
            std::vector<Placement> pl;
            for (auto &m : s.pops[pi].models) collect(m, I3, z, 1.0, s, pl, st);
            {
                std::lock_guard<std::mutex> lk(mu);
                progress = (pi + 0.5) / s.pops.size();
            }
            if (stop_flag) throw Error(kErrStopped, "stopped");
            DockSet D(pl, st);
            std::vector<double> I(q.size());
            azimuthal_average(D, (int)q.size(), q.data(), th.data(), nphi, s.seed, I.data(), st);
            timing_collect();
            for (size_t k = 0; k < at.size(); k++) total[at[k]] += s.pops[pi].size * I[k];
        }
        for (auto &v : total) v = s.scale * v + s.constant;
        return total;
    }

    void run(Spec s) {
        hipStream_t st = Device::get().stream;
        try {
            Device::get().require();
            if (s.two_d) {
                std::vector<double> m = run_2d(s, st);
                std::lock_guard<std::mutex> lk(mu);
                graph2d.swap(m);
                have_results_2d = true;
                code = 0;
                message = "";
                progress = 1.0;
                running = false;
                return;
            }
            std::vector<double> total(s.q.size(), 0.0);
            const double I3[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, z[3] = {0, 0, 0};
            for (size_t pi = 0; pi < s.pops.size(); pi++) {
                std::vector<Placement> pl;
                for (auto &m : s.pops[pi].models) collect(m, I3, z, 1.0, s, pl, st);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    progress = (pi + 0.5) / s.pops.size();
                }
                DockSet D(pl, st);
                dplus_oa_params p{};
                p.method = s.method;
                p.seed = s.seed;
                p.max_dirs = s.iters;
                const bool single = pl.size() == 1 && pl[0].t[0] == 0 && pl[0].t[1] == 0 && pl[0].t[2] == 0;
                p.chunk = single ? 8192 : 4096;  // OAJacob.cu:147 / GPUHybridCalc.cu:204
                p.eps = s.eps;
                p.vegas_iterations = (int)std::max<long long>(1, (s.iters + 4095) / 4096);
                p.vegas_points_per_bin = 4;  // 32x32 bins x 4 points (GPUHybridCalc.cu:607-618)
                p.gk_max_depth = s.iters;
                std::vector<double> I(s.q.size());
                if (stop_flag) throw Error(kErrStopped, "stopped");
                orientation_average(D, (int)s.q.size(), s.q.data(), &p, I.data(), nullptr, st);
                timing_collect();
                for (size_t k = 0; k < I.size(); k++) total[k] += s.pops[pi].size * I[k];
            }
            for (auto &v : total) v = s.scale * v + s.constant;
            std::lock_guard<std::mutex> lk(mu);
            graph.swap(total);
            have_results = true;
            code = 0;
            message = "";
            progress = 1.0;
        } catch (const Error &e) {
            std::lock_guard<std::mutex> lk(mu);
            code = e.code;
            message = e.what();
        } catch (const std::exception &e) {
            std::lock_guard<std::mutex> lk(mu);
            code = kErrGeneral;
            message = e.what();
        }
        std::lock_guard<std::mutex> lk(mu);
        running = false;
    }

    void start(const std::string &state, bool two_d = false) {
        // PythonBackendWrapper::StartGenerate: ERROR_ILLEGAL_JSON first, then
        // CheckCapabilities (BackendWrapper.cpp:317), then the parameter tree
        const json::Value doc = json::parse(state);
        Device::get().require();
        Spec s = parse_state(doc);
        s.two_d = two_d;
        // one start at a time: the running check, the join of the finished
        // worker and the assignment of the new one happen under start_mu, so a
        // concurrent start sees running == true and fails with ERROR_JOBRUNNING
        std::lock_guard<std::mutex> start_lk(start_mu);
        {
            std::lock_guard<std::mutex> lk(mu);
            if (running) throw Error(kErrJobRunning, "a job is already running");  // JobManager.cpp:150-153
        }
        if (worker.joinable()) worker.join();
        std::lock_guard<std::mutex> lk(mu);
        running = true;
        progress = 0.0;
        code = -1;  // running (JobManager.cpp:164-165)
        message = "";
        have_results = false;
        have_results_2d = false;
        stop_flag = false;
        model_ptrs = s.ptrs;
        job_spec = s;
        have_job = true;
        headers.clear();
        model_grid.clear();
        job_grids.clear();
        worker = std::thread([this, s]() { run(s); });
    }

    std::string status() {
        std::lock_guard<std::mutex> lk(mu);
        return "{\"isRunning\":" + std::string(running ? "true" : "false") + ",\"progress\":" + json::num(progress) +
               ",\"code\":" + std::to_string(code) + ",\"message\":" + json::quote(message) + "}";
    }

    std::string results() {
        std::lock_guard<std::mutex> lk(mu);
        if (running) throw Error(kErrJobRunning, "the job is still running");
        if (!have_results) throw Error(code > 0 ? code : kErrInvalidArgs, message.empty() ? "no results" : message);
        std::string o = "{\"Headers\":[";
        bool first = true;
        for (auto &h : headers) {
            if (!first) o += ",";
            first = false;
            o += "{\"ModelPtr\":" + std::to_string(h.first) + ",\"Header\":" + json::quote(h.second) + "}";
        }
        o += "],\"Graph\":[";
        for (size_t i = 0; i < graph.size(); i++) o += (i ? "," : "") + json::num(graph[i]);
        return o + "]}";
    }

    // GetGenerate2DResults (BackendWrapper.cpp:583-624, Write2DGraph :514-536):
    // {"Headers": [...], "2DGraph": [[row q_z = x[0]], [row q_z = x[1]], ...]}
    std::string results_2d() {
        std::lock_guard<std::mutex> lk(mu);
        if (running) throw Error(kErrJobRunning, "the job is still running");
        if (!have_results_2d)
            throw Error(code > 0 ? code : kErrInvalidArgs, message.empty() ? "no 2-D results" : message);
        std::string o = "{\"Headers\":[";
        bool first = true;
        for (auto &h : headers) {
            if (!first) o += ",";
            first = false;
            o += "{\"ModelPtr\":" + std::to_string(h.first) + ",\"Header\":" + json::quote(h.second) + "}";
        }
        o += "],\"2DGraph\":[";
        const size_t n = (size_t)std::llround(std::sqrt((double)graph2d.size()));
        for (size_t i = 0; i < n; i++) {
            o += i ? ",[" : "[";
            for (size_t j = 0; j < n; j++) o += (j ? "," : "") + json::num(graph2d[i * n + j]);
            o += "]";
        }
        return o + "]}";
    }

    // ---- GetPDB (LocalBackend::HandleGetPDB, LocalBackend.cpp:969-1015)
    struct AtomAt {
        const PdbAtoms *A;
        size_t i;
        PdbPoint x;
    };

    static const ModelNode *find_node(const std::vector<ModelNode> &ms, unsigned ptr) {
        for (auto &m : ms) {
            if (m.ptr == ptr) return &m;
            if (const ModelNode *c = find_node(m.children, ptr)) return c;
        }
        return nullptr;
    }

    const PdbAtoms &atoms_of(const ModelNode &n) {
        const std::string fk = file_key(n.filename);
        auto it = pdb_cache.find(fk);
        if (it == pdb_cache.end()) it = pdb_cache.emplace(fk, read_pdb(n.filename)).first;
        return it->second;
    }

    // PDBAmplitude::AssemblePDBFile (PDBAmplitude.cpp:1504-1538) and
    // Symmetry::AssemblePDBFile (Symmetry.cpp:278-351): atoms in the model's own
    // frame (rotation R(alpha, beta, gamma) in float, then its translation); a
    // symmetry repeats its children at every sub-location, then applies its own
    // frame.  Amplitude-file leaves hold no atoms.
    void assemble(const ModelNode &n, std::vector<AtomAt> &out) {
        float R[9], t[3] = {(float)n.loc[0], (float)n.loc[1], (float)n.loc[2]};
        euler_float((float)n.loc[3], (float)n.loc[4], (float)n.loc[5], R);
        if (is_pdb(n)) {
            const PdbAtoms &A = atoms_of(n);
            const std::vector<float> xf = centred_coords(A, n.centered);
            for (size_t i = 0; i < A.species.size(); i++) {
                AtomAt a{&A, i, {{xf[3 * i], xf[3 * i + 1], xf[3 * i + 2]}}};
                rigid(R, t, a.x);
                out.push_back(a);
            }
            return;
        }
        if (!is_symmetry(n)) return;
        const std::vector<SubLoc> subs = sub_locations(n);
        for (size_t k = 0; k < subs.size(); k++) {
            float Rs[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            const float ts[3] = {(float)subs[k].t[0], (float)subs[k].t[1], (float)subs[k].t[2]};
            if (!is_space_filling(n))
                euler_float((float)(n.rows[k][3] * kPi / 180.0), (float)(n.rows[k][4] * kPi / 180.0),
                            (float)(n.rows[k][5] * kPi / 180.0), Rs);
            const size_t first = out.size();
            for (auto &c : n.children) assemble(c, out);
            for (size_t a = first; a < out.size(); a++) {
                rigid(Rs, ts, out[a].x);
                rigid(R, t, out[a].x);
            }
        }
    }

    std::string pdb_text(unsigned ptr) {
        std::lock_guard<std::mutex> lk(mu);
        if (!have_job) throw Error(kErrJobNotFound, "no job has been started");
        if (running) throw Error(kErrJobRunning, "the job is still running");
        std::vector<ModelNode> roots;
        for (auto &p : job_spec.pops)
            for (auto &m : p.models) roots.push_back(m);
        const ModelNode *n = find_node(roots, ptr);
        if (!n || !(is_pdb(*n) || is_symmetry(*n)))
            throw Error(kErrModelNotFound, "model " + std::to_string(ptr) + " has no atoms in this job");
        std::string out;
        if (is_symmetry(*n)) {
            // Symmetry::SavePDBFile (Symmetry.cpp:234-276): the model header as REMARK lines
            std::string st = "REMARK This is the header of the D+ state that was used to create the PDB below.\n";
            for (auto &h : headers)
                if (h.first == ptr) st += h.second;
            std::string r;
            for (char c : st) r += c == '\n' ? std::string("\nREMARK ") : std::string(1, c);
            out = r + "\n";
        }
        std::vector<AtomAt> atoms;
        assemble(*n, atoms);
        out.reserve(out.size() + atoms.size() * 81);
        for (auto &a : atoms) out += pdb_record(*a.A, a.i, a.x) + "\n";
        return out;
    }

    Grid *grid_of(unsigned ptr) {
        std::lock_guard<std::mutex> lk(mu);
        if (running) throw Error(kErrJobRunning, "the job is still running");
        auto it = model_grid.find(ptr);
        if (it == model_grid.end()) throw Error(kErrModelNotFound, "no amplitude grid for model " + std::to_string(ptr));
        return it->second;
    }
};

}  // namespace dplus

using namespace dplus;

static char *dup_string(const std::string &s) {
    char *p = (char *)malloc(s.size() + 1);
    if (p) memcpy(p, s.c_str(), s.size() + 1);
    return p;
}

#define BACKEND_GUARD(...)                                  \
    try {                                                   \
        __VA_ARGS__;                                        \
        return kOk;                                         \
    } catch (const Error &e) {                              \
        return set_error(e.code, e.what());                 \
    } catch (const std::exception &e) {                     \
        return set_error(kErrGeneral, e.what());            \
    } catch (...) {                                         \
        return set_error(kErrGeneral, "unknown exception"); \
    }

extern "C" {

void dplus_free(void *p) { free(p); }

int dplus_backend_create(dplus_backend **out) {
    BACKEND_GUARD({
        if (!out) throw Error(kErrInvalidArgs, "null output");
        *out = (dplus_backend *)new Backend();
    })
}

void dplus_backend_destroy(dplus_backend *b) { delete (Backend *)b; }

int dplus_backend_check_capabilities(dplus_backend *b, int check_tdr) {
    (void)check_tdr;
    BACKEND_GUARD({
        if (!b) throw Error(kErrInvalidArgs, "null backend");
        Device::get().require();
    })
}

int dplus_backend_get_all_model_metadata(dplus_backend *b, char **out) {
    BACKEND_GUARD({
        if (!b || !out) throw Error(kErrInvalidArgs, "null argument");
        *out = dup_string(
            "{\"Models\":[{\"Type\":\"PDB\",\"Name\":\"PDB\",\"ExtraParameters\":[\"Scale\",\"Solvent ED\",\"C1\","
            "\"Solvent Voxel Size\",\"Solvent Probe Radius\",\"Solvation Thickness\",\"Outer Solvent ED\","
            "\"Fill Holes\",\"Solvent Only\",\"Solvent method\"]},"
            "{\"Type\":\"AMP\",\"Name\":\"Amplitude Grid (AMP)\",\"ExtraParameters\":[\"Scale\"]},"
            "{\"Type\":\",26\",\"Name\":\"Manual Symmetry\",\"LayerParameters\":[\"X\",\"Y\",\"Z\",\"Alpha\",\"Beta\","
            "\"Gamma\"],\"ExtraParameters\":[\"Scale\"]},"
            "{\"Type\":\",25\",\"Name\":\"Space-filling Symmetry\",\"LayerParameters\":[\"Distance\",\"Angle\","
            "\"Repetitions\"],\"ExtraParameters\":[\"Scale\"]},"
            "{\"Type\":\"Scripted Symmetry\",\"Name\":\"Scripted Symmetry\",\"LayerParameters\":[\"Parameter\"],"
            "\"ExtraParameters\":[\"Scale\"]}],"
            "\"OrientationMethods\":[\"Monte Carlo (Mersenne Twister)\",\"Adaptive (VEGAS) Monte Carlo\","
            "\"Adaptive Gauss Kronrod\"]}");
    })
}

int dplus_backend_initialize_cache(dplus_backend *b, const char *dir) {
    BACKEND_GUARD({
        if (!b) throw Error(kErrInvalidArgs, "null backend");
        ((Backend *)b)->cache_dir = dir ? dir : "";
    })
}

int dplus_backend_start_generate(dplus_backend *b, const char *state, int use_gpu) {
    (void)use_gpu;  // the library is the GPU path
    BACKEND_GUARD({
        if (!b || !state) throw Error(kErrInvalidArgs, "null argument");
        ((Backend *)b)->start(state);
    })
}

int dplus_backend_start_generate_2d(dplus_backend *b, const char *state, int use_gpu) {
    (void)use_gpu;
    BACKEND_GUARD({
        if (!b || !state) throw Error(kErrInvalidArgs, "null argument");
        ((Backend *)b)->start(state, true);
    })
}

int dplus_backend_get_generate_2d_results(dplus_backend *b, char **out) {
    BACKEND_GUARD({
        if (!b || !out) throw Error(kErrInvalidArgs, "null argument");
        *out = dup_string(((Backend *)b)->results_2d());
    })
}

int dplus_backend_get_job_status(dplus_backend *b, char **out) {
    BACKEND_GUARD({
        if (!b || !out) throw Error(kErrInvalidArgs, "null argument");
        *out = dup_string(((Backend *)b)->status());
    })
}

int dplus_backend_get_generate_results(dplus_backend *b, char **out) {
    BACKEND_GUARD({
        if (!b || !out) throw Error(kErrInvalidArgs, "null argument");
        *out = dup_string(((Backend *)b)->results());
    })
}

int dplus_backend_get_model_ptrs(dplus_backend *b, unsigned *ptrs, int capacity, int *n) {
    BACKEND_GUARD({
        if (!b || !n) throw Error(kErrInvalidArgs, "null argument");
        Backend *B = (Backend *)b;
        std::lock_guard<std::mutex> lk(B->mu);
        *n = (int)B->model_ptrs.size();
        for (int i = 0; i < *n && i < capacity; i++) ptrs[i] = B->model_ptrs[i];
    })
}

int dplus_backend_get_amplitude(dplus_backend *b, unsigned model_ptr, char **bytes, long long *len) {
    BACKEND_GUARD({
        if (!b || !bytes || !len) throw Error(kErrInvalidArgs, "null argument");
        Backend *B = (Backend *)b;
        Grid *g = B->grid_of(model_ptr);
        Device::get().require();
        std::string hdr;
        for (auto &h : B->headers)
            if (h.first == model_ptr) hdr = h.second;
        const std::string s = amp_legacy_bytes(*g, hdr, Device::get().stream);
        *bytes = (char *)malloc(s.size());
        if (!*bytes) throw Error(kErrOutOfMemory, "out of host memory");
        memcpy(*bytes, s.data(), s.size());
        *len = (long long)s.size();
    })
}

int dplus_backend_get_pdb(dplus_backend *b, unsigned model_ptr, char **text, long long *len) {
    BACKEND_GUARD({
        if (!b || !text || !len) throw Error(kErrInvalidArgs, "null argument");
        const std::string s = ((Backend *)b)->pdb_text(model_ptr);
        *text = (char *)malloc(s.size() + 1);
        if (!*text) throw Error(kErrOutOfMemory, "out of host memory");
        memcpy(*text, s.data(), s.size());
        (*text)[s.size()] = '\0';
        *len = (long long)s.size();
    })
}

int dplus_backend_save_amplitude(dplus_backend *b, unsigned model_ptr, const char *path) {
    BACKEND_GUARD({
        if (!b || !path) throw Error(kErrInvalidArgs, "null argument");
        Backend *B = (Backend *)b;
        Grid *g = B->grid_of(model_ptr);
        Device::get().require();
        std::string hdr;
        for (auto &h : B->headers)
            if (h.first == model_ptr) hdr = h.second;
        amp_write(*g, path, hdr, Device::get().stream);
    })
}

int dplus_backend_stop(dplus_backend *b) {
    BACKEND_GUARD({
        if (!b) throw Error(kErrInvalidArgs, "null backend");
        ((Backend *)b)->stop_flag = true;
    })
}

}  // extern "C"
