// export.cpp -- the output formats of a solve (manager/source/file_manager.cpp):
//
//   axs_save_solution_npy   saveSolutionNpy (file_manager.cpp:2074-2309): the
//                           cell-centred fields as a NumPy table, the cell
//                           polygons, one table per transient frame, meta.json
//   axs_save_solver_file /  the .axislv solver payload, format version 5 (and 4 read)
//   axs_load_solver_file    (saveFromPathSolver / loadFromPathSolver,
//                           file_manager.cpp:1482-1549)
//
// Both are plain host code over the solver's host-side results; the byte layout
// of each file is the reference's, so files move between the two programs.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "axs.h"
#include "handles.hpp"
#include "solver.hpp"

namespace axs {
namespace {

// ---------------- NumPy ----------------
// An NPY 1.0 file: magic, version, little-endian header length, the header dict
// padded with spaces to a 64-byte boundary (newline last), then the row-major
// payload (writeNpyRaw, file_manager.cpp:1800-1858).
void writeNpy(const std::filesystem::path& path, const char* descr, const void* data, size_t elem, size_t rows,
              size_t cols) {
    std::string header = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (" +
                         std::to_string(rows) + ", " + std::to_string(cols) + "), }";
    const size_t unpadded = 10 + header.size() + 1;
    const size_t total = (unpadded + 63) / 64 * 64;
    header.append(total - unpadded, ' ');
    header.push_back('\n');
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("cannot open " + path.string());
    const size_t hl = header.size();
    const char pre[10] = {'\x93', 'N', 'U', 'M', 'P', 'Y', 1, 0, (char)(hl & 0xff), (char)((hl >> 8) & 0xff)};
    out.write(pre, 10);
    out.write(header.data(), (std::streamsize)hl);
    if (rows * cols) out.write((const char*)data, (std::streamsize)(rows * cols * elem));
    if (!out) throw std::runtime_error("failed while writing " + path.string());
}

// jsonNumber: 17 significant digits, null for a non-finite value
std::string jsonNumber(double v) {
    if (!std::isfinite(v)) return "null";
    char buf[40];
    std::snprintf(buf, sizeof(buf), "%.17g", v);
    return buf;
}

std::string jsonString(const std::string& s) {
    std::string out = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') out.push_back('\\');
        out.push_back(c);
    }
    return out + "\"";
}

// The fields of a solution in the results panel's order (addFieldType,
// solver.cpp:333-362); pressure gradients are solutions but not listed there.
std::vector<std::pair<std::string, int>> solutionColumns(const axs_config& c) {
    std::vector<std::pair<std::string, int>> cols = {{"Axial Velocity", AXS_FIELD_U},
                                                     {"Radial Velocity", AXS_FIELD_V},
                                                     {"Velocity Magnitude", AXS_FIELD_VMAG},
                                                     {"Pressure", AXS_FIELD_P}};
    if (c.solve_energy)
        for (auto e : {std::pair<std::string, int>{"Temperature", AXS_FIELD_T}, {"dT/dz", AXS_FIELD_GRAD_TZ},
                       {"dT/dr", AXS_FIELD_GRAD_TR}})
            cols.push_back(e);
    if (c.solve_concentration)
        for (auto e : {std::pair<std::string, int>{"Concentration", AXS_FIELD_C}, {"dC/dz", AXS_FIELD_GRAD_CZ},
                       {"dC/dr", AXS_FIELD_GRAD_CR}})
            cols.push_back(e);
    for (auto e : {std::pair<std::string, int>{"Continuity", AXS_FIELD_CONTINUITY},
                   {"Cell Reynolds Number", AXS_FIELD_CELL_RE}, {"dU/dz", AXS_FIELD_GRAD_UZ},
                   {"dU/dr", AXS_FIELD_GRAD_UR}, {"dV/dz", AXS_FIELD_GRAD_VZ}, {"dV/dr", AXS_FIELD_GRAD_VR}})
        cols.push_back(e);
    return cols;
}

// Quads -> one shared point list + 4 indices per cell: corners closer than
// 1e-9 of the geometry's extent (at least 1) are one point, found by rounding to
// that grid and probing the neighbouring cells of it (weldCellCorners,
// file_manager.cpp:1906-1973).
void weldQuads(const double* qz, const double* qr, size_t nCells, std::vector<double>& xy,
               std::vector<int32_t>& idx) {
    double extent = 1.0;
    for (size_t i = 0; i < 4 * nCells; i++) extent = std::max(extent, std::max(std::fabs(qz[i]), std::fabs(qr[i])));
    const double tol = extent * 1e-9;
    struct KeyHash {
        size_t operator()(const std::pair<long long, long long>& k) const {
            const size_t h = std::hash<long long>{}(k.first);
            return h ^ (std::hash<long long>{}(k.second) + 0x9e3779b97f4a7c15ULL + (h << 6) + (h >> 2));
        }
    };
    std::unordered_map<std::pair<long long, long long>, int32_t, KeyHash> seen;
    seen.reserve(nCells * 2);
    xy.clear();
    idx.clear();
    idx.reserve(nCells * 4);
    for (size_t i = 0; i < 4 * nCells; i++) {
        const long long kz = std::llround(qz[i] / tol), kr = std::llround(qr[i] / tol);
        int32_t id = -1;
        for (int dz = -1; dz <= 1 && id < 0; dz++)
            for (int dr = -1; dr <= 1 && id < 0; dr++) {
                auto it = seen.find({kz + dz, kr + dr});
                if (it != seen.end()) id = it->second;
            }
        if (id < 0) {
            id = (int32_t)(xy.size() / 2);
            xy.push_back(qz[i]);
            xy.push_back(qr[i]);
            seen.emplace(std::make_pair(kz, kr), id);
        }
        idx.push_back(id);
    }
}

void saveSolutionNpy(Solver& s, const axs_cell_polygons* poly, const std::filesystem::path& dir) {
    const HostMesh& m = s.hostMesh();
    const axs_config& c = s.config();
    const size_t N = (size_t)m.nCells;
    if (N == 0) throw std::logic_error("no cells to export: mesh and solve first");
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) throw std::runtime_error("cannot create " + dir.string() + ": " + ec.message());

    const std::vector<std::pair<std::string, int>> cols = solutionColumns(c);
    const size_t nGeom = 4, nCols = nGeom + cols.size();
    std::vector<double> table(N * nCols, 0.0), col(N);
    for (size_t k = 0; k < cols.size(); k++) {
        s.getField(cols[k].second, col.data(), N);
        for (size_t i = 0; i < N; i++) table[i * nCols + nGeom + k] = col[i];
    }
    for (size_t i = 0; i < N; i++) {
        double* row = table.data() + i * nCols;
        row[0] = m.cellCenterZ[i];
        row[1] = m.cellCenterR[i];
        row[2] = m.cellVolume[i];
        row[3] = m.cellArea2D[i];
    }
    writeNpy(dir / "solution.npy", "<f8", table.data(), 8, N, nCols);

    size_t nPoints = 0, corners = 0;
    if (poly && poly->corners == 4 && poly->quad_z && poly->quad_r) {
        std::vector<double> xy;
        std::vector<int32_t> idx;
        weldQuads(poly->quad_z, poly->quad_r, N, xy, idx);
        nPoints = xy.size() / 2;
        corners = 4;
        writeNpy(dir / "points.npy", "<f8", xy.data(), 8, nPoints, 2);
        writeNpy(dir / "cells.npy", "<i4", idx.data(), 4, N, 4);
    } else if (poly && poly->corners == 3 && poly->point_z && poly->point_r && poly->tris && poly->n_points > 0) {
        std::vector<double> xy((size_t)poly->n_points * 2);
        for (int i = 0; i < poly->n_points; i++) {
            xy[2 * (size_t)i] = poly->point_z[i];
            xy[2 * (size_t)i + 1] = poly->point_r[i];
        }
        std::vector<int32_t> idx(poly->tris, poly->tris + 3 * N);
        nPoints = (size_t)poly->n_points;
        corners = 3;
        writeNpy(dir / "points.npy", "<f8", xy.data(), 8, nPoints, 2);
        writeNpy(dir / "cells.npy", "<i4", idx.data(), 4, N, 3);
    }

    // frames: the field columns only, each frame's primary fields filled in
    std::vector<std::pair<std::string, double>> frameIndex;
    for (int i = 0; i < s.frameCount(); i++) {
        const Solver::Frame& fr = s.frame(i);
        std::vector<double> ft(N * cols.size(), 0.0);
        for (size_t k = 0; k < cols.size(); k++) {
            const int id = cols[k].second;
            const std::vector<double>* src = id == AXS_FIELD_U ? &fr.u : id == AXS_FIELD_V ? &fr.v
                                           : id == AXS_FIELD_P ? &fr.p : id == AXS_FIELD_T ? &fr.t
                                           : id == AXS_FIELD_C ? &fr.c : nullptr;
            if (!src || src->size() != N) continue;
            for (size_t j = 0; j < N; j++) ft[j * cols.size() + k] = (*src)[j];
        }
        char name[32];
        std::snprintf(name, sizeof(name), "frame_%04d.npy", i);
        writeNpy(dir / name, "<f8", ft.data(), 8, N, cols.size());
        frameIndex.emplace_back(name, fr.time);
    }

    std::ofstream meta(dir / "meta.json");
    if (!meta) throw std::runtime_error("cannot open " + (dir / "meta.json").string());
    const bool structured = poly ? poly->structured != 0 : m.quadZ.size() == 4 * N;
    const bool multiblock = poly ? poly->multiblock != 0 : structured;
    meta << "{\n";
    meta << "  \"app\": \"AxiSim\",\n";
    meta << "  \"format\": 1,\n";
    meta << "  \"note\": \"cell-centered values, base SI units\",\n";
    meta << "  \"cells\": " << N << ",\n";
    meta << "  \"meshType\": \"" << (structured ? "Structured" : "Unstructured") << "\",\n";
    meta << "  \"multiBlock\": " << (multiblock ? "true" : "false") << ",\n";
    meta << "  \"structured\": false,\n";  // no single nr x nz raster: cells are numbered per block
    if (corners) {
        meta << "  \"points\": " << nPoints << ",\n";
        meta << "  \"cellCorners\": " << corners << ",\n";
    }
    meta << "  \"fluid\": {\"rho\": " << jsonNumber(c.rho) << ", \"mu\": " << jsonNumber(c.mu)
         << ", \"cp\": " << jsonNumber(c.cp) << ", \"k\": " << jsonNumber(c.k) << ", \"D\": "
         << jsonNumber(c.diffusivity) << "},\n";
    meta << "  \"transient\": " << (c.transient ? "true" : "false") << ",\n";
    meta << "  \"dt\": " << jsonNumber(c.dt) << ",\n";
    meta << "  \"columns\": [\"z\", \"r\", \"volume\", \"area2D\"";
    for (const auto& e : cols) meta << ", " << jsonString(e.first);
    meta << "],\n";
    meta << "  \"frameColumns\": [";
    for (size_t k = 0; k < cols.size(); k++) meta << (k ? ", " : "") << jsonString(cols[k].first);
    meta << "],\n";
    meta << "  \"frames\": [";
    for (size_t i = 0; i < frameIndex.size(); i++)
        meta << (i ? ",\n    " : "\n    ") << "{\"file\": \"" << frameIndex[i].first
             << "\", \"time\": " << jsonNumber(frameIndex[i].second) << "}";
    meta << (frameIndex.empty() ? "]\n" : "\n  ]\n");
    meta << "}\n";
    if (!meta) throw std::runtime_error("failed while writing meta.json");
}

// ---------------- .axislv solver payload, version 5 ----------------
constexpr uint32_t kSolverMagic = 0x53585641u;  // "AXVS" little-endian
constexpr uint32_t kSolverVersion = 5u;
constexpr uint32_t kSolverVersionSplitConfig = 4u;  // still read (readSolverPayloadLegacy)

// Field-by-field little-endian byte image of the payload, each struct at its
// reference size with zeroed padding: ConfigSolver 40 bytes (five one-byte enums,
// linearMaxIter at 8, four flags at 12-15, dt 16, tEnd 24, saveKeyFrameIter 32),
// the fluid properties 40, ConfigSimple 24, VariableUnits 15, SolverFieldOption 2.
struct Bytes {
    std::vector<uint8_t> b;
    template <class T>
    void put(T v) {
        const size_t o = b.size();
        b.resize(o + sizeof(T));
        std::memcpy(b.data() + o, &v, sizeof(T));
    }
    void pad(size_t n) { b.resize(b.size() + n, 0); }
};
struct Reader {
    const std::vector<uint8_t>& b;
    size_t at = 0;
    bool ok = true;
    template <class T>
    T get() {
        T v{};
        if (at + sizeof(T) > b.size()) {
            ok = false;
            return v;
        }
        std::memcpy(&v, b.data() + at, sizeof(T));
        at += sizeof(T);
        return v;
    }
    void skip(size_t n) {
        if (at + n > b.size()) ok = false;
        at += n;
    }
};

// sanitizeSolverConfig, file_manager.cpp:235-301: counts and tolerances back to
// their defaults, enums to what the solver runs, the fluid reset as a whole
void sanitize(axs_config& c, axs_solver_file& x) {
    if (c.linear_max_iter < 1) c.linear_max_iter = 20;
    if (c.save_key_frame_iter < 1) c.save_key_frame_iter = 2;
    if (c.max_iter < 1) c.max_iter = 50;
    if (c.check_conv < 1) c.check_conv = 1;
    if (!std::isfinite(x.mom_tol) || x.mom_tol <= 0.0) x.mom_tol = 1e-8;
    if (!std::isfinite(x.pp_tol) || x.pp_tol <= 0.0) x.pp_tol = 1e-5;
    if (c.linear_solver < 0 || c.linear_solver > AXS_LINEAR_GS_RB) c.linear_solver = AXS_LINEAR_JACOBI;
    if (x.velocity_solver != 0) x.velocity_solver = 0;
    if (c.gradient_scheme < 0 || c.gradient_scheme > AXS_GRAD_LSQ) c.gradient_scheme = AXS_GRAD_LSQ;
    for (int i = 0; i < AXS_RES_COUNT; i++) {  // clampResidualSettings, :131-141
        if (x.residual_type[i] < 0 || x.residual_type[i] > 2) x.residual_type[i] = 1;
        if (c.residual[i].norm < 0 || c.residual[i].norm > AXS_NORM_LINF) c.residual[i].norm = AXS_NORM_LINF;
        if (c.residual[i].scaling < 0 || c.residual[i].scaling > AXS_SCALE_DIAGONAL) c.residual[i].scaling = AXS_SCALE_NONE;
    }
    if (c.convection_scheme < 0 || c.convection_scheme > AXS_CONV_SOU) c.convection_scheme = AXS_CONV_UPWIND;
    if (c.time_scheme < 0 || c.time_scheme > AXS_TIME_BDF2) c.time_scheme = AXS_TIME_BDF1;
    const bool badFluid = !std::isfinite(c.rho) || c.rho < 1.0e-12 || !std::isfinite(c.mu) || c.mu < 1.0e-12 ||
                          !std::isfinite(c.cp) || c.cp <= 0.0 || !std::isfinite(c.k) || c.k < 0.0 ||
                          !std::isfinite(c.diffusivity) || c.diffusivity < 0.0;
    if (badFluid) {  // FluidPropertyConfig defaults, src/header/setting.cuh:61-71
        c.rho = 998.0;
        c.mu = 0.0010518;
        c.cp = 4180.0;
        c.k = 0.6;
        c.diffusivity = 3.0277e-9;
    }
}

axs_solver_file defaultExtra() {
    axs_solver_file x{};
    x.velocity_solver = 0;
    x.mom_tol = 1e-8;  // ConfigSimple, solver_struct.h:158-163
    x.pp_tol = 1e-5;
    for (int i = 0; i < AXS_RES_COUNT; i++) x.residual_type[i] = 0;  // RESIDUAL_SCALED
    return x;
}

std::vector<uint8_t> encodeSolverFile(axs_config c, axs_solver_file x) {
    sanitize(c, x);
    Bytes o;
    o.put<uint32_t>(kSolverMagic);
    o.put<uint32_t>(kSolverVersion);
    for (int i = 0; i < 15; i++) o.put<uint8_t>(x.var_units[i]);
    o.put<uint8_t>(c.solve_energy ? 1 : 0);
    o.put<uint8_t>(c.solve_concentration ? 1 : 0);
    // ConfigSolver
    o.put<uint8_t>((uint8_t)x.velocity_solver);
    o.put<uint8_t>((uint8_t)c.convection_scheme);
    o.put<uint8_t>((uint8_t)c.gradient_scheme);
    o.put<uint8_t>((uint8_t)c.linear_solver);
    o.put<uint8_t>((uint8_t)c.time_scheme);
    o.pad(3);
    o.put<int32_t>(c.linear_max_iter);
    o.put<uint8_t>(c.use_multigrid ? 1 : 0);
    o.put<uint8_t>(c.add_convection ? 1 : 0);
    o.put<uint8_t>(c.transient ? 1 : 0);
    o.put<uint8_t>(c.use_non_orth ? 1 : 0);
    o.put<double>(c.dt);
    o.put<double>(c.t_end);
    o.put<int32_t>(c.save_key_frame_iter);
    o.pad(4);
    // FluidPropertyConfig
    for (double v : {c.rho, c.mu, c.cp, c.k, c.diffusivity}) o.put<double>(v);
    // ConfigSimple
    o.put<int32_t>(c.max_iter);
    o.put<int32_t>(c.check_conv);
    o.put<double>(x.mom_tol);
    o.put<double>(x.pp_tol);
    // writeResidualConfigs, file_manager.cpp:151-173
    for (int i = 0; i < AXS_RES_COUNT; i++) {
        o.put<uint8_t>((uint8_t)x.residual_type[i]);
        o.put<uint8_t>((uint8_t)c.residual[i].norm);
        o.put<uint8_t>((uint8_t)c.residual[i].scaling);
        o.put<uint8_t>(c.residual[i].enabled ? 1 : 0);
        o.put<double>(c.residual[i].tol);
    }
    return o.b;
}

// readSolverPayloadLegacy (file_manager.cpp:321-373): the version-4 payload,
// whose ConfigSolver was 32 bytes of int-width enums (linear solver, linear
// iterations, convection and transient flags, the time scheme, dt, tEnd) with
// the velocity solver, convection scheme and keyframe interval as loose int32s
// after it; ConfigSimple 32 bytes with the corrector as a trailing byte (a
// pass count in old files: any non-zero is on); the gradient scheme as an
// int32 after it; residual records with int32 enums.  The settings v4 did not
// hold take the reference's ConfigSolver / ConfigSimple defaults (the
// multigrid on).
void decodeSolverFileV4(Reader& r, axs_config& c, axs_solver_file& x) {
    for (int i = 0; i < 15; i++) x.var_units[i] = r.get<uint8_t>();
    c.solve_energy = r.get<uint8_t>() != 0;
    c.solve_concentration = r.get<uint8_t>() != 0;
    c.linear_solver = r.get<int32_t>();
    c.linear_max_iter = r.get<int32_t>();
    c.add_convection = r.get<uint8_t>() != 0;
    c.transient = r.get<uint8_t>() != 0;
    c.time_scheme = r.get<uint8_t>();
    r.skip(5);
    c.dt = r.get<double>();
    c.t_end = r.get<double>();
    x.velocity_solver = r.get<int32_t>();
    c.convection_scheme = r.get<int32_t>();
    c.save_key_frame_iter = r.get<int32_t>();
    c.rho = r.get<double>();
    c.mu = r.get<double>();
    c.cp = r.get<double>();
    c.k = r.get<double>();
    c.diffusivity = r.get<double>();
    c.max_iter = r.get<int32_t>();
    c.check_conv = r.get<int32_t>();
    x.mom_tol = r.get<double>();
    x.pp_tol = r.get<double>();
    c.use_non_orth = r.get<uint8_t>() != 0;
    r.skip(7);
    c.gradient_scheme = r.get<int32_t>();
    c.use_multigrid = 1;
    for (int i = 0; i < AXS_RES_COUNT; i++) {  // readResidualConfigsLegacy, :195-219
        x.residual_type[i] = r.get<int32_t>();
        c.residual[i].norm = r.get<int32_t>();
        c.residual[i].scaling = r.get<int32_t>();
        c.residual[i].enabled = r.get<uint8_t>() != 0;
        c.residual[i].tol = r.get<double>();
    }
}

// loadFromPathSolver (file_manager.cpp:1518-1546): readSolverPayload
// (file_manager.cpp:300-314) for version 5, readSolverPayloadLegacy for 4,
// into a copy of cfg; throws on a foreign, other-version or short file
void decodeSolverFile(const std::vector<uint8_t>& bytes, axs_config& cfg, axs_solver_file& x) {
    Reader r{bytes};
    const uint32_t magic = r.get<uint32_t>();
    const uint32_t version = r.get<uint32_t>();
    if (!r.ok) throw std::invalid_argument("solver file too short");
    if (magic != kSolverMagic) throw std::domain_error("not an AxiSim solver file (magic)");
    if (version == kSolverVersionSplitConfig) {
        axs_config c = cfg;
        decodeSolverFileV4(r, c, x);
        if (!r.ok) throw std::invalid_argument("solver file truncated");
        sanitize(c, x);
        cfg = c;
        return;
    }
    if (version != kSolverVersion)
        throw std::domain_error("solver file version " + std::to_string(version) + " (this reads versions 4 and 5)");
    axs_config c = cfg;
    for (int i = 0; i < 15; i++) x.var_units[i] = r.get<uint8_t>();
    c.solve_energy = r.get<uint8_t>() != 0;
    c.solve_concentration = r.get<uint8_t>() != 0;
    x.velocity_solver = r.get<uint8_t>();
    c.convection_scheme = r.get<uint8_t>();
    c.gradient_scheme = r.get<uint8_t>();
    c.linear_solver = r.get<uint8_t>();
    c.time_scheme = r.get<uint8_t>();
    r.skip(3);
    c.linear_max_iter = r.get<int32_t>();
    c.use_multigrid = r.get<uint8_t>() != 0;
    c.add_convection = r.get<uint8_t>() != 0;
    c.transient = r.get<uint8_t>() != 0;
    c.use_non_orth = r.get<uint8_t>() != 0;
    c.dt = r.get<double>();
    c.t_end = r.get<double>();
    c.save_key_frame_iter = r.get<int32_t>();
    r.skip(4);
    c.rho = r.get<double>();
    c.mu = r.get<double>();
    c.cp = r.get<double>();
    c.k = r.get<double>();
    c.diffusivity = r.get<double>();
    c.max_iter = r.get<int32_t>();
    c.check_conv = r.get<int32_t>();
    x.mom_tol = r.get<double>();
    x.pp_tol = r.get<double>();
    for (int i = 0; i < AXS_RES_COUNT; i++) {
        x.residual_type[i] = r.get<uint8_t>();
        c.residual[i].norm = r.get<uint8_t>();
        c.residual[i].scaling = r.get<uint8_t>();
        c.residual[i].enabled = r.get<uint8_t>() != 0;
        c.residual[i].tol = r.get<double>();
    }
    if (!r.ok) throw std::invalid_argument("solver file truncated");
    sanitize(c, x);
    cfg = c;
}

}  // namespace
}  // namespace axs

extern "C" {

axs_status axs_save_solution_npy(axs_solver* h, const axs_cell_polygons* poly, const char* dir) {
    if (!h || !dir) return AXS_ERR_INVALID;
    return axs::guarded(&h->s->lastError, [&] { axs::saveSolutionNpy(*h->s, poly, std::filesystem::path(dir)); });
}

axs_status axs_save_solver_file(const char* path, const axs_config* cfg, const axs_solver_file* extra) {
    if (!path || !cfg) return AXS_ERR_INVALID;
    return axs::guarded(nullptr, [&] {
        const std::vector<uint8_t> b = axs::encodeSolverFile(*cfg, extra ? *extra : axs::defaultExtra());
        std::ofstream out(path, std::ios::binary);
        if (!out) throw std::invalid_argument(std::string("cannot open ") + path);
        out.write((const char*)b.data(), (std::streamsize)b.size());
        if (!out) throw std::runtime_error(std::string("failed while writing ") + path);
    });
}

axs_status axs_load_solver_file(const char* path, axs_config* cfg, axs_solver_file* extra) {
    if (!path || !cfg) return AXS_ERR_INVALID;
    return axs::guarded(nullptr, [&] {
        std::ifstream in(path, std::ios::binary);
        if (!in) throw std::invalid_argument(std::string("cannot open ") + path);
        const std::vector<uint8_t> b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
        axs_solver_file x = axs::defaultExtra();
        axs::decodeSolverFile(b, *cfg, x);
        if (extra) *extra = x;
    });
}

}  // extern "C"
