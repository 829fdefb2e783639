// Result export of the host C++ layer: CBModelExporterVTK
// (mechanics/src/CBModel/CBModelExporterVTK.cpp:29-499) with the fields of
// CBSolver::Export (mechanics/src/CBSolver/CBSolver.cpp:538-1164) and the Newmark node
// fields (CBSolverNewmarkBeta.cpp:800-827).  The per-element state comes from the GPU
// (cmg_element_export); the derived fields follow the reference's formulas (see
// cardiomechanics_amd/export.py, which writes the same files):
//   F_loc = B0^T F B0, J = det F_loc, Lambda = column norms of F_loc, C = F_loc^T F_loc,
//   E = (C - I)/2, f/s/n = F_loc B0[:, k], S/E/C _ii = (00, 11, 22), _ij = (01, 02, 12).
#include <cmath>
#include <cstring>
#include <filesystem>
#include <fstream>

#include "cm_host.h"

namespace cmhost {

namespace {

std::string Base64(const unsigned char* data, size_t n) {
    static const char* tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve(4 * ((n + 2) / 3));
    size_t i = 0;
    for (; i + 2 < n; i += 3) {
        const unsigned v = (data[i] << 16) | (data[i + 1] << 8) | data[i + 2];
        out += tab[(v >> 18) & 63];
        out += tab[(v >> 12) & 63];
        out += tab[(v >> 6) & 63];
        out += tab[v & 63];
    }
    if (i < n) {
        unsigned v = data[i] << 16;
        if (i + 1 < n) v |= data[i + 1] << 8;
        out += tab[(v >> 18) & 63];
        out += tab[(v >> 12) & 63];
        out += (i + 1 < n) ? tab[(v >> 6) & 63] : '=';
        out += '=';
    }
    return out;
}

// a binary DataArray: base64 of (UInt64 byte count + little-endian values)
template <class T>
void DataArray(std::ostream& o, const char* type, const std::string& name, int ncomp, const std::vector<T>& v) {
    const uint64_t bytes = v.size() * sizeof(T);
    std::vector<unsigned char> raw(8 + bytes);
    std::memcpy(raw.data(), &bytes, 8);
    if (bytes) std::memcpy(raw.data() + 8, v.data(), bytes);
    o << "        <DataArray type=\"" << type << "\" Name=\"" << name << "\" NumberOfComponents=\"" << ncomp
      << "\" format=\"binary\">\n          " << Base64(raw.data(), raw.size()) << "\n        </DataArray>\n";
}

}  // namespace

bool CBModelExporterVTK::Option(const std::string& name) const { return options_.at(name); }

void CBModelExporterVTK::Init(const ParameterMap& p, const TetgenModel& m, cmg_ctx* ctx) {
    const std::string prefix = p.Get<std::string>("Export.Prefix", "");
    if (prefix.empty()) return;
    active_ = true;
    ctx_ = ctx;
    timeStep_ = p.Get<double>("Export.TimeStep", 1e-3);
    const size_t pos = prefix.find_last_of("/\\");
    dir_ = pos == std::string::npos ? std::string(".") : prefix.substr(0, pos);
    prefix_ = pos == std::string::npos ? prefix : prefix.substr(pos + 1);
    // Export.Options defaults of the reference
    const std::pair<const char*, bool> defaults[] = {
        {"Material", true}, {"Fixation", true}, {"Fiber", true}, {"Sheet", true}, {"SheetNormal", true},
        {"Lambda", true}, {"Jacobian", false}, {"Deformation", false}, {"PK2Stress", true},
        {"GreenLagrangeStrain", true}, {"Velocity", true}, {"Acceleration", true}, {"Displacement", true},
        {"AbsDisplacement", true}};
    for (const auto& d : defaults) options_[d.first] = p.Get<bool>(std::string("Export.Options.") + d.first, d.second);
    std::error_code ec;
    std::filesystem::create_directories(dir_ + "/" + prefix_ + "_vtu", ec);
    if (ec) throw std::runtime_error("cannot create " + dir_ + "/" + prefix_ + "_vtu: " + ec.message());
    std::ofstream pvd(dir_ + "/" + prefix_ + ".pvd");
    if (!pvd) throw std::runtime_error("cannot create " + dir_ + "/" + prefix_ + ".pvd");
    pvd << "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n\t<Collection>\n"
        << "\t</Collection>\n</VTKFile>\n";
    nNodes_ = m.nNodes;
    nSolid_ = m.nElements;
    refCoords_ = m.coords;
    const int nb = m.npe == 10 ? 5 : 1;
    basis0_.resize(9 * nSolid_);
    for (int64_t e = 0; e < nSolid_; e++)
        for (int r = 0; r < 9; r++) basis0_[9 * e + r] = m.bases[(e * nb) * 9 + r];
    material_ = m.material;
    fixation_.assign(m.bc.begin(), m.bc.end());
    // cells: solid elements with their own node order, then the surface elements
    for (int64_t e = 0; e < nSolid_; e++) {
        for (int k = 0; k < m.npe; k++) connectivity_.push_back(m.elements[e * m.npe + k]);
        offsets_.push_back((int64_t)connectivity_.size());
        types_.push_back(m.npe == 10 ? 24 : 10);
    }
    const int64_t S = (int64_t)m.surfaceIndex.size();
    for (int64_t s = 0; s < S; s++) {
        for (int k = 0; k < m.nps; k++) connectivity_.push_back(m.surfaces[s * m.nps + k]);
        offsets_.push_back((int64_t)connectivity_.size());
        types_.push_back(m.nps == 6 ? 22 : 5);
    }
}

bool CBModelExporterVTK::Due(double time) const {
    return active_ && (!exported_ || time - lastTime_ >= (1 - 1e-12) * timeStep_);
}

void CBModelExporterVTK::Export(double time) {
    if (!active_) return;
    const int64_t n3 = 3 * nNodes_, E = nSolid_, ncell = (int64_t)types_.size();
    std::vector<double> nodes(n3), vel(n3), acc(n3), inc(n3);
    if (cmg_get_state(ctx_, nodes.data(), vel.data(), acc.data(), inc.data()) != CMG_SUCCESS)
        throw std::runtime_error(std::string("cmg_get_state: ") + cmg_last_error());
    std::vector<double> st(18 * E);
    if (cmg_element_export(ctx_, nullptr, st.data()) != CMG_SUCCESS)
        throw std::runtime_error(std::string("cmg_element_export: ") + cmg_last_error());
    // element fields, zero on the surface cells
    std::vector<double> J(ncell, 0.0), lambda(3 * ncell, 0.0), fr(3 * ncell, 0.0), sr(3 * ncell, 0.0),
        nr(3 * ncell, 0.0), Sii(3 * ncell, 0.0), Sij(3 * ncell, 0.0), Eii(3 * ncell, 0.0), Eij(3 * ncell, 0.0),
        Cii(3 * ncell, 0.0), Cij(3 * ncell, 0.0);
    const int dg[3][2] = {{0, 0}, {1, 1}, {2, 2}}, od[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int64_t e = 0; e < E; e++) {
        double F[9], B[9], T1[9], Fl[9], C[9];
        for (int r = 0; r < 9; r++) {
            F[r] = st[r * E + e];
            B[r] = basis0_[9 * e + r];
        }
        for (int i = 0; i < 3; i++)        // B^T F
            for (int j = 0; j < 3; j++) T1[3 * i + j] = B[i] * F[j] + B[3 + i] * F[3 + j] + B[6 + i] * F[6 + j];
        for (int i = 0; i < 3; i++)        // (B^T F) B
            for (int j = 0; j < 3; j++)
                Fl[3 * i + j] = T1[3 * i] * B[j] + T1[3 * i + 1] * B[3 + j] + T1[3 * i + 2] * B[6 + j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) C[3 * i + j] = Fl[i] * Fl[j] + Fl[3 + i] * Fl[3 + j] + Fl[6 + i] * Fl[6 + j];
        J[e] = Fl[0] * Fl[4] * Fl[8] + Fl[3] * Fl[7] * Fl[2] + Fl[6] * Fl[1] * Fl[5] - Fl[2] * Fl[4] * Fl[6] -
               Fl[5] * Fl[7] * Fl[0] - Fl[8] * Fl[3] * Fl[1];
        for (int c = 0; c < 3; c++) {
            lambda[3 * e + c] = std::sqrt(Fl[c] * Fl[c] + Fl[3 + c] * Fl[3 + c] + Fl[6 + c] * Fl[6 + c]);
            for (int i = 0; i < 3; i++) {
                fr[3 * e + i] += Fl[3 * i + c] * B[3 * c];
                sr[3 * e + i] += Fl[3 * i + c] * B[3 * c + 1];
                nr[3 * e + i] += Fl[3 * i + c] * B[3 * c + 2];
            }
            const int a = dg[c][0], b = dg[c][1], o0 = od[c][0], o1 = od[c][1];
            Sii[3 * e + c] = st[(9 + 3 * a + b) * E + e];
            Sij[3 * e + c] = st[(9 + 3 * o0 + o1) * E + e];
            Cii[3 * e + c] = C[3 * a + b];
            Cij[3 * e + c] = C[3 * o0 + o1];
            Eii[3 * e + c] = 0.5 * (C[3 * a + b] - 1.0);
            Eij[3 * e + c] = 0.5 * C[3 * o0 + o1];
        }
    }
    const std::string name = prefix_ + "." + std::to_string(counter_) + ".vtu";
    std::ofstream o(dir_ + "/" + prefix_ + "_vtu/" + name);
    if (!o) throw std::runtime_error("cannot create " + name);
    o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" "
         "header_type=\"UInt64\">\n  <UnstructuredGrid>\n    <Piece NumberOfPoints=\""
      << nNodes_ << "\" NumberOfCells=\"" << ncell << "\">\n      <PointData>\n";
    std::vector<int32_t> ids(nNodes_);
    for (int64_t i = 0; i < nNodes_; i++) ids[i] = (int32_t)i;
    DataArray(o, "Int32", "PointID", 1, ids);
    if (Option("Fixation")) DataArray(o, "Int32", "Fixation", 1, fixation_);
    if (Option("Displacement")) DataArray(o, "Float64", "Displacement", 3, inc);
    if (Option("AbsDisplacement")) {
        std::vector<double> u(n3);
        for (int64_t i = 0; i < n3; i++) u[i] = nodes[i] - refCoords_[i];
        DataArray(o, "Float64", "AbsDisplacement", 3, u);
    }
    if (Option("Velocity")) DataArray(o, "Float64", "Velocity", 3, vel);
    if (Option("Acceleration")) DataArray(o, "Float64", "Acceleration", 3, acc);
    o << "      </PointData>\n      <CellData>\n";
    std::vector<int32_t> cid(ncell), mat(ncell, 0);
    for (int64_t i = 0; i < ncell; i++) cid[i] = (int32_t)i;
    for (int64_t e = 0; e < E; e++) mat[e] = material_[e];
    DataArray(o, "Int32", "CellID", 1, cid);
    if (Option("Material")) DataArray(o, "Int32", "Material", 1, mat);
    if (Option("Jacobian")) DataArray(o, "Float64", "J", 1, J);
    if (Option("Lambda")) DataArray(o, "Float64", "Lambda", 3, lambda);
    if (Option("Fiber")) DataArray(o, "Float64", "f", 3, fr);
    if (Option("Sheet")) DataArray(o, "Float64", "s", 3, sr);
    if (Option("SheetNormal")) DataArray(o, "Float64", "n", 3, nr);
    if (Option("PK2Stress")) {
        DataArray(o, "Float64", "S_ii", 3, Sii);
        DataArray(o, "Float64", "S_ij", 3, Sij);
    }
    if (Option("GreenLagrangeStrain")) {
        DataArray(o, "Float64", "E_ii", 3, Eii);
        DataArray(o, "Float64", "E_ij", 3, Eij);
    }
    if (Option("Deformation")) {
        DataArray(o, "Float64", "C_ii", 3, Cii);
        DataArray(o, "Float64", "C_ij", 3, Cij);
    }
    o << "      </CellData>\n      <Points>\n";
    DataArray(o, "Float64", "Points", 3, nodes);
    o << "      </Points>\n      <Cells>\n";
    DataArray(o, "Int64", "connectivity", 1, connectivity_);
    DataArray(o, "Int64", "offsets", 1, offsets_);
    DataArray(o, "UInt8", "types", 1, types_);
    o << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
    o.close();
    // the collection: a DataSet entry before </Collection> (CBModelExporterVTK.cpp:53-76)
    const std::string pvdName = dir_ + "/" + prefix_ + ".pvd";
    std::ifstream in(pvdName);
    std::vector<std::string> lines;
    for (std::string ln; std::getline(in, ln);) lines.push_back(ln);
    in.close();
    std::ofstream pvd(pvdName);
    for (const auto& ln : lines) {
        if (ln.find("</Collection>") != std::string::npos)
            pvd << "\t\t<DataSet timestep=\"" << time << "\" part=\"0\" group=\"\" file=\"" << prefix_ << "_vtu/"
                << name << "\"/>\n";
        pvd << ln << "\n";
    }
    counter_++;
    lastTime_ = time;
    exported_ = true;
}

}  // namespace cmhost
