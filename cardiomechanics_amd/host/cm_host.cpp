// Host-side C++ layer above include/cmg.h (see cm_host.h for the reference interfaces
// each class mirrors).
#include "cm_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <numeric>

namespace cmhost {

std::string CBStatusToStr(CBStatus s) {
    switch (s) {
    case CBStatus::SUCCESS: return "SUCCESS";
    case CBStatus::FAILED: return "FAILED";
    case CBStatus::REPEAT: return "REPEAT";
    case CBStatus::CORRUPT_ELEMENT: return "CORRUPT_ELEMENT";
    case CBStatus::INFINITIVE: return "INFINITIVE";
    case CBStatus::NOT_A_NUMBER: return "NOT_A_NUMBER";
    case CBStatus::NOTHING_DONE: return "NOTHING_DONE";
    default: return "STATUS(" + std::to_string((int)s) + ")";
    }
}

static void check(int rc, const char* what) {
    if (rc != CMG_SUCCESS)
        throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + cmg_last_error());
}

// ===========================================================================
// ParameterMap: nested XML elements flattened to dotted keys
// ===========================================================================
std::string ParameterMap::Trim(const std::string& s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

void ParameterMap::ParseXML(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open settings file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    ParseXMLString(ss.str());
}

void ParameterMap::ParseXMLString(const std::string& t) {
    std::vector<std::string> stack;
    std::string text;
    size_t i = 0;
    auto key = [&]() {
        std::string k;
        for (size_t q = 0; q < stack.size(); q++) k += (q ? "." : "") + stack[q];
        return k;
    };
    while (i < t.size()) {
        if (t[i] != '<') {
            text += t[i++];
            continue;
        }
        if (t.compare(i, 4, "<!--") == 0) {   // comment
            size_t e = t.find("-->", i);
            i = (e == std::string::npos) ? t.size() : e + 3;
            continue;
        }
        if (t.compare(i, 2, "<?") == 0) {     // declaration
            size_t e = t.find("?>", i);
            i = (e == std::string::npos) ? t.size() : e + 2;
            continue;
        }
        size_t e = t.find('>', i);
        if (e == std::string::npos) throw std::runtime_error("unterminated XML tag");
        std::string tag = t.substr(i + 1, e - i - 1);
        i = e + 1;
        if (!tag.empty() && tag[0] == '/') {      // closing tag
            std::string tv = Trim(text);
            if (!tv.empty() && !stack.empty()) params_[key()] = tv;
            text.clear();
            if (!stack.empty()) stack.pop_back();
            continue;
        }
        bool self_closing = !tag.empty() && tag.back() == '/';
        if (self_closing) tag.pop_back();
        std::string name = Trim(tag.substr(0, tag.find_first_of(" \t\r\n")));
        text.clear();
        if (self_closing) {
            stack.push_back(name);
            params_[key()] = "";
            stack.pop_back();
        } else {
            stack.push_back(name);
        }
    }
    // the reference's files wrap everything in a root tag (e.g. <settings>): drop it
    std::map<std::string, std::string> flat;
    for (auto& kv : params_) {
        size_t dot = kv.first.find('.');
        flat[dot == std::string::npos ? kv.first : kv.first.substr(dot + 1)] = kv.second;
    }
    params_.swap(flat);
}

std::vector<std::string> ParameterMap::GetChildNodes(const std::string& parent) const {
    std::set<std::string> out;
    const std::string pre = parent + ".";
    for (auto& kv : params_)
        if (kv.first.compare(0, pre.size(), pre) == 0) {
            size_t d = kv.first.find('.', pre.size());
            out.insert(kv.first.substr(0, d));
        }
    return std::vector<std::string>(out.begin(), out.end());
}

// ===========================================================================
// analytic functions (CBData/CBAnalyticFunction.h:39-176)
// ===========================================================================
CBAnalyticFunction::CBAnalyticFunction(const ParameterMap& p, const std::string& key) {
    start_ = p.Get<double>(key + ".StartTime", 0);
    stop_ = p.Get<double>(key + ".StopTime", 0);
    keepMax_ = p.Get<bool>(key + ".KeepMaxValue", true);
}

void CBAnalyticFunction::DeviceParams(double* fn) const {
    for (int i = 0; i < 12; i++) fn[i] = 0.0;
    fn[0] = start_;
    fn[1] = stop_;
    fn[2] = keepMax_ ? 1.0 : 0.0;
}

namespace {
class Linear : public CBAnalyticFunction {
public:
    using CBAnalyticFunction::CBAnalyticFunction;
    double operator()(double time) override {
        double val;
        if (time < start_)
            val = 0.0;
        else if (time < stop_ + 0.5 * (time - prev_))
            val = (time - start_) * 1.0 / (stop_ - start_);
        else
            val = keepMax_ ? 1.0 : 0.0;
        prev_ = time;
        return val;
    }
    int DeviceType() const override { return CMG_FN_LINEAR; }

private:
    double prev_ = 0.0;
};

class Sinus : public CBAnalyticFunction {
public:
    Sinus(const ParameterMap& p, const std::string& key) : CBAnalyticFunction(p, key) {
        omega_ = p.Get<double>(key + ".Sinus.Omega", 0);
        phi_ = p.Get<double>(key + ".Sinus.Phi", 0);
    }
    double operator()(double time) override {
        if (time >= start_ && time <= stop_) return 0.5 - 0.5 * std::cos(omega_ * (time - start_) + phi_);
        return 0;
    }
    int DeviceType() const override { return CMG_FN_SINUS; }
    void DeviceParams(double* fn) const override {
        CBAnalyticFunction::DeviceParams(fn);
        fn[3] = omega_;
        fn[4] = phi_;
    }

private:
    double omega_, phi_;
};

class SinusDriver : public CBAnalyticFunction {
public:
    using CBAnalyticFunction::CBAnalyticFunction;
    double operator()(double time) override {
        if (time >= start_ && time <= stop_) return std::sin(3.14159 * (time - start_) / (stop_ - start_));
        return 0;
    }
    int DeviceType() const override { return CMG_FN_SINUSDRIVER; }
};

class DoubleHill : public CBAnalyticFunction {
public:
    DoubleHill(const ParameterMap& p, const std::string& key) : CBAnalyticFunction(p, key) {
        period_ = p.Get<double>(key + ".DoubleHill.Period", 0.8);
        contrOff_ = p.Get<double>(key + ".DoubleHill.ContrTimeOffset", 0.215);
        relaxOff_ = p.Get<double>(key + ".DoubleHill.RelaxTimeOffset", 0.362);
        contrRate_ = p.Get<double>(key + ".DoubleHill.ContrRateConst", 1.32);
        relaxRate_ = p.Get<double>(key + ".DoubleHill.RelaxRateConst", 21.9);
        onset_ = p.Get<double>(key + ".DoubleHill.OnsetTime", 0.0);
        driverMax_ = -INFINITY;
        for (int i = 0; i <= 1000; i++) driverMax_ = std::max(driverMax_, Raw(i * period_ / 1000));
    }
    double operator()(double time) override {
        if (!(time >= start_ && time <= stop_)) return 0.0;
        double t = time - onset_;
        t = (t > 0.0) ? std::fmod(t, period_) : 0.0;
        return Raw(t) / driverMax_;
    }
    int DeviceType() const override { return CMG_FN_DOUBLEHILL; }
    void DeviceParams(double* fn) const override {
        CBAnalyticFunction::DeviceParams(fn);
        fn[5] = period_;
        fn[6] = contrOff_;
        fn[7] = relaxOff_;
        fn[8] = contrRate_;
        fn[9] = relaxRate_;
        fn[10] = onset_;
        fn[11] = driverMax_;
    }

private:
    double Raw(double t) const {
        double c = std::pow(t / contrOff_, contrRate_);
        c = c / (1.0 + c);
        double r = std::pow(t / relaxOff_, relaxRate_);
        r = 1.0 / (1.0 + r);
        return c * r;
    }
    double period_, contrOff_, relaxOff_, contrRate_, relaxRate_, onset_, driverMax_;
};
}  // namespace

std::unique_ptr<CBAnalyticFunction> CBAnalyticFunction::Create(const std::string& type, const ParameterMap& p,
                                                                const std::string& key) {
    if (type == "Linear") return std::make_unique<Linear>(p, key);
    if (type == "Sinus") return std::make_unique<Sinus>(p, key);
    if (type == "SinusDriver") return std::make_unique<SinusDriver>(p, key);
    if (type == "DoubleHill") return std::make_unique<DoubleHill>(p, key);
    throw std::runtime_error("CBAnalyticFunction: Type " + type + " is unknown");
}

// ===========================================================================
// Tetgen model
// ===========================================================================
namespace {
std::vector<std::vector<std::string>> ReadRows(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::vector<std::vector<std::string>> rows;
    std::string line;
    while (std::getline(in, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line.resize(h);
        std::istringstream ss(line);
        std::vector<std::string> r;
        std::string tok;
        while (ss >> tok) r.push_back(tok);
        if (!r.empty()) rows.push_back(r);
    }
    return rows;
}
}  // namespace

void TetgenModel::Load(const std::string& nodeFile, const std::string& eleFile, const std::string& basesFile,
                       const std::string& surFile, double unit) {
    auto nr = ReadRows(nodeFile);
    nNodes = std::stoll(nr[0][0]);
    const bool attr = nr[0].size() > 2 && std::stoi(nr[0][2]) > 0;
    coords.resize(3 * nNodes);
    bc.assign(nNodes, 0);
    for (int64_t i = 0; i < nNodes; i++) {
        const auto& r = nr[i + 1];
        if (std::stoll(r[0]) != i + 1) throw std::runtime_error("node ids must be 1-based and consecutive");
        for (int c = 0; c < 3; c++) coords[3 * i + c] = std::stod(r[1 + c]) * unit;
        if (attr && r.size() > 4) bc[i] = (uint8_t)std::stoi(r[4]);
    }
    auto er = ReadRows(eleFile);
    nElements = std::stoll(er[0][0]);
    npe = std::stoi(er[0][1]);
    elements.resize(nElements * npe);
    material.assign(nElements, 0);
    for (int64_t e = 0; e < nElements; e++) {
        const auto& r = er[e + 1];
        for (int k = 0; k < npe; k++) elements[e * npe + k] = std::stoi(r[1 + k]) - 1;
        if ((int)r.size() > npe + 1) material[e] = std::stoi(r[npe + 1]);
    }
    const int nb = (npe == 10) ? 5 : 1;
    bases.assign(nElements * nb * 9, 0.0);
    for (int64_t e = 0; e < nElements * nb; e++) bases[9 * e] = bases[9 * e + 4] = bases[9 * e + 8] = 1.0;
    if (!basesFile.empty()) {
        auto br = ReadRows(basesFile);
        const int nq = std::stoi(br[0][1]);
        for (int64_t e = 0; e < nElements; e++) {
            const auto& r = br[e + 1];
            for (int q = 0; q < nb; q++) {
                const int src = (nq == 1) ? 0 : q;
                for (int col = 0; col < 3; col++)
                    for (int row = 0; row < 3; row++)   // stored f, s, n: columns of the Matrix3
                        bases[(e * nb + q) * 9 + 3 * row + col] = std::stod(r[1 + src * 9 + 3 * col + row]);
            }
        }
    }
    if (!surFile.empty()) {
        auto sr = ReadRows(surFile);
        const int64_t S = std::stoll(sr[0][0]);
        nps = std::stoi(sr[0][1]);
        surfaces.resize(S * nps);
        surfaceIndex.resize(S);
        for (int64_t s = 0; s < S; s++) {
            const auto& r = sr[s + 1];
            for (int k = 0; k < nps; k++) surfaces[s * nps + k] = std::stoi(r[1 + k]) - 1;
            surfaceIndex[s] = std::stoi(r[nps + 2]);
        }
    }
    activation.assign(nElements, 0.0);
}

void TetgenModel::CheckNodeSorting() {
    if (npe != 10) return;
    for (int64_t e = 0; e < nElements; e++) {
        int32_t* n = &elements[e * 10];
        const double* p[4];
        for (int k = 0; k < 4; k++) p[k] = &coords[3 * (int64_t)n[k]];
        double d1[3], d2[3], d3[3];
        for (int c = 0; c < 3; c++) {
            d1[c] = p[1][c] - p[0][c];
            d2[c] = p[2][c] - p[0][c];
            d3[c] = p[3][c] - p[0][c];
        }
        const double v = d1[0] * (d2[1] * d3[2] - d3[1] * d2[2]) + d1[1] * (d3[0] * d2[2] - d2[0] * d3[2]) +
                         d1[2] * (d2[0] * d3[1] - d3[0] * d2[1]);
        if (v < 0) {
            std::swap(n[1], n[2]);
            std::swap(n[6], n[4]);
            std::swap(n[8], n[9]);
            for (int r = 0; r < 9; r++) std::swap(bases[(e * 5 + 2) * 9 + r], bases[(e * 5 + 3) * 9 + r]);
        }
    }
}

void TetgenModel::LoadActivationTimes(const std::string& lat) {
    std::ifstream in(lat);
    if (!in) throw std::runtime_error("cannot open " + lat);
    std::string line;
    std::getline(in, line);   // header skipped (CBFileManager.cpp:120-194)
    int64_t e = 0;
    while (std::getline(in, line) && e < nElements) {
        std::istringstream ss(line);
        std::string id, t;
        if (!(ss >> id >> t)) continue;
        double v = (t == "nan" || t == "NaN") ? 0.0 : std::stod(t);
        activation[e++] = std::isnan(v) ? 0.0 : v;
    }
}

// ===========================================================================
// materials (CBConstitutiveModelGuccione.cpp:20-71, CBConstitutiveModelHolzapfel.cpp:18-85,
// CBMaterial.cpp:19-74, CBTensionFromFunction.cpp:36-56)
// ===========================================================================
std::vector<cmg_material> MaterialsFromParameters(const ParameterMap& p, const std::vector<int32_t>& ids,
                                                  std::vector<std::vector<double>>* land17) {
    std::vector<cmg_material> out;
    if (land17) land17->assign(ids.size(), std::vector<double>());
    for (int32_t id : ids) {
        const std::string own = "Materials.Mat_" + std::to_string(id) + ".";
        const std::string def = "Materials.Mat_Default.";
        auto key = [&](const std::string& k) { return (!p.IsAvailable(own + k) && p.IsAvailable(def + k)) ? def + k : own + k; };
        cmg_material m;
        std::memset(&m, 0, sizeof(m));
        const std::string type = p.Get<std::string>(key("Type"), "Guccione");
        m.density = p.Get<double>(key("Density"), 0.0);
        m.ignore_corrupt = p.Get<bool>(key("IgnoreCorruptElements"), false) ? 1 : 0;
        if (type == "Guccione") {
            m.law = CMG_LAW_GUCCIONE;
            const double as = p.Get<double>(key("Guccione.aScale"), 1.0), bs = p.Get<double>(key("Guccione.bScale"), 1.0);
            m.params[0] = p.Get<double>(key("Guccione.C")) * as;
            m.params[1] = p.Get<double>(key("Guccione.bf")) * bs;
            m.params[2] = p.Get<double>(key("Guccione.bt")) * bs;
            m.params[3] = p.Get<double>(key("Guccione.bfs")) * bs;
            m.params[4] = p.Get<double>(key("Guccione.K"));
        } else if (type == "Holzapfel") {
            m.law = CMG_LAW_HOLZAPFEL;
            const char* names[9] = {"a", "b", "af", "bf", "as", "bs", "afs", "bfs", "kappa"};
            for (int k = 0; k < 9; k++) m.params[k] = p.Get<double>(key(std::string("Holzapfel.") + names[k]));
            m.params[9] = p.Get<double>(key("Holzapfel.k"), 100.0);
        } else {
            throw std::runtime_error("material law " + type + " is not on this path (Guccione, Holzapfel)");
        }
        const std::string tm = p.Get<std::string>(key("TensionModel"), "None");
        if (tm == "FromFunction") {
            m.tension_max = p.Get<double>(own + "TensionMax");   // no Mat_Default fallback (:44)
            const std::string ftype = p.Get<std::string>(key("FromFunction.Type"), "None");
            auto fn = CBAnalyticFunction::Create(ftype, p, own + "FromFunction");
            m.tension_fn = fn->DeviceType();
            fn->DeviceParams(m.fn);
        } else if (tm == "TanH") {
            // CBTensionModelTanH (CBTensionModelTanH.cpp:27-125)
            m.tension_max = p.Get<double>(key("TensionMax"), 0.0);
            if (m.tension_max != 1000.0)
                throw std::runtime_error("CBTensionModelTanH: Parameter TensionMax != 1000 is not compatible with "
                                         "this model to ensure model output in Pa!");
            auto tkey = [&](const std::string& tail, double dflt) {
                const std::string o = own + "TanH." + tail, d = def + "TanH." + tail;
                // the reference reads the fallback key without its tail (InitKey, :18-24): no value
                if (!p.IsAvailable(o) && p.IsAvailable(d)) return p.Get<double>(def + "TanH");
                return p.Get<double>(o, dflt);
            };
            const char* names[9] = {"cycleLength", "MinFiberStretch", "EMD", "PeakTension", "Duration",
                                    "ContractionTime", "DegreeOfLengthDependence", "LengthDependenceUpstrokeTime",
                                    "RelaxationTime"};
            const double dflts[9] = {1.0, 0.7, 0.015, 120.0, 0.3, 0.1, 5.0, 0.5, 0.1};
            for (int k = 0; k < 9; k++) m.fn[k] = tkey(names[k], dflts[k]);
            m.fn[9] = tkey("kff", 1.0);
            m.fn[10] = tkey("startTime", 0.0);
            if (tkey("kss", 0.0) != 0.0 || tkey("knn", 0.0) != 0.0 || tkey("ksn", 0.0) != 0.0)
                throw std::runtime_error("TanH stress coefficients kss/knn/ksn are not on this path (only kff)");
            m.tension_fn = CMG_FN_TANH;
        } else if (tm == "Land17") {
            // CBTensionModelLand17::InitParamsFromXml (CBTensionModelLand17.cpp:432-500); Tmax is the
            // material's TensionMax with its Mat_Default fallback (CBMaterial.cpp:39-42)
            m.tension_max = p.Get<double>(key("TensionMax"), 0.0);
            auto lkey = [&](const std::string& tail, double dflt) {
                const std::string o = own + "Land17." + tail, d = def + "Land17." + tail;
                // InitKey reads the fallback key without its tail (:289-295): no value
                if (!p.IsAvailable(o) && p.IsAvailable(d)) return p.Get<double>(def + "Land17");
                return p.Get<double>(o, dflt);
            };
            if (lkey("kss", 0.0) != 0.0 || lkey("knn", 0.0) != 0.0 || lkey("ksn", 0.0) != 0.0)
                throw std::runtime_error("Land17 stress coefficients kss/knn/ksn are not on this path (only kff)");
            // block order of cmg_set_land17_parameters; the two switches have no fallback (:434, :445)
            const char* names[22] = {"TRPNn", "TRPNk", "Ca50", "ku", "nTM", "TRPN50", "kuw", "kws",
                                     "rs", "rw", "gammas", "gammaw", "phi", "Aeff", "beta0", "beta1",
                                     "Tref", "a", "k", "eta_l", "eta_s", "xi"};
            const double dflts[22] = {2.0, 0.1, 0.805, 1.0, 5.0, 0.35, 0.182, 0.012, 0.25, 0.5, 0.0085,
                                      0.615, 2.23, 25.0, 2.3, -2.4, 120.0, 2.1, 7.0, 200.0, 20.0, 1.0};
            std::vector<double> b(CMG_LAND17_NPARAMS, 0.0);
            for (int k = 0; k < 22; k++) b[k] = lkey(names[k], dflts[k]);
            b[22] = lkey("startTime", 0.0);
            b[23] = lkey("cycleLength", 1.0);
            const std::string ca = p.Get<std::string>(own + "Land17.CalciumTransientType", "Coppini");
            if (ca == "Coppini") b[24] = 0.0;
            else if (ca == "Brixius") b[24] = 1.0;
            else if (ca == "External" || ca == "Elphy")
                throw std::runtime_error("Land17: calcium transient " + ca + " needs a calcium source this build "
                                         "does not have (Coppini, Brixius)");
            else
                throw std::runtime_error("CBTensionModelLand17::InitParamsFromXml: no suitable calcium transient "
                                         "was defined by the user");
            // any value other than OFF switches the rate dependence on (:368)
            b[25] = p.Get<std::string>(own + "Land17.rateDependancy", "OFF") == "OFF" ? 0.0 : 1.0;
            b[26] = lkey("kff", 1.0);
            const char* st[7] = {"XS", "XW", "TRPN", "TmBlocked", "ZetaS", "ZetaW", "Cd"};
            const double st0[7] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < 7; k++) b[27 + k] = lkey(st[k], st0[k]);
            m.tension_fn = CMG_FN_LAND17;
            if (!land17) throw std::runtime_error("Land17 materials need the parameter-block output");
            (*land17)[out.size()] = b;
        } else if (tm != "None") {
            throw std::runtime_error("tension model " + tm + " is not on this path (FromFunction, TanH, Land17)");
        }
        out.push_back(m);
    }
    return out;
}

// ===========================================================================
// CBTiming (CBTiming.cpp:25-162)
// ===========================================================================
void CBTiming::Init(const ParameterMap& p) {
    startTime_ = p.Get<double>("Solver.StartTime", 0);
    stopTime_ = p.Get<double>("Solver.StopTime", 1);
    firstTimeStep_ = p.Get<double>("Solver.TimeStep", 1e-4);
    minTimeStep_ = p.Get<double>("Solver.MinTimeStep", 1e-9);
    maxSimTimeStep_ = p.Get<double>("Solver.MaxTimeStep", firstTimeStep_);
    maxInitTimeStep_ = p.Get<double>("Solver.MaxInitTimeStep", maxSimTimeStep_);
    fastRelaxation_ = p.Get<bool>("Solver.FastRelaxation", false);
    minNumCnt_ = p.Get<int>("Solver.MinSteps", 4);
    ResetForInit();
}

void CBTiming::ResetForInit() {
    maxTimeStep_ = maxInitTimeStep_;
    timeStep_ = std::min(firstTimeStep_, maxTimeStep_);
    ResetCounter();
}

void CBTiming::ResetForSim() {
    maxTimeStep_ = maxSimTimeStep_;
    timeStep_ = std::min(firstTimeStep_, maxTimeStep_);
    ResetCounter();
}

CBStatus CBTiming::ComputeNextTimeStep() {
    if (status_ == CBStatus::FAILED) {
        lowestLevelCounter_ = 0;
        counter_ *= 2;
        timeStep_ /= 2;
        return timeStep_ < minTimeStep_ ? CBStatus::FAILED : CBStatus::SUCCESS;
    }
    if (status_ == CBStatus::SUCCESS) {
        counter_++;
        lowestLevelCounter_++;
        while (counter_ % 2 == 0 && counter_ % minNumCnt_ == 0 && lowestLevelCounter_ >= minNumCnt_) {
            counter_ = counter_ / 2;
            if (timeStep_ < maxInitTimeStep_) {
                timeStep_ *= 2;
                if (!fastRelaxation_) break;
            }
        }
    }
    return CBStatus::SUCCESS;
}

// ===========================================================================
// CBApplyPressureFromFunction (CBApplyPressureFromFunction.cpp:20-111)
// ===========================================================================
void CBApplyPressureFromFunction::Init(const ParameterMap& p, const std::vector<int32_t>& groupSurfaceIds) {
    pressure_.assign(std::max<size_t>(1, groupSurfaceIds.size()), 0.0);
    if (!p.Get<bool>("Solver.Plugins.ApplyPressureFromFunction", false)) return;
    for (const auto& gk : p.GetChildNodes("Plugins.ApplyPressureFromFunction.Groups")) {
        std::vector<int> surf = p.GetArray<int>(gk + ".Surfaces");
        std::vector<Interval> ivs;
        double prevStop = 0.0;
        for (const auto& ik : p.GetChildNodes(gk + ".Intervals")) {
            Interval iv;
            iv.function = CBAnalyticFunction::Create(p.Get<std::string>(ik + ".Type", "Linear"), p, ik);
            iv.invert = p.Get<bool>(ik + ".Invert", false);
            iv.startTime = p.Get<double>(ik + ".StartTime", 0.0);
            if (iv.startTime < prevStop) throw std::runtime_error("interval StartTime before previous StopTime");
            iv.stopTime = p.Get<double>(ik + ".StopTime", 0.0);
            prevStop = iv.stopTime;
            iv.offset = p.Get<double>(ik + ".Offset", 0.0);
            iv.amplitude = p.Get<double>(ik + ".Amplitude", 0.0);
            ivs.push_back(std::move(iv));
        }
        if (ivs.empty()) throw std::runtime_error("no Intervals specified in " + gk);
        for (int sid : surf) {
            auto it = std::find(groupSurfaceIds.begin(), groupSurfaceIds.end(), sid);
            if (it == groupSurfaceIds.end())
                throw std::runtime_error("no cavity with surface index " + std::to_string(sid));
            // each surface gets its own copy of the interval functions (stateful Linear)
            std::vector<Interval> copy;
            for (const auto& ik : p.GetChildNodes(gk + ".Intervals")) {
                Interval iv;
                iv.function = CBAnalyticFunction::Create(p.Get<std::string>(ik + ".Type", "Linear"), p, ik);
                iv.invert = p.Get<bool>(ik + ".Invert", false);
                iv.startTime = p.Get<double>(ik + ".StartTime", 0.0);
                iv.stopTime = p.Get<double>(ik + ".StopTime", 0.0);
                iv.offset = p.Get<double>(ik + ".Offset", 0.0);
                iv.amplitude = p.Get<double>(ik + ".Amplitude", 0.0);
                copy.push_back(std::move(iv));
            }
            groups_[(int)(it - groupSurfaceIds.begin())] = std::move(copy);
        }
    }
}

void CBApplyPressureFromFunction::Apply(double time, double dt) {
    for (auto& g : groups_)
        for (auto& iv : g.second)
            if (time >= iv.startTime && time < iv.stopTime + 0.5 * dt) {
                const double f = (*iv.function)(time - iv.offset);   // CBDataFromFunction.cpp:45-49
                pressure_[g.first] = iv.offset + iv.amplitude * (iv.invert ? (1.0 - f) : f);
                break;
            }
}

// ===========================================================================
// formulation and solver
// ===========================================================================
CBStatus CBFormulationTotalLagrangianGPU::CalcNodalForces() {
    CBSolverGPU* s = solver_;
    const auto& x = s->Nodes();
    s->forces_.resize(s->nDofs_);
    int rc = cmg_residual(s->ctx_, x.data(), s->currentTime_, 1, s->forces_.data());
    return rc == CMG_SUCCESS ? CBStatus::SUCCESS : (CBStatus)std::max(rc, -8);
}

CBStatus CBFormulationTotalLagrangianGPU::CalcNodalForcesJacobian() {
    CBSolverGPU* s = solver_;
    const auto& x = s->Nodes();
    s->jacobian_.resize(s->nnzb_ * 9);
    int rc = cmg_tangent(s->ctx_, x.data(), s->currentTime_, 1, s->jacobian_.data());
    return rc == CMG_SUCCESS ? CBStatus::SUCCESS : (CBStatus)std::max(rc, -8);
}

CBSolverGPU::~CBSolverGPU() {
    if (ctx_) cmg_destroy(ctx_);
}

void CBSolverGPU::Init(const ParameterMap& p, TetgenModel& m) {
    type_ = p.Get<std::string>("Solver.Type", "Static");
    if (type_ != "Static" && type_ != "NewmarkBeta")
        throw std::runtime_error("Solver.Type " + type_ + " is not on this path (Static, NewmarkBeta)");
    if (p.Get<std::string>("Solver.Formulation", "TotalLagrangian") != "TotalLagrangian")
        throw std::runtime_error("only Solver.Formulation = TotalLagrangian exists");
    timing_.Init(p);
    // materials: ids in order of first appearance, elements mapped to table positions
    std::vector<int32_t> ids;
    std::vector<int32_t> matidx(m.nElements);
    for (int64_t e = 0; e < m.nElements; e++) {
        auto it = std::find(ids.begin(), ids.end(), m.material[e]);
        if (it == ids.end()) {
            ids.push_back(m.material[e]);
            it = ids.end() - 1;
        }
        matidx[e] = (int32_t)(it - ids.begin());
    }
    std::vector<std::vector<double>> land17;
    std::vector<cmg_material> mats = MaterialsFromParameters(p, ids, &land17);
    // pressure surfaces: surface indices named by the plugin groups
    std::vector<int32_t> sel_conn, sel_group;
    std::set<int> wanted;
    for (const auto& gk : p.GetChildNodes("Plugins.ApplyPressureFromFunction.Groups"))
        for (int sid : p.GetArray<int>(gk + ".Surfaces")) wanted.insert(sid);
    const bool circulation = p.Get<bool>("Solver.Plugins.Circulation", false);
    if (circulation)
        for (const auto& ck : p.GetChildNodes("Plugins.Circulation.Circs"))
            if (p.Get<bool>(ck + ".Active", false)) wanted.insert(p.Get<int>(ck + ".Cavities.VentrSurface"));
    groupIds_.assign(wanted.begin(), wanted.end());
    const int64_t S = (int64_t)m.surfaceIndex.size();
    for (int64_t s = 0; s < S; s++) {
        auto it = std::find(groupIds_.begin(), groupIds_.end(), m.surfaceIndex[s]);
        if (it == groupIds_.end()) continue;
        for (int k = 0; k < m.nps; k++) sel_conn.push_back(m.surfaces[s * m.nps + k]);
        sel_group.push_back((int32_t)(it - groupIds_.begin()));
    }
    pressure_.Init(p, groupIds_);
    if (circulation) {
        if (type_ != "NewmarkBeta") throw std::runtime_error("the circulation plugin needs Solver.Type = NewmarkBeta");
        circulation_.Init(p, groupIds_);
    }
    cmg_options o;
    cmg_default_options(&o);
    const double prec = p.Get<double>("Solver.Precision", 1e-8);
    o.newton_atol = o.newton_rtol = o.newton_stol = prec;
    o.newton_maxit = p.Get<int>("Solver.MaxIterations", 50);
    o.newmark_beta = p.Get<double>("Solver.NewmarkBeta.Beta", 0.25);
    o.newmark_gamma = p.Get<double>("Solver.NewmarkBeta.Gamma", 0.5);
    o.consistent_mass = p.Get<bool>("Solver.NewmarkBeta.ConsistentMassMatrix", true) ? 1 : 0;
    o.fd_epsilon = p.Get<double>("Solver.Epsilon", 1e-12);
    {
        // CBSolverNewmarkBeta::InitDampingParameter (:274-288): the type is read under
        // "Material.Global", the coefficients under "Materials.Global"
        const std::string dtype = p.Get<std::string>("Material.Global.Damping.Type", "Rayleigh");
        if (dtype != "Rayleigh") throw std::runtime_error("Global damping type " + dtype + " is unknown");
        o.rayleigh_alpha = p.Get<double>("Materials.Global.Damping.Rayleigh.Alpha", 0.0);
        o.rayleigh_beta = p.Get<double>("Materials.Global.Damping.Rayleigh.Beta", 0.0);
    }
    o.precision = p.Get<std::string>("Solver.GPU.Precision", "FP64") == "Mixed" ? CMG_PREC_MIXED : CMG_PREC_FP64;
    o.krylov_rtol = p.Get<double>("Solver.GPU.KrylovTolerance", o.krylov_rtol);
    o.krylov_restart = p.Get<int>("Solver.GPU.KrylovRestart", o.krylov_restart);
    o.krylov_maxit = p.Get<int>("Solver.GPU.KrylovMaxIterations", o.krylov_maxit);
    {
        const std::string km = p.Get<std::string>("Solver.GPU.KrylovMethod", "BiCGStab");
        if (km == "GMRES") o.krylov_method = 0;
        else if (km == "BiCGStabGraph") o.krylov_method = 1;
        else if (km == "BiCGStab") o.krylov_method = 2;
        else throw std::runtime_error("Solver.GPU.KrylovMethod must be GMRES, BiCGStab or BiCGStabGraph, not " + km);
    }
    {
        const std::string pc = p.Get<std::string>("Solver.GPU.Preconditioner", "Schwarz");
        if (pc == "Schwarz") o.preconditioner = 0;   // vertex-patch Schwarz where it applies (T10)
        else if (pc == "BlockJacobi") o.preconditioner = 1;
        else throw std::runtime_error("Solver.GPU.Preconditioner must be Schwarz or BlockJacobi, not " + pc);
    }
    o.precond_reuse = p.Get<bool>("Solver.GPU.PreconditionerReuse", true) ? 1 : 0;
    o.newton_stall = p.Get<int>("Solver.GPU.NewtonStall", o.newton_stall);
    o.spmv_choice = p.Get<int>("Solver.GPU.SpMVLayout", 0);
    o.transpose_tangent = p.Get<bool>("Solver.GPU.TransposeTangent", false) ? 1 : 0;
    o.pressure_tangent = p.Get<bool>("Solver.GPU.PressureTangent", true) ? 1 : 0;
    o.device = p.Get<int>("Solver.GPU.Device", 0);
    cmg_mesh mesh;
    std::memset(&mesh, 0, sizeof(mesh));
    mesh.n_nodes = m.nNodes;
    mesh.n_elements = m.nElements;
    mesh.n_surfaces = (int64_t)sel_group.size();
    mesh.nodes_per_element = m.npe;
    mesh.nodes_per_surface = mesh.n_surfaces ? m.nps : 0;
    mesh.n_groups = std::max<int>(1, (int)groupIds_.size());
    mesh.ref_coords = m.coords.data();
    mesh.elements = m.elements.data();
    mesh.bases = m.bases.data();
    mesh.material_index = matidx.data();
    mesh.activation_time = m.activation.data();
    mesh.bc = m.bc.data();
    mesh.surfaces = sel_conn.empty() ? nullptr : sel_conn.data();
    mesh.surface_group = sel_group.empty() ? nullptr : sel_group.data();
    check(cmg_create(&mesh, mats.data(), (int32_t)mats.size(), &o, &ctx_), "cmg_create");
    for (size_t k = 0; k < land17.size(); k++)
        if (!land17[k].empty())
            check(cmg_set_land17_parameters(ctx_, (int32_t)k, land17[k].data()), "cmg_set_land17_parameters");
    int64_t nn, ne;
    check(cmg_sizes(ctx_, &nn, &ne, &nDofs_, &nnzb_), "cmg_sizes");
    formulation_ = std::make_unique<CBFormulationTotalLagrangianGPU>(this);
    exportFile_ = p.Get<std::string>("Plugins.ApplyPressureFromFunction.ExportFile", "");
    appliedPressures_.assign(std::max<size_t>(1, groupIds_.size()), 0.0);
    if (circulation_.Active()) circulation_.InitVolumes(CavityVolumes(), timing_.GetTimeStep());
    exporter_.Init(p, m, ctx_);
}

// Apply of every plugin before a solve (CBSolverNewmarkBeta.cpp:839-840): the
// prescribed pressures, then the circulation's estimates for its cavities.
void CBSolverGPU::ApplyPlugins(double time) {
    if (!pressure_.Active() && !circulation_.Active()) return;
    const double dt = timing_.GetTimeStep();
    std::fill(appliedPressures_.begin(), appliedPressures_.end(), 0.0);
    if (pressure_.Active()) {
        pressure_.Apply(time, dt);
        appliedPressures_ = pressure_.Pressures();
    }
    if (circulation_.Active()) {
        circulation_.Apply(time, dt);
        for (const auto& gp : circulation_.Pressures()) appliedPressures_[gp.first] = gp.second;
        if (circulation_.TakeZeroVelocityRequest()) {
            std::vector<double> z(nDofs_, 0.0);
            check(cmg_set_state(ctx_, nullptr, z.data(), z.data()), "cmg_set_state");
        }
    }
    check(cmg_set_pressures(ctx_, appliedPressures_.data()), "cmg_set_pressures");
}

const std::vector<double>& CBSolverGPU::Nodes() {
    nodes_.resize(nDofs_);
    check(cmg_get_state(ctx_, nodes_.data(), nullptr, nullptr, nullptr), "cmg_get_state");
    return nodes_;
}

std::vector<double> CBSolverGPU::CavityVolumes() {
    std::vector<double> v(std::max<size_t>(1, groupIds_.size()));
    check(cmg_cavity_volumes(ctx_, nullptr, v.data()), "cmg_cavity_volumes");
    return v;
}

CBStatus CBSolverGPU::SolverStep(double time) {
    return SolverStep(time, false, nullptr);
}

// CBSolverNewmarkBeta::SolverStep / CBSolverEquilibrium::SolverStep with the plugins:
// with the circulation the converged trial is analysed before it is accepted
// (CBSolverNewmarkBeta.cpp:928-976); REPEAT leaves it pending for the next solve.
CBStatus CBSolverGPU::SolverStep(double time, bool repeat, cmg_step_stats* stats) {
    currentTime_ = time;
    cmg_step_stats local;
    cmg_step_stats* st = stats ? stats : &local;
    std::memset(st, 0, sizeof(*st));
    ApplyPlugins(time);
    const double dt = timing_.GetTimeStep();
    int rc;
    if (type_ == "Static") rc = cmg_static_step(ctx_, time, dt, st);
    else if (!circulation_.Active()) rc = cmg_newmark_step(ctx_, time, dt, st);
    else rc = cmg_newmark_solve(ctx_, time, dt, repeat ? 1 : 0, st);
    if (rc != CMG_SUCCESS && rc != CMG_FAILED)
        throw std::runtime_error(std::string("solver step failed: ") + cmg_last_error());
    if (rc == CMG_FAILED) return CBStatus::FAILED;
    if (circulation_.Active()) {
        std::vector<double> vol(appliedPressures_.size());
        check(cmg_trial_volumes(ctx_, vol.data()), "cmg_trial_volumes");
        circulation_.AnalyzeResults(vol);
        const CBStatus f = circulation_.GetStatus();
        if (f == CBStatus::FAILED || f == CBStatus::REPEAT) return f;
        check(cmg_commit_step(ctx_), "cmg_commit_step");
    }
    return CBStatus::SUCCESS;
}

// CBSolver::Run (CBSolver.cpp:1724-1858): first pass at t = StartTime, then steps until StopTime
void CBSolverGPU::Run(std::vector<StepRecord>* records, int maxSteps) {
    std::ofstream out;
    if (!exportFile_.empty()) {
        out.open(exportFile_);
        out << std::setw(16) << "time";
        for (int id : groupIds_) out << std::setw(16) << ("pressure" + std::to_string(id)) << std::setw(16)
                                     << ("volume" + std::to_string(id));
        out << "\n";
    }
    double time = timing_.GetStartTime();
    bool first = true;
    int steps = 0;
    timing_.ResetForSim();
    while (timing_.IsValid(time) || first) {
        bool repeat = false;
        int repeats = 0;
        while (true) {
            double t = time + timing_.GetTimeStep();
            if (first) {
                time = timing_.GetStartTime() - timing_.GetTimeStep();
                t = timing_.GetStartTime();
            }
            cmg_step_stats st;
            const auto w0 = std::chrono::steady_clock::now();
            const CBStatus rc = SolverStep(t, repeat, &st);
            const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
            // plugins not asking for the repeat go back to their last accepted state
            if (rc == CBStatus::FAILED || (rc == CBStatus::REPEAT && circulation_.GetStatus() != CBStatus::REPEAT))
                circulation_.StepBack();
            if (rc == CBStatus::REPEAT) repeats++;
            repeat = rc == CBStatus::REPEAT;
            timing_.SetSolverStatus(rc);
            const CBStatus next = timing_.ComputeNextTimeStep();
            if (rc == CBStatus::SUCCESS) {
                std::vector<double> vol = CavityVolumes();
                if (records)
                    records->push_back({t, first ? 0.0 : t - time, st.newton_iterations, st.krylov_iterations,
                                        st.residual_norm, wall, vol, repeats, appliedPressures_});
                if (out) {
                    // pressure in mmHg and volume in ml (CBApplyPressureFromFunction.cpp:167)
                    out << std::setprecision(7) << std::fixed << std::setw(16) << t;
                    for (size_t g = 0; g < groupIds_.size(); g++)
                        out << std::setw(16) << appliedPressures_[g] / 133.322 << std::setw(16) << vol[g] * 1e6;
                    out << "\n";
                }
                time = t;
                break;
            }
            if (next != CBStatus::SUCCESS)
                throw std::runtime_error("minimum time step size has been reached with no success");
        }
        if (first) {
            first = false;
            timing_.ResetCounter();
        }
        if (exporter_.Due(time)) exporter_.Export(time);   // CBSolver.cpp:1808-1818
        circulation_.WriteToFile(time);
        const bool exitNow = circulation_.ExitCheck(time);
        if (maxSteps > 0 && ++steps >= maxSteps) break;
        if (exitNow) break;
    }
}

}  // namespace cmhost
