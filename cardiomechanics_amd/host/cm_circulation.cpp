// 0D circulation coupled to the cavity volumes (host C++ layer):
//   CBCircModelBase    CBCircModel<nC, nS, nR> (mechanics/src/CBCavity/CBCircModel.hpp)
//   CBCircWindkessel3  mechanics/src/CBCavity/CBCircWindkessel3.cpp
//   CBCirculation      mechanics/src/CBSolverPlugins/CBCirculation.cpp
// The same algorithm as cardiomechanics_amd/circulation.py; the two are compared step
// for step on a linear cavity by the test suite.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <limits>

#include "cm_host.h"

namespace cmhost {

namespace {
// in-place inverse of an n x n row-major matrix (Gauss-Jordan with partial pivoting)
std::vector<double> Inverse(std::vector<double> a, int n) {
    std::vector<double> inv(n * n, 0.0);
    for (int i = 0; i < n; i++) inv[i * n + i] = 1.0;
    for (int c = 0; c < n; c++) {
        int piv = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(a[r * n + c]) > std::fabs(a[piv * n + c])) piv = r;
        if (a[piv * n + c] == 0.0) throw std::runtime_error("CBCirculation: singular system compliance");
        for (int k = 0; k < n; k++) {
            std::swap(a[c * n + k], a[piv * n + k]);
            std::swap(inv[c * n + k], inv[piv * n + k]);
        }
        const double d = a[c * n + c];
        for (int k = 0; k < n; k++) {
            a[c * n + k] /= d;
            inv[c * n + k] /= d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c) continue;
            const double f = a[r * n + c];
            for (int k = 0; k < n; k++) {
                a[r * n + k] -= f * a[c * n + k];
                inv[r * n + k] -= f * inv[c * n + k];
            }
        }
    }
    return inv;
}

std::vector<double> MatVec(const std::vector<double>& A, const std::vector<double>& x) {
    const int n = (int)x.size();
    std::vector<double> y(n, 0.0);
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) y[r] += A[r * n + c] * x[c];
    return y;
}
}  // namespace

// ===========================================================================
// CBCircModelBase
// ===========================================================================
CBCircModelBase::CBCircModelBase(const ParameterMap& p, const std::string& prefix, int nC, int nS, int nR)
    : nC_(nC), nS_(nS), nR_(nR) {
    name_ = prefix.substr(prefix.find_last_of('.') + 1);
    filename_ = p.Get<std::string>(prefix + ".ExportFile", "");
    maxTimeStep_ = p.Get<double>(prefix + ".MaxIntegrationTimeStep", 1e-4);
    state_.assign(nS, 0.0);
    prevState_.assign(nS, 0.0);
    estState_.assign(nS, 0.0);
    results_.assign(nR, 0.0);
    cavityPressures_.assign(nC, 0.0);
    prevCavityPressures_.assign(nC, 0.0);
    estCavityPressures_.assign(nC, 0.0);
    cavitySurfaces_.assign(nC, 0);
    preloading_.assign(2 * nC, 0.0);
}

void CBCircModelBase::Integrate(const double* p, double dt) {
    const int n = (int)std::ceil(dt / maxTimeStep_);
    const double h = dt / n;
    for (int i = 0; i < nC_; i++) estCavityPressures_[i] = p[i];
    estState_ = state_;
    std::vector<double> k1(nS_), k2(nS_), k3(nS_), k4(nS_), u(nS_);
    for (int s = 0; s < n; s++) {
        Rates(p, estState_.data(), k1.data());
        for (int i = 0; i < nS_; i++) u[i] = estState_[i] + h * k1[i] / 2.0;
        Rates(p, u.data(), k2.data());
        for (int i = 0; i < nS_; i++) u[i] = estState_[i] + h * k2[i] / 2.0;
        Rates(p, u.data(), k3.data());
        for (int i = 0; i < nS_; i++) u[i] = estState_[i] + h * k3[i];
        Rates(p, u.data(), k4.data());
        for (int i = 0; i < nS_; i++) estState_[i] += h * (k1[i] + 2.0 * k2[i] + 2.0 * k3[i] + k4[i]) / 6.0;
    }
    Algebraics(p, estState_.data());
}

void CBCircModelBase::AcceptEstState() {
    prevState_ = state_;
    state_ = estState_;
    prevCavityPressures_ = cavityPressures_;
    cavityPressures_ = estCavityPressures_;
}

void CBCircModelBase::RevertPrevState() {
    state_ = prevState_;
    cavityPressures_ = prevCavityPressures_;
}

double CBCircModelBase::Result(const std::string& name) const {
    for (int i = 0; i < nR_; i++)
        if (resultNames_[i] == name) return results_[i];
    for (int i = 0; i < nS_; i++)
        if (stateNames_[i] == name) return state_[i];
    throw std::runtime_error(name + " is not a state or result of " + name_);
}

void CBCircModelBase::InitSteadyStateCheck(const ParameterMap& p, const std::string& prefix) {
    steadyActive_ = p.Get<bool>(prefix + ".SteadyStateCheck.Active", false);
    if (!steadyActive_) return;
    steadyStart_ = p.Get<double>(prefix + ".SteadyStateCheck.StartTime");
    steadyPeriod_ = p.Get<double>(prefix + ".SteadyStateCheck.Period");
    steadyThresh_ = p.Get<double>(prefix + ".SteadyStateCheck.Threshold");
    if (p.Get<std::string>(prefix + ".SteadyStateCheck.Mode", "TemporalDifference") != "TemporalDifference")
        throw std::runtime_error("only TemporalDifference steady-state checks are available for " + name_);
    steadyParam_ = p.Get<std::string>(prefix + ".SteadyStateCheck.Parameter");
    (void)Result(steadyParam_);
}

bool CBCircModelBase::SteadyStateCheck(double time) {
    bool steady = false;
    if (steadyActive_) {
        const double half = 0.5 * (time - prevCheckTime_);
        if (std::fabs(time - steadyStart_) < half) {
            prevCheckValue_ = Result(steadyParam_);
            prevCheckTimeActive_ = steadyStart_;
        } else if (time > steadyStart_ && std::fabs(time - prevCheckTimeActive_ - steadyPeriod_) < half) {
            const double v = std::fabs(Result(steadyParam_) - prevCheckValue_);
            prevCheckValue_ = Result(steadyParam_);
            steady = v < steadyThresh_;
            prevCheckTimeActive_ = time;
        }
    }
    prevCheckTime_ = time;
    return steady;
}

void CBCircModelBase::WriteHeaderToFile() {
    if (filename_.empty()) return;
    std::ofstream f(filename_);
    if (!f.good()) throw std::runtime_error("cannot create " + filename_);
    f << std::setw(18) << "time";
    for (const auto& s : stateNames_) f << std::setw(18) << s;
    for (const auto& s : resultNames_) f << std::setw(18) << s;
}

void CBCircModelBase::WriteResultsToFile(double time) {
    if (filename_.empty()) return;
    std::ofstream f(filename_, std::ios::app);
    if (!f.good()) throw std::runtime_error("cannot open " + filename_);
    f << "\n" << std::setprecision(5) << std::fixed << std::setw(18) << time << std::setprecision(9);
    for (double v : estState_) f << std::setw(18) << v;
    for (double v : results_) f << std::setw(18) << v;
}

// ===========================================================================
// CBCircWindkessel3
// ===========================================================================
CBCircWindkessel3::CBCircWindkessel3(const ParameterMap& p, const std::string& prefix)
    : CBCircModelBase(p, prefix, 1, 2, 5) {
    stateNames_ = {"VentrVolume", "ArtVolume"};
    resultNames_ = {"VentrPressure", "ArtPressure", "ArtFlow", "PerFlow", "VenFlow"};
    artResist_ = p.Get<double>(prefix + ".CircParameters.ArtResist");
    artCompli_ = p.Get<double>(prefix + ".CircParameters.ArtCompli");
    perResist_ = p.Get<double>(prefix + ".CircParameters.PerResist");
    venResist_ = p.Get<double>(prefix + ".CircParameters.VenResist");
    venPressure_ = p.Get<double>(prefix + ".CircParameters.VenPressure");
    initialArtVolume_ = p.Get<double>(prefix + ".InitialConditions.ArtVolume");
    resetArtVolume_ = p.Get<bool>(prefix + ".InitialConditions.ResetArtVolume", false);
    state_[1] = initialArtVolume_;
    cavitySurfaces_[0] = p.Get<int>(prefix + ".Cavities.VentrSurface");
    preloading_[0] = p.Get<double>(prefix + ".Cavities.VentrPreloading", 0.0);
    preloading_[1] = p.Get<double>(prefix + ".InitialConditions.VentrPressure", 0.0);
    InitSteadyStateCheck(p, prefix);
    WriteHeaderToFile();
}

void CBCircWindkessel3::GetEstCavityVolumes(const double* p, double dt, double* v) {
    Integrate(p, dt);
    v[0] = estState_[0];
}

void CBCircWindkessel3::Algebraics(const double* p, const double* u) {
    const double pv = p[0];
    const double pc = u[1] / artCompli_;   // pressure of the arterial compliance
    double& artPressure = results_[1];
    double& artFlow = results_[2];
    if (pv >= pc) {                        // aortic valve open
        artFlow = (pv - pc) / artResist_;
        artPressure = pv;
    } else {
        artFlow = 0.0;
        artPressure = pc;
        if (resetArtVolume_) {
            estState_[1] = std::max(initialArtVolume_, artCompli_ * pv);
            artPressure = estState_[1] / artCompli_;
        }
    }
    results_[3] = pc / perResist_;
    results_[4] = venPressure_ >= pv ? (venPressure_ - pv) / venResist_ : 0.0;
    results_[0] = pv;
}

void CBCircWindkessel3::Rates(const double* p, const double* u, double* du) {
    Algebraics(p, u);
    du[0] = results_[4] - results_[2];
    du[1] = results_[2] - results_[3];
}

// ===========================================================================
// CBCirculation
// ===========================================================================
void CBCirculation::Init(const ParameterMap& p, const std::vector<int32_t>& groupSurfaceIds) {
    const std::string r = "Plugins.Circulation";
    startTime_ = p.Get<double>(r + ".StartTime", std::numeric_limits<double>::lowest());
    status_ = CBStatus::DACCORD;
    tolerance_ = p.Get<double>(r + ".Tolerance", 1e-4);
    maxIterations_ = p.Get<int>(r + ".MaxIterations", 100);
    nNonNewton_ = p.Get<int>(r + ".MaxNonNewtonIterations", 10);
    nSecant_ = p.Get<int>(r + ".SecantIterations", 3);
    perturbation_ = p.Get<bool>(r + ".Perturbation", true);
    preloadingTime1_ = p.Get<double>(r + ".PreloadingTime1", 0.0);
    preloadingTime2_ = p.Get<double>(r + ".PreloadingTime2", 0.0);
    if (preloadingTime1_ != 0.0)
        throw std::runtime_error("Plugins.Circulation.PreloadingTime1 > 0 relaxes the reference configuration, "
                                 "which this build does not provide");
    dampingFactor_ = std::min(p.Get<double>(r + ".CouplingDamping.InitialFactor", 0.0), 0.9);
    dampingActive_ = dampingFactor_ >= 1e-5;
    dampingDeclineFactor_ = p.Get<double>(r + ".CouplingDamping.DeclineFactor", 0.0);
    const std::string up = p.Get<std::string>(r + ".PressureUnit");
    if (up == "mmHg") uP_ = 1.33322e2;
    else if (up == "hPa") uP_ = 1e2;
    else if (up == "Pa") uP_ = 1.0;
    else throw std::runtime_error("Plugins.Circulation.PressureUnit " + up + " unknown");
    const std::string uv = p.Get<std::string>(r + ".VolumeUnit");
    if (uv == "ml") uV_ = 1e6;
    else if (uv == "m3") uV_ = 1.0;
    else throw std::runtime_error("Plugins.Circulation.VolumeUnit " + uv + " unknown");
    maxP_ = 1.33322e3 / uP_;
    deltaP_ = 1.33322 / uP_;
    dP_ = 1.33322e-6 / uP_;
    itsFile_ = p.Get<std::string>(r + ".IterationsFile", "");
    for (const auto& key : p.GetChildNodes(r + ".Circs")) {
        if (!p.Get<bool>(key + ".Active", false)) continue;
        const std::string type = p.Get<std::string>(key + ".Type", "");
        if (type == "CircWindkessel3") models_.push_back(std::make_unique<CBCircWindkessel3>(p, key));
        else throw std::runtime_error("circulation model type " + type + " is not available");
    }
    for (const auto& m : models_)
        for (int s : m->CavitySurfaceIndices()) {
            if (std::find(surfaces_.begin(), surfaces_.end(), s) != surfaces_.end())
                throw std::runtime_error("cavity " + std::to_string(s) + " is used more than once");
            auto it = std::find(groupSurfaceIds.begin(), groupSurfaceIds.end(), s);
            if (it == groupSurfaceIds.end())
                throw std::runtime_error("no cavity with surface index " + std::to_string(s));
            surfaces_.push_back(s);
            groups_.push_back((int)(it - groupSurfaceIds.begin()));
        }
    nC_ = (int)surfaces_.size();
    const std::vector<double> z(nC_, 0.0);
    est_ = prevEst_ = fine_ = prevFine_ = circ_ = prevCirc_ = residual_ = prevResidual_ = z;
    pressure_.assign(N_, z);
    for (const auto& m : models_) {
        const auto& pl = m->PreloadingPressures();
        const int k = m->GetNumberOfCavities();
        preload1_.insert(preload1_.end(), pl.begin(), pl.begin() + k);
        preload2_.insert(preload2_.end(), pl.begin() + k, pl.end());
    }
    fineCompli_.assign(nC_ * nC_, 0.0);
    circCompli_.assign(nC_ * nC_, 0.0);
    sysElast_.assign(nC_ * nC_, 0.0);
    for (int i = 0; i < nC_; i++) sysElast_[i * nC_ + i] = 1.0;
}

void CBCirculation::InitVolumes(const std::vector<double>& volumes, double dt) {
    dt_ = dt;
    for (int k = 0; k < nC_; k++) fine_[k] = uV_ * volumes[groups_[k]];
    circ_ = fine_;
    int pos = 0;
    for (auto& m : models_) {
        m->SetInitialCavityVolumes(fine_.data() + pos);
        pos += m->GetNumberOfCavities();
    }
    if (!itsFile_.empty()) {
        std::ofstream f(itsFile_);
        if (!f.good()) throw std::runtime_error("cannot create " + itsFile_);
    }
}

std::map<int, double> CBCirculation::Pressures() const {
    std::map<int, double> out;
    if (!hasStarted_) return out;
    for (int k = 0; k < nC_; k++) out[groups_[k]] = uP_ * est_[k];
    return out;
}

std::vector<double> CBCirculation::CouplingResidual() const {
    std::vector<double> r(nC_);
    for (int k = 0; k < nC_; k++) r[k] = std::fabs(fine_[k] - circ_[k]);
    return r;
}

void CBCirculation::GetCircVolume(const std::vector<double>& p, std::vector<double>& v) {
    int pos = 0;
    for (auto& m : models_) {
        m->GetEstCavityVolumes(p.data() + pos, dt_, v.data() + pos);
        pos += m->GetNumberOfCavities();
    }
}

std::vector<double> CBCirculation::Extrapolate() const {
    // Adams-Bashforth-type extrapolation over the last five accepted pressures with
    // their own step sizes (CBCirculation.cpp:518-525)
    auto P = [&](unsigned long k) -> const std::vector<double>& { return pressure_[k % N_]; };
    auto T = [&](unsigned long k) { return timeStep_[k % N_]; };
    const unsigned long n = n_;
    std::vector<double> out(nC_);
    for (int c = 0; c < nC_; c++) {
        const double d1 = (P(n - 1)[c] - P(n - 2)[c]) / T(n - 1);
        const double d2 = (P(n - 2)[c] - P(n - 3)[c]) / T(n - 2);
        const double d3 = (P(n - 3)[c] - P(n - 4)[c]) / T(n - 3);
        const double d4 = (P(n - 4)[c] - P(n - 5)[c]) / T(n - 4);
        out[c] = P(n - 1)[c] + dt_ / 24.0 * (55.0 * d1 - 59.0 * d2 + 37.0 * d3 - 9.0 * d4);
    }
    return out;
}

void CBCirculation::ExtrapolatePressure() {
    prevEst_ = est_;
    est_ = Extrapolate();
    if (dampingActive_)
        for (int c = 0; c < nC_; c++)
            est_[c] = (est_[c] - dampingFactor_ * pressure_[(n_ - 1) % N_][c]) / (1.0 - dampingFactor_);
}

void CBCirculation::IncreasePressure() {
    prevEst_ = est_;
    for (double& e : est_) e += deltaP_;
}

void CBCirculation::PerturbPressure(int i) {
    prevEst_[i] = est_[i];
    est_[i] += dP_;
}

void CBCirculation::AssembleComplianceMatrices(int i) {
    const double dp = est_[i] - prevEst_[i];
    for (int r = 0; r < nC_; r++) {
        fineCompli_[r * nC_ + i] = (fine_[r] - prevFine_[r]) / dp;
        circCompli_[r * nC_ + i] = (circ_[r] - prevCirc_[r]) / dp;
    }
}

void CBCirculation::EstimatePressureUsingSecant() {
    std::vector<double> dpdr(nC_);
    for (int c = 0; c < nC_; c++) {
        const double dp = est_[c] - prevEst_[c];
        dpdr[c] = 1.0 / ((fine_[c] - prevFine_[c]) / dp - (circ_[c] - prevCirc_[c]) / dp);
    }
    prevEst_ = est_;
    for (int c = 0; c < nC_; c++) est_[c] -= dpdr[c] * (fine_[c] - circ_[c]);
}

void CBCirculation::EstimatePressureUsingQuasiNewton() {
    // Broyden update of the inverse system compliance (CBCirculation.cpp:582-600)
    prevResidual_ = residual_;
    for (int c = 0; c < nC_; c++) residual_[c] = fine_[c] - circ_[c];
    if (i_ > 1 || n_ < 10) {
        std::vector<double> dF(nC_), dX(nC_);
        for (int c = 0; c < nC_; c++) {
            dF[c] = residual_[c] - prevResidual_[c];
            dX[c] = est_[c] - prevEst_[c];
        }
        const std::vector<double> v = MatVec(sysElast_, dF);
        double dxv = 0.0;
        for (int c = 0; c < nC_; c++) dxv += dX[c] * v[c];
        // sysElast += ((dX - v) / (dX . v)) (dX^T sysElast)
        std::vector<double> xtH(nC_, 0.0);
        for (int c = 0; c < nC_; c++)
            for (int r = 0; r < nC_; r++) xtH[c] += dX[r] * sysElast_[r * nC_ + c];
        for (int r = 0; r < nC_; r++)
            for (int c = 0; c < nC_; c++) sysElast_[r * nC_ + c] += (dX[r] - v[r]) / dxv * xtH[c];
    }
    prevEst_ = est_;
    const std::vector<double> step = MatVec(sysElast_, residual_);
    for (int c = 0; c < nC_; c++) est_[c] -= step[c];
}

void CBCirculation::EstimatePressureUsingNewton() {
    for (int c = 0; c < nC_; c++) residual_[c] = fine_[c] - circ_[c];
    std::vector<double> sys(nC_ * nC_);
    for (int k = 0; k < nC_ * nC_; k++) sys[k] = fineCompli_[k] - circCompli_[k];
    sysElast_ = Inverse(sys, nC_);
    prevEst_ = est_;
    const std::vector<double> step = MatVec(sysElast_, residual_);
    for (int c = 0; c < nC_; c++) est_[c] -= step[c];
    j_ = 0;
}

void CBCirculation::Apply(double time, double dt) {
    if (!hasStarted_) {
        if (time > startTime_) hasStarted_ = true;
        else return;
    }
    prevDt_ = dt_;
    dt_ = dt;
    if (preloadingFinished_) {
        const unsigned nS = (unsigned)nSecant_, nC = (unsigned)nC_, nNN = (unsigned)nNonNewton_;
        if (dampingIteration_) {
            for (int c = 0; c < nC_; c++)
                est_[c] = dampingFactor_ * pressure_[(n_ - 1) % N_][c] + (1.0 - dampingFactor_) * est_[c];
            dampingFactor_ *= dampingDeclineFactor_;
            if (dampingFactor_ < 1e-5) dampingActive_ = false;
        } else if (status_ == CBStatus::SUCCESS) {
            if (!stepBack_) {
                for (auto& m : models_) m->AcceptEstState();
                pressure_[n_ % N_] = est_;
                timeStep_[n_ % N_] = prevDt_;
                n_++;
                i_ = 0;
                j_ = 0;
            }
            if (n_ > (unsigned long)N_) ExtrapolatePressure();
            else IncreasePressure();
        } else {
            if (i_ == 0) {
                if (n_ > (unsigned long)N_) ExtrapolatePressure();
                else IncreasePressure();
            } else if (perturbation_ && i_ <= nS) {
                EstimatePressureUsingSecant();
            } else if (perturbation_ && i_ <= nS + nC) {
                PerturbPressure((int)(i_ - nS - 1));
            } else if (j_ > nNN && j_ <= nNN + nC) {
                PerturbPressure((int)(j_ - nNN - 1));
            } else if ((!perturbation_ || i_ > nS + nC + 1) && j_ <= nNN) {
                EstimatePressureUsingQuasiNewton();
            }
            if (perturbation_ || j_ > nNN + 1) {
                if (i_ > nS + 1 && i_ <= nS + nC + 1) AssembleComplianceMatrices((int)(i_ - nS - 2));
                else if (j_ > nNN + 1) AssembleComplianceMatrices((int)(j_ - nNN - 2));
                if (i_ == nS + nC + 1 || j_ == nNN + nC + 1) EstimatePressureUsingNewton();
            }
        }
        for (int c = 0; c < nC_; c++) {
            const double diff = est_[c] - prevEst_[c];
            if (diff > maxP_) est_[c] = prevEst_[c] + maxP_;
            else if (diff < -maxP_) est_[c] = prevEst_[c] - maxP_;
        }
        stepBack_ = false;
        i_++;
        j_++;
        iSum_++;
    } else {
        if (lastPreloadingStep1_) {
            zeroVelocity_ = true;   // SetZeroVelocityAndAcceleration before this step's solve
            lastPreloadingStep1_ = false;
        }
        if (time < preloadingTime1_) {
            for (int c = 0; c < nC_; c++) est_[c] = -time / preloadingTime1_ * preload1_[c];
        } else if (std::fabs(time - preloadingTime1_) < 0.5 * dt_) {
            for (int c = 0; c < nC_; c++) est_[c] = -1.0 * preload1_[c];
            lastPreloadingStep1_ = true;
        } else if (time <= preloadingTime1_ + preloadingTime2_) {
            for (int c = 0; c < nC_; c++) est_[c] = (time - preloadingTime1_) / preloadingTime2_ * preload2_[c];
        } else {
            est_ = preload2_;
            pressure_[0] = est_;
            lastPreloadingStep2_ = true;
        }
    }
    status_ = CBStatus::SUCCESS;
}

void CBCirculation::AnalyzeResults(const std::vector<double>& volumes) {
    if (!hasStarted_) return;
    std::vector<double> fe(nC_);
    for (int k = 0; k < nC_; k++) fe[k] = uV_ * volumes[groups_[k]];
    if (!preloadingFinished_) {
        if (lastPreloadingStep2_) {
            fine_ = fe;
            circ_ = fe;
            int pos = 0;
            for (auto& m : models_) {
                m->SetInitialCavityVolumes(fe.data() + pos);
                pos += m->GetNumberOfCavities();
            }
            preloadingFinished_ = true;
            status_ = CBStatus::DACCORD;
        } else {
            status_ = CBStatus::SUCCESS;
        }
        return;
    }
    prevFine_ = fine_;
    fine_ = fe;
    prevCirc_ = circ_;
    GetCircVolume(est_, circ_);
    status_ = CBStatus::SUCCESS;
    if (dampingIteration_) {
        dampingIteration_ = false;
        return;
    }
    for (int k = 0; k < nC_; k++) {
        if (std::fabs(fine_[k] - circ_[k]) > tolerance_) {
            if (i_ > (unsigned)maxIterations_) {
                status_ = CBStatus::FAILED;
            } else {
                status_ = CBStatus::REPEAT;
                break;
            }
        }
    }
    if (dampingActive_ && status_ == CBStatus::SUCCESS) {
        status_ = CBStatus::REPEAT;
        dampingIteration_ = true;
    }
}

void CBCirculation::StepBack() {
    if (!hasStarted_) return;
    est_ = pressure_[(n_ - 1) % N_];
    stepBack_ = true;
    i_ = 0;
    j_ = 0;
}

void CBCirculation::WriteToFile(double time) {
    if (!hasStarted_) return;
    if (time > preloadingTime1_ + preloadingTime2_ + dt_) {
        for (auto& m : models_) m->WriteResultsToFile(time);
        if (!itsFile_.empty()) {
            std::ofstream f(itsFile_, std::ios::app);
            f << std::setprecision(3) << std::fixed << std::setw(14) << time << std::setw(14) << i_ << "\n";
        }
    }
}

bool CBCirculation::ExitCheck(double time) {
    if (!hasStarted_ || models_.empty()) return false;
    bool all = true;
    for (auto& m : models_)
        if (!m->SteadyStateCheck(time)) all = false;
    return all;
}

}  // namespace cmhost
