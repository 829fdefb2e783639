// Host-side C++ layer above the C ABI (include/cmg.h).  It mirrors the reference's
// interfaces on the hot path so that a caller of the reference finds the same
// names, keys and return codes:
//   ParameterMap                      settings as dotted keys (ParameterMap/ParameterMap.h:29-152)
//   CBAnalyticFunction family         Linear / Sinus / SinusDriver / DoubleHill (CBData/CBAnalyticFunction.h)
//   TetgenModel                       .node/.ele/.bases/.sur + LAT files (CBModelLoaderTetgen.cpp,
//                                     CBFileManager.cpp:120-194)
//   CBFormulationTotalLagrangianGPU   CalcNodalForces / CalcNodalForcesJacobian (CBFormulation.h:30-41)
//   CBTiming                          adaptive time step (CBSolver/CBTiming.cpp:25-162)
//   CBApplyPressureFromFunction       prescribed cavity pressure (CBApplyPressureFromFunction.cpp:20-132)
//   CBCirculation, CBCircWindkessel3  cavity pressure from a 0D circulation (CBCirculation.cpp,
//                                     CBCircWindkessel3.cpp, CBCircModel.hpp)
//   CBModelExporterVTK                .pvd / .vtu result export (CBModelExporterVTK.cpp)
//   CBSolverGPU                       Init / SolverStep / Run (CBSolver.h:68-215) with the static
//                                     (CBSolverEquilibrium) and Newmark (CBSolverNewmarkBeta) steps
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cmg.h"

namespace cmhost {

// CBStatus (CBStatus/CBStatus.h:20-37)
enum class CBStatus : int {
    DACCORD = 6, WAITING = 5, NOT_INITIALIZED = 4, PREPARING_SIMULATION = 3, SUCCESS = 2, RUNNING = 1,
    NOTHING_DONE = 0, INTERRUPTED = -1, REPEAT = -2, FAILED = -3, INFINITIVE = -4, NOT_A_NUMBER = -5,
    CORRUPT_ELEMENT = -6, CRITICAL_STRESS = -7, SYSTEM_WRONG_STATE = -8
};
std::string CBStatusToStr(CBStatus s);

// ---------------------------------------------------------------------------
class ParameterMap {
public:
    void ParseXML(const std::string& path);
    void ParseXMLString(const std::string& text);
    void Set(const std::string& key, const std::string& value) { params_[key] = value; }
    bool IsAvailable(const std::string& key) const { return params_.count(key) > 0; }
    std::vector<std::string> GetChildNodes(const std::string& parent) const;
    template <class T> T Get(const std::string& key) const {
        auto it = params_.find(key);
        if (it == params_.end()) throw std::runtime_error(key + " not found in the parameter map");
        return Convert<T>(it->second);
    }
    template <class T> T Get(const std::string& key, T dflt) const {
        auto it = params_.find(key);
        return it == params_.end() ? dflt : Convert<T>(it->second);
    }
    template <class T> std::vector<T> GetArray(const std::string& key) const {
        std::vector<T> out;
        std::istringstream ss(Get<std::string>(key));
        std::string tok;
        while (std::getline(ss, tok, ',')) out.push_back(Convert<T>(tok));
        return out;
    }
    const std::map<std::string, std::string>& All() const { return params_; }

private:
    template <class T> static T Convert(const std::string& s) {
        std::string t = Trim(s);
        if constexpr (std::is_same<T, std::string>::value) {
            return t;
        } else if constexpr (std::is_same<T, bool>::value) {
            return t == "true" || t == "1" || t == "True" || t == "TRUE";
        } else {
            std::istringstream ss(t);
            T v{};
            ss >> v;
            if (ss.fail()) throw std::runtime_error("cannot convert '" + t + "'");
            return v;
        }
    }
    static std::string Trim(const std::string& s);
    std::map<std::string, std::string> params_;
};

// ---------------------------------------------------------------------------
class CBAnalyticFunction {
public:
    CBAnalyticFunction(const ParameterMap& p, const std::string& key);
    virtual ~CBAnalyticFunction() = default;
    virtual double operator()(double time) = 0;
    static std::unique_ptr<CBAnalyticFunction> Create(const std::string& type, const ParameterMap& p,
                                                      const std::string& key);
    // flat description used by the device tension model: type + fn[12] (cmg_material.fn)
    virtual int DeviceType() const = 0;
    virtual void DeviceParams(double* fn) const;

protected:
    double start_, stop_;
    bool keepMax_;
};

// ---------------------------------------------------------------------------
struct TetgenModel {
    int64_t nNodes = 0, nElements = 0;
    int npe = 10;
    std::vector<double> coords;         // [N*3] metres
    std::vector<uint8_t> bc;            // [N]
    std::vector<int32_t> elements;      // [E*npe] 0-based
    std::vector<int32_t> material;      // [E] material id as in the file
    std::vector<double> bases;          // [E*nb*9] row-major, columns f,s,n
    std::vector<int32_t> surfaces;      // [S*nps]
    std::vector<int32_t> surfaceIndex;  // [S]
    int nps = 3;
    std::vector<double> activation;     // [E]
    void Load(const std::string& nodes, const std::string& elements, const std::string& bases,
              const std::string& surfaces, double unit);
    void CheckNodeSorting();            // CBElementSolidT10::CheckNodeSorting (:573-618)
    void LoadActivationTimes(const std::string& latFile);
};

// ---------------------------------------------------------------------------
class CBTiming {
public:
    void Init(const ParameterMap& p);
    double GetTimeStep() const { return timeStep_; }
    double GetStartTime() const { return startTime_; }
    double GetStopTime() const { return stopTime_; }
    void SetSolverStatus(CBStatus s) { status_ = s; }
    CBStatus ComputeNextTimeStep();
    void ResetForInit();
    void ResetForSim();
    void ResetCounter() { counter_ = lowestLevelCounter_ = 0; }
    bool IsValid(double t) const { return t < stopTime_ && stopTime_ - t > 1e-14; }

private:
    double startTime_ = 0, stopTime_ = 1, firstTimeStep_ = 1e-4, minTimeStep_ = 1e-9;
    double maxSimTimeStep_ = 1e-4, maxInitTimeStep_ = 1e-4, maxTimeStep_ = 1e-4, timeStep_ = 1e-4;
    bool fastRelaxation_ = false;
    int minNumCnt_ = 4, counter_ = 0, lowestLevelCounter_ = 0;
    CBStatus status_ = CBStatus::SUCCESS;
};

// ---------------------------------------------------------------------------
class CBApplyPressureFromFunction {
public:
    // groups of Plugins.ApplyPressureFromFunction.Groups.*; surface index -> pressure group
    void Init(const ParameterMap& p, const std::vector<int32_t>& groupSurfaceIds);
    void Apply(double time, double dt);
    const std::vector<double>& Pressures() const { return pressure_; }
    bool Active() const { return !groups_.empty(); }

private:
    struct Interval {
        std::unique_ptr<CBAnalyticFunction> function;
        double startTime, stopTime, offset, amplitude;
        bool invert;
    };
    std::map<int, std::vector<Interval>> groups_;   // pressure group -> intervals
    std::vector<double> pressure_;
};

// ---------------------------------------------------------------------------
// 0D circulation (mechanics/src/CBCavity/CBCircModel.h, CBCircWindkessel3.cpp) and the
// plugin that couples it to the cavity volumes (CBSolverPlugins/CBCirculation.cpp).
// Pressures are in Plugins.Circulation.PressureUnit, volumes in VolumeUnit.
class CBCircModelBase {
public:
    CBCircModelBase(const ParameterMap& p, const std::string& prefix, int nC, int nS, int nR);
    virtual ~CBCircModelBase() = default;
    int GetNumberOfCavities() const { return nC_; }
    const std::vector<int>& CavitySurfaceIndices() const { return cavitySurfaces_; }
    const std::vector<double>& PreloadingPressures() const { return preloading_; }   // [2 nC]
    virtual void SetInitialCavityVolumes(const double* v) = 0;
    virtual void GetEstCavityVolumes(const double* p, double dt, double* v) = 0;
    void AcceptEstState();
    void RevertPrevState();
    bool SteadyStateCheck(double time);
    void WriteHeaderToFile();
    void WriteResultsToFile(double time);
    double Result(const std::string& name) const;
    const std::vector<double>& State() const { return state_; }

protected:
    void Integrate(const double* p, double dt);   // RK4 at constant cavity pressures (CBCircModel.hpp:187-236)
    virtual void Algebraics(const double* p, const double* u) = 0;
    virtual void Rates(const double* p, const double* u, double* du) = 0;
    std::string name_, filename_;
    int nC_, nS_, nR_;
    double maxTimeStep_;
    std::vector<double> state_, prevState_, estState_;
    std::vector<std::string> stateNames_, resultNames_;
    std::vector<double> results_;
    std::vector<double> cavityPressures_, prevCavityPressures_, estCavityPressures_;
    std::vector<int> cavitySurfaces_;
    std::vector<double> preloading_;
    bool steadyActive_ = false;
    double steadyStart_ = 0, steadyPeriod_ = 0, steadyThresh_ = 0;
    std::string steadyParam_;
    double prevCheckValue_ = 0, prevCheckTime_ = 0, prevCheckTimeActive_ = 0;
    void InitSteadyStateCheck(const ParameterMap& p, const std::string& prefix);
};

class CBCircWindkessel3 : public CBCircModelBase {
public:
    CBCircWindkessel3(const ParameterMap& p, const std::string& prefix);
    void SetInitialCavityVolumes(const double* v) override { state_[0] = v[0]; }
    void GetEstCavityVolumes(const double* p, double dt, double* v) override;

private:
    void Algebraics(const double* p, const double* u) override;
    void Rates(const double* p, const double* u, double* du) override;
    double artResist_, artCompli_, perResist_, venResist_, venPressure_, initialArtVolume_;
    bool resetArtVolume_;
};

class CBCirculation {
public:
    // groupSurfaceIds: surface index of each pressure group of the context
    void Init(const ParameterMap& p, const std::vector<int32_t>& groupSurfaceIds);
    // InitCircModels' volume part: FE cavity volumes (m^3 per pressure group) at the start
    void InitVolumes(const std::vector<double>& volumes, double dt);
    void Apply(double time, double dt);
    // cavity pressure in Pa per pressure group this plugin drives
    std::map<int, double> Pressures() const;
    void AnalyzeResults(const std::vector<double>& volumes);
    void StepBack();
    CBStatus GetStatus() const { return status_; }
    void WriteToFile(double time);
    bool ExitCheck(double time);
    bool TakeZeroVelocityRequest() { bool z = zeroVelocity_; zeroVelocity_ = false; return z; }
    bool Active() const { return !models_.empty(); }
    const std::vector<int>& CavitySurfaces() const { return surfaces_; }
    const std::vector<double>& EstimatedPressures() const { return est_; }
    std::vector<double> CouplingResidual() const;
    const std::vector<std::unique_ptr<CBCircModelBase>>& Models() const { return models_; }

private:
    static constexpr int N_ = 5;   // past pressures kept for the extrapolation
    void GetCircVolume(const std::vector<double>& p, std::vector<double>& v);
    std::vector<double> Extrapolate() const;
    void ExtrapolatePressure();
    void IncreasePressure();
    void PerturbPressure(int i);
    void AssembleComplianceMatrices(int i);
    void EstimatePressureUsingSecant();
    void EstimatePressureUsingQuasiNewton();
    void EstimatePressureUsingNewton();
    std::vector<std::unique_ptr<CBCircModelBase>> models_;
    std::vector<int> surfaces_, groups_;
    int nC_ = 0;
    CBStatus status_ = CBStatus::DACCORD;
    double uV_ = 1e6, uP_ = 133.322, maxP_ = 0, deltaP_ = 0, dP_ = 0;
    double tolerance_ = 1e-4, startTime_ = 0, preloadingTime1_ = 0, preloadingTime2_ = 0;
    double dampingFactor_ = 0, dampingDeclineFactor_ = 0;
    bool dampingActive_ = false, dampingIteration_ = false, perturbation_ = true;
    int maxIterations_ = 100, nNonNewton_ = 10, nSecant_ = 3;
    bool hasStarted_ = false, stepBack_ = false, preloadingFinished_ = false;
    bool lastPreloadingStep1_ = false, lastPreloadingStep2_ = false, zeroVelocity_ = false;
    double dt_ = 0, prevDt_ = 0;
    double timeStep_[N_] = {0, 0, 0, 0, 0};
    std::vector<std::vector<double>> pressure_;
    std::vector<double> est_, prevEst_, fine_, prevFine_, circ_, prevCirc_, residual_, prevResidual_;
    std::vector<double> preload1_, preload2_;
    std::vector<double> fineCompli_, circCompli_, sysElast_;   // nC x nC row-major
    unsigned long n_ = 1;
    unsigned i_ = 0, j_ = 0;
    unsigned long iSum_ = 0;
    std::string itsFile_;
};

// ---------------------------------------------------------------------------
// CBModelExporterVTK (CBModel/CBModelExporterVTK.cpp:29-499) with the fields of
// CBSolver::Export (CBSolver.cpp:538-1164) and CBSolverNewmarkBeta::Export
// (:800-827): <dir>/<prefix>.pvd and <dir>/<prefix>_vtu/<prefix>.<n>.vtu, binary
// base64 arrays with UInt64 headers; Export.Prefix, Export.TimeStep, Export.Options.*.
// The same files as cardiomechanics_amd/export.py writes.
class CBModelExporterVTK {
public:
    void Init(const ParameterMap& p, const TetgenModel& m, cmg_ctx* ctx);
    bool Active() const { return active_; }
    bool Due(double time) const;     // the start, then every Export.TimeStep (CBSolver.cpp:1808-1810)
    void Export(double time);
    int Counter() const { return counter_; }

private:
    bool Option(const std::string& name) const;
    bool active_ = false;
    cmg_ctx* ctx_ = nullptr;
    std::string dir_, prefix_;
    double timeStep_ = 1e-3, lastTime_ = 0.0;
    bool exported_ = false;
    int counter_ = 0;
    std::map<std::string, bool> options_;
    int64_t nNodes_ = 0, nSolid_ = 0;
    std::vector<double> refCoords_, basis0_;    // reference coordinates, centroid fibre frames [E][9]
    std::vector<int32_t> material_, fixation_;
    std::vector<int64_t> connectivity_, offsets_;
    std::vector<uint8_t> types_;
};

// ---------------------------------------------------------------------------
class CBSolverGPU;

class CBFormulationTotalLagrangianGPU {
public:
    explicit CBFormulationTotalLagrangianGPU(CBSolverGPU* solver) : solver_(solver) {}
    // element loop on the GPU at the solver's current coordinates and time
    CBStatus CalcNodalForces();           // fills solver nodal forces [3N]
    CBStatus CalcNodalForcesJacobian();   // fills solver Jacobian values [nnzb*9]

private:
    CBSolverGPU* solver_;   // no ownership
};

struct StepRecord {
    double time, dt;
    int newton, krylov;
    double residual, wall;
    std::vector<double> volumes;
    int repeats = 0;                 // plugin-requested repeats (circulation coupling iterations - 1)
    std::vector<double> pressures;   // cavity pressures applied (Pa per pressure group)
};

class CBSolverGPU {
public:
    ~CBSolverGPU();
    // Solver.Type = Static | NewmarkBeta; materials from Materials.Mat_<id>.*
    void Init(const ParameterMap& p, TetgenModel& model);
    CBStatus SolverStep(double time);
    // repeat: the previous call of this step answered REPEAT (start from its trial increment)
    CBStatus SolverStep(double time, bool repeat, cmg_step_stats* stats);
    void Run(std::vector<StepRecord>* records = nullptr, int maxSteps = -1);
    cmg_ctx* Context() { return ctx_; }
    const std::vector<double>& Nodes();   // current coordinates
    std::vector<double>& NodalForces() { return forces_; }
    std::vector<double>& JacobianValues() { return jacobian_; }
    double CurrentTime() const { return currentTime_; }
    CBTiming& GetTiming() { return timing_; }
    CBCirculation& GetCirculation() { return circulation_; }
    std::vector<double> CavityVolumes();
    const std::string& Type() const { return type_; }

private:
    friend class CBFormulationTotalLagrangianGPU;
    std::string type_;
    cmg_ctx* ctx_ = nullptr;
    CBTiming timing_;
    CBApplyPressureFromFunction pressure_;
    CBCirculation circulation_;
    CBModelExporterVTK exporter_;
    std::vector<double> appliedPressures_;
    void ApplyPlugins(double time);
    std::unique_ptr<CBFormulationTotalLagrangianGPU> formulation_;
    std::vector<double> nodes_, forces_, jacobian_;
    std::vector<int32_t> groupIds_;
    double currentTime_ = 0.0;
    int64_t nnzb_ = 0, nDofs_ = 0;
    std::string exportFile_;
};

// land17 (optional) receives, per material, the Land17 parameter block of
// cmg_set_land17_parameters, empty for materials of another tension model
std::vector<cmg_material> MaterialsFromParameters(const ParameterMap& p, const std::vector<int32_t>& ids,
                                                  std::vector<std::vector<double>>* land17 = nullptr);

}  // namespace cmhost
