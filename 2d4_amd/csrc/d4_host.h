// d4_host.h -- the reference's host-side classes (namespace sdddstCore), kept
// with their names, members and call structure so a caller of the reference's
// Simulation / SimulationData / Field finds the same interface; the pair passes
// underneath run on the GPU through include/d4.h.
//
// Reference interfaces: Field (include/Fields/Field.h:34-50), StressProtocol
// (include/stress_protocol.h), PrecisionHandler (include/precision_handler.h),
// SimulationData (include/simulation_data.h:78-328, the members run() uses),
// Simulation (include/simulation.h:122-240).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/d4.h"
#include "d4_solver.h"

namespace sdddstCore {

struct DislwoB {  // include/dislocation.h:23-34: b implied by the index
    double x = 0, y = 0;
    DislwoB() = default;
    DislwoB(double x_, double y_) : x(x_), y(y_) {}
};

// Field::xy / xy_diff_x (Field.h:40-49), evaluated by lib2d4 on the GPU
class Field {
public:
    explicit Field(d4_ctx* ctx = nullptr) : ctx_(ctx) {}
    void bind(d4_ctx* ctx) { ctx_ = ctx; }
    double xy(double dx, double dy) const;
    double xy_diff_x(double dx, double dy) const;
    double xy(double dx, double dy, double exp2pix, double cos2piy) const;
    double xy_diff_x(double dx, double dy, double exp2pix, double cos2piy) const;

private:
    d4_ctx* ctx_;
};

// stress_protocol.cpp:13-73
class StressProtocol {
public:
    explicit StressProtocol(double init = 0) : initExtStress(init) {}
    virtual ~StressProtocol() = default;
    virtual double extStress(double) const { return 0; }

protected:
    double initExtStress;
};
class FixedRateProtocol : public StressProtocol {
public:
    FixedRateProtocol(double init, double rate) : StressProtocol(init), m_rate(rate) {}
    double extStress(double t) const override { return t * m_rate + initExtStress; }

protected:
    double m_rate;
};
class CyclicLoadProtocol : public FixedRateProtocol {
public:
    CyclicLoadProtocol(double init, double rate, double period) : FixedRateProtocol(init, rate), m_timePeriod(period) {}
    double extStress(double t) const override;

protected:
    double m_timePeriod;
};
class IncreasingCyclicLoadProtocol : public CyclicLoadProtocol {
public:
    IncreasingCyclicLoadProtocol(double init, double rate, double period, double amp)
        : CyclicLoadProtocol(init, rate, period), m_amplitudeRate(amp) {}
    double extStress(double t) const override { return CyclicLoadProtocol::extStress(t) * m_amplitudeRate * t; }

protected:
    double m_amplitudeRate;
};

// precision_handler.cpp:19-95
class PrecisionHandler {
public:
    void setSize(unsigned size) { if (use_dipoleCrit) tolerance.assign(size, minPrecisitySq); }
    void reset();
    void updateMaxErrorToleranceRatioSq(double error, unsigned ID);
    double getNewStepSize(double oldStepSize) const;
    void setMinPrecisity(double v) { minPrecisity = v; minPrecisitySq = v * v; }
    void setDipolePrecisity(double v);
    double getMaxErrorRatioSqr() const { return maxErrorToleranceRatioSq / dipolePrecisitySq; }
    unsigned maxErrorRatioID() const { return selectedID; }
    // the pair passes update `tolerance` on the GPU when the dipole criterion is on
    double* toleranceData() { return use_dipoleCrit ? tolerance.data() : nullptr; }
    double toleranceCap2() const { return use_dipoleCrit ? minPrecisitySq : 0.0; }

private:
    std::vector<double> tolerance;
    double minPrecisity = 0, minPrecisitySq = 0, maxErrorToleranceRatioSq = 0;
    unsigned selectedID = 0;
    bool use_dipoleCrit = false;
    double dipolePrecisity = 0, dipolePrecisitySq = 0;
};

class SimulationData {
public:
    // from a .dconf file (simulation_data.cpp:12-73, 76-154)
    explicit SimulationData(const std::string& path);
    // from arrays (same parse checks, sort and numbering)
    SimulationData(int n, const double* x, const double* y, const double* b);

    void writeDislocationDataToFile(const std::string& path) const;  // simulation_data.cpp:177-191
    void updateCutOff();                                              // simulation_data.cpp:253-262
    int b(unsigned ID) const { return ID < dc / 2 ? 1 : -1; }        // simulation_data.cpp:265-270

    std::vector<DislwoB> disl_sorted;
    std::vector<unsigned> disl_order;
    std::vector<double> g, initSpeed, initSpeed2, speed, speed2, dVec;
    std::vector<DislwoB> bigStep_sorted, firstSmall_sorted, secondSmall_sorted;
    std::vector<int> Ap, Ai, indexes;   // CSC of the step matrix (grown as needed, not realloc'd by N)
    std::vector<double> Ax, x;
    Field tau;
    double cutOffMultiplier = 0, cutOff = 0, cutOffSqr = 0, onePerCutOffSqr = 0;
    bool heavisideCutoff = false;
    double prec = 0, dipole_prec = 0;
    char weightFunc = 'c';
    unsigned dc = 0;
    double timeLimit = 0, stepSize = 0, initStepSize = 0, simTime = 0;
    size_t succesfulSteps = 0, failedSteps = 0;
    double totalAccumulatedStrainIncrease = 0;
    bool isStrainIncreaseLimit = false, isMaxStepSizeLimit = false, isTimeLimit = false, isStepCountLimit = false;
    double totalAccumulatedStrainIncreaseLimit = 0, maxStepSizeLimit = 0;
    int64_t stepCountLimit = 0;
    bool calculateStrainDuringSimulation = true;
    std::unique_ptr<StressProtocol> externalStressProtocol;
    double speedThresholdForCutoffChange = 0;
    bool isSpeedThresholdForCutoffChange = false;
    int solverKind = 0;  // 0 host sparse LU, 1 GPU dense LU
    std::string error;  // set when the input is rejected (the reference prints and exits)

private:
    void build(std::vector<double> x, std::vector<double> y, std::vector<double> b);
};

class Simulation;

// One of the pair passes a step needs (simulation.cpp:98-152 makes six GPU
// calls per step: two Jacobian passes, three speed passes and the stage-II
// rescale; run() adds one speed pass before the loop).
struct PassJob {
    enum Kind { None = 0, Speed, Jacobian, Rescale, Factor, Solve };
    Kind kind = None;
    const std::vector<DislwoB>* xy = nullptr;  // Speed, Jacobian
    double tau = 0;                            // Speed: extStress(t)
    std::vector<double>* v = nullptr;          // Speed: out
    d4_jac_params jp{};                        // Jacobian
    std::vector<double>* vA = nullptr;         // Jacobian: speeds at tau_a, tau_b
    std::vector<double>* vB = nullptr;
    double factor = 0, dtau = 0;               // Rescale (simulation.cpp:546-608)
    std::vector<double>* speed = nullptr;      // Rescale: speeds moved by dtau b
    int nnz = 0;                               // Jacobian: entries written to the step matrix
    // Factor: LU of the step matrix (the simulation's Ap, Ai, Ax; simulation.cpp:635-641);
    // the engine reports a zero pivot and whether rows were left in place
    bool singular = false, unpivoted = false;
    // Solve: sol = A^{-1} rhs with the last Factor's matrix (simulation.cpp:643-659)
    const std::vector<double>* rhs = nullptr;
    std::vector<double>* sol = nullptr;
};

// Runs PassJobs.  The default engine runs each job at once: pair passes through
// include/d4.h on the simulation's context, Factor / Solve with the simulation's
// own linear solver; a batching engine collects the same job from many
// simulations and runs them together.
class PassEngine {
public:
    virtual ~PassEngine() = default;
    virtual int run(Simulation& sim, PassJob& job);  // run_job, timed into the simulation's counters

protected:
    int run_job(Simulation& sim, PassJob& job);
};

class Simulation {
public:
    Simulation(std::shared_ptr<SimulationData> sD, d4_ctx* ctx);
    void run();           // simulation.cpp:30-281 (positions and counters; no log file)
    void run_prologue();  // the initial speed pass of run() (simulation.cpp:46)
    bool step();          // one loop iteration (stages I-IV); true if accepted
    double lastRatio() const { return lastRatio_; }
    bool interrupt() const;  // simulation.cpp:722-750 (limits that exist here)
    int status() const { return status_; }
    const std::string& error() const { return error_; }
    int64_t pairPasses = 0, factorizations = 0, singularSolves = 0, unpivotedFactorizations = 0;
    double passMs = 0.0, solveMs = 0.0;  // host wall time in pair passes and in linear solves

    // the step split at its pair passes and linear solves: begin_step() (false
    // when interrupt() holds) sets the first job, then each after_pass()
    // consumes job() and sets the next; job().kind == None when the step is
    // complete
    void begin_prologue();
    bool begin_step();
    void after_pass();
    PassJob& job() { return job_; }
    bool lastAccepted() const { return lastAccepted_; }
    SimulationData& data() { return *sD; }
    PrecisionHandler& precision() { return *pH; }
    d4_ctx* ctx() const { return ctx_; }
    d4host::LinearSolver& solver() { return *lu_; }
    bool check(int rc, const char* where);

private:
    std::shared_ptr<SimulationData> sD;
    std::unique_ptr<PrecisionHandler> pH;
    d4_ctx* ctx_;
    std::unique_ptr<d4host::LinearSolver> lu_;
    PassEngine default_engine_;
    int status_ = 0;
    double lastRatio_ = 0;
    bool lastAccepted_ = false;
    std::string error_;
    PassJob job_;
    int phase_ = 0;
    double t0_ = 0, t1_ = 0, th_ = 0;
    void set_speed_job(const std::vector<DislwoB>& dis, std::vector<double>& forces, double simTime);
    void set_jacobian_job(double stepsize, const std::vector<DislwoB>& dislocs, std::vector<double>& forces_A,
                          std::vector<double>& forces_B, double simTime_A, double simTime_B);
    void set_rescale_job(double baseTime, double newTime, double oldTime);
    bool singular_ = false;  // the last factorization met a zero pivot
    void set_factor_job();
    void after_factor();
    void set_solve_job();
    void after_solve();
    void calcG(const std::vector<DislwoB>& new_disloc, const std::vector<DislwoB>& old_config, double stepSize,
               const std::vector<double>& endSpeed, const std::vector<double>& initSpeed);
    void calcGFromSpeeds(double stepSize, const std::vector<double>& endSpeed, const std::vector<double>& initSpeed);
    void calculateXError();
    double calculateStrainIncrement(const std::vector<DislwoB>& old, const std::vector<DislwoB>& newD) const;
    bool finish_step();
};

}  // namespace sdddstCore
