// d4_sim.cpp -- host simulation (d4_host.h) and its C ABI (include/d4_sim.h).
#include "../../include/d4_sim.h"
#include "d4_host.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <numeric>

namespace sdddstCore {

// ------------------------------------------------------------------ Field

static double field_one(d4_ctx* ctx, double dx, double dy, const double* e, const double* c, bool deriv)
{
    double s = NAN, d = NAN;
    if (!ctx) return NAN;
    d4_field_eval(ctx, 1, &dx, &dy, e, c, deriv ? nullptr : &s, deriv ? &d : nullptr, 0);
    return deriv ? d : s;
}
double Field::xy(double dx, double dy) const { return field_one(ctx_, dx, dy, nullptr, nullptr, false); }
double Field::xy_diff_x(double dx, double dy) const { return field_one(ctx_, dx, dy, nullptr, nullptr, true); }
double Field::xy(double dx, double dy, double e, double c) const { return field_one(ctx_, dx, dy, &e, &c, false); }
double Field::xy_diff_x(double dx, double dy, double e, double c) const
{
    return field_one(ctx_, dx, dy, &e, &c, true);
}

// ------------------------------------------------------------------ stress

double CyclicLoadProtocol::extStress(double t) const  // stress_protocol.cpp:47-56
{
    const double pt = t - std::floor((t / m_timePeriod + 1. / 4)) * m_timePeriod;
    double s = m_rate * pt + initExtStress;
    if (pt > m_timePeriod / 4) s = m_rate * m_timePeriod / 2 - s + initExtStress;
    return s;
}

// ------------------------------------------------------------------ precision

void PrecisionHandler::reset()  // precision_handler.cpp:26-31
{
    maxErrorToleranceRatioSq = 0;
    if (use_dipoleCrit) tolerance.assign(tolerance.size(), minPrecisitySq);
}

void PrecisionHandler::updateMaxErrorToleranceRatioSq(double error, unsigned ID)  // precision_handler.cpp:41-54
{
    const double tmp = use_dipoleCrit ? error * error / tolerance[ID] : error * error / minPrecisitySq;
    if (tmp > maxErrorToleranceRatioSq) {
        maxErrorToleranceRatioSq = tmp;
        selectedID = ID;
    }
}

double PrecisionHandler::getNewStepSize(double oldStepSize) const  // precision_handler.cpp:57-62
{
    const double factor = std::min(2., std::pow(getMaxErrorRatioSqr(), -1. / 6));
    return 0.9 * oldStepSize * factor;
}

void PrecisionHandler::setDipolePrecisity(double v)  // precision_handler.cpp:72-89
{
    if (v == 0) {
        use_dipoleCrit = false;
        dipolePrecisity = 1;
        dipolePrecisitySq = 1;
    } else {
        use_dipoleCrit = true;
        dipolePrecisity = v;
        minPrecisity /= v;
        dipolePrecisitySq = v * v;
        minPrecisitySq /= v * v;
    }
}

// ------------------------------------------------------------------ data

static void normalize(double& v)  // sim_utils.cpp:6-13
{
    while (v < -0.5) v += 1;
    while (v >= 0.5) v -= 1;
}

SimulationData::SimulationData(const std::string& path)
{
    std::ifstream in(path);
    if (!in) {
        error = "cannot open dislocation data file " + path;
        return;
    }
    std::vector<double> xs, ys, bs;
    for (double x; in >> x;) {
        double y, b;
        if (!(in >> y && in >> b)) {
            error = "cannot read the y coordinate and Burgers vector of a dislocation in " + path;
            return;
        }
        xs.push_back(x);
        ys.push_back(y);
        bs.push_back(b);
    }
    build(xs, ys, bs);
}

SimulationData::SimulationData(int n, const double* x, const double* y, const double* b)
{
    if (n < 0 || (n > 0 && (!x || !y || !b))) {
        error = "invalid arrays";
        return;
    }
    build(std::vector<double>(x, x + n), std::vector<double>(y, y + n), std::vector<double>(b, b + n));
}

// simulation_data.cpp:98-154: integer b, sum b == 0, sort by y + b descending
void SimulationData::build(std::vector<double> xs, std::vector<double> ys, std::vector<double> bs)
{
    const size_t n = xs.size();
    double sum_b = 0;
    for (size_t i = 0; i < n; ++i) {
        if (std::fabs(bs[i] - std::rint(bs[i])) > 1e-5) {
            error = "Burgers vector supposed to be an integer";
            return;
        }
        sum_b += bs[i];
    }
    if (sum_b != 0) {
        error = "the sum of the Burgers vectors supposed to be 0";
        return;
    }
    std::vector<unsigned> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::stable_sort(idx.begin(), idx.end(), [&](unsigned l, unsigned r) { return ys[l] + bs[l] > ys[r] + bs[r]; });
    dc = (unsigned)n;
    disl_order.resize(n);
    disl_sorted.clear();
    for (size_t i = 0; i < n; ++i) {
        disl_sorted.emplace_back(xs[idx[i]], ys[idx[i]]);
        disl_order[idx[i]] = (unsigned)i;
    }
    g.resize(n);
    initSpeed.resize(n);
    initSpeed2.resize(n);
    speed.resize(n);
    speed2.resize(n);
    dVec.resize(n);
    bigStep_sorted.resize(n);
    firstSmall_sorted.resize(n);
    secondSmall_sorted.resize(n);
    Ap.assign(n + 1, 0);
    indexes.resize(n);
    x.resize(n);
    updateCutOff();
}

void SimulationData::updateCutOff()
{
    double m = cutOffMultiplier;
    if (!heavisideCutoff && cutOffMultiplier >= std::sqrt(2.0 * dc) / 12) m = 1e20;
    cutOff = m / std::sqrt((double)dc);
    cutOffSqr = cutOff * cutOff;
    onePerCutOffSqr = 1 / cutOffSqr;
}

void SimulationData::writeDislocationDataToFile(const std::string& path) const
{
    std::ofstream of(path);
    if (!of) return;
    of << std::setprecision(14);
    for (auto id : disl_order) of << disl_sorted[id].x << "\t" << disl_sorted[id].y << "\t" << b(id) << "\n";
}

// ------------------------------------------------------------------ simulation

Simulation::Simulation(std::shared_ptr<SimulationData> s, d4_ctx* ctx) : sD(std::move(s)), pH(new PrecisionHandler), ctx_(ctx)
{
    pH->setMinPrecisity(sD->prec);  // simulation.cpp:24-26
    pH->setDipolePrecisity(sD->dipole_prec);
    pH->setSize(sD->dc);
    sD->tau.bind(ctx);
    if (sD->solverKind == 1) {
        lu_.reset(d4host::new_gpu_dense_lu(d4_ctx_device(ctx)));
        if (!lu_) {
            status_ = D4_ERR_HIP;
            error_ = "GPU dense LU unavailable";
        }
    }
    if (!lu_) lu_.reset(new d4host::SparseLU());
}

bool Simulation::check(int rc, const char* where)
{
    if (rc == D4_OK) return true;
    if (!status_) {
        status_ = rc;
        error_ = std::string(where) + ": " + d4_last_error(ctx_);
    }
    return false;
}

void Simulation::set_speed_job(const std::vector<DislwoB>& dis, std::vector<double>& forces, double simTime)
{
    job_ = PassJob();  // calculateSpeedsAtTime, simulation.cpp:287-333
    job_.kind = PassJob::Speed;
    job_.xy = &dis;
    job_.tau = sD->externalStressProtocol->extStress(simTime);
    job_.v = &forces;
}

void Simulation::set_jacobian_job(double stepsize, const std::vector<DislwoB>& dislocs, std::vector<double>& forces_A,
                                  std::vector<double>& forces_B, double simTime_A, double simTime_B)
{
    job_ = PassJob();  // calcJacobianAndSpeedsAtTimes, simulation.cpp:345-527
    job_.kind = PassJob::Jacobian;
    job_.xy = &dislocs;
    d4_jac_params& p = job_.jp;
    p.h = stepsize;
    p.tau_a = sD->externalStressProtocol->extStress(simTime_A);
    p.tau_b = sD->externalStressProtocol->extStress(simTime_B);
    p.cutoff = sD->cutOff;
    p.cut_sq = sD->cutOffSqr;
    p.one_per_cut_sq = sD->onePerCutOffSqr;
    p.heaviside = sD->heavisideCutoff ? 1 : 0;
    p.wfunc = sD->weightFunc;
    p.tol_cap2 = pH->toleranceCap2();
    job_.vA = &forces_A;
    job_.vB = &forces_B;
}

void Simulation::set_rescale_job(double baseTime, double newTime, double oldTime)
{
    job_ = PassJob();  // calcJacobianAndSpeedsFromPrev, simulation.cpp:546-608
    job_.kind = PassJob::Rescale;
    job_.factor = (newTime - baseTime) / (oldTime - baseTime);
    job_.dtau = sD->externalStressProtocol->extStress(newTime) - sD->externalStressProtocol->extStress(oldTime);
    job_.speed = &sD->speed;
}

int PassEngine::run(Simulation& sim, PassJob& job)
{
    // host wall time of the job, pair passes and linear solves apart
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = run_job(sim, job);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    (job.kind == PassJob::Factor || job.kind == PassJob::Solve ? sim.solveMs : sim.passMs) += ms;
    return rc;
}

int PassEngine::run_job(Simulation& sim, PassJob& job)
{
    SimulationData& d = sim.data();
    PrecisionHandler& pH = sim.precision();
    d4_ctx* ctx = sim.ctx();
    const int n = (int)d.dc;
    switch (job.kind) {
    case PassJob::Speed:
        sim.pairPasses++;
        return d4_speed_pass(ctx, n, &(*job.xy)[0].x, job.tau, job.v->data(), pH.toleranceData(), pH.toleranceCap2());
    case PassJob::Jacobian: {
        sim.pairPasses++;
        if (d.Ai.empty()) {
            d.Ai.resize((size_t)n * 16 + 16);
            d.Ax.resize(d.Ai.size());
        }
        int nnz = 0;
        int rc = d4_jacobian_pass(ctx, n, &(*job.xy)[0].x, &job.jp, job.vA->data(), job.vB->data(), d.Ap.data(),
                                  d.Ai.data(), d.Ax.data(), (int)d.Ai.size(), &nnz, d.indexes.data(), d.dVec.data(),
                                  pH.toleranceData());
        if (rc == D4_ERR_CAPACITY) {  // grow (simulation_data.cpp:157-175 grows by dc) and run the pass again
            d.Ai.resize((size_t)nnz + n);
            d.Ax.resize(d.Ai.size());
            rc = d4_jacobian_pass(ctx, n, &(*job.xy)[0].x, &job.jp, job.vA->data(), job.vB->data(), d.Ap.data(),
                                  d.Ai.data(), d.Ax.data(), (int)d.Ai.size(), &nnz, d.indexes.data(), d.dVec.data(),
                                  pH.toleranceData());
        }
        job.nnz = nnz;
        return rc;
    }
    case PassJob::Rescale:
        return d4_jacobian_rescale(ctx, n, d.Ap.data(), d.Ai.data(), d.Ax.data(), d.dVec.data(), d.weightFunc,
                                   job.factor, job.dtau, job.speed->data());
    case PassJob::Factor: {
        d4host::LinearSolver& lu = sim.solver();
        if (lu.factor(n, d.Ap.data(), d.Ai.data(), d.Ax.data()) < 0) return D4_ERR_HIP;
        job.singular = lu.singular();
        job.unpivoted = lu.unpivoted();
        return D4_OK;
    }
    case PassJob::Solve: return sim.solver().solve(job.rhs->data(), job.sol->data()) < 0 ? D4_ERR_HIP : D4_OK;
    default: return D4_OK;
    }
}

void Simulation::set_factor_job()  // simulation.cpp:635-641
{
    job_ = PassJob();
    job_.kind = PassJob::Factor;
    factorizations++;
}

void Simulation::after_factor()
{
    singular_ = job_.singular;
    if (job_.unpivoted) unpivotedFactorizations++;
}

void Simulation::set_solve_job()  // solveEQSys, simulation.cpp:643-659
{
    job_ = PassJob();
    job_.kind = PassJob::Solve;
    job_.rhs = &sD->g;
    job_.sol = &sD->x;
}

void Simulation::after_solve()  // simulation.cpp:651-657: the singular factorization's clamp
{
    if (!singular_) return;
    singularSolves++;
    for (unsigned i = 0; i < sD->dc; ++i)
        if (!((-1 < sD->x[i]) && (sD->x[i] < 1))) sD->x[i] = 0;
}

void Simulation::calcG(const std::vector<DislwoB>& new_disloc, const std::vector<DislwoB>& old_config,
                       double stepSize, const std::vector<double>& endSpeed, const std::vector<double>& initSpeed)
{
    for (unsigned i = 0; i < sD->dc; i++)  // simulation.cpp:620-624
        sD->g[i] = new_disloc[i].x - old_config[i].x -
                   stepSize * ((1 + sD->dVec[i]) * endSpeed[i] + (1 - sD->dVec[i]) * initSpeed[i]) / 2;
}

void Simulation::calcGFromSpeeds(double stepSize, const std::vector<double>& endSpeed,
                                 const std::vector<double>& initSpeed)
{
    for (unsigned i = 0; i < sD->dc; i++)  // simulation.cpp:102-105, 121-124, 142-145
        sD->g[i] = -stepSize * ((1 + sD->dVec[i]) * endSpeed[i] + (1 - sD->dVec[i]) * initSpeed[i]) / 2;
}

void Simulation::calculateXError()  // simulation.cpp:692-700
{
    for (unsigned i = 0; i < sD->dc; i++)
        pH->updateMaxErrorToleranceRatioSq(std::fabs(sD->bigStep_sorted[i].x - sD->secondSmall_sorted[i].x), i);
}

double Simulation::calculateStrainIncrement(const std::vector<DislwoB>& old, const std::vector<DislwoB>& newD) const
{
    double ret = 0;  // simulation.cpp:713-720
    for (size_t i = 0; i < old.size(); i++) ret += sD->b((unsigned)i) * (newD[i].x - old[i].x);
    return ret;
}

bool Simulation::interrupt() const  // simulation.cpp:722-750 (time, strain and step-count limits)
{
    if (sD->isTimeLimit && sD->simTime > sD->timeLimit) return true;
    if (sD->isStrainIncreaseLimit && sD->totalAccumulatedStrainIncrease > sD->totalAccumulatedStrainIncreaseLimit)
        return true;
    if (sD->isStepCountLimit && (int64_t)sD->succesfulSteps > sD->stepCountLimit) return true;
    return false;
}

void Simulation::begin_prologue()
{
    set_speed_job(sD->disl_sorted, sD->initSpeed, sD->simTime);  // simulation.cpp:46
    phase_ = -1;
}

bool Simulation::begin_step()  // simulation.cpp:88-94
{
    job_ = PassJob();
    lastAccepted_ = false;
    if (status_ || interrupt()) return false;
    SimulationData& d = *sD;
    t0_ = d.simTime;
    t1_ = d.simTime + d.stepSize;
    th_ = d.simTime + d.stepSize / 2;
    // stage I
    set_jacobian_job(d.stepSize, d.disl_sorted, d.initSpeed, d.speed, t0_, t1_);
    phase_ = 1;
    return true;
}

void Simulation::after_pass()
{
    SimulationData& d = *sD;
    const unsigned n = d.dc;
    const int done = phase_;
    if (!check(status_ ? status_ : D4_OK, "pass")) {  // an earlier failure ends the step
        job_ = PassJob();
        return;
    }
    // each stage: pass, Factor, Solve (first Newton iterate), speed pass, Solve
    // (second iterate, simulation.cpp:611-633), then the next stage's pass
    switch (done) {
    case -1:  // the initial speed pass
        job_ = PassJob();
        break;
    case 1:  // stage I, after the Jacobian pass (simulation.cpp:98-100)
    case 6:  // stage II, after the rescale (simulation.cpp:118-120)
    case 11:  // stage III, after the Jacobian pass (simulation.cpp:139-141)
        set_factor_job();
        break;
    case 2:  // simulation.cpp:101-105
        after_factor();
        calcGFromSpeeds(d.stepSize, d.speed, d.initSpeed);
        set_solve_job();
        break;
    case 3:  // simulation.cpp:106-109
        after_solve();
        for (unsigned i = 0; i < n; i++) d.bigStep_sorted[i] = DislwoB(d.disl_sorted[i].x - d.x[i], d.disl_sorted[i].y);
        set_speed_job(d.bigStep_sorted, d.speed2, t1_);
        break;
    case 4:  // simulation.cpp:111-113
        calcG(d.bigStep_sorted, d.disl_sorted, d.stepSize, d.speed2, d.initSpeed);
        set_solve_job();
        break;
    case 5:  // then stage II's rescale
        after_solve();
        for (unsigned i = 0; i < n; i++) d.bigStep_sorted[i].x -= d.x[i];
        set_rescale_job(t0_, th_, t1_);
        break;
    case 7:  // simulation.cpp:121-124
        after_factor();
        calcGFromSpeeds(d.stepSize / 2, d.speed, d.initSpeed);
        set_solve_job();
        break;
    case 8:  // simulation.cpp:125-130
        after_solve();
        for (unsigned i = 0; i < n; i++)
            d.firstSmall_sorted[i] = DislwoB(d.disl_sorted[i].x - d.x[i], d.disl_sorted[i].y);
        set_speed_job(d.firstSmall_sorted, d.speed2, th_);
        break;
    case 9:  // simulation.cpp:132-134
        calcG(d.firstSmall_sorted, d.disl_sorted, d.stepSize / 2, d.speed2, d.initSpeed);
        set_solve_job();
        break;
    case 10:  // then stage III's Jacobian pass
        after_solve();
        for (unsigned i = 0; i < n; i++) d.firstSmall_sorted[i].x -= d.x[i];
        set_jacobian_job(d.stepSize / 2, d.firstSmall_sorted, d.initSpeed2, d.speed2, th_, t1_);
        break;
    case 12:  // simulation.cpp:142-145
        after_factor();
        calcGFromSpeeds(d.stepSize / 2, d.speed2, d.initSpeed2);
        set_solve_job();
        break;
    case 13:  // simulation.cpp:146-150
        after_solve();
        for (unsigned i = 0; i < n; i++)
            d.secondSmall_sorted[i] = DislwoB(d.firstSmall_sorted[i].x - d.x[i], d.firstSmall_sorted[i].y);
        set_speed_job(d.secondSmall_sorted, d.speed2, t1_);
        break;
    case 14:  // simulation.cpp:152-154
        calcG(d.secondSmall_sorted, d.firstSmall_sorted, d.stepSize / 2, d.speed2, d.initSpeed2);
        set_solve_job();
        break;
    case 15:  // stage IV
        after_solve();
        for (unsigned i = 0; i < n; i++) d.secondSmall_sorted[i].x -= d.x[i];
        job_ = PassJob();
        lastAccepted_ = finish_step();
        break;
    default: job_ = PassJob(); break;
    }
    phase_ = done < 0 ? 0 : done + 1;
}

bool Simulation::finish_step()  // stage IV, simulation.cpp:160-276
{
    SimulationData& d = *sD;
    const unsigned n = d.dc;
    calculateXError();
    const bool accept = pH->getMaxErrorRatioSqr() < 1;
    if (accept) {
        d.simTime += d.stepSize;
        d.succesfulSteps++;
        if (d.calculateStrainDuringSimulation) {
            d.totalAccumulatedStrainIncrease += calculateStrainIncrement(d.disl_sorted, d.firstSmall_sorted);
            d.totalAccumulatedStrainIncrease += calculateStrainIncrement(d.firstSmall_sorted, d.secondSmall_sorted);
        }
        d.disl_sorted.swap(d.secondSmall_sorted);
        for (auto& p : d.disl_sorted) normalize(p.x);
        if (d.isSpeedThresholdForCutoffChange) {  // simulation.cpp:177, 215-219
            double avg = 0;
            for (double v : d.initSpeed2) avg += std::fabs(v);
            avg /= n;
            if (d.speedThresholdForCutoffChange > avg) {
                d.cutOffMultiplier = 1e20;
                d.updateCutOff();
            }
        }
    } else {
        d.failedSteps++;
    }
    lastRatio_ = pH->getMaxErrorRatioSqr();
    d.stepSize = pH->getNewStepSize(d.stepSize);  // simulation.cpp:253
    pH->reset();                                   // simulation.cpp:271 (no write-out times here)
    if (d.isMaxStepSizeLimit && d.maxStepSizeLimit < d.stepSize) d.stepSize = d.maxStepSizeLimit;
    return accept;
}

bool Simulation::step()  // one iteration of simulation.cpp:88-277
{
    if (!begin_step()) return false;
    while (job_.kind != PassJob::None) {
        check(default_engine_.run(*this, job_), job_.kind == PassJob::Factor  ? "linear solver factorization"
                                                : job_.kind == PassJob::Solve ? "linear solve"
                                                                              : "pair pass");
        after_pass();
    }
    return status_ ? false : lastAccepted_;
}

void Simulation::run_prologue()
{
    begin_prologue();
    check(default_engine_.run(*this, job_), "speed pass");
    after_pass();
}

void Simulation::run()
{
    run_prologue();
    while (!status_ && !interrupt()) step();
}

}  // namespace sdddstCore

// ------------------------------------------------------------------ C ABI

struct d4_sim {
    std::shared_ptr<sdddstCore::SimulationData> data;
    std::unique_ptr<sdddstCore::Simulation> sim;
    bool started = false;
    double last_ratio = 0;
    std::string err;
};

namespace sdddstCore {
// the parameters of a run as project_parser.cpp stores them in SimulationData
int apply_sim_params(SimulationData& d, const d4_sim_params* p)
{
    if (p->wfunc != 'c' && p->wfunc != 'p' && p->wfunc != 'm') return D4_ERR_INVALID;
    if (p->stress_kind < 0 || p->stress_kind > 3) return D4_ERR_INVALID;
    d.initStepSize = d.stepSize = p->init_step_size;
    d.isMaxStepSizeLimit = p->max_step_size > 0;
    d.maxStepSizeLimit = p->max_step_size;
    d.prec = p->prec;
    d.dipole_prec = p->dipole_prec;
    d.cutOffMultiplier = p->cutoff_mult;
    d.heavisideCutoff = p->heaviside != 0;
    d.weightFunc = p->wfunc;
    d.isTimeLimit = p->time_limit > 0;
    d.timeLimit = p->time_limit;
    d.isStepCountLimit = p->step_limit >= 0;
    d.stepCountLimit = p->step_limit;
    d.isStrainIncreaseLimit = p->strain_limit > 0;
    d.totalAccumulatedStrainIncreaseLimit = p->strain_limit;
    d.isSpeedThresholdForCutoffChange = p->speed_threshold > 0;
    d.speedThresholdForCutoffChange = p->speed_threshold;
    if (p->solver < 0 || p->solver > 1) return D4_ERR_INVALID;
    d.solverKind = p->solver;
    switch (p->stress_kind) {  // project_parser.cpp:287-326
    case 1: d.externalStressProtocol.reset(new FixedRateProtocol(p->stress_init, p->stress_rate)); break;
    case 2: d.externalStressProtocol.reset(new CyclicLoadProtocol(p->stress_init, p->stress_rate, p->stress_period)); break;
    case 3:
        d.externalStressProtocol.reset(
            new IncreasingCyclicLoadProtocol(p->stress_init, p->stress_rate, p->stress_period, p->stress_amp));
        break;
    default: d.externalStressProtocol.reset(new StressProtocol(p->stress_init)); break;
    }
    d.updateCutOff();
    return D4_OK;
}
}  // namespace sdddstCore

static int sim_setup(d4_ctx* ctx, std::shared_ptr<sdddstCore::SimulationData> sd, const d4_sim_params* p, d4_sim** out)
{
    if (!sd->error.empty()) return D4_ERR_INVALID;
    const int rc = sdddstCore::apply_sim_params(*sd, p);
    if (rc) return rc;
    d4_sim* s = new d4_sim();
    s->data = sd;
    s->sim.reset(new sdddstCore::Simulation(sd, ctx));
    *out = s;
    return D4_OK;
}

extern "C" {

void d4_sim_default_params(d4_sim_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->init_step_size = 1e-6;
    p->prec = 1e-5;
    p->dipole_prec = 0.05;
    p->cutoff_mult = 1e20;
    p->wfunc = 'c';
    p->step_limit = -1;
}

int d4_sim_create(d4_ctx* ctx, int n, const double* x, const double* y, const double* b, const d4_sim_params* p,
                  d4_sim** out)
{
    if (!ctx || !p || !out || n <= 0) return D4_ERR_INVALID;
    *out = nullptr;
    return sim_setup(ctx, std::make_shared<sdddstCore::SimulationData>(n, x, y, b), p, out);
}

int d4_sim_create_from_file(d4_ctx* ctx, const char* path, const d4_sim_params* p, d4_sim** out)
{
    if (!ctx || !p || !out || !path) return D4_ERR_INVALID;
    *out = nullptr;
    return sim_setup(ctx, std::make_shared<sdddstCore::SimulationData>(std::string(path)), p, out);
}

void d4_sim_destroy(d4_sim* s) { delete s; }

const char* d4_sim_last_error(const d4_sim* s)
{
    if (!s) return "null simulation";
    return s->sim->error().c_str();
}

int d4_sim_step(d4_sim* s, int* accepted, int* stop)
{
    if (!s) return D4_ERR_INVALID;
    sdddstCore::Simulation& sim = *s->sim;
    if (!s->started) {
        s->started = true;
        sim.run_prologue();
    }
    if (accepted) *accepted = 0;
    if (sim.interrupt()) {
        if (stop) *stop = 1;
        return sim.status();
    }
    if (stop) *stop = 0;
    const bool acc = sim.step();
    if (accepted) *accepted = acc ? 1 : 0;
    return sim.status();
}

int d4_sim_run(d4_sim* s)
{
    if (!s) return D4_ERR_INVALID;
    int stop = 0, acc = 0;
    while (true) {
        int rc = d4_sim_step(s, &acc, &stop);
        if (rc || stop) return rc;
    }
}

int d4_sim_positions(const d4_sim* s, double* xy, int original_order)
{
    if (!s || !xy) return D4_ERR_INVALID;
    const auto& d = *s->data;
    for (unsigned i = 0; i < d.dc; ++i) {
        const unsigned k = original_order ? d.disl_order[i] : i;
        xy[2 * i] = d.disl_sorted[k].x;
        xy[2 * i + 1] = d.disl_sorted[k].y;
    }
    return D4_OK;
}

int d4_sim_get_stats(const d4_sim* s, d4_sim_stats* st)
{
    if (!s || !st) return D4_ERR_INVALID;
    const auto& d = *s->data;
    st->sim_time = d.simTime;
    st->step_size = d.stepSize;
    st->strain = d.totalAccumulatedStrainIncrease;
    st->last_max_error_ratio_sq = s->sim->lastRatio();
    st->successful_steps = (int64_t)d.succesfulSteps;
    st->failed_steps = (int64_t)d.failedSteps;
    st->pair_passes = s->sim->pairPasses;
    st->factorizations = s->sim->factorizations;
    st->singular_solves = s->sim->singularSolves;
    st->unpivoted_factorizations = s->sim->unpivotedFactorizations;
    st->pass_ms = s->sim->passMs;
    st->solve_ms = s->sim->solveMs;
    return D4_OK;
}

int d4_sim_write_dconf(const d4_sim* s, const char* path)
{
    if (!s || !path) return D4_ERR_INVALID;
    s->data->writeDislocationDataToFile(path);
    return D4_OK;
}

}  // extern "C"
