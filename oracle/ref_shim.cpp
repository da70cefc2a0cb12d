// ref_shim.cpp -- extern "C" entry points over the reference's own compiled
// classes, for checking the oracle.  TEST INFRASTRUCTURE ONLY.
//
// oracle/Makefile compiles this file together with the reference sources where
// they lie (/root/reference/src/{Fields/Field,sim_utils,precision_handler,
// simulation_data,stress_protocol,dislocation}.cpp) into oracle/_ref/.  Nothing
// here restates the reference: every call goes to the reference's code.
// src/simulation.cpp is not part of the build: it includes
// <suitesparse/umfpack.h> (simulation.cpp:6), which this image lacks.
#include "Fields/Field.h"
#include "precision_handler.h"
#include "sim_utils.h"
#include "simulation_data.h"

#include <cmath>
#include <memory>
#include <string>
#include <vector>

using sdddstCore::Field;

extern "C" {

// Field::xy / xy_diff_x, 2-arg and 4-arg (include/Fields/Field.h:40-49)
double ref_xy(double dx, double dy) { return Field().xy(dx, dy); }
double ref_dxy(double dx, double dy) { return Field().xy_diff_x(dx, dy); }
double ref_xy4(double dx, double dy, double e, double c) { return Field().xy(dx, dy, e, c); }
double ref_dxy4(double dx, double dy, double e, double c) { return Field().xy_diff_x(dx, dy, e, c); }

// the two transcendental inputs exactly as the Jacobian pass forms them
// (simulation.cpp:456-457), evaluated with this library's libm
void ref_inputs(double dx, double dy, double* e, double* c)
{
    *e = exp2(PI2LN2INV * dx);
    *c = cos(2 * M_PI * dy);
}

void ref_field_batch(long n, const double* dx, const double* dy, const double* e, const double* c,
                     double* xy2, double* dxy2, double* xy4, double* dxy4)
{
    Field f;
    for (long k = 0; k < n; ++k) {
        if (xy2) xy2[k] = f.xy(dx[k], dy[k]);
        if (dxy2) dxy2[k] = f.xy_diff_x(dx[k], dy[k]);
        if (xy4) xy4[k] = f.xy(dx[k], dy[k], e[k], c[k]);
        if (dxy4) dxy4[k] = f.xy_diff_x(dx[k], dy[k], e[k], c[k]);
    }
}

// sim_utils.cpp:6-13, 36-75
double ref_normalize(double v) { normalize(v); return v; }
double ref_weight(double s, char t) { return weight(s, t); }
double ref_weight_inv(double w, char t) { return weightInv(w, t); }

// PrecisionHandler (precision_handler.cpp): replays a sequence of tolerance
// updates for `size` particles and reports the final tolerance vector via a
// max-error probe of unit error per particle.
int ref_tolerance_replay(double prec, double dipole, unsigned size, long m, const double* d2, const unsigned* id,
                         double* ratio_out)
{
    sdddstCore::PrecisionHandler ph;
    ph.setMinPrecisity(prec);
    ph.setDipolePrecisity(dipole);
    ph.setSize(size);
    for (long k = 0; k < m; ++k) ph.updateTolerance(d2[k], id[k]);
    // error^2 / tolerance for error = 1, one particle at a time
    for (unsigned i = 0; i < size; ++i) {
        sdddstCore::PrecisionHandler probe = ph;
        probe.updateMaxErrorToleranceRatioSq(1.0, i);
        ratio_out[i] = probe.getMaxErrorRatioSqr();
    }
    return 0;
}

double ref_new_step_size(double prec, double dipole, unsigned size, double err, double old_h)
{
    sdddstCore::PrecisionHandler ph;
    ph.setMinPrecisity(prec);
    ph.setDipolePrecisity(dipole);
    ph.setSize(size);
    ph.updateMaxErrorToleranceRatioSq(err, 0);
    return ph.getNewStepSize(old_h);
}

// SimulationData: read + sort a .dconf (simulation_data.cpp:76-154), cutoff
// (simulation_data.cpp:253-262) and b(ID) (simulation_data.cpp:265-270).
int ref_read_dconf(const char* path, double mult, int heaviside, double* xy, unsigned* order, int* b, int cap,
                   double* cutoff3)
{
    sdddstCore::SimulationData sd(path, "");
    if ((int)sd.dc > cap) return -1;
    sd.heavisideCutoff = heaviside != 0;
    sd.cutOffMultiplier = mult;
    sd.updateCutOff();
    for (unsigned i = 0; i < sd.dc; ++i) {
        xy[2 * i] = sd.disl_sorted[i].x;
        xy[2 * i + 1] = sd.disl_sorted[i].y;
        order[i] = sd.disl_order[i];
        b[i] = sd.b(i);
    }
    cutoff3[0] = sd.cutOff;
    cutoff3[1] = sd.cutOffSqr;
    cutoff3[2] = sd.onePerCutOffSqr;
    return (int)sd.dc;
}

// writeDislocationDataToFile (simulation_data.cpp:177-191)
int ref_write_dconf(const char* in_path, const double* xy_sorted, const char* out_path)
{
    sdddstCore::SimulationData sd(in_path, "");
    for (unsigned i = 0; i < sd.dc; ++i) {
        sd.disl_sorted[i].x = xy_sorted[2 * i];
        sd.disl_sorted[i].y = xy_sorted[2 * i + 1];
    }
    sd.writeDislocationDataToFile(out_path);
    return (int)sd.dc;
}

}  // extern "C"

#include "stress_protocol.h"

extern "C" {

// stress_protocol.cpp:13-73, via the reference classes; kind as orc_ext_stress
double ref_ext_stress(int kind, double init, double rate, double period, double amp, double t)
{
    std::unique_ptr<sdddstCore::StressProtocol> p;
    if (kind == 1) p.reset(new sdddstCore::FixedRateProtocol(init, rate));
    else if (kind == 2) p.reset(new sdddstCore::CyclicLoadProtocol(init, rate, period));
    else if (kind == 3) p.reset(new sdddstCore::IncreasingCyclicLoadProtocol(init, rate, period, amp));
    else p.reset(new sdddstCore::StressProtocol(init));
    return p->extStress(t);
}

}  // extern "C"
