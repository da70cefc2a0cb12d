// capi_driver.cpp -- a compiled C++ caller of lib2d4.so that sees only the
// C ABI (include/d4.h, include/d4_sim.h): no Python, no torch.  It does what
// the reference's program would do after the drop-in patch, and the GPU tests
// compare its outputs with the same calls made through the Python binding.
//
//   capi_driver sim <in.dconf> <out.dconf> <out.bin> [key=value ...]
//     d4_sim_create_from_file, d4_sim_step until the loop stops (the
//     reference's Simulation::run, simulation.cpp:30-281), d4_sim_write_dconf;
//     out.bin holds the n {x, y} of d4_sim_positions (sorted order).  Keys are
//     d4_sim_params fields (init_step_size, max_step_size, prec, dipole_prec,
//     cutoff_mult, heaviside, step_limit, solver).  Prints one line of stats.
//
//   capi_driver passes <in.dconf> <out.bin> <cutoff> <heaviside> <h> <tau_a> <tau_b>
//     The configuration sorted as SimulationData does (through a d4_sim), then
//     d4_speed_pass at tau_a and d4_jacobian_pass with the reference's storage
//     growth (simulation_data.cpp:157-175: start at n entries, grow by n while
//     the pass answers D4_ERR_CAPACITY).  out.bin: int32 n, nnz, attempts;
//     doubles xy[2n], v[n], vA[n], vB[n], dvec[n]; int32 Ap[n+1], indexes[n],
//     Ai[nnz]; doubles Ax[nnz].
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "d4_sim.h"

namespace {

int fail(const char* what, int rc, const char* msg)
{
    std::fprintf(stderr, "capi_driver: %s failed (%d): %s\n", what, rc, msg ? msg : "");
    return 1;
}

bool set_param(d4_sim_params& p, const char* kv)
{
    const char* eq = std::strchr(kv, '=');
    if (!eq) return false;
    const std::string k(kv, eq - kv);
    const double v = std::strtod(eq + 1, nullptr);
    if (k == "init_step_size") p.init_step_size = v;
    else if (k == "max_step_size") p.max_step_size = v;
    else if (k == "prec") p.prec = v;
    else if (k == "dipole_prec") p.dipole_prec = v;
    else if (k == "cutoff_mult") p.cutoff_mult = v;
    else if (k == "heaviside") p.heaviside = (int)v;
    else if (k == "step_limit") p.step_limit = (int64_t)v;
    else if (k == "solver") p.solver = (int)v;
    else return false;
    return true;
}

template <class T>
void put(FILE* f, const T* a, size_t count)
{
    if (count) std::fwrite(a, sizeof(T), count, f);
}

int run_sim(d4_ctx* ctx, int argc, char** argv)
{
    d4_sim_params p;
    d4_sim_default_params(&p);
    for (int i = 5; i < argc; ++i)
        if (!set_param(p, argv[i])) return fail("parameter", D4_ERR_INVALID, argv[i]);
    d4_sim* s = nullptr;
    int rc = d4_sim_create_from_file(ctx, argv[2], &p, &s);
    if (rc) return fail("d4_sim_create_from_file", rc, "unreadable file or not a configuration");
    for (;;) {
        int accepted = 0, stop = 0;
        rc = d4_sim_step(s, &accepted, &stop);
        if (rc) return fail("d4_sim_step", rc, d4_sim_last_error(s));
        if (stop) break;
    }
    rc = d4_sim_write_dconf(s, argv[3]);
    if (rc) return fail("d4_sim_write_dconf", rc, d4_sim_last_error(s));
    d4_sim_stats st;
    d4_sim_get_stats(s, &st);
    int n = 0;
    {
        // the file's line count is the configuration's size
        FILE* f = std::fopen(argv[2], "r");
        double x, y, b;
        while (f && std::fscanf(f, "%lf %lf %lf", &x, &y, &b) == 3) ++n;
        if (f) std::fclose(f);
    }
    std::vector<double> xy(2 * (size_t)n);
    rc = d4_sim_positions(s, xy.data(), 0);
    if (rc) return fail("d4_sim_positions", rc, d4_sim_last_error(s));
    FILE* o = std::fopen(argv[4], "wb");
    if (!o) return fail("open", D4_ERR_INVALID, argv[4]);
    put(o, xy.data(), xy.size());
    std::fclose(o);
    std::printf("n %d successful %lld failed %lld passes %lld factorizations %lld sim_time %.17g\n", n,
                (long long)st.successful_steps, (long long)st.failed_steps, (long long)st.pair_passes,
                (long long)st.factorizations, st.sim_time);
    d4_sim_destroy(s);
    return 0;
}

int run_passes(d4_ctx* ctx, char** argv)
{
    std::vector<double> fx, fy, fb;
    {
        FILE* f = std::fopen(argv[2], "r");
        if (!f) return fail("open", D4_ERR_INVALID, argv[2]);
        double x, y, b;
        while (std::fscanf(f, "%lf %lf %lf", &x, &y, &b) == 3) {
            fx.push_back(x);
            fy.push_back(y);
            fb.push_back(b);
        }
        std::fclose(f);
    }
    const int n = (int)fx.size();
    d4_sim_params sp;
    d4_sim_default_params(&sp);
    d4_sim* s = nullptr;
    int rc = d4_sim_create(ctx, n, fx.data(), fy.data(), fb.data(), &sp, &s);
    if (rc) return fail("d4_sim_create", rc, "not a configuration (integer Burgers vectors summing to 0)");
    std::vector<double> xy(2 * (size_t)n);
    rc = d4_sim_positions(s, xy.data(), 0);
    d4_sim_destroy(s);
    if (rc) return fail("d4_sim_positions", rc, "");

    d4_jac_params jp{};
    jp.cutoff = std::strtod(argv[4], nullptr);
    jp.cut_sq = jp.cutoff * jp.cutoff;
    jp.one_per_cut_sq = 1.0 / jp.cut_sq;
    jp.heaviside = std::atoi(argv[5]);
    jp.h = std::strtod(argv[6], nullptr);
    jp.tau_a = std::strtod(argv[7], nullptr);
    jp.tau_b = std::strtod(argv[8], nullptr);
    jp.wfunc = 'c';

    std::vector<double> v(n), vA(n), vB(n), dvec(n);
    rc = d4_speed_pass(ctx, n, xy.data(), jp.tau_a, v.data(), nullptr, 0.0);
    if (rc) return fail("d4_speed_pass", rc, d4_last_error(ctx));

    // SimulationData's storage: n entries at first, grown by n at a time
    std::vector<int> Ap(n + 1), idx(n);
    size_t cap = (size_t)n;
    std::vector<int> Ai(cap);
    std::vector<double> Ax(cap);
    int nnz = 0, attempts = 0;
    for (;;) {
        ++attempts;
        rc = d4_jacobian_pass(ctx, n, xy.data(), &jp, vA.data(), vB.data(), Ap.data(), Ai.data(), Ax.data(),
                              (int)cap, &nnz, idx.data(), dvec.data(), nullptr);
        if (rc != D4_ERR_CAPACITY) break;
        while (cap < (size_t)nnz + (size_t)n) cap += (size_t)n;
        Ai.resize(cap);
        Ax.resize(cap);
    }
    if (rc) return fail("d4_jacobian_pass", rc, d4_last_error(ctx));

    FILE* o = std::fopen(argv[3], "wb");
    if (!o) return fail("open", D4_ERR_INVALID, argv[3]);
    put(o, &n, 1);
    put(o, &nnz, 1);
    put(o, &attempts, 1);
    put(o, xy.data(), xy.size());
    put(o, v.data(), (size_t)n);
    put(o, vA.data(), (size_t)n);
    put(o, vB.data(), (size_t)n);
    put(o, dvec.data(), (size_t)n);
    put(o, Ap.data(), (size_t)n + 1);
    put(o, idx.data(), (size_t)n);
    put(o, Ai.data(), (size_t)nnz);
    put(o, Ax.data(), (size_t)nnz);
    std::fclose(o);
    std::printf("n %d nnz %d attempts %d\n", n, nnz, attempts);
    return 0;
}

}  // namespace

int main(int argc, char** argv)
{
    const bool sim = argc >= 5 && std::strcmp(argv[1], "sim") == 0;
    const bool passes = argc == 9 && std::strcmp(argv[1], "passes") == 0;
    if (!sim && !passes) {
        std::fprintf(stderr,
                     "usage: capi_driver sim <in.dconf> <out.dconf> <out.bin> [key=value ...]\n"
                     "       capi_driver passes <in.dconf> <out.bin> <cutoff> <heaviside> <h> <tau_a> <tau_b>\n");
        return 2;
    }
    d4_ctx* ctx = nullptr;
    int rc = d4_ctx_create(0, &ctx);
    if (rc) return fail("d4_ctx_create", rc, "no usable device 0");
    const int r = sim ? run_sim(ctx, argc, argv) : run_passes(ctx, argv);
    d4_ctx_destroy(ctx);
    return r;
}
