// mph_explicit: command-line host driver with the reference's interface
//   mph_explicit data grid prof%03d vtk%03d log [nthreads] [-2d] [-mixed] [-graph]
// (source/main.cpp:290-452): reads .data/.grid, runs the time loop on the GPU through the
// C ABI, writes .prof and .vtk files at the same times and in the same layout.  The
// thread-count argument is accepted for compatibility and ignored (one GPU per process).
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <vector>

#include "mph.h"

static FILE* g_log = nullptr;

static void logf(const char* fmt, ...)
{
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    if (g_log) vfprintf(g_log, fmt, a);
    vfprintf(stderr, fmt, b);
    va_end(b);
    va_end(a);
}

static int die(const char* what)
{
    logf("error: %s: %s\n", what, mph_last_error());
    if (g_log) std::fclose(g_log);
    return 1;
}

int main(int argc, char** argv)
{
    std::string data = "sample.data", grid = "sample.grid", prof = "sample%03d.prof",
                vtk = "sample%03d.vtk", logname = "sample.log";
    int dim = 3, precision = MPH_FP64, use_graph = 0;
    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "-2d")) dim = 2;
        else if (!std::strcmp(argv[i], "-3d")) dim = 3;
        else if (!std::strcmp(argv[i], "-mixed")) precision = MPH_MIXED;
        else if (!std::strcmp(argv[i], "-graph")) use_graph = 1;
        else pos.push_back(argv[i]);
    }
    if (pos.size() > 0) data = pos[0];
    if (pos.size() > 1) grid = pos[1];
    if (pos.size() > 2) prof = pos[2];
    if (pos.size() > 3) vtk = pos[3];
    if (pos.size() > 4) logname = pos[4];
    g_log = std::fopen(logname.c_str(), "w");
    {
        time_t t = time(nullptr);
        logf("start reading files at %s\n", ctime(&t));
    }
    mph_params p;
    mph_params_default(&p);
    if (mph_read_data_file(data.c_str(), &p)) return die("reading data file");
    int n = 0;
    if (mph_read_grid_header(grid.c_str(), &p, &n)) return die("reading grid header");
    std::vector<int> type(n);
    std::vector<double> x(3 * (size_t)n), v(3 * (size_t)n);
    if (mph_read_grid_particles(grid.c_str(), n, type.data(), x.data(), v.data())) return die("reading grid");
    p.dim = dim;
    p.precision = precision;
    mph_ctx* ctx = mph_create(&p, n, type.data(), x.data(), v.data());
    if (!ctx) return die("creating the solver");
    std::vector<int> ft(n), nc(n);
    std::vector<double> fx(3 * (size_t)n), fv(3 * (size_t)n), ff(3 * (size_t)n), da(n), vs(n), vp(n, 0.0);
    mph_fields out;
    std::memset(&out, 0, sizeof(out));
    out.type = ft.data(); out.position = fx.data(); out.velocity = fv.data(); out.force = ff.data();
    out.density_a = da.data(); out.vol_strain_p = vs.data(); out.neighbor_count = nc.data();
    out.virial_pressure = vp.data();
    auto write_vtk = [&](const char* name, bool virial) -> int {
        if (virial && mph_compute_virial(ctx)) return die("virial");
        if (mph_get_fields(ctx, 0, MPH_ORDER_CURRENT, &out)) return die("fields");
        if (!virial) std::fill(vp.begin(), vp.end(), 0.0);
        if (mph_write_vtk_file(name, n, ft.data(), fx.data(), fv.data(), da.data(), vs.data(), vp.data(),
                               nc.data(), ff.data()))
            return die("writing vtk");
        return 0;
    };
    if (write_vtk("output.vtk", false)) return 1;       // main.cpp:365
    {
        time_t t = time(nullptr);
        logf("start main roop at %s\n", ctime(&t));
    }
    double Time = p.time, Dt = p.dt;
    double OutputNext = 0.0, VtkOutputNext = 0.0;
    int iStep = (int)(Time / Dt);                        // main.cpp:371
    const clock_t c0 = clock();
    while (Time < p.end_time + 1.0e-5 * Dt) {            // main.cpp:374
        if (Time + 1.0e-5 * Dt >= OutputNext) {          // main.cpp:376-382
            char fn[1024];
            std::snprintf(fn, sizeof(fn), prof.c_str(), iStep);
            if (mph_get_fields(ctx, 0, MPH_ORDER_CURRENT, &out)) return die("fields");
            double tnow = 0.0;
            mph_get_time(ctx, 0, &tnow);
            if (mph_write_prof_file(fn, &p, tnow, n, ft.data(), fx.data(), fv.data())) return die("writing prof");
            logf("@ Prof Output Time : %e\n", Time);
            OutputNext += p.output_interval;
        }
        // run steps until the next step that writes output: the reference's loop
        // condition and output tests use the same running Time (main.cpp:436, 449)
        int k = 0;
        double T = Time;
        bool vtk_due = false;
        while (T < p.end_time + 1.0e-5 * Dt) {
            ++k;
            if (T + 1.0e-5 * Dt >= VtkOutputNext) { vtk_due = true; break; }
            T += Dt;
            if (T + 1.0e-5 * Dt >= OutputNext) break;
        }
        if (k == 0) break;
        const int rc = use_graph ? mph_step_graph(ctx, k) : mph_step(ctx, k);
        if (rc) return die("step");
        for (int s = 0; s < k; ++s) {
            if (s == k - 1 && vtk_due) {
                char fn[1024];
                std::snprintf(fn, sizeof(fn), vtk.c_str(), iStep);
                if (write_vtk(fn, true)) return 1;
                logf("@ Vtk Output Time : %e\n", Time);
                VtkOutputNext += p.vtk_output_interval;
            }
            Time += Dt;
            ++iStep;
        }
    }
    mph_sync(ctx);
    const double sec = (double)(clock() - c0) / CLOCKS_PER_SEC;
    {
        time_t t = time(nullptr);
        logf("end main roop at %s\n", ctime(&t));
        logf("total:                   %lf [CPU sec]\n", sec);
    }
    mph_stats st;
    if (mph_get_stats(ctx, &st) == MPH_OK)
        logf("steps %lld, neighbour rebuilds %lld, list width %d\n", st.steps, st.rebuilds, st.list_width);
    mph_destroy(ctx);
    if (g_log) std::fclose(g_log);
    return 0;
}
