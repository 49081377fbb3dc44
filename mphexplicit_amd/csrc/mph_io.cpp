// Host-side file formats of the reference: .data (main.cpp:489-539), .grid
// (main.cpp:541-670), .prof (main.cpp:672-696) and legacy VTK (main.cpp:698-834).
// The keyword patterns and output layouts are format definitions and are kept identical
// to the reference's; the reader reports errors instead of exiting.
#include <cstdio>
#include <cstring>
#include <string>

#include "mph.h"
#include "mph_error.h"

namespace mph {
std::string& last_error()
{
    static thread_local std::string msg;
    return msg;
}
}  // namespace mph

using mph::fail;

extern "C" const char* mph_last_error(void) { return mph::last_error().c_str(); }

extern "C" void mph_params_default(mph_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->dim = 3;
    p->precision = MPH_FP64;
    p->dt = 1.0e100;   // main.cpp:64
}

extern "C" int mph_read_data_file(const char* path, mph_params* p)
{
    if (!path || !p) return fail(MPH_ERR_ARG, "mph_read_data_file: null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(MPH_ERR_IO, "cannot open data file '%s'", path);
    char buf[1024];
    double (*ir)[4] = p->interaction_ratio;
    while (!std::feof(fp) && !std::ferror(fp)) {
        if (std::fgets(buf, sizeof(buf), fp) == nullptr) continue;
        if (buf[0] == '#') continue;
        if (std::sscanf(buf, " Dt %lf", &p->dt) == 1) continue;
        if (std::sscanf(buf, " OutputInterval %lf", &p->output_interval) == 1) continue;
        if (std::sscanf(buf, " VtkOutputInterval %lf", &p->vtk_output_interval) == 1) continue;
        if (std::sscanf(buf, " EndTime %lf", &p->end_time) == 1) continue;
        if (std::sscanf(buf, " RadiusRatioA %lf", &p->radius_ratio_a) == 1) continue;
        if (std::sscanf(buf, " RadiusRatioP %lf", &p->radius_ratio_p) == 1) continue;
        if (std::sscanf(buf, " RadiusRatioV %lf", &p->radius_ratio_v) == 1) continue;
        double* d = p->density;
        if (std::sscanf(buf, " Density %lf %lf %lf %lf", d, d + 1, d + 2, d + 3) == 4) continue;
        d = p->bulk_modulus;
        if (std::sscanf(buf, " BulkModulus %lf %lf %lf %lf", d, d + 1, d + 2, d + 3) == 4) continue;
        d = p->bulk_viscosity;
        if (std::sscanf(buf, " BulkViscosity %lf %lf %lf %lf", d, d + 1, d + 2, d + 3) == 4) continue;
        d = p->shear_viscosity;
        if (std::sscanf(buf, " ShearViscosity %lf %lf %lf %lf", d, d + 1, d + 2, d + 3) == 4) continue;
        d = p->surface_tension;
        if (std::sscanf(buf, " SurfaceTension %lf %lf %lf %lf", d, d + 1, d + 2, d + 3) == 4) continue;
        bool hit = false;
        for (int t = 0; t < 4 && !hit; ++t) {
            char pat[64];
            std::snprintf(pat, sizeof(pat), " InteractionRatio(Type%d) %%lf %%lf %%lf %%lf", t);
            hit = std::sscanf(buf, pat, &ir[t][0], &ir[t][1], &ir[t][2], &ir[t][3]) == 4;
        }
        if (hit) continue;
        double* g = p->gravity;
        if (std::sscanf(buf, " Gravity %lf %lf %lf", g, g + 1, g + 2) == 3) continue;
        for (int w = 2; w < 4 && !hit; ++w) {
            char pat[128];
            std::snprintf(pat, sizeof(pat),
                          " Wall%d  Center %%lf %%lf %%lf Velocity %%lf %%lf %%lf Omega %%lf %%lf %%lf", w);
            double* c = p->wall_center[w];
            double* v = p->wall_velocity[w];
            double* o = p->wall_omega[w];
            hit = std::sscanf(buf, pat, c, c + 1, c + 2, v, v + 1, v + 2, o, o + 1, o + 2) == 9;
        }
        // any other line is ignored, as the reference does after logging it (main.cpp:522-524)
    }
    std::fclose(fp);
    return MPH_OK;
}

extern "C" int mph_read_grid_header(const char* path, mph_params* p, int* count)
{
    if (!path || !p || !count) return fail(MPH_ERR_ARG, "mph_read_grid_header: null argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(MPH_ERR_IO, "cannot open grid file '%s'", path);
    char buf[1024];
    int rc = MPH_OK;
    if (!std::fgets(buf, sizeof(buf), fp) || std::sscanf(buf, "%lf", &p->time) != 1) {
        rc = fail(MPH_ERR_FORMAT, "grid file '%s': missing time line", path);
    } else if (!std::fgets(buf, sizeof(buf), fp) ||
               std::sscanf(buf, "%d  %lf  %lf %lf %lf  %lf %lf %lf", count, &p->particle_spacing,
                           &p->domain_min[0], &p->domain_max[0], &p->domain_min[1],
                           &p->domain_max[1], &p->domain_min[2], &p->domain_max[2]) != 8) {
        rc = fail(MPH_ERR_FORMAT, "grid file '%s': malformed header line", path);
    }
    std::fclose(fp);
    return rc;
}

extern "C" int mph_read_grid_particles(const char* path, int count, int* type, double* x, double* v)
{
    if (!path || count < 0 || !type || !x || !v) return fail(MPH_ERR_ARG, "mph_read_grid_particles: bad argument");
    FILE* fp = std::fopen(path, "r");
    if (!fp) return fail(MPH_ERR_IO, "cannot open grid file '%s'", path);
    char buf[1024];
    int rc = MPH_OK;
    for (int k = 0; k < 2 && rc == MPH_OK; ++k)
        if (!std::fgets(buf, sizeof(buf), fp)) rc = fail(MPH_ERR_FORMAT, "grid file '%s': short header", path);
    for (int i = 0; i < count && rc == MPH_OK; ++i) {
        if (!std::fgets(buf, sizeof(buf), fp) ||
            std::sscanf(buf, "%d  %lf %lf %lf  %lf %lf %lf", &type[i], &x[3 * i], &x[3 * i + 1],
                        &x[3 * i + 2], &v[3 * i], &v[3 * i + 1], &v[3 * i + 2]) != 7)
            rc = fail(MPH_ERR_FORMAT, "grid file '%s': particle line %d missing or malformed", path, i);
    }
    std::fclose(fp);
    return rc;
}

extern "C" int mph_write_prof_file(const char* path, const mph_params* p, double time, int n,
                                   const int* type, const double* x, const double* v)
{
    if (!path || !p || n < 0 || (n > 0 && (!type || !x || !v)))
        return fail(MPH_ERR_ARG, "mph_write_prof_file: bad argument");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(MPH_ERR_IO, "cannot open '%s' for writing", path);
    std::fprintf(fp, "%e\n", time);
    std::fprintf(fp, "%d %e %e %e %e %e %e %e\n", n, p->particle_spacing, p->domain_min[0],
                 p->domain_max[0], p->domain_min[1], p->domain_max[1], p->domain_min[2],
                 p->domain_max[2]);
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%d %e %e %e  %e %e %e\n", type[i], x[3 * i], x[3 * i + 1], x[3 * i + 2],
                     v[3 * i], v[3 * i + 1], v[3 * i + 2]);
    std::fclose(fp);
    return MPH_OK;
}

extern "C" int mph_write_vtk_file(const char* path, int n, const int* type, const double* x,
                                  const double* v, const double* dens_a, const double* vstrain,
                                  const double* virial, const int* ncount, const double* force)
{
    if (!path || n < 0 || (n > 0 && (!type || !x || !v || !dens_a || !vstrain || !virial || !ncount || !force)))
        return fail(MPH_ERR_ARG, "mph_write_vtk_file: bad argument");
    FILE* fp = std::fopen(path, "w");
    if (!fp) return fail(MPH_ERR_IO, "cannot open '%s' for writing", path);
    std::fprintf(fp, "# vtk DataFile Version 2.0\n");
    std::fprintf(fp, "Unstructured Grid Example\n");
    std::fprintf(fp, "ASCII\n");
    std::fprintf(fp, "DATASET UNSTRUCTURED_GRID\n");
    std::fprintf(fp, "POINTS %d float\n", n);
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)x[3 * i], (float)x[3 * i + 1], (float)x[3 * i + 2]);
    std::fprintf(fp, "CELLS %d %d\n", n, 2 * n);
    for (int i = 0; i < n; ++i) std::fprintf(fp, "1 %d ", i);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "CELL_TYPES %d\n", n);
    for (int i = 0; i < n; ++i) std::fprintf(fp, "1 ");
    std::fprintf(fp, "\n");
    std::fprintf(fp, "\n");
    std::fprintf(fp, "POINT_DATA %d\n", n);
    std::fprintf(fp, "SCALARS label float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%d\n", type[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS DensityA float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%e\n", (float)dens_a[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS VolStrainP float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%e\n", (float)vstrain[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS VirialPressureAtParticle float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%e\n", (float)virial[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "SCALARS neighbor float 1\n");
    std::fprintf(fp, "LOOKUP_TABLE default\n");
    for (int i = 0; i < n; ++i) std::fprintf(fp, "%d\n", ncount[i]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "VECTORS velocity float\n");
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)v[3 * i], (float)v[3 * i + 1], (float)v[3 * i + 2]);
    std::fprintf(fp, "\n");
    std::fprintf(fp, "VECTORS force float\n");
    for (int i = 0; i < n; ++i)
        std::fprintf(fp, "%e %e %e\n", (float)force[3 * i], (float)force[3 * i + 1], (float)force[3 * i + 2]);
    std::fprintf(fp, "\n");
    std::fclose(fp);
    return MPH_OK;
}
