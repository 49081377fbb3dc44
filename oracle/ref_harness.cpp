// Test infrastructure only: builds the reference solver's own source (source/main.cpp,
// compiled unmodified where it lies) into a shared library so that the oracle and the
// HIP path can be checked against it.  Nothing in the product links or loads this file.
//
// The reference has no library API: all state is file-scope statics of main.cpp and the
// step is the call sequence inside main()'s loop (source/main.cpp:386-431).  This harness
// includes main.cpp into the same translation unit (renaming its main), so the static
// functions are callable, and exposes a small extern "C" surface:
//   mphref_load   = main.cpp:324-364 (read .data/.grid, initialise, first bucketing and
//                   densities; the "output.vtk" write at :365 is skipped)
//   mphref_step   = main.cpp:386-431 + :449 (one loop body, Time += Dt), n times
//   mphref_get    = copies of the per-particle arrays (current, i.e. sorted, order)
//   mphref_neighbors = NeighborPtr / NeighborInd copies
//   mphref_virial = calculateVirialStressAtParticle (main.cpp:2057-2302)
//
// Build: oracle/Makefile (g++ -O3 -fopenmp, the flags of source/makefile:43).

#define main mphref_reference_main
#include "main.cpp"
#undef main

#include <cstring>

#define MPHREF_API extern "C" __attribute__((visibility("default")))

static int g_last_rebuild = 0;

static void reset_globals()
{
    // Values the statics hold at program start (main.cpp:56-211), so that a second load
    // in the same process behaves like a fresh run.
    ParticleSpacing = 0.0; ParticleVolume = 0.0; OutputInterval = 0.0; OutputNext = 0.0;
    VtkOutputInterval = 0.0; VtkOutputNext = 0.0; EndTime = 0.0; Time = 0.0; Dt = 1.0e100;
    for (int d = 0; d < DIM; ++d) { DomainMin[d] = DomainMax[d] = DomainWidth[d] = 0.0; Gravity[d] = 0.0; }
    for (int t = 0; t < TYPE_COUNT; ++t) {
        Density[t] = BulkModulus[t] = BulkViscosity[t] = ShearViscosity[t] = SurfaceTension[t] = 0.0;
        CofA[t] = 0.0;
        for (int u = 0; u < TYPE_COUNT; ++u) InteractionRatio[t][u] = 0.0;
    }
    for (int w = 0; w < WALL_END; ++w)
        for (int d = 0; d < DIM; ++d) { WallCenter[w][d] = WallVelocity[w][d] = WallOmega[w][d] = 0.0; }
    CofK = 0.0;
    RadiusRatioA = RadiusRatioG = RadiusRatioP = RadiusRatioV = 0.0;
    MaxRadius = RadiusA = RadiusG = RadiusP = RadiusV = 0.0;
    Swa = Swg = Swp = Swv = 1.0; N0a = N0p = R2g = 1.0;
    TotalCellCount = 0;
    NeighborInd = NULL;
    NeighborIndCount = 0;
}

MPHREF_API int mphref_dim(void) {
#ifdef TWO_DIMENSIONAL
    return 2;
#else
    return 3;
#endif
}

MPHREF_API int mphref_load(const char* data, const char* grid, const char* logfile, int nthreads)
{
    reset_globals();
    char d[1024], g[1024];
    strncpy(d, data, sizeof(d) - 1); d[sizeof(d) - 1] = 0;
    strncpy(g, grid, sizeof(g) - 1); g[sizeof(g) - 1] = 0;
    log_open(logfile ? logfile : "/dev/null");
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#endif
    readDataFile(d);
    readGridFile(g);
    initializeWeight();
    initializeFluid();
    initializeWall();
    initializeDomain();
    calculateCellParticle();
    calculateNeighbor();
    calculateDensityA();
    calculateGravityCenter();
    calculateDensityP();
    // DivergenceP, PressureP, PressureA, Force are not computed before the first step in
    // the reference; zero them so copies are deterministic.
    for (int i = 0; i < ParticleCount; ++i) {
        DivergenceP[i] = PressureP[i] = PressureA[i] = 0.0;
        for (int k = 0; k < DIM; ++k) Force[i][k] = 0.0;
    }
    g_last_rebuild = 1;
    return ParticleCount;
}

MPHREF_API int mphref_step(int nsteps)
{
    int rebuilds = 0;
    for (int s = 0; s < nsteps; ++s) {
        calculateConvection();
        calculateWall();
        calculatePeriodicBoundary();
        resetForce();
        g_last_rebuild = 0;
        if (neighborCalculation() == 1) {
            freeNeighbor();
            calculateCellParticle();
            calculateNeighbor();
            g_last_rebuild = 1;
            ++rebuilds;
        }
        calculateDensityA();
        calculateGravityCenter();
        calculateDensityP();
        calculateDivergenceP();
        calculatePhysicalCoefficients();
        calculatePressureP();
        calculatePressureA();
        calculateDiffuseInterface();
        calculateViscosityV();
        calculateGravity();
        calculateAcceleration();
        Time += Dt;
    }
    return rebuilds;
}

MPHREF_API int mphref_last_rebuild(void) { return g_last_rebuild; }
MPHREF_API int mphref_count(void) { return ParticleCount; }
MPHREF_API double mphref_time(void) { return Time; }
MPHREF_API double mphref_dt(void) { return Dt; }
MPHREF_API long mphref_neighbor_total(void) { return NeighborIndCount; }

MPHREF_API void mphref_scalars(double* out /* [16] */)
{
    out[0] = N0a; out[1] = N0p; out[2] = R2g; out[3] = MaxRadius;
    out[4] = CellWidth[0]; out[5] = CellWidth[1]; out[6] = CellWidth[2];
    out[7] = CellCount[0]; out[8] = CellCount[1]; out[9] = CellCount[2];
    out[10] = Swa; out[11] = Swg; out[12] = CofK; out[13] = CofA[0];
    out[14] = FluidParticleEnd; out[15] = WallParticleEnd;
}

// Any pointer may be NULL.  Vectors are [n][3].
MPHREF_API void mphref_get(int* index, int* type, double* x, double* v, double* force,
                           double* dens_a, double* vol_strain, double* divergence,
                           double* press_p, double* press_a, double* gcenter, int* ncount,
                           double* mass)
{
    const int n = ParticleCount;
    if (index) memcpy(index, ParticleIndex, sizeof(int) * n);
    if (type) memcpy(type, Property, sizeof(int) * n);
    if (x) memcpy(x, Position, sizeof(double) * 3 * n);
    if (v) memcpy(v, Velocity, sizeof(double) * 3 * n);
    if (force) memcpy(force, Force, sizeof(double) * 3 * n);
    if (dens_a) memcpy(dens_a, DensityA, sizeof(double) * n);
    if (vol_strain) memcpy(vol_strain, VolStrainP, sizeof(double) * n);
    if (divergence) memcpy(divergence, DivergenceP, sizeof(double) * n);
    if (press_p) memcpy(press_p, PressureP, sizeof(double) * n);
    if (press_a) memcpy(press_a, PressureA, sizeof(double) * n);
    if (gcenter) memcpy(gcenter, GravityCenter, sizeof(double) * 3 * n);
    if (ncount) memcpy(ncount, NeighborCount, sizeof(int) * n);
    if (mass) memcpy(mass, Mass, sizeof(double) * n);
}

MPHREF_API void mphref_neighbors(long* ptr /* [n+1] */, int* ind /* [total] */)
{
    for (int i = 0; i <= ParticleCount; ++i) ptr[i] = NeighborPtr[i];
    memcpy(ind, NeighborInd, sizeof(int) * NeighborIndCount);
}

MPHREF_API void mphref_virial(double* vp /* [n] */)
{
    calculateVirialStressAtParticle();
    memcpy(vp, VirialPressureAtParticle, sizeof(double) * ParticleCount);
}

// the reference's own main() (main.cpp:290-470): reads the files named on its command line,
// runs the whole time loop and writes its .prof/.vtk/log outputs in the working directory
MPHREF_API int mphref_main(int argc, char** argv)
{
    reset_globals();
    return mphref_reference_main(argc, argv);
}
