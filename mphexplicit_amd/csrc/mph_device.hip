// Device context and C ABI of the MPH-WC step on gfx950 (declarations in include/mph.h).
//
// Set-up follows the reference's initialise* functions (main.cpp:836-1084) on the host
// with the same arithmetic, so every constant a kernel sees equals the reference's.  A
// step is a fixed sequence of launches on one stream:
//   move -> [keys -> radix sort -> cell starts -> gather -> scatter back -> lists] -> A -> B
// The bracketed rebuild launches exit at once unless the move kernel raised the Verlet
// flag, so the host never waits inside a step and the sequence can be captured in a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <memory>
#include <vector>

#include "mph.h"
#include "mph_case.h"
#include "mph_error.h"
#include "mph_kernels.h"

using namespace mph;

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t _e = (expr);                                                         \
        if (_e != hipSuccess)                                                           \
            return fail(MPH_ERR_DEVICE, "%s failed: %s", #expr, hipGetErrorString(_e)); \
    } while (0)

namespace {

constexpr int kBlock = 256;   // particle kernels (block maps are built for this size)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    int alloc(size_t count)
    {
        release();
        if (count == 0) count = 1;
        hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            return fail(MPH_ERR_DEVICE, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
        }
        n = count;
        return MPH_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

struct CaseHost {
    mph_params prm;
    CaseConst k;
    double time;
    double wall_center0[kTypes][3];
};

}  // namespace

struct mph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncases = 0, n = 0, stride = 0, width = 0;
    int nkeys = 0, sort_passes = 0, sort_blocks = 0, any_surface = 0;
    long long steps = 0;
    std::vector<CaseHost> cases;
    std::vector<int2> bmap_host;
    DevBuf<CaseConst> d_cases;
    DevBuf<int2> d_bmap;
    DevBuf<double4> pos, pos_alt, vel, vel_alt, vnext, aux, force, snap;
    DevBuf<int> pidx, pidx_alt, ncnt, nout, nbr, start;
    DevBuf<double> dens_a, vstrain, divp, wallc, vir;
    DevBuf<unsigned> keys_a, keys_b, hist;
    DevBuf<int> perm_a, perm_b, rank, cell_of_rank;
    DevBuf<int> flags, any_flag, max_len;
    DevBuf<long long> rebuilds;
    DevBuf<unsigned long long> pairs, pairs_scratch;
    DevBuf<double4> tune_v, tune_f;      // sweep B outputs while the autotuner times variants
    int variant_a = 1, variant_b = 1;    // indices into the sweep kernel variant tables
    double4* vnext_out(int real) { return real ? vnext.p : tune_v.p; }
    double4* force_out(int real) { return real ? force.p : tune_f.p; }
    int* h_scalar = nullptr;   // pinned host word for small readbacks
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;

    ~mph_ctx()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_src) (void)hipGraphDestroy(graph_src);
        if (h_scalar) (void)hipHostFree(h_scalar);
        if (stream) (void)hipStreamDestroy(stream);
    }

    unsigned* sorted_keys() { return (sort_passes & 1) ? keys_b.p : keys_a.p; }
    int* sorted_perm() { return (sort_passes & 1) ? perm_b.p : perm_a.p; }
    int nblocks() const { return (int)bmap_host.size(); }
};

// ---------------------------------------------------------------------------------------
// host-side constants, main.cpp:836-1084

static double hd(int dim, double h) { return dim == 2 ? h * h : h * h * h; }

static double lattice_sum(int dim, double l0, double radius, double sw, bool attractive)
{
    // main.cpp:861-949: sum of the kernel over lattice neighbours within the radius
    const double ratio = radius / l0;
    const int range = (int)(ratio + 3.0);
    const int zr = dim == 2 ? 0 : range;
    double sum = 0.0;
    for (int ix = -range; ix <= range; ++ix)
        for (int iy = -range; iy <= range; ++iy)
            for (int iz = -zr; iz <= zr; ++iz) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                const double x = l0 * ((double)ix), y = l0 * ((double)iy), z = l0 * ((double)iz);
                const double r2 = dim == 2 ? x * x + y * y : x * x + y * y + z * z;
                if (r2 <= radius * radius) {
                    const double r = std::sqrt(r2);
                    const double w = attractive
                        ? 1.0 / sw * 1.0 / hd(dim, radius) * (r / radius) * (1.0 - (r / radius)) * (1.0 - (r / radius))
                        : 1.0 / sw * 1.0 / hd(dim, radius) * ((1.0 - r / radius) * (1.0 - r / radius));
                    sum += w;
                }
            }
    return sum;
}

// Morton order of a case's cells: rank[cell id] and its inverse
static void morton_ranks(const CaseConst& k, std::vector<int>& rank, std::vector<int>& cell_of_rank)
{
    auto spread = [](unsigned long long v) {
        unsigned long long x = v & 0x1fffffull;
        x = (x | x << 32) & 0x1f00000000ffffull;
        x = (x | x << 16) & 0x1f0000ff0000ffull;
        x = (x | x << 8) & 0x100f00f00f00f00full;
        x = (x | x << 4) & 0x10c30c30c30c30c3ull;
        x = (x | x << 2) & 0x1249249249249249ull;
        return x;
    };
    const int T = k.tcc;
    std::vector<std::pair<unsigned long long, int>> code(T);
    for (int ix = 0, id = 0; ix < k.cc[0]; ++ix)
        for (int iy = 0; iy < k.cc[1]; ++iy)
            for (int iz = 0; iz < k.cc[2]; ++iz, ++id)
                code[id] = {spread(ix) << 2 | spread(iy) << 1 | spread(iz), id};
    std::sort(code.begin(), code.end());
    rank.resize(T);
    cell_of_rank.resize(T);
    for (int r = 0; r < T; ++r) {
        cell_of_rank[r] = code[r].second;
        rank[code[r].second] = r;
    }
}

static int derive_case(const mph_params& p, int begin, int end, int nfluid, int key_base, CaseConst& k)
{
    std::memset(&k, 0, sizeof(k));
    if (p.dim != 2 && p.dim != 3) return fail(MPH_ERR_ARG, "dim must be 2 or 3 (got %d)", p.dim);
    if (!(p.particle_spacing > 0.0)) return fail(MPH_ERR_ARG, "particle spacing must be positive");
    if (!(p.dt > 0.0) || !std::isfinite(p.dt)) return fail(MPH_ERR_ARG, "dt must be positive and finite");
    const double l0 = p.particle_spacing;
    k.begin = begin; k.end = end; k.nfluid = nfluid; k.dim = p.dim; k.key_base = key_base;
    k.l0 = l0;
    k.dt = p.dt;
    k.vol = p.dim == 2 ? l0 * l0 : l0 * l0 * l0;                      // main.cpp:558-562
    const double ha = p.radius_ratio_a * l0, hg = p.radius_ratio_a * l0;   // main.cpp:838-843
    const double hp = p.radius_ratio_p * l0, hv = p.radius_ratio_v * l0;
    if (!(ha > 0.0 && hp > 0.0 && hv > 0.0)) return fail(MPH_ERR_ARG, "radius ratios must be positive");
    double swa, swg, swp, swv, r2g;
    if (p.dim == 2) {                                                   // main.cpp:846-851
        swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / l0 / l0;
        swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        r2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * hg * hg / l0 / l0 / swg;
    } else {                                                            // main.cpp:853-857
        swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / l0 / l0 / l0;
        swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        r2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * hg * hg / l0 / l0 / l0 / swg;
    }
    k.n0a = lattice_sum(p.dim, l0, ha, swa, true);
    k.n0p = lattice_sum(p.dim, l0, hp, swp, false);
    k.ha2 = ha * ha; k.hp2 = hp * hp; k.hv2 = hv * hv;
    k.inv_ha = 1.0 / ha; k.inv_hp = 1.0 / hp; k.inv_hv = 1.0 / hv;
    k.ca = 1.0 / swa * 1.0 / hd(p.dim, ha);
    k.dca = k.ca * (1.0 / ha);
    k.cg = 1.0 / swg * 1.0 / hd(p.dim, hg);
    k.dcg = k.cg * (-2.0 / hg);
    k.cp = 1.0 / swp * 1.0 / hd(p.dim, hp);
    k.dcp = k.cp * (-2.0 / hp);
    k.dcv = 1.0 / swv * 1.0 / hd(p.dim, hv) * (-2.0 / hv);
    k.g_scale = 1.0 / r2g * hg;
    k.s_scale = 1.0 / r2g * hg * (k.vol / l0);
    k.visc_c = p.dim == 2 ? 8.0 : 10.0;
    // surface-tension constants, main.cpp:971-983
    const double cofk = p.dim == 2 ? 0.350778153 : 0.326976006;
    const double integ_n = p.dim == 2 ? 0.024679383 : 0.021425779;
    const double integ_x = p.dim == 2 ? 0.226126699 : 0.233977488;
    k.surface = 0;
    k.ratio_one = 1;
    for (int t = 0; t < kTypes; ++t) {
        k.cofa[t] = p.surface_tension[t] / ((hg / l0) * (integ_n + cofk * cofk * integ_x));
        k.adiff[t] = k.cofa[t] * (cofk) * (cofk);
        if (p.surface_tension[t] != 0.0) k.surface = 1;
        k.mass[t] = p.density[t] * k.vol;
        k.kbulk[t] = p.bulk_modulus[t];
        k.lbulk[t] = p.bulk_viscosity[t];
        k.mu[t] = p.shear_viscosity[t];
        for (int u = 0; u < kTypes; ++u) {
            k.ir[t][u] = p.interaction_ratio[t][u];
            if (k.ir[t][u] != 1.0) k.ratio_one = 0;
        }
    }
    for (int t = 0; t < kTypes; ++t)
        for (int u = 0; u < kTypes; ++u)
            k.muij[t][u] = 2.0 * (k.mu[t] * k.mu[u]) / (k.mu[t] + k.mu[u]);
    k.mu_uniform = 1;
    for (int t = 1; t < kTypes; ++t)
        if (k.mu[t] != k.mu[0]) k.mu_uniform = 0;
    for (int d = 0; d < 3; ++d) k.grav[d] = p.gravity[d];
    // rigid wall motion per dt, main.cpp:994-1033
    for (int w = 2; w < kTypes; ++w) {
        double nrm[3] = {0.0, 0.0, 0.0}, q[4];
        const double* om = p.wall_omega[w];
        const double theta = std::fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int d = 0; d < 3; ++d) nrm[d] = om[d] / theta;
        q[0] = nrm[0] * std::sin(theta * p.dt / 2.0);
        q[1] = nrm[1] * std::sin(theta * p.dt / 2.0);
        q[2] = nrm[2] * std::sin(theta * p.dt / 2.0);
        q[3] = std::cos(theta * p.dt / 2.0);
        double (*R)[3] = k.wrot[w];
        R[0][0] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[0][1] = 2.0 * (q[0] * q[1] - q[2] * q[3]);
        R[0][2] = 2.0 * (q[0] * q[2] + q[1] * q[3]);
        R[1][0] = 2.0 * (q[0] * q[1] + q[2] * q[3]);
        R[1][1] = -q[0] * q[0] + q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[1][2] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
        R[2][0] = 2.0 * (q[0] * q[2] - q[1] * q[3]);
        R[2][1] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
        R[2][2] = -q[0] * q[0] - q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
        for (int d = 0; d < 3; ++d) { k.womega[w][d] = om[d]; k.wvel[w][d] = p.wall_velocity[w][d]; }
    }
    // cell grid, main.cpp:1035-1068
    double maxr = 0.0;
    maxr = ha > maxr ? ha : maxr;
    maxr = hg > maxr ? hg : maxr;
    maxr = hp > maxr ? hp : maxr;
    maxr = hv > maxr ? hv : maxr;
    const double margin = 0.1 * l0;                                    // main.cpp:89
    double cnt[3];
    for (int d = 0; d < 3; ++d) {
        k.dmin[d] = p.domain_min[d];
        k.dmax[d] = p.domain_max[d];
        k.dwid[d] = p.domain_max[d] - p.domain_min[d];
        k.hwid[d] = 0.5 * k.dwid[d];
        if (!(k.dwid[d] > 0.0)) return fail(MPH_ERR_ARG, "domain width on axis %d must be positive", d);
    }
    cnt[0] = std::floor((p.domain_max[0] - p.domain_min[0]) / (maxr + margin));
    cnt[1] = std::floor((p.domain_max[1] - p.domain_min[1]) / (maxr + margin));
    cnt[2] = p.dim == 2 ? 1 : std::floor((p.domain_max[2] - p.domain_min[2]) / (maxr + margin));
    for (int d = 0; d < 3; ++d) {
        if (!(cnt[d] >= 1.0)) return fail(MPH_ERR_ARG, "domain axis %d is narrower than the interaction radius", d);
        if (cnt[d] > 1.0e6) return fail(MPH_ERR_ARG, "too many cells on axis %d", d);
        k.cc[d] = (int)cnt[d];
    }
    const double tcc = cnt[0] * cnt[1] * cnt[2];
    if (tcc > 2.0e8) return fail(MPH_ERR_ARG, "cell grid too large (%g cells)", tcc);
    k.tcc = (int)tcc;
    for (int d = 0; d < 3; ++d) k.cw[d] = k.dwid[d] / k.cc[d];
    k.cut2 = (maxr + margin) * (maxr + margin);                        // main.cpp:1408
    // list partners are within the list cut-off at the last rebuild and each particle has
    // moved less than margin/2 since (main.cpp:1102)
    const double reach = maxr + 2.5 * margin;
    for (int d = 0; d < 3; ++d) k.guard[d] = k.hwid[d] > 2.0 * reach ? reach : HUGE_VAL;
    k.hmax2 = maxr * maxr;
    k.shift2 = 0.5 * margin * 0.5 * margin;                             // main.cpp:1102
    return MPH_OK;
}

// ---------------------------------------------------------------------------------------

static int launch_rebuild(mph_ctx* c, int count_only)
{
    const int nb = c->nblocks();
    hipStream_t s = c->stream;
    unsigned* kf = c->sorted_keys();
    hipLaunchKernelGGL(k_keys, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->pos.p,
                       c->flags.p, c->any_flag.p, c->rank.p, kf, c->keys_a.p, c->perm_a.p);
    for (int ps = 0; ps < c->sort_passes; ++ps) {
        unsigned* kin = (ps & 1) ? c->keys_b.p : c->keys_a.p;
        unsigned* kout = (ps & 1) ? c->keys_a.p : c->keys_b.p;
        int* vin = (ps & 1) ? c->perm_b.p : c->perm_a.p;
        int* vout = (ps & 1) ? c->perm_a.p : c->perm_b.p;
        const int shift = ps * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(c->sort_blocks), dim3(kSortThreads), 0, s, kin, c->n,
                           shift, c->hist.p, c->any_flag.p);
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, c->hist.p,
                           kRadix * c->sort_blocks, c->any_flag.p);
        hipLaunchKernelGGL(k_radix_scatter, dim3(c->sort_blocks), dim3(kSortThreads), 0, s, kin, vin,
                           kout, vout, c->n, shift, c->hist.p, c->any_flag.p);
    }
    hipLaunchKernelGGL(k_cell_start, dim3((c->nkeys + 1 + 255) / 256), dim3(256), 0, s, kf, c->n,
                       c->nkeys, c->start.p, c->any_flag.p);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->flags.p,
                       c->any_flag.p, c->sorted_perm(), c->pos.p, c->vel.p, c->pidx.p, c->pos_alt.p,
                       c->vel_alt.p, c->pidx_alt.p);
    hipLaunchKernelGGL(k_scatter_back, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p,
                       c->flags.p, c->any_flag.p, c->pos_alt.p, c->vel_alt.p, c->pidx_alt.p, c->pos.p,
                       c->vel.p, c->pidx.p, c->snap.p);
    hipLaunchKernelGGL(k_build_lists, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p,
                       c->flags.p, c->any_flag.p, c->pos.p, kf, c->start.p, c->rank.p,
                       c->cell_of_rank.p, c->nbr.p, c->stride,
                       count_only ? 0 : c->width, c->ncnt.p, c->nout.p, c->max_len.p);
    return hipGetLastError() == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "rebuild launch failed");
}

// sweep kernel variants: minimum waves per SIMD the register allocation must allow
// (the autotuner picks one per context by timing them on the case itself)
constexpr int kSweepVariants = 4;
static const int kSweepMinWaves[kSweepVariants] = {4, 5, 6, 8};
using SweepAFn = decltype(&k_sweep_a<0, 4>);
using SweepBFn = decltype(&k_sweep_b<0, 4>);
static const SweepAFn kSweepA[2][kSweepVariants] = {
    {k_sweep_a<0, 4>, k_sweep_a<0, 5>, k_sweep_a<0, 6>, k_sweep_a<0, 8>},
    {k_sweep_a<1, 4>, k_sweep_a<1, 5>, k_sweep_a<1, 6>, k_sweep_a<1, 8>}};
static const SweepBFn kSweepB[2][kSweepVariants] = {
    {k_sweep_b<0, 4>, k_sweep_b<0, 5>, k_sweep_b<0, 6>, k_sweep_b<0, 8>},
    {k_sweep_b<1, 4>, k_sweep_b<1, 5>, k_sweep_b<1, 6>, k_sweep_b<1, 8>}};

// advance_walls: this launch is a real step (advances wall centres, counts pairs); the
// autotuner launches with 0 so that repeated launches leave the state unchanged
static int launch_sweep_a(mph_ctx* c, int advance_walls, int variant = -1)
{
    const int v = variant >= 0 ? variant : c->variant_a;
    const int ncw = advance_walls ? c->ncases : 0;
    hipLaunchKernelGGL(kSweepA[c->any_surface ? 1 : 0][v], dim3(c->nblocks()), dim3(kBlock), 0, c->stream,
                       c->d_cases.p, c->d_bmap.p, c->pos.p, c->vel.p, c->aux.p, c->dens_a.p, c->vstrain.p,
                       c->divp.p, c->nbr.p, c->ncnt.p, c->stride, c->wallc.p, ncw,
                       advance_walls ? c->pairs.p : c->pairs_scratch.p);
    return hipGetLastError() == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "sweep A launch failed");
}

static int launch_move(mph_ctx* c)
{
    hipLaunchKernelGGL(k_move, dim3(c->nblocks()), dim3(kBlock), 0, c->stream, c->d_cases.p, c->d_bmap.p,
                       c->pos.p, c->vel.p, c->vnext.p, c->snap.p, c->wallc.p, c->flags.p, c->any_flag.p);
    return hipGetLastError() == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "move launch failed");
}

static int launch_sweep_b(mph_ctx* c, int bookkeeping = 1, int variant = -1)
{
    const int v = variant >= 0 ? variant : c->variant_b;
    hipLaunchKernelGGL(kSweepB[c->any_surface ? 1 : 0][v], dim3(c->nblocks()), dim3(kBlock), 0, c->stream,
                       c->d_cases.p, c->d_bmap.p, c->pos.p, c->vel.p, c->aux.p, c->vnext_out(bookkeeping),
                       c->force_out(bookkeeping), c->nbr.p, c->ncnt.p, c->stride, c->flags.p, c->any_flag.p,
                       c->rebuilds.p, bookkeeping ? c->ncases : 0);
    return hipGetLastError() == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "sweep B launch failed");
}

static int launch_step(mph_ctx* c)
{
    int rc = launch_move(c);
    if (!rc) rc = launch_rebuild(c, 0);
    if (!rc) rc = launch_sweep_a(c, 1);
    if (!rc) rc = launch_sweep_b(c);
    return rc;
}

static int read_int(mph_ctx* c, const int* dptr, int* out)
{
    HIP_TRY(hipMemcpyAsync(c->h_scalar, dptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *out = *c->h_scalar;
    return MPH_OK;
}

static int check_capacity(mph_ctx* c)
{
    int mx = 0;
    int rc = read_int(c, c->max_len.p, &mx);
    if (rc) return rc;
    if (mx > c->width)
        return fail(MPH_ERR_CAPACITY, "a neighbour list reached %d entries, above the reserved width %d; "
                    "results of this call are invalid", mx, c->width);
    return MPH_OK;
}

// ---------------------------------------------------------------------------------------
// autotuner: time every sweep variant on the context's own lists and keep the fastest.
// Results are cached in a text file (MPH_TUNE_CACHE, default ~/.cache/mphexplicit_amd/
// sweep_variants.txt) keyed by device, dimension, surface tension and problem size;
// MPH_TUNE=0 keeps the built-in defaults, MPH_TUNE=force ignores the cache.

static std::string tune_key(mph_ctx* c)
{
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
    int dim = 3;
    for (auto& ch : c->cases) dim = std::min(dim, ch.k.dim);
    int nb = 0;
    while ((1ll << (nb + 1)) <= (long long)c->n) ++nb;
    char buf[256];
    std::snprintf(buf, sizeof(buf), "%s|dim%d|surf%d|n2^%d|w%d", arch.c_str(), dim, c->any_surface, nb, c->width / 16);
    return buf;
}

static std::string tune_path()
{
    const char* p = std::getenv("MPH_TUNE_CACHE");
    if (p && *p) return p;
    const char* h = std::getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/mphexplicit_amd/sweep_variants.txt";
}

static int autotune(mph_ctx* c)
{
    const char* mode = std::getenv("MPH_TUNE");
    if (mode && !std::strcmp(mode, "0")) return MPH_OK;
    const std::string key = tune_key(c), path = tune_path();
    if (!(mode && !std::strcmp(mode, "force"))) {
        if (FILE* f = std::fopen(path.c_str(), "r")) {
            char k[256];
            int a, b;
            while (std::fscanf(f, "%255s %d %d", k, &a, &b) == 3)
                if (key == k && a >= 0 && a < kSweepVariants && b >= 0 && b < kSweepVariants) {
                    c->variant_a = a;
                    c->variant_b = b;
                    std::fclose(f);
                    return MPH_OK;
                }
            std::fclose(f);
        }
    }
    int rc = 0;
    if ((rc = c->tune_v.alloc(c->n)) || (rc = c->tune_f.alloc(c->n))) return rc;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    float best_a = 1e30f, best_b = 1e30f;
    for (int v = 0; v < kSweepVariants && rc == MPH_OK; ++v) {
        for (int kind = 0; kind < 2 && rc == MPH_OK; ++kind) {
            float total = 0.0f;
            for (int rep = 0; rep < 4 && rc == MPH_OK; ++rep) {
                (void)hipEventRecord(e0, c->stream);
                rc = kind == 0 ? launch_sweep_a(c, 0, v) : launch_sweep_b(c, 0, v);
                (void)hipEventRecord(e1, c->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                if (rep > 0) total += ms;           // the first launch warms caches
            }
            if (kind == 0 && total < best_a) { best_a = total; c->variant_a = v; }
            if (kind == 1 && total < best_b) { best_b = total; c->variant_b = v; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->tune_v.release();
    c->tune_f.release();
    if (rc) return rc;
    const std::string dir = path.substr(0, path.rfind('/'));
    std::string mk = "mkdir -p '" + dir + "' 2>/dev/null";
    (void)!std::system(mk.c_str());
    if (FILE* f = std::fopen(path.c_str(), "a")) {
        std::fprintf(f, "%s %d %d\n", key.c_str(), c->variant_a, c->variant_b);
        std::fclose(f);
    }
    return MPH_OK;
}

static int setup(mph_ctx* c, int ncases, const mph_params* prm, const int* counts, const int* type,
                 const double* x, const double* v)
{
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count <= 0)
        return fail(MPH_ERR_DEVICE, "no GPU available");
    HIP_TRY(hipGetDevice(&c->device));
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipHostMalloc((void**)&c->h_scalar, sizeof(long long) * 4, hipHostMallocDefault));
    c->ncases = ncases;
    c->cases.resize(ncases);
    // particles of each case, fluid first (the reference's order after its first sort)
    long long total = 0;
    for (int k = 0; k < ncases; ++k) {
        if (counts[k] <= 0) return fail(MPH_ERR_ARG, "case %d has no particles", k);
        total += counts[k];
    }
    if (total > (1ll << 30)) return fail(MPH_ERR_ARG, "too many particles");
    c->n = (int)total;
    c->stride = (c->n + 63) / 64 * 64;
    std::vector<double4> hpos(c->n), hvel(c->n);
    std::vector<int> hidx(c->n);
    std::vector<double> hwall((size_t)ncases * kTypes * 3, 0.0);
    int begin = 0, key_base = 0;
    std::vector<int> hrank, hcell;
    std::vector<int2> logical;
    c->any_surface = 0;
    for (int k = 0; k < ncases; ++k) {
        const int nk = counts[k];
        int nfluid = 0;
        for (int i = 0; i < nk; ++i) {
            const int t = type[begin + i];
            if (t < 0 || t > 3) return fail(MPH_ERR_ARG, "case %d particle %d: type %d outside 0..3", k, i, t);
            if (t < 2) ++nfluid;
        }
        CaseHost& ch = c->cases[k];
        ch.prm = prm[k];
        int rc = derive_case(prm[k], begin, begin + nk, nfluid, key_base, ch.k);
        if (rc) return fail(rc, "case %d: %s", k, mph::last_error().c_str());
        if ((long long)key_base + 2ll * ch.k.tcc > (1ll << 31) - 2) return fail(MPH_ERR_ARG, "cell keys overflow");
        key_base += 2 * ch.k.tcc;
        ch.time = prm[k].time;
        c->any_surface |= ch.k.surface;
        for (int t = 0; t < kTypes; ++t)
            for (int d = 0; d < 3; ++d) hwall[((size_t)k * kTypes + t) * 3 + d] = prm[k].wall_center[t][d];
        for (int i = 0; i < nk; ++i) {
            const int g = begin + i;
            const long long tag = (long long)(type[g] | (k << 2));
            double w;
            std::memcpy(&w, &tag, sizeof(w));
            hpos[g] = make_double4(x[3 * g], x[3 * g + 1], x[3 * g + 2], w);
            hvel[g] = make_double4(v[3 * g], v[3 * g + 1], v[3 * g + 2], 0.0);
            hidx[g] = i;
        }
        for (int i = 0; i < nk; ++i)
            for (int d = 0; d < 3; ++d) {
                const double xv = x[3 * (size_t)(begin + i) + d];
                if (!(xv >= ch.k.dmin[d] - 0.25 * ch.k.dwid[d] && xv <= ch.k.dmax[d] + 0.25 * ch.k.dwid[d]))
                    return fail(MPH_ERR_ARG, "case %d particle %d lies outside the periodic domain", k, i);
            }
        std::vector<int> rk, cr;
        morton_ranks(ch.k, rk, cr);
        ch.k.rank_base = (int)hrank.size();
        hrank.insert(hrank.end(), rk.begin(), rk.end());
        hcell.insert(hcell.end(), cr.begin(), cr.end());
        for (int b = 0; b < nk; b += kBlock) logical.push_back(make_int2(k, begin + b));
        begin += nk;
    }
    // XCD-aware order: physical workgroup w runs on XCD w % kXcds; give each XCD one
    // contiguous run of the logical (particle-ordered) blocks
    {
        const int nb = (int)logical.size();
        const int per = (nb + kXcds - 1) / kXcds;
        c->bmap_host.assign((size_t)per * kXcds, make_int2(-1, 0));
        for (int w = 0; w < per * kXcds; ++w) {
            const int L = (w % kXcds) * per + w / kXcds;
            if (L < nb) c->bmap_host[w] = logical[L];
        }
    }
    c->nkeys = key_base;
    int bits = 1;
    while ((1ll << bits) < (long long)c->nkeys) ++bits;
    c->sort_passes = (bits + kRadixBits - 1) / kRadixBits;
    c->sort_blocks = (c->n + kSortTile - 1) / kSortTile;

    std::vector<CaseConst> hk(ncases);
    for (int k = 0; k < ncases; ++k) hk[k] = c->cases[k].k;
    int rc = 0;
    const size_t n = c->n;
    if ((rc = c->d_cases.alloc(ncases)) || (rc = c->d_bmap.alloc(c->bmap_host.size())) ||
        (rc = c->pos.alloc(n)) || (rc = c->pos_alt.alloc(n)) || (rc = c->vel.alloc(n)) ||
        (rc = c->vel_alt.alloc(n)) || (rc = c->vnext.alloc(n)) || (rc = c->aux.alloc(n)) ||
        (rc = c->force.alloc(n)) || (rc = c->snap.alloc(n)) || (rc = c->pidx.alloc(n)) ||
        (rc = c->pidx_alt.alloc(n)) || (rc = c->ncnt.alloc(n)) || (rc = c->nout.alloc(n)) ||
        (rc = c->start.alloc((size_t)c->nkeys + 1)) || (rc = c->dens_a.alloc(n)) ||
        (rc = c->vstrain.alloc(n)) || (rc = c->divp.alloc(n)) || (rc = c->vir.alloc(n)) || (rc = c->wallc.alloc(hwall.size())) ||
        (rc = c->keys_a.alloc(n)) || (rc = c->keys_b.alloc(n)) ||
        (rc = c->hist.alloc((size_t)kRadix * c->sort_blocks)) || (rc = c->perm_a.alloc(n)) ||
        (rc = c->perm_b.alloc(n)) || (rc = c->flags.alloc(ncases)) ||
        (rc = c->rank.alloc(hrank.size())) || (rc = c->cell_of_rank.alloc(hcell.size())) || (rc = c->any_flag.alloc(1)) ||
        (rc = c->max_len.alloc(1)) || (rc = c->rebuilds.alloc(ncases)) || (rc = c->pairs.alloc(1)) ||
        (rc = c->pairs_scratch.alloc(1)))
        return rc;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(c->d_cases.p, hk.data(), sizeof(CaseConst) * ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_bmap.p, c->bmap_host.data(), sizeof(int2) * c->bmap_host.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->pos.p, hpos.data(), sizeof(double4) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->vel.p, hvel.data(), sizeof(double4) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->pidx.p, hidx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->wallc.p, hwall.data(), sizeof(double) * hwall.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->rank.p, hrank.data(), sizeof(int) * hrank.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->cell_of_rank.p, hcell.data(), sizeof(int) * hcell.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(c->aux.p, 0, sizeof(double4) * n, s));
    HIP_TRY(hipMemsetAsync(c->force.p, 0, sizeof(double4) * n, s));
    HIP_TRY(hipMemsetAsync(c->vir.p, 0, sizeof(double) * n, s));
    HIP_TRY(hipMemsetAsync(c->rebuilds.p, 0, sizeof(long long) * ncases, s));
    HIP_TRY(hipMemsetAsync(c->ncnt.p, 0, sizeof(int) * n, s));
    HIP_TRY(hipMemsetAsync(c->max_len.p, 0, sizeof(int), s));
    HIP_TRY(hipMemsetAsync(c->pairs.p, 0, sizeof(unsigned long long), s));
    // first bucketing and lists for every case (main.cpp:358-359): count pass to size the
    // list width, then the real build
    std::vector<int> ones(ncases, 1);
    HIP_TRY(hipMemcpyAsync(c->flags.p, ones.data(), sizeof(int) * ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->any_flag.p, ones.data(), sizeof(int), hipMemcpyHostToDevice, s));
    // the list buffer is needed by the count pass only as a valid pointer
    if ((rc = c->nbr.alloc(1))) return rc;
    if ((rc = launch_rebuild(c, 1))) return rc;
    int mx = 0;
    if ((rc = read_int(c, c->max_len.p, &mx))) return rc;
    c->width = ((int)std::ceil(mx * 1.25) + 8 + 3) / 4 * 4;
    if ((rc = c->nbr.alloc((size_t)c->width * c->stride))) return rc;
    HIP_TRY(hipMemsetAsync(c->max_len.p, 0, sizeof(int), s));
    // the lists again with room for entries (the sort is the identity now)
    if ((rc = launch_rebuild(c, 0))) return rc;
    if ((rc = launch_sweep_a(c, 0))) return rc;                          // main.cpp:362-364
    HIP_TRY(hipMemcpyAsync(c->vnext.p, c->vel.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(c->flags.p, 0, sizeof(int) * ncases, s));
    HIP_TRY(hipMemsetAsync(c->any_flag.p, 0, sizeof(int), s));
    std::vector<long long> one_ll(ncases, 1);
    HIP_TRY(hipMemcpyAsync(c->rebuilds.p, one_ll.data(), sizeof(long long) * ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    if ((rc = check_capacity(c))) return rc;
    return autotune(c);
}

// ---------------------------------------------------------------------------------------
// C ABI

extern "C" int mph_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" mph_ctx* mph_create_batch(int ncases, const mph_params* p, const int* counts, const int* type,
                                     const double* x, const double* v)
{
    if (ncases <= 0 || !p || !counts || !type || !x || !v) {
        fail(MPH_ERR_ARG, "mph_create_batch: bad argument");
        return nullptr;
    }
    mph::last_error().clear();
    std::unique_ptr<mph_ctx> c(new mph_ctx());
    if (setup(c.get(), ncases, p, counts, type, x, v) != MPH_OK) return nullptr;
    return c.release();
}

extern "C" mph_ctx* mph_create(const mph_params* p, int n, const int* type, const double* x, const double* v)
{
    return mph_create_batch(1, p, &n, type, x, v);
}

extern "C" int mph_step(mph_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MPH_ERR_ARG, "mph_step: bad argument");
    for (int k = 0; k < nsteps; ++k) {
        int rc = launch_step(c);
        if (rc) return rc;
        for (auto& ch : c->cases) ch.time += ch.prm.dt;   // Time += Dt, main.cpp:449
        ++c->steps;
    }
    return check_capacity(c);
}

extern "C" int mph_step_graph(mph_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MPH_ERR_ARG, "mph_step_graph: bad argument");
    if (!c->graph) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = launch_step(c);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return fail(MPH_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
        c->graph_src = g;
        HIP_TRY(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
    }
    for (int k = 0; k < nsteps; ++k) {
        HIP_TRY(hipGraphLaunch(c->graph, c->stream));
        for (auto& ch : c->cases) ch.time += ch.prm.dt;
        ++c->steps;
    }
    return check_capacity(c);
}

extern "C" int mph_step_timed(mph_ctx* c, int nsteps, mph_timing* out)
{
    if (!c || nsteps < 0 || !out) return fail(MPH_ERR_ARG, "mph_step_timed: bad argument");
    std::memset(out, 0, sizeof(*out));
    std::vector<long long> rb0(c->ncases), rb1(c->ncases);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(rb0.data(), c->rebuilds.p, sizeof(long long) * c->ncases, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemsetAsync(c->pairs.p, 0, sizeof(unsigned long long), c->stream));
    std::vector<hipEvent_t> ev((size_t)nsteps * 5, nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    int rc = MPH_OK;
    for (int k = 0; k < nsteps && rc == MPH_OK; ++k) {
        hipEvent_t* e = &ev[(size_t)k * 5];
        (void)hipEventRecord(e[0], c->stream);
        rc = launch_move(c);
        (void)hipEventRecord(e[1], c->stream);
        if (!rc) rc = launch_rebuild(c, 0);
        (void)hipEventRecord(e[2], c->stream);
        if (!rc) rc = launch_sweep_a(c, 1);
        (void)hipEventRecord(e[3], c->stream);
        if (!rc) rc = launch_sweep_b(c);
        (void)hipEventRecord(e[4], c->stream);
        for (auto& ch : c->cases) ch.time += ch.prm.dt;
        ++c->steps;
    }
    hipError_t se = hipStreamSynchronize(c->stream);
    double* slot[4] = {&out->move_ms, &out->rebuild_ms, &out->sweep_a_ms, &out->sweep_b_ms};
    for (int k = 0; k < nsteps && rc == MPH_OK && se == hipSuccess; ++k)
        for (int ph = 0; ph < 4; ++ph) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, ev[(size_t)k * 5 + ph], ev[(size_t)k * 5 + ph + 1]) == hipSuccess)
                *slot[ph] += ms;
        }
    for (auto& e : ev) if (e) (void)hipEventDestroy(e);
    if (rc) return rc;
    if (se != hipSuccess) return fail(MPH_ERR_DEVICE, "step failed: %s", hipGetErrorString(se));
    unsigned long long pairs = 0;
    HIP_TRY(hipMemcpy(&pairs, c->pairs.p, sizeof(pairs), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(rb1.data(), c->rebuilds.p, sizeof(long long) * c->ncases, hipMemcpyDeviceToHost));
    out->pairs = (long long)pairs;
    for (int k = 0; k < c->ncases; ++k) out->rebuilds = std::max(out->rebuilds, rb1[k] - rb0[k]);
    return check_capacity(c);
}

extern "C" int mph_sync(mph_ctx* c)
{
    if (!c) return fail(MPH_ERR_ARG, "mph_sync: null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPH_OK;
}

extern "C" int mph_particle_count(const mph_ctx* c, int case_id)
{
    if (!c) return fail(MPH_ERR_ARG, "null context");
    if (case_id == -1) return c->n;
    if (case_id < 0 || case_id >= c->ncases) return fail(MPH_ERR_ARG, "case %d out of range", case_id);
    return c->cases[case_id].k.end - c->cases[case_id].k.begin;
}

extern "C" int mph_get_time(const mph_ctx* c, int case_id, double* t)
{
    if (!c || !t || case_id < 0 || case_id >= c->ncases) return fail(MPH_ERR_ARG, "mph_get_time: bad argument");
    *t = c->cases[case_id].time;
    return MPH_OK;
}

template <class T>
static int fetch(mph_ctx* c, const T* dptr, int begin, int count, std::vector<T>& out)
{
    out.resize(count);
    HIP_TRY(hipMemcpyAsync(out.data(), dptr + begin, sizeof(T) * count, hipMemcpyDeviceToHost, c->stream));
    return MPH_OK;
}

extern "C" int mph_get_fields(mph_ctx* c, int case_id, int order, mph_fields* out)
{
    if (!c || !out || case_id < 0 || case_id >= c->ncases || (order != MPH_ORDER_CURRENT && order != MPH_ORDER_ORIGINAL))
        return fail(MPH_ERR_ARG, "mph_get_fields: bad argument");
    const CaseConst& k = c->cases[case_id].k;
    const int b = k.begin, m = k.end - k.begin;
    std::vector<double4> p4, v4, vn4, a4, f4;
    std::vector<int> idx, nc;
    std::vector<double> da, vs, dv, vp;
    int rc = 0;
    if ((rc = fetch(c, c->pos.p, b, m, p4)) || (rc = fetch(c, c->vel.p, b, m, v4)) ||
        (rc = fetch(c, c->vnext.p, b, m, vn4)) || (rc = fetch(c, c->aux.p, b, m, a4)) ||
        (rc = fetch(c, c->force.p, b, m, f4)) || (rc = fetch(c, c->pidx.p, b, m, idx)) ||
        (rc = fetch(c, c->nout.p, b, m, nc)) || (rc = fetch(c, c->dens_a.p, b, m, da)) ||
        (rc = fetch(c, c->vstrain.p, b, m, vs)) || (rc = fetch(c, c->divp.p, b, m, dv)) ||
        (rc = fetch(c, c->vir.p, b, m, vp)))
        return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < m; ++i) {
        const int o = order == MPH_ORDER_CURRENT ? i : idx[i];
        if (o < 0 || o >= m) return fail(MPH_ERR_STATE, "corrupt particle index");
        long long tg;
        std::memcpy(&tg, &p4[i].w, sizeof(tg));
        if (out->particle_index) out->particle_index[o] = idx[i];
        if (out->type) out->type[o] = (int)(tg & 3);
        if (out->position) { out->position[3 * o] = p4[i].x; out->position[3 * o + 1] = p4[i].y; out->position[3 * o + 2] = p4[i].z; }
        if (out->velocity) { out->velocity[3 * o] = vn4[i].x; out->velocity[3 * o + 1] = vn4[i].y; out->velocity[3 * o + 2] = vn4[i].z; }
        if (out->force) { out->force[3 * o] = f4[i].x; out->force[3 * o + 1] = f4[i].y; out->force[3 * o + 2] = f4[i].z; }
        if (out->density_a) out->density_a[o] = da[i];
        if (out->gravity_center) { out->gravity_center[3 * o] = a4[i].x; out->gravity_center[3 * o + 1] = a4[i].y; out->gravity_center[3 * o + 2] = a4[i].z; }
        if (out->pressure_a) out->pressure_a[o] = a4[i].w;
        if (out->vol_strain_p) out->vol_strain_p[o] = vs[i];
        if (out->divergence_p) out->divergence_p[o] = dv[i];
        if (out->pressure_p) out->pressure_p[o] = v4[i].w;
        if (out->virial_pressure) out->virial_pressure[o] = vp[i];
        if (out->neighbor_count) out->neighbor_count[o] = nc[i];
    }
    return MPH_OK;
}

extern "C" int mph_compute_virial(mph_ctx* c)
{
    if (!c) return fail(MPH_ERR_ARG, "null context");
    hipLaunchKernelGGL(k_virial, dim3(c->nblocks()), dim3(kBlock), 0, c->stream, c->d_cases.p, c->d_bmap.p,
                       c->pos.p, c->vel.p, c->vnext.p, c->aux.p, c->nbr.p, c->ncnt.p, c->stride, c->vir.p);
    HIP_TRY(hipGetLastError());
    return MPH_OK;
}

extern "C" int mph_get_neighbors(mph_ctx* c, int case_id, long long* ptr, int* ind)
{
    if (!c || !ptr || case_id < 0 || case_id >= c->ncases) return fail(MPH_ERR_ARG, "mph_get_neighbors: bad argument");
    const CaseConst& k = c->cases[case_id].k;
    const int b = k.begin, m = k.end - k.begin;
    std::vector<int> idx, cnt, nc, ell((size_t)c->width * m);
    int rc = 0;
    if ((rc = fetch(c, c->pidx.p, b, m, idx)) || (rc = fetch(c, c->ncnt.p, b, m, cnt)) ||
        (rc = fetch(c, c->nout.p, b, m, nc)))
        return rc;
    // groups of four entries: rows of the 2D copy are groups, columns the case's particles
    HIP_TRY(hipMemcpy2DAsync(ell.data(), sizeof(int) * 4 * m, c->nbr.p + (size_t)4 * b, sizeof(int) * 4 * c->stride,
                             sizeof(int) * 4 * m, c->width / 4, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // rows by original id: reference lists include self (when the particle has a list)
    std::vector<long long> rowlen(m, 0);
    for (int i = 0; i < m; ++i) rowlen[idx[i]] = nc[i];
    ptr[0] = 0;
    for (int o = 0; o < m; ++o) ptr[o + 1] = ptr[o] + rowlen[o];
    if (!ind) return MPH_OK;
    for (int i = 0; i < m; ++i) {
        const int o = idx[i];
        long long w = ptr[o];
        std::vector<int> row;
        for (int e = 0; e < cnt[i]; ++e) {
            const int j = ell[(((size_t)(e >> 2) * m + i) << 2) + (e & 3)] - b;
            if (j < 0 || j >= m) return fail(MPH_ERR_STATE, "neighbour index outside its case");
            row.push_back(idx[j]);
        }
        if (nc[i] > cnt[i]) row.push_back(o);
        std::sort(row.begin(), row.end());
        if ((long long)row.size() != rowlen[o]) return fail(MPH_ERR_STATE, "neighbour row length mismatch");
        for (int j : row) ind[w++] = j;
    }
    return MPH_OK;
}

extern "C" int mph_get_stats(const mph_ctx* c, mph_stats* out)
{
    if (!c || !out) return fail(MPH_ERR_ARG, "mph_get_stats: bad argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<long long> rb(c->ncases);
    std::vector<int> cnt(c->n);
    int mx = 0;
    HIP_TRY(hipMemcpy(rb.data(), c->rebuilds.p, sizeof(long long) * c->ncases, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cnt.data(), c->ncnt.p, sizeof(int) * c->n, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&mx, c->max_len.p, sizeof(int), hipMemcpyDeviceToHost));
    out->steps = c->steps;
    out->rebuilds = 0;
    for (long long r : rb) out->rebuilds = std::max(out->rebuilds, r);
    out->list_entries = 0;
    for (int v : cnt) out->list_entries += v;
    out->list_width = c->width;
    out->max_list_len = mx;
    out->sweep_a_min_waves = kSweepMinWaves[c->variant_a];
    out->sweep_b_min_waves = kSweepMinWaves[c->variant_b];
    return MPH_OK;
}

extern "C" void mph_destroy(mph_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}
