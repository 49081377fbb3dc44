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
#include <cerrno>
#include <cmath>
#include <filesystem>
#include <cstdlib>
#include <cstring>
#include <string>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "mph.h"
#include "mph_case.h"
#include "mph_error.h"
#include "mph_kernels.h"
#include "mph_mixed.h"
#include "mph_surrogate.h"

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

// cluster sizes the sweeps are instantiated for; the autotuner picks one per context
constexpr int kClusterSizes[] = {2, 4, 8};
constexpr int kNumClusterSizes = 3;

struct mph_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int ncases = 0, n = 0, ncells = 0;
    int nkeys = 0, sort_passes = 0, sort_blocks = 0, any_surface = 0;
    int mixed = 0;                 // MPH_MIXED: FP32 pair sweeps (all cases of a context)
    int same_pv = 0;               // every case has equal pressure and viscosity radii
    int same_ap = 0;               // every case has equal attractive and pressure radii
    int ratio_one = 0;             // every case has all InteractionRatio == 1
    int dens_stale = 0;            // DensityA was not summed by the last step (no surface tension)
    int graph_stale = 0;           // ... nor by the captured step
    int K = 4;                 // cluster size (rows per wavefront)
    int lcap = 0;              // largest grid of the list build (workgroups)
    int width = 0;             // reserved entries per cluster list
    int nslots = 0;            // cluster slots over all cases (capacity)
    long long steps = 0;
    std::vector<CaseHost> cases;
    std::vector<int2> bmap_host;   // particle blocks (XCD-interleaved)
    std::vector<int2> cmap_host;   // cluster blocks of 4 slots (XCD-interleaved)
    DevBuf<CaseConst> d_cases;
    DevBuf<int2> d_bmap, d_cmap;
    DevBuf<double4> pos, pos_alt, vel, vel_alt, vnext, aux, force, snap;
    DevBuf<int> pidx, pidx_alt, nout, start;
    DevBuf<int> selfin;            // 1: the particle's own index is in its cluster's union list
    DevBuf<double> dens_a, vstrain, divp, wallc, vir;
    DevBuf<unsigned> keys_a, keys_b, hist, ccount, scan_sums;
    DevBuf<unsigned> kfull;        // batches: sorted keys of all particles
    DevBuf<int> pfull, coff, nsort;  // batches: sort sources, compacted case offsets, counts
    DevBuf<int> perm_a, perm_b, rank, cell_case, slot_case;
    DevBuf<int4> cl_info;
    DevBuf<int4> stencil;          // per cell: its stencil cells' ranges in ascending rank (27 slots)
    DevBuf<int> stencil_rank;      // per cell: its stencil cells' ranks, ascending (fixed grid)
    DevBuf<int> cl_count, cl_nbr, cl_len;
    DevBuf<int> flags, any_flag, max_len;
    DevBuf<long long> rebuilds;
    DevBuf<unsigned long long> pairs, pairs_scratch;
    DevBuf<double4> tune_v, tune_f;      // sweep B outputs while the autotuner times tiles
    int* h_scalar = nullptr;   // pinned host word for small readbacks
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_src = nullptr;
    hipEvent_t ev_flag = nullptr;     // the eager step's flag readback has landed
    hipEvent_t ev_cap = nullptr;      // the eager step's list-overflow readback has landed

    void drop_graph()
    {
        if (graph) (void)hipGraphExecDestroy(graph);
        if (graph_src) (void)hipGraphDestroy(graph_src);
        graph = nullptr;
        graph_src = nullptr;
    }

    ~mph_ctx()
    {
        drop_graph();
        if (ev_flag) (void)hipEventDestroy(ev_flag);
        if (ev_cap) (void)hipEventDestroy(ev_cap);
        if (h_scalar) (void)hipHostFree(h_scalar);
        if (stream) (void)hipStreamDestroy(stream);
    }

    unsigned* sorted_keys() { return (sort_passes & 1) ? keys_b.p : keys_a.p; }
    int* sorted_perm() { return (sort_passes & 1) ? perm_b.p : perm_a.p; }
    int nblocks() const { return (int)bmap_host.size(); }
    int ncblocks() const { return (int)cmap_host.size(); }
    // sweep A without surface tension: eight-wavefront workgroups, each two consecutive entries
    // of one XCD's run
    int nablocks() const { return ((ncblocks() - kXcds) / kXcds + 1) / 2 * kXcds; }
    double4* vnext_out(int real) { return real ? vnext.p : tune_v.p; }
    double4* force_out(int real) { return real ? force.p : tune_f.p; }
};

// XCD-aware order: physical workgroup w runs on XCD w % kXcds; give each XCD one
// contiguous run of the logical (space-ordered) blocks, padding with case -1
static std::vector<int2> xcd_order(const std::vector<int2>& logical)
{
    const int nb = (int)logical.size();
    const int per = (nb + kXcds - 1) / kXcds;
    std::vector<int2> out((size_t)per * kXcds, make_int2(-1, 0));
    for (int w = 0; w < per * kXcds; ++w) {
        const int L = (w % kXcds) * per + w / kXcds;
        if (L < nb) out[w] = logical[L];
    }
    return out;
}

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

// The stencil of every cell in storage order: the Morton ranks of the cells of its 3^d window
// ({i-1, i, i+1} per axis, periodic, in ascending cell index; one cell on an axis of range 0,
// the 2D z axis; repeated cells where an axis has fewer than 3 cells), in ascending rank with
// repeats in window order (main.cpp:1365-1391 visits them in ascending cell id, and cell ids
// and ranks order a window alike).  27 entries per cell (9 used in 2D).  The grid is fixed, so
// this is computed once; each rebuild looks up the cells' particle ranges (k_stencil).
static void stencil_table(const CaseConst& k, const std::vector<int>& rank, const std::vector<int>& cell_of_rank,
                          std::vector<int>& out)
{
    auto window = [](int ic, int c, int range, int* w) {
        if (range == 0) { w[0] = ic; return 1; }
        int a = (ic - 1 + c) % c, b = ic, d = (ic + 1) % c;
        if (a > b) std::swap(a, b);
        if (b > d) std::swap(b, d);
        if (a > b) std::swap(a, b);
        w[0] = a; w[1] = b; w[2] = d;
        return 3;
    };
    const size_t base = out.size();
    out.resize(base + (size_t)27 * k.tcc, 0);
    for (int r = 0; r < k.tcc; ++r) {
        const int cid = cell_of_rank[r];
        const int iz = cid % k.cc[2], iy = (cid / k.cc[2]) % k.cc[1], ix = cid / (k.cc[1] * k.cc[2]);
        int wx[3], wy[3], wz[3];
        const int nx = window(ix, k.cc[0], 1, wx), ny = window(iy, k.cc[1], 1, wy);
        const int nz = window(iz, k.cc[2], k.dim == 2 ? 0 : 1, wz);
        std::vector<std::pair<int, int>> cells;
        for (int a = 0; a < nx; ++a)
            for (int b = 0; b < ny; ++b)
                for (int d = 0; d < nz; ++d)
                    cells.push_back({rank[(wx[a] * k.cc[1] + wy[b]) * k.cc[2] + wz[d]], (int)cells.size()});
        std::sort(cells.begin(), cells.end());
        for (size_t q = 0; q < cells.size(); ++q) out[base + (size_t)27 * r + q] = cells[q].first;
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
    // the members of a cluster shared a cell at the last rebuild and have moved less than
    // margin/2 since, and an interacting partner is within maxr of its row: a cell plus reach
    // below half the width makes the image near the first member the minimum image
    k.img_axes = 0;
    for (int d = 0; d < 3; ++d)
        if (k.hwid[d] > k.cw[d] + reach) k.img_axes |= 1 << d;
    // plain differences round within 2^-52 (W/2 + cut-off) of the reference's periodic form,
    // under 1e-9 of the cut-off for W/2 below 1e6 cut-offs: the list build's plain test
    // redoes the reference's arithmetic for q^2 within 1e-8 of the cut-off.  Images x + kW
    // round within 2^-53 M more (M: the largest coordinate magnitude plus W); below
    // 4e6 cut-offs q^2 of image differences stays within 7e-9 of the reference's
    k.list_plain = 1;
    k.list_img = 1;
    for (int d = 0; d < 3; ++d) {
        if (!(k.hwid[d] <= 1.0e6 * (maxr + margin))) k.list_plain = 0;
        const double m = std::max(std::fabs(k.dmin[d]), std::fabs(k.dmax[d])) + k.dwid[d];
        if (!(m <= 4.0e6 * (maxr + margin))) k.list_img = 0;
    }
    if (!k.list_plain) k.list_img = 0;
    k.hmax2 = maxr * maxr;
    k.shift2 = 0.5 * margin * 0.5 * margin;                             // main.cpp:1102
    return MPH_OK;
}

// ---------------------------------------------------------------------------------------
// launches

#define LAUNCH_K(kern, K, ...)                                                        \
    do {                                                                              \
        switch (K) {                                                                  \
        case 2: hipLaunchKernelGGL((kern<2>), __VA_ARGS__); break;                    \
        case 4: hipLaunchKernelGGL((kern<4>), __VA_ARGS__); break;                    \
        default: hipLaunchKernelGGL((kern<8>), __VA_ARGS__); break;                   \
        }                                                                             \
    } while (0)

#define LAUNCH_SKC(kern, SURF, K, COUNT, ...)                                         \
    do {                                                                              \
        if (COUNT) LAUNCH_SK_(kern, SURF, K, true, __VA_ARGS__);                      \
        else LAUNCH_SK_(kern, SURF, K, false, __VA_ARGS__);                           \
    } while (0)

#define LAUNCH_SKCC(kern, SURF, K, B1, B2, ...)                                      \
    do {                                                                              \
        if (B1) {                                                                     \
            if (B2) LAUNCH_SK__(kern, SURF, K, true, true, __VA_ARGS__);              \
            else LAUNCH_SK__(kern, SURF, K, true, false, __VA_ARGS__);                \
        } else {                                                                      \
            if (B2) LAUNCH_SK__(kern, SURF, K, false, true, __VA_ARGS__);             \
            else LAUNCH_SK__(kern, SURF, K, false, false, __VA_ARGS__);               \
        }                                                                             \
    } while (0)

#define LAUNCH_SK__(kern, SURF, K, CB, CB2, ...)                                      \
    do {                                                                              \
        if (SURF) {                                                                   \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<1, 2, CB, CB2>), __VA_ARGS__); break;    \
            case 4: hipLaunchKernelGGL((kern<1, 4, CB, CB2>), __VA_ARGS__); break;    \
            default: hipLaunchKernelGGL((kern<1, 8, CB, CB2>), __VA_ARGS__); break;   \
            }                                                                         \
        } else {                                                                      \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<0, 2, CB, CB2>), __VA_ARGS__); break;    \
            case 4: hipLaunchKernelGGL((kern<0, 4, CB, CB2>), __VA_ARGS__); break;    \
            default: hipLaunchKernelGGL((kern<0, 8, CB, CB2>), __VA_ARGS__); break;   \
            }                                                                         \
        }                                                                             \
    } while (0)

// surface-tension-free instantiations only
#define LAUNCH_S0__(kern, K, CB, CB2, ...)                                            \
    do {                                                                              \
        switch (K) {                                                                  \
        case 2: hipLaunchKernelGGL((kern<0, 2, CB, CB2>), __VA_ARGS__); break;        \
        case 4: hipLaunchKernelGGL((kern<0, 4, CB, CB2>), __VA_ARGS__); break;        \
        default: hipLaunchKernelGGL((kern<0, 8, CB, CB2>), __VA_ARGS__); break;       \
        }                                                                             \
    } while (0)

#define LAUNCH_SK_(kern, SURF, K, CB, ...)                                            \
    do {                                                                              \
        if (SURF) {                                                                   \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<1, 2, CB>), __VA_ARGS__); break;         \
            case 4: hipLaunchKernelGGL((kern<1, 4, CB>), __VA_ARGS__); break;         \
            default: hipLaunchKernelGGL((kern<1, 8, CB>), __VA_ARGS__); break;        \
            }                                                                         \
        } else {                                                                      \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<0, 2, CB>), __VA_ARGS__); break;         \
            case 4: hipLaunchKernelGGL((kern<0, 4, CB>), __VA_ARGS__); break;         \
            default: hipLaunchKernelGGL((kern<0, 8, CB>), __VA_ARGS__); break;        \
            }                                                                         \
        }                                                                             \
    } while (0)

#define LAUNCH_SK(kern, SURF, K, ...)                                                 \
    do {                                                                              \
        if (SURF) {                                                                   \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<1, 2>), __VA_ARGS__); break;             \
            case 4: hipLaunchKernelGGL((kern<1, 4>), __VA_ARGS__); break;             \
            default: hipLaunchKernelGGL((kern<1, 8>), __VA_ARGS__); break;            \
            }                                                                         \
        } else {                                                                      \
            switch (K) {                                                              \
            case 2: hipLaunchKernelGGL((kern<0, 2>), __VA_ARGS__); break;             \
            case 4: hipLaunchKernelGGL((kern<0, 4>), __VA_ARGS__); break;             \
            default: hipLaunchKernelGGL((kern<0, 8>), __VA_ARGS__); break;            \
            }                                                                         \
        }                                                                             \
    } while (0)

static int launch_ok(const char* what)
{
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "%s launch failed: %s", what, hipGetErrorString(e));
}

// exclusive scan of `total` counters in place: one workgroup for short arrays, the
// three-phase multi-block scan for long ones (scan_sums holds the tile sums)
static void launch_scan(mph_ctx* c, unsigned* data, int total)
{
    hipStream_t s = c->stream;
    if (total <= 4 * kScanTile) {
        hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, data, total, c->any_flag.p);
        return;
    }
    const int tiles = (total + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, data, total, c->scan_sums.p,
                       c->any_flag.p, nullptr);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, c->scan_sums.p, tiles, c->any_flag.p);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, s, data, total, c->scan_sums.p,
                       c->any_flag.p, nullptr);
}

// the same for a length counted on the device (dtotal[0] <= max_total)
static void launch_scan_dev(mph_ctx* c, unsigned* data, int max_total, const int* dtotal)
{
    hipStream_t s = c->stream;
    const int tiles = std::max(1, (max_total + kScanTile - 1) / kScanTile);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(tiles), dim3(kScanThreads), 0, s, data, max_total, c->scan_sums.p,
                       c->any_flag.p, dtotal);
    hipLaunchKernelGGL(k_scan_exclusive, dim3(1), dim3(1024), 0, s, c->scan_sums.p, tiles, c->any_flag.p);
    hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanThreads), 0, s, data, max_total, c->scan_sums.p,
                       c->any_flag.p, dtotal);
}

// work items per XCD chunk of a single case's list build (0: no XCD chunks)
constexpr int kListChunk = 256;

// workgroups of the list build: the autotuner's choice (c->lcap), else the default cap
static int list_grid(const mph_ctx* c)
{
    const int lcap = c->lcap > 0 ? c->lcap : (c->ncases == 1 ? 20480 : 4096);
    const int g = std::max(1, std::min((c->nslots + 3) / 4, lcap));
    return (g + kXcds - 1) / kXcds * kXcds;          // the kernel's XCD chunks want a multiple of 8
}

// the union lists of the cases flagged for a rebuild (count_only: only their lengths)
static int launch_lists(mph_ctx* c, int count_only)
{
    // 4-wavefront blocks looping over the cluster slots.  A single case rebuilds all its
    // clusters, and more resident rounds balance their uneven cost (318k clusters: 2.45 ms at
    // 4096 blocks, 2.30 at 10240, 2.27 at 20480 = 16 rounds at 5 wavefronts per SIMD).  In a
    // batch a rebuild usually covers a few cases, and the blocks that only skip slots cost
    // more than the balance gains (64-case batch: rebuild phase 0.60 ms at 4096, 0.73 at
    // 20480).  A step without a rebuild exits in about 5 microseconds either way.  For a
    // single case the autotuner times the grids of kListGrids and keeps the fastest.
    const int lgrid = list_grid(c);
    LAUNCH_K(k_cluster_lists, c->K, dim3((unsigned)lgrid), dim3(256), 0, c->stream, c->d_cases.p,
             c->slot_case.p, c->nslots, c->flags.p, c->any_flag.p, c->pos.p, c->stencil.p,
             c->cl_info.p, c->cl_count.p, count_only ? 0 : c->width, c->cl_nbr.p, c->cl_len.p, c->selfin.p, c->nout.p,
             c->max_len.p, c->ncases == 1 ? kListChunk : 0);
    return launch_ok("lists");
}

static int launch_rebuild(mph_ctx* c, int count_only)
{
    const int nb = c->nblocks();
    hipStream_t s = c->stream;
    // a single case sorts all its particles in place (the keys of the last sort stay in the
    // buffer the passes end in); a batch sorts only its flagged cases, compacted (nsort)
    const bool batch = c->ncases > 1;
    unsigned* kf = batch ? c->kfull.p : c->sorted_keys();
    int* pf = batch ? c->pfull.p : c->sorted_perm();
    const int* dn = batch ? c->nsort.p : nullptr;
    if (batch) {
        hipLaunchKernelGGL(k_sort_offsets, dim3(1), dim3(256), 0, s, c->d_cases.p, c->ncases, c->flags.p,
                           c->any_flag.p, c->coff.p, c->nsort.p);
        hipLaunchKernelGGL(k_keys_compact, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->pos.p,
                           c->flags.p, c->any_flag.p, c->rank.p, c->coff.p, c->keys_a.p, c->perm_a.p);
    } else {
        hipLaunchKernelGGL(k_keys, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->pos.p,
                           c->flags.p, c->any_flag.p, c->rank.p, kf, c->keys_a.p, c->perm_a.p);
    }
    for (int ps = 0; ps < c->sort_passes; ++ps) {
        unsigned* kin = (ps & 1) ? c->keys_b.p : c->keys_a.p;
        unsigned* kout = (ps & 1) ? c->keys_a.p : c->keys_b.p;
        int* vin = (ps & 1) ? c->perm_b.p : c->perm_a.p;
        int* vout = (ps & 1) ? c->perm_a.p : c->perm_b.p;
        const int shift = ps * kRadixBits;
        hipLaunchKernelGGL(k_radix_hist, dim3(c->sort_blocks), dim3(kSortThreads), 0, s, kin, c->n,
                           shift, c->hist.p, c->any_flag.p, dn);
        if (batch)
            launch_scan_dev(c, c->hist.p, kRadix * c->sort_blocks, c->nsort.p + 1);
        else
            launch_scan(c, c->hist.p, kRadix * c->sort_blocks);
        hipLaunchKernelGGL(k_radix_scatter, dim3(c->sort_blocks), dim3(kSortThreads), 0, s, kin, vin,
                           kout, vout, c->n, shift, c->hist.p, c->any_flag.p, dn);
    }
    if (batch)
        hipLaunchKernelGGL(k_expand_sorted, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->flags.p,
                           c->any_flag.p, c->coff.p, c->sorted_keys(), c->sorted_perm(), kf, pf);
    hipLaunchKernelGGL(k_cell_start, dim3((c->nkeys + 1 + 255) / 256), dim3(256), 0, s, kf, c->n,
                       c->nkeys, c->start.p, c->any_flag.p);
    hipLaunchKernelGGL(k_stencil, dim3((c->ncells + 7) / 8), dim3(256), 0, s, c->d_cases.p, c->cell_case.p,
                       c->ncells, c->stencil_rank.p, c->start.p, c->flags.p, c->any_flag.p, c->stencil.p);
    hipLaunchKernelGGL(k_gather, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p, c->flags.p,
                       c->any_flag.p, pf, c->pos.p, c->vel.p, c->pidx.p, c->pos_alt.p,
                       c->vel_alt.p, c->pidx_alt.p);
    hipLaunchKernelGGL(k_scatter_back, dim3(nb), dim3(kBlock), 0, s, c->d_cases.p, c->d_bmap.p,
                       c->flags.p, c->any_flag.p, c->pos_alt.p, c->vel_alt.p, c->pidx_alt.p, c->pos.p,
                       c->vel.p, c->pidx.p, c->snap.p);
    const int gcells = (c->ncells + 255) / 256;
    hipLaunchKernelGGL(k_cluster_count, dim3(gcells), dim3(256), 0, s, c->d_cases.p, c->cell_case.p,
                       c->ncells, c->start.p, c->K, c->ccount.p, c->any_flag.p);
    launch_scan(c, c->ccount.p, c->ncells);
    hipLaunchKernelGGL(k_cluster_fill, dim3(gcells), dim3(256), 0, s, c->d_cases.p, c->cell_case.p,
                       c->ncells, c->start.p, c->K, c->ccount.p, c->cl_info.p,
                       c->cl_count.p, c->any_flag.p);
    return launch_lists(c, count_only);
}

// real: a step (advances wall centres); the autotuner launches with 0 so that repeated
// launches leave the state unchanged.  count: count the interacting pairs (timed steps).
// speculative: queued before the host knows whether the lists are rebuilt (returns at once
// on the device if they are).  need_na: sum DensityA also where nothing in the step needs it
// (without surface tension the FP64 step skips it and marks it stale; ensure_dens_a sums it
// when the fields are read)
static int launch_sweep_a(mph_ctx* c, int real, bool count = false, bool speculative = false,
                          bool need_na = false)
{
    const int spec = speculative ? 1 : 0;
    if (c->mixed) {
#define SWEEP_AM_ARGS                                                                                       \
    dim3(c->ncblocks()), dim3(256), 0, c->stream, c->d_cases.p, c->d_cmap.p, c->cl_info.p, c->cl_count.p,  \
        c->cl_nbr.p, c->cl_len.p, c->width, c->nout.p, c->pos.p, c->vel.p, c->aux.p, c->dens_a.p,          \
        c->vstrain.p, c->divp.p, real ? c->pairs.p : c->pairs_scratch.p, c->selfin.p, c->any_flag.p, spec
        const bool cntm = count && real;
        if (!c->any_surface && !need_na) {
            if (cntm) LAUNCH_S0__(k_sweep_a_mixed, c->K, true, false, SWEEP_AM_ARGS);
            else LAUNCH_S0__(k_sweep_a_mixed, c->K, false, false, SWEEP_AM_ARGS);
            c->dens_stale = 1;
        } else {
            if (cntm) LAUNCH_SK__(k_sweep_a_mixed, c->any_surface, c->K, true, true, SWEEP_AM_ARGS);
            else LAUNCH_SK__(k_sweep_a_mixed, c->any_surface, c->K, false, true, SWEEP_AM_ARGS);
            c->dens_stale = 0;
        }
#undef SWEEP_AM_ARGS
        return launch_ok("sweep A");
    }
    const int mode = (c->same_ap ? 1 : 0) | (c->ratio_one ? 2 : 0);
#define SWEEP_A_ARGS                                                                                        \
    dim3(c->any_surface ? c->ncblocks() : c->nablocks()), dim3(c->any_surface ? 256 : 512), 0, c->stream, c->d_cases.p, c->d_cmap.p, c->cl_info.p, c->cl_count.p,  \
        c->cl_nbr.p, c->cl_len.p, c->width, c->nout.p, c->pos.p, c->vel.p, c->aux.p, c->dens_a.p,          \
        c->vstrain.p, c->divp.p, real ? c->pairs.p : c->pairs_scratch.p, c->selfin.p, c->any_flag.p, spec
    const bool cnt = count && real;
#define SWEEP_A_MODE(M)                                                                            \
    do {                                                                                           \
        if (cnt) LAUNCH_SK__(k_sweep_a, c->any_surface, c->K, true, M, SWEEP_A_ARGS);              \
        else LAUNCH_SK__(k_sweep_a, c->any_surface, c->K, false, M, SWEEP_A_ARGS);                 \
    } while (0)
#define SWEEP_A_MODE0(M)                                                                           \
    do {                                                                                           \
        if (cnt) LAUNCH_S0__(k_sweep_a, c->K, true, M, SWEEP_A_ARGS);                              \
        else LAUNCH_S0__(k_sweep_a, c->K, false, M, SWEEP_A_ARGS);                                 \
    } while (0)
    if (!c->any_surface && !need_na) {
        switch (mode) {
        case 0: SWEEP_A_MODE0(4); break;
        case 1: SWEEP_A_MODE0(5); break;
        case 2: SWEEP_A_MODE0(6); break;
        default: SWEEP_A_MODE0(7); break;
        }
        c->dens_stale = 1;
    } else {
        switch (mode) {
        case 0: SWEEP_A_MODE(0); break;
        case 1: SWEEP_A_MODE(1); break;
        case 2: SWEEP_A_MODE(2); break;
        default: SWEEP_A_MODE(3); break;
        }
        c->dens_stale = 0;
    }
#undef SWEEP_A_MODE0
#undef SWEEP_A_MODE
#undef SWEEP_A_ARGS
    return launch_ok("sweep A");
}

// DensityA of the current positions and lists, where the last step did not sum it: sweep A
// again with n^A (its other outputs are rewritten with the same values)
static int ensure_dens_a(mph_ctx* c)
{
    return c->dens_stale ? launch_sweep_a(c, 0, false, false, true) : MPH_OK;
}

static int launch_sweep_b(mph_ctx* c, int real)
{
    if (c->mixed) {
        LAUNCH_SKC(k_sweep_b_mixed, c->any_surface, c->K, c->same_pv, dim3(c->ncblocks()), dim3(256), 0,
                   c->stream, c->d_cases.p, c->d_cmap.p, c->cl_info.p, c->cl_count.p, c->cl_nbr.p, c->cl_len.p,
                   c->width, c->nout.p, c->pos.p, c->vel.p, c->aux.p, c->vnext_out(real), c->force_out(real),
                   c->flags.p, c->any_flag.p, c->rebuilds.p, c->wallc.p, real ? c->ncases : 0);
        return launch_ok("sweep B");
    }
    LAUNCH_SKC(k_sweep_b, c->any_surface, c->K, c->same_pv, dim3(c->ncblocks()), dim3(256), 0, c->stream,
               c->d_cases.p,
              c->d_cmap.p, c->cl_info.p, c->cl_count.p, c->cl_nbr.p, c->cl_len.p, c->width, c->nout.p,
              c->pos.p, c->vel.p, c->aux.p, c->vnext_out(real), c->force_out(real), c->flags.p,
              c->any_flag.p, c->rebuilds.p, c->wallc.p, real ? c->ncases : 0);
    return launch_ok("sweep B");
}

static int launch_move(mph_ctx* c)
{
    hipLaunchKernelGGL(k_move, dim3(c->nblocks()), dim3(kBlock), 0, c->stream, c->d_cases.p, c->d_bmap.p,
                       c->pos.p, c->vel.p, c->vnext.p, c->snap.p, c->wallc.p, c->flags.p, c->any_flag.p);
    return launch_ok("move");
}

static int launch_step(mph_ctx* c)
{
    int rc = launch_move(c);
    if (!rc) rc = launch_rebuild(c, 0);
    if (!rc) rc = launch_sweep_a(c, 1);
    if (!rc) rc = launch_sweep_b(c, 1);
    return rc;
}

static int read_int(mph_ctx* c, const int* dptr, int* out)
{
    HIP_TRY(hipMemcpyAsync(c->h_scalar, dptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *out = *c->h_scalar;
    return MPH_OK;
}

static int read_ll(mph_ctx* c, const long long* dptr, long long* out)
{
    long long* h = reinterpret_cast<long long*>(c->h_scalar) + 2;
    HIP_TRY(hipMemcpyAsync(h, dptr, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *out = *h;
    return MPH_OK;
}

// A list build of this step cut a list at the reserved width (any_flag[1]; every later
// kernel of the step has returned at once).  Widen the table geometrically, to 1.25x the
// longest list, carry the other cases' lists over, build the flagged cases' lists again and
// clear the halt.  The reference sizes NeighborInd at every rebuild (main.cpp:1487-1495,
// 1598-1602); here the table only grows.
static int grow_lists(mph_ctx* c)
{
    for (int tries = 0; tries < 8; ++tries) {
        int mx = 0;
        int rc = read_int(c, c->max_len.p, &mx);
        if (rc) return rc;
        const int nw = std::max(c->width + 16, ((int)std::ceil(mx * 1.25) + 16 + 15) / 16 * 16);
        DevBuf<int> wider;
        if ((rc = wider.alloc((size_t)nw * c->nslots))) return rc;
        const long long total = (long long)c->width * c->nslots;
        const unsigned grid = (unsigned)std::max(1ll, std::min((total + 255) / 256, 65536ll));
        hipLaunchKernelGGL(k_widen_lists, dim3(grid), dim3(256), 0, c->stream, c->cl_nbr.p, c->width, wider.p, nw,
                           total);
        if ((rc = launch_ok("widen lists"))) return rc;
        std::swap(c->cl_nbr.p, wider.p);
        std::swap(c->cl_nbr.n, wider.n);    // `wider` frees the old table (hipFree waits for the copy)
        c->width = nw;
        c->drop_graph();                    // captured launches hold the old table and width
        HIP_TRY(hipMemsetAsync(c->any_flag.p + 1, 0, sizeof(int), c->stream));
        HIP_TRY(hipMemsetAsync(c->max_len.p, 0, sizeof(int), c->stream));
        if ((rc = launch_lists(c, 0))) return rc;
        int halt = 0;
        if ((rc = read_int(c, c->any_flag.p + 1, &halt))) return rc;
        if (!halt) return MPH_OK;
    }
    return fail(MPH_ERR_CAPACITY, "the cluster neighbour lists could not be widened enough");
}

// finish a step whose list build overflowed: wider lists, then its sweeps
static int finish_halted_step(mph_ctx* c, bool count = false)
{
    int rc = grow_lists(c);
    if (!rc) rc = launch_sweep_a(c, 1, count);
    if (!rc) rc = launch_sweep_b(c, 1);
    return rc;
}

// launch_step, then finish it if its list build overflowed (synchronises)
static int launch_step_checked(mph_ctx* c)
{
    int rc = launch_step(c);
    int halt = 0;
    if (!rc) rc = read_int(c, c->any_flag.p + 1, &halt);
    if (!rc && halt) rc = finish_halted_step(c);
    return rc;
}

// launch_rebuild, then widen the lists if it overflowed (synchronises)
static int launch_rebuild_checked(mph_ctx* c)
{
    int rc = launch_rebuild(c, 0);
    int halt = 0;
    if (!rc) rc = read_int(c, c->any_flag.p + 1, &halt);
    if (!rc && halt) rc = grow_lists(c);
    return rc;
}

static void advance_clock(mph_ctx* c, int nsteps)
{
    for (int k = 0; k < nsteps; ++k)
        for (auto& ch : c->cases) ch.time += ch.prm.dt;   // Time += Dt, main.cpp:449
    c->steps += nsteps;
}

// The eager step decides the rebuild on the host, as the reference's loop does
// (main.cpp:386-395): the Verlet flag of the move is read back and the 25 rebuild launches
// are issued only when it is set.  Launched unconditionally they return at once without a
// rebuild, but each costs about 4.4 us on the GPU (0.11 ms per step at 1.27M particles).
// Sweep A is queued before the flag is known (speculative): it runs unless this step
// rebuilds the lists, when it returns at once on the device and is queued again after the
// rebuild, so the GPU does not wait for the host.  After a rebuild the host reads whether a
// list overflowed the width while the sweeps run, and if one did (the sweeps then returned
// at once) widens the lists and finishes the step.  Graph replay keeps the device-side
// decision.
static int wait_event(hipEvent_t e)
{
    hipError_t q;
    while ((q = hipEventQuery(e)) == hipErrorNotReady) {
    }
    return q == hipSuccess ? MPH_OK : fail(MPH_ERR_DEVICE, "step: %s", hipGetErrorString(q));
}

static int launch_step_host(mph_ctx* c)
{
    int rc = launch_move(c);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(c->h_scalar, c->any_flag.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_flag, c->stream));
    if ((rc = launch_sweep_a(c, 1, false, true))) return rc;
    if ((rc = wait_event(c->ev_flag))) return rc;
    if (!c->h_scalar[0]) return launch_sweep_b(c, 1);
    if ((rc = launch_rebuild(c, 0))) return rc;
    HIP_TRY(hipMemcpyAsync(c->h_scalar + 1, c->any_flag.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipEventRecord(c->ev_cap, c->stream));
    if ((rc = launch_sweep_a(c, 1)) || (rc = launch_sweep_b(c, 1))) return rc;
    if ((rc = wait_event(c->ev_cap))) return rc;
    return c->h_scalar[1] ? finish_halted_step(c) : MPH_OK;
}

// end-of-call check: every overflow is resolved inside the step that had it, so a halt
// left over here is an internal error
static int check_capacity(mph_ctx* c)
{
    int halt = 0;
    int rc = read_int(c, c->any_flag.p + 1, &halt);
    if (rc) return rc;
    if (halt)
        return fail(MPH_ERR_CAPACITY, "a cluster neighbour list overflowed the reserved width %d and was not "
                    "widened; results of this call are invalid", c->width);
    return MPH_OK;
}

// Cluster slots, block map and lists for cluster size K: each case reserves
// ceil(n/K) + cells slots (every cell adds at most one partial cluster), rounded up to a
// multiple of 4.  All cases are
// flagged for a rebuild, the union lists are counted to size their width, then built.
static int prepare_clusters(mph_ctx* c, int K)
{
    c->K = K;
    std::vector<int> hslot_case;
    std::vector<int2> logical;
    int base = 0;
    for (int k = 0; k < c->ncases; ++k) {
        CaseConst& kc = c->cases[k].k;
        const int nk = kc.end - kc.begin;
        kc.cl_base = base;
        // a multiple of 4: the four slots of a map entry stay inside the case (cluster_frame
        // loads a slot's entry before it tests the slot against the count)
        kc.cl_cap = ((nk + K - 1) / K + kc.tcc + 3) / 4 * 4;
        for (int q = 0; q < kc.cl_cap; ++q) hslot_case.push_back(k);
        for (int q = 0; q < kc.cl_cap; q += 4) logical.push_back(make_int2(k, base + q));   // global slot
        base += kc.cl_cap;
    }
    c->nslots = base;
    c->cmap_host = xcd_order(logical);
    // eight padding entries: sweep A's workgroups take entries in pairs (k_sweep_a)
    for (int q = 0; q < kXcds; ++q) c->cmap_host.push_back(make_int2(-1, 0));
    std::vector<CaseConst> hk(c->ncases);
    for (int k = 0; k < c->ncases; ++k) hk[k] = c->cases[k].k;
    int rc = 0;
    if ((rc = c->slot_case.alloc(c->nslots)) || (rc = c->cl_info.alloc(c->nslots)) ||
        (rc = c->cl_len.alloc(c->nslots)) || (rc = c->d_cmap.alloc(c->cmap_host.size())))
        return rc;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemcpyAsync(c->d_cases.p, hk.data(), sizeof(CaseConst) * c->ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->slot_case.p, hslot_case.data(), sizeof(int) * c->nslots, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->d_cmap.p, c->cmap_host.data(), sizeof(int2) * c->cmap_host.size(),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(c->cl_len.p, 0, sizeof(int) * c->nslots, s));
    std::vector<int> ones(c->ncases, 1);
    HIP_TRY(hipMemcpyAsync(c->flags.p, ones.data(), sizeof(int) * c->ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->any_flag.p, ones.data(), sizeof(int), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(c->max_len.p, 0, sizeof(int), s));
    if ((rc = c->cl_nbr.alloc(1))) return rc;
    c->width = 0;
    if ((rc = launch_rebuild(c, 1))) return rc;
    int mx = 0;
    if ((rc = read_int(c, c->max_len.p, &mx))) return rc;
    c->width = ((int)std::ceil(mx * 1.25) + 16 + 15) / 16 * 16;
    if ((rc = c->cl_nbr.alloc((size_t)c->width * c->nslots))) return rc;
    HIP_TRY(hipMemsetAsync(c->max_len.p, 0, sizeof(int), s));
    return launch_rebuild(c, 0);      // the sort is the identity now
}

// ---------------------------------------------------------------------------------------
// autotuner: time the sweeps for every cluster size on the context's own particles and
// keep the fastest; for a single case, then time full list builds at several grid sizes
// and keep the fastest.  Results are cached in a text file (MPH_TUNE_CACHE, default
// ~/.cache/mphexplicit_amd/tiles.txt, lines "key K grid") keyed by device, dimension, surface
// tension, precision, case count and problem size; MPH_TUNE=0 keeps K = 4, MPH_TUNE=<2|4|8>
// forces a size (the grid keeps its default) and MPH_TUNE=force ignores the cache.

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
    std::snprintf(buf, sizeof(buf), "%s|dim%d|surf%d|mixed%d|cases%d|n2^%d", arch.c_str(), dim, c->any_surface,
                  c->mixed, c->ncases, nb);
    std::string k = buf;
    for (auto& ch : k) if (ch == ' ') ch = '_';
    return k;
}

static std::string tune_path()
{
    const char* p = std::getenv("MPH_TUNE_CACHE");
    if (p && *p) return p;
    const char* h = std::getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/mphexplicit_amd/tiles.txt";
}

static int time_sweeps(mph_ctx* c, float* ms_out)
{
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = MPH_OK;
    float total = 0.0f;
    for (int rep = 0; rep < 4 && rc == MPH_OK; ++rep) {
        (void)hipEventRecord(e0, c->stream);
        rc = launch_sweep_a(c, 0);
        if (!rc) rc = launch_sweep_b(c, 0);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) total += ms;           // the first launch warms caches
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = total;
    return rc;
}

// a full build of every list, timed (the cases are flagged for a rebuild, so each launch
// rebuilds the same lists)
static int time_lists(mph_ctx* c, float* ms_out)
{
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = MPH_OK;
    float total = 0.0f;
    for (int rep = 0; rep < 4 && rc == MPH_OK; ++rep) {
        (void)hipEventRecord(e0, c->stream);
        rc = launch_lists(c, 0);
        (void)hipEventRecord(e1, c->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (rep > 0) total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_out = total;
    return rc;
}

// list-build grids the tuner tries for a single case (4 wavefronts per workgroup looping
// over the cluster slots); batches keep 4096 workgroups, where a rebuild usually covers a
// few cases and blocks that only skip slots cost more than the balance gains
constexpr int kListGrids[] = {4096, 10240, 20480, 40960};

static int choose_cluster_size(mph_ctx* c, int* K)
{
    const char* mode = std::getenv("MPH_TUNE");
    *K = 4;
    if (mode && !std::strcmp(mode, "0")) return MPH_OK;
    if (mode && (!std::strcmp(mode, "2") || !std::strcmp(mode, "4") || !std::strcmp(mode, "8"))) {
        *K = std::atoi(mode);
        return MPH_OK;
    }
    const std::string key = tune_key(c), path = tune_path();
    if (!(mode && !std::strcmp(mode, "force"))) {
        if (FILE* f = std::fopen(path.c_str(), "r")) {
            char k[256];
            int v, g;
            while (std::fscanf(f, "%255s %d %d", k, &v, &g) == 3)
                if (key == k && (v == 2 || v == 4 || v == 8) && g >= 0) {
                    *K = v;
                    c->lcap = g;
                    std::fclose(f);
                    return MPH_OK;
                }
            std::fclose(f);
        }
    }
    int rc = 0;
    if ((rc = c->tune_v.alloc(c->n)) || (rc = c->tune_f.alloc(c->n))) return rc;
    float best = 1e30f;
    for (int q = 0; q < kNumClusterSizes && rc == MPH_OK; ++q) {
        rc = prepare_clusters(c, kClusterSizes[q]);
        if (!rc) rc = launch_sweep_a(c, 0);     // densities and pressures for sweep B's inputs
        float ms = 0.0f;
        if (!rc) rc = time_sweeps(c, &ms);
        if (!rc && ms < best) { best = ms; *K = kClusterSizes[q]; }
    }
    c->tune_v.release();
    c->tune_f.release();
    // the list build's grid for the chosen size, timed on full rebuilds of this case
    if (!rc && c->ncases == 1) {
        rc = prepare_clusters(c, *K);
        float bl = 1e30f;
        int bg = 0;
        for (int q = 0; q < (int)(sizeof(kListGrids) / sizeof(kListGrids[0])) && rc == MPH_OK; ++q) {
            c->lcap = kListGrids[q];
            float ms = 0.0f;
            rc = time_lists(c, &ms);
            if (!rc && ms < bl) { bl = ms; bg = kListGrids[q]; }
        }
        c->lcap = bg;
    }
    if (rc) return rc;
    // the cache keeps later contexts of the same shape on the same size (sizes round
    // differently); failing to write it is reported, and this context keeps its choice
    const size_t cut = path.rfind('/');
    std::error_code ec;
    if (cut != std::string::npos && cut > 0) std::filesystem::create_directories(path.substr(0, cut), ec);
    FILE* f = ec ? nullptr : std::fopen(path.c_str(), "a");
    if (f) {
        std::fprintf(f, "%s %d %d\n", key.c_str(), *K, c->lcap);
        std::fclose(f);
    } else {
        std::fprintf(stderr, "mphexplicit_amd: cannot write the tile cache %s (%s)\n", path.c_str(),
                     ec ? ec.message().c_str() : std::strerror(errno));
    }
    return MPH_OK;
}

static int resolve_scratch(int device);

static int setup(mph_ctx* c, int ncases, const mph_params* prm, const int* counts, const int* type,
                 const double* x, const double* v)
{
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count <= 0)
        return fail(MPH_ERR_DEVICE, "no GPU available");
    HIP_TRY(hipGetDevice(&c->device));
    int rs = resolve_scratch(c->device);
    if (rs) return rs;
    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_flag, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&c->ev_cap, hipEventDisableTiming));
    HIP_TRY(hipHostMalloc((void**)&c->h_scalar, sizeof(long long) * 4, hipHostMallocDefault));
    c->ncases = ncases;
    c->cases.resize(ncases);
    long long total = 0;
    for (int k = 0; k < ncases; ++k) {
        if (counts[k] <= 0) return fail(MPH_ERR_ARG, "case %d has no particles", k);
        total += counts[k];
    }
    if (total > (1ll << 30)) return fail(MPH_ERR_ARG, "too many particles");
    c->n = (int)total;
    std::vector<double4> hpos(c->n), hvel(c->n);
    std::vector<int> hidx(c->n);
    std::vector<double> hwall((size_t)ncases * kTypes * 3, 0.0);
    int begin = 0, key_base = 0;
    std::vector<int> hrank, hcell_case, hstencil;
    std::vector<int2> logical;
    c->any_surface = 0;
    c->same_pv = 1;
    c->same_ap = 1;
    c->ratio_one = 1;
    c->mixed = prm[0].precision == MPH_MIXED;
    for (int k = 0; k < ncases; ++k) {
        const int nk = counts[k];
        if (prm[k].precision != MPH_FP64 && prm[k].precision != MPH_MIXED)
            return fail(MPH_ERR_ARG, "case %d: unknown precision %d", k, prm[k].precision);
        if ((prm[k].precision == MPH_MIXED) != (c->mixed != 0))
            return fail(MPH_ERR_ARG, "case %d: all cases of a batch must use the same precision", k);
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
        if (((long long)key_base + 2ll * ch.k.tcc + 1) << kSubBits >= (1ll << 32))
            return fail(MPH_ERR_ARG, "too many cells for 32-bit sort keys");
        key_base += 2 * ch.k.tcc;
        ch.time = prm[k].time;
        c->any_surface |= ch.k.surface;
        if (ch.k.inv_hp != ch.k.inv_hv) c->same_pv = 0;
        if (ch.k.inv_ha != ch.k.inv_hp) c->same_ap = 0;
        if (!ch.k.ratio_one) c->ratio_one = 0;
        for (int t = 0; t < kTypes; ++t)
            for (int d = 0; d < 3; ++d) hwall[((size_t)k * kTypes + t) * 3 + d] = prm[k].wall_center[t][d];
        for (int i = 0; i < nk; ++i) {
            const int g = begin + i;
            for (int d = 0; d < 3; ++d) {
                const double xv = x[3 * (size_t)g + d];
                if (!(xv >= ch.k.dmin[d] - 0.25 * ch.k.dwid[d] && xv <= ch.k.dmax[d] + 0.25 * ch.k.dwid[d]))
                    return fail(MPH_ERR_ARG, "case %d particle %d lies outside the periodic domain", k, i);
            }
            const long long tag = (long long)(type[g] | (k << 2));
            double w;
            std::memcpy(&w, &tag, sizeof(w));
            hpos[g] = make_double4(x[3 * g], x[3 * g + 1], x[3 * g + 2], w);
            hvel[g] = make_double4(v[3 * g], v[3 * g + 1], v[3 * g + 2], 0.0);
            hidx[g] = i;
        }
        std::vector<int> rk, cr;
        morton_ranks(ch.k, rk, cr);
        stencil_table(ch.k, rk, cr, hstencil);
        ch.k.rank_base = (int)hrank.size();
        hrank.insert(hrank.end(), rk.begin(), rk.end());
        hcell_case.insert(hcell_case.end(), (size_t)ch.k.tcc, k);
        for (int b = 0; b < nk; b += kBlock) logical.push_back(make_int2(k, begin + b));
        begin += nk;
    }
    c->bmap_host = xcd_order(logical);
    c->ncells = (int)hrank.size();
    c->nkeys = key_base;
    int bits = 1;
    while ((1ll << bits) < ((long long)c->nkeys << kSubBits)) ++bits;
    c->sort_passes = (bits + kRadixBits - 1) / kRadixBits;
    c->sort_blocks = (c->n + kSortTile - 1) / kSortTile;

    int rc = 0;
    const size_t n = c->n;
    if ((rc = c->d_cases.alloc(ncases)) || (rc = c->d_bmap.alloc(c->bmap_host.size())) ||
        (rc = c->pos.alloc(n)) || (rc = c->pos_alt.alloc(n)) || (rc = c->vel.alloc(n)) ||
        (rc = c->vel_alt.alloc(n)) || (rc = c->vnext.alloc(n)) || (rc = c->aux.alloc(n)) ||
        (rc = c->force.alloc(n)) || (rc = c->snap.alloc(n)) || (rc = c->pidx.alloc(n)) ||
        (rc = c->pidx_alt.alloc(n)) || (rc = c->nout.alloc(n)) || (rc = c->selfin.alloc(n)) ||
        (rc = c->start.alloc((size_t)c->nkeys + 1)) || (rc = c->dens_a.alloc(n)) ||
        (rc = c->vstrain.alloc(n)) || (rc = c->divp.alloc(n)) || (rc = c->vir.alloc(n)) ||
        (rc = c->wallc.alloc(hwall.size())) || (rc = c->keys_a.alloc(n)) || (rc = c->keys_b.alloc(n)) ||
        (rc = c->hist.alloc((size_t)kRadix * c->sort_blocks)) || (rc = c->perm_a.alloc(n)) ||
        (rc = c->perm_b.alloc(n)) || (rc = c->flags.alloc(ncases)) || (rc = c->any_flag.alloc(2)) ||
        (rc = c->rank.alloc(hrank.size())) ||
        (rc = c->cell_case.alloc(c->ncells)) || (rc = c->ccount.alloc(c->ncells)) ||
        (rc = c->stencil.alloc((size_t)27 * c->ncells)) || (rc = c->stencil_rank.alloc(hstencil.size())) ||
        (rc = c->scan_sums.alloc((size_t)std::max(kRadix * c->sort_blocks, c->ncells) / kScanTile + 2)) ||
        (rc = c->cl_count.alloc(ncases)) || (rc = c->max_len.alloc(1)) ||
        (rc = c->rebuilds.alloc(ncases + 1)) || (rc = c->pairs.alloc(kPairSlots * kPairStride)) ||
        (rc = c->pairs_scratch.alloc(kPairSlots * kPairStride)))
        return rc;
    if (ncases > 1 && ((rc = c->kfull.alloc(n)) || (rc = c->pfull.alloc(n)) || (rc = c->coff.alloc(ncases + 1)) ||
                       (rc = c->nsort.alloc(2))))
        return rc;
    hipStream_t s = c->stream;
    HIP_TRY(hipMemsetAsync(c->any_flag.p, 0, 2 * sizeof(int), s));
    HIP_TRY(hipMemcpyAsync(c->d_bmap.p, c->bmap_host.data(), sizeof(int2) * c->bmap_host.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->pos.p, hpos.data(), sizeof(double4) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->vel.p, hvel.data(), sizeof(double4) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->pidx.p, hidx.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->wallc.p, hwall.data(), sizeof(double) * hwall.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->rank.p, hrank.data(), sizeof(int) * hrank.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(c->stencil_rank.p, hstencil.data(), sizeof(int) * hstencil.size(), hipMemcpyHostToDevice,
                           s));
    HIP_TRY(hipMemcpyAsync(c->cell_case.p, hcell_case.data(), sizeof(int) * hcell_case.size(), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemsetAsync(c->aux.p, 0, sizeof(double4) * n, s));
    HIP_TRY(hipMemsetAsync(c->force.p, 0, sizeof(double4) * n, s));
    HIP_TRY(hipMemsetAsync(c->vir.p, 0, sizeof(double) * n, s));
    HIP_TRY(hipMemsetAsync(c->rebuilds.p, 0, sizeof(long long) * (ncases + 1), s));
    HIP_TRY(hipMemsetAsync(c->pairs.p, 0, sizeof(unsigned long long) * kPairSlots * kPairStride, s));
    // first bucketing and lists (main.cpp:358-359), for the cluster size the autotuner
    // picks (it builds and times every size, the last build is then redone for the winner)
    int K = 4;
    if ((rc = choose_cluster_size(c, &K))) return rc;
    if ((rc = prepare_clusters(c, K))) return rc;
    if ((rc = launch_sweep_a(c, 0))) return rc;                          // main.cpp:362-364
    HIP_TRY(hipMemcpyAsync(c->vnext.p, c->vel.p, sizeof(double4) * n, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(c->flags.p, 0, sizeof(int) * ncases, s));
    HIP_TRY(hipMemsetAsync(c->any_flag.p, 0, 2 * sizeof(int), s));
    std::vector<long long> one_ll(ncases, 1);
    HIP_TRY(hipMemcpyAsync(c->rebuilds.p, one_ll.data(), sizeof(long long) * ncases, hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return check_capacity(c);
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
        int rc = launch_step_host(c);
        if (rc) return rc;
        advance_clock(c, 1);
    }
    return check_capacity(c);
}

// Steps launched on the device-side decision (graph replays, timed steps) run on without the
// host; a list overflow halts the rest of the call's launches on the device.  Afterwards the
// completed-step counter (rebuilds[ncases], counted by sweep B) tells how many steps ran to
// the end; the halted step is finished here and the caller launches the remaining ones.
static int steps_done(mph_ctx* c, long long* out) { return read_ll(c, c->rebuilds.p + c->ncases, out); }

static int resolve_halt(mph_ctx* c, long long done0, int launched, int* ran, bool count = false)
{
    int halt = 0, rc = read_int(c, c->any_flag.p + 1, &halt);
    if (rc) return rc;
    *ran = launched;
    if (!halt) return MPH_OK;
    long long done1 = 0;
    if ((rc = steps_done(c, &done1))) return rc;
    *ran = (int)(done1 - done0) + 1;
    return finish_halted_step(c, count);
}

extern "C" int mph_step_graph(mph_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return fail(MPH_ERR_ARG, "mph_step_graph: bad argument");
    int left = nsteps;
    while (left > 0) {
        if (!c->graph) {
            HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = launch_step(c);
            c->graph_stale = c->dens_stale;      // the captured step sums DensityA or not
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (e != hipSuccess) return fail(MPH_ERR_DEVICE, "graph capture failed: %s", hipGetErrorString(e));
            c->graph_src = g;
            HIP_TRY(hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0));
        }
        long long done0 = 0;
        int rc = steps_done(c, &done0), ran = 0;
        if (rc) return rc;
        for (int k = 0; k < left; ++k) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
        c->dens_stale = c->graph_stale;
        if ((rc = resolve_halt(c, done0, left, &ran))) return rc;
        advance_clock(c, ran);
        left -= ran;
    }
    return check_capacity(c);
}

static int step_timed_once(mph_ctx* c, int nsteps, mph_timing* out, int* ran)
{
    long long done0 = 0;
    int rc = steps_done(c, &done0);
    if (rc) return rc;
    std::vector<hipEvent_t> ev((size_t)nsteps * 5, nullptr);
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    for (int k = 0; k < nsteps && rc == MPH_OK; ++k) {
        hipEvent_t* e = &ev[(size_t)k * 5];
        (void)hipEventRecord(e[0], c->stream);
        rc = launch_move(c);
        (void)hipEventRecord(e[1], c->stream);
        if (!rc) rc = launch_rebuild(c, 0);
        (void)hipEventRecord(e[2], c->stream);
        if (!rc) rc = launch_sweep_a(c, 1, true);
        (void)hipEventRecord(e[3], c->stream);
        if (!rc) rc = launch_sweep_b(c, 1);
        (void)hipEventRecord(e[4], c->stream);
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
    return resolve_halt(c, done0, nsteps, ran, true);
}

extern "C" int mph_step_timed(mph_ctx* c, int nsteps, mph_timing* out)
{
    if (!c || nsteps < 0 || !out) return fail(MPH_ERR_ARG, "mph_step_timed: bad argument");
    std::memset(out, 0, sizeof(*out));
    std::vector<long long> rb0(c->ncases), rb1(c->ncases);
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(rb0.data(), c->rebuilds.p, sizeof(long long) * c->ncases, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemsetAsync(c->pairs.p, 0, sizeof(unsigned long long) * kPairSlots * kPairStride, c->stream));
    for (int left = nsteps; left > 0;) {
        int ran = 0, rc = step_timed_once(c, left, out, &ran);
        if (rc) return rc;
        advance_clock(c, ran);
        left -= ran;
    }
    std::vector<unsigned long long> slots((size_t)kPairSlots * kPairStride);
    HIP_TRY(hipMemcpy(slots.data(), c->pairs.p, sizeof(unsigned long long) * slots.size(), hipMemcpyDeviceToHost));
    unsigned long long pairs = 0;
    for (int k = 0; k < kPairSlots; ++k) pairs += slots[(size_t)k * kPairStride];
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
    int rc = ensure_dens_a(c);
    if (rc) return rc;
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

extern "C" int mph_set_time_step(mph_ctx* c, int case_id, double dt)
{
    if (!c || case_id < 0 || case_id >= c->ncases || !(dt > 0.0) || !std::isfinite(dt))
        return fail(MPH_ERR_ARG, "mph_set_time_step: bad argument");
    c->cases[case_id].prm.dt = dt;
    c->cases[case_id].k.dt = dt;
    HIP_TRY(hipMemcpyAsync(c->d_cases.p + case_id, &c->cases[case_id].k, sizeof(CaseConst), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPH_OK;
}

extern "C" int mph_compute_virial(mph_ctx* c)
{
    if (!c) return fail(MPH_ERR_ARG, "null context");
    LAUNCH_K(k_virial, c->K, dim3(c->ncblocks()), dim3(256), 0, c->stream, c->d_cases.p, c->d_cmap.p,
             c->cl_info.p, c->cl_count.p, c->cl_nbr.p, c->cl_len.p, c->width, c->nout.p, c->pos.p, c->vel.p,
             c->vnext.p, c->aux.p, c->vir.p);
    return launch_ok("virial");
}

// the reference's periodic difference (main.cpp:70, 1398) on the host, for list export
static double host_rel(double xj, double xi, double hw, double w)
{
    const double a = xj - xi + hw;
    return (a - w * std::floor(a / w)) - hw;
}

extern "C" int mph_get_neighbors(mph_ctx* c, int case_id, long long* ptr, int* ind)
{
    if (!c || !ptr || case_id < 0 || case_id >= c->ncases) return fail(MPH_ERR_ARG, "mph_get_neighbors: bad argument");
    const CaseConst& k = c->cases[case_id].k;
    const int b = k.begin, m = k.end - k.begin;
    std::vector<int> idx, nc, len;
    std::vector<double4> sn;
    std::vector<int4> info;
    int rc = 0;
    if ((rc = fetch(c, c->pidx.p, b, m, idx)) || (rc = fetch(c, c->nout.p, b, m, nc)) ||
        (rc = fetch(c, c->snap.p, b, m, sn)) || (rc = fetch(c, c->cl_info.p, k.cl_base, k.cl_cap, info)) ||
        (rc = fetch(c, c->cl_len.p, k.cl_base, k.cl_cap, len)))
        return rc;
    int ncl = 0;
    HIP_TRY(hipMemcpyAsync(c->h_scalar, c->cl_count.p + case_id, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    ncl = *c->h_scalar;
    std::vector<int> lists((size_t)std::max(ncl, 1) * c->width);
    if (ncl > 0)
        HIP_TRY(hipMemcpy(lists.data(), c->cl_nbr.p + (size_t)k.cl_base * c->width,
                          sizeof(int) * (size_t)ncl * c->width, hipMemcpyDeviceToHost));
    // rows by original id: the reference's lists include self (when the particle has one)
    std::vector<long long> rowlen(m, 0);
    for (int i = 0; i < m; ++i) rowlen[idx[i]] = nc[i];
    ptr[0] = 0;
    for (int o = 0; o < m; ++o) ptr[o + 1] = ptr[o] + rowlen[o];
    if (!ind) return MPH_OK;
    for (int g = 0; g < ncl; ++g) {
        const int4 in = info[g];
        for (int r = 0; r < in.y; ++r) {
            const int il = in.x + r - b;
            if (il < 0 || il >= m) return fail(MPH_ERR_STATE, "cluster member outside its case");
            const int o = idx[il];
            if (nc[il] == 0) continue;
            std::vector<int> row;
            for (int e = 0; e < len[g]; ++e) {
                const int jl = lists[(size_t)g * c->width + e] - b;
                if (jl < 0 || jl >= m) return fail(MPH_ERR_STATE, "neighbour index outside its case");
                if (jl == il) continue;
                const double dx = host_rel(sn[jl].x, sn[il].x, k.hwid[0], k.dwid[0]);
                const double dy = host_rel(sn[jl].y, sn[il].y, k.hwid[1], k.dwid[1]);
                const double dz = host_rel(sn[jl].z, sn[il].z, k.hwid[2], k.dwid[2]);
                if (dx * dx + dy * dy + dz * dz <= k.cut2) row.push_back(idx[jl]);
            }
            // the particle itself, once per visit of its own cell (more than once where the
            // stencil repeats cells on an axis of fewer than three)
            const long long nself = rowlen[o] - (long long)row.size();
            if (nself < 1)
                return fail(MPH_ERR_STATE, "neighbour row length mismatch (%zu others vs %lld)", row.size(), rowlen[o]);
            row.insert(row.end(), (size_t)nself, o);
            std::sort(row.begin(), row.end());
            std::copy(row.begin(), row.end(), ind + ptr[o]);
        }
    }
    return MPH_OK;
}

// ---------------------------------------------------------------------------------------
// Allocation-free entry point: mph_get_neighbors on the device, into the caller's device
// buffers.  Its scratch (the scan's tile sums and an error counter) is this file-scope
// array; every GPU holds its own copy until the process ends, and the host addresses it only
// through the pointer hipGetSymbolAddress gives for each device (resolved once per device,
// at the first mph_create on it).  Calls that overlap on one GPU share it and race.
constexpr int kScratchBytes = 64 * 1024;
constexpr int kScratchHead = 256;                // error counter, then the tile sums
constexpr int kScratchTiles = (kScratchBytes - kScratchHead) / (int)sizeof(long long);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static int resolve_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(MPH_ERR_DEVICE, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return MPH_OK;
    void* p = nullptr;
    HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    return MPH_OK;
}

static bool is_device_pointer(const void* p)
{
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();            // an unregistered host pointer leaves an error behind
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

extern "C" int mph_get_neighbors_noalloc(mph_ctx* c, int case_id, long long* ptr, int* ind, long long ind_capacity,
                                         long long* total)
{
    if (!c || !ptr || case_id < 0 || case_id >= c->ncases)
        return fail(MPH_ERR_ARG, "mph_get_neighbors_noalloc: bad argument");
    if (!is_device_pointer(ptr) || (ind && !is_device_pointer(ind)))
        return fail(MPH_ERR_ARG, "mph_get_neighbors_noalloc: ptr and ind must be device memory");
    unsigned char* sp = g_scratch_ptr[c->device];
    if (!sp) return fail(MPH_ERR_STATE, "mph_get_neighbors_noalloc: scratch not resolved for device %d", c->device);
    const CaseConst& k = c->cases[case_id].k;
    const int n = k.end - k.begin;
    const int tiles = (n + kCsrTile - 1) / kCsrTile;
    if (tiles > kScratchTiles)
        return fail(MPH_ERR_ARG, "mph_get_neighbors_noalloc: a case of %d particles exceeds the scratch limit of %d",
                    n, kScratchTiles * kCsrTile);
    hipStream_t s = c->stream;
    int* err = reinterpret_cast<int*>(sp);
    long long* sums = reinterpret_cast<long long*>(sp + kScratchHead);
    hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(256), 0, s, err, 1);
    hipLaunchKernelGGL(k_csr_counts, dim3((n + 255) / 256), dim3(256), 0, s, k.begin, n, c->pidx.p, c->nout.p, ptr);
    hipLaunchKernelGGL(k_csr_tile_sums, dim3(tiles), dim3(kCsrThreads), 0, s, ptr + 1, n, sums);
    hipLaunchKernelGGL(k_csr_scan_sums, dim3(1), dim3(1024), 0, s, sums, tiles);
    hipLaunchKernelGGL(k_csr_tile_scan, dim3(tiles), dim3(kCsrThreads), 0, s, ptr + 1, n, sums);
    int rc = launch_ok("neighbour rows: counts");
    if (rc) return rc;
    long long tot = 0;
    HIP_TRY(hipMemcpyAsync(&tot, ptr + n, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (total) *total = tot;
    if (!ind) return MPH_OK;
    if (tot > ind_capacity)
        return fail(MPH_ERR_ARG, "mph_get_neighbors_noalloc: ind holds %lld entries, the lists need %lld",
                    ind_capacity, tot);
    int pad = 1;
    while (pad < c->width + 32) pad <<= 1;      // the union's entries and up to 27 of the particle itself
    const int waves = std::max(1, std::min(4, (int)((64 * 1024) / ((size_t)pad * sizeof(int)))));
    int ncl = 0;
    HIP_TRY(hipMemcpyAsync(c->h_scalar, c->cl_count.p + case_id, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    ncl = *c->h_scalar;
    if (ncl > 0) {
        const int grid = std::max(1, std::min((ncl + waves - 1) / waves, 20480));
        hipLaunchKernelGGL(k_csr_rows, dim3(grid), dim3(64 * waves), (size_t)pad * sizeof(int) * waves, s,
                           c->d_cases.p, case_id, c->cl_info.p, c->cl_count.p, c->cl_nbr.p, c->cl_len.p, c->width,
                           c->snap.p, c->pidx.p, c->nout.p, ptr, ind, pad, err);
        if ((rc = launch_ok("neighbour rows"))) return rc;
    }
    int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, err, sizeof(int), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (bad) return fail(MPH_ERR_STATE, "mph_get_neighbors_noalloc: %d rows differ from their NeighborCount", bad);
    return MPH_OK;
}

extern "C" int mph_get_stats(const mph_ctx* cc, mph_stats* out)
{
    mph_ctx* c = const_cast<mph_ctx*>(cc);
    if (!c || !out) return fail(MPH_ERR_ARG, "mph_get_stats: bad argument");
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<long long> rb(c->ncases);
    std::vector<int> len(c->nslots), cnt(c->ncases);
    int mx = 0;
    HIP_TRY(hipMemcpy(rb.data(), c->rebuilds.p, sizeof(long long) * c->ncases, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(len.data(), c->cl_len.p, sizeof(int) * c->nslots, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cnt.data(), c->cl_count.p, sizeof(int) * c->ncases, hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(&mx, c->max_len.p, sizeof(int), hipMemcpyDeviceToHost));
    out->steps = c->steps;
    out->rebuilds = 0;
    for (long long r : rb) out->rebuilds = std::max(out->rebuilds, r);
    out->list_entries = 0;
    out->eval_slots = 0;
    // a sweep's wavefront walks a list in blocks of 64 entries, K steps of 64/K per full
    // block and ceil(rest / (64/K)) for the last, each step 64 lane evaluations
    const int S = 64 / c->K;
    for (int k = 0; k < c->ncases; ++k)
        for (int g = 0; g < cnt[k]; ++g) {
            const int l = len[c->cases[k].k.cl_base + g];
            out->list_entries += l;
            out->eval_slots += 64ll * ((long long)c->K * (l / 64) + (l % 64 + S - 1) / S);
        }
    out->list_width = c->width;
    out->max_list_len = mx;
    out->cluster_size = c->K;
    out->list_grid = list_grid(c);
    out->clusters = 0;
    for (int v : cnt) out->clusters += v;
    return MPH_OK;
}

extern "C" void mph_destroy(mph_ctx* c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c;
}

// ---------------------------------------------------------------------------------------
// Pairwise-message surrogate (no counterpart in the reference; see mph_surrogate.h)

namespace {

struct SurCkpt {
    DevBuf<double4> pos, vel, vnext, snap;
    DevBuf<int> pidx;
    DevBuf<double> wallc;
    std::vector<double> time;
    long long steps = 0;
};

}  // namespace

struct mph_surrogate {
    mph_ctx* c = nullptr;
    mph_sur_params prm{};
    long long t_adam = 0;
    int wh = 0;
    DevBuf<float> w, m1, m2, grad, wr;
    DevBuf<bf16x8> pk;
    DevBuf<int> pcnt, pnbr;
    DevBuf<float4> apred;                 // by original id (inference) or list entry (training)
    DevBuf<float4> msum;                  // message sums of the inference, sorted order
    // training set of the current sample
    DevBuf<int> plist, toff, towner;
    int nlist = 0, ntiles = 0;
    long nslots = 0, cap_slots = 0;
    DevBuf<__bf16> X;
    DevBuf<float> Z1, Z2, Z3, dz1, dz2, dz3, dm, Bs;
    DevBuf<float4> target, dmk;
    DevBuf<double> loss, sums;
    // snapshot of a training state and original-order scratch
    DevBuf<double4> spos, svel, v_orig, v_ref;
    DevBuf<double> sdens, n_ref;
    DevBuf<int> spidx;
    SurCkpt ck[2];
};

extern "C" void mph_sur_params_default(mph_sur_params* p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->u_scale = 0.1;
    p->a_scale = 10.0;
    p->dens_scale = 50.0;
    p->lr = 1e-3;
    p->beta1 = 0.9;
    p->beta2 = 0.999;
    p->eps = 1e-8;
    p->train_stride = 1;
    p->seed = 12345;
}

static SurScales sur_scales(const mph_surrogate* s)
{
    const CaseConst& k = s->c->cases[0].k;
    SurScales sc;
    sc.inv_u = (float)(1.0 / s->prm.u_scale);
    sc.a_scale = (float)s->prm.a_scale;
    sc.dens_scale = (float)s->prm.dens_scale;
    const double r = s->prm.radius_ratio > 0.0 ? s->prm.radius_ratio * k.l0 : std::sqrt(k.hmax2);
    sc.cut2 = r * r;
    sc.inv_h = (float)(1.0 / r);
    sc.inv_n0p = (float)(1.0 / k.n0p);
    sc.dt = (float)k.dt;
    sc.gx = (float)k.grav[0];
    sc.gy = (float)k.grav[1];
    sc.gz = (float)k.grav[2];
    return sc;
}

static int sur_pack(mph_surrogate* s)
{
    hipLaunchKernelGGL(k_sur_pack, dim3((kPkTiles * 64 + 255) / 256), dim3(256), 0, s->c->stream, s->w.p, s->pk.p,
                       s->wr.p);
    return launch_ok("surrogate pack");
}

// pairs within the interaction radius for every fluid particle, from the current lists
static int sur_pairs(mph_surrogate* s)
{
    mph_ctx* c = s->c;
    if (c->width > s->wh) {        // the lists were widened since: so is the pair table
        int rc = s->pnbr.alloc((size_t)c->n * c->width);
        if (rc) return rc;
        s->wh = c->width;
    }
    const long threads = (long)c->nslots * c->K;
    hipLaunchKernelGGL(k_sur_pairs, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, c->stream, c->d_cases.p,
                       c->slot_case.p, c->nslots, c->cl_info.p, c->cl_count.p, c->cl_nbr.p, c->cl_len.p, c->width,
                       c->K, c->pos.p, sur_scales(s).cut2, s->wh, s->pcnt.p, s->pnbr.p);
    return launch_ok("surrogate pairs");
}

static int sur_grid(const mph_ctx* c)
{
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) cus = prop.multiProcessorCount;
    return cus * 5;     // persistent: five 256-thread blocks per CU (the LDS and register limit)
}

// inference forward at (pos, velf, vstrain) over all fluid particles, pair search fused:
// s->apred by original id
static int sur_infer(mph_surrogate* s, const double4* velf)
{
    mph_ctx* c = s->c;
    const SurScales sc = sur_scales(s);
    HIP_TRY(hipMemsetAsync(s->msum.p, 0, sizeof(float4) * c->n, c->stream));
    hipLaunchKernelGGL(k_sur_infer, dim3(sur_grid(c)), dim3(256), 0, c->stream, c->d_cases.p, c->cl_info.p,
                       c->cl_count.p, c->cl_nbr.p, c->cl_len.p, c->width, c->pos.p, velf, c->vstrain.p, s->pk.p,
                       s->w.p, sc, s->msum.p);
    hipLaunchKernelGGL(k_sur_finish, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->d_cases.p, c->n,
                       c->pos.p, c->pidx.p, s->msum.p, sc, s->apred.p);
    return launch_ok("surrogate forward");
}

static int sur_kick(mph_surrogate* s)
{
    mph_ctx* c = s->c;
    hipLaunchKernelGGL(k_sur_kick, dim3((c->n + 255) / 256), dim3(256), 0, c->stream, c->d_cases.p, c->n, c->pos.p,
                       c->vel.p, c->pidx.p, s->apred.p, c->vnext.p, c->flags.p, c->any_flag.p, c->rebuilds.p);
    return launch_ok("surrogate kick");
}

// one surrogate step (2 dt): move, lists, densities, prediction, kick, move, lists, kick
// (a list overflow is resolved at once: the rebuilds are checked on the host)
static int sur_step(mph_surrogate* s)
{
    mph_ctx* c = s->c;
    int rc = 0;
    if ((rc = launch_move(c)) || (rc = launch_rebuild_checked(c)) || (rc = launch_sweep_a(c, 1)) ||
        (rc = sur_infer(s, c->vel.p)) || (rc = sur_kick(s)) || (rc = launch_move(c)) ||
        (rc = launch_rebuild_checked(c)) || (rc = sur_kick(s)))
        return rc;
    for (auto& ch : c->cases) ch.time += 2.0 * ch.prm.dt;
    c->steps += 2;
    return MPH_OK;
}

static int sur_ckpt_save(mph_surrogate* s, SurCkpt& k)
{
    mph_ctx* c = s->c;
    const size_t n = c->n;
    int rc = 0;
    if (!k.pos.p && ((rc = k.pos.alloc(n)) || (rc = k.vel.alloc(n)) || (rc = k.vnext.alloc(n)) || (rc = k.snap.alloc(n)) ||
                     (rc = k.pidx.alloc(n)) || (rc = k.wallc.alloc(c->wallc.n))))
        return rc;
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(k.pos.p, c->pos.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(k.vel.p, c->vel.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(k.vnext.p, c->vnext.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(k.snap.p, c->snap.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(k.pidx.p, c->pidx.p, n * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(k.wallc.p, c->wallc.p, c->wallc.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    k.time.clear();
    for (auto& ch : c->cases) k.time.push_back(ch.time);
    k.steps = c->steps;
    return MPH_OK;
}

// restore a checkpoint; the next step rebuilds the cells and lists from the restored state
static int sur_ckpt_load(mph_surrogate* s, const SurCkpt& k)
{
    mph_ctx* c = s->c;
    const size_t n = c->n;
    hipStream_t st = c->stream;
    HIP_TRY(hipMemcpyAsync(c->pos.p, k.pos.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->vel.p, k.vel.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->vnext.p, k.vnext.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->snap.p, k.snap.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->pidx.p, k.pidx.p, n * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->wallc.p, k.wallc.p, c->wallc.n * sizeof(double), hipMemcpyDeviceToDevice, st));
    std::vector<int> ones(c->ncases, 1);
    HIP_TRY(hipMemcpyAsync(c->flags.p, ones.data(), sizeof(int) * c->ncases, hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(c->any_flag.p, ones.data(), sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));      // the host vector above must outlive the copies
    for (int q = 0; q < c->ncases; ++q) c->cases[q].time = k.time[q];
    c->steps = k.steps;
    return MPH_OK;
}

extern "C" mph_surrogate* mph_sur_create(mph_ctx* c, const mph_sur_params* p)
{
    if (!c || !p) {
        fail(MPH_ERR_ARG, "mph_sur_create: bad argument");
        return nullptr;
    }
    if (c->ncases != 1 || c->mixed) {
        fail(MPH_ERR_ARG, "the surrogate needs a single-case FP64 context");
        return nullptr;
    }
    const CaseConst& k = c->cases[0].k;
    for (int t = 2; t < kTypes; ++t)
        for (int d = 0; d < 3; ++d)
            if (k.wvel[t][d] != 0.0 || k.womega[t][d] != 0.0) {
                fail(MPH_ERR_ARG, "the surrogate supports static walls only");
                return nullptr;
            }
    if (!(p->u_scale > 0.0 && p->a_scale > 0.0 && p->train_stride >= 1)) {
        fail(MPH_ERR_ARG, "mph_sur_create: scales must be positive and train_stride >= 1");
        return nullptr;
    }
    // the pairs come from the neighbour lists, which reach the largest kernel radius
    if (!(p->radius_ratio >= 0.0) || p->radius_ratio * c->cases[0].k.l0 > std::sqrt(c->cases[0].k.hmax2) * (1.0 + 1e-12)) {
        fail(MPH_ERR_ARG, "mph_sur_create: radius_ratio must be in [0, largest kernel radius / spacing]");
        return nullptr;
    }
    std::unique_ptr<mph_surrogate> s(new mph_surrogate());
    s->c = c;
    s->prm = *p;
    s->wh = c->width;
    int rc = 0;
    if ((rc = s->w.alloc(kSurNP)) || (rc = s->m1.alloc(kSurNP)) || (rc = s->m2.alloc(kSurNP)) ||
        (rc = s->grad.alloc(kSurNP)) || (rc = s->wr.alloc(kSurNP)) || (rc = s->pk.alloc(kPkTiles * 64)) ||
        (rc = s->pcnt.alloc(c->n)) || (rc = s->pnbr.alloc((size_t)c->n * s->wh)) || (rc = s->apred.alloc(c->n)) ||
        (rc = s->msum.alloc(c->n)) ||
        (rc = s->loss.alloc(1)) || (rc = s->sums.alloc(4)) || (rc = s->v_orig.alloc(c->n)) ||
        (rc = s->v_ref.alloc(c->n)) || (rc = s->n_ref.alloc(c->n)))
        return nullptr;
    // uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights from a splitmix64 stream, zero
    // biases, output layer scaled by 0.1
    std::vector<float> hw(kSurNP, 0.0f);
    unsigned long long st = p->seed;
    auto uni = [&]() {
        st += 0x9E3779B97F4A7C15ull;
        unsigned long long z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    for (int q = 0; q < 16 * kSurHid; ++q) hw[kOffW1 + q] = (float)(uni() / std::sqrt(16.0));
    for (int q = 0; q < kSurHid * kSurHid; ++q) hw[kOffW2 + q] = (float)(uni() / 8.0);
    for (int q = 0; q < kSurHid * kSurHid; ++q) hw[kOffW3 + q] = (float)(uni() / 8.0);
    for (int q = 0; q < kSurHid * kSurOut; ++q) hw[kOffW4 + q] = (float)(0.1 * uni() / 8.0);
    if (mph_sur_set_weights(s.get(), hw.data()) != MPH_OK) return nullptr;
    return s.release();
}

extern "C" int mph_sur_set_weights(mph_surrogate* s, const float* w)
{
    if (!s || !w) return fail(MPH_ERR_ARG, "mph_sur_set_weights: bad argument");
    hipStream_t st = s->c->stream;
    HIP_TRY(hipMemcpyAsync(s->w.p, w, kSurNP * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(s->m1.p, 0, kSurNP * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(s->m2.p, 0, kSurNP * sizeof(float), st));
    s->t_adam = 0;
    int rc = sur_pack(s);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(st));
    return MPH_OK;
}

extern "C" int mph_sur_get_weights(mph_surrogate* s, float* w)
{
    if (!s || !w) return fail(MPH_ERR_ARG, "mph_sur_get_weights: bad argument");
    HIP_TRY(hipMemcpyAsync(w, s->w.p, kSurNP * sizeof(float), hipMemcpyDeviceToHost, s->c->stream));
    HIP_TRY(hipStreamSynchronize(s->c->stream));
    return MPH_OK;
}

extern "C" int mph_sur_predict(mph_surrogate* s, double* accel)
{
    if (!s || !accel) return fail(MPH_ERR_ARG, "mph_sur_predict: bad argument");
    mph_ctx* c = s->c;
    int rc = 0;
    if ((rc = sur_infer(s, c->vnext.p))) return rc;
    std::vector<float4> h(c->n);
    HIP_TRY(hipMemcpyAsync(h.data(), s->apred.p, sizeof(float4) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->n; ++i) {
        accel[3 * i] = h[i].x;
        accel[3 * i + 1] = h[i].y;
        accel[3 * i + 2] = h[i].z;
    }
    return MPH_OK;
}

// training list of the current pair lists: fluid particles with sorted index % stride == 0,
// with their tile offsets (the total rounded up to an even number of tiles)
static int sur_make_list(mph_surrogate* s)
{
    mph_ctx* c = s->c;
    std::vector<int> cnt(c->n), typ(c->n);
    std::vector<double4> hp(c->n);
    HIP_TRY(hipMemcpyAsync(cnt.data(), s->pcnt.p, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hp.data(), c->pos.p, sizeof(double4) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<int> list, off, owner;
    int tiles = 0;
    for (int i = 0; i < c->n; ++i) {
        long long tg;
        std::memcpy(&tg, &hp[i].w, sizeof(tg));
        if ((tg & 3) >= 2 || i % s->prm.train_stride) continue;
        list.push_back(i);
        off.push_back(tiles);
        const int nt = (cnt[i] + 15) / 16;
        for (int q = 0; q < nt; ++q) owner.push_back((int)list.size() - 1);
        tiles += nt;
    }
    if (list.empty()) return fail(MPH_ERR_STATE, "no fluid particles to train on");
    if (tiles & 1) {
        owner.push_back((int)list.size() - 1);      // padding tile: all its slots are invalid
        ++tiles;
    }
    s->nlist = (int)list.size();
    s->ntiles = tiles;
    s->nslots = 16l * tiles;
    int rc = 0;
    if ((size_t)s->nlist > s->plist.n &&
        ((rc = s->plist.alloc(s->nlist)) || (rc = s->toff.alloc(s->nlist)) || (rc = s->target.alloc(s->nlist)) ||
         (rc = s->dmk.alloc(s->nlist))))
        return rc;
    if ((size_t)tiles > s->towner.n && (rc = s->towner.alloc(tiles))) return rc;
    if (s->nslots > s->cap_slots) {
        const long cap = s->nslots + s->nslots / 4;
        if ((rc = s->X.alloc((size_t)cap * 32)) || (rc = s->Z1.alloc((size_t)cap * 64)) ||
            (rc = s->Z2.alloc((size_t)cap * 64)) || (rc = s->Z3.alloc((size_t)cap * 64)) ||
            (rc = s->dz1.alloc((size_t)cap * 64)) || (rc = s->dz2.alloc((size_t)cap * 64)) ||
            (rc = s->dz3.alloc((size_t)cap * 64)) || (rc = s->dm.alloc((size_t)cap * 4)) ||
            (rc = s->Bs.alloc((size_t)cap * 8)))
            return rc;
        s->cap_slots = cap;
        // slots of the padding tile are never written by the forward: start from zeros
        HIP_TRY(hipMemsetAsync(s->X.p, 0, sizeof(__bf16) * cap * 32, c->stream));
        HIP_TRY(hipMemsetAsync(s->Z1.p, 0, sizeof(float) * cap * 64, c->stream));
        HIP_TRY(hipMemsetAsync(s->Z2.p, 0, sizeof(float) * cap * 64, c->stream));
        HIP_TRY(hipMemsetAsync(s->Z3.p, 0, sizeof(float) * cap * 64, c->stream));
        HIP_TRY(hipMemsetAsync(s->Bs.p, 0, sizeof(float) * cap * 8, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(s->plist.p, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(s->toff.p, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(s->towner.p, owner.data(), sizeof(int) * owner.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPH_OK;
}

// forward (storing activations), loss, backward and weight gradients over the training
// list at (pos, velf, strain) with s->target set; gradient in s->grad, loss in s->loss
static int sur_fwd_bwd(mph_surrogate* s, const double4* pos, const double4* velf, const double* strain)
{
    mph_ctx* c = s->c;
    hipStream_t st = c->stream;
    HIP_TRY(hipMemsetAsync(s->loss.p, 0, sizeof(double), st));
    HIP_TRY(hipMemsetAsync(s->grad.p, 0, sizeof(float) * kSurNP, st));
    hipLaunchKernelGGL(k_sur_forward_train, dim3(sur_grid(c)), dim3(256), 0, st, c->d_cases.p, s->plist.p,
                       s->nlist, s->toff.p, pos, velf, strain, s->pcnt.p, s->pnbr.p, s->wh, s->pk.p, s->w.p,
                       sur_scales(s), s->apred.p, s->X.p, s->Z1.p, s->Z2.p, s->Z3.p, s->Bs.p);
    hipLaunchKernelGGL(k_sur_loss, dim3((s->nlist + 255) / 256), dim3(256), 0, st, s->plist.p, s->nlist, s->apred.p,
                       s->target.p, (float)(1.0 / s->prm.a_scale), s->dmk.p, s->loss.p);
    const unsigned gb = (unsigned)((s->nslots + 255) / 256);
    hipLaunchKernelGGL((k_sur_backward<kSurOut, true>), dim3(gb), dim3(256), 0, st, s->nslots, s->towner.p, s->toff.p,
                       s->plist.p, s->pcnt.p, s->dmk.p, s->wr.p, kOffW4, nullptr, s->Z3.p, s->dz3.p, s->dm.p,
                       s->Bs.p);
    hipLaunchKernelGGL((k_sur_backward<kSurHid, false>), dim3(gb), dim3(256), 0, st, s->nslots, s->towner.p,
                       s->toff.p, s->plist.p, s->pcnt.p, s->dmk.p, s->wr.p, kOffW3, s->dz3.p, s->Z2.p, s->dz2.p,
                       nullptr, nullptr);
    hipLaunchKernelGGL((k_sur_backward<kSurHid, false>), dim3(gb), dim3(256), 0, st, s->nslots, s->towner.p,
                       s->toff.p, s->plist.p, s->pcnt.p, s->dmk.p, s->wr.p, kOffW2, s->dz2.p, s->Z1.p, s->dz1.p,
                       nullptr, nullptr);
    const long nslices = s->nslots / 32;
    const int waves = 1024;
    const int per = (int)((nslices + waves - 1) / waves);
    const unsigned gw = (unsigned)((nslices + per - 1) / per + 3) / 4;
    hipLaunchKernelGGL((k_sur_wgrad<0, 2, 4>), dim3(gw), dim3(256), 0, st, s->nslots, per, s->X.p, nullptr, s->dz1.p,
                       kSurHid, kSurHid, kSurIn, s->grad.p, kOffW1, -1);
    hipLaunchKernelGGL((k_sur_wgrad<1, 4, 4>), dim3(gw), dim3(256), 0, st, s->nslots, per, nullptr, s->Z1.p, s->dz2.p,
                       kSurHid, kSurHid, kSurHid, s->grad.p, kOffW2, kOffB2);
    hipLaunchKernelGGL((k_sur_wgrad<1, 4, 4>), dim3(gw), dim3(256), 0, st, s->nslots, per, nullptr, s->Z2.p, s->dz3.p,
                       kSurHid, kSurHid, kSurHid, s->grad.p, kOffW3, kOffB3);
    hipLaunchKernelGGL((k_sur_wgrad<1, 4, 1>), dim3(gw), dim3(256), 0, st, s->nslots, per, nullptr, s->Z3.p, s->dm.p, 4,
                       kSurOut, kSurHid, s->grad.p, kOffW4, kOffB4);
    return launch_ok("surrogate backward");
}

static int sur_adam(mph_surrogate* s)
{
    ++s->t_adam;
    const double b1 = s->prm.beta1, b2 = s->prm.beta2;
    const float c1 = (float)(1.0 / (1.0 - std::pow(b1, (double)s->t_adam)));
    const float c2 = (float)(1.0 / (1.0 - std::pow(b2, (double)s->t_adam)));
    hipLaunchKernelGGL(k_sur_adam, dim3((kSurNP + 255) / 256), dim3(256), 0, s->c->stream, s->w.p, s->m1.p, s->m2.p,
                       s->grad.p, (float)s->prm.lr, (float)b1, (float)b2, (float)s->prm.eps, c1, c2);
    int rc = launch_ok("adam");
    return rc ? rc : sur_pack(s);
}

extern "C" int mph_sur_gradient(mph_surrogate* s, const double* target, double* loss, float* grad)
{
    if (!s || !target || !loss) return fail(MPH_ERR_ARG, "mph_sur_gradient: bad argument");
    mph_ctx* c = s->c;
    int rc = 0;
    if ((rc = sur_pairs(s)) || (rc = sur_make_list(s))) return rc;
    std::vector<int> pidx(c->n), list(s->nlist);
    HIP_TRY(hipMemcpyAsync(pidx.data(), c->pidx.p, sizeof(int) * c->n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(list.data(), s->plist.p, sizeof(int) * s->nlist, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<float4> t(s->nlist);
    for (int k = 0; k < s->nlist; ++k) {
        const int o = pidx[list[k]];
        t[k] = make_float4((float)target[3 * o], (float)target[3 * o + 1], (float)target[3 * o + 2], 0.f);
    }
    HIP_TRY(hipMemcpyAsync(s->target.p, t.data(), sizeof(float4) * t.size(), hipMemcpyHostToDevice, c->stream));
    if ((rc = sur_fwd_bwd(s, c->pos.p, c->vnext.p, c->vstrain.p))) return rc;
    HIP_TRY(hipMemcpyAsync(loss, s->loss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad) HIP_TRY(hipMemcpyAsync(grad, s->grad.p, sizeof(float) * kSurNP, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MPH_OK;
}

extern "C" int mph_sur_train(mph_surrogate* s, int samples, int iters, double* losses)
{
    if (!s || samples < 0 || iters < 0) return fail(MPH_ERR_ARG, "mph_sur_train: bad argument");
    mph_ctx* c = s->c;
    hipStream_t st = c->stream;
    const size_t n = c->n;
    int rc = 0;
    if (!s->spos.p && ((rc = s->spos.alloc(n)) || (rc = s->svel.alloc(n)) || (rc = s->sdens.alloc(n)) ||
                       (rc = s->spidx.alloc(n))))
        return rc;
    for (int smp = 0; smp < samples; ++smp) {
        // plain step n+1: state (x_{n+1}, v_n) and volumetric strain at x_{n+1}
        if ((rc = launch_step_checked(c))) return rc;
        for (auto& ch : c->cases) ch.time += ch.prm.dt;
        ++c->steps;
        if ((rc = sur_pairs(s)) || (rc = sur_make_list(s))) return rc;
        HIP_TRY(hipMemcpyAsync(s->spos.p, c->pos.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->svel.p, c->vel.p, n * sizeof(double4), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->sdens.p, c->vstrain.p, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(s->spidx.p, c->pidx.p, n * sizeof(int), hipMemcpyDeviceToDevice, st));
        // plain step n+2: v_{n+2}, by original id
        if ((rc = launch_step_checked(c))) return rc;
        for (auto& ch : c->cases) ch.time += ch.prm.dt;
        ++c->steps;
        hipLaunchKernelGGL(k_sur_to_orig, dim3((c->n + 255) / 256), dim3(256), 0, st, c->n, c->pidx.p, c->vnext.p,
                           nullptr, s->v_orig.p, nullptr);
        hipLaunchKernelGGL(k_sur_target, dim3((s->nlist + 255) / 256), dim3(256), 0, st, c->d_cases.p, s->plist.p,
                           s->nlist, s->spidx.p, s->svel.p, s->v_orig.p, s->target.p);
        if ((rc = launch_ok("surrogate targets"))) return rc;
        for (int it = 0; it < iters; ++it) {
            // the snapshot's sorted order is the one its pair lists and training list use
            if ((rc = sur_fwd_bwd(s, s->spos.p, s->svel.p, s->sdens.p)) || (rc = sur_adam(s))) return rc;
            if (losses) {
                HIP_TRY(hipMemcpyAsync(&losses[(size_t)smp * iters + it], s->loss.p, sizeof(double),
                                       hipMemcpyDeviceToHost, st));
            }
        }
        HIP_TRY(hipStreamSynchronize(st));
    }
    return check_capacity(c);
}

extern "C" int mph_sur_step(mph_surrogate* s, int nsteps)
{
    if (!s || nsteps < 0) return fail(MPH_ERR_ARG, "mph_sur_step: bad argument");
    for (int k = 0; k < nsteps; ++k) {
        int rc = sur_step(s);
        if (rc) return rc;
    }
    return check_capacity(s->c);
}

static int plain_steps(mph_ctx* c, int k)
{
    for (int q = 0; q < k; ++q) {
        int rc = launch_step_host(c);
        if (rc) return rc;
        for (auto& ch : c->cases) ch.time += ch.prm.dt;
        ++c->steps;
    }
    return MPH_OK;
}

extern "C" int mph_sur_advance(mph_surrogate* s, int blocks, int k, double tol, mph_sur_stats* out)
{
    if (!s || blocks < 0 || k < 1 || !(tol > 0.0) || !out) return fail(MPH_ERR_ARG, "mph_sur_advance: bad argument");
    mph_ctx* c = s->c;
    hipStream_t st = c->stream;
    const double t0 = c->cases[0].time;
    int rc = 0;
    for (int b = 0; b < blocks; ++b) {
        if ((rc = sur_ckpt_save(s, s->ck[0]))) return rc;
        for (int q = 0; q < k; ++q)
            if ((rc = sur_step(s))) return rc;
        // judge: one surrogate step and two plain steps from the same state
        if ((rc = sur_ckpt_save(s, s->ck[1])) || (rc = sur_step(s)) || (rc = launch_sweep_a(c, 0, false, false, true)))
            return rc;
        hipLaunchKernelGGL(k_sur_to_orig, dim3((c->n + 255) / 256), dim3(256), 0, st, c->n, c->pidx.p, c->vnext.p,
                           c->dens_a.p, s->v_ref.p, s->n_ref.p);
        if ((rc = launch_ok("judge")) || (rc = sur_ckpt_load(s, s->ck[1])) || (rc = plain_steps(c, 2)) ||
            (rc = ensure_dens_a(c)))
            return rc;
        HIP_TRY(hipMemsetAsync(s->sums.p, 0, 4 * sizeof(double), st));
        hipLaunchKernelGGL(k_sur_compare, dim3((c->n + 255) / 256), dim3(256), 0, st, c->n, c->pos.p, c->pidx.p,
                           c->vnext.p, c->dens_a.p, s->v_ref.p, s->n_ref.p, s->sums.p);
        double h[4];
        HIP_TRY(hipMemcpyAsync(h, s->sums.p, sizeof(h), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        const double ev = std::sqrt(h[0] / std::max(h[1], 1e-300));
        const double en = std::sqrt(h[2] / std::max(h[3], 1e-300));
        out->blocks += 1;
        out->last_err_v = ev;
        out->last_err_n = en;
        out->max_err_v = std::max(out->max_err_v, ev);
        out->max_err_n = std::max(out->max_err_n, en);
        if (ev > tol || en > tol || !(ev == ev) || !(en == en)) {
            // roll back the block and replay it with plain steps
            if ((rc = sur_ckpt_load(s, s->ck[0])) || (rc = plain_steps(c, 2 * k))) return rc;
            out->rollbacks += 1;
            out->plain_steps += 2 * k;
        } else {
            out->sur_steps += k;
            out->plain_steps += 2;
        }
    }
    out->sim_time += c->cases[0].time - t0;
    return check_capacity(c);
}

extern "C" void mph_sur_destroy(mph_surrogate* s)
{
    // touches no context state: the context may already be destroyed (hipFree synchronises)
    delete s;
}
