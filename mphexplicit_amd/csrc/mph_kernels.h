// gfx950 kernels of the MPH-WC step.  Included by mph_device.hip only.
//
// Layout in HBM (all arrays in the current, cell-sorted particle order; cases are
// contiguous ranges):
//   pos   double4 {x, y, z, tag}    tag = type | case<<2 stored as the bits of the double
//   vel   double4 {vx, vy, vz, P}   velocity used by this step's sweeps + pressure P
//   vnext double4 {vx, vy, vz, 0}   velocity after the step (input of the next move)
//   aux   double4 {Gx, Gy, Gz, PA}  gravity centre and attractive pressure
//   force double4 {Fx, Fy, Fz, 0}
//   snap  double4                   positions at the last list build
//   nbr   int [width/4][stride][4]  ELL neighbour lists in groups of four: entry k of particle i
//                                   at ((k/4)*stride + i)*4 + k%4, one 16-byte load per group
//
// Every particle kernel is launched over a block map: block b handles particles
// [map[b].y, min(map[b].y + blockDim, case.end)) of case map[b].x, so the case constants
// are uniform per workgroup.  The map is laid out for the 8 XCDs of the MI355X: the
// hardware hands workgroup b to XCD b % 8, and the map gives each XCD one contiguous run
// of particle blocks (entries with case -1 are padding), so a spatially compact region
// shares one XCD's L2.
//
// Cells are stored in Morton order: the sort key of a particle is
//   key_base + 2*rank(cell) + (wall ? 1 : 0)
// so a cell's fluid particles are followed by its wall particles and consecutive particles
// are close in space in all three directions.  The reference's own cell ids (main.cpp:96-100)
// still define the grid, the stencil and the order in which lists are filled.
#pragma once

#include <hip/hip_runtime.h>
#include "mph_case.h"

namespace mph {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kXcds = 8;   // MI355X: 8 XCDs x 32 CUs, workgroups dispatched round-robin

__device__ __forceinline__ int tag_of(double w) { return (int)__double_as_longlong(w); }

// Mod(a, w) = a - w*floor(a/w) of main.cpp:70, returning the identical double.
// For 0 <= a < w the IEEE quotient a/w rounds to a value below 1, so floor gives 0 and
// the reference computes a - w*(+-0) (= a, with -0 mapped to +0); for a in [-w, 0) or
// [w, 2w) floor gives -1 or 1 exactly and w*floor is exact, so a single add/sub rounds
// like the reference.  Only the remaining range needs the division.
__device__ __forceinline__ double mod_w(double a, double w)
{
    if (a >= 0.0 && a < w) return a + 0.0;
    if (a >= -w && a < w + w) return a < 0.0 ? a + w : a - w;
    return a - w * floor(a / w);
}

// periodic x_j - x_i per axis, main.cpp:1680 (bit-identical to the reference)
__device__ __forceinline__ void rel3(const CaseConst& C, const double4& pi, const double4& pj,
                                     double& dx, double& dy, double& dz)
{
    dx = mod_w(pj.x - pi.x + C.hwid[0], C.dwid[0]) - C.hwid[0];
    dy = mod_w(pj.y - pi.y + C.hwid[1], C.dwid[1]) - C.hwid[1];
    dz = mod_w(pj.z - pi.z + C.hwid[2], C.dwid[2]) - C.hwid[2];
}

// The same difference for a list partner, without the division: with positions inside
// the periodic box, a = x_j - x_i + W/2 lies in [-W, 2W), where floor(a/W) is -1, 0 or 1.
__device__ __forceinline__ double rel_wrap(double xj, double xi, double hw, double w)
{
    const double a = xj - xi + hw;
    const double s = a < 0.0 ? -w : (a >= w ? w : 0.0);
    return (a - s) - hw;
}

// ... and for a particle far enough from the periodic faces that no partner wraps
__device__ __forceinline__ double rel_in(double xj, double xi, double hw) { return (xj - xi + hw) - hw; }

// 1/sqrt(q2): hardware estimate and two Newton steps (about 1 ulp)
__device__ __forceinline__ double rsqrt_nr(double q2)
{
    double y = __builtin_amdgcn_rsq(q2);
    const double h = 0.5 * q2;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

// true if no neighbour-list partner of a particle at p can need a periodic image on any
// axis: p is at least `guard` (list cut-off + Verlet skin) inside every face
__device__ __forceinline__ bool away_from_faces(const CaseConst& C, const double4& p)
{
    return p.x - C.dmin[0] >= C.guard[0] && C.dmax[0] - p.x >= C.guard[0] &&
           p.y - C.dmin[1] >= C.guard[1] && C.dmax[1] - p.y >= C.guard[1] &&
           p.z - C.dmin[2] >= C.guard[2] && C.dmax[2] - p.z >= C.guard[2];
}

// -------------------------------------------------------------------------------------
// move: convection (fluid), rigid wall motion, periodic wrap, Verlet trigger
// main.cpp:1605-1615, 2013-2053, 2306-2317, 1086-1108

__global__ void __launch_bounds__(256)
k_move(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
       double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ vnext,
       const double4* __restrict__ snap, const double* __restrict__ wallc,
       int* __restrict__ flags, int* __restrict__ any_flag)
{
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    bool exceed = false;
    if (i < C.end) {
        double4 p = pos[i];
        const double4 vn = vnext[i];
        double vx = vn.x, vy = vn.y, vz = vn.z;
        const int t = tag_of(p.w) & 3;
        if (t < 2) {
            p.x += vx * C.dt;
            p.y += vy * C.dt;
            p.z += vz * C.dt;
        } else {
            const double* c = wallc + (m.x * kTypes + t) * 3;
            const double (*R)[3] = C.wrot[t];
            const double* w = C.womega[t];
            const double* V = C.wvel[t];
            double r0 = p.x - c[0], r1 = p.y - c[1], r2 = p.z - c[2];
            // components are overwritten in this order, as main.cpp:2033-2035 does
            r0 = R[0][0] * r0 + R[0][1] * r1 + R[0][2] * r2;
            r1 = R[1][0] * r0 + R[1][1] * r1 + R[1][2] * r2;
            r2 = R[2][0] * r0 + R[2][1] * r1 + R[2][2] * r2;
            vx = w[1] * r2 - w[2] * r1 + V[0];
            vy = w[2] * r0 - w[0] * r2 + V[1];
            vz = w[0] * r1 - w[1] * r0 + V[2];
            p.x = r0 + c[0] + V[0] * C.dt;
            p.y = r1 + c[1] + V[1] * C.dt;
            p.z = r2 + c[2] + V[2] * C.dt;
        }
        p.x = mod_w(p.x - C.dmin[0], C.dwid[0]) + C.dmin[0];
        p.y = mod_w(p.y - C.dmin[1], C.dwid[1]) + C.dmin[1];
        p.z = mod_w(p.z - C.dmin[2], C.dwid[2]) + C.dmin[2];
        pos[i] = p;
        vel[i] = make_double4(vx, vy, vz, 0.0);
        const double4 s = snap[i];
        const double ex = mod_w(p.x - s.x + C.hwid[0], C.dwid[0]) - C.hwid[0];
        const double ey = mod_w(p.y - s.y + C.hwid[1], C.dwid[1]) - C.hwid[1];
        const double ez = mod_w(p.z - s.z + C.hwid[2], C.dwid[2]) - C.hwid[2];
        exceed = ex * ex + ey * ey + ez * ez > C.shift2;
    }
    const unsigned long long b = __ballot(exceed);
    if (b != 0ull && (threadIdx.x & 63) == 0) {
        atomicOr(&flags[m.x], 1);
        atomicOr(any_flag, 1);
    }
}

// -------------------------------------------------------------------------------------
// bucketing: cell key, LSD radix sort, cell starts, reorder (main.cpp:1111-1326)

// CellId of main.cpp:96-100 (the 3D form; 2D cases have cc[2] == 1) -> storage key
__device__ __forceinline__ int cell_key(const CaseConst& C, const int* __restrict__ rank,
                                        const double4& p, int type)
{
    int ix = ((int)floor((p.x - C.dmin[0]) / C.cw[0])) % C.cc[0];
    int iy = ((int)floor((p.y - C.dmin[1]) / C.cw[1])) % C.cc[1];
    int iz = ((int)floor((p.z - C.dmin[2]) / C.cw[2])) % C.cc[2];
    ix = (ix % C.cc[0] + C.cc[0]) % C.cc[0];
    iy = (iy % C.cc[1] + C.cc[1]) % C.cc[1];
    iz = (iz % C.cc[2] + C.cc[2]) % C.cc[2];
    const int id = (ix * C.cc[1] + iy) * C.cc[2] + iz;
    return C.key_base + 2 * rank[C.rank_base + id] + (type >= 2 ? 1 : 0);
}

// keys of rebuilding cases from positions; other cases keep their sorted keys
__global__ void __launch_bounds__(256)
k_keys(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
       const double4* __restrict__ pos, const int* __restrict__ flags,
       const int* __restrict__ any_flag, const int* __restrict__ rank, const unsigned* kprev,
       unsigned* kout, int* vout)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    unsigned key;
    if (flags[m.x]) {
        const double4 p = pos[i];
        key = (unsigned)cell_key(C, rank, p, tag_of(p.w) & 3);
    } else {
        key = kprev[i];
    }
    kout[i] = key;
    vout[i] = i;
}

__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const unsigned* __restrict__ keys, int n, int shift, unsigned* __restrict__ hist,
             const int* __restrict__ any_flag)
{
    if (*any_flag == 0) return;
    __shared__ unsigned cnt[kRadix];
    const int tid = threadIdx.x;
    cnt[tid] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        const int e = base + it * kSortThreads + tid;
        if (e < n) atomicAdd(&cnt[(keys[e] >> shift) & (kRadix - 1)], 1u);
    }
    __syncthreads();
    hist[tid * gridDim.x + blockIdx.x] = cnt[tid];
}

// exclusive scan of `total` counters in place, one workgroup of 1024 threads
__global__ void __launch_bounds__(1024)
k_scan_exclusive(unsigned* __restrict__ data, int total, const int* __restrict__ any_flag)
{
    if (any_flag && *any_flag == 0) return;
    __shared__ unsigned part[1024];
    const int tid = threadIdx.x;
    const int chunk = (total + 1023) / 1024;
    const int lo = min(total, tid * chunk), hi = min(total, lo + chunk);
    unsigned s = 0;
    for (int k = lo; k < hi; ++k) s += data[k];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {   // Hillis-Steele inclusive scan of the partials
        const unsigned v = tid >= off ? part[tid - off] : 0u;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    unsigned run = part[tid] - s;
    for (int k = lo; k < hi; ++k) {
        const unsigned v = data[k];
        data[k] = run;
        run += v;
    }
}

// stable scatter of one radix digit; ranks inside a tile come from 64-lane ballots
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin,
                unsigned* __restrict__ kout, int* __restrict__ vout, int n, int shift,
                const unsigned* __restrict__ hist, const int* __restrict__ any_flag)
{
    if (*any_flag == 0) return;
    __shared__ unsigned base[kRadix];
    __shared__ unsigned wcnt[kSortThreads / 64][kRadix];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    base[tid] = hist[tid * gridDim.x + blockIdx.x];
    const int tile = blockIdx.x * kSortTile;
    for (int it = 0; it < kSortItems; ++it) {
        for (int w = 0; w < kSortThreads / 64; ++w) wcnt[w][tid] = 0;
        __syncthreads();
        const int e = tile + it * kSortThreads + tid;
        const bool valid = e < n;
        unsigned key = 0;
        int val = 0;
        if (valid) { key = kin[e]; val = vin[e]; }
        const unsigned d = (key >> shift) & (kRadix - 1);
        unsigned long long peers = __ballot(valid);
        for (int b = 0; b < kRadixBits; ++b) {
            const bool bit = (d >> b) & 1u;
            const unsigned long long mb = __ballot(bit);
            peers &= bit ? mb : ~mb;
        }
        const int rank = __popcll(peers & lt);
        if (valid && rank == 0) wcnt[wave][d] = (unsigned)__popcll(peers);
        __syncthreads();
        if (valid) {
            unsigned off = base[d] + (unsigned)rank;
            for (int w = 0; w < wave; ++w) off += wcnt[w][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        unsigned add = 0;
        for (int w = 0; w < kSortThreads / 64; ++w) add += wcnt[w][tid];
        base[tid] += add;
        __syncthreads();
    }
}

// start[k] = first sorted slot with key >= k, for k in [0, nkeys]
__global__ void __launch_bounds__(256)
k_cell_start(const unsigned* __restrict__ keys, int n, int nkeys, int* __restrict__ start,
             const int* __restrict__ any_flag)
{
    if (*any_flag == 0) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (unsigned)k) lo = mid + 1; else hi = mid;
    }
    start[k] = lo;
}

// gather id, position (tag), velocity by the sort permutation (main.cpp:1262-1325) into
// the alternate buffers; the snapshot of the rebuilt cases is the new position
__global__ void __launch_bounds__(256)
k_gather(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
         const int* __restrict__ flags, const int* __restrict__ any_flag,
         const int* __restrict__ perm, const double4* __restrict__ pos,
         const double4* __restrict__ vel, const int* __restrict__ pidx,
         double4* __restrict__ pos_alt, double4* __restrict__ vel_alt, int* __restrict__ pidx_alt)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end || flags[m.x] == 0) return;
    const int s = perm[i];
    pos_alt[i] = pos[s];
    vel_alt[i] = vel[s];
    pidx_alt[i] = pidx[s];
}

__global__ void __launch_bounds__(256)
k_scatter_back(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
               const int* __restrict__ flags, const int* __restrict__ any_flag,
               const double4* __restrict__ pos_alt, const double4* __restrict__ vel_alt,
               const int* __restrict__ pidx_alt, double4* __restrict__ pos,
               double4* __restrict__ vel, int* __restrict__ pidx, double4* __restrict__ snap)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end || flags[m.x] == 0) return;
    const double4 p = pos_alt[i];
    pos[i] = p;
    snap[i] = p;
    vel[i] = vel_alt[i];
    pidx[i] = pidx_alt[i];
}

// -------------------------------------------------------------------------------------
// neighbour lists (main.cpp:1329-1596): for each particle the fluid candidates of the
// 3^d periodic cells, then (unless it is a wall particle with no fluid neighbour,
// main.cpp:1416) the wall candidates; kept if q^2 <= (MaxRadius+MARGIN)^2, with the
// reference's exact periodic difference.  Cells are visited in ascending reference cell
// id, as the reference fills its lists.  Self is not stored; count_out carries the
// reference's NeighborCount (list length including self).

__device__ __forceinline__ size_t ell_slot(int k, int stride, int i)
{
    return ((size_t)(k >> 2) * stride + i) * 4 + (k & 3);
}

__device__ __forceinline__ int axis_window(int ic, int c, int range, int* out)
{
    // the periodic window {ic-1, ic, ic+1} (or {ic} for range 0) in ascending order,
    // repeats included when the axis has fewer than 3 cells (as main.cpp:1365-1391)
    if (range == 0) { out[0] = ic; return 1; }
    int a = (ic - 1 + c) % c, b = ic, d = (ic + 1) % c;
    if (a > b) { int t = a; a = b; b = t; }
    if (b > d) { int t = b; b = d; d = t; }
    if (a > b) { int t = a; a = b; b = t; }
    out[0] = a; out[1] = b; out[2] = d;
    return 3;
}

__global__ void __launch_bounds__(256)
k_build_lists(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
              const int* __restrict__ flags, const int* __restrict__ any_flag,
              const double4* __restrict__ pos, const unsigned* __restrict__ keys,
              const int* __restrict__ start, const int* __restrict__ rank,
              const int* __restrict__ cell_of_rank, int* __restrict__ nbr, int stride, int width,
              int* __restrict__ ncnt, int* __restrict__ count_out, int* __restrict__ max_len)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    if (flags[m.x] == 0) return;          // uniform: a workgroup belongs to one case
    const int i = m.y + (int)threadIdx.x;
    int count = 0;                        // stored entries (self excluded)
    if (i < C.end) {
        const double4 pi = pos[i];
        const int kk = (int)keys[i] - C.key_base;
        const bool wall = kk & 1;
        const int cid = cell_of_rank[C.rank_base + (kk >> 1)];
        const int iz = cid % C.cc[2];
        const int iy = (cid / C.cc[2]) % C.cc[1];
        const int ix = cid / (C.cc[1] * C.cc[2]);
        int wx[3], wy[3], wz[3];
        const int nx = axis_window(ix, C.cc[0], 1, wx);
        const int ny = axis_window(iy, C.cc[1], 1, wy);
        const int nz = axis_window(iz, C.cc[2], C.dim == 2 ? 0 : 1, wz);
        int seen_self = 0;
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1 && wall && count == 0) break;
            for (int a = 0; a < nx; ++a)
                for (int b = 0; b < ny; ++b)
                    for (int c = 0; c < nz; ++c) {
                        const int cell = (wx[a] * C.cc[1] + wy[b]) * C.cc[2] + wz[c];
                        const int k0 = C.key_base + 2 * rank[C.rank_base + cell] + pass;
                        const int lo = start[k0], hi = start[k0 + 1];
                        for (int j = lo; j < hi; ++j) {
                            const double4 pj = pos[j];
                            double dx, dy, dz;
                            rel3(C, pi, pj, dx, dy, dz);
                            const double q2 = dx * dx + dy * dy + dz * dz;
                            if (q2 <= C.cut2) {
                                if (j == i) { seen_self = 1; continue; }
                                if (count < width) nbr[ell_slot(count, stride, i)] = j;
                                ++count;
                            }
                        }
                    }
        }
        ncnt[i] = count < width ? count : width;
        count_out[i] = count + seen_self;
    }
    int mx = count;
    for (int off = 32; off > 0; off >>= 1) mx = max(mx, __shfl_xor(mx, off));
    if ((threadIdx.x & 63) == 0) atomicMax(max_len, mx);
}

// -------------------------------------------------------------------------------------
// Partner loop shared by the sweeps.  List entries arrive four at a time (one 16-byte
// load per group); the data of the next partner is loaded while the current one is
// evaluated, alternating two register sets, so a wavefront always has loads in flight.
// Entries past a particle's count are never dereferenced.

template <class Load, class Eval>
__device__ __forceinline__ void for_partners(const int* __restrict__ nbr, int stride, int i, int cnt,
                                             Load load, Eval eval)
{
    const int4* __restrict__ e4 = reinterpret_cast<const int4*>(nbr);
    const int ng = (cnt + 3) >> 2;
    int4 cur = ng > 0 ? e4[i] : make_int4(i, i, i, i);
    auto A = load(cnt > 0 ? cur.x : i);
    for (int g = 0; g < ng; ++g) {
        const int k = 4 * g;
        const int4 nxt = g + 1 < ng ? e4[(size_t)(g + 1) * stride + i] : make_int4(i, i, i, i);
        auto B = load(k + 1 < cnt ? cur.y : i);
        eval(A);                                   // k < cnt holds inside the loop
        A = load(k + 2 < cnt ? cur.z : i);
        if (k + 1 < cnt) eval(B);
        B = load(k + 3 < cnt ? cur.w : i);
        if (k + 2 < cnt) eval(A);
        A = load(k + 4 < cnt ? nxt.x : i);
        if (k + 3 < cnt) eval(B);
        cur = nxt;
    }
}

// -------------------------------------------------------------------------------------
// sweep A: number densities, gravity centre, volumetric strain, divergence
// (main.cpp:1665-1727, 1831-1896) with the pointwise coefficients and pressures of
// main.cpp:1631-1662, 1898-1906, 1729-1740 fused into the epilogue.
//
// Per pair: exact periodic difference (the division-free forms above), q^2, one
// reciprocal square root, then the kernel polynomials with their constant prefactors
// pulled out of the sums.  The kernels are evaluated branch-free with the factor
// max(1 - r/h, 0), which is zero outside the support exactly as the reference's cut-offs.

struct PairA {
    double4 p;
    double vx, vy, vz;
};

template <int SURF, bool WRAP>
__device__ __forceinline__ void sweep_a_loop(const CaseConst& C, const double* tab, int ti,
                                             const double4& pi, double vix, double viy, double viz,
                                             const double4* __restrict__ pos, const double4* vel,
                                             const int* __restrict__ nbr, int stride, int i, int cnt,
                                             double& na, double& gx, double& gy, double& gz,
                                             double& wsum, double& dsum, int& npair)
{
    auto load = [&](int j) {
        PairA q;
        q.p = pos[j];
        q.vx = vel[j].x;          // P (vel.w) of partners is being written by this kernel
        q.vy = vel[j].y;
        q.vz = vel[j].z;
        return q;
    };
    auto eval = [&](const PairA& q) {
        const double dx = WRAP ? rel_wrap(q.p.x, pi.x, C.hwid[0], C.dwid[0]) : rel_in(q.p.x, pi.x, C.hwid[0]);
        const double dy = WRAP ? rel_wrap(q.p.y, pi.y, C.hwid[1], C.dwid[1]) : rel_in(q.p.y, pi.y, C.hwid[1]);
        const double dz = WRAP ? rel_wrap(q.p.z, pi.z, C.hwid[2], C.dwid[2]) : rel_in(q.p.z, pi.z, C.hwid[2]);
        const double q2 = fma(dz, dz, fma(dy, dy, dx * dx));
        npair += q2 <= C.hmax2;
        const double ir = rsqrt_nr(q2);
        const double r = q2 * ir;
        const double sa = r * C.inv_ha;
        const double oma = fmax(1.0 - sa, 0.0);
        const double ratio = C.ratio_one ? 1.0 : tab[16 + ti * 4 + (tag_of(q.p.w) & 3)];
        const double wa = ratio * (oma * oma);
        na = fma(wa, sa, na);
        if (SURF) {
            gx = fma(dx, wa, gx);
            gy = fma(dy, wa, gy);
            gz = fma(dz, wa, gz);
        }
        const double omp = fmax(fma(-r, C.inv_hp, 1.0), 0.0);
        wsum = fma(omp, omp, wsum);
        const double ud = fma(q.vz - viz, dz, fma(q.vy - viy, dy, (q.vx - vix) * dx));
        dsum = fma(ud, omp * ir, dsum);
    };
    for_partners(nbr, stride, i, cnt, load, eval);
}

template <int SURF, int MINW>
__global__ void __launch_bounds__(256, MINW)
k_sweep_a(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
          const double4* __restrict__ pos, double4* vel, double4* __restrict__ aux,
          double* __restrict__ dens_a, double* __restrict__ vstrain, double* __restrict__ divp,
          const int* __restrict__ nbr, const int* __restrict__ ncnt, int stride,
          double* __restrict__ wallc, int ncases, unsigned long long* __restrict__ pair_count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // wall centres advance once per step (main.cpp:2048-2052); the move kernel of
        // this step has already used the old value
        for (int c = 0; c < ncases; ++c)
            for (int t = 2; t < kTypes; ++t)
                for (int d = 0; d < 3; ++d)
                    wallc[(c * kTypes + t) * 3 + d] += cases[c].wvel[t][d] * cases[c].dt;
    }
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    __shared__ double tab[32];                // [0,16): mu_ij, [16,32): InteractionRatio
    if (threadIdx.x < 16) {
        tab[threadIdx.x] = C.muij[threadIdx.x >> 2][threadIdx.x & 3];
        tab[16 + threadIdx.x] = C.ir[threadIdx.x >> 2][threadIdx.x & 3];
    }
    __syncthreads();
    const int i = m.y + (int)threadIdx.x;
    int npair = 0;
    if (i < C.end) {
        const double4 pi = pos[i];
        const int ti = tag_of(pi.w) & 3;
        const double vix = vel[i].x, viy = vel[i].y, viz = vel[i].z;
        double na = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, wsum = 0.0, dsum = 0.0;
        const int cnt = ncnt[i];
        if (away_from_faces(C, pi))
            sweep_a_loop<SURF, false>(C, tab, ti, pi, vix, viy, viz, pos, vel, nbr, stride, i, cnt,
                                      na, gx, gy, gz, wsum, dsum, npair);
        else
            sweep_a_loop<SURF, true>(C, tab, ti, pi, vix, viy, viz, pos, vel, nbr, stride, i, cnt,
                                     na, gx, gy, gz, wsum, dsum, npair);
        na *= C.ca;
        const double ep = C.cp * wsum - C.n0p;
        const double dv = -C.dcp * dsum;
        const double kap = ep < 0.0 ? 0.0 : C.kbulk[ti];
        const double p = -C.lbulk[ti] * dv + kap * ep;
        double pa = C.cofa[ti] * (na - C.n0a) / C.l0;
        if (C.n0a <= na) pa = 0.0;
        const double gs = C.cg * C.g_scale;
        vel[i].w = p;
        aux[i] = make_double4(gx * gs, gy * gs, gz * gs, pa);
        dens_a[i] = na;
        vstrain[i] = ep;
        divp[i] = dv;
    }
    for (int off = 32; off > 0; off >>= 1) npair += __shfl_xor(npair, off);
    if ((threadIdx.x & 63) == 0 && npair) atomicAdd(pair_count, (unsigned long long)npair);
}

// -------------------------------------------------------------------------------------
// sweep B: pressure, attractive-pressure, diffuse-interface and viscous forces
// (main.cpp:1908-1938, 1742-1775, 1778-1829, 1941-1985), gravity and the velocity update
// (main.cpp:1988-2010).  Each force term is summed separately with its constant factors
// pulled out, and the terms are combined in the reference's order.

struct PairB {
    double4 p, v, a;
};

template <int SURF, bool WRAP>
__device__ __forceinline__ void sweep_b_loop(const CaseConst& C, const double* tab, int ti,
                                             const double4& pi, const double4& vi, const double4& ai,
                                             const double4* __restrict__ pos, const double4* __restrict__ vel,
                                             const double4* __restrict__ aux, const int* __restrict__ nbr,
                                             int stride, int i, int cnt, double3& fp, double3& fv,
                                             double3& fa, double3& fs)
{
    const double mu_i = tab[ti * 4 + ti];
    auto load = [&](int j) {
        PairB q;
        q.p = pos[j];
        q.v = vel[j];
        if (SURF) q.a = aux[j];
        return q;
    };
    auto eval = [&](const PairB& q) {
        const double dx = WRAP ? rel_wrap(q.p.x, pi.x, C.hwid[0], C.dwid[0]) : rel_in(q.p.x, pi.x, C.hwid[0]);
        const double dy = WRAP ? rel_wrap(q.p.y, pi.y, C.hwid[1], C.dwid[1]) : rel_in(q.p.y, pi.y, C.hwid[1]);
        const double dz = WRAP ? rel_wrap(q.p.z, pi.z, C.hwid[2], C.dwid[2]) : rel_in(q.p.z, pi.z, C.hwid[2]);
        const double q2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double ir = rsqrt_nr(q2);
        const double r = q2 * ir;
        const int tj = tag_of(q.p.w) & 3;
        // pressure
        const double cp = (vi.w + q.v.w) * fmax(fma(-r, C.inv_hp, 1.0), 0.0) * ir;
        fp.x = fma(cp, dx, fp.x);
        fp.y = fma(cp, dy, fp.y);
        fp.z = fma(cp, dz, fp.z);
        // viscosity with the harmonic-mean coefficient of the type pair
        const double mu = C.mu_uniform ? mu_i : tab[ti * 4 + tj];
        const double ud = fma(q.v.z - vi.z, dz, fma(q.v.y - vi.y, dy, (q.v.x - vi.x) * dx));
        const double cv = mu * ud * fmax(fma(-r, C.inv_hv, 1.0), 0.0) * (ir * ir * ir);
        fv.x = fma(cv, dx, fv.x);
        fv.y = fma(cv, dy, fv.y);
        fv.z = fma(cv, dz, fv.z);
        if (SURF) {
            const double rij = C.ratio_one ? 1.0 : tab[16 + ti * 4 + tj];
            const double rji = C.ratio_one ? 1.0 : tab[16 + tj * 4 + ti];
            const double s = r * C.inv_ha;
            const double om = fmax(1.0 - s, 0.0);
            const double ca = fma(q.a.w, rji, ai.w * rij) * om * fma(-3.0, s, 1.0) * ir;
            fa.x = fma(ca, dx, fa.x);
            fa.y = fma(ca, dy, fa.y);
            fa.z = fma(ca, dz, fa.z);
            // diffuse interface with Q = R_ji G_j - R_ij G_i; a_j uses the type of i
            // (main.cpp:1792), so a_i factors out of the sum
            const double qx = fma(rji, q.a.x, -rij * ai.x);
            const double qy = fma(rji, q.a.y, -rij * ai.y);
            const double qz = fma(rji, q.a.z, -rij * ai.z);
            const double qd = fma(qz, dz, fma(qy, dy, qx * dx)) * (C.dcg * ir);
            const double w = C.cg * om;
            fs.x = fma(-om, fma(qd, dx, w * qx), fs.x);
            fs.y = fma(-om, fma(qd, dy, w * qy), fs.y);
            fs.z = fma(-om, fma(qd, dz, w * qz), fs.z);
        }
    };
    for_partners(nbr, stride, i, cnt, load, eval);
}

template <int SURF, int MINW>
__global__ void __launch_bounds__(256, MINW)
k_sweep_b(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
          const double4* __restrict__ pos, const double4* __restrict__ vel,
          const double4* __restrict__ aux, double4* __restrict__ vnext,
          double4* __restrict__ force, const int* __restrict__ nbr,
          const int* __restrict__ ncnt, int stride, int* __restrict__ flags,
          int* __restrict__ any_flag, long long* __restrict__ rebuilds, int ncases)
{
    if (blockIdx.x == 0 && threadIdx.x == 0 && ncases > 0) {
        for (int c = 0; c < ncases; ++c) {
            if (flags[c]) rebuilds[c] += 1;
            flags[c] = 0;
        }
        *any_flag = 0;
    }
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    __shared__ double tab[32];
    if (threadIdx.x < 16) {
        tab[threadIdx.x] = C.muij[threadIdx.x >> 2][threadIdx.x & 3];
        tab[16 + threadIdx.x] = C.ir[threadIdx.x >> 2][threadIdx.x & 3];
    }
    __syncthreads();
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    const double4 pi = pos[i];
    const double4 vi = vel[i];
    const int ti = tag_of(pi.w) & 3;
    const double4 ai = SURF ? aux[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    double3 fp = make_double3(0.0, 0.0, 0.0), fv = fp, fa = fp, fs = fp;
    const int cnt = ncnt[i];
    if (away_from_faces(C, pi))
        sweep_b_loop<SURF, false>(C, tab, ti, pi, vi, ai, pos, vel, aux, nbr, stride, i, cnt, fp, fv, fa, fs);
    else
        sweep_b_loop<SURF, true>(C, tab, ti, pi, vi, ai, pos, vel, aux, nbr, stride, i, cnt, fp, fv, fa, fs);
    const double cp = C.dcp * C.vol;                      // pressure
    const double cv = C.visc_c * (-C.dcv) * C.vol;        // viscosity
    double fx = fp.x * cp, fy = fp.y * cp, fz = fp.z * cp;
    if (SURF) {
        const double ca = C.dca * C.vol;
        const double cs = C.adiff[ti] * C.s_scale;
        fx += fa.x * ca; fy += fa.y * ca; fz += fa.z * ca;
        fx += fs.x * cs; fy += fs.y * cs; fz += fs.z * cs;
    }
    fx += fv.x * cv; fy += fv.y * cv; fz += fv.z * cv;
    double4 vn = make_double4(vi.x, vi.y, vi.z, 0.0);
    if (ti < 2) {
        const double ms = C.mass[ti];
        fx += ms * C.grav[0];
        fy += ms * C.grav[1];
        fz += ms * C.grav[2];
        vn.x += fx / ms * C.dt;
        vn.y += fy / ms * C.dt;
        vn.z += fz / ms * C.dt;
    }
    force[i] = make_double4(fx, fy, fz, 0.0);
    vnext[i] = vn;
}

// -------------------------------------------------------------------------------------
// virial pressure at output steps (main.cpp:2057-2302).  Only the trace of the per-particle
// virial stress is reported (VirialPressureAtParticle = -trace/d), so each pair adds
// f_ij . x_ij / V0 of the four terms directly:
//   pressure      P_i dw_P r
//   attractive    PA_i R_ij dw_A r
//   viscous       0.5 c_d mu_ij (u.e) (-dw_V)
//   diffuse       a_i (G_i . x) R_ij w_G /l0 * (h_G/R2g)  and  a_i (G_i . x) R_ij dw_G r /l0 * (h_G/R2g)
// with the reference's per-term cut-offs.  Velocities are those after the step's update.

__global__ void __launch_bounds__(256)
k_virial(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
         const double4* __restrict__ pos, const double4* __restrict__ vel,
         const double4* __restrict__ vnext, const double4* __restrict__ aux,
         const int* __restrict__ nbr, const int* __restrict__ ncnt, int stride,
         double* __restrict__ vir)
{
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    const double4 pi = pos[i];
    const double4 vi = vnext[i];
    const double P = vel[i].w;
    const double4 ai = aux[i];
    const int ti = tag_of(pi.w) & 3;
    const double a = C.adiff[ti];
    const double gsc = C.s_scale / C.vol;      // h_G / R2g / l0
    double tp = 0.0, ta = 0.0, tv = 0.0, ts = 0.0;
    const int cnt = ncnt[i];
    for (int k = 0; k < cnt; ++k) {
        const int j = nbr[ell_slot(k, stride, i)];
        const double4 pj = pos[j];
        double dx, dy, dz;
        rel3(C, pi, pj, dx, dy, dz);
        const double q2 = dx * dx + dy * dy + dz * dz;
        const int tj = tag_of(pj.w) & 3;
        const double r = sqrt(q2);
        if (C.hp2 - q2 > 0.0) tp += P * (C.dcp * (1.0 - r * C.inv_hp)) * r;
        if (C.ha2 - q2 > 0.0) {
            const double s = r * C.inv_ha;
            const double om = 1.0 - s;
            const double rij = C.ir[ti][tj];
            ta += ai.w * rij * (C.dca * om * (1.0 - 3.0 * s)) * r;
            const double gx = ai.x * dx + ai.y * dy + ai.z * dz;
            ts += a * gx * rij * (C.cg * om * om) * gsc;
            ts -= a * (-gx) * rij * (C.dcg * om) * r * gsc;
        }
        if (C.hv2 - q2 > 0.0) {
            const double4 vj = vnext[j];
            const double ue = ((vj.x - vi.x) * dx + (vj.y - vi.y) * dy + (vj.z - vi.z) * dz) / r;
            tv += 0.5 * C.visc_c * C.muij[ti][tj] * ue * (-(C.dcv * (1.0 - r * C.inv_hv)));
        }
    }
    const double tr = ((tp + ta) + tv) + ts;
    vir[i] = -tr / (C.dim == 2 ? 2.0 : 3.0);
}

}  // namespace mph
