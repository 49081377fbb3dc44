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
//   nbr   int [width][stride]       ELL neighbour lists, entry k of particle i at k*stride+i
//
// Every particle kernel is launched over a block map: block b handles particles
// [map[b].y, min(map[b].y + blockDim, case.end)) of case map[b].x, so the case constants
// are uniform per workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include "mph_case.h"

namespace mph {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;

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

__device__ __forceinline__ double sel4(int t, double a0, double a1, double a2, double a3)
{
    return t == 0 ? a0 : (t == 1 ? a1 : (t == 2 ? a2 : a3));
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

// CellId of main.cpp:96-100 (the 3D form; 2D cases have cc[2] == 1)
__device__ __forceinline__ int cell_key(const CaseConst& C, const double4& p, int type)
{
    int ix = ((int)floor((p.x - C.dmin[0]) / C.cw[0])) % C.cc[0];
    int iy = ((int)floor((p.y - C.dmin[1]) / C.cw[1])) % C.cc[1];
    int iz = ((int)floor((p.z - C.dmin[2]) / C.cw[2])) % C.cc[2];
    ix = (ix % C.cc[0] + C.cc[0]) % C.cc[0];
    iy = (iy % C.cc[1] + C.cc[1]) % C.cc[1];
    iz = (iz % C.cc[2] + C.cc[2]) % C.cc[2];
    const int id = (ix * C.cc[1] + iy) * C.cc[2] + iz;
    return C.key_base + id + (type >= 2 ? C.tcc : 0);
}

// keys of rebuilding cases from positions; other cases keep their sorted keys
__global__ void __launch_bounds__(256)
k_keys(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
       const double4* __restrict__ pos, const int* __restrict__ flags,
       const int* __restrict__ any_flag, const unsigned* kprev, unsigned* kout, int* vout)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    unsigned key;
    if (flags[m.x]) {
        const double4 p = pos[i];
        key = (unsigned)cell_key(C, p, tag_of(p.w) & 3);
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
// main.cpp:1416) the wall candidates; kept if q^2 <= (MaxRadius+MARGIN)^2.  Cells are
// visited in ascending id, so each list is ascending in sorted index like the
// reference's.  Self is not stored; count_out carries the reference's NeighborCount.

__device__ __forceinline__ int axis_window(int ic, int c, int range, int* out)
{
    // the periodic window {ic-1, ic, ic+1} (or {ic} for range 0) in ascending order
    if (range == 0) { out[0] = ic; return 1; }
    int a = (ic - 1 + c) % c, b = ic, d = (ic + 1) % c;
    // sort three values
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
              const int* __restrict__ start, int* __restrict__ nbr, int stride, int width,
              int* __restrict__ ncnt, int* __restrict__ count_out, int* __restrict__ max_len)
{
    if (*any_flag == 0) return;
    const int2 m = bmap[blockIdx.x];
    const CaseConst& C = cases[m.x];
    if (flags[m.x] == 0) return;          // uniform: a workgroup belongs to one case
    const int i = m.y + (int)threadIdx.x;
    int count = 0;                        // stored entries (self excluded)
    if (i < C.end) {
        const double4 pi = pos[i];
        int cid = (int)keys[i] - C.key_base;
        const bool wall = cid >= C.tcc;
        if (wall) cid -= C.tcc;
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
            const int off = C.key_base + (pass ? C.tcc : 0);
            for (int a = 0; a < nx; ++a)
                for (int b = 0; b < ny; ++b)
                    for (int c = 0; c < nz; ++c) {
                        const int cell = (wx[a] * C.cc[1] + wy[b]) * C.cc[2] + wz[c];
                        const int lo = start[off + cell], hi = start[off + cell + 1];
                        for (int j = lo; j < hi; ++j) {
                            const double4 pj = pos[j];
                            double dx, dy, dz;
                            rel3(C, pi, pj, dx, dy, dz);
                            const double q2 = dx * dx + dy * dy + dz * dz;
                            if (q2 <= C.cut2) {
                                if (j == i) { seen_self = 1; continue; }
                                if (count < width) nbr[(size_t)count * stride + i] = j;
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
// sweep A: number densities, gravity centre, volumetric strain, divergence
// (main.cpp:1665-1727, 1831-1896) with the pointwise coefficients and pressures of
// main.cpp:1631-1662, 1898-1906, 1729-1740 fused into the epilogue.

template <int SURF>
__global__ void __launch_bounds__(256)
k_sweep_a(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
          const double4* __restrict__ pos, double4* vel, double4* __restrict__ aux,
          double* __restrict__ dens_a, double* __restrict__ vstrain, double* __restrict__ divp,
          const int* __restrict__ nbr, const int* __restrict__ ncnt, int stride,
          double* __restrict__ wallc, int ncases, unsigned long long* __restrict__ pair_count)
{
    const int2 m = bmap[blockIdx.x];
    const CaseConst& C = cases[m.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // wall centres advance once per step (main.cpp:2048-2052); the move kernel of
        // this step has already used the old value
        for (int c = 0; c < ncases; ++c)
            for (int t = 2; t < kTypes; ++t)
                for (int d = 0; d < 3; ++d)
                    wallc[(c * kTypes + t) * 3 + d] += cases[c].wvel[t][d] * cases[c].dt;
    }
    const int i = m.y + (int)threadIdx.x;
    int npair = 0;
    if (i < C.end) {
        const double4 pi = pos[i];
        const int ti = tag_of(pi.w) & 3;
        const double vix = vel[i].x, viy = vel[i].y, viz = vel[i].z;
        const double r0 = C.ir[ti][0], r1 = C.ir[ti][1], r2 = C.ir[ti][2], r3 = C.ir[ti][3];
        double na = 0.0, gx = 0.0, gy = 0.0, gz = 0.0, wsum = 0.0, dsum = 0.0;
        const int cnt = ncnt[i];
        for (int k = 0; k < cnt; ++k) {
            const int j = nbr[(size_t)k * stride + i];
            const double4 pj = pos[j];
            double dx, dy, dz;
            rel3(C, pi, pj, dx, dy, dz);
            const double q2 = dx * dx + dy * dy + dz * dz;
            npair += q2 <= C.hmax2;
            const bool in_a = C.ha2 - q2 >= 0.0;
            const bool in_p = C.hp2 - q2 >= 0.0;
            if (!(in_a || in_p)) continue;
            const double r = sqrt(q2);
            if (in_a) {
                const double ratio = C.ratio_one ? 1.0 : sel4(tag_of(pj.w) & 3, r0, r1, r2, r3);
                const double s = r * C.inv_ha;
                const double om = 1.0 - s;
                na += ratio * (C.ca * s * om * om);
                if (SURF) {
                    const double wg = ratio * (C.cg * om * om);
                    gx += dx * wg;
                    gy += dy * wg;
                    gz += dz * wg;
                }
            }
            if (in_p) {
                const double s = r * C.inv_hp;
                const double om = 1.0 - s;
                wsum += C.cp * om * om;
                const double ux = vel[j].x - vix, uy = vel[j].y - viy, uz = vel[j].z - viz;
                dsum -= (ux * dx + uy * dy + uz * dz) * (C.dcp * om / r);
            }
        }
        const double ep = wsum - C.n0p;
        const double kap = ep < 0.0 ? 0.0 : C.kbulk[ti];
        const double p = -C.lbulk[ti] * dsum + kap * ep;
        double pa = C.cofa[ti] * (na - C.n0a) / C.l0;
        if (C.n0a <= na) pa = 0.0;
        vel[i].w = p;
        aux[i] = make_double4(gx * C.g_scale, gy * C.g_scale, gz * C.g_scale, pa);
        dens_a[i] = na;
        vstrain[i] = ep;
        divp[i] = dsum;
    }
    for (int off = 32; off > 0; off >>= 1) npair += __shfl_xor(npair, off);
    if ((threadIdx.x & 63) == 0 && npair) atomicAdd(pair_count, (unsigned long long)npair);
}

// -------------------------------------------------------------------------------------
// sweep B: pressure, attractive-pressure, diffuse-interface and viscous forces
// (main.cpp:1908-1938, 1742-1775, 1778-1829, 1941-1985), gravity and the velocity update
// (main.cpp:1988-2010).  Force is summed per term and combined in the reference's order.

template <int SURF>
__global__ void __launch_bounds__(256)
k_sweep_b(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
          const double4* __restrict__ pos, const double4* __restrict__ vel,
          const double4* __restrict__ aux, double4* __restrict__ vnext,
          double4* __restrict__ force, const int* __restrict__ nbr,
          const int* __restrict__ ncnt, int stride, int* __restrict__ flags,
          int* __restrict__ any_flag, long long* __restrict__ rebuilds, int ncases)
{
    const int2 m = bmap[blockIdx.x];
    const CaseConst& C = cases[m.x];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int c = 0; c < ncases; ++c) {
            if (flags[c]) rebuilds[c] += 1;
            flags[c] = 0;
        }
        *any_flag = 0;
    }
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    const double4 pi = pos[i];
    const double4 vi = vel[i];
    const int ti = tag_of(pi.w) & 3;
    const double mu0 = C.muij[ti][0], mu1 = C.muij[ti][1], mu2 = C.muij[ti][2], mu3 = C.muij[ti][3];
    double fpx = 0.0, fpy = 0.0, fpz = 0.0;
    double fvx = 0.0, fvy = 0.0, fvz = 0.0;
    double fax = 0.0, fay = 0.0, faz = 0.0;
    double fsx = 0.0, fsy = 0.0, fsz = 0.0;
    double4 ai = make_double4(0.0, 0.0, 0.0, 0.0);
    double ri0 = 1.0, ri1 = 1.0, ri2 = 1.0, ri3 = 1.0;
    if (SURF) {
        ai = aux[i];
        ri0 = C.ir[ti][0]; ri1 = C.ir[ti][1]; ri2 = C.ir[ti][2]; ri3 = C.ir[ti][3];
    }
    const int cnt = ncnt[i];
    for (int k = 0; k < cnt; ++k) {
        const int j = nbr[(size_t)k * stride + i];
        const double4 pj = pos[j];
        double dx, dy, dz;
        rel3(C, pi, pj, dx, dy, dz);
        const double q2 = dx * dx + dy * dy + dz * dz;
        const bool in_p = C.hp2 - q2 > 0.0;
        const bool in_v = C.hv2 - q2 > 0.0;
        const bool in_a = SURF && (C.ha2 - q2 > 0.0);
        if (!(in_p || in_v || in_a)) continue;
        const double r = sqrt(q2);
        const double ir = 1.0 / r;
        const double4 vj = vel[j];
        const int tj = tag_of(pj.w) & 3;
        if (in_p) {
            const double om = 1.0 - r * C.inv_hp;
            const double c = (vi.w + vj.w) * (C.dcp * om) * ir;
            fpx += c * dx;
            fpy += c * dy;
            fpz += c * dz;
        }
        if (in_v) {
            const double om = 1.0 - r * C.inv_hv;
            const double muij = sel4(tj, mu0, mu1, mu2, mu3);
            const double ue = ((vj.x - vi.x) * dx + (vj.y - vi.y) * dy + (vj.z - vi.z) * dz) * ir;
            const double c = C.visc_c * muij * ue * (-(C.dcv * om)) * ir * ir;
            fvx += c * dx;
            fvy += c * dy;
            fvz += c * dz;
        }
        if (SURF && in_a) {
            const double4 aj = aux[j];
            const double rij = C.ratio_one ? 1.0 : sel4(tj, ri0, ri1, ri2, ri3);
            const double rji = C.ratio_one ? 1.0 : C.ir[tj][ti];
            const double s = r * C.inv_ha;
            const double om = 1.0 - s;
            const double dwa = C.dca * om * (1.0 - 3.0 * s);
            const double ca = (ai.w * rij + aj.w * rji) * dwa * ir;
            fax += ca * dx;
            fay += ca * dy;
            faz += ca * dz;
            // diffuse interface; a_j uses the type of i (main.cpp:1792)
            const double wg = C.cg * om * om;
            const double dwg = C.dcg * om;
            const double wij = rij * wg, wji = rji * wg;
            const double dwij = rij * dwg, dwji = rji * dwg;
            const double gr = ((aj.x * dwji - ai.x * dwij) * dx + (aj.y * dwji - ai.y * dwij) * dy +
                               (aj.z * dwji - ai.z * dwij) * dz) * ir;
            // gr = (G_j dw_ji - G_i dw_ij) . x_ij / r, times x_ij gives gr_ref * e_ij
            // (main.cpp:1813-1821)
            fsx -= (aj.x * wji - ai.x * wij) + gr * dx;
            fsy -= (aj.y * wji - ai.y * wij) + gr * dy;
            fsz -= (aj.z * wji - ai.z * wij) + gr * dz;
        }
    }
    const double V0 = C.vol;
    double fx = fpx * V0, fy = fpy * V0, fz = fpz * V0;
    if (SURF) {
        const double sa = C.adiff[ti] * C.s_scale;
        fx += fax * V0; fy += fay * V0; fz += faz * V0;
        fx += fsx * sa; fy += fsy * sa; fz += fsz * sa;
    }
    fx += fvx * V0; fy += fvy * V0; fz += fvz * V0;
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
        const int j = nbr[(size_t)k * stride + i];
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
