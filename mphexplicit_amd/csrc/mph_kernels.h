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
//   cl_nbr int [clusters][width]    union neighbour list of each cluster (see "Cluster tiles")
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
#include <type_traits>

#include "mph_case.h"

namespace mph {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortThreads * kSortItems;
constexpr int kRadixBits = 8;
constexpr int kRadix = 1 << kRadixBits;
constexpr int kXcds = 8;        // MI355X: 8 XCDs x 32 CUs, workgroups dispatched round-robin
constexpr int kSubBits = 6;     // sub-cell bits of a sort key
// pair counters: one atomic per wavefront would serialise on a single address, so
// wavefronts add into kPairSlots counters 128 bytes apart, summed on the host
constexpr int kPairSlots = 1024;
constexpr int kPairStride = 16;

__device__ __forceinline__ int tag_of(double w) { return (int)__double_as_longlong(w); }

// native 4-wide vectors: records built from them stay in registers (the HIP double4
// wrapper type, copied as a whole, can keep a record in private memory)
typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

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

// rel3 on one axis with the common case branch-free: for a = x_j - x_i + W/2 in [-W, 2W)
// (positions inside the box, which the move keeps) the shift is -W, 0 or W and one
// subtraction rounds like the reference; a - 0.0 differs from mod_w's a + 0.0 only in the
// sign of a zero, which the following - W/2 removes.  Other a take the division, in lanes
// that need it only.
__device__ __forceinline__ double rel_exact(double xj, double xi, double hw, double w)
{
    const double a = xj - xi + hw;
    double m = a - (a < 0.0 ? -w : (a >= w ? w : 0.0));
    if (!(a >= -w && a < w + w)) m = a - w * floor(a / w);
    return m - hw;
}

// The same difference for a list partner, without the division: with positions inside
// the periodic box, a = x_j - x_i + W/2 lies in [-W, 2W), where floor(a/W) is -1, 0 or 1.
__device__ __forceinline__ double rel_wrap(double xj, double xi, double hw, double w)
{
    const double a = xj - xi + hw;
    const double s = a < 0.0 ? -w : (a >= w ? w : 0.0);
    return (a - s) - hw;
}

// The image x - W, x or x + W of a coordinate nearest to a reference point xf (a cluster's
// first member).  On an axis where a cell plus the reach of a list pair stays below W/2,
// the images of a row and of its interacting partners near the same xf differ by their
// minimum-image difference, and a row's own entry gives exactly zero: the periodic path of
// the sweeps takes the images once per union entry when it is fetched (shared by the K rows)
// and once per row, and then uses the plain difference.
__device__ __forceinline__ double image_near(double x, double xf, double hw, double w)
{
    const double a = x - xf + hw;
    return a < 0.0 ? x + w : (a >= w ? x - w : x);
}

// a position's images near xf on the given axes
__device__ __forceinline__ void image_axes(const CaseConst& C, int axes, const double4& xf, double& x, double& y,
                                           double& z)
{
    if (axes & 1) x = image_near(x, xf.x, C.hwid[0], C.dwid[0]);
    if (axes & 2) y = image_near(y, xf.y, C.hwid[1], C.dwid[1]);
    if (axes & 4) z = image_near(z, xf.z, C.hwid[2], C.dwid[2]);
}

// For a particle far enough from the periodic faces that no list partner wraps, the pair
// sweeps use the plain difference x_j - x_i: the exact minimum image, which differs from
// the reference's Mod form only by the rounding of the +-W/2 round trip.  Discrete
// decisions (list cut-off, cell keys) keep the reference's form (rel3).

// 1/sqrt(q2) for the pair sweeps, from the hardware estimate y (measured on MI355X: at most
// 5.2e-8 relative error).  With e = 1 - q2 y^2,
//   1/sqrt(q2) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + O(e^3)).
// rsqrt_pair (sweep B's forces): one Newton step, y (1 + e/2), at most 1.5 e^2 ~ 4.1e-15
// relative (37 ulp).  rsqrt_pair3 (sweep A's densities): the third-order correction, below
// 0.3 |e|^3 ~ 4e-23 before rounding, within 1 ulp (one instruction more): VolStrainP =
// sum w_P - N0p is a difference of sums near N0p, and the pressure it sets is sensitive to
// every rounding of r.  On the 1M-particle dam break the GPU's VolStrainP is within 1.1e-15
// of the exact sum (the reference's own FP64 sum: 7.6e-15).  In sweep B the third-order step
// measured 6% slower and the forces do not need it.
__device__ __forceinline__ double rsqrt_pair(double q2)
{
    const double y = __builtin_amdgcn_rsq(q2);
    const double e = fma(-(0.5 * q2) * y, y, 0.5);
    return fma(y, e, y);
}

__device__ __forceinline__ double rsqrt_pair3(double q2)
{
    const double y = __builtin_amdgcn_rsq(q2);
    const double e = fma(-(q2 * y), y, 1.0);
    const double t = e * fma(e, 0.375, 0.5);
    return fma(y, t, y);
}

// The pair sweeps evaluate every union entry without a branch.  Padding entries carry a
// record far outside every support (kFar), and a row's own entry has q^2 == 0 exactly (the
// same double on both sides).  far_if_self maps that q^2 to one outside every support, so
// both contribute exact zeros to every sum (0 * finite); sweeps A and B instead keep q^2 off
// zero with one max and handle the own entry's terms themselves.  Two distinct particles at
// the same position would make the reference divide by zero; the sweeps stay finite.
constexpr double kFar = 1.0e30;
__device__ __forceinline__ double far_if_self(double q2) { return q2 > 0.0 ? q2 : 1.0e300; }

// a per-type case constant for a lane's type t in 0..3: four scalar loads and selects instead
// of a vector load indexed by the lane's type (one round trip fewer at a wavefront's end)
__device__ __forceinline__ double by_type(const double (&a)[kTypes], int t)
{
    static_assert(kTypes == 4, "four particle types");
    double a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
    // opaque to the compiler, which would otherwise fold the selects into one vector load
    // from a selected address
    asm("" : "+s"(a0), "+s"(a1), "+s"(a2), "+s"(a3));
    return t == 0 ? a0 : (t == 1 ? a1 : (t == 2 ? a2 : a3));
}

// The block-map entry of a sweep B workgroup, which walks its XCD's run of the map backwards:
// the grid is the whole map (8 runs of `per` entries, XCD-interleaved, then 8 padding entries),
// so sweep B starts on the clusters whose lists and records sweep A touched last (still in
// that XCD's L2) and ends where the next step's sweep A begins.  On the 1.27M deck: FP64
// 582-587 -> 589-591 steps/s (both sweeps about 1% faster), mixed precision 689-690 -> 712.
__device__ __forceinline__ int reversed_entry()
{
    const int per = (int)gridDim.x / kXcds - 1;
    const int j = (int)(blockIdx.x / kXcds);
    return j < per ? (int)(blockIdx.x % kXcds) + kXcds * (per - 1 - j) : (int)blockIdx.x;
}

// The step's device flags, any_flag[0..1]:
//   [0] some case rebuilds its lists this step (the Verlet trigger, raised by the move)
//   [1] a list build of this step outgrew the reserved list width.  Every later kernel then
//       returns at once, so the state stays that of the overflowing step after its rebuild
//       (positions sorted, lists incomplete, rebuild flags still set); the host widens the
//       lists, builds them again and finishes the step (the reference sizes NeighborInd at
//       every rebuild, main.cpp:1487-1495, and never runs out of room).
__device__ __forceinline__ bool halted(const int* any_flag) { return any_flag[1] != 0; }
__device__ __forceinline__ bool no_rebuild(const int* any_flag) { return any_flag[0] == 0 || any_flag[1] != 0; }


// -------------------------------------------------------------------------------------
// move: convection (fluid), rigid wall motion, periodic wrap, Verlet trigger
// main.cpp:1605-1615, 2013-2053, 2306-2317, 1086-1108

__global__ void __launch_bounds__(256)
k_move(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
       double4* __restrict__ pos, double4* __restrict__ vel, const double4* __restrict__ vnext,
       const double4* __restrict__ snap, const double* __restrict__ wallc,
       int* __restrict__ flags, int* __restrict__ any_flag)
{
    if (halted(any_flag)) return;
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

// CellId of main.cpp:96-100 (the 3D form; 2D cases have cc[2] == 1) -> storage key.
// The key is (cell key << kSubBits) | sub-cell, where the cell key is
// key_base + 2*rank(cell) + (wall ? 1 : 0) and the sub-cell code orders the particles of a
// cell in 4x4x4 Morton bins, so that consecutive particles of a cell are close in space.
__device__ __forceinline__ int sub_bin(double t)
{
    const double f = (t - floor(t)) * 4.0;
    const int b = (int)f;
    return b < 0 ? 0 : (b > 3 ? 3 : b);
}

__device__ __forceinline__ unsigned cell_key(const CaseConst& C, const int* __restrict__ rank,
                                             const double4& p, int type)
{
    // clamped so that any position, non-finite ones included, maps to a cell of the grid
    // (fmax/fmin return the non-NaN operand); positions inside the domain are unchanged
    const double tx = fmin(fmax((p.x - C.dmin[0]) / C.cw[0], -1.0e9), 1.0e9);
    const double ty = fmin(fmax((p.y - C.dmin[1]) / C.cw[1], -1.0e9), 1.0e9);
    const double tz = fmin(fmax((p.z - C.dmin[2]) / C.cw[2], -1.0e9), 1.0e9);
    int ix = ((int)floor(tx)) % C.cc[0];
    int iy = ((int)floor(ty)) % C.cc[1];
    int iz = ((int)floor(tz)) % C.cc[2];
    ix = (ix % C.cc[0] + C.cc[0]) % C.cc[0];
    iy = (iy % C.cc[1] + C.cc[1]) % C.cc[1];
    iz = (iz % C.cc[2] + C.cc[2]) % C.cc[2];
    const int id = (ix * C.cc[1] + iy) * C.cc[2] + iz;
    const int sx = sub_bin(tx), sy = sub_bin(ty), sz = sub_bin(tz);
    const unsigned sub = (sx & 1) | (sy & 1) << 1 | (sz & 1) << 2 | (sx >> 1) << 3 | (sy >> 1) << 4 | (sz >> 1) << 5;
    const unsigned ck = (unsigned)(C.key_base + 2 * rank[C.rank_base + id] + (type >= 2 ? 1 : 0));
    return ck << kSubBits | sub;
}

// keys of rebuilding cases from positions; other cases keep their sorted keys
__global__ void __launch_bounds__(256)
k_keys(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
       const double4* __restrict__ pos, const int* __restrict__ flags,
       const int* __restrict__ any_flag, const int* __restrict__ rank, const unsigned* kprev,
       unsigned* kout, int* vout)
{
    if (no_rebuild(any_flag)) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end) return;
    unsigned key;
    if (flags[m.x]) {
        const double4 p = pos[i];
        key = cell_key(C, rank, p, tag_of(p.w) & 3);
    } else {
        key = kprev[i];
    }
    kout[i] = key;
    vout[i] = i;
}

// Batches sort only the cases flagged for a rebuild.  Their particles are compacted into one
// array in case order (coff[c]: case c's first slot there; nsort[0]: the number of keys,
// nsort[1]: the length of their radix histogram), sorted there and expanded back into the
// sorted keys and the permutation of all particles.  A case's keys lie in a key range of its
// own above those of the cases before it, so a stable sort of all keys orders each flagged
// case's particles exactly as this does and leaves the other cases in place.
__global__ void __launch_bounds__(256)
k_sort_offsets(const CaseConst* __restrict__ cases, int ncases, const int* __restrict__ flags,
               const int* __restrict__ any_flag, int* __restrict__ coff, int* __restrict__ nsort)
{
    if (no_rebuild(any_flag)) return;
    __shared__ int part[256];
    const int t = threadIdx.x;
    const int chunk = (ncases + 255) / 256;
    const int lo = min(ncases, t * chunk), hi = min(ncases, lo + chunk);
    int s = 0;
    for (int c = lo; c < hi; ++c)
        if (flags[c]) s += cases[c].end - cases[c].begin;
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int c = lo; c < hi; ++c) {
        coff[c] = run;
        if (flags[c]) run += cases[c].end - cases[c].begin;
    }
    if (t == 255) {
        const int nf = part[255];
        nsort[0] = nf;
        nsort[1] = kRadix * ((nf + kSortTile - 1) / kSortTile);
    }
}

// keys of the flagged cases' particles, compacted
__global__ void __launch_bounds__(256)
k_keys_compact(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
               const double4* __restrict__ pos, const int* __restrict__ flags,
               const int* __restrict__ any_flag, const int* __restrict__ rank, const int* __restrict__ coff,
               unsigned* __restrict__ kout, int* __restrict__ vout)
{
    if (no_rebuild(any_flag)) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end || flags[m.x] == 0) return;
    const double4 p = pos[i];
    const int e = coff[m.x] + (i - C.begin);
    kout[e] = cell_key(C, rank, p, tag_of(p.w) & 3);
    vout[e] = i;
}

// the sorted compacted keys and sources back at the flagged cases' particle slots
__global__ void __launch_bounds__(256)
k_expand_sorted(const CaseConst* __restrict__ cases, const int2* __restrict__ bmap,
                const int* __restrict__ flags, const int* __restrict__ any_flag, const int* __restrict__ coff,
                const unsigned* __restrict__ ks, const int* __restrict__ vs, unsigned* __restrict__ kfull,
                int* __restrict__ vfull)
{
    if (no_rebuild(any_flag)) return;
    const int2 m = bmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    const int i = m.y + (int)threadIdx.x;
    if (i >= C.end || flags[m.x] == 0) return;
    const int e = coff[m.x] + (i - C.begin);
    kfull[i] = ks[e];
    vfull[i] = vs[e];
}

// dn (batches, may be null): the number of keys to sort is dn[0], counted on the device; the
// blocks past it return, and the histogram has one column per block that sorts
__global__ void __launch_bounds__(kSortThreads)
k_radix_hist(const unsigned* __restrict__ keys, int n, int shift, unsigned* __restrict__ hist,
             const int* __restrict__ any_flag, const int* __restrict__ dn)
{
    if (no_rebuild(any_flag)) return;
    if (dn) n = dn[0];
    const int nb = dn ? (n + kSortTile - 1) / kSortTile : (int)gridDim.x;
    if ((int)blockIdx.x >= nb) return;
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
    hist[tid * nb + blockIdx.x] = cnt[tid];
}

// exclusive scan of `total` counters in place, one workgroup of 1024 threads
__global__ void __launch_bounds__(1024)
k_scan_exclusive(unsigned* __restrict__ data, int total, const int* __restrict__ any_flag)
{
    if (any_flag && no_rebuild(any_flag)) return;
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

// Multi-block exclusive scan for long arrays (the cell counts of large batches): tile sums,
// a single-workgroup scan of the tile sums (k_scan_exclusive), then a scan of each tile
// from LDS with its tile's offset.  Tiles of kScanTile counters, coalesced loads.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void __launch_bounds__(kScanThreads)
k_scan_tile_sums(const unsigned* __restrict__ data, int total, unsigned* __restrict__ sums,
                 const int* __restrict__ any_flag, const int* __restrict__ dtotal = nullptr)
{
    if (any_flag && no_rebuild(any_flag)) return;
    if (dtotal) total = dtotal[0];      // counted on the device (tiles past it sum to zero)
    const int base = blockIdx.x * kScanTile;
    unsigned s = 0;
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads) s += base + k < total ? data[base + k] : 0u;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ unsigned ws[kScanThreads / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int w = 0; w < kScanThreads / 64; ++w) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_tiles(unsigned* __restrict__ data, int total, const unsigned* __restrict__ offsets,
             const int* __restrict__ any_flag, const int* __restrict__ dtotal = nullptr)
{
    if (any_flag && no_rebuild(any_flag)) return;
    if (dtotal) total = dtotal[0];
    // padded so that the 16 consecutive counters of one thread fall in different banks
    __shared__ unsigned tile[kScanTile + kScanTile / kScanItems];
    __shared__ unsigned part[kScanThreads];
    const int base = blockIdx.x * kScanTile;
    auto at = [](int k) { return k + k / kScanItems; };
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        tile[at(k)] = base + k < total ? data[base + k] : 0u;
    __syncthreads();
    const int t = threadIdx.x;
    unsigned s = 0;
    for (int q = 0; q < kScanItems; ++q) s += tile[at(t * kScanItems + q)];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {   // inclusive scan of the thread sums
        const unsigned v = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned run = offsets[blockIdx.x] + part[t] - s;
    for (int q = 0; q < kScanItems; ++q) {
        const int k = at(t * kScanItems + q);
        const unsigned v = tile[k];
        tile[k] = run;
        run += v;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kScanTile; k += kScanThreads)
        if (base + k < total) data[base + k] = tile[at(k)];
}

// stable scatter of one radix digit; ranks inside a tile come from 64-lane ballots
__global__ void __launch_bounds__(kSortThreads)
k_radix_scatter(const unsigned* __restrict__ kin, const int* __restrict__ vin,
                unsigned* __restrict__ kout, int* __restrict__ vout, int n, int shift,
                const unsigned* __restrict__ hist, const int* __restrict__ any_flag, const int* __restrict__ dn)
{
    if (no_rebuild(any_flag)) return;
    if (dn) n = dn[0];
    const int nb = dn ? (n + kSortTile - 1) / kSortTile : (int)gridDim.x;
    if ((int)blockIdx.x >= nb) return;
    __shared__ unsigned base[kRadix];
    __shared__ unsigned wcnt[kSortThreads / 64][kRadix];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    base[tid] = hist[tid * nb + blockIdx.x];
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

// start[k] = first sorted slot whose cell key is >= k, for cell keys k in [0, nkeys]
__global__ void __launch_bounds__(256)
k_cell_start(const unsigned* __restrict__ keys, int n, int nkeys, int* __restrict__ start,
             const int* __restrict__ any_flag)
{
    if (no_rebuild(any_flag)) return;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nkeys) return;
    const unsigned kk = (unsigned)k << kSubBits;
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < kk) lo = mid + 1; else hi = mid;
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
    if (no_rebuild(any_flag)) return;
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
    if (no_rebuild(any_flag)) return;
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
// Cluster tiles.  After each rebuild the particles of every cell (fluid then wall, in
// sub-cell Morton order) are cut into clusters of K consecutive particles.  A cluster
// shares the cell's 3^d stencil, and its list is the union, in ascending sorted index, of
// the exact neighbour lists of its members (the reference's lists, main.cpp:1329-1596:
// q^2 <= (MaxRadius+MARGIN)^2 with the periodic difference of main.cpp:1398, and no list
// for a wall particle without fluid neighbours, main.cpp:1416).
//
// A wavefront evaluates one cluster: lane = row * S + slot with K rows (the members) and
// S = 64/K slots.  At each step the S slots take S consecutive union entries, so one
// 16-byte load per lane fetches S distinct partners and each is used by K rows; the
// per-member sums are reduced across the S slot lanes with shuffles at the end.  Pairs
// outside a member's cut-offs contribute exactly zero (the kernel factors vanish there),
// so each member's sums cover exactly the pairs of its reference list that interact.
//
// Cluster table entry: int4 {first particle, size, storage slot of its cell, case}.

// the periodic window {ic-1, ic, ic+1} (or {ic} for range 0) in ascending order,
// repeats included when the axis has fewer than 3 cells (as main.cpp:1365-1391)
__device__ __forceinline__ int axis_window(int ic, int c, int range, int* out)
{
    if (range == 0) { out[0] = ic; return 1; }
    int a = (ic - 1 + c) % c, b = ic, d = (ic + 1) % c;
    if (a > b) { int t = a; a = b; b = t; }
    if (b > d) { int t = b; b = d; d = t; }
    if (a > b) { int t = a; a = b; b = t; }
    out[0] = a; out[1] = b; out[2] = d;
    return 3;
}

// ranks of the stencil cells of reference cell `cid`, ascending (storage order)
__device__ __forceinline__ int stencil_ranks(const CaseConst& C, const int* __restrict__ rank, int cid,
                                             int* rk)
{
    const int iz = cid % C.cc[2];
    const int iy = (cid / C.cc[2]) % C.cc[1];
    const int ix = cid / (C.cc[1] * C.cc[2]);
    int wx[3], wy[3], wz[3];
    const int nx = axis_window(ix, C.cc[0], 1, wx);
    const int ny = axis_window(iy, C.cc[1], 1, wy);
    const int nz = axis_window(iz, C.cc[2], C.dim == 2 ? 0 : 1, wz);
    int n = 0;
    for (int a = 0; a < nx; ++a)
        for (int b = 0; b < ny; ++b)
            for (int c = 0; c < nz; ++c) {
                int v = rank[C.rank_base + (wx[a] * C.cc[1] + wy[b]) * C.cc[2] + wz[c]];
                int q = n++;
                while (q > 0 && rk[q - 1] > v) { rk[q] = rk[q - 1]; --q; }
                rk[q] = v;
            }
    return n;
}

// clusters per cell: ceil(particles / K), for every cell slot (storage order)
__global__ void __launch_bounds__(256)
k_cluster_count(const CaseConst* __restrict__ cases, const int* __restrict__ cell_case, int ncells,
                const int* __restrict__ start, int K, unsigned* __restrict__ ccount,
                const int* __restrict__ any_flag)
{
    if (no_rebuild(any_flag)) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ncells) return;
    const CaseConst& C = cases[cell_case[s]];
    const int k0 = C.key_base + 2 * (s - C.rank_base);
    const int n = start[k0 + 2] - start[k0];
    ccount[s] = (unsigned)((n + K - 1) / K);
}

// cluster table from the scanned counts; non-rebuilt cases are rewritten with the same values
__global__ void __launch_bounds__(256)
k_cluster_fill(const CaseConst* __restrict__ cases, const int* __restrict__ cell_case, int ncells,
               const int* __restrict__ start, int K,
               const unsigned* __restrict__ coff, int4* __restrict__ cl_info, int* __restrict__ cl_count,
               const int* __restrict__ any_flag)
{
    if (no_rebuild(any_flag)) return;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ncells) return;
    const int c = cell_case[s];
    const CaseConst& C = cases[c];
    const int r = s - C.rank_base;
    const int k0 = C.key_base + 2 * r;
    const int first = start[k0], n = start[k0 + 2] - first;
    const int base = (int)(coff[s] - coff[C.rank_base]);
    const int nq = (n + K - 1) / K;
    for (int q = 0; q < nq; ++q) {
        const int g = base + q;
        if (g < C.cl_cap) cl_info[C.cl_base + g] = make_int4(first + q * K, min(K, n - q * K), s, c);
    }
    if (r == C.tcc - 1) cl_count[c] = base + nq;
}

// Union lists, one wavefront per cluster, or per two clusters of one cell when K <= 4
// (main.cpp:1329-1596 per member, united in ascending index).  The candidates are the
// particles of the cell's 3^d stencil cells in ascending rank and index (the cell's record
// from k_stencil).  The wavefront takes
// 64 at a time (one load per lane, the next block's loads in flight), drops those farther than
// the cut-off from the members' bounding box (conservatively: cut-off^2 (1 + 1e-8) against a
// box built from the same differences as the pair test), and appends the rest to a ring of
// 128 entries in LDS.  Every 64 ring entries are tested one entry per lane against each member
// in turn (a member's hits are one ballot), and the block's union is appended to the list in
// one pass.  A candidate joins a cluster's union if one of its members has it within the
// cut-off and is not it; the union keeps candidate order.  Per member: NeighborCount as the
// reference counts it (main.cpp:1408-1420).
// Wavefronts loop over the cluster slots: a step without a rebuild costs a small grid, and a
// rebuild of a few clusters runs them in parallel.  Full rebuilds of the 1.27M-particle deck:
// one thread per member walking the candidates with a dependent load each, 5.5-6 ms; rows
// (members) x slots testing each 64-candidate block as soon as it was compacted (partly
// filled steps, the union appended per step), 1.84 ms; this kernel, 1.07 ms at six wavefronts
// per SIMD with the stencil records.
//
// Pair decisions.  A wavefront whose members are all at least `guard` inside both faces of
// an axis has no list pair that wraps on that axis.  On such axes, and on the axes near a face
// where the images near the first member are the minimum images (img_axes), pairs are decided
// by the plain difference of the (image) coordinates; where that q^2 is within 1e-8 of the
// cut-off, where it can round apart from the reference's periodic form, the lane redoes the
// reference's arithmetic on the original positions.  list_img (host) bounds the rounding of
// both forms below that band.  Other wavefronts use the reference's arithmetic per pair.

// Stencil records, once per rebuilt cell: the particle ranges of the cell's stencil cells in
// ascending rank (the table stencil_rank, fixed with the grid), as {first index, end of the
// fluid part, count, exclusive prefix of the counts}, at stencil[slot * 27 + q] for the cell's
// storage slot.  The list build's wavefronts of a cell (about six per cell) read them instead
// of each ranking the stencil and loading its ranges.  Two cells per wavefront (lanes 0-31 and
// 32-63); cells of cases without a rebuild keep their records (their ranges do not move).
__global__ void __launch_bounds__(256)
k_stencil(const CaseConst* __restrict__ cases, const int* __restrict__ cell_case, int ncells,
          const int* __restrict__ stencil_rank, const int* __restrict__ start, const int* __restrict__ flags,
          const int* __restrict__ any_flag, int4* __restrict__ stencil)
{
    if (no_rebuild(any_flag)) return;
    const int q = threadIdx.x & 31;
    const int slot = blockIdx.x * 8 + (threadIdx.x >> 5);
    if (slot >= ncells) return;
    const int c = cell_case[slot];
    if (flags[c] == 0) return;
    const CaseConst& C = cases[c];
    const int nc = C.dim == 2 ? 9 : 27;
    const int r = stencil_rank[(size_t)slot * 27 + (q < nc ? q : 0)];
    const int k0 = C.key_base + 2 * r;
    const int lo = start[k0], mid = start[k0 + 1], cnt = q < nc ? start[k0 + 2] - lo : 0;
    int pre = cnt;                               // inclusive prefix over the half-wavefront
    for (int off = 1; off < 32; off <<= 1) {
        const int t = __shfl_up(pre, off, 32);
        if (q >= off) pre += t;
    }
    if (q < nc) stencil[(size_t)slot * 27 + q] = make_int4(lo, mid, cnt, pre - cnt);
}

// Ring of the candidates that passed the box test, in planes (x, y | z | index with the
// fluid flag in bit 31), and the members of the wavefront's clusters (x, y, z, index): the
// ring's planes are read and written at consecutive entries, the members by broadcast
struct ListLds {
    v2d xy[128];
    double z[128];
    int jf[128];
    double4 mem[8];
    int clo[32], cmid[32], cpre[33];
};

// wave-wide ballot of a lane predicate straight into a scalar mask
__device__ __forceinline__ unsigned long long ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// P clusters of one cell per wavefront (P = 2: member slots 0..K-1 are cluster g's, K..2K-1
// cluster g1's): the candidates of the shared stencil are loaded, located and box-tested
// once for both, and each cluster's members unite into its own list.  Each lane holds one
// ring entry and tests it against the members in turn (their positions broadcast from LDS), so
// a member's hits are one ballot and the union of a cluster the OR of its members' ballots.
template <int K, int P>
__device__ __forceinline__ int cluster_list_one(const CaseConst& C, int g, const int4 info, int g1, const int4 info1,
                                                 const double4* __restrict__ pos, const int4* __restrict__ stencil,
                                                 int width, int* __restrict__ cl_nbr,
                                                 int* __restrict__ cl_len, int* __restrict__ selfin,
                                                 int* __restrict__ count_out, ListLds& L)
{
    constexpr int M = K * P;                           // member slots
    const int lane = threadIdx.x & 63;
    const int cell = info.z;                           // storage slot of the cluster's cell
    const int nc = C.dim == 2 ? 9 : 27;
    // lane l < M takes member slot l; the other lanes repeat the slots (l mod M), and slots
    // of absent members repeat the cluster's first member (which does not change the box)
    const int ms = lane % M;
    const bool second = P == 2 && ms >= K;
    const int4 inf = second ? info1 : info;
    const int mk = second ? ms - K : ms;
    const bool present = mk < inf.y;
    const int i = inf.x + (present ? mk : 0);
    const double4 pio = pos[i];
    const bool wall_i = (tag_of(pio.w) & 3) >= 2;
    const double4 pf = pos[info.x];
    const bool nx = !(pio.x - C.dmin[0] >= C.guard[0] && C.dmax[0] - pio.x >= C.guard[0]);
    const bool ny = !(pio.y - C.dmin[1] >= C.guard[1] && C.dmax[1] - pio.y >= C.guard[1]);
    const bool nz = !(pio.z - C.dmin[2] >= C.guard[2] && C.dmax[2] - pio.z >= C.guard[2]);
    const int ax = (ballot64(nx) != 0ull ? 1 : 0) | (ballot64(ny) != 0ull ? 2 : 0) | (ballot64(nz) != 0ull ? 4 : 0);
    const bool exact = ax == 0 ? !C.list_plain : !(C.list_img && (ax & ~C.img_axes) == 0);   // uniform
    const bool img = !exact && ax != 0;
    // the member's position: its images near the first member on the axes near a face
    double xi = pio.x, yi = pio.y, zi = pio.z;
    if (img) image_axes(C, ax, pf, xi, yi, zi);
    if (lane < M) L.mem[lane] = make_double4(xi, yi, zi, __longlong_as_double((long long)i));
    // members' bounding box relative to the first member (reduced over the M slots)
    double bx0, by0, bz0;
    if (exact) {
        bx0 = rel_exact(pio.x, pf.x, C.hwid[0], C.dwid[0]);
        by0 = rel_exact(pio.y, pf.y, C.hwid[1], C.dwid[1]);
        bz0 = rel_exact(pio.z, pf.z, C.hwid[2], C.dwid[2]);
    } else {
        bx0 = xi - pf.x;
        by0 = yi - pf.y;
        bz0 = zi - pf.z;
    }
    double bx1 = bx0, by1 = by0, bz1 = bz0;
    for (int off = 1; off < M; off <<= 1) {
        bx0 = fmin(bx0, __shfl_xor(bx0, off)); bx1 = fmax(bx1, __shfl_xor(bx1, off));
        by0 = fmin(by0, __shfl_xor(by0, off)); by1 = fmax(by1, __shfl_xor(by1, off));
        bz0 = fmin(bz0, __shfl_xor(bz0, off)); bz1 = fmax(bz1, __shfl_xor(bz1, off));
    }
    // the box as centre and half-extents (widened by 1e-12 of the extent and a few ulp of the
    // coordinates, far inside the test's margin): |d - c| - h per axis
    const double bcx = 0.5 * (bx0 + bx1), bcy = 0.5 * (by0 + by1), bcz = 0.5 * (bz0 + bz1);
    const double slack = 1.0e-12 * (bx1 - bx0 + by1 - by0 + bz1 - bz0) + 1.0e-15 * (fabs(bx0) + fabs(bx1) +
                         fabs(by0) + fabs(by1) + fabs(bz0) + fabs(bz1));
    const double bhx = 0.5 * (bx1 - bx0) + slack, bhy = 0.5 * (by1 - by0) + slack, bhz = 0.5 * (bz1 - bz0) + slack;
    const double band = 1.0e-8 * C.cut2;
    const double bound = C.cut2 + band;
    const double cut_lo = C.cut2 - band;
    // the box test needs one minimum image for all members: |x_j - x_first| < W/2 on every
    // axis for candidates within the cut-off of a member (members share a cell narrower than
    // two cut-offs), which eight cut-offs per width guarantee; narrower boxes test everything
    const double w8 = 64.0 * C.cut2;
    const bool use_box = C.dwid[0] * C.dwid[0] >= w8 && C.dwid[1] * C.dwid[1] >= w8 &&
                         (C.dim == 2 || C.dwid[2] * C.dwid[2] >= w8);
    // the stencil cells' particle ranges, ascending in rank, from the cell's record
    // (k_stencil): lane q < nc holds cell q's first index, fluid end, count and prefix.  Loaded
    // here, not earlier: held across the members' box it pushed the kernel into spilling
    __builtin_amdgcn_sched_barrier(0);
    const int4 srec = stencil[(size_t)cell * 27 + (lane < nc ? lane : 0)];
    if (lane < nc) {
        L.clo[lane] = srec.x;
        L.cmid[lane] = srec.y;
        L.cpre[lane] = srec.w;
    }
    if (lane == nc - 1) L.cpre[nc] = srec.w + srec.z;
    __builtin_amdgcn_wave_barrier();
    const int total = L.cpre[nc];
    // a block's lanes e = base + lane take the cell q with cpre[q] <= e: from the block's first
    // cell q0, the cells starting inside the block are applied in ascending order
    int q0 = 0;
    auto locate = [&](int base, int& j, int& fl) {
        int off = L.clo[q0] - L.cpre[q0], mid = L.cmid[q0];
        for (int q = q0 + 1; q < nc; ++q) {
            const int b = L.cpre[q];
            if (b >= base + 64) break;
            if (base + lane >= b) {
                off = L.clo[q] - b;
                mid = L.cmid[q];
            }
            q0 = q;
        }
        j = base + lane + off;
        fl = j < mid;
    };
    const unsigned long long below = (1ull << lane) - 1ull;
    int ulen = 0, ulen1 = 0;
    // per member slot m (uniform): its count in 32-bit field m mod 2 of cnt01 (m < 2), cnt23,
    // cnt45 or cnt67, and in bit m of flbits whether it has a fluid partner; bit m of pres:
    // slot m holds a member
    unsigned long long cnt01 = 0, cnt23 = 0, cnt45 = 0, cnt67 = 0;
    unsigned flbits = 0;
    const unsigned pres = ((1u << info.y) - 1u) | (P == 2 ? ((1u << info1.y) - 1u) << K : 0u);
    // per member m of each cluster, the number of its entries in the cluster's union (8 bits
    // each; more than one only where a stencil repeats cells)
    unsigned long long own = 0, own1 = 0;
    int* out = cl_nbr + (size_t)g * width;
    int* out1 = cl_nbr + (size_t)(P == 2 ? g1 : g) * width;
    // the walk over the candidates, specialised for the wavefront's path (EX: the reference's
    // periodic arithmetic per pair)
    auto walk = [&](auto ex_tag) {
        constexpr bool EX = decltype(ex_tag)::value;
        // ring write and read counts (uniform); blocks are read 64 at a time, so a block never
        // wraps around the ring's end
        unsigned wp = 0, rp = 0;
        // ring entries rp .. rp + nb - 1 against every member, then the block's union appended
        auto test_block = [&](int nb) {
            const int b0 = (int)(rp & 127u);
            const bool ok = lane < nb;
            // lanes past the block's end hold stale LDS: the plain path gives them a far
            // position (no hit, no exact re-test, no load), the exact path masks them
            v2d cxy = L.xy[b0 + lane];
            double cz = L.z[b0 + lane];
            if (!EX && !ok) {
                cxy = v2d{kFar, kFar};
                cz = kFar;
            }
            const int cjf = L.jf[b0 + lane];
            const int cj = cjf & 0x7fffffff;
            const unsigned long long okm = EX ? ballot64(ok) : ~0ull;
            const unsigned long long fm = ballot64(cjf < 0);
            unsigned long long ub = 0, ub1 = 0;
            // hits of member slot m on the block's entries, and those that are not the member
            auto member_hits = [&](int m, unsigned long long& hit, unsigned long long& other) {
                const double4 pm = L.mem[m];
                const int im = (int)__double_as_longlong(pm.w);
                double q2;
                if constexpr (!EX) {
                    const double dx = cxy.x - pm.x, dy = cxy.y - pm.y, dz = cz - pm.z;
                    q2 = fma(dz, dz, fma(dy, dy, dx * dx));
                    // within 1e-8 of the cut-off: the reference's arithmetic on the original
                    // positions
                    if (q2 >= cut_lo && q2 <= bound) {
                        const double4 pj = pos[cj], pi = pos[im];
                        const double ex = rel_exact(pj.x, pi.x, C.hwid[0], C.dwid[0]);
                        const double ey = rel_exact(pj.y, pi.y, C.hwid[1], C.dwid[1]);
                        const double ez = rel_exact(pj.z, pi.z, C.hwid[2], C.dwid[2]);
                        q2 = ex * ex + ey * ey + ez * ez;
                    }
                } else {
                    const double dx = rel_exact(cxy.x, pm.x, C.hwid[0], C.dwid[0]);
                    const double dy = rel_exact(cxy.y, pm.y, C.hwid[1], C.dwid[1]);
                    const double dz = rel_exact(cz, pm.z, C.hwid[2], C.dwid[2]);
                    q2 = dx * dx + dy * dy + dz * dz;
                }
                hit = ballot64(q2 <= C.cut2) & okm;
                other = hit & ~ballot64(cj == im);
            };
            // the plain path unrolls the members (static slots), the exact path does not
            auto account = [&](int m, unsigned long long hit, unsigned long long other) {
                const unsigned long long pc = (unsigned long long)__popcll(hit) << (32 * (m & 1));
                if (m < 2)
                    cnt01 += pc;
                else if (m < 4)
                    cnt23 += pc;
                else if (m < 6)
                    cnt45 += pc;
                else
                    cnt67 += pc;
                if ((hit & fm) != 0ull) flbits |= 1u << m;
                if (P == 2 && m >= K)
                    ub1 |= other;
                else
                    ub |= other;
            };
            if constexpr (!EX) {
#pragma unroll
                for (int m = 0; m < M; ++m) {
                    if (!((pres >> m) & 1u)) continue;      // absent member (uniform)
                    unsigned long long hit, other;
                    member_hits(m, hit, other);
                    account(m, hit, other);
                }
            } else {
#pragma unroll 1
                for (int m = 0; m < M; ++m) {
                    if (!((pres >> m) & 1u)) continue;
                    unsigned long long hit, other;
                    member_hits(m, hit, other);
                    account(m, hit, other);
                }
            }
            if ((ub >> lane) & 1ull) {
                const int at = ulen + (int)__popcll(ub & below);
                if (at < width) out[at] = cj;
                const unsigned m = (unsigned)(cj - info.x);
                if (m < (unsigned)info.y) own += 1ull << (8 * m);
            }
            ulen += (int)__popcll(ub);
            if (P == 2) {
                if ((ub1 >> lane) & 1ull) {
                    const int at = ulen1 + (int)__popcll(ub1 & below);
                    if (at < width) out1[at] = cj;
                    const unsigned m = (unsigned)(cj - info1.x);
                    if (m < (unsigned)info1.y) own1 += 1ull << (8 * m);
                }
                ulen1 += (int)__popcll(ub1);
            }
            rp += (unsigned)nb;
        };
        int jn = -1, fn = 0;
        q0 = 0;
        if (total > 0) locate(0, jn, fn);
        double4 pn = lane < total ? pos[jn] : make_double4(0.0, 0.0, 0.0, 0.0);
        // one loop for the blocks and the ring's last entries (one copy of the test code)
        for (int base = 0;; base += 64) {
          if (base < total) {
            // this block's candidates near the box, appended to the ring in candidate order
            double px = pn.x, py = pn.y, pz = pn.z;
            if (!EX && img) image_axes(C, ax, pf, px, py, pz);
            bool kp = base + lane < total;
            if (kp && use_box) {
                double dx, dy, dz;
                if constexpr (EX) {
                    dx = rel_exact(px, pf.x, C.hwid[0], C.dwid[0]);
                    dy = rel_exact(py, pf.y, C.hwid[1], C.dwid[1]);
                    dz = rel_exact(pz, pf.z, C.hwid[2], C.dwid[2]);
                } else {
                    dx = px - pf.x;
                    dy = py - pf.y;
                    dz = pz - pf.z;
                }
                const double ex = fmax(fabs(dx - bcx) - bhx, 0.0);
                const double ey = fmax(fabs(dy - bcy) - bhy, 0.0);
                const double ez = fmax(fabs(dz - bcz) - bhz, 0.0);
                kp = fma(ez, ez, fma(ey, ey, ex * ex)) <= bound;
            }
            const unsigned long long kb = ballot64(kp);
            if (kp) {
                const unsigned at = (wp + (unsigned)__popcll(kb & below)) & 127u;
                L.xy[at] = v2d{px, py};
                L.z[at] = pz;
                L.jf[at] = fn ? (jn | (int)0x80000000u) : jn;
            }
            wp += (unsigned)__popcll(kb);
            // the next block's loads are in flight while the ring is tested
            if (base + 64 < total) {
                locate(base + 64, jn, fn);
                if (base + 64 + lane < total) pn = pos[jn];
            }
          }
            const bool last = base + 64 >= total;
            __builtin_amdgcn_wave_barrier();
            if (wp - rp >= 64u || (last && wp != rp)) test_block(wp - rp >= 64u ? 64 : (int)(wp - rp));
            __builtin_amdgcn_wave_barrier();
            if (last && wp == rp) break;
        }
    };
    if (exact)
        walk(std::true_type{});
    else
        walk(std::false_type{});
    for (int off = 32; off > 0; off >>= 1) {
        own += __shfl_xor(own, off);
        if (P == 2) own1 += __shfl_xor(own1, off);
    }
    // NeighborCount: the member's own entry included; a wall particle with no fluid partner
    // has none (main.cpp:1416)
    const unsigned long long cpair = lane < 2 ? cnt01 : (lane < 4 ? cnt23 : (lane < 6 ? cnt45 : cnt67));
    const int cnt = (int)((cpair >> (32 * (lane & 1))) & 0xffffffffull);
    if (lane < M && present) count_out[i] = (wall_i && ((flbits >> lane) & 1u) == 0u) ? 0 : cnt;
    if (lane < info.y) selfin[info.x + lane] = (int)((own >> (8 * lane)) & 0xffull);
    if (P == 2 && lane < info1.y) selfin[info1.x + lane] = (int)((own1 >> (8 * lane)) & 0xffull);
    if (lane == 0) cl_len[g] = ulen < width ? ulen : width;
    if (P == 2 && lane == 0) cl_len[g1] = ulen1 < width ? ulen1 : width;
    __builtin_amdgcn_wave_barrier();       // this cluster's LDS reads precede the next one's writes
    return ulen > ulen1 ? ulen : ulen1;    // the caller keeps the largest for the capacity check
}

// six wavefronts per SIMD (at most 80 VGPRs, 32 bytes per lane spilled outside the candidate
// and ring loops): full rebuilds of the 1.27M-particle deck 1.21 -> 1.14 ms against five
// (96 VGPRs, no spills); seven measured the same as six
template <int K>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6, 8)))
k_cluster_lists(const CaseConst* __restrict__ cases, const int* __restrict__ slot_case, int nslots,
                const int* __restrict__ flags, int* __restrict__ any_flag,
                const double4* __restrict__ pos, const int4* __restrict__ stencil,
                const int4* __restrict__ cl_info,
                const int* __restrict__ cl_count, int width, int* __restrict__ cl_nbr,
                int* __restrict__ cl_len, int* __restrict__ selfin, int* __restrict__ count_out,
                int* __restrict__ max_len, int chunk)
{
    if (no_rebuild(any_flag)) return;
    __shared__ ListLds lds[4];
    // the wavefront's slot and case are uniform: readfirstlane lets the compiler keep the
    // case's constants in scalar registers (scalar loads) instead of vector loads in the loop
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    // chunk > 0 (and a grid of a multiple of 8): the work items (pairs of slots, or slots) go
    // in chunks of `chunk` consecutive items to the XCDs in turn, and workgroup b, which runs
    // on XCD b % 8, walks that XCD's chunks in order: the stencil cells' positions a wavefront
    // reads are then mostly in its XCD's L2 already.  Otherwise the workgroups stride over all
    // items in turn (a batch's flagged slots can lie in one region: chunks would idle XCDs).
    const bool xm = chunk > 0 && gridDim.x % kXcds == 0;
    const int nt = K <= 4 ? (nslots + 1) / 2 : nslots;
    const int C = xm ? chunk : max(nt, 1);
    const int nx = xm ? kXcds : 1;
    const int x = xm ? (int)(blockIdx.x % kXcds) : 0;
    const int lb = xm ? (int)(blockIdx.x / kXcds) : (int)blockIdx.x;
    const int step = 4 * (int)gridDim.x / nx;
    const int uend = ((nt + C - 1) / C - x + nx - 1) / nx * C;        // items of this XCD's chunks
    int longest = 0;
    const int4 none = make_int4(0, 0, 0, 0);
#define MPH_LIST_ONE(PP, CC, G, I, G1, I1)                                                                  \
    cluster_list_one<K, PP>(cases[CC], G, I, G1, I1, pos, stencil, width, cl_nbr, cl_len, selfin,          \
                            count_out, lds[wv])
    if (K <= 4) {
        // slots in pairs (2t, 2t + 1): two clusters of one cell share a wavefront
        for (int u = lb * 4 + wv; u < uend; u += step) {
            const int t = (u / C * nx + x) * C + u % C;     // grows with u: past the end, stop
            if (t >= nt) break;
            const int g0 = 2 * t, g1 = g0 + 1;
            // the cluster entries are loaded with the slots' cases, not after them
            const int4 i0 = cl_info[g0], i1 = cl_info[g1 < nslots ? g1 : g0];
            const int c0 = __builtin_amdgcn_readfirstlane(slot_case[g0]);
            const bool v0 = flags[c0] != 0 && g0 - cases[c0].cl_base < cl_count[c0];      // wave-uniform
            int c1 = c0;
            bool v1 = false;
            if (g1 < nslots) {
                c1 = __builtin_amdgcn_readfirstlane(slot_case[g1]);
                v1 = flags[c1] != 0 && g1 - cases[c1].cl_base < cl_count[c1];
            }
            if (v0 && v1 && c0 == c1 && i0.z == i1.z) {
                longest = max(longest, MPH_LIST_ONE(2, c0, g0, i0, g1, i1));
                continue;
            }
            if (v0) longest = max(longest, MPH_LIST_ONE(1, c0, g0, i0, 0, none));
            if (v1) longest = max(longest, MPH_LIST_ONE(1, c1, g1, i1, 0, none));
        }
    } else {
        for (int u = lb * 4 + wv; u < uend; u += step) {
            const int g = (u / C * nx + x) * C + u % C;
            if (g >= nt) break;
            const int c = __builtin_amdgcn_readfirstlane(slot_case[g]);
            if (flags[c] == 0 || g - cases[c].cl_base >= cl_count[c]) continue;     // wave-uniform
            longest = max(longest, MPH_LIST_ONE(1, c, g, cl_info[g], 0, none));
        }
    }
#undef MPH_LIST_ONE
    // one atomic per wavefront: an atomic per cluster on the one counter serialised the
    // kernel (about 4 ms for 318k clusters)
    if ((threadIdx.x & 63) == 0 && longest > 0) {
        atomicMax(max_len, longest);
        // a list longer than the width was cut: the rest of the step waits for the host
        // (width 0 is the counting pass that sizes the lists)
        if (width > 0 && longest > width) atomicOr(any_flag + 1, 1);
    }
}

// copy every cluster list into a wider table (the lists of cases that did not rebuild stay
// valid; the rebuilt ones are built again into the new table)
__global__ void __launch_bounds__(256)
k_widen_lists(const int* __restrict__ src, int w_src, int* __restrict__ dst, int w_dst, long long total)
{
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
        const long long g = e / w_src;
        dst[g * w_dst + (e - g * w_src)] = src[e];
    }
}

// -------------------------------------------------------------------------------------
// Cluster frame of a wavefront in the sweep kernels: which cluster, which row/slot.  The
// pair sweeps walk the union list through LDS blocks (walk_union_lds); the virial, run at
// output steps only, uses the register-pipelined walk_union (indices two steps ahead,
// partner data one step ahead in alternating register sets).

template <int K>
struct Frame {
    static constexpr int S = 64 / K;
    int g;          // global cluster slot
    int len;        // union length
    int row, slot;
    int i;          // member particle of this lane's row
    int first;      // first member (reference point of the mixed-precision sweeps)
    bool valid;     // row < cluster size
    const int* list;
    int width;      // entries reserved for the list
    int j0;         // list entry `lane` as loaded (the length not yet applied)
};

// The wavefront's cluster.  first_slot is the global slot of the block's first cluster (the
// block map's entry).  PIN: the slot's entry, union length and first 64 list entries (j0) are
// loaded together with the case's cluster count, and only then is the slot tested against the
// count; the four slots of a map entry lie inside the case's capacity (a multiple of 4), so
// the loads of an unused slot stay in bounds.  The empty asm takes the loaded values, so the
// compiler issues every load before the test instead of sinking them behind it: one round
// trip instead of two (sweep A without surface tension: 0.745 -> 0.733 ms on the 1.27M deck).
// Otherwise the loads follow the test and j0 is not loaded: sweep B and the mixed-precision
// sweeps, at their register budgets, measured slower with the loads held across the test.
template <int K, bool PIN = false>
__device__ __forceinline__ bool cluster_frame(const CaseConst& C, int first_slot, const int4* __restrict__ cl_info,
                                              const int* __restrict__ cl_count, int case_id,
                                              const int* __restrict__ cl_nbr, const int* __restrict__ cl_len,
                                              int width, Frame<K>& f, int wave = -1)
{
    const int lane = threadIdx.x & 63;
    f.g = first_slot + (wave >= 0 ? wave : (int)(threadIdx.x >> 6));   // the wavefront's slot of the entry
    f.width = width;
    f.list = cl_nbr + (size_t)f.g * width;
    int4 info;
    if (PIN) {
        const int count = cl_count[case_id];
        info = cl_info[f.g];
        f.len = cl_len[f.g];
        f.j0 = lane < width ? f.list[lane] : -1;
        const int base = C.cl_base;
        asm volatile("" ::"s"(count), "s"(base), "s"(info.x), "s"(info.y), "s"(f.len), "v"(f.j0));
        if (f.g - base >= count) return false;          // wave-uniform
    } else {
        if (f.g - C.cl_base >= cl_count[case_id]) return false;
        info = cl_info[f.g];
        f.len = cl_len[f.g];
        f.j0 = -1;
    }
    f.row = lane / Frame<K>::S;
    f.slot = lane % Frame<K>::S;
    f.valid = f.row < info.y;
    f.i = info.x + (f.valid ? f.row : 0);
    f.first = info.x;
    return true;
}

template <int K, class Load, class Eval>
__device__ __forceinline__ void walk_union(const Frame<K>& f, Load load, Eval eval)
{
    constexpr int S = Frame<K>::S;
    auto idx = [&](int step) { const int e = step * S + f.slot; return e < f.len ? f.list[e] : -1; };
    const int steps = (f.len + S - 1) / S;
    int j0 = idx(0), j1 = idx(1);
    auto A = load(j0 >= 0 ? j0 : f.i);
    for (int k = 0; k < steps; k += 2) {
        const int j2 = idx(k + 2);
        auto B = load(j1 >= 0 ? j1 : f.i);
        if (j0 >= 0 && j0 != f.i && f.valid) eval(A);
        const int j3 = idx(k + 3);
        A = load(j2 >= 0 ? j2 : f.i);
        if (j1 >= 0 && j1 != f.i && f.valid) eval(B);
        j0 = j2;
        j1 = j3;
    }
}

__device__ __forceinline__ v4d ld4(const double4* p) { return *reinterpret_cast<const v4d*>(p); }

struct Rec {
    v4d p, v, a;            // position+tag, velocity+P, gravity centre+PA (a only with surface tension)
};

// The same walk with the partner data staged through LDS in blocks of 64 union entries:
// all 64 lanes of the wavefront load one entry each (ascending sorted indices, so the
// loads of a block mostly share cache lines), store the records to this wavefront's LDS
// slice, and then the K rows sweep the block in K steps of S slots, each row reading the
// slot's record from LDS (K lanes read each address, a broadcast).  Every union entry is
// fetched from global memory once per cluster instead of once per member, and the next
// block's loads are in flight while the current block is evaluated.  Entries are evaluated
// without a branch: padding slots hold the far record and a row's own entry is caught by
// far_if_self in the evaluators, so the loop over a block is straight-line code.
// NREC = 3 adds the gravity-centre record (surface tension).  PL3: records without the tag
// and P, in three planes (x, y | vx, vy | z, vz): 48 bytes of LDS per entry instead of 64.
// Applied to each fetched record just before it is stored to LDS (its loads have landed by
// then): the periodic path's images near the cluster's first member, or nothing.
struct NoPrep {
    __device__ void operator()(Rec&) const {}
};

template <int K, int NREC, class Fetch, class Eval, bool PL3 = false, class Prep = NoPrep, bool J0 = false>
__device__ __forceinline__ void walk_union_lds(const Frame<K>& f, v2d* __restrict__ lds, Fetch fetch, Eval eval,
                                               Prep prep = Prep())
{
    // LDS slice: 2*NREC planes of 64 16-byte halves (x,y | z,w of each record), so the
    // 16 or 32 distinct entries one read touches are contiguous and free of bank conflicts
    constexpr int S = Frame<K>::S;
    const int lane = (int)(threadIdx.x & 63);
    const int len = __builtin_amdgcn_readfirstlane(f.len);      // one cluster per wavefront
    const int blocks = (len + 63) >> 6;
    auto idx = [&](int blk) { const int e = (blk << 6) + lane; return e < len ? f.list[e] : -1; };
    const Rec far = {v4d{kFar, kFar, kFar, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}, v4d{0.0, 0.0, 0.0, 0.0}};
    Rec r = far;
    // indices run two blocks ahead of the records, so a block's record loads never wait
    // on its index load
    // J0: block 0's entries came with the (pinned) frame; otherwise they are loaded here
    int jr = J0 ? (lane < len ? f.j0 : -1) : idx(0), jn = idx(1);
    if (jr >= 0) fetch(jr, r);
    for (int blk = 0; blk < blocks; ++blk) {
        // LDS operations of one wavefront complete in issue order: the reads of the
        // previous block are done before these stores land, and the stores are done
        // before the reads below; the barriers keep the compiler from moving LDS
        // accesses across the block boundary
        __builtin_amdgcn_wave_barrier();
        prep(r);
        if (PL3) {
            lds[0 * 64 + lane] = r.p.xy;
            lds[1 * 64 + lane] = r.v.xy;
            lds[2 * 64 + lane] = v2d{r.p.z, r.v.z};
        } else {
            lds[0 * 64 + lane] = r.p.xy;
            lds[1 * 64 + lane] = r.p.zw;
            lds[2 * 64 + lane] = r.v.xy;
            lds[3 * 64 + lane] = r.v.zw;
        }
        if (NREC > 2) {
            lds[4 * 64 + lane] = r.a.xy;
            lds[5 * 64 + lane] = r.a.zw;
        }
        __builtin_amdgcn_wave_barrier();
        if (blk + 1 < blocks) {
            jr = jn;
            jn = idx(blk + 2);
            if (jr >= 0)
                fetch(jr, r);
            else
                r = far;
        }
        const int rest = len - (blk << 6);
        const int nsteps = rest >= 64 ? K : (rest + S - 1) / S;
        // not unrolled: one pair per iteration keeps the sweeps within 80 VGPRs (six
        // wavefronts per SIMD), which measures faster than the unrolled forms
#pragma unroll 1
        for (int t = 0; t < nsteps; ++t) {
            const int e = t * S + f.slot;
            Rec q;
            if (PL3) {
                q.p.xy = lds[0 * 64 + e];
                q.v.xy = lds[1 * 64 + e];
                const v2d zz = lds[2 * 64 + e];
                q.p.z = zz.x;
                q.v.z = zz.y;
                q.p.w = 0.0;
                q.v.w = 0.0;
            } else {
                q.p.xy = lds[0 * 64 + e];
                q.p.zw = lds[1 * 64 + e];
                q.v.xy = lds[2 * 64 + e];
                q.v.zw = lds[3 * 64 + e];
            }
            if (NREC > 2) {
                q.a.xy = lds[4 * 64 + e];
                q.a.zw = lds[5 * 64 + e];
            }
            eval(q);
        }
    }
}

template <int K>
__device__ __forceinline__ double slot_sum(double v)
{
    for (int off = Frame<K>::S / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, Frame<K>::S);
    return v;
}

template <int K>
__device__ __forceinline__ int slot_sum_i(int v)
{
    for (int off = Frame<K>::S / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, Frame<K>::S);
    return v;
}

// Periodic path of a wavefront: 0 if no row is near a face (plain differences), 1 if the
// axes its rows are near all take images near the first member, 2 otherwise (the exact
// periodic difference per row and pair).  axes: the axes that need images (uniform).  A
// particle at least `guard` (list cut-off + Verlet skin) inside both faces of an axis has
// no list partner that wraps on that axis.
__device__ __forceinline__ int wrap_path(const CaseConst& C, const double4& p, int& axes)
{
    const bool nx = !(p.x - C.dmin[0] >= C.guard[0] && C.dmax[0] - p.x >= C.guard[0]);
    const bool ny = !(p.y - C.dmin[1] >= C.guard[1] && C.dmax[1] - p.y >= C.guard[1]);
    const bool nz = !(p.z - C.dmin[2] >= C.guard[2] && C.dmax[2] - p.z >= C.guard[2]);
    axes = (__ballot(nx) != 0ull ? 1 : 0) | (__ballot(ny) != 0ull ? 2 : 0) | (__ballot(nz) != 0ull ? 4 : 0);
    return axes == 0 ? 0 : ((axes & ~C.img_axes) == 0 ? 1 : 2);
}

template <bool ON>
struct ImagePrep {
    const CaseConst& C;
    const double4& xf;
    int axes;
    __device__ void operator()(Rec& r) const
    {
        if (!ON) return;
        double x = r.p.x, y = r.p.y, z = r.p.z;
        image_axes(C, axes, xf, x, y, z);
        r.p.x = x;
        r.p.y = y;
        r.p.z = z;
    }
};

// -------------------------------------------------------------------------------------
// sweep A: number densities, gravity centre, volumetric strain, divergence
// (main.cpp:1665-1727, 1831-1896) with the pointwise coefficients and pressures of
// main.cpp:1631-1662, 1898-1906, 1729-1740 fused into the epilogue.  Per pair: exact
// periodic difference, q^2, one reciprocal square root, kernel polynomials with their
// constant prefactors pulled out of the sums, branch-free with max(1 - r/h, 0).

struct PairA {
    double4 p;
    double vx, vy, vz;
};

struct AccA {
    double na, gx, gy, gz, ws, ds;
    int np;
};

// WRAP: the periodic path of wrap_path (pi holds the row's images on path 1)
template <int SURF, int K, int WRAP, bool COUNT, int MODE>
__device__ __forceinline__ void sweep_a_pairs(const CaseConst& C, const double* tab, const Frame<K>& f, int ti,
                                              const double4& pi, double vix, double viy, double viz,
                                              const double4* __restrict__ pos, const double4* vel,
                                              v2d* lds, AccA& acc, const double4& pf, int axes)
{
    constexpr bool SAMEAP = (MODE & 1) != 0;
    constexpr bool R1 = (MODE & 2) != 0 && !SURF;     // every InteractionRatio is 1: no tag needed
    constexpr bool NA = (MODE & 4) == 0 || SURF;      // the number density n^A is summed
    auto fetch = [&](int j, Rec& r) {
        if (R1) {
            // no tag needed: x, y, z only (two fewer registers per record in flight)
            const double4 p = pos[j];
            r.p = v4d{p.x, p.y, p.z, 0.0};
        } else {
            r.p = ld4(pos + j);
        }
        // P (vel.w) of partners is being written by this kernel: read x, y, z only
        const double4 u = vel[j];
        r.v = v4d{u.x, u.y, u.z, 0.0};
    };
    auto eval = [&](const Rec& q) {
        const double dx = WRAP == 2 ? rel_wrap(q.p.x, pi.x, C.hwid[0], C.dwid[0]) : q.p.x - pi.x;
        const double dy = WRAP == 2 ? rel_wrap(q.p.y, pi.y, C.hwid[1], C.dwid[1]) : q.p.y - pi.y;
        const double dz = WRAP == 2 ? rel_wrap(q.p.z, pi.z, C.hwid[2], C.dwid[2]) : q.p.z - pi.z;
        // a row's own entry (x_ij = 0 exactly) is kept off q^2 = 0: it then adds 1 to the
        // sum of w_P and to the pair count, which the epilogue takes back (selfin), r of
        // about 1e-150 to the n^A sum, below every other term's last bit, and exact zeros
        // elsewhere (x_ij = 0, u.x_ij = 0)
        const double q2 = fmax(fma(dz, dz, fma(dy, dy, dx * dx)), 1.0e-300);
        if (COUNT) acc.np += q2 <= C.hmax2;
        const double ir = rsqrt_pair3(q2);
        const double r = q2 * ir;
        const double oma = fmax(fma(-r, C.inv_ha, 1.0), 0.0);
        if (NA || SURF) {
            const double ratio = (R1 || C.ratio_one) ? 1.0 : tab[16 + ti * 4 + (tag_of(q.p.w) & 3)];
            const double wa = ratio * (oma * oma);
            acc.na = fma(wa, r, acc.na);      // sum of w_A r; 1/h_A joins the prefactor
            if (SURF) {
                acc.gx = fma(dx, wa, acc.gx);
                acc.gy = fma(dy, wa, acc.gy);
                acc.gz = fma(dz, wa, acc.gz);
            }
        }
        const double omp = SAMEAP ? oma : fmax(fma(-r, C.inv_hp, 1.0), 0.0);
        acc.ws = fma(omp, omp, acc.ws);
        const double ud = fma(q.v.z - viz, dz, fma(q.v.y - viy, dy, (q.v.x - vix) * dx));
        acc.ds = fma(ud, omp * ir, acc.ds);
    };
    walk_union_lds<K, 2, decltype(fetch), decltype(eval), R1, ImagePrep<WRAP == 1>, true>(f, lds, fetch, eval,
                                                                                        ImagePrep<WRAP == 1>{C, pf, axes});
}

// Without surface tension both sweeps run at six wavefronts per SIMD (at most 80 VGPRs, a
// few bytes spilled outside the pair loop), measured faster than five and seven, also with
// the three-plane records of MODE bit 1; with surface tension sweep A is held to five (92
// VGPRs).
// COUNT: also count the interacting pairs (r <= largest radius) into pair_count; only the
// timed steps that report pair rates use it.  MODE bit 0: the attractive and pressure
// kernels have the same radius (one weight per pair); bit 1: every InteractionRatio is 1, so
// without surface tension the partner records need no tag (three LDS planes); bit 2 (no
// surface tension only): n^A is not summed and DensityA not written.  Without surface
// tension the attractive pressure is zero (its coefficient a_t is), so n^A changes nothing
// in the step and is an output field only; the host computes it when the fields are read,
// by this kernel without bit 2 on the same positions and lists (the same sums, bit for bit).
//
// Without surface tension, workgroups of eight wavefronts: workgroup b (on XCD b % 8) takes the
// block map's entries (b % 8) + 8 (2 (b / 8) + h), h = 0, 1, consecutive entries of that XCD's
// run (the map holds eight padding entries, so the last pair stays in bounds).  Eight
// neighbouring clusters on one CU share more of their partner records in its L1: 655-656 ->
// 664-665 steps/s at the start of the 1.27M deck's flow, the developed flow within 0.3%;
// sweep B measured slower so (0.804 -> 0.818 ms), and one-wavefront workgroups slower for both.
// With surface tension, four (five wavefronts per SIMD are not a multiple of two).
template <int SURF, int K, bool COUNT, int MODE>
__global__ void __launch_bounds__(SURF ? 256 : 512) __attribute__((amdgpu_waves_per_eu(SURF ? 5 : 6, 8)))
k_sweep_a(const CaseConst* __restrict__ cases, const int2* __restrict__ cmap,
          const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
          const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
          const int* __restrict__ count_out,
          const double4* __restrict__ pos, double4* vel, double4* __restrict__ aux,
          double* __restrict__ dens_a, double* __restrict__ vstrain, double* __restrict__ divp,
          unsigned long long* __restrict__ pair_count, const int* __restrict__ selfin,
          const int* __restrict__ any_flag, int speculative)
{
    // no side effect beyond its outputs: the eager step queues it before it knows whether the
    // lists are rebuilt (speculative) and, if they are, once more after the rebuild; the
    // speculative launch returns at once when the move raised the rebuild flag.  The block's
    // map entry is loaded with the flags (one round trip for both).
    constexpr int WPB = SURF ? 4 : 8;         // wavefronts per workgroup
    const int entry = WPB == 4 ? (int)blockIdx.x
                               : __builtin_amdgcn_readfirstlane((int)((blockIdx.x & (kXcds - 1)) +
                                                                      kXcds * (2 * (blockIdx.x / kXcds) + (threadIdx.x >> 8))));
    const int wave = __builtin_amdgcn_readfirstlane((int)((threadIdx.x >> 6) & 3));
    const int2 m = cmap[entry];
    const int flag0 = any_flag[0], flag1 = any_flag[1];
    asm volatile("" ::"s"(m.x), "s"(m.y), "s"(flag0), "s"(flag1));
    if (flag1 != 0 || (speculative && flag0 != 0)) return;      // the same for the whole workgroup
    // a padding entry (m.x < 0) can share a workgroup with a real one: its wavefronts wait at
    // the table's barrier below with the others and leave after it
    const bool pad = m.x < 0;
    const CaseConst& C = cases[pad ? 0 : m.x];
    // the InteractionRatio table: only the n^A sums weigh pairs by it
    constexpr bool TAB = ((MODE & 4) == 0 || SURF) && !((MODE & 2) != 0 && !SURF);
    __shared__ double tab[32];                // [0,16): mu_ij, [16,32): InteractionRatio
    constexpr int S = Frame<K>::S;
    __shared__ v2d lds[WPB][4 * 64];      // per wavefront: {pos, vel} halves x 64 entries
    // the frame's loads are in flight while the table is filled; wavefronts of unused slots
    // leave after the barrier
    Frame<K> f;
    const bool live = !pad && cluster_frame<K, true>(C, m.y, cl_info, cl_count, m.x, cl_nbr, cl_len, width, f, wave);
    if (TAB) {
        if (threadIdx.x < 16) {
            tab[threadIdx.x] = C.muij[threadIdx.x >> 2][threadIdx.x & 3];
            tab[16 + threadIdx.x] = C.ir[threadIdx.x >> 2][threadIdx.x & 3];
        }
        __syncthreads();
    }
    if (!live) return;
    const int w = threadIdx.x >> 6;
    const double4 pi = pos[f.i];
    const int ti = tag_of(pi.w) & 3;
    const double vix = vel[f.i].x, viy = vel[f.i].y, viz = vel[f.i].z;
    // the epilogue's per-particle words, loaded with the row's own data (landed by the end)
    const int own_i = selfin[f.i], listed_i = count_out[f.i];
    AccA acc = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    // a wavefront takes a periodic path if any of its rows needs it (uniform branch)
    int axes;
    const int path = wrap_path(C, pi, axes);
    if (path == 0) {
        sweep_a_pairs<SURF, K, 0, COUNT, MODE>(C, tab, f, ti, pi, vix, viy, viz, pos, vel, lds[w], acc, pi, 0);
    } else if (path == 1) {
        const double4 pf = pos[__builtin_amdgcn_readfirstlane(f.first)];   // uniform
        double4 pim = pi;
        image_axes(C, axes, pf, pim.x, pim.y, pim.z);
        sweep_a_pairs<SURF, K, 1, COUNT, MODE>(C, tab, f, ti, pim, vix, viy, viz, pos, vel, lds[w], acc, pf,
                                               axes);
    } else {
        sweep_a_pairs<SURF, K, 2, COUNT, MODE>(C, tab, f, ti, pi, vix, viy, viz, pos, vel, lds[w], acc, pi, 0);
    }
    constexpr bool NA = (MODE & 4) == 0 || SURF;
    if (NA) acc.na = slot_sum<K>(acc.na);
    acc.ws = slot_sum<K>(acc.ws);
    acc.ds = slot_sum<K>(acc.ds);
    if (SURF) {
        acc.gx = slot_sum<K>(acc.gx);
        acc.gy = slot_sum<K>(acc.gy);
        acc.gz = slot_sum<K>(acc.gz);
    }
    int npair = COUNT ? slot_sum_i<K>(acc.np) : 0;
    if (f.slot == 0 && f.valid) {
        const int i = f.i;
        const int own = own_i;                 // the row's own entry: w_P(0) = 1 and one pair
        acc.ws -= (double)own;
        npair -= own;
        if (listed_i == 0)                     // no list (main.cpp:1416): all sums are zero
            acc = AccA{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
        const double na = acc.na * C.inv_ha * C.ca;
        const double ep = C.cp * acc.ws - C.n0p;
        const double dv = -C.dcp * acc.ds;
        const double kap = ep < 0.0 ? 0.0 : by_type(C.kbulk, ti);
        const double p = -by_type(C.lbulk, ti) * dv + kap * ep;
        double pa = 0.0;
        if (NA) {
            pa = by_type(C.cofa, ti) * (na - C.n0a) / C.l0;
            if (C.n0a <= na) pa = 0.0;
        }
        const double gs = C.cg * C.g_scale;
        vel[i].w = p;
        aux[i] = make_double4(acc.gx * gs, acc.gy * gs, acc.gz * gs, pa);
        if (NA) dens_a[i] = na;
        vstrain[i] = ep;
        divp[i] = dv;
    } else {
        npair = 0;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) npair += __shfl_xor(npair, off);
        if ((threadIdx.x & 63) == 0 && npair) {
            const int slot = (int)((blockIdx.x * (unsigned)WPB + (threadIdx.x >> 6)) & (kPairSlots - 1));
            atomicAdd(pair_count + slot * kPairStride, (unsigned long long)npair);
        }
    }
}

// -------------------------------------------------------------------------------------
// sweep B: pressure, attractive-pressure, diffuse-interface and viscous forces
// (main.cpp:1908-1938, 1742-1775, 1778-1829, 1941-1985), gravity and the velocity update
// (main.cpp:1988-2010).  Each force term is summed separately with its constant factors
// pulled out, and the terms are combined in the reference's order.

struct PairB {
    double4 p, v, a;
};

struct AccB {
    double px, py, pz;      // pressure + viscosity, with their constant factors
    double vx, vy, vz;      // (unused)
    double ax, ay, az;      // attractive pressure
    double sx, sy, sz;      // diffuse interface
};

template <int SURF, int K, int WRAP, bool SAMEPV>
__device__ __forceinline__ void sweep_b_pairs(const CaseConst& C, const double* tab, const Frame<K>& f, int ti,
                                              const double4& pi, const double4& vi, const double4& ai,
                                              const double4* __restrict__ pos, const double4* __restrict__ vel,
                                              const double4* __restrict__ aux, v2d* lds, AccB& acc,
                                              const double4& pf, int axes)
{
    const double mu_i = tab[ti * 4 + ti];
    const double kp = C.dcp * C.vol, kv = C.visc_c * (-C.dcv) * C.vol;
    const double kp_pi = kp * vi.w;
    auto fetch = [&](int j, Rec& r) {
        r.p = ld4(pos + j);
        r.v = ld4(vel + j);
        if (SURF) r.a = ld4(aux + j);
    };
    auto eval = [&](const Rec& q) {
        const double dx = WRAP == 2 ? rel_wrap(q.p.x, pi.x, C.hwid[0], C.dwid[0]) : q.p.x - pi.x;
        const double dy = WRAP == 2 ? rel_wrap(q.p.y, pi.y, C.hwid[1], C.dwid[1]) : q.p.y - pi.y;
        const double dz = WRAP == 2 ? rel_wrap(q.p.z, pi.z, C.hwid[2], C.dwid[2]) : q.p.z - pi.z;
        // a row's own entry (x_ij = 0 exactly): with one radius and no surface tension every
        // term below is (finite coefficient) * x_ij once q^2 is kept off zero (1/r = 1e150,
        // ud / r^2 = 0 * 1e300 = 0), so it adds exact zeros; otherwise it is mapped outside
        // every support
        const double q0 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double q2 = (SAMEPV && !SURF) ? fmax(q0, 1.0e-300) : far_if_self(q0);
        const double ir = rsqrt_pair(q2);
        const double r = q2 * ir;
        const int tj = tag_of(q.p.w) & 3;
        // pressure, and viscosity with the harmonic-mean coefficient of the type pair, both
        // along x_ij: one coefficient with their constant factors (kp, kv)
        const double ud = fma(q.v.z - vi.z, dz, fma(q.v.y - vi.y, dy, (q.v.x - vi.x) * dx));
        double c;
        if (SAMEPV) {
            // kp (P_i + P_j) w / r + kv mu ud w / r^3 = (w / r) (kp (P_i + P_j) + kv mu ud / r^2),
            // kv mu_ij from the table, and w / r = (1 - r/h) / r = max(1/r - 1/h, 0)
            // kv mu_ij of the type pair (with one viscosity for every type all entries are kv mu_i)
            const double kvmu = tab[32 + ti * 4 + tj];
            c = fmax(ir - C.inv_hp, 0.0) * fma(kvmu, ud * (ir * ir), fma(kp, q.v.w, kp_pi));
        } else {
            const double wp = fmax(fma(-r, C.inv_hp, 1.0), 0.0);
            const double wv = fmax(fma(-r, C.inv_hv, 1.0), 0.0);
            const double cp = (vi.w + q.v.w) * wp * ir;
            const double mu = C.mu_uniform ? mu_i : tab[ti * 4 + tj];
            const double cv = mu * ud * wv * (ir * ir * ir);
            c = fma(kp, cp, kv * cv);
        }
        acc.px = fma(c, dx, acc.px);
        acc.py = fma(c, dy, acc.py);
        acc.pz = fma(c, dz, acc.pz);
        if (SURF) {
            const double rij = C.ratio_one ? 1.0 : tab[16 + ti * 4 + tj];
            const double rji = C.ratio_one ? 1.0 : tab[16 + tj * 4 + ti];
            const double s = r * C.inv_ha;
            const double om = fmax(1.0 - s, 0.0);
            const double ca = fma(q.a.w, rji, ai.w * rij) * om * fma(-3.0, s, 1.0) * ir;
            acc.ax = fma(ca, dx, acc.ax);
            acc.ay = fma(ca, dy, acc.ay);
            acc.az = fma(ca, dz, acc.az);
            // diffuse interface with Q = R_ji G_j - R_ij G_i; a_j uses the type of i
            // (main.cpp:1792), so a_i factors out of the sum
            const double qx = fma(rji, q.a.x, -rij * ai.x);
            const double qy = fma(rji, q.a.y, -rij * ai.y);
            const double qz = fma(rji, q.a.z, -rij * ai.z);
            const double qd = fma(qz, dz, fma(qy, dy, qx * dx)) * (C.dcg * ir);
            const double w = C.cg * om;
            acc.sx = fma(-om, fma(qd, dx, w * qx), acc.sx);
            acc.sy = fma(-om, fma(qd, dy, w * qy), acc.sy);
            acc.sz = fma(-om, fma(qd, dz, w * qz), acc.sz);
        }
    };
    walk_union_lds<K, SURF ? 3 : 2, decltype(fetch), decltype(eval), false>(f, lds, fetch, eval,
                                                                          ImagePrep<WRAP == 1>{C, pf, axes});
}

// SAMEPV: the pressure and viscosity kernels have the same radius (one weight per pair).
// Six wavefronts per SIMD without surface tension, as sweep A; four with it (128 VGPRs
// instead of the compiler's 142: 1.045 -> 1.0 ms on the 64-case batch).
template <int SURF, int K, bool SAMEPV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SURF ? 4 : 6, 8)))
k_sweep_b(const CaseConst* __restrict__ cases, const int2* __restrict__ cmap,
          const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
          const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
          const int* __restrict__ count_out,
          const double4* __restrict__ pos, const double4* __restrict__ vel,
          const double4* __restrict__ aux, double4* __restrict__ vnext,
          double4* __restrict__ force, int* __restrict__ flags,
          int* __restrict__ any_flag, long long* __restrict__ rebuilds, double* __restrict__ wallc, int ncases)
{
    if (halted(any_flag)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && ncases > 0) {
        for (int c = 0; c < ncases; ++c) {
            if (flags[c]) rebuilds[c] += 1;
            flags[c] = 0;
            // wall centres advance once per step (main.cpp:2048-2052); the move kernel of
            // this step has already used the old value
            for (int t = 2; t < kTypes; ++t)
                for (int d = 0; d < 3; ++d)
                    wallc[(c * kTypes + t) * 3 + d] += cases[c].wvel[t][d] * cases[c].dt;
        }
        rebuilds[ncases] += 1;          // completed steps
        *any_flag = 0;
    }
    const int2 m = cmap[reversed_entry()];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    __shared__ double tab[48];                 // mu_ij, InteractionRatio, kv mu_ij
    constexpr int S = Frame<K>::S;
    constexpr int NREC = SURF ? 3 : 2;
    __shared__ v2d lds[4][2 * NREC * 64]; // per wavefront: record halves x 64 entries
    if (threadIdx.x < 16) {
        tab[threadIdx.x] = C.muij[threadIdx.x >> 2][threadIdx.x & 3];
        tab[16 + threadIdx.x] = C.ir[threadIdx.x >> 2][threadIdx.x & 3];
        tab[32 + threadIdx.x] = C.visc_c * (-C.dcv) * C.vol * C.muij[threadIdx.x >> 2][threadIdx.x & 3];
    }
    __syncthreads();
    Frame<K> f;
    if (!cluster_frame<K>(C, m.y, cl_info, cl_count, m.x, cl_nbr, cl_len, width, f)) return;
    const int w = threadIdx.x >> 6;
    const double4 pi = pos[f.i];
    const double4 vi = vel[f.i];
    const int ti = tag_of(pi.w) & 3;
    const double4 ai = SURF ? aux[f.i] : make_double4(0.0, 0.0, 0.0, 0.0);
    AccB acc = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    int axes;
    const int path = wrap_path(C, pi, axes);
    if (path == 0) {
        sweep_b_pairs<SURF, K, 0, SAMEPV>(C, tab, f, ti, pi, vi, ai, pos, vel, aux, lds[w], acc, pi, 0);
    } else if (path == 1) {
        const double4 pf = pos[__builtin_amdgcn_readfirstlane(f.first)];   // uniform
        double4 pim = pi;
        image_axes(C, axes, pf, pim.x, pim.y, pim.z);
        sweep_b_pairs<SURF, K, 1, SAMEPV>(C, tab, f, ti, pim, vi, ai, pos, vel, aux, lds[w], acc, pf, axes);
    } else {
        sweep_b_pairs<SURF, K, 2, SAMEPV>(C, tab, f, ti, pi, vi, ai, pos, vel, aux, lds[w], acc, pi, 0);
    }
    acc.px = slot_sum<K>(acc.px); acc.py = slot_sum<K>(acc.py); acc.pz = slot_sum<K>(acc.pz);
    if (SURF) {
        acc.ax = slot_sum<K>(acc.ax); acc.ay = slot_sum<K>(acc.ay); acc.az = slot_sum<K>(acc.az);
        acc.sx = slot_sum<K>(acc.sx); acc.sy = slot_sum<K>(acc.sy); acc.sz = slot_sum<K>(acc.sz);
    }
    if (f.slot != 0 || !f.valid) return;
    const int i = f.i;
    if (count_out[i] == 0)                     // no list (main.cpp:1416): all sums are zero
        acc = AccB{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double fx = acc.px, fy = acc.py, fz = acc.pz;       // pressure + viscosity
    if (SURF) {
        const double ca = C.dca * C.vol;
        const double cs = by_type(C.adiff, ti) * C.s_scale;
        fx += acc.ax * ca; fy += acc.ay * ca; fz += acc.az * ca;
        fx += acc.sx * cs; fy += acc.sy * cs; fz += acc.sz * cs;
    }
    double4 vn = make_double4(vi.x, vi.y, vi.z, 0.0);
    if (ti < 2) {
        const double ms = by_type(C.mass, ti);
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
//   diffuse       a_i (G_i . x) R_ij w_G (h_G/R2g)/l0  and  a_i (G_i . x) R_ij dw_G r (h_G/R2g)/l0
// with the reference's per-term cut-offs.  Velocities are those after the step's update.

template <int K>
__global__ void __launch_bounds__(256)
k_virial(const CaseConst* __restrict__ cases, const int2* __restrict__ cmap,
         const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
         const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
         const int* __restrict__ count_out, const double4* __restrict__ pos,
         const double4* __restrict__ vel, const double4* __restrict__ vnext,
         const double4* __restrict__ aux, double* __restrict__ vir)
{
    const int2 m = cmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    Frame<K> f;
    if (!cluster_frame<K>(C, m.y, cl_info, cl_count, m.x, cl_nbr, cl_len, width, f)) return;
    const double4 pi = pos[f.i];
    const double4 vi = vnext[f.i];
    const double P = vel[f.i].w;
    const double4 ai = aux[f.i];
    const int ti = tag_of(pi.w) & 3;
    const double a = C.adiff[ti];
    const double gsc = C.s_scale / C.vol;      // h_G / R2g / l0
    double tp = 0.0, ta = 0.0, tv = 0.0, ts = 0.0;
    auto load = [&](int j) { return j; };
    auto eval = [&](int j) {
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
    };
    walk_union<K>(f, load, eval);
    tp = slot_sum<K>(tp);
    ta = slot_sum<K>(ta);
    tv = slot_sum<K>(tv);
    ts = slot_sum<K>(ts);
    if (f.slot != 0 || !f.valid) return;
    const double tr = count_out[f.i] == 0 ? 0.0 : ((tp + ta) + tv) + ts;
    vir[f.i] = -tr / (C.dim == 2 ? 2.0 : 3.0);
}

}  // namespace mph

namespace mph {

// -------------------------------------------------------------------------------------
// Neighbour lists in CSR form on the device, keyed by original particle id: the reference's
// NeighborPtr / NeighborInd (main.cpp:1329-1596) mapped through ParticleIndex, each row
// ascending with self included (mph_get_neighbors_noalloc).  Row lengths are the lists'
// NeighborCount; a scan gives the row starts; each member of each cluster then rebuilds its
// row from the cluster's union list with the reference's exact cut-off test at the rebuild
// positions, and sorts it in LDS.  The scan's tile sums and an error counter live in a
// caller-free scratch buffer.

constexpr int kCsrThreads = 256;
constexpr int kCsrItems = 16;
constexpr int kCsrTile = kCsrThreads * kCsrItems;

__global__ void __launch_bounds__(256)
k_scratch_clear(int* __restrict__ w, int count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) w[k] = 0;
}

// ptr[o + 1] = NeighborCount of the particle with original id o, ptr[0] = 0
__global__ void __launch_bounds__(256)
k_csr_counts(int begin, int n, const int* __restrict__ pidx, const int* __restrict__ nout,
             long long* __restrict__ ptr)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k == 0) ptr[0] = 0;
    if (k < n) ptr[pidx[begin + k] + 1] = nout[begin + k];
}

__global__ void __launch_bounds__(kCsrThreads)
k_csr_tile_sums(const long long* __restrict__ v, int n, long long* __restrict__ sums)
{
    const int base = blockIdx.x * kCsrTile;
    long long s = 0;
    for (int k = threadIdx.x; k < kCsrTile; k += kCsrThreads) s += base + k < n ? v[base + k] : 0ll;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    __shared__ long long ws[kCsrThreads / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        long long t = 0;
        for (int w = 0; w < kCsrThreads / 64; ++w) t += ws[w];
        sums[blockIdx.x] = t;
    }
}

// exclusive scan of `total` tile sums in place, one workgroup of 1024 threads
__global__ void __launch_bounds__(1024)
k_csr_scan_sums(long long* __restrict__ data, int total)
{
    __shared__ long long part[1024];
    const int tid = threadIdx.x;
    const int chunk = (total + 1023) / 1024;
    const int lo = min(total, tid * chunk), hi = min(total, lo + chunk);
    long long s = 0;
    for (int k = lo; k < hi; ++k) s += data[k];
    part[tid] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long v = tid >= off ? part[tid - off] : 0ll;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long run = part[tid] - s;
    for (int k = lo; k < hi; ++k) {
        const long long v = data[k];
        data[k] = run;
        run += v;
    }
}

// inclusive scan of each tile in place, plus the tile's offset
__global__ void __launch_bounds__(kCsrThreads)
k_csr_tile_scan(long long* __restrict__ v, int n, const long long* __restrict__ offsets)
{
    __shared__ long long part[kCsrThreads];
    const int base = blockIdx.x * kCsrTile + threadIdx.x * kCsrItems;
    long long x[kCsrItems], s = 0;
#pragma unroll
    for (int q = 0; q < kCsrItems; ++q) {
        x[q] = base + q < n ? v[base + q] : 0ll;
        s += x[q];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < kCsrThreads; off <<= 1) {
        const long long t = (int)threadIdx.x >= off ? part[threadIdx.x - off] : 0ll;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    long long run = offsets[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
    for (int q = 0; q < kCsrItems; ++q) {
        run += x[q];
        if (base + q < n) v[base + q] = run;
    }
}

// one wavefront per cluster of case `cc`, member by member: the row of a member with a list
// is itself plus every union entry within the cut-off at the rebuild positions (the
// reference's test, main.cpp:1398-1408), as original ids, sorted ascending in LDS (bitonic,
// rows padded to `pad`, a power of two above the list width); a row whose length differs
// from NeighborCount raises err
__global__ void __launch_bounds__(256)
k_csr_rows(const CaseConst* __restrict__ cases, int cc, const int4* __restrict__ cl_info,
           const int* __restrict__ cl_count, const int* __restrict__ cl_nbr, const int* __restrict__ cl_len,
           int width, const double4* __restrict__ snap, const int* __restrict__ pidx,
           const int* __restrict__ nout, const long long* __restrict__ ptr, int* __restrict__ ind, int pad,
           int* __restrict__ err)
{
    extern __shared__ int rows[];
    const CaseConst& C = cases[cc];
    const int nwv = (int)(blockDim.x >> 6);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    int* row = rows + (size_t)wv * pad;
    const int ncl = cl_count[cc];
    for (int q = blockIdx.x * nwv + wv; q < ncl; q += gridDim.x * nwv) {
        const int g = C.cl_base + q;
        const int4 info = cl_info[g];
        const int len = cl_len[g];
        const int* list = cl_nbr + (size_t)g * width;
        for (int m = 0; m < info.y; ++m) {
            const int i = info.x + m;
            const int want = nout[i];
            if (want == 0) continue;                       // no list (main.cpp:1416)
            const double4 pi = snap[i];
            int cnt = 0;
            for (int base = 0; base < len; base += 64) {
                const int e = base + lane;
                bool in = false;
                int o = 0;
                if (e < len) {
                    const int j = list[e];
                    const double4 pj = snap[j];
                    const double dx = rel_exact(pj.x, pi.x, C.hwid[0], C.dwid[0]);
                    const double dy = rel_exact(pj.y, pi.y, C.hwid[1], C.dwid[1]);
                    const double dz = rel_exact(pj.z, pi.z, C.hwid[2], C.dwid[2]);
                    in = j != i && dx * dx + dy * dy + dz * dz <= C.cut2;
                    o = pidx[j];
                }
                const unsigned long long b = __ballot(in);
                const int at = cnt + (int)__popcll(b & below);
                if (in && at < pad) row[at] = o;
                cnt += (int)__popcll(b);
            }
            // the particle itself, once per visit of its own cell (more than once where the
            // stencil repeats cells on an axis of fewer than three)
            const int nself = want - cnt;
            if (nself < 1 || cnt + nself > pad) {            // cannot happen with consistent lists
                if (lane == 0) atomicAdd(err, 1);
                continue;
            }
            for (int k = lane; k < nself; k += 64) row[cnt + k] = pidx[i];
            const int L = cnt + nself;
            int P = 1;
            while (P < L) P <<= 1;
            for (int k = L + lane; k < P; k += 64) row[k] = 0x7fffffff;
            __builtin_amdgcn_wave_barrier();
            for (int k = 2; k <= P; k <<= 1)
                for (int j = k >> 1; j > 0; j >>= 1) {
                    for (int a = lane; a < P; a += 64) {
                        const int b2 = a ^ j;
                        if (b2 > a) {
                            const int x = row[a], y = row[b2];
                            const bool up = (a & k) == 0;
                            if (up ? x > y : x < y) {
                                row[a] = y;
                                row[b2] = x;
                            }
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                }
            const long long start = ptr[pidx[i]];
            for (int k = lane; k < L; k += 64) ind[start + k] = row[k];
            __builtin_amdgcn_wave_barrier();                 // the row's reads precede the next row's writes
        }
    }
}

}  // namespace mph
