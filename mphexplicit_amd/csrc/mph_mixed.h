// Mixed-precision pair sweeps (MPH_MIXED): pair geometry, kernel weights and force sums
// in FP32; positions, velocity integration and the density/pressure state in FP64.
//
// The weakly compressible pressure P = -lambda D + kappa e^P with e^P = sum w_P - N0p is a
// difference of two numbers near N0p, so sweep A accumulates its density sums (n^A,
// sum w_P, the divergence sum) in FP64 from FP32 pair weights; sweep B's force sums stay
// FP32 and are combined with their FP64 prefactors in the epilogue.  The rest of the step
// (move, wrap, rebuild, lists, update) is the FP64 path.
//
// Geometry: the loader lanes convert each union entry to an FP32 offset from the
// cluster's first member x_f, r_j = (float)(x_j - x_f); each row holds o_i = (float)(x_i
// - x_f), and x_ij = r_j - o_i.  Offsets are a few cut-off radii, so the FP32 rounding is
// about 1e-7 of the interaction radius.  On the periodic path both offsets are minimum
// images relative to x_f and x_ij is folded once more into [-W/2, W/2) in FP32.  Records
// are 16-byte float4, half the LDS and register footprint of the FP64 sweep.
#pragma once

#include "mph_kernels.h"

namespace mph {

typedef float v4f __attribute__((ext_vector_type(4)));

struct RecF {
    v4f p, v, a;            // offset+tag bits, velocity+P, gravity centre+PA (a with surface tension)
};

__device__ __forceinline__ float fold_f(float d, float hw, float w)
{
    return d >= hw ? d - w : (d < -hw ? d + w : d);
}

// FP32 offset of x from the reference point xf on one axis: the plain difference (WRAP 0),
// the image near xf on the axes that need images (WRAP 1, `on`) or on every axis (WRAP 2).
// On WRAP 1 those axes take exact images near xf (wrap_path), so the offsets of a row and
// its partners differ by the minimum image and need no fold per pair.
template <int WRAP>
__device__ __forceinline__ float offset_f(double x, double xf, double hw, double w, bool on = true)
{
    return (float)((WRAP == 2 || (WRAP == 1 && on)) ? rel_wrap(x, xf, hw, w) : x - xf);
}

// FP32 counterparts of kFar and far_if_self: squares of the far offset stay finite in FP32
constexpr float kFarF = 1.0e18f;
__device__ __forceinline__ float far_if_self_f(float q2) { return q2 > 0.f ? q2 : 1.0e36f; }


// The FP32 counterpart of walk_union_lds (branch-free entries, the far record in padding
// slots), with NREC planes of 64 float4 records.  Each lane loads its union entry for the
// next block (load, into a Raw) while the current block is evaluated, and conv turns it into
// the FP32 record when the lane stores it to LDS.  With Raw = Rec (FP64 data) the loads stay
// in flight during the evaluation; with Raw = RecF load converts at once (fewer registers)
// and conv passes the record through.
template <int K, int NREC, class Raw, class Load, class Conv, class Eval>
__device__ __forceinline__ void walk_union_lds_f(const Frame<K>& f, v4f* __restrict__ lds, Load load, Conv conv,
                                                 Eval eval)
{
    constexpr int S = Frame<K>::S;
    const int lane = (int)(threadIdx.x & 63);
    const int len = __builtin_amdgcn_readfirstlane(f.len);
    const int blocks = (len + 63) >> 6;
    auto idx = [&](int blk) { const int e = (blk << 6) + lane; return e < len ? f.list[e] : -1; };
    const RecF far = {v4f{kFarF, kFarF, kFarF, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}, v4f{0.f, 0.f, 0.f, 0.f}};
    Raw r = {};
    int jr = idx(0), jn = idx(1);
    bool have = jr >= 0;
    if (have) load(jr, r);
    for (int blk = 0; blk < blocks; ++blk) {
        // in-order LDS operations of one wavefront; the barriers only stop the compiler
        // from moving LDS accesses across the block boundary
        __builtin_amdgcn_wave_barrier();
        const RecF s = have ? conv(r) : far;
        lds[0 * 64 + lane] = s.p;
        lds[1 * 64 + lane] = s.v;
        if (NREC > 2) lds[2 * 64 + lane] = s.a;
        __builtin_amdgcn_wave_barrier();
        if (blk + 1 < blocks) {
            jr = jn;
            jn = idx(blk + 2);
            have = jr >= 0;
            if (have) load(jr, r);
        }
        const int rest = len - (blk << 6);
        const int nsteps = rest >= 64 ? K : (rest + S - 1) / S;
#pragma unroll 1
        for (int t = 0; t < nsteps; ++t) {
            const int e = t * S + f.slot;
            RecF q;
            q.p = lds[0 * 64 + e];
            q.v = lds[1 * 64 + e];
            if (NREC > 2) q.a = lds[2 * 64 + e];
            eval(q);
        }
    }
}

// -------------------------------------------------------------------------------------
// sweep A, mixed: same sums and epilogue as k_sweep_a (main.cpp:1665-1727, 1831-1906,
// 1729-1740)

template <int SURF, int K, int WRAP, bool COUNT, bool NA>
__device__ __forceinline__ void sweep_a_pairs_f(const CaseConst& C, const float* tabf, const Frame<K>& f, int ti,
                                                const double4& pf, float ox, float oy, float oz,
                                                float vix, float viy, float viz,
                                                const double4* __restrict__ pos, const double4* vel,
                                                v4f* lds, AccA& acc, float& gxf, float& gyf,
                                                float& gzf, int axes)
{
    const float hwx = (float)C.hwid[0], hwy = (float)C.hwid[1], hwz = (float)C.hwid[2];
    const float wx = (float)C.dwid[0], wy = (float)C.dwid[1], wz = (float)C.dwid[2];
    const float inv_ha = (float)C.inv_ha, inv_hp = (float)C.inv_hp, hmax2 = (float)C.hmax2;
    auto load = [&](int j, Rec& r) {
        r.p = ld4(pos + j);
        // P (vel.w) of partners is being written by this kernel: read x, y, z only
        const double4 u = vel[j];
        r.v = v4d{u.x, u.y, u.z, 0.0};
    };
    auto conv = [&](const Rec& r) {
        RecF s;
        s.p = v4f{offset_f<WRAP>(r.p.x, pf.x, C.hwid[0], C.dwid[0], axes & 1),
                  offset_f<WRAP>(r.p.y, pf.y, C.hwid[1], C.dwid[1], axes & 2),
                  offset_f<WRAP>(r.p.z, pf.z, C.hwid[2], C.dwid[2], axes & 4), __int_as_float(tag_of(r.p.w) & 3)};
        s.v = v4f{(float)r.v.x, (float)r.v.y, (float)r.v.z, 0.f};
        return s;
    };
    auto eval = [&](const RecF& q) {
        float dx = q.p.x - ox, dy = q.p.y - oy, dz = q.p.z - oz;
        if (WRAP == 2) {
            dx = fold_f(dx, hwx, wx);
            dy = fold_f(dy, hwy, wy);
            dz = fold_f(dz, hwz, wz);
        }
        // a row's own entry (zero offset difference) is kept off q^2 = 0 as in k_sweep_a; its
        // w_P(0) = 1 and pair count are taken back in the epilogue (selfin)
        const float q2 = fmaxf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)), 1.0e-36f);
        if (COUNT) acc.np += q2 <= hmax2;
        const float ir = __builtin_amdgcn_rsqf(q2);
        const float r = q2 * ir;
        if (NA || SURF) {
            const float sa = r * inv_ha;
            const float oma = fmaxf(1.f - sa, 0.f);
            const float ratio = C.ratio_one ? 1.f : tabf[16 + ti * 4 + __float_as_int(q.p.w)];
            const float wa = ratio * (oma * oma);
            acc.na += (double)(wa * sa);
            if (SURF) {
                gxf = fmaf(dx, wa, gxf);
                gyf = fmaf(dy, wa, gyf);
                gzf = fmaf(dz, wa, gzf);
            }
        }
        const float omp = fmaxf(fmaf(-r, inv_hp, 1.f), 0.f);
        acc.ws += (double)(omp * omp);
        const float ud = fmaf(q.v.z - viz, dz, fmaf(q.v.y - viy, dy, (q.v.x - vix) * dx));
        acc.ds += (double)(ud * (omp * ir));
    };
    walk_union_lds_f<K, 2, Rec>(f, lds, load, conv, eval);
}

// Without surface tension the mixed sweeps are held to seven wavefronts per SIMD (72
// VGPRs, a few bytes spilled outside the pair loop): measured faster than six and five and
// than the compiler's choice (four).  COUNT: count the interacting pairs (timed steps); NA:
// sum n^A (without surface tension only an output field, as in k_sweep_a).
template <int SURF, int K, bool COUNT, bool NA>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SURF ? 1 : 7, 8)))
k_sweep_a_mixed(const CaseConst* __restrict__ cases, const int2* __restrict__ cmap,
                const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
                const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
                const int* __restrict__ count_out,
                const double4* __restrict__ pos, double4* vel, double4* __restrict__ aux,
                double* __restrict__ dens_a, double* __restrict__ vstrain, double* __restrict__ divp,
                unsigned long long* __restrict__ pair_count, const int* __restrict__ selfin,
                const int* __restrict__ any_flag, int speculative)
{
    // as k_sweep_a: a speculative launch returns at once when the lists are being rebuilt
    if (halted(any_flag) || (speculative && any_flag[0] != 0)) return;
    const int2 m = cmap[blockIdx.x];
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    __shared__ float tabf[32];                 // [16,32): InteractionRatio
    __shared__ v4f lds[4][2 * 64];             // per wavefront: {offset, vel} x 64 entries
    if (threadIdx.x < 16) {
        tabf[threadIdx.x] = (float)C.muij[threadIdx.x >> 2][threadIdx.x & 3];
        tabf[16 + threadIdx.x] = (float)C.ir[threadIdx.x >> 2][threadIdx.x & 3];
    }
    __syncthreads();
    Frame<K> f;
    if (!cluster_frame<K>(C, m.y, cl_info, cl_count, m.x, cl_nbr, cl_len, width, f)) return;
    const int w = threadIdx.x >> 6;
    const double4 pi = pos[f.i];
    const double4 pf = pos[f.first];
    const int ti = tag_of(pi.w) & 3;
    const float vix = (float)vel[f.i].x, viy = (float)vel[f.i].y, viz = (float)vel[f.i].z;
    AccA acc = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
    float gxf = 0.f, gyf = 0.f, gzf = 0.f;
    int axes;
    const int path = wrap_path(C, pi, axes);
#define MPH_SWEEP_A_F(P)                                                                              \
    sweep_a_pairs_f<SURF, K, P, COUNT, NA || SURF>(C, tabf, f, ti, pf, offset_f<P>(pi.x, pf.x, C.hwid[0], C.dwid[0], axes & 1), \
                                offset_f<P>(pi.y, pf.y, C.hwid[1], C.dwid[1], axes & 2),              \
                                offset_f<P>(pi.z, pf.z, C.hwid[2], C.dwid[2], axes & 4), vix, viy, viz, \
                                pos, vel, lds[w], acc, gxf, gyf, gzf, axes)
    if (path == 0)
        MPH_SWEEP_A_F(0);
    else if (path == 1)
        MPH_SWEEP_A_F(1);
    else
        MPH_SWEEP_A_F(2);
#undef MPH_SWEEP_A_F
    if (NA || SURF) acc.na = slot_sum<K>(acc.na);
    acc.ws = slot_sum<K>(acc.ws);
    acc.ds = slot_sum<K>(acc.ds);
    if (SURF) {
        acc.gx = slot_sum<K>((double)gxf);
        acc.gy = slot_sum<K>((double)gyf);
        acc.gz = slot_sum<K>((double)gzf);
    }
    int npair = COUNT ? slot_sum_i<K>(acc.np) : 0;
    if (f.slot == 0 && f.valid) {
        const int i = f.i;
        const int own = selfin[i];
        acc.ws -= (double)own;
        npair -= own;
        if (count_out[i] == 0)                 // no list (main.cpp:1416): all sums are zero
            acc = AccA{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0};
        const double na = acc.na * C.ca;
        const double ep = C.cp * acc.ws - C.n0p;
        const double dv = -C.dcp * acc.ds;
        const double kap = ep < 0.0 ? 0.0 : C.kbulk[ti];
        const double p = -C.lbulk[ti] * dv + kap * ep;
        double pa = 0.0;
        if (NA || SURF) {
            pa = C.cofa[ti] * (na - C.n0a) / C.l0;
            if (C.n0a <= na) pa = 0.0;
        }
        const double gs = C.cg * C.g_scale;
        vel[i].w = p;
        aux[i] = make_double4(acc.gx * gs, acc.gy * gs, acc.gz * gs, pa);
        if (NA || SURF) dens_a[i] = na;
        vstrain[i] = ep;
        divp[i] = dv;
    } else {
        npair = 0;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) npair += __shfl_xor(npair, off);
        if ((threadIdx.x & 63) == 0 && npair) {
            const int slot = (int)((blockIdx.x * 4u + (threadIdx.x >> 6)) & (kPairSlots - 1));
            atomicAdd(pair_count + slot * kPairStride, (unsigned long long)npair);
        }
    }
}

// -------------------------------------------------------------------------------------
// sweep B, mixed: force terms of k_sweep_b (main.cpp:1908-1938, 1742-1829, 1941-1985),
// gravity and the FP64 velocity update (main.cpp:1988-2010)

struct AccBF {
    float px, py, pz, vx, vy, vz, ax, ay, az, sx, sy, sz;
};

template <int SURF, int K, int WRAP, bool SAMEPV>
__device__ __forceinline__ void sweep_b_pairs_f(const CaseConst& C, const float* tabf, const Frame<K>& f, int ti,
                                                const double4& pf, float ox, float oy, float oz,
                                                const v4f& vi, const v4f& ai,
                                                const double4* __restrict__ pos, const double4* __restrict__ vel,
                                                const double4* __restrict__ aux, v4f* lds, AccBF& acc,
                                                int axes)
{
    const float hwx = (float)C.hwid[0], hwy = (float)C.hwid[1], hwz = (float)C.hwid[2];
    const float wx = (float)C.dwid[0], wy = (float)C.dwid[1], wz = (float)C.dwid[2];
    const float inv_ha = (float)C.inv_ha, inv_hp = (float)C.inv_hp, inv_hv = (float)C.inv_hv;
    const float dcg = (float)C.dcg, cg = (float)C.cg;
    const float mu_i = tabf[ti * 4 + ti];
    // SAMEPV: one FP32 sum of the pressure and viscosity terms with their constant factors
    const float kpf = (float)(C.dcp * C.vol), kp_pi = kpf * vi.w;
    // converted as loaded: measured faster here than keeping the FP64 loads in flight
    auto load = [&](int j, RecF& r) {
        const double4 p = pos[j];
        r.p = v4f{offset_f<WRAP>(p.x, pf.x, C.hwid[0], C.dwid[0], axes & 1),
                  offset_f<WRAP>(p.y, pf.y, C.hwid[1], C.dwid[1], axes & 2),
                  offset_f<WRAP>(p.z, pf.z, C.hwid[2], C.dwid[2], axes & 4), __int_as_float(tag_of(p.w) & 3)};
        const double4 u = vel[j];
        r.v = v4f{(float)u.x, (float)u.y, (float)u.z, (float)u.w};
        if (SURF) {
            const double4 a = aux[j];
            r.a = v4f{(float)a.x, (float)a.y, (float)a.z, (float)a.w};
        }
    };
    auto conv = [](const RecF& r) { return r; };
    auto eval = [&](const RecF& q) {
        float dx = q.p.x - ox, dy = q.p.y - oy, dz = q.p.z - oz;
        if (WRAP == 2) {
            dx = fold_f(dx, hwx, wx);
            dy = fold_f(dy, hwy, wy);
            dz = fold_f(dz, hwz, wz);
        }
        const float q2 = far_if_self_f(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        const float ir = __builtin_amdgcn_rsqf(q2);
        const float r = q2 * ir;
        const int tj = __float_as_int(q.p.w);
        const float ud = fmaf(q.v.z - vi.z, dz, fmaf(q.v.y - vi.y, dy, (q.v.x - vi.x) * dx));
        if (SAMEPV) {
            // (w / r) (kp (P_i + P_j) + kv mu ud / r^2) with w / r = max(1/r - 1/h, 0), as the
            // FP64 sweep
            const float kvmu = tabf[32 + ti * 4 + tj];     // kv mu_ij of the type pair
            const float c = fmaxf(ir - inv_hp, 0.f) * fmaf(kvmu, ud * (ir * ir), fmaf(kpf, q.v.w, kp_pi));
            acc.px = fmaf(c, dx, acc.px);
            acc.py = fmaf(c, dy, acc.py);
            acc.pz = fmaf(c, dz, acc.pz);
        } else {
            const float cp = (vi.w + q.v.w) * fmaxf(fmaf(-r, inv_hp, 1.f), 0.f) * ir;
            acc.px = fmaf(cp, dx, acc.px);
            acc.py = fmaf(cp, dy, acc.py);
            acc.pz = fmaf(cp, dz, acc.pz);
            const float mu = C.mu_uniform ? mu_i : tabf[ti * 4 + tj];
            const float cv = mu * ud * fmaxf(fmaf(-r, inv_hv, 1.f), 0.f) * (ir * ir * ir);
            acc.vx = fmaf(cv, dx, acc.vx);
            acc.vy = fmaf(cv, dy, acc.vy);
            acc.vz = fmaf(cv, dz, acc.vz);
        }
        if (SURF) {
            const float rij = C.ratio_one ? 1.f : tabf[16 + ti * 4 + tj];
            const float rji = C.ratio_one ? 1.f : tabf[16 + tj * 4 + ti];
            const float s = r * inv_ha;
            const float om = fmaxf(1.f - s, 0.f);
            const float ca = fmaf(q.a.w, rji, ai.w * rij) * om * fmaf(-3.f, s, 1.f) * ir;
            acc.ax = fmaf(ca, dx, acc.ax);
            acc.ay = fmaf(ca, dy, acc.ay);
            acc.az = fmaf(ca, dz, acc.az);
            // Q = R_ji G_j - R_ij G_i; a_j uses the type of i (main.cpp:1792)
            const float qx = fmaf(rji, q.a.x, -rij * ai.x);
            const float qy = fmaf(rji, q.a.y, -rij * ai.y);
            const float qz = fmaf(rji, q.a.z, -rij * ai.z);
            const float qd = fmaf(qz, dz, fmaf(qy, dy, qx * dx)) * (dcg * ir);
            const float wg = cg * om;
            acc.sx = fmaf(-om, fmaf(qd, dx, wg * qx), acc.sx);
            acc.sy = fmaf(-om, fmaf(qd, dy, wg * qy), acc.sy);
            acc.sz = fmaf(-om, fmaf(qd, dz, wg * qz), acc.sz);
        }
    };
    walk_union_lds_f<K, SURF ? 3 : 2, RecF>(f, lds, load, conv, eval);
}

template <int K>
__device__ __forceinline__ float slot_sum_f(float v)
{
    for (int off = Frame<K>::S / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, Frame<K>::S);
    return v;
}

template <int SURF, int K, bool SAMEPV>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(SURF ? 1 : 7, 8)))
k_sweep_b_mixed(const CaseConst* __restrict__ cases, const int2* __restrict__ cmap,
                const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
                const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
                const int* __restrict__ count_out,
                const double4* __restrict__ pos, const double4* __restrict__ vel,
                const double4* __restrict__ aux, double4* __restrict__ vnext,
                double4* __restrict__ force, int* __restrict__ flags,
                int* __restrict__ any_flag, long long* __restrict__ rebuilds, double* __restrict__ wallc,
                int ncases)
{
    if (halted(any_flag)) return;
    if (blockIdx.x == 0 && threadIdx.x == 0 && ncases > 0) {
        for (int c = 0; c < ncases; ++c) {
            if (flags[c]) rebuilds[c] += 1;
            flags[c] = 0;
            // wall centres advance once per step (main.cpp:2048-2052)
            for (int t = 2; t < kTypes; ++t)
                for (int d = 0; d < 3; ++d)
                    wallc[(c * kTypes + t) * 3 + d] += cases[c].wvel[t][d] * cases[c].dt;
        }
        rebuilds[ncases] += 1;          // completed steps
        *any_flag = 0;
    }
    const int2 m = cmap[reversed_entry()];     // the map walked backwards, as k_sweep_b does
    if (m.x < 0) return;
    const CaseConst& C = cases[m.x];
    __shared__ float tabf[48];                 // mu_ij, InteractionRatio, kv mu_ij
    constexpr int NREC = SURF ? 3 : 2;
    __shared__ v4f lds[4][NREC * 64];          // per wavefront: records x 64 entries
    if (threadIdx.x < 16) {
        tabf[threadIdx.x] = (float)C.muij[threadIdx.x >> 2][threadIdx.x & 3];
        tabf[16 + threadIdx.x] = (float)C.ir[threadIdx.x >> 2][threadIdx.x & 3];
        tabf[32 + threadIdx.x] = (float)(C.visc_c * (-C.dcv) * C.vol * C.muij[threadIdx.x >> 2][threadIdx.x & 3]);
    }
    __syncthreads();
    Frame<K> f;
    if (!cluster_frame<K>(C, m.y, cl_info, cl_count, m.x, cl_nbr, cl_len, width, f)) return;
    const int w = threadIdx.x >> 6;
    const double4 pi = pos[f.i];
    const double4 pf = pos[f.first];
    const double4 vi = vel[f.i];
    const int ti = tag_of(pi.w) & 3;
    const double4 ai = SURF ? aux[f.i] : make_double4(0.0, 0.0, 0.0, 0.0);
    const v4f vif = v4f{(float)vi.x, (float)vi.y, (float)vi.z, (float)vi.w};
    const v4f aif = v4f{(float)ai.x, (float)ai.y, (float)ai.z, (float)ai.w};
    AccBF acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int axes;
    const int path = wrap_path(C, pi, axes);
#define MPH_SWEEP_B_F(P)                                                                              \
    sweep_b_pairs_f<SURF, K, P, SAMEPV>(C, tabf, f, ti, pf, offset_f<P>(pi.x, pf.x, C.hwid[0], C.dwid[0], axes & 1), \
                                        offset_f<P>(pi.y, pf.y, C.hwid[1], C.dwid[1], axes & 2),      \
                                        offset_f<P>(pi.z, pf.z, C.hwid[2], C.dwid[2], axes & 4), vif, aif, \
                                        pos, vel, aux, lds[w], acc, axes)
    if (path == 0)
        MPH_SWEEP_B_F(0);
    else if (path == 1)
        MPH_SWEEP_B_F(1);
    else
        MPH_SWEEP_B_F(2);
#undef MPH_SWEEP_B_F
    acc.px = slot_sum_f<K>(acc.px); acc.py = slot_sum_f<K>(acc.py); acc.pz = slot_sum_f<K>(acc.pz);
    acc.vx = slot_sum_f<K>(acc.vx); acc.vy = slot_sum_f<K>(acc.vy); acc.vz = slot_sum_f<K>(acc.vz);
    if (SURF) {
        acc.ax = slot_sum_f<K>(acc.ax); acc.ay = slot_sum_f<K>(acc.ay); acc.az = slot_sum_f<K>(acc.az);
        acc.sx = slot_sum_f<K>(acc.sx); acc.sy = slot_sum_f<K>(acc.sy); acc.sz = slot_sum_f<K>(acc.sz);
    }
    if (f.slot != 0 || !f.valid) return;
    const int i = f.i;
    if (count_out[i] == 0)                     // no list (main.cpp:1416): all sums are zero
        acc = AccBF{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const double cp = SAMEPV ? 1.0 : C.dcp * C.vol;      // SAMEPV sums carry their factors
    const double cv = C.visc_c * (-C.dcv) * C.vol;
    double fx = (double)acc.px * cp, fy = (double)acc.py * cp, fz = (double)acc.pz * cp;
    if (SURF) {
        const double ca = C.dca * C.vol;
        const double cs = C.adiff[ti] * C.s_scale;
        fx += (double)acc.ax * ca; fy += (double)acc.ay * ca; fz += (double)acc.az * ca;
        fx += (double)acc.sx * cs; fy += (double)acc.sy * cs; fz += (double)acc.sz * cs;
    }
    if (!SAMEPV) {
        fx += (double)acc.vx * cv; fy += (double)acc.vy * cv; fz += (double)acc.vz * cv;
    }
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

}  // namespace mph
