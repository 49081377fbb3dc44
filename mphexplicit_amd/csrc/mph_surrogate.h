// Pairwise-message surrogate: BF16 MFMA forward, FP32 backward data path, MFMA weight
// gradients, Adam.  The network and its training have no counterpart in the reference;
// the features are built from the reference's quantities (periodic difference
// main.cpp:1680, neighbour set main.cpp:1408, VolStrainP e^P = sum w_P - N0p with N0p
// main.cpp:1850-1856, 948).
//
// Matrix-core tiling (v_mfma_f32_16x16x32_bf16, one wavefront): the pairs are the N
// dimension.  A tile is 16 pairs of one particle; layer outputs are D = W^T (hidden x in)
// * X^T (in x 16 pairs).  With the operand layouts of this instruction (lane l holds, for
// column/pair l%16, the eight K positions 8*(l/16)..+7 of B, and rows 4*(l/16)..+3 of D),
// the four 16-row output tiles a lane holds for its pair are exactly the eight K values
// it must supply for each 32-wide K slice of the next layer, up to a fixed permutation of
// K: position q of lane group g in slice s is hidden unit 32 s + 16 (q/4) + 4 g + q%4.
// The weights are packed once per update with the same permutation, so activations never
// leave their registers between layers.
#pragma once

#include "mph_kernels.h"

namespace mph {

constexpr int kSurIn = 17, kSurHid = 64, kSurOut = 3;
constexpr int kSurNP = 9603;
constexpr int kOffW1 = 0;                          // [17][64], row 16 = bias of layer 1
constexpr int kOffW2 = kOffW1 + kSurIn * kSurHid;  // [64][64]
constexpr int kOffB2 = kOffW2 + kSurHid * kSurHid;
constexpr int kOffW3 = kOffB2 + kSurHid;
constexpr int kOffB3 = kOffW3 + kSurHid * kSurHid;
constexpr int kOffW4 = kOffB3 + kSurHid;           // [64][3]
constexpr int kOffB4 = kOffW4 + kSurHid * kSurOut;
static_assert(kOffB4 + kSurOut == kSurNP, "parameter layout");

// packed A operands: layer 1 (4 M tiles), layers 2 and 3 (4 M tiles x 2 K slices),
// layer 4 (1 M tile x 2 K slices), 64 lanes x 8 bf16 each
constexpr int kPkL1 = 0, kPkL2 = 4, kPkL3 = 12, kPkL4 = 20, kPkTiles = 22;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct SurScales {
    float inv_u, a_scale, dens_scale, inv_h, inv_n0p;
    float dt, gx, gy, gz;
    double cut2;            // message radius squared (pairs with r^2 <= cut2)
};

__device__ __forceinline__ int k_hidden(int s, int g, int q) { return 32 * s + 16 * (q >> 2) + 4 * g + (q & 3); }

__device__ __forceinline__ float sur_sigmoid(float z) { return __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }
__device__ __forceinline__ float sur_silu(float z) { return z * sur_sigmoid(z); }
__device__ __forceinline__ float sur_dsilu(float z)
{
    const float s = sur_sigmoid(z);
    return s * (1.0f + z * (1.0f - s));
}

// FP32 master weights -> packed bf16 operands, and bf16-rounded FP32 copies of W2, W3, W4
// for the backward data path (wr has the layout of w)
__global__ void __launch_bounds__(256)
k_sur_pack(const float* __restrict__ w, bf16x8* __restrict__ pk, float* __restrict__ wr)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kPkTiles * 64) {
        const int T = t >> 6, lane = t & 63, row = lane & 15, g = lane >> 4;
        bf16x8 v;
        for (int q = 0; q < 8; ++q) {
            float x = 0.0f;
            if (T < kPkL2) {
                const int k = 8 * g + q;
                if (k < kSurIn) x = w[kOffW1 + k * kSurHid + 16 * T + row];
            } else if (T < kPkL4) {
                const int L = T < kPkL3 ? 0 : 1, u = T - (L ? kPkL3 : kPkL2);
                const int mt = u >> 1, s = u & 1;
                x = w[(L ? kOffW3 : kOffW2) + k_hidden(s, g, q) * kSurHid + 16 * mt + row];
            } else {
                const int s = T - kPkL4;
                if (row < kSurOut) x = w[kOffW4 + k_hidden(s, g, q) * kSurOut + row];
            }
            v[q] = (__bf16)x;
        }
        pk[t] = v;
    }
    for (int p = t; p < kSurNP; p += gridDim.x * blockDim.x) {
        const bool bias = (p >= kOffB2 && p < kOffW3) || (p >= kOffB3 && p < kOffW4) || p >= kOffB4;
        wr[p] = bias ? w[p] : (float)(__bf16)w[p];
    }
}

// pairs within the message radius (r^2 <= cut2), per fluid particle, from the cluster union
// lists: thread (cluster slot, member), like the list build
__global__ void __launch_bounds__(256)
k_sur_pairs(const CaseConst* __restrict__ cases, const int* __restrict__ slot_case, int nslots,
            const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
            const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width, int K,
            const double4* __restrict__ pos, double cut2, int wh, int* __restrict__ pcnt, int* __restrict__ pnbr)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int g = (int)(t / K), m = (int)(t % K);
    if (g >= nslots) return;
    const int c = slot_case[g];
    const CaseConst& C = cases[c];
    if (g - C.cl_base >= cl_count[c]) return;
    const int4 info = cl_info[g];
    if (m >= info.y) return;
    const int i = info.x + m;
    const double4 pi = pos[i];
    int cnt = 0;
    if ((tag_of(pi.w) & 3) < 2) {
        const int* list = cl_nbr + (size_t)g * width;
        const int len = cl_len[g];
        int* out = pnbr + (size_t)i * wh;
        for (int e = 0; e < len; ++e) {
            const int j = list[e];
            if (j == i) continue;
            double dx, dy, dz;
            rel3(C, pi, pos[j], dx, dy, dz);
            if (dx * dx + dy * dy + dz * dz <= cut2 && cnt < wh) out[cnt++] = j;
        }
    }
    pcnt[i] = cnt;
}

__device__ __forceinline__ int sur_class(int t) { return t < 2 ? t : 2; }

// Matrix-core operands of the network staged in LDS once per block and shared by its
// wavefronts (22.5 KB): held in registers they took 124 VGPRs and left two wavefronts per
// SIMD for a kernel that waits on gathers
struct SurW {
    bf16x8 A[kPkTiles * 64];            // packed A operands, [tile][lane]
    float b2[kSurHid], b3[kSurHid], b4[4];
    // operand of this lane
    __device__ __forceinline__ bf16x8 a(int tile, int lane) const { return A[tile * 64 + lane]; }
    // rows 4 g .. 4 g + 3 of the 16-row output tile t of layer 2 or 3
    __device__ __forceinline__ f32x4 bias(const float* b, int t, int g) const
    {
        return *reinterpret_cast<const f32x4*>(b + 16 * t + 4 * g);
    }
};

// all threads of the block: stage the operands, then a block barrier
__device__ __forceinline__ void sur_stage_weights(SurW& W, const bf16x8* __restrict__ pk,
                                                  const float* __restrict__ w)
{
    for (int t = threadIdx.x; t < kPkTiles * 64; t += blockDim.x) W.A[t] = pk[t];
    for (int t = threadIdx.x; t < kSurHid; t += blockDim.x) {
        W.b2[t] = w[kOffB2 + t];
        W.b3[t] = w[kOffB3 + t];
    }
    if (threadIdx.x < 4) W.b4[threadIdx.x] = threadIdx.x < kSurOut ? w[kOffB4 + threadIdx.x] : 0.0f;
    __syncthreads();
}

// One tile of 16 pairs of particle i (lane l: pair l%16, lane group g = l/16): the
// features, symmetric under i <-> j, the three layers on the matrix cores and the
// antisymmetric message m = c0 e + c1 u + c2 (u.e) e.  Adds the tile's message sum to
// (ax, ay, az) in every lane.  TRAIN stores X, Z1..Z3 and the basis at pair slot P.
// SCATTER also adds m_ji = -m_ij to msum[j] (atomically) for the pairs with scat set.
template <bool TRAIN, bool SCATTER = false>
__device__ __forceinline__ void sur_tile(const SurW& W, const SurScales& sc, int lane, bool valid, float dx, float dy,
                                         float dz, int ci, int cj, float di, float dj, float ux, float uy, float uz,
                                         size_t P, __bf16* __restrict__ X, float* __restrict__ Z1,
                                         float* __restrict__ Z2, float* __restrict__ Z3, float* __restrict__ Bs,
                                         float& ax, float& ay, float& az, float4* __restrict__ msum = nullptr,
                                         int j = 0, bool scat = false)
{
    const int g = lane >> 4;
    float f[16];
    float ex, ey, ez, ue;
    {
        const float r = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
        const float ir = valid ? 1.0f / r : 0.0f;
        ex = dx * ir; ey = dy * ir; ez = dz * ir;
        ue = fmaf(uz, ez, fmaf(uy, ey, ux * ex));
        const float s = r * sc.inv_h;
        f[0] = s;
        f[1] = s * s;
        f[2] = ue;
        f[3] = fmaf(uz, uz, fmaf(uy, uy, ux * ux));
        const int lo = ci < cj ? ci : cj, hi = ci < cj ? cj : ci;
        const int cp = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);   // unordered class pair
        for (int q = 0; q < 6; ++q) f[4 + q] = cp == q ? 1.0f : 0.0f;
        f[10] = di + dj;
        f[11] = (di - dj) * (di - dj);
        f[12] = di * dj;
        f[13] = 0.0f; f[14] = 0.0f; f[15] = 0.0f;
        if (!valid) {
            for (int q = 0; q < 16; ++q) f[q] = 0.0f;
            ex = ey = ez = ux = uy = uz = ue = 0.0f;
        }
    }
    bf16x8 b1;
    for (int q = 0; q < 8; ++q) {
        const float x = g == 0 ? f[q] : (g == 1 ? f[8 + q] : ((g == 2 && q == 0) ? 1.0f : 0.0f));
        b1[q] = (__bf16)x;
    }
    if (TRAIN) *reinterpret_cast<bf16x8*>(X + P * 32 + 8 * g) = b1;
    f32x4 z[4];
    for (int t = 0; t < 4; ++t) {
        z[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL1 + t, lane), b1, z[t], 0, 0, 0);
        if (TRAIN) *reinterpret_cast<f32x4*>(Z1 + P * 64 + 16 * t + 4 * g) = z[t];
    }
    bf16x8 bh[2];
    for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 8; ++q) bh[s][q] = (__bf16)sur_silu(z[2 * s + (q >> 2)][q & 3]);
    for (int t = 0; t < 4; ++t) {
        z[t] = W.bias(W.b2, t, g);
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL2 + 2 * t, lane), bh[0], z[t], 0, 0, 0);
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL2 + 2 * t + 1, lane), bh[1], z[t], 0, 0, 0);
        if (TRAIN) *reinterpret_cast<f32x4*>(Z2 + P * 64 + 16 * t + 4 * g) = z[t];
    }
    for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 8; ++q) bh[s][q] = (__bf16)sur_silu(z[2 * s + (q >> 2)][q & 3]);
    for (int t = 0; t < 4; ++t) {
        z[t] = W.bias(W.b3, t, g);
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL3 + 2 * t, lane), bh[0], z[t], 0, 0, 0);
        z[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL3 + 2 * t + 1, lane), bh[1], z[t], 0, 0, 0);
        if (TRAIN) *reinterpret_cast<f32x4*>(Z3 + P * 64 + 16 * t + 4 * g) = z[t];
    }
    for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 8; ++q) bh[s][q] = (__bf16)sur_silu(z[2 * s + (q >> 2)][q & 3]);
    // output layer: rows 0..2 of lane group 0 are the coefficients of pair p
    f32x4 o = g == 0 ? *reinterpret_cast<const f32x4*>(W.b4) : f32x4{0.f, 0.f, 0.f, 0.f};
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL4, lane), bh[0], o, 0, 0, 0);
    o = __builtin_amdgcn_mfma_f32_16x16x32_bf16(W.a(kPkL4 + 1, lane), bh[1], o, 0, 0, 0);
    // m_ij = c0 e + c1 u + c2 (u.e) e: with symmetric coefficients m_ji = -m_ij
    const float c0 = (g == 0 && valid) ? o[0] : 0.f, c1 = (g == 0 && valid) ? o[1] : 0.f;
    const float c2 = (g == 0 && valid) ? o[2] : 0.f;
    const float ce = fmaf(c2, ue, c0);
    float mx = fmaf(c1, ux, ce * ex), my = fmaf(c1, uy, ce * ey), mz = fmaf(c1, uz, ce * ez);
    if (TRAIN && g == 0) {
        *reinterpret_cast<float4*>(Bs + P * 8) = make_float4(ex, ey, ez, ux);
        *reinterpret_cast<float4*>(Bs + P * 8 + 4) = make_float4(uy, uz, ue, 0.f);
    }
    if (SCATTER && g == 0 && scat) {
        atomicAdd(&msum[j].x, -mx);
        atomicAdd(&msum[j].y, -my);
        atomicAdd(&msum[j].z, -mz);
    }
    for (int off = 8; off > 0; off >>= 1) {
        mx += __shfl_xor(mx, off, 16);
        my += __shfl_xor(my, off, 16);
        mz += __shfl_xor(mz, off, 16);
    }
    ax += mx; ay += my; az += mz;
}

// Training forward over a list of particles with precomputed pair lists (k_sur_pairs).
// Persistent wavefronts, one particle at a time; stores the bf16 inputs X [P][32], the FP32
// pre-activations Z1..Z3 [P][64] and the message basis Bs [P][8] of every pair slot
// P = 16 * toff[k] + e (e = pair number).  apred[k] = a_i, with .w = a_scale * rho_0/rho_i.
__global__ void __launch_bounds__(256)
k_sur_forward_train(const CaseConst* __restrict__ cases, const int* __restrict__ plist, int nlist,
                    const int* __restrict__ toff, const double4* __restrict__ pos,
                    const double4* __restrict__ velf, const double* __restrict__ strain,
                    const int* __restrict__ pcnt, const int* __restrict__ pnbr, int wh,
                    const bf16x8* __restrict__ pk, const float* __restrict__ w, SurScales sc,
                    float4* __restrict__ apred, __bf16* __restrict__ X, float* __restrict__ Z1,
                    float* __restrict__ Z2, float* __restrict__ Z3, float* __restrict__ Bs)
{
    const CaseConst& C = cases[0];
    const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4;
    __shared__ SurW W;
    sur_stage_weights(W, pk, w);
    const int nw = gridDim.x * (blockDim.x >> 6);
    const int wid = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    for (int k = wid; k < nlist; k += nw) {
        const int i = plist[k];
        const double4 pi = pos[i];
        const int ti = tag_of(pi.w) & 3;
        const double4 vi = velf[i];
        const float di = (float)strain[i] * sc.inv_n0p * sc.dens_scale;
        const int c = pcnt[i];
        const int ntile = (c + 15) >> 4;
        float ax = 0.f, ay = 0.f, az = 0.f;
        for (int tt = 0; tt < ntile; ++tt) {
            const int e = 16 * tt + p;
            const bool valid = e < c;
            const int j = valid ? pnbr[(size_t)i * wh + e] : i;
            const double4 pj = pos[j];
            double dxd, dyd, dzd;
            rel3(C, pi, pj, dxd, dyd, dzd);
            const double4 vj = velf[j];
            const float dj = (float)strain[j] * sc.inv_n0p * sc.dens_scale;
            sur_tile<true>(W, sc, lane, valid, (float)dxd, (float)dyd, (float)dzd, sur_class(ti),
                           sur_class(tag_of(pj.w) & 3), di, dj, (float)(vj.x - vi.x) * sc.inv_u,
                           (float)(vj.y - vi.y) * sc.inv_u, (float)(vj.z - vi.z) * sc.inv_u,
                           (size_t)16 * toff[k] + e, X, Z1, Z2, Z3, Bs, ax, ay, az);
        }
        if (lane == 0) {
            // messages are forces per unit volume of the reference density: a_i scales
            // with rho_0 / rho_i
            const float rr = (float)(C.mass[0] / C.mass[ti]) * sc.a_scale;
            apred[k] = make_float4(ax * rr, ay * rr, az * rr, rr);
        }
    }
}

// Inference forward fused with the pair search: persistent wavefronts over the clusters;
// for each fluid member the 64 lanes test 64 union-list entries at a time (exact periodic
// difference, r <= h), ballot-compact the pairs with their FP32 offsets into this
// wavefront's LDS queue, and every 16 queued pairs run one matrix tile.
// Each unordered fluid pair is evaluated once, by its lower sorted index i: the features
// are symmetric and the message antisymmetric, so m_ji = -m_ij goes to j by an atomic
// add.  Fluid-wall pairs are evaluated by the fluid particle.  Sums go to msum (sorted
// order, zeroed by the caller); k_sur_finish turns them into a_i.  Five wavefronts per
// SIMD (92 VGPRs; 29.4 KB of LDS per block, five blocks per CU): measured 10% faster than
// the compiler's four.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5, 8)))
k_sur_infer(const CaseConst* __restrict__ cases, const int4* __restrict__ cl_info, const int* __restrict__ cl_count,
            const int* __restrict__ cl_nbr, const int* __restrict__ cl_len, int width,
            const double4* __restrict__ pos, const double4* __restrict__ velf, const double* __restrict__ strain,
            const bf16x8* __restrict__ pk, const float* __restrict__ w, SurScales sc, float4* __restrict__ msum)
{
    constexpr int kQ = 80;                       // queue: 15 left over + 64 new at most
    __shared__ int qj[4][kQ];
    __shared__ float4 qd[4][kQ];                 // dx, dy, dz (FP32 offsets), class of j
    const CaseConst& C = cases[0];
    const int lane = threadIdx.x & 63, p = lane & 15, g = lane >> 4, wv = threadIdx.x >> 6;
    __shared__ SurW W;
    sur_stage_weights(W, pk, w);
    const int nw = gridDim.x * (blockDim.x >> 6);
    const int wid = blockIdx.x * (blockDim.x >> 6) + wv;
    const int ncl = cl_count[0];
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int local = wid; local < ncl; local += nw) {
        const int gsl = C.cl_base + local;
        const int4 info = cl_info[gsl];
        const int* list = cl_nbr + (size_t)gsl * width;
        const int len = cl_len[gsl];
        for (int m = 0; m < info.y; ++m) {
            const int i = info.x + m;
            const double4 pi = pos[i];
            const int ti = tag_of(pi.w) & 3;
            if (ti >= 2) continue;
            const double4 vi = velf[i];
            const float di = (float)strain[i] * sc.inv_n0p * sc.dens_scale;
            const int ci = sur_class(ti);
            float ax = 0.f, ay = 0.f, az = 0.f;
            int count = 0;
            auto run_tile = [&](int nvalid) {
                __builtin_amdgcn_wave_barrier();
                const bool valid = p < nvalid;
                const int j = valid ? qj[wv][p] : i;
                const float4 d = qd[wv][p];
                const double4 vj = velf[j];
                const float dj = (float)strain[j] * sc.inv_n0p * sc.dens_scale;
                const int cj = __float_as_int(d.w);
                sur_tile<false, true>(W, sc, lane, valid, d.x, d.y, d.z, ci, cj, di, dj,
                                      (float)(vj.x - vi.x) * sc.inv_u, (float)(vj.y - vi.y) * sc.inv_u,
                                      (float)(vj.z - vi.z) * sc.inv_u, 0, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, ax, ay, az, msum, j, valid && cj < 2);
            };
            for (int e0 = 0; e0 < len; e0 += 64) {
                const int e = e0 + lane;
                const int j = e < len ? list[e] : -1;
                bool ok = false;
                double dxd = 0.0, dyd = 0.0, dzd = 0.0;
                int cj = 0;
                if (j >= 0 && j != i) {
                    const double4 pj = pos[j];
                    cj = sur_class(tag_of(pj.w) & 3);
                    rel3(C, pi, pj, dxd, dyd, dzd);
                    // fluid partners with a lower index evaluate this pair themselves
                    ok = dxd * dxd + dyd * dyd + dzd * dzd <= sc.cut2 && (j > i || cj == 2);
                }
                const unsigned long long mask = __ballot(ok);
                if (ok) {
                    const int at = count + (int)__popcll(mask & below);
                    qj[wv][at] = j;
                    qd[wv][at] = make_float4((float)dxd, (float)dyd, (float)dzd, __int_as_float(cj));
                }
                count += (int)__popcll(mask);
                while (count >= 16) {
                    run_tile(16);
                    // move the rest of the queue down by one tile
                    __builtin_amdgcn_wave_barrier();
                    int mj = 0;
                    float4 md = make_float4(0.f, 0.f, 0.f, 0.f);
                    const bool mv = lane < count - 16;
                    if (mv) {
                        mj = qj[wv][lane + 16];
                        md = qd[wv][lane + 16];
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (mv) {
                        qj[wv][lane] = mj;
                        qd[wv][lane] = md;
                    }
                    count -= 16;
                }
            }
            if (count > 0) run_tile(count);
            if (lane == 0) {
                atomicAdd(&msum[i].x, ax);
                atomicAdd(&msum[i].y, ay);
                atomicAdd(&msum[i].z, az);
            }
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// a_i = a_scale rho_0/rho_i sum_j m_ij by original id (0 for walls), from the sums of
// k_sur_infer; .w = a_scale rho_0/rho_i
__global__ void __launch_bounds__(256)
k_sur_finish(const CaseConst* __restrict__ cases, int n, const double4* __restrict__ pos,
             const int* __restrict__ pidx, const float4* __restrict__ msum, SurScales sc, float4* __restrict__ apred)
{
    const CaseConst& C = cases[0];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ti = tag_of(pos[i].w) & 3;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    if (ti < 2) {
        const float rr = (float)(C.mass[0] / C.mass[ti]) * sc.a_scale;
        const float4 m = msum[i];
        a = make_float4(m.x * rr, m.y * rr, m.z * rr, rr);
    }
    apred[pidx[i]] = a;
}

// loss over the training list and dL/dm per particle (FP32), with the loss sum in FP64
__global__ void __launch_bounds__(256)
k_sur_loss(const int* __restrict__ plist, int nlist, const float4* __restrict__ apred,
           const float4* __restrict__ target, float inv_a, float4* __restrict__ dmk, double* __restrict__ loss)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (k < nlist) {
        const float4 a = apred[k], t = target[k];
        const float dx = (a.x - t.x) * inv_a, dy = (a.y - t.y) * inv_a, dz = (a.z - t.z) * inv_a;
        l = (double)dx * dx + (double)dy * dy + (double)dz * dz;
        // L = (1/N) sum |a - t|^2 / a_s^2 and a = s_i sum_j m_ij with s_i = apred.w:
        // dL/dm = 2 (a - t) s_i / (N a_s^2)
        const float c = 2.0f * a.w * inv_a / (float)nlist;
        dmk[k] = make_float4(c * dx, c * dy, c * dz, 0.f);
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / (double)nlist);
}

// backward data path of one layer, one lane per pair slot (FP32):
//   dz_in[a] = dsilu(Z[a]) * sum_b W[a][b] dz_out[b]
// The first (output) layer takes dz_out = dL/dc, the message coefficients' gradient from
// dL/dm of the slot's particle and the slot's basis (0 for padding slots), and also stores
// it as the output gradient of the last layer ([P][4]).
template <int NOUT, bool FIRST>
__global__ void __launch_bounds__(256)
k_sur_backward(long nslots, const int* __restrict__ tile_owner, const int* __restrict__ toff,
               const int* __restrict__ plist, const int* __restrict__ pcnt, const float4* __restrict__ dmk,
               const float* __restrict__ wr, int off_w, const float* __restrict__ dz_out,
               const float* __restrict__ Z, float* __restrict__ dz_in, float* __restrict__ dm_store,
               const float* __restrict__ Bs)
{
    __shared__ float W[kSurHid * NOUT];
    for (int q = threadIdx.x; q < kSurHid * NOUT; q += blockDim.x) W[q] = wr[off_w + q];
    __syncthreads();
    const long P = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (P >= nslots) return;
    float d[NOUT];
    if (FIRST) {
        // dL/dc of the message coefficients: m = c0 e + c1 u + c2 (u.e) e
        const int T = (int)(P >> 4), e0 = (int)(P & 15);
        const int k = tile_owner[T];
        const bool valid = 16 * (T - toff[k]) + e0 < pcnt[plist[k]];
        const float4 m = dmk[k];
        const float4 b0 = *reinterpret_cast<const float4*>(Bs + P * 8);
        const float4 b1 = *reinterpret_cast<const float4*>(Bs + P * 8 + 4);
        const float ge = fmaf(m.z, b0.z, fmaf(m.y, b0.y, m.x * b0.x));
        const float gu = fmaf(m.z, b1.y, fmaf(m.y, b1.x, m.x * b0.w));
        d[0] = valid ? ge : 0.f;
        d[1] = valid ? gu : 0.f;
        d[2] = valid ? ge * b1.z : 0.f;
        *reinterpret_cast<float4*>(dm_store + P * 4) = make_float4(d[0], d[1], d[2], 0.f);
    } else {
        for (int b = 0; b < NOUT; ++b) d[b] = dz_out[P * NOUT + b];
    }
    for (int a = 0; a < kSurHid; ++a) {
        float s = 0.f;
#pragma unroll
        for (int b = 0; b < NOUT; ++b) s = fmaf(W[a * NOUT + b], d[b], s);
        dz_in[P * kSurHid + a] = s * sur_dsilu(Z[P * kSurHid + a]);
    }
}

// Weight gradient of one layer as a reduction over pair slots on the matrix cores:
//   dW[a][b] += sum_P in[P][a] * dout[P][b],  db[b] += sum_P dout[P][b]
// A = in^T (rows a, K = slots), B = dout (K = slots, columns b); both rounded to bf16.
// IN: 0 = X (bf16 [P][32], 17 used), 1 = silu(Z) of the previous layer (FP32 [P][64]).
// Each wavefront reduces a contiguous range of 32-slot slices, then adds its tiles to the
// FP32 gradient with atomics.  MT/NT: 16-row tiles of a and 16-column tiles of b.
template <int IN, int MT, int NT>
__global__ void __launch_bounds__(256)
k_sur_wgrad(long nslots, int slices_per_wave, const __bf16* __restrict__ X, const float* __restrict__ Zin,
            const float* __restrict__ dout, int ld_out, int nout, int kin, float* __restrict__ grad, int off_w,
            int off_b)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const long wid = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const long nslices = nslots >> 5;
    const long s0 = wid * slices_per_wave;
    const long s1 = s0 + slices_per_wave < nslices ? s0 + slices_per_wave : nslices;
    if (s0 >= s1) return;
    f32x4 acc[MT][NT], accb[NT];
    for (int ta = 0; ta < MT; ++ta)
        for (int tb = 0; tb < NT; ++tb) acc[ta][tb] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int tb = 0; tb < NT; ++tb) accb[tb] = f32x4{0.f, 0.f, 0.f, 0.f};
    bf16x8 ones;
    for (int q = 0; q < 8; ++q) ones[q] = (__bf16)1.0f;
    for (long sl = s0; sl < s1; ++sl) {
        const long P0 = sl * 32 + 8 * g;
        bf16x8 A[MT], B[NT];
        for (int ta = 0; ta < MT; ++ta)
            for (int q = 0; q < 8; ++q) {
                const int a = 16 * ta + r;
                if (IN == 0) A[ta][q] = X[(P0 + q) * 32 + a];
                else A[ta][q] = (__bf16)sur_silu(Zin[(P0 + q) * kSurHid + a]);
            }
        for (int tb = 0; tb < NT; ++tb)
            for (int q = 0; q < 8; ++q) {
                const int b = 16 * tb + r;
                B[tb][q] = (__bf16)(b < nout ? dout[(P0 + q) * ld_out + b] : 0.f);
            }
        for (int ta = 0; ta < MT; ++ta)
            for (int tb = 0; tb < NT; ++tb)
                acc[ta][tb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A[ta], B[tb], acc[ta][tb], 0, 0, 0);
        if (off_b >= 0)
            for (int tb = 0; tb < NT; ++tb)
                accb[tb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ones, B[tb], accb[tb], 0, 0, 0);
    }
    // D layout: lane holds column b = 16 tb + lane%16 and rows a = 16 ta + 4 g + q
    for (int ta = 0; ta < MT; ++ta)
        for (int tb = 0; tb < NT; ++tb) {
            const int b = 16 * tb + r;
            if (b >= nout) continue;
            for (int q = 0; q < 4; ++q) {
                const int a = 16 * ta + 4 * g + q;
                if (a < kin) atomicAdd(grad + off_w + a * nout + b, acc[ta][tb][q]);
            }
        }
    if (off_b >= 0 && g == 0)
        for (int tb = 0; tb < NT; ++tb) {
            const int b = 16 * tb + r;
            if (b < nout) atomicAdd(grad + off_b + b, accb[tb][0]);
        }
}

__global__ void __launch_bounds__(256)
k_sur_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ grad,
           float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kSurNP) return;
    const float gr = grad[p];
    const float mm = b1 * m[p] + (1.f - b1) * gr;
    const float vv = b2 * v[p] + (1.f - b2) * gr * gr;
    m[p] = mm;
    v[p] = vv;
    // c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
    w[p] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

// velocity kick of the surrogate step: v_next = v + dt (a + g) for fluid (acceleration by
// original id), walls keep the velocity the move gave them; end-of-step bookkeeping of the
// rebuild flags as sweep B does it
__global__ void __launch_bounds__(256)
k_sur_kick(const CaseConst* __restrict__ cases, int n, const double4* __restrict__ pos,
           const double4* __restrict__ vel, const int* __restrict__ pidx, const float4* __restrict__ acc,
           double4* __restrict__ vnext, int* __restrict__ flags, int* __restrict__ any_flag,
           long long* __restrict__ rebuilds)
{
    const CaseConst& C = cases[0];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (flags[0]) rebuilds[0] += 1;
        flags[0] = 0;
        *any_flag = 0;
    }
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 v = vel[i];
    double4 vn = make_double4(v.x, v.y, v.z, 0.0);
    if ((tag_of(pos[i].w) & 3) < 2) {
        const float4 a = acc[pidx[i]];
        vn.x += C.dt * ((double)a.x + C.grav[0]);
        vn.y += C.dt * ((double)a.y + C.grav[1]);
        vn.z += C.dt * ((double)a.z + C.grav[2]);
    }
    vnext[i] = vn;
}

// per-particle values into original-id order
__global__ void __launch_bounds__(256)
k_sur_to_orig(int n, const int* __restrict__ pidx, const double4* __restrict__ v, const double* __restrict__ dens,
              double4* __restrict__ v_orig, double* __restrict__ n_orig)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int o = pidx[i];
    v_orig[o] = v[i];
    if (dens) n_orig[o] = dens[i];
}

// relative L2 differences of fluid velocities and densities against stored original-order
// values: sums[0..3] += |dv|^2, |v|^2, dn^2, n^2
__global__ void __launch_bounds__(256)
k_sur_compare(int n, const double4* __restrict__ pos, const int* __restrict__ pidx,
              const double4* __restrict__ v, const double* __restrict__ dens, const double4* __restrict__ v_ref,
              const double* __restrict__ n_ref, double* __restrict__ sums)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (i < n && (tag_of(pos[i].w) & 3) < 2) {
        const int o = pidx[i];
        const double4 a = v[i], b = v_ref[o];
        const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
        s0 = dx * dx + dy * dy + dz * dz;
        s1 = a.x * a.x + a.y * a.y + a.z * a.z;
        const double dn = dens[i] - n_ref[o];
        s2 = dn * dn;
        s3 = dens[i] * dens[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_xor(s0, off);
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
        s3 += __shfl_xor(s3, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(sums + 0, s0);
        atomicAdd(sums + 1, s1);
        atomicAdd(sums + 2, s2);
        atomicAdd(sums + 3, s3);
    }
}

// training targets in snapshot order: a = (v_{n+2} - v_n) / (2 dt) - g
__global__ void __launch_bounds__(256)
k_sur_target(const CaseConst* __restrict__ cases, const int* __restrict__ plist, int nlist,
             const int* __restrict__ spidx, const double4* __restrict__ svel, const double4* __restrict__ v2_orig,
             float4* __restrict__ target)
{
    const CaseConst& C = cases[0];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nlist) return;
    const int i = plist[k];
    const double4 a = svel[i], b = v2_orig[spidx[i]];
    const double s = 0.5 / C.dt;
    target[k] = make_float4((float)((b.x - a.x) * s - C.grav[0]), (float)((b.y - a.y) * s - C.grav[1]),
                            (float)((b.z - a.z) * s - C.grav[2]), 0.f);
}

}  // namespace mph
