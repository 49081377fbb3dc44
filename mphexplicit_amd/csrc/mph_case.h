// Per-case constants shared by host set-up and the gfx950 kernels.
#pragma once

#include <cstdint>

namespace mph {

constexpr int kTypes = 4;

// Everything a kernel needs to know about one case, derived on the host from mph_params
// with the reference's own arithmetic (main.cpp:836-1084).  Kernels read it through
// uniform (scalar) loads: every workgroup belongs to exactly one case.
struct CaseConst {
    // particle range of this case in the global arrays [begin, end); fluid first
    int begin, end, nfluid, dim;
    // cell grid (main.cpp:1035-1068); keys of this case are key_base + [0, 2*tcc)
    int cc[3], tcc, key_base;
    int surface;            // any surface tension -> densities A/G and forces A/S needed
    int ratio_one;          // every InteractionRatio == 1
    int rank_base;          // offset of this case in the cell rank tables
    int mu_uniform;         // every type has the same shear viscosity
    int cl_base, cl_cap;    // this case's cluster slots [cl_base, cl_base + cl_cap)
    int img_axes;           // bit d: images near a cluster's first member are exact on axis d
    int list_plain;         // the list build may decide pairs away from the faces by plain differences
    int list_img;           // ... and near the faces by plain differences of images (on img_axes)
    double dmin[3], dmax[3], dwid[3], hwid[3];   // domain min, max, width, 0.5*width
    double guard[3];        // distance from a face beyond which no list partner wraps
    double cw[3];                       // cell widths
    double cut2;            // (MaxRadius + MARGIN)^2, list cutoff (main.cpp:1408)
    double hmax2;           // MaxRadius^2: pairs closer than this interact (pair counter)
    double shift2;          // 0.5*MARGIN*0.5*MARGIN, Verlet trigger (main.cpp:1102)
    double dt, vol, l0;
    double ha2, hp2, hv2;   // squared radii (G uses h_A, main.cpp:838)
    double inv_ha, inv_hp, inv_hv;
    // kernel prefactors 1/Sw * 1/h^d and derivative prefactors (main.cpp:214-284)
    double ca, dca;         // w_A = ca s (1-s)^2,        dw_A = dca (1-s)(1-3s)
    double cg, dcg;         // w_G = cg (1-s)^2,          dw_G = dcg (1-s)
    double cp, dcp;         // w_P, dw_P likewise
    double dcv;             // dw_V = dcv (1-s)
    double n0a, n0p;
    double g_scale;         // h_G / R2g          (gravity centre, main.cpp:1718)
    double s_scale;         // h_G / R2g * V0/l0  (diffuse interface, main.cpp:1808)
    double visc_c;          // 10 (3D) or 8 (2D), main.cpp:1973-1975
    double grav[3];
    double mass[kTypes], kbulk[kTypes], lbulk[kTypes], cofa[kTypes], adiff[kTypes];
    double mu[kTypes];
    double muij[kTypes][kTypes];        // 2 mu_i mu_j / (mu_i + mu_j) (main.cpp:1968)
    double ir[kTypes][kTypes];
    double wrot[kTypes][3][3], womega[kTypes][3], wvel[kTypes][3];
};

}  // namespace mph
