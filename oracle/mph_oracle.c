/*
 * CPU oracle for the MPH-WC explicit step -- TEST INFRASTRUCTURE ONLY (see mph_oracle.h).
 *
 * Restates source/main.cpp of the reference in plain C with a runtime dimension flag
 * (the reference selects 2D with the TWO_DIMENSIONAL macro, main.cpp:38).  Each function
 * names the reference lines it follows.  Expressions are written in the reference's
 * operation order (C evaluates * and / left to right), and the file is compiled with
 * -ffp-contract=off, because bit-for-bit agreement with the reference is what pins the
 * oracle.  The unstable bitonic sort of the reference is reproduced step for step, since
 * it fixes the order of particles inside a cell and therefore the summation order.
 */
#include "mph_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double vec3[3];

struct orc_state {
    int dim, n, p2, p2pow;
    double l0, vol, dt, time;
    double dmin[3], dmax[3], dwid[3];
    /* type tables */
    double rho[4], kbulk[4], lbulk[4], mshear[4], sigma[4], cofa[4], ir[4][4];
    double cofk, grav[3];
    double wc[4][3], wv[4][3], wo[4][3], wr[4][3][3];
    /* kernel constants (main.cpp:193-211) */
    double ha, hg, hp, hv, maxr, swa, swg, swp, swv, n0a, n0p, r2g;
    /* cells (main.cpp:102-112) */
    double cw[3];
    int cc[3], tcc;
    int *cbeg_f, *cend_f, *cbeg_w, *cend_w, *ckey, *cpart;
    int fbeg, fend, wbeg, wend;
    /* particles */
    int *pidx, *ptype, *tmpi;
    vec3 *x, *v, *f, *xs, *gc, *tmpv;
    double *mass, *dens_a, *press_a, *vstrain, *div, *press_p, *mu, *lam, *kap, *vir;
    int *ncnt;
    long *nptr, ntotal;
    int *nind;
    int last_rebuild;
};

/* main.cpp:70 */
static inline double mod_w(double x, double w) { return x - w * floor(x / w); }

/* weight functions, main.cpp:214-284 (2D: h^2, 3D: h^3) */
static inline double hd(const orc_state *s, double h) { return s->dim == 2 ? h * h : h * h * h; }
static inline double w_a(const orc_state *s, double r, double h)
{ return 1.0 / s->swa * 1.0 / hd(s, h) * (r / h) * (1.0 - (r / h)) * (1.0 - (r / h)); }
static inline double dw_a(const orc_state *s, double r, double h)
{ return 1.0 / s->swa * 1.0 / hd(s, h) * (1.0 - (r / h)) * (1.0 - 3.0 * (r / h)) * (1.0 / h); }
static inline double w_g(const orc_state *s, double r, double h)
{ return 1.0 / s->swg * 1.0 / hd(s, h) * ((1.0 - r / h) * (1.0 - r / h)); }
static inline double dw_g(const orc_state *s, double r, double h)
{ return 1.0 / s->swg * 1.0 / hd(s, h) * (-2.0 / h * (1.0 - r / h)); }
static inline double w_p(const orc_state *s, double r, double h)
{ return 1.0 / s->swp * 1.0 / hd(s, h) * ((1.0 - r / h) * (1.0 - r / h)); }
static inline double dw_p(const orc_state *s, double r, double h)
{ return 1.0 / s->swp * 1.0 / hd(s, h) * (-2.0 / h * (1.0 - r / h)); }
static inline double dw_v(const orc_state *s, double r, double h)
{ return 1.0 / s->swv * 1.0 / hd(s, h) * (-2.0 / h * (1.0 - r / h)); }

/* periodic relative position x_j - x_i, main.cpp:1680 (same expression in every pass) */
static inline void rel(const orc_state *s, const double *xi, const double *xj, double *o)
{
    for (int d = 0; d < 3; ++d)
        o[d] = mod_w(xj[d] - xi[d] + 0.5 * s->dwid[d], s->dwid[d]) - 0.5 * s->dwid[d];
}

static inline int is_fluid(int t) { return 0 <= t && t < 2; }
static inline int is_wall(int t) { return 2 <= t && t < 4; }
#define MARGIN(s) (0.1 * (s)->l0)

/* ---------------------------------------------------------------- initialisation ---- */

/* lattice sum of a kernel over the neighbours of a lattice site, main.cpp:861-949 */
static double lattice_sum(const orc_state *s, double radius, int kind)
{
    const double ratio = radius / s->l0;
    const int range = (int)(ratio + 3.0);
    double sum = 0.0;
    const int zr = s->dim == 2 ? 0 : range;
    for (int ix = -range; ix <= range; ++ix)
        for (int iy = -range; iy <= range; ++iy)
            for (int iz = -zr; iz <= zr; ++iz) {
                if (ix == 0 && iy == 0 && iz == 0) continue;
                const double x = s->l0 * ((double)ix);
                const double y = s->l0 * ((double)iy);
                const double z = s->l0 * ((double)iz);
                const double r2 = s->dim == 2 ? x * x + y * y : x * x + y * y + z * z;
                if (r2 <= radius * radius) {
                    const double r = sqrt(r2);
                    sum += kind == 0 ? w_a(s, r, radius) : w_p(s, r, radius);
                }
            }
    return sum;
}

static void init_weight(orc_state *s, const mph_params *p)   /* main.cpp:836-954 */
{
    const double l0 = s->l0;
    s->ha = p->radius_ratio_a * l0;
    s->hg = p->radius_ratio_a * l0;
    s->hp = p->radius_ratio_p * l0;
    s->hv = p->radius_ratio_v * l0;
    if (s->dim == 2) {
        s->swa = 1.0 / 2.0 * 2.0 / 15.0 * M_PI / l0 / l0;
        s->swg = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        s->swp = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        s->swv = 1.0 / 2.0 * 1.0 / 3.0 * M_PI / l0 / l0;
        s->r2g = 1.0 / 2.0 * 1.0 / 30.0 * M_PI * s->hg * s->hg / l0 / l0 / s->swg;
    } else {
        s->swa = 1.0 / 3.0 * 1.0 / 5.0 * M_PI / l0 / l0 / l0;
        s->swg = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        s->swp = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        s->swv = 1.0 / 3.0 * 2.0 / 5.0 * M_PI / l0 / l0 / l0;
        s->r2g = 1.0 / 3.0 * 4.0 / 105.0 * M_PI * s->hg * s->hg / l0 / l0 / l0 / s->swg;
    }
    s->n0a = lattice_sum(s, s->ha, 0);
    s->n0p = lattice_sum(s, s->hp, 1);
}

static void init_fluid(orc_state *s)   /* main.cpp:957-990 (per-type part) */
{
    double integ_n, integ_x;
    if (s->dim == 2) { s->cofk = 0.350778153; integ_n = 0.024679383; integ_x = 0.226126699; }
    else { s->cofk = 0.326976006; integ_n = 0.021425779; integ_x = 0.233977488; }
    for (int t = 0; t < 4; ++t)
        s->cofa[t] = s->sigma[t] / ((s->hg / s->l0) * (integ_n + s->cofk * s->cofk * integ_x));
}

static void init_wall(orc_state *s)   /* main.cpp:994-1033 */
{
    for (int w = 2; w < 4; ++w) {
        double nrm[3] = {0.0, 0.0, 0.0}, q[4];
        const double *om = s->wo[w];
        const double theta = fabs(om[0] * om[0] + om[1] * om[1] + om[2] * om[2]);
        if (theta != 0.0)
            for (int d = 0; d < 3; ++d) nrm[d] = om[d] / theta;
        q[0] = nrm[0] * sin(theta * s->dt / 2.0);
        q[1] = nrm[1] * sin(theta * s->dt / 2.0);
        q[2] = nrm[2] * sin(theta * s->dt / 2.0);
        q[3] = cos(theta * s->dt / 2.0);
        double (*R)[3] = s->wr[w];
        R[0][0] = q[0] * q[0] - q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[0][1] = 2.0 * (q[0] * q[1] - q[2] * q[3]);
        R[0][2] = 2.0 * (q[0] * q[2] + q[1] * q[3]);
        R[1][0] = 2.0 * (q[0] * q[1] + q[2] * q[3]);
        R[1][1] = -q[0] * q[0] + q[1] * q[1] - q[2] * q[2] + q[3] * q[3];
        R[1][2] = 2.0 * (q[1] * q[2] - q[0] * q[3]);
        R[2][0] = 2.0 * (q[0] * q[2] - q[1] * q[3]);
        R[2][1] = 2.0 * (q[1] * q[2] + q[0] * q[3]);
        R[2][2] = -q[0] * q[0] - q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
    }
}

static int init_domain(orc_state *s)   /* main.cpp:1035-1084 */
{
    double m = 0.0;
    m = s->ha > m ? s->ha : m;
    m = s->hg > m ? s->hg : m;
    m = s->hp > m ? s->hp : m;
    m = s->hv > m ? s->hv : m;
    s->maxr = m;
    double cnt[3];
    for (int d = 0; d < 3; ++d) s->dwid[d] = s->dmax[d] - s->dmin[d];
    cnt[0] = floor((s->dmax[0] - s->dmin[0]) / (s->maxr + MARGIN(s)));
    cnt[1] = floor((s->dmax[1] - s->dmin[1]) / (s->maxr + MARGIN(s)));
    cnt[2] = s->dim == 2 ? 1 : floor((s->dmax[2] - s->dmin[2]) / (s->maxr + MARGIN(s)));
    for (int d = 0; d < 3; ++d) s->cc[d] = (int)cnt[d];
    if (s->cc[0] < 1 || s->cc[1] < 1 || s->cc[2] < 1) return -1;
    s->tcc = (int)(cnt[0] * cnt[1] * cnt[2]);
    for (int d = 0; d < 3; ++d) s->cw[d] = s->dwid[d] / s->cc[d];
    s->cbeg_f = calloc(s->tcc, sizeof(int));
    s->cend_f = calloc(s->tcc, sizeof(int));
    s->cbeg_w = calloc(s->tcc, sizeof(int));
    s->cend_w = calloc(s->tcc, sizeof(int));
    return 0;
}

/* --------------------------------------------------------------- bucketing --------- */

static inline int wrapi(int i, int c) { return (i % c + c) % c; }
static inline int cell_id(const orc_state *s, int ix, int iy, int iz)   /* main.cpp:96-100 */
{
    if (s->dim == 2) return wrapi(ix, s->cc[0]) * s->cc[1] + wrapi(iy, s->cc[1]);
    return wrapi(ix, s->cc[0]) * s->cc[1] * s->cc[2] + wrapi(iy, s->cc[1]) * s->cc[2] + wrapi(iz, s->cc[2]);
}

static void cell_particle(orc_state *s)   /* main.cpp:1111-1326 */
{
    const int n = s->n, P = s->p2, T = s->tcc;
    for (int i = 0; i < P; ++i) {
        if (i < n) {
            const int ix = ((int)floor((s->x[i][0] - s->dmin[0]) / s->cw[0])) % s->cc[0];
            const int iy = ((int)floor((s->x[i][1] - s->dmin[1]) / s->cw[1])) % s->cc[1];
            const int iz = ((int)floor((s->x[i][2] - s->dmin[2]) / s->cw[2])) % s->cc[2];
            s->ckey[i] = cell_id(s, ix, iy, iz) + (is_wall(s->ptype[i]) ? T : 0);
            s->cpart[i] = i;
        } else {
            s->ckey[i] = 2 * T;
            s->cpart[i] = n;
        }
    }
    /* the reference's bitonic network (main.cpp:1137-1158): not stable, so it is followed
     * exactly to obtain the same order inside each cell */
    for (int stage = 0; stage < s->p2pow; ++stage)
        for (int sub = 0; sub <= stage; ++sub) {
            const int dist = 1 << (stage - sub);
            for (int i = 0; i < P; ++i) {
                if (i & dist) continue;
                const int asc = ((i >> stage) & 2) == 0;
                const int a = s->ckey[i], b = s->ckey[i | dist];
                if ((a > b) == asc) {
                    const int cp = s->cpart[i];
                    s->ckey[i] = b; s->cpart[i] = s->cpart[i | dist];
                    s->ckey[i | dist] = a; s->cpart[i | dist] = cp;
                }
            }
        }
    /* per-cell begin/end: counts then an inclusive prefix over [fluid cells | wall cells]
     * (main.cpp:1160-1250) */
    int *cnt = calloc(2 * (size_t)T, sizeof(int));
    for (int i = 0; i < n; ++i) {
        const int k = s->ckey[i];
        if (k < 2 * T) cnt[k]++;
    }
    long acc = 0;
    for (int k = 0; k < 2 * T; ++k) { acc += cnt[k]; cnt[k] = (int)acc; }
    for (int c = 0; c < T; ++c) {
        s->cbeg_f[c] = c == 0 ? 0 : cnt[c - 1];
        s->cend_f[c] = cnt[c];
        s->cbeg_w[c] = cnt[c - 1 + T];
        s->cend_w[c] = cnt[c + T];
    }
    free(cnt);
    s->fbeg = s->cbeg_f[0];
    s->fend = s->cend_f[T - 1];
    s->wbeg = s->cbeg_w[0];
    s->wend = s->cend_w[T - 1];
    /* gather-reorder of id, type, position, velocity (main.cpp:1262-1325) */
    for (int i = 0; i < n; ++i) s->tmpi[i] = s->pidx[s->cpart[i]];
    memcpy(s->pidx, s->tmpi, sizeof(int) * n);
    for (int i = 0; i < n; ++i) s->tmpi[i] = s->ptype[s->cpart[i]];
    memcpy(s->ptype, s->tmpi, sizeof(int) * n);
    for (int i = 0; i < n; ++i) memcpy(s->tmpv[i], s->x[s->cpart[i]], sizeof(vec3));
    memcpy(s->x, s->tmpv, sizeof(vec3) * n);
    for (int i = 0; i < n; ++i) memcpy(s->tmpv[i], s->v[s->cpart[i]], sizeof(vec3));
    memcpy(s->v, s->tmpv, sizeof(vec3) * n);
}

/* the periodic 3-cell window along one axis, in ascending cell order; the reference's
 * jXmin rotation (main.cpp:1365-1391) yields exactly this sequence, repeats included */
static int stencil_axis(int ic, int range, int c, int *out)
{
    const int m = 2 * range + 1;
    for (int k = 0; k < m; ++k) out[k] = wrapi(ic - range + k, c);
    for (int a = 1; a < m; ++a)            /* insertion sort of <= 3 entries */
        for (int b = a; b > 0 && out[b - 1] > out[b]; --b) { int t = out[b]; out[b] = out[b - 1]; out[b - 1] = t; }
    return m;
}

typedef void (*visit_fn)(orc_state *, int i, int j, void *ctx);

/* walk the list candidates of particle i: fluid cells, then (unless the wall rule of
 * main.cpp:1416 applies) wall cells; `emit` gets each j within the list cutoff */
static int scan_candidates(orc_state *s, int i, int *out, int cap_or_neg)
{
    const int T = s->tcc;
    const int key = s->ckey[i] >= T ? s->ckey[i] - T : s->ckey[i];   /* CellIndex of the sorted slot */
    const int icx = (s->ckey[i] / (s->cc[1] * s->cc[2])) % T;        /* main.cpp:1356-1358 */
    const int icy = (s->ckey[i] / s->cc[2]) % s->cc[1];
    const int icz = s->ckey[i] % s->cc[2];
    (void)key;
    const int rx = (int)ceil((s->maxr + MARGIN(s)) / s->cw[0]);
    const int ry = (int)ceil((s->maxr + MARGIN(s)) / s->cw[1]);
    const int rz = s->dim == 2 ? 0 : (int)ceil((s->maxr + MARGIN(s)) / s->cw[2]);
    int sx[3], sy[3], sz[3];
    const int mx = stencil_axis(icx % s->cc[0], rx, s->cc[0], sx);
    const int my = stencil_axis(icy, ry, s->cc[1], sy);
    const int mz = stencil_axis(icz, rz, s->cc[2], sz);
    const double cut = (s->maxr + MARGIN(s)) * (s->maxr + MARGIN(s));
    int count = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && s->wbeg <= i && i < s->wend && count == 0) break;
        for (int a = 0; a < mx; ++a)
            for (int b = 0; b < my; ++b)
                for (int c = 0; c < mz; ++c) {
                    const int jc = cell_id(s, sx[a], sy[b], sz[c]);
                    const int lo = pass == 0 ? s->cbeg_f[jc] : s->cbeg_w[jc];
                    const int hi = pass == 0 ? s->cend_f[jc] : s->cend_w[jc];
                    for (int j = lo; j < hi; ++j) {
                        double q[3];
                        rel(s, s->x[i], s->x[j], q);
                        const double q2 = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
                        if (q2 <= cut) {
                            if (out) out[count] = j;
                            ++count;
                        }
                    }
                }
    }
    (void)cap_or_neg;
    return count;
}

static void neighbor(orc_state *s)   /* main.cpp:1329-1596 */
{
    const int n = s->n;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) s->ncnt[i] = scan_candidates(s, i, NULL, -1);
    s->nptr[0] = 0;
    for (int i = 0; i < n; ++i) s->nptr[i + 1] = s->nptr[i] + s->ncnt[i];
    s->ntotal = s->nptr[n];
    free(s->nind);
    s->nind = malloc(sizeof(int) * (s->ntotal > 0 ? s->ntotal : 1));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) scan_candidates(s, i, s->nind + s->nptr[i], -1);
    memcpy(s->xs, s->x, sizeof(vec3) * n);
}

/* main.cpp:1086-1108 */
static int need_rebuild(const orc_state *s)
{
    double mx = 0.0;
    for (int i = 0; i < s->n; ++i) {
        double d[3];
        for (int k = 0; k < 3; ++k)
            d[k] = mod_w(s->x[i][k] - s->xs[i][k] + 0.5 * s->dwid[k], s->dwid[k]) - 0.5 * s->dwid[k];
        const double sh = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
        if (sh > mx) mx = sh;
    }
    return mx > 0.5 * MARGIN(s) * 0.5 * MARGIN(s);
}

/* ---------------------------------------------------------------- per-step passes ---- */

#define FOR_NEIGHBORS(s, i, j)                                                          \
    for (long _k = (s)->nptr[i]; _k < (s)->nptr[i] + (s)->ncnt[i]; ++_k)                 \
        for (int j = (s)->nind[_k], _once = 1; _once; _once = 0)                         \
            if (j != (i))

static void density_a(orc_state *s)   /* main.cpp:1665-1692 */
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double sum = 0.0;
        FOR_NEIGHBORS(s, i, j) {
            const double ratio = s->ir[s->ptype[i]][s->ptype[j]];
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->ha * s->ha - r2 >= 0) {
                const double r = sqrt(r2);
                sum += ratio * w_a(s, r, s->ha);
            }
        }
        s->dens_a[i] = sum;
    }
}

static void gravity_center(orc_state *s)   /* main.cpp:1695-1727 */
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double sum[3] = {0.0, 0.0, 0.0};
        FOR_NEIGHBORS(s, i, j) {
            const double ratio = s->ir[s->ptype[i]][s->ptype[j]];
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hg * s->hg - r2 >= 0) {
                const double r = sqrt(r2);
                const double w = ratio * w_g(s, r, s->hg);
                for (int d = 0; d < 3; ++d) sum[d] += q[d] * w / s->r2g * s->hg;
            }
        }
        for (int d = 0; d < 3; ++d) s->gc[i][d] = sum[d];
    }
}

static void density_p(orc_state *s)   /* main.cpp:1831-1858 */
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double sum = 0.0;
        FOR_NEIGHBORS(s, i, j) {
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hp * s->hp - r2 >= 0) {
                const double r = sqrt(r2);
                sum += w_p(s, r, s->hp);
            }
        }
        s->vstrain[i] = (sum - s->n0p);
    }
}

static void divergence_p(orc_state *s)   /* main.cpp:1860-1896 */
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double sum = 0.0;
        FOR_NEIGHBORS(s, i, j) {
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hp * s->hp - r2 >= 0) {
                const double r = sqrt(r2);
                const double dw = dw_p(s, r, s->hp);
                const double e[3] = {q[0] / r, q[1] / r, q[2] / r};
                double u[3];
                for (int d = 0; d < 3; ++d) u[d] = s->v[j][d] - s->v[i][d];
                for (int d = 0; d < 3; ++d) sum -= u[d] * e[d] * dw;
            }
        }
        s->div[i] = sum;
    }
}

static void coefficients(orc_state *s)   /* main.cpp:1631-1662 */
{
    for (int i = 0; i < s->n; ++i) {
        const int t = s->ptype[i];
        s->mass[i] = s->rho[t] * s->vol;
        s->kap[i] = s->kbulk[t];
        if (s->vstrain[i] < 0.0) s->kap[i] = 0.0;
        s->lam[i] = s->lbulk[t];
        s->mu[i] = s->mshear[t];
    }
}

static void pressure_p(orc_state *s)   /* main.cpp:1898-1939 */
{
    for (int i = 0; i < s->n; ++i) s->press_p[i] = -s->lam[i] * s->div[i] + s->kap[i] * s->vstrain[i];
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double fo[3] = {0.0, 0.0, 0.0};
        FOR_NEIGHBORS(s, i, j) {
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hp * s->hp - r2 > 0) {
                const double r = sqrt(r2);
                const double dw = dw_p(s, r, s->hp);
                const double g[3] = {dw * q[0] / r, dw * q[1] / r, dw * q[2] / r};
                for (int d = 0; d < 3; ++d) fo[d] += (s->press_p[i] + s->press_p[j]) * g[d] * s->vol;
            }
        }
        for (int d = 0; d < 3; ++d) s->f[i][d] += fo[d];
    }
}

static void pressure_a(orc_state *s)   /* main.cpp:1729-1776 */
{
    for (int i = 0; i < s->n; ++i) {
        s->press_a[i] = s->cofa[s->ptype[i]] * (s->dens_a[i] - s->n0a) / s->l0;
        if (s->n0a <= s->dens_a[i]) s->press_a[i] = 0.0;
    }
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double fo[3] = {0.0, 0.0, 0.0};
        FOR_NEIGHBORS(s, i, j) {
            const double rij_ = s->ir[s->ptype[i]][s->ptype[j]];
            const double rji_ = s->ir[s->ptype[j]][s->ptype[i]];
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->ha * s->ha - r2 > 0) {
                const double r = sqrt(r2);
                const double dwij = rij_ * dw_a(s, r, s->ha);
                const double dwji = rji_ * dw_a(s, r, s->ha);
                const double e[3] = {q[0] / r, q[1] / r, q[2] / r};
                for (int d = 0; d < 3; ++d)
                    fo[d] += (s->press_a[i] * dwij + s->press_a[j] * dwji) * e[d] * s->vol;
            }
        }
        for (int d = 0; d < 3; ++d) s->f[i][d] += fo[d];
    }
}

static void diffuse_interface(orc_state *s)   /* main.cpp:1778-1829 */
{
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        const double ai = s->cofa[s->ptype[i]] * (s->cofk) * (s->cofk);
        double fo[3] = {0.0, 0.0, 0.0};
        FOR_NEIGHBORS(s, i, j) {
            const double aj = s->cofa[s->ptype[i]] * (s->cofk) * (s->cofk);   /* type of i, as in main.cpp:1792 */
            const double rij_ = s->ir[s->ptype[i]][s->ptype[j]];
            const double rji_ = s->ir[s->ptype[j]][s->ptype[i]];
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hg * s->hg - r2 > 0) {
                const double r = sqrt(r2);
                const double wij = rij_ * w_g(s, r, s->hg);
                const double wji = rji_ * w_g(s, r, s->hg);
                for (int d = 0; d < 3; ++d)
                    fo[d] -= (aj * s->gc[j][d] * wji - ai * s->gc[i][d] * wij) / s->r2g * s->hg * (s->vol / s->l0);
                const double dwij = rij_ * dw_g(s, r, s->hg);
                const double dwji = rji_ * dw_g(s, r, s->hg);
                const double e[3] = {q[0] / r, q[1] / r, q[2] / r};
                double gr = 0.0;
                for (int d = 0; d < 3; ++d)
                    gr += (aj * s->gc[j][d] * dwji - ai * s->gc[i][d] * dwij) * q[d];
                for (int d = 0; d < 3; ++d)
                    fo[d] -= (gr) * e[d] / s->r2g * s->hg * (s->vol / s->l0);
            }
        }
        for (int d = 0; d < 3; ++d) s->f[i][d] += fo[d];
    }
}

static void viscosity_v(orc_state *s)   /* main.cpp:1941-1985 */
{
    const double cd = s->dim == 2 ? 8.0 : 10.0;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double fo[3] = {0.0, 0.0, 0.0};
        FOR_NEIGHBORS(s, i, j) {
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (s->hv * s->hv - r2 > 0) {
                const double r = sqrt(r2);
                const double dwij = -dw_v(s, r, s->hv);
                const double e[3] = {q[0] / r, q[1] / r, q[2] / r};
                double u[3];
                for (int d = 0; d < 3; ++d) u[d] = s->v[j][d] - s->v[i][d];
                const double muij = 2.0 * (s->mu[i] * s->mu[j]) / (s->mu[i] + s->mu[j]);
                for (int d = 0; d < 3; ++d)
                    fo[d] += cd * muij * (u[0] * e[0] + u[1] * e[1] + u[2] * e[2]) * e[d] * dwij / r * s->vol;
            }
        }
        for (int d = 0; d < 3; ++d) s->f[i][d] += fo[d];
    }
}

/* ------------------------------------------------------------- integration & walls --- */

static void convection(orc_state *s)   /* main.cpp:1605-1615 */
{
    for (int i = s->fbeg; i < s->fend; ++i)
        for (int d = 0; d < 3; ++d) s->x[i][d] += s->v[i][d] * s->dt;
}

static void walls(orc_state *s)   /* main.cpp:2013-2053 */
{
    for (int i = s->wbeg; i < s->wend; ++i) {
        const int w = s->ptype[i];
        const double *c = s->wc[w], *om = s->wo[w], *vel = s->wv[w];
        const double (*R)[3] = (const double (*)[3])s->wr[w];
        double r[3] = {s->x[i][0] - c[0], s->x[i][1] - c[1], s->x[i][2] - c[2]};
        /* components are overwritten in place, in this order, as in main.cpp:2033-2035 */
        r[0] = R[0][0] * r[0] + R[0][1] * r[1] + R[0][2] * r[2];
        r[1] = R[1][0] * r[0] + R[1][1] * r[1] + R[1][2] * r[2];
        r[2] = R[2][0] * r[0] + R[2][1] * r[1] + R[2][2] * r[2];
        s->v[i][0] = om[1] * r[2] - om[2] * r[1] + vel[0];
        s->v[i][1] = om[2] * r[0] - om[0] * r[2] + vel[1];
        s->v[i][2] = om[0] * r[1] - om[1] * r[0] + vel[2];
        s->x[i][0] = r[0] + c[0] + vel[0] * s->dt;
        s->x[i][1] = r[1] + c[1] + vel[1] * s->dt;
        s->x[i][2] = r[2] + c[2] + vel[2] * s->dt;
    }
    for (int w = 2; w < 4; ++w)
        for (int d = 0; d < 3; ++d) s->wc[w][d] += s->wv[w][d] * s->dt;
}

static void periodic(orc_state *s)   /* main.cpp:2306-2317 */
{
    for (int i = 0; i < s->n; ++i)
        for (int d = 0; d < 3; ++d) s->x[i][d] = mod_w(s->x[i][d] - s->dmin[d], s->dwid[d]) + s->dmin[d];
}

static void gravity_accel(orc_state *s)   /* main.cpp:1988-2010 */
{
    for (int i = s->fbeg; i < s->fend; ++i)
        for (int d = 0; d < 3; ++d) s->f[i][d] += s->mass[i] * s->grav[d];
    for (int i = s->fbeg; i < s->fend; ++i)
        for (int d = 0; d < 3; ++d) s->v[i][d] += s->f[i][d] / s->mass[i] * s->dt;
}

/* ------------------------------------------------------------------ public API ------- */

orc_state *orc_create(const mph_params *p, int n, const int *type, const double *x,
                      const double *v, int nthreads)
{
#ifdef _OPENMP
    omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#else
    (void)nthreads;
#endif
    if (n <= 0 || (p->dim != 2 && p->dim != 3)) return NULL;
    orc_state *s = calloc(1, sizeof(orc_state));
    s->dim = p->dim; s->n = n; s->l0 = p->particle_spacing; s->dt = p->dt; s->time = p->time;
    s->vol = s->dim == 2 ? s->l0 * s->l0 : s->l0 * s->l0 * s->l0;   /* main.cpp:558-562 */
    for (int d = 0; d < 3; ++d) { s->dmin[d] = p->domain_min[d]; s->dmax[d] = p->domain_max[d]; s->grav[d] = p->gravity[d]; }
    for (int t = 0; t < 4; ++t) {
        s->rho[t] = p->density[t]; s->kbulk[t] = p->bulk_modulus[t]; s->lbulk[t] = p->bulk_viscosity[t];
        s->mshear[t] = p->shear_viscosity[t]; s->sigma[t] = p->surface_tension[t];
        for (int u = 0; u < 4; ++u) s->ir[t][u] = p->interaction_ratio[t][u];
        for (int d = 0; d < 3; ++d) { s->wc[t][d] = p->wall_center[t][d]; s->wv[t][d] = p->wall_velocity[t][d]; s->wo[t][d] = p->wall_omega[t][d]; }
    }
    /* main.cpp:612-617: smallest power of two strictly above n */
    s->p2pow = 0;
    while ((n >> s->p2pow) != 0) ++s->p2pow;
    s->p2 = 1 << s->p2pow;
    s->pidx = malloc(sizeof(int) * n); s->ptype = malloc(sizeof(int) * n); s->tmpi = malloc(sizeof(int) * n);
    s->x = malloc(sizeof(vec3) * n); s->v = malloc(sizeof(vec3) * n); s->f = calloc(n, sizeof(vec3));
    s->xs = malloc(sizeof(vec3) * n); s->gc = calloc(n, sizeof(vec3)); s->tmpv = malloc(sizeof(vec3) * n);
    s->mass = calloc(n, 8); s->dens_a = calloc(n, 8); s->press_a = calloc(n, 8); s->vstrain = calloc(n, 8);
    s->div = calloc(n, 8); s->press_p = calloc(n, 8); s->mu = calloc(n, 8); s->lam = calloc(n, 8);
    s->kap = calloc(n, 8); s->vir = calloc(n, 8);
    s->ncnt = calloc(n, sizeof(int)); s->nptr = calloc(n + 1, sizeof(long));
    s->ckey = malloc(sizeof(int) * s->p2); s->cpart = malloc(sizeof(int) * s->p2);
    for (int i = 0; i < n; ++i) {
        s->pidx[i] = i; s->ptype[i] = type[i];
        for (int d = 0; d < 3; ++d) { s->x[i][d] = x[3 * i + d]; s->v[i][d] = v[3 * i + d]; }
    }
    init_weight(s, p);
    init_fluid(s);
    init_wall(s);
    if (init_domain(s) != 0) { orc_destroy(s); return NULL; }
    for (int i = 0; i < n; ++i) s->mass[i] = s->rho[s->ptype[i]] * s->vol;
    /* main.cpp:358-364 */
    cell_particle(s);
    neighbor(s);
    density_a(s);
    gravity_center(s);
    density_p(s);
    s->last_rebuild = 1;
    return s;
}

int orc_step(orc_state *s, int nsteps)   /* main.cpp:386-431, 449 */
{
    int rebuilds = 0;
    for (int k = 0; k < nsteps; ++k) {
        convection(s);
        walls(s);
        periodic(s);
        memset(s->f, 0, sizeof(vec3) * s->n);
        s->last_rebuild = 0;
        if (need_rebuild(s)) {
            cell_particle(s);
            neighbor(s);
            s->last_rebuild = 1;
            ++rebuilds;
        }
        density_a(s);
        gravity_center(s);
        density_p(s);
        divergence_p(s);
        coefficients(s);
        pressure_p(s);
        pressure_a(s);
        diffuse_interface(s);
        viscosity_v(s);
        gravity_accel(s);
        s->time += s->dt;
    }
    return rebuilds;
}

int orc_last_rebuild(const orc_state *s) { return s->last_rebuild; }
int orc_count(const orc_state *s) { return s->n; }
double orc_time(const orc_state *s) { return s->time; }
long orc_neighbor_total(const orc_state *s) { return s->ntotal; }

void orc_scalars(const orc_state *s, double *o)
{
    o[0] = s->n0a; o[1] = s->n0p; o[2] = s->r2g; o[3] = s->maxr;
    o[4] = s->cw[0]; o[5] = s->cw[1]; o[6] = s->cw[2];
    o[7] = s->cc[0]; o[8] = s->cc[1]; o[9] = s->cc[2];
    o[10] = s->swa; o[11] = s->swg; o[12] = s->cofk; o[13] = s->cofa[0];
    o[14] = s->fend; o[15] = s->wend;
}

void orc_get(const orc_state *s, int *index, int *type, double *x, double *v, double *force,
             double *dens_a, double *vol_strain, double *divergence, double *press_p,
             double *press_a, double *gcenter, int *ncount, double *mass)
{
    const int n = s->n;
    if (index) memcpy(index, s->pidx, sizeof(int) * n);
    if (type) memcpy(type, s->ptype, sizeof(int) * n);
    if (x) memcpy(x, s->x, sizeof(vec3) * n);
    if (v) memcpy(v, s->v, sizeof(vec3) * n);
    if (force) memcpy(force, s->f, sizeof(vec3) * n);
    if (dens_a) memcpy(dens_a, s->dens_a, 8 * n);
    if (vol_strain) memcpy(vol_strain, s->vstrain, 8 * n);
    if (divergence) memcpy(divergence, s->div, 8 * n);
    if (press_p) memcpy(press_p, s->press_p, 8 * n);
    if (press_a) memcpy(press_a, s->press_a, 8 * n);
    if (gcenter) memcpy(gcenter, s->gc, sizeof(vec3) * n);
    if (ncount) memcpy(ncount, s->ncnt, sizeof(int) * n);
    if (mass) memcpy(mass, s->mass, 8 * n);
}

void orc_neighbors(const orc_state *s, long *ptr, int *ind)
{
    for (int i = 0; i <= s->n; ++i) ptr[i] = s->nptr[i];
    memcpy(ind, s->nind, sizeof(int) * s->ntotal);
}

/* calculateVirialStressAtParticle, main.cpp:2057-2302 */
/* Not part of the restatement: VolStrainP of main.cpp:1831-1858 over the same lists and
 * cut-off decisions, evaluated exactly up to about 1e-18: the pair differences as the
 * exact minimum image x_j - x_i (the reference's Mod(x_j - x_i + W/2, W) - W/2 rounds in
 * the +-W/2 round trip), the weights and their sum in long double (x87, 64-bit
 * significand).  The tests use it to measure how far the FP64 VolStrainP of the reference
 * and of the GPU each are from the exact value for the same positions: as a difference of
 * sums near N0p it makes the pressure sensitive to every rounding of the pair geometry and
 * of the summation order. */
void orc_vol_strain_exact(const orc_state *s, double *out)
{
    /* the weight's constant factor as the reference forms it in double (main.cpp:214-284) */
    const long double c = (long double)(1.0 / s->swp * 1.0 / hd(s, s->hp));
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        long double sum = 0.0L;
        FOR_NEIGHBORS(s, i, j) {
            double q[3];
            rel(s, s->x[i], s->x[j], q);
            const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
            if (i != j && s->hp * s->hp - r2 >= 0) {
                long double e2 = 0.0L;
                for (int d = 0; d < 3; ++d) {
                    const long double w = (long double)s->dwid[d];
                    long double a = (long double)s->x[j][d] - (long double)s->x[i][d];
                    if (a > 0.5L * w) a -= w;
                    else if (a < -0.5L * w) a += w;
                    e2 += a * a;
                }
                const long double om = 1.0L - sqrtl(e2) / (long double)s->hp;
                sum += c * (om * om);
            }
        }
        out[i] = (double)(sum - (long double)s->n0p);
    }
}

void orc_virial(orc_state *s, double *vp)
{
    const double cd = s->dim == 2 ? 8.0 : 10.0;
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < s->n; ++i) {
        double st[3][3] = {{0}};
        double acc[3][3];
        /* four passes, each summed separately and then added (main.cpp:2076-2288) */
        for (int pass = 0; pass < 4; ++pass) {
            memset(acc, 0, sizeof(acc));
            FOR_NEIGHBORS(s, i, j) {
                double q[3];
                rel(s, s->x[i], s->x[j], q);
                const double r2 = (q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                if (pass == 0 && s->hp * s->hp - r2 > 0) {
                    const double r = sqrt(r2);
                    const double dw = dw_p(s, r, s->hp);
                    const double g[3] = {dw * q[0] / r, dw * q[1] / r, dw * q[2] / r};
                    double fij[3];
                    for (int d = 0; d < 3; ++d) fij[d] = (s->press_p[i]) * g[d] * s->vol;
                    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) acc[a][b] += 1.0 * fij[a] * q[b] / s->vol;
                }
                if (pass == 1 && s->ha * s->ha - r2 > 0) {
                    const double ratio = s->ir[s->ptype[i]][s->ptype[j]];
                    const double r = sqrt(r2);
                    const double dw = ratio * dw_a(s, r, s->ha);
                    const double g[3] = {dw * q[0] / r, dw * q[1] / r, dw * q[2] / r};
                    double fij[3];
                    for (int d = 0; d < 3; ++d) fij[d] = (s->press_a[i]) * g[d] * s->vol;
                    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) acc[a][b] += 1.0 * fij[a] * q[b] / s->vol;
                }
                if (pass == 2 && s->hv * s->hv - r2 > 0) {
                    const double r = sqrt(r2);
                    const double dw = -dw_v(s, r, s->hv);
                    const double e[3] = {q[0] / r, q[1] / r, q[2] / r};
                    const double u[3] = {s->v[j][0] - s->v[i][0], s->v[j][1] - s->v[i][1], s->v[j][2] - s->v[i][2]};
                    const double muij = 2.0 * (s->mu[i] * s->mu[j]) / (s->mu[i] + s->mu[j]);
                    double fij[3];
                    for (int d = 0; d < 3; ++d)
                        fij[d] = cd * muij * (u[0] * e[0] + u[1] * e[1] + u[2] * e[2]) * e[d] * dw / r * s->vol;
                    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) acc[a][b] += 0.5 * fij[a] * q[b] / s->vol;
                }
                if (pass == 3) {
                    if (s->hg * s->hg - r2 > 0) {
                        const double a_ = s->cofa[s->ptype[i]] * (s->cofk) * (s->cofk);
                        const double ratio = s->ir[s->ptype[i]][s->ptype[j]];
                        const double r = sqrt(r2);
                        const double w = ratio * w_g(s, r, s->hg);
                        double fij[3];
                        for (int d = 0; d < 3; ++d)
                            fij[d] = -a_ * (-s->gc[i][d]) * w / s->r2g * s->hg * (s->vol / s->l0);
                        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) acc[a][b] += 1.0 * fij[a] * q[b] / s->vol;
                    }
                    if (s->hg * s->hg - r2 > 0.0) {
                        const double a_ = s->cofa[s->ptype[i]] * (s->cofk) * (s->cofk);
                        const double ratio = s->ir[s->ptype[i]][s->ptype[j]];
                        const double r = sqrt(r2);
                        const double dw = ratio * dw_g(s, r, s->hg);
                        const double g[3] = {dw * q[0] / r, dw * q[1] / r, dw * q[2] / r};
                        double gr = 0.0;
                        for (int d = 0; d < 3; ++d) gr += (-s->gc[i][d]) * q[d];
                        double fij[3];
                        for (int d = 0; d < 3; ++d) fij[d] = -a_ * (gr) * g[d] / s->r2g * s->hg * (s->vol / s->l0);
                        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) acc[a][b] += 1.0 * fij[a] * q[b] / s->vol;
                    }
                }
            }
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) st[a][b] += acc[a][b];
        }
        s->vir[i] = s->dim == 2 ? -1.0 / 2.0 * (st[0][0] + st[1][1])
                                : -1.0 / 3.0 * (st[0][0] + st[1][1] + st[2][2]);
    }
    memcpy(vp, s->vir, 8 * s->n);
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    free(s->pidx); free(s->ptype); free(s->tmpi); free(s->x); free(s->v); free(s->f); free(s->xs);
    free(s->gc); free(s->tmpv); free(s->mass); free(s->dens_a); free(s->press_a); free(s->vstrain);
    free(s->div); free(s->press_p); free(s->mu); free(s->lam); free(s->kap); free(s->vir);
    free(s->ncnt); free(s->nptr); free(s->nind); free(s->ckey); free(s->cpart);
    free(s->cbeg_f); free(s->cend_f); free(s->cbeg_w); free(s->cend_w);
    free(s);
}
