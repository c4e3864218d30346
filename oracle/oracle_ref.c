/*
 * oracle_ref.c -- FP64 restatement of the reference's cpumd path (walled box,
 * N/O/Ne species).  TEST INFRASTRUCTURE ONLY: the checker for the HIP engine;
 * never linked into the product.
 *
 * Every parameter below is a float literal in common/Constants.h, so the FP64
 * build of the reference (cpumd is compiled with EQUI, real = double,
 * cpu/cpumd/cpumd.vcproj:46) sees the float-rounded value widened to double.
 * The same widening is done here by writing the constants as float literals.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------- parameter table (common/Constants.h:36-79) ---------------- */
#define P_RE_N 1.097f
#define P_RE_O 1.2075f
#define P_RE_NO 1.15f
#define P_R0V_N 3.72f
#define P_R0V_O 3.31f
#define P_R0V_NO 3.51f
#define P_SIG_NG 2.72f
#define P_SIG_NGN 3.015f
#define P_SIG_NGO 2.835f
#define P_EPS_NG 0.391e-4f
#define P_EPS_NGN 0.328e-4f
#define P_EPS_NGO 0.447e-4f
#define P_DE_N 941.4e-4f
#define P_DE_O 493.55309e-4f
#define P_DE_NO 626.734e-4f
#define P_BETA_N 2.71009f
#define P_BETA_O 2.2478f
#define P_BETA_NO 2.24068f
#define P_M_N 14.00674f
#define P_M_O 15.9994f
#define P_M_NG 20.1797f
/* bond-order pair cutoff: 15 * sigma_vdW(N), evaluated in float (Constants.h:60-61) */
#define P_CUT_RIJ (15 * 3.31f)
#define P_CUT_RIJSQ (225 * 3.31f * 3.31f)
/* Gear corrector coefficients 19/120, 3/4, 1/2, 1/12 (Constants.h:64-67) */
#define P_GEAR1 0.15833333333333333333333333333333f
#define P_GEAR2 0.75f
#define P_GEAR3 0.5f
#define P_GEAR4 0.08333333333333333333333333333333f
/* Morse/vdW/bond-order shape constants (Constants.h:69-79) */
#define P_S 2.0f
#define P_GM 0.3f
#define P_GMSQ 0.09f
#define P_ALFAV 3.0f
#define P_ALFAVSQ 9.0f
#define P_ALFA 2.0f
#define P_R0 1.3f
#define P_LAMDA3 1.33f
#define P_C 25.f
#define P_D 4.0417f
#define P_H 0.0f

void orc_free(void* p) { free(p); }

/* common/Getters.cpp:13-78: mass by species, pair constants by (ti,tj) where
 * equal species pick the N or O value and unlike ones the NO value. */
double orc_ref_param(int32_t which, int32_t ti, int32_t tj) {
    int same = (ti == tj), isn = (ti == ORC_N);
    switch (which) {
    case 0: return ti == ORC_N ? (double)P_M_N : ti == ORC_O ? (double)P_M_O : (double)P_M_NG;
    case 1: return same ? (isn ? (double)P_RE_N : (double)P_RE_O) : (double)P_RE_NO;
    case 2: return same ? (isn ? (double)P_DE_N : (double)P_DE_O) : (double)P_DE_NO;
    case 3: return same ? (isn ? (double)P_BETA_N : (double)P_BETA_O) : (double)P_BETA_NO;
    case 4: return same ? (isn ? (double)P_R0V_N : (double)P_R0V_O) : (double)P_R0V_NO;
    }
    return 0.0;
}

static double mass_of(double w) { return orc_ref_param(0, (int32_t)w, 0); }

/* ---------------- arvdwfc (common/arvdwfc.cpp:16-59) ---------------- */
/* V_A = S De/(S-1) exp(-beta sqrt(2/S) (r-re)); dva = -dV_A/dr */
void orc_ref_attractive(double r, double de, double beta, double re, double* va, double* dva) {
    double k = beta * sqrt((double)(2.0f / P_S));
    *va = (double)P_S * de / (double)(P_S - 1.0f) * exp(-k * (r - re));
    *dva = k * *va;
}
/* V_R = De/(S-1) exp(-beta sqrt(2S)(r-re)) + 7e-8 exp(-10 beta (r-re)); the
 * reference's "derivative" scales the whole sum by beta sqrt(2S) */
void orc_ref_repulsive(double r, double de, double beta, double re, double* vr, double* dvr) {
    double k = beta * sqrt((double)(P_S * 2.0f));
    *vr = de / (double)(P_S - 1.0f) * exp(-k * (r - re)) + (double)7.0e-8f * exp(-10.0f * beta * (r - re));
    *dvr = k * *vr;
}
/* vdW = 1e-4 (P0 + 10(GM-1) a'(r)) / a(r)^2, a = e^{av(r-r0)} + GM e^{-av(r-r0)};
 * dvdw = -d(vdW)/dr.  P0 = 1 + 50 GM + GM^2 is evaluated in float there. */
void orc_ref_vdw(double r, double r0v, double* vdw, double* dvdw) {
    const float p0 = 1.0f + 50.0f * P_GM + P_GMSQ;
    const float k10 = 10.0f * (P_GM - 1.0f);
    double x = (double)P_ALFAV * (r - r0v);
    double ep = exp(x), em = (double)P_GM * exp(-x);
    double a = ep + em;             /* a(r) */
    double ap = (double)P_ALFAV * (ep - em);  /* a'(r) */
    double app = (double)P_ALFAVSQ * a;       /* a''(r) */
    double num = (double)p0 + (double)k10 * ap;
    *dvdw = -1.0e-4f * (((double)k10 * app * a - 2.0 * ap * num) / (a * a * a));
    *vdw = 1.0e-4f * num / (a * a);
}
/* f_c = (1 - tanh(ALFA (r - R0)))/2, dfc = d f_c / dr */
void orc_ref_cutoff(double r, double* fc, double* dfc) {
    double t = tanh((double)P_ALFA * (r - (double)P_R0));
    *dfc = -0.5f * (double)P_ALFA * (1.0 - t * t);
    *fc = 0.5f * (1.0 - t);
}

/* ---------------- hashed 27-cell candidate list (common/NeighborList.cpp) ---------------- */
typedef struct { int32_t cx, cy, cz, head; } cell_slot;

static uint32_t cell_hash(int32_t cx, int32_t cy, int32_t cz) {
    uint64_t h = (uint64_t)(uint32_t)cx * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)(uint32_t)cy * 0xC2B2AE3D27D4EB4Full + (h >> 29);
    h ^= (uint64_t)(uint32_t)cz * 0x165667B19E3779F9ull + (h >> 31);
    return (uint32_t)(h ^ (h >> 32));
}

/* open addressing; returns slot index, inserting if asked */
static int32_t cell_find(cell_slot* tab, int32_t cap, int32_t cx, int32_t cy, int32_t cz, int insert) {
    uint32_t s = cell_hash(cx, cy, cz) % (uint32_t)cap;
    for (;;) {
        cell_slot* e = &tab[s];
        if (e->head == -2) {
            if (!insert) return -1;
            e->cx = cx; e->cy = cy; e->cz = cz; e->head = -1;
            return (int32_t)s;
        }
        if (e->cx == cx && e->cy == cy && e->cz == cz) return (int32_t)s;
        s = (s + 1) % (uint32_t)cap;
    }
}

/*
 * Restates common/NeighborList.cpp:14-156: the cell edge is rs, the number of
 * cells per axis is (int)(box/rs + 0.99) (at least 1), an atom's cell is
 * (int)((x + box/2)/rs) per axis, cells at index < 0 or >= cells are skipped
 * when gathering, the row of atom i lists the members of the 27 cells around
 * its own in (dx,dy,dz) lexicographic order (dz fastest) with members in
 * ascending index and i itself left out, padded with -1 to
 * nlistsize = min(27 * max occupancy, n - 1).  The bucket layout of the
 * reference's hash does not affect rows, so a different hash is used.
 */
int32_t* orc_ref_build_neighbor_list(const orc_real4* pos, orc_list_settings* ls,
                                     double box_size, int32_t buckets, int32_t n) {
    (void)buckets;
    double rs = (double)ls->rs;
    int32_t cells = (int32_t)(box_size / rs + 0.99f);
    if (cells < 1) cells = 1;
    double npos = box_size / 2;
    int32_t cap = 3 * (n > 0 ? n : 1) + 7;
    cell_slot* tab = (cell_slot*)malloc(sizeof(cell_slot) * (size_t)cap);
    int32_t* next = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* tail = (int32_t*)malloc(sizeof(int32_t) * (size_t)cap);
    int32_t* count = (int32_t*)calloc((size_t)cap, sizeof(int32_t));
    int32_t* cc = (int32_t*)malloc(sizeof(int32_t) * 3 * (size_t)(n > 0 ? n : 1));
    if (!tab || !next || !tail || !count || !cc) return NULL;
    for (int32_t s = 0; s < cap; s++) tab[s].head = -2;
    int32_t largest = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t cx = (int32_t)((pos[i].x + npos) / rs);
        int32_t cy = (int32_t)((pos[i].y + npos) / rs);
        int32_t cz = (int32_t)((pos[i].z + npos) / rs);
        cc[3 * i] = cx; cc[3 * i + 1] = cy; cc[3 * i + 2] = cz;
        int32_t s = cell_find(tab, cap, cx, cy, cz, 1);
        next[i] = -1;
        if (tab[s].head < 0) tab[s].head = i; else next[tail[s]] = i;  /* ascending order */
        tail[s] = i;
        if (++count[s] > largest) largest = count[s];
    }
    int32_t nls = 27 * largest;
    if (nls > n - 1) nls = n - 1;
    ls->nlistsize = nls;
    int32_t* out = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nls * (size_t)n + 1));
    if (!out) return NULL;
    int32_t longest = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t* row = out + (size_t)nls * (size_t)i;
        int32_t k = 0;
        for (int32_t dx = -1; dx <= 1; dx++)
            for (int32_t dy = -1; dy <= 1; dy++)
                for (int32_t dz = -1; dz <= 1; dz++) {
                    int32_t x = cc[3 * i] + dx, y = cc[3 * i + 1] + dy, z = cc[3 * i + 2] + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= cells || y >= cells || z >= cells) continue;
                    int32_t s = cell_find(tab, cap, x, y, z, 0);
                    if (s < 0) continue;
                    for (int32_t j = tab[s].head; j >= 0; j = next[j])
                        if (j != i) row[k++] = j;
                }
        if (k > longest) longest = k;
        for (; k < nls; k++) row[k] = -1;
    }
    ls->nlargestsize = longest;
    free(tab); free(next); free(tail); free(count); free(cc);
    return out;
}

/* ---------------- list schedule (cpu/cpumd/Host.cpp:245-291) ---------------- */
void orc_ref_ready_list(const orc_real4* pos, const orc_real3* vel, int32_t n,
                        double* vmax, double* box_size_list) {
    double vm = 0, bs = 0;
    for (int32_t i = 0; i < n; i++) {
        double s = sqrt(vel[i].x * vel[i].x + vel[i].y * vel[i].y + vel[i].z * vel[i].z);
        if (s > vm) vm = s;
        if (fabs(pos[i].x) > bs) bs = fabs(pos[i].x);
        if (fabs(pos[i].y) > bs) bs = fabs(pos[i].y);
        if (fabs(pos[i].z) > bs) bs = fabs(pos[i].z);
    }
    *vmax = vm;
    *box_size_list = 2 * bs;
}

/* next = t + (int)((maxnlmove / vmax) / |dt|).  The reference divides by zero
 * for vmax == 0 and overflows the int cast for huge quotients; both make the
 * list never rebuild again, which INT32_MAX expresses. */
int32_t orc_ref_next_build(float maxnlmove, double vmax, double dt, int32_t t) {
    double q = ((double)maxnlmove / vmax) / fabs(dt);
    if (!(q < 2147483647.0 - (double)t)) return INT32_MAX;
    return (int32_t)q + t;
}

/* ---------------- Gear predictor-corrector (cpu/common/Device.cpp) ---------------- */
static void gear_coeffs(double dt, double c[4]) {
    c[0] = dt;
    c[1] = c[0] * dt / 2.0f;
    c[2] = c[1] * dt / 3.0f;
    c[3] = c[2] * dt / 4.0f;
}

/* Taylor predictor on r, v, a, b (c is unchanged), Device.cpp:17-39 */
void orc_ref_predict(orc_real4* r, orc_real3* v, orc_real3* a, orc_real3* b, orc_real3* c,
                     double dt, int32_t n) {
    double k[4];
    gear_coeffs(dt, k);
    for (int32_t i = 0; i < n; i++) {
        double* R = &r[i].x; double* V = &v[i].x; double* A = &a[i].x;
        double* B = &b[i].x; const double* Cc = &c[i].x;
        for (int d = 0; d < 3; d++) {
            R[d] = R[d] + k[0] * V[d] + k[1] * A[d] + k[2] * B[d] + k[3] * Cc[d];
            V[d] = V[d] + k[0] * A[d] + k[1] * B[d] + k[2] * Cc[d];
            A[d] = A[d] + k[0] * B[d] + k[1] * Cc[d];
            B[d] = B[d] + k[0] * Cc[d];
        }
    }
}

/* Corrector with the walled box of Device.cpp:44-120.  Walls at +-npos; the
 * +x wall reflects with restitution energy_loss for y > 0 and teleports the
 * atom to -npos for y < 0 (y == 0 exactly: no action); -x, +-y, +-z reflect.
 * Any wall action sets *flag. */
/* hits (may be NULL): per wall branch, the atoms it acted on, in the order
 * x+ reflect, x+ teleport, x-, y+, y-, z+, z- */
static void correct_counted(orc_real4* r, orc_real3* v, const orc_real3* f, orc_real3* a,
                            orc_real3* b, orc_real3* c, double dt, int32_t n, double npos,
                            int32_t* flag, float energy_loss, int64_t* hits) {
    double k[4];
    gear_coeffs(dt, k);
    double cr = P_GEAR1 * k[1], cv = P_GEAR2 * k[1] / k[0];
    double cb = P_GEAR3 * k[1] / k[2], cc = P_GEAR4 * k[1] / k[3];
    int32_t hit = 0;
    for (int32_t i = 0; i < n; i++) {
        double m = mass_of(r[i].w);
        double* R = &r[i].x; double* V = &v[i].x; double* A = &a[i].x;
        double* B = &b[i].x; double* Cc = &c[i].x; const double* F = &f[i].x;
        for (int d = 0; d < 3; d++) {
            double anew = F[d] / m;
            double corr = anew - A[d];
            R[d] = R[d] + cr * corr;
            V[d] = V[d] + cv * corr;
            A[d] = anew;
            B[d] = B[d] + cb * corr;
            Cc[d] = Cc[d] + cc * corr;
        }
        if (R[0] > npos && R[1] > 0) { hit = 1; V[0] = -V[0] * energy_loss; R[0] = npos; if (hits) hits[0]++; }
        if (R[0] > npos && R[1] < 0) { hit = 1; R[0] = -npos; if (hits) hits[1]++; }
        if (R[0] < -npos) { hit = 1; V[0] = -V[0]; R[0] = -npos; if (hits) hits[2]++; }
        for (int d = 1; d < 3; d++)
            if (fabs(R[d]) > npos) {
                hit = 1;
                if (hits) hits[2 * d + (R[d] > 0 ? 1 : 2)]++;
                V[d] = -V[d];
                R[d] = npos * (R[d] / fabs(R[d]));
            }
    }
    *flag = hit;
}

void orc_ref_correct(orc_real4* r, orc_real3* v, const orc_real3* f, orc_real3* a,
                     orc_real3* b, orc_real3* c, double dt, int32_t n, double npos,
                     int32_t* flag, float energy_loss) {
    correct_counted(r, v, f, a, b, c, dt, n, npos, flag, energy_loss, NULL);
}

void orc_ref_accelerations(const orc_real4* r, orc_real3* acc, const orc_real3* f, int32_t n) {
    for (int32_t i = 0; i < n; i++) {
        double m = mass_of(r[i].w);
        acc[i].x = f[i].x / m; acc[i].y = f[i].y / m; acc[i].z = f[i].z / m;
    }
}

/* ---------------- Lennard-Jones (cpu/common/LennardJones.h) ---------------- */
/* LJ acts only on pairs that involve Ne; sigma/eps are chosen by the partner
 * species of the Ne atom (LennardJones.h:63-88). */
static void lj_pair_params(int32_t ti, int32_t tj, double* sig, double* eps) {
    int32_t sel = (ti == ORC_NG) ? tj : ti;
    if (sel == ORC_N) { *sig = P_SIG_NGN; *eps = P_EPS_NGN; }
    else if (sel == ORC_O) { *sig = P_SIG_NGO; *eps = P_EPS_NGO; }
    else { *sig = P_SIG_NG; *eps = P_EPS_NG; }
}

/* F_i = sum_j 24 eps/sig^2 [2 (sig/r)^14 - (sig/r)^8] (r_i - r_j) over list
 * slots with r^2 < rcutsq (LennardJones.h:12-30, 35-93); f is assigned. */
void orc_ref_lj_forces(const orc_real4* r, orc_real3* f, int32_t n, int32_t nlargestsize,
                       int32_t nlistsize, const int32_t* list, double rcutsq) {
    for (int32_t i = 0; i < n; i++) {
        double fx = 0, fy = 0, fz = 0;
        int32_t ti = (int32_t)r[i].w;
        const int32_t* row = list + (size_t)nlistsize * (size_t)i;
        for (int32_t s = 0; s < nlargestsize && row[s] != -1; s++) {
            int32_t j = row[s];
            double dx = r[i].x - r[j].x, dy = r[i].y - r[j].y, dz = r[i].z - r[j].z;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rcutsq) continue;
            int32_t tj = (int32_t)r[j].w;
            if (ti != ORC_NG && tj != ORC_NG) continue;
            double sig, eps;
            lj_pair_params(ti, tj, &sig, &eps);
            double q = r2 / (sig * sig);          /* (r/sig)^2 */
            double iq4 = 1.0 / (q * q * q * q);   /* (sig/r)^8 */
            double iq7 = iq4 / (q * q * q);       /* (sig/r)^14 */
            double s_ = (2 * iq7 - iq4) * (24.0f * eps / (sig * sig));
            fx += dx * s_; fy += dy * s_; fz += dz * s_;
        }
        f[i].x = fx; f[i].y = fy; f[i].z = fz;
    }
}

/* ---------------- bond order (cpu/common/BondOrderFuncs.h) ---------------- */
typedef struct { double x, y, z; } v3;
static v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 vscale(v3 a, double s) { v3 r = {a.x * s, a.y * s, a.z * s}; return r; }
static v3 vaxpy(v3 y, double s, v3 x) { v3 r = {y.x + s * x.x, y.y + s * x.y, y.z + s * x.z}; return r; }
static double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 P(const orc_real4* r, int32_t i) { v3 p = {r[i].x, r[i].y, r[i].z}; return p; }

/* angular factor g(cos) = 1 + C/D - C/(D + (H - cos)^2) and dg/dcos */
static void ang_g(double cs, double* g, double* dg) {
    double u = (double)P_H - cs;
    double den = (double)P_D + u * u;
    *g = 1.0f + (double)(P_C / P_D) - (double)P_C / den;
    *dg = -2.0 * (double)P_C * u / (den * den);
}

/*
 * bondOrder1 (BondOrderFuncs.h:12-143) for the ordered pair (i,j):
 *   xi_ij = sum_m fc(r_im) g(cos theta_jim) e,   xi_ji = sum_m fc(r_jm) g(cos theta_ijm) e,
 *   e = exp(LAMDA3 (r_ij - re)), m over i's row, m != j, m not Ne, r_im^2 < rcutsq.
 *   b = (1/(1+xi_ij) + 1/(1+xi_ji))/2; dbo = -db/dx_i.
 */
static void bo_pair(const orc_real4* r, int32_t i, int32_t j, double rij, double re,
                    const int32_t* row, int32_t nlargest, double rcutsq,
                    double* bo, v3* dbo, double* xi_ij, double* xi_ji) {
    v3 pi = P(r, i), pj = P(r, j);
    v3 eij = vscale(vsub(pi, pj), 1.0 / rij);       /* d r_ij / d x_i */
    double e = exp((double)P_LAMDA3 * (rij - re));
    double xi = 0, xj = 0;
    v3 dxi = {0, 0, 0}, dxj = {0, 0, 0};
    for (int32_t s = 0; s < nlargest && row[s] != -1; s++) {
        int32_t m = row[s];
        if ((int32_t)r[m].w == ORC_NG || m == j) continue;
        v3 pm = P(r, m);
        v3 im = vsub(pi, pm);
        double rim2 = vdot(im, im);
        if (rim2 >= rcutsq) continue;
        double rim = sqrt(rim2);
        v3 jm = vsub(pj, pm);
        double rjm = sqrt(vdot(jm, jm));
        double fci, dfci, fcj, dfcj;
        orc_ref_cutoff(rim, &fci, &dfci);
        orc_ref_cutoff(rjm, &fcj, &dfcj);
        v3 eim = vscale(im, 1.0 / rim);             /* d r_im / d x_i */
        if (fci > 0) {
            double cs = (rij * rij + rim * rim - rjm * rjm) / (2.0 * rij * rim);
            double g, dg;
            ang_g(cs, &g, &dg);
            xi += fci * g * e;
            /* d cos / d x_i = (eim - cos eij)/rij + (eij - cos eim)/rim */
            v3 dcs = vaxpy(vscale(vaxpy(eim, -cs, eij), 1.0 / rij), 1.0 / rim, vaxpy(eij, -cs, eim));
            dxi = vaxpy(dxi, dfci * g * e, eim);
            dxi = vaxpy(dxi, fci * dg * e, dcs);
            dxi = vaxpy(dxi, fci * g * (double)P_LAMDA3 * e, eij);
        }
        if (fcj > 0) {
            double cs = (rij * rij + rjm * rjm - rim * rim) / (2.0 * rij * rjm);
            double g, dg;
            ang_g(cs, &g, &dg);
            xj += fcj * g * e;
            /* angle at j: d cos / d x_i = (-(x_j - x_m)/rjm - cos eij)/rij */
            v3 dcs = vscale(vaxpy(vscale(jm, -1.0 / rjm), -cs, eij), 1.0 / rij);
            dxj = vaxpy(dxj, fcj * dg * e, dcs);
            dxj = vaxpy(dxj, fcj * g * (double)P_LAMDA3 * e, eij);
        }
    }
    double qi = 1.0 / ((1.0 + xi) * (1.0 + xi)), qj = 1.0 / ((1.0 + xj) * (1.0 + xj));
    *bo = 0.5 * (1.0 / (1.0 + xi) + 1.0 / (1.0 + xj));
    *dbo = vscale(vaxpy(vscale(dxi, qi), qj, dxj), 0.5);
    *xi_ij = xi;
    *xi_ji = xj;
}

/*
 * bondOrder2 (BondOrderFuncs.h:148-237): derivative of b_ij with respect to
 * the third atom k that sits in xi_ij (through fc(r_ik), angle at i) and in
 * xi_ji (through fc(r_jk), angle at j); dbo = -db_ij/dx_k.
 */
static v3 bo_third(const orc_real4* r, int32_t k, int32_t i, int32_t j, double rij, double rik,
                   double re, double xi, double xj) {
    v3 pk = P(r, k), pi = P(r, i), pj = P(r, j);
    v3 jk = vsub(pj, pk);
    double rjk = sqrt(vdot(jk, jk));
    v3 ejk = vscale(vsub(pk, pj), 1.0 / rjk);    /* d r_jk / d x_k */
    v3 eik = vscale(vsub(pk, pi), 1.0 / rik);    /* d r_ik / d x_k */
    v3 uij = vscale(vsub(pj, pi), 1.0 / rij);    /* unit i -> j */
    double e = exp((double)P_LAMDA3 * (rij - re));
    double fci, dfci, fcj, dfcj;
    orc_ref_cutoff(rik, &fci, &dfci);
    orc_ref_cutoff(rjk, &fcj, &dfcj);
    v3 d = {0, 0, 0};
    if (fci > 0) {
        double cs = (rij * rij + rik * rik - rjk * rjk) / (2.0 * rij * rik);
        double g, dg;
        ang_g(cs, &g, &dg);
        v3 dcs = vscale(vaxpy(uij, -cs, eik), 1.0 / rik);
        v3 dx = vaxpy(vscale(eik, dfci * g * e), fci * dg * e, dcs);
        d = vaxpy(d, 0.5 / ((1.0 + xi) * (1.0 + xi)), dx);
    }
    if (fcj > 0) {
        double cs = (rij * rij + rjk * rjk - rik * rik) / (2.0 * rij * rjk);
        double g, dg;
        ang_g(cs, &g, &dg);
        v3 dcs = vscale(vaxpy(vscale(uij, -1.0), -cs, ejk), 1.0 / rjk);
        v3 dx = vaxpy(vscale(ejk, dfcj * g * e), fcj * dg * e, dcs);
        d = vaxpy(d, 0.5 / ((1.0 + xj) * (1.0 + xj)), dx);
    }
    return d;
}

/*
 * manyBodyForces (BondOrderFuncs.h:283-397).
 * Pass 1, for each i and row slot j with r_ij^2 < rcutsq, r_ij^2 <= CUTOFF_RIJSQ
 * and j not Ne:  F_i += (dV_R - b dV_A + dV_vdW) (r_i - r_j)/r_ij - dbo V_A,
 * with the pair terms gated by f_c(r_ij/10) > 0 (bondForces, :242-278).
 * Pass 2, for each k and row pairs (i, j), i != j, r_ik^2 < rcutsq,
 * r_ij^2 < rcutsq, r_ij^2 <= CUTOFF_RIJSQ:  F_k -= dbo_k V_A / 2, using the
 * xi values pass 1 stored for (i, j).  The reference keeps xi in N x N
 * scratch; here it is kept per (i, row slot) and looked up by j.  A pair that
 * pass 1 never visited (the reference reads uninitialised scratch there)
 * uses xi = 0.
 */
void orc_ref_mb_forces(const orc_real4* r, orc_real3* f, int32_t n, int32_t nlargestsize,
                       int32_t nlistsize, const int32_t* list, double rcutsq, int32_t ljexist) {
    size_t cap = (size_t)n * (size_t)(nlistsize > 0 ? nlistsize : 1);
    double* xis = (double*)malloc(sizeof(double) * 2 * cap);
    uint8_t* have = (uint8_t*)calloc(cap, 1);
    if (!xis || !have) { free(xis); free(have); return; }
    const double cut2 = (double)P_CUT_RIJSQ;
    for (int32_t i = 0; i < n; i++) {
        int32_t ti = (int32_t)r[i].w;
        const int32_t* row = list + (size_t)nlistsize * (size_t)i;
        v3 F = {0, 0, 0};
        for (int32_t s = 0; s < nlargestsize && row[s] != -1; s++) {
            int32_t j = row[s];
            v3 ij = vsub(P(r, i), P(r, j));
            double r2 = vdot(ij, ij);
            if (r2 >= rcutsq || r2 > cut2) continue;
            double rij = sqrt(r2);
            int32_t tj = (int32_t)r[j].w;
            if (tj == ORC_NG) continue;
            double va = 0, dva = 0, vr = 0, dvr = 0, vdw, dvdw, bo = 0;
            v3 dbo = {0, 0, 0};
            double fcut, dfcut;
            orc_ref_cutoff(rij / 10, &fcut, &dfcut);
            if (fcut > 0) {
                double de = orc_ref_param(2, ti, tj), re = orc_ref_param(1, ti, tj);
                double be = orc_ref_param(3, ti, tj);
                orc_ref_attractive(rij, de, be, re, &va, &dva);
                orc_ref_repulsive(rij, de, be, re, &vr, &dvr);
                size_t slot = (size_t)nlistsize * (size_t)i + (size_t)s;
                bo_pair(r, i, j, rij, re, row, nlargestsize, rcutsq, &bo, &dbo,
                        &xis[2 * slot], &xis[2 * slot + 1]);
                have[slot] = 1;
            }
            orc_ref_vdw(rij, orc_ref_param(4, ti, tj), &vdw, &dvdw);
            double fij = dvr - bo * dva + dvdw;
            F = vaxpy(F, fij / rij, ij);
            F = vaxpy(F, -va, dbo);
        }
        if (ljexist) { f[i].x += F.x; f[i].y += F.y; f[i].z += F.z; }
        else { f[i].x = F.x; f[i].y = F.y; f[i].z = F.z; }
    }
    for (int32_t k = 0; k < n; k++) {
        const int32_t* row = list + (size_t)nlistsize * (size_t)k;
        for (int32_t a = 0; a < nlargestsize && row[a] != -1; a++) {
            int32_t i = row[a];
            v3 ik = vsub(P(r, k), P(r, i));
            double rik2 = vdot(ik, ik);
            if (rik2 >= rcutsq) continue;
            double rik = sqrt(rik2);
            int32_t ti = (int32_t)r[i].w;
            const int32_t* irow = list + (size_t)nlistsize * (size_t)i;
            for (int32_t b = 0; b < nlargestsize && row[b] != -1; b++) {
                int32_t j = row[b];
                if (i == j) continue;
                v3 ij = vsub(P(r, i), P(r, j));
                double r2 = vdot(ij, ij);
                if (r2 >= rcutsq || r2 > cut2) continue;
                double rij = sqrt(r2);
                int32_t tj = (int32_t)r[j].w;
                double de = orc_ref_param(2, ti, tj), re = orc_ref_param(1, ti, tj);
                double be = orc_ref_param(3, ti, tj);
                double va, dva;
                orc_ref_attractive(rij, de, be, re, &va, &dva);
                double xi = 0, xj = 0;
                for (int32_t s = 0; s < nlargestsize && irow[s] != -1; s++)
                    if (irow[s] == j) {
                        size_t slot = (size_t)nlistsize * (size_t)i + (size_t)s;
                        if (have[slot]) { xi = xis[2 * slot]; xj = xis[2 * slot + 1]; }
                        break;
                    }
                v3 d = bo_third(r, k, i, j, rij, rik, re, xi, xj);
                f[k].x -= 0.5 * d.x * va;
                f[k].y -= 0.5 * d.y * va;
                f[k].z -= 0.5 * d.z * va;
            }
        }
    }
    free(xis);
    free(have);
}

/* ---------------- diagnostics (common/DebugCalculations.cpp) ---------------- */
/* out8 = {EK, T, COMx, COMy, COMz, Px, Py, Pz}; T = sum m|v - v_cm|^2 / (3 N kB),
 * kB = 8.314e-7 in g A^2 fs^-2 mol^-1 K^-1 (DebugCalculations.cpp:46-124). */
void orc_ref_kinetics(const orc_real4* r, const orc_real3* v, int32_t n, double* out) {
    double ek = 0, px = 0, py = 0, pz = 0, mt = 0, cx = 0, cy = 0, cz = 0;
    for (int32_t i = 0; i < n; i++) {
        double m = mass_of(r[i].w);
        ek += m * (v[i].x * v[i].x + v[i].y * v[i].y + v[i].z * v[i].z);
        px += m * v[i].x; py += m * v[i].y; pz += m * v[i].z;
        mt += m;
        cx += m * r[i].x; cy += m * r[i].y; cz += m * r[i].z;
    }
    double vx = px / mt, vy = py / mt, vz = pz / mt, t = 0;
    for (int32_t i = 0; i < n; i++) {
        double m = mass_of(r[i].w);
        double ux = v[i].x - vx, uy = v[i].y - vy, uz = v[i].z - vz;
        t += m * (ux * ux + uy * uy + uz * uz);
    }
    out[0] = ek / 2;
    out[1] = t / (3.0f * (double)n * (double)8.314e-7f);
    out[2] = cx / mt; out[3] = cy / mt; out[4] = cz / mt;
    out[5] = px; out[6] = py; out[7] = pz;
}

/* Potential energy by brute force over all ordered pairs
 * (DebugCalculations.cpp:129-270): pairs with a Ne atom add 2 eps
 * [(s/r)^12 - (s/r)^6] with no cutoff; other pairs with r <= CUTOFF_RIJ add
 * (V_R - b V_A + vdW)/2 with xi summed over every non-Ne atom m != i, j. */
void orc_ref_potential(const orc_real4* r, int32_t n, double* out2) {
    double elj = 0, emb = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t ti = (int32_t)r[i].w;
        for (int32_t j = 0; j < n; j++) {
            if (i == j) continue;
            int32_t tj = (int32_t)r[j].w;
            v3 ij = vsub(P(r, i), P(r, j));
            double r2 = vdot(ij, ij);
            if (ti == ORC_NG || tj == ORC_NG) {
                double sig, eps;
                lj_pair_params(ti, tj, &sig, &eps);
                double s2 = sig * sig / r2, s6 = s2 * s2 * s2;
                elj += 2.0f * eps * (s6 * s6 - s6);
                continue;
            }
            double rij = sqrt(r2);
            if (rij > (double)P_CUT_RIJ) continue;
            double va = 0, dva, vr = 0, dvr, vdw, dvdw, bo = 0, fcut, dfcut;
            orc_ref_cutoff(rij / 10, &fcut, &dfcut);
            if (fcut > 0) {
                double de = orc_ref_param(2, ti, tj), re = orc_ref_param(1, ti, tj);
                double be = orc_ref_param(3, ti, tj);
                orc_ref_attractive(rij, de, be, re, &va, &dva);
                orc_ref_repulsive(rij, de, be, re, &vr, &dvr);
                double e = exp((double)P_LAMDA3 * (rij - re)), xi = 0, xj = 0;
                for (int32_t m = 0; m < n; m++) {
                    if ((int32_t)r[m].w == ORC_NG || m == i || m == j) continue;
                    v3 im = vsub(P(r, i), P(r, m)), jm = vsub(P(r, j), P(r, m));
                    double rim = sqrt(vdot(im, im)), rjm = sqrt(vdot(jm, jm));
                    double fci, dfci, fcj, dfcj, g, dg;
                    orc_ref_cutoff(rim, &fci, &dfci);
                    orc_ref_cutoff(rjm, &fcj, &dfcj);
                    if (fci > 0) {
                        ang_g((rij * rij + rim * rim - rjm * rjm) / (2.0 * rij * rim), &g, &dg);
                        xi += fci * g * e;
                    }
                    if (fcj > 0) {
                        ang_g((rij * rij + rjm * rjm - rim * rim) / (2.0 * rij * rjm), &g, &dg);
                        xj += fcj * g * e;
                    }
                }
                bo = 0.5 * (1.0 / (1.0 + xi) + 1.0 / (1.0 + xj));
            }
            orc_ref_vdw(rij, orc_ref_param(4, ti, tj), &vdw, &dvdw);
            emb += 0.5 * ((vr - bo * va) + vdw);
        }
    }
    out2[0] = elj;
    out2[1] = emb;
}

/* ---------------- the cpumd step loop (cpu/cpumd/Host.cpp:14-216) ---------------- */
static void record(double* rec, const orc_real4* r, const orc_real3* v, int32_t n, int32_t t, double dt) {
    double k[8], u[2];
    orc_ref_kinetics(r, v, n, k);
    orc_ref_potential(r, n, u);
    rec[0] = t * dt; rec[1] = k[1]; rec[3] = k[0]; rec[4] = u[1] + u[0]; rec[2] = rec[3] + rec[4];
    for (int q = 0; q < 6; q++) rec[5 + q] = k[2 + q];
}

typedef struct { orc_list_settings ls; int32_t* list; int32_t next; int32_t n; int32_t on; } ref_list;

static void list_settings(orc_list_settings* ls, float rs, float rc) {
    ls->maxnlmove = (rs - rc) * 0.5f;
    ls->maxnlmovesq = (float)pow(ls->maxnlmove, 2);
    ls->rcutsq = (float)pow(rc, 2);
    ls->rcut = rc;
    ls->rs = rs;
    ls->nlistsize = ls->nlargestsize = 0;
}

static int rebuild(ref_list* L, const orc_real4* r, double box, int32_t buckets, double vmax, double dt, int32_t t) {
    free(L->list);
    L->list = orc_ref_build_neighbor_list(r, &L->ls, box, buckets, L->n);
    if (!L->list) return -1;
    L->next = orc_ref_next_build(L->ls.maxnlmove, vmax, dt, t);
    return 0;
}

int32_t orc_ref_run(const orc_ref_run_cfg* cfg, orc_real4* r, orc_real3* v,
                    double* records, int32_t max_records, double* elapsed_s) {
    int32_t n = cfg->n_lj + cfg->n_mb;
    double dt = (double)cfg->dt;
    double box = pow((double)n, 1.0 / 3.0) * 52.8;
    const double npos = cfg->wall > 0 ? cfg->wall : box;   /* Host.cpp:166: the walls sit at +-boxSize */
    float eloss = sqrtf(cfg->energy_loss);
    int64_t* st = cfg->stats;
    if (st) memset(st, 0, sizeof(int64_t) * ORC_RUN_STATS);
    int32_t teleported = 0;   /* a teleport happened since the last LJ list build */
    orc_real3* f = (orc_real3*)calloc((size_t)n, sizeof(orc_real3));
    orc_real3* a = (orc_real3*)calloc((size_t)n, sizeof(orc_real3));
    orc_real3* b = (orc_real3*)calloc((size_t)n, sizeof(orc_real3));
    orc_real3* c = (orc_real3*)calloc((size_t)n, sizeof(orc_real3));
    if (!f || !a || !b || !c) return -1;
    ref_list lj = {{0}, NULL, 0, n, cfg->use_lj}, mb = {{0}, NULL, 0, cfg->n_mb, cfg->use_mb};
    if (lj.on) list_settings(&lj.ls, cfg->lj_rs, cfg->lj_rcut);
    if (mb.on) list_settings(&mb.ls, cfg->mb_rs, cfg->mb_rcut);
    int32_t buckets = 3 * n, nrec = 0, rc = 0;
    double vmax, bsl, half = box / 2;
    /* Host.cpp:94-106: clamp into +-box/2 and add the centre-of-mass x velocity */
    for (int32_t i = 0; i < n; i++) {
        double* R = &r[i].x;
        for (int d = 0; d < 3; d++)
            if (fabs(R[d]) > half) R[d] = half * (R[d] / fabs(R[d]));
        v[i].x += cfg->vxcm;
    }
    orc_ref_ready_list(r, v, n, &vmax, &bsl);
    if (lj.on && rebuild(&lj, r, bsl, buckets, vmax, dt, 0)) rc = -1;
    if (mb.on && rebuild(&mb, r, bsl, buckets, vmax, dt, 0)) rc = -1;
    if (nrec < max_records) record(records + 11 * nrec++, r, v, n, 0, dt);
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
#define FORCES()                                                                              \
    do {                                                                                      \
        if (lj.on) orc_ref_lj_forces(r, f, n, lj.ls.nlargestsize, lj.ls.nlistsize, lj.list,   \
                                     (double)lj.ls.rcutsq);                                   \
        if (mb.on) orc_ref_mb_forces(r, f, cfg->n_mb, mb.ls.nlargestsize, mb.ls.nlistsize,    \
                                     mb.list, (double)mb.ls.rcutsq, lj.on);                   \
    } while (0)
    FORCES();
    orc_ref_accelerations(r, a, f, n);
    int32_t printout = cfg->output_timesteps, t = 0;
    while (rc == 0 && t < cfg->timesteps) {
        if (lj.next == t || mb.next == t) {
            orc_ref_ready_list(r, v, n, &vmax, &bsl);
            if (lj.on && lj.next == t) {
                if (rebuild(&lj, r, bsl, buckets, vmax, dt, t)) rc = -1;
                if (st) { st[8]++; st[9] += teleported; }
                teleported = 0;
            }
            if (mb.on && mb.next == t && rebuild(&mb, r, bsl, buckets, vmax, dt, t)) rc = -1;
        }
        orc_ref_predict(r, v, a, b, c, dt, n);
        FORCES();
        int32_t flag = 0;
        const int64_t tele0 = st ? st[1] : 0;
        correct_counted(r, v, f, a, b, c, dt, n, npos, &flag, eloss, st);
        if (st && st[1] > tele0) teleported = 1;
        if (flag) {
            if (st) st[7]++;
            FORCES();
            orc_ref_accelerations(r, a, f, n);
        }
        if (printout == t) {
            if (nrec < max_records) record(records + 11 * nrec++, r, v, n, t, dt);
            printout += cfg->output_timesteps;
        }
        t++;
    }
#undef FORCES
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (elapsed_s) *elapsed_s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    if (nrec < max_records) record(records + 11 * nrec++, r, v, n, t, dt);
    free(f); free(a); free(b); free(c); free(lj.list); free(mb.list);
    return rc ? -1 : nrec;
}
