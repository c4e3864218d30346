/*
 * oracle_pbc.c -- FP64 CPU restatement for the periodic benchmark systems:
 * reduced-unit Lennard-Jones with the reference's candidate-list algorithm
 * carried over to a periodic cell grid, the reference's Gear 5th-order
 * predictor-corrector without walls, and Tersoff silicon.
 * TEST INFRASTRUCTURE ONLY: the checker and CPU baseline for the HIP engine.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static double wrap01(double x, double L) {
    double y = x - L * floor(x / L);
    return y >= L ? y - L : y;
}

#ifndef ORC_CFLAGS
#define ORC_CFLAGS "?"
#endif
const char* orc_build_info(void) { return "gcc " __VERSION__ " " ORC_CFLAGS; }

static double min_image(double d, double L) { return d - L * nearbyint(d / L); }
static double sq3(double x, double y, double z) { return x * x + y * y + z * z; }

/* ---------------- periodic candidate list ---------------- */
/*
 * Same row layout as common/NeighborList.cpp:88-145 (27 cells in (dx,dy,dz)
 * lexicographic order from -1, members ascending, self excluded, -1 padding,
 * nlistsize = min(27 * max occupancy, n - 1)) on a periodic grid of
 * floor(L/rs) cells per axis; neighbour cells wrap around.
 */
static int32_t* build_list_threads(const double* pos3, int32_t n, double L, double rs,
                                   int32_t* nlistsize, int32_t* nlargestsize, int32_t threads) {
    int32_t nc = (int32_t)floor(L / rs);
    if (nc < 3) return NULL;
    double h = L / nc;
    size_t ncell = (size_t)nc * nc * nc;
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* start = (int32_t*)calloc(ncell + 1, sizeof(int32_t));
    int32_t* members = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!cell || !start || !members) return NULL;
    for (int32_t i = 0; i < n; i++) {
        int32_t c[3];
        for (int d = 0; d < 3; d++) {
            c[d] = (int32_t)(wrap01(pos3[3 * i + d], L) / h);
            if (c[d] >= nc) c[d] = nc - 1;
        }
        cell[i] = (c[0] * nc + c[1]) * nc + c[2];
        start[cell[i] + 1]++;
    }
    int32_t largest = 0;
    for (size_t c = 0; c < ncell; c++) {
        if (start[c + 1] > largest) largest = start[c + 1];
        start[c + 1] += start[c];
    }
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ncell);
    if (!fill) return NULL;
    memcpy(fill, start, sizeof(int32_t) * ncell);
    for (int32_t i = 0; i < n; i++) members[fill[cell[i]]++] = i;  /* ascending within a cell */
    int32_t nls = 27 * largest;
    if (nls > n - 1) nls = n - 1;
    int32_t* out = (int32_t*)malloc(sizeof(int32_t) * ((size_t)nls * (size_t)n + 1));
    if (!out) return NULL;
    int32_t longest = 0;
#pragma omp parallel for num_threads(threads) if (threads > 1) reduction(max : longest) schedule(static)
    for (int32_t i = 0; i < n; i++) {
        int32_t cz = cell[i] % nc, cy = (cell[i] / nc) % nc, cx = cell[i] / (nc * nc);
        int32_t* row = out + (size_t)nls * (size_t)i;
        int32_t k = 0;
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int32_t x = (cx + dx + nc) % nc, y = (cy + dy + nc) % nc, z = (cz + dz + nc) % nc;
                    int32_t c = (x * nc + y) * nc + z;
                    for (int32_t p = start[c]; p < start[c + 1]; p++)
                        if (members[p] != i) row[k++] = members[p];
                }
        if (k > longest) longest = k;
        for (; k < nls; k++) row[k] = -1;
    }
    *nlistsize = nls;
    *nlargestsize = longest;
    free(cell); free(start); free(members); free(fill);
    return out;
}

int32_t* orc_pbc_build_list(const double* pos3, int32_t n, double L, double rs,
                            int32_t* nlistsize, int32_t* nlargestsize) {
    return build_list_threads(pos3, n, L, rs, nlistsize, nlargestsize, 1);
}

/* Check of an exported pair list (rows of width ints, counts[i] partners of
 * atom i) against the exact FP64 sets at the build positions: every j != i
 * with |x_i - x_j|^2 < rs^2 (minimum image) must be in row i; no entry may
 * lie at rs^2 + slack2 or beyond; no entry may repeat, be i or be out of
 * range.  The exact sets come from a periodic cell grid of floor(L/rs) cells
 * per axis (the 27-cell stencil holds every atom within rs).  stats:
 * [0] entries, [1] pairs within rs missing from the list, [2] entries at or
 * beyond rs^2 + slack2, [3] entries in [rs^2, rs^2 + slack2), [4] repeated,
 * self or out-of-range entries, [5] pairs within rc, [6] of those in the
 * list, [7] pairs within rs.  Returns 0, or -1 (grid narrower than 3 cells,
 * allocation failure). */
int32_t orc_pbc_pair_list_check(const double* pos3, int32_t n, double L, double rs, double rc, double slack2,
                                const int32_t* rows, const int32_t* counts, int32_t width, int64_t* stats) {
    int32_t nc = (int32_t)floor(L / rs);
    if (nc < 3) return -1;
    double h = L / nc;
    size_t ncell = (size_t)nc * nc * nc;
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* start = (int32_t*)calloc(ncell + 1, sizeof(int32_t));
    int32_t* members = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ncell);
    int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * ((size_t)width + 1));
    if (!cell || !start || !members || !fill || !sorted) return -1;
    for (int32_t i = 0; i < n; i++) {
        int32_t c[3];
        for (int d = 0; d < 3; d++) {
            c[d] = (int32_t)(wrap01(pos3[3 * i + d], L) / h);
            if (c[d] >= nc) c[d] = nc - 1;
        }
        cell[i] = (c[0] * nc + c[1]) * nc + c[2];
        start[cell[i] + 1]++;
    }
    for (size_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    memcpy(fill, start, sizeof(int32_t) * ncell);
    for (int32_t i = 0; i < n; i++) members[fill[cell[i]]++] = i;
    for (int k = 0; k < 8; k++) stats[k] = 0;
    const double rs2 = rs * rs, rc2 = rc * rc;
#define R2(i, j) sq3(min_image(pos3[3 * (i)] - pos3[3 * (j)], L), min_image(pos3[3 * (i) + 1] - pos3[3 * (j) + 1], L), \
                     min_image(pos3[3 * (i) + 2] - pos3[3 * (j) + 2], L))
    for (int32_t i = 0; i < n; i++) {
        const int32_t m = counts[i];
        if (m < 0 || m > width) { stats[4]++; continue; }
        /* the row's entries, sorted (insertion sort: rows are short) */
        for (int32_t k = 0; k < m; k++) {
            int32_t v = rows[(size_t)i * width + k], q = k;
            while (q > 0 && sorted[q - 1] > v) { sorted[q] = sorted[q - 1]; q--; }
            sorted[q] = v;
        }
        stats[0] += m;
        for (int32_t k = 0; k < m; k++) {
            int32_t j = sorted[k];
            if (j < 0 || j >= n || j == i || (k > 0 && sorted[k - 1] == j)) { stats[4]++; continue; }
            double r2 = R2(i, j);
            if (r2 >= rs2 + slack2) stats[2]++;
            else if (r2 >= rs2) stats[3]++;
        }
        int32_t cz = cell[i] % nc, cy = (cell[i] / nc) % nc, cx = cell[i] / (nc * nc);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int32_t x = (cx + dx + nc) % nc, y = (cy + dy + nc) % nc, z = (cz + dz + nc) % nc;
                    int32_t c = (x * nc + y) * nc + z;
                    for (int32_t p = start[c]; p < start[c + 1]; p++) {
                        int32_t j = members[p];
                        if (j == i) continue;
                        double r2 = R2(i, j);
                        if (!(r2 < rs2)) continue;
                        stats[7]++;
                        int32_t lo = 0, hi = m;   /* binary search in the sorted row */
                        while (lo < hi) { int32_t mid = (lo + hi) / 2; if (sorted[mid] < j) lo = mid + 1; else hi = mid; }
                        int found = lo < m && sorted[lo] == j;
                        if (!found) stats[1]++;
                        if (r2 < rc2) { stats[5]++; if (found) stats[6]++; }
                    }
                }
    }
#undef R2
    free(cell); free(start); free(members); free(fill); free(sorted);
    return 0;
}

/* 12-6 LJ over candidate rows, minimum image, r < rc; energy shifted so the
 * pair energy is zero at rc.  Each unordered pair is visited from both ends,
 * so energy and virial take half of each visit. */
/* threads > 1 splits the atoms over OpenMP threads (each atom's force is
 * still summed in row order; only the energy and virial sums change order) */
static double lj_forces_threads(const double* pos3, double* f3, int32_t n, double L,
                                double sigma, double eps, double rc, int32_t nlistsize,
                                const int32_t* list, double* virial, int32_t threads) {
    double rc2 = rc * rc, s2 = sigma * sigma;
    double coeff = 24.0 * eps / s2, e4 = 4 * eps;
    double qc = rc2 / s2, iqc6 = 1.0 / (qc * qc * qc);
    double shift = e4 * iqc6 * (iqc6 - 1);
    double e = 0, w = 0;
#pragma omp parallel for num_threads(threads) if (threads > 1) reduction(+ : e, w) schedule(static)
    for (int32_t i = 0; i < n; i++) {
        double fx = 0, fy = 0, fz = 0;
        const int32_t* row = list + (size_t)nlistsize * (size_t)i;
        for (int32_t s = 0; s < nlistsize && row[s] != -1; s++) {
            int32_t j = row[s];
            double dx = min_image(pos3[3 * i] - pos3[3 * j], L);
            double dy = min_image(pos3[3 * i + 1] - pos3[3 * j + 1], L);
            double dz = min_image(pos3[3 * i + 2] - pos3[3 * j + 2], L);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 >= rc2) continue;
            /* F/r = 24 eps/sig^2 [2 (sig/r)^14 - (sig/r)^8], the reference's form */
            double q = r2 / s2;
            double iq4 = 1.0 / (q * q * q * q);
            double iq7 = iq4 / (q * q * q);
            double fr = (2 * iq7 - iq4) * coeff;
            fx += dx * fr; fy += dy * fr; fz += dz * fr;
            double iq6 = 1.0 / (q * q * q);
            e += 0.5 * (e4 * iq6 * (iq6 - 1) - shift);
            w += 0.5 * fr * r2;
        }
        f3[3 * i] = fx; f3[3 * i + 1] = fy; f3[3 * i + 2] = fz;
    }
    if (virial) *virial = w;
    return e;
}

double orc_pbc_lj_forces(const double* pos3, double* f3, int32_t n, double L,
                         double sigma, double eps, double rc, int32_t nlistsize,
                         const int32_t* list, double* virial) {
    return lj_forces_threads(pos3, f3, n, L, sigma, eps, rc, nlistsize, list, virial, 1);
}

/* ---------------- Gear 5th-order PC without walls ---------------- */
/* coefficients as cpu/common/Device.cpp:19-22 and :48-56 */
static void coeffs(double dt, double k[4], double kc[4]) {
    k[0] = dt;
    k[1] = k[0] * dt / 2.0f;
    k[2] = k[1] * dt / 3.0f;
    k[3] = k[2] * dt / 4.0f;
    kc[0] = 0.15833333333333333333333333333333f * k[1];
    kc[1] = 0.75f * k[1] / k[0];
    kc[2] = 0.5f * k[1] / k[2];
    kc[3] = 0.08333333333333333333333333333333f * k[1] / k[3];
}

static void gear_predict_threads(double* r, double* v, double* a, double* b, const double* c,
                                 double dt, int32_t n, int32_t threads) {
    double k[4], kc[4];
    coeffs(dt, k, kc);
#pragma omp parallel for num_threads(threads) if (threads > 1) schedule(static)
    for (size_t q = 0; q < 3 * (size_t)n; q++) {
        r[q] = r[q] + k[0] * v[q] + k[1] * a[q] + k[2] * b[q] + k[3] * c[q];
        v[q] = v[q] + k[0] * a[q] + k[1] * b[q] + k[2] * c[q];
        a[q] = a[q] + k[0] * b[q] + k[1] * c[q];
        b[q] = b[q] + k[0] * c[q];
    }
}

void orc_gear_predict(double* r, double* v, double* a, double* b, const double* c,
                      double dt, int32_t n) {
    gear_predict_threads(r, v, a, b, c, dt, n, 1);
}

static void gear_correct_threads(double* r, double* v, const double* f, double* a, double* b,
                                 double* c, double dt, int32_t n, double mass, double scale, int32_t threads) {
    double k[4], kc[4];
    coeffs(dt, k, kc);
#pragma omp parallel for num_threads(threads) if (threads > 1) schedule(static)
    for (size_t q = 0; q < 3 * (size_t)n; q++) {
        double an = f[q] / mass * scale;
        double corr = an - a[q];
        r[q] = r[q] + kc[0] * corr;
        v[q] = v[q] + kc[1] * corr;
        a[q] = an;
        b[q] = b[q] + kc[2] * corr;
        c[q] = c[q] + kc[3] * corr;
    }
}

void orc_gear_correct(double* r, double* v, const double* f, double* a, double* b,
                      double* c, double dt, int32_t n, double mass, double scale) {
    gear_correct_threads(r, v, f, a, b, c, dt, n, mass, scale, 1);
}

static double kinetic(const double* v, int32_t n) {
    double s = 0;
    for (size_t q = 0; q < 3 * (size_t)n; q++) s += v[q] * v[q];
    return 0.5 * s;
}

int32_t orc_pbc_lj_run(double* r, double* v, int32_t n, double L, double rc, double rs,
                       double dt, int32_t nsteps, int32_t rebuild_every_step,
                       double* energies, double* elapsed_s) {
    return orc_pbc_lj_run_threads(r, v, n, L, rc, rs, dt, nsteps, rebuild_every_step, energies, elapsed_s, 1);
}

int32_t orc_pbc_lj_run_threads(double* r, double* v, int32_t n, double L, double rc, double rs,
                               double dt, int32_t nsteps, int32_t rebuild_every_step,
                               double* energies, double* elapsed_s, int32_t threads) {
    size_t m = 3 * (size_t)n;
    double* f = (double*)calloc(m, sizeof(double));
    double* a = (double*)calloc(m, sizeof(double));
    double* b = (double*)calloc(m, sizeof(double));
    double* c = (double*)calloc(m, sizeof(double));
    double* r0 = (double*)malloc(m * sizeof(double));
    if (!f || !a || !b || !c || !r0) return -1;
    double half_skin2 = 0.25 * (rs - rc) * (rs - rc);
    int32_t nls, nlg, builds = 1;
    int32_t* list = build_list_threads(r, n, L, rs, &nls, &nlg, threads);
    if (!list) return -1;
    memcpy(r0, r, m * sizeof(double));
    double t0 = now_s();
    double u = lj_forces_threads(r, f, n, L, 1.0, 1.0, rc, nls, list, NULL, threads);
    for (size_t q = 0; q < m; q++) a[q] = f[q];
    if (energies) { energies[0] = u; energies[1] = kinetic(v, n); energies[2] = u + energies[1]; }
    for (int32_t t = 1; t <= nsteps; t++) {
        gear_predict_threads(r, v, a, b, c, dt, n, threads);
        /* the list must hold every pair within rc at the positions the force
         * uses: rebuild when any atom moved skin/2 or more since the build */
        int need = rebuild_every_step;
        if (!need) {
#pragma omp parallel for num_threads(threads) if (threads > 1) reduction(| : need) schedule(static)
            for (int32_t i = 0; i < n; i++) {
                double dx = r[3 * i] - r0[3 * i], dy = r[3 * i + 1] - r0[3 * i + 1], dz = r[3 * i + 2] - r0[3 * i + 2];
                if (dx * dx + dy * dy + dz * dz >= half_skin2) need |= 1;
            }
        }
        if (need) {
            free(list);
            list = build_list_threads(r, n, L, rs, &nls, &nlg, threads);
            if (!list) return -1;
            memcpy(r0, r, m * sizeof(double));
            builds++;
        }
        u = lj_forces_threads(r, f, n, L, 1.0, 1.0, rc, nls, list, NULL, threads);
        gear_correct_threads(r, v, f, a, b, c, dt, n, 1.0, 1.0, threads);
        if (energies) {
            double k = kinetic(v, n);
            energies[3 * t] = u; energies[3 * t + 1] = k; energies[3 * t + 2] = u + k;
        }
    }
    if (elapsed_s) *elapsed_s = now_s() - t0;
    free(f); free(a); free(b); free(c); free(r0); free(list);
    return builds;
}

/* ---------------- Tersoff silicon ---------------- */
/* Tersoff, PRB 38, 9902 (1988), Si(C) parameters; cutoff R = 2.7 A, D = 0.3 A */
void orc_tersoff_si_params(orc_tersoff_params* p) {
    p->A = 1830.8; p->B = 471.18;
    p->lambda1 = 2.4799; p->lambda2 = 1.7322; p->lambda3 = 0.0;
    p->beta = 1.1e-6; p->n = 0.78734;
    p->c = 100390.0; p->d = 16.217; p->h = -0.59825;
    p->R = 2.7; p->D = 0.3; p->m = 3;
}

static void ts_fc(const orc_tersoff_params* p, double r, double* fc, double* dfc) {
    if (r < p->R - p->D) { *fc = 1; *dfc = 0; return; }
    if (r > p->R + p->D) { *fc = 0; *dfc = 0; return; }
    double x = M_PI_2 * (r - p->R) / p->D;
    *fc = 0.5 - 0.5 * sin(x);
    *dfc = -0.25 * M_PI / p->D * cos(x);
}

static void ts_g(const orc_tersoff_params* p, double cs, double* g, double* dg) {
    double c2 = p->c * p->c, d2 = p->d * p->d, u = p->h - cs;
    double den = d2 + u * u;
    *g = 1 + c2 / d2 - c2 / den;
    *dg = -2 * c2 * u / (den * den);   /* dg/dcos */
}

/* exp(lambda3^m (rij - rik)^m) and its derivative with respect to (rij - rik) */
static void ts_ex(const orc_tersoff_params* p, double x, double* e, double* de) {
    double l = p->lambda3;
    if (l == 0) { *e = 1; *de = 0; return; }
    if (p->m == 3) {
        double l3 = l * l * l;
        *e = exp(l3 * x * x * x);
        *de = 3 * l3 * x * x * *e;
    } else {
        *e = exp(l * x);
        *de = l * *e;
    }
}

typedef struct { int32_t j; double d[3], r; } ts_nb;

/* neighbours within R + D of each atom, appended in candidate order j */
static int ts_push(ts_nb** nb, size_t* cap, size_t* cnt, const double* x, int32_t i, int32_t j, double L,
                   double rc2) {
    double d[3];
    for (int q = 0; q < 3; q++) {
        d[q] = x[3 * j + q] - x[3 * i + q];
        if (L > 0) d[q] = min_image(d[q], L);
    }
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 >= rc2) return 0;
    if (*cnt == *cap) {
        *cap *= 2;
        *nb = (ts_nb*)realloc(*nb, *cap * sizeof(ts_nb));
        if (!*nb) return -1;
    }
    (*nb)[*cnt].j = j;
    memcpy((*nb)[*cnt].d, d, sizeof d);
    (*nb)[*cnt].r = sqrt(r2);
    (*cnt)++;
    return 0;
}

/* Large periodic systems: cells of edge >= R + D (at least 3 per axis), the
 * 27 surrounding cells scanned per atom; the same neighbour set as the brute
 * force, in a different order. */
static ts_nb* ts_neighbors_cells(const orc_tersoff_params* p, const double* x, int32_t n, double L,
                                 int32_t nc, int32_t** start_out) {
    double rc = p->R + p->D, rc2 = rc * rc, h = L / nc;
    int32_t ncell = nc * nc * nc;
    int32_t* head = (int32_t*)malloc(sizeof(int32_t) * ((size_t)ncell + 1));
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* start = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    size_t cap = 16 * (size_t)n + 16, cnt = 0;
    ts_nb* nb = (ts_nb*)malloc(cap * sizeof(ts_nb));
    if (!head || !cell || !order || !start || !nb) return NULL;
    memset(head, 0, sizeof(int32_t) * ((size_t)ncell + 1));
    for (int32_t i = 0; i < n; i++) {
        int32_t c[3];
        for (int q = 0; q < 3; q++) {
            double y = x[3 * i + q] - L * floor(x[3 * i + q] / L);
            c[q] = (int32_t)(y / h);
            if (c[q] >= nc) c[q] = nc - 1;
        }
        cell[i] = (c[0] * nc + c[1]) * nc + c[2];
        head[cell[i] + 1]++;
    }
    for (int32_t c = 0; c < ncell; c++) head[c + 1] += head[c];
    int32_t* fill = (int32_t*)calloc((size_t)ncell, sizeof(int32_t));
    if (!fill) return NULL;
    for (int32_t i = 0; i < n; i++) order[head[cell[i]] + fill[cell[i]]++] = i;
    free(fill);
    for (int32_t i = 0; i < n; i++) {
        start[i] = (int32_t)cnt;
        int32_t ci = cell[i], cz = ci % nc, cy = (ci / nc) % nc, cx = ci / (nc * nc);
        for (int dx = -1; dx <= 1; dx++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dz = -1; dz <= 1; dz++) {
                    int32_t c = (((cx + dx + nc) % nc) * nc + (cy + dy + nc) % nc) * nc + (cz + dz + nc) % nc;
                    for (int32_t q = head[c]; q < head[c + 1]; q++) {
                        int32_t j = order[q];
                        if (j != i && ts_push(&nb, &cap, &cnt, x, i, j, L, rc2)) return NULL;
                    }
                }
    }
    start[n] = (int32_t)cnt;
    free(head);
    free(cell);
    free(order);
    *start_out = start;
    return nb;
}

/* neighbours within R + D by brute force over all pairs (minimum image when L > 0);
 * periodic systems above 2048 atoms use the cell search */
static ts_nb* ts_neighbors(const orc_tersoff_params* p, const double* x, int32_t n, double L,
                           int32_t** start_out) {
    double rc = p->R + p->D, rc2 = rc * rc;
    if (L > 0 && n > 2048 && (int32_t)floor(L / rc) >= 3)
        return ts_neighbors_cells(p, x, n, L, (int32_t)floor(L / rc), start_out);
    int32_t* start = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    size_t cap = 16 * (size_t)n + 16, cnt = 0;
    ts_nb* nb = (ts_nb*)malloc(cap * sizeof(ts_nb));
    if (!start || !nb) return NULL;
    for (int32_t i = 0; i < n; i++) {
        start[i] = (int32_t)cnt;
        for (int32_t j = 0; j < n; j++) {
            if (j == i) continue;
            double d[3];
            for (int q = 0; q < 3; q++) {
                d[q] = x[3 * j + q] - x[3 * i + q];
                if (L > 0) d[q] = min_image(d[q], L);
            }
            double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            if (r2 >= rc2) continue;
            if (cnt == cap) { cap *= 2; nb = (ts_nb*)realloc(nb, cap * sizeof(ts_nb)); if (!nb) return NULL; }
            nb[cnt].j = j;
            memcpy(nb[cnt].d, d, sizeof d);
            nb[cnt].r = sqrt(r2);
            cnt++;
        }
    }
    start[n] = (int32_t)cnt;
    *start_out = start;
    return nb;
}

/*
 * E = 1/2 sum_i sum_{j != i} fc(r_ij) [A e^{-l1 r_ij} - b_ij B e^{-l2 r_ij}],
 * b_ij = (1 + beta^n zeta_ij^n)^{-1/(2n)},
 * zeta_ij = sum_{k != i,j} fc(r_ik) g(theta_ijk) exp(l3^m (r_ij - r_ik)^m).
 * Forces are the analytic negative gradient.
 */
double orc_tersoff_energy_forces(const orc_tersoff_params* p, const double* x, int32_t n,
                                 double L, double* f) {
    return orc_tersoff_centres(p, x, n, L, NULL, f, NULL);
}

/*
 * The same sum restricted to the centres i with centre[i] != 0 (all centres
 * when centre is NULL): the energy sum_{i in centres} e_i with
 * e_i = 1/2 sum_j V_ij, its negative gradient in f and e_i in e_atom.
 */
double orc_tersoff_centres(const orc_tersoff_params* p, const double* x, int32_t n, double L,
                           const uint8_t* centre, double* f, double* e_atom) {
    int32_t* st;
    ts_nb* nb = ts_neighbors(p, x, n, L, &st);
    if (!nb) return NAN;
    if (f) memset(f, 0, sizeof(double) * 3 * (size_t)n);
    if (e_atom) memset(e_atom, 0, sizeof(double) * (size_t)n);
    double E = 0;
    for (int32_t i = 0; i < n; i++) {
        if (centre && !centre[i]) continue;
        for (int32_t a = st[i]; a < st[i + 1]; a++) {
            const ts_nb* J = &nb[a];
            double rij = J->r, u[3] = {J->d[0] / rij, J->d[1] / rij, J->d[2] / rij};
            double fc, dfc;
            ts_fc(p, rij, &fc, &dfc);
            double fr = p->A * exp(-p->lambda1 * rij), dfr = -p->lambda1 * fr;
            double fa = -p->B * exp(-p->lambda2 * rij), dfa = -p->lambda2 * fa;
            /* zeta and its pieces */
            double zeta = 0;
            for (int32_t b = st[i]; b < st[i + 1]; b++) {
                if (b == a) continue;
                const ts_nb* K = &nb[b];
                double fk, dfk, g, dg, ex, dex;
                ts_fc(p, K->r, &fk, &dfk);
                double cs = (J->d[0] * K->d[0] + J->d[1] * K->d[1] + J->d[2] * K->d[2]) / (rij * K->r);
                ts_g(p, cs, &g, &dg);
                ts_ex(p, rij - K->r, &ex, &dex);
                zeta += fk * g * ex;
            }
            double bz = 1, dbz = 0;
            if (zeta > 0) {
                double t = pow(p->beta * zeta, p->n);
                bz = pow(1 + t, -0.5 / p->n);
                dbz = -0.5 * bz / (1 + t) * t / zeta;   /* db/dzeta */
            }
            E += 0.5 * fc * (fr + bz * fa);
            if (e_atom) e_atom[i] += 0.5 * fc * (fr + bz * fa);
            if (!f) continue;
            /* pair part: d/d r_ij at fixed zeta */
            double dEdr = 0.5 * (dfc * (fr + bz * fa) + fc * (dfr + bz * dfa));
            double pref = 0.5 * fc * fa * dbz;   /* dE/dzeta */
            double gi[3] = {0, 0, 0}, gj[3] = {0, 0, 0};
            for (int q = 0; q < 3; q++) { gi[q] -= dEdr * u[q]; gj[q] += dEdr * u[q]; }
            if (pref != 0) {
                for (int32_t b = st[i]; b < st[i + 1]; b++) {
                    if (b == a) continue;
                    const ts_nb* K = &nb[b];
                    double rik = K->r, w[3] = {K->d[0] / rik, K->d[1] / rik, K->d[2] / rik};
                    double fk, dfk, g, dg, ex, dex;
                    ts_fc(p, rik, &fk, &dfk);
                    double cs = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
                    ts_g(p, cs, &g, &dg);
                    ts_ex(p, rij - rik, &ex, &dex);
                    /* dzeta contributions */
                    double d_rij = fk * g * dex, d_rik = dfk * g * ex - fk * g * dex, d_cs = fk * dg * ex;
                    double gk[3];
                    for (int q = 0; q < 3; q++) {
                        double dcs_dj = (w[q] - cs * u[q]) / rij;   /* d cos / d x_j */
                        double dcs_dk = (u[q] - cs * w[q]) / rik;   /* d cos / d x_k */
                        double gjq = pref * (d_rij * u[q] + d_cs * dcs_dj);
                        gk[q] = pref * (d_rik * w[q] + d_cs * dcs_dk);
                        gj[q] += gjq;
                        gi[q] -= gjq + gk[q];
                    }
                    for (int q = 0; q < 3; q++) f[3 * K->j + q] -= gk[q];
                }
            }
            for (int q = 0; q < 3; q++) { f[3 * i + q] -= gi[q]; f[3 * J->j + q] -= gj[q]; }
        }
    }
    free(nb);
    free(st);
    return E;
}
