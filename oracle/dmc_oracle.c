/*
 * dmc_oracle.c -- sequential CPU restatement of the reference DMC path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker).  The engine in tmc_amd/
 * never links or calls this file.
 *
 * Build with -O2 -ffp-contract=off: the reference's float/double mixture is
 * reproduced operation by operation, so no FMA contraction is allowed.
 * Reference citations are to cpp_dmc/ in the reference checkout.
 */
#include "dmc_oracle.h"
#include "../tmc_amd/csrc/dmc_tables.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* tables                                                              */
/* ------------------------------------------------------------------ */
static const uint64_t k_case[256] = DMC_CASE_CONTOURS_INIT;     /* LookupTables.h:1101 */
static const uint64_t k_ambig[4] = DMC_AMBIGUOUS_MASK_INIT;     /* LookupTables.h:1877 */

/* edge end points, DualMarchingCubes.cpp:110 (l_edges_) */
static const int E_V0[12] = {0, 1, 2, 0, 4, 5, 6, 4, 0, 1, 3, 2};
static const int E_V1[12] = {1, 3, 3, 2, 5, 7, 7, 6, 4, 5, 7, 6};
/* axis of each edge, DualMarchingCubes.h:396-397 (e_ pattern) */
static const int E_AXIS[12] = {0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2};
/* oriented face vertices / edges, DualMarchingCubes.h:301-303 */
static const int F_V[6][4] = {{0, 2, 1, 3}, {5, 7, 4, 6}, {4, 0, 5, 1},
                              {2, 6, 3, 7}, {4, 6, 0, 2}, {1, 3, 5, 7}};
static const int F_E[6][4] = {{3, 2, 1, 0}, {5, 6, 7, 4}, {8, 0, 9, 4},
                              {11, 6, 10, 2}, {7, 11, 3, 8}, {1, 10, 5, 9}};
/* global edge key: point offset and direction, DualMarchingCubes.h:215-223 */
static const int E_KEY[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1},
                                 {0, 0, 1, 0}, {1, 0, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1},
                                 {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
/* quad slot of an edge by orientation bit, DualMarchingCubes.h:226-230 */
static const int E_SLOT[12][2] = {{0, 0}, {0, 0}, {3, 1}, {3, 1}, {1, 3}, {1, 3},
                                  {2, 2}, {2, 2}, {0, 0}, {3, 1}, {2, 2}, {1, 3}};
/* corner order of the bilinear faces per projection, DualMarchingCubes.cpp:200-202 */
static const int PRJ_W[8] = {0, 1, 2, 3, 4, 5, 6, 7};
static const int PRJ_V[8] = {0, 1, 4, 5, 2, 3, 6, 7};
static const int PRJ_U[8] = {0, 2, 4, 6, 1, 3, 5, 7};

/* ------------------------------------------------------------------ */
/* small growable arrays                                               */
/* ------------------------------------------------------------------ */
typedef struct { void* p; int64_t n, cap; size_t esz; } vec;
static void vec_init(vec* v, size_t esz) { v->p = NULL; v->n = 0; v->cap = 0; v->esz = esz; }
static void* vec_push(vec* v) {
    if (v->n == v->cap) {
        v->cap = v->cap ? 2 * v->cap : 1024;
        v->p = realloc(v->p, (size_t)v->cap * v->esz);
    }
    return (char*)v->p + (size_t)(v->n++) * v->esz;
}
static void* xmalloc(size_t n) { return malloc(n ? n : 1); }

void dmco_free(void* p) { free(p); }
void dmco_mesh_free(dmco_mesh* m) {
    free(m->v); free(m->n); free(m->quads); free(m->colors); free(m->tris);
    memset(m, 0, sizeof(*m));
}
void dmco_halfedges_free(dmco_halfedges* h) {
    free(h->he); free(h->he_v); free(h->he_f); free(h->bnd);
    memset(h, 0, sizeof(*h));
}

/* ------------------------------------------------------------------ */
/* UniformGrid                                                         */
/* ------------------------------------------------------------------ */
static int64_t gidx(const dmco_grid* g, int64_t i, int64_t j, int64_t k) {
    return (k * g->ny + j) * g->nx + i;           /* UniformGrid.h:263 */
}

/* UniformGrid.cpp:156-184 (estimateGradient) and :186-197 (flip_gradient).
 * The index lambda maps i+1 == n to plane 0 (max - i), the lower side clamps. */
void dmco_estimate_gradient(const dmco_grid* g, const double* s, double* grad, int flip) {
    const int64_t nr = (int64_t)g->nx * g->ny * g->nz;
    for (int64_t q = 0; q < nr; q++) {
        const int i = (int)(q % g->nx), j = (int)((q / g->nx) % g->ny), k = (int)(q / ((int64_t)g->nx * g->ny));
#define IDX(a, m) (((a) < 0) ? 0 : ((a) >= (m)) ? (m) - (a) : (a))
        const int i0 = IDX(i - 1, g->nx), i1 = IDX(i + 1, g->nx);
        const int j0 = IDX(j - 1, g->ny), j1 = IDX(j + 1, g->ny);
        const int k0 = IDX(k - 1, g->nz), k1 = IDX(k + 1, g->nz);
#undef IDX
        const double gx = (s[gidx(g, i1, j, k)] - s[gidx(g, i0, j, k)]) / (2 * g->dx);
        const double gy = (s[gidx(g, i, j1, k)] - s[gidx(g, i, j0, k)]) / (2 * g->dy);
        const double gz = (s[gidx(g, i, j, k1)] - s[gidx(g, i, j, k0)]) / (2 * g->dz);
        grad[3 * q + 0] = gx;
        grad[3 * q + 1] = gy;
        grad[3 * q + 2] = gz;
    }
    if (flip) {
        const double sgn = -1.;
        for (int64_t q = 0; q < 3 * nr; q++) grad[q] *= sgn;
    }
}

/* ------------------------------------------------------------------ */
/* MC polygons                                                         */
/* ------------------------------------------------------------------ */
typedef struct { int cnt; int size[4]; int e[4][12]; } polys;

/* DualMarchingCubes.cpp:463-481: unambiguous case from the table. */
static void polygons_table(int icase, polys* P) {
    const uint64_t w = k_case[icase];
    int pos = 0;
    P->cnt = 0;
    for (int t = 0; t < 4; t++) {
        const int s = (int)((w >> (4 * t)) & 0xF);
        if (s == 0) break;
        P->size[t] = s;
        for (int i = 0; i < s; i++) P->e[t][i] = (int)((w >> (16 + 4 * (pos++))) & 0xF);
        P->cnt++;
    }
}

/* DualMarchingCubes.h:330-332 */
static double asymptotic_decider(double f0, double f1, double f2, double f3) {
    return (f0 * f3 - f1 * f2) / (f0 + f3 - f1 - f2);
}

/* DualMarchingCubes.cpp:328-461: oriented segments per face, ambiguous
 * faces resolved by the asymptotic decider, then linked into closed
 * contours starting from the lowest intersected edge. */
static void polygons_decider(double iso, const double F[8], polys* P) {
    int out[12], in_[12], set[12];
    for (int e = 0; e < 12; e++) { out[e] = 15; in_[e] = 15; set[e] = 0; }
#define SEG(a, b) do { out[a] = (b); set[a] = 1; in_[b] = (a); set[b] = 1; } while (0)
    for (int f = 0; f < 6; f++) {
        const double f0 = F[F_V[f][0]], f1 = F[F_V[f][1]], f2 = F[F_V[f][2]], f3 = F[F_V[f][3]];
        const int e0 = F_E[f][0], e1 = F_E[f][1], e2 = F_E[f][2], e3 = F_E[f][3];
        unsigned fc = 0;
        if (f0 >= iso) fc |= 1;
        if (f1 >= iso) fc |= 2;
        if (f2 >= iso) fc |= 4;
        if (f3 >= iso) fc |= 8;
        switch (fc) {
        case 1: SEG(e0, e3); break;
        case 2: SEG(e1, e0); break;
        case 3: SEG(e1, e3); break;
        case 4: SEG(e3, e2); break;
        case 5: SEG(e0, e2); break;
        case 6: {
            const double val = asymptotic_decider(f0, f1, f2, f3);
            if (val >= iso) { SEG(e3, e0); SEG(e1, e2); }
            else if (val < iso) { SEG(e1, e0); SEG(e3, e2); }
        } break;
        case 7: SEG(e1, e2); break;
        case 8: SEG(e2, e1); break;
        case 9: {
            const double val = asymptotic_decider(f0, f1, f2, f3);
            if (val >= iso) { SEG(e0, e1); SEG(e2, e3); }
            else if (val < iso) { SEG(e0, e3); SEG(e2, e1); }
        } break;
        case 10: SEG(e2, e0); break;
        case 11: SEG(e2, e3); break;
        case 12: SEG(e3, e1); break;
        case 13: SEG(e0, e1); break;
        case 14: SEG(e3, e0); break;
        default: break;
        }
    }
#undef SEG
    (void)in_;
    P->cnt = 0;
    for (int e = 0; e < 12; e++) {
        if (!set[e]) continue;
        if (P->cnt >= 4) break;                  /* cannot happen for valid input */
        const int start = e;
        int to = out[e], pos = 0, guard = 0;
        P->e[P->cnt][pos] = start;
        while (to != start && to < 12 && guard++ < 12) {
            pos++;
            P->e[P->cnt][pos] = to;
            const int cur = to;
            to = out[cur];
            set[cur] = 0;
        }
        P->size[P->cnt] = pos + 1;
        P->cnt++;
    }
}

static int is_ambiguous(int icase) { return (int)((k_ambig[icase >> 6] >> (icase & 63)) & 1); }

static void cell_polygons(int icase, double iso, const double f[8], polys* P) {
    if (is_ambiguous(icase)) polygons_decider(iso, f, P);    /* DualMarchingCubes.cpp:103-109 */
    else polygons_table(icase, P);
}

/* DualMarchingCubes.h:391-407: local coordinates of the edge intersection.
 * The parameter is rounded to float and the coordinate assembled in float. */
static void local_coords(int e, const double f[8], double iso, double u[3]) {
    const int v0 = E_V0[e], v1 = E_V1[e];
    const float l = (iso - f[v0]) / (f[v1] - f[v0]);
    for (int c = 0; c < 3; c++) {
        const float val = (E_AXIS[e] == c) ? 1.f : 0.f;
        const float con = (float)((v0 >> c) & 1);
        u[c] = l * val + con;
    }
}

/* DualMarchingCubes.h:590-597 */
static void trilinear_gradient(const double f[8], double u, double v, double w, double g[3]) {
    g[0] = (1 - w) * ((1 - v) * (f[1] - f[0]) + v * (f[3] - f[2])) + w * ((1 - v) * (f[5] - f[4]) + v * (f[7] - f[6]));
    g[1] = (1 - w) * ((1 - u) * (f[2] - f[0]) + u * (f[3] - f[1])) + w * ((1 - u) * (f[6] - f[4]) + u * (f[7] - f[5]));
    g[2] = (1 - v) * ((1 - u) * (f[4] - f[0]) + u * (f[5] - f[1])) + v * ((1 - u) * (f[6] - f[2]) + u * (f[7] - f[3]));
}

/* DualMarchingCubes.h:529-538 (1 = W, 2 = V, 3 = U) */
static int l_projection(const double n[3]) {
    const double dx = fabs(n[0]), dy = fabs(n[1]), dz = fabs(n[2]);
    if (dz >= dx && dz >= dy) return 1;
    if (dy >= dx && dy >= dz) return 2;
    if (dx >= dy && dx >= dz) return 3;
    return 1;
}

/* DualMarchingCubes.h:546-552 */
static int min_projection(double u, double v, double w) {
    if (w <= u && w <= v) return 1;
    if (v <= u && v <= w) return 2;
    if (u <= v && u <= w) return 3;
    return 1;
}

typedef double bbox4[4][2];

/* DualMarchingCubes.h:500-519 */
static int in_neighbor_bbox(double u, double v, double w, const polys* P, int t,
                            const bbox4 bA, const bbox4 bB, const bbox4 bC) {
    for (int i = 0; i < P->cnt; i++) {
        if (P->size[i] == 3 && t != i) {
            if (u < bA[i][0] || bA[i][1] < u) continue;
            if (v < bB[i][0] || bB[i][1] < v) continue;
            if (w < bC[i][0] || bC[i][1] < w) continue;
            return 1;
        }
    }
    return 0;
}

/* DualMarchingCubes.cpp:274-326: sample the bilinear faces on the contour's
 * bounding box (interior (ns-2)^2 samples), keep the closest admissible
 * iso-surface point. */
static int projection(const int r[8], double iso, const double f[8], const polys* P, int t,
                      double p[3], const bbox4 bA, const bbox4 bB, const bbox4 bC, int ns) {
    const double umin = bA[t][0], umax = bA[t][1];
    const double vmin = bB[t][0], vmax = bB[t][1];
    const double du = (umax - umin) / (ns - 1);
    const double dv = (vmax - vmin) / (ns - 1);
    double minDistance = 100;
    double pt[3] = {0, 0, -1};
    double wmin = bC[t][0], wmax = bC[t][1];
    const double eps = 1e-5;
    if (fabs(wmax - wmin) < eps) { wmin -= eps; wmax += eps; }
    const int cnt_size = P->size[t];
    for (int i = 1; i < (ns - 1); i++) {
        const float u = umin + i * du;
        for (int j = 1; j < (ns - 1); j++) {
            const float v = vmin + j * dv;
            const float g1 = (1 - v) * ((1 - u) * f[0] + u * f[r[1]]) + v * ((1 - u) * f[r[2]] + u * f[r[3]]);
            const float g2 = (1 - v) * ((1 - u) * f[r[4]] + u * f[r[5]]) + v * ((1 - u) * f[r[6]] + u * f[r[7]]);
            if (g1 == g2) continue;
            const float w = (iso - g1) / (g2 - g1);
            if (wmin <= w && w <= wmax) {
                if (P->cnt == 1 || cnt_size == 3 || !in_neighbor_bbox(u, v, w, P, t, bA, bB, bC)) {
                    const float d = (p[0] - u) * (p[0] - u) + (p[1] - v) * (p[1] - v) + (p[2] - w) * (p[2] - w);
                    if (minDistance > d) {
                        minDistance = d;
                        pt[0] = u; pt[1] = v; pt[2] = w;
                    }
                }
            }
        }
    }
    if (pt[2] > -1) {
        p[0] = pt[0]; p[1] = pt[1]; p[2] = pt[2];
        return 1;
    }
    return 0;
}

/* DualMarchingCubes.cpp:194-272 */
static void representative(double iso, const double f[8], const polys* P, int t, double p[3],
                           const bbox4 bU, const bbox4 bV, const bbox4 bW,
                           uint32_t* c1, uint32_t* c2) {
    const int ns = 9;
    int prj;
    if (P->size[t] >= 6) {
        double g[3];
        trilinear_gradient(f, p[0], p[1], p[2], g);
        prj = l_projection(g);
    } else {
        prj = min_projection(bU[t][1] - bU[t][0], bV[t][1] - bV[t][0], bW[t][1] - bW[t][0]);
    }
    double pt[3];
    if (prj == 1) {
        pt[0] = p[0]; pt[1] = p[1]; pt[2] = p[2];
        if (!projection(PRJ_W, iso, f, P, t, pt, bU, bV, bW, ns)) {
            (*c1)++;
            if (!projection(PRJ_W, iso, f, P, t, pt, bU, bV, bW, ns * 3 + 1)) (*c2)++;
        }
        p[0] = pt[0]; p[1] = pt[1]; p[2] = pt[2];
    } else if (prj == 2) {
        pt[0] = p[0]; pt[1] = p[2]; pt[2] = p[1];
        if (!projection(PRJ_V, iso, f, P, t, pt, bU, bW, bV, ns)) {
            (*c1)++;
            if (!projection(PRJ_V, iso, f, P, t, pt, bU, bW, bV, ns * 3 + 1)) (*c2)++;
        }
        p[0] = pt[0]; p[1] = pt[2]; p[2] = pt[1];
    } else {
        pt[0] = p[1]; pt[1] = p[2]; pt[2] = p[0];
        if (!projection(PRJ_U, iso, f, P, t, pt, bV, bW, bU, ns)) {
            (*c1)++;
            if (!projection(PRJ_U, iso, f, P, t, pt, bV, bW, bU, ns * 3 + 1)) (*c2)++;
        }
        p[0] = pt[2]; p[1] = pt[0]; p[2] = pt[1];
    }
}

/* The per-contour part of slice (DualMarchingCubes.cpp:113-158): bounding
 * boxes, orientation bits, mean point and representative, in local
 * coordinates.  Returns the orientation bits e_. */
static unsigned cell_representatives(double iso, const double f[8], const polys* P,
                                     double loc[4][3], uint32_t* c1, uint32_t* c2) {
    bbox4 bU, bV, bW;
    unsigned e_ = 0;
    for (int t = 0; t < 4; t++) { bU[t][0] = bU[t][1] = bV[t][0] = bV[t][1] = bW[t][0] = bW[t][1] = 0; }
    for (int t = 0; t < P->cnt; t++) {
        bU[t][0] = 1; bU[t][1] = 0;
        bV[t][0] = 1; bV[t][1] = 0;
        bW[t][0] = 1; bW[t][1] = 0;
        for (int i = 0; i < P->size[t]; i++) {
            const int e = P->e[t][i];
            double ui[3];
            local_coords(e, f, iso, ui);
            bU[t][0] = bU[t][0] > ui[0] ? ui[0] : bU[t][0];
            bU[t][1] = bU[t][1] < ui[0] ? ui[0] : bU[t][1];
            bV[t][0] = bV[t][0] > ui[1] ? ui[1] : bV[t][0];
            bV[t][1] = bV[t][1] < ui[1] ? ui[1] : bV[t][1];
            bW[t][0] = bW[t][0] > ui[2] ? ui[2] : bW[t][0];
            bW[t][1] = bW[t][1] < ui[2] ? ui[2] : bW[t][1];
            if (f[E_V0[e]] < iso) e_ |= (1u << e);
        }
    }
    for (int t = 0; t < P->cnt; t++) {
        double p[3] = {0, 0, 0};
        for (int i = 0; i < P->size[t]; i++) {
            double ui[3];
            local_coords(P->e[t][i], f, iso, ui);
            p[0] += ui[0]; p[1] += ui[1]; p[2] += ui[2];
        }
        p[0] /= P->size[t]; p[1] /= P->size[t]; p[2] /= P->size[t];
        representative(iso, f, P, t, p, bU, bV, bW, c1, c2);
        loc[t][0] = p[0]; loc[t][1] = p[1]; loc[t][2] = p[2];
    }
    return e_;
}

/* UniformGrid.h:286-299 */
static void grid_normal(const dmco_grid* g, const double* grad, int i, int j, int k,
                        double u, double v, double w, double n[3]) {
    const double* G[8];
    for (int c = 0; c < 8; c++) G[c] = grad + 3 * gidx(g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1));
    for (int c = 0; c < 3; c++) {
        const double a = G[0][c] * (1 - u) + G[1][c] * u;
        const double b = G[2][c] * (1 - u) + G[3][c] * u;
        const double cc = G[4][c] * (1 - u) + G[5][c] * u;
        const double d = G[6][c] * (1 - u) + G[7][c] * u;
        const double lo = a * (1 - v) + b * v;
        const double hi = cc * (1 - v) + d * v;
        n[c] = lo * (1 - w) + hi * w;
    }
    const double s = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);   /* Vector.h:12-17 */
    n[0] /= s; n[1] /= s; n[2] /= s;
}

/* ------------------------------------------------------------------ */
/* cell sweep + quad map + unused vertex removal                        */
/* ------------------------------------------------------------------ */
typedef struct { int64_t key; int64_t seq; int32_t pos, vaddr, color; } qrec;

static int cmp_qrec(const void* a, const void* b) {
    const qrec* x = (const qrec*)a; const qrec* y = (const qrec*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq);
}

int dmco_cells(const dmco_grid* g, const double* s, const double* grad, double iso, dmco_mesh* out) {
    memset(out, 0, sizeof(*out));
    vec V, N, R;
    vec_init(&V, 3 * sizeof(double));
    vec_init(&N, 3 * sizeof(double));
    vec_init(&R, sizeof(qrec));
    /* DualMarchingCubes.cpp:16-50 */
    for (int k = 0; k < g->nz - 1; k++)
        for (int j = 0; j < g->ny - 1; j++)
            for (int i = 0; i < g->nx - 1; i++) {
                double u[8];
                for (int c = 0; c < 8; c++) u[c] = s[gidx(g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1))];
                unsigned icase = 0;
                for (int c = 0; c < 8; c++) icase += ((unsigned)(u[c] >= iso)) << c;
                if (icase == 0 || icase == 255) continue;
                /* slice, DualMarchingCubes.cpp:97-192 */
                polys P;
                cell_polygons((int)icase, iso, u, &P);
                double loc[4][3];
                const unsigned e_ = cell_representatives(iso, u, &P, loc, &out->nr_proj1, &out->nr_proj2);
                for (int t = 0; t < P.cnt; t++) {
                    double ni[3], pi[3];
                    grid_normal(g, grad, i, j, k, loc[t][0], loc[t][1], loc[t][2], ni);
                    pi[0] = g->x0 + (i + loc[t][0]) * g->dx;       /* UniformGrid.h:300-305 */
                    pi[1] = g->y0 + (j + loc[t][1]) * g->dy;
                    pi[2] = g->z0 + (k + loc[t][2]) * g->dz;
                    const int32_t vaddr = (int32_t)V.n;
                    memcpy(vec_push(&V), pi, sizeof(pi));
                    memcpy(vec_push(&N), ni, sizeof(ni));
                    for (int m = 0; m < P.size[t]; m++) {
                        const int e = P.e[t][m];
                        qrec* r = (qrec*)vec_push(&R);
                        r->key = 3 * gidx(g, i + E_KEY[e][0], j + E_KEY[e][1], k + E_KEY[e][2]) + E_KEY[e][3];
                        r->seq = R.n - 1;
                        r->pos = E_SLOT[e][(e_ >> e) & 1];
                        r->vaddr = vaddr;
                        const int par = (i & 1) | (j & 1) << 1 | (k & 1) << 2;
                        r->color = (e == 0) ? 3 * par : (e == 3) ? 3 * par + 1 : (e == 8) ? 3 * par + 2 : -1;
                    }
                }
            }
    /* collectQuadrilaterals, DualMarchingCubes.h:663-676 (map order = key order) */
    qsort(R.p, (size_t)R.n, sizeof(qrec), cmp_qrec);
    vec Q, C;
    vec_init(&Q, 4 * sizeof(int32_t));
    vec_init(&C, sizeof(int32_t));
    const qrec* rr = (const qrec*)R.p;
    for (int64_t a = 0; a < R.n;) {
        int64_t b = a;
        int32_t q[5] = {-1, -1, -1, -1, -1};
        while (b < R.n && rr[b].key == rr[a].key) {
            q[rr[b].pos] = rr[b].vaddr;                  /* addVertex, DualMarchingCubes.h:607-652 */
            if (rr[b].color != -1) q[4] = rr[b].color;
            b++;
        }
        if (q[0] != -1 && q[1] != -1 && q[2] != -1 && q[3] != -1) {
            memcpy(vec_push(&Q), q, 4 * sizeof(int32_t));
            *(int32_t*)vec_push(&C) = q[4];
        }
        a = b;
    }
    free(R.p);
    /* removeUnusedVertices, DualMarchingCubes.h:685-729 */
    const int64_t nv = V.n, nq = Q.n;
    int32_t* quads = (int32_t*)Q.p;
    uint8_t* flag = (uint8_t*)calloc((size_t)nv + 1, 1);
    for (int64_t f = 0; f < 4 * nq; f++) flag[quads[f]] = 1;
    int32_t* map = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)(nv + 1));
    int64_t nnv = 0;
    double* vv = (double*)V.p;
    double* nn = (double*)N.p;
    for (int64_t a = 0; a < nv; a++) {
        map[a] = -1;
        if (flag[a]) {
            map[a] = (int32_t)nnv;
            memmove(vv + 3 * nnv, vv + 3 * a, 3 * sizeof(double));
            memmove(nn + 3 * nnv, nn + 3 * a, 3 * sizeof(double));
            nnv++;
        }
    }
    for (int64_t f = 0; f < 4 * nq; f++) quads[f] = map[quads[f]];
    free(flag);
    free(map);
    out->nv_before_unused = nv;
    out->nv = nnv;
    out->v = vv ? vv : (double*)xmalloc(8);
    out->n = nn ? nn : (double*)xmalloc(8);
    out->nq = nq;
    out->quads = quads ? quads : (int32_t*)xmalloc(4);
    out->colors = C.p ? (int32_t*)C.p : (int32_t*)xmalloc(4);
    return 0;
}

/* ------------------------------------------------------------------ */
/* halfedges                                                           */
/* ------------------------------------------------------------------ */
typedef struct { uint64_t key; int32_t he; } hrec;
static int cmp_hrec(const void* a, const void* b) {
    const hrec* x = (const hrec*)a; const hrec* y = (const hrec*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->he < y->he ? -1 : (x->he > y->he);
}
/* setKey, DualMarchingCubes.cpp:498-504 */
static uint64_t he_key(int32_t v0, int32_t v1) {
    if (v0 < v1) return ((uint64_t)(int64_t)v0 << 32) | (uint64_t)((int64_t)v1 & 0xffffffffLL);
    return ((uint64_t)(int64_t)v1 << 32) | (uint64_t)((int64_t)v0 & 0xffffffffLL);
}

/* DualMarchingCubes.cpp:488-671 */
int dmco_halfedges_build(int64_t nv, int64_t nq, const int32_t* quads, dmco_halfedges* H) {
    memset(H, 0, sizeof(*H));
    H->nhe = 4 * nq;
    H->he = (int32_t*)xmalloc(sizeof(int32_t) * 5 * (size_t)H->nhe);
    H->he_v = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nv);
    H->he_f = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nq);
    for (int64_t a = 0; a < nv; a++) H->he_v[a] = -1;
    hrec* R = (hrec*)xmalloc(sizeof(hrec) * (size_t)H->nhe);
    int32_t* he = H->he;
    for (int64_t i = 0; i < nq; i++) {
        const int32_t* q = quads + 4 * i;
        for (int k = 0; k < 4; k++) {
            int32_t* h = he + 5 * (4 * i + k);
            h[0] = q[k];
            h[1] = (int32_t)i;
            h[2] = (int32_t)(4 * i + (k + 1) % 4);
            h[3] = -1;
            h[4] = 0;
        }
        H->he_f[i] = (int32_t)(4 * i);
        for (int k = 0; k < 4; k++) H->he_v[q[k]] = (int32_t)(4 * i + k);
        for (int k = 0; k < 4; k++) {
            R[4 * i + k].key = he_key(q[k], q[(k + 1) % 4]);
            R[4 * i + k].he = (int32_t)(4 * i + k);
        }
    }
    qsort(R, (size_t)H->nhe, sizeof(hrec), cmp_hrec);
    vec B;
    vec_init(&B, sizeof(int32_t));
#define TW(x) he[5 * (x) + 3]
#define FL(x) he[5 * (x) + 4]
#define OR(x) he[5 * (x) + 0]
    for (int64_t a = 0; a < H->nhe;) {
        int64_t b = a;
        while (b < H->nhe && R[b].key == R[a].key) b++;
        const int64_t c = b - a;
        const int32_t e0 = R[a].he;
        if (c == 1) {
            TW(e0) = -1;
        } else if (c == 2) {
            const int32_t e1 = R[a + 1].he;
            TW(e0) = e1; TW(e1) = e0;
        } else if (c == 3) {
            const int32_t e1 = R[a + 1].he, e2 = R[a + 2].he;
            const int32_t v0 = OR(e0), v1 = OR(e1), v2 = OR(e2);
            FL(e0) = 2; FL(e1) = 2; FL(e2) = 2;
            if (v0 != v1) {
                *(int32_t*)vec_push(&B) = v0; *(int32_t*)vec_push(&B) = v1;
                TW(e0) = e1; TW(e1) = e0; TW(e2) = -1;
                H->he_v[v2] = e2;
            } else if (v0 != v2) {
                *(int32_t*)vec_push(&B) = v0; *(int32_t*)vec_push(&B) = v2;
                TW(e0) = e2; TW(e2) = e0; TW(e1) = -1;
                H->he_v[v1] = e1;
            } else {
                *(int32_t*)vec_push(&B) = v1; *(int32_t*)vec_push(&B) = v2;
                TW(e1) = e2; TW(e2) = e1; TW(e0) = -1;
                H->he_v[v0] = e0;
            }
            H->nonmanifold++;
        } else if (c == 4) {
            H->nonmanifold++;
            const int32_t e1 = R[a + 1].he, e2 = R[a + 2].he, e3 = R[a + 3].he;
            if (OR(e0) != OR(e1)) {
                TW(e0) = e1; TW(e1) = e0; TW(e2) = e3; TW(e3) = e2;
            } else {
                TW(e0) = e2; TW(e2) = e0; TW(e1) = e3; TW(e3) = e1;
            }
            FL(e0) = 1; FL(e1) = 1; FL(e2) = 1; FL(e3) = 1;
        } else {
            H->errors++;
        }
        a = b;
    }
#undef TW
#undef FL
#undef OR
    free(R);
    H->nbnd = B.n;
    H->bnd = B.p ? (int32_t*)B.p : (int32_t*)xmalloc(4);
    return 0;
}

/* DualMarchingCubes.cpp:673-703 */
static void neighbors(int64_t q, const int32_t* he, const int32_t* he_f, int32_t nb[4]) {
    const int32_t e0 = he_f[q];
    const int32_t e1 = he[5 * e0 + 2], e2 = he[5 * e1 + 2], e3 = he[5 * e2 + 2];
    const int32_t t[4] = {he[5 * e0 + 3], he[5 * e1 + 3], he[5 * e2 + 3], he[5 * e3 + 3]};
    for (int k = 0; k < 4; k++) nb[k] = (t[k] != -1) ? he[5 * t[k] + 1] : -1;
}

/* DualMarchingCubes.cpp:714-844 */
void dmco_color_faces(int64_t nq, const int32_t* he, const int32_t* he_f, int32_t* colors) {
    int32_t* initial = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)nq);
    memcpy(initial, colors, sizeof(int32_t) * (size_t)nq);
    for (int cls = 5; cls < 24; cls++) {
        for (int64_t f = 0; f < nq; f++) {
            if (initial[f] != cls) continue;
            int32_t nb[4], c[4];
            neighbors(f, he, he_f, nb);
            for (int k = 0; k < 4; k++) c[k] = (nb[k] != -1) ? colors[nb[k]] : -1;
            for (int col = 0; col < 5; col++) {
                if (c[0] != col && c[1] != col && c[2] != col && c[3] != col) { colors[f] = col; break; }
            }
        }
    }
    for (int64_t f = 0; f < nq; f++) {
        if (initial[f] != 4) continue;
        int32_t nb[4], c[4];
        neighbors(f, he, he_f, nb);
        for (int k = 0; k < 4; k++) c[k] = (nb[k] != -1) ? colors[nb[k]] : -1;
        for (int col = 0; col < 4; col++) {
            if (c[0] != col && c[1] != col && c[2] != col && c[3] != col) { colors[f] = col; break; }
        }
    }
    free(initial);
}

/* DualMarchingCubes.cpp:847-871; returns 1 when some vertex has valence 0 */
int dmco_vertex_valence(int64_t nv, int64_t nhe, const int32_t* he, int32_t* vV) {
    for (int64_t a = 0; a < nv; a++) vV[a] = 0;
    for (int64_t e = 0; e < nhe; e++) {
        vV[he[5 * e]]++;
        if (he[5 * e + 3] == -1) vV[he[5 * he[5 * e + 2]]]++;
    }
    int zero = 0;
    for (int64_t a = 0; a < nv; a++) zero |= (vV[a] == 0);
    return zero;
}

/* DualMarchingCubes.cpp:873-883 */
static int non_manifold(int64_t f, const int32_t* he, const int32_t* he_f) {
    const int32_t e0 = he_f[f];
    const int32_t e1 = he[5 * e0 + 2], e2 = he[5 * e1 + 2], e3 = he[5 * e2 + 2];
    return he[5 * e0 + 4] == 1 || he[5 * e1 + 4] == 1 || he[5 * e2 + 4] == 1 || he[5 * e3 + 4] == 1;
}

/* DualMarchingCubes.cpp:889-913 */
void dmco_mark3x3y(int64_t nq, const int32_t* quads, const int32_t* vV, uint8_t* p) {
    for (int64_t f = 0; f < nq; f++) {
        const int32_t a = vV[quads[4 * f]], b = vV[quads[4 * f + 1]], c = vV[quads[4 * f + 2]], d = vV[quads[4 * f + 3]];
        p[f] = (a == 3 && b >= 5 && c == 3 && d >= 5) || (a >= 5 && b == 3 && c >= 5 && d == 3);
    }
}

static void midpoint_normalized(double* dst, const double* a, const double* b) {
    for (int c = 0; c < 3; c++) dst[c] = a[c] + 0.5 * (b[c] - a[c]);
    const double s = sqrt(dst[0] * dst[0] + dst[1] * dst[1] + dst[2] * dst[2]);
    dst[0] /= s; dst[1] /= s; dst[2] /= s;
}

/* DualMarchingCubes.cpp:916-1012 (colors != NULL) and :1014-1094 (colors == NULL) */
void dmco_merge3x3y(int64_t nv, int64_t nq, double* v, double* n, const uint8_t* p,
                    const int32_t* vV, const int32_t* colors, const int32_t* he,
                    const int32_t* he_f, uint8_t* vm_first, int32_t* vm_second, uint8_t* em) {
    for (int64_t a = 0; a < nv; a++) { vm_first[a] = 0; vm_second[a] = -1; }
    for (int64_t f = 0; f < nq; f++) em[f] = 0;
    for (int64_t f = 0; f < nq; f++) {
        if (!p[f]) continue;
        const int32_t e0 = he_f[f];
        const int32_t e1 = he[5 * e0 + 2], e2 = he[5 * e1 + 2], e3 = he[5 * e2 + 2];
        int32_t nb[4];
        neighbors(f, he, he_f, nb);
        if (colors) {
            const int c = colors[f];
            int skip = 0;
            for (int k = 0; k < 4; k++) {
                if (nb[k] >= 0 && non_manifold(nb[k], he, he_f)) skip = 1;
                if (nb[k] < 0) skip = 1;
            }
            if (skip) continue;
            int flag = 1;
            for (int k = 0; k < 4; k++) {
                const int nc = (nb[k] != -1) ? colors[nb[k]] : -1;
                const int np = (nb[k] != -1) ? p[nb[k]] : 0;
                if (np && nc <= c) flag = 0;
            }
            if (!flag) continue;
        } else {
            int skip = 0;
            for (int k = 0; k < 4; k++)
                if (nb[k] != -1 && non_manifold(nb[k], he, he_f)) skip = 1;
            if (skip) continue;
            for (int k = 0; k < 4; k++)
                if (nb[k] != -1 && p[nb[k]]) skip = 1;
            if (skip) continue;
        }
        const int32_t v0 = he[5 * e0], v1 = he[5 * e1], v2 = he[5 * e2], v3 = he[5 * e3];
        const int32_t a = vV[v0], b = vV[v1], c = vV[v2], d = vV[v3];
        if (a == 3 && b >= 5 && c == 3 && d >= 5) {
            double t[3];
            for (int q = 0; q < 3; q++) t[q] = v[3 * v1 + q] + 0.5 * (v[3 * v3 + q] - v[3 * v1 + q]);
            memcpy(v + 3 * v0, t, sizeof(t));
            midpoint_normalized(t, n + 3 * v1, n + 3 * v3);
            memcpy(n + 3 * v0, t, sizeof(t));
            vm_first[v2] = 1;
            vm_second[v2] = v0;
            em[f] = 1;
        } else if (a >= 5 && b == 3 && c >= 5 && d == 3) {
            double t[3];
            for (int q = 0; q < 3; q++) t[q] = v[3 * v0 + q] + 0.5 * (v[3 * v2 + q] - v[3 * v0 + q]);
            memcpy(v + 3 * v1, t, sizeof(t));
            midpoint_normalized(t, n + 3 * v0, n + 3 * v2);
            memcpy(n + 3 * v1, t, sizeof(t));
            vm_first[v3] = 1;
            vm_second[v3] = v1;
            em[f] = 1;
        }
    }
}

/* removeVertices3X3Y (DualMarchingCubes.cpp:1096-1112) and
 * removeQuadrilaterals3X3Y (:1114-1137), applied to m in place. */
static void remove3x3y(dmco_mesh* m, uint8_t* vm_first, int32_t* vm_second, const uint8_t* em) {
    int64_t nn = 0;
    for (int64_t a = 0; a < m->nv; a++) {
        if (!vm_first[a]) {
            memmove(m->v + 3 * nn, m->v + 3 * a, 3 * sizeof(double));
            memmove(m->n + 3 * nn, m->n + 3 * a, 3 * sizeof(double));
            vm_second[a] = (int32_t)nn;
            nn++;
        }
    }
    m->nv = nn;
    int64_t nq = 0;
    for (int64_t f = 0; f < m->nq; f++) {
        if (em[f]) continue;
        int32_t r[4];
        for (int k = 0; k < 4; k++) {
            const int32_t x = m->quads[4 * f + k];
            r[k] = vm_first[x] ? vm_second[vm_second[x]] : vm_second[x];
        }
        memcpy(m->quads + 4 * nq, r, sizeof(r));
        nq++;
    }
    m->nq = nq;
}

/* One round of simplify3X3Y (DualMarchingCubes.cpp:1152-1158 colored,
 * :1172-1178 isolated). */
int dmco_simplify3x3y_round(dmco_mesh* m, int colored) {
    dmco_halfedges H;
    dmco_halfedges_build(m->nv, m->nq, m->quads, &H);
    if (colored) dmco_color_faces(m->nq, H.he, H.he_f, m->colors);
    int32_t* vV = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m->nv);
    dmco_vertex_valence(m->nv, H.nhe, H.he, vV);
    uint8_t* p = (uint8_t*)xmalloc((size_t)m->nq);
    dmco_mark3x3y(m->nq, m->quads, vV, p);
    uint8_t* vf = (uint8_t*)xmalloc((size_t)m->nv);
    int32_t* vs = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m->nv);
    uint8_t* em = (uint8_t*)xmalloc((size_t)m->nq);
    dmco_merge3x3y(m->nv, m->nq, m->v, m->n, p, vV, colored ? m->colors : NULL, H.he, H.he_f, vf, vs, em);
    remove3x3y(m, vf, vs, em);
    free(vV); free(p); free(vf); free(vs); free(em);
    dmco_halfedges_free(&H);
    return 0;
}

/* DualMarchingCubes.cpp:1139-1183 */
int dmco_simplify3x3y(dmco_mesh* m) {
    dmco_simplify3x3y_round(m, 1);
    dmco_simplify3x3y_round(m, 0);
    return 0;
}

/* DualMarchingCubes.cpp:1190-1301 */
void dmco_mark3333(int64_t nv, int64_t nq, const int32_t* he, const int32_t* he_f,
                   const int32_t* vV, uint8_t* vm_flag, int32_t* vm_twin,
                   uint8_t* p3333, uint8_t* rflag) {
    for (int64_t a = 0; a < nv; a++) { vm_flag[a] = 0; vm_twin[a] = -1; }
    for (int64_t f = 0; f < nq; f++) { p3333[f] = 0; rflag[f] = 0; }
    for (int64_t f = 0; f < nq; f++) {
        if (non_manifold(f, he, he_f)) continue;
        const int32_t e[4] = {he_f[f], he[5 * he_f[f] + 2], he[5 * he[5 * he_f[f] + 2] + 2],
                              he[5 * he[5 * he[5 * he_f[f] + 2] + 2] + 2]};
        int32_t vq[4], va[4];
        for (int k = 0; k < 4; k++) { vq[k] = he[5 * e[k]]; va[k] = vV[vq[k]]; }
        if (va[0] != 3 || va[1] != 3 || va[2] != 3 || va[3] != 3) continue;
        int32_t fn[4], vo[4];
        int boundary = 0;
        for (int k = 0; k < 4; k++) {
            const int32_t tw = he[5 * e[k] + 3];
            if (tw == -1) { boundary = 1; break; }
            fn[k] = he[5 * tw + 1];
            const int32_t nx1 = he[5 * tw + 2];
            const int32_t nx2 = he[5 * nx1 + 2];
            vo[k] = he[5 * nx2];
        }
        if (boundary) continue;
        int nm = 0;
        for (int k = 0; k < 4; k++) nm |= non_manifold(fn[k], he, he_f);
        if (nm) continue;
        const int32_t a4 = vV[vo[0]], a5 = vV[vo[1]], a6 = vV[vo[2]], a7 = vV[vo[3]];
        if ((a4 == 3 && a5 == 3) || (a5 == 3 && a6 == 3) || (a6 == 3 && a7 == 3) || (a7 == 3 && a4 == 3)) continue;
        if (a4 == 3 || a5 == 3 || a6 == 3 || a7 == 3) continue;
        if (a4 == 4 && a5 == 4 && a6 == 4 && a7 == 4) continue;
        for (int k = 0; k < 4; k++) { vm_flag[vq[k]] = 1; vm_twin[vq[k]] = vo[k]; }
        p3333[f] = 1;
        for (int k = 0; k < 4; k++) rflag[fn[k]] = 1;
    }
}

/* DualMarchingCubes.cpp:1362-1385 with :1303-1318 and :1320-1360 */
int dmco_simplify3333(dmco_mesh* m) {
    dmco_halfedges H;
    dmco_halfedges_build(m->nv, m->nq, m->quads, &H);
    int32_t* vV = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m->nv);
    dmco_vertex_valence(m->nv, H.nhe, H.he, vV);
    uint8_t* vf = (uint8_t*)xmalloc((size_t)m->nv);
    int32_t* vt = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m->nv);
    int32_t* vn = (int32_t*)xmalloc(sizeof(int32_t) * (size_t)m->nv);
    uint8_t* p = (uint8_t*)xmalloc((size_t)m->nq);
    uint8_t* r = (uint8_t*)xmalloc((size_t)m->nq);
    dmco_mark3333(m->nv, m->nq, H.he, H.he_f, vV, vf, vt, p, r);
    int64_t nn = 0;
    for (int64_t a = 0; a < m->nv; a++) {
        vn[a] = -1;
        if (!vf[a]) {
            vn[a] = (int32_t)nn;
            memmove(m->v + 3 * nn, m->v + 3 * a, 3 * sizeof(double));
            memmove(m->n + 3 * nn, m->n + 3 * a, 3 * sizeof(double));
            nn++;
        }
    }
    m->nv = nn;
    int64_t nq = 0;
    for (int64_t f = 0; f < m->nq; f++) {
        if (r[f]) continue;
        int32_t q[4];
        for (int k = 0; k < 4; k++) {
            const int32_t x = m->quads[4 * f + k];
            q[k] = p[f] ? vn[vt[x]] : vn[x];
        }
        memcpy(m->quads + 4 * nq, q, sizeof(q));
        nq++;
    }
    m->nq = nq;
    free(vV); free(vf); free(vt); free(vn); free(p); free(r);
    dmco_halfedges_free(&H);
    return 0;
}

/* Vector.h:153-156 */
static double vdistance(const double* a, const double* b) {
    const double d = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
    return sqrt(d);
}
static float fminf_ref(float a, float b) { return (b < a) ? b : a; }   /* std::min */

/* DualMarchingCubes.h:746-756 */
static double min_angle(const double* v0, const double* v1, const double* v2) {
    const float a = vdistance(v0, v1);
    const float b = vdistance(v1, v2);
    const float c = vdistance(v2, v0);
    const float A = acosf((b * b + c * c - a * a) / (2 * b * c));
    const float B = acosf((a * a + c * c - b * b) / (2 * a * c));
    const float C = acosf((b * b + a * a - c * c) / (2 * b * a));
    return fminf_ref(fminf_ref(A, B), C);
}

/* DualMarchingCubes.cpp:1387-1460 */
int dmco_triangles(dmco_mesh* m) {
    free(m->tris);
    m->nt = 2 * m->nq;
    m->tris = (int32_t*)xmalloc(sizeof(int32_t) * 3 * (size_t)m->nt);
    for (int64_t i = 0; i < m->nq; i++) {
        const int32_t v0 = m->quads[4 * i], v1 = m->quads[4 * i + 1], v2 = m->quads[4 * i + 2], v3 = m->quads[4 * i + 3];
        const double* V = m->v;
        double a1 = min_angle(V + 3 * v0, V + 3 * v1, V + 3 * v2);
        double a2 = min_angle(V + 3 * v0, V + 3 * v2, V + 3 * v3);
        const double b1 = (a2 < a1) ? a2 : a1, b2 = (a1 < a2) ? a2 : a1;
        a1 = min_angle(V + 3 * v1, V + 3 * v3, V + 3 * v0);
        a2 = min_angle(V + 3 * v1, V + 3 * v2, V + 3 * v3);
        const double c1 = (a2 < a1) ? a2 : a1, c2 = (a1 < a2) ? a2 : a1;
        int32_t* t = m->tris + 6 * i;
        if (b1 < c1 || (b1 == c1 && b2 <= c2)) {
            t[0] = v1; t[1] = v3; t[2] = v0; t[3] = v1; t[4] = v2; t[5] = v3;
        } else {
            t[0] = v0; t[1] = v1; t[2] = v2; t[3] = v0; t[4] = v2; t[5] = v3;
        }
    }
    return 0;
}

/* DualMarchingCubes.cpp:5-88.  flags: 1 = 3X3Y colored round, 2 = 3X3Y
 * isolated round, 4 = 3333 (sFlag == true is all three); 8 = triangles. */
int dmco_dualmc_flags(const dmco_grid* g, const double* s, const double* grad, double iso,
                      int flags, dmco_mesh* out, dmco_halfedges* he_out) {
    dmco_cells(g, s, grad, iso, out);
    if (flags & 1) dmco_simplify3x3y_round(out, 1);
    if (flags & 2) dmco_simplify3x3y_round(out, 0);
    if (flags & 4) dmco_simplify3333(out);
    dmco_halfedges H;
    dmco_halfedges_build(out->nv, out->nq, out->quads, &H);
    if (he_out) *he_out = H;
    else dmco_halfedges_free(&H);
    if (flags & 8) dmco_triangles(out);
    return 0;
}

int dmco_dualmc(const dmco_grid* g, const double* s, const double* grad, double iso,
                int sflag, dmco_mesh* out, dmco_halfedges* he_out) {
    return dmco_dualmc_flags(g, s, grad, iso, (sflag ? 7 : 0) | 8, out, he_out);
}

/* ------------------------------------------------------------------ */
/* single cell helpers                                                 */
/* ------------------------------------------------------------------ */
int dmco_cell_polygons(const double f[8], double iso, int32_t sizes[4], int32_t edges[12]) {
    unsigned icase = 0;
    for (int c = 0; c < 8; c++) icase += ((unsigned)(f[c] >= iso)) << c;
    for (int t = 0; t < 4; t++) sizes[t] = 0;
    for (int e = 0; e < 12; e++) edges[e] = -1;
    if (icase == 0 || icase == 255) return 0;
    polys P;
    cell_polygons((int)icase, iso, f, &P);
    int pos = 0;
    for (int t = 0; t < P.cnt; t++) {
        sizes[t] = P.size[t];
        for (int i = 0; i < P.size[t]; i++) edges[pos++] = P.e[t][i];
    }
    return P.cnt;
}

int dmco_cell_vertices(const double f[8], double iso, double* local, uint32_t* fail1, uint32_t* fail2) {
    unsigned icase = 0;
    for (int c = 0; c < 8; c++) icase += ((unsigned)(f[c] >= iso)) << c;
    if (icase == 0 || icase == 255) return 0;
    polys P;
    cell_polygons((int)icase, iso, f, &P);
    double loc[4][3];
    cell_representatives(iso, f, &P, loc, fail1, fail2);
    for (int t = 0; t < P.cnt; t++)
        for (int c = 0; c < 3; c++) local[3 * t + c] = loc[t][c];
    return P.cnt;
}
