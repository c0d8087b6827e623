/* tmc_oracle.c -- C restatement of the reference's topologically correct
 * Marching Cubes (mc/MarchingCubes.cpp: t_mc :150-318, p_slice :322-1303,
 * UniformGrid::gradient :1450-1512).  TEST INFRASTRUCTURE ONLY: the engine
 * never links it; tests and the benchmark's CPU leg use it as the checker.
 *
 * Pinned (tests/test_tmc_oracle.py) against the reference compiled from its
 * own sources (oracle/_ref/libtmcref.so): vertices, normals and triangles
 * bit for bit on fields rich in ambiguous cells, tunnels and 12-vertex
 * contours.
 *
 * Sequential over cells in the reference's loop order (k, j, i).  A vertex
 * on a grid edge is numbered when the first cell that cuts the edge meets it
 * (the reference's std::map keyed by 6 * point index + axis); here a dense
 * array of the same keys replaces the map. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "dmc_oracle.h"
#include "../tmc_amd/csrc/tmc_mc_tables.h"

static const uint64_t kTri[256] = TMC_CASE_TRIANGLES_INIT;
static const uint64_t kAmb[4] = TMC_AMBIGUOUS_MASK_INIT;

/* edges: end corners (corner c = i + 2 j + 4 k) and, for the global key,
 * the offset of the edge's first point from the cell's and its axis */
static const int kEdgeV[12][2] = {{0, 1}, {1, 3}, {2, 3}, {0, 2}, {4, 5}, {5, 7},
                                  {6, 7}, {4, 6}, {0, 4}, {1, 5}, {3, 7}, {2, 6}};
static const int kEdgeKey[12][4] = {{0, 0, 0, 0}, {1, 0, 0, 1}, {0, 1, 0, 0}, {0, 0, 0, 1},
                                    {0, 0, 1, 0}, {1, 0, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1},
                                    {0, 0, 0, 2}, {1, 0, 0, 2}, {1, 1, 0, 2}, {0, 1, 0, 2}};
/* faces seen from outside: corners and edges in the reference's oriented
 * order (MarchingCubes.h:84-107) */
static const int kFaceV[6][4] = {{0, 2, 1, 3}, {5, 7, 4, 6}, {4, 0, 5, 1}, {2, 6, 3, 7}, {4, 6, 0, 2}, {1, 3, 5, 7}};
static const int kFaceE[6][4] = {{3, 2, 1, 0}, {5, 6, 7, 4}, {8, 0, 9, 4}, {11, 6, 10, 2}, {7, 11, 3, 8}, {1, 10, 5, 9}};
/* local coordinate of each edge's points per axis: 0, 1, or 3 = along the edge */
static const unsigned kLCoord[3] = {1324855u, 5299420u, 16733440u};

typedef struct {
    int nt;                 /* triangles, local refs: 0..11 edge, 12..17 inner */
    unsigned char t[64][3];
    int ninner;             /* inner vertices (0, 1 or 6) */
    double ip[6][3], in[6][3];
} cell_out;

static void push_tri(cell_out* o, int a, int b, int c) {
    o->t[o->nt][0] = (unsigned char)a;
    o->t[o->nt][1] = (unsigned char)b;
    o->t[o->nt][2] = (unsigned char)c;
    o->nt++;
}

/* trilinear blend in the reference's operation order */
static double tri_blend(const double (*q)[3], int x, double u, double v, double w) {
    return (1 - w) * ((1 - v) * (q[0][x] + u * (q[1][x] - q[0][x])) + v * (q[2][x] + u * (q[3][x] - q[2][x]))) +
           w * ((1 - v) * (q[4][x] + u * (q[5][x] - q[4][x])) + v * (q[6][x] + u * (q[7][x] - q[6][x])));
}

static void inner_vertex(const double P[8][3], const double N[8][3], double u, double v, double w, double ip[3],
                         double in[3]) {
    for (int x = 0; x < 3; x++) ip[x] = tri_blend(P, x, u, v, w);
    for (int x = 0; x < 3; x++) in[x] = tri_blend(N, x, u, v, w);
    const double f = sqrt(in[0] * in[0] + in[1] * in[1] + in[2] * in[2]);
    in[0] = in[0] / f;
    in[1] = in[1] / f;
    in[2] = in[2] / f;
}

static double decider(double f0, double f1, double f2, double f3) {
    return (f0 * f3 - f1 * f2) / (f0 + f3 - f1 - f2);
}

/* p_slice (mc/MarchingCubes.cpp:322-1303): contours of an ambiguous cell
 * from the face deciders, then triangles with inner vertices where the
 * cell's interior has saddle points.  Returns 0 when a singular face cannot
 * be triangulated (the reference then emits no triangles for the cell). */
static int slice_cell(const double F[8], const double P[8][3], const double N[8][3], double i0, int icase,
                      cell_out* o) {
    double ecoord[12];
    for (int e = 0; e < 12; e++) {
        ecoord[e] = 0;
        const int a = kEdgeV[e][0], b = kEdgeV[e][1];
        if (((icase >> a) ^ (icase >> b)) & 1) ecoord[e] = (i0 - F[a]) / (F[b] - F[a]);
    }
    /* oriented segments: seg_to[e] = next edge, seg_in[e] = previous */
    int seg_to[12], seg_in[12];
    for (int e = 0; e < 12; e++) seg_to[e] = seg_in[e] = -1;
#define SEG(a, b) (seg_to[a] = (b), seg_in[b] = (a))
    int f_flag[6] = {0, 0, 0, 0, 0, 0};
    for (int f = 0; f < 6; f++) {
        const double f0 = F[kFaceV[f][0]], f1 = F[kFaceV[f][1]], f2 = F[kFaceV[f][2]], f3 = F[kFaceV[f][3]];
        const int e0 = kFaceE[f][0], e1 = kFaceE[f][1], e2 = kFaceE[f][2], e3 = kFaceE[f][3];
        const int fc = (f0 >= i0) | (f1 >= i0) << 1 | (f2 >= i0) << 2 | (f3 >= i0) << 3;
        switch (fc) {
        case 1: SEG(e0, e3); break;
        case 2: SEG(e1, e0); break;
        case 3: SEG(e1, e3); break;
        case 4: SEG(e3, e2); break;
        case 5: SEG(e0, e2); break;
        case 6:
        case 9: {
            const double val = decider(f0, f1, f2, f3);
            int sep;   /* 1: the "val > i0" pairing, 0: the "val < i0" one */
            if (val > i0) sep = 1;
            else if (val < i0) sep = 0;
            else {
                f_flag[f] = 1;
                const unsigned short e_flag = 0x218;
                double c0 = ecoord[e0], c1 = ecoord[e1], c2 = ecoord[e2], c3 = ecoord[e3];
                if ((e_flag >> (f * 2)) & 1) { c0 = 1 - c0; c2 = 1 - c2; }
                if ((e_flag >> (f * 2)) & 2) { c1 = 1 - c1; c3 = 1 - c3; }
                if (c1 < c3 && c0 > c2) sep = (fc == 6) ? 0 : 1;
                else if (c1 > c3 && c0 < c2) sep = (fc == 6) ? 1 : 0;
                else return 0;
            }
            if (fc == 6) {
                if (sep) { SEG(e3, e0); SEG(e1, e2); }
                else { SEG(e1, e0); SEG(e3, e2); }
            } else {
                if (sep) { SEG(e0, e1); SEG(e2, e3); }
                else { SEG(e0, e3); SEG(e2, e1); }
            }
        } break;
        case 7: SEG(e1, e2); break;
        case 8: SEG(e2, e1); break;
        case 10: SEG(e2, e0); break;
        case 11: SEG(e2, e3); break;
        case 12: SEG(e3, e1); break;
        case 13: SEG(e0, e1); break;
        case 14: SEG(e3, e0); break;
        default: break;
        }
    }
#undef SEG
    /* closed contours, each started at its smallest edge (the start edge
     * stays set; the scan has moved past it) */
    int cnt = 0, csz[4] = {0, 0, 0, 0}, cedge[4][12];
    for (int e = 0; e < 12; e++) {
        if (seg_to[e] < 0) continue;
        int pos = 0, to = seg_to[e];
        cedge[cnt][pos] = e;
        while (to != e) {
            pos++;
            cedge[cnt][pos] = to;
            const int nx = seg_to[to];
            seg_to[to] = -1;
            to = nx;
        }
        csz[cnt] = pos + 1;
        cnt++;
    }

    /* saddle points of the interior: pairs of intersections of opposite
     * faces' hyperbolas */
    double ui[2] = {0, 0}, vi[2] = {0, 0}, wi[2] = {0, 0};
    unsigned q_sol = 0;
    const double a = (F[0] - F[1]) * (-F[6] + F[7] + F[4] - F[5]) - (F[4] - F[5]) * (-F[2] + F[3] + F[0] - F[1]);
    const double b = (i0 - F[0]) * (-F[6] + F[7] + F[4] - F[5]) + (F[0] - F[1]) * (F[6] - F[4]) -
                     (i0 - F[4]) * (-F[2] + F[3] + F[0] - F[1]) - (F[4] - F[5]) * (F[2] - F[0]);
    const double c = (i0 - F[0]) * (F[6] - F[4]) - (i0 - F[4]) * (F[2] - F[0]);
    double d = b * b - 4 * a * c;
    if (d > 0) {
        d = sqrt(d);
        ui[0] = (-b - d) / (2 * a);
        ui[1] = (-b + d) / (2 * a);
        for (int s = 0; s < 2; s++) {
            double g1 = F[0] * (1 - ui[s]) + F[1] * ui[s];
            double g2 = F[2] * (1 - ui[s]) + F[3] * ui[s];
            vi[s] = (i0 - g1) / (g2 - g1);
            if (isnan(vi[s]) || isinf(vi[s])) vi[s] = -1.f;
        }
        for (int s = 0; s < 2; s++) {
            double g1 = F[0] * (1 - ui[s]) + F[1] * ui[s];
            double g2 = F[4] * (1 - ui[s]) + F[5] * ui[s];
            wi[s] = (i0 - g1) / (g2 - g1);
            if (isnan(wi[s]) || isinf(wi[s])) wi[s] = -1.f;
        }
        if (f_flag[0]) { if (wi[0] < wi[1]) wi[0] = 0; else wi[1] = 0; }
        if (f_flag[1]) { wi[1] = 1; }   /* both branches of the reference set wi[1] */
        if (f_flag[2]) { if (vi[0] < vi[1]) vi[0] = 0; else vi[1] = 0; }
        if (f_flag[3]) { if (vi[0] > vi[1]) vi[0] = 1; else vi[1] = 1; }
        if (f_flag[4]) { if (ui[0] < ui[1]) ui[0] = 0; else ui[1] = 0; }
        if (f_flag[5]) { if (ui[0] > ui[1]) ui[0] = 1; else ui[1] = 1; }
        if (0 < ui[0] && ui[0] < 1) q_sol |= 1;
        if (0 < ui[1] && ui[1] < 1) q_sol |= 2;
        if (0 < vi[0] && vi[0] < 1) q_sol |= 4;
        if (0 < vi[1] && vi[1] < 1) q_sol |= 8;
        if (0 < wi[0] && wi[0] < 1) q_sol |= 16;
        if (0 < wi[1] && wi[1] < 1) q_sol |= 32;
    }
#define EV(e, x) ((((kLCoord[x] >> (2 * (e))) & 3u) == 3u) ? ecoord[e] : (double)((kLCoord[x] >> (2 * (e))) & 3u))
    if (__builtin_popcount(q_sol) == 6) {
        /* tunnel or a single 12-vertex contour: inner hexagon */
        double hvt[6][3] = {{ui[0], vi[0], wi[0]}, {ui[0], vi[0], wi[1]}, {ui[1], vi[0], wi[1]},
                            {ui[1], vi[1], wi[1]}, {ui[1], vi[1], wi[0]}, {ui[0], vi[1], wi[0]}};
        int not_tunnel = 0xF;
        if (cnt == 3) {
            const double uc_min = (ui[0] < ui[1]) ? ui[0] : ui[1];
            const double uc_max = (ui[0] < ui[1]) ? ui[1] : ui[0];
            for (int t = 0; t < cnt; t++) {
                if (csz[t] != 3) continue;
                double umin = 2, umax = -2;
                const int x0 = cedge[t][0], x1 = cedge[t][1], x2 = cedge[t][2];
                const double u0 = EV(x0, 0), u1 = EV(x1, 0), u2 = EV(x2, 0);
                umin = (u0 < umin) ? u0 : umin;
                umin = (u1 < umin) ? u1 : umin;
                umin = (u2 < umin) ? u2 : umin;
                umax = (u0 > umax) ? u0 : umax;
                umax = (u1 > umax) ? u1 : umax;
                umax = (u2 > umax) ? u1 : umax;   /* the reference takes u1 here */
                if (uc_min > umax || uc_max < umin) {
                    not_tunnel = t;
                    push_tri(o, x0, x1, x2);
                }
            }
        }
        o->ninner = 6;
        for (int s = 0; s < 6; s++) inner_vertex(P, N, hvt[s][0], hvt[s][1], hvt[s][2], o->ip[s], o->in[s]);
        unsigned char tcon[12];
        for (int t = 0; t < cnt; t++) {
            if (not_tunnel == t) continue;
            const int sz = csz[t];
            for (int r = 0; r < sz; r++) {
                int index = -1;
                double dist = 1000.;
                const int ci = cedge[t][r];
                const double ue = EV(ci, 0), ve = EV(ci, 1), we = EV(ci, 2);
                for (int s = 0; s < 6; s++) {
                    const double du = ue - hvt[s][0], dv = ve - hvt[s][1], dw = we - hvt[s][2];
                    const double val = du * du + dv * dv + dw * dw;
                    if (dist > val) {
                        index = s;
                        dist = val;
                    }
                }
                tcon[ci] = (unsigned char)index;
            }
            for (int r = 0; r < sz; r++) {
                const int t1 = cedge[t][r], t2 = cedge[t][(r + 1) % sz];
                const int c1 = tcon[t1], c2 = tcon[t2];
                int dst = c1 - c2 < 0 ? c2 - c1 : c1 - c2;
                if (dst > 2) dst = 6 - dst;
                if (dst == 0) {
                    push_tri(o, t1, t2, 12 + c1);
                } else if (dst == 1) {
                    const double l1 = (EV(t1, 0) - hvt[c2][0]) * (EV(t1, 0) - hvt[c2][0]) +
                                      (EV(t1, 1) - hvt[c2][1]) * (EV(t1, 1) - hvt[c2][1]) +
                                      (EV(t1, 2) - hvt[c2][2]) * (EV(t1, 2) - hvt[c2][2]);
                    const double l2 = (EV(t2, 0) - hvt[c1][0]) * (EV(t2, 0) - hvt[c1][0]) +
                                      (EV(t2, 1) - hvt[c1][1]) * (EV(t2, 1) - hvt[c1][1]) +
                                      (EV(t2, 2) - hvt[c1][2]) * (EV(t2, 2) - hvt[c1][2]);
                    if (l1 < l2) {
                        push_tri(o, t1, t2, 12 + c2);
                        push_tri(o, t1, 12 + c2, 12 + c1);
                    } else {
                        push_tri(o, t1, t2, 12 + c1);
                        push_tri(o, t2, 12 + c2, 12 + c1);
                    }
                } else if (dst == 2) {
                    const int dmax = c1 > c2 ? c1 : c2, dmin = c1 < c2 ? c1 : c2;
                    const int cm = ((dmax + 2) % 6 == dmin) ? (dmax + 1) % 6 : (dmax + dmin) / 2;
                    push_tri(o, t1, t2, 12 + cm);
                    push_tri(o, t1, 12 + cm, 12 + c1);
                    push_tri(o, t2, 12 + c2, 12 + cm);
                }
            }
        }
        if (cnt == 1) {
            const int s_ = decider(F[0], F[1], F[2], F[3]) <= i0;
            const int of_ = (wi[1] < wi[0]) ? s_ : !s_;
            if (!of_) {
                push_tri(o, 12, 14, 13);
                push_tri(o, 14, 16, 15);
                push_tri(o, 12, 17, 16);
                push_tri(o, 12, 16, 14);
            } else {
                push_tri(o, 12, 13, 14);
                push_tri(o, 14, 15, 16);
                push_tri(o, 12, 16, 17);
                push_tri(o, 12, 14, 16);
            }
        }
        return 1;
    }
    /* saddle points counted per pair of opposite faces */
    const unsigned nr_u = ((q_sol & 0x5) == 0x5) + ((q_sol & 0xA) == 0xA);
    const unsigned nr_v = ((q_sol & 0x11) == 0x11) + ((q_sol & 0x22) == 0x22);
    const unsigned nr_w = ((q_sol & 0x18) == 0x18) + ((q_sol & 0x24) == 0x24);
    const unsigned nr_t = nr_u + nr_v + nr_w;
    if (nr_t == nr_u || nr_t == nr_v || nr_t == nr_w) {
        /* no saddle point inside: fans of the contours (sizes 3..6) */
        for (int t = 0; t < cnt; t++) {
            const int* ce = cedge[t];
            switch (csz[t]) {
            case 3: push_tri(o, ce[0], ce[1], ce[2]); break;
            case 4:
                push_tri(o, ce[0], ce[1], ce[2]);
                push_tri(o, ce[0], ce[2], ce[3]);
                break;
            case 5:
                push_tri(o, ce[0], ce[1], ce[2]);
                push_tri(o, ce[0], ce[2], ce[3]);
                push_tri(o, ce[0], ce[3], ce[4]);
                break;
            case 6:
                push_tri(o, ce[0], ce[1], ce[3]);
                push_tri(o, ce[1], ce[2], ce[3]);
                push_tri(o, ce[0], ce[3], ce[4]);
                push_tri(o, ce[0], ce[4], ce[5]);
                break;
            default: break;
            }
        }
        return 1;
    }
    /* one inner vertex joined to the contours with more than three edges */
    const double fs[3][2] = {{(double)(q_sol & 1), (double)((q_sol >> 1) & 1)},
                             {(double)((q_sol >> 2) & 1), (double)((q_sol >> 3) & 1)},
                             {(double)((q_sol >> 4) & 1), (double)((q_sol >> 5) & 1)}};
    const unsigned char ufs[3][2] = {{q_sol & 1, (q_sol >> 1) & 1}, {(q_sol >> 2) & 1, (q_sol >> 3) & 1},
                                     {(q_sol >> 4) & 1, (q_sol >> 5) & 1}};
    const int fc1 = ufs[0][0] * ufs[1][0] + ufs[0][1] * ufs[1][1];
    const int fc2 = ufs[0][0] * ufs[2][0] + ufs[0][1] * ufs[2][1];
    const int fc3 = ufs[1][0] * ufs[2][1] + ufs[1][1] * ufs[2][0];
    const int c_faces = fc1 + fc2 + fc3;
    double uc = 0, vc = 0, wc = 0;
    switch (c_faces) {
    case 2:
        if (fc1 == 0) {
            uc = fs[0][0] * ui[0] + fs[0][1] * ui[1];
            vc = fs[1][0] * vi[0] + fs[1][1] * vi[1];
            wc = fs[1][0] * wi[1] + fs[1][1] * wi[0];
        } else if (fc2 == 0) {
            uc = fs[0][0] * ui[0] + fs[0][1] * ui[1];
            vc = fs[0][0] * vi[0] + fs[0][1] * vi[1];
            wc = fs[0][0] * wi[1] + fs[0][1] * wi[0];
        } else if (fc3 == 0) {
            uc = fs[1][0] * ui[0] + fs[1][1] * ui[1];
            vc = fs[1][0] * vi[0] + fs[1][1] * vi[1];
            wc = fs[1][0] * wi[0] + fs[1][1] * wi[1];
        }
        break;
    case 3:
        uc = (fs[0][0] * ui[0] + fs[0][1] * ui[1]) / (double)(ufs[0][0] + ufs[0][1]);
        vc = (fs[1][0] * vi[0] + fs[1][1] * vi[1]) / (double)(ufs[1][0] + ufs[1][1]);
        wc = (fs[2][0] * wi[0] + fs[2][1] * wi[1]) / (double)(ufs[2][0] + ufs[2][1]);
        break;
    case 4: {
        const int n_u = ufs[0][0] + ufs[0][1], n_v = ufs[1][0] + ufs[1][1], n_w = ufs[2][0] + ufs[2][1];
        if (n_w == 1) {
            uc = fs[2][0] * ui[0] + fs[2][1] * ui[1];
            vc = fs[2][1] * vi[0] + fs[2][0] * vi[1];
            wc = fs[2][0] * wi[0] + fs[2][1] * wi[1];
        } else if (n_v == 1) {
            uc = fs[1][0] * ui[0] + fs[1][1] * ui[1];
            vc = fs[1][0] * vi[0] + fs[1][1] * vi[1];
            wc = fs[1][1] * wi[0] + fs[1][0] * wi[1];
        } else if (n_u == 1) {
            uc = fs[0][0] * ui[0] + fs[0][1] * ui[1];
            vc = fs[0][0] * vi[0] + fs[0][1] * vi[1];
            wc = fs[0][0] * wi[0] + fs[0][1] * wi[1];
        }
    } break;
    default: break;
    }
    int used = 0;
    for (int t = 0; t < cnt; t++) {
        const int sz = csz[t];
        if (sz == 3) {
            push_tri(o, cedge[t][0], cedge[t][1], cedge[t][2]);
        } else {
            used = 1;
            for (int r = 0; r < sz; r++) push_tri(o, cedge[t][r], cedge[t][(r + 1) % sz], 12);
        }
    }
    if (used) {
        o->ninner = 1;
        inner_vertex(P, N, uc, vc, wc, o->ip[0], o->in[0]);
    }
    return 1;
#undef EV
}

/* central differences inside, one-sided at the border
 * (mc/MarchingCubes.cpp:1450-1512) */
static void gradient(const dmco_grid* g, const double* s, double* n) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#define GI(i, j, k) (((int64_t)(k) * ny + (j)) * nx + (i))
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                double fx, fy, fz;
                if (i == 0) fx = (s[GI(i + 1, j, k)] - s[GI(i, j, k)]) / g->dx;
                else if (i == nx - 1) fx = (s[GI(i, j, k)] - s[GI(i - 1, j, k)]) / g->dx;
                else fx = (s[GI(i + 1, j, k)] - s[GI(i - 1, j, k)]) / (2 * g->dx);
                if (j == 0) fy = (s[GI(i, j + 1, k)] - s[GI(i, j, k)]) / g->dy;
                else if (j == ny - 1) fy = (s[GI(i, j, k)] - s[GI(i, j - 1, k)]) / g->dy;
                else fy = (s[GI(i, j + 1, k)] - s[GI(i, j - 1, k)]) / (2 * g->dy);
                if (k == 0) fz = (s[GI(i, j, k + 1)] - s[GI(i, j, k)]) / g->dz;
                else if (k == nz - 1) fz = (s[GI(i, j, k)] - s[GI(i, j, k - 1)]) / g->dz;
                else fz = (s[GI(i, j, k + 1)] - s[GI(i, j, k - 1)]) / (2 * g->dz);
                double* o = n + 3 * GI(i, j, k);
                o[0] = fx;
                o[1] = fy;
                o[2] = fz;
            }
#undef GI
}

typedef struct {
    int64_t nv, nt, cap_v, cap_t;
    double *v, *n;
    int32_t* t;
} tmc_lists;

static void add_vertex(tmc_lists* L, const double p[3], const double q[3]) {
    if (L->nv == L->cap_v) {
        L->cap_v = L->cap_v ? 2 * L->cap_v : 1024;
        L->v = (double*)realloc(L->v, 3 * L->cap_v * sizeof(double));
        L->n = (double*)realloc(L->n, 3 * L->cap_v * sizeof(double));
    }
    memcpy(L->v + 3 * L->nv, p, 3 * sizeof(double));
    memcpy(L->n + 3 * L->nv, q, 3 * sizeof(double));
    L->nv++;
}

static void add_tri(tmc_lists* L, int32_t a, int32_t b, int32_t c) {
    if (L->nt == L->cap_t) {
        L->cap_t = L->cap_t ? 2 * L->cap_t : 1024;
        L->t = (int32_t*)realloc(L->t, 3 * L->cap_t * sizeof(int32_t));
    }
    L->t[3 * L->nt] = a;
    L->t[3 * L->nt + 1] = b;
    L->t[3 * L->nt + 2] = c;
    L->nt++;
}

/* t_mc (mc/MarchingCubes.cpp:150-318) on a field of nx*ny*nz doubles, x
 * fastest; point (i, j, k) sits at (x0 + i dx, y0 + j dy, z0 + k dz).
 * Returns malloc'd positions and normals (3 per vertex) and triangles (3 ids
 * each); stats[0] = ambiguous cells, stats[1] = cells whose singular face
 * could not be triangulated. */
int tmco_tmc(const dmco_grid* g, const double* scalars, double i0, int64_t* nv, double** v, double** n,
             int64_t* nt, int32_t** tris, int64_t* stats) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int64_t npts = (int64_t)nx * ny * nz;
    double* grad = (double*)malloc(3 * npts * sizeof(double));
    int32_t* vid = (int32_t*)malloc(6 * npts * sizeof(int32_t));   /* vertex of key 6 * point + axis */
    if (!grad || !vid) {
        free(grad);
        free(vid);
        return -1;
    }
    gradient(g, scalars, grad);
    for (int64_t x = 0; x < 6 * npts; x++) vid[x] = -1;
    tmc_lists L = {0, 0, 0, 0, NULL, NULL, NULL};
    stats[0] = stats[1] = 0;
    for (int k = 0; k < nz - 1; k++)
        for (int j = 0; j < ny - 1; j++)
            for (int i = 0; i < nx - 1; i++) {
                double F[8], P[8][3], N[8][3];
                int icase = 0;
                for (int c = 0; c < 8; c++) {
                    const int ci = i + (c & 1), cj = j + ((c >> 1) & 1), ck = k + (c >> 2);
                    const int64_t gi = ((int64_t)ck * ny + cj) * nx + ci;
                    F[c] = scalars[gi];
                    P[c][0] = g->x0 + ci * g->dx;
                    P[c][1] = g->y0 + cj * g->dy;
                    P[c][2] = g->z0 + ck * g->dz;
                    memcpy(N[c], grad + 3 * gi, 3 * sizeof(double));
                    if (F[c] >= i0) icase |= 1 << c;
                }
                if (icase == 0 || icase == 255) continue;
                /* vertices of the cut edges, numbered at first sight */
                int32_t ev[12];
                for (int e = 0; e < 12; e++) {
                    ev[e] = -1;
                    const int a = kEdgeV[e][0], b = kEdgeV[e][1];
                    if (!(((icase >> a) ^ (icase >> b)) & 1)) continue;
                    const int64_t key = 6 * (((int64_t)(k + kEdgeKey[e][2]) * ny + (j + kEdgeKey[e][1])) * nx +
                                             (i + kEdgeKey[e][0])) + kEdgeKey[e][3];
                    if (vid[key] < 0) {
                        const double l = (i0 - F[a]) / (F[b] - F[a]);
                        double p[3], q[3];
                        for (int x = 0; x < 3; x++) p[x] = (1 - l) * P[a][x] + l * P[b][x];
                        for (int x = 0; x < 3; x++) q[x] = (1 - l) * N[a][x] + l * N[b][x];
                        const double s = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
                        q[0] = q[0] / s;
                        q[1] = q[1] / s;
                        q[2] = q[2] / s;
                        vid[key] = (int32_t)L.nv;
                        add_vertex(&L, p, q);
                    }
                    ev[e] = vid[key];
                }
                if ((kAmb[icase >> 6] >> (icase & 63)) & 1) {
                    stats[0]++;
                    cell_out o;
                    o.nt = 0;
                    o.ninner = 0;
                    if (!slice_cell(F, P, N, i0, icase, &o)) {
                        stats[1]++;
                        continue;
                    }
                    int32_t base = (int32_t)L.nv;
                    for (int s = 0; s < o.ninner; s++) add_vertex(&L, o.ip[s], o.in[s]);
                    for (int t = 0; t < o.nt; t++) {
                        int32_t id[3];
                        for (int x = 0; x < 3; x++) {
                            const int r = o.t[t][x];
                            id[x] = r < 12 ? ev[r] : base + (r - 12);
                        }
                        add_tri(&L, id[0], id[1], id[2]);
                    }
                } else {
                    const uint64_t w = kTri[icase];
                    const int ntri = (int)(w >> 60);
                    for (int t = 0; t < ntri; t++) {
                        const int e0 = (int)((w >> (12 * t)) & 0xF), e1 = (int)((w >> (12 * t + 4)) & 0xF),
                                  e2 = (int)((w >> (12 * t + 8)) & 0xF);
                        add_tri(&L, ev[e2], ev[e1], ev[e0]);   /* reversed: positive corners outside */
                    }
                }
            }
    free(grad);
    free(vid);
    *nv = L.nv;
    *v = L.v;
    *n = L.n;
    *nt = L.nt;
    *tris = L.t;
    return 0;
}
