/*
 * dmc_oracle.h -- CPU restatement of the reference Dual Marching Cubes path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * engine in tmc_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the engine never links or calls
 * it.
 *
 * It restates, sequentially and in plain C, the semantics of the reference
 * implementation cpp_dmc (DualMarchingCubes.cpp / DualMarchingCubes.h /
 * UniformGrid.{h,cpp}), including its float-rounding points.  Every function
 * cites the reference lines it follows.  Differences from the reference:
 *   - keys are 64-bit (the reference's int edge key overflows once a grid has
 *     more than 715,827,882 points; below that the order is identical);
 *   - the reference's OpenMP pragmas are ignored (its sequential build is
 *     the defined behaviour);
 *   - an edge shared by more than four faces is reported as an error and left
 *     unpaired (the reference writes past its 5-slot array there).
 *
 * Pinning: see tests/test_oracle.py -- every stage is compared with the
 * compiled reference (oracle/_ref) and with committed golden fixtures.
 */
#ifndef TMC_DMC_ORACLE_H
#define TMC_DMC_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nx, ny, nz;          /* grid points per axis */
    double x0, y0, z0;           /* position of point (0,0,0) */
    double dx, dy, dz;           /* spacing */
} dmco_grid;

/* A mesh as the reference holds it between stages.  All arrays are malloc'd
 * by the oracle and released with dmco_mesh_free. */
typedef struct {
    int64_t nv;
    double* v;          /* 3*nv positions */
    double* n;          /* 3*nv normals */
    int64_t nq;
    int32_t* quads;     /* 4*nq vertex ids */
    int32_t* colors;    /* nq face colors (valid after dmco_cells) */
    int64_t nt;
    int32_t* tris;      /* 3*nt */
    uint32_t nr_proj1;  /* failed projections, first level */
    uint32_t nr_proj2;  /* failed projections, second level */
    int64_t nv_before_unused;  /* vertices emitted by the cell sweep */
} dmco_mesh;

/* Halfedges: 5 ints per halfedge {origin, face, next, twin, flag}. */
typedef struct {
    int64_t nhe;
    int32_t* he;        /* 5*nhe */
    int32_t* he_v;      /* nv */
    int32_t* he_f;      /* nq */
    int64_t nbnd;
    int32_t* bnd;       /* boundary vertices of 3-face edges, in key order */
    int64_t nonmanifold;
    int64_t errors;     /* edges with more than 4 faces */
} dmco_halfedges;

void dmco_mesh_free(dmco_mesh* m);
void dmco_halfedges_free(dmco_halfedges* h);
void dmco_free(void* p);

/* UniformGrid::estimateGradient (+ flip_gradient when flip != 0). */
void dmco_estimate_gradient(const dmco_grid* g, const double* scalars, double* grad, int flip);

/* Cell sweep + collectQuadrilaterals + removeUnusedVertices. */
int dmco_cells(const dmco_grid* g, const double* scalars, const double* grad,
               double iso, dmco_mesh* out);

int dmco_halfedges_build(int64_t nv, int64_t nq, const int32_t* quads, dmco_halfedges* out);
void dmco_color_faces(int64_t nq, const int32_t* he, const int32_t* he_f, int32_t* colors);
int dmco_vertex_valence(int64_t nv, int64_t nhe, const int32_t* he, int32_t* valence);
void dmco_mark3x3y(int64_t nq, const int32_t* quads, const int32_t* valence, uint8_t* p3x3y);

/* mergeVertices3X3Y, colored (colors != NULL) or isolated (colors == NULL).
 * v/n are modified in place; vm_first/vm_second (nv) and em (nq) are outputs. */
void dmco_merge3x3y(int64_t nv, int64_t nq, double* v, double* n, const uint8_t* p3x3y,
                    const int32_t* valence, const int32_t* colors, const int32_t* he,
                    const int32_t* he_f, uint8_t* vm_first, int32_t* vm_second, uint8_t* em);

int dmco_simplify3x3y(dmco_mesh* m);
int dmco_simplify3x3y_round(dmco_mesh* m, int colored);
int dmco_simplify3333(dmco_mesh* m);
void dmco_mark3333(int64_t nv, int64_t nq, const int32_t* he, const int32_t* he_f,
                   const int32_t* valence, uint8_t* vm_flag, int32_t* vm_twin,
                   uint8_t* p3333, uint8_t* rflag);
int dmco_triangles(dmco_mesh* m);

/* Full reference dualMC: cells, (3X3Y color, 3X3Y isolated, 3333 if sflag),
 * final halfedges (returned in he_out when non-NULL), triangles. */
int dmco_dualmc(const dmco_grid* g, const double* scalars, const double* grad, double iso,
                int sflag, dmco_mesh* out, dmco_halfedges* he_out);
/* Same with the engine's flag bits: 1 = 3X3Y colored, 2 = 3X3Y isolated,
 * 4 = 3333, 8 = triangles. */
int dmco_dualmc_flags(const dmco_grid* g, const double* scalars, const double* grad, double iso,
                      int flags, dmco_mesh* out, dmco_halfedges* he_out);

/* Single-cell helpers used by the known-answer tests. */
int dmco_cell_polygons(const double f[8], double iso, int32_t sizes[4], int32_t edges[12]);
int dmco_cell_vertices(const double f[8], double iso, double* local /* 4*3 */,
                       uint32_t* fail1, uint32_t* fail2);

#ifdef __cplusplus
}
#endif
#endif
