/*
 * mvoro_oracle.h -- CPU oracle for the per-cell Voronoi construction path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the algorithm
 * of the meshless_voronoi crate (v0.7.2), used by the tests as the checker
 * and by bench.py as the timed CPU baseline.  The product library (libmvoro)
 * never links, loads or calls it.
 *
 * Function-by-function citations (reference file:line) are in
 * mvoro_oracle.c.
 */
#ifndef MVORO_ORACLE_H
#define MVORO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: one per panic site of the reference, same numbering as
 * include/mvoro.h */
enum {
  ORC_OK = 0,
  ORC_E_ARG = 1,
  ORC_E_DEGENERATE_POINTS = 2, /* convex_cell.rs:342 */
  ORC_E_SELF_NOT_FIRST = 3,    /* convex_cell.rs:325-329 */
  ORC_E_PLANE_DET_ZERO = 4,    /* geometry.rs:70 */
  ORC_E_NAN = 5,               /* convex_cell.rs:448 */
  ORC_E_BOUNDARY_WALK = 6,     /* convex_cell.rs:428 */
  ORC_E_OOM = 8,
  ORC_E_FACE_WALK = 10,        /* convex_cell.rs:538-541 */
  ORC_E_NOT_3D = 11            /* convex_cell.rs:462-466 */
};

typedef struct orc_voronoi {
  int status;
  int64_t status_cell;
  double anchor[3], width[3];
  int dim, periodic;
  size_t n_cells, n_faces, n_links;
  /* cells (SoA) -- VoronoiCell, voronoi_cell.rs:12-20 */
  double *cell_loc;           /* 3N */
  double *cell_centroid;      /* 3N */
  double *cell_volume;        /* N */
  double *cell_safety_radius; /* N */
  uint64_t *cell_face_offset; /* N */
  uint64_t *cell_face_count;  /* N */
  uint64_t *cell_idx;         /* N (0 for inactive cells, as Default) */
  /* faces (SoA) -- VoronoiFace, voronoi_face.rs:11-15, integrals.rs:205-213 */
  uint64_t *face_left;
  int64_t *face_right; /* -1 = None */
  double *face_area;
  double *face_centroid; /* 3F */
  double *face_normal;   /* 3F */
  uint8_t *face_has_shift;
  double *face_shift; /* 3F (zero when has_shift == 0) */
  uint64_t *links;    /* cell_face_connections */
  /* per-cell statistics of the construction (not in the reference) */
  uint32_t *cell_n_popped; /* neighbour-stream entries consumed, incl. self
                              and the entry that triggered the stop */
  uint32_t *cell_n_planes; /* clipping planes at the end (6 box + accepted) */
  uint32_t *cell_n_vertices;
  uint8_t *cell_exhausted; /* stream ran out before the security test */
  uint64_t flops;          /* FP64 add/sub/mul/div/sqrt (counting build) */
  uint64_t exact_tests;
} orc_voronoi;

/* Voronoi::build / build_partial (voronoi.rs:151-190).
 * xyz: N x 3 doubles (AoS).  mask: NULL = all active.  dim: 1, 2 or 3.
 * nthreads <= 0: all hardware threads. */
int orc_build(const double *xyz, size_t n, const uint8_t *mask,
              const double anchor[3], const double width[3], int dim,
              int periodic, int nthreads, orc_voronoi **out);
void orc_free(orc_voronoi *v);

/* decomposition export (VoronoiIntegrator path, voronoi.rs:549-618,
 * convex_cell.rs:131-187): for every active cell, the oriented tets
 * (v0, v1, v2, plane index) in decomposition order, plus each cell's
 * clipping planes.  Arrays are owned by the returned object. */
typedef struct orc_tets {
  int status;
  size_t n_cells;
  uint64_t *tet_offset; /* N+1 */
  double *tet_v;        /* 9 per tet: v0, v1, v2 */
  int32_t *tet_plane;   /* per tet */
  uint64_t *plane_offset; /* N+1 */
  double *plane_n;        /* 3 per plane */
  double *plane_p;        /* 3 per plane */
  int64_t *plane_right;   /* -1 = None */
  uint8_t *plane_has_shift;
  double *plane_shift;    /* 3 per plane */
  uint64_t *vert_offset;  /* N+1 */
  double *vert_loc;       /* 3 per vertex */
  int32_t *vert_dual;     /* 3 per vertex */
} orc_tets;

int orc_build_tets(const double *xyz, size_t n, const uint8_t *mask,
                   const double anchor[3], const double width[3], int dim,
                   int periodic, int nthreads, orc_tets **out);
void orc_free_tets(orc_tets *t);

/* ConvexCell::with_faces (convex_cell.rs:461-505, sort_face_vertices
 * :508-543) for every cell of a decomposition export, and the fan
 * decomposition that replaces it (DecompositionWithFaces :89-129).
 * Faces of cell i: face_offset[i] .. face_offset[i+1], in half-space order
 * (half-spaces holding at least one vertex); face f has face_plane[f],
 * face_count[f] vertices at conn[conn_offset[i] + face_vertex_offset[f]]
 * (vertex indices within the cell, the reference's vertex_offset being
 * relative to the cell).  Fan tets like orc_tets' tet arrays.  dim must be
 * 3 (ORC_E_NOT_3D otherwise, where the reference panics). */
typedef struct orc_faces {
  int status;
  size_t n_cells;
  uint64_t *face_offset;        /* N+1 */
  int32_t *face_plane;
  uint32_t *face_count;
  uint64_t *face_vertex_offset; /* within the cell's connection list */
  uint64_t *conn_offset;        /* N+1 */
  uint32_t *conn;
  uint64_t *tet_offset;         /* N+1 */
  double *tet_v;                /* 9 per tet */
  int32_t *tet_plane;
} orc_faces;

int orc_with_faces(const orc_tets *t, int dim, orc_faces **out);
void orc_free_faces(orc_faces *f);

/* primitives exported for unit tests */
int orc_in_sphere_exact(const int64_t a[3], const int64_t b[3],
                        const int64_t c[3], const int64_t d[3],
                        const int64_t v[3]); /* -1, 0, +1 */
void orc_iloc(const double anchor[3], const double width[3], int periodic,
              int dim, const double loc[3], int64_t out[3]);
double orc_signed_volume_tet(const double v0[3], const double v1[3],
                             const double v2[3], const double v3[3]);
double orc_signed_area_tri(const double v0[3], const double v1[3],
                           const double v2[3], const double t[3]);
int orc_intersect_planes(const double n0[3], const double p0[3],
                         const double n1[3], const double p1[3],
                         const double n2[3], const double p2[3],
                         double out[3]);
/* half-space clip test (half_space.rs:40-47): -1, 0, +1 */
int orc_halfspace_clip(const double n[3], const double p[3],
                       const double v[3]);
/* SimpleCycle driver for the known-answer test of simple_cycle.rs:238-263:
 * tris: n_tris x 3 indices, capacity = number of planes.  Writes the cycle
 * (length returned) into out. */
int orc_simple_cycle(const int32_t *tris, int n_tris, int capacity,
                     int32_t *out);
/* neighbour stream (rtree_nn.rs) for one query cell: first `max_out`
 * entries in ascending (key, idx, image) order. Returns count written. */
size_t orc_neighbour_stream(const double *xyz, size_t n, const double width[3],
                            int dim, int periodic, size_t query,
                            size_t max_out, uint64_t *out_idx,
                            int32_t *out_image, double *out_key);

#ifdef __cplusplus
}
#endif
#endif
