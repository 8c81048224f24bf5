/*
 * mvoro.h -- C ABI of the MI355X-native per-cell Voronoi construction
 * (libmvoro.so, gfx950 HIP kernels).
 *
 * Drop-in boundary for the construction path of the meshless_voronoi crate
 * (v0.7.2).  Every entry point names the reference item it replaces.  The
 * reference has no FFI of its own (it is a Rust crate); a Rust caller binds
 * these functions with an extern "C" block over the same signatures.
 *
 * Conventions
 *  - positions are f64 AoS [x, y, z] with a 24-byte stride, which is the
 *    memory layout of glam's DVec3, so a &[DVec3] can be passed as-is;
 *  - masks are one byte per generator (0 = inactive), as Rust's &[bool];
 *  - every call blocks until its results are complete;
 *  - errors are returned as mv_status, one code per panic site of the
 *    reference (it panics where this library returns an error);
 *  - a handle is immutable after the build returns and may be read from any
 *    thread; builds on different handles may run concurrently.
 */
#ifndef MVORO_H
#define MVORO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum mv_status {
  MV_OK = 0,
  MV_E_ARG = 1,               /* bad argument (also: mask shorter than N,
                                 reference src/voronoi.rs:286) */
  MV_E_DEGENERATE_POINTS = 2, /* zero or non-finite generator distance,
                                 src/voronoi/convex_cell.rs:342 */
  MV_E_SELF_NOT_FIRST = 3,    /* duplicate generators,
                                 src/voronoi/convex_cell.rs:325-329 */
  MV_E_PLANE_DET_ZERO = 4,    /* src/geometry.rs:70 */
  MV_E_NAN = 5,               /* src/voronoi/convex_cell.rs:448 */
  MV_E_BOUNDARY_WALK = 6,     /* src/voronoi/convex_cell.rs:428 */
  MV_E_DEVICE = 7,            /* HIP runtime error or no GPU */
  MV_E_OOM = 8,               /* device or host allocation failed */
  MV_E_CAPACITY = 9,          /* a cell exceeded the largest per-cell
                                 capacity (planes / vertices / neighbours) */
  MV_E_FACE_WALK = 10,        /* no next vertex around a face,
                                 src/voronoi/convex_cell.rs:538-541 */
  MV_E_NOT_3D = 11,           /* with_faces outside 3D,
                                 src/voronoi/convex_cell.rs:462-466 */
  MV_E_SCRATCH = 12,          /* the allocation-free build's static scratch
                                 cannot hold this input */
  MV_E_OUT_SMALL = 13         /* a caller's output array is too short (the
                                 counts needed are returned) */
} mv_status;

/* Construction modes. */
enum {
  MV_MODE_PLAIN = 0,  /* fixed k neighbours per pass, k doubled for the cells
                         whose security radius was not reached */
  MV_MODE_LEARNED = 1 /* per-cell k proposed by the learned model, cells
                         bucketed by k and run in one persistent kernel;
                         the security test still decides every cell */
};

struct mv_model;

typedef struct mv_options {
  uint32_t mode;     /* MV_MODE_*; default MV_MODE_PLAIN */
  uint32_t k0;       /* neighbours in the first pass, self included
                        (plain mode); default 27 */
  uint32_t k_max;    /* largest neighbour list tried; default and
                        maximum 4194304 (MV_K_MAX) */
  uint32_t lanes;    /* thread-per-cell lanes in flight; 0 = auto */
  uint32_t grid_occupancy16; /* generators per grid cell x 16; 0 = 32 */
  uint32_t flags;    /* MV_FLAG_* */
  const struct mv_model *model; /* MV_MODE_LEARNED: proposal weights */
  uint32_t clip_blocks_per_cu;  /* one-cell-per-wavefront clip launch:
                                   blocks of 4 wavefronts per CU; 0 = 4 */
  uint32_t coop_max_planes;     /* large-cell kernel: planes and vertices a */
  uint32_t coop_max_vertices;   /* cell may use there before it moves to the
                                   wide kernel (device-memory arrays without
                                   fixed capacity); 0 = the kernel's 255 and
                                   1024.  Lower values are for testing. */
  uint32_t cells_per_wave;      /* main lists: 1 cell per wavefront (0 or 1,
                                   the default), 2 (two cells per wavefront,
                                   one per 32-lane half; cells over 32
                                   vertices or planes continue on the
                                   one-cell kernel) or 64 (one cell per
                                   lane) */
  uint32_t first_list_cap;      /* learned mode: longest first-pass list,
                                   multiple of 8 in [16, 1024]; 0 = 256 */
  uint32_t knn_min_tile;        /* neighbour gathers: smallest LDS tile
                                   (512, 1024 or 2048 entries); 0 = 512 */
  uint32_t pipeline;            /* first pass: bits 0-7 = chunks C (0 or 1:
                                   off); the lists of chunk k + 1 are built
                                   on a second stream with bits 8-15 (0 = 1)
                                   blocks per CU while chunk k is clipped */
} mv_options;

/* Run every cell on the thread-per-cell kernel, which sums each integral in
 * the reference's exact order (bit-identical volumes, centroids and face
 * areas); the default one-cell-per-wavefront kernel gives the same
 * connectivity with integrals equal up to rounding. */
#define MV_FLAG_EXACT_ORDER 1u

/* Largest neighbour list (entries of the neighbour stream one cell may
 * need: all generators within its security radius, every periodic image
 * counted). */
#define MV_K_MAX 4194304u

/* VoronoiIntegrator::build (src/voronoi.rs:549-621): also keep every active
 * cell's half-spaces and vertices on the device, for custom cell and face
 * integrals (mv_cell_integrals, mv_face_integrals, mv_tets and the device
 * templates of mvoro_integrals.hpp). */
#define MV_FLAG_KEEP_CELLS 2u

/* Learned mode: keep the cells in grid order instead of ordering them by
 * proposed list length (longest first). */
#define MV_FLAG_NO_LENGTH_ORDER 4u

/* Fill *opt with the defaults. */
void mv_default_options(mv_options *opt);

typedef struct mv_tess mv_tess;

/* Voronoi::build (src/voronoi.rs:151-159) when mask == NULL and
 * Voronoi::build_partial (src/voronoi.rs:181-190) otherwise.
 * xyz, mask: host pointers.  dimensionality: 1, 2 or 3.  Results are
 * downloaded to host memory owned by *out. */
mv_status mv_build(const double *xyz, size_t n, const uint8_t *mask,
                   const double anchor[3], const double width[3],
                   uint32_t dimensionality, uint32_t periodic,
                   const mv_options *opt, mv_tess **out);

/* Same as mv_build with device-resident inputs (d_xyz, d_mask are device
 * pointers on the current HIP device); the results stay in device memory
 * (see mv_device_view) and are downloaded on the first host accessor
 * call. */
mv_status mv_build_device(const double *d_xyz, size_t n,
                          const uint8_t *d_mask, const double anchor[3],
                          const double width[3], uint32_t dimensionality,
                          uint32_t periodic, const mv_options *opt,
                          mv_tess **out);

/* Many independent tessellations in one pass (the batched form of
 * Voronoi::build / build_partial; no counterpart in the reference, which
 * builds one at a time).  Problem p owns generators [offsets[p],
 * offsets[p+1]) of xyz (and of mask), its box anchors[3p..3p+2] /
 * widths[3p..3p+2] and periodic[p] (NULL: none periodic); offsets has P + 1
 * entries, offsets[0] = 0.  out[p] receives problem p's handle, equal to
 * what mv_build of that problem alone returns; each handle is freed with
 * mv_free (the device memory is released with the last one).  All problems
 * run in the same kernel launches: one grid build, one neighbour-list and
 * one clip launch per pass over every cell of every problem. */
mv_status mv_build_batch(const double *xyz, const uint64_t *offsets, uint32_t P,
                         const uint8_t *mask, const double *anchors,
                         const double *widths, const uint32_t *periodic,
                         uint32_t dimensionality, const mv_options *opt,
                         mv_tess **out);
/* Same with device-resident xyz and mask (offsets, anchors, widths and
 * periodic are host arrays); results stay on the device. */
mv_status mv_build_batch_device(const double *d_xyz, const uint64_t *offsets,
                                uint32_t P, const uint8_t *d_mask,
                                const double *anchors, const double *widths,
                                const uint32_t *periodic,
                                uint32_t dimensionality,
                                const mv_options *opt, mv_tess **out);

void mv_free(mv_tess *t);

/* Voronoi::anchor / width / dimensionality / periodic
 * (src/voronoi.rs:317-359); anchor and width after the 1D/2D normalisation
 * of src/voronoi.rs:200-210. */
void mv_anchor(const mv_tess *t, double out[3]);
void mv_width(const mv_tess *t, double out[3]);
uint32_t mv_dimensionality(const mv_tess *t);
uint32_t mv_periodic(const mv_tess *t);

/* sizes: Voronoi::cells().len(), faces().len(),
 * cell_face_connections().len() */
size_t mv_n_cells(const mv_tess *t);
size_t mv_n_faces(const mv_tess *t);
size_t mv_n_links(const mv_tess *t);

/* VoronoiCell accessors (src/voronoi/voronoi_cell.rs:107-165), SoA, host
 * memory valid until mv_free.  Inactive cells of a partial build are all
 * zero (VoronoiCell::default), including loc and idx. */
const double *mv_cell_loc(mv_tess *t);           /* N x 3 */
const double *mv_cell_centroid(mv_tess *t);      /* N x 3 */
const double *mv_cell_volume(mv_tess *t);        /* N */
const double *mv_cell_safety_radius(mv_tess *t); /* N */
const uint64_t *mv_cell_face_offset(mv_tess *t); /* face_connections_offset */
const uint64_t *mv_cell_face_count(mv_tess *t);  /* face_count */
const uint64_t *mv_cell_idx(mv_tess *t);         /* N (0 when inactive) */

/* VoronoiFace accessors (src/voronoi/voronoi_face.rs:241-300). */
const uint64_t *mv_face_left(mv_tess *t);
const int64_t *mv_face_right(mv_tess *t);   /* -1 = None (boundary face) */
const double *mv_face_area(mv_tess *t);
const double *mv_face_centroid(mv_tess *t); /* F x 3 */
const double *mv_face_normal(mv_tess *t);   /* F x 3 (plane normal of the
                                               left cell's half-space) */
const uint8_t *mv_face_has_shift(mv_tess *t);
const double *mv_face_shift(mv_tess *t);    /* F x 3 (0 when no shift) */

/* Voronoi::cell_face_connections (src/voronoi.rs:346-349). */
const uint64_t *mv_cell_face_connections(mv_tess *t);

/* Device-resident results (same layouts as the host accessors). */
typedef struct mv_device_view {
  const double *cell_loc, *cell_centroid, *cell_volume, *cell_safety_radius;
  const uint64_t *cell_face_offset, *cell_face_count, *cell_idx;
  const uint64_t *face_left;
  const int64_t *face_right;
  const double *face_area, *face_centroid, *face_normal;
  const uint8_t *face_has_shift;
  const double *face_shift;
  const uint64_t *cell_face_connections;
} mv_device_view;
void mv_get_device_view(const mv_tess *t, mv_device_view *v);

/* Construction statistics (no counterpart in the reference). */
typedef struct mv_stats {
  uint32_t passes;          /* neighbour-list passes run */
  uint32_t k_last;          /* list length of the last pass */
  uint64_t cells_redone;    /* cells rebuilt after an insufficient k */
  uint64_t exact_tests;     /* exact in-sphere evaluations */
  uint64_t big_cells;       /* cells run with the large-capacity kernel */
  double ms_grid, ms_knn, ms_clip, ms_finalize, ms_total;
  /* per-kernel figures of the two hot kernels (device time from HIP events
   * on the build's stream, summed over launches) */
  uint32_t clip_wave_launches;  /* one-cell-per-wavefront clip launches */
  uint32_t knn_launches;        /* neighbour-list launches (all variants) */
  uint64_t clip_wave_cells;     /* cells the wavefront clip completed */
  uint64_t knn_queries;         /* neighbour lists built */
  uint64_t knn_entries;         /* list entries written */
  double ms_clip_wave;          /* device time of those clip launches */
  double ms_knn_kernels;        /* device time of the neighbour-list launches */
  double ms_model;              /* features + proposal inference */
  uint64_t wide_cells;          /* cells moved to the wide kernel */
  uint64_t pair_handovers;      /* cells_per_wave = 2: cells continued on the
                                   one-cell kernel from their saved state */
  uint64_t knn_big_items;       /* lists built by the long-list kernel */
  uint64_t knn_big_gathers;     /* its ball gathers (more than one per list
                                   when the first radius held too few or too
                                   many entries) */
  uint64_t knn_big_entries;     /* entries in its last ball of each list */
  uint64_t knn_gathers;         /* ball gathers of the other list kernels
                                   (knn_queries when every first radius
                                   held the list) */
  uint64_t knn_reads;           /* seeds those gathers read (tested against
                                   the ball) */
  uint64_t dense_cells;         /* grid cells of more than 256 seeds split
                                   into sub-columns for the gathers */
} mv_stats;
void mv_get_stats(const mv_tess *t, mv_stats *s);

/* Per-cell construction record (no counterpart in the reference):
 * number of neighbour-stream entries consumed, self and the entry that
 * triggered the security test included (0 for inactive cells). */
const uint32_t *mv_cell_n_popped(mv_tess *t);

const char *mv_status_string(mv_status s);

/* ---------------------------------------------------------------------
 * Custom integrals over kept cells (builds with MV_FLAG_KEEP_CELLS).
 * Outputs follow VoronoiIntegrator (src/voronoi.rs:661-767): one entry per
 * active cell in generator order; faces per active cell in half-space
 * order. */

/* Number of active (built) cells. */
size_t mv_n_active(const mv_tess *t);

/* compute_cell_integrals::<I> for the reference's example integrals
 * (src/voronoi/integrals.rs:69-142): MV_CELL_VOLUME writes 1 double per
 * active cell, MV_CELL_VOLUME_CENTROID 4 (volume, centroid x, y, z). */
enum { MV_CELL_VOLUME = 0, MV_CELL_VOLUME_CENTROID = 1 };
mv_status mv_cell_integrals(mv_tess *t, uint32_t kind, double *out);

/* compute_face_integrals::<I> (symmetric = 0) and
 * compute_face_integrals_sym::<I> (symmetric = 1) for AreaIntegral
 * (MV_FACE_AREA: 1 value) and AreaCentroidIntegral (MV_FACE_AREA_CENTROID:
 * area, centroid x, y, z), with FaceIntegrator's left / right (-1 = None) /
 * shift and the face's half-space index within its cell.  Call with
 * left == NULL to get *count; then pass arrays of that many entries (any of
 * right, has_shift, shift, plane, values may be NULL). */
enum { MV_FACE_AREA = 0, MV_FACE_AREA_CENTROID = 1 };
mv_status mv_face_integrals(mv_tess *t, uint32_t kind, uint32_t symmetric,
                            size_t *count, uint64_t *left, int64_t *right,
                            uint8_t *has_shift, double *shift, int32_t *plane,
                            double *values);

/* The oriented tets that CellIntegral::collect receives, for every active
 * cell in order (ConvexCell::decompose, src/voronoi/convex_cell.rs:131-187,
 * 576-592): v012 holds count x 9 doubles (v0, v1, v2), plane the half-space
 * index of each tet, cell_offset n_active + 1 tet offsets.  Call with
 * v012 == NULL to get *count.  For host-side custom integrals. */
mv_status mv_tets(mv_tess *t, size_t *count, double *v012, int32_t *plane,
                  uint64_t *cell_offset);

/* Device view of the kept cells, the input of the templates in
 * mvoro_integrals.hpp (mv::CellGeom has this layout).  Pointers are device
 * memory owned by the handle. */
typedef struct mv_cell_geometry {
  const double *pl_a;       /* 4 per half-space: n.xyz, d */
  const double *pl_b;       /* 4 per half-space: p.xyz, reference word */
  const double *vx;         /* 4 per vertex: loc.xyz, packed dual bits */
  const uint64_t *pbase, *vbase;  /* per generator */
  const uint32_t *np, *nv;        /* per generator (0: not built) */
  const double *cell_loc;         /* per generator, 3 */
  double width[3];
  int32_t dim;
  uint64_t idx_base;  /* subtracted from the generator index of reference
                         words (non-zero for the problems of a batch) */
  /* after mv_with_faces (NULL before and after mv_discard_faces): */
  const uint64_t *fbase;      /* per generator: its first face */
  const uint32_t *nf;         /* per generator: face count */
  const int32_t *face_plane;  /* per face: half-space index in the cell */
  const uint32_t *face_nvert; /* per face: vertex count */
  const uint64_t *face_voff;  /* per face: offset of its vertices in face_vtx */
  const uint32_t *face_vtx;   /* vertex indices within the cell, CCW */
} mv_cell_geometry;
mv_status mv_get_cell_geometry(const mv_tess *t, mv_cell_geometry *g);

/* The kept cells' ConvexCell fields, copied to host: clipping_planes
 * (HalfSpace, src/voronoi/half_space.rs:7-19: plane n and p, right_idx
 * (-1 = None), shift) and vertices (Vertex, src/voronoi/convex_cell.rs:14-26:
 * loc and dual, the dual's indices into the cell's half-spaces), for every
 * active cell in generator order: half-spaces plane_offset[i] ..
 * plane_offset[i+1], vertices vertex_offset[i] .. vertex_offset[i+1]
 * (n_active + 1 offsets each).  Call with plane_n == NULL to get the
 * totals; then pass arrays of that size (3 doubles per n, p, shift, loc;
 * 3 int32 per dual). */
mv_status mv_kept_cells(mv_tess *t, size_t *n_planes, size_t *n_vertices,
                        uint64_t *plane_offset, uint64_t *vertex_offset,
                        double *plane_n, double *plane_p, int64_t *plane_right,
                        uint8_t *plane_has_shift, double *plane_shift,
                        double *vertex_loc, int32_t *vertex_dual);

/* VoronoiIntegrator::with_faces (src/voronoi.rs:625-643) and
 * ConvexCell::with_faces (src/voronoi/convex_cell.rs:461-505): for every
 * kept cell, the vertices of each half-space holding any, sorted
 * counter-clockwise as seen from the generator by the reference's walk
 * (sort_face_vertices, :508-543).  Afterwards mv_tets, mv_cell_integrals,
 * mv_face_integrals, mv_voronoi_from_integrator and the templates of
 * mvoro_integrals.hpp decompose each face as the fan (v0, vk, vk+1) of
 * DecompositionWithFaces (:89-129).  MV_E_NOT_3D outside 3D. */
mv_status mv_with_faces(mv_tess *t);

/* ConvexCell::discard_faces (src/voronoi/convex_cell.rs:667-673): back to
 * the decomposition without faces. */
mv_status mv_discard_faces(mv_tess *t);

/* The faces of mv_with_faces (ConvexCell::face_count / clipping_plane /
 * face_vertex_count / face_vertices, src/voronoi/convex_cell.rs:675-722):
 * per active cell i its faces cell_face_offset[i] .. cell_face_offset[i+1]
 * (n_active + 1 offsets), in half-space order; per face its half-space
 * index, vertex count and offset into face_vertices, which holds vertex
 * indices within the cell.  Call with face_plane == NULL to get *n_faces
 * and *n_entries.  MV_E_ARG when the handle is not in the with-faces
 * state. */
mv_status mv_face_vertices(mv_tess *t, size_t *n_faces, size_t *n_entries,
                           uint64_t *cell_face_offset, int32_t *face_plane,
                           uint32_t *face_vertex_count, uint64_t *face_vertex_offset,
                           uint32_t *face_vertices);

/* From<&VoronoiIntegrator<M>> for Voronoi (src/voronoi.rs:493-512,
 * build_voronoi_cells :811-827): a new handle (free with mv_free) whose
 * cells and faces are integrated from the kept cells' current
 * decomposition (fans after mv_with_faces); connectivity, links, normals
 * and shifts are those of the build. */
mv_status mv_voronoi_from_integrator(const mv_tess *t, mv_tess **out);

/* ---------------------------------------------------------------------
 * Learned neighbour-count proposal (no counterpart in the reference; it
 * only sizes each cell's first neighbour list in MV_MODE_LEARNED, the
 * security test still decides every cell).
 * A 16 -> 64 -> 64 -> 2 ReLU MLP with BF16 weights; output 0 = log of the
 * list length a cell needs (the neighbour-stream entries it clips with, self
 * included: a list that long plus the distance of the next entry decides the
 * cell), output 1 = log of its final safety radius over the local spacing
 * (the start radius of its neighbour gather).  This struct is also the
 * on-disk checkpoint format (little-endian, version 1). */
#define MV_MODEL_MAGIC 0x4b4c564du /* "MVLK" */
typedef struct mv_model {
  uint32_t magic, version;
  float mu[16], sigma[16];        /* feature normalisation */
  uint16_t w1[16 * 64];           /* bf16, [input][output] */
  uint16_t w2[64 * 64];
  uint16_t w3[64 * 2];
  float b1[64], b2[64], b3[2];
  float tau;                      /* quantile of the training loss */
  float kscale;                   /* multiplier on the proposed k */
  float rscale;                   /* multiplier on the proposed gather
                                     radius; 0 = start from the grid
                                     density instead */
} mv_model;

typedef struct mv_train_config {
  uint32_t steps;   /* Adam steps */
  uint32_t batch;   /* rows per step (multiple of 16) */
  uint32_t seed;    /* initialisation and batch sampling */
  float lr;         /* Adam learning rate */
  float tau;        /* pinball-loss quantile */
} mv_train_config;

/* The 16 proposal features of every generator (host in/out; features
 * n x 16 and local spacing n, in generator order). */
mv_status mv_model_features(const double *xyz, size_t n,
                            const double anchor[3], const double width[3],
                            uint32_t dimensionality, uint32_t periodic,
                            float *features, float *hloc);

/* Train on the GPU (BF16 MFMA forward and backward kernels, FP32 Adam):
 * features n x 16, targets n x 2 (host).  *model receives the weights. */
mv_status mv_model_train(const float *features, const float *targets,
                         size_t n, const mv_train_config *cfg,
                         mv_model *model, float *final_loss);

/* Inference on the GPU: out n x 2 (host). */
mv_status mv_model_predict(const mv_model *model, const float *features,
                           size_t n, float *out);

/* One forward and backward pass of the training kernels on the given rows
 * (n multiple of 16): mean pinball loss and its gradient with respect to
 * the FP32 parameters, laid out w1 (16x64), b1, w2 (64x64), b2, w3 (64x2),
 * b3 -- 5378 floats. */
mv_status mv_model_loss_grad(const mv_model *model, const float *features,
                             const float *targets, size_t n, float tau,
                             float *loss, float *grad);

/* Exact in-sphere sign on snapped integer coordinates, evaluated on the
 * device for `count` 5-tuples (a, b, c, d, v; 15 int64 per tuple, host
 * memory).  Replaces in_sphere_test_exact (src/geometry.rs:271-314);
 * writes -1, 0 or +1 per tuple. */
mv_status mv_in_sphere_exact(const int64_t *abcdv, size_t count,
                             int8_t *out_sign);

/* ---- allocation-free build ------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The same construction as
 * mv_build_device (same kernels, same results bit for bit), but every
 * scratch buffer -- grid, neighbour lists, face staging, counters, the
 * learned model's weights -- is carved from one static device array of
 * MV_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until
 * the process ends, and the results go to device arrays the caller passes.
 * The call allocates nothing (no hipMalloc, hipFree, stream or event
 * creation, no pinned host memory), runs on the null stream and returns
 * when the results are written.  Every call on one GPU shares that GPU's
 * scratch: calls from one process are serialised by the library, but a
 * call that overlaps another on the same GPU (from another thread's own
 * copy of the library, say) races.  Limits: n <= MV_NOALLOC_MAX_SEEDS
 * (larger inputs return MV_E_SCRATCH at once); an input whose lists or
 * faces outgrow the scratch returns MV_E_SCRATCH too; MV_FLAG_KEEP_CELLS is
 * refused (MV_E_ARG: kept cells need a handle). */
#define MV_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define MV_NOALLOC_MAX_SEEDS 32768u

/* Caller-owned device output arrays (layouts as mv_device_view; any
 * pointer may be NULL to skip that array).  Cell arrays hold n entries
 * (x3 where noted), face arrays face_cap (x3), links link_cap.  On return
 * n_faces / n_links hold the counts; when they exceed face_cap / link_cap
 * nothing is written and the call returns MV_E_OUT_SMALL. */
typedef struct mv_noalloc_out {
  double *cell_loc, *cell_centroid, *cell_volume, *cell_safety_radius; /* 3n, 3n, n, n */
  uint64_t *cell_face_offset, *cell_face_count, *cell_idx;            /* n */
  uint64_t *face_left;                                                /* face_cap */
  int64_t *face_right;
  double *face_area, *face_centroid, *face_normal;                    /* F, 3F, 3F */
  uint8_t *face_has_shift;
  double *face_shift;                                                 /* 3F */
  uint64_t *cell_face_connections;                                    /* link_cap */
  size_t face_cap, link_cap;
  size_t n_faces, n_links;                                            /* out */
  size_t scratch_used;                                                /* out: bytes */
} mv_noalloc_out;

mv_status mv_build_device_noalloc(const double *d_xyz, size_t n,
                                  const uint8_t *d_mask, const double anchor[3],
                                  const double width[3], uint32_t dimensionality,
                                  uint32_t periodic, const mv_options *opt,
                                  mv_noalloc_out *out);

/* ---------------------------------------------------------------------
 * Captured builds.  No counterpart in the reference, which rebuilds from
 * scratch on every call (src/voronoi.rs:151-238); for callers that rebuild
 * the same problems again and again as their seeds move (a simulation
 * step, a serving loop).  mv_graph_create runs one ordinary build of the
 * problems (as mv_build_batch_device, P = 1 for a single build), plans
 * every pass from it with headroom (cells per pass, list lengths and
 * entries, faces) and records the whole build -- grid, neighbour lists,
 * clipping, redo passes, faces and links -- as one hipGraph whose counts
 * stay on the device.  mv_graph_launch replays it on the current contents
 * of d_xyz (and d_mask): when the replay needed no more than was recorded,
 * *replayed = 1 and the P handles view the graph's output arrays (the next
 * launch overwrites them; they stay readable until then, and their memory
 * until the graph and the handles are freed); otherwise (more cells to
 * redo, longer lists, a seed outside the recorded grid, an error) the
 * ordinary build runs instead, *replayed = 0, and the handles are its own.
 * Either way the results are those of mv_build_batch_device on the same
 * data, bit for bit.  d_xyz, d_mask, the problem layout and the options
 * are fixed at creation; d_xyz and d_mask must stay valid until
 * mv_graph_free.  Refused (MV_E_ARG): MV_FLAG_KEEP_CELLS,
 * MV_FLAG_EXACT_ORDER, cells_per_wave other than 0 or 1, no seeds.
 * A launch starts after the work queued before it on the null stream (as
 * the ordinary build, which runs on it: writes to d_xyz queued there, or
 * on a stream that synchronises with it, are seen) and returns when the
 * results are complete.  Launches of one graph must not overlap. */
typedef struct mv_graph mv_graph;
mv_status mv_graph_create(const double *d_xyz, const uint64_t *offsets, uint32_t P,
                          const uint8_t *d_mask, const double *anchors, const double *widths,
                          const uint32_t *periodic, uint32_t dimensionality,
                          const mv_options *opt, mv_graph **out);
mv_status mv_graph_launch(mv_graph *g, mv_tess **out, uint32_t *replayed);
void mv_graph_free(mv_graph *g);
/* Why the last launch did not replay (0: it did): bit 0 a seed outside the
 * recorded grid, 1 more cells for a pass than recorded, 2 more list
 * entries, 3 an error status or full face staging, 4 cells left after the
 * last recorded pass, 5 a cell over the large-cell kernel's capacity. */
uint32_t mv_graph_last_violation(const mv_graph *g);

/* Test hook, no counterpart in the reference: the number of hipMalloc calls
 * the library has made in this process. */
uint64_t mv_test_device_mallocs(void);

/* Test hook, no counterpart in the reference: runs the host download's
 * thread fan-out over `threads` shares, with the start of thread
 * `fail_after` and later ones made to fail (-1: none), each share writing
 * out[w] = w + 1.  Shows that shares whose thread could not start run on
 * the calling thread and that no exception leaves the library.  Needs no
 * GPU. */
mv_status mv_test_run_shares(int32_t threads, int32_t fail_after, int32_t *out);

#ifdef __cplusplus
}
#endif
#endif
