/*
 * brille_amd.h -- C ABI of the MI355X implementation of brille's batched
 * Q-point path: BrillouinZone::ir_moveinto -> PolyTrellis point location ->
 * DualInterpolator weighted sums -> Interpolator::rotate_in_place.
 *
 * Everything crossing this boundary is a plain pointer, a size or a POD
 * struct; no HIP, torch or C++ types appear in the signatures.  Streams are
 * passed as opaque `void*` (a hipStream_t; NULL = the default stream).
 *
 * The grid is handed over as a "frozen" set of tables (brille_grid_desc):
 * the numbers the reference keeps inside BrillouinZone, PolyTrellis,
 * NodeContainer and PointSymmetry objects, flattened to arrays.  The library
 * copies them to HBM once (brille_grid_create) and every call afterwards only
 * streams Q in and results out.
 *
 * Reference interfaces replaced (paths relative to the brille source tree):
 *   brille_grid_create          BrillouinZoneTrellis3(const BrillouinZone&, double, bool)
 *                               src/bz_trellis.hpp:56-59 (table upload only: the
 *                               construction itself is host code, see the package)
 *   brille_grid_fill            BZTrellisQdc.fill(values, values_elements, vectors,
 *                               vectors_elements, sort=False)
 *                               wrap/_common_grid.hpp:26-124, wrap/_interpolator.hpp:15-51,
 *                               src/interpolatordual.hpp:189-195
 *   brille_ir_interpolate_at    BrillouinZoneTrellis3<double,std::complex<double>,double>
 *                               ::ir_interpolate_at(const LVec<double>&, int)
 *                               src/bz_trellis.hpp:153-196, Python binding
 *                               wrap/_common_grid.hpp:276-301
 *   brille_ir_moveinto          BrillouinZone::ir_moveinto  src/bz_move.cpp:165-296
 *   brille_bytes_per_point      bytes_per_point  src/interpolatordual.hpp:297-299
 *   brille_last_error           std::runtime_error / std::logic_error text that the
 *                               reference raises (messages listed with the status codes)
 *   brille_batch_*              no reference counterpart (many grids in one launch)
 *   brille_proposal_*           no reference counterpart (learned tetrahedron proposal)
 *   brille_device_* / stream    plumbing for callers without a GPU runtime of their own
 */
#ifndef BRILLE_AMD_H
#define BRILLE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRILLE_AMD_ABI_VERSION 1

/* Status codes.  Each maps onto one reference error message. */
enum brille_status {
  BRILLE_OK = 0,
  BRILLE_ERR_INVALID_ARGUMENT = 1, /* "Only (n,3) two-dimensional Q vectors supported ..." src/trellis_poly.hpp:261-262 */
  BRILLE_ERR_NOT_FILLED = 2,       /* "The interpolation data must be filled before interpolating." src/trellis_poly.hpp:259-260 */
  BRILLE_ERR_OUTSIDE_BZ = 3,       /* "Not all points inside Brillouin zone" src/bz_move.cpp:158 */
  BRILLE_ERR_OUTSIDE_IR = 4,       /* "Q = ... is outside of the irreducible BrillouinZone" src/bz_move.cpp:288-294 */
  BRILLE_ERR_NULL_NODE = 5,        /* "Attempting to access null node at index ..." src/trellis_node.hpp:593-610 */
  BRILLE_ERR_NOT_FOUND = 6,        /* "interpolate_at failed to find N points" src/trellis_poly.hpp:345-349 */
  BRILLE_ERR_ROTATION = 7,         /* "LengthUnit, RotatesLike combination not implemented" src/interpolator.hpp:400-426 */
  BRILLE_ERR_DEVICE = 8,           /* a HIP runtime call failed */
  BRILLE_ERR_NO_DEVICE = 9,        /* no usable gfx950 device */
  BRILLE_ERR_UNSUPPORTED = 10,     /* a size beyond the kernels' documented limits */
  BRILLE_ERR_NUMERIC = 11          /* "Complex Hermitian angle cos(theta)=... " src/utilities.tpp:362-378 */
};

/* Node types, numbered as the reference's NodeType enum (src/enums.hpp:43). */
enum brille_node_type {
  BRILLE_NODE_ASSUMED_NULL = 0,
  BRILLE_NODE_FOUND_NULL = 1,
  BRILLE_NODE_NULL = 2,
  BRILLE_NODE_CUBE = 3,
  BRILLE_NODE_POLY = 4
};

/* RotatesLike and LengthUnit values as accepted by fill (wrap/_interpolator.hpp:31-47). */
enum brille_rotates_like { BRILLE_ROT_VECTOR = 0, BRILLE_ROT_PSEUDOVECTOR = 1, BRILLE_ROT_GAMMA = 2 };
enum brille_length_unit {
  BRILLE_LU_NONE = 0, BRILLE_LU_ANGSTROM = 1, BRILLE_LU_INVERSE_ANGSTROM = 2,
  BRILLE_LU_REAL_LATTICE = 3, BRILLE_LU_RECIPROCAL_LATTICE = 4
};

/* Per-point status byte written by the kernels (aux->status). */
enum brille_point_status {
  BRILLE_PT_OK = 0,
  BRILLE_PT_OUTSIDE_BZ = 1,  /* moveinto left q outside the first zone (or Q not finite) */
  BRILLE_PT_OUTSIDE_IR = 2,  /* no point-group operation put q in the wedge */
  BRILLE_PT_NULL_NODE = 3,   /* located node is null / outside the trellis */
  BRILLE_PT_NOT_FOUND = 4    /* no (index, weight) pair survived */
};

/*
 * Frozen grid tables.  All matrices are 3x3 row-major (m[3*i+j]).  All
 * arrays are host memory owned by the caller for the duration of
 * brille_grid_create only.  "rlu" = components in a reciprocal basis,
 * "xyz" = Cartesian inverse angstrom.
 */
typedef struct brille_grid_desc {
  /* conventional ("outer") reciprocal lattice: Lattice::to_xyz, metric  src/lattice_dual.hpp:566-573, 672-685 */
  double outer_to_xyz[9];        /* B, basis vectors as columns */
  double outer_metric_recip[9];  /* B^T B */
  double outer_metric_real[9];   /* A^T A */
  double outer_cross_scale;      /* det(B) / (2 pi), the LVec cross product scale src/array_functions.hpp:196 */
  /* primitive ("inner") reciprocal lattice used by moveinto  src/bz_move.cpp:103-163, src/primitive.hpp:112-180 */
  int32_t transform_needed;      /* PrimitiveTransform(bravais).does_anything() */
  int32_t to_primitive_6Pt[9];   /* 6 P^T */
  int32_t from_primitive_invPt[9]; /* (P^-1)^T */
  double prim_metric_recip[9];
  double prim_metric_real[9];
  double prim_cross_scale;
  /* first Brillouin zone, F faces */
  uint32_t n_faces;
  const double* prim_plane_a;    /* F x 3: first three vertices of each face, primitive rlu */
  const double* prim_plane_b;
  const double* prim_plane_c;
  const double* prim_normals;    /* F x 3: face normals normalised in the primitive metric */
  const int32_t* prim_taus;      /* F x 3: the lattice vector each face bisects, primitive rlu */
  const double* prim_tau_lens;   /* F: norm(taus) in the primitive metric */
  const double* outer_plane_a;   /* F x 3: the same planes in conventional rlu (isinside) */
  const double* outer_plane_b;
  const double* outer_plane_c;
  /* irreducible wedge  src/bz.hpp:757-763 */
  uint32_t n_wedge;
  const double* wedge_normals;   /* W x 3, conventional rlu */
  int32_t no_ir_mirroring;
  int32_t has_wedge_seed;        /* 1 if the wedge is the Dirichlet domain of wedge_seed */
  double wedge_seed[3];          /* conventional rlu */
  /* point group in PointSymmetry order  src/pointsymmetry.cpp:131-188 */
  uint32_t n_ops;
  const int32_t* rotations;      /* |G| x 9, real-space matrices W */
  const uint32_t* inverse_index; /* |G|: get_inverse_index */
  uint32_t identity_index;
  /* tolerances: approx_float::Config reciprocal tolerance and digits */
  double float_tolerance;
  int32_t approx_digits;
  /* trellis  src/trellis_poly.hpp:123-129, src/trellis_node.hpp:82, 210-212, 441-443 */
  uint32_t n_knots[3];
  const double* knots[3];        /* bin edges along x, y, z (xyz) */
  uint64_t n_nodes;              /* (n_knots[0]-1)(n_knots[1]-1)(n_knots[2]-1), x fastest */
  const uint8_t* node_type;      /* brille_node_type */
  const uint32_t* node_index;    /* index into cubes or polys */
  uint64_t n_cubes;
  const uint32_t* cube_vertices; /* n_cubes x 8: (000),(100),(110),(010),(101),(001),(011),(111) */
  uint64_t n_polys;
  const uint32_t* poly_tet_offset; /* n_polys + 1: CSR offsets into the tetrahedra */
  uint64_t n_tets;
  const uint32_t* tet_vertices;  /* n_tets x 4 */
  const double* tet_circumsphere;/* n_tets x 4: centre xyz, radius */
  const double* tet_volume;      /* n_tets */
  uint64_t n_vertices;
  const double* vertices;        /* V x 3 xyz */
} brille_grid_desc;

/*
 * Interpolation payload for one of the two Interpolator objects
 * (src/interpolator.hpp:94-101): data is (V, branches, span) flattened,
 * span = elements[0] + elements[1] + elements[2].
 */
typedef struct brille_data_desc {
  uint64_t n_vertices;
  uint32_t branches;
  uint32_t elements[3];          /* scalars, vector elements (3N), matrix elements (9M) */
  int32_t rotates_like;          /* brille_rotates_like */
  int32_t length_unit;           /* brille_length_unit */
  int32_t is_complex;            /* 0: double, 1: interleaved complex double */
  const double* data;
} brille_data_desc;

/* GammaTable (src/phonon.hpp:91-194), needed when vectors rotate like Gamma. */
typedef struct brille_gamma_desc {
  uint32_t n_atoms;
  uint32_t n_ops;                /* must equal the grid's n_ops */
  const uint32_t* F0;            /* n_atoms x n_ops: l_mapping */
  const uint32_t* vector_index;  /* n_atoms x n_ops: v_mapping */
  uint32_t n_vectors;
  const double* vectors;         /* n_vectors x 3: R^-1 r_l - r_k, direct lattice units */
  const double* cart_rotations;  /* n_ops x 9 Cartesian rotations (LengthUnit::angstrom), or NULL */
} brille_gamma_desc;

/* PermutationTable (src/permutation_table.hpp:63-214); NULL means identity everywhere. */
typedef struct brille_perm_desc {
  uint64_t n_keys;
  const uint64_t* keys;          /* sorted i*V+j */
  const uint32_t* perm_index;    /* n_keys: index into perms */
  uint32_t n_perms;
  const uint32_t* perms;         /* n_perms x branches */
} brille_perm_desc;

enum brille_precision { BRILLE_FP64 = 0, BRILLE_FP32 = 1 };
enum brille_memory { BRILLE_MEM_DEVICE = 0, BRILLE_MEM_HOST = 1 };
/* AUTO and TUNED pick TILE or BINNED by size (TILE when any aux output is
 * requested); PLAIN is the literal FP64 restatement (exact fold-in, atan2
 * antiphases); TILE is the fused one-launch kernel; BINNED buckets points by
 * trellis node range first so the payload gathers hit in cache (device work
 * space; reports every aux output but tau). */
enum brille_kernel {
  BRILLE_KERNEL_AUTO = 0,
  BRILLE_KERNEL_PLAIN = 1,
  BRILLE_KERNEL_TUNED = 2,
  BRILLE_KERNEL_TILE = 3,
  BRILLE_KERNEL_BINNED = 4
};

typedef struct brille_run_opts {
  int32_t no_move;       /* do_not_move_points: q is already irreducible */
  int32_t precision;     /* brille_precision of the value sums and outputs */
  int32_t use_proposal;  /* 1: learned tetrahedron proposal (needs a trained proposal) */
  int32_t memory;        /* brille_memory of Q and of every output pointer */
  int32_t kernel;        /* brille_kernel */
  int32_t check;         /* 1: reduce per-point status and return the first error (default) */
} brille_run_opts;

/* Optional per-point outputs; any pointer may be NULL. */
typedef struct brille_aux_out {
  double* ir_q;          /* N x 3 conventional rlu */
  int32_t* tau;          /* N x 3 */
  uint32_t* rot;         /* N: Ridx */
  uint32_t* invrot;      /* N: invRidx */
  uint32_t* node;        /* N: linear trellis node index */
  uint32_t* tet;         /* N: global tetrahedron id, 0xFFFFFFFF for cube nodes */
  uint8_t* n_iw;         /* N: number of (index, weight) pairs */
  uint32_t* iw_index;    /* N x 8 */
  double* iw_weight;     /* N x 8 */
  uint8_t* status;       /* N: brille_point_status */
} brille_aux_out;

typedef struct brille_grid brille_grid;
typedef struct brille_batch brille_batch;

/* ---- library / device ---- */
int brille_abi_version(void);
int brille_device_count(void);
size_t brille_last_error(char* buffer, size_t length);

/* ---- grid lifetime ---- */
int brille_grid_create(const brille_grid_desc* desc, brille_grid** out);
int brille_grid_fill(brille_grid* grid, const brille_data_desc* values,
                     const brille_data_desc* vectors, const brille_gamma_desc* gamma,
                     const brille_perm_desc* perm);
int brille_grid_destroy(brille_grid* grid);
/* brille_grid_fill for payloads that are not host float64 / complex128 (no
 * reference counterpart: the reference's fill takes float64 and complex128
 * arrays, wrap/_common_grid.hpp:26-124).  values->data and vectors->data point
 * to float / complex float when data_type is BRILLE_FP32 and may live in device
 * memory; only the precisions named in `keep` are packed (a call at another
 * precision then fails), so a payload of tens of GB needs no host copy and no
 * second device copy.  Checks run before the previous payload is released: a
 * failed fill leaves it in place unless packing itself fails. */
enum brille_keep { BRILLE_KEEP_BOTH = 0, BRILLE_KEEP_FP64 = 1, BRILLE_KEEP_FP32 = 2 };
typedef struct brille_fill_opts {
  int32_t memory;     /* brille_memory of values->data and vectors->data */
  int32_t data_type;  /* BRILLE_FP64: double / complex double; BRILLE_FP32: float / complex float */
  int32_t keep;       /* brille_keep: which payload precisions to pack */
} brille_fill_opts;
int brille_grid_fill_ex(brille_grid* grid, const brille_data_desc* values, const brille_data_desc* vectors,
                        const brille_gamma_desc* gamma, const brille_perm_desc* perm,
                        const brille_fill_opts* opts);
/* ---- mode sorting ---- */
/* The mode-assignment cost of one interpolator (Interpolator::set_cost_info,
 * src/interpolator.hpp:246-305, and its _costmult): scalar_cost 0 is the sum
 * of |a - b| over the scalar elements (the only kind); vector_cost 0 (and any
 * value above 4) is sin^2 of the Hermitian angle, 1 the vector distance, 2
 * one minus the vector product, 3 the vector angle, 4 the Hermitian angle;
 * weights scale the scalar, vector-element and matrix-element terms. */
typedef struct brille_cost_info {
  int32_t scalar_cost;
  int32_t vector_cost;
  double weights[3];
} brille_cost_info;
/* DualInterpolator::sort  src/interpolatordual.hpp:398-434 (called by the
 * wrapper's fill(..., sort=True), wrap/_common_grid.hpp:39-43): for every
 * pair of vertices that share a cube node or a tetrahedron
 * (PolyTrellis::collect_keys, src/trellis_poly_parallel.tpp:487-524), the
 * assignment of the second vertex's modes to the first's that minimises the
 * summed cost (src/interpolatordual.hpp:384-396, src/interpolator_cost.tpp:21-58),
 * stored for (i, j) and its inverse for (j, i).  Replaces the grid's
 * permutation table; needs a filled grid with complex vectors.  NULL cost
 * infos take the reference's defaults ({0, 0, {1, 1, 1}}).  The costs use the
 * FP64 payload rows when they were kept, else the FP32 rows. */
int brille_grid_sort(brille_grid* grid, const brille_cost_info* values_cost, const brille_cost_info* vectors_cost);
/* The grid's permutation table in brille_perm_desc form (keys sorted, one
 * permutation per key, index 0 the identity).  Call with NULL arrays to read
 * n_keys and n_perms, then with arrays of those sizes (perms: n_perms x
 * branches).  No reference counterpart: the reference keeps the table inside
 * the grid (src/permutation_table.hpp:63-214). */
int brille_grid_permutations(const brille_grid* grid, uint64_t* n_keys, uint32_t* n_perms, uint64_t* keys,
                             uint32_t* perm_index, uint32_t* perms);

/* bytes of interpolation output per Q at a precision  src/interpolatordual.hpp:297-299 */
uint64_t brille_bytes_per_point(const brille_grid* grid, int32_t precision);

/* ---- the hot path ---- */
int brille_ir_interpolate_at(brille_grid* grid, const double* Q, uint64_t n,
                             const brille_run_opts* opts, void* values_out,
                             void* vectors_out, const brille_aux_out* aux,
                             void* stream);
/* BrillouinZoneTrellis3::interpolate_at  src/bz_trellis.hpp:111-128 (Python
 * BZTrellisQdc.interpolate_at, wrap/_common_grid.hpp:412-440): Q is moved into
 * the first Brillouin zone only (BrillouinZone::moveinto, no wedge search) and
 * interpolated with no rotation; the trellis covers the irreducible wedge, so
 * first-zone images outside it fail as in the reference.  Same arguments. */
int brille_interpolate_at(brille_grid* grid, const double* Q, uint64_t n,
                          const brille_run_opts* opts, void* values_out,
                          void* vectors_out, const brille_aux_out* aux, void* stream);
/* No counterpart in the reference.  brille_ir_interpolate_at for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it allocates nothing.  Its error
 * record and the binned pipeline's work space live in one static device array
 * of 64 MiB per GPU (held from module load until the process ends); passes
 * are sized to fit it (2^19 points) and run on `stream` alone, and the row
 * table is used only if an earlier call built it.  Q, the outputs and aux must
 * be device memory (opts->memory = BRILLE_MEM_DEVICE); everything else is as
 * brille_ir_interpolate_at, results bit for bit the same.  Calls share their
 * GPU's array: calls whose work overlaps on one GPU race (use one stream, or
 * synchronise between them). */
int brille_ir_interpolate_at_noalloc(brille_grid* grid, const double* Q, uint64_t n,
                                     const brille_run_opts* opts, void* values_out, void* vectors_out,
                                     const brille_aux_out* aux, void* stream);
/* stage (a) alone: q, tau, Ridx, invRidx  src/bz_move.cpp:165-296 */
int brille_ir_moveinto(brille_grid* grid, const double* Q, uint64_t n, int32_t memory,
                       double* q_out, int32_t* tau_out, uint32_t* rot_out,
                       uint32_t* invrot_out, uint8_t* status_out, void* stream);

/* ---- many grids in one persistent launch (no reference counterpart) ---- */
/* A batch holds filled grids (not owned; they must outlive it).  A run is one
 * persistent kernel whose blocks take tiles of points from a work counter
 * across all grids, captured once into a HIP graph per precision and schedule
 * and replayed, followed on the same stream by one binned call per grid
 * given at least `binned_min` points; the tile shape, block count, tiles per
 * grab, grid order and binned_min come from brille_batch_autotune (defaults
 * before it runs: no binned grids). */
int brille_batch_create(brille_grid* const* grids, uint32_t n_grids, brille_batch** out);
/* Q: all points, grid g owns rows [offsets[g], offsets[g+1]); values_out[g] and
 * vectors_out[g] receive that grid's results.  Device memory only.  kernel =
 * BRILLE_KERNEL_PLAIN or BRILLE_KERNEL_TILE run the grids one call after
 * another instead (the unbatched baselines). */
int brille_batch_run(brille_batch* batch, const double* Q, const uint64_t* offsets,
                     const brille_run_opts* opts, void* const* values_out,
                     void* const* vectors_out, void* stream);
/* times every candidate schedule on these inputs and keeps the fastest */
int brille_batch_autotune(brille_batch* batch, const double* Q, const uint64_t* offsets,
                          const brille_run_opts* opts, void* const* values_out,
                          void* const* vectors_out, void* stream);
/* the schedule in use: points per tile (= threads per block), persistent
 * blocks, tiles per grab, grid order (0 as given, 1 costliest first), and
 * the point count from which a grid runs its own binned pipeline after the
 * persistent kernel instead (UINT64_MAX: none) */
int brille_batch_schedule(const brille_batch* batch, uint32_t* tile_points, uint32_t* blocks,
                          uint32_t* tiles_per_grab, uint32_t* order, uint64_t* binned_min);
/* sets the schedule (a saved autotune result, or a chosen one): tile_points
 * 128 or 256, blocks per CU and tiles per grab 1-64, order 0 or 1,
 * binned_min as above */
int brille_batch_set_schedule(brille_batch* batch, uint32_t tile_points, uint32_t blocks_per_cu,
                              uint32_t tiles_per_grab, uint32_t order, uint64_t binned_min);
int brille_batch_destroy(brille_batch* batch);

/* ---- learned tetrahedron proposal (no reference counterpart) ---- */
typedef struct brille_train_opts {
  uint64_t n_train;      /* (irreducible q, tetrahedron) pairs drawn from the exact locator */
  uint64_t seed;
  uint32_t epochs;
  uint32_t batch;        /* minibatch size */
  float learning_rate;
} brille_train_opts;
typedef struct brille_train_report {
  uint64_t n_pairs;      /* pairs that landed in polyhedron nodes */
  float final_loss;
  float train_hit_rate;  /* fraction of held-out pairs whose first proposal was right */
  uint32_t n_classes;    /* K: tetrahedron slots per node */
  float ms;
} brille_train_report;
int brille_proposal_train(brille_grid* grid, const brille_train_opts* opts,
                          brille_train_report* report);
/* hit statistics of the last use_proposal call: proposals tried, proposals verified */
int brille_proposal_stats(const brille_grid* grid, uint64_t* tried, uint64_t* hits);
/* forward/backward of the proposal network on caller data, for gradient checks:
 * x (n x 3 float), labels (n), returns loss and writes gradients (n_params float) */
int brille_proposal_loss_grad(brille_grid* grid, const float* x, const int32_t* labels,
                              uint32_t n, float* loss, float* grads, uint32_t* n_params);
int brille_proposal_get_params(const brille_grid* grid, float* params, uint32_t* n_params);
int brille_proposal_set_params(brille_grid* grid, const float* params, uint32_t n_params);
/* the network's proposal for points given by their features (n x 3 float: the
 * Cartesian coordinates inside the point's trellis node, scaled to [-1, 1])
 * and any tetrahedron of their polyhedron node (in_tet selects the
 * candidates); host memory */
int brille_proposal_predict(brille_grid* grid, const float* x, const int32_t* in_tet, uint32_t n,
                            int32_t* tet_out);

/* ---- timing and device plumbing (used by the benchmark and tests) ---- */
/* run the hot path `iters` times on `stream` and report mean times (ms)
 * measured with HIP events: total_ms is one whole call (events on `stream`
 * around all iterations); ms[] are the kernels (fused path) or the stages of
 * the binned pipeline, timed on the stream each ran on (with several passes
 * the binned stages overlap and add up to more than total_ms) */
typedef struct brille_kernel_times {
  uint32_t n_kernels;
  char names[4][32];
  float ms[4];
  float total_ms;        /* mean wall time of one whole call */
} brille_kernel_times;
int brille_time_ir_interpolate_at(brille_grid* grid, const double* Q, uint64_t n,
                                  const brille_run_opts* opts, void* values_out,
                                  void* vectors_out, uint32_t iters, void* stream,
                                  brille_kernel_times* times);
int brille_time_batch_run(brille_batch* batch, const double* Q, const uint64_t* offsets,
                          const brille_run_opts* opts, void* const* values_out,
                          void* const* vectors_out, uint32_t iters, void* stream,
                          brille_kernel_times* times);
int brille_device_alloc(uint64_t bytes, void** ptr);
int brille_device_free(void* ptr);
int brille_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
int brille_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
int brille_stream_create(void** stream);
int brille_stream_destroy(void* stream);
int brille_stream_sync(void* stream);
int brille_device_sync(void);

#ifdef __cplusplus
}
#endif
#endif /* BRILLE_AMD_H */
