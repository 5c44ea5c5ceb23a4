/*
 * ef_hps.h -- C ABI of the MI355X HPS factor-and-solve library (libefhps.so).
 *
 * The library replaces, for one GPU (one MPI rank), the numerical core under the
 * reference's C++ API:
 *   HPSAlgorithm::buildStage / upwardsStage / solveStage   (src/HPSAlgorithm.hpp:120-445)
 *   FiniteVolumeSolver::solve / mapD2N / buildD2N / particularNeumannData
 *                                   (src/Patches/FiniteVolume/FiniteVolumeSolver.cpp:25-462)
 *   Matrix/Vector dgesv, dgemm, dgemv wrappers on that path (src/Matrix.hpp:789-953)
 *   L21 / L12 interpolation operators                   (src/SpecialMatrices.hpp:93-173)
 * The C++ drop-in headers (include/EllipticForest.hpp) and the Python host mirror
 * (ellipticforest_amd/_lib.py) call only these entry points.
 *
 * Conventions
 *  - Every function returns 0 on success or a nonzero EF_ERR_* code; no exception
 *    crosses the ABI.  ef_last_error() returns the message of the calling thread's
 *    last failure.
 *  - Host buffers are borrowed for the duration of a call.  Device arenas (operators,
 *    vectors) are owned by the mesh.  Device-pointer variants (*_device) take memory
 *    the caller allocated on the mesh's device and are ordered on the mesh's stream.
 *  - A leaf cell vector is indexed j + i*N (x-major, y fastest).  Boundary vectors are
 *    WESN blocks, each ordered by increasing coordinate.
 *  - Multi-RHS arrays are right-hand-side major: f and u are [rhs][leaf][cell] with
 *    leaves in z-order (pre-order), r is [rhs][4M] where M is the root side.
 *  - Thread-safe across distinct meshes, not within one mesh.
 */
#ifndef EF_HPS_H_
#define EF_HPS_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EF_OK 0
#define EF_ERR_ARG 1        /* invalid argument (maps to std::invalid_argument)  */
#define EF_ERR_STATE 2      /* call out of order, e.g. solve before build         */
#define EF_ERR_DEVICE 3     /* HIP runtime failure or out of device memory        */
#define EF_ERR_SINGULAR 4   /* zero pivot in an interface or root-boundary solve  */
#define EF_ERR_INTERNAL 5

/* Leaf patch solver (FiniteVolumeSolverType, FiniteVolumeSolver.hpp:17-20). */
#define EF_SOLVER_FIVE_POINT 0   /* alpha*div(beta grad u) + lambda u = f, per-cell coefficients   */
#define EF_SOLVER_FISHPACK 1     /* u_xx + u_yy + lambda0 u = f (alpha, beta ignored, :224-290)      */

/* Arithmetic of the operators. */
#define EF_PRECISION_FP64 0
#define EF_PRECISION_MIXED 1     /* FP32 merge products, FP64 interface inverse, refined solves */
/* ef_mesh_footprint only: or'ed into precision, counts the per-leaf storage of variable
 * coefficients (per-cell alpha, beta, lambda: samples, stencil, block-LU pivots, leaf DtN) */
#define EF_FOOTPRINT_VARIABLE 0x100

typedef struct ef_ctx_s* ef_ctx;
typedef struct ef_mesh_s* ef_mesh;

/* Context: one per device. */
int ef_ctx_create(int device, ef_ctx* out);
int ef_ctx_destroy(ef_ctx ctx);
const char* ef_last_error(void);
int ef_version(void);

/* Quadtree mesh given by its leaves in z-order (pre-order, children SW, SE, NW, NE),
 * each leaf an N x N cell patch; child grids are parent quadrants with the same N
 * (FiniteVolumeNodeFactory.cpp:15-67).  Replaces the p4est-backed
 * Quadtree/Mesh storage of the path (src/Quadtree.hpp:118-196, src/Mesh.hpp:111-180). */
typedef struct {
    int nleaves;
    const int* level;   /* leaf level (root = 0) */
    const int* ix;      /* leaf x index at its level, 0 .. 2^level - 1 */
    const int* iy;
    int N;              /* cells per leaf side (>= 4, even) */
    double xl, xu, yl, yu;
} ef_mesh_desc;

/* Mesh::refineByFunction (src/Mesh.hpp:111-180) without p4est: uniform quadrants at
 * min_level, recursive refinement while level < max_level and fn marks any of the
 * quadrant's nx*ny cell centres, then 2:1 face+corner balance (P4EST_CONNECT_CORNER).
 * fn receives one quadrant's cell centres (index j + i*ny) and returns nonzero to refine.
 * Leaves come back in z-order.  With capacity == 0 only *nleaves is set; with a
 * capacity smaller than the leaf count EF_ERR_ARG is returned and *nleaves is set. */
typedef int (*ef_refine_fn)(const double* x, const double* y, int npts, void* user);
int ef_refine_by_function(ef_refine_fn fn, void* user, int nx, int ny, double xl, double xu, double yl, double yu,
                          int min_level, int max_level, int capacity, int* level, int* ix, int* iy, int* nleaves);

int ef_mesh_create(ef_ctx ctx, const ef_mesh_desc* desc, ef_mesh* out);
int ef_mesh_destroy(ef_mesh mesh);
/* Device bytes the mesh needs for operators plus vectors of max_rhs right-hand sides
 * (precision | EF_FOOTPRINT_VARIABLE for variable-coefficient leaves). */
int ef_mesh_footprint(const ef_mesh_desc* desc, int precision, int max_rhs, size_t* bytes);
/* info[0..7] = nnodes, nleaves, root side M, max depth, N, internal nodes, max interface n, 0 */
int ef_mesh_info(ef_mesh mesh, int* info);
/* Node table in pre-order: per node 12 ints = level, ix, iy, is_leaf, leaf_index,
 * n_coarsens, grid nx, child0..3, parent; and 4 doubles of grid bounds. */
int ef_mesh_nodes(ef_mesh mesh, int* info, double* bounds);
/* Use a caller-owned hipStream_t (passed as void*); NULL restores the mesh's own stream. */
int ef_mesh_set_stream(ef_mesh mesh, void* hip_stream);
/* The stream the mesh's stages run on (its own, or the one given to ef_mesh_set_stream), for
 * timing queued stages with ef_timer_* (no counterpart in the reference, whose stages run on
 * the host). */
int ef_mesh_stream(ef_mesh mesh, void** hip_stream);

/* Stage options of EllipticForestApp (HPSAlgorithm.hpp:134, 532, 587, 1202). */
int ef_mesh_set_options(ef_mesh mesh, int cache_operators, int homogeneous_rhs);

/* Leaf coefficients for ef_build.
 *  FISHPACK:   lambda0 = lambda(0,0) (FiniteVolumeSolver.cpp:262).
 *  FIVE_POINT: cell != NULL -> per leaf 6 arrays of N*N doubles, in z-order:
 *              alpha(x_i,y_j), beta(x_i -+ dx/2, y_j) (W, E), beta(x_i, y_j -+ dy/2) (S, N),
 *              lambda(x_i,y_j), sampled as FiniteVolumeSolver.cpp:64-78 samples them;
 *              cell == NULL -> constant alpha0, beta0, lambda0. */
typedef struct {
    int solver_type;
    double alpha0, beta0, lambda0;
    const double* cell;
} ef_coeffs;

/* buildStage: leaf DtN operators and all 4-to-1 merges, level by level. */
int ef_build(ef_mesh mesh, const ef_coeffs* coeffs, int precision);

/* upwardsStage with load data f ([rhs][leaf][N*N], host) for nrhs right-hand sides. */
int ef_upwards(ef_mesh mesh, const double* f, int nrhs);
int ef_upwards_device(ef_mesh mesh, const double* f_dev, int nrhs);

/* solveStage with the general boundary condition a u + b du/dn = r at root face
 * points (HPSAlgorithm.hpp:343-445).  a and b have 4M entries shared by the batch,
 * r is [rhs][4M]; u receives [rhs][leaf][N*N].  b == NULL means Dirichlet (b = 0). */
int ef_solve(ef_mesh mesh, const double* r, const double* a, const double* b, double* u, int nrhs);
/* Device variant: f_dev is the load of the preceding upwards call (the leaf solves
 * use it, as leafSolve uses vectorF), all pointers on the mesh's device. */
int ef_solve_device(ef_mesh mesh, const double* f_dev, const double* r_dev, const double* a_dev,
                    const double* b_dev, double* u_dev, int nrhs);

/* Copy one stored operator or vector of a node to the host, as the reference patch
 * holds it after the stages: which = 0 T, 1 S, 2 X, 3 H, 4 u, 5 g, 6 f, 7 h, 8 w
 * (FiniteVolumePatch accessors, FiniteVolumePatch.cpp:30-68); rhs selects the column.
 * With out == NULL only rows/cols are returned. */
int ef_get_operator(ef_mesh mesh, int node, int which, int rhs, double* out, int* rows, int* cols);

/* Wait for all work on the mesh's stream. */
int ef_mesh_sync(ef_mesh mesh);

/* Device memory of the context's GPU for the *_device entry points (callers that do not
 * bring their own allocator).  kind: 0 host->device, 1 device->host, 2 device->device
 * (whole, aligned doubles run on the library's copy kernel).  All three return when the
 * copy is complete. */
int ef_device_alloc(ef_ctx ctx, size_t bytes, void** ptr);
int ef_device_free(ef_ctx ctx, void* ptr);
int ef_memcpy(ef_ctx ctx, void* dst, const void* src, size_t bytes, int kind);

/* Tile autotuner of the sweeps (no counterpart in the reference): at the first upwards or solve
 * stage of each number of right-hand sides, every level's GEMMs are timed with 16x16, 16x32 and
 * 32x32 blocks per wave and the fastest is kept (results do not depend on the choice).  The
 * timed passes overwrite the upward state, so the solve stage after a newly tuned width needs its
 * own upwards stage (as it does anyway).  mode 1 (default) tunes; 0 (or EF_AUTOTUNE=0) keeps the
 * fixed rule (16x16 up to 16 right-hand sides, else 32x32, 16x32 for operators of <= 16 rows);
 * 11, 12 or 22 forces one tile (rows x 10 + columns, in units of 16) everywhere.  Setting it drops
 * the sweep plans and the upward state.
 * ef_mesh_sweep_tiles: the tiles for nrhs right-hand sides, per depth (root first) the w, h and u
 * products as rows x 10 + columns (0: no product at that depth); tiles holds 3 x *levels ints and
 * may be NULL to query *levels. */
int ef_mesh_set_autotune(ef_mesh mesh, int mode);
int ef_mesh_sweep_tiles(ef_mesh mesh, int nrhs, int* tiles, int* levels);

/* Mixed precision (EF_PRECISION_MIXED at ef_build; no counterpart in the reference, which is
 * FP64 throughout): the S and T products of every merge on FP32 matrix cores (FP32 operands and
 * accumulation), X^-1 and the leaf operators in FP64; each solve stage then runs `steps` rounds
 * of iterative refinement: the FP64 residual of the five-point system, a correction solve with
 * it as the load and zero Dirichlet data, u += correction.  Uniform meshes, Dirichlet root data,
 * homogeneous-rhs off; a mixed solve consumes its upwards stage.  Default 2 rounds. */
int ef_mesh_set_refinement(ef_mesh mesh, int steps);

/* Serving data path (no counterpart in the reference: the request loop keeps factorised
 * meshes resident and feeds them batches of right-hand sides).
 * ef_host_alloc / ef_host_free: page-locked host memory, so ef_mesh_copy_async really is
 * asynchronous.  ef_mesh_copy_async: hipMemcpyAsync on the mesh's stream (kind as ef_memcpy),
 * ordered with the mesh's stages.  ef_mesh_record / ef_mesh_query / ef_mesh_wait: an event
 * per slot (0..7) recorded on the mesh's stream, polled, or waited for.
 * ef_mesh_eval_load: the serving workload's loads for nrhs requests from 64 parameters each
 * (params_dev, device): f_dev[rhs][leaf][N*N] = sum of 8 Gaussians (amp, cx, cy, w; amp 0
 * unused) amp exp(-|p - c|^2 / (2 w^2)) at the leaf cell centres, r_dev[rhs][4M] = sum of 8
 * modes (a, kx, ky, phi) a cos(kx x + ky y + phi) at the root boundary points; queued on the
 * mesh's stream.  Each Gaussian is evaluated as the product of its x and y factors (equal to
 * the exponential of the sum up to rounding), computed once per distinct cell-centre coordinate
 * of the mesh and request in a work buffer the mesh keeps (8 (X + Y) doubles per request, X and
 * Y the distinct centre coordinates). */
int ef_host_alloc(ef_ctx ctx, size_t bytes, void** ptr);
int ef_host_free(ef_ctx ctx, void* ptr);
int ef_mesh_copy_async(ef_mesh mesh, void* dst, const void* src, size_t bytes, int kind);
int ef_mesh_record(ef_mesh mesh, int slot);
int ef_mesh_query(ef_mesh mesh, int slot, int* done);
int ef_mesh_wait(ef_mesh mesh, int slot);
int ef_mesh_eval_load(ef_mesh mesh, const double* params_dev, int nrhs, double* f_dev, double* r_dev);

/* Diagnostic: the tree GEMMs of one upwards + solve step with nrhs right-hand sides alone
 * (w = X^-1 hDiff and h = H w per level up, u = S g per level down: mergeW_/mergeH_/applyS_,
 * HPSAlgorithm.hpp:1059-1150, 1165-1230), reps passes between two HIP events on the mesh's
 * stream over whatever the vector arena holds; ms per pass and kernel launches per pass. */
int ef_mesh_time_sweep_gemms(ef_mesh mesh, int nrhs, int reps, double* ms_per_pass, int* launches);

/* Diagnostic: the batched inversion every merge uses (blocked Gauss-Jordan with partial
 * pivoting, tournament pivoting over 512-row chunks), applied to B host matrices of order m
 * (row-major, B x m x m).  Ainv receives the inverses; ipiv (B x m, may be NULL) the
 * LAPACK-order row interchanges (0-based: position c was exchanged with ipiv[c]), equal to
 * dgetrf's wherever no step spans more than one chunk (m <= 512).  Replaces the LU of the
 * interface system in mergeX_/mergeS_ (HPSAlgorithm.hpp:906-929) for testing only.
 * EF_ERR_SINGULAR when a pivot is zero. */
int ef_debug_invert(ef_ctx ctx, int m, int B, const double* A, double* Ainv, int* ipiv);

/* Stage timers measured with HIP events on the mesh's stream, in milliseconds, of the
 * last build, upwards and solve calls: ms[0] build-stage, ms[1] upwards-stage,
 * ms[2] solve-stage (the timer names of HPSAlgorithm.hpp:125, 183, 295). */
int ef_mesh_stage_ms(ef_mesh mesh, double* ms);

/* ------------------------------------------------------------------------------------
 * Batched leaf patches: the FiniteVolumeSolver operations on one grid
 * (FiniteVolumeSolver.hpp:122-149, FiniteVolumeSolver.cpp:25-462) applied to many
 * independent N x N patches at once.  Patches with constant coefficients (FISHPACK
 * semantics, or five-point with alpha, beta, lambda constant within the patch) use fast
 * diagonalisation (in registers for N <= 32, as batched GEMMs beyond); otherwise the
 * five-point matrix is factorised once at creation (block-tridiagonal, N <= 32).
 * Layouts: g and h are [rhs][patch][4N] (WESN), f and u are [rhs][patch][N*N].
 * ------------------------------------------------------------------------------------ */
typedef struct ef_patches_s* ef_patches;

typedef struct ef_patch_batch {
    int npatches;
    int N;                  /* cells per side, 4 <= N <= 1024 (N <= 32 when the
                               coefficients vary within a patch)             */
    const double* bounds;   /* [npatches][4]: xl, xu, yl, yu of each patch   */
} ef_patch_batch;

#define EF_PATCH_SOLVE 0     /* FiniteVolumeSolver::solve (:25-294): u from g, f            */
#define EF_PATCH_MAP_D2N 1   /* FiniteVolumeSolver::mapD2N (:296-353): h from g, f;
                                with g == NULL, particularNeumannData (:457-462)            */

/* Set up a batch.  Coefficients as for ef_build; per-cell samples (coeffs->cell) are
 * [npatches][6][N*N] (alpha, beta W/E/S/N faces, lambda). */
int ef_patches_create(ef_ctx ctx, const ef_patch_batch* batch, const ef_coeffs* coeffs, ef_patches* out);
int ef_patches_destroy(ef_patches p);
int ef_patches_set_stream(ef_patches p, void* stream);
/* info[0] npatches, info[1] N, info[2] 0 for the fast-diagonalisation path (constant coefficients
 * per patch), 1 for the factorised (variable-coefficient) path, 2 for the factorised path with
 * packed symmetric inverse pivot blocks (symmetric five-point matrices, N in {8, 16, 32}) */
int ef_patches_info(ef_patches p, int* info);
/* op = EF_PATCH_SOLVE (out = u) or EF_PATCH_MAP_D2N (out = h); g or f may be NULL (zero). */
int ef_patches_apply(ef_patches p, int op, const double* g, const double* f, double* out, int nrhs);
int ef_patches_apply_device(ef_patches p, int op, const double* g_dev, const double* f_dev, double* out_dev,
                            int nrhs);
/* ef_patches_apply_device without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Work space (the forward rows of packed variable-coefficient
 * Neumann maps, the GEMM buffers of patches with N > 32) comes from one fixed array of 2^22
 * doubles (32 MiB) per GPU, kept until the process ends; a batch needing more is refused with
 * EF_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that overlap on one GPU
 * (different streams, different threads) race.  Nothing in the array needs to start at zero.
 * The call records no timer events (ef_patches_last_ms does not see it). */
int ef_patches_apply_device_noalloc(ef_patches p, int op, const double* g_dev, const double* f_dev,
                                    double* out_dev, int nrhs);
/* FiniteVolumeSolver::buildD2N (:355-455) of every patch: T is [npatches][4N][4N]. */
int ef_patches_dtn(ef_patches p, double* T);
int ef_patches_sync(ef_patches p);
/* HIP-event time (ms) of the kernels of the last apply or dtn call on the batch's stream. */
int ef_patches_last_ms(ef_patches p, double* ms);
/* The stream the batch's kernels run on (its own, or the one given to ef_patches_set_stream). */
int ef_patches_stream(ef_patches p, void** stream);

/* Device timers: a pair of HIP events recorded on a stream (NULL: the null stream), for
 * timing a sequence of launches on the GPU's clock, as the reference's Timer objects
 * time its stages on the host (EllipticForestApp timers, src/EllipticForest.hpp). */
typedef struct ef_timer_s* ef_timer;
int ef_timer_create(ef_ctx ctx, ef_timer* out);
int ef_timer_destroy(ef_timer t);
/* which = 0: start event, 1: stop event */
int ef_timer_record(ef_timer t, void* stream, int which);
/* Waits for the stop event; milliseconds between the two events. */
int ef_timer_elapsed_ms(ef_timer t, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* EF_HPS_H_ */
