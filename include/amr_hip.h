/*
 * amr_hip.h -- C ABI of the MI355X-native patch-AMR advection step.
 *
 * One opaque context (amr_ctx) owns one periodic quad-/octree of leaf patches,
 * all of its device memory and one HIP stream.  Every entry point returns an
 * int status (AMR_OK on success) and never throws; amr_last_error() gives the
 * message of the last failure on the calling thread.  Arguments are plain
 * pointers and sizes; host arrays are owned by the caller.
 *
 * Each entry point names the reference interface it replaces
 * (bgce-cse/gpu-amr, paths relative to the reference root).
 *
 * Data conventions (the reference's):
 *   - key = (interleave(x, y[, z]) << 6) | level, x in bit 0 of each group
 *     (include/morton/morton_id.hpp:132-140, :350-359); leaves are kept
 *     sorted by key (include/ndtree/ndtree.hpp:1335-1414).
 *   - a patch is n_0 x .. x n_(rank-1) interior cells (amr_tree_config), row-
 *     major with the last index fastest; physical x is the last index
 *     (include/containers/static_layout.hpp:28-37,
 *     include/solver/physics_system.hpp:126-135).  The padded view adds
 *     halo_width ghost layers per side (prod (n_d + 2h) doubles per patch).
 *   - direction k addresses layout dim k/2, negative when k is even
 *     (include/ndtree/neighbor.hpp:156-211).
 */
#ifndef AMR_HIP_H
#define AMR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference throws std::runtime_error on device errors
 * (src/cuda/device_buffer.cu:17-27) and leaves capacity unchecked
 * (include/ndtree/ndtree.hpp:1304-1316); here both become codes. */
enum amr_status
{
    AMR_OK              = 0,
    AMR_ERR_HIP         = 1, /* a HIP runtime call failed                      */
    AMR_ERR_CAPACITY    = 2, /* the leaf count would exceed the capacity        */
    AMR_ERR_ARGUMENT    = 3, /* bad pointer / size / configuration              */
    AMR_ERR_UNSUPPORTED = 4, /* configuration or operation not provided         */
    AMR_ERR_TOPOLOGY    = 5  /* leaves do not tile the domain or break 2:1      */
};

typedef struct amr_ctx amr_ctx;

/* Template parameters of ndtree<Cell, morton_id<Depth, Rank>,
 * patch_layout<static_layout<static_shape<N, M[, O]>>, Halo>,
 * linear_interpolator> and physics_system<..., Domain_Sizes>
 * (include/ndtree/ndtree.hpp:94-99, include/containers/static_layout.hpp:28-37,
 * include/solver/physics_system.hpp:13-17), as run-time values.
 *
 * The patch shape is patch_extents[0..rank-1] = N, M(, O): the interior
 * cells of layout dims 0..rank-1, the last being physical x; or, with all
 * three extents 0, patch_size cells in every dim.  Extents must be even,
 * within 2..4096 and at least 2 * halo_width (AMR_ERR_ARGUMENT otherwise).
 * Any such shape runs; cubic 16^2, 8^3, 10^2 and 50^2 (halo 2), 64^2, 4^2,
 * 8^2, 4^3 have kernels specialised for their size.
 *
 * Face ghosts from finer neighbours take the fine cells at (2 v) mod n(d)
 * along every dim d, as the reference's device path does
 * (src/cuda/halo_exchange.cu:174-176); its CPU path reduces every dim mod
 * the face normal's extent (include/ndtree/patch_utils.hpp:358), which
 * agrees for square patches only and otherwise addresses cells outside the
 * fine patch. */
typedef struct amr_tree_config
{
    int32_t  rank;             /* 2 or 3                                        */
    int32_t  max_depth;        /* Morton depth (7 in 2D, 5 in 3D in the configs) */
    int32_t  patch_size;       /* interior cells per dimension (square / cubic)  */
    int32_t  halo_width;       /* ghost layers of the padded view (1..8)        */
    uint64_t capacity;         /* maximum number of leaf patches                */
    double   domain_length[3]; /* physical lengths of x, y, z                   */
    int32_t  patch_extents[3]; /* N, M(, O) per layout dim; all 0: patch_size   */
    int32_t  reserved;         /* 0                                             */
} amr_tree_config;

/* One (leaf, direction) neighbour record in the reference's flattened form
 * (include/cuda/halo_exchange.hpp:11-26): relation 0 none, 1 same, 2 finer,
 * 3 coarser; finer ids ordered little-endian over the non-normal layout
 * dims; quadrant = contact quadrant in the coarser leaf, per layout dim. */
typedef struct amr_halo_direction_metadata
{
    int32_t neighbor;
    int32_t finer_ids[4];
    int32_t quadrant[3];
    int8_t  relation;
    int8_t  padding[3];
} amr_halo_direction_metadata;

/* Refinement criterion evaluated on the device (the fill_refine_flags form,
 * include/ndtree/ndtree.hpp:889-894).  kind 0: maximum cell value over the
 * interior and face halos, starting from 0, as the reference's CPU criterion
 * functors compute it (examples/fvm_solver_advection.e.cpp:75-90,
 * tests/advection_equation_amr.t.cpp:155-170).  The reference's CUDA kernel
 * (src/cuda/fvm_refinement_criterion.cu:27-67) starts from -DBL_MAX instead
 * and also scans the corner cells, which no halo exchange writes; its flags
 * differ from these only on patches whose values are all negative, with a
 * threshold between their maximum and 0.  kind 1: maximum undivided
 * gradient max(|u[c+s]-u[c]|, |u[c]-u[c-s]|) over interior cells and dims
 * (no reference counterpart).  Refine if value > refine_threshold and
 * level < max_level, else Coarsen if value < coarsen_threshold and
 * level > min_level, else Stable (fvm_refinement_criterion.cu:53-66). */
typedef struct amr_criterion
{
    int32_t kind;
    int32_t min_level;
    int32_t max_level;
    int32_t reserved;
    double  refine_threshold;
    double  coarsen_threshold;
} amr_criterion;

enum amr_criterion_kind
{
    AMR_CRITERION_MAX_VALUE = 0,
    AMR_CRITERION_GRADIENT  = 1
};

/* Message of the last failed call on this thread (empty if none). */
const char* amr_last_error(void);

/* ---- tree (include/ndtree/ndtree.hpp) ---------------------------------- */

/* ndtree(capacity): allocates the capacity and appends the periodic root
 * (ndtree.hpp:320-425). */
int amr_tree_create(const amr_tree_config* cfg, amr_ctx** out);
/* ~ndtree (ndtree.hpp:427-468). */
int amr_tree_destroy(amr_ctx* ctx);
/* ndtree::size (ndtree.hpp:471-475). */
int amr_tree_size(const amr_ctx* ctx, uint64_t* n);
/* ndtree::get_node_index_at for every leaf, in linear order
 * (ndtree.hpp:1290-1295). */
int amr_tree_get_keys(amr_ctx* ctx, uint64_t* host_keys, uint64_t n);
/* ndtree::rebuild_halo_exchange_metadata + device copy-out
 * (ndtree.hpp:1606-1660); count = size * 2 * rank records. */
int amr_tree_get_neighbor_metadata(amr_ctx* ctx, amr_halo_direction_metadata* host, uint64_t count);
/* ndtree::reconstruct_tree(fn) with the per-leaf functor already evaluated
 * into host_flags (0 Stable, 1 Refine, 2 Coarsen; ndtree.hpp:886-902,
 * 1249-1271).  *changed is set to 1 when the topology changed. */
int amr_tree_reconstruct_flags(amr_ctx* ctx, const int8_t* host_flags, uint64_t n, int* changed);
/* ndtree::reconstruct_tree(criterion) with fill_refine_flags on the device
 * (ndtree.hpp:889-894; src/cuda/fvm_refinement_criterion.cu:71-108). */
int amr_tree_reconstruct_criterion(amr_ctx* ctx, const amr_criterion* crit, int* changed);
/* The flags the criterion produces for the current leaves (copied out;
 * the reference's sync_refine_status_from_device, ndtree.hpp:639-656). */
int amr_tree_compute_flags(amr_ctx* ctx, const amr_criterion* crit, int8_t* host_flags, uint64_t n);
/* amr_tree_compute_flags without any device allocation: no counterpart in
 * the reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-leaf flags are staged in a fixed 1 MiB device array that each GPU
 * holds until the process ends, so trees of up to 1,048,576 leaves are
 * accepted; larger trees return AMR_ERR_CAPACITY.  All calls on one GPU
 * share that array: calls that overlap on one GPU (two threads, or two
 * contexts' streams) race.  Results are those of amr_tree_compute_flags. */
int amr_tree_compute_flags_noalloc(amr_ctx* ctx, const amr_criterion* crit, int8_t* host_flags, uint64_t n);

/* ---- fields ------------------------------------------------------------ */

/* Write / read the interiors of all leaves (fields * size * n^rank doubles:
 * field after field, leaf after leaf within a field; fields = 1 for
 * advection, rank + 2 for Euler): the host side of amr_solver::initialize
 * (amr_solver.hpp:105-145) and sync_current_to_device /
 * sync_current_from_device (ndtree.hpp:583-591). */
int amr_tree_set_interior(amr_ctx* ctx, const double* host, uint64_t count);
int amr_tree_get_interior(amr_ctx* ctx, double* host, uint64_t count);
/* Padded patches after ndtree::halo_exchange_update (ndtree.hpp:1862-1877):
 * fields * size * (n+2h)^rank doubles, field-major; corner cells are 0
 * (never written). */
int amr_tree_get_patches(amr_ctx* ctx, double* host, uint64_t count);
/* ndtree::halo_exchange_update (ndtree.hpp:1862-1877): fills the face halos
 * of the device padded view.  Time steps gather ghosts from neighbour
 * interiors directly, so stepping never needs this call. */
int amr_tree_halo_exchange_update(amr_ctx* ctx);
/* ndtree::swap_buffers (ndtree.hpp:1558-1579). */
int amr_tree_swap_buffers(amr_ctx* ctx);
/* Device pointer of the current interiors (size * n^rank doubles) and of the
 * context's stream, for callers that share device memory (PyTorch). */
int amr_tree_device_interior(amr_ctx* ctx, double** dev_ptr);
int amr_tree_stream(amr_ctx* ctx, void** hip_stream);

/* vtk_print<physics_system>::print (include/ndtree/vtk_print.hpp:53-254):
 * writes the legacy-VTK binary (big-endian) unstructured grid of the current
 * leaves to `path` -- one quad / voxel per interior cell, cell_index,
 * is_halo (as the reference computes it) and one double scalar per field
 * ("Scalar"; Euler: "Rho", "Rhou", "Rhov", ("Rhow"), "E2D" / "E3D"). */
int amr_tree_write_vtk(amr_ctx* ctx, const char* path);

/* Number of fields per cell: 1 for advection, rank + 2 for Euler (the
 * FieldTags of the equation, include/solver/EulerPhysics.hpp:22-30). */
int amr_tree_fields(const amr_ctx* ctx, int32_t* fields);

/* ---- solver (include/solver/amr_solver.hpp) ---------------------------- */

/* amr_solver(capacity, gamma, cfl) parameters plus the advection velocity
 * that AdvectionPhysics hard-codes (AdvectionPhysics.hpp:24,54,82). */
int amr_solver_configure(amr_ctx* ctx, const double velocity[3], double cfl);
/* compute_time_step_cpu / compute_dt_kernel: cfl * min dx_d / |v_d| over the
 * leaves present (amr_solver.hpp:355-413). */
int amr_solver_compute_dt(amr_ctx* ctx, double* dt);
/* advance_batch_async(step_count, remaining_time) (amr_solver.hpp:155-243):
 * enqueues up to step_count steps on the context stream, each clamped to the
 * remaining time; returns without waiting. */
int amr_solver_advance_batch_async(amr_ctx* ctx, uint64_t step_count, double remaining_time);
/* finish_advance_batch (amr_solver.hpp:245-262): waits for the batch and
 * returns the summed dt and the executed step count. */
int amr_solver_finish_advance_batch(amr_ctx* ctx, double* dt_sum, uint64_t* executed_steps);
/* advance() = advance_batch_async(1) + finish_advance_batch
 * (amr_solver.hpp:149-153). */
int amr_solver_advance(amr_ctx* ctx, double* dt);

/* ---- Euler physics (include/solver/EulerPhysics.hpp) ------------------- */

/* amr_solver<tree, physics, EulerPhysics<rank>, rank>(capacity, gamma, cfl)
 * (amr_solver.hpp:28-97) with its tree: a context whose cells carry the
 * rank + 2 conservative variables [rho, rho u, rho v, (rho w), E].  The tree
 * entry points (regrid by flags or criterion -- evaluated on rho, as the
 * reference's Euler drivers do, benchmark/bench_fvm_solver_integration3D.b.cpp:86-107
 * -- keys, tables, fields, padded view) apply to it; amr_solver_compute_dt,
 * amr_solver_advance_batch_async, amr_solver_finish_advance_batch and
 * amr_solver_advance run the Euler step: Rusanov fluxes
 * (EulerPhysics::rusanovFlux, EulerPhysics.hpp:72-125) and, every step, the
 * CFL dt from the maximum wave speed (compute_time_step_cpu,
 * amr_solver.hpp:355-413; getMaxSpeed, EulerPhysics.hpp:132-155) clamped to
 * the remaining time, computed on the device.  Step variants, fusion, batches
 * and the adjoint tape are advection-only (AMR_ERR_UNSUPPORTED). */
int amr_euler_create(const amr_tree_config* cfg, double gamma, double cfl, amr_ctx** out);
int amr_euler_configure(amr_ctx* ctx, double gamma, double cfl);
/* amr_solver::initialize's conversion (amr_solver.hpp:130-134,
 * EulerPhysics::primitiveToConservative, EulerPhysics.hpp:39-62): primitive
 * values [rho, u, v, (w), p], field-major like amr_tree_set_interior, are
 * converted per cell and written as the current state. */
int amr_euler_set_primitive(amr_ctx* ctx, const double* host_prim, uint64_t count);

/* ---- step-kernel autotuner (no counterpart in the reference) ----------- */

/* Number / name of the step-kernel variants built for this patch shape.  All
 * variants compute bit-identical results; they differ in the leaf-to-
 * workgroup mapping and in how many workgroups share a CU. */
int amr_solver_step_variants(amr_ctx* ctx, int32_t* count);
int amr_solver_step_variant_name(amr_ctx* ctx, int32_t variant, const char** name);
int amr_solver_set_step_variant(amr_ctx* ctx, int32_t variant);
int amr_solver_get_step_variant(amr_ctx* ctx, int32_t* variant);
/* Times every variant for trial_steps launches on the current tree (reading
 * the current state, writing only the scratch buffer: the solution is not
 * changed), selects the fastest and returns its index and time per step. */
int amr_solver_autotune(amr_ctx* ctx, uint32_t trial_steps, int32_t* chosen, float* best_us);

/* Two-step fusion (no counterpart in the reference, which runs one halo
 * exchange and one update per step, amr_solver.hpp:221-241).  On a tree
 * whose leaves all sit on one level, advance_batch_async runs consecutive
 * steps in pairs, each pair as one launch that reads u^n once and writes
 * u^{n+2} once; the results are bit-identical to single steps.  Enabled by
 * default.  `applies` reports whether the current tree qualifies (whether
 * or not fusion is enabled). */
int amr_solver_set_step_fusion(amr_ctx* ctx, int32_t enable);
int amr_solver_get_step_fusion(amr_ctx* ctx, int32_t* enabled, int32_t* applies);

/* Waits for all work queued on the context stream. */
int amr_ctx_synchronize(amr_ctx* ctx);

/* ---- batched instances (no counterpart in the reference, whose drivers run
 * one tree per process) -------------------------------------------------- */

typedef struct amr_batch amr_batch;

/* Concatenates the current state of `count` trees of the same rank, depth,
 * patch size and halo into one leaf list whose time steps run as one launch
 * over all leaves; every instance keeps its own velocity, cfl and dt (fixed
 * at creation from its finest level).  The trees are not modified until
 * amr_batch_scatter. */
int amr_batch_create(amr_ctx* const* trees, uint32_t count, amr_batch** out);
int amr_batch_destroy(amr_batch* batch);
int amr_batch_size(amr_batch* batch, uint64_t* total_leaves);
/* Every instance advances `steps` steps.  use_graph != 0 captures the steps
 * in a hipGraph once (per step count, buffer parity and kernel variant) and
 * replays it as a single graph launch.  Returns without waiting. */
int amr_batch_advance(amr_batch* batch, uint64_t steps, int use_graph);
int amr_batch_synchronize(amr_batch* batch);
/* Writes each instance's current interiors back into its source tree. */
int amr_batch_scatter(amr_batch* batch);
int amr_batch_stream(amr_batch* batch, void** hip_stream);
int amr_batch_set_step_variant(amr_batch* batch, int32_t variant);
/* As amr_solver_autotune, over the whole batch. */
int amr_batch_autotune(amr_batch* batch, uint32_t trial_steps, int32_t* chosen, float* best_us);

/* ---- discrete adjoint (no counterpart in the reference, which has no
 * differentiation; the forward pass it differentiates is the solve loop of
 * include/solver/amr_solver.hpp:149-262 with reconstruct_tree calls,
 * include/ndtree/ndtree.hpp:1249-1271, between steps) -------------------- */

typedef struct amr_tape amr_tape;

typedef struct amr_tape_info
{
    uint64_t steps;            /* recorded time steps                         */
    uint64_t checkpoint_every; /* state kept before every k-th step           */
    uint64_t checkpoints;
    uint64_t regrids;          /* regrid calls (changed or not)               */
    uint64_t topologies;       /* distinct trees the solve went through       */
    uint64_t initial_leaves;
    uint64_t final_leaves;
    uint64_t max_leaves;
    uint64_t cells_per_leaf;
} amr_tape_info;

/* Advances the context `steps` steps (dt from the cfl condition of the
 * current finest level, as amr_solver_advance), calling the regrid of
 * amr_tree_reconstruct_criterion with `criterion` before every step k > 0
 * with k % regrid_every == 0 (criterion NULL or regrid_every 0: no regrid),
 * and records what the adjoint needs: each topology, the dt of every step,
 * every regrid's flags and leaf maps, and the state before every
 * checkpoint_every-th step.  The context ends at the final state. */
int amr_tape_record(amr_ctx* ctx, const amr_criterion* criterion, uint64_t steps, uint64_t regrid_every,
                    uint64_t checkpoint_every, amr_tape** out);
int amr_tape_destroy(amr_tape* tape);
int amr_tape_get_info(const amr_tape* tape, amr_tape_info* info);
/* The dt of every recorded step (count >= steps). */
int amr_tape_get_dts(const amr_tape* tape, double* dts, uint64_t count);
/* Regrid `index` (0-based, in step order): the step it preceded, the leaf
 * count before it, whether it changed the tree, and its refine flags
 * (flags_count >= *leaves_before; flags may be NULL). */
int amr_tape_get_regrid(const amr_tape* tape, uint64_t index, uint64_t* step, uint64_t* leaves_before, int* changed,
                        int8_t* flags, uint64_t flags_count);
/* Keys of topology `index` (0 = initial tree); *leaves gets its leaf count. */
int amr_tape_get_keys(const amr_tape* tape, uint64_t index, uint64_t* keys, uint64_t count, uint64_t* leaves);
/* u(T) from initial interiors u0 (initial_leaves * cells_per_leaf doubles)
 * and velocity v, over the recorded topologies and dts (dt does not follow
 * v).  Device pointers; the call waits for all prior device work before
 * reading and for its own work before returning.  Refreshes the checkpoints:
 * the next amr_tape_adjoint differentiates this pass.  The pass runs as a
 * hipGraph captured on the first call (amr_tape_adjoint likewise); the
 * velocity-dependent constants are copied to the device before each replay,
 * so a new velocity needs no new capture. */
int amr_tape_forward(amr_tape* tape, const double* u0_device, const double velocity[3], double* uT_device);
/* grad_u0 = (du(T)/du0)^T lambda_T and grad_velocity = (du(T)/dv)^T lambda_T
 * of the most recent forward pass (the recording or amr_tape_forward), with
 * the states between checkpoints recomputed.  lambda_T and grad_u0 are
 * device pointers (final_leaves and initial_leaves patches); grad_velocity
 * is a host array of 3 (unused dimensions 0). */
int amr_tape_adjoint(amr_tape* tape, const double* lambda_T_device, double* grad_u0_device, double grad_velocity[3]);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* AMR_HIP_H */
