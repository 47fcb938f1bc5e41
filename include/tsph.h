/*
 * tsph.h -- C ABI of libtsph, the MI355X (gfx950) particle pipeline of
 * tensorsph_amd.
 *
 * Plain C: status codes, POD parameter structs, raw device pointers, counts
 * and a hipStream_t.  No C++ types or exceptions cross this boundary.  Every
 * call is asynchronous on the given stream unless its comment says it
 * synchronises (the *_search / *_export calls that return a count, and
 * tsph_sim_step's Verlet-skin decision).
 *
 * Layouts.  Coordinates and velocities are FP64 (N, D) row-major, the layout
 * of an (N, D) torch tensor.  Per-particle scalars (mass, density, pressure)
 * and accelerations are FP64 in TSPH_F64 mode and FP32 in TSPH_F32 mode
 * ("FP32 compensated": pair deltas formed in FP64, kernel arithmetic in FP32,
 * per-particle sums Kahan-compensated).  Particle sets are fluid first, wall
 * after, as the reference orders them.
 *
 * The reference (cxs503/TensorSPH) is pure Python; the functions below
 * replace the Python call sites cited next to each declaration.  The Python
 * package tensorsph_amd binds them with ctypes (INTEGRATION.md shows the
 * binding).
 */
#ifndef TSPH_H
#define TSPH_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSPH_ABI_VERSION 1

/* status codes; the Python layer maps them to the reference's exceptions */
#define TSPH_OK 0
#define TSPH_ERR_ARG 1      /* ValueError */
#define TSPH_ERR_HIP 2      /* RuntimeError (HIP failure) */
#define TSPH_ERR_OOM 3      /* MemoryError */
#define TSPH_ERR_NUMERIC 4  /* FloatingPointError */
#define TSPH_ERR_TYPE 5     /* TypeError */

/* precision / dtype selectors */
#define TSPH_F64 0
#define TSPH_F32 1

typedef struct tsph_ctx tsph_ctx;

/* Message of the last failing call on this thread. */
const char* tsph_last_error(void);
int tsph_abi_version(void);

/* Layout of the parameter structs below as this library was compiled: entry
 * `index` names a struct and one of its fields with the field's offset and
 * size and the struct's size (bytes).  Indices run from 0 until the call
 * returns TSPH_ERR_ARG.  No counterpart in the reference (which has no C
 * ABI): it lets a binding that restates the structs check its copy. */
int tsph_abi_field(int index, const char** struct_name, const char** field_name, int64_t* offset,
                   int64_t* size, int64_t* struct_size);

/* Context: owns device scratch (cell grid, neighbour lists) that grows on
 * demand; one per simulation / stream.  Not thread-safe across threads. */
int tsph_ctx_create(tsph_ctx** out);
int tsph_ctx_destroy(tsph_ctx* ctx);

/* ---------------------------------------------------------------- pair search
 * Replaces neighbor_search.cell_linked_pairs (neighbor_search.py:44-120) and,
 * with box != NULL, threed.neighbors3.periodic_pairs (neighbors3.py:71-142).
 * pos: device (n, dim) in `dtype` (TSPH_F64 / TSPH_F32; the search runs in that
 * precision, as the reference does).  box: host array of dim edges or NULL.
 * fluid_count >= 0 drops pairs with both ends >= fluid_count (the
 * PairListCache prune, neighbor_search.py:208-217); -1 keeps all.
 * Synchronises once and returns the number of unordered pairs in *npairs. */
int tsph_pairs_search(tsph_ctx* ctx, const void* pos, int dtype, int64_t n, int dim,
                      double cutoff, const double* box, int64_t fluid_count,
                      hipStream_t stream, int64_t* npairs);

/* The O(N^2) search of neighbor_search.brute_force_pairs (neighbor_search.py:
 * 29-41) and, with box != NULL, threed.neighbors3.brute_force_pairs_min_image
 * (neighbors3.py:145-170): every pair i < j is tested with the same squared
 * distance as the linked-cell search, and nothing else of that search is
 * used, so the two can check each other.  At most 2^20 particles.
 * Synchronises once; tsph_pairs_export writes the pairs. */
int tsph_pairs_brute(tsph_ctx* ctx, const void* pos, int dtype, int64_t n, int dim, double cutoff,
                     const double* box, hipStream_t stream, int64_t* npairs);

/* Writes the pairs of the last search as (npairs, 2) int64, i < j, sorted
 * lexicographically: the reference's output format. */
int tsph_pairs_export(tsph_ctx* ctx, int64_t* pairs, hipStream_t stream);

/* Cell bookkeeping of the last search (neighbor_search.py:65-79): cells
 * (n, dim) int64 normalised by the per-axis minimum, row-major keys (n),
 * stable cell order (n) = argsort(keys, stable).  Any pointer may be NULL.
 * *ncells receives the number of occupied cells (host). */
int tsph_cells_export(tsph_ctx* ctx, int64_t* cells, int64_t* keys, int64_t* order,
                      int64_t* ncells, hipStream_t stream);

/* max_i |pos_i - anchor_i| (PairListCache.displacement, neighbor_search.py:
 * 203-206), written to *out (device scalar of `dtype`). */
int tsph_max_displacement(const void* pos, const void* anchor, int dtype, int64_t n,
                          int dim, void* out, hipStream_t stream);

/* ---------------------------------------------------------------- operators
 * Neighbour lists for the operators below, from a reference pair list
 * (P, 2) int64 (any order, i != j).  Targets are the fluid particles
 * [0, n_fluid) (all their neighbours) and the wall particles
 * [n_fluid, n_total) (their fluid neighbours).  Each list is sorted by
 * neighbour index, so every gather below is deterministic. */
int tsph_nlist_from_pairs(tsph_ctx* ctx, const int64_t* pairs, int64_t npairs,
                          int64_t n_fluid, int64_t n_total, int all_targets,
                          hipStream_t stream);

typedef struct tsph_wcsph_params {
  int32_t dim;               /* 2 or 3 */
  int32_t precision;         /* TSPH_F64 or TSPH_F32 */
  int64_t n_fluid, n_wall;
  double h, rho0, sound_speed, gamma, gravity, viscosity, wall_friction, dt;
  double diffusion_scale;    /* 2 delta c0 h; 0 disables density diffusion */
  int32_t diffusion_form;    /* 0 Molteni-Colagrossi, 1 Antuono (renormalised) */
  double eigenvalue_floor;   /* Antuono eigenvalue floor (fraction of trace/D) */
  int32_t reaction_style;    /* 0 none; 1 2-D split loads; 2 rhs_3d wall force */
  int32_t has_wall_accel;
  /* opt-in branches of droplet_impact.rhs (zero = off; dim 2 only, except the
   * surface-pressure closure, which rhs_3d also runs: dam_break_3d.py:331-335) */
  double wrap_width;                   /* > 0: x separations modulo wrap_width (:921-927) */
  int32_t density_renormalization;     /* Shepard: rho lattice_kernel_sum / (sum_j W_ij + W(0)) (:954-964) */
  double lattice_kernel_sum;           /* lattice_kernel_sum(h, dx) (:577-598) */
  int32_t boundary_consistent_density; /* wall mass scaled by rho_w / rho0 in the fluid density (:1090-1112) */
  int32_t surface_pressure_closure;    /* free-surface pressure extrapolation (:1031-1038, :429-492) */
  double surface_pressure_threshold;   /* receivers: rho < threshold * rho0 */
  int32_t surface_pressure_donors_all; /* 1 "all" (one pass), 0 "interior" */
  int32_t surface_pressure_layers;     /* passes of "interior" */
} tsph_wcsph_params;

/* Replaces droplet_impact.rhs (droplet_impact.py:810-1228; dim 2) and
 * dam_break_3d.rhs_3d (dam_break_3d.py:290-422; dim 3) on the lists of the
 * last tsph_nlist_from_pairs call.  pos/vel: (n_fluid + n_wall, dim) FP64
 * (wall rows of vel = wall velocity).  mass: (n_fluid + n_wall) scalar.
 * wall_accel: (n_wall, dim) FP64 or NULL.  Outputs (scalar type): rho, p
 * (n_fluid), acc (n_fluid, dim), wall_p, wall_rho (n_wall, may be NULL),
 * reaction (n_wall, dim; NULL unless reaction_style != 0). */
int tsph_wcsph_rhs(tsph_ctx* ctx, const tsph_wcsph_params* prm, const double* pos,
                   const double* vel, const void* mass, const double* wall_accel,
                   void* rho, void* p, void* acc, void* wall_p, void* wall_rho,
                   void* reaction, hipStream_t stream);

/* droplet_impact.rhs with its per-particle inputs (droplet_impact.py:810-
 * 1228, dim 2): density_offset (n_fluid, scalar type; :936-943) is added to
 * the clamped summation density, nu_field (n_fluid, scalar type;
 * :1164-1178) replaces the scalar viscosity per pair.  Either may be NULL;
 * with both NULL this is tsph_wcsph_rhs.  The caller passes no extras when
 * the pair list is empty, as the reference returns before them. */
int tsph_wcsph_rhs_ex(tsph_ctx* ctx, const tsph_wcsph_params* prm, const double* pos,
                      const double* vel, const void* mass, const double* wall_accel,
                      const void* density_offset, const void* nu_field, void* rho, void* p,
                      void* acc, void* wall_p, void* wall_rho, void* reaction,
                      hipStream_t stream);

/* Tait EOS p = max(0, B((max(rho/rho0, ratio_floor))^gamma - 1)), B = rho0 c0^2/gamma:
 * droplet_impact.tait_pressure (droplet_impact.py:415-427; ratio_floor 0.01)
 * and threed.operators3.tait_pressure (operators3.py:82-86; ratio_floor 0). */
int tsph_eos_tait(const void* rho, int64_t n, int precision, double rho0, double sound_speed,
                  double gamma, double ratio_floor, void* p, hipStream_t stream);

/* Kernel values for radii r: kind 0 = 2-D cubic spline, 1 = 3-D cubic spline
 * (g and the radial Laplacian lap), 2 = 2-D Wendland C2.  g is the gradient
 * scale (grad_i W = delta * g).  Replaces cubic_spline_kernel/_gradient
 * (droplet_impact.py:276-304), wendland_c2_* (:320-346) and
 * threed.kernel3.pair_kernels (kernel3.py:167-181).  g, lap may be NULL. */
int tsph_kernel_eval(int kind, const void* r, int64_t n, int precision, double h, void* w,
                     void* g, void* lap, hipStream_t stream);

/* Safety clamp of apply_boundary (droplet_impact.py:1293-1312, 3-D:
 * dam_break_3d.py:429-452), in place on FP64 (n, dim) pos/vel: positions
 * clamped to [lo, hi], the velocity component zeroed where a particle was
 * beyond the box by more than tol.  *count (device uint64) receives the
 * number of particles found beyond the tolerance. */
int tsph_apply_boundary(double* pos, double* vel, int64_t n, int dim, const double* lo,
                        const double* hi, double tol, unsigned long long* count,
                        hipStream_t stream);

/* clamp_out_of_obstacle (kleefsman_obstacle_3d.py:208-266), in place on FP64
 * (n, 3) pos/vel: a particle strictly inside the box (lo, hi) leaves through
 * the nearest face (first of -x, +x, -y, +y, -z, +z on a tie) to that face
 * -/+ eps, and that velocity component is zeroed.  *count (device uint64)
 * receives the number moved.  mass (n, scalar type of `precision`) may be
 * NULL; otherwise the momentum removed, sum m (v_before - v_after), is added
 * to impulse (3 device doubles, accumulated across calls). */
int tsph_obstacle_clamp(double* pos, double* vel, const void* mass, int precision, int64_t n,
                        const double* lo, const double* hi, double eps, unsigned long long* count,
                        double* impulse, hipStream_t stream);

/* ---------------------------------------------------------------- 3-D operators
 * threed.operators3.accelerations (operators3.py:212-256) with
 * summation_density (:70-79), tait_pressure (:82-86), wall_pressure_adami
 * (:173-209) + tait_density (:89-93), pressure_acceleration (:96-113),
 * viscous_acceleration with order 2 (:116-154) and density_diffusion_rate_3d
 * (:259-295).  Every particle is a target: build the lists with
 * tsph_nlist_from_pairs(..., all_targets = 1). */
typedef struct tsph_threed_params {
  int32_t precision;         /* TSPH_F64 or TSPH_F32 */
  int64_t n, n_fluid;        /* all particles; fluid rows come first */
  double h, rho0, sound_speed, gamma, viscosity;
  double gravity[3];
  int32_t has_gravity;       /* 0: gravity=None */
  double box[3];
  int32_t periodic;          /* 1: minimum-image pair deltas in box */
  int32_t renormalise;       /* divide the viscous kernel by K_ij / 2 */
  double diffusion_scale;    /* 2 delta c0 h; 0 disables density diffusion */
  int32_t diffusion_form;    /* 0 Molteni-Colagrossi, 1 Antuono (renormalised) */
  double eigenvalue_floor;
  double dt;                 /* used by density diffusion only */
  int32_t wall_model;        /* 0 "adami", 1 "tait" (walls keep their own Tait state) */
  int32_t viscous_order;     /* 2 (0 reads as 2) or 4: L - L(beta L), operators3.py:146-151 */
} tsph_threed_params;

/* pos, vel: (n, 3) FP64; mass (n) scalar.  Outputs: acc (n, 3), rho and p (n)
 * in the scalar type (rho, p may be NULL).  Wall rows of acc are computed
 * like fluid rows, as the reference does. */
int tsph_threed_accelerations(tsph_ctx* ctx, const tsph_threed_params* prm, const double* pos,
                              const double* vel, const void* mass, void* acc, void* rho, void* p,
                              hipStream_t stream);

/* One operator of threed.operators3 with the caller's rho / p (scalar type),
 * on the same all-target lists:
 *   op 0 summation_density (operators3.py:70-79)         -> out (n)
 *   op 1 pressure_acceleration (:96-113), rho and p       -> out (n, 3)
 *   op 2 viscous_acceleration, order 2 or 4 (:116-154), rho -> out (n, 3)
 *   op 3 wall_pressure_adami (:173-209), rho and p        -> out (n - n_fluid)
 *   op 4 density_diffusion_rate_3d (:259-295), rho        -> out (n)
 * gravity enters op 3 only; viscosity / renormalise enter op 2 only;
 * diffusion_scale / diffusion_form / eigenvalue_floor enter op 4 only. */
int tsph_threed_operator(tsph_ctx* ctx, const tsph_threed_params* prm, int op, const double* pos,
                         const double* vel, const void* mass, const void* rho, const void* p,
                         void* out, hipStream_t stream);

/* ---------------------------------------------------------------- delta-plus step
 * delta_plus.delta_plus_step (delta_plus.py:105-229): one 2-D δ⁺-SPH step,
 * FP64, on the lists of tsph_nlist_from_pairs(pairs, n_fluid = n,
 * n_total = n + n_wall, all_targets = 0) built from the pair list over
 * cat(pos, wall) (skin pairs allowed; the strict d^2 < (2h)^2 test, h = 2 dx,
 * is applied per pair).  pos, vel (n, 2); rho, mass (n); wall, wall_vel,
 * wall_normal (n_wall, 2) with unit normals.  Outputs: pos_out, vel_out,
 * acc_out, shift (n, 2); rho_out (n); reaction (n_wall, 2).  Synchronises and
 * returns TSPH_ERR_NUMERIC for a nonpositive wall density or a nonpositive /
 * nonfinite new density, as the reference raises FloatingPointError. */
typedef struct tsph_dp_params {
  int32_t precision;         /* TSPH_F64 */
  int64_t n, n_wall;
  double dx, sound_speed, reference_speed, dt, rho0, gravity, viscosity, diffusion, alpha;
} tsph_dp_params;

int tsph_deltaplus_step(tsph_ctx* ctx, const tsph_dp_params* prm, const double* pos,
                        const double* vel, const double* rho, const double* mass,
                        const double* wall, const double* wall_vel, const double* wall_normal,
                        double* pos_out, double* vel_out, double* rho_out, double* acc_out,
                        double* reaction, double* shift, hipStream_t stream);

/* ---------------------------------------------------------------- served batch of delta-plus runs
 * No counterpart in the reference (which runs one simulation per Python
 * loop): the batch engine that advances many independent δ⁺-SPH runs, each
 * the loop of benchmarks/delta_plus/elliptical_drop.py:34-57 (adaptive
 * dt = courant 2 dx / (c0 + max|v|), fresh cell lists every step) with the
 * step of delta_plus.py:105-229, packed together in HBM.  Capacities are
 * fixed at creation; one captured step graph advances every active run, so
 * runs can be admitted and released between steps.  A run's results are
 * bitwise the same alone and in any mix.  Calls on one batch must not
 * overlap; the stream must not be the legacy default stream (graph capture). */
typedef struct tsph_batch tsph_batch;

typedef struct tsph_batch_caps {
  int64_t fluid, wall, sims;       /* rows and run slots */
  int64_t cells;                   /* cells of all per-run grids together (cell edge 4 dx) */
  int64_t fluid_entries, wall_entries;  /* neighbour-list entries: each row gets entries / rows
                                           slots rounded down to a multiple of 4 (32..256, 64
                                           is ample for h = 2 dx); a row with more neighbours
                                           fails its run (state 3) */
  int32_t sorted_lists;            /* 1: lists in ascending row order (the order of
                                      tsph_deltaplus_step, for bitwise checks); 0: cell order */
} tsph_batch_caps;

typedef struct tsph_dp_case {
  int64_t n, n_wall, steps;        /* fluid rows, wall stations, steps to run */
  double dx, sound_speed, reference_speed, rho0, gravity, viscosity, diffusion, alpha;
  double courant;                  /* adaptive step factor, <= 0.3 */
} tsph_dp_case;

int tsph_batch_create(const tsph_batch_caps* caps, hipStream_t stream, tsph_batch** out);
/* Copies one run's initial state (device pointers, layouts as for
 * tsph_deltaplus_step) into free capacity.  *slot receives its id.  Returns
 * TSPH_ERR_OOM when no slot or no contiguous room is free (a scheduler
 * retries after a release).  Synchronises. */
int tsph_batch_admit(tsph_batch* b, const tsph_dp_case* c, const double* pos, const double* vel,
                     const double* rho, const double* mass, const double* wall,
                     const double* wall_vel, const double* wall_normal, int32_t* slot,
                     hipStream_t stream);
/* Advances every active run by nsteps steps (runs stop at their own step
 * count); synchronises at the end. */
int tsph_batch_step(tsph_batch* b, int64_t nsteps, hipStream_t stream);
/* Launch shape of the captured step: rows per workgroup of the four delta-plus
 * pair sweeps (sweep_block) and of the list build (list_block), each 64, 128,
 * 192 or 256, and xcd_chunks 1 (consecutive workgroups of one XCD take
 * neighbouring rows, which share that XCD's L2) or 0 (round-robin).  Results
 * do not depend on it; the step graph is captured again at the next step.
 * For the autotuner (serving.BatchEngine.autotune); no counterpart in the
 * reference, which runs one case per Python loop. */
int tsph_batch_launch(tsph_batch* b, int32_t sweep_block, int32_t list_block, int32_t xcd_chunks);
/* Finished runs since their admission: slots and states (2 done, 3 failed:
 * a nonpositive wall or fluid density, the reference's FloatingPointError). */
int tsph_batch_poll(tsph_batch* b, int32_t* slots, int32_t* states, int32_t max, int32_t* n);
/* A run's state: pos, vel (n, 2), rho (n) device pointers (may be NULL);
 * *steps the steps it has taken. */
int tsph_batch_read(tsph_batch* b, int32_t slot, double* pos, double* vel, double* rho,
                    int64_t* steps, hipStream_t stream);
int tsph_batch_release(tsph_batch* b, int32_t slot, hipStream_t stream);
/* counters: batch steps, resident runs, resident fluid rows, wall rows */
int tsph_batch_stats(tsph_batch* b, int64_t* counters);
int tsph_batch_destroy(tsph_batch* b);

/* ---------------------------------------------------------------- conservation ledger
 * threed.conservation.diagnostics (conservation.py:106-121) and the shear-wave
 * mode projections (operators3.py:401-408) as one FP64 row of
 * TSPH_CONSERVATION_ROW host doubles:
 *   [0] sum m, [1..3] momentum, [4..6] angular momentum about the origin,
 *   [7] kinetic energy, [8] sum m e(rho) (Tait internal energy),
 *   [9..11] sum m sin(2 pi x_d / L_d), [12..14] sum m cos(2 pi x_d / L_d)
 *   (L = box, or 1 when box is NULL), [15] sum u_x sin(k z),
 *   [16] sum u_x sin(k y) sin(k z), [17] sum m |v|_1, [18..20] sum m x_d.
 * pos, vel: (n, 3) FP64; mass, rho: (n) in the scalar type of `precision`.
 * Synchronises. */
#define TSPH_CONSERVATION_ROW 21
int tsph_conservation_row(tsph_ctx* ctx, const double* pos, const double* vel, const void* mass,
                          const void* rho, int64_t n, int precision, double rho0, double sound_speed, double gamma,
                          const double* box, double wavenumber, double* out, hipStream_t stream);

/* tsph_conservation_row without a context and without allocating: no
 * counterpart in the reference; for callers that cannot allocate device
 * memory (HIP graph capture, tight serving loops).  Its per-block partials
 * live in a fixed 8 MiB device array of the library, so n is at most
 * 12,782,336 rows (larger n: TSPH_ERR_ARG).  Same kernels and grid as
 * tsph_conservation_row, so the same row bit for bit.  Calls that overlap on
 * one GPU share that array and race; each device has its own.
 * Synchronises. */
int tsph_conservation_row_noalloc(const double* pos, const double* vel, const void* mass, const void* rho,
                                  int64_t n, int precision, double rho0, double sound_speed, double gamma,
                                  const double* box, double wavenumber, double* out, hipStream_t stream);

/* ---------------------------------------------------------------- periodic stepper
 * The step loop of threed.solver.run_shear_wave (solver.py:211-308) with the
 * state resident in HBM: periodic_pairs at 2h + skin (neighbors3.py:71-142)
 * kept until the largest minimum-image displacement exceeds skin / 2,
 * accelerations as above, kick-drift-kick leapfrog with the positions wrapped
 * by torch.remainder (:285-299).  prm->periodic must be 1 and n_fluid == n. */
typedef struct tsph_psim tsph_psim;

int tsph_psim_create(const tsph_threed_params* prm, double skin, const double* pos,
                     const double* vel, const void* mass, hipStream_t stream, tsph_psim** out);
/* Advances nsteps leapfrog steps (threed/solver.py:285-299).  The skin test
 * runs on the device and the host reads it two steps later: a step whose
 * drift went past half the skin has its evaluation (and the queued steps
 * after it) skipped on the device, and the host rebuilds the lists from that
 * step's drifted positions, evaluates it and queues the rest again, so the
 * rebuild schedule is the sequential loop's.  Returns when the last step has
 * finished. */
int tsph_psim_step(tsph_psim* sim, int64_t nsteps, hipStream_t stream);
/* The conservation row (TSPH_CONSERVATION_ROW doubles, see
 * tsph_conservation_row) of the current state.  Synchronises. */
int tsph_psim_diagnostics(tsph_psim* sim, double wavenumber, double* out, hipStream_t stream);
/* The n masses in the caller's order, FP64 (device).  An FP32-compensated
 * periodic stepper keeps each as two floats and sums densities with the FP64
 * value: near rho0 the Tait law turns a density error into
 * c0^2 rho0 (dm / m), so a mass rounded to FP32 would shift every pressure
 * by about 7 B 2^-24.  Any time; takes effect at the next evaluation;
 * synchronises the stream (it re-checks whether every mass is equal).  The
 * reference holds masses in the run's dtype (threed/solver.py:84-90); its
 * FP64 runs see the FP64 masses this keeps. */
int tsph_psim_set_masses_f64(tsph_psim* sim, const double* mass, hipStream_t stream);
/* State in the caller's order: pos, vel (n, 3) FP64; rho (n) and acc (n, 3)
 * of the last evaluation in the scalar type.  Any pointer may be NULL. */
int tsph_psim_read(tsph_psim* sim, double* pos, double* vel, void* rho, void* acc,
                   hipStream_t stream);
/* counters: steps done, list builds (cell binnings), entries the
 * periodic_pairs lists at the build radius would hold (each pair twice).
 * Host pointer. */
int tsph_psim_stats(tsph_psim* sim, int64_t* counters);
int tsph_psim_destroy(tsph_psim* sim);

/* ---------------------------------------------------------------- resident stepper
 * The step loop of droplet_impact.run (droplet_impact.py:1674-1763) and
 * dam_break.run_case (dam_break.py:374-465): Verlet-skin pair cache
 * (neighbor_search.py:166-227), rhs, symplectic Euler (:1745-1756) and the
 * apply_boundary clamp (:1293-1312), with the state resident in HBM.  The
 * caller's particle order is restored by tsph_sim_read. */
typedef struct tsph_sim tsph_sim;

typedef struct tsph_sim_params {
  tsph_wcsph_params op;      /* reaction_style and has_wall_accel must be 0 */
  double skin;               /* Verlet skin (absolute length, > 0) */
  double lo[3], hi[3], tol;  /* clamp box and crossing tolerance */
  int32_t clamp;             /* 1: apply the boundary clamp after every step */
  /* Rigid obstacle (kleefsman_obstacle_3d.py:553-726; dim 3): wall rows
   * [obstacle_begin, obstacle_end) of the caller's wall order.  When the range
   * is nonempty every step records the obstacle's load F = sum of the wall
   * reactions of rhs_3d(return_wall_force=True) (dam_break_3d.py:410-422) and
   * its moment M = sum (x_w - centre) x F_w (:652-673).  obstacle_clamp = 1
   * applies clamp_out_of_obstacle (:208-266) to the box [obstacle_lo,
   * obstacle_hi] after the boundary clamp, with eps = 0.01 dx. */
  int64_t obstacle_begin, obstacle_end;
  double obstacle_centre[3];
  int32_t obstacle_clamp;
  double obstacle_lo[3], obstacle_hi[3], obstacle_eps;
  /* Adaptive step (adaptive_dt; droplet_impact.py:495-552, :1655-1763,
   * dam_break.py:364-465).  adaptive = 1 integrates to target_time: after step
   * k, dt = min(dt_safety * min(courant_num / (c0 + v_k), dt_limit),
   * dt * dt_growth, op.dt) with v_k the largest speed at the start of step k,
   * courant_num = max_courant * h_factor * dx and dt_limit the smallest
   * speed-independent limit (viscous, gravity, diffusion); step k integrates
   * with min(dt, target_time - elapsed), and the step that starts within
   * 1e-12 max(1, target_time) of target_time is the final sample: it is
   * evaluated and metered but not moved, and tsph_sim_step returns after it.
   * adaptive = 0: every step integrates with op.dt. */
  int32_t adaptive;
  double target_time, courant_num, dt_limit, dt_safety, dt_growth;
  /* adaptive_free_last = 1: steps are never shortened onto target_time and
   * the final sample is the first step starting at elapsed >= target_time -
   * tol (kleefsman_obstacle_3d.py:644-713); adaptive_tol > 0 sets tol, 0
   * keeps 1e-12 max(1, target_time). */
  int32_t adaptive_free_last;
  double adaptive_tol;
  /* Time state a resumed run starts from (kleefsman_obstacle_3d.py:597-628):
   * the simulated time at the start of the first step and, with adaptive =
   * 1, the step-size variable (<= 0: op.dt).  Zero for a fresh run. */
  double initial_time, initial_dt;
} tsph_sim_params;

/* pos: (n_fluid + n_wall, dim) FP64 device, fluid then wall; vel: (n_fluid,
 * dim) FP64; mass: (n_fluid + n_wall) in the scalar type. */
int tsph_sim_create(const tsph_sim_params* prm, const double* pos, const double* vel,
                    const void* mass, hipStream_t stream, tsph_sim** out);
/* Advances nsteps steps (adaptive runs: fewer once the final sample is done).
 * meter_every > 0 records a meter row on the steps whose index is a multiple
 * of it.  Returns when the last step has finished. */
int tsph_sim_step(tsph_sim* sim, int64_t nsteps, int64_t meter_every, hipStream_t stream);
/* State in the caller's order: pos, vel (n_fluid, dim) FP64; rho, p (n_fluid)
 * and acc (n_fluid, dim) of the last step.  Any pointer may be NULL. */
int tsph_sim_read(tsph_sim* sim, double* pos, double* vel, void* rho, void* p, void* acc,
                  hipStream_t stream);
/* counters (6): steps done, list builds (cell binnings), boundary crossings,
 * cells of the dense cell table the list builds walk, meter rows, 1 once an
 * adaptive run has done its final sample; peaks: max density and max speed
 * over all steps.  Host pointers. */
int tsph_sim_stats(tsph_sim* sim, int64_t* counters, double* peaks);
/* Meter rows (host): step, sums[mass, momentum x/y/z, kinetic energy, sum rho,
 * potential energy], maxima[rho, p, speed, |acc|, y, x, displacement, and the
 * step's density and speed maxima tracked on every step (tsph_sim_stats keeps
 * the run's peaks)], minima[rho, p, y], time[elapsed at the step's start, the
 * previous step's size, this step's size, the step-size variable]: 24
 * doubles per row.  The meters are of the state at the step's start. */
int tsph_sim_meters(tsph_sim* sim, double* rows, int64_t max_rows, int64_t* nrows);
/* Time state of the coming step (host, 4 doubles): simulated time at its
 * start, the step-size variable, the size it integrates with (0 for an
 * adaptive run's final sample) and the previous step's size.  What a
 * checkpoint of an adaptive run records.  Synchronises. */
int tsph_sim_time(tsph_sim* sim, double* out, hipStream_t stream);
/* Obstacle loads of every completed step (host), 12 doubles per row: step,
 * F x/y/z, M x/y/z (about obstacle_centre), cumulative clamp impulse x/y/z
 * (sum of m (v_before - v_after) of clamped particles), the cumulative
 * number of clamped particles and the step's size. */
int tsph_sim_loads(tsph_sim* sim, double* rows, int64_t max_rows, int64_t* nrows);
/* Sweep layout.  tiles = 1 runs the LDS cell-tile sweeps with tile_x x
 * tile_y cells per workgroup ((tile_x+2)(tile_y+2) <= 108, tile_x tile_y <=
 * 64; 2-D only); tiles = 0 runs the neighbour-list sweeps from the global
 * arrays; tiles = 2 (the default) runs the neighbour-list sweeps with each
 * 256-row workgroup's neighbourhood rows staged in LDS (up to 1,024 rows in
 * 2-D, 2,048 in 3-D; a tile with more reads the global arrays), with the same
 * bits as tiles = 0.
 * Takes effect at the next step, which rebuilds the layout.  No counterpart
 * in the reference, which has one layout (its neighbour pairs). */
int tsph_sim_configure(tsph_sim* sim, int tiles, int tile_x, int tile_y);
/* Masses of all n_fluid + n_wall rows in the caller's order, FP64 (device).
 * An FP32-compensated stepper keeps each as two floats (the FP32 value and
 * the part it dropped) and sums densities with the FP64 value, so its
 * density matches an FP64 run's to FP64 rounding: at depth H the pressure
 * contrast between neighbours is about h / H of the pressure, so the 2^-24
 * rounding of a mass would reach the force about H / h times amplified.
 * An FP64 stepper stores them as given.  Any time; takes effect at the next
 * step.  The reference holds masses in the run's dtype (droplet_impact.py:
 * 618-700); its FP64 runs see the FP64 masses this keeps. */
int tsph_sim_set_masses_f64(tsph_sim* sim, const double* mass, hipStream_t stream);
/* Slot width of the per-row neighbour lists (host): *current receives the
 * width of the last build; width >= 0 sets the width the next build starts
 * from (rounded up to a multiple of 4; 0: count first).  A build widens the
 * slots by itself when a list does not fit: in the stepping loop the steps
 * queued on such a build change nothing and are run again on a wider build,
 * which is not counted as a list rebuild.  No counterpart in the reference;
 * for callers that plan memory ahead. */
int tsph_sim_slot_width(tsph_sim* sim, int64_t width, int64_t* current);
/* Layout of the last build (host, 4 int64): the layout, the 256-row tiles
 * on the LDS halo path, all tiles (both 0 outside layout 2) and the largest
 * halo in rows.  Diagnostic; no counterpart in the reference. */
int tsph_sim_layout(tsph_sim* sim, int64_t* out);
/* Sweep timing with HIP events on the stepping stream.  Returns the summed
 * milliseconds of the density, wall and force+integrate sweeps and the number
 * of timed steps since the last reset; enable = k > 0 resets the sums and
 * times every k-th step (1: every step), 0 resets and turns timing off,
 * -1 only reads. */
int tsph_sim_timing(tsph_sim* sim, int enable, double* sweep_ms, int64_t* steps);
/* Number of unordered pairs with separation <= radius in the current state
 * (radius <= the list's build radius); synchronises. */
int tsph_sim_count_pairs(tsph_sim* sim, double radius, int64_t* pairs, hipStream_t stream);
int tsph_sim_destroy(tsph_sim* sim);

/* ---------------------------------------------------------------- particle shifting
 * particle_shifting.shift_positions (particle_shifting.py:88-224) applied to
 * the 2-D stepper's fluid after its last step, as dam_break.run_case does
 * every shift_every steps (dam_break.py:424-433): concentration C from the
 * last evaluation's volumes m / rho over the fluid pairs within 2h, shift
 * -strength h v_ref dt grad C, tangential where C < free_surface C_ref,
 * capped at limit h; dt = 0 takes the last step's size (adaptive runs).
 * C_ref is c_reference when >= 0 (0: no free surface), else the percentile
 * of C (linear interpolation).  v_ref < 0 takes max(largest speed at the
 * start of the last step, v_floor); v_ref = 0 moves nothing.  info (may be NULL): C_ref, largest shift,
 * fraction of rows capped, v_ref.  Synchronises. */
typedef struct tsph_shift_params {
  double dt, v_ref, v_floor, strength, limit, free_surface, percentile, c_reference;
} tsph_shift_params;
int tsph_sim_shift(tsph_sim* sim, const tsph_shift_params* prm, double* info, hipStream_t stream);

/* transport_velocity_correction (transport_velocity_correction.py:285-513)
 * applied to the 2-D stepper's fluid after its last step, as
 * dam_break.run_case does every tvc_every steps (dam_break.py:434-460):
 * g_i = -2 sum_j V_j grad_i W_ij over the fluid and wall neighbours within
 * 2h (fluid V = m / rho of the last evaluation, wall V = wall_volume),
 * shift = alpha h^2 L(h^2 |g|^2) g with L = min(limiter_slope x, 1) when
 * limiter_slope > 0 (TruncatedLinear) else 1; scope 1 keeps rows with
 * C >= bulk_c C_ref (C, C_ref as for tsph_sim_shift), scope 2 rows with
 * |g| h <= bulk_g; max_shift >= 0 caps |shift| (0 clips every shift to
 * zero), max_shift < 0 leaves it uncapped; then the boundary clamp.
 * info (may be NULL): largest |shift|, fraction capped, boundary crossings
 * (counted into the stepper's escapes).  Synchronises. */
typedef struct tsph_tvc_params {
  double alpha, limiter_slope, max_shift, wall_volume;
  int32_t scope;
  double bulk_c, bulk_g, percentile;
} tsph_tvc_params;
int tsph_sim_tvc(tsph_sim* sim, const tsph_tvc_params* prm, double* info, hipStream_t stream);

/* ---------------------------------------------------------------- gauges
 * The water-height gauges and obstacle pressure taps of the 3-D dam break
 * against a box (kleefsman_obstacle_3d.py:273-318, marin_probes.py:25-79).
 * Slab: station_x - aperture/2 < x < station_x + aperture/2 and
 * obstacle_lo[2] < z < obstacle_hi[2]; the centreline gauge is centred on
 * (station_x, tank_depth / 2); taps: 8 sensors (4 on the upstream face, 4 on
 * the top face) of per_tap quadrature nodes each (1: the point, 7: the disk
 * cubature), nodes[(8 per_tap) x 3].  Output row (TSPH_GAUGE_ROW doubles):
 * slab height, peak height in the slab (0 if empty), front x, bypass count,
 * centreline height H2, then (pressure, support) per node (56 slots, unused
 * slots 0).  Synchronises. */
#define TSPH_GAUGE_ROW 117
typedef struct tsph_gauge_spec {
  double station_x, aperture, tank_depth, dx, h, gravity;
  double obstacle_lo[3], obstacle_hi[3];   /* (x0, 0, z0), (x1, height, z1) */
  int32_t per_tap;
  double nodes[8 * 7 * 3];
} tsph_gauge_spec;

/* On caller arrays: pos (n, 3) FP64; vol = m / rho, rho, p (n) scalar type. */
int tsph_gauges(tsph_ctx* ctx, const tsph_gauge_spec* spec, int precision, int64_t n,
                const double* pos, const void* vol, const void* rho, const void* p, double* out,
                hipStream_t stream);
/* On the 3-D stepper's fluid rows at the start of its last step: the
 * positions that step began from and the density, pressure and volume its
 * evaluation produced (the state the reference samples at that step). */
int tsph_sim_gauges(tsph_sim* sim, const tsph_gauge_spec* spec, double* out, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* TSPH_H */
