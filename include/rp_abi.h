/*
 * rp_abi.h — C ABI of the MI355X DEM time-step engine (librp_hip.so).
 *
 * This is the drop-in boundary for the GPU path of rusty_particles.  The
 * reference binds its GPU backend through three hooks:
 *   - GpuSimulation::new(&[Particle], dt, bounds_min, bounds_max, periodic,
 *     p_mat, w_mat, normal_model, tangential_model)   src/gpu.rs:61-71,
 *     called from Simulation::enable_gpu               src/simulation.rs:129-146
 *   - GpuSimulation::step(&self)                       src/gpu.rs:153-166,
 *     called from Simulation::step                     src/simulation.rs:156-158
 *   - GpuSimulation::read_particles(&self)             src/gpu.rs:168-213,
 *     called from Simulation::sync_particles           src/simulation.rs:564-568
 * Each entry point below names the reference item it replaces.
 *
 * Conventions
 *   - Plain C types only.  No exception or panic crosses the ABI; every call
 *     returns rp_status (0 = OK, < 0 = error) and rp_last_error() returns the
 *     message of the last failing call on the calling thread.
 *   - The library owns device memory; the caller owns every host array.
 *   - Particle order at the ABI is always the reference's API order: insertion
 *     order minus removed particles (the reference's retain + re-index,
 *     src/simulation.rs:200-219).  The device keeps its own cell-sorted order
 *     internally.
 *   - One engine = one HIP stream on the current device.  An engine is not
 *     thread-safe; use one engine per thread.
 *   - All state is FP64.  The reference computes in f32 (glam Vec3, Cargo.toml:7);
 *     FP64 is this engine's documented precision.
 */
#ifndef RP_ABI_H
#define RP_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rp_status;

enum {
  RP_OK = 0,
  RP_ERR_INVALID = -1,  /* bad argument or parameter */
  RP_ERR_DEVICE = -2,   /* no HIP device, or a HIP runtime error */
  RP_ERR_NOMEM = -3,    /* device allocation failed */
  RP_ERR_STATE = -4,    /* call not valid in the engine's current state */
  RP_ERR_IO = -5        /* file could not be read or parsed */
};

/* Normal force models.  The values are the reference's enum discriminants,
 * which its GPU path already treats as an ABI (src/physics.rs:6-17,
 * src/gpu.rs:108). */
enum {
  RP_NORMAL_LINEAR = 0,      /* LinearSpringDashpot  physics.rs:49-70   */
  RP_NORMAL_HERTZ = 1,       /* Hertzian             physics.rs:86-129  */
  RP_NORMAL_HYSTERETIC = 2,  /* Hysteretic           physics.rs:334-380 */
  RP_NORMAL_JKR = 3,         /* JKR                  physics.rs:132-219 */
  RP_NORMAL_SJKR = 4         /* SimplifiedJKR        physics.rs:234-325 */
};

/* Tangential force models (src/physics.rs:20-28). */
enum {
  RP_TANGENTIAL_COULOMB = 0, /* physics.rs:385-397 */
  RP_TANGENTIAL_LSC = 1,     /* LinearSpringCoulomb physics.rs:400-434 */
  RP_TANGENTIAL_MINDLIN = 2  /* physics.rs:443-459 */
};

/* Semantics mode.  PARITY reproduces the reference's step_cpu exactly,
 * quirks included (shear modulus placeholder 0, material.rs:49).
 * PHYSICS_FIX (no counterpart in the reference) corrects what the reference
 * leaves as placeholders or gets wrong, keeping everything else:
 *   - shear modulus G = E / (2 (1 + nu)) (material.rs:52-54) and the Mindlin
 *     effective 1/G* = (2 - nu1)/G1 + (2 - nu2)/G2 (material.rs:39-50 returns 0);
 *   - Mindlin kt = 8 G* sqrt(R* delta) only for delta > 0 (physics.rs:452
 *     takes the root of a negative overlap: NaN with JKR/sJKR margins);
 *   - tangential history per directed contact, kept while the spring is
 *     loaded and dropped when the contact opens (the reference keys one entry
 *     per unordered pair that both particles add to, never removed,
 *     simulation.rs:397-399), projected onto the current tangent plane each
 *     step; slip direction for any non-zero trial force (physics.rs:425's
 *     1e-6 N threshold);
 *   - the closest point of a triangle with Eberly's normalised edge
 *     parameters (mesh.rs:63-113 clamps unnormalised ones to [0, a] / [0, c]);
 *   - every pair inside the cutoff acts once: no reference-grid visibility
 *     or double visits (grid.rs:48-87 depends on a 1 m cell).
 * Walls keep Coulomb friction (simulation.rs:512-514).  Batched engines
 * (rp_engine_upload_batch) keep a history per instance, keyed by the stable
 * ids that are unique across the batch. */
enum { RP_MODE_PARITY = 0, RP_MODE_PHYSICS_FIX = 1 };

/* Material, field order as Material::new (src/material.rs:12) and the
 * Python setter (src/lib.rs:109). */
typedef struct rp_material {
  double youngs_modulus;
  double poissons_ratio;
  double density;
  double friction_coefficient;
  double restitution_coefficient;
  double surface_energy;
} rp_material;

/* Everything GpuSimulation::new receives apart from the particles
 * (src/gpu.rs:61-71), plus the items the reference hard-codes. */
typedef struct rp_params {
  double dt;                 /* s */
  double bounds_min[3];      /* m */
  double bounds_max[3];      /* m */
  int32_t periodic[3];       /* 0/1 per axis (Simulation.periodic) */
  int32_t normal_model;      /* RP_NORMAL_*     */
  int32_t tangential_model;  /* RP_TANGENTIAL_* */
  int32_t mode;              /* RP_MODE_*       */
  rp_material particle;      /* Simulation.particle_material */
  rp_material wall;          /* Simulation.wall_material     */
  double gravity[3];         /* (0,-9.81,0) = simulation.rs:267 */
  double ref_cell_size;      /* 1.0 = simulation.rs:66; decides which particles
                                the reference grid can see (grid.rs:84-94) */
  double cell_size;          /* engine's own broad-phase cell, 0 = automatic */
  double skin;               /* Verlet-list skin distance, 0 = automatic */
  int32_t cell_order;        /* RP_CELL_ORDER_*: numbering of the broad-phase
                                cells (no effect on results) */
  int32_t reserved0;         /* 0 */
} rp_params;

/* Numbering of the engine's broad-phase cells (rp_params.cell_order; results
 * never depend on it): AUTO picks pencils (8 x 8 cells of the two fast axes
 * running along the slow axis) when a layer of the slow axis holds more than
 * 8,192 particles, else row-major. */
enum { RP_CELL_ORDER_AUTO = 0, RP_CELL_ORDER_ROW = 1, RP_CELL_ORDER_PENCIL = 2 };

/* Counters of the last step of the most recent rp_engine_step call
 * (integers only). */
typedef struct rp_stats {
  uint64_t steps;              /* steps taken since create */
  uint64_t rebuilds;           /* neighbour-list rebuilds since create */
  uint64_t contacts;           /* particle-particle directed contacts, last step */
  uint64_t wall_contacts;      /* particle-triangle contacts, last step */
  uint64_t removed;            /* particles removed since create */
  uint64_t jkr_failures;       /* JKR roots rejected by the residual check, last step
                                  (physics.rs:193-199) */
  int32_t excessive_overlap;   /* simulation.rs:361-370,469-473,528-530 */
  int32_t neighbour_capacity;  /* current per-particle list capacity */
  int32_t max_neighbours;      /* largest list length seen at a rebuild */
  int32_t reserved;
  uint64_t overlap_steps;      /* steps since create whose excessive_overlap
                                  was set: the reference prints its warning
                                  once per such step (simulation.rs:528-530) */
  uint64_t history_contacts;   /* directed contacts that carried tangential
                                  history in the last step (physics-fix mode) */
} rp_stats;

typedef struct rp_engine rp_engine;

/* Operating envelope of the JKR surrogate: d = delta / t with
 * s = (R*^2 8 pi gamma / E*)^(1/3), t = s^2 / R* (the scaled overlap), and the
 * ranges of R*, E* and gamma.  Contacts outside it use the exact solve. */
typedef struct rp_jkr_envelope {
  double d_lo, d_hi;
  double rstar_lo, rstar_hi;  /* m */
  double estar_lo, estar_hi;  /* Pa */
  double gamma_lo, gamma_hi;  /* J/m^2 */
} rp_jkr_envelope;

/* Fills *p with the reference defaults of Simulation::new
 * (src/simulation.rs:62-91): particle material (1e7,0.3,2500,0.5,0.5,0),
 * wall material (1e9,0.3,7800,0.5,0.5,0), Hertzian + Mindlin, no periodic
 * axis, gravity (0,-9.81,0), reference cell 1.0 m. */
void rp_params_default(rp_params* p, double dt, const double bounds_min[3],
                       const double bounds_max[3]);

/* Number of HIP devices visible (0 on a machine without a GPU). */
int32_t rp_device_count(void);

/* Replaces GpuSimulation::new (src/gpu.rs:61-151) minus the particle upload.
 * Fails with RP_ERR_DEVICE when no HIP device is present, the condition the
 * reference maps to RuntimeError in enable_gpu (src/lib.rs:146-148). */
rp_status rp_engine_create(const rp_params* params, rp_engine** out);

/* Live parameter update.  The reference snapshots parameters at enable_gpu
 * (src/gpu.rs:100-115); this engine applies later setter calls too. */
rp_status rp_engine_set_params(rp_engine* e, const rp_params* params);

/* Replaces the particle upload of GpuSimulation::new (src/gpu.rs:92-98).
 * Arrays hold n particles in API order; x3/v3/w3/a3/tau3 are n*3 doubles,
 * radius/mass n doubles, fixed n bytes (0/1).  Any of v3, w3, a3, tau3 may
 * be NULL (zero).  Kahan residuals start at zero, as in Particle::new
 * (src/particle.rs:22-37).  Replaces all particles held by the engine. */
rp_status rp_engine_upload(rp_engine* e, size_t n, const double* x3,
                           const double* v3, const double* w3, const double* a3,
                           const double* tau3, const double* radius,
                           const double* mass, const uint8_t* fixed);

/* As rp_engine_upload, but also sets the three Kahan residual vectors
 * (position, velocity, angular velocity; src/particle.rs:11-17). */
rp_status rp_engine_upload_full(rp_engine* e, size_t n, const double* x3,
                                const double* v3, const double* w3,
                                const double* a3, const double* tau3,
                                const double* rx3, const double* rv3,
                                const double* rw3, const double* radius,
                                const double* mass, const uint8_t* fixed);

/* Appends one STL mesh, as Simulation::add_mesh (src/simulation.rs:98-100)
 * after Mesh::load_stl (src/mesh.rs:125-143): ntri triangles, v9 = ntri*9
 * floats (v0,v1,v2), n3 = ntri*3 floats (facet normal), STL float32. */
rp_status rp_engine_add_mesh(rp_engine* e, size_t ntri, const float* v9,
                             const float* n3);

/* Reads a binary or ASCII STL file and appends it as rp_engine_add_mesh.
 * RP_ERR_IO when the file cannot be read (the reference's IOError,
 * src/lib.rs:87-91). */
rp_status rp_engine_add_mesh_stl(rp_engine* e, const char* path);

/* Surface motion of mesh `mesh` (0-based, in the order added): a rigid
 * rotation with angular velocity omega[3] (rad/s) about the point origin[3].
 * The triangles stay where they are; a wall contact's relative velocity
 * subtracts omega x (c - origin) at the contact point c (the closest point on
 * the triangle).  Exact for surfaces of revolution about that axis (drums,
 * cylinders, cones), which the rotation maps onto themselves.  omega = 0 makes
 * the mesh static again.  At most 4 meshes move at once; RP_ERR_INVALID for a
 * fifth, an unknown mesh or non-finite values (the previous motion is kept).
 * No counterpart in the reference, whose walls are at rest
 * (src/simulation.rs:475-479): this is for rotating drums. */
rp_status rp_engine_set_mesh_motion(rp_engine* e, int32_t mesh, const double* origin, const double* omega);

/* Replaces GpuSimulation::step (src/gpu.rs:153-166): advances nsteps time
 * steps of Simulation::step_cpu semantics (src/simulation.rs:173-562).
 * Enqueues work on the engine's stream and returns without a host sync per
 * step. */
rp_status rp_engine_step(rp_engine* e, uint64_t nsteps);

/* Waits for all queued work of the engine. */
rp_status rp_engine_synchronize(rp_engine* e);

/* Number of particles after removals (Simulation.particles.len()). */
rp_status rp_engine_count(rp_engine* e, size_t* n);

/* Replaces read_particles (src/gpu.rs:168-213): synchronises and copies the
 * state back in API order.  Any pointer may be NULL (skipped).  Unlike the
 * reference, angular velocity is returned, not zeroed (gpu.rs:205-210). */
rp_status rp_engine_download(rp_engine* e, double* x3, double* v3, double* w3,
                             double* a3, double* tau3);

/* Kahan residuals, radius, mass and fixed flags in API order (NULL skips). */
rp_status rp_engine_download_aux(rp_engine* e, double* rx3, double* rv3,
                                 double* rw3, double* radius, double* mass,
                                 uint8_t* fixed);

/* Batched ensemble (no counterpart in the reference, where each Simulation
 * is its own object, src/simulation.rs:35-60): ninst independent
 * simulations that share this engine's parameters (dt, bounds, periodic
 * axes, models, gravity, meshes) but have their own particle and wall
 * materials, stepped together by every rp_engine_step -- one set of kernel
 * launches over all instances' particles.  counts[k] particles of instance k
 * follow those of instance k-1 in the arrays (the state arrays as in
 * rp_engine_upload_full; any but x3, radius and mass may be NULL); wall may
 * be NULL (the engine's wall material for all).  Particles of different instances never
 * interact.  Each instance evolves exactly as a single engine given that
 * instance alone.  Downloads list the instances in order, each in API
 * order.  Not available with pair recording.  The JKR surrogate serves the
 * instances whose E* and gamma lie in its envelope. */
rp_status rp_engine_upload_batch(rp_engine* e, int32_t ninst, const size_t* counts,
                                 const rp_material* particle, const rp_material* wall,
                                 const double* x3, const double* v3, const double* w3,
                                 const double* a3, const double* tau3, const double* rx3,
                                 const double* rv3, const double* rw3,
                                 const double* radius, const double* mass, const uint8_t* fixed);

/* Alive particles per instance (ninst entries; 1 for a single engine). */
rp_status rp_engine_instance_counts(rp_engine* e, size_t* counts);

/* Turns recording of the pair list on or off (off by default).  When on,
 * the last step of each rp_engine_step call records its contact pairs for
 * rp_engine_pairs. */
rp_status rp_engine_set_record(rp_engine* e, int on);

/* Reads an STL file (binary or ASCII, as Mesh::load_stl via stl_io,
 * src/mesh.rs:125-143) without an engine.  Call with v9 = n3 = NULL to get
 * the triangle count in *ntri, then with arrays of ntri*9 and ntri*3 floats.
 * RP_ERR_IO when the file cannot be read. */
rp_status rp_stl_read(const char* path, size_t* ntri, float* v9, float* n3);

/* Particle-particle contact list of the most recent step, in API order and,
 * per particle, in the reference's neighbour order (grid.rs:102-141 cell
 * walk, ascending index inside a cell; simulation.rs:317-357).  A pair the
 * reference's grid walk visits twice is listed twice.  offsets has n+1
 * entries; js receives up to cap entries; *total receives the full count.
 * Returns RP_ERR_INVALID (with *total set) when cap is too small. */
rp_status rp_engine_pairs(rp_engine* e, int64_t* offsets, int64_t* js,
                          size_t cap, size_t* total);

/* Tangential history of the physics-fix mode (no counterpart in the
 * reference, whose Contact map is internal, src/contact.rs:3-18): per
 * particle in API order, its loaded directed contacts -- partner's API index
 * (ascending; in a batched engine the index over all instances, as in the
 * downloads) and stored displacement xi (3 doubles) -- as the next step
 * will read them.  offsets has n+1 entries; partner and xi3 receive up to cap
 * entries (xi3: 3 per entry); *total receives the count.  Empty when no
 * history is kept (parity mode or Coulomb friction). */
rp_status rp_engine_history(rp_engine* e, int64_t* offsets, int64_t* partner, double* xi3, size_t cap,
                            size_t* total);

/* Loads a tangential history in the layout rp_engine_history returns (no
 * counterpart in the reference, whose contacts map never leaves its
 * Simulation, src/simulation.rs:30): the particles' loaded contacts, per particle in API order by
 * ascending partner index, as the next step will read them.  Lets a caller
 * that re-uploads a state (an edit between runs, an ensemble formed from
 * simulations that have run) keep the history it read before.  Replaces the
 * current tables; entries with partner -1 (a removed particle) are skipped.
 * RP_ERR_INVALID for offsets that do not run from 0 to total without
 * decreasing, a partner out of range or not ascending, or more entries for a
 * particle than its neighbour list can hold; RP_ERR_STATE for entries when no
 * history is kept (parity mode or Coulomb friction). */
rp_status rp_engine_set_history(rp_engine* e, const int64_t* offsets, const int64_t* partner,
                                const double* xi3, size_t total);

/* Cells of the broad phase, for checking them (no counterpart in the
 * reference, whose Grid keeps them private, src/grid.rs:30-46).  With
 * rebuild != 0 the neighbour lists are first rebuilt from the current
 * positions (results never depend on when that happens).  Per particle in
 * API order: ref_cell = the reference grid's cell index of its position
 * (grid.rs:30-40: coords = floor((x - bounds_min) / ref_cell_size), range
 * check against ceil(size / ref_cell_size) + 1 cells per axis, index
 * cx + cy dx + cz dx dy; -1 outside the grid), fine_key = the engine's own
 * sort key of the current position (instance * cells + cell number).
 * slot_api lists, in device slot order, the API index of each particle: the
 * permutation of the last rebuild, which sorts the particles by the key of
 * their position at that rebuild, ties by API index.  grid (22 doubles):
 * origin[3], 1/cell[3], cells[3], and per axis the cell-number terms
 * shift[3], hi[3], lo[3] (cell number of fine coordinate c on axis k =
 * (c >> shift) * hi + (c & (2^shift - 1)) * lo, summed over the axes),
 * periodic[3], cells per instance.  Any pointer may be NULL. */
rp_status rp_engine_cells(rp_engine* e, int32_t rebuild, int64_t* ref_cell, uint32_t* fine_key,
                          int64_t* slot_api, double* grid);

/* Autotuner (no counterpart in the reference): from the current state,
 * times each candidate Verlet skin (the current one, then 0.2 .. 0.005 of
 * the largest radius, never below the automatic skin's motion bound nor the
 * fall of a particle released at rest over 1,000 steps) and broad-phase cell
 * (automatic or half the largest cutoff + skin) over calls of `steps` steps
 * (at least 100) until it has rebuilt its lists 3 times (at most 10 calls),
 * keeps the
 * fastest in the engine's parameters (the current one unless beaten by 3 %)
 * and returns it (0 = automatic).  Where the force kernel has a zero-force
 * pass (JKR with the shortcut or surrogate, sJKR; Coulomb or spring
 * tangential law) it then times the other tiling of that pass
 * (rp_engine_set_tiling) with the chosen skin and cell and keeps the faster.
 * The particle state is restored, the physics-fix mode's tangential history
 * with it, so the simulation continues as if the call had not been made
 * (skin, cell size and tiling never change a result).  A
 * tuned skin is widened later, up to the automatic skin, when more than a
 * quarter of the rebuild checks of 200 steps rebuilt the lists. */
rp_status rp_engine_autotune(rp_engine* e, uint32_t steps, double* skin_out, double* cell_out);

/* Tiling of the force kernel's zero-force pass (no counterpart in the
 * reference): RP_TILING_LANE, one lane per particle walks its list;
 * RP_TILING_WAVE (default), a wavefront deals the list entries of its 64
 * particles over its lanes in turn.  Results are the same either way. */
#define RP_TILING_LANE 0
#define RP_TILING_WAVE 1
rp_status rp_engine_set_tiling(rp_engine* e, int32_t tiling);
rp_status rp_engine_get_tiling(rp_engine* e, int32_t* tiling);

/* Step loop in HIP graphs (on by default; no counterpart in the reference):
 * rp_engine_step replays a captured loop of 10 launches (a rebuild-chain
 * slot, the force kernels, the surrogate's kernels when active) where the
 * step count allows; the results are the same either way.  Timing runs
 * (rp_engine_set_timing) launch every kernel individually. */
rp_status rp_engine_set_graph(rp_engine* e, int on);

/* Turns per-kernel timing on or off (off by default).  When on, every
 * force-kernel launch, every rebuild chain, every first-half kernel, every
 * surrogate feature collection and every surrogate inference launch is
 * bracketed by HIP events on the engine's stream. */
rp_status rp_engine_set_timing(rp_engine* e, int on);

/* Synchronises, then returns the summed event times in milliseconds and the
 * number of timed launches since the last call, in arrays of RP_TIMING_KINDS:
 * index 0 = force kernel, 1 = rebuild chain (including its no-op launches),
 * 2 = first half kernel, 3 = JKR surrogate feature collection, 4 = JKR
 * surrogate network inference (BF16 MFMA). */
#define RP_TIMING_KINDS 5
rp_status rp_engine_timing(rp_engine* e, double* ms, uint64_t* count);

/* Diagnostics, no counterpart in the reference: copies `bytes` bytes
 * between two fresh device buffers with a plain streaming kernel (16-byte
 * non-temporal accesses) and returns the fastest of 5 launches.  A counter profile of the same process uses it to scale
 * FETCH_SIZE / WRITE_SIZE by a kernel of known traffic. */
rp_status rp_probe_copy(size_t bytes, double* ms);

/* ---- JKR contact-radius surrogate (no counterpart in the reference) ----
 * A 4 -> 64 -> 64 -> 2 MLP (tanh; BF16 MFMA, FP32 accumulate) mapping a
 * contact's scaled (d, ln R*, ln E*, ln gamma) to its scaled contact radius
 * a / s and force F / (pi gamma R*); it replaces the initial guess of the
 * Newton solve of compute_jkr_force (physics.rs:132-219).  Parameters are a
 * packed FP32 vector of RP_MLP_PARAMS floats: W1[4x64] b1[64] W2[64x64]
 * b2[64] W3[64x2] b3[2] (row-major, input index first). */
#define RP_MLP_PARAMS 4610

/* Trains a network on the GPU: nsamples samples drawn uniformly over the
 * envelope, labelled by the FP64 Newton solve; epochs of mini-batch Adam
 * (batch a multiple of 16), cosine-decayed learning rate lr.  Writes the
 * parameters and, when epoch_loss is not NULL, the last batch's mean squared
 * error (scaled outputs) of every epoch. */
rp_status rp_jkr_surrogate_train(const rp_jkr_envelope* env, uint64_t seed, size_t nsamples, int epochs,
                                 int batch, float lr, float* params_out, double* epoch_loss);

/* Labelled samples as used by the training (device generated): x_scaled
 * rows*4, t_scaled rows*2, raw8 rows*8 = (delta, R*, E*, gamma, a, F at a
 * without the pull-off clip, d, 1 if the Newton solve converged).  Any output
 * may be NULL. */
rp_status rp_jkr_samples(const rp_jkr_envelope* env, uint64_t seed, size_t n, float* x_scaled,
                         float* t_scaled, double* raw8);

/* Input and output scaling of an envelope (in: d, ln R*, ln E*, ln gamma;
 * out: a / s, F / (pi gamma R*)); values map linearly to [-1, 1]. */
rp_status rp_jkr_envelope_scale(const rp_jkr_envelope* env, float* in_lo4, float* in_hi4, float* out_lo2,
                                float* out_hi2);

/* Inference (the engine's fused MFMA kernel) on rows scaled inputs. */
rp_status rp_mlp_infer(const float* params, const float* x_scaled, size_t rows, float* y_scaled);

/* One forward + backward pass on rows (a multiple of 16) samples: gradient
 * of the mean squared error of the scaled outputs, packed like the
 * parameters, and the loss. */
rp_status rp_mlp_gradient(const float* params, const float* x_scaled, const float* t_scaled, size_t rows,
                          float* grad, double* loss);

/* Inference that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate device memory -- code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  The same
 * network and kernel as rp_mlp_infer, but every pointer is device memory the
 * caller owns: params_dev (RP_MLP_PARAMS floats), x_scaled_dev (rows x 4),
 * y_scaled_dev (rows x 2).  The packed network (BF16 weights in MFMA
 * fragment order, FP32 biases) is written into a static device array of
 * RP_MLP_NOALLOC_SCRATCH bytes that each GPU holds until the process ends;
 * the work is enqueued on `stream` (a hipStream_t, NULL = the null stream)
 * and the call returns without waiting.  rows above RP_MLP_NOALLOC_MAX_ROWS
 * are refused (RP_ERR_INVALID).  Every call on a GPU shares that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race.
 * The first call on a device looks the array's address up
 * (hipGetSymbolAddress); make it outside a stream capture. */
#define RP_MLP_NOALLOC_SCRATCH 16384
#define RP_MLP_NOALLOC_MAX_ROWS (1 << 28)
rp_status rp_mlp_infer_noalloc(const float* params_dev, const float* x_scaled_dev, size_t rows,
                               float* y_scaled_dev, void* stream);

/* One 16x16x16 BF16 MFMA tile C = A B (row-major 16x16 FP32 in and out; A and
 * B rounded to BF16) through the library's product kernel. */
rp_status rp_mfma_probe(const float* a16x16, const float* b16x16, float* c16x16);

/* Message of the last failed surrogate call on this thread. */
const char* rp_surrogate_last_error(void);

/* JKR solver of an engine (no counterpart in the reference, whose JKR force
 * is always its Newton solve, physics.rs:132-219):
 *   RP_JKR_EXACT      the reference's solve as written (default);
 *   RP_JKR_SHORTCUT   the same, except that a contact whose overlap lies
 *                     below t d_min - 1e-3 R* (no iterate can pass the
 *                     residual check) returns the zero force and rejected
 *                     root at once -- identical results;
 *   RP_JKR_SURROGATE  the shortcut, plus the network for particle-particle
 *                     contacts inside the envelope (away from
 *                     0 < d < 0.2, where the reference's start lies on the
 *                     wrong branch): its radius is polished by FP64 Newton
 *                     steps; contacts outside the envelope, wall contacts
 *                     and failed polishes use the exact solve.
 * check != 0 also runs the exact solve for every served contact and records
 * the relative force errors (rp_engine_jkr_stats). */
enum { RP_JKR_EXACT = 0, RP_JKR_SHORTCUT = 1, RP_JKR_SURROGATE = 2 };
rp_status rp_engine_set_jkr_solver(rp_engine* e, int32_t solver, const rp_jkr_envelope* env,
                                   const float* params, int32_t check);

/* Surrogate counters of the most recent force launch. */
typedef struct rp_jkr_stats {
  uint64_t served;         /* contacts whose radius came from the network */
  uint64_t fallback;       /* served-envelope contacts that took the exact solve */
  uint64_t rows_max;       /* most rows one of the 16 claim ranges was asked for in a step */
  uint64_t rows_capacity;  /* feature rows the engine holds */
  double err_max, err_mean;  /* check mode: |F - F_exact| / max(|F_exact|, (2/3) pi gamma R*) */
  double raw_max, raw_mean;  /* check mode: the same for the network's force output */
} rp_jkr_stats;
rp_status rp_engine_jkr_stats(rp_engine* e, rp_jkr_stats* out);

/* Counters (rp_stats). */
rp_status rp_engine_stats(rp_engine* e, rp_stats* out);

/* Releases every device resource of the engine. */
void rp_engine_destroy(rp_engine* e);

/* Message of the last failed call on this thread ("" if none). */
const char* rp_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RP_ABI_H */
