/*
 * psum_hip.h -- C ABI of the MI355X (gfx950) electrostatic PIC step engine.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory;
 * `stream` is a hipStream_t (NULL = the legacy default stream).  Every entry
 * point returns a psum_status; nothing throws across this boundary.
 *
 * Each entry point replaces a reference (PSuM) interface; the citation is
 * given beside it.  The field-solver plugin ABI that PSuM's Poisson
 * frontends bind (register_solver / get_solver_funcs and the func_table) is
 * declared in psum_field_solver.h and implemented by the same library.
 */
#ifndef PSUM_HIP_H
#define PSUM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int psum_status;
#define PSUM_OK 0
#define PSUM_ERR_ARG 1          /* bad argument (null pointer, size, dtype) */
#define PSUM_ERR_HIP 2          /* a HIP runtime call failed */
#define PSUM_ERR_UNSUPPORTED 3  /* input outside what the engine implements */
#define PSUM_ERR_NOMEM 4
#define PSUM_ERR_STATE 5        /* call order violated (e.g. solve before set_matrix) */

#define PSUM_F64 0
#define PSUM_F32 1
/* adjoint kernels only: FP32 positions with the adjoint state (lx, lv),
 * phi / lrho and the arithmetic in FP64 */
#define PSUM_F32_ADJ64 2

#define PSUM_KEY_INVALID 0xFFFFFFFFu

/* Node grid, reference semantics of simple_grid<D> (src/field/simple_grid.hpp:16-307):
 * cells[d] cells and cells[d]+1 nodes per dimension, row-major node index
 * with the last dimension fastest (n2i, :107-112), delta = (ub-lb)/cells and
 * delta_r = 1/delta computed once (:46-47).  periodic[d] picks the particle
 * boundary: 1 wraps into [plo, plo+plen) (TSI/main.cpp:71-72), 0 absorbs a
 * particle leaving [lb, ub) (inGrid, :168-174; validator.hpp:20-22). */
typedef struct psum_grid {
    int32_t dim;
    int32_t cells[3];
    double lb[3];
    double ub[3];
    double delta[3];
    double delta_r[3];
    int32_t periodic[3];
    int32_t pad_;
    double plo[3];
    double plen[3];
} psum_grid;

/* Last error message of the calling thread (static storage). */
const char* psum_hip_last_error(void);

/* ---- particle container ------------------------------------------------
 * Particles are SoA: pos[d][i], vel[d][i] for d < dim, all of one dtype.
 * Invalid particles have non-finite pos[0][i] (pos_x_nan_is_invalid,
 * src/particle_container/validator.hpp:18-26). */

/* Cell binning: keys[i] = n2i of the wrapped base node cN(x_i)
 * (simple_grid.hpp:89-95,107-112,338-343), PSUM_KEY_INVALID for invalid
 * particles.  The index arithmetic is FP64 for both dtypes.
 * No reference counterpart (the reference never bins); it feeds the sort. */
psum_status psum_hip_bin(const psum_grid* g, int dtype, void* const* pos, int64_t n,
                         uint32_t* keys, void* stream);

/* Bytes of device workspace psum_hip_sort needs for n keys. */
size_t psum_hip_sort_workspace(int64_t n);

/* Stable LSD radix sort of keys in [0, nkeys) plus PSUM_KEY_INVALID (last).
 * perm[r] = index of the r-th particle; keys_out = sorted keys.
 * No reference counterpart; defines the sort order the engine uses. */
psum_status psum_hip_sort(const uint32_t* keys, int64_t n, uint32_t nkeys, uint32_t* keys_out,
                          uint32_t* perm, void* work, size_t work_bytes, void* stream);

/* dst[c][i] = src[c][perm[i]] for c < ncomp arrays of `dtype`. */
psum_status psum_hip_permute(int dtype, int ncomp, void* const* src, void* const* dst,
                             const uint32_t* perm, int64_t n, void* stream);

/* Stable stream compaction of valid particles into out_pos/out_vel;
 * *count_host receives the survivor count.  Replaces
 * particle_group::compress (particle_container/particle_group.hpp:223-238),
 * whose atomic push_back order is nondeterministic; here the order is the
 * original index order. */
size_t psum_hip_compact_workspace(int64_t n);
psum_status psum_hip_compact(int dtype, int dim, void* const* pos, void* const* vel, int64_t n,
                             void* const* out_pos, void* const* out_vel, int64_t* count_host,
                             void* work, size_t work_bytes, void* stream);

/* psum_hip_compact without a workspace argument and without allocating:
 * its scratch (block counts and scan partial sums, 8.06 MiB) is a static
 * device array of the library, one per GPU, held until the process ends.
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  n <= 2^31 (PSUM_ERR_ARG beyond).  Calls that overlap
 * on one GPU share the scratch and race: order them on one stream. */
psum_status psum_hip_compact_noalloc(int dtype, int dim, void* const* pos, void* const* vel, int64_t n,
                                     void* const* out_pos, void* const* out_vel, int64_t* count_host,
                                     void* stream);

/* ---- field: CIC deposit and gather ------------------------------------
 * Deposit (add_back, src/field/device_field.hpp:295-310 ->
 * field_interp.hpp:63-75 -> atomic_add.hpp:34-48): for every valid
 * particle and every support node with |coef| > 1e-10,
 *     acc[node] += llrint((w * coef) * scale)          (int64, exact)
 * scale must be a power of two with |w| * scale <= 2^50.  Integer
 * accumulation makes the result independent of particle order and thread
 * schedule; no float atomics.  acc is accumulated into (zero it first).
 * counters (device, 4 entries, may be NULL): [0] particles whose support was
 * out of range, [1] blocks that fell back from the LDS tile to global
 * integer atomics, [3] adds to acc whose int64 partial sum wrapped (the
 * accumulator's overflow flag: zero means every node's sum stayed in range,
 * so acc is exact; a node overflows after about 2^63 / (|w| * scale) full
 * terms).  Entries are added to, never cleared. */
psum_status psum_hip_deposit_cic(const psum_grid* g, int dtype, void* const* pos, int64_t n,
                                 double w, double scale, int64_t* acc, int64_t* counters,
                                 void* stream);

/* rho[i] = (double)acc[i] * inv_scale + offset */
psum_status psum_hip_fixed_to_double(const int64_t* acc, int64_t nn, double inv_scale,
                                     double offset, double* rho, void* stream);

/* Gather + push (the push lambda of TSI/main.cpp:59-75 and
 * big_bang/stage_6.cpp:48-65 with interp_diff, device_field.hpp:270-284):
 *     E = -grad(phi)(x)   (gradient of the multilinear interpolant, terms
 *                          with |coef| <= 1e-10 skipped, field_interp.hpp:108)
 *     v = v + (E*qm)*dt ;  x = x + v*dt
 * then periodic wrap / absorbing wall per psum_grid.  phi has the particle
 * dtype.  When acc != NULL the new positions are also deposited as in
 * psum_hip_deposit_cic with (w_dep, scale) -- the next step's charge.
 * diag (device, may be NULL) receives per-block partial sums of
 * 1/2 m|v|^2 and m v_d over the velocities entering the step; reduce them
 * with psum_hip_reduce_diag.  counters (device, 4 entries, may be NULL):
 * [0] support out of range, [1] tile fallbacks, [2] particles absorbed this
 * call, [3] int64 wraps of the deposit (as psum_hip_deposit_cic). */
int64_t psum_hip_push_blocks(int64_t n);
psum_status psum_hip_gather_push(const psum_grid* g, int dtype, void* const* pos,
                                 void* const* vel, int64_t n, const void* phi, double qm,
                                 double dt, double mass, double w_dep, double scale, int64_t* acc,
                                 double* diag, int64_t* counters, void* stream);

/* Higher-order coupling (interp_B_spline / add_back_B_spline of
 * higher_order_interp.hpp:196-248, selected by TSI's interp_order,
 * TSI/main.cpp:29,54,64): quadratic (order 2), cubic (3) and quartic (4)
 * B-spline weights (simple_interpolation.hpp:105-184) on order+1 nodes per
 * dimension around the nearest node (odd support) or the corner node (even
 * support), wrapped once (simple_grid.hpp:316-385).  Weights are FP64 for
 * both dtypes.  order 1 forwards to the CIC entry points.  2D and 3D.
 * Deposit: as psum_hip_deposit_cic with these weights (exact int64 sums,
 * scale a power of two, |w| * scale <= 2^50; counters [0] out of range,
 * [1] tile fallbacks, [3] int64 wraps).
 * Gather + push: E = -grad(phi) summed in FP64 and rounded to the dtype,
 * then the kick, drift and boundaries of psum_hip_gather_push (no fused
 * deposit); diag has psum_hip_push_blocks(n) blocks as there. */
psum_status psum_hip_deposit_bspline(const psum_grid* g, int dtype, int order, void* const* pos, int64_t n,
                                     double w, double scale, int64_t* acc, int64_t* counters, void* stream);
psum_status psum_hip_gather_push_bspline(const psum_grid* g, int dtype, int order, void* const* pos,
                                         void* const* vel, int64_t n, const void* phi, double qm, double dt,
                                         double mass, double* diag, int64_t* counters, void* stream);

/* ---- Monte Carlo collisions --------------------------------------------
 * One call applies a sequence of collision processes (incident species
 * against a background target each) to every valid particle, between the
 * push and the next deposit.  Replaces the collision_processor /
 * sequence_executor of particle_collision/foundation.hpp:17-107 run by
 * execute_mcc_model (mcc_model.hpp:184-218), with the actions of the He
 * CCP model (application/benchmark/CCP/custom_mccm.hpp:29-104) as channel
 * kinds.  Per process and particle:
 *   dens  = CIC interpolation of the process's node field (|coef| > 1e-10,
 *           0 outside the grid) or the process's constant density
 *   v2 = |v|^2, r = R(), sigma_k = cross section of channel k at v2
 *   chosen = first k with r*sum <= sigma_0 + .. + sigma_k (sum > 0)
 *   collide when chosen and R() < ((sum*dt)*dens)*|v|
 * R() advances the particle's uint32 seed as the LCG
 * x = x*1664525 + 1013904223 and returns x / (2^32-1)
 * (random/rander.hpp:40-74).  The collision action draws from a copy of
 * the stream; the seed keeps the state after the process's own draws, as
 * in the reference, whose process wrapper writes its state back last.
 * Velocities always have three components (vel[0..2]); positions have
 * g->dim.  Arithmetic is FP64 for both dtypes.
 *
 * Cross sections (the form of CCP/HeCCS.h): piecewise polynomials of degree
 * 4 in x = ln(v2 * arg_scale), segment k while x < seg_hi[k] (the last
 * segment otherwise), c4 x^4 + c3 x^3 + c2 x^2 + c1 x + c0 summed left to
 * right with x^k as repeated products, exponentiated when exp_form; nseg 0
 * is the constant coef[0][4]; arg_scale > 0.
 *
 * Ionization products (a secondary electron with half the energy left
 * after the threshold, an ion with a Maxwellian target velocity, each with
 * a seed R()*(2^31-1)) are appended to `electrons` and `ions` at their
 * current counts in (particle, process) order -- deterministic, unlike the
 * reference's atomic push_back.  `electrons` may be the incident species
 * itself (products land after its n particles and do not collide in this
 * call).  When the products would exceed a capacity, nothing is modified
 * and PSUM_ERR_ARG is returned with *produced = the number needed. */
#define PSUM_MCC_MAX_SEG 12
#define PSUM_MCC_MAX_CHAN 8
#define PSUM_MCC_MAX_PROC 4
#define PSUM_MCC_ELASTIC 0        /* col_ele_to_atom_elastic: v = |v| u, u isotropic */
#define PSUM_MCC_EXCITATION 1     /* |v|^2 -= 2 e thr / m_incident, new direction (if > 0) */
#define PSUM_MCC_IONIZATION 2     /* E' = (E - thr e)/2 for the incident and a new electron, new ion */
#define PSUM_MCC_ION_ISOTROPIC 3  /* isotropic in the centre-of-mass frame of a Maxwellian atom */
#define PSUM_MCC_ION_BACKWARD 4   /* charge exchange: v = Maxwellian target velocity */

typedef struct {
    int32_t kind;
    int32_t nseg;
    int32_t exp_form;
    int32_t pad_;
    double arg_scale;
    double threshold_ev;
    double seg_hi[PSUM_MCC_MAX_SEG];
    double coef[PSUM_MCC_MAX_SEG][5];
} psum_mcc_channel;

typedef struct {
    int32_t nchan;
    int32_t pad_;
    double density;             /* when the process has no density field */
    double unit_charge;         /* J per eV */
    double incident_mass;       /* kg */
    double target_mass;         /* kg, background atoms (ion products, Maxwellians) */
    double target_temperature;  /* K */
    double boltzmann;           /* J/K */
    psum_mcc_channel chan[PSUM_MCC_MAX_CHAN];
} psum_mcc_process;

typedef struct {
    void* pos[3];               /* dim arrays */
    void* vel[3];               /* 3 arrays */
    uint32_t* seed;
    int64_t n;                  /* particles present */
    int64_t cap;                /* capacity of every array */
} psum_mcc_species;

/* Bytes of device workspace psum_hip_mcc needs for n incident particles. */
size_t psum_hip_mcc_workspace(int64_t n, int nproc);
/* density: nproc device pointers to FP64 node fields of g (or NULL), or
 * NULL for all constant.  electrons / ions may be NULL when no channel is
 * an ionization.  *produced (host) = product pairs appended (the caller
 * adds it to both counts).  outcome (device, may be NULL): n*nproc int32,
 * the channel that collided per (particle, process), -1 none, -2 invalid
 * particle.  Synchronizes the stream when an ionization channel exists
 * (the product count is checked against the capacities first). */
psum_status psum_hip_mcc(const psum_grid* g, int dtype, const psum_mcc_process* procs, int nproc,
                         const double* const* density, const psum_mcc_species* incident,
                         const psum_mcc_species* electrons, const psum_mcc_species* ions, double dt,
                         int64_t* produced, int32_t* outcome, void* work, size_t work_bytes,
                         void* stream);

/* ---- triangle-mesh particle boundary -----------------------------------
 * triangle_mesh_trigger (particle_boundary/triangle_mesh_trigger.hpp:
 * 30-437) and boundary_router + basic_material_set (boundary_router.hpp:
 * 20-92, basic_material_set.hpp:15-119).  psum_mesh_create builds the
 * partition on the host: tri = ntri x 9 doubles (p1, p2, p3), ids[ntri]
 * (the router reads an id as the triangle's material), an I x J x K
 * uniform partition of `box` (x0 y0 z0 x1 y1 z1; NULL = the triangles'
 * bounding box widened by the reference's fractions, :378-392).  Normals
 * are (p2-p1) x (p3-p1) normalised; zero-area triangles are dropped.  A
 * triangle is listed in each cell of its clamped bounding-box range whose
 * cell box it overlaps (geometry.hpp:292-350), in triangle order.
 * psum_hip_mesh_detect traces segments p1 -> p2 (3 coordinate arrays of
 * dtype each): the cells along the segment from its start (restrict mode,
 * geometry.hpp:40-206; fast mode: start and end cells only), the first
 * cell with a hit giving the smallest Moller-Trumbore parameter k in
 * (0, 1] among its triangles (first on ties); k, normal (3 per segment)
 * and id (-1: no hit) out (k / normal may be NULL).  Segments with a
 * non-finite end are not traced.  A traversal cell outside the partition
 * (rounding at the entry face) is empty here; the reference reads outside
 * its arrays there. */
#define PSUM_MAT_DIFFUSE 0          /* cosine-law direction, |v| kept */
#define PSUM_MAT_SPECULAR 1         /* v -= 2 (v.n) n */
#define PSUM_MAT_MAXWELL_REFLECT 2  /* half-Maxwellian re-emission at (temperature, mass) */
#define PSUM_MAT_ABSORB 3           /* position x = NaN (invalid) */
void* psum_mesh_create(const double* tri, const int32_t* ids, int64_t ntri, int I, int J, int K,
                       const double* box);
/* info[0] triangles kept, [1] partition entries, [2] cells, [3] largest
 * id; box[6] (either may be NULL) */
psum_status psum_mesh_info(void* m, int64_t* info, double* box);
/* host copies of the partition: cell_start[cells+1], cell_tri[entries],
 * tri12[12*ntri] (vertices then unit normal; zeros for dropped triangles) */
psum_status psum_mesh_partition(void* m, uint32_t* cell_start, uint32_t* cell_tri, double* tri12);
/* triangle_mesh_trigger::set_fast / set_restrict (the router's default) */
psum_status psum_mesh_set_fast(void* m, int fast);
psum_status psum_hip_mesh_detect(void* m, int dtype, const void* const* p1, const void* const* p2, int64_t n,
                                 double* k, double* normal, int32_t* id, void* stream);
/* boundary_router::deal for every valid particle: segment old -> pos; on a
 * hit at h = old + k (pos - old) the material of the triangle's id acts on
 * pos, vel and the particle's seed (RandV_Cosine / RandV_halfMaxw draws,
 * random/rand_functions.hpp:31-53).  Triangle ids must be 0..3.
 * hit_id (device, may be NULL) per particle; *hits (device, may be NULL)
 * is incremented by the number of hits. */
psum_status psum_hip_mesh_boundary(void* m, int dtype, const void* const* old, void* const* pos,
                                   void* const* vel, uint32_t* seed, int64_t n, double temperature,
                                   double mass, int32_t* hit_id, unsigned long long* hits, void* stream);
void psum_mesh_destroy(void* m);
/* Device view of a partition for kernels compiled elsewhere (the header-only
 * C++ API, include/psum/particle_boundary.hpp): valid until
 * psum_mesh_destroy. */
typedef struct {
    int32_t I, J, K, fast;
    double box[6];
    double rd[3];                 /* cells per unit length */
    const uint32_t* cell_start;   /* I*J*K + 1 */
    const uint32_t* cell_tri;
    const double* tri12;          /* 12 per triangle: vertices, unit normal */
    const int32_t* ids;
    const float* udf;             /* I*J*K distance field, NULL = no skip */
} psum_mesh_view;
psum_status psum_mesh_get_view(void* m, psum_mesh_view* out);
/* triangle_mesh_trigger::set_acceleration_UDF (triangle_mesh_trigger.hpp:
 * 383-394, geometry.hpp:343-460): enabled, a step that starts in a cell
 * whose distance field (cell units, from the occupied cells; computed on
 * the host at the first enable) exceeds one cell and is shorter than it is
 * not traced, and a longer one starts its traversal that far along.  Only
 * the full traversal (not fast mode) uses it. */
psum_status psum_mesh_set_udf(void* m, int enabled);

/* out[k] = sum over blocks of diag[4*b + k], k < 4, in a fixed order
 * (deterministic).  work: psum_hip_reduce_workspace() bytes of device
 * memory (may be NULL when nblocks <= 1024). */
size_t psum_hip_reduce_workspace(void);
psum_status psum_hip_reduce_diag(const double* diag, int64_t nblocks, double* out, void* work,
                                 void* stream);

/* Field energy 1/2 eps sum_d sum_edges ((phi[n+e_d]-phi[n]) * delta_r_d)^2 * dV
 * (TSI/main.cpp:138-141 in D dimensions; periodic dims include the wrap
 * edge).  out: 4 doubles on the device, out[0] = energy (out[1..3] scratch);
 * work: device workspace of psum_hip_field_energy_workspace(g) bytes. */
size_t psum_hip_field_energy_workspace(const psum_grid* g);
psum_status psum_hip_field_energy(const psum_grid* g, const double* phi, double eps, double* out,
                                  void* work, size_t work_bytes, void* stream);

/* Element-wise precision conversions on the device. */
psum_status psum_hip_f64_to_f32(const double* in, float* out, int64_t n, void* stream);

/* ---- step engine -------------------------------------------------------
 * One step of the reference user loop (TSI/main.cpp:121-146): deposit ->
 * Poisson_solver::solve through the multigrid backend -> gather + push ->
 * boundaries, with sort every `sort_every` steps.  Particle state is
 * resident on the device; nothing crosses PCIe inside psum_stepper_run. */
typedef struct psum_species {
    double q, m, w;          /* charge, mass, macro-particle weight */
    int64_t n;               /* particle count */
    int64_t capacity;        /* allocated slots (>= n) */
} psum_species;

typedef struct psum_step_config {
    psum_grid grid;
    int32_t dtype;           /* particle state: PSUM_F64 or PSUM_F32 */
    int32_t mixed_mg;        /* 0: FP64 V-cycles; 1: FP32 V-cycles inside FP64 refinement */
    int32_t nspecies;        /* 1..4 */
    int32_t depth, pre, post, max_cycles;
    int32_t sort_every;      /* 0 = never re-sort */
    int32_t mg_graph;        /* 1: replay each V-cycle / refinement iteration from a hipGraph */
    int32_t extrapolate;     /* 1: warm start phi by linear extrapolation 2 phi_n - phi_{n-1} */
    double tol;              /* relative residual target ||b-A phi|| / ||b|| */
    double eps, dt, rho_bg;
    psum_species species[4];
    const uint8_t* fixed;    /* host mask of Dirichlet nodes (NULL = none); all-periodic
                                grids pin node 0 */
    const double* fixed_val; /* host values for fixed nodes (NULL = 0) */
} psum_step_config;

void* psum_stepper_create(const psum_step_config* cfg);
/* copy species s in (d < dim arrays of n particles, dtype of cfg; host or
 * device memory).  Starts a new run: the step counter that sets the sort
 * cadence (a sort before step k when k % sort_every == 0) restarts at 0. */
psum_status psum_stepper_load(void* h, int s, const void* const* pos, const void* const* vel,
                              int64_t n);
/* device pointers of species s state (for zero-copy wrapping) */
psum_status psum_stepper_state(void* h, int s, void** pos, void** vel, int64_t* n);
/* run `steps` steps; hist (host, may be NULL) gets 5 doubles per step:
 * field energy, kinetic energy, momentum x, y, z, as in orc_pic_run_f64 */
psum_status psum_stepper_run(void* h, int steps, double* hist);
/* copy phi (FP64, host) */
psum_status psum_stepper_phi(void* h, double* phi_host);
/* copy rho of the current positions (FP64, host, may be NULL): the charge
 * density the next step's solve starts from, acc * 2^-S + rho_bg before
 * neutralisation and fixed-node values, as add_back leaves it
 * (device_field.hpp:295-310); *scale (may be NULL) = 2^S of the int64
 * fixed-point accumulator */
psum_status psum_stepper_rho(void* h, double* rho_host, double* scale);
/* total V-cycles so far, last relative residual; counters (host, 4
 * entries, may be NULL): [0] support out of range, [1] tile fallbacks,
 * [2] particles absorbed, [3] wrapped adds to the int64 charge accumulator
 * since the last load (psum_stepper_run and psum_batch_run fail with an
 * overflow message when it is non-zero) */
psum_status psum_stepper_stats(void* h, int64_t* cycles, double* last_rel_res, int64_t* counters);
/* steps taken since creation (a sort runs before step k when
 * k % sort_every == 0); restore() sets it from the checkpoint */
psum_status psum_stepper_step_count(void* h, int64_t* steps);
psum_status psum_stepper_sync(void* h);
/* HIP-event timing of the fused gather/push/deposit launches on the engine's
 * stream (reset by set_timing): total ms, launches, particles pushed. */
psum_status psum_stepper_set_timing(void* h, int on);
psum_status psum_stepper_push_timing(void* h, double* ms, int64_t* launches, int64_t* particles);
/* Loss L = field energy (psum_hip_field_energy) of phi solved from the
 * positions after `steps` steps from the current state (no re-sorting). */
psum_status psum_stepper_forward_loss(void* h, int steps, double* loss);
/* Reverse mode over the step (no reference counterpart: the reference has
 * no adjoint), for L = field energy after `steps` steps from the current
 * state.
 * psum_stepper_record: forward with a checkpoint (x, v, phi, solver state)
 *   every `ckpt` steps; *loss = L; *t_ms = HIP-event time.  Sorting is
 *   suspended until psum_stepper_reverse returns.
 * psum_stepper_reverse: seed * dL/dx0 and seed * dL/dv0 into grad_x / grad_v
 *   (device arrays of the particle dtype, index s*3 + d for species s,
 *   dimension d; NULL entries are skipped), recomputing each segment from its
 *   checkpoint.  Restores the state of the psum_stepper_record call.
 *   *t_ms = HIP-event time of the sweep (recompute included);
 *   *solve_cycles = V-cycles of its adjoint solves.
 * psum_stepper_adjoint: record + reverse with seed 1. */
psum_status psum_stepper_record(void* h, int steps, int ckpt, double* loss, double* t_ms);
psum_status psum_stepper_reverse(void* h, double seed, void* const* grad_x, void* const* grad_v,
                                 double* t_ms, int64_t* solve_cycles);
psum_status psum_stepper_adjoint(void* h, int steps, int ckpt, double* loss, void* const* grad_x,
                                 void* const* grad_v, double* t_fwd_ms, double* t_bwd_ms);
void psum_stepper_destroy(void* h);

/* Sort every species now, as the step does every sort_every steps: stable
 * LSD radix sort by the key bin * 2^dim + sum_d [v_d >= 0] 2^(dim-1-d), bin =
 * the 2^sort_shift-cell bin of x (the cell key of psum_hip_bin at shift 0);
 * absorbed particles are dropped.  No reference counterpart. */
psum_status psum_stepper_sort(void* h);

/* Interpolation order of the step (TSI's interp_order, TSI/main.cpp:29,54,
 * 64): 1 = CIC (default, fused push + deposit), 2..4 = B-splines through
 * psum_hip_gather_push_bspline / psum_hip_deposit_bspline.  The adjoint
 * entry points (record / reverse / adjoint) accept order 1 only. */
psum_status psum_stepper_set_interp_order(void* h, int order);

/* Multigrid kernel-shape options that may change after creation, as
 * whitespace "key value" pairs: "fuse 0|1" (2D fused tile smoothing and
 * residual + restriction, default 1), "tail 0|1" (the coarse end of each
 * V-cycle, levels of at most 33^2 / 9^3 nodes down to a coarsest level of
 * at most 10^2 / 5^3, in one workgroup; default 0) and "line 0|1" (3D 7-point
 * levels in chunks of four nodes per thread, 3D prolongation over planes,
 * default 1).  Results are bit-identical either way.  A batch instance's
 * change makes the batch capture its step graph again at its next run.  No
 * reference counterpart (tuning knobs). */
psum_status psum_stepper_mg_tuning(void* h, const char* options);
/* Smoothing sweeps per level before and after the coarse-grid correction
 * (the backend's pre_relax / post_relax options, backend.hpp:84-92),
 * changed after creation: 0 <= pre, post <= 64, pre + post >= 1.  Changes
 * the iterates (and so the cycles a tolerance needs), not the fixed point. */
psum_status psum_stepper_set_smoothing(void* h, int pre, int post);
/* Autotuner for the smoothing sweeps: for each candidate (pre[c], post[c])
 * a sort, one untimed step (residual checked after every cycle) and `steps`
 * steps timed with HIP events (no sorts, tolerance-controlled solves); the
 * fastest is kept and returned in *best_pre / *best_post, ms (may be NULL)
 * gets n per-step times.  The state advances.  No reference counterpart. */
psum_status psum_stepper_tune_smoothing(void* h, const int* pre, const int* post, int n, int steps,
                                        int* best_pre, int* best_post, double* ms);

/* Checkpoint files in the reference's MAS format (src/serialization/mas.hpp:
 * 49-453; particle_group / device_field save(), particle_group.hpp:272-275,
 * device_field.hpp:327-331): species<q>.data.content.{position,velocity}
 * (n x dim x 1), phi.{location,grid.*,content.*} and psum.* scalars (step
 * counter, solver cycle predictors, extrapolation state).  restore() needs a
 * stepper of the same grid, dtype and species count; the run continues
 * bit-identically to one that was never interrupted. */
psum_status psum_stepper_save(void* h, const char* path);
psum_status psum_stepper_restore(void* h, const char* path);

/* Solver and sort schedule of a stepper: fixed_cycles > 0 switches its
 * multigrid to that many V-cycles per solve (one per refinement iteration,
 * no tolerance check); sorts every sort_every steps (0 = never) on
 * 2^sort_shift-cell bins.  sort_every is the interval of the species
 * without one of their own (psum_stepper_set_species_sort); those keep
 * theirs.  No reference counterpart (tuning knobs). */
psum_status psum_stepper_set_schedule(void* h, int fixed_cycles, int sort_every, int sort_shift);
/* Species s sorts every `every` steps (every > 0; a sort before step k when
 * k % every == 0) instead of on the stepper's sort_every, or on sort_every
 * again (every = 0).  psum_stepper_tune sets every species' interval (slow
 * species, such as ions, keep their order much longer than fast ones).  No
 * reference counterpart (tuning knob). */
psum_status psum_stepper_set_species_sort(void* h, int species, int every);
/* The interval species s sorts on: its own, else sort_every (0: never). */
psum_status psum_stepper_species_sort(void* h, int species, int* every);

/* FD Poisson matrix A = -div(eps grad) of the reference frontends
 * (Poisson_solver_2d.hpp:182-410, Poisson_solver_3d.hpp:178-431, Cartesian,
 * constant eps): identity rows at nodes with fixed[i] != 0 (NULL: none),
 * cyclic neighbours over cells+1 nodes otherwise, as host COO in row-major
 * order (set_matrix takes any order).  With rows == NULL only *nnz is
 * returned; else up to cap entries are written. */
psum_status psum_poisson_assemble(const psum_grid* g, double eps, const uint8_t* fixed, int64_t cap,
                                  unsigned long long* rows, unsigned long long* cols, double* vals,
                                  int64_t* nnz);

/* Mixed (Robin/Neumann) boundary node (mixed_boundary.hpp:217-221):
 * value_coef * phi + grad_coef * dphi/dn = value at node `element` on side
 * `direction` -- 2D: N=0 S=1 (axis 1, +/-), E=2 W=3 (axis 0, +/-); 3D:
 * Xpos=0 Xneg=1 Ypos=2 Yneg=3 Zpos=4 Zneg=5 (foundation.hpp:43-51). */
typedef struct psum_mixed_node {
    unsigned long long element;
    int32_t direction;
    int32_t pad_;
    double value_coef;
    double grad_coef;
} psum_mixed_node;

/* permittivity at a face midpoint; pos holds dim coordinates */
typedef double (*psum_eps_fn)(const double* pos, void* user);

typedef struct psum_poisson_spec {
    int32_t cylindrical;            /* 2D only: axis 0 = z, axis 1 = r */
    int32_t pad_;
    double eps;                     /* used when eps_fn is NULL */
    psum_eps_fn eps_fn;
    void* eps_user;
    const uint8_t* fixed;           /* node mask or NULL */
    const psum_mixed_node* mixed;   /* n_mixed nodes or NULL */
    int64_t n_mixed;
} psum_poisson_spec;

/* General frontend assembly (Poisson_solver_2d.hpp:182-410 with the
 * ghost_element_manager elimination, ghost_element_manager.hpp:56-117):
 * each mixed node adds a ghost value beyond its side, defined by the
 * central-difference boundary equation and eliminated from the system:
 * A' = A - B D^-1 C (rows/cols/vals, row-major) and the
 * right-hand-side map M = -B D^-1 (mrows/mcols/mvals, row-major, column =
 * index into spec->mixed).  A solve then uses b = source, b[r] +=
 * sum_m M(r,m) value(m) (set_source_addback), b[fixed] = value.  Mixed
 * nodes at fixed elements are ignored (fixed covers mixed).  NULL output
 * pointers return the counts only. */
psum_status psum_poisson_assemble_ex(const psum_grid* g, const psum_poisson_spec* spec, int64_t cap,
                                     unsigned long long* rows, unsigned long long* cols, double* vals,
                                     int64_t* nnz, int64_t mcap, unsigned long long* mrows,
                                     unsigned long long* mcols, double* mvals, int64_t* mnnz);

/* ---- batched ensemble (no reference counterpart: the reference steps one
 * problem per process).  Independent instances, each configured like a
 * stepper, advance together: one step of all of them is captured into one
 * hipGraph and replayed.  Inside the graph each solver runs a fixed V-cycle
 * count; sorts run between replays at per-instance intervals, on 2^shift-cell
 * bins.  psum_batch_tune is the autotuner: per instance, `warmup`
 * tolerance-controlled steps calibrate the cycle count (most needed after
 * the first, cold-started step, + 1: replayed solves start from the
 * previous step's potential),
 * then each bin shift 0..2 gets a timed sort and `probe` timed steps, and the
 * (shift, interval) minimising the modelled sort + push cost is chosen.
 * psum_batch_configure sets the same three values by hand.
 * psum_batch_instance returns the instance's stepper handle for
 * psum_stepper_state / _phi / _stats (do not destroy it). */
typedef struct {
    int fixed_cycles;
    int sort_every;
    int sort_shift;
    int pad_;
    double sort_ms;            /* measured sort time at the chosen shift */
    double push_ms;            /* push time of the first step after that sort */
    double push_sort_cost_ms;  /* modelled (sort + push) cost per step */
} psum_batch_tuning;

/* The autotuner of psum_batch_tune for one stepper: returns the calibrated
 * cycle count and the chosen sort interval / bin shift.  Each species gets
 * its own interval (timed sort and push per species); sort_every returns
 * the shortest.  The bin shift and the per-species intervals are applied
 * (psum_stepper_species_sort reads them); apply the cycle count with
 * psum_stepper_set_schedule.  The state advances by the tuning steps. */
psum_status psum_stepper_tune(void* h, int warmup, int probe, psum_batch_tuning* out);

void* psum_batch_create(const psum_step_config* cfgs, int n);
psum_status psum_batch_load(void* b, int inst, int s, const void* const* pos, const void* const* vel,
                            int64_t n);
psum_status psum_batch_configure(void* b, int inst, int fixed_cycles, int sort_every, int sort_shift);
psum_status psum_batch_tune(void* b, int warmup, int probe, psum_batch_tuning* out);
/* history rows (5 doubles per step, as psum_stepper_run) restart at 0 and
 * are kept for up to `rows` steps */
psum_status psum_batch_reset_history(void* b, int64_t rows);
/* Split the instances into `lanes` (1..64) groups balanced by particle
 * count; each group's step is its own graph, replayed on its own stream, so
 * groups run concurrently.  Takes effect at the next psum_batch_run (the
 * graphs are captured again). */
psum_status psum_batch_set_lanes(void* b, int lanes);
/* Solver batching (default on): with one lane and every instance's solve a
 * 2D fixed-cycle solve through the fused tile kernels, the captured step
 * runs the solver cycles of all instances as batched launches, one per
 * phase (cycle, level, kernel variant) over the blocks of every instance,
 * instead of one launch per instance and phase.  Results are bit-identical
 * either way.  No reference counterpart. */
psum_status psum_batch_set_solver_batching(void* b, int on);
/* batched solver launches per step, and the per-instance launches they
 * replace (0, 0 when the step does not batch its solves) */
psum_status psum_batch_solver_launches(void* b, int64_t* launches, int64_t* ops);
/* `steps` replays; *t_ms = HIP-event time of the whole call */
psum_status psum_batch_run(void* b, int steps, double* t_ms);
psum_status psum_batch_history(void* b, int inst, double* out, int64_t max_rows, int64_t* rows);
void* psum_batch_instance(void* b, int inst);
void psum_batch_destroy(void* b);

/* ---- adjoint kernels (used by psum_stepper_adjoint) --------------------
 * dtype PSUM_F64 / PSUM_F32: positions, lx, lv and phi / lrho of that type;
 * PSUM_F32_ADJ64: FP32 positions, FP64 lx, lv, phi / lrho and arithmetic
 * (the FP32 stepper's reverse sweep).
 * Reverse of gather + kick + drift for one species, in place on (lx, lv):
 *   lvt = lv + dt lx;  lx -= qm dt H(x)^T lvt;  lv = lvt;
 *   acc[idx_k] += llrint((-qm dt sum_e lvt_e d_e c_k(x)) * scale_dev[0])
 * (the transpose of the gather, deposited in int64 fixed point). */
psum_status psum_hip_adjoint_push(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                  void* const* lv, int64_t n, const void* phi, double qm, double dt,
                                  const double* scale_dev, int64_t* acc, void* stream);
int64_t psum_hip_adjoint_blocks(int64_t n);
/* Reverse of the deposit: lx += w_dep grad(lrho)(x) (the transpose of the
 * deposit); bmax[b] = per-block max of sum_e |lv_e + dt lx_e|. */
psum_status psum_hip_adjoint_gather(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                    void* const* lv, int64_t n, const void* lrho, double w_dep,
                                    double dt, double* bmax, void* stream);
/* scale_dev[0..2] = 2^(40-e), 2^(e-40), bound with 2^(e-1) <= bound < 2^e,
 * bound = factor * max(bmax) (max with the previous bound when accumulate). */
psum_status psum_hip_adjoint_scale(const double* bmax, int64_t nb, double factor, double* scale_dev,
                                   int accumulate, void* stream);
psum_status psum_hip_fixed_to_double_dev(const int64_t* acc, int64_t nn, const double* scale_dev,
                                         double* out, void* stream);
/* dW/dphi of the field energy: eps dV sum over existing edges of
 * delta_r_d^2 (phi_i - phi_j). */
psum_status psum_hip_field_energy_grad(const psum_grid* g, const double* phi, double eps, double* out,
                                       void* stream);
/* x_p = xi_p + sum_{k=1..nm} a_k L/(2 pi k) sin(2 pi k (xi_p - lo)/L), wrapped
 * into [lo, lo+L) (a_dev: nm doubles on the device), and its transpose
 * out_k = sum_p lx_p L/(2 pi k) sin(2 pi k (xi_p - lo)/L). */
psum_status psum_hip_mode_displace(int dtype, const void* xi, void* x, int64_t n, const double* a_dev,
                                   int nm, double lo, double L, void* stream);
size_t psum_hip_mode_project_workspace(int64_t n, int nm);
psum_status psum_hip_mode_project(int dtype, const void* xi, const void* lx, int64_t n, int nm,
                                  double lo, double L, double* out_dev, void* work, void* stream);

#ifdef __cplusplus
}
#endif

#endif
