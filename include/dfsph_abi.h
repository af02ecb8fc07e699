/* C-ABI of the MI355X DFSPH liquid step (libdfsph_hip.so).
 *
 * The entry points replace the Warp launches of alloy/solvers/sph that SolverDFSPH.step issues
 * (alloy/solvers/sph/solver.py:218-268). Python binds them with ctypes (alloy_amd/sph/_abi.py), the
 * binding a host that loads the library directly would copy.
 *
 * Conventions
 *  - Status 0 is success. A negative status comes with a thread-local message (dfsph_last_error):
 *      DFSPH_EINVAL    -> ValueError     (bad argument, bad layout, dt <= 0)
 *      DFSPH_ECAPACITY -> CapacityError  (a size beyond what the handle was created for)
 *      DFSPH_EDEVICE   -> RuntimeError   (a HIP call failed)
 *  - Particle, world and state pointers are device pointers (torch data_ptr() values). Shape tables and
 *    the world layout are host arrays copied at create time. The library never frees caller memory.
 *  - Every launch goes to the stream passed in (NULL = the default stream). The only host syncs are
 *    dfsph_report, dfsph_read_buffer/dfsph_write_buffer and, per iteration of a pressure solve, the
 *    per-world convergence count (the device-side loop replacing wp.capture_while).
 *  - A handle is thread-compatible, not thread-safe.
 */
#ifndef DFSPH_ABI_H
#define DFSPH_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DFSPH_ABI_VERSION 5

#define DFSPH_OK 0
#define DFSPH_EINVAL (-1)
#define DFSPH_ECAPACITY (-2)
#define DFSPH_EDEVICE (-3)

/* Solver parameters. Mirrors DFSPHConfig (alloy/solvers/sph/config.py:46-131) plus the constants the
 * reference derives from it at construction (solver.py:92-94, layout.py:195-249, solve_driver.py:91).
 * Host doubles; each kernel parameter is rounded to float32 exactly where the reference passes a Python
 * float to a float32 Warp kernel argument. */
typedef struct DfsphParams {
    double particle_radius;          /* r [m] */
    double h;                        /* support radius = support_factor * r */
    double spacing;                  /* 2 r */
    int32_t kernel;                  /* 0 cubic, 1 wendland_c2 (kernels.py:27-29) */
    int32_t pressure_solver;         /* 0 projected CG (krylov.py), 1 Jacobi (solve_driver.py:171-186) */
    double max_density_error;
    double max_divergence_error;
    double min_density_error;
    double max_error_factor;
    double rest_jitter_speed;        /* <= 0: None (rule off, solve_driver.py:108-119) */
    double rest_jitter_flow_ratio;
    int32_t min_iterations;
    int32_t max_iterations;
    double min_neighbors_divergence;
    int32_t divergence_solve;
    int32_t divergence_two_sided;
    double jacobi_omega;
    double warm_start_density;
    double warm_start_divergence;
    double xsph_c;                   /* 0: no XSPH */
    int32_t neighbor_capacity;       /* 8..64 */
    int32_t wall_friction;
    double n_full;                   /* full_neighbor_count (config.py:230-236) */
    double tot_min;                  /* diagonal_floor * lattice_diagonal (solver.py:94) */
    double lattice_grad;             /* sqrt(lattice_diagonal) (solve_driver.py:91) */
    double compression_relief_speed; /* <= 0: None */
    double contact_separation_speed;
    double wall_offset;              /* boundary_fit delta (layout.py:243) */
    double wall_amplitude;           /* boundary_fit A */
    double taper0;                   /* r (layout.py:244) */
    double taper1;                   /* 3 r */
    double reach;                    /* 1.5 spacing (layout.py:195-199) */
    int32_t curvature;               /* boundary_curvature_order (0 or 1) */
    double project_eps;              /* PROJECT_EPS * r (solver.py:483) */
    double sweep_len;                /* integrate.sweep_length(h, thinnest) */
    int32_t one_way;                 /* coupling == "one_way": no shape meters, no particle-wall reactions
                                        (bodies.py:58-59, particle_walls.py:156-157) */
    double skin;                     /* Verlet skin [m] = reneighbor_skin x h, 0 off (neighbors_skin.py); the cell
                                        grid's side is then h + skin */
    int32_t skin_rows;               /* superset row capacity (skin_capacity, neighbors_skin.py:149-158), at most 64 */
    int32_t viscosity_implicit;      /* 1: Weiler 2018 implicit viscosity after the non-pressure update (viscosity.py) */
    double visc_d;                   /* LAPLACE_D / laplacian_lattice_factor (solver.py:157-160) */
    double visc_tol;                 /* viscosity_tolerance: stop at |r| <= tol |v*| per world */
    int32_t visc_max_iterations;
    double boundary_viscosity;       /* mu_b [Pa s] of the wall term, 0 = free-slip walls (config.py:98,
                                        wall_viscosity.py); needs viscosity_implicit */
    double visc_wall_coeff;          /* mu_b x wall_coefficient_scale (solver.py:160-162) */
    int32_t csf;                     /* 1: surface_tension "csf" (csf.py), a stage of the non-pressure update */
    double csf_c_surface;            /* c_s of flat_lattice_constants (csf.py:252-278) */
    double csf_inv_i;                /* 1 / I of flat_lattice_constants */
    double csf_band_lo, csf_band_hi; /* CSF_BAND (csf_sparse.py:40) */
    double csf_k_cohesion;           /* 32 / (pi h^9), Akinci's cohesion spline factor (tension.py:81-92) */
} DfsphParams;

/* Static description of the particle pack and its walls (layout.py:47-72, boundary.py:79-95). */
typedef struct DfsphLayout {
    int32_t n;                       /* particles in the pack (liquid, spares, wall particles, others), < 2^28 */
    int32_t world_count;             /* W */
    const int32_t* world_start;      /* host [W+1], contiguous world ranges */
    /* per-world cell grid of the neighbour search (cell side = h * (1 + 1e-5)), host arrays */
    const float* grid_origin;        /* [W*3] */
    const int32_t* grid_dims;        /* [W*3], each >= 1 */
    /* shapes (boundary.py:134-157): S model shapes; per world CSR start[W+1] / ids */
    int32_t shape_count;
    const int32_t* shape_start;      /* host [W+1] */
    const int32_t* shape_ids;        /* host [shape_start[W]] */
    const int32_t* shape_geo;        /* host [S] GeoType: 1 plane, 3 sphere, 7 box */
    const float* shape_scale;        /* host [S*3] */
    const float* shape_xform;        /* host [S*7] p(3) q(4, xyzw) */
    const int32_t* shape_body;       /* host [S], -1 static (only static shapes are accepted) */
    const float* shape_radius;       /* host [S] origin radius (layout.py:151-163) */
    const float* shape_mu;           /* host [S] Coulomb friction of the wall projection */
    int32_t wall_particle_count;     /* particles flagged ACTIVE|PROXY acting as particle walls */
    const float* shape_cos;          /* host [S] cos of sph:shape_contact_angle (csf.py:281-296); NULL: all free */
    const float* shape_sin;          /* host [S] its sin, -1 for a free contact line */
} DfsphLayout;

/* Model arrays (device pointers, length n unless stated). */
typedef struct DfsphModel {
    const float* mass;               /* particle_mass */
    const float* rest_density;       /* sph:rest_density */
    int32_t* flags;                  /* particle_flags (ACTIVE = 1, PROXY = 2) */
    int8_t* phase;                   /* sph:phase (1 liquid, 0 spare, -1 none, -2 retired) */
    const int32_t* world;            /* particle_world */
    const float* gravity;            /* [W*3] */
    const float* wall_volume;        /* V_j = 8 r^3 of particle-wall particles, 0 otherwise (particle_walls.py:227) */
    const float* viscosity;          /* sph:viscosity [Pa s] (read by the implicit viscosity; may be NULL without it) */
    const float* surface_tension;    /* sph:surface_tension sigma [N/m] (read by csf; may be NULL without it) */
    const float* csf_gamma;          /* the small-piece pair force's gamma per particle (csf_sparse.py:91-99) */
} DfsphModel;

/* Host copy of the per-world report (types.py:11-28 SPHReport); arrays of length W owned by the caller. */
typedef struct DfsphReport {
    float* density_err;
    float* divergence_err;
    int32_t* iters_density;
    int32_t* iters_divergence;
    int32_t* neighbor_overflow;
    float* max_speed;
    float* density_err_max;
    float* density_tol;
    int32_t* projected;
    int32_t* swept;
    int32_t* nonfinite;
    double* n_active;
} DfsphReport;

/* Buffer ids for dfsph_get_buffer / dfsph_read_buffer / dfsph_write_buffer. */
enum DfsphBuffer {
    DFSPH_BUF_NBR = 0,        /* uint16 [cap, n] world-local offsets, column-major (neighbors.py:13-18) */
    DFSPH_BUF_NBR_HI,         /* uint8  [cap, n] high byte when a world has > 65,536 particles, else [1] */
    DFSPH_BUF_COUNT,          /* int32  [n] */
    DFSPH_BUF_GROUP,          /* int32  [n] world, -2 - world (particle wall), -1 (neighbors.py:36-51) */
    DFSPH_BUF_SLOT_SHAPE,     /* int32  [n, 4] */
    DFSPH_BUF_SLOT_PSI,       /* f32    [n, 4] */
    DFSPH_BUF_SLOT_GPSI,      /* f32    [n, 4, 3] */
    DFSPH_BUF_SLOT_XB,        /* f32    [n, 4, 3] */
    DFSPH_BUF_SLOT_VB,        /* f32    [n, 4, 3] */
    DFSPH_BUF_SLOT_PHI,       /* f32    [n, 4] */
    DFSPH_BUF_SLOT_XBL,       /* f32    [n, 4, 3] */
    DFSPH_BUF_SLOT_NBL,       /* f32    [n, 4, 3] */
    DFSPH_BUF_SLOT_IMPULSE,   /* f32    [n, 4, 3] */
    DFSPH_BUF_RATIO,          /* f32    [n] */
    DFSPH_BUF_ALPHA,
    DFSPH_BUF_NEFF,
    DFSPH_BUF_ACTIVE,
    DFSPH_BUF_REG,
    DFSPH_BUF_SRC,
    DFSPH_BUF_KAPPA,
    DFSPH_BUF_ACC,            /* f32    [n, 3] */
    DFSPH_BUF_ERR,
    DFSPH_BUF_LOWER,
    DFSPH_BUF_KV,             /* warm start k dt (divergence) */
    DFSPH_BUF_KD,             /* warm start k dt^2 (density) */
    DFSPH_BUF_VTMP,           /* f32    [n, 3] the liquid's velocities after the last non-pressure update (the rows
                                 of other particles are not written) */
    DFSPH_BUF_CONVERGED,      /* int32  [W] */
    DFSPH_BUF_ITERS,          /* int32  [W] */
    DFSPH_BUF_ERR_MEAN,       /* f32    [W] */
    DFSPH_BUF_ERR_MAX,        /* f32    [W] */
    DFSPH_BUF_N_ACTIVE,       /* f64    [W] */
    DFSPH_BUF_TOL_SCALE,      /* f32    [W] */
    DFSPH_BUF_OVERFLOW,       /* int32  [W] */
    DFSPH_BUF_MAX_SPEED,      /* f32    [W] */
    DFSPH_BUF_ITERS_DEN,      /* int32  [W] */
    DFSPH_BUF_ITERS_DIV,      /* int32  [W] */
    DFSPH_BUF_DEN_ERR,        /* f32    [W] */
    DFSPH_BUF_DIV_ERR,        /* f32    [W] */
    DFSPH_BUF_SHAPE_METER,    /* int64  [S, 3] impulse per shape, 2^50 fixed point (bodies.py:21-25) */
    DFSPH_BUF_FEATURES,       /* f32    [n, 16] learned warm-start features (solve inputs) */
    /* mass ledger and emitter state (runtime.py:93-107, emitters.py:200-216); masses are int64 at 2^40 per kg
     * (kernels.py:142-144), so active + out == initial + in holds exactly */
    DFSPH_BUF_MASS_INITIAL,   /* int64  [W] active liquid mass at construction (init_mass) */
    DFSPH_BUF_MASS_IN,        /* int64  [W] emitted */
    DFSPH_BUF_MASS_OUT,       /* int64  [W] retired by sinks, the escape box or dfsph_retire */
    DFSPH_BUF_MASS_ACTIVE,    /* int64  [W] as of the last dfsph_active_mass */
    DFSPH_BUF_ESCAPED,        /* int32  [W] particles retired by the escape box */
    DFSPH_BUF_EMIT_OVERFLOW,  /* int32  [W] layer sites an exhausted pool could not fill */
    DFSPH_BUF_FLUX_E,         /* int64  [W] energy brought in minus taken out, m (v^2/2 - g.x), 2^50 per J */
    DFSPH_BUF_FLUX_G,         /* int64  [W] its slope m g_z for the monitor's z_ref, 2^40 per N */
    DFSPH_BUF_EMIT_T,         /* f32    [E] layer clock of each emitter */
    DFSPH_BUF_EMIT_CURSOR,    /* int32  [E] spares each emitter has used */
    DFSPH_BUF_EMIT_ENABLED,   /* int32  [E] */
    DFSPH_BUF_FLAGS,          /* int32  [n] the model's particle_flags (the emitter pass writes them) */
    DFSPH_BUF_PHASE,          /* int8   [n] the model's sph:phase */
    DFSPH_BUF_PW_EXCH,        /* int64  [NW, 3] particle-wall impulse of the exchange window, 2^50 per N s, rows in
                                 ascending id order of the wall particles (particle_walls.py:138-177) */
    DFSPH_BUF_PW_WORK,        /* int64  [W] the particle walls' work on the liquid in the last step, 2^50 per J */
    DFSPH_BUF_SKIN_FLAG,      /* int32  [W] worlds whose Verlet-skin rows were rebuilt in the last list build */
    DFSPH_BUF_VISC_ITERS,     /* int32  [W] iterations of the last implicit viscosity solve */
    DFSPH_BUF_CSF_SURF,       /* int32  [n] csf: 1 on the surface particles of the last step */
    DFSPH_BUF_CSF_KAPPA,      /* float  [n] csf: the curvature applied at the interface, 0 off the surface */
    DFSPH_BUF_PW_ROW_COUNT,   /* int32  [n] with particle walls: the wall particles the last slot pass's colour sum of
                                 each liquid particle visited, -1 when more than its 32-entry row holds (the reaction
                                 split then walks the cells again) */
    DFSPH_BUF_HALO_COUNT,     /* int32  [chunks] with DFSPH_OPT_HALO: distinct neighbours of each 256-particle chunk at
                                 the last list build, -1 where the halo did not fit (direct gathers) */
    DFSPH_BUF_COUNT_
};

/* Box emitters, sinks and the escape box (emitters.py:32-255, runtime.py:197-304). Host arrays, copied by
 * dfsph_set_emitters; E emitters, K sinks. An emitter emits a layer (its offsets, through its region transform,
 * with its velocity) every `interval` seconds of simulated time, activating the next spares of its pool in order. */
typedef struct DfsphEmitters {
    int32_t emitter_count;           /* E */
    const int32_t* world;            /* [E] */
    const float* xform;              /* [E*7] region p(3) q(4, xyzw) */
    const float* velocity;           /* [E*3] */
    const float* interval;           /* [E] n_layer V / rate [s] (runtime.py:266) */
    const int32_t* offset_start;     /* [E+1] CSR into offsets */
    const float* offsets;            /* [offset_start[E]*3] layer lattice in the region frame (emitters.py:46-64) */
    const int32_t* pool_start;       /* [E+1] CSR into pool_ids */
    const int32_t* pool_ids;         /* spare particle ids of each emitter, in emission order */
    int32_t sink_count;              /* K */
    const int32_t* sink_world;       /* [K] */
    const float* sink_xform;         /* [K*7] */
    const float* sink_half;          /* [K*3] half extents */
    int32_t has_bounds;              /* escape box (DFSPHConfig.world_bounds) */
    float bounds_lo[3];
    float bounds_hi[3];
} DfsphEmitters;

/* Library */
int32_t dfsph_abi_version(void);
const char* dfsph_last_error(void);

/* Construction: SolverDFSPH.__init__ / _alloc (solver.py:72-168). Allocates every work buffer once. */
int32_t dfsph_create(const DfsphParams* params, const DfsphLayout* layout, const DfsphModel* model,
                     void** out_handle);
int32_t dfsph_destroy(void* handle);

/* One DFSPH step on the state (q, qd [n*3] f32, in place; f [n*3] or NULL): SolverDFSPH.step
 * (solver.py:218-268). dt in seconds (> 0). */
int32_t dfsph_step(void* handle, float* q, float* qd, const float* f, double dt, void* stream);

/* Phase entry points (the same buffers as dfsph_step) for parity tests and callers that interleave work. */
int32_t dfsph_build_neighbors(void* handle, const float* q, void* stream);              /* neighbors.py:256-283 */
int32_t dfsph_boundary_slots(void* handle, const float* q, const float* qd, void* stream); /* boundary.py:399-414 */
int32_t dfsph_density_alpha(void* handle, const float* q, void* stream);                 /* pressure.py:144-193 + reduce */
/* mode 0 divergence, 1 density (solve_driver.py:121-162). fixed_iters >= 0 forces that iteration count
 * (teacher-forced parity); -1 iterates to convergence. */
int32_t dfsph_solve(void* handle, int32_t mode, float* q, float* qd, double dt, int32_t fixed_iters, void* stream);
int32_t dfsph_nonpressure(void* handle, const float* q, float* qd, const float* f, double dt, void* stream); /* integrate.py:150-178 */
int32_t dfsph_advect_project(void* handle, float* q, float* qd, double dt, void* stream);  /* integrate.py:181-309 + speed */
/* Mean-velocity sums of the resting-jitter rule at the step's start state (solve_driver.py:100-106). */
int32_t dfsph_flow_speed(void* handle, const float* qd, void* stream);
/* Hydrostatic prime (prime.py:53-115): guess [n] k = p/rho0 (host-computed, device pointer). Returns
 * iterations used through *iters. */
int32_t dfsph_prime(void* handle, float* q, float* qd, const float* guess, double dt, double tol,
                    int32_t max_iterations, int32_t* iters, void* stream);

/* Learned warm start: an MLP (13 -> 64 -> 64 -> 64 -> 2, SiLU, bf16 on the MFMA units) predicting, per particle,
 * the correction y to solve_init's warm-start rule (pressure.py:254, 261): kappa0 = ks sinh(asinh(rule / ks) + y),
 * ks = |g| h, within the rule's clamps. weights: the network's float32 parameters on the device (the layout of
 * dfsph_mlp_param_count); NULL switches the learned start off. */
int32_t dfsph_set_warm_start_model(void* handle, const void* weights, void* stream);
/* Learned warm start given as initial kappa fields [n] (device; NULL = the reference warm-start rule) of the
 * divergence and density solves, read by solve_init (pressure.py:254, 261) in place of warm x stored kappa; the
 * solve still iterates to the same tolerance. */
int32_t dfsph_set_warm_start_kappa(void* handle, const float* kappa_div, const float* kappa_den);
/* Host convergence check every `every` iterations (1..64); converged worlds' blocks exit at once in between. */
int32_t dfsph_set_check_interval(void* handle, int32_t every);
/* Tuning options. Every option except 4 gives bit-identical results and only trades time; option 4 chooses which
 * pressure solves start from the learned warm start (both still iterate to the same tolerance). */
enum DfsphOption {
    DFSPH_OPT_CHECK_INTERVAL = 1, /* as dfsph_set_check_interval */
    DFSPH_OPT_GRAPHS = 2,         /* as dfsph_set_graphs */
    DFSPH_OPT_PAIR_CACHE = 3,     /* mask of DFSPH_PAIR_CACHE_*: gather passes that read the density pass's stored
                                     kernel-gradient scale of each listed pair instead of re-evaluating it (4 B of
                                     memory traffic per pair traded for a square root and two divisions) */
    DFSPH_OPT_LEARNED_SOLVES = 4, /* with a warm-start model set: bit 0 the divergence solve, bit 1 the density solve
                                     start from its prediction (default 3); the others keep the reference's rule */
    DFSPH_OPT_SPARSE_CORRECTION = 5, /* 1 (default): the CG iteration's correction product (krylov.py:238, A corr
                                        on clamped particles) evaluates only particles with a nonzero term and writes
                                        the others' exact zero results; 0 evaluates every particle. Bit-identical. */
    DFSPH_OPT_BUILD_BLOCK = 6,       /* threads per block of the neighbour-list build: 64, 128 (default) or 256
                                        (256 only without 24-bit rows). Bit-identical. */
    DFSPH_OPT_PACKED_ACCEL = 7,      /* 1 (default when every world's SPH particles share one mass): the pressure
                                        acceleration gathers one 16 B record {x, y, z, kappa / rho0} per neighbour
                                        from a packed copy, the mass the world's; 0: positions and quotients apart.
                                        Bit-identical. */
    DFSPH_OPT_HALO = 8,              /* 1: each list build also lists every 256-particle chunk's distinct neighbours
                                        (its halo, up to 1,536), and the pressure acceleration and operator apply
                                        stage the halo's records in LDS and gather from there; chunks whose halo does
                                        not fit gather directly. 0: every chunk gathers directly. Bit-identical. */
    DFSPH_OPT_DEVICE_LOOP = 9,       /* 1: the projected-CG iterations of each solve run as one cooperative launch
                                        that loops on the device until every world's stop test passed (grid barriers
                                        between the phases, direct gathers), so a step makes no host round trip;
                                        0: iteration batches with a host check of the unconverged-world count.
                                        Jacobi and profiled steps use the batches. Bit-identical. */
    DFSPH_OPT_BUILD_STAGED = 10,     /* 1: the plain neighbour-list build stages each block's box of grid cells in
                                        LDS (64 or 128 threads by DFSPH_OPT_BUILD_BLOCK, large boxes in z-slabs) and
                                        every lane walks its 27 cells there; 0 (default): every lane walks its cells
                                        from global memory. Bit-identical. */
    DFSPH_OPT_LEARNED_GATE = 11      /* with the warm-start network set (dfsph_set_warm_start_model): 1 (default): a
                                        world takes the network's start at a step when the last step's solve that
                                        the network feeds needed more than min_iterations iterations, and only such
                                        worlds pay for the features and the network; the others (and every world at
                                        its first step) start from the reference's rule. 0: every world takes the
                                        network's start. Like option 4 it chooses starts, so results differ in bits;
                                        every solve still iterates to the same tolerance. */
};
enum DfsphPairCache {
    DFSPH_PAIR_CACHE_ACCEL = 1, /* pressure acceleration (k_accel, post-solve) */
    DFSPH_PAIR_CACHE_APPLY = 2, /* operator apply (divergence of the acceleration) */
    DFSPH_PAIR_CACHE_OTHER = 4, /* solve source, learned-start features */
    DFSPH_PAIR_CACHE_ALL = 7
};
/* No counterpart in the reference (Warp picks launch parameters itself); used by the autotuner. */
int32_t dfsph_set_option(void* handle, int32_t option, int32_t value);
int32_t dfsph_get_option(void* handle, int32_t option, int32_t* value);
/* dfsph_step replays its segments (pre-solve, solve head, iteration batch, solve tail) as hipGraphs captured once
 * per (dt, state pointers, batch size); 0 launches every kernel eagerly. Results are identical either way. */
int32_t dfsph_set_graphs(void* handle, int32_t on);

/* Per-launch timing of the kernels selected by `mask` (bit k = kernel id k: 0 pressure acceleration, 1 operator
 * apply, 2 neighbour rows, 3 density/alpha, 4 boundary slots, 5 non-pressure, 6 advect/project, 7 solve init,
 * 8 post-solve; 0 stops), of a pseudo-random 1 / `sample` of their launches (a timed launch costs the host more).
 * The events are part of each timed kernel's dispatch and skip the system-scope fence, so durations agree with a
 * profiler's kernel trace; timed k_accel and k_apply launches also count the algorithmic bytes of the particles
 * they process. No counterpart in the reference (its timing is wall clock around graph replays,
 * alloy/testing/sph.py:334-364); for measurement only. Enabling resets the totals. */
int32_t dfsph_set_profiling(void* handle, int32_t mask, int32_t sample);
/* launches, summed duration [ms] and algorithmic bytes (may be NULL) of kernel id `kernel_id` (synchronizes) */
int32_t dfsph_profile_read(void* handle, int32_t kernel_id, int32_t* count, double* total_ms, double* alg_bytes);

/* Learned warm start network (13 -> 64 -> 64 -> 64 -> 2, SiLU; float32 master parameters, bf16 on the MFMA units).
 * No counterpart in the reference: these train and evaluate the model that dfsph_set_warm_start_model installs.
 * All pointers are device pointers; X [n][16] float32 (13 features used), Y/T [n][2]. */
int32_t dfsph_mlp_param_count(void);
int32_t dfsph_mlp_forward(const float* params, const float* X, int32_t n, float* Y, void* stream);
/* Floats of scratch dfsph_mlp_grad needs for n samples: one (P + 1)-float partial per 16-sample wave of each
 * 64-sample block, 4 * ceil(n / 64) * (P + 1). */
int64_t dfsph_mlp_grad_scratch_floats(int32_t n);
/* gradient of sum((Y - T)^2) / (2 n_total) into grad [P] and the squared-error sum into grad[P];
 * scratch: scratch_floats >= dfsph_mlp_grad_scratch_floats(n) floats of per-wave partials, reduced in a fixed
 * order (deterministic); a smaller scratch is rejected with DFSPH_EINVAL. */
int32_t dfsph_mlp_grad(const float* params, const float* X, const float* T, int32_t n, int32_t n_total,
                       float* scratch, int64_t scratch_floats, float* grad, void* stream);
/* No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight training loops, harnesses that forbid hipMalloc): dfsph_mlp_grad with its per-wave partials in a static
 * device array of the library (about 10.7 MB per GPU, held until the process ends) instead of a caller workspace.
 * n at most DFSPH_MLP_NOALLOC_MAX_SAMPLES, else DFSPH_ECAPACITY. It allocates nothing; calls that overlap on one GPU
 * share that array and race. Same results, bit for bit, as dfsph_mlp_grad. */
#define DFSPH_MLP_NOALLOC_MAX_SAMPLES 4096
int32_t dfsph_mlp_grad_noalloc(const float* params, const float* X, const float* T, int32_t n, int32_t n_total,
                               float* grad, void* stream);
int32_t dfsph_mlp_adam(float* params, const float* grad, float* m, float* v, double lr, double beta1,
                       double beta2, double eps, int32_t t, void* stream);
/* Training-data logging of dfsph_step: the step's start-of-step features [n][16] and the converged kappa of its
 * two solves as targets asinh(kappa / (|g| h)) [n][2] (device pointers; NULL stops logging). */
int32_t dfsph_set_logging(void* handle, float* features, float* targets);

/* Buffers */
int32_t dfsph_get_buffer(void* handle, int32_t which, void** ptr, int64_t* n_bytes);
int32_t dfsph_read_buffer(void* handle, int32_t which, void* host_dst, int64_t n_bytes);   /* syncs */
int32_t dfsph_write_buffer(void* handle, int32_t which, const void* host_src, int64_t n_bytes); /* syncs */
int32_t dfsph_zero_stats(void* handle, void* stream);   /* WorldStats.reset counters of every world */

/* Per-world report (runtime.py:136-147): the sync point of a step. */
int32_t dfsph_report(void* handle, DfsphReport* out, void* stream);

/* Emitters, sinks and the escape box (runtime.py:231-304 add_emitter / add_sink / _upload, EmitterDevice.carry_from):
 * replaces the tables; the first min(old, new) emitters keep their clock, cursor and enabled flag. With any emitter,
 * sink or escape box, dfsph_step ends with the emitter pass (runtime.py:318-347): sinks and the escape box retire
 * liquid, then each due emitter emits a layer; the step's speed statistics are taken after it. */
int32_t dfsph_set_emitters(void* handle, const DfsphEmitters* table);
/* SolverDFSPH.set_emitter_enabled (solver.py:438-447); an unknown emitter is DFSPH_EINVAL */
int32_t dfsph_set_emitter_enabled(void* handle, int32_t emitter, int32_t on);
/* the emitter pass alone (after dfsph_advect_project, whose speed statistics it then supplies) */
int32_t dfsph_emitters_step(void* handle, float* q, float* qd, double dt, void* stream);
/* SolverDFSPH.retire (solver.py:428-431, ports.py:66-93): the liquid particles with mask[i] != 0 (device int32 [n])
 * leave for good (ACTIVE cleared, phase -2, velocity 0); their mass and energy leave the ledgers as at a sink */
int32_t dfsph_retire(void* handle, const int32_t* mask, float* q, float* qd, void* stream);
/* active liquid mass per world into DFSPH_BUF_MASS_ACTIVE (runtime.py:126-130) */
int32_t dfsph_active_mass(void* handle, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* DFSPH_ABI_H */
