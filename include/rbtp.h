/*
 * rbtp.h — C-ABI of the MI355X (gfx950) D3Q27 thermal lattice-Boltzmann solver with
 * finite-size thermal spheres.
 *
 * The reference solver (Summer536/RB-ThermoParticles) has no plugin layer: its hot path is a set
 * of free host functions over global state, called from main() in a fixed order
 * (main.cu:32-86).  This header is the drop-in boundary that replaces them.  Every entry point
 * takes plain pointers and sizes, returns 0 on success and a negative RBTP_E* code on failure
 * (the message is kept in rbtp_last_error()), and never calls exit().  Device memory is owned by
 * the rbtp_sim handle; host buffers passed in are owned by the caller.  A handle is not
 * thread-safe; distinct handles are independent and each runs on its own HIP stream.
 *
 * Array layouts follow the reference: a scalar field is N = lx*ly*lz doubles indexed
 * idx = iz*lx*ly + iy*lx + ix; a population field is 27*N doubles indexed ip*N + idx with the
 * D3Q27 ordering of initial.cu:276-279; particle arrays are P doubles indexed by particle id.
 */
#ifndef RBTP_H
#define RBTP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RBTP_ABI_VERSION 1

/* Status codes. */
#define RBTP_OK 0
#define RBTP_E_ARG (-1)      /* invalid argument or case description */
#define RBTP_E_HIP (-2)      /* HIP runtime error (message names the call) */
#define RBTP_E_NOMEM (-3)    /* device or host allocation failed */
#define RBTP_E_IO (-4)       /* file missing, short or malformed */
#define RBTP_E_STATE (-5)    /* call not valid in the current state */
#define RBTP_E_CAPACITY (-6) /* a device-side capacity bound was exceeded */
#define RBTP_E_PLACEMENT (-7)/* particle placement gave up (pinit.cu:126-131) */
#define RBTP_E_DIVERGED (-8) /* the divergence watchdog flagged a node (rbtp_set_watchdog) */

/* Storage/arithmetic forms of the lattice. */
#define RBTP_PREC_FP64 0   /* FP64 storage and arithmetic, operation order of the reference */
#define RBTP_PREC_FP32 1   /* FP32 arithmetic, populations stored as FP32 (shifted, 1/w scaled) */
#define RBTP_PREC_FP16S 2  /* FP32 arithmetic, populations stored as FP16 (shifted, 1/w scaled) */
#define RBTP_PREC_BF16S 3  /* FP32 arithmetic, populations stored as BF16 (shifted, 1/w scaled) */

/*
 * Case description: every macro of parameters.h:1-55 as a runtime field (same names, lower
 * case), plus the extensions at the end.  rbtp_case_default() fills the values parameters.h
 * ships with.  BLOCK_X/Y/Z (parameters.h:51-53) have no counterpart: launch shapes are chosen by
 * the library.
 */
typedef struct rbtp_case {
    /* parameters.h:5 and :8-20 */
    int continuous_steps;
    int lx, ly, lz;
    int nend, ndiag, nflowout, nnuout, nout_p;
    /* parameters.h:23-30 */
    double rayleigh, prandtl, t_hot, t_cold, t_ref, beta, visc, rho0;
    /* parameters.h:33-48 */
    int activate_particles;
    double particle_fraction, particle_diam_frac, cp_particle_ratio, cp_fluid;
    double rho_particle_ratio, repforce_scale, particle_init_temp;
    int particle_rand_seed;
    double rep_min_gap, rep_min_gap_w, rep_stf0, rep_stf1, rep_stf0_w, rep_stf1_w;
    /* Extensions. */
    int npart_override; /* >=0: use this particle count instead of the fraction rule */
    int precision;      /* RBTP_PREC_* */
    int reserved[6];    /* zero */
} rbtp_case;

/* Constants derived from the case (globals.cu:18-33, globals.cu:126-169). */
typedef struct rbtp_derived {
    double grav0, tau, u0, diff, tauc, tauci;
    double particle_radius, particle_mass, particle_inertia, particle_heat_capacity;
    int npart;
    int reserved;
} rbtp_derived;

typedef struct rbtp_sim rbtp_sim;

/* Last error message of the calling thread ("" if none). */
const char *rbtp_last_error(void);
int rbtp_abi_version(void);

/* parameters.h defaults. */
int rbtp_case_default(rbtp_case *c);
/* Read a parameters.h file at run time (parameters.h:1-55): every #define is evaluated with C
 * arithmetic (integer literals stay integers), known macro names fill the case, the rest of the
 * fields keep their defaults.  Two optional extra macros: NPART_OVERRIDE and PRECISION. */
int rbtp_case_from_parameters_h(const char *path, rbtp_case *out);
/* globals.cu:18-33 (grav0, tau, u0, diff, tauc, tauci), globals.cu:126-148 (particle constants)
 * and update_particle_count (globals.cu:150-169). */
int rbtp_case_derive(const rbtp_case *c, rbtp_derived *out);

/* Allocation part of initialize() (initial.cu:240-263) and allocate_particle_arrays
 * (pinit.cu:8-95).  Fails with RBTP_E_ARG for an invalid case. */
int rbtp_create(const rbtp_case *c, rbtp_sim **out);
void rbtp_destroy(rbtp_sim *s);
int rbtp_get_case(const rbtp_sim *s, rbtp_case *c, rbtp_derived *d);

/*
 * Initial state: init_f/init_g (initial.cu:335-397, called at initial.cu:318-327), particle
 * placement (pinit.cu:97-144, glibc srand/rand) and the first build_links (initial.cu:330).
 *   temperature: N values of the initial temperature field phi, or NULL for the reference's
 *                uniform t_ref.  g is set from phi by init_g's formula with rho = rho0, u = 0.
 *   ppos:        3*P values (x[P], y[P], z[P]) or NULL for the reference's rejection placement.
 *   ptemp:       P initial particle temperatures, or NULL for particle_init_temp.
 */
int rbtp_init(rbtp_sim *s, const double *temperature, const double *ppos, const double *ptemp);

/* Host helper: the initial temperature used by the benchmark cases, T(ix) linear from t_hot at
 * x=0 to t_cold at x=lx at cell centres, plus amp * U[-1,1) from splitmix64(seed) in idx order. */
int rbtp_linear_temperature(const rbtp_case *c, uint64_t seed, double amp, double *out_n);

/* generate_initial_positions (pinit.cu:97-144) alone, on the host (no device is touched): the
 * rejection placement with glibc srand(particle_rand_seed)/rand that rbtp_init runs when ppos is
 * NULL.  out_3p receives x[P], y[P], z[P] for P = rbtp_case_derive's npart.  Returns
 * RBTP_E_PLACEMENT where the reference prints "Can't distribute the particles" and exits
 * (pinit.cu:126-131), after 10^6 rejections. */
int rbtp_place_particles(const rbtp_case *c, double *out_3p);

/* The particle part of load_particles_from_restart (initial.cu:57-192) alone, on the host: read
 * np lines of a pos_%09d.txt file (pid followed by the 33 values of output_part.cu:101-113) into
 * out (RBTP_P_COUNT arrays of np doubles, indexed by RBTP_P_*; arrays the file does not hold are
 * zero).  A missing file, a short or unparsable line, or a pid outside [0, np) is RBTP_E_IO (the
 * reference exits, initial.cu:63-66,99-101, or writes out of bounds). */
int rbtp_read_particle_file(const char *path, int np, double *out);

/*
 * Advance nsteps time steps of main.cu:32-86 (macrovar -> forcing -> collision_BGK -> streaming
 * -> bounce_back_particles -> collision_BGK_scalar -> streaming_scalar ->
 * bounce_back_thermal_particles -> prepare_particle_forces_step -> apply_repulsive_forces ->
 * compute_particle_forces -> compute_particle_heat -> update_particles -> build_links ->
 * refill_nodes).  Work is enqueued on the handle's stream; the call returns after the stream
 * has drained and device-side error flags were checked.
 *   flags: RBTP_STEP_MACRO_LAST makes the last step leave rho/ux/uy/uz/phi as macrovar (and
 *          refill) left them, as the reference's output functions read them.
 */
#define RBTP_STEP_MACRO_LAST 1u
int rbtp_step(rbtp_sim *s, int nsteps, unsigned flags);
/* Same, without waiting for the stream (errors are reported by the next synchronising call). */
int rbtp_step_async(rbtp_sim *s, int nsteps, unsigned flags);
int rbtp_synchronize(rbtp_sim *s);
/* Optional divergence watchdog; no reference counterpart (the reference checks neither NaN nor
 * velocity).  With umax > 0 every lattice step flags a node whose density or velocity is not
 * finite or whose |u| exceeds umax, at no cost to the values computed; rbtp_step (and
 * rbtp_synchronize after rbtp_step_async) then returns RBTP_E_DIVERGED, naming the first
 * flagged step and the number of flags, and clears them.  The state is left as the steps made
 * it.  umax = 0 (the default) turns the check off; batch slots are not watched. */
int rbtp_set_watchdog(rbtp_sim *s, double umax);
/* Number of completed time steps since rbtp_init (restart_step counts). */
int rbtp_current_step(const rbtp_sim *s, int *step);

/*
 * One stage of the time step at a time, as main.cu:35-84 launches them (declared at
 * include/lbm.h:59-78 and include/particle.h:110-121), for callers that drive the stages
 * themselves and for localising a divergence to one stage.  FP64 form only (RBTP_E_STATE for the
 * reduced-precision forms and for serving-batch slots).  The stages must be called in main.cu
 * order, starting at a step boundary (RBTP_E_STATE otherwise); each runs as its own kernels on
 * the reference's arrays (d_f, d_f_collide, d_g, d_g_collide, rho/u/phi, F_x, ibnode/owner and
 * their previous generation, the particle arrays), which the accessors then return.  After
 * RBTP_STAGE_REFILL the step counter advances.  rbtp_step continues from the staged state once
 * the staged step is complete.  The particle stages do nothing without particles.
 */
#define RBTP_STAGE_MACROVAR 0     /* macrovar (macrovar.cu:3-34) */
#define RBTP_STAGE_FORCING 1      /* forcing (macrovar.cu:36-47) */
#define RBTP_STAGE_COLLIDE_F 2    /* collision_BGK (collision.cu:3-33) */
#define RBTP_STAGE_STREAM_F 3     /* d_f_collide <- d_f, streaming (streaming.cu:3-23), main.cu:45-48 */
#define RBTP_STAGE_IBB_F 4        /* bounce_back_particles (ibb.cu:5-82, 209-221) */
#define RBTP_STAGE_COLLIDE_G 5    /* collision_BGK_scalar (collision.cu:35-89) */
#define RBTP_STAGE_STREAM_G 6     /* d_g_collide <- d_g, streaming_scalar (streaming.cu:25-63) */
#define RBTP_STAGE_IBB_G 7        /* bounce_back_thermal_particles (ibb.cu:84-136, 223-235) */
#define RBTP_STAGE_PREPARE 8      /* prepare_particle_forces_step (pforces.cu:80-116) */
#define RBTP_STAGE_REPULSION 9    /* apply_repulsive_forces (pforces.cu:6-76, 118-148) */
#define RBTP_STAGE_FORCES 10      /* compute_particle_forces (ibb.cu:138-188, 237-256) */
#define RBTP_STAGE_HEAT 11        /* compute_particle_heat (ibb.cu:190-205, 258-271) */
#define RBTP_STAGE_MOVE 12        /* update_particles (move.cu:5-103) */
#define RBTP_STAGE_BUILD_LINKS 13 /* build_links (links.cu:143-256) */
#define RBTP_STAGE_REFILL 14      /* refill_nodes (refill.cu:244-329) */
#define RBTP_STAGE_COUNT 15
int rbtp_stage(rbtp_sim *s, int stage);
/* The stage the next rbtp_stage call must run (0 at a step boundary). */
int rbtp_next_stage(const rbtp_sim *s, int *stage);

/*
 * Parity accessors (reference semantics at a step boundary).  Each copies to caller-owned host
 * memory.  Fields: which = RBTP_FIELD_*.  Populations: which = RBTP_POP_* (27*N doubles).
 */
#define RBTP_FIELD_RHO 0
#define RBTP_FIELD_UX 1
#define RBTP_FIELD_UY 2
#define RBTP_FIELD_UZ 3
#define RBTP_FIELD_PHI 4
#define RBTP_FIELD_FX 5
#define RBTP_FIELD_FY 6
#define RBTP_FIELD_FZ 7
int rbtp_get_field(rbtp_sim *s, int which, double *out_n);
#define RBTP_POP_F 0          /* f after streaming, IBB and refill (d_f) */
#define RBTP_POP_G 1          /* g after streaming, IBB and refill (d_g) */
#define RBTP_POP_F_COLLIDE 2  /* post-collision f of the last step (d_f_collide) */
#define RBTP_POP_G_COLLIDE 3  /* post-collision g of the last step (d_g_collide) */
int rbtp_get_populations(rbtp_sim *s, int which, double *out_27n);
/* Overwrite f and g (post-stream state, as restart loads them: initial.cu:7-30). */
int rbtp_set_populations(rbtp_sim *s, const double *f_27n, const double *g_27n);

/* ibnode / ibnode_owner after the last build_links (pinit.cu:146-179, links.cu:143-191). */
int rbtp_get_flags(rbtp_sim *s, int32_t *ibnode_n, int32_t *owner_n);

/* Particle state (P doubles each).  which = RBTP_P_*. */
#define RBTP_P_POS_X 0
#define RBTP_P_POS_Y 1
#define RBTP_P_POS_Z 2
#define RBTP_P_VEL_X 3
#define RBTP_P_VEL_Y 4
#define RBTP_P_VEL_Z 5
#define RBTP_P_OMEGA_X 6
#define RBTP_P_OMEGA_Y 7
#define RBTP_P_OMEGA_Z 8
#define RBTP_P_THETA_X 9
#define RBTP_P_THETA_Y 10
#define RBTP_P_THETA_Z 11
#define RBTP_P_FORCE_X 12   /* d_particle_force_accum_* */
#define RBTP_P_FORCE_Y 13
#define RBTP_P_FORCE_Z 14
#define RBTP_P_TORQUE_X 15  /* d_particle_torque_accum_* */
#define RBTP_P_TORQUE_Y 16
#define RBTP_P_TORQUE_Z 17
#define RBTP_P_FORCE_PREV_X 18
#define RBTP_P_FORCE_PREV_Y 19
#define RBTP_P_FORCE_PREV_Z 20
#define RBTP_P_FORCE_PREV2_X 21
#define RBTP_P_FORCE_PREV2_Y 22
#define RBTP_P_FORCE_PREV2_Z 23
#define RBTP_P_TORQUE_PREV_X 24
#define RBTP_P_TORQUE_PREV_Y 25
#define RBTP_P_TORQUE_PREV_Z 26
#define RBTP_P_TORQUE_PREV2_X 27
#define RBTP_P_TORQUE_PREV2_Y 28
#define RBTP_P_TORQUE_PREV2_Z 29
#define RBTP_P_TEMP 30
#define RBTP_P_HEAT 31
#define RBTP_P_HEAT_PREV 32
#define RBTP_P_COUNT 33
int rbtp_get_particles(rbtp_sim *s, int which, double *out_p);
int rbtp_set_particles(rbtp_sim *s, int which, const double *in_p);

/*
 * The IBB link list of the last build_links in the reference's record layout
 * (particle.h:19-29: int cell_i, cell_j, cell_k, dir, pid; double q, rx, ry, rz; 56 bytes),
 * sorted by (idx, dir) as link_fill_kernel emits it (links.cu:89-139).
 * Pass out=NULL to query the count.
 */
typedef struct rbtp_link {
    int32_t cell_i, cell_j, cell_k, dir, pid;
    int32_t pad;
    double q, rx, ry, rz;
} rbtp_link;
int rbtp_get_links(rbtp_sim *s, rbtp_link *out, int64_t capacity, int64_t *count);
/* New-fluid nodes and their previous owners of the last refill_nodes (refill.cu:244-302),
 * sorted by idx.  Pass NULL buffers to query the count. */
int rbtp_get_new_fluid(rbtp_sim *s, int32_t *nodes, int32_t *pids, int64_t capacity,
                       int64_t *count);

/*
 * Diagnostics computed on the device in FP64 (output_statis.cu).  They read the macroscopic
 * fields left by the last step, so call them after rbtp_step(..., RBTP_STEP_MACRO_LAST).
 *   rbtp_nu:    output_nu (output_statis.cu:111-223): out[7] = {step, volume_fraction, Nu_bulk,
 *               Nu_top, Nu_bottom, Nu_bulk_fluid, Nu_bulk_part}.
 *   rbtp_means: diag_flow (output_statis.cu:64-109) after overwrite_particle_nodes:
 *               out[4] = {umean, vmean, wmean, tmean}.
 *   rbtp_kinetic_energy: mean of 0.5*(ux^2+uy^2+uz^2) from the last macro fields over the nodes
 *                        that are fluid in the current and the previous geometry (all nodes
 *                        without particles); no reference counterpart.
 */
int rbtp_nu(rbtp_sim *s, double out[7]);
/* rbtp_nu with no device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate (tight serving loops, harnesses that forbid hipMalloc).  Same kernels and the
 * same result as rbtp_nu, with the partial sums in a fixed static device array of the library
 * (4*lx + 11 doubles, lx at most 8192; RBTP_E_CAPACITY beyond) that every call on a device
 * shares: calls that overlap on one device race. */
int rbtp_nu_noalloc(rbtp_sim *s, double out[7]);
int rbtp_means(rbtp_sim *s, double out[4]);
int rbtp_kinetic_energy(rbtp_sim *s, double *out);
/* Achievable HBM bandwidth on the current device: a streaming copy of `bytes` (read + write,
 * 16-byte vectors) timed over `reps` launches; *gbs = 2*bytes*reps/time in GB/s.  No reference
 * counterpart; used to put the roofline fractions in context. */
int rbtp_bandwidth_probe(size_t bytes, int reps, double *gbs);
/* The HBM access pattern of the FP64 lattice step without its arithmetic, on an lx x ly x lz
 * lattice: the same pull of 54 populations under the wall rules, F_x read and written, and 54
 * stores, timed over `reps` launches; *gbs = 880 B x nodes x reps / time in GB/s.  No reference
 * counterpart; it is the ceiling of the access pattern itself on this device. */
int rbtp_stencil_probe(int lx, int ly, int lz, int reps, double *gbs);
/* The FP64 divisions and square root of the strict kernels on the device, for bitwise checks
 * against IEEE host arithmetic (no reference counterpart: the reference divides with the CUDA
 * compiler's own sequence, which this library replaces by div_rn, a correctly rounded quotient
 * from a double-word reciprocal in three fused multiply-adds).  out[j] = a[j] op b[j] for n
 * host operand pairs (b unused for the one-operand ops):
 *   RBTP_ARITH_DIV            a / b, reciprocal pair formed on the device (grad_denom,
 *                             collision.cu:73-77)
 *   RBTP_ARITH_DIV_HOSTRECIP  a / b, reciprocal pair formed on the host (RT, 3, tau, tauc)
 *   RBTP_ARITH_DIV6           a / 6 (compute_link_q for a diagonal direction, links.cu:21)
 *   RBTP_ARITH_DIV_RT         a / (1.0/3.0) (IBB momentum term, ibb.cu:48-49)
 *   RBTP_ARITH_SQRT           sqrt(a) (links.cu:21, refill.cu:38, pforces.cu:20) */
#define RBTP_ARITH_DIV 0
#define RBTP_ARITH_DIV_HOSTRECIP 1
#define RBTP_ARITH_DIV6 2
#define RBTP_ARITH_DIV_RT 3
#define RBTP_ARITH_SQRT 4
int rbtp_arith_probe(int op, const double *a, const double *b, int64_t n, double *out);
/* Streaming schemes as access patterns on the lattice kernels' layout, no arithmetic, same sizes
 * as rbtp_stencil_probe.  FP64 (GB/s at 880 B per node): 0 the pull the FP64 kernel uses
 * (rbtp_stencil_probe: the row pull where lx is a multiple of 64), 1 and 2 the even and odd
 * steps of in-place AA streaming on one buffer per distribution, 3 the plain pull (every
 * population one 8-byte load).  16-bit storage (GB/s at 224 B per node: 54 2-byte populations
 * and a 4-byte F_x, read and written): 4 the plain pull the FP16/BF16 kernels use, 5 the row pull
 * with lane shifts (lx a multiple of 64), 6 two nodes per thread with 4-byte words (lx even,
 * nodes a multiple of 512).  No reference counterpart; it measures candidate patterns on this
 * device. */
int rbtp_streaming_probe(int scheme, int lx, int ly, int lz, int reps, double *gbs);

/*
 * Output-side fields and files (output/output_flow.cu, output/output_part.cu,
 * output/output_statis.cu).  They read the state left by rbtp_step(..., RBTP_STEP_MACRO_LAST).
 *   rbtp_get_output_fields: the seven fields of %09d.dat in file order (rho, ux, uy, uz, phi,
 *       ted, Ked; 7*N doubles) after overwrite_particle_nodes (output_flow.cu:64-110), with ted
 *       and Ked from compute_ted / compute_Ked (collision.cu:91-198) on the device.
 *   rbtp_profile: output_profile's per-x-plane means of ux, uy, uz, phi (4*lx doubles,
 *       output_statis.cu:225-285).
 *   rbtp_write_fg: output_fg, f then g after streaming/IBB/refill (output_flow.cu:112-127).
 *   rbtp_write_flow: output_flow's %09d.dat (output_flow.cu:97-109).
 *   rbtp_write_particles: output_particles' pos_%09d.txt lines (output_part.cu:91-115).
 *   rbtp_restart_load: load_restart_state (initial.cu:194-236) from <dir>/fg%09d.dat,
 *       <dir>/pos/pos_%09d.txt and <dir>/%09d.dat; fails with RBTP_E_IO on a missing, short or
 *       malformed file (the reference does not check the %09d.dat size).
 */
int rbtp_get_output_fields(rbtp_sim *s, double *out_7n);
int rbtp_profile(rbtp_sim *s, double *out_4lx);
int rbtp_write_fg(rbtp_sim *s, const char *path);
int rbtp_write_flow(rbtp_sim *s, const char *path);
int rbtp_write_particles(rbtp_sim *s, const char *path);
int rbtp_restart_load(rbtp_sim *s, const char *dir, int step);

/*
 * Serving (no counterpart in the reference, which runs one case per process): a batch of
 * independent cavities of any size, Ra, Pr and particle count, all in one precision form,
 * resident in HBM and advanced together.  Each admitted case is a slot; rbtp_batch_slot gives
 * an rbtp_sim view for the accessors and diagnostics above (do not step it directly).
 *   rbtp_batch_create: max_slots concurrent cases; admission keeps the batch under hbm_fraction
 *       of the device memory and leaves 1 GiB free.
 *   rbtp_batch_admit: initialise a case (as rbtp_init) and run its first step; RBTP_E_CAPACITY
 *       when no slot or not enough HBM is free (the caller retries after a release).
 *   rbtp_batch_step: nsteps steps of every active slot: one batched persistent lattice launch per
 *       step plus each slot's particle work, replayed from a hipGraph captured per composition.
 *   rbtp_batch_tune: choose the lattice kernel's tile size and persistent grid by timing the
 *       candidates on the current composition (no slot's state changes).
 *   rbtp_batch_set_launch: fix tile size (64/128/256), grid and graph length (even; 0 = no graph).
 */
typedef struct rbtp_batch rbtp_batch;
int rbtp_batch_create(int precision, int max_slots, double hbm_fraction, rbtp_batch **out);
void rbtp_batch_destroy(rbtp_batch *b);
int rbtp_batch_admit(rbtp_batch *b, const rbtp_case *c, const double *temperature,
                     const double *ppos, const double *ptemp, int *slot);
int rbtp_batch_release(rbtp_batch *b, int slot);
int rbtp_batch_slot(rbtp_batch *b, int slot, rbtp_sim **sim);
int rbtp_batch_step(rbtp_batch *b, int nsteps, unsigned flags);
int rbtp_batch_tune(rbtp_batch *b, int *block, int *grid, double *ms);
int rbtp_batch_set_launch(rbtp_batch *b, int block, int grid, int graph_steps);
int rbtp_batch_info(rbtp_batch *b, int *active, int64_t *used_bytes, int64_t *free_bytes,
                    double *last_ms, int *block, int *grid);

/* Device time of the kernels enqueued by the last rbtp_step* call, from HIP events on the
 * handle's stream (milliseconds). */
int rbtp_last_step_ms(rbtp_sim *s, double *ms);
/* Average duration (ms) of the fused lattice kernel over the last rbtp_step* call, measured
 * with HIP events around every launch when enabled (0 disables, the default).  The events are
 * recorded without waiting and read at the call's synchronisation (rbtp_step, rbtp_synchronize,
 * rbtp_kernel_ms), so timed steps are enqueued as untimed ones are. */
int rbtp_set_kernel_timing(rbtp_sim *s, int enable);
int rbtp_kernel_ms(rbtp_sim *s, double *avg_ms, int64_t *launches);

/* Reduced-precision storage keeps g_i/w_i relative to a per-x-column shift profile, the
 * conduction profile after rbtp_init and rbtp_restart_load.  Every `steps` steps of rbtp_step*
 * the profile is moved to the current plane-mean temperature and the stored g re-encoded
 * (0 never; default 500; ignored for FP64 storage and for batch slots).  No counterpart in the
 * reference, whose FP64 storage has no encoding. */
int rbtp_set_shift_interval(rbtp_sim *s, int steps);

#ifdef __cplusplus
}
#endif

#endif /* RBTP_H */
