// sim.hip — the rbtp C-ABI (include/rbtp.h): handle, device buffers, the time loop and the
// parity accessors.
//
// Time-step organisation.  The reference runs ~15 separate passes per step over global arrays
// (main.cu:35-85) and keeps three copies of f and of g.  Here the state between steps is the
// post-collision pair (f*, g*) in one buffer of a ping-pong pair, and one step is
//   1. the fused lattice kernel: pull-stream of the previous step (wall rules included), then
//      macrovar, forcing and both collisions of this step, into the other buffer; at special
//      nodes (IBB link nodes, refilled nodes) the post-stream populations come from the
//      special-node list instead of the pull;
//   2. the particle phase (particles.hip): force/torque/heat sums, contacts, motion, new
//      geometry, new-fluid detection, and the special-node list with the corrected post-stream
//      populations for the next step.
// Every reference observable at a step boundary is reconstructed by the accessors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rbtp.h"
#include "lattice.h"
#include "particles.h"

using namespace rbtp;

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(call)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return set_error(e_ == hipErrorOutOfMemory ? RBTP_E_NOMEM : RBTP_E_HIP,           \
                             "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, \
                             __LINE__);                                                       \
    } while (0)

#define RB_TRY(call)              \
    do {                          \
        int r_ = (call);          \
        if (r_ != RBTP_OK) return r_; \
    } while (0)

size_t storage_bytes(int prec) {
    switch (prec) {
        case RBTP_PREC_FP64: return 8;
        case RBTP_PREC_FP32: return 4;
        default: return 2;
    }
}
size_t aux_bytes(int prec) { return prec == RBTP_PREC_FP64 ? 8 : 4; }

}  // namespace

struct rbtp_sim {
    rbtp_case c;
    rbtp_derived d;
    Dims dims;
    FluidK k;
    int prec;
    bool strict;
    int device;
    hipStream_t stream;
    bool own_stream;
    hipEvent_t ev_start, ev_stop;
    bool step_events;  // ev_start/ev_stop bracket an asynchronous run not yet timed
    // Per-launch timing of the lattice kernel (rbtp_set_kernel_timing): a pair of events around
    // each launch, recorded without waiting and read after the run's own synchronisation (or when
    // all KT_CAP pairs are in use), so that the timed steps run as they do untimed.
    bool kernel_timing;
    std::vector<hipEvent_t> kt_ev;  // 2 * KT_CAP events, created on first use
    int kt_pending;                 // pairs recorded and not yet read
    double kernel_ms_sum;
    int64_t kernel_launches;
    double last_step_ms;
    // divergence watchdog: watch_dev[0] the first flagged step (INT_MAX: none), [1] the flags
    double watch_umax;
    int32_t *watch_dev;
    // lattice state
    void *f[2], *g[2];
    void *fx;        // F_x (aux type)
    void *rho_wall;  // 2*ly*lz (aux type)
    double *macro[5];  // rho, ux, uy, uz, phi
    // rho of the last step at every node for the particle coupling (aux type): macro[0] in FP64;
    // an array of its own in the FP32-arithmetic forms, whose rho is an FP32 value (4 B per node
    // and step less than FP64).  Null without particles.
    void *rho_full;
    double *temp_dev;  // initial temperature upload
    // Shift profile of the reduced-precision g encoding ([lx], the conduction profile after
    // init/restart), refreshed to the plane-mean temperature every shift_interval steps
    // (0: never); shift_partial [lx*lz] and shift_delta [lx] are the refresh's scratch.
    float *gshift;
    double *shift_partial;
    float *shift_delta;
    int shift_interval;
    int cur;
    bool poststream;   // cur holds post-stream populations (after init / restart / set)
    bool macro_valid;
    int step;
    bool initialised;
    // particles
    PartDev pd;
    bool particles;
    PartTables ptab;  // this lattice's particle-phase tables (one slot)
    bool ptab_valid;
    int ptab_rel;     // cur ^ pd.geo_cur at the particle phase the tables describe
    bool batch_slot;  // owned by a serving batch (stepped by it, never staged)
    // Per-stage execution (rbtp_stage).  While st_active the state is the reference's arrays:
    // d_f and d_g in buffers 1 - cur, d_f_collide and d_g_collide in buffers cur (FP64, blocked
    // layout), the fields and the particle arrays as they are; st_next is the stage the next
    // call must run, st_built that build_links has run and refill not yet (the newest geometry
    // is then owner[1 - geo_cur]).  st_tab are the particle tables of those arrays.
    bool st_active;
    int st_next;
    bool st_built;
    PartTables st_tab;
};

// ---------------------------------------------------------------------------------------------
// Case description
// ---------------------------------------------------------------------------------------------

extern "C" const char *rbtp_last_error(void) { return g_last_error.c_str(); }
extern "C" int rbtp_abi_version(void) { return RBTP_ABI_VERSION; }

// parameters.h:1-55
extern "C" int rbtp_case_default(rbtp_case *c) {
    if (!c) return set_error(RBTP_E_ARG, "rbtp_case_default: null case");
    std::memset(c, 0, sizeof(*c));
    c->continuous_steps = 0;
    c->lx = 320; c->ly = 320; c->lz = 160;
    c->nend = 1000000; c->ndiag = 125; c->nflowout = 25000; c->nnuout = 125; c->nout_p = 125;
    c->rayleigh = 1.0e7; c->prandtl = 0.71;
    c->t_hot = 1.0; c->t_cold = 0.0; c->t_ref = 0.5;
    c->beta = 0.0000625; c->visc = 0.01206; c->rho0 = 1.0;
    c->activate_particles = 1;
    c->particle_fraction = 0.05;
    c->particle_diam_frac = (1.0 / 20.0);
    c->cp_particle_ratio = 10.0;
    c->cp_fluid = 1.0;
    c->rho_particle_ratio = 1.0;
    c->repforce_scale = 2.0e-4;
    c->particle_init_temp = (0.0 + 1.0) / 2;
    c->particle_rand_seed = 86143;
    c->rep_min_gap = 2.0; c->rep_min_gap_w = 2.0;
    c->rep_stf0 = 0.025; c->rep_stf1 = 0.002;
    c->rep_stf0_w = 0.025; c->rep_stf1_w = 0.002;
    c->npart_override = -1;
    c->precision = RBTP_PREC_FP64;
    return RBTP_OK;
}

// globals.cu:18-33, init_particle_constants (globals.cu:126-148), update_particle_count
// (globals.cu:150-169).
extern "C" int rbtp_case_derive(const rbtp_case *c, rbtp_derived *d) {
    if (!c || !d) return set_error(RBTP_E_ARG, "rbtp_case_derive: null argument");
    std::memset(d, 0, sizeof(*d));
    const double rayl = c->rayleigh, prand = c->prandtl, beta = c->beta, visc = c->visc;
    const double tHot = c->t_hot, tCold = c->t_cold;
    const int LX = c->lx;
    d->grav0 = (beta != 0.0 && (tHot - tCold) != 0.0)
                   ? rayl * visc * visc /
                         (beta * (tHot - tCold) * (double)LX * (double)LX * (double)LX * prand)
                   : 0.0;
    d->tau = 3.0 * visc + 0.5;
    d->u0 = std::sqrt(std::fabs(d->grav0 * beta * (tHot - tCold) * (double)LX));
    d->diff = visc / prand;
    d->tauc = 3.0 * d->diff + 0.5;
    d->tauci = 1.0 / d->tauc;
    constexpr double pi = 3.14159265358979323846;
    const double radius = 0.5 * c->particle_diam_frac * static_cast<double>(LX);
    const double volume = 4.0 * pi * radius * radius * radius / 3.0;
    const double mass = c->rho0 * c->rho_particle_ratio * volume;
    d->particle_radius = radius;
    d->particle_mass = mass;
    d->particle_inertia = 0.4 * mass * radius * radius;
    d->particle_heat_capacity = mass * c->cp_particle_ratio * c->cp_fluid;
    if (!c->activate_particles) {
        d->npart = 0;
    } else if (c->npart_override >= 0) {
        d->npart = c->npart_override;
    } else if (c->particle_fraction <= 0.0 || volume <= 0.0) {
        d->npart = 0;
    } else {
        const int lxyz = c->lx * c->ly * c->lz;
        const double raw = lxyz * c->particle_fraction / volume;
        d->npart = static_cast<int>(std::floor(raw < 0.0 ? 0.0 : raw));
    }
    return RBTP_OK;
}

static int validate_case(const rbtp_case *c) {
    if (c->lx < 3 || c->ly < 3 || c->lz < 3)
        return set_error(RBTP_E_ARG, "lattice %dx%dx%d: every extent must be >= 3", c->lx, c->ly, c->lz);
    const long long n = (long long)c->lx * c->ly * c->lz;
    if (n >= (1ll << 31) / 1)
        return set_error(RBTP_E_ARG, "lattice of %lld nodes exceeds 2^31", n);
    if (c->precision < RBTP_PREC_FP64 || c->precision > RBTP_PREC_BF16S)
        return set_error(RBTP_E_ARG, "unknown precision %d", c->precision);
    if (!(c->visc > 0.0) || !(c->prandtl > 0.0))
        return set_error(RBTP_E_ARG, "visc and prandtl must be positive");
    if (c->activate_particles && c->particle_diam_frac < 0.0)
        return set_error(RBTP_E_ARG, "particle_diam_frac must be >= 0");
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Create / destroy
// ---------------------------------------------------------------------------------------------

extern "C" void rbtp_destroy(rbtp_sim *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (int b = 0; b < 2; b++) (void)hipFree(s->f[b]);  // g[b] lies in the same pair buffer
    (void)hipFree(s->fx);
    (void)hipFree(s->rho_wall);
    if (s->rho_full != (void *)s->macro[0]) (void)hipFree(s->rho_full);
    for (double *m : s->macro) (void)hipFree(m);
    (void)hipFree(s->temp_dev);
    (void)hipFree(s->gshift);
    (void)hipFree(s->shift_partial);
    (void)hipFree(s->shift_delta);
    (void)hipFree(s->watch_dev);
    part_free(s->pd);
    part_tables_free(s->ptab);
    part_tables_free(s->st_tab);
    if (s->ev_start) (void)hipEventDestroy(s->ev_start);
    if (s->ev_stop) (void)hipEventDestroy(s->ev_stop);
    for (hipEvent_t e : s->kt_ev) (void)hipEventDestroy(e);
    if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

// Steps between refreshes of the shift profile of a reduced-precision sim (rbtp_step).
constexpr int SHIFT_INTERVAL = 500;

// The shift profile back to the conduction profile T(ix) = tl0 + tls*ix (init and restart
// encode against it).
static int reset_shift(rbtp_sim *s) {
    std::vector<float> h((size_t)s->c.lx);
    for (int ix = 0; ix < s->c.lx; ix++) h[ix] = float(s->k.tl0 + s->k.tls * double(ix));
    HIP_TRY(hipMemcpyAsync(s->gshift, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice,
                           s->stream));
    HIP_TRY(hipStreamSynchronize(s->stream));
    return RBTP_OK;
}

static int create_impl(const rbtp_case *c, rbtp_sim *s, hipStream_t stream) {
    s->c = *c;
    rbtp_case_derive(c, &s->d);
    s->dims = Dims{c->lx, c->ly, c->lz, c->lx * c->ly, c->lx * c->ly * c->lz};
    const double tau = s->d.tau, tauc = s->d.tauc;
    s->k.tau = tau;
    s->k.tauc = tauc;
    s->k.gb = s->d.grav0 * c->beta;
    s->k.t0 = c->t_ref;
    s->k.thot = c->t_hot;
    s->k.tcold = c->t_cold;
    s->k.force_coeff = (1.0 - 0.5 / tau);
    s->k.lambda0 = 1.0 - weight(0);
    s->k.ralph_coeff = (1.0 - 1.0 / (2.0 * tauc));
    s->k.c2t = 2.0 * tauc / 3.0;
    s->k.inv_tau = 1.0 / tau;
    s->k.inv_tauc = 1.0 / tauc;
    dw_recip(1.0 / 3.0, s->k.rt_h, s->k.rt_l);
    dw_recip(3.0, s->k.three_h, s->k.three_l);
    dw_recip(tau, s->k.tau_h, s->k.tau_l);
    dw_recip(tauc, s->k.tauc_h, s->k.tauc_l);
    s->k.tls = (c->t_cold - c->t_hot) / (double)c->lx;
    s->k.tl0 = c->t_hot + 0.5 * s->k.tls;
    s->prec = c->precision;
    s->strict = (c->precision == RBTP_PREC_FP64);
    HIP_TRY(hipGetDevice(&s->device));
    if (stream) {
        s->stream = stream;
        s->own_stream = false;
    } else {
        HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        s->own_stream = true;
    }
    HIP_TRY(hipEventCreate(&s->ev_start));
    HIP_TRY(hipEventCreate(&s->ev_stop));
    {
        const int32_t none[2] = {INT_MAX, 0};
        HIP_TRY(hipMalloc(&s->watch_dev, sizeof(none)));
        HIP_TRY(hipMemcpy(s->watch_dev, none, sizeof(none), hipMemcpyHostToDevice));
    }
    const size_t n = (size_t)s->dims.n;
    // one pair buffer (f and g interleaved by runs) per time level
    const size_t pair_bytes = pop_buffer_len(n) * storage_bytes(s->prec);
    for (int b = 0; b < 2; b++) {
        HIP_TRY(hipMalloc(&s->f[b], pair_bytes));
        s->g[b] = (char *)s->f[b] + G_OFFSET * storage_bytes(s->prec);
    }
    HIP_TRY(hipMalloc(&s->fx, n * aux_bytes(s->prec)));
    HIP_TRY(hipMalloc(&s->rho_wall, 2 * (size_t)c->ly * c->lz * aux_bytes(s->prec)));
    for (double *&m : s->macro) HIP_TRY(hipMalloc(&m, n * sizeof(double)));
    HIP_TRY(hipMalloc(&s->gshift, (size_t)c->lx * sizeof(float)));
    HIP_TRY(hipMalloc(&s->shift_partial, (size_t)c->lx * c->lz * sizeof(double)));
    HIP_TRY(hipMalloc(&s->shift_delta, (size_t)c->lx * sizeof(float)));
    RB_TRY(reset_shift(s));
    s->shift_interval = s->prec == RBTP_PREC_FP64 ? 0 : SHIFT_INTERVAL;
    s->particles = c->activate_particles && s->d.npart > 0;
    if (s->particles) {
        RB_TRY(part_alloc(s->pd, *c, s->d, s->dims));
        s->pd.gshift = s->gshift;
        if (s->prec == RBTP_PREC_FP64) s->rho_full = s->macro[0];
        else HIP_TRY(hipMalloc(&s->rho_full, n * sizeof(float)));
    }
    return RBTP_OK;
}

static int create_on_stream(const rbtp_case *c, rbtp_sim **out, hipStream_t stream) {
    if (!c || !out) return set_error(RBTP_E_ARG, "rbtp_create: null argument");
    *out = nullptr;
    RB_TRY(validate_case(c));
    rbtp_sim *s = new rbtp_sim();
    std::memset((void *)s, 0, sizeof(*s));
    const int r = create_impl(c, s, stream);
    if (r != RBTP_OK) {
        std::string msg = g_last_error;
        rbtp_destroy(s);
        g_last_error = msg;
        return r;
    }
    *out = s;
    return RBTP_OK;
}

extern "C" int rbtp_create(const rbtp_case *c, rbtp_sim **out) {
    return create_on_stream(c, out, nullptr);
}

extern "C" int rbtp_get_case(const rbtp_sim *s, rbtp_case *c, rbtp_derived *d) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (c) *c = s->c;
    if (d) *d = s->d;
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Typed dispatch helpers
// ---------------------------------------------------------------------------------------------

template <class F>
static hipError_t dispatch(int prec, F &&fn) {
    switch (prec) {
        case RBTP_PREC_FP64: return fn((double *)nullptr);
        case RBTP_PREC_FP32: return fn((float *)nullptr);
        case RBTP_PREC_FP16S: return fn((__half *)nullptr);
        default: return fn((__hip_bfloat16 *)nullptr);
    }
}

template <class S>
static StepArgs<S> step_args(rbtp_sim *s, bool macro_out) {
    StepArgs<S> a;
    a.fin = (const S *)s->f[s->cur];
    a.gin = (const S *)s->g[s->cur];
    a.fout = (S *)s->f[1 - s->cur];
    a.gout = (S *)s->g[1 - s->cur];
    a.fx = (typename Enc<S>::A *)s->fx;
    a.rho_wall = (typename Enc<S>::A *)s->rho_wall;
    a.rho_full = (typename Enc<S>::A *)s->rho_full;
    a.gshift = s->gshift;
    a.sp_map = s->particles ? s->pd.sp.map : nullptr;
    a.sp_pop = s->particles ? (const typename Enc<S>::A *)s->pd.sp.pop : nullptr;
    a.sp_mask = s->particles ? s->pd.sp.mask : nullptr;
    if (macro_out) {
        a.macro_rho = s->macro[0];
        a.macro_ux = s->macro[1];
        a.macro_uy = s->macro[2];
        a.macro_uz = s->macro[3];
        a.macro_phi = s->macro[4];
    } else {
        a.macro_rho = a.macro_ux = a.macro_uy = a.macro_uz = a.macro_phi = nullptr;
    }
    a.watch = (s->watch_umax > 0.0 && !s->batch_slot) ? s->watch_dev : nullptr;
    a.watch_step = s->step + 1;
    a.watch_u2 = s->watch_umax * s->watch_umax;
    a.d = s->dims;
    a.k = s->k;
    return a;
}

static int stage_exit(rbtp_sim *s);
static int current_gen(const rbtp_sim *s);
static int watch_check(rbtp_sim *s);

// ---------------------------------------------------------------------------------------------
// Initial state
// ---------------------------------------------------------------------------------------------

static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

extern "C" int rbtp_linear_temperature(const rbtp_case *c, uint64_t seed, double amp, double *out) {
    if (!c || !out) return set_error(RBTP_E_ARG, "rbtp_linear_temperature: null argument");
    const int LX = c->lx;
    const long n = (long)c->lx * c->ly * c->lz;
    uint64_t st = seed;
    for (long idx = 0; idx < n; idx++) {
        const int ix = (int)(idx % LX);
        const double xi = (static_cast<double>(ix) + 0.5) / static_cast<double>(LX);
        const double lin = c->t_hot + (c->t_cold - c->t_hot) * xi;
        const double u = static_cast<double>(splitmix64(st) >> 11) * 0x1.0p-53;
        out[idx] = lin + amp * (2.0 * u - 1.0);
    }
    return RBTP_OK;
}

extern "C" int rbtp_init(rbtp_sim *s, const double *temperature, const double *ppos,
                         const double *ptemp) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->dims.n;
    const double *tdev = nullptr;
    if (temperature) {
        if (!s->temp_dev) HIP_TRY(hipMalloc(&s->temp_dev, n * sizeof(double)));
        HIP_TRY(copy_to_device(s->temp_dev, temperature, n * sizeof(double), s->stream));
        tdev = s->temp_dev;
    }
    s->cur = 0;
    s->st_active = false;
    s->ptab_valid = false;
    RB_TRY(reset_shift(s));
    HIP_TRY(dispatch(s->prec, [&](auto *tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        return launch_init<S>((S *)s->f[0], (S *)s->g[0], (typename Enc<S>::A *)s->fx,
                              (typename Enc<S>::A *)s->rho_wall, tdev, s->c.rho0, s->c.t_ref,
                              s->d.grav0, s->c.beta, s->gshift, s->dims, s->macro[0], s->macro[1],
                              s->macro[2], s->macro[3], s->macro[4], s->stream);
    }));
    if (s->particles) {
        // init_particles (pinit.cu:183-295) and the first build_links (initial.cu:330).
        RB_TRY(part_init(s->pd, s->c, s->d, s->dims, ppos, ptemp, s->stream));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->poststream = true;
    s->macro_valid = true;
    s->step = 0;
    s->initialised = true;
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Time loop
// ---------------------------------------------------------------------------------------------

// The particle phase's view of this lattice when its post-collision populations are in buffer
// `cur` and the geometry of the step's IBB is generation `geo`.
static PartSlot part_slot(const rbtp_sim *s, int cur, int geo) {
    PartSlot sl;
    sl.pd = s->pd;
    sl.d = s->dims;
    sl.owner_prev = s->pd.owner[geo];
    sl.box_prev = s->pd.box[geo];
    sl.owner_curr = s->pd.owner[1 - geo];
    sl.box_curr = s->pd.box[1 - geo];
    sl.fcur = s->f[cur];
    sl.gcur = s->g[cur];
    sl.fx = s->fx;
    sl.rho_wall = s->rho_wall;
    for (int k = 0; k < 5; k++) sl.macro[k] = s->macro[k];
    sl.rho_full = s->rho_full;
    return sl;
}

// Everything of a time step after the lattice kernel: the buffer flip and the particle phase.
static int after_lattice(rbtp_sim *s) {
    s->cur = 1 - s->cur;
    s->poststream = false;
    if (s->particles) {
        // cur and geo_cur alternate together, so two tables (one per geometry generation) cover
        // every step until a reload changes their relation
        const int rel = s->cur ^ s->pd.geo_cur;
        if (!s->ptab_valid || rel != s->ptab_rel) {
            std::vector<PartSlot> v[2];
            for (int g = 0; g < 2; g++) v[g].push_back(part_slot(s, g ^ rel, g));
            RB_TRY(part_tables_build(s->ptab, v, s->stream));
            s->ptab_valid = true;
            s->ptab_rel = rel;
        }
        RB_TRY(part_phase(s->ptab.g[s->pd.geo_cur], s->prec, s->stream));
        s->pd.geo_cur ^= 1;
        s->pd.special_pending = true;
    }
    s->step++;
    if (s->shift_interval > 0 && s->step % s->shift_interval == 0) {
        // after the particle phase, which read this state under the old profile
        HIP_TRY(dispatch(s->prec, [&](auto *tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            return launch_shift_refresh<S>((S *)s->g[s->cur], s->dims, s->gshift,
                                           s->shift_partial, s->shift_delta, s->stream);
        }));
    }
    return RBTP_OK;
}

// Pairs of timing events in the pool of a sim.
constexpr int KT_CAP = 1024;

// Adds the recorded launch times to the sums (waits for the last recorded launch).
static int kt_drain(rbtp_sim *s) {
    if (s->kt_pending == 0) return RBTP_OK;
    HIP_TRY(hipEventSynchronize(s->kt_ev[2 * (size_t)s->kt_pending - 1]));
    for (int j = 0; j < s->kt_pending; j++) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->kt_ev[2 * (size_t)j], s->kt_ev[2 * (size_t)j + 1]));
        s->kernel_ms_sum += ms;
        s->kernel_launches++;
    }
    s->kt_pending = 0;
    return RBTP_OK;
}

static int one_step(rbtp_sim *s, bool macro_out) {
    const bool nostream = s->poststream;
    const hipError_t pending = hipGetLastError();
    if (pending != hipSuccess)
        return set_error(RBTP_E_HIP, "step %d: HIP error left by earlier work: %s", s->step,
                         hipGetErrorString(pending));
    hipEvent_t k0 = nullptr, k1 = nullptr;
    if (s->kernel_timing) {
        if (s->kt_pending == KT_CAP) RB_TRY(kt_drain(s));
        k0 = s->kt_ev[2 * (size_t)s->kt_pending];
        k1 = s->kt_ev[2 * (size_t)s->kt_pending + 1];
    }
    HIP_TRY(dispatch(s->prec, [&](auto *tag) -> hipError_t {
        using S = std::remove_pointer_t<decltype(tag)>;
        const StepArgs<S> a = step_args<S>(s, macro_out);
        if (k0) {
            hipError_t e = hipEventRecord(k0, s->stream);
            if (e != hipSuccess) return e;
        }
        hipError_t e = launch_lattice_step<S>(a, s->strict, nostream, s->stream);
        if (e != hipSuccess) return e;
        if (k1) e = hipEventRecord(k1, s->stream);
        return e;
    }));
    if (k0) s->kt_pending++;
    return after_lattice(s);
}

static int step_impl(rbtp_sim *s, int nsteps, unsigned flags, bool sync) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (!s->initialised) return set_error(RBTP_E_STATE, "rbtp_step before rbtp_init");
    if (nsteps < 0) return set_error(RBTP_E_ARG, "nsteps < 0");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(stage_exit(s));
    if (s->kernel_timing) {
        RB_TRY(kt_drain(s));
        s->kernel_ms_sum = 0.0;
        s->kernel_launches = 0;
    }
    HIP_TRY(hipEventRecord(s->ev_start, s->stream));
    for (int n = 0; n < nsteps; n++) {
        const bool macro_out = (flags & RBTP_STEP_MACRO_LAST) && n == nsteps - 1;
        RB_TRY(one_step(s, macro_out));
        if (nsteps > 0) s->macro_valid = macro_out;
    }
    HIP_TRY(hipEventRecord(s->ev_stop, s->stream));
    s->step_events = !sync;
    if (sync) {
        HIP_TRY(hipEventSynchronize(s->ev_stop));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_start, s->ev_stop));
        s->last_step_ms = ms;
        RB_TRY(kt_drain(s));
        if (s->particles) RB_TRY(part_check_errors(s->pd));
        RB_TRY(watch_check(s));
    }
    return RBTP_OK;
}

// The watchdog's verdict on the steps run so far (after a synchronisation); clears it.
static int watch_check(rbtp_sim *s) {
    if (!(s->watch_umax > 0.0) || s->batch_slot) return RBTP_OK;
    int32_t w[2];
    HIP_TRY(hipMemcpy(w, s->watch_dev, sizeof(w), hipMemcpyDeviceToHost));
    if (w[0] == INT_MAX) return RBTP_OK;
    const int32_t none[2] = {INT_MAX, 0};
    HIP_TRY(hipMemcpy(s->watch_dev, none, sizeof(none), hipMemcpyHostToDevice));
    return set_error(RBTP_E_DIVERGED,
                     "step %d: %d node update(s) with a non-finite density or velocity or "
                     "|u| > %g", w[0], w[1], s->watch_umax);
}

extern "C" int rbtp_set_watchdog(rbtp_sim *s, double umax) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (!(umax >= 0.0) || !std::isfinite(umax))
        return set_error(RBTP_E_ARG, "watchdog umax must be finite and >= 0");
    s->watch_umax = umax;
    return RBTP_OK;
}

extern "C" int rbtp_step(rbtp_sim *s, int nsteps, unsigned flags) {
    return step_impl(s, nsteps, flags, true);
}
extern "C" int rbtp_step_async(rbtp_sim *s, int nsteps, unsigned flags) {
    return step_impl(s, nsteps, flags, false);
}
extern "C" int rbtp_synchronize(rbtp_sim *s) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->step_events) {
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev_start, s->ev_stop));
        s->last_step_ms = ms;
        s->step_events = false;
    }
    RB_TRY(kt_drain(s));
    if (s->particles) RB_TRY(part_check_errors(s->pd));
    RB_TRY(watch_check(s));
    return RBTP_OK;
}
extern "C" int rbtp_current_step(const rbtp_sim *s, int *step) {
    if (!s || !step) return set_error(RBTP_E_ARG, "null argument");
    *step = s->step;
    return RBTP_OK;
}
extern "C" int rbtp_last_step_ms(rbtp_sim *s, double *ms) {
    if (!s || !ms) return set_error(RBTP_E_ARG, "null argument");
    *ms = s->last_step_ms;
    return RBTP_OK;
}
extern "C" int rbtp_set_kernel_timing(rbtp_sim *s, int enable) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(kt_drain(s));
    if (enable && s->kt_ev.empty()) {
        s->kt_ev.assign(2 * (size_t)KT_CAP, nullptr);
        for (hipEvent_t &e : s->kt_ev) HIP_TRY(hipEventCreate(&e));
    }
    s->kernel_timing = enable != 0;
    return RBTP_OK;
}
extern "C" int rbtp_set_shift_interval(rbtp_sim *s, int steps) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (steps < 0) return set_error(RBTP_E_ARG, "negative shift interval %d", steps);
    s->shift_interval = s->prec == RBTP_PREC_FP64 ? 0 : steps;
    return RBTP_OK;
}
extern "C" int rbtp_kernel_ms(rbtp_sim *s, double *avg_ms, int64_t *launches) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(kt_drain(s));
    if (avg_ms) *avg_ms = s->kernel_launches ? s->kernel_ms_sum / s->kernel_launches : 0.0;
    if (launches) *launches = s->kernel_launches;
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Per-stage execution
// ---------------------------------------------------------------------------------------------

static const char *const kStageNames[RBTP_STAGE_COUNT] = {
    "macrovar", "forcing", "collide_f", "stream_f", "ibb_f", "collide_g", "stream_g", "ibb_g",
    "prepare", "repulsion", "forces", "heat", "move", "build_links", "refill"};

// Geometry generation that holds the reference's current ibnode/owner.
static int current_gen(const rbtp_sim *s) {
    return (s->st_active && s->st_built) ? 1 - s->pd.geo_cur : s->pd.geo_cur;
}

static StageFluid stage_fluid(rbtp_sim *s) {
    StageFluid a;
    a.f = (double *)s->f[1 - s->cur];
    a.fc = (double *)s->f[s->cur];
    a.g = (double *)s->g[1 - s->cur];
    a.gc = (double *)s->g[s->cur];
    a.rho = s->macro[0];
    a.ux = s->macro[1];
    a.uy = s->macro[2];
    a.uz = s->macro[3];
    a.phi = s->macro[4];
    a.fx = (double *)s->fx;
    a.rho_wall = (const double *)s->rho_wall;
    a.d = s->dims;
    a.k = s->k;
    return a;
}

// Particle tables over the staged arrays: d_f_collide/d_g_collide in buffer cur, the links of
// geometry generation geo_cur, build_links writing the other one.
static int stage_tables(rbtp_sim *s) {
    std::vector<PartSlot> v[2];
    for (int g = 0; g < 2; g++) v[g].push_back(part_slot(s, s->cur, s->pd.geo_cur));
    return part_tables_build(s->st_tab, v, s->stream);
}

// From the fused state at a step boundary to the reference's arrays: d_f_collide and
// d_g_collide are the post-collision buffers (or, right after an initialisation, the post-stream
// ones, which initial.cu:326 copies into them), d_f and d_g the post-stream populations pulled
// from them with the special nodes' corrected values.
static int stage_enter(rbtp_sim *s) {
    if (!s->initialised) return set_error(RBTP_E_STATE, "rbtp_stage before rbtp_init");
    if (s->prec != RBTP_PREC_FP64)
        return set_error(RBTP_E_STATE, "per-stage execution runs the FP64 form only");
    if (s->batch_slot)
        return set_error(RBTP_E_STATE, "a serving-batch slot is stepped by its batch");
    const StageFluid a = stage_fluid(s);
    HIP_TRY(launch_stage_enter(a, s->poststream, s->stream));
    if (s->particles && !s->poststream)
        HIP_TRY(part_overwrite_special_blocked(s->pd, a.f, a.g, s->dims, s->stream));
    if (s->particles) RB_TRY(stage_tables(s));
    HIP_TRY(hipStreamSynchronize(s->stream));
    s->st_active = true;
    s->st_next = 0;
    s->st_built = false;
    return RBTP_OK;
}

// Back to the fused state at a step boundary: d_f and d_g become the post-stream state the next
// lattice kernel reads in place, and the special-node list is empty.
static int stage_exit(rbtp_sim *s) {
    if (!s->st_active) return RBTP_OK;
    if (s->st_next != 0)
        return set_error(RBTP_E_STATE, "the staged step is not finished (next stage: %s)",
                         kStageNames[s->st_next]);
    s->cur = 1 - s->cur;
    s->poststream = true;
    s->ptab_valid = false;
    if (s->particles) {
        s->pd.special_pending = false;
        HIP_TRY(hipMemsetAsync(s->pd.sp.count, 0, 2 * sizeof(int), s->stream));
        HIP_TRY(hipMemsetAsync(s->pd.sp.map, 0xFF, (size_t)s->dims.n * sizeof(int32_t), s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    s->st_active = false;
    return RBTP_OK;
}

extern "C" int rbtp_stage(rbtp_sim *s, int stage) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (stage < 0 || stage >= RBTP_STAGE_COUNT) return set_error(RBTP_E_ARG, "unknown stage %d", stage);
    HIP_TRY(hipSetDevice(s->device));
    if (!s->st_active) {
        if (stage != 0)
            return set_error(RBTP_E_STATE, "the staged step starts with macrovar, not %s",
                             kStageNames[stage]);
        RB_TRY(stage_enter(s));
    }
    if (stage != s->st_next)
        return set_error(RBTP_E_STATE, "stage %s called where main.cu runs %s", kStageNames[stage],
                         kStageNames[s->st_next]);
    const StageFluid a = stage_fluid(s);
    switch (stage) {
        case RBTP_STAGE_MACROVAR:
        case RBTP_STAGE_FORCING:
        case RBTP_STAGE_COLLIDE_F:
        case RBTP_STAGE_STREAM_F:
        case RBTP_STAGE_COLLIDE_G:
        case RBTP_STAGE_STREAM_G:
            HIP_TRY(launch_stage_fluid(stage, a, s->stream));
            break;
        default:
            if (!s->particles) break;
            RB_TRY(part_stage(stage, s->st_tab.g[0], a.f, a.g, s->stream));
            if (stage == RBTP_STAGE_BUILD_LINKS) s->st_built = true;
            if (stage == RBTP_STAGE_REFILL) {
                // ibnode_prev <- ibnode, owner_prev <- owner (refill.cu:327-328)
                s->pd.geo_cur ^= 1;
                s->st_built = false;
                s->pd.special_pending = true;  // the list holds this step's new-fluid nodes
                RB_TRY(stage_tables(s));
            }
            break;
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->particles) RB_TRY(part_check_errors(s->pd));
    if (stage == RBTP_STAGE_MACROVAR) s->macro_valid = true;
    s->st_next = (stage + 1) % RBTP_STAGE_COUNT;
    if (s->st_next == 0) s->step++;
    return RBTP_OK;
}

extern "C" int rbtp_next_stage(const rbtp_sim *s, int *stage) {
    if (!s || !stage) return set_error(RBTP_E_ARG, "null argument");
    *stage = s->st_active ? s->st_next : 0;
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Accessors
// ---------------------------------------------------------------------------------------------

extern "C" int rbtp_get_field(rbtp_sim *s, int which, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->initialised) return set_error(RBTP_E_STATE, "not initialised");
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    const size_t n = (size_t)s->dims.n;
    if (which >= RBTP_FIELD_RHO && which <= RBTP_FIELD_PHI) {
        if (!s->macro_valid)
            return set_error(RBTP_E_STATE,
                             "macroscopic fields are only kept when the last step ran with "
                             "RBTP_STEP_MACRO_LAST");
        HIP_TRY(copy_to_host(out, s->macro[which], n * sizeof(double), s->stream));
        return RBTP_OK;
    }
    if (which == RBTP_FIELD_FX) {
        if (s->prec == RBTP_PREC_FP64) {
            HIP_TRY(copy_to_host(out, s->fx, n * sizeof(double), s->stream));
        } else {
            std::vector<float> tmp(n);
            HIP_TRY(copy_to_host(tmp.data(), s->fx, n * sizeof(float), s->stream));
            for (size_t i = 0; i < n; i++) out[i] = tmp[i];
        }
        return RBTP_OK;
    }
    if (which == RBTP_FIELD_FY || which == RBTP_FIELD_FZ) {
        std::fill(out, out + n, 0.0);  // forcing writes 0 (macrovar.cu:45-46)
        return RBTP_OK;
    }
    return set_error(RBTP_E_ARG, "unknown field %d", which);
}

extern "C" int rbtp_get_populations(rbtp_sim *s, int which, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->initialised) return set_error(RBTP_E_STATE, "not initialised");
    if (which < RBTP_POP_F || which > RBTP_POP_G_COLLIDE) return set_error(RBTP_E_ARG, "unknown population set %d", which);
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->dims.n;
    double *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, 2 * 27 * n * sizeof(double)));
    int rc = RBTP_OK;
    hipError_t e = hipSuccess;
    const bool want_g = (which == RBTP_POP_G || which == RBTP_POP_G_COLLIDE);
    if (s->st_active) {
        // the staged arrays: d_f, d_g in buffers 1 - cur, d_f_collide, d_g_collide in cur
        const bool collide = (which == RBTP_POP_F_COLLIDE || which == RBTP_POP_G_COLLIDE);
        const int b = collide ? s->cur : 1 - s->cur;
        e = launch_decode<double>((const double *)(want_g ? s->g[b] : s->f[b]), dev, n,
                                  want_g ? 1 : 0, s->gshift, s->dims.lx, s->stream);
    } else if (which == RBTP_POP_F_COLLIDE || which == RBTP_POP_G_COLLIDE || s->poststream) {
        // d_f_collide holds f after init (initial.cu:326) and the last collision otherwise.
        e = dispatch(s->prec, [&](auto *tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            const void *src = want_g ? s->g[s->cur] : s->f[s->cur];
            return launch_decode<S>((const S *)src, dev, n, want_g ? 1 : 0, s->gshift,
                                    s->dims.lx, s->stream);
        });
    } else {
        e = dispatch(s->prec, [&](auto *tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            return launch_stream_only<S>((const S *)s->f[s->cur], (const S *)s->g[s->cur],
                                         (const typename Enc<S>::A *)s->rho_wall, dev,
                                         dev + 27 * n, s->dims, s->gshift, s->c.t_hot,
                                         s->c.t_cold, false, s->stream);
        });
        if (e == hipSuccess && s->particles)
            e = part_overwrite_special(s->pd, dev, dev + 27 * n, s->dims, s->stream);
        if (want_g && e == hipSuccess) {
            e = hipMemcpyAsync(dev, dev + 27 * n, 27 * n * sizeof(double), hipMemcpyDeviceToDevice, s->stream);
        }
    }
    if (e == hipSuccess) e = copy_to_host(out, dev, 27 * n * sizeof(double), s->stream);
    if (e != hipSuccess) rc = set_error(RBTP_E_HIP, "rbtp_get_populations: %s", hipGetErrorString(e));
    (void)hipFree(dev);
    return rc;
}

extern "C" int rbtp_set_populations(rbtp_sim *s, const double *f, const double *g) {
    if (!s || !f || !g) return set_error(RBTP_E_ARG, "null argument");
    if (!s->initialised) return set_error(RBTP_E_STATE, "not initialised");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(stage_exit(s));
    const size_t n = (size_t)s->dims.n;
    double *dev = nullptr;
    HIP_TRY(hipMalloc(&dev, 27 * n * sizeof(double)));
    hipError_t e = hipSuccess;
    for (int which = 0; which < 2 && e == hipSuccess; which++) {
        e = copy_to_device(dev, which ? g : f, 27 * n * sizeof(double), s->stream);
        if (e != hipSuccess) break;
        e = dispatch(s->prec, [&](auto *tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            return launch_encode<S>(dev, (S *)(which ? s->g[s->cur] : s->f[s->cur]), n, which,
                                    s->gshift, s->dims.lx, s->stream);
        });
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return set_error(RBTP_E_HIP, "rbtp_set_populations: %s", hipGetErrorString(e));
    s->poststream = true;
    if (s->particles) {
        // the buffered special-node populations no longer belong to this state
        s->pd.special_pending = false;
        HIP_TRY(hipMemsetAsync(s->pd.sp.count, 0, 2 * sizeof(int), s->stream));
        HIP_TRY(hipMemsetAsync(s->pd.sp.map, 0xFF, (size_t)s->dims.n * sizeof(int32_t), s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    return RBTP_OK;
}

extern "C" int rbtp_get_flags(rbtp_sim *s, int32_t *ibnode, int32_t *owner) {
    if (!s) return set_error(RBTP_E_ARG, "null handle");
    if (!s->initialised) return set_error(RBTP_E_STATE, "not initialised");
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->dims.n;
    std::vector<int32_t> own(n, -1);
    if (s->particles) {
        HIP_TRY(hipStreamSynchronize(s->stream));
        HIP_TRY(copy_to_host(own.data(), s->pd.owner[current_gen(s)], n * sizeof(int32_t), s->stream));
    }
    if (owner) std::copy(own.begin(), own.end(), owner);
    if (ibnode)
        for (size_t i = 0; i < n; i++) ibnode[i] = own[i] >= 0 ? 1 : 0;
    return RBTP_OK;
}

extern "C" int rbtp_get_particles(rbtp_sim *s, int which, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (which < 0 || which >= RBTP_P_COUNT) return set_error(RBTP_E_ARG, "unknown particle array %d", which);
    if (!s->particles) return RBTP_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(copy_to_host(out, s->pd.p[which], s->pd.np * sizeof(double), s->stream));
    return RBTP_OK;
}

extern "C" int rbtp_set_particles(rbtp_sim *s, int which, const double *in) {
    if (!s || !in) return set_error(RBTP_E_ARG, "null argument");
    if (which < 0 || which >= RBTP_P_COUNT) return set_error(RBTP_E_ARG, "unknown particle array %d", which);
    if (!s->particles) return RBTP_OK;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(copy_to_device(s->pd.p[which], in, s->pd.np * sizeof(double), s->stream));
    return RBTP_OK;
}

extern "C" int rbtp_get_links(rbtp_sim *s, rbtp_link *out, int64_t capacity, int64_t *count) {
    if (!s || !count) return set_error(RBTP_E_ARG, "null argument");
    *count = 0;
    if (!s->particles) return RBTP_OK;
    HIP_TRY(hipSetDevice(s->device));
    std::vector<rbtp_link> links;
    // between update_particles and build_links of a staged step the list is still the one built
    // from the positions before the motion
    const bool snapshot = s->st_active && s->st_next == RBTP_STAGE_BUILD_LINKS;
    RB_TRY(part_collect_links(s->pd, current_gen(s), snapshot, s->dims, s->stream, links));
    std::sort(links.begin(), links.end(), [](const rbtp_link &a, const rbtp_link &b) {
        if (a.cell_k != b.cell_k) return a.cell_k < b.cell_k;
        if (a.cell_j != b.cell_j) return a.cell_j < b.cell_j;
        if (a.cell_i != b.cell_i) return a.cell_i < b.cell_i;
        return a.dir < b.dir;
    });
    *count = (int64_t)links.size();
    if (out) {
        if (capacity < *count) return set_error(RBTP_E_ARG, "link buffer too small (%lld < %lld)", (long long)capacity, (long long)*count);
        std::copy(links.begin(), links.end(), out);
    }
    return RBTP_OK;
}

extern "C" int rbtp_get_new_fluid(rbtp_sim *s, int32_t *nodes, int32_t *pids, int64_t capacity,
                                  int64_t *count) {
    if (!s || !count) return set_error(RBTP_E_ARG, "null argument");
    *count = 0;
    if (!s->particles) return RBTP_OK;
    HIP_TRY(hipSetDevice(s->device));
    std::vector<std::pair<int32_t, int32_t>> nf;
    RB_TRY(part_collect_new_fluid(s->pd, s->stream, nf));
    std::sort(nf.begin(), nf.end());
    *count = (int64_t)nf.size();
    if (nodes || pids) {
        if (capacity < *count) return set_error(RBTP_E_ARG, "new-fluid buffer too small");
        for (size_t i = 0; i < nf.size(); i++) {
            if (nodes) nodes[i] = nf[i].first;
            if (pids) pids[i] = nf[i].second;
        }
    }
    return RBTP_OK;
}

// ---------------------------------------------------------------------------------------------
// Diagnostics
// ---------------------------------------------------------------------------------------------

extern "C" int rbtp_nu(rbtp_sim *s, double out[7]) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->macro_valid) return set_error(RBTP_E_STATE, "run the last step with RBTP_STEP_MACRO_LAST first");
    HIP_TRY(hipSetDevice(s->device));
    const int32_t *owner = s->particles ? s->pd.owner[current_gen(s)] : nullptr;
    RB_TRY(diag_nu(s->c, s->d, s->dims, s->macro[1], s->macro[4], owner, s->step, s->stream, out));
    return RBTP_OK;
}

extern "C" int rbtp_nu_noalloc(rbtp_sim *s, double out[7]) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->macro_valid) return set_error(RBTP_E_STATE, "run the last step with RBTP_STEP_MACRO_LAST first");
    HIP_TRY(hipSetDevice(s->device));
    const int32_t *owner = s->particles ? s->pd.owner[s->pd.geo_cur] : nullptr;
    RB_TRY(diag_nu_noalloc(s->c, s->d, s->dims, s->macro[1], s->macro[4], owner, s->step,
                           s->stream, out));
    return RBTP_OK;
}

extern "C" int rbtp_means(rbtp_sim *s, double out[4]) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->macro_valid) return set_error(RBTP_E_STATE, "run the last step with RBTP_STEP_MACRO_LAST first");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(diag_means(s->dims, s->macro, s->particles ? &s->pd : nullptr, s->stream, out));
    return RBTP_OK;
}

extern "C" int rbtp_bandwidth_probe(size_t bytes, int reps, double *gbs) {
    if (!gbs) return set_error(RBTP_E_ARG, "null argument");
    return diag_copy_probe(bytes, reps, gbs);
}

extern "C" int rbtp_stencil_probe(int lx, int ly, int lz, int reps, double *gbs) {
    if (!gbs) return set_error(RBTP_E_ARG, "null argument");
    return lattice_stencil_probe(lx, ly, lz, reps, gbs);
}

extern "C" int rbtp_arith_probe(int op, const double *a, const double *b, int64_t n,
                                double *out) {
    if (n < 0) return set_error(RBTP_E_ARG, "negative operand count");
    return lattice_arith_probe(op, a, b, (size_t)n, out);
}

extern "C" int rbtp_streaming_probe(int scheme, int lx, int ly, int lz, int reps, double *gbs) {
    if (!gbs) return set_error(RBTP_E_ARG, "null argument");
    return lattice_streaming_probe(scheme, lx, ly, lz, reps, gbs);
}

extern "C" int rbtp_kinetic_energy(rbtp_sim *s, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    if (!s->macro_valid) return set_error(RBTP_E_STATE, "run the last step with RBTP_STEP_MACRO_LAST first");
    HIP_TRY(hipSetDevice(s->device));
    RB_TRY(diag_kinetic_energy(s->dims, s->macro,
                               s->particles ? s->pd.owner[s->pd.geo_cur] : nullptr,
                               s->particles ? s->pd.owner[1 - s->pd.geo_cur] : nullptr, s->stream,
                               out));
    return RBTP_OK;
}


// ---------------------------------------------------------------------------------------------
// Output fields and files
// ---------------------------------------------------------------------------------------------

static OutK out_consts(const rbtp_sim *s) {
    OutK k;
    k.tau = s->d.tau;
    k.tauc = s->d.tauc;
    k.visc = s->c.visc;
    k.diff = s->d.diff;
    k.t0 = s->c.t_ref;
    k.rho0 = s->c.rho0;
    k.gshift = s->gshift;
    return k;
}

// The seven %09d.dat fields on the device (caller frees *dev).
static int output_fields_dev(rbtp_sim *s, double **dev) {
    if (!s->macro_valid)
        return set_error(RBTP_E_STATE, "run the last step with RBTP_STEP_MACRO_LAST first");
    const size_t n = (size_t)s->dims.n;
    HIP_TRY(hipMalloc(dev, 7 * n * sizeof(double)));
    const int r = diag_output_fields(s->prec, s->f[s->cur], s->g[s->cur], s->fx, s->macro,
                                     s->particles ? &s->pd : nullptr, out_consts(s), s->dims,
                                     s->stream, *dev);
    if (r != RBTP_OK) {
        (void)hipFree(*dev);
        *dev = nullptr;
    }
    return r;
}

extern "C" int rbtp_get_output_fields(rbtp_sim *s, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    double *dev = nullptr;
    RB_TRY(output_fields_dev(s, &dev));
    const hipError_t e = copy_to_host(out, dev, 7 * (size_t)s->dims.n * sizeof(double), s->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return set_error(RBTP_E_HIP, "rbtp_get_output_fields: %s", hipGetErrorString(e));
    return RBTP_OK;
}

extern "C" int rbtp_profile(rbtp_sim *s, double *out) {
    if (!s || !out) return set_error(RBTP_E_ARG, "null argument");
    HIP_TRY(hipSetDevice(s->device));
    double *dev = nullptr;
    RB_TRY(output_fields_dev(s, &dev));
    const int r = diag_profile(dev, s->dims, s->stream, out);
    (void)hipFree(dev);
    return r;
}

static int write_all(FILE *fp, const double *p, size_t count, const char *path) {
    if (fwrite(p, sizeof(double), count, fp) != count)
        return set_error(RBTP_E_IO, "short write to %s", path);
    return RBTP_OK;
}

extern "C" int rbtp_write_fg(rbtp_sim *s, const char *path) {
    if (!s || !path) return set_error(RBTP_E_ARG, "null argument");
    const size_t n = 27 * (size_t)s->dims.n;
    std::vector<double> buf(n);
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_error(RBTP_E_IO, "cannot open %s", path);
    int r = rbtp_get_populations(s, RBTP_POP_F, buf.data());
    if (r == RBTP_OK) r = write_all(fp, buf.data(), n, path);
    if (r == RBTP_OK) r = rbtp_get_populations(s, RBTP_POP_G, buf.data());
    if (r == RBTP_OK) r = write_all(fp, buf.data(), n, path);
    if (fclose(fp) != 0 && r == RBTP_OK) r = set_error(RBTP_E_IO, "close failed on %s", path);
    return r;
}

extern "C" int rbtp_write_flow(rbtp_sim *s, const char *path) {
    if (!s || !path) return set_error(RBTP_E_ARG, "null argument");
    const size_t n = (size_t)s->dims.n;
    std::vector<double> buf(7 * n);
    RB_TRY(rbtp_get_output_fields(s, buf.data()));
    FILE *fp = fopen(path, "wb");
    if (!fp) return set_error(RBTP_E_IO, "cannot open %s", path);
    int r = write_all(fp, buf.data(), 7 * n, path);
    if (fclose(fp) != 0 && r == RBTP_OK) r = set_error(RBTP_E_IO, "close failed on %s", path);
    return r;
}

// Column order of pos_%09d.txt (output_part.cu:101-113).
static const int kPosColumns[33] = {
    RBTP_P_POS_X, RBTP_P_POS_Y, RBTP_P_POS_Z, RBTP_P_VEL_X, RBTP_P_VEL_Y, RBTP_P_VEL_Z,
    RBTP_P_FORCE_X, RBTP_P_FORCE_Y, RBTP_P_FORCE_Z, RBTP_P_FORCE_PREV_X, RBTP_P_FORCE_PREV_Y,
    RBTP_P_FORCE_PREV_Z, RBTP_P_FORCE_PREV2_X, RBTP_P_FORCE_PREV2_Y, RBTP_P_FORCE_PREV2_Z,
    RBTP_P_THETA_X, RBTP_P_THETA_Y, RBTP_P_THETA_Z, RBTP_P_OMEGA_X, RBTP_P_OMEGA_Y,
    RBTP_P_OMEGA_Z, RBTP_P_TORQUE_X, RBTP_P_TORQUE_Y, RBTP_P_TORQUE_Z, RBTP_P_TORQUE_PREV_X,
    RBTP_P_TORQUE_PREV_Y, RBTP_P_TORQUE_PREV_Z, RBTP_P_TORQUE_PREV2_X, RBTP_P_TORQUE_PREV2_Y,
    RBTP_P_TORQUE_PREV2_Z, RBTP_P_TEMP, RBTP_P_HEAT, RBTP_P_HEAT_PREV};

extern "C" int rbtp_write_particles(rbtp_sim *s, const char *path) {
    if (!s || !path) return set_error(RBTP_E_ARG, "null argument");
    if (!s->particles) return RBTP_OK;
    const int np = s->pd.np;
    HIP_TRY(hipSetDevice(s->device));
    HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<double> h((size_t)RBTP_P_COUNT * np);
    HIP_TRY(copy_to_host(h.data(), s->pd.base, h.size() * sizeof(double), s->stream));
    FILE *fp = fopen(path, "w");
    if (!fp) return set_error(RBTP_E_IO, "cannot open %s", path);
    for (int pid = 0; pid < np; ++pid) {
        fprintf(fp, "%d", pid);
        for (int c = 0; c < 33; c++)
            fprintf(fp, c == 32 ? " %14.6e\n" : " %14.6e", h[(size_t)kPosColumns[c] * np + pid]);
    }
    if (fclose(fp) != 0) return set_error(RBTP_E_IO, "close failed on %s", path);
    return RBTP_OK;
}

// The particle loop of load_particles_from_restart (initial.cu:57-192): np lines of pid and 33
// values in the pos_%09d.txt column order.  The reference exits on a short or unreadable line
// and indexes its arrays with the pid as read; here both, and a pid outside [0, np), are
// RBTP_E_IO.  Host only.
extern "C" int rbtp_read_particle_file(const char *path, int np, double *out) {
    if (!path || (np > 0 && !out)) return set_error(RBTP_E_ARG, "null argument");
    if (np < 0) return set_error(RBTP_E_ARG, "negative particle count %d", np);
    FILE *fp = fopen(path, "r");
    if (!fp) return set_error(RBTP_E_IO, "Failed to open restart particle file %s", path);
    std::fill(out, out + (size_t)RBTP_P_COUNT * np, 0.0);
    for (int line = 0; line < np; line++) {
        int pid = -1;
        double v[33];
        bool ok = fscanf(fp, "%d", &pid) == 1;
        for (int c = 0; ok && c < 33; c++) ok = fscanf(fp, "%lf", &v[c]) == 1;
        if (!ok || pid < 0 || pid >= np) {
            fclose(fp);
            return set_error(RBTP_E_IO, "Failed to read particle restart data from %s (line %d)",
                             path, line + 1);
        }
        for (int c = 0; c < 33; c++) out[(size_t)kPosColumns[c] * np + pid] = v[c];
    }
    fclose(fp);
    return RBTP_OK;
}

// generate_initial_positions (pinit.cu:97-144) alone, on the host: the positions rbtp_init
// places when it is given none.
extern "C" int rbtp_place_particles(const rbtp_case *c, double *out_3p) {
    if (!c || !out_3p) return set_error(RBTP_E_ARG, "null argument");
    rbtp_derived d;
    rbtp_case_derive(c, &d);
    const int np = d.npart;
    std::vector<double> X(np), Y(np), Z(np);
    RB_TRY(place_particles(*c, d, X, Y, Z));
    std::copy(X.begin(), X.end(), out_3p);
    std::copy(Y.begin(), Y.end(), out_3p + np);
    std::copy(Z.begin(), Z.end(), out_3p + 2 * (size_t)np);
    return RBTP_OK;
}

// load_restart_state (initial.cu:194-236).
extern "C" int rbtp_restart_load(rbtp_sim *s, const char *dir, int step) {
    if (!s || !dir) return set_error(RBTP_E_ARG, "null argument");
    if (step < 0) return set_error(RBTP_E_ARG, "negative restart step");
    if (!s->initialised) {
        // allocate the state the way rbtp_init would, then replace it
        RB_TRY(rbtp_init(s, nullptr, nullptr, nullptr));
    }
    s->st_active = false;  // the loaded state replaces any staged one
    HIP_TRY(hipSetDevice(s->device));
    const size_t n = (size_t)s->dims.n;
    char path[4096];
    // load_distributions_from_restart (initial.cu:7-30)
    snprintf(path, sizeof(path), "%s/fg%09d.dat", dir, step);
    std::vector<double> f(27 * n), g(27 * n);
    {
        FILE *fp = fopen(path, "rb");
        if (!fp) return set_error(RBTP_E_IO, "Failed to open restart distribution file %s", path);
        const size_t rf = fread(f.data(), sizeof(double), f.size(), fp);
        const size_t rg = fread(g.data(), sizeof(double), g.size(), fp);
        fclose(fp);
        if (rf != f.size() || rg != g.size())
            return set_error(RBTP_E_IO, "Restart distribution file %s is corrupted (read %zu/%zu for f, %zu/%zu for g)",
                             path, rf, f.size(), rg, g.size());
    }
    // load_particles_from_restart (initial.cu:57-192)
    if (s->particles) {
        const int np = s->pd.np;
        snprintf(path, sizeof(path), "%s/pos/pos_%09d.txt", dir, step);
        std::vector<double> h((size_t)RBTP_P_COUNT * np, 0.0);
        RB_TRY(rbtp_read_particle_file(path, np, h.data()));
        RB_TRY(part_load_state(s->pd, h.data(), s->dims, s->stream));
    }
    // the encodings restart from the conduction profile, as after rbtp_init
    RB_TRY(reset_shift(s));
    RB_TRY(rbtp_set_populations(s, f.data(), g.data()));
    // load_flow_fields (initial.cu:32-55) and forcing on the loaded phi (initial.cu:207)
    snprintf(path, sizeof(path), "%s/%09d.dat", dir, step);
    {
        std::vector<double> fields(7 * n);
        FILE *fp = fopen(path, "rb");
        if (!fp) return set_error(RBTP_E_IO, "Failed to open flow restart file %s", path);
        const size_t rd = fread(fields.data(), sizeof(double), fields.size(), fp);
        fclose(fp);
        if (rd != fields.size())
            return set_error(RBTP_E_IO, "flow restart file %s is short (%zu of %zu values)", path, rd, fields.size());
        for (int k = 0; k < 5; k++)
            HIP_TRY(copy_to_device(s->macro[k], fields.data() + k * n, n * sizeof(double), s->stream));
        HIP_TRY(dispatch(s->prec, [&](auto *tag) {
            using S = std::remove_pointer_t<decltype(tag)>;
            return launch_forcing_field<S>((typename Enc<S>::A *)s->fx,
                                           (typename Enc<S>::A *)s->rho_wall, s->macro[0],
                                           s->macro[4], s->k.gb, s->c.t_ref, s->dims, s->stream);
        }));
        HIP_TRY(hipStreamSynchronize(s->stream));
    }
    s->poststream = true;
    s->macro_valid = true;
    s->step = step;
    return RBTP_OK;
}


// ---------------------------------------------------------------------------------------------
// Serving: a batch of independent cavities advanced together
// ---------------------------------------------------------------------------------------------
//
// Each admitted request is an rbtp_sim (a slot) whose state stays resident in HBM.  All slots
// share one stream; one batched, persistent lattice kernel advances every slot's lattice per
// step (tiles numbered across slots), followed by each slot's special-node and particle work.
// A step sequence of fixed length is captured once per batch composition into a hipGraph and
// replayed; admissions and releases happen between replays and invalidate the graph.  The
// lattice kernel's tile size and persistent grid come from rbtp_batch_tune, which times the
// candidates on the current composition without changing any slot's state.

struct rbtp_batch {
    int prec;
    bool strict;
    int max_slots;
    double hbm_fraction;
    int device;
    int num_cu;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    std::vector<rbtp_sim *> slot;
    std::vector<size_t> slot_bytes;
    size_t used_bytes;
    void *args_dev[4];  // StepArgs<S>[max_slots] for (parity, macro)
    int *tile_start_dev;
    std::vector<int> active;
    std::vector<int> cur_at_build;  // each active slot's buffer parity when the tables were built
    std::vector<int> geo_at_build;  // each active slot's geometry generation then (0 without particles)
    std::vector<int> part_slots;    // positions in `active` of the slots with particles
    PartTables ptab;                // particle-phase tables over part_slots, per phase
    int total_tiles;
    int block, grid;
    bool tables_dirty;
    int phase;
    bool use_graph;
    int graph_steps;
    hipGraph_t graph;
    hipGraphExec_t exec;
    bool graph_valid;
    bool capturing;  // batch_step_eager is being recorded into the graph
    double last_ms;
    int64_t graph_launches;
};

static size_t args_size(int prec) {
    switch (prec) {
        case RBTP_PREC_FP64: return sizeof(StepArgs<double>);
        case RBTP_PREC_FP32: return sizeof(StepArgs<float>);
        case RBTP_PREC_FP16S: return sizeof(StepArgs<__half>);
        default: return sizeof(StepArgs<__hip_bfloat16>);
    }
}

// Device bytes one slot allocates (create_impl, part_alloc).
static size_t sim_device_bytes(const rbtp_case &c, const rbtp_derived &d) {
    const size_t n = (size_t)c.lx * c.ly * c.lz;
    size_t b = 2 * pop_buffer_len(n) * storage_bytes(c.precision) + n * aux_bytes(c.precision) +
               2 * (size_t)c.ly * c.lz * aux_bytes(c.precision) + 5 * n * sizeof(double);
    if (c.activate_particles && d.npart > 0) {
        const size_t np = (size_t)d.npart;
        if (c.precision != RBTP_PREC_FP64) b += n * sizeof(float);  // rho_full
        const long long w = (long long)std::ceil(2.0 * d.particle_radius) + 1;
        const long long cap = special_capacity((int)np, (int)w, (long long)n);
        b += (RBTP_P_COUNT + OLD_COUNT) * np * sizeof(double) + 3 * n * sizeof(int32_t) +
             12 * np * sizeof(int32_t) +
             (size_t)cap * (2 * 27 * aux_bytes(c.precision) + sizeof(double) + 2 * sizeof(int32_t)) +
             np * (size_t)link_stride((int)w) * (7 * sizeof(double) + 1);  // shares + bits (rounded up)
    }
    return b;
}

static void batch_invalidate_graph(rbtp_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->graph_valid = false;
}

extern "C" void rbtp_batch_destroy(rbtp_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_invalidate_graph(b);
    for (rbtp_sim *s : b->slot)
        if (s) rbtp_destroy(s);
    for (void *p : b->args_dev) (void)hipFree(p);
    part_tables_free(b->ptab);
    (void)hipFree(b->tile_start_dev);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int rbtp_batch_create(int precision, int max_slots, double hbm_fraction, rbtp_batch **out) {
    if (!out) return set_error(RBTP_E_ARG, "null argument");
    *out = nullptr;
    if (precision < RBTP_PREC_FP64 || precision > RBTP_PREC_BF16S) return set_error(RBTP_E_ARG, "unknown precision %d", precision);
    if (max_slots < 1 || max_slots > 4096) return set_error(RBTP_E_ARG, "max_slots must be in [1, 4096]");
    if (!(hbm_fraction > 0.0 && hbm_fraction <= 1.0)) return set_error(RBTP_E_ARG, "hbm_fraction must be in (0, 1]");
    rbtp_batch *b = new rbtp_batch();
    b->prec = precision;
    b->strict = precision == RBTP_PREC_FP64;
    b->max_slots = max_slots;
    b->hbm_fraction = hbm_fraction;
    b->slot.assign(max_slots, nullptr);
    b->slot_bytes.assign(max_slots, 0);
    b->block = 256;
    b->use_graph = true;
    b->graph_steps = 20;
    b->tables_dirty = true;
    auto fail = [&](hipError_t e, const char *what) {
        const int r = set_error(RBTP_E_HIP, "%s: %s", what, hipGetErrorString(e));
        std::string msg = g_last_error;
        rbtp_batch_destroy(b);
        g_last_error = msg;
        return r;
    };
    hipError_t e = hipGetDevice(&b->device);
    if (e != hipSuccess) return fail(e, "hipGetDevice");
    e = hipDeviceGetAttribute(&b->num_cu, hipDeviceAttributeMultiprocessorCount, b->device);
    if (e != hipSuccess) return fail(e, "hipDeviceGetAttribute");
    b->grid = 4 * b->num_cu;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return fail(e, "hipEventCreate");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return fail(e, "hipEventCreate");
    for (void *&p : b->args_dev)
        if ((e = hipMalloc(&p, args_size(precision) * max_slots)) != hipSuccess) return fail(e, "hipMalloc");
    if ((e = hipMalloc(&b->tile_start_dev, sizeof(int) * (max_slots + 1))) != hipSuccess) return fail(e, "hipMalloc");
    *out = b;
    return RBTP_OK;
}

extern "C" int rbtp_batch_admit(rbtp_batch *b, const rbtp_case *c, const double *temperature,
                                const double *ppos, const double *ptemp, int *slot_out) {
    if (!b || !c || !slot_out) return set_error(RBTP_E_ARG, "null argument");
    *slot_out = -1;
    if (c->precision != b->prec)
        return set_error(RBTP_E_ARG, "case precision %d differs from the batch's %d", c->precision, b->prec);
    RB_TRY(validate_case(c));
    int k = -1;
    for (int i = 0; i < b->max_slots; i++)
        if (!b->slot[i]) { k = i; break; }
    if (k < 0) return set_error(RBTP_E_CAPACITY, "no free slot (%d in use)", b->max_slots);
    rbtp_derived d;
    rbtp_case_derive(c, &d);
    const size_t need = sim_device_bytes(*c, d);
    HIP_TRY(hipSetDevice(b->device));
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    const size_t reserve = (size_t)1 << 30;
    if (b->used_bytes + need > (size_t)(b->hbm_fraction * (double)total_b) || need + reserve > free_b)
        return set_error(RBTP_E_CAPACITY, "not enough HBM for %zu bytes (batch uses %zu, %zu free)",
                         need, b->used_bytes, free_b);
    HIP_TRY(hipStreamSynchronize(b->stream));
    rbtp_sim *s = nullptr;
    RB_TRY(create_on_stream(c, &s, b->stream));
    s->batch_slot = true;
    int r = rbtp_init(s, temperature, ppos, ptemp);
    if (r == RBTP_OK) r = step_impl(s, 1, 0, true);  // first step reads the post-stream start
    if (r != RBTP_OK) {
        std::string msg = g_last_error;
        rbtp_destroy(s);
        g_last_error = msg;
        return r;
    }
    b->slot[k] = s;
    b->slot_bytes[k] = need;
    b->used_bytes += need;
    b->tables_dirty = true;
    batch_invalidate_graph(b);
    *slot_out = k;
    return RBTP_OK;
}

extern "C" int rbtp_batch_release(rbtp_batch *b, int k) {
    if (!b) return set_error(RBTP_E_ARG, "null handle");
    if (k < 0 || k >= b->max_slots || !b->slot[k]) return set_error(RBTP_E_ARG, "slot %d is not in use", k);
    HIP_TRY(hipStreamSynchronize(b->stream));
    rbtp_destroy(b->slot[k]);
    b->slot[k] = nullptr;
    b->used_bytes -= b->slot_bytes[k];
    b->slot_bytes[k] = 0;
    b->tables_dirty = true;
    batch_invalidate_graph(b);
    return RBTP_OK;
}

extern "C" int rbtp_batch_slot(rbtp_batch *b, int k, rbtp_sim **out) {
    if (!b || !out) return set_error(RBTP_E_ARG, "null argument");
    if (k < 0 || k >= b->max_slots || !b->slot[k]) return set_error(RBTP_E_ARG, "slot %d is not in use", k);
    *out = b->slot[k];
    return RBTP_OK;
}

// Device tables of the batched kernel for the current composition and tile size.
static int batch_build_tables(rbtp_batch *b) {
    b->active.clear();
    for (int i = 0; i < b->max_slots; i++)
        if (b->slot[i]) b->active.push_back(i);
    const int na = (int)b->active.size();
    std::vector<int> tiles(na + 1, 0);
    for (int j = 0; j < na; j++) {
        const rbtp_sim *s = b->slot[b->active[j]];
        tiles[j + 1] = tiles[j] + (s->dims.n + b->block - 1) / b->block;
    }
    b->total_tiles = tiles[na];
    b->cur_at_build.clear();
    b->geo_at_build.clear();
    b->part_slots.clear();
    for (int j = 0; j < na; j++) {
        const rbtp_sim *s = b->slot[b->active[j]];
        b->cur_at_build.push_back(s->cur);
        b->geo_at_build.push_back(s->particles ? s->pd.geo_cur : 0);
        if (s->particles) b->part_slots.push_back(j);
    }
    // kernels already queued may still read the old tables
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (na > 0) HIP_TRY(hipMemcpy(b->tile_start_dev, tiles.data(), sizeof(int) * na, hipMemcpyHostToDevice));
    const int nps = (int)b->part_slots.size();
    HIP_TRY(dispatch(b->prec, [&](auto *tag) -> hipError_t {
        using S = std::remove_pointer_t<decltype(tag)>;
        std::vector<StepArgs<S>> h(na > 0 ? na : 1);
        for (int par = 0; par < 2; par++)
            for (int mac = 0; mac < 2; mac++) {
                for (int j = 0; j < na; j++) {
                    rbtp_sim *s = b->slot[b->active[j]];
                    if (par) s->cur = 1 - s->cur;
                    h[j] = step_args<S>(s, mac != 0);
                    if (par) s->cur = 1 - s->cur;
                }
                if (na > 0) {
                    hipError_t e = hipMemcpy(b->args_dev[par * 2 + mac], h.data(), sizeof(StepArgs<S>) * na, hipMemcpyHostToDevice);
                    if (e != hipSuccess) return e;
                }
            }
        return hipSuccess;
    }));
    // Particle phase of a step taken at batch phase ph: populations in buffer cur ^ ph ^ 1
    // (after the lattice kernel's flip), IBB geometry geo_cur ^ ph.
    std::vector<PartSlot> pv[2];
    for (int ph = 0; ph < 2; ph++)
        for (int q = 0; q < nps; q++) {
            const rbtp_sim *s = b->slot[b->active[b->part_slots[q]]];
            pv[ph].push_back(part_slot(s, s->cur ^ ph ^ 1, s->pd.geo_cur ^ ph));
        }
    RB_TRY(part_tables_build(b->ptab, pv, b->stream));
    b->phase = 0;
    b->tables_dirty = false;
    batch_invalidate_graph(b);
    return RBTP_OK;
}

static hipError_t batch_lattice(rbtp_batch *b, bool macro, int block, int grid) {
    const int na = (int)b->active.size();
    return dispatch(b->prec, [&](auto *tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        return launch_lattice_batched<S>((const StepArgs<S> *)b->args_dev[b->phase * 2 + (macro ? 1 : 0)],
                                         b->tile_start_dev, na, b->total_tiles, block, grid,
                                         b->strict, b->stream);
    });
}

// One step of every slot: the batched lattice kernel (which collides each slot's special nodes
// from its special-node list), then one launch of each particle kernel for all slots with
// particles.
static int batch_step_eager(rbtp_batch *b, bool macro) {
    HIP_TRY(batch_lattice(b, macro, b->block, b->grid));
    const int nps = (int)b->part_slots.size();
    const PartGrid &pg = b->ptab.g[b->phase];
    for (int k : b->active) {
        rbtp_sim *s = b->slot[k];
        s->cur = 1 - s->cur;
        s->poststream = false;
    }
    if (nps > 0) {
        RB_TRY(part_phase(pg, b->prec, b->stream));
        for (int q : b->part_slots) {
            rbtp_sim *s = b->slot[b->active[q]];
            s->pd.geo_cur ^= 1;
            s->pd.special_pending = true;
        }
    }
    for (int k : b->active) {
        rbtp_sim *s = b->slot[k];
        s->step++;
        // each slot's own shift refresh, at the same steps and in the same place of the step as
        // in rbtp_step (after the particle phase); never inside a captured graph, whose replays
        // rbtp_batch_step keeps clear of every slot's refresh steps
        if (!b->capturing && s->shift_interval > 0 && s->step % s->shift_interval == 0)
            HIP_TRY(dispatch(s->prec, [&](auto *tag) {
                using S = std::remove_pointer_t<decltype(tag)>;
                return launch_shift_refresh<S>((S *)s->g[s->cur], s->dims, s->gshift,
                                               s->shift_partial, s->shift_delta, b->stream);
            }));
    }
    b->phase ^= 1;
    return RBTP_OK;
}

// Batched steps from now until the first one after which some slot refreshes its shift profile
// (that step included); INT_MAX when no slot refreshes.
static int steps_to_refresh(const rbtp_batch *b) {
    int room = INT_MAX;
    for (int k : b->active) {
        const rbtp_sim *s = b->slot[k];
        if (s->shift_interval > 0) room = std::min(room, s->shift_interval - s->step % s->shift_interval);
    }
    return room;
}

static int batch_capture(rbtp_batch *b) {
    batch_invalidate_graph(b);
    std::vector<int> steps;
    for (int k : b->active) steps.push_back(b->slot[k]->step);
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int r = RBTP_OK;
    b->capturing = true;
    for (int j = 0; j < b->graph_steps && r == RBTP_OK; j++) r = batch_step_eager(b, false);
    b->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->stream, &g);
    for (size_t j = 0; j < b->active.size(); j++) b->slot[b->active[j]]->step = steps[j];
    if (r != RBTP_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return set_error(RBTP_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->graph_valid = true;
    return RBTP_OK;
}

extern "C" int rbtp_batch_step(rbtp_batch *b, int nsteps, unsigned flags) {
    if (!b) return set_error(RBTP_E_ARG, "null handle");
    if (nsteps < 0) return set_error(RBTP_E_ARG, "nsteps < 0");
    HIP_TRY(hipSetDevice(b->device));
    if (!b->tables_dirty)
        for (size_t j = 0; j < b->active.size(); j++) {
            const rbtp_sim *s = b->slot[b->active[j]];
            if (s->cur != (b->cur_at_build[j] ^ b->phase)) b->tables_dirty = true;
            if (s->particles && s->pd.geo_cur != (b->geo_at_build[j] ^ b->phase)) b->tables_dirty = true;
        }
    if (b->tables_dirty) RB_TRY(batch_build_tables(b));
    if (b->active.empty() || nsteps == 0) return RBTP_OK;
    const bool macro_last = (flags & RBTP_STEP_MACRO_LAST) != 0;
    int body = macro_last ? nsteps - 1 : nsteps;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    // Graph replays (captured from phase 0, an even number of steps) where a whole one fits and
    // no slot refreshes its shift profile inside it; single eager steps elsewhere.
    while (body > 0) {
        if (b->use_graph && b->phase == 0 && body >= b->graph_steps &&
            steps_to_refresh(b) > b->graph_steps) {
            if (!b->graph_valid) RB_TRY(batch_capture(b));
            HIP_TRY(hipGraphLaunch(b->exec, b->stream));
            b->graph_launches++;
            for (int k : b->active) b->slot[k]->step += b->graph_steps;
            body -= b->graph_steps;
        } else {
            RB_TRY(batch_step_eager(b, false));
            body--;
        }
    }
    if (macro_last) RB_TRY(batch_step_eager(b, true));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
    HIP_TRY(hipEventSynchronize(b->ev1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->last_ms = ms;
    for (int k : b->active) {
        rbtp_sim *s = b->slot[k];
        s->macro_valid = macro_last;
        if (s->particles) RB_TRY(part_check_errors(s->pd));
    }
    return RBTP_OK;
}

// Times the batched lattice kernel on the current composition for every tile size (64, 128,
// 256 nodes) and persistent grid (1, 2, 4, 8, 16 workgroups per CU, or one per tile) and keeps
// the fastest.  F_x and the wall rho (updated in place by the kernel) are saved and restored,
// and the kernel's other outputs are rewritten by the next step, so no slot's state changes.
extern "C" int rbtp_batch_tune(rbtp_batch *b, int *block_out, int *grid_out, double *ms_out) {
    if (!b) return set_error(RBTP_E_ARG, "null handle");
    HIP_TRY(hipSetDevice(b->device));
    if (b->tables_dirty) RB_TRY(batch_build_tables(b));
    if (b->active.empty()) return set_error(RBTP_E_STATE, "no active slot to tune on");
    HIP_TRY(hipStreamSynchronize(b->stream));
    std::vector<void *> saved;
    for (int k : b->active) {
        rbtp_sim *s = b->slot[k];
        const size_t fxb = (size_t)s->dims.n * aux_bytes(s->prec);
        const size_t rwb = 2 * (size_t)s->c.ly * s->c.lz * aux_bytes(s->prec);
        void *p = nullptr;
        HIP_TRY(hipMalloc(&p, fxb + rwb));
        saved.push_back(p);
        HIP_TRY(hipMemcpyAsync(p, s->fx, fxb, hipMemcpyDeviceToDevice, b->stream));
        HIP_TRY(hipMemcpyAsync((char *)p + fxb, s->rho_wall, rwb, hipMemcpyDeviceToDevice, b->stream));
    }
    const int phase0 = b->phase;
    int best_block = b->block, best_grid = b->grid;
    double best = 1e30;
    int rc = RBTP_OK;
    const int blocks[3] = {64, 128, 256};
    for (int bi = 0; bi < 3 && rc == RBTP_OK; bi++) {
        b->block = blocks[bi];
        rc = batch_build_tables(b);
        if (rc != RBTP_OK) break;
        const int grids[6] = {b->num_cu, 2 * b->num_cu, 4 * b->num_cu, 8 * b->num_cu,
                              16 * b->num_cu, b->total_tiles};
        for (int gi = 0; gi < 6 && rc == RBTP_OK; gi++) {
            const int g = std::min(grids[gi], b->total_tiles);
            hipError_t e = batch_lattice(b, false, b->block, g);
            if (e == hipSuccess) e = hipEventRecord(b->ev0, b->stream);
            for (int rep = 0; rep < 3 && e == hipSuccess; rep++) e = batch_lattice(b, false, b->block, g);
            if (e == hipSuccess) e = hipEventRecord(b->ev1, b->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b->ev1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, b->ev0, b->ev1);
            if (e != hipSuccess) {
                rc = set_error(RBTP_E_HIP, "tuning launch: %s", hipGetErrorString(e));
                break;
            }
            if (ms / 3.0 < best) {
                best = ms / 3.0;
                best_block = b->block;
                best_grid = g;
            }
        }
    }
    for (size_t j = 0; j < b->active.size(); j++) {
        rbtp_sim *s = b->slot[b->active[j]];
        const size_t fxb = (size_t)s->dims.n * aux_bytes(s->prec);
        const size_t rwb = 2 * (size_t)s->c.ly * s->c.lz * aux_bytes(s->prec);
        (void)hipMemcpyAsync(s->fx, saved[j], fxb, hipMemcpyDeviceToDevice, b->stream);
        (void)hipMemcpyAsync(s->rho_wall, (char *)saved[j] + fxb, rwb, hipMemcpyDeviceToDevice, b->stream);
        s->macro_valid = false;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (void *p : saved) (void)hipFree(p);
    if (rc != RBTP_OK) return rc;
    b->block = best_block;
    b->grid = best_grid;
    RB_TRY(batch_build_tables(b));
    (void)phase0;
    if (block_out) *block_out = best_block;
    if (grid_out) *grid_out = best_grid;
    if (ms_out) *ms_out = best;
    return RBTP_OK;
}

extern "C" int rbtp_batch_set_launch(rbtp_batch *b, int block, int grid, int graph_steps) {
    if (!b) return set_error(RBTP_E_ARG, "null handle");
    if (block != 64 && block != 128 && block != 256) return set_error(RBTP_E_ARG, "block must be 64, 128 or 256");
    if (grid < 1) return set_error(RBTP_E_ARG, "grid must be >= 1");
    if (graph_steps < 0 || graph_steps % 2) return set_error(RBTP_E_ARG, "graph_steps must be even and >= 0 (0 disables graphs)");
    b->block = block;
    b->grid = grid;
    b->use_graph = graph_steps > 0;
    b->graph_steps = graph_steps > 0 ? graph_steps : 2;
    b->tables_dirty = true;
    batch_invalidate_graph(b);
    return RBTP_OK;
}

extern "C" int rbtp_batch_info(rbtp_batch *b, int *active, int64_t *used_bytes, int64_t *free_bytes,
                               double *last_ms, int *block, int *grid) {
    if (!b) return set_error(RBTP_E_ARG, "null handle");
    int na = 0;
    for (rbtp_sim *s : b->slot) na += s != nullptr;
    if (active) *active = na;
    if (used_bytes) *used_bytes = (int64_t)b->used_bytes;
    if (free_bytes) {
        size_t f = 0, t = 0;
        HIP_TRY(hipMemGetInfo(&f, &t));
        *free_bytes = (int64_t)f;
    }
    if (last_ms) *last_ms = b->last_ms;
    if (block) *block = b->block;
    if (grid) *grid = b->grid;
    return RBTP_OK;
}

// Error helper for the other translation units.
namespace rbtp {
int report(int code, const char *msg) { return set_error(code, "%s", msg); }
int report_hip(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? RBTP_E_NOMEM : RBTP_E_HIP, "%s: %s", what,
                     hipGetErrorString(e));
}
}  // namespace rbtp
