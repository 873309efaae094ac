// oracle.cpp — FP64 host restatement of the reference kernels (TEST INFRASTRUCTURE ONLY).
//
// This file is the checker for the MI355X library: only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  The product never links or calls it.
//
// Every function restates one reference kernel or host routine (cited file:line, relative to the
// reference repository root) with the same arithmetic order and branch precedence, compiled
// with -ffp-contract=off so that no multiply-add is fused.  Where the reference leaves an order
// unspecified (FP64 atomicAdd into per-particle accumulators: ibb.cu:182-187,204 and
// pforces.cu:32-37) the oracle fixes one deterministic order and documents it at the function:
//   * pair repulsion: partners j in ascending order, then the six walls, then the link sum;
//   * link sums (force, torque, heat): per particle, each node's links are summed first (from 0,
//     in increasing direction) into the node's share; the shares are dealt round-robin to 64
//     lane sums by the node's position k in the particle's link box (k mod 64, added from 0 in
//     increasing k), and the 64 lane sums are combined by a pairwise tree (stride 32, 16, ...,
//     1).  The link box is the particle's marking box floor(p -/+ R - 0.5) (links.cu:164-177)
//     grown by one node per side, clamped to the domain.
//
// Parity status of each function (how it is pinned independently of this restatement) is listed
// in the header of oracle/oracle.h.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

// Planted mistakes.  ORACLE_MUTANT = k > 0 builds the oracle with mistake k of the list below;
// tests/test_oracle_mutants.py builds each one and shows that the pin meant to catch it fails.
// Every normal build has ORACLE_MUTANT = 0, and every MUT(k) is then a constant false.
//    1-4  bounce_back_particles: x and x - c swapped (q < 1/2), i and opp swapped (q < 1/2),
//         i and opp swapped (q >= 1/2), the momentum term dropped
//    5-8  bounce_back_thermal: x and x - c swapped (q <= 1/2), i and opp swapped (q > 1/2),
//         the (1 + rho) factor dropped, ibnode in place of ibnode_prev
//    9-11 new-fluid predicate: < for <=, the current owner for the previous one, the previous
//         and current flags swapped
//   12-14 output_nu: the previous flags, the wall sums normalised by the fluid count, the
//         actual solid fraction in place of PARTICLE_FRACTION
//   15-17 output fields: compute_Ked without the force term, compute_ted's denominator with
//         rho + 1, the particle-node overwrite without the rotation
//      18 macrovar: u without F_old/2
//   19-20 collision_BGK: the Guo term's (c_x - u_x) reversed, its (1 - 1/(2 tau)) left out
//   21-23 collision_BGK_scalar: the lambda0 term of geqnew_0 with + for -, grad_denom with
//         rho + 1, the 3 phi c.F source term dropped
//      24 streaming: a wall returns population i instead of opp(i)
//   25-27 streaming_scalar: the x walls without (1 + rho), the hot and cold walls swapped, the
//         x walls before the z and y walls
//   28-29 geometry: compute_link_q's larger root, the highest pid keeping a shared node
//   30-31 momentum exchange with (c_opp + u_w), heat with g*_i + g_opp
//      32 repulsion: the gap measured from one radius
//   33-34 update_particles: force history weights 1, 1, 0, buoyancy with the wrong sign
//   35-37 refill: e.u divided by RT in the fallback, first-order samples only, F kept
//      38 init_g: the lambda0 term of g_0 with + for -
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif
#define MUT(k) (ORACLE_MUTANT == (k))

namespace {

constexpr int NPOP = 27;

// D3Q27 tables (initial.cu:276-287).
const int CIX[NPOP] = {0, 1, 0, -1, 0, 0, 0, 1, -1, -1, 1, 1, 0, -1, 0, 1, 0, -1, 0, 1, -1, -1, 1, 1, -1, -1, 1};
const int CIY[NPOP] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1, 1, 1, -1, -1, 1, 1, -1, -1};
const int CIZ[NPOP] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1};
const int OPP[NPOP] = {0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12, 25, 26, 23, 24, 21, 22, 19, 20};

double weight(int i) {
    if (i == 0) return 8.0 / 27.0;
    if (i <= 6) return 2.0 / 27.0;
    if (i <= 18) return 1.0 / 54.0;
    return 1.0 / 216.0;
}

constexpr int RED_LANES = 64;  // lane sums per particle for the link reductions

// Lattice-sized arrays: an allocator that leaves new elements uninitialised, so that the arrays
// are first touched (and zeroed) by all threads in parallel instead of page by page on one
// thread (58 GB of populations at 512x512x256).
template <class T>
struct UninitAlloc : std::allocator<T> {
    template <class U>
    struct rebind {
        using other = UninitAlloc<U>;
    };
    UninitAlloc() = default;
    template <class U>
    UninitAlloc(const UninitAlloc<U> &) {}
    template <class U>
    void construct(U *) noexcept {}
    template <class U, class... A>
    void construct(U *p, A &&...a) {
        ::new ((void *)p) U(std::forward<A>(a)...);
    }
};
using Field = std::vector<double, UninitAlloc<double>>;

void fill_par(Field &v, size_t n, double value) {
    v.resize(n);
    const long m = (long)n;
#pragma omp parallel for schedule(static)
    for (long i = 0; i < m; i++) v[i] = value;
}

}  // namespace

struct OrState {
    rbtp_case c;
    rbtp_derived d;
    int LX, LY, LZ, LXY;
    long N;
    double tp[NPOP];
    // fluid (initial.cu:245-263); the streaming temporaries f_temp/g_temp are not needed (the
    // pull goes from the f_collide copy straight into f)
    Field f, g, f_collide, g_collide;
    Field rho, ux, uy, uz, phi, fx, fy, fz;
    // geometry (globals.cu:74-75)
    std::vector<int> ibnode, ibnode_prev, owner, owner_prev;
    // particles (globals.cu:64-103): indexed by RBTP_P_*
    int np;
    std::vector<double> p[RBTP_P_COUNT];
    std::vector<rbtp_link> links;
    std::vector<int> nf_nodes, nf_pids;
    long unhandled_f, error_g;
    int step;
    bool particles() const { return c.activate_particles && np > 0; }
};

extern "C" {

int or_abi_version(void) { return RBTP_ABI_VERSION; }

// parameters.h:1-55
void or_case_default(rbtp_case *c) {
    std::memset(c, 0, sizeof(*c));
    c->continuous_steps = 0;
    c->lx = 320; c->ly = 320; c->lz = 160;
    c->nend = 1000000; c->ndiag = 125; c->nflowout = 25000; c->nnuout = 125; c->nout_p = 125;
    c->rayleigh = 1.0e7; c->prandtl = 0.71; c->t_hot = 1.0; c->t_cold = 0.0; c->t_ref = 0.5;
    c->beta = 0.0000625; c->visc = 0.01206; c->rho0 = 1.0;
    c->activate_particles = 1;
    c->particle_fraction = 0.05;
    c->particle_diam_frac = (1.0 / 20.0);
    c->cp_particle_ratio = 10.0;
    c->cp_fluid = 1.0;
    c->rho_particle_ratio = 1.0;
    c->repforce_scale = 2.0e-4;
    c->particle_init_temp = (0.0 + 1.0) / 2;  // (T_COLD + T_HOT)/2
    c->particle_rand_seed = 86143;
    c->rep_min_gap = 2.0; c->rep_min_gap_w = 2.0;
    c->rep_stf0 = 0.025; c->rep_stf1 = 0.002; c->rep_stf0_w = 0.025; c->rep_stf1_w = 0.002;
    c->npart_override = -1;
    c->precision = RBTP_PREC_FP64;
}

// globals.cu:18-33, globals.cu:126-148 (init_particle_constants), globals.cu:150-169
// (update_particle_count).
void or_case_derive(const rbtp_case *c, rbtp_derived *d) {
    std::memset(d, 0, sizeof(*d));
    const double rayl = c->rayleigh, prand = c->prandtl, beta = c->beta;
    const double tHot = c->t_hot, tCold = c->t_cold, visc = c->visc;
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
    const double fluid_density = c->rho0;
    const double mass = fluid_density * c->rho_particle_ratio * volume;
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
        const double clamped = (raw < 0.0) ? 0.0 : raw;
        d->npart = static_cast<int>(std::floor(clamped));
    }
}

OrState *or_create(const rbtp_case *c) {
    if (c->lx < 3 || c->ly < 3 || c->lz < 3) return nullptr;
    OrState *s = new OrState();
    s->c = *c;
    or_case_derive(c, &s->d);
    s->LX = c->lx; s->LY = c->ly; s->LZ = c->lz; s->LXY = c->lx * c->ly;
    s->N = (long)s->LXY * c->lz;
    for (int i = 0; i < NPOP; i++) s->tp[i] = weight(i);
    const long N = s->N;
    for (auto *v : {&s->f, &s->g, &s->f_collide, &s->g_collide}) fill_par(*v, NPOP * N, 0.0);
    for (auto *v : {&s->rho, &s->ux, &s->uy, &s->uz, &s->phi, &s->fx, &s->fy, &s->fz})
        fill_par(*v, N, 0.0);
    s->ibnode.assign(N, 0); s->ibnode_prev.assign(N, 0);
    s->owner.assign(N, -1); s->owner_prev.assign(N, -1);
    s->np = c->activate_particles ? s->d.npart : 0;
    for (int k = 0; k < RBTP_P_COUNT; k++) s->p[k].assign(s->np > 0 ? s->np : 1, 0.0);
    s->unhandled_f = s->error_g = 0;
    s->step = 0;
    return s;
}

void or_destroy(OrState *s) { delete s; }

int or_npart(const OrState *s) { return s->np; }
long or_nodes(const OrState *s) { return s->N; }
int or_current_step(const OrState *s) { return s->step; }
void or_get_derived(const OrState *s, rbtp_derived *d) { *d = s->d; }
long or_unhandled_links(const OrState *s, int thermal) { return thermal ? s->error_g : s->unhandled_f; }

// Raw array access for the tests (numpy views over oracle memory).
// which: 0 f, 1 g, 2 f_collide, 3 g_collide, 10+RBTP_FIELD_* scalar fields,
//        100+RBTP_P_* particle arrays.
double *or_array(OrState *s, int which) {
    switch (which) {
        case 0: return s->f.data();
        case 1: return s->g.data();
        case 2: return s->f_collide.data();
        case 3: return s->g_collide.data();
        case 10 + RBTP_FIELD_RHO: return s->rho.data();
        case 10 + RBTP_FIELD_UX: return s->ux.data();
        case 10 + RBTP_FIELD_UY: return s->uy.data();
        case 10 + RBTP_FIELD_UZ: return s->uz.data();
        case 10 + RBTP_FIELD_PHI: return s->phi.data();
        case 10 + RBTP_FIELD_FX: return s->fx.data();
        case 10 + RBTP_FIELD_FY: return s->fy.data();
        case 10 + RBTP_FIELD_FZ: return s->fz.data();
        default:
            if (which >= 100 && which < 100 + RBTP_P_COUNT) return s->p[which - 100].data();
            return nullptr;
    }
}
int *or_iarray(OrState *s, int which) {
    switch (which) {
        case 0: return s->ibnode.data();
        case 1: return s->owner.data();
        case 2: return s->ibnode_prev.data();
        case 3: return s->owner_prev.data();
        default: return nullptr;
    }
}
long or_nlinks(const OrState *s) { return (long)s->links.size(); }
const rbtp_link *or_links(const OrState *s) { return s->links.data(); }
long or_nnew_fluid(const OrState *s) { return (long)s->nf_nodes.size(); }
const int *or_new_fluid_nodes(const OrState *s) { return s->nf_nodes.data(); }
const int *or_new_fluid_pids(const OrState *s) { return s->nf_pids.data(); }

// splitmix64 (Steele, Lea, Flood 2014), used for the seeded noise of the initial temperature.
static uint64_t splitmix64(uint64_t &x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Initial temperature of the benchmark cases (no reference counterpart): linear between the
// hot (x=0) and cold (x=LX) walls at cell centres plus amp*U[-1,1) in idx order.
void or_linear_temperature(const rbtp_case *c, uint64_t seed, double amp, double *out) {
    const int LX = c->lx, LXY = c->lx * c->ly;
    const long N = (long)LXY * c->lz;
    uint64_t st = seed;
    for (long idx = 0; idx < N; idx++) {
        const int ix = (int)(idx % LX);
        const double xi = (static_cast<double>(ix) + 0.5) / static_cast<double>(LX);
        const double lin = c->t_hot + (c->t_cold - c->t_hot) * xi;
        const double u = static_cast<double>(splitmix64(st) >> 11) * 0x1.0p-53;
        out[idx] = lin + amp * (2.0 * u - 1.0);
    }
}

// ---------------------------------------------------------------------------------------------
// Fluid kernels
// ---------------------------------------------------------------------------------------------

// The per-node loops below are written population-outer over chunks of CH consecutive nodes so
// that the compiler vectorises across nodes; each node's own arithmetic, and the order in which
// its sums accumulate over the populations, is exactly the reference's.
constexpr long CH = 512;

// macrovar (macrovar.cu:3-34)
void or_macrovar(OrState *s) {
    const long N = s->N;
#pragma omp parallel for schedule(static)
    for (long base = 0; base < N; base += CH) {
        const long m = std::min(CH, N - base);
        double rho_local[CH], ux_local[CH], uy_local[CH], uz_local[CH], phi_local[CH];
        for (long j = 0; j < m; j++) {
            rho_local[j] = 0.0; ux_local[j] = 0.0; uy_local[j] = 0.0; uz_local[j] = 0.0;
            phi_local[j] = 0.0;
        }
        for (int ip = 0; ip < NPOP; ip++) {
            const double *fp = &s->f[ip * N + base];
            const double *gp = &s->g[ip * N + base];
            const int cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip];
            for (long j = 0; j < m; j++) {
                const double f_val = fp[j];
                const double g_val = gp[j];
                rho_local[j] += f_val;
                ux_local[j] += cx * f_val;
                uy_local[j] += cy * f_val;
                uz_local[j] += cz * f_val;
                phi_local[j] += g_val;
            }
        }
        for (long j = 0; j < m; j++) {
            const long idx = base + j;
            s->rho[idx] = rho_local[j];
            s->ux[idx] = ux_local[j] + (MUT(18) ? 0.0 : 0.5 * s->fx[idx]);
            s->uy[idx] = uy_local[j] + 0.5 * s->fy[idx];
            s->uz[idx] = uz_local[j] + 0.5 * s->fz[idx];
            s->phi[idx] = phi_local[j];
        }
    }
}

// forcing (macrovar.cu:36-47)
void or_forcing(OrState *s) {
    const long N = s->N;
    const double grav0 = s->d.grav0, beta = s->c.beta, t0 = s->c.t_ref;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        s->fx[idx] = grav0 * beta * (s->phi[idx] - t0);
        s->fy[idx] = 0.0;
        s->fz[idx] = 0.0;
    }
}

// collision_BGK (collision.cu:3-33)
void or_collision_f(OrState *s) {
    const long N = s->N;
    const double tau = s->d.tau;
    const double RT = 1.0 / 3.0;
    const double force_coeff = MUT(20) ? 1.0 : (1.0 - 0.5 / tau);
#pragma omp parallel for schedule(static)
    for (long base = 0; base < N; base += CH) {
        const long m = std::min(CH, N - base);
        const double *fxv = &s->fx[base], *rho_local = &s->rho[base];
        const double *ux_local = &s->ux[base], *uy_local = &s->uy[base], *uz_local = &s->uz[base];
        double uv[CH];
        for (long j = 0; j < m; j++) {
            const double u2 = ux_local[j] * ux_local[j] + uy_local[j] * uy_local[j] + uz_local[j] * uz_local[j];
            uv[j] = u2 / RT;
        }
        for (int ip = 0; ip < NPOP; ip++) {
            const int cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip];
            const double tp = s->tp[ip];
            double *fp = &s->f[ip * N + base];
            for (long j = 0; j < m; j++) {
                double eu = (cx * ux_local[j] + cy * uy_local[j] + cz * uz_local[j]) / RT;
                double feq = tp * (rho_local[j] + eu + 0.5 * (eu * eu - uv[j]));
                double fpop = fp[j];
                double fchange = -(fpop - feq) / tau;
                double feq_rho0 = tp * (1.0 + eu + 0.5 * (eu * eu - uv[j]));
                double cuF = (MUT(19) ? ux_local[j] - cx : cx - ux_local[j]) * fxv[j];
                fchange += force_coeff * cuF / RT * feq_rho0;
                fp[j] = fpop + fchange;
            }
        }
    }
}

// collision_BGK_scalar (collision.cu:35-89)
void or_collision_g(OrState *s) {
    const long N = s->N;
    const double tauc = s->d.tauc;
    const double RT = 1.0 / 3.0;
    const double lambda0 = 1.0 - s->tp[0];
    const double Ralph_coeff = (1.0 - 1.0 / (2.0 * tauc));
#pragma omp parallel for schedule(static)
    for (long base = 0; base < N; base += CH) {
        const long m = std::min(CH, N - base);
        const double *fxv = &s->fx[base], *fyv = &s->fy[base], *fzv = &s->fz[base];
        const double *rho_local = &s->rho[base], *phi_local = &s->phi[base];
        const double *ux_local = &s->ux[base], *uy_local = &s->uy[base], *uz_local = &s->uz[base];
        double uv[CH], gneqx[CH], gneqy[CH], gneqz[CH], grad_denom[CH];
        for (long j = 0; j < m; j++) {
            const double u2 = ux_local[j] * ux_local[j] + uy_local[j] * uy_local[j] + uz_local[j] * uz_local[j];
            uv[j] = u2 / RT;
            gneqx[j] = 0.0; gneqy[j] = 0.0; gneqz[j] = 0.0;
            grad_denom[j] = (MUT(22) ? rho_local[j] + 1.0 : rho_local[j]) / 3.0 + 2.0 * tauc / 3.0;
        }
        for (int ip = 0; ip < NPOP; ip++) {
            const int cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip];
            const double tp = s->tp[ip];
            const double *gp = &s->g[ip * N + base];
            for (long j = 0; j < m; j++) {
                double eu = (cx * ux_local[j] + cy * uy_local[j] + cz * uz_local[j]) / RT;
                double geq = tp * phi_local[j] * (1.0 + eu + 0.5 * (eu * eu - uv[j]));
                double geqnew = (ip == 0) ? (MUT(21) ? geq + lambda0 * phi_local[j] * rho_local[j]
                                                     : geq - lambda0 * phi_local[j] * rho_local[j])
                                          : (geq + tp * phi_local[j] * rho_local[j]);
                double diff = gp[j] - geqnew;
                gneqx[j] += cx * diff;
                gneqy[j] += cy * diff;
                gneqz[j] += cz * diff;
            }
        }
        for (int ip = 0; ip < NPOP; ip++) {
            const int cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip];
            const double tp = s->tp[ip];
            double *gp = &s->g[ip * N + base];
            for (long j = 0; j < m; j++) {
                double eb = cx * fxv[j] + cy * fyv[j] + cz * fzv[j];
                double gneq = cx * gneqx[j] + cy * gneqy[j] + cz * gneqz[j];
                double gradt = -(2.0 * gneq + eb * phi_local[j]) / grad_denom[j];
                double Ralph = tp * Ralph_coeff * (gradt * rho_local[j] + (MUT(23) ? 0.0 : 3.0 * phi_local[j] * eb));
                double eu = (cx * ux_local[j] + cy * uy_local[j] + cz * uz_local[j]) / RT;
                double geq = tp * phi_local[j] * (1.0 + eu + 0.5 * (eu * eu - uv[j]));
                double geqnew = (ip == 0) ? (MUT(21) ? geq + lambda0 * phi_local[j] * rho_local[j]
                                                     : geq - lambda0 * phi_local[j] * rho_local[j])
                                          : (geq + tp * phi_local[j] * rho_local[j]);
                double gpop = gp[j];
                double gchange = -(gpop - geqnew) / tauc + Ralph;
                gp[j] = gpop + gchange;
            }
        }
    }
}

// streaming (streaming.cu:3-23) and streaming_scalar (streaming.cu:25-63), pulling from src into
// dst, row by row.  A row whose source row lies outside in z or y takes the opposite population
// of the node itself (the z/y walls come first in streaming_scalar's precedence); inside rows
// copy with the x shift, and only their end nodes can meet the x walls.  thermal selects the
// x-wall rule of streaming_scalar (g) over plain bounce-back (f).
static void stream_into(OrState *s, const Field &src, Field &dst,
                        bool thermal) {
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    const double tHot = s->c.t_hot, tCold = s->c.t_cold;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ip = 0; ip < NPOP; ++ip) {
        for (int iz = 0; iz < LZ; iz++) {
            const int cx = CIX[ip], cy = CIY[ip], cz = CIZ[ip], opp = OPP[ip];
            const double *sp = &src[ip * N];
            const double *so = &src[opp * N];
            double *dp = &dst[ip * N];
            const double tpw = s->tp[ip];
            for (int iy = 0; iy < LY; iy++) {
                const long row = (long)iz * LXY + (long)iy * LX;
                const int sy = iy - cy, sz = iz - cz;
                if (!(sy >= 0 && sy < LY && sz >= 0 && sz < LZ)) {
                    for (int ix = 0; ix < LX; ix++) dp[row + ix] = so[row + ix];
                    if (MUT(27) && thermal && cx != 0) {
                        const int iw = cx > 0 ? 0 : LX - 1;
                        dp[row + iw] = -so[row + iw] + 2.0 * tpw * (cx > 0 ? tHot : tCold) *
                                                           (1.0 + s->rho[row + iw]);
                    }
                    continue;
                }
                const long srow = (long)sz * LXY + (long)sy * LX;
                for (int ix = 0; ix < LX; ix++) {
                    const int sx = ix - cx;
                    if (sx >= 0 && sx < LX) {
                        dp[row + ix] = sp[srow + sx];
                    } else if (!thermal) {
                        dp[row + ix] = MUT(24) ? sp[row + ix] : so[row + ix];
                    } else {
                        const double density = MUT(25) ? 0.0 : s->rho[row + ix];
                        const double t_in = MUT(26) ? tCold : tHot, t_out = MUT(26) ? tHot : tCold;
                        if (sx < 0 && cx > 0)
                            dp[row + ix] = -so[row + ix] + 2.0 * tpw * t_in * (1.0 + density);
                        else if (sx >= LX && cx < 0)
                            dp[row + ix] = -so[row + ix] + 2.0 * tpw * t_out * (1.0 + density);
                        else
                            dp[row + ix] = so[row + ix];
                    }
                }
            }
        }
    }
}

static void copy_par(const Field &src, Field &dst) {
    const long n = (long)src.size();
#pragma omp parallel for schedule(static)
    for (long i = 0; i < n; i++) dst[i] = src[i];
}

// main.cu:45-48: f_collide <- f, stream f into f_temp, f <- f_temp.  Pulling from the
// f_collide copy straight into f gives the same values without the temporary.
void or_stream_f(OrState *s) {
    copy_par(s->f, s->f_collide);
    stream_into(s, s->f_collide, s->f, false);
}
// main.cu:61-64: the same for g with streaming_scalar.
void or_stream_g(OrState *s) {
    copy_par(s->g, s->g_collide);
    stream_into(s, s->g_collide, s->g, true);
}

// init_f (initial.cu:335-360) and init_g (initial.cu:362-397).  With temperature != NULL the
// uniform tempT = t0 of init_g is replaced by the given field (the benchmark initial state); the
// formula is otherwise unchanged.
static void init_fluid(OrState *s, const double *temperature) {
    const long N = s->N;
    const double t0 = s->c.t_ref, grav0 = s->d.grav0, beta = s->c.beta, RHO0 = s->c.rho0;
    // every node is independent: the loops run in parallel
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        // init_f
        const double r = 0.0, u = 0.0, v = 0.0, w = 0.0;
        const double RT = 1.0 / 3.0;
        const double uv = (u * u + v * v + w * w) / RT;
        s->rho[idx] = RHO0;
        s->ux[idx] = u; s->uy[idx] = v; s->uz[idx] = w;
        for (int ip = 0; ip < NPOP; ip++) {
            const double eu = (CIX[ip] * u + CIY[ip] * v + CIZ[ip] * w) / RT;
            const double feq = s->tp[ip] * (r + eu + 0.5 * (eu * eu - uv));
            s->f[ip * N + idx] = feq;
        }
    }
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        // init_g
        const double r = s->rho[idx];
        const double u = s->ux[idx], v = s->uy[idx], w = s->uz[idx];
        const double RT = 1.0 / 3.0;
        const double tempT = temperature ? temperature[idx] : t0;
        const double uv = (u * u + v * v + w * w) / RT;
        const double lambda0 = 1.0 - s->tp[0];
        s->phi[idx] = tempT;
        s->fx[idx] = grav0 * beta * (tempT - t0);
        s->fy[idx] = 0.0;
        s->fz[idx] = 0.0;
        for (int ip = 0; ip < NPOP; ip++) {
            const double eu = (CIX[ip] * u + CIY[ip] * v + CIZ[ip] * w) / RT;
            double geq = s->tp[ip] * tempT * (1.0 + eu + 0.5 * (eu * eu - uv));
            if (ip == 0) {
                if (MUT(38)) geq += lambda0 * tempT * r;
                else geq -= lambda0 * tempT * r;
            } else {
                geq += s->tp[ip] * tempT * r;
            }
            s->g[ip * N + idx] = geq;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Particle geometry
// ---------------------------------------------------------------------------------------------

// compute_link_q (links.cu:7-23)
static bool compute_link_q(int ip, double x0, double y0, double z0, double px, double py,
                           double pz, double radius_sq, double &q) {
    const double cx_dir = static_cast<double>(CIX[ip]);
    const double cy_dir = static_cast<double>(CIY[ip]);
    const double cz_dir = static_cast<double>(CIZ[ip]);
    const double dx = x0 - px;
    const double dy = y0 - py;
    const double dz = z0 - pz;
    const double a = cx_dir * cx_dir + cy_dir * cy_dir + cz_dir * cz_dir;
    const double b = 2.0 * (cx_dir * dx + cy_dir * dy + cz_dir * dz);
    const double c = dx * dx + dy * dy + dz * dz - radius_sq;
    const double disc = b * b - 4.0 * a * c;
    if (disc <= 0.0) return false;
    q = (MUT(28) ? -b + std::sqrt(disc) : -b - std::sqrt(disc)) / (2.0 * a);
    return (q >= 0.0 && q <= 1.0);
}

// Marking box of a particle (links.cu:164-177).  Returns false when it is empty.
static bool mark_box(const OrState *s, double px, double py, double pz, int lo[3], int hi[3]) {
    const double R = s->d.particle_radius;
    lo[0] = static_cast<int>(std::floor(px - R - 0.5));
    hi[0] = static_cast<int>(std::floor(px + R - 0.5));
    lo[1] = static_cast<int>(std::floor(py - R - 0.5));
    hi[1] = static_cast<int>(std::floor(py + R - 0.5));
    lo[2] = static_cast<int>(std::floor(pz - R - 0.5));
    hi[2] = static_cast<int>(std::floor(pz + R - 0.5));
    const int L[3] = {s->LX, s->LY, s->LZ};
    for (int a = 0; a < 3; a++) {
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] > L[a] - 1) hi[a] = L[a] - 1;
    }
    return !(lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]);
}

// build_links (links.cu:143-256): clear flags, mark every particle's box in pid order
// (mark_ibnode_kernel, links.cu:25-48: the first particle to claim a node keeps it), count links
// per node (links.cu:50-87), exclusive scan, fill (links.cu:89-139).
void or_build_links(OrState *s) {
    if (!s->particles()) return;
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    std::fill(s->ibnode.begin(), s->ibnode.end(), 0);
    std::fill(s->owner.begin(), s->owner.end(), -1);
    const double R = s->d.particle_radius;
    const double radius_sq = R * R;
    const double *PX = s->p[RBTP_P_POS_X].data(), *PY = s->p[RBTP_P_POS_Y].data(),
                 *PZ = s->p[RBTP_P_POS_Z].data();
    for (int p = 0; p < s->np; ++p) {
        int lo[3], hi[3];
        if (!mark_box(s, PX[p], PY[p], PZ[p], lo, hi)) continue;
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    const double cx = static_cast<double>(ix) + 0.5;
                    const double cy = static_cast<double>(iy) + 0.5;
                    const double cz = static_cast<double>(iz) + 0.5;
                    const double dx = cx - PX[p], dy = cy - PY[p], dz = cz - PZ[p];
                    if (dx * dx + dy * dy + dz * dz > radius_sq) continue;
                    const long idx = (long)iz * LXY + (long)iy * LX + ix;
                    if (!MUT(29) && s->owner[idx] != -1) continue;
                    s->owner[idx] = p;
                    s->ibnode[idx] = 1;
                }
    }
    // link_count_kernel + ExclusiveSum + link_fill_kernel: the list in (idx, dir) order.  Each
    // thread lists the links of one contiguous range of nodes, and the ranges are joined in
    // order, which is what the scan's offsets do.
    s->links.clear();
    const int nchunk = 256;
    std::vector<std::vector<rbtp_link>> part(nchunk);
#pragma omp parallel for schedule(dynamic, 1)
    for (int ch = 0; ch < nchunk; ch++) {
        const long idx_lo = N * ch / nchunk, idx_hi = N * (ch + 1) / nchunk;
        std::vector<rbtp_link> &out = part[ch];
        for (long idx = idx_lo; idx < idx_hi; idx++) {
            if (s->ibnode[idx] == 1) continue;
            const int iz = (int)(idx / LXY);
            const int rem = (int)(idx - (long)iz * LXY);
            const int iy = rem / LX;
            const int ix = rem - iy * LX;
            const double x0 = static_cast<double>(ix) + 0.5;
            const double y0 = static_cast<double>(iy) + 0.5;
            const double z0 = static_cast<double>(iz) + 0.5;
            for (int ip = 1; ip < NPOP; ++ip) {
                const int nx = ix + CIX[ip], ny = iy + CIY[ip], nz = iz + CIZ[ip];
                if (nx < 0 || nx >= LX || ny < 0 || ny >= LY || nz < 0 || nz >= LZ) continue;
                const long nidx = (long)nz * LXY + (long)ny * LX + nx;
                const int pid = s->owner[nidx];
                if (pid < 0) continue;
                double q = 0.0;
                if (!compute_link_q(ip, x0, y0, z0, PX[pid], PY[pid], PZ[pid], radius_sq, q)) continue;
                rbtp_link L;
                L.cell_i = ix; L.cell_j = iy; L.cell_k = iz; L.dir = ip; L.pid = pid; L.pad = 0;
                L.q = q;
                L.rx = q * static_cast<double>(CIX[ip]);
                L.ry = q * static_cast<double>(CIY[ip]);
                L.rz = q * static_cast<double>(CIZ[ip]);
                out.push_back(L);
            }
        }
    }
    for (const auto &v : part) s->links.insert(s->links.end(), v.begin(), v.end());
}

// build_initial_ibnode (pinit.cu:146-179).  The reference walks every node and gives it to the
// first pid (in increasing order) whose sphere contains it, O(N P).  Walking the pids in
// increasing order over their marking boxes floor(p -/+ R - 0.5) (links.cu:164-177) and keeping
// a node's first owner gives the same map: a node centre i + 1/2 within R of p lies in
// [p - R, p + R], so i lies in [p - R - 1/2, p + R - 1/2] and inside the box.  The distance test
// is the reference's expression.
static void initial_ibnode(OrState *s) {
    const long N = s->N;
    const int LX = s->LX, LXY = s->LXY;
    const double R = s->d.particle_radius;
    const double radius_sq = R * R;
    for (long idx = 0; idx < N; ++idx) {
        s->ibnode[idx] = 0;
        s->owner[idx] = -1;
    }
    for (int p = 0; p < s->np; ++p) {
        const double px = s->p[RBTP_P_POS_X][p], py = s->p[RBTP_P_POS_Y][p],
                     pz = s->p[RBTP_P_POS_Z][p];
        int lo[3], hi[3];
        if (!mark_box(s, px, py, pz, lo, hi)) continue;
        for (int iz = lo[2]; iz <= hi[2]; iz++)
            for (int iy = lo[1]; iy <= hi[1]; iy++)
                for (int ix = lo[0]; ix <= hi[0]; ix++) {
                    const long idx = (long)iz * LXY + (long)iy * LX + ix;
                    if (!MUT(29) && s->owner[idx] != -1) continue;
                    const double dx = (static_cast<double>(ix) + 0.5) - px;
                    const double dy = (static_cast<double>(iy) + 0.5) - py;
                    const double dz = (static_cast<double>(iz) + 0.5) - pz;
                    if (dx * dx + dy * dy + dz * dz <= radius_sq) {
                        s->ibnode[idx] = 1;
                        s->owner[idx] = p;
                    }
                }
    }
    s->ibnode_prev = s->ibnode;
    s->owner_prev = s->owner;
}

// generate_initial_positions (pinit.cu:97-144), glibc srand/rand.  Returns -1 on give-up.
static int place_particles(OrState *s) {
    if (s->np == 0) return 0;
    const double radius = s->d.particle_radius;
    std::srand(s->c.particle_rand_seed);
    const double min_pos_x = 1.5 * radius, min_pos_y = 1.5 * radius, min_pos_z = 1.5 * radius;
    const double span_x = static_cast<double>(s->LX) - 3.0 * radius;
    const double span_y = static_cast<double>(s->LY) - 3.0 * radius;
    const double span_z = static_cast<double>(s->LZ) - 3.0 * radius;
    long counter = 0;
    double *X = s->p[RBTP_P_POS_X].data(), *Y = s->p[RBTP_P_POS_Y].data(),
           *Z = s->p[RBTP_P_POS_Z].data();
    for (int p = 0; p < s->np; ++p) {
        while (true) {
            double rx = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            double ry = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            double rz = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            double x = min_pos_x + rx * span_x;
            double y = min_pos_y + ry * span_y;
            double z = min_pos_z + rz * span_z;
            bool ok = true;
            for (int q = 0; q < p; ++q) {
                double dx = x - X[q], dy = y - Y[q], dz = z - Z[q];
                double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (dist - 2.0 * radius < radius / 3.0) {
                    counter++;
                    if (counter > 1000000) return -1;
                    ok = false;
                    break;
                }
            }
            if (ok) {
                X[p] = x; Y[p] = y; Z[p] = z;
                break;
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Particle coupling kernels
// ---------------------------------------------------------------------------------------------

// bounce_back_particles_kernel (ibb.cu:5-82)
void or_ibb_f(OrState *s) {
    if (!s->particles()) return;
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    const double RHO0 = s->c.rho0;
    for (const rbtp_link &link : s->links) {
        const int ip = link.dir;
        const int opp = OPP[ip];
        const long idx = (long)link.cell_k * LXY + (long)link.cell_j * LX + link.cell_i;
        const int pid = link.pid;
        const double q = link.q;
        const int ix = link.cell_i, iy = link.cell_j, iz = link.cell_k;
        const int ix1 = ix - CIX[ip], iy1 = iy - CIY[ip], iz1 = iz - CIZ[ip];
        const double cx = static_cast<double>(ix) + 0.5 + q * CIX[ip];
        const double cy = static_cast<double>(iy) + 0.5 + q * CIY[ip];
        const double cz = static_cast<double>(iz) + 0.5 + q * CIZ[ip];
        const double px = s->p[RBTP_P_POS_X][pid], py = s->p[RBTP_P_POS_Y][pid],
                     pz = s->p[RBTP_P_POS_Z][pid];
        const double relx = cx - px, rely = cy - py, relz = cz - pz;
        const double *V[3] = {s->p[RBTP_P_VEL_X].data(), s->p[RBTP_P_VEL_Y].data(), s->p[RBTP_P_VEL_Z].data()};
        const double *W[3] = {s->p[RBTP_P_OMEGA_X].data(), s->p[RBTP_P_OMEGA_Y].data(), s->p[RBTP_P_OMEGA_Z].data()};
        const double uwx = V[0][pid] + (relz * W[1][pid] - rely * W[2][pid]);
        const double uwy = V[1][pid] + (relx * W[2][pid] - relz * W[0][pid]);
        const double uwz = V[2][pid] + (rely * W[0][pid] - relx * W[1][pid]);
        const double evel = CIX[ip] * uwx + CIY[ip] * uwy + CIZ[ip] * uwz;
        const double RT = 1.0 / 3.0;
        const double momentum_term = MUT(4) ? 0.0 : 2.0 * s->tp[ip] * RHO0 * evel / RT;
        bool handled = false;
        if (ix1 >= 0 && ix1 < LX && iy1 >= 0 && iy1 < LY && iz1 >= 0 && iz1 < LZ) {
            const long idx1 = (long)iz1 * LXY + (long)iy1 * LX + ix1;
            if (s->ibnode[idx1] == 0) {
                const double f_here = s->f_collide[ip * N + idx];
                if (q < 0.5) {
                    const double f_upstream = s->f_collide[(MUT(2) ? opp : ip) * N + idx1];
                    s->f[opp * N + idx] = MUT(1) ? 2.0 * q * f_upstream + (1.0 - 2.0 * q) * f_here - momentum_term
                                                 : 2.0 * q * f_here + (1.0 - 2.0 * q) * f_upstream - momentum_term;
                } else {
                    const double f_opposite_collide = s->f_collide[opp * N + idx];
                    const double inv_2q = 0.5 / q;
                    s->f[opp * N + idx] = MUT(3) ? (f_opposite_collide - momentum_term) * inv_2q +
                                                       (2.0 * q - 1.0) * inv_2q * f_here
                                                 : (f_here - momentum_term) * inv_2q +
                                                       (2.0 * q - 1.0) * inv_2q * f_opposite_collide;
                }
                handled = true;
            }
        }
        if (!handled && s->ibnode[idx] == 0) {
            const double f_here = s->f_collide[ip * N + idx];
            s->f[opp * N + idx] = f_here - momentum_term;
            handled = true;
        }
        if (!handled) {
            s->f[opp * N + idx] = s->f_collide[ip * N + idx];
            s->unhandled_f++;
        }
    }
}

// bounce_back_thermal_kernel (ibb.cu:84-136)
void or_ibb_g(OrState *s) {
    if (!s->particles()) return;
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    for (const rbtp_link &link : s->links) {
        const int ip = link.dir;
        const int opp = OPP[ip];
        const int ix = link.cell_i, iy = link.cell_j, iz = link.cell_k;
        const long idx = (long)iz * LXY + (long)iy * LX + ix;
        const int pid = link.pid;
        const double q = link.q;
        const int ix1 = ix - CIX[ip], iy1 = iy - CIY[ip], iz1 = iz - CIZ[ip];
        const long idx1 = (long)iz1 * LXY + (long)iy1 * LX + ix1;
        const double density = s->rho[idx];
        const double Tp = s->p[RBTP_P_TEMP][pid];
        const double amp = 2.0 * s->tp[ip] * Tp * (MUT(7) ? 1.0 : 1.0 + density);
        double value = 0.0;
        if (ix1 >= 0 && ix1 < LX && iy1 >= 0 && iy1 < LY && iz1 >= 0 && iz1 < LZ &&
            (MUT(8) ? s->ibnode[idx1] : s->ibnode_prev[idx1]) == 0) {
            if (q <= 0.5) {
                const double g_here = s->g_collide[ip * N + (MUT(5) ? idx1 : idx)];
                const double g_upstream = s->g_collide[ip * N + (MUT(5) ? idx : idx1)];
                value = -(2.0 * q * g_here + (1.0 - 2.0 * q) * g_upstream) + amp;
            } else {
                const double g_here = s->g_collide[(MUT(6) ? opp : ip) * N + idx];
                const double g_opp_here = s->g_collide[(MUT(6) ? ip : opp) * N + idx];
                const double inv_2q = 0.5 / q;
                value = -g_here * inv_2q + amp * inv_2q + (2.0 * q - 1.0) * inv_2q * g_opp_here;
            }
        } else if (s->ibnode[idx] == 0) {
            const double g_here = s->g_collide[ip * N + idx];
            value = -g_here + amp;
        } else {
            s->error_g++;
            continue;
        }
        s->g[opp * N + idx] = value;
    }
}

// prepare_particle_forces_step (pforces.cu:80-116)
void or_prepare_forces(OrState *s) {
    if (!s->particles()) return;
    for (int a = 0; a < 3; a++) {
        s->p[RBTP_P_FORCE_PREV2_X + a] = s->p[RBTP_P_FORCE_PREV_X + a];
        s->p[RBTP_P_TORQUE_PREV2_X + a] = s->p[RBTP_P_TORQUE_PREV_X + a];
    }
    for (int a = 0; a < 3; a++) {
        s->p[RBTP_P_FORCE_PREV_X + a] = s->p[RBTP_P_FORCE_X + a];
        s->p[RBTP_P_TORQUE_PREV_X + a] = s->p[RBTP_P_TORQUE_X + a];
    }
    for (int a = 0; a < 3; a++) {
        std::fill(s->p[RBTP_P_FORCE_X + a].begin(), s->p[RBTP_P_FORCE_X + a].end(), 0.0);
        std::fill(s->p[RBTP_P_TORQUE_X + a].begin(), s->p[RBTP_P_TORQUE_X + a].end(), 0.0);
    }
}

// apply_repulsive_forces (pforces.cu:118-148) with repulsive_pair_kernel (pforces.cu:6-38) and
// repulsive_wall_kernel (pforces.cu:40-76).  Pair terms are added per particle in ascending
// partner order; the term on the higher pid of a pair is computed from its own side, which is
// the exact negative of the reference's -lubforce*n.
void or_repulsion(OrState *s) {
    if (!s->particles()) return;
    const int np = s->np;
    const double radius = s->d.particle_radius;
    const double coeff = s->d.particle_mass * s->c.repforce_scale;
    const double min_gap = s->c.rep_min_gap, min_gap_w = s->c.rep_min_gap_w;
    const double *X = s->p[RBTP_P_POS_X].data(), *Y = s->p[RBTP_P_POS_Y].data(),
                 *Z = s->p[RBTP_P_POS_Z].data();
    double *FX = s->p[RBTP_P_FORCE_X].data(), *FY = s->p[RBTP_P_FORCE_Y].data(),
           *FZ = s->p[RBTP_P_FORCE_Z].data();
    if (np > 1) {
        for (int i = 0; i < np; i++) {
            for (int j = 0; j < np; j++) {
                if (j == i) continue;
                double dx = X[i] - X[j], dy = Y[i] - Y[j], dz = Z[i] - Z[j];
                double dist_sq = dx * dx + dy * dy + dz * dz;
                if (dist_sq <= 1.0e-12) continue;
                double dist = std::sqrt(dist_sq);
                double hgap = dist - (MUT(32) ? 1.0 : 2.0) * radius;
                if (hgap >= min_gap) continue;
                double overlap = (hgap < 0.0) ? -hgap : 0.0;
                double factor = (min_gap - hgap) / min_gap;
                double lubforce = coeff * (factor * factor / s->c.rep_stf0 +
                                           overlap / (min_gap * s->c.rep_stf1));
                double nx = dx / dist, ny = dy / dist, nz = dz / dist;
                FX[i] += lubforce * nx;
                FY[i] += lubforce * ny;
                FZ[i] += lubforce * nz;
            }
        }
    }
    const double L[3] = {static_cast<double>(s->LX), static_cast<double>(s->LY), static_cast<double>(s->LZ)};
    for (int p = 0; p < np; p++) {
        double fx = 0.0, fy = 0.0, fz = 0.0;
        auto apply_wall_force = [&](double gap, double nx, double ny, double nz) {
            if (gap >= min_gap_w) return;
            double overlap = (gap < 0.0) ? -gap : 0.0;
            double factor = (min_gap_w - gap) / min_gap_w;
            double lubforce = coeff * (factor * factor / s->c.rep_stf0_w +
                                       overlap / (min_gap_w * s->c.rep_stf1_w));
            fx += lubforce * nx;
            fy += lubforce * ny;
            fz += lubforce * nz;
        };
        const double x = X[p], y = Y[p], z = Z[p];
        apply_wall_force(x - radius, 1.0, 0.0, 0.0);
        apply_wall_force((L[0] - x) - radius, -1.0, 0.0, 0.0);
        apply_wall_force(y - radius, 0.0, 1.0, 0.0);
        apply_wall_force((L[1] - y) - radius, 0.0, -1.0, 0.0);
        apply_wall_force(z - radius, 0.0, 0.0, 1.0);
        apply_wall_force((L[2] - z) - radius, 0.0, 0.0, -1.0);
        FX[p] += fx;
        FY[p] += fy;
        FZ[p] += fz;
    }
}

// Position of a link in its particle's link box (see file header), or -1.
static long link_slot(const OrState *s, const rbtp_link &L) {
    const int pid = L.pid;
    int lo[3], hi[3];
    if (!mark_box(s, s->p[RBTP_P_POS_X][pid], s->p[RBTP_P_POS_Y][pid], s->p[RBTP_P_POS_Z][pid], lo, hi))
        return -1;
    const int Lm[3] = {s->LX, s->LY, s->LZ};
    int elo[3], ehi[3];
    for (int a = 0; a < 3; a++) {
        elo[a] = std::max(lo[a] - 1, 0);
        ehi[a] = std::min(hi[a] + 1, Lm[a] - 1);
    }
    const int nx = ehi[0] - elo[0] + 1, ny = ehi[1] - elo[1] + 1;
    return ((long)(L.cell_k - elo[2]) * ny + (L.cell_j - elo[1])) * nx + (L.cell_i - elo[0]);
}

static void tree_reduce(double *lanes) {
    for (int stride = RED_LANES / 2; stride >= 1; stride >>= 1)
        for (int t = 0; t < stride; t++) lanes[t] = lanes[t] + lanes[t + stride];
}

// compute_particle_forces_kernel (ibb.cu:138-188) with the deterministic summation order of the
// file header.
void or_forces(OrState *s) {
    if (!s->particles() || s->links.empty()) return;
    const long N = s->N;
    const int LX = s->LX, LXY = s->LXY;
    const int np = s->np;
    std::vector<double> part((size_t)np * 6 * RED_LANES, 0.0);
    std::map<std::pair<int, long>, std::array<double, 6>> share;  // (pid, box position) -> node share
    for (const rbtp_link &link : s->links) {
        const int ip = link.dir;
        const int opp = OPP[ip];
        const long idx = (long)link.cell_k * LXY + (long)link.cell_j * LX + link.cell_i;
        const int pid = link.pid;
        const double contact_x = link.cell_i + 0.5 + link.rx;
        const double contact_y = link.cell_j + 0.5 + link.ry;
        const double contact_z = link.cell_k + 0.5 + link.rz;
        const double px = s->p[RBTP_P_POS_X][pid], py = s->p[RBTP_P_POS_Y][pid],
                     pz = s->p[RBTP_P_POS_Z][pid];
        const double relx = contact_x - px, rely = contact_y - py, relz = contact_z - pz;
        const double wx = s->p[RBTP_P_OMEGA_X][pid], wy = s->p[RBTP_P_OMEGA_Y][pid],
                     wz = s->p[RBTP_P_OMEGA_Z][pid];
        const double uwx = s->p[RBTP_P_VEL_X][pid] + (relz * wy - rely * wz);
        const double uwy = s->p[RBTP_P_VEL_Y][pid] + (relx * wz - relz * wx);
        const double uwz = s->p[RBTP_P_VEL_Z][pid] + (rely * wx - relx * wy);
        const double f_in = s->f_collide[ip * N + idx];
        const double f_out = s->f[opp * N + idx];
        const double sw = MUT(30) ? -1.0 : 1.0;
        const double Fx = (CIX[ip] - uwx) * f_in - (CIX[opp] - sw * uwx) * f_out;
        const double Fy = (CIY[ip] - uwy) * f_in - (CIY[opp] - sw * uwy) * f_out;
        const double Fz = (CIZ[ip] - uwz) * f_in - (CIZ[opp] - sw * uwz) * f_out;
        std::array<double, 6> &v = share[{pid, link_slot(s, link)}];
        v[0] += Fx;
        v[1] += Fy;
        v[2] += Fz;
        v[3] += rely * Fz - relz * Fy;
        v[4] += relz * Fx - relx * Fz;
        v[5] += relx * Fy - rely * Fx;
    }
    for (const auto &kv : share) {
        double *b = &part[((size_t)kv.first.first * 6) * RED_LANES];
        const int lane = (int)(kv.first.second % RED_LANES);
        for (int k = 0; k < 6; k++) b[k * RED_LANES + lane] += kv.second[k];
    }
    const int dst[6] = {RBTP_P_FORCE_X, RBTP_P_FORCE_Y, RBTP_P_FORCE_Z,
                        RBTP_P_TORQUE_X, RBTP_P_TORQUE_Y, RBTP_P_TORQUE_Z};
    for (int pid = 0; pid < np; pid++)
        for (int k = 0; k < 6; k++) {
            double *b = &part[((size_t)pid * 6 + k) * RED_LANES];
            tree_reduce(b);
            s->p[dst[k]][pid] += b[0];
        }
}

// compute_particle_heat (ibb.cu:258-271) and compute_particle_heat_kernel (ibb.cu:190-205) with
// the deterministic summation order of the file header.
void or_heat(OrState *s) {
    if (!s->particles()) return;
    const long N = s->N;
    const int LX = s->LX, LXY = s->LXY;
    const int np = s->np;
    s->p[RBTP_P_HEAT_PREV] = s->p[RBTP_P_HEAT];
    std::fill(s->p[RBTP_P_HEAT].begin(), s->p[RBTP_P_HEAT].end(), 0.0);
    if (s->links.empty()) return;
    std::vector<double> part((size_t)np * RED_LANES, 0.0);
    std::map<std::pair<int, long>, double> share;  // (pid, box position) -> node share
    for (const rbtp_link &link : s->links) {
        const int ip = link.dir;
        const int opp = OPP[ip];
        const long idx = (long)link.cell_k * LXY + (long)link.cell_j * LX + link.cell_i;
        const double Gi = s->g_collide[ip * N + idx];
        const double Go = s->g[opp * N + idx];
        share[{link.pid, link_slot(s, link)}] += MUT(31) ? Gi + Go : Gi - Go;
    }
    for (const auto &kv : share)
        part[(size_t)kv.first.first * RED_LANES + kv.first.second % RED_LANES] += kv.second;
    for (int pid = 0; pid < np; pid++) {
        double *b = &part[(size_t)pid * RED_LANES];
        tree_reduce(b);
        s->p[RBTP_P_HEAT][pid] += b[0];
    }
}

// update_particles_kernel (move.cu:5-103)
void or_move(OrState *s) {
    if (!s->particles()) return;
    auto P = [&](int k) { return s->p[k].data(); };
    const double mass = s->d.particle_mass, inertia = s->d.particle_inertia;
    const double heat_capacity = s->d.particle_heat_capacity, rho_ratio = s->c.rho_particle_ratio;
    for (int pid = 0; pid < s->np; pid++) {
        const double dt = 1.0;
        const double forceX = P(RBTP_P_FORCE_X)[pid], forceY = P(RBTP_P_FORCE_Y)[pid],
                     forceZ = P(RBTP_P_FORCE_Z)[pid];
        const double torqueX = P(RBTP_P_TORQUE_X)[pid], torqueY = P(RBTP_P_TORQUE_Y)[pid],
                     torqueZ = P(RBTP_P_TORQUE_Z)[pid];
        const double fx_prev = P(RBTP_P_FORCE_PREV_X)[pid], fy_prev = P(RBTP_P_FORCE_PREV_Y)[pid],
                     fz_prev = P(RBTP_P_FORCE_PREV_Z)[pid];
        const double fx_prev2 = P(RBTP_P_FORCE_PREV2_X)[pid], fy_prev2 = P(RBTP_P_FORCE_PREV2_Y)[pid],
                     fz_prev2 = P(RBTP_P_FORCE_PREV2_Z)[pid];
        const double tpx = P(RBTP_P_TORQUE_PREV_X)[pid], tpy = P(RBTP_P_TORQUE_PREV_Y)[pid],
                     tpz = P(RBTP_P_TORQUE_PREV_Z)[pid];
        const double tp2x = P(RBTP_P_TORQUE_PREV2_X)[pid], tp2y = P(RBTP_P_TORQUE_PREV2_Y)[pid],
                     tp2z = P(RBTP_P_TORQUE_PREV2_Z)[pid];
        const double buoyancy = (MUT(34) ? -1.0 : 1.0) * (1.0 - 1.0 / rho_ratio) * s->d.grav0;
        const double vx_old = P(RBTP_P_VEL_X)[pid], vy_old = P(RBTP_P_VEL_Y)[pid],
                     vz_old = P(RBTP_P_VEL_Z)[pid];
        const double omega_x_old = P(RBTP_P_OMEGA_X)[pid], omega_y_old = P(RBTP_P_OMEGA_Y)[pid],
                     omega_z_old = P(RBTP_P_OMEGA_Z)[pid];
        const double combo_fx = MUT(33) ? forceX + fx_prev : 0.5 * forceX + fx_prev + 0.5 * fx_prev2;
        const double combo_fy = 0.5 * forceY + fy_prev + 0.5 * fy_prev2;
        const double combo_fz = 0.5 * forceZ + fz_prev + 0.5 * fz_prev2;
        const double combo_torque_x = 0.5 * torqueX + tpx + 0.5 * tp2x;
        const double combo_torque_y = 0.5 * torqueY + tpy + 0.5 * tp2y;
        const double combo_torque_z = 0.5 * torqueZ + tpz + 0.5 * tp2z;
        double vx_new = vx_old + 0.5 * dt * combo_fx / mass + dt * buoyancy;
        double vy_new = vy_old + 0.5 * dt * combo_fy / mass;
        double vz_new = vz_old + 0.5 * dt * combo_fz / mass;
        double omega_x_new = omega_x_old + 0.5 * dt * combo_torque_x / inertia;
        double omega_y_new = omega_y_old + 0.5 * dt * combo_torque_y / inertia;
        double omega_z_new = omega_z_old + 0.5 * dt * combo_torque_z / inertia;
        const double x_old = P(RBTP_P_POS_X)[pid], y_old = P(RBTP_P_POS_Y)[pid],
                     z_old = P(RBTP_P_POS_Z)[pid];
        P(RBTP_P_POS_X)[pid] = x_old + 0.5 * dt * (vx_new + vx_old);
        P(RBTP_P_POS_Y)[pid] = y_old + 0.5 * dt * (vy_new + vy_old);
        P(RBTP_P_POS_Z)[pid] = z_old + 0.5 * dt * (vz_new + vz_old);
        P(RBTP_P_VEL_X)[pid] = vx_new;
        P(RBTP_P_VEL_Y)[pid] = vy_new;
        P(RBTP_P_VEL_Z)[pid] = vz_new;
        P(RBTP_P_OMEGA_X)[pid] = omega_x_new;
        P(RBTP_P_OMEGA_Y)[pid] = omega_y_new;
        P(RBTP_P_OMEGA_Z)[pid] = omega_z_new;
        double theta_x = P(RBTP_P_THETA_X)[pid], theta_y = P(RBTP_P_THETA_Y)[pid],
               theta_z = P(RBTP_P_THETA_Z)[pid];
        theta_x += 0.5 * dt * (omega_x_new + omega_x_old);
        theta_y += 0.5 * dt * (omega_y_new + omega_y_old);
        theta_z += 0.5 * dt * (omega_z_new + omega_z_old);
        P(RBTP_P_THETA_X)[pid] = theta_x;
        P(RBTP_P_THETA_Y)[pid] = theta_y;
        P(RBTP_P_THETA_Z)[pid] = theta_z;
        const double heat = P(RBTP_P_HEAT)[pid], heat_prev = P(RBTP_P_HEAT_PREV)[pid];
        double T = P(RBTP_P_TEMP)[pid];
        T += 0.5 * dt * (heat + heat_prev) / heat_capacity;
        P(RBTP_P_TEMP)[pid] = T;
    }
}

// refill_nodes (refill.cu:244-329): new_fluid_count_kernel (refill.cu:8-41), ExclusiveSum,
// new_fluid_fill_kernel (refill.cu:43-76), refill_new_fluid_kernel (refill.cu:78-240), then
// ibnode_prev <- ibnode and owner_prev <- owner (refill.cu:327-328).
void or_refill(OrState *s) {
    if (!s->particles()) return;
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    const int np = s->np;
    const double R = s->d.particle_radius;
    const double *X = s->p[RBTP_P_POS_X].data(), *Y = s->p[RBTP_P_POS_Y].data(),
                 *Z = s->p[RBTP_P_POS_Z].data();
    s->nf_nodes.clear();
    s->nf_pids.clear();
    for (long idx = 0; idx < N; idx++) {
        if (MUT(11) ? (s->ibnode[idx] != 1 || s->ibnode_prev[idx] == 1)
                    : (s->ibnode_prev[idx] != 1 || s->ibnode[idx] == 1))
            continue;
        const int pid = MUT(10) ? s->owner[idx] : s->owner_prev[idx];
        if (pid < 0 || pid >= np) continue;
        const int iz = (int)(idx / LXY);
        const int rem = (int)(idx - (long)iz * LXY);
        const int iy = rem / LX;
        const int ix = rem - iy * LX;
        const double x0 = static_cast<double>(ix) + 0.5;
        const double y0 = static_cast<double>(iy) + 0.5;
        const double z0 = static_cast<double>(iz) + 0.5;
        const double dx = x0 - X[pid], dy = y0 - Y[pid], dz = z0 - Z[pid];
        const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double close_thresh = 1.7606816861659007;
        if (!(MUT(9) ? dist - R < close_thresh : dist - R <= close_thresh)) continue;
        s->nf_nodes.push_back((int)idx);
        s->nf_pids.push_back(pid);
    }
    const double t0 = s->c.t_ref, RHO0 = s->c.rho0;
    const size_t M = s->nf_nodes.size();
    for (size_t tid = 0; tid < M; tid++) {
        const long idx = s->nf_nodes[tid];
        const int iz = (int)(idx / LXY);
        const int iy = (int)((idx - (long)iz * LXY) / LX);
        const int ix = (int)(idx - (long)iz * LXY - (long)iy * LX);
        const double lambda0 = 1.0 - s->tp[0];
        const double RT = 1.0 / 3.0;
        int pid = s->nf_pids[tid];
        if (pid < 0) {
            int prev_pid = s->owner_prev[idx];
            if (prev_pid >= 0) pid = prev_pid;
        }
        double rho_loc = 0.0, mom_x = 0.0, mom_y = 0.0, mom_z = 0.0, phi_loc = 0.0;
        for (int ip = 0; ip < NPOP; ++ip) {
            double temp_f = 0.0, temp_g = 0.0;
            int sample_count = 0;
            for (int ip2 = 1; ip2 < NPOP; ++ip2) {
                int ix1 = ix + CIX[ip2], iy1 = iy + CIY[ip2], iz1 = iz + CIZ[ip2];
                if (ix1 < 0 || ix1 >= LX || iy1 < 0 || iy1 >= LY || iz1 < 0 || iz1 >= LZ) continue;
                const long idx1 = (long)iz1 * LXY + (long)iy1 * LX + ix1;
                if (s->ibnode_prev[idx1] != 0) continue;
                int ix2 = ix + 2 * CIX[ip2], iy2 = iy + 2 * CIY[ip2], iz2 = iz + 2 * CIZ[ip2];
                bool inside2 = (ix2 >= 0 && ix2 < LX && iy2 >= 0 && iy2 < LY && iz2 >= 0 && iz2 < LZ);
                if (inside2 && !MUT(36)) {
                    const long idx2 = (long)iz2 * LXY + (long)iy2 * LX + ix2;
                    if (s->ibnode_prev[idx2] == 0) {
                        temp_f += 2.0 * s->f_collide[ip * N + idx1] - s->f_collide[ip * N + idx2];
                        temp_g += 2.0 * s->g_collide[ip * N + idx1] - s->g_collide[ip * N + idx2];
                        ++sample_count;
                        continue;
                    }
                }
                temp_f += s->f_collide[ip * N + idx1];
                temp_g += s->g_collide[ip * N + idx1];
                ++sample_count;
            }
            if (sample_count == 0) {
                double px = 0.0, py = 0.0, pz = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
                double omega_x = 0.0, omega_y = 0.0, omega_z = 0.0;
                double tempT = t0;
                if (pid >= 0) {
                    px = X[pid]; py = Y[pid]; pz = Z[pid];
                    vx = s->p[RBTP_P_VEL_X][pid]; vy = s->p[RBTP_P_VEL_Y][pid]; vz = s->p[RBTP_P_VEL_Z][pid];
                    omega_x = s->p[RBTP_P_OMEGA_X][pid]; omega_y = s->p[RBTP_P_OMEGA_Y][pid];
                    omega_z = s->p[RBTP_P_OMEGA_Z][pid];
                    tempT = s->p[RBTP_P_TEMP][pid];
                }
                const double x = static_cast<double>(ix) + 0.5;
                const double y = static_cast<double>(iy) + 0.5;
                const double z = static_cast<double>(iz) + 0.5;
                const double relx = x - px, rely = y - py, relz = z - pz;
                double u = vx + (relz * omega_y - rely * omega_z);
                double v = vy + (relx * omega_z - relz * omega_x);
                double w = vz + (rely * omega_x - relx * omega_y);
                if (pid < 0) { u = 0.0; v = 0.0; w = 0.0; }
                const double r = RHO0;
                const double eu = (CIX[ip] * u + CIY[ip] * v + CIZ[ip] * w) / (MUT(35) ? RT : 1.0);
                const double uv = (u * u + v * v + w * w) / RT;
                temp_f = s->tp[ip] * (r + eu + 0.5 * (eu * eu - uv));
                temp_g = s->tp[ip] * tempT * (1.0 + eu + 0.5 * (eu * eu - uv));
                if (ip == 0) {
                    temp_g -= lambda0 * tempT * r;
                } else {
                    temp_g += s->tp[ip] * tempT * r;
                }
                sample_count = 1;
            }
            const double inv_count = 1.0 / static_cast<double>(sample_count);
            const double f_val = temp_f * inv_count;
            const double g_val = temp_g * inv_count;
            s->f[ip * N + idx] = f_val;
            s->g[ip * N + idx] = g_val;
            rho_loc += f_val;
            mom_x += CIX[ip] * f_val;
            mom_y += CIY[ip] * f_val;
            mom_z += CIZ[ip] * f_val;
            phi_loc += g_val;
        }
        s->rho[idx] = rho_loc;
        if (rho_loc > 1e-12) {
            const double inv_rho = 1.0 / rho_loc;
            s->ux[idx] = mom_x * inv_rho;
            s->uy[idx] = mom_y * inv_rho;
            s->uz[idx] = mom_z * inv_rho;
        } else {
            s->ux[idx] = 0.0; s->uy[idx] = 0.0; s->uz[idx] = 0.0;
        }
        s->phi[idx] = phi_loc;
        if (!MUT(37)) { s->fx[idx] = 0.0; s->fy[idx] = 0.0; s->fz[idx] = 0.0; }
    }
    s->ibnode_prev = s->ibnode;
    s->owner_prev = s->owner;
}

// initialize() (initial.cu:240-333) for a fresh start, with init_particles (pinit.cu:183-295).
int or_init(OrState *s, const double *temperature, const double *ppos, const double *ptemp) {
    if (s->c.activate_particles && s->np > 0) {
        for (int k = 0; k < RBTP_P_COUNT; k++) std::fill(s->p[k].begin(), s->p[k].end(), 0.0);
        if (ppos) {
            for (int a = 0; a < 3; a++)
                std::copy(ppos + (size_t)a * s->np, ppos + (size_t)(a + 1) * s->np,
                          s->p[RBTP_P_POS_X + a].begin());
        } else if (place_particles(s) != 0) {
            return RBTP_E_PLACEMENT;
        }
        for (int p = 0; p < s->np; p++)
            s->p[RBTP_P_TEMP][p] = ptemp ? ptemp[p] : s->c.particle_init_temp;
        initial_ibnode(s);
    }
    init_fluid(s, temperature);
    copy_par(s->f, s->f_collide);
    copy_par(s->g, s->g_collide);
    if (s->particles()) or_build_links(s);
    s->step = 0;
    return 0;
}

// Stages of one time step in main.cu order (main.cu:35-85).
enum {
    ST_MACROVAR = 0, ST_FORCING, ST_COLLIDE_F, ST_STREAM_F, ST_IBB_F, ST_COLLIDE_G, ST_STREAM_G,
    ST_IBB_G, ST_PREPARE, ST_REPULSION, ST_FORCES, ST_HEAT, ST_MOVE, ST_BUILD_LINKS, ST_REFILL,
    ST_COUNT
};

int or_stage(OrState *s, int stage) {
    switch (stage) {
        case ST_MACROVAR: or_macrovar(s); break;
        case ST_FORCING: or_forcing(s); break;
        case ST_COLLIDE_F: or_collision_f(s); break;
        case ST_STREAM_F: or_stream_f(s); break;
        case ST_IBB_F: or_ibb_f(s); break;
        case ST_COLLIDE_G: or_collision_g(s); break;
        case ST_STREAM_G: or_stream_g(s); break;
        case ST_IBB_G: or_ibb_g(s); break;
        case ST_PREPARE: or_prepare_forces(s); break;
        case ST_REPULSION: or_repulsion(s); break;
        case ST_FORCES: or_forces(s); break;
        case ST_HEAT: or_heat(s); break;
        case ST_MOVE: or_move(s); break;
        case ST_BUILD_LINKS: or_build_links(s); break;
        case ST_REFILL: or_refill(s); break;
        default: return RBTP_E_ARG;
    }
    return 0;
}

// One or more time steps of main.cu:32-86.
int or_step(OrState *s, int nsteps) {
    for (int n = 0; n < nsteps; n++) {
        for (int st = 0; st < ST_COUNT; st++) or_stage(s, st);
        s->step++;
    }
    return 0;
}

// output_nu (output_statis.cu:111-223) on the current fields and flags.
void or_nu(const OrState *s, double out[7]) {
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    const bool parts = s->particles();
    const std::vector<int> &ibnode = MUT(12) ? s->ibnode_prev : s->ibnode;
    double volume_fraction = s->c.particle_fraction;
    if (MUT(14) && parts) {
        volume_fraction = 0.0;
        for (int v : ibnode) volume_fraction += v != 0;
        volume_fraction /= static_cast<double>(s->N);
    }
    const double fluid_fraction = std::fmax(0.0, 1.0 - volume_fraction);
    double plane_avg_uT_fluid = 0.0, plane_avg_uT_part = 0.0;
    for (int ix = 0; ix < LX; ++ix) {
        double plane_sum_fluid = 0.0, plane_sum_part = 0.0;
        int fluid_count = 0, part_count = 0;
        for (int iz = 0; iz < LZ; ++iz)
            for (int iy = 0; iy < LY; ++iy) {
                const long idx = (long)iz * LXY + (long)iy * LX + ix;
                bool is_fluid = true;
                if (parts) is_fluid = (ibnode[idx] == 0);
                if (is_fluid) {
                    plane_sum_fluid += s->ux[idx] * s->phi[idx];
                    fluid_count++;
                } else {
                    plane_sum_part += s->ux[idx] * s->phi[idx];
                    part_count++;
                }
            }
        if (fluid_count > 0) plane_avg_uT_fluid += plane_sum_fluid / static_cast<double>(fluid_count);
        if (part_count > 0) plane_avg_uT_part += plane_sum_part / static_cast<double>(part_count);
    }
    plane_avg_uT_fluid /= static_cast<double>(LX);
    plane_avg_uT_part /= static_cast<double>(LX);
    const double tHot = s->c.t_hot, tCold = s->c.t_cold;
    const double conductive_flux = s->d.diff * (tHot - tCold) / static_cast<double>(LX);
    double Nu_bulk_fluid = fluid_fraction, Nu_bulk_part = 0.0;
    if (conductive_flux != 0.0) {
        Nu_bulk_fluid += fluid_fraction * (plane_avg_uT_fluid / conductive_flux);
        Nu_bulk_part = volume_fraction * s->c.cp_particle_ratio * (plane_avg_uT_part / conductive_flux);
    }
    const double Nu_bulk = Nu_bulk_fluid + Nu_bulk_part;
    double bottom = 0.0, top = 0.0;
    int bottom_count = 0, top_count = 0;
    for (int iz = 0; iz < LZ; ++iz)
        for (int iy = 0; iy < LY; ++iy) {
            const long idx_bottom = (long)iz * LXY + (long)iy * LX;
            bool bottom_fluid = true;
            if (parts) bottom_fluid = (ibnode[idx_bottom] == 0);
            if (bottom_fluid) {
                bottom += (tHot - s->phi[idx_bottom]) / 0.5;
                bottom_count++;
            }
            const long idx_top = (long)iz * LXY + (long)iy * LX + (LX - 1);
            bool top_fluid = true;
            if (parts) top_fluid = (ibnode[idx_top] == 0);
            if (top_fluid) {
                top += (s->phi[idx_top] - tCold) / 0.5;
                top_count++;
            }
        }
    const double section_area = static_cast<double>(LY) * static_cast<double>(LZ);
    const double wall_scale = static_cast<double>(LX) / (tHot - tCold);
    double Tbottom = 0.0, Ttop = 0.0;
    if (bottom_count > 0) Tbottom = (bottom / (MUT(13) ? bottom_count : section_area)) * wall_scale;
    if (top_count > 0) Ttop = (top / (MUT(13) ? top_count : section_area)) * wall_scale;
    out[0] = (double)s->step;
    out[1] = volume_fraction;
    out[2] = Nu_bulk;
    out[3] = Ttop;
    out[4] = Tbottom;
    out[5] = Nu_bulk_fluid;
    out[6] = Nu_bulk_part;
}

// diag_flow (output_statis.cu:64-109) after overwrite_particle_nodes (output_statis.cu:5-60).
void or_means(const OrState *s, double out[4]) {
    const long N = s->N;
    const int LX = s->LX, LXY = s->LXY;
    double umean = 0.0, vmean = 0.0, wmean = 0.0, tmean = 0.0;
    for (long i = 0; i < N; i++) {
        double u = s->ux[i], v = s->uy[i], w = s->uz[i], t = s->phi[i];
        if (s->particles() && s->ibnode[i] != 0) {
            const int pid = s->owner[i];
            if (pid >= 0 && pid < s->np) {
                const int iz = (int)(i / LXY);
                const int iy = (int)((i - (long)iz * LXY) / LX);
                const int ix = (int)(i - (long)iz * LXY - (long)iy * LX);
                const double relx = (static_cast<double>(ix) + 0.5) - s->p[RBTP_P_POS_X][pid];
                const double rely = (static_cast<double>(iy) + 0.5) - s->p[RBTP_P_POS_Y][pid];
                const double relz = (static_cast<double>(iz) + 0.5) - s->p[RBTP_P_POS_Z][pid];
                const double wx = s->p[RBTP_P_OMEGA_X][pid], wy = s->p[RBTP_P_OMEGA_Y][pid],
                             wz = s->p[RBTP_P_OMEGA_Z][pid];
                u = s->p[RBTP_P_VEL_X][pid] + (relz * wy - rely * wz);
                v = s->p[RBTP_P_VEL_Y][pid] + (relx * wz - relz * wx);
                w = s->p[RBTP_P_VEL_Z][pid] + (rely * wx - relx * wy);
                t = s->p[RBTP_P_TEMP][pid];
            }
        }
        umean += u; vmean += v; wmean += w; tmean += t;
    }
    out[0] = umean / N; out[1] = vmean / N; out[2] = wmean / N; out[3] = tmean / N;
}


// The seven %09d.dat fields of output_flow (output_flow.cu:64-110): compute_Ked
// (collision.cu:91-152) on f_collide, compute_ted (collision.cu:154-198) on g_collide, then
// overwrite_particle_nodes (output_flow.cu:5-60).  out: 7*N doubles in file order.
void or_output_fields(const OrState *s, double *out) {
    const long N = s->N;
    const int LX = s->LX, LXY = s->LXY;
    const double tau = s->d.tau, tauc = s->d.tauc, visc = s->c.visc, diff = s->d.diff;
#pragma omp parallel for schedule(static)
    for (long idx = 0; idx < N; idx++) {
        const double fxv = s->fx[idx], fyv = s->fy[idx], fzv = s->fz[idx];
        const double rho_local = s->rho[idx], phi_local = s->phi[idx];
        const double ux_local = s->ux[idx], uy_local = s->uy[idx], uz_local = s->uz[idx];
        const double RT = 1.0 / 3.0;
        const double u2 = ux_local * ux_local + uy_local * uy_local + uz_local * uz_local;
        const double uv = u2 / RT;
        const double bu = fxv * ux_local + fyv * uy_local + fzv * uz_local;
        double tau_xx = 0.0, tau_xy = 0.0, tau_xz = 0.0, tau_yy = 0.0, tau_yz = 0.0, tau_zz = 0.0;
        for (int ip = 0; ip < NPOP; ip++) {
            double eu = (CIX[ip] * ux_local + CIY[ip] * uy_local + CIZ[ip] * uz_local) / RT;
            double feq = s->tp[ip] * (rho_local + eu + 0.5 * (eu * eu - uv));
            double fneq = s->f_collide[ip * N + idx] - feq;
            double cidotF = CIX[ip] * fxv + CIY[ip] * fyv + CIZ[ip] * fzv;
            double feq_term = MUT(15) ? 0.0 : 0.5 * (cidotF - bu) * feq / RT;
            tau_xx += CIX[ip] * CIX[ip] * (fneq + feq_term);
            tau_xy += CIX[ip] * CIY[ip] * (fneq + feq_term);
            tau_xz += CIX[ip] * CIZ[ip] * (fneq + feq_term);
            tau_yy += CIY[ip] * CIY[ip] * (fneq + feq_term);
            tau_yz += CIY[ip] * CIZ[ip] * (fneq + feq_term);
            tau_zz += CIZ[ip] * CIZ[ip] * (fneq + feq_term);
        }
        double coeff = -(1.0 - 0.5 / tau);
        tau_xx *= coeff; tau_xy *= coeff; tau_xz *= coeff;
        tau_yy *= coeff; tau_yz *= coeff; tau_zz *= coeff;
        double S_xx = tau_xx / (2.0 * visc), S_xy = tau_xy / (2.0 * visc), S_xz = tau_xz / (2.0 * visc);
        double S_yy = tau_yy / (2.0 * visc), S_yz = tau_yz / (2.0 * visc), S_zz = tau_zz / (2.0 * visc);
        double Ked = 0.5 * visc * (4.0 * (S_xx * S_xx + S_yy * S_yy + S_zz * S_zz) +
                                   8.0 * (S_xy * S_xy + S_xz * S_xz + S_yz * S_yz));
        const double lambda0 = 1.0 - s->tp[0];
        double gneqx = 0.0, gneqy = 0.0, gneqz = 0.0;
        for (int ip = 0; ip < NPOP; ip++) {
            double eu = (CIX[ip] * ux_local + CIY[ip] * uy_local + CIZ[ip] * uz_local) / RT;
            double geq = s->tp[ip] * phi_local * (1.0 + eu + 0.5 * (eu * eu - uv));
            double geqnew = (ip == 0) ? (geq - lambda0 * phi_local * rho_local)
                                      : (geq + s->tp[ip] * phi_local * rho_local);
            double dg = s->g_collide[ip * N + idx] - geqnew;
            gneqx += CIX[ip] * dg;
            gneqy += CIY[ip] * dg;
            gneqz += CIZ[ip] * dg;
        }
        double p = rho_local * RT;
        double denominator = 2.0 * tauc * RT + (MUT(16) ? (1.0 + rho_local) * RT : p);
        double tx = -(2.0 * gneqx + phi_local * fxv) / denominator;
        double ty = -(2.0 * gneqy + phi_local * fyv) / denominator;
        double tz = -(2.0 * gneqz + phi_local * fzv) / denominator;
        double ted = (diff > 0.0) ? diff * (tx * tx + ty * ty + tz * tz) : 0.0;
        double r = rho_local, u = ux_local, v = uy_local, w = uz_local, T = phi_local;
        if (s->particles() && s->ibnode[idx] != 0) {
            const int pid = s->owner[idx];
            if (pid >= 0 && pid < s->np) {
                const int iz = (int)(idx / LXY);
                const int iy = (int)((idx - (long)iz * LXY) / LX);
                const int ix = (int)(idx - (long)iz * LXY - (long)iy * LX);
                const double relx = (static_cast<double>(ix) + 0.5) - s->p[RBTP_P_POS_X][pid];
                const double rely = (static_cast<double>(iy) + 0.5) - s->p[RBTP_P_POS_Y][pid];
                const double relz = (static_cast<double>(iz) + 0.5) - s->p[RBTP_P_POS_Z][pid];
                const double wx = s->p[RBTP_P_OMEGA_X][pid], wy = s->p[RBTP_P_OMEGA_Y][pid],
                             wz = s->p[RBTP_P_OMEGA_Z][pid];
                r = s->c.rho0;
                const double rot = MUT(17) ? 0.0 : 1.0;
                u = s->p[RBTP_P_VEL_X][pid] + rot * (relz * wy - rely * wz);
                v = s->p[RBTP_P_VEL_Y][pid] + rot * (relx * wz - relz * wx);
                w = s->p[RBTP_P_VEL_Z][pid] + rot * (rely * wx - relx * wy);
                T = s->p[RBTP_P_TEMP][pid];
                ted = 0.0;
                Ked = 0.0;
            }
        }
        out[idx] = r; out[N + idx] = u; out[2 * N + idx] = v; out[3 * N + idx] = w;
        out[4 * N + idx] = T; out[5 * N + idx] = ted; out[6 * N + idx] = Ked;
    }
}

// output_profile (output_statis.cu:251-281) on the output fields: out[4*LX].
void or_profile(const OrState *s, const double *fields, double *out) {
    const long N = s->N;
    const int LX = s->LX, LY = s->LY, LZ = s->LZ, LXY = s->LXY;
    for (int ix = 0; ix < LX; ix++) {
        double sum_ux = 0.0, sum_uy = 0.0, sum_uz = 0.0, sum_phi = 0.0;
        int count = 0;
        for (int iz = 0; iz < LZ; iz++)
            for (int iy = 0; iy < LY; iy++) {
                const long idx = (long)iz * LXY + (long)iy * LX + ix;
                sum_ux += fields[N + idx];
                sum_uy += fields[2 * N + idx];
                sum_uz += fields[3 * N + idx];
                sum_phi += fields[4 * N + idx];
                count++;
            }
        const double inv_count = (count > 0) ? (1.0 / static_cast<double>(count)) : 0.0;
        out[4 * ix] = sum_ux * inv_count;
        out[4 * ix + 1] = sum_uy * inv_count;
        out[4 * ix + 2] = sum_uz * inv_count;
        out[4 * ix + 3] = sum_phi * inv_count;
    }
}

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
