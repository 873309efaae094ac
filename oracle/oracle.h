/*
 * oracle.h — FP64 host restatement of the reference kernels.  TEST INFRASTRUCTURE ONLY: loaded
 * by tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py, never by the product.
 *
 * How each function is pinned independently of its own formula (tests in tests/test_oracle.py
 * unless noted).  The reference ships no tests, fixtures or golden vectors and cannot be built
 * here (CUDA + CUB, no nvcc), so every pin below comes from mathematics or brute force:
 *
 *   or_case_derive        closed forms of tau, tauc, grav0, u0, radius, mass, npart (hand values
 *                         for parameters.h defaults and the benchmark cases).
 *   D3Q27 tables          opp∘opp = id, c_opp = -c, sum w = 1, sum w c c = I/3, 4th-order
 *                         isotropy, 27 distinct velocities.
 *   or_macrovar           moments of random populations against numpy sums (rho, u+F/2, phi).
 *   or_collision_f        mass conservation, momentum change = Guo source first moment, rest
 *                         state with u=F=0 is a fixed point (f=0), feq moments.
 *   or_collision_g        zeroth moment of g after collision equals sum of geqnew (= phi); the
 *                         first moment against its closed form G(1 - 1/tauc) + Rc[phi F -
 *                         rho(2G + phi F)/(3D)] on random states; per population, what remains
 *                         after the relaxation is w_i c_i.A with A three times that first-moment
 *                         source, so every population of the source term (collision.cu:55-87)
 *                         is fixed; rest state g = w*T (uniform T, u=F=0) stays put bitwise
 *                         over many full steps.
 *   or_stream_f/g         single-population delta moves to x+c, reflects into opp at walls, the
 *                         x-wall anti-bounce-back value 2 w T_wall (1+rho) - g_opp, z/y/x
 *                         precedence at edges; mass conservation of f over a full step.
 *   init_fluid            uniform T: g_0 = T(2 w_0 - 1), g_i = 2 w_i T; phi = T recovered.
 *   compute_link_q        axis-aligned links against the analytic intersection; brute-force
 *                         sampling of the segment on random spheres.
 *   or_build_links        brute-force enumeration on tiny grids; sphere centred on a cell centre
 *                         gives a symmetric link set; lowest pid keeps shared nodes.
 *   or_ibb_f              Bouzidi's linear interpolation on post-collision fields linear in
 *                         position with 27 distinct coefficients (f*_i != f*_opp), moving and
 *                         rotating spheres: q < 1/2 is f*_i at x - (1-2q) c_i minus mt, the
 *                         other branches their closed forms; every branch is hit, including
 *                         upstream solid (between two spheres) and outside the domain
 *                         (tests/test_oracle_pins.py).
 *   or_forces             a sphere at rest in fluid at rest: zero force and torque by symmetry.
 *                         A sphere moving through fluid at rest (f* = 0): each link writes
 *                         -mt or -mt/(2q), so force and torque are closed-form sums over the
 *                         link list, which exercise both q branches.
 *   or_ibb_g              the same on linear g* fields with a non-uniform density deviation
 *                         (the (1 + rho(x)) factor), per-sphere T_p, and a previous geometry that
 *                         differs from the current one (ibnode_prev decides 'upstream fluid').
 *   or_heat               a sphere at T_p in fluid held at the fixed point g = w T_f: closed-form
 *                         heat over the link list, both q branches including the g*_opp term.
 *   or_repulsion          two spheres and a sphere at a wall at known gaps (closed form).
 *   or_move               constant force and heat: closed-form trapezoid sequences; free fall.
 *   new-fluid predicate   brute force over every node of a lattice where one sphere jumps 3
 *                         nodes and another leaves a node exactly on close_thresh (radius
 *                         1.75 + 2^-52 so that the distance minus R is exactly the literal).
 *   or_refill             a new-fluid node with no prev-fluid neighbour (a sphere that jumps 3
 *                         nodes): the fallback's moments are rho = 1 - 4|u|^2/3, momentum u/3
 *                         (e.u not divided by RT), phi = T_p (1 - 4|u|^2/3), F = 0.  Linear
 *                         pre-stream fields make every second-order extrapolation exact, so the
 *                         average (refill.cu:121-153) has a closed form in the first-order
 *                         directions.  rho, u, phi written consistently with the populations.
 *   or_nu                 exact linear conduction profile with u=0: wall Nu = 1 and
 *                         Nu_bulk = 1 - phi_p; with particle nodes, fields constant on the fluid
 *                         and on the particle nodes of each plane give closed-form plane averages,
 *                         bulk and wall Nu (PARTICLE_FRACTION, current flags, full-section
 *                         wall normalisation).
 *   or_output_fields      compute_Ked / compute_ted on equilibrium plus a known non-equilibrium
 *                         part with a force: closed forms from the D3Q27 fourth moments;
 *                         overwrite_particle_nodes: rigid-body velocity, T_p, RHO0, zeros.
 *   or_means, or_profile  lattice and plane means of the overwritten fields.
 *   place_particles       non-overlap and window constraints; same glibc rand sequence as the
 *                         library (same call, same glibc).
 *
 * Each pin in tests/test_oracle_pins.py also asserts that the mistakes it targets (x and x - c
 * or i and opp swapped, the (1 + rho) factor dropped, ibnode for ibnode_prev, < for <=, the old
 * centre or the current owner in the new-fluid test, the previous flags or a fluid-count wall
 * normalisation in Nu, the force term of compute_Ked, rho + 1 in compute_ted, the rotation in the
 * overwrite) move the result far beyond its tolerance.  Thirty-eight plausible mistakes, these
 * and others across macrovar, both collisions, both streamings, link q, marking, the momentum
 * exchange, heat, repulsion, the integrator, refill and init_g, are planted in oracle.cpp behind
 * ORACLE_MUTANT (off in every normal build); tests/test_oracle_mutants.py builds the oracle once
 * per mistake and shows that the targeted pin fails on that build, while the unmutated build
 * passes.
 *
 * Summation order: see the header of oracle.cpp (deterministic per-particle reductions).
 *
 * Drift: tests/golden/ holds the oracle's own output on three small cases (fluid only, one
 * sphere, two moving spheres with refill), written by tests/golden/make_fixtures.py; the oracle
 * must reproduce them bit for bit (tests/test_golden.py).  These fixtures pin the oracle against
 * its own past, not against the reference.
 */
#ifndef RBTP_ORACLE_H
#define RBTP_ORACLE_H

#include "../include/rbtp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrState OrState;

int or_abi_version(void);
void or_case_default(rbtp_case *c);
void or_case_derive(const rbtp_case *c, rbtp_derived *d);
OrState *or_create(const rbtp_case *c);
void or_destroy(OrState *s);
int or_init(OrState *s, const double *temperature, const double *ppos, const double *ptemp);
int or_stage(OrState *s, int stage);
int or_step(OrState *s, int nsteps);
int or_npart(const OrState *s);
long or_nodes(const OrState *s);
int or_current_step(const OrState *s);
void or_get_derived(const OrState *s, rbtp_derived *d);
long or_unhandled_links(const OrState *s, int thermal);
double *or_array(OrState *s, int which);
int *or_iarray(OrState *s, int which);
long or_nlinks(const OrState *s);
const rbtp_link *or_links(const OrState *s);
long or_nnew_fluid(const OrState *s);
const int *or_new_fluid_nodes(const OrState *s);
const int *or_new_fluid_pids(const OrState *s);
void or_linear_temperature(const rbtp_case *c, uint64_t seed, double amp, double *out);
void or_nu(const OrState *s, double out[7]);
void or_means(const OrState *s, double out[4]);
void or_output_fields(const OrState *s, double *out);
void or_profile(const OrState *s, const double *fields, double *out);
int or_threads(void);

#ifdef __cplusplus
}
#endif

#endif
