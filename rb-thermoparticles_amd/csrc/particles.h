// particles.h — particle coupling and diagnostics (particles.hip, diag.hip).
#pragma once

#include <algorithm>
#include <utility>
#include <vector>

#include "../../include/rbtp.h"
#include "rbtp_internal.h"

namespace rbtp {

// Indices of the snapshot of the particle state before the motion of the current step.
enum { OLD_PX = 0, OLD_PY, OLD_PZ, OLD_VX, OLD_VY, OLD_VZ, OLD_WX, OLD_WY, OLD_WZ, OLD_T, OLD_COUNT };

// Error flags raised by device code (checked after every synchronising call).  The reference's
// other device-side report, a link whose node is solid (ibb.cu:79,130), cannot occur: links
// start only at fluid nodes.
enum { ERR_SPECIAL_CAP = 0, ERR_COUNT = 1 };

struct PartDev {
    int np;
    // Case constants (globals.cu:126-148, pforces.cu:121-124).
    double radius, radius_sq, mass, inertia, heat_capacity, rho_ratio, grav0, coeff;
    double min_gap, min_gap_w, stf0, stf1, stf0_w, stf1_w, rho0, t0, thot, tcold;
    const float *gshift;     // shift profile of the reduced-precision g encoding, [lx] (the sim's)
    double *base;            // one allocation for every P-length FP64 array
    double *p[RBTP_P_COUNT]; // particle state, indexed by RBTP_P_*
    double *old[OLD_COUNT];  // state before this step's motion
    int32_t *owner[2];       // solid-node owner (-1: fluid), two geometry generations
    int32_t *box[2];         // marking box per particle and generation: lo xyz, hi xyz
    int geo_cur;             // generation holding the newest geometry
    SpecialBuf sp;
    int32_t *sp_kind;        // -1: IBB link node; >= 0: refilled node, value = previous owner
    size_t sp_elem;          // bytes per value of sp.pop (the precision form's auxiliary type)
    bool special_pending;    // sp holds corrected populations for the next lattice step
    int32_t *err;            // ERR_COUNT device flags
    int box_w;               // upper bound of marking-box nodes per axis
    // Uniform cell grid of the pair search: cells at least as wide as the pair range
    // 2R + REP_MIN_GAP, each holding up to CELL_CAP particle ids.
    int ncell[3];
    double cell_scale[3];    // cells per lattice unit along each axis
    int32_t *cell_cnt;       // [ncells] particles binned into each cell (may exceed CELL_CAP)
    int32_t *cell_pid;       // [ncells][CELL_CAP]
    bool use_cells;          // pair search through the cell grid (else every pair)
    // Each link-box node's share of its particle's force, torque and heat sums: lnk_stride
    // entries per particle (the largest link box, padded to whole waves), 7 planes of np *
    // lnk_stride values, and one bit per entry (64-bit words) telling whether the node has links.
    int lnk_stride;
    double *lnk_part;
    unsigned long long *lnk_bits;
};

// Link-box node threads per particle: the largest link box ((box_w + 2)^3 nodes, box_w the
// largest marking box per axis) rounded up to whole 64-lane waves.
inline int link_stride(int box_w) {
    const int b = box_w + 2;
    return (b * b * b + 63) / 64 * 64;
}

// Entries of the special-node list: a head of one entry per link-box node of every particle
// (IBB link nodes, at fixed positions) and a tail for the refilled nodes, which lie in the
// marking boxes and are distinct lattice nodes.
inline long long special_capacity(int np, int box_w, long long n) {
    const long long mb = (long long)box_w * box_w * box_w;
    const long long tail = std::min((long long)np * mb, n);
    return std::max((long long)np * link_stride(box_w) + tail, 1LL);
}


// generate_initial_positions (pinit.cu:97-144) on the host: d.npart positions into X, Y, Z
// (resized by the caller); RBTP_E_PLACEMENT after 10^6 rejections.
int place_particles(const rbtp_case &c, const rbtp_derived &d, std::vector<double> &X,
                    std::vector<double> &Y, std::vector<double> &Z);
int part_alloc(PartDev &pd, const rbtp_case &c, const rbtp_derived &d, const Dims &dims);
void part_free(PartDev &pd);
int part_init(PartDev &pd, const rbtp_case &c, const rbtp_derived &d, const Dims &dims,
              const double *ppos, const double *ptemp, hipStream_t st);
int part_load_state(PartDev &pd, const double *host, const Dims &dims, hipStream_t st);
// One lattice's share of a particle phase: the particle state and the lattice arrays the
// phase reads and writes at one time step.  Kept in device memory, so that one launch of each
// particle kernel serves every lattice of a serving batch (or the one lattice of an rbtp_sim).
struct PartSlot {
    PartDev pd;
    Dims d;
    const int32_t *owner_prev, *box_prev;  // geometry of this step's IBB and refill
    int32_t *owner_curr, *box_curr;        // geometry built from the moved particles
    const void *fcur, *gcur;               // this step's post-collision populations
    void *fx;                              // F_x (Enc<S>::A)
    const void *rho_wall;                  // wall rho (Enc<S>::A)
    double *macro[5];                      // output fields rho, ux, uy, uz, phi
    const void *rho_full;                  // rho of the last lattice step (Enc<S>::A)
};

// Launch shape of a particle phase over a table of slots.  The prefix arrays map a flat index
// (a particle, a pair-search cell, a fix-up workgroup) to its slot.
struct PartGrid {
    const PartSlot *tab;
    const int *pstart;  // [nslots + 1] particles
    const int *cstart;  // [nslots + 1] pair-search cells
    const int *sstart;  // [nslots + 1] workgroups of the refill (4 refilled nodes each)
    const int *lstart;  // [nslots + 1] workgroups over the special list's head (128 entries
                        // each): the unmap and the link-node fix-up
    const int *bstart;  // [nslots + 1] link-box node threads (np * lnk_stride per slot)
    int nslots, total_p, total_c, total_s, total_l, total_b;
};

// Device tables for the two phases of a time step (the geometry generations and the
// population buffers alternate every step).
struct PartTables {
    char *mem = nullptr;
    size_t bytes = 0;
    PartGrid g[2] = {};
};

int part_tables_build(PartTables &t, const std::vector<PartSlot> (&slots)[2], hipStream_t st);
void part_tables_free(PartTables &t);
// Workgroups of 128 threads the special-list kernels give to one lattice.  Their loops stride
// over the grid, so these counts set only the parallelism: the list head (one entry per
// link-box node) at 128 entries per workgroup, and the refilled nodes of a step at one
// workgroup of four per sphere (a sphere uncovers a few nodes per step at most).
int part_head_blocks(const PartDev &pd);
int part_refill_blocks(const PartDev &pd);
// The particle phase (main.cu:52-84 minus the lattice passes) of every slot of g, in precision
// form prec.  The caller then advances each slot's geo_cur and sets special_pending.
int part_phase(const PartGrid &g, int prec, hipStream_t st);
int part_check_errors(PartDev &pd);
// One particle stage of the per-stage execution (RBTP_STAGE_IBB_F, IBB_G and PREPARE ...
// REFILL) over the slots of g, whose fcur/gcur are d_f_collide/d_g_collide and whose geometry
// generations are those of the step's links (owner_prev) and of build_links (owner_curr);
// fo/go are d_f/d_g (FP64, blocked layout).
int part_stage(int stage, const PartGrid &g, double *fo, double *go, hipStream_t st);
// part_overwrite_special into populations in the blocked layout of pop_at.
hipError_t part_overwrite_special_blocked(PartDev &pd, double *f, double *g, const Dims &dims,
                                          hipStream_t st);
hipError_t part_overwrite_special(PartDev &pd, double *f, double *g, const Dims &dims,
                                  hipStream_t st);
// Link records of geometry generation gen (pd.owner[gen]) in the reference layout, from the
// current positions or (snapshot) from those before the last update_particles.
int part_collect_links(PartDev &pd, int gen, bool snapshot, const Dims &dims, hipStream_t st,
                       std::vector<rbtp_link> &out);
int part_collect_new_fluid(PartDev &pd, hipStream_t st,
                           std::vector<std::pair<int32_t, int32_t>> &out);

// Constants of the output-field kernels (compute_Ked / compute_ted, overwrite_particle_nodes).
struct OutK {
    double tau, tauc, visc, diff, t0, rho0;
    const float *gshift;  // shift profile of the reduced-precision g encoding, [lx]
};

int diag_output_fields(int prec, const void *fc, const void *gc, const void *fx,
                       double *const *macro, const PartDev *pd, const OutK &k, const Dims &dims,
                       hipStream_t st, double *dev_out);
int diag_profile(const double *dev_fields, const Dims &dims, hipStream_t st, double *out);
int diag_nu(const rbtp_case &c, const rbtp_derived &d, const Dims &dims, const double *ux,
            const double *phi, const int32_t *owner, int step, hipStream_t st, double out[7]);
int diag_nu_noalloc(const rbtp_case &c, const rbtp_derived &d, const Dims &dims, const double *ux,
                    const double *phi, const int32_t *owner, int step, hipStream_t st,
                    double out[7]);
int diag_means(const Dims &dims, double *const *macro, const PartDev *pd, hipStream_t st,
               double out[4]);
int diag_copy_probe(size_t bytes, int reps, double *gbs);
int diag_kinetic_energy(const Dims &dims, double *const *macro, const int32_t *owner,
                        const int32_t *owner_prev, hipStream_t st, double *out);

}  // namespace rbtp
