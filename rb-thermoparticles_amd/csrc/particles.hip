// particles.hip — particle coupling on gfx950: geometry, IBB, force/torque/heat sums, contacts,
// motion, new-fluid detection and refill (src_part/*.cu of the reference).
//
// Links are never stored.  A link (x, i, p) exists when x is a fluid node, x + c_i is a solid
// node owned by p, and the ray from x's centre along c_i meets sphere p at q in [0, 1]
// (link_count_kernel, links.cu:50-87).  Each consumer re-derives the links it needs from the two
// owner maps with the same bit-exact test:
//   * the link-node pass (in k_contacts_links): one thread per node of each particle's link box
//     finds the node's links to the particle and stores the node's share of force, torque and
//     heat; k_link_reduce sums the shares of each particle in a fixed order (64 lane partials +
//     pairwise tree),
//     replacing the reference's FP64 atomics (ibb.cu:182-187,204) with a reproducible reduction;
//   * the link-node pass also lists the IBB link nodes (one fixed entry per link-box node), and
//     k_refill_find appends the refilled nodes: the list of the nodes whose post-stream
//     populations the coupling replaces;
//   * k_fixups (the shared link nodes, then the refilled nodes): those nodes' corrected
//     post-stream populations (pull + IBB, or refill), buffered for the next lattice step.
// Geometry is two owner maps (previous and current generation); each generation is cleared only
// inside the marking boxes it was built from, so a step touches O(P R^3) nodes instead of the
// reference's full-lattice memsets, scans and host round trips (links.cu:147-256,
// refill.cu:244-329).  Nothing is read back to the host inside a step.
//
// The step kernels take a PartGrid: a table of slots (one lattice each) in device memory plus
// prefix arrays that map a flat particle, cell or workgroup index to its slot, so one launch
// serves one lattice or every lattice of a serving batch.
//
// The translation unit is compiled with -ffp-contract=off: every expression is evaluated in the
// reference's operation order.
#include "particles.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "lattice_device.h"

namespace rbtp {

namespace {

constexpr int PB = 256;   // threads per particle workgroup of the box clear/mark/collect kernels
constexpr int CELL_CAP = 16;  // particle ids per cell of the pair-search grid
// Pair candidates per particle held in LDS.  A 3x3x3 block of cells, each 2R + REP_MIN_GAP
// wide, holds at most about 74 non-overlapping spheres, so clustered spheres stay on the cell
// path; beyond 96 a particle checks every partner.
constexpr int CAND_MAX = 96;
constexpr int CELLS_MIN_NP = 32;  // particle count from which the pair search uses the grid
constexpr int LB = 1024;  // threads per particle workgroup of the special-node search

__device__ inline double wt(int i) {
    return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
}

// compute_link_q (links.cu:7-23)
__device__ inline bool link_q(int ip, double x0, double y0, double z0, double px, double py,
                              double pz, double radius_sq, double &q) {
    const double cx_dir = static_cast<double>(dir_cx(ip));
    const double cy_dir = static_cast<double>(dir_cy(ip));
    const double cz_dir = static_cast<double>(dir_cz(ip));
    const double dx = x0 - px;
    const double dy = y0 - py;
    const double dz = z0 - pz;
    const double a = cx_dir * cx_dir + cy_dir * cy_dir + cz_dir * cz_dir;
    const double b = 2.0 * (cx_dir * dx + cy_dir * dy + cz_dir * dz);
    const double c = dx * dx + dy * dy + dz * dz - radius_sq;
    const double disc = b * b - 4.0 * a * c;
    if (disc <= 0.0) return false;
    // q = (-b - sqrt(disc)) / (2a), a in {1, 2, 3}: the divisions by 2 and 4 are exact scalings
    // and the one by 6 is div_rn with RN(1/6) + RN(1/6 - RN(1/6)), the same correctly rounded
    // quotient in three FMAs
    const double num = -b - sqrt(disc);
    q = (a == 3.0) ? div_rn(num, 6.0, RECIP6_HI, RECIP6_LO) : num * (a == 1.0 ? 0.5 : 0.25);
    return (q >= 0.0 && q <= 1.0);
}

// Whether link_q finds a link from a node x0 outside the sphere to its neighbour x0 + c_ip
// inside it.  The quadratic of link_q is then positive at q = 0 and negative at q = 1, so its
// smaller root lies in (0, 1); when both values clear a margin far above the rounding of
// link_q (1e-6 against errors near 1e-14 for spheres of a few lattice units), the root link_q
// computes does too, and no square root is needed.  Otherwise link_q itself decides.
__device__ inline bool link_exists(int ip, double x0, double y0, double z0, double px,
                                   double py, double pz, double radius_sq) {
    const double cx_dir = static_cast<double>(dir_cx(ip));
    const double cy_dir = static_cast<double>(dir_cy(ip));
    const double cz_dir = static_cast<double>(dir_cz(ip));
    const double dx = x0 - px;
    const double dy = y0 - py;
    const double dz = z0 - pz;
    const double a = cx_dir * cx_dir + cy_dir * cy_dir + cz_dir * cz_dir;
    const double b = 2.0 * (cx_dir * dx + cy_dir * dy + cz_dir * dz);
    const double c = dx * dx + dy * dy + dz * dz - radius_sq;
    if (c > 1e-6 && (a + b) + c < -1e-6) return true;
    double q;
    return link_q(ip, x0, y0, z0, px, py, pz, radius_sq, q);
}

// Marking box of a particle (links.cu:164-177); false when empty.
__host__ __device__ inline bool marking_box(double px, double py, double pz, double R,
                                            const Dims &d, int lo[3], int hi[3]) {
    lo[0] = static_cast<int>(floor(px - R - 0.5));
    hi[0] = static_cast<int>(floor(px + R - 0.5));
    lo[1] = static_cast<int>(floor(py - R - 0.5));
    hi[1] = static_cast<int>(floor(py + R - 0.5));
    lo[2] = static_cast<int>(floor(pz - R - 0.5));
    hi[2] = static_cast<int>(floor(pz + R - 0.5));
    const int L[3] = {d.lx, d.ly, d.lz};
    for (int a = 0; a < 3; a++) {
        if (lo[a] < 0) lo[a] = 0;
        if (hi[a] > L[a] - 1) hi[a] = L[a] - 1;
    }
    return !(lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]);
}

__device__ inline bool load_box(const int32_t *box, int p, int lo[3], int hi[3]) {
    for (int a = 0; a < 3; a++) {
        lo[a] = box[p * 6 + a];
        hi[a] = box[p * 6 + 3 + a];
    }
    return !(lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]);
}

// The link box: marking box grown by one node per side, clamped.
__device__ inline void grow_box(const Dims &d, int lo[3], int hi[3]) {
    const int L[3] = {d.lx, d.ly, d.lz};
    for (int a = 0; a < 3; a++) {
        lo[a] = lo[a] - 1 < 0 ? 0 : lo[a] - 1;
        hi[a] = hi[a] + 1 > L[a] - 1 ? L[a] - 1 : hi[a] + 1;
    }
}

__device__ inline int box_node(const Dims &d, const int lo[3], const int hi[3], int k, int &ix,
                               int &iy, int &iz) {
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
    const int kz = k / (nx * ny);
    const int r = k - kz * nx * ny;
    const int ky = r / nx;
    const int kx = r - ky * nx;
    ix = lo[0] + kx;
    iy = lo[1] + ky;
    iz = lo[2] + kz;
    return iz * d.lxy + iy * d.lx + ix;
}

__device__ inline int box_count(const int lo[3], const int hi[3]) {
    return (hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
}

__device__ inline bool in_domain(const Dims &d, int x, int y, int z) {
    return x >= 0 && x < d.lx && y >= 0 && y < d.ly && z >= 0 && z < d.lz;
}

// Owners of the 26 neighbours of (ix, iy, iz), all loads issued together (addresses clamped into
// the domain; -2 where the neighbour is outside it).  With `interior` (the node and its
// neighbours inside the domain) there is no clamping at all.
__device__ inline void neighbour_owners(gptr<const int32_t> owner, const Dims &d, int ix,
                                        int iy, int iz, bool interior, int (&own)[NPOP]) {
    own[0] = -2;
    if (interior) {
        const int idx = iz * d.lxy + iy * d.lx + ix;
#pragma unroll
        for (int ip = 1; ip < NPOP; ip++)
            own[ip] = owner[idx + D3Q27::cx[ip] + D3Q27::cy[ip] * d.lx + D3Q27::cz[ip] * d.lxy];
        return;
    }
#pragma unroll
    for (int ip = 1; ip < NPOP; ip++) {
        const int nx = ix + D3Q27::cx[ip], ny = iy + D3Q27::cy[ip], nz = iz + D3Q27::cz[ip];
        const int cx = min(max(nx, 0), d.lx - 1);
        const int cy = min(max(ny, 0), d.ly - 1);
        const int cz = min(max(nz, 0), d.lz - 1);
        own[ip] = owner[cz * d.lxy + cy * d.lx + cx];
    }
    // An empty statement that takes the 26 values as read and written: the compiler cannot
    // move a load past it (into the branch of the select below), so the loads stay one batch
    // with a single wait.
    asm volatile(""
                 : "+v"(own[1]), "+v"(own[2]), "+v"(own[3]), "+v"(own[4]), "+v"(own[5]),
                   "+v"(own[6]), "+v"(own[7]), "+v"(own[8]), "+v"(own[9]), "+v"(own[10]),
                   "+v"(own[11]), "+v"(own[12]), "+v"(own[13]), "+v"(own[14]), "+v"(own[15]),
                   "+v"(own[16]), "+v"(own[17]), "+v"(own[18]), "+v"(own[19]), "+v"(own[20]),
                   "+v"(own[21]), "+v"(own[22]), "+v"(own[23]), "+v"(own[24]), "+v"(own[25]),
                   "+v"(own[26]));
#pragma unroll
    for (int ip = 1; ip < NPOP; ip++)
        if (!in_domain(d, ix + D3Q27::cx[ip], iy + D3Q27::cy[ip], iz + D3Q27::cz[ip])) own[ip] = -2;
}

// ---- slot lookup -------------------------------------------------------------------------------

// Slot that owns flat index u under the prefix array start[0..n] (start[0] = 0, non-decreasing;
// slots with an empty range own nothing).
__device__ inline int slot_of(const int *start, int n, int u) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= u) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- geometry --------------------------------------------------------------------------------

// The owner map generation that is about to be rebuilt goes back to -1 inside particle p's box
// (nthreads threads share the box; this is thread tid).
__device__ inline void geo_clear_body(int32_t *owner, const int32_t *box, const Dims &d, int p,
                                      int tid, int nthreads) {
    int lo[3], hi[3];
    if (!load_box(box, p, lo, hi)) return;
    const int cnt = box_count(lo, hi);
    for (int k = tid; k < cnt; k += nthreads) {
        int ix, iy, iz;
        owner[box_node(d, lo, hi, k, ix, iy, iz)] = -1;
    }
}

__global__ __launch_bounds__(PB) void k_geo_clear(PartGrid g) {
    const int s = slot_of(g.pstart, g.nslots, blockIdx.x);
    const PartSlot &sl = g.tab[s];
    geo_clear_body(sl.owner_curr, sl.box_curr, sl.d, blockIdx.x - g.pstart[s], threadIdx.x, PB);
}

// mark_ibnode_kernel (links.cu:25-48) for every particle at once.  The reference launches the
// particles one after another so the lowest pid claims a shared node; atomicMin on the owner
// (-1 read as the largest unsigned value) gives the same result in one launch.
__device__ inline void geo_mark_body(int32_t *owner, int32_t *box, const double *px,
                                     const double *py, const double *pz, double R,
                                     double radius_sq, const Dims &d, int p) {
    int lo[3], hi[3];
    const double X = px[p], Y = py[p], Z = pz[p];
    const bool ok = marking_box(X, Y, Z, R, d, lo, hi);
    if (threadIdx.x == 0) {
        for (int a = 0; a < 3; a++) {
            box[p * 6 + a] = ok ? lo[a] : 1;
            box[p * 6 + 3 + a] = ok ? hi[a] : 0;
        }
    }
    if (!ok) return;
    const int cnt = box_count(lo, hi);
    for (int k = threadIdx.x; k < cnt; k += PB) {
        int ix, iy, iz;
        const int idx = box_node(d, lo, hi, k, ix, iy, iz);
        const double cx = static_cast<double>(ix) + 0.5;
        const double cy = static_cast<double>(iy) + 0.5;
        const double cz = static_cast<double>(iz) + 0.5;
        const double dx = cx - X, dy = cy - Y, dz = cz - Z;
        if (dx * dx + dy * dy + dz * dz > radius_sq) continue;
        atomicMin(reinterpret_cast<unsigned int *>(owner + idx), static_cast<unsigned int>(p));
    }
}

// Initial geometry of one lattice.
__global__ __launch_bounds__(PB) void k_geo_mark_init(int32_t *owner, int32_t *box,
                                                      const double *px, const double *py,
                                                      const double *pz, double R, double radius_sq,
                                                      Dims d) {
    geo_mark_body(owner, box, px, py, pz, R, radius_sq, d, blockIdx.x);
}

__global__ __launch_bounds__(PB) void k_geo_mark(PartGrid g) {
    const int s = slot_of(g.pstart, g.nslots, blockIdx.x);
    const PartSlot &sl = g.tab[s];
    const PartDev &pd = sl.pd;
    geo_mark_body(sl.owner_curr, sl.box_curr, pd.p[RBTP_P_POS_X], pd.p[RBTP_P_POS_Y],
                  pd.p[RBTP_P_POS_Z], pd.radius, pd.radius_sq, sl.d, blockIdx.x - g.pstart[s]);
}

// ---- contacts and force history ----------------------------------------------------------------

// ---- pair search grid -----------------------------------------------------------------------

__device__ inline int cell_coord(double x, double scale, int n) {
    const double t = fmin(fmax(x * scale, 0.0), static_cast<double>(n - 1));  // NaN -> 0
    return static_cast<int>(t);
}

__global__ __launch_bounds__(256) void k_cell_clear(PartGrid g) {
    for (int u = blockIdx.x * 256 + threadIdx.x; u < g.total_c; u += gridDim.x * 256) {
        const int s = slot_of(g.cstart, g.nslots, u);
        g.tab[s].pd.cell_cnt[u - g.cstart[s]] = 0;
    }
}

// Bins particle u (flat index) by its centre.  The order inside a cell is arbitrary: the pair sum
// sorts its candidates by pid.
__device__ inline void cell_bin_body(const PartGrid &g, int u) {
    if (u >= g.total_p) return;
    const int s = slot_of(g.pstart, g.nslots, u);
    const PartDev &pd = g.tab[s].pd;
    if (!pd.use_cells) return;
    const int i = u - g.pstart[s];
    const int cx = cell_coord(pd.p[RBTP_P_POS_X][i], pd.cell_scale[0], pd.ncell[0]);
    const int cy = cell_coord(pd.p[RBTP_P_POS_Y][i], pd.cell_scale[1], pd.ncell[1]);
    const int cz = cell_coord(pd.p[RBTP_P_POS_Z][i], pd.cell_scale[2], pd.ncell[2]);
    const int c = (cz * pd.ncell[1] + cy) * pd.ncell[0] + cx;
    const int slot = atomicAdd(pd.cell_cnt + c, 1);
    if (slot < CELL_CAP) pd.cell_pid[(size_t)c * CELL_CAP + slot] = i;
}

// Lubrication/overlap force of partner j on particle i (repulsive_pair_kernel, pforces.cu:6-38);
// false where the reference skips the pair.
__device__ inline bool pair_term(const PartDev &pd, const double *X, const double *Y,
                                 const double *Z, double xi, double yi, double zi, int j,
                                 double &fx, double &fy, double &fz) {
    const double dx = xi - X[j], dy = yi - Y[j], dz = zi - Z[j];
    const double dist_sq = dx * dx + dy * dy + dz * dz;
    if (dist_sq <= 1.0e-12) return false;
    const double dist = sqrt(dist_sq);
    const double min_gap = pd.min_gap;
    const double hgap = dist - 2.0 * pd.radius;
    if (hgap >= min_gap) return false;
    const double overlap = (hgap < 0.0) ? -hgap : 0.0;
    const double factor = (min_gap - hgap) / min_gap;
    const double lubforce = pd.coeff * (factor * factor / pd.stf0 + overlap / (min_gap * pd.stf1));
    fx = lubforce * (dx / dist);
    fy = lubforce * (dy / dist);
    fz = lubforce * (dz / dist);
    return true;
}

// prepare_particle_forces_step (pforces.cu:80-116), the heat history of compute_particle_heat
// (ibb.cu:260-261), and apply_repulsive_forces (pforces.cu:118-148), one wave per particle.
// Pair terms are summed per particle in ascending partner order; the term on the higher pid of a
// pair is evaluated from its own side, which is bit-for-bit the negation the reference applies
// (pforces.cu:35-37).  With the cell grid, the partners are the particles binned in the 27 cells
// around i's cell, ranked by pid; every pair within range is among them, and the pairs left out
// are ones the reference skips, so the sum is the same.  When a cell overflowed or i has more
// than CAND_MAX candidates, i falls back to every partner.  The lanes evaluate the terms of 64
// consecutive partners at a time; every lane then adds them in partner order (shuffles), so the
// sum is the sequential one.
constexpr int CW = 64;  // lanes per particle in the contact sums (one wave)

// Orders a wave's LDS stores before its lanes' later LDS loads: the wave's rows of the candidate
// lists are its own, so no workgroup barrier is needed.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Particle u (flat index) on one wave (lane): cand and sorted are the wave's own LDS rows of
// CAND_MAX ids (candidates in cell order, then ranked by pid).
__device__ inline void contacts_body(const PartGrid &g, int u, int lane, int32_t *cand,
                                     int32_t *sorted) {
    const int s = slot_of(g.pstart, g.nslots, u);
    const PartDev &pd = g.tab[s].pd;
    const Dims d = g.tab[s].d;
    const int i = u - g.pstart[s];
    double *const *P = pd.p;
    // The owner map generation rebuilt at the end of this step (k_geo_mark) is cleared here in
    // particle i's old box: nothing reads that generation before k_geo_mark, and the stores need
    // no launch of their own.
    geo_clear_body(g.tab[s].owner_curr, g.tab[s].box_curr, d, i, lane, CW);
    if (lane == 0) {
        // the special-node list of this step is rebuilt from here on (the unmap in k_bin_unmap,
        // the last reader of the previous list, ran before this launch): its head is one entry
        // per link-box node, which the link-node pass fills or marks empty; refilled nodes follow
        if (i == 0) {
            pd.sp.count[0] = pd.np * pd.lnk_stride;
            pd.sp.count[1] = pd.np * pd.lnk_stride;
        }
        for (int a = 0; a < 3; a++) {
            P[RBTP_P_FORCE_PREV2_X + a][i] = P[RBTP_P_FORCE_PREV_X + a][i];
            P[RBTP_P_TORQUE_PREV2_X + a][i] = P[RBTP_P_TORQUE_PREV_X + a][i];
            P[RBTP_P_FORCE_PREV_X + a][i] = P[RBTP_P_FORCE_X + a][i];
            P[RBTP_P_TORQUE_PREV_X + a][i] = P[RBTP_P_TORQUE_X + a][i];
            P[RBTP_P_TORQUE_X + a][i] = 0.0;
        }
        P[RBTP_P_HEAT_PREV][i] = P[RBTP_P_HEAT][i];
        P[RBTP_P_HEAT][i] = 0.0;
    }
    const double *X = P[RBTP_P_POS_X], *Y = P[RBTP_P_POS_Y], *Z = P[RBTP_P_POS_Z];
    const double xi = X[i], yi = Y[i], zi = Z[i];
    double FX = 0.0, FY = 0.0, FZ = 0.0;
    const double radius = pd.radius, coeff = pd.coeff;
    if (pd.np > 1) {
        int count = pd.np;     // partners to walk
        bool ranked = false;   // partners taken from `sorted`
        if (pd.use_cells) {
            const int cx = cell_coord(xi, pd.cell_scale[0], pd.ncell[0]);
            const int cy = cell_coord(yi, pd.cell_scale[1], pd.ncell[1]);
            const int cz = cell_coord(zi, pd.cell_scale[2], pd.ncell[2]);
            // the (up to) 27 neighbour cells, in z, y, x order: their counts, then their ids
            // into `cand` one cell after another
            const int z0 = max(cz - 1, 0), z1 = min(cz + 1, pd.ncell[2] - 1);
            const int y0 = max(cy - 1, 0), y1 = min(cy + 1, pd.ncell[1] - 1);
            const int x0 = max(cx - 1, 0), x1 = min(cx + 1, pd.ncell[0] - 1);
            bool full = false;
            int total = 0;
            for (int z = z0; z <= z1; z++)
                for (int y = y0; y <= y1; y++)
                    for (int x = x0; x <= x1; x++) {
                        const int c = pd.cell_cnt[(z * pd.ncell[1] + y) * pd.ncell[0] + x];
                        full = full || c > CELL_CAP;
                        total += c;
                    }
            // every lane read the same counts: this branch is uniform over the wave
            if (!full && total <= CAND_MAX) {
                int base = 0;
                for (int z = z0; z <= z1; z++)
                    for (int y = y0; y <= y1; y++)
                        for (int x = x0; x <= x1; x++) {
                            const int cell = (z * pd.ncell[1] + y) * pd.ncell[0] + x;
                            const int c = pd.cell_cnt[cell];
                            for (int t = lane; t < c; t += CW)
                                cand[base + t] = pd.cell_pid[(size_t)cell * CELL_CAP + t];
                            base += c;
                        }
                wave_lds_sync();
                // rank by pid (every particle is binned once: the ids are distinct)
                for (int t = lane; t < total; t += CW) {
                    const int32_t v = cand[t];
                    int r = 0;
                    for (int m = 0; m < total; m++) r += cand[m] < v ? 1 : 0;
                    sorted[r] = v;
                }
                wave_lds_sync();
                count = total;
                ranked = true;
            }
        }
        for (int k0 = 0; k0 < count; k0 += CW) {
            const int k = k0 + lane;
            double fx = 0.0, fy = 0.0, fz = 0.0;
            bool ok = false;
            if (k < count) {
                const int j = ranked ? sorted[k] : k;
                if (j != i) ok = pair_term(pd, X, Y, Z, xi, yi, zi, j, fx, fy, fz);
            }
            const unsigned long long okm = __ballot(ok);
            for (int m = 0; m < CW; m++) {
                const double tx = __shfl(fx, m), ty = __shfl(fy, m), tz = __shfl(fz, m);
                if ((okm >> m) & 1ull) {
                    FX += tx;
                    FY += ty;
                    FZ += tz;
                }
            }
        }
    }
    if (lane != 0) return;
    // repulsive_wall_kernel (pforces.cu:40-76)
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const double mgw = pd.min_gap_w;
    auto wall = [&](double gap, double nx, double ny, double nz) {
        if (gap >= mgw) return;
        const double overlap = (gap < 0.0) ? -gap : 0.0;
        const double factor = (mgw - gap) / mgw;
        const double lubforce = coeff * (factor * factor / pd.stf0_w + overlap / (mgw * pd.stf1_w));
        fx += lubforce * nx;
        fy += lubforce * ny;
        fz += lubforce * nz;
    };
    wall(xi - radius, 1.0, 0.0, 0.0);
    wall((static_cast<double>(d.lx) - xi) - radius, -1.0, 0.0, 0.0);
    wall(yi - radius, 0.0, 1.0, 0.0);
    wall((static_cast<double>(d.ly) - yi) - radius, 0.0, -1.0, 0.0);
    wall(zi - radius, 0.0, 0.0, 1.0);
    wall((static_cast<double>(d.lz) - zi) - radius, 0.0, 0.0, -1.0);
    P[RBTP_P_FORCE_X][i] = FX + fx;
    P[RBTP_P_FORCE_Y][i] = FY + fy;
    P[RBTP_P_FORCE_Z][i] = FZ + fz;
}

// ---- special-node list -------------------------------------------------------------------------

// Appends the nodes of the calling wave's lanes that have `want` set to the list's tail (the
// refilled nodes), in lane order, with one atomic per wave.  Every active lane of the wave must
// call it.  Returns the lane's list position, or -1 (not appended, or the list is full).  The
// IBB link nodes do not come through here: a counter that every wave of the link-node pass
// incremented cost more than the rest of that kernel at 5000 spheres, so they have fixed
// positions in the list's head.
__device__ inline int push_special(const PartDev &pd, bool want, int idx, int kind) {
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return -1;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    int base = 0;
    if (lane == leader)
        base = __hip_atomic_fetch_add(gbl(pd.sp.count), __popcll(mask), __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
    base = __shfl(base, leader);
    if (!want) return -1;
    const int j = base + __popcll(mask & ((1ull << lane) - 1ull));
    if (j < pd.sp.cap) {
        gbl(pd.sp.node)[j] = idx;
        gbl(pd.sp_kind)[j] = kind;
        return j;
    }
    atomicOr(pd.err + ERR_SPECIAL_CAP, 1);
    return -1;
}

// ---- link evaluation -----------------------------------------------------------------------------

// Particle state used by the IBB of a link.
struct PView {
    gptr<const double> px, py, pz, vx, vy, vz, wx, wy, wz, temp;
};
// The view of the state arrays `a` (pd.p or pd.old) with the given indices.
template <class Arr>
__device__ inline PView pview(const Arr &a, int ipx, int ipy, int ipz, int ivx, int ivy, int ivz,
                              int iwx, int iwy, int iwz, int it) {
    return PView{gbl((const double *)a[ipx]), gbl((const double *)a[ipy]),
                 gbl((const double *)a[ipz]), gbl((const double *)a[ivx]),
                 gbl((const double *)a[ivy]), gbl((const double *)a[ivz]),
                 gbl((const double *)a[iwx]), gbl((const double *)a[iwy]),
                 gbl((const double *)a[iwz]), gbl((const double *)a[it])};
}
__device__ inline PView pview_now(const PartDev &pd) {
    return pview(pd.p, RBTP_P_POS_X, RBTP_P_POS_Y, RBTP_P_POS_Z, RBTP_P_VEL_X, RBTP_P_VEL_Y,
                 RBTP_P_VEL_Z, RBTP_P_OMEGA_X, RBTP_P_OMEGA_Y, RBTP_P_OMEGA_Z, RBTP_P_TEMP);
}
__device__ inline PView pview_old(const PartDev &pd) {
    return pview(pd.old, OLD_PX, OLD_PY, OLD_PZ, OLD_VX, OLD_VY, OLD_VZ, OLD_WX, OLD_WY, OLD_WZ,
                 OLD_T);
}

// The state of one particle that the bounce-back of its links reads, loaded once per node
// rather than once per link (the stores of the link values could alias these arrays, so the
// compiler would otherwise reload them after every link).
struct PState {
    double px, py, pz, vx, vy, vz, wx, wy, wz, temp;
};
__device__ inline PState pstate(const PView &pv, int pid) {
    return PState{pv.px[pid], pv.py[pid], pv.pz[pid], pv.vx[pid], pv.vy[pid],
                  pv.vz[pid], pv.wx[pid], pv.wy[pid], pv.wz[pid], pv.temp[pid]};
}
// A double that every active lane holds the same value of, moved to scalar registers.
__device__ inline double wave_uniform(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)b);
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
// pstate of a particle that is the same for the whole wave.
__device__ inline PState pstate_uniform(const PView &pv, int pid) {
    const PState s = pstate(pv, pid);
    return PState{wave_uniform(s.px), wave_uniform(s.py), wave_uniform(s.pz),
                  wave_uniform(s.vx), wave_uniform(s.vy), wave_uniform(s.vz),
                  wave_uniform(s.wx), wave_uniform(s.wy), wave_uniform(s.wz),
                  wave_uniform(s.temp)};
}

template <class S>
struct PopView {
    gptr<const S> f, g;
    size_t n;
    gptr<const float> gshift;  // shift profile of the reduced-precision g encoding, [lx]
    int lx;
    __device__ inline double F(int i, int idx) const {
        return dec_f<S, double>(Enc<S>::ld(f + pop_at(i, idx)), i);
    }
    __device__ inline double G(int i, int idx) const {
        return dec_g<S, double>(Enc<S>::ld(g + pop_at(i, idx)), i,
                                Enc<S>::scaled ? double(gshift[idx % lx]) : 0.0);
    }
};

struct LinkOut {
    double f_in, f_out, g_in, g_out;
    double uwx, uwy, uwz;
};

// The lattice values the bounce-back of link (x, ip) reads.  Links start only at fluid nodes,
// so the last branch of bounce_back_particles_kernel and the no-write branch of
// bounce_back_thermal_kernel (ibb.cu:77-80,128-131) are never taken, and each distribution
// needs two values: its own post-collision value at x, and the partner of the branch the link
// takes, which is known before any population is read (q from the geometry, the upstream
// node's owner from the owner map):
//   upstream x - c_ip inside and fluid, q <  1/2 (g: q <= 1/2)  ->  f*_ip(x - c_ip)
//   upstream x - c_ip inside and fluid, q >= 1/2 (g: q >  1/2)  ->  f*_opp(x)
//   otherwise                                                   ->  none (x's own value again)
// All four loads are unconditional, with the address chosen, so that the loads of several links
// can be in flight together; a repeated read of x's own value touches no further cache line.
struct LinkLoads {
    double f_here, f_pair, g_here, g_pair;
};

// A population value as stored (the 2-byte forms as their bits), loaded without decoding.
template <class S>
__device__ inline auto raw_pop(gptr<const S> p) {
    if constexpr (sizeof(S) == 2) return uint32_t(*(gptr<const unsigned short>)p);
    else return *p;
}
template <class S, class R>
__device__ inline S raw_as(R r) {
    if constexpr (std::is_same<S, __half>::value) return __ushort_as_half((unsigned short)r);
    else if constexpr (std::is_same<S, __hip_bfloat16>::value)
        return __ushort_as_bfloat16((unsigned short)r);
    else return r;
}

// The two loads of x's own values, which do not depend on q (issued before link_q's square
// root), and the pair loads of the branch q selects.
template <class S> struct LinkRaw {
    decltype(raw_pop<S>(nullptr)) fh, fp, gh, gp;
    float sh_here, sh_pair;  // shift profile at x and at the g partner (reduced-precision g)
};
template <class S>
__device__ inline void link_raw_here(LinkRaw<S> &r, int ip, int idx, const PopView<S> &pop) {
    r.fh = raw_pop<S>(pop.f + pop_at(ip, idx));
    r.gh = raw_pop<S>(pop.g + pop_at(ip, idx));
    r.sh_here = 0.0f;
    if constexpr (Enc<S>::scaled) r.sh_here = pop.gshift[idx % pop.lx];
}
template <class S>
__device__ inline LinkLoads link_raw_pair(LinkRaw<S> &r, int ip, int idx, int idx_up,
                                          bool up_fluid, double q, const PopView<S> &pop) {
    const int opp = dir_opp(ip);
    const int fi = up_fluid ? (q < 0.5 ? ip : opp) : ip;
    const int fn = up_fluid && q < 0.5 ? idx_up : idx;
    const int gi = up_fluid ? (q <= 0.5 ? ip : opp) : ip;
    const int gn = up_fluid && q <= 0.5 ? idx_up : idx;
    r.fp = raw_pop<S>(pop.f + pop_at(fi, fn));
    r.gp = raw_pop<S>(pop.g + pop_at(gi, gn));
    r.sh_pair = 0.0f;
    if constexpr (Enc<S>::scaled) r.sh_pair = pop.gshift[gn % pop.lx];
    // All four values are in flight here and the link waits for them once: without this point
    // the compiler sinks each pair load into the branch of eval_link_at that reads it, and the
    // link then waits for memory two or three times in turn.
    asm volatile("" ::"v"(r.fh), "v"(r.fp), "v"(r.gh), "v"(r.gp));
    LinkLoads L;
    L.f_here = dec_f<S, double>(raw_as<S>(r.fh), ip);
    L.f_pair = dec_f<S, double>(raw_as<S>(r.fp), fi);
    L.g_here = dec_g<S, double>(raw_as<S>(r.gh), ip, double(r.sh_here));
    L.g_pair = dec_g<S, double>(raw_as<S>(r.gp), gi, double(r.sh_pair));
    return L;
}

template <class S>
__device__ inline LinkLoads link_loads(int ip, int idx, int idx_up, bool up_fluid, double q,
                                       const PopView<S> &pop) {
    LinkRaw<S> r;
    link_raw_here<S>(r, ip, idx, pop);
    return link_raw_pair<S>(r, ip, idx, idx_up, up_fluid, q, pop);
}

// bounce_back_particles_kernel (ibb.cu:19-81) and bounce_back_thermal_kernel (ibb.cu:94-135)
// for the link of fluid node (ix,iy,iz) in direction ip, from the post-collision populations in
// L; returns the values they write into f[opp][idx] and g[opp][idx] plus what the force and
// heat sums read (ibb.cu:160-180,202-203).
__device__ inline LinkOut eval_link_at(const LinkLoads &L, int ip, int ix, int iy, int iz,
                                       bool up_fluid, double q, const PState &ps,
                                       double density, double rho0) {
    LinkOut o;
    const double tp = wt(ip);
    // wall velocity at the boundary point (ibb.cu:33-45)
    const double cx = static_cast<double>(ix) + 0.5 + q * dir_cx(ip);
    const double cy = static_cast<double>(iy) + 0.5 + q * dir_cy(ip);
    const double cz = static_cast<double>(iz) + 0.5 + q * dir_cz(ip);
    const double relx = cx - ps.px, rely = cy - ps.py, relz = cz - ps.pz;
    const double wx = ps.wx, wy = ps.wy, wz = ps.wz;
    const double uwx = ps.vx + (relz * wy - rely * wz);
    const double uwy = ps.vy + (relx * wz - relz * wx);
    const double uwz = ps.vz + (rely * wx - relx * wy);
    o.uwx = uwx; o.uwy = uwy; o.uwz = uwz;
    const double evel = dir_cx(ip) * uwx + dir_cy(ip) * uwy + dir_cz(ip) * uwz;
    const double RT = 1.0 / 3.0;
    // / RT as div_rn with RN(1/RT) = 3 and RN(1/RT - 3) = 3 * 2^-54
    const double momentum_term = div_rn(2.0 * tp * rho0 * evel, RT, RECIP_RT_HI, RECIP_RT_LO);
    const double f_here = L.f_here;
    o.f_in = f_here;
    if (up_fluid) {
        if (q < 0.5) {
            o.f_out = 2.0 * q * f_here + (1.0 - 2.0 * q) * L.f_pair - momentum_term;
        } else {
            const double inv_2q = 0.5 / q;
            o.f_out = (f_here - momentum_term) * inv_2q + (2.0 * q - 1.0) * inv_2q * L.f_pair;
        }
    } else {
        o.f_out = f_here - momentum_term;
    }
    // thermal (ibb.cu:109-135)
    const double Tp = ps.temp;
    const double amp = 2.0 * tp * Tp * (1.0 + density);
    const double g_here = L.g_here;
    o.g_in = g_here;
    if (up_fluid) {
        if (q <= 0.5) {
            o.g_out = -(2.0 * q * g_here + (1.0 - 2.0 * q) * L.g_pair) + amp;
        } else {
            const double inv_2q = 0.5 / q;
            o.g_out = -g_here * inv_2q + amp * inv_2q + (2.0 * q - 1.0) * inv_2q * L.g_pair;
        }
    } else {
        o.g_out = -g_here + amp;
    }
    return o;
}

// Hydrodynamic force, torque and heat of every particle (compute_particle_forces_kernel,
// ibb.cu:138-188; compute_particle_heat_kernel, ibb.cu:190-205), in two launches over every
// particle of every slot.  The link-node pass gives one thread to each node of a particle's link
// box: it finds the node's links to the particle, and stores the node's share of the seven sums
// (its links' terms in direction order, summed from 0) and a bit telling that the node has
// links.  When the particle is the node's lowest linked pid, it also lists the node as a special
// node and writes the bounce-back values of its links into the node's entry (ibb.cu:5-136); a
// node that also links to a sphere of higher pid is marked SP_MULTI and completed by the
// shared-link fix-up (k_fixups).
// k_link_reduce then gives one wave to each particle: lane t adds the shares of the box nodes
// k = t, t+64, ... in increasing k, and the 64 lane sums are combined pairwise (stride 32 ...
// 1) and added to the sums left by the contact sums.
constexpr int RL = 64;  // lanes of the per-particle reduction (one wave)
constexpr uint32_t SP_MULTI = 1u << 31;  // special-node mask flag: links to several spheres

// Link-box node threads u = blk * 256 + threadIdx.x of one workgroup of k_contacts_links.
template <class S>
__device__ inline void link_nodes_body(const PartGrid &g, int blk) {
    const int u = blk * 256 + threadIdx.x;
    if (u >= g.total_b) return;  // total_b is a whole number of waves
    // every wave lies in one particle of one slot: slot and particle are wave-uniform
    const int u0 = __builtin_amdgcn_readfirstlane(u);
    const int slot = slot_of(g.bstart, g.nslots, u0);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const Dims d = sl.d;
    const int p = (u0 - g.bstart[slot]) / pd.lnk_stride;
    const int k = u - g.bstart[slot] - p * pd.lnk_stride;
    const gptr<const int32_t> owner_prev = gbl(sl.owner_prev);
    const PView pv = pview_now(pd);
    int lo[3], hi[3];
    bool in_box = load_box(sl.box_prev, p, lo, hi);
    bool interior = false;  // the box and its neighbours lie inside the domain
    if (in_box) {
        grow_box(d, lo, hi);
        in_box = k < box_count(lo, hi);
        interior = lo[0] > 0 && lo[1] > 0 && lo[2] > 0 && hi[0] < d.lx - 1 &&
                   hi[1] < d.ly - 1 && hi[2] < d.lz - 1;
    }
    const double X = wave_uniform(pv.px[p]), Y = wave_uniform(pv.py[p]), Z = wave_uniform(pv.pz[p]);
    int ix = 0, iy = 0, iz = 0, idx = 0;
    // links of this node to p, as a direction mask (q is evaluated again below), whether it also
    // links to a sphere of lower pid (a node linked to several spheres is listed once, by the
    // lowest linked pid) or of higher pid, and which neighbours are fluid
    unsigned mask = 0u, fluid_nb = 0u;
    bool lower = false, higher = false;
    double x0 = 0.0, y0 = 0.0, z0 = 0.0;
    if (in_box) {
        idx = box_node(d, lo, hi, k, ix, iy, iz);
        x0 = static_cast<double>(ix) + 0.5;
        y0 = static_cast<double>(iy) + 0.5;
        z0 = static_cast<double>(iz) + 0.5;
        // a link x0 -> x0 + c_ip (|c_ip| <= sqrt 3) reaches into the sphere only from within
        // R + sqrt 3 of its centre; the margin to R + 1.75 is far above any rounding
        const double dx = x0 - X, dy = y0 - Y, dz = z0 - Z;
        const double reach = pd.radius + 1.75;
        if (dx * dx + dy * dy + dz * dz <= reach * reach) {
            // the node's owner and its 26 neighbours' in one batch of loads
            const int self = owner_prev[idx];
            int own[NPOP];
            neighbour_owners(owner_prev, d, ix, iy, iz, interior, own);
            if (self == -1) {
                unsigned other = 0u;
#pragma unroll
                for (int ip = 1; ip < NPOP; ip++) {
                    if (own[ip] == -1) fluid_nb |= 1u << ip;
                    if (own[ip] >= 0 && own[ip] != p) other |= 1u << ip;
                    // direction constants fold into the unrolled code; no loads
                    if (own[ip] == p && link_exists(ip, x0, y0, z0, X, Y, Z, pd.radius_sq))
                        mask |= 1u << ip;
                }
                // neighbours owned by other spheres (only near another sphere)
                while (other) {
                    const int ip = __builtin_ctz(other);
                    other &= other - 1u;
                    const int q_owner = own[ip];
                    double q;
                    if (link_q(ip, x0, y0, z0, pv.px[q_owner], pv.py[q_owner], pv.pz[q_owner],
                               pd.radius_sq, q)) {
                        if (q_owner < p) lower = true;
                        else higher = true;
                    }
                }
            }
        }
    }
    // The node's entry in the special list is its box-node slot p * lnk_stride + k (the list
    // head is one entry per link-box node, no counter): a listed node writes it below, every
    // other slot is marked empty.
    const int e = p * pd.lnk_stride + k;
    const int j = (mask != 0u && !lower) ? e : -1;
    if (j >= 0) {
        gbl(pd.sp.node)[j] = idx;
        gbl(pd.sp_kind)[j] = -1;
        gbl(pd.sp.map)[idx] = j;
    } else {
        gbl(pd.sp.mask)[e] = 0u;
    }
    // every lane of the wave is here
    const unsigned long long has = __ballot(mask != 0u);
    if (__lane_id() == 0) gbl(pd.lnk_bits)[(size_t)p * (pd.lnk_stride / 64) + k / 64] = has;
    if (has == 0ull) return;  // the whole wave
    // the populations a listed node replaces are its links' opposites
    if (j >= 0) {
        uint32_t replaced = 0u;
        for (uint32_t m = mask; m; m &= m - 1u) replaced |= 1u << dir_opp(__builtin_ctz(m));
        gbl(pd.sp.mask)[j] = replaced | (higher ? SP_MULTI : 0u);
    }
    // The wave's links are numbered in lane order, and in ascending direction order within a
    // lane: lane L holds links off .. off + cnt - 1.  They are evaluated one per lane, 64 at a
    // time (a wave holds about 90), instead of each node walking its own links while the lanes
    // of nodes with fewer links wait.
    const int lane = __lane_id();
    const int cnt = __popc(mask);
    int incl = cnt;
#pragma unroll
    for (int s = 1; s < 64; s <<= 1) {
        const int t = __shfl_up(incl, s);
        if (lane >= s) incl += t;
    }
    const int off = incl - cnt;
    const int total = __shfl(incl, 63);
    const PopView<S> pop{gbl((const S *)sl.fcur), gbl((const S *)sl.gcur), (size_t)d.n,
                         gbl(pd.gshift), d.lx};
    using A = typename Enc<S>::A;
    const gptr<A> out = gbl((A *)pd.sp.pop);
    const double density =
        mask != 0u ? double(gbl((const typename Enc<S>::A *)sl.rho_full)[idx]) : 0.0;
    const PState ps = pstate_uniform(pv, p);  // p is the same for the whole wave
    const double radius_sq = wave_uniform(pd.radius_sq), rho0 = wave_uniform(pd.rho0);
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < total; base += 64) {
        const int t = base + lane;
        const bool act = t < total;
        // the node of link t: the last lane whose first link is at or before t (that lane has
        // links, since the next lane's first link is past t)
        int n = 0;
#pragma unroll
        for (int st = 32; st >= 1; st >>= 1)
            if (__shfl(off, n + st) <= t) n += st;
        uint32_t m = __shfl(mask, n);
        const int kth = t - __shfl(off, n);
        for (int r = 0; r < kth; r++) m &= m - 1u;  // drop the node's earlier links
        const int nidx = __shfl(idx, n), nj = __shfl(j, n);
        const uint32_t nfluid = __shfl(fluid_nb, n);
        const int nix = __shfl(ix, n), niy = __shfl(iy, n), niz = __shfl(iz, n);
        const double ndens = __shfl(density, n);
        double c[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        if (act) {
            const int ip = __builtin_ctz(m);
            const int opp = dir_opp(ip);
            const int cx = dir_cx(ip), cy = dir_cy(ip), cz = dir_cz(ip);
            const bool up_fluid = (nfluid >> opp) & 1u;  // x - c_ip = x + c_opp
            const int idx_up = nidx - (cx + cy * d.lx + cz * d.lxy);
            const double lx0 = static_cast<double>(nix) + 0.5;
            const double ly0 = static_cast<double>(niy) + 0.5;
            const double lz0 = static_cast<double>(niz) + 0.5;
            LinkRaw<S> raw;
            link_raw_here<S>(raw, ip, nidx, pop);
            double q = 0.0;
            link_q(ip, lx0, ly0, lz0, X, Y, Z, radius_sq, q);
            const LinkLoads Ld = link_raw_pair<S>(raw, ip, nidx, idx_up, up_fluid, q, pop);
            const LinkOut L = eval_link_at(Ld, ip, nix, niy, niz, up_fluid, q, ps, ndens, rho0);
            if (nj >= 0) {
                out[sp_at(nj, opp)] = A(L.f_out);
                out[sp_at(nj, NPOP + opp)] = A(L.g_out);
            }
            // Force uses the link record's r = q c (links.cu:134-136, ibb.cu:160-168).
            const double rx = q * static_cast<double>(cx);
            const double ry = q * static_cast<double>(cy);
            const double rz = q * static_cast<double>(cz);
            const double relx = nix + 0.5 + rx - X;
            const double rely = niy + 0.5 + ry - Y;
            const double relz = niz + 0.5 + rz - Z;
            const double Fx = (cx - L.uwx) * L.f_in - (-cx - L.uwx) * L.f_out;
            const double Fy = (cy - L.uwy) * L.f_in - (-cy - L.uwy) * L.f_out;
            const double Fz = (cz - L.uwz) * L.f_in - (-cz - L.uwz) * L.f_out;
            c[0] = Fx;
            c[1] = Fy;
            c[2] = Fz;
            c[3] = rely * Fz - relz * Fy;
            c[4] = relz * Fx - relx * Fz;
            c[5] = relx * Fy - rely * Fx;
            c[6] = L.g_in - L.g_out;
        }
        // each node adds its links' shares of this batch in ascending direction order, the
        // order of its own walk over them
        const int first = off > base ? off : base;
        const int last = (off + cnt < base + 64) ? off + cnt : base + 64;
        const int here = last - first;
        int rounds = here;
#pragma unroll
        for (int s = 32; s >= 1; s >>= 1) {
            const int o = __shfl_xor(rounds, s);
            rounds = o > rounds ? o : rounds;
        }
        for (int r = 0; r < rounds; r++) {
            const int src = (first - base + r) & 63;
            double sh[7];
#pragma unroll
            for (int cc = 0; cc < 7; cc++) sh[cc] = __shfl(c[cc], src);
            if (r < here) {
#pragma unroll
                for (int cc = 0; cc < 7; cc++) v[cc] += sh[cc];
            }
        }
    }
    if (mask == 0u) return;
    const size_t plane = (size_t)pd.np * pd.lnk_stride;
    const gptr<double> part = gbl(pd.lnk_part);
    for (int cc = 0; cc < 7; cc++) part[cc * plane + e] = v[cc];
}

// Contact sums and link nodes in one launch: neither reads what the other writes (the contacts
// write the force, torque and heat history, the force of the contacts, the owner generation
// being rebuilt and the special list's counters; the link nodes read positions, velocities,
// temperatures, the other owner generation and the populations, and write the node shares and
// the list's head), and the wave-per-particle contact sums, latency-bound like the link nodes,
// then run beside them instead of as a launch of their own.  The first nc workgroups (a multiple
// of 8) hold four particles' contact sums each, one per wave; the rest are the link nodes',
// remapped so that a particle's consecutive workgroups share an XCD.  Four waves per SIMD (at
// most 128 VGPRs) measured 1.7 % faster on the whole step with 5000 spheres in FP16 storage than
// the compiler's own choice of three for the link nodes; five waves spill far more.
template <class S>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_contacts_links(
    PartGrid g, int nc) {
    __shared__ int32_t cand[4][CAND_MAX];
    __shared__ int32_t sorted[4][CAND_MAX];
    if ((int)blockIdx.x < nc) {
        const int w = threadIdx.x >> 6;
        const int u = (int)blockIdx.x * 4 + w;
        if (u < g.total_p) contacts_body(g, u, threadIdx.x & 63, cand[w], sorted[w]);
        return;
    }
    link_nodes_body<S>(g, xcd_remap((int)blockIdx.x - nc, (int)gridDim.x - nc));
}

__device__ inline void move_particle(const PartDev &pd, int pid);

// Planes C0 .. C0+NC-1 of the node shares (0-2 force, 3-5 torque, 6 heat) onto the particle's
// sums; the particle phase reduces all seven and, with MOVE, steps the particle right after
// (update_particles reads only its own sums); the per-stage execution reduces the forces or
// the heat and moves the particles in a launch of its own.
template <int C0, int NC, bool MOVE = false>
__global__ __launch_bounds__(256) void k_link_reduce(PartGrid g) {
    const int w = blockIdx.x * (256 / RL) + (threadIdx.x / RL);
    if (w >= g.total_p) return;  // whole waves
    const int lane = threadIdx.x % RL;
    const int w0 = __builtin_amdgcn_readfirstlane(w);
    const int slot = slot_of(g.pstart, g.nslots, w0);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const int p = w0 - g.pstart[slot];
    double s[NC];
    for (int c = 0; c < NC; c++) s[c] = 0.0;
    int lo[3], hi[3];
    if (load_box(sl.box_prev, p, lo, hi)) {
        grow_box(sl.d, lo, hi);
        const int nw = (box_count(lo, hi) + RL - 1) / RL;  // words of node bits
        const unsigned long long *bits = pd.lnk_bits + (size_t)p * (pd.lnk_stride / 64);
        const size_t base = (size_t)p * pd.lnk_stride + lane;
        const size_t plane = (size_t)pd.np * pd.lnk_stride;
        constexpr int U = 4;  // shares of up to U nodes loaded together, then added in order
        // up to 64 bit words at once, one per lane, handed to every lane by shuffles
        for (int wb = 0; wb < nw; wb += RL) {
            const unsigned long long mine = (wb + lane < nw) ? bits[wb + lane] : 0ull;
            const int nwb = min(RL, nw - wb);
            for (int w1 = 0; w1 < nwb; w1 += U) {
                bool has[U];
                double sh[U][NC];
#pragma unroll
                for (int i = 0; i < U; i++) {
                    const unsigned long long b = __shfl(mine, (w1 + i) & (RL - 1));
                    has[i] = w1 + i < nwb && ((b >> lane) & 1ull);
                    for (int c = 0; c < NC; c++)
                        sh[i][c] = has[i] ? pd.lnk_part[(C0 + c) * plane + base +
                                                        (size_t)(wb + w1 + i) * RL]
                                          : 0.0;
                }
#pragma unroll
                for (int i = 0; i < U; i++)
                    if (has[i])
                        for (int c = 0; c < NC; c++) s[c] += sh[i][c];
            }
        }
    }
    for (int stride = RL / 2; stride >= 1; stride >>= 1)
        for (int c = 0; c < NC; c++) s[c] = s[c] + __shfl_down(s[c], stride, RL);
    if (lane == 0) {
        const int dst[7] = {RBTP_P_FORCE_X, RBTP_P_FORCE_Y, RBTP_P_FORCE_Z, RBTP_P_TORQUE_X,
                            RBTP_P_TORQUE_Y, RBTP_P_TORQUE_Z, RBTP_P_HEAT};
        for (int c = 0; c < NC; c++) pd.p[dst[C0 + c]][p] += s[c];
        // the particle step needs only this particle's sums: the same lane takes it
        if (MOVE) move_particle(pd, p);
    }
}

// ---- motion -------------------------------------------------------------------------------------

// Entries of the previous step's special-node list leave the node map (the lattice kernel that
// read them ran before this kernel); slot s uses workgroups lstart[s] .. lstart[s+1]-1.  Empty
// entries of the list head have a zero mask.  The list itself is rebuilt from k_contacts_links on.
// Workgroup b of 128 threads (thread tid) of that walk.
__device__ inline void special_unmap_body(const PartGrid &g, int b, int tid) {
    const int slot = slot_of(g.lstart, g.nslots, b);
    const PartDev &pd = g.tab[slot].pd;
    const int blk = b - g.lstart[slot];
    const int nblk = g.lstart[slot + 1] - g.lstart[slot];
    const int count = min(pd.sp.count[0], pd.sp.cap);
    for (int j = blk * 128 + tid; j < count; j += nblk * 128)
        if (pd.sp.mask[j] != 0u) pd.sp.map[pd.sp.node[j]] = -1;
}

// The cell binning (first nbin workgroups) and the unmap (two of its 128-thread workgroups per
// workgroup of 256) in one launch: they touch different arrays.
__global__ __launch_bounds__(256) void k_bin_unmap(PartGrid g, int nbin) {
    if ((int)blockIdx.x < nbin) {
        cell_bin_body(g, (int)blockIdx.x * 256 + threadIdx.x);
        return;
    }
    const int b = ((int)blockIdx.x - nbin) * 2 + (threadIdx.x >> 7);
    if (b < g.total_l) special_unmap_body(g, b, threadIdx.x & 127);
}

// update_particles_kernel (move.cu:5-103) for particle pid of lattice pd, after saving the state
// the IBB of this step used.
__device__ inline void move_particle(const PartDev &pd, int pid) {
    double *const *P = pd.p;
    const int src[OLD_COUNT] = {RBTP_P_POS_X, RBTP_P_POS_Y, RBTP_P_POS_Z, RBTP_P_VEL_X,
                                RBTP_P_VEL_Y, RBTP_P_VEL_Z, RBTP_P_OMEGA_X, RBTP_P_OMEGA_Y,
                                RBTP_P_OMEGA_Z, RBTP_P_TEMP};
    for (int k = 0; k < OLD_COUNT; k++) pd.old[k][pid] = P[src[k]][pid];
    const double dt = 1.0;
    const double mass = pd.mass, inertia = pd.inertia, heat_capacity = pd.heat_capacity;
    const double forceX = P[RBTP_P_FORCE_X][pid], forceY = P[RBTP_P_FORCE_Y][pid], forceZ = P[RBTP_P_FORCE_Z][pid];
    const double torqueX = P[RBTP_P_TORQUE_X][pid], torqueY = P[RBTP_P_TORQUE_Y][pid], torqueZ = P[RBTP_P_TORQUE_Z][pid];
    const double fx_prev = P[RBTP_P_FORCE_PREV_X][pid], fy_prev = P[RBTP_P_FORCE_PREV_Y][pid], fz_prev = P[RBTP_P_FORCE_PREV_Z][pid];
    const double fx_prev2 = P[RBTP_P_FORCE_PREV2_X][pid], fy_prev2 = P[RBTP_P_FORCE_PREV2_Y][pid], fz_prev2 = P[RBTP_P_FORCE_PREV2_Z][pid];
    const double tpx = P[RBTP_P_TORQUE_PREV_X][pid], tpy = P[RBTP_P_TORQUE_PREV_Y][pid], tpz = P[RBTP_P_TORQUE_PREV_Z][pid];
    const double tp2x = P[RBTP_P_TORQUE_PREV2_X][pid], tp2y = P[RBTP_P_TORQUE_PREV2_Y][pid], tp2z = P[RBTP_P_TORQUE_PREV2_Z][pid];
    const double buoyancy = (1.0 - 1.0 / pd.rho_ratio) * pd.grav0;
    const double vx_old = P[RBTP_P_VEL_X][pid], vy_old = P[RBTP_P_VEL_Y][pid], vz_old = P[RBTP_P_VEL_Z][pid];
    const double omega_x_old = P[RBTP_P_OMEGA_X][pid], omega_y_old = P[RBTP_P_OMEGA_Y][pid], omega_z_old = P[RBTP_P_OMEGA_Z][pid];
    const double combo_fx = 0.5 * forceX + fx_prev + 0.5 * fx_prev2;
    const double combo_fy = 0.5 * forceY + fy_prev + 0.5 * fy_prev2;
    const double combo_fz = 0.5 * forceZ + fz_prev + 0.5 * fz_prev2;
    const double combo_torque_x = 0.5 * torqueX + tpx + 0.5 * tp2x;
    const double combo_torque_y = 0.5 * torqueY + tpy + 0.5 * tp2y;
    const double combo_torque_z = 0.5 * torqueZ + tpz + 0.5 * tp2z;
    const double vx_new = vx_old + 0.5 * dt * combo_fx / mass + dt * buoyancy;
    const double vy_new = vy_old + 0.5 * dt * combo_fy / mass;
    const double vz_new = vz_old + 0.5 * dt * combo_fz / mass;
    const double omega_x_new = omega_x_old + 0.5 * dt * combo_torque_x / inertia;
    const double omega_y_new = omega_y_old + 0.5 * dt * combo_torque_y / inertia;
    const double omega_z_new = omega_z_old + 0.5 * dt * combo_torque_z / inertia;
    P[RBTP_P_POS_X][pid] = P[RBTP_P_POS_X][pid] + 0.5 * dt * (vx_new + vx_old);
    P[RBTP_P_POS_Y][pid] = P[RBTP_P_POS_Y][pid] + 0.5 * dt * (vy_new + vy_old);
    P[RBTP_P_POS_Z][pid] = P[RBTP_P_POS_Z][pid] + 0.5 * dt * (vz_new + vz_old);
    P[RBTP_P_VEL_X][pid] = vx_new;
    P[RBTP_P_VEL_Y][pid] = vy_new;
    P[RBTP_P_VEL_Z][pid] = vz_new;
    P[RBTP_P_OMEGA_X][pid] = omega_x_new;
    P[RBTP_P_OMEGA_Y][pid] = omega_y_new;
    P[RBTP_P_OMEGA_Z][pid] = omega_z_new;
    double theta_x = P[RBTP_P_THETA_X][pid], theta_y = P[RBTP_P_THETA_Y][pid], theta_z = P[RBTP_P_THETA_Z][pid];
    theta_x += 0.5 * dt * (omega_x_new + omega_x_old);
    theta_y += 0.5 * dt * (omega_y_new + omega_y_old);
    theta_z += 0.5 * dt * (omega_z_new + omega_z_old);
    P[RBTP_P_THETA_X][pid] = theta_x;
    P[RBTP_P_THETA_Y][pid] = theta_y;
    P[RBTP_P_THETA_Z][pid] = theta_z;
    const double heat = P[RBTP_P_HEAT][pid], heat_prev = P[RBTP_P_HEAT_PREV][pid];
    double T = P[RBTP_P_TEMP][pid];
    T += 0.5 * dt * (heat + heat_prev) / heat_capacity;
    P[RBTP_P_TEMP][pid] = T;
}

__global__ __launch_bounds__(128) void k_move(PartGrid g) {
    const int u = blockIdx.x * 128 + threadIdx.x;
    if (u >= g.total_p) return;
    const int s = slot_of(g.pstart, g.nslots, u);
    move_particle(g.tab[s].pd, u - g.pstart[s]);
}

// ---- special nodes ------------------------------------------------------------------------------

// New-fluid nodes of one particle's marking box (new_fluid_count_kernel, refill.cu:8-41): solid
// in the previous geometry, fluid now, previous owner valid, and within close_thresh of that
// owner's new surface.  (IBB link nodes were listed by the link-node pass.)
__global__ __launch_bounds__(LB) void k_refill_find(PartGrid g) {
    const int slot = slot_of(g.pstart, g.nslots, blockIdx.x);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const Dims d = sl.d;
    const int32_t *owner_prev = sl.owner_prev;
    const int32_t *owner_curr = sl.owner_curr;
    const int p = blockIdx.x - g.pstart[slot];
    int lo[3], hi[3];
    if (!load_box(sl.box_prev, p, lo, hi)) return;
    const double PXv = pd.p[RBTP_P_POS_X][p], PYv = pd.p[RBTP_P_POS_Y][p], PZv = pd.p[RBTP_P_POS_Z][p];
    const int cnt = box_count(lo, hi);
    for (int k = threadIdx.x; k < cnt; k += LB) {
        int ix, iy, iz;
        const int idx = box_node(d, lo, hi, k, ix, iy, iz);
        bool want = false;
        if (owner_prev[idx] == p && owner_curr[idx] == -1) {
            const double x0 = static_cast<double>(ix) + 0.5;
            const double y0 = static_cast<double>(iy) + 0.5;
            const double z0 = static_cast<double>(iz) + 0.5;
            const double dx = x0 - PXv, dy = y0 - PYv, dz = z0 - PZv;
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            const double close_thresh = 1.7606816861659007;
            want = dist - pd.radius <= close_thresh;
        }
        push_special(pd, want, idx, p);
    }
}

// Bounce-back values of the IBB link nodes that link to more than one sphere (SP_MULTI entries in
// the head of the special list; the link-node pass wrote every other link node): for every link
// x -> x + c_ip (x + c_ip solid and linked, with the state before this step's motion) the values
// bounce_back_particles / bounce_back_thermal_particles write into f[opp][x] and g[opp][x]
// (ibb.cu:5-136), and the mask of the replaced populations.  The node's other populations are
// plain pulls, which the next lattice kernel makes itself.  Each wave reads the masks of 64
// consecutive entries at once and completes their SP_MULTI entries one after another, one lane
// per direction; the links of a node are independent (no sum), so their order does not matter.
template <class S>
__device__ inline void fixup_links_body(const PartGrid &g, int b) {
    const int slot = slot_of(g.lstart, g.nslots, b);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const Dims d = sl.d;
    using A = typename Enc<S>::A;
    const PopView<S> pop{gbl((const S *)sl.fcur), gbl((const S *)sl.gcur), (size_t)d.n,
                         gbl(pd.gshift), d.lx};
    const gptr<const int32_t> owner_prev = gbl(sl.owner_prev);
    const gptr<const typename Enc<S>::A> rho_full = gbl((const typename Enc<S>::A *)sl.rho_full);
    const int lane = __lane_id();
    const int wave = (b - g.lstart[slot]) * 2 + (threadIdx.x >> 6);
    const int nwave = (g.lstart[slot + 1] - g.lstart[slot]) * 2;
    const int nlink = min(pd.sp.count[1], pd.sp.cap);
    const gptr<A> out = gbl((A *)pd.sp.pop);
    const gptr<uint32_t> spmask = gbl(pd.sp.mask);
    const PView pv = pview_old(pd);
    const int ip = lane;  // the direction of this lane (1..26 take part)
    for (int w0 = wave * 64; w0 < nlink; w0 += nwave * 64) {
        const int jl = w0 + lane;
        const uint32_t ml = jl < nlink ? spmask[jl] : 0u;
        unsigned long long multi = __ballot((ml & SP_MULTI) != 0u);
        while (multi) {
            const int j = w0 + __builtin_ctzll(multi);
            multi &= multi - 1ull;
            const int idx = pd.sp.node[j];
            const int iz = idx / d.lxy;
            const int rem = idx - iz * d.lxy;
            const int iy = rem / d.lx;
            const int ix = rem - iy * d.lx;
            // this lane's neighbour owner (-2 outside the domain and for lanes without a
            // direction), and the upstream neighbour's (lane opp) by a shuffle
            int own = -2;
            if (ip >= 1 && ip < NPOP) {
                const int nx = ix + dir_cx(ip), ny = iy + dir_cy(ip), nz = iz + dir_cz(ip);
                if (in_domain(d, nx, ny, nz)) own = owner_prev[nz * d.lxy + ny * d.lx + nx];
            }
            const int own_up = __shfl(own, (ip >= 1 && ip < NPOP) ? dir_opp(ip) : 0);
            uint32_t bit = 0u;
            if (own >= 0) {
                const double x0 = static_cast<double>(ix) + 0.5;
                const double y0 = static_cast<double>(iy) + 0.5;
                const double z0 = static_cast<double>(iz) + 0.5;
                double q = 0.0;
                if (link_q(ip, x0, y0, z0, pv.px[own], pv.py[own], pv.pz[own], pd.radius_sq, q)) {
                    const int opp = dir_opp(ip);
                    const bool up_fluid = own_up == -1;  // x - c_ip = x + c_opp, inside and fluid
                    const int idx_up = idx - (dir_cx(ip) + dir_cy(ip) * d.lx + dir_cz(ip) * d.lxy);
                    const double density = double(rho_full[idx]);
                    const LinkLoads Ld = link_loads<S>(ip, idx, idx_up, up_fluid, q, pop);
                    const LinkOut L = eval_link_at(Ld, ip, ix, iy, iz, up_fluid, q, pstate(pv, own),
                                                   density, pd.rho0);
                    out[sp_at(j, opp)] = A(L.f_out);
                    out[sp_at(j, NPOP + opp)] = A(L.g_out);
                    bit = 1u << opp;
                }
            }
            // OR of the lanes' bits: the replaced populations of entry j
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) bit |= (uint32_t)__shfl_xor((int)bit, o);
            if (lane == 0) spmask[j] = bit;
        }
    }
}

// Populations of the refilled nodes (the tail of the special list), one 32-lane group per node
// and one lane per population j: refill_new_fluid_kernel (refill.cu:121-215) from the
// post-collision state at neighbours that were fluid in the previous geometry, with the moved
// state for the fallback equilibrium; lane 0 then forms rho, u, phi (refill.cu:217-235) in
// population order and sets F = 0.  The particle phase stores the populations in the special
// list; the per-stage execution (STAGED) writes them into d_f and d_g (fo, go: FP64, blocked).
template <class S, bool STAGED>
__device__ inline void fixup_refill_body(const PartGrid &g, int b, double *fo, double *go) {
    const int slot = slot_of(g.sstart, g.nslots, b);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const Dims d = sl.d;
    using A = typename Enc<S>::A;
    const PopView<S> pop{gbl((const S *)sl.fcur), gbl((const S *)sl.gcur), (size_t)d.n,
                         gbl(pd.gshift), d.lx};
    const int32_t *owner_prev = sl.owner_prev;
    A *fx_field = (A *)sl.fx;
    double *macro_rho = sl.macro[0], *macro_ux = sl.macro[1], *macro_uy = sl.macro[2];
    double *macro_uz = sl.macro[3], *macro_phi = sl.macro[4];
    const int blk = b - g.sstart[slot];
    const int nblk = g.sstart[slot + 1] - g.sstart[slot];
    const int grp = threadIdx.x >> 5, lane = threadIdx.x & 31;
    const int count = min(pd.sp.count[0], pd.sp.cap);
    const int first = min(pd.sp.count[1], pd.sp.cap);
    A *out = (A *)pd.sp.pop;
    const double lambda0 = 1.0 - wt(0);
    const double RT = 1.0 / 3.0;
    // each 32-lane group walks its own nodes: no workgroup barrier couples the groups
    for (int j = first + blk * 4 + grp; j < count; j += nblk * 4) {
        const int idx = pd.sp.node[j];
        const int pid = pd.sp_kind[j];
        const int iz = idx / d.lxy;
        const int rem = idx - iz * d.lxy;
        const int iy = rem / d.lx;
        const int ix = rem - iy * d.lx;
        // Which directions d give a sample (refill.cu:121-153): x + c_d inside and fluid in the
        // previous geometry, and whether x + 2c_d is too (the second-order sample).  This depends
        // only on the geometry, so lane d decides direction d for the whole group, with its two
        // owner loads in flight together, and the group shares the answers as bit masks.
        bool s1 = false, s2 = false;
        if (lane >= 1 && lane < NPOP) {
            const int cx = dir_cx(lane), cy = dir_cy(lane), cz = dir_cz(lane);
            const int off = cx + cy * d.lx + cz * d.lxy;
            const bool in1 = in_domain(d, ix + cx, iy + cy, iz + cz);
            const bool in2 = in_domain(d, ix + 2 * cx, iy + 2 * cy, iz + 2 * cz);
            const int o1 = in1 ? owner_prev[idx + off] : 0;
            const int o2 = in2 ? owner_prev[idx + 2 * off] : 0;
            s1 = in1 && o1 == -1;
            s2 = s1 && in2 && o2 == -1;
        }
        const int half = threadIdx.x & 32;  // the group's half of its wave
        const uint32_t samp = (uint32_t)(__ballot(s1) >> half);  // bit d: x + c_d is sampled
        const uint32_t ext = (uint32_t)(__ballot(s2) >> half);   // bit d: and x + 2c_d
        double fv = 0.0, gv = 0.0;
        if (lane < NPOP) {
            const int ipop = lane;
            double temp_f = 0.0, temp_g = 0.0;
            int sample_count = __popc(samp);
            // The samples in increasing d, as the reference adds them.  The loads of a chunk of
            // directions go out before its sums, so a node waits for memory a few times rather
            // than once or twice per direction.
            constexpr int CHUNK = 9;
#pragma unroll
            for (int d0 = 1; d0 < NPOP; d0 += CHUNK) {
                double f1[CHUNK], f2[CHUNK], g1[CHUNK], g2[CHUNK];
#pragma unroll
                for (int k = 0; k < CHUNK; k++) {
                    const int dd = d0 + k;
                    f1[k] = f2[k] = g1[k] = g2[k] = 0.0;
                    if (dd < NPOP && ((samp >> dd) & 1u)) {
                        const int off = dir_cx(dd) + dir_cy(dd) * d.lx + dir_cz(dd) * d.lxy;
                        f1[k] = pop.F(ipop, idx + off);
                        g1[k] = pop.G(ipop, idx + off);
                        if ((ext >> dd) & 1u) {
                            f2[k] = pop.F(ipop, idx + 2 * off);
                            g2[k] = pop.G(ipop, idx + 2 * off);
                        }
                    }
                }
#pragma unroll
                for (int k = 0; k < CHUNK; k++) {
                    const int dd = d0 + k;
                    if (dd < NPOP && ((samp >> dd) & 1u)) {
                        if ((ext >> dd) & 1u) {
                            temp_f += 2.0 * f1[k] - f2[k];
                            temp_g += 2.0 * g1[k] - g2[k];
                        } else {
                            temp_f += f1[k];
                            temp_g += g1[k];
                        }
                    }
                }
            }
            if (sample_count == 0) {
                const double px = pd.p[RBTP_P_POS_X][pid], py = pd.p[RBTP_P_POS_Y][pid], pz = pd.p[RBTP_P_POS_Z][pid];
                const double vx = pd.p[RBTP_P_VEL_X][pid], vy = pd.p[RBTP_P_VEL_Y][pid], vz = pd.p[RBTP_P_VEL_Z][pid];
                const double omega_x = pd.p[RBTP_P_OMEGA_X][pid], omega_y = pd.p[RBTP_P_OMEGA_Y][pid], omega_z = pd.p[RBTP_P_OMEGA_Z][pid];
                const double tempT = pd.p[RBTP_P_TEMP][pid];
                const double x = static_cast<double>(ix) + 0.5;
                const double y = static_cast<double>(iy) + 0.5;
                const double z = static_cast<double>(iz) + 0.5;
                const double relx = x - px, rely = y - py, relz = z - pz;
                const double u = vx + (relz * omega_y - rely * omega_z);
                const double v = vy + (relx * omega_z - relz * omega_x);
                const double w = vz + (rely * omega_x - relx * omega_y);
                const double r = pd.rho0;
                const double eu = dir_cx(ipop) * u + dir_cy(ipop) * v + dir_cz(ipop) * w;
                const double uv = (u * u + v * v + w * w) / RT;
                temp_f = wt(ipop) * (r + eu + 0.5 * (eu * eu - uv));
                temp_g = wt(ipop) * tempT * (1.0 + eu + 0.5 * (eu * eu - uv));
                if (ipop == 0) temp_g -= lambda0 * tempT * r;
                else temp_g += wt(ipop) * tempT * r;
                sample_count = 1;
            }
            const double inv_count = 1.0 / static_cast<double>(sample_count);
            fv = temp_f * inv_count;
            gv = temp_g * inv_count;
            if constexpr (STAGED) {
                fo[pop_at(lane, idx)] = fv;
                go[pop_at(lane, idx)] = gv;
            } else {
                out[sp_at(j, lane)] = A(fv);
                out[sp_at(j, NPOP + lane)] = A(gv);
            }
        }
        // the group's 27 populations, read across its lanes in population order (all 32 lanes
        // of the group are active here)
        double rho_loc = 0.0, mom_x = 0.0, mom_y = 0.0, mom_z = 0.0, phi_loc = 0.0;
        for (int ip = 0; ip < NPOP; ip++) {
            const double f_val = __shfl(fv, ip, 32);
            const double g_val = __shfl(gv, ip, 32);
            rho_loc += f_val;
            mom_x += dir_cx(ip) * f_val;
            mom_y += dir_cy(ip) * f_val;
            mom_z += dir_cz(ip) * f_val;
            phi_loc += g_val;
        }
        if (lane == 0) {
            macro_rho[idx] = rho_loc;
            if (rho_loc > 1e-12) {
                const double inv_rho = 1.0 / rho_loc;
                macro_ux[idx] = mom_x * inv_rho;
                macro_uy[idx] = mom_y * inv_rho;
                macro_uz[idx] = mom_z * inv_rho;
            } else {
                macro_ux[idx] = 0.0;
                macro_uy[idx] = 0.0;
                macro_uz[idx] = 0.0;
            }
            macro_phi[idx] = phi_loc;
            fx_field[idx] = A(0.0);
            pd.sp.mask[j] = (1u << NPOP) - 1u;
            if constexpr (!STAGED) pd.sp.map[idx] = j;
        }
    }
}

// The replaced populations of the special nodes over a plain pull of the whole lattice (the
// post-stream accessor): f and g are FP64 in the reference's SoA order, or in the blocked order
// of pop_at (BLOCKED, the start of the per-stage execution).
template <class S, bool STAGED = false>
__global__ __launch_bounds__(128) void k_fixup_refill(PartGrid g, double *fo = nullptr,
                                                      double *go = nullptr) {
    fixup_refill_body<S, STAGED>(g, blockIdx.x, fo, go);
}

// Both fix-ups in one launch (the link nodes' in the first total_l workgroups, the refilled
// nodes' in the next total_s): they complete different entries of the special list (the head's
// shared link nodes, the tail's refilled nodes, which were solid in the step's geometry and so
// are no link nodes) and read only what earlier kernels wrote, so their latency-bound work
// overlaps.  Four waves per SIMD, as the refill alone runs (the two bodies together would take
// one register more than that allows).
template <class S>
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4))) void k_fixups(PartGrid g) {
    if ((int)blockIdx.x < g.total_l) fixup_links_body<S>(g, blockIdx.x);
    else fixup_refill_body<S, false>(g, (int)blockIdx.x - g.total_l, nullptr, nullptr);
}

template <class A, bool BLOCKED = false>
__global__ __launch_bounds__(256) void k_overwrite_special(PartDev pd, double *f, double *g, Dims d) {
    const A *pop = (const A *)pd.sp.pop;
    const int count = min(*pd.sp.count, pd.sp.cap);
    const size_t n = (size_t)d.n;
    for (int j = blockIdx.x * 256 + threadIdx.x; j < count; j += gridDim.x * 256) {
        const int idx = pd.sp.node[j];
        const uint32_t m = pd.sp.mask[j];
        for (int i = 0; i < NPOP; i++) {
            if (!(m & (1u << i))) continue;
            const size_t at = BLOCKED ? pop_at(i, idx) : (size_t)i * n + idx;
            f[at] = double(pop[sp_at(j, i)]);
            g[at] = double(pop[sp_at(j, NPOP + i)]);
        }
    }
}

// ---- per-stage execution (rbtp_stage) -------------------------------------------------------
//
// The particle stages one at a time on the FP64 arrays of the staged mode: fcur/gcur of the slot
// are d_f_collide/d_g_collide, fo/go are d_f/d_g (blocked layout), owner_prev/box_prev the
// geometry of this step's links (ibnode = ibnode_prev until build_links), owner_curr/box_curr the
// geometry build_links writes.  Links are derived per link-box node as in the link-node pass.

enum { STL_IBB_F = 0, STL_IBB_G = 1, STL_FORCES = 2, STL_HEAT = 3 };

// IBB_F: bounce_back_particles_kernel (ibb.cu:5-82) writes f_opp(x) into d_f.
// IBB_G: bounce_back_thermal_kernel (ibb.cu:84-136) writes g_opp(x) into d_g.
// FORCES: the node's share of compute_particle_forces_kernel (ibb.cu:138-188), reading the
//         bounced-back f_opp(x) from d_f; HEAT: of compute_particle_heat_kernel (ibb.cu:190-205).
// The shares are summed by k_link_reduce in the order of the particle phase.
template <int MODE>
__global__ __launch_bounds__(256) void k_st_links(PartGrid g, double *fo, double *go) {
    const int u = blockIdx.x * 256 + threadIdx.x;
    if (u >= g.total_b) return;  // total_b is a whole number of waves
    const int u0 = __builtin_amdgcn_readfirstlane(u);
    const int slot = slot_of(g.bstart, g.nslots, u0);
    const PartSlot &sl = g.tab[slot];
    const PartDev &pd = sl.pd;
    const Dims d = sl.d;
    const int p = (u0 - g.bstart[slot]) / pd.lnk_stride;
    const int k = u - g.bstart[slot] - p * pd.lnk_stride;
    const int32_t *owner = sl.owner_prev;
    const PView pv = pview_now(pd);
    int lo[3], hi[3];
    bool in_box = load_box(sl.box_prev, p, lo, hi);
    if (in_box) {
        grow_box(d, lo, hi);
        in_box = k < box_count(lo, hi);
    }
    const double X = wave_uniform(pv.px[p]), Y = wave_uniform(pv.py[p]), Z = wave_uniform(pv.pz[p]);
    unsigned mask = 0u, fluid_nb = 0u;
    int ix = 0, iy = 0, iz = 0, idx = 0;
    double x0 = 0.0, y0 = 0.0, z0 = 0.0;
    if (in_box) {
        idx = box_node(d, lo, hi, k, ix, iy, iz);
        x0 = static_cast<double>(ix) + 0.5;
        y0 = static_cast<double>(iy) + 0.5;
        z0 = static_cast<double>(iz) + 0.5;
        if (owner[idx] == -1) {
            for (int ip = 1; ip < NPOP; ip++) {
                const int nx = ix + dir_cx(ip), ny = iy + dir_cy(ip), nz = iz + dir_cz(ip);
                if (!in_domain(d, nx, ny, nz)) continue;
                const int own = owner[nz * d.lxy + ny * d.lx + nx];
                if (own == -1) fluid_nb |= 1u << ip;
                if (own == p && link_exists(ip, x0, y0, z0, X, Y, Z, pd.radius_sq))
                    mask |= 1u << ip;
            }
        }
    }
    if (MODE == STL_FORCES || MODE == STL_HEAT) {
        const unsigned long long has = __ballot(mask != 0u);
        if (__lane_id() == 0) pd.lnk_bits[(size_t)p * (pd.lnk_stride / 64) + k / 64] = has;
    }
    if (mask == 0u) return;
    const PopView<double> pop{gbl((const double *)sl.fcur), gbl((const double *)sl.gcur),
                              (size_t)d.n, gbl(pd.gshift), d.lx};
    const double density = sl.macro[0][idx];
    const PState ps = pstate_uniform(pv, p);  // p is the same for the whole wave
    double v[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    while (mask) {
        const int ip = __builtin_ctz(mask);
        mask &= mask - 1u;
        const int opp = dir_opp(ip);
        const int cx = dir_cx(ip), cy = dir_cy(ip), cz = dir_cz(ip);
        const bool up_fluid = (fluid_nb >> opp) & 1u;  // x - c_ip = x + c_opp
        const int idx_up = idx - (cx + cy * d.lx + cz * d.lxy);
        double q = 0.0;
        link_q(ip, x0, y0, z0, X, Y, Z, pd.radius_sq, q);
        const LinkLoads Ld = link_loads<double>(ip, idx, idx_up, up_fluid, q, pop);
        const LinkOut L = eval_link_at(Ld, ip, ix, iy, iz, up_fluid, q, ps, density, pd.rho0);
        if (MODE == STL_IBB_F) fo[pop_at(opp, idx)] = L.f_out;
        if (MODE == STL_IBB_G) go[pop_at(opp, idx)] = L.g_out;
        if (MODE == STL_FORCES) {
            const double f_out = fo[pop_at(opp, idx)];
            const double rx = q * static_cast<double>(cx);
            const double ry = q * static_cast<double>(cy);
            const double rz = q * static_cast<double>(cz);
            const double relx = ix + 0.5 + rx - X;
            const double rely = iy + 0.5 + ry - Y;
            const double relz = iz + 0.5 + rz - Z;
            const double Fx = (cx - L.uwx) * L.f_in - (-cx - L.uwx) * f_out;
            const double Fy = (cy - L.uwy) * L.f_in - (-cy - L.uwy) * f_out;
            const double Fz = (cz - L.uwz) * L.f_in - (-cz - L.uwz) * f_out;
            v[0] += Fx;
            v[1] += Fy;
            v[2] += Fz;
            v[3] += rely * Fz - relz * Fy;
            v[4] += relz * Fx - relx * Fz;
            v[5] += relx * Fy - rely * Fx;
        }
        if (MODE == STL_HEAT) v[6] += L.g_in - go[pop_at(opp, idx)];
    }
    const size_t plane = (size_t)pd.np * pd.lnk_stride;
    const size_t e = (size_t)p * pd.lnk_stride + k;
    if (MODE == STL_FORCES)
        for (int c = 0; c < 6; c++) pd.lnk_part[c * plane + e] = v[c];
    if (MODE == STL_HEAT) pd.lnk_part[6 * plane + e] = v[6];
}

// prepare_particle_forces_step (pforces.cu:80-116).
__global__ __launch_bounds__(128) void k_st_prepare(PartGrid g) {
    const int u = blockIdx.x * 128 + threadIdx.x;
    if (u >= g.total_p) return;
    const int s = slot_of(g.pstart, g.nslots, u);
    double *const *P = g.tab[s].pd.p;
    const int i = u - g.pstart[s];
    for (int a = 0; a < 3; a++) {
        P[RBTP_P_FORCE_PREV2_X + a][i] = P[RBTP_P_FORCE_PREV_X + a][i];
        P[RBTP_P_TORQUE_PREV2_X + a][i] = P[RBTP_P_TORQUE_PREV_X + a][i];
    }
    for (int a = 0; a < 3; a++) {
        P[RBTP_P_FORCE_PREV_X + a][i] = P[RBTP_P_FORCE_X + a][i];
        P[RBTP_P_TORQUE_PREV_X + a][i] = P[RBTP_P_TORQUE_X + a][i];
    }
    for (int a = 0; a < 3; a++) {
        P[RBTP_P_FORCE_X + a][i] = 0.0;
        P[RBTP_P_TORQUE_X + a][i] = 0.0;
    }
}

// apply_repulsive_forces (pforces.cu:118-148): every partner in ascending order
// (repulsive_pair_kernel, pforces.cu:6-38), then the six walls (repulsive_wall_kernel,
// pforces.cu:40-76), added onto the accumulated force.
__global__ __launch_bounds__(128) void k_st_repulsion(PartGrid g) {
    const int u = blockIdx.x * 128 + threadIdx.x;
    if (u >= g.total_p) return;
    const int s = slot_of(g.pstart, g.nslots, u);
    const PartDev &pd = g.tab[s].pd;
    const Dims d = g.tab[s].d;
    const int i = u - g.pstart[s];
    double *const *P = pd.p;
    const double *X = P[RBTP_P_POS_X], *Y = P[RBTP_P_POS_Y], *Z = P[RBTP_P_POS_Z];
    const double xi = X[i], yi = Y[i], zi = Z[i];
    double FX = P[RBTP_P_FORCE_X][i], FY = P[RBTP_P_FORCE_Y][i], FZ = P[RBTP_P_FORCE_Z][i];
    if (pd.np > 1)
        for (int j = 0; j < pd.np; j++) {
            if (j == i) continue;
            double fx, fy, fz;
            if (pair_term(pd, X, Y, Z, xi, yi, zi, j, fx, fy, fz)) {
                FX += fx;
                FY += fy;
                FZ += fz;
            }
        }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const double mgw = pd.min_gap_w, radius = pd.radius, coeff = pd.coeff;
    auto wall = [&](double gap, double nx, double ny, double nz) {
        if (gap >= mgw) return;
        const double overlap = (gap < 0.0) ? -gap : 0.0;
        const double factor = (mgw - gap) / mgw;
        const double lubforce = coeff * (factor * factor / pd.stf0_w + overlap / (mgw * pd.stf1_w));
        fx += lubforce * nx;
        fy += lubforce * ny;
        fz += lubforce * nz;
    };
    wall(xi - radius, 1.0, 0.0, 0.0);
    wall((static_cast<double>(d.lx) - xi) - radius, -1.0, 0.0, 0.0);
    wall(yi - radius, 0.0, 1.0, 0.0);
    wall((static_cast<double>(d.ly) - yi) - radius, 0.0, -1.0, 0.0);
    wall(zi - radius, 0.0, 0.0, 1.0);
    wall((static_cast<double>(d.lz) - zi) - radius, 0.0, 0.0, -1.0);
    P[RBTP_P_FORCE_X][i] = FX + fx;
    P[RBTP_P_FORCE_Y][i] = FY + fy;
    P[RBTP_P_FORCE_Z][i] = FZ + fz;
}

// The host part of compute_particle_heat (ibb.cu:260-261): heat_prev <- heat, heat <- 0.
__global__ __launch_bounds__(128) void k_st_heat_reset(PartGrid g) {
    const int u = blockIdx.x * 128 + threadIdx.x;
    if (u >= g.total_p) return;
    const int s = slot_of(g.pstart, g.nslots, u);
    double *const *P = g.tab[s].pd.p;
    const int i = u - g.pstart[s];
    P[RBTP_P_HEAT_PREV][i] = P[RBTP_P_HEAT][i];
    P[RBTP_P_HEAT][i] = 0.0;
}

// An empty special-node list for each slot (the new-fluid nodes of refill_nodes are listed next).
__global__ void k_st_special_reset(PartGrid g) {
    for (int s = threadIdx.x; s < g.nslots; s += blockDim.x) {
        g.tab[s].pd.sp.count[0] = 0;
        g.tab[s].pd.sp.count[1] = 0;
    }
}

// Link records of the newest geometry in the reference layout (links.cu:89-139), unordered.
// A record is listed when the hot path's own test (link_exists, as in the link-node pass) accepts the
// link, so the list compared with the reference criterion is the set the coupling uses.
__global__ __launch_bounds__(PB) void k_collect_links(PartDev pd, const int32_t *owner,
                                                      const int32_t *box, Dims d, int snapshot,
                                                      rbtp_link *out, int cap, int *count) {
    const int p = blockIdx.x;
    int lo[3], hi[3];
    if (!load_box(box, p, lo, hi)) return;
    grow_box(d, lo, hi);
    // the positions the geometry was built from: the current ones, or the snapshot k_move took
    // before moving the particles (between update_particles and build_links)
    const double *PX = snapshot ? pd.old[OLD_PX] : pd.p[RBTP_P_POS_X];
    const double *PY = snapshot ? pd.old[OLD_PY] : pd.p[RBTP_P_POS_Y];
    const double *PZ = snapshot ? pd.old[OLD_PZ] : pd.p[RBTP_P_POS_Z];
    const double X = PX[p], Y = PY[p], Z = PZ[p];
    const int cnt = box_count(lo, hi);
    for (int k = threadIdx.x; k < cnt; k += PB) {
        int ix, iy, iz;
        const int idx = box_node(d, lo, hi, k, ix, iy, iz);
        if (owner[idx] != -1) continue;
        const double x0 = static_cast<double>(ix) + 0.5;
        const double y0 = static_cast<double>(iy) + 0.5;
        const double z0 = static_cast<double>(iz) + 0.5;
        for (int ip = 1; ip < NPOP; ip++) {
            const int nx = ix + dir_cx(ip), ny = iy + dir_cy(ip), nz = iz + dir_cz(ip);
            if (!in_domain(d, nx, ny, nz)) continue;
            if (owner[nz * d.lxy + ny * d.lx + nx] != p) continue;
            // the decision the link-node pass takes (link_exists), then q as it computes it
            if (!link_exists(ip, x0, y0, z0, X, Y, Z, pd.radius_sq)) continue;
            double q = 0.0;
            link_q(ip, x0, y0, z0, X, Y, Z, pd.radius_sq, q);
            const int j = atomicAdd(count, 1);
            if (j >= cap) continue;
            rbtp_link L;
            L.cell_i = ix; L.cell_j = iy; L.cell_k = iz; L.dir = ip; L.pid = p; L.pad = 0;
            L.q = q;
            L.rx = q * static_cast<double>(dir_cx(ip));
            L.ry = q * static_cast<double>(dir_cy(ip));
            L.rz = q * static_cast<double>(dir_cz(ip));
            out[j] = L;
        }
    }
}

template <class F>
hipError_t dispatch_prec(int prec, F &&fn) {
    switch (prec) {
        case RBTP_PREC_FP64: return fn((double *)nullptr);
        case RBTP_PREC_FP32: return fn((float *)nullptr);
        case RBTP_PREC_FP16S: return fn((__half *)nullptr);
        default: return fn((__hip_bfloat16 *)nullptr);
    }
}

#define PHIP(call)                                   \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) return report_hip(e_, #call); \
    } while (0)

int grid_for(int count, int block, int maxg) {
    int g = (count + block - 1) / block;
    if (g < 1) g = 1;
    return g > maxg ? maxg : g;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------

int part_alloc(PartDev &pd, const rbtp_case &c, const rbtp_derived &d, const Dims &dims) {
    std::memset(&pd, 0, sizeof(pd));
    pd.np = d.npart;
    pd.radius = d.particle_radius;
    pd.radius_sq = d.particle_radius * d.particle_radius;
    pd.mass = d.particle_mass;
    pd.inertia = d.particle_inertia;
    pd.heat_capacity = d.particle_heat_capacity;
    pd.rho_ratio = c.rho_particle_ratio;
    pd.grav0 = d.grav0;
    pd.coeff = d.particle_mass * c.repforce_scale;
    pd.min_gap = c.rep_min_gap;
    pd.min_gap_w = c.rep_min_gap_w;
    pd.stf0 = c.rep_stf0;
    pd.stf1 = c.rep_stf1;
    pd.stf0_w = c.rep_stf0_w;
    pd.stf1_w = c.rep_stf1_w;
    pd.rho0 = c.rho0;
    pd.t0 = c.t_ref;
    pd.thot = c.t_hot;
    pd.tcold = c.t_cold;
    const size_t np = (size_t)pd.np;
    PHIP(hipMalloc(&pd.base, (RBTP_P_COUNT + OLD_COUNT) * np * sizeof(double)));
    PHIP(hipMemset(pd.base, 0, (RBTP_P_COUNT + OLD_COUNT) * np * sizeof(double)));
    for (int k = 0; k < RBTP_P_COUNT; k++) pd.p[k] = pd.base + (size_t)k * np;
    for (int k = 0; k < OLD_COUNT; k++) pd.old[k] = pd.base + (size_t)(RBTP_P_COUNT + k) * np;
    for (int b = 0; b < 2; b++) {
        PHIP(hipMalloc(&pd.owner[b], (size_t)dims.n * sizeof(int32_t)));
        PHIP(hipMalloc(&pd.box[b], 6 * np * sizeof(int32_t)));
    }
    // Marking boxes span at most ceil(2R)+1 nodes per axis; link boxes two more.
    pd.box_w = (int)std::ceil(2.0 * pd.radius) + 1;
    pd.lnk_stride = link_stride(pd.box_w);
    if ((long long)pd.np * pd.lnk_stride > INT_MAX)
        return report(RBTP_E_CAPACITY, "too many particles for the link-box thread grid");
    const long long cap = special_capacity(pd.np, pd.box_w, dims.n);
    if (cap > INT_MAX) return report(RBTP_E_CAPACITY, "special-node list beyond 2^31 entries");
    pd.sp.cap = (int)cap;
    PHIP(hipMalloc(&pd.sp.node, cap * sizeof(int32_t)));
    PHIP(hipMemset(pd.sp.node, 0, cap * sizeof(int32_t)));
    PHIP(hipMalloc(&pd.sp_kind, cap * sizeof(int32_t)));
    pd.sp_elem = c.precision == RBTP_PREC_FP64 ? sizeof(double) : sizeof(float);
    PHIP(hipMalloc(&pd.sp.pop, (size_t)2 * NPOP * cap * pd.sp_elem));
    PHIP(hipMalloc(&pd.sp.mask, cap * sizeof(uint32_t)));
    PHIP(hipMemset(pd.sp.mask, 0, cap * sizeof(uint32_t)));
    PHIP(hipMalloc(&pd.sp.count, 2 * sizeof(int)));
    PHIP(hipMemset(pd.sp.count, 0, 2 * sizeof(int)));
    PHIP(hipMalloc(&pd.sp.map, (size_t)dims.n * sizeof(int32_t)));
    PHIP(hipMemset(pd.sp.map, 0xFF, (size_t)dims.n * sizeof(int32_t)));
    PHIP(hipMalloc(&pd.err, ERR_COUNT * sizeof(int32_t)));
    PHIP(hipMemset(pd.err, 0, ERR_COUNT * sizeof(int32_t)));
    PHIP(hipMalloc(&pd.lnk_part, 7 * np * (size_t)pd.lnk_stride * sizeof(double)));
    PHIP(hipMalloc(&pd.lnk_bits, np * (size_t)(pd.lnk_stride / 64) * sizeof(unsigned long long)));
    // Pair-search grid: cells 1% wider than the pair range so that rounding in the binning
    // never separates a pair in range by more than one cell; coarsened when it would hold more
    // than 8 cells per particle.
    pd.use_cells = pd.np >= CELLS_MIN_NP;
    if (pd.use_cells) {
        const int L[3] = {dims.lx, dims.ly, dims.lz};
        double w = 1.01 * (2.0 * pd.radius + std::fmax(pd.min_gap, 0.0)) + 1e-9;
        const double limit = 8.0 * (double)np + 4096.0;
        for (int attempt = 0; attempt < 64; attempt++) {
            double cells = 1.0;
            for (int a = 0; a < 3; a++) {
                pd.ncell[a] = std::max(1, (int)std::floor((double)L[a] / w));
                cells *= pd.ncell[a];
            }
            if (cells <= limit) break;
            w *= 1.1;
        }
        size_t ncells = 1;
        for (int a = 0; a < 3; a++) {
            pd.cell_scale[a] = (double)pd.ncell[a] / (double)L[a];
            ncells *= (size_t)pd.ncell[a];
        }
        PHIP(hipMalloc(&pd.cell_cnt, ncells * sizeof(int32_t)));
        PHIP(hipMalloc(&pd.cell_pid, ncells * CELL_CAP * sizeof(int32_t)));
    }
    return RBTP_OK;
}

void part_free(PartDev &pd) {
    (void)hipFree(pd.base);
    for (int b = 0; b < 2; b++) {
        (void)hipFree(pd.owner[b]);
        (void)hipFree(pd.box[b]);
    }
    (void)hipFree(pd.cell_cnt);
    (void)hipFree(pd.cell_pid);
    (void)hipFree(pd.sp.node);
    (void)hipFree(pd.sp_kind);
    (void)hipFree(pd.sp.pop);
    (void)hipFree(pd.sp.mask);
    (void)hipFree(pd.sp.count);
    (void)hipFree(pd.sp.map);
    (void)hipFree(pd.err);
    (void)hipFree(pd.lnk_part);
    (void)hipFree(pd.lnk_bits);
    std::memset(&pd, 0, sizeof(pd));
}

// generate_initial_positions (pinit.cu:97-144): glibc srand/rand rejection sampling.
int place_particles(const rbtp_case &c, const rbtp_derived &d, std::vector<double> &X,
                    std::vector<double> &Y, std::vector<double> &Z) {
    const int np = d.npart;
    const double radius = d.particle_radius;
    std::srand(c.particle_rand_seed);
    const double min_pos_x = 1.5 * radius, min_pos_y = 1.5 * radius, min_pos_z = 1.5 * radius;
    const double span_x = static_cast<double>(c.lx) - 3.0 * radius;
    const double span_y = static_cast<double>(c.ly) - 3.0 * radius;
    const double span_z = static_cast<double>(c.lz) - 3.0 * radius;
    long counter = 0;
    for (int p = 0; p < np; ++p) {
        while (true) {
            const double rx = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            const double ry = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            const double rz = static_cast<double>(std::rand()) / static_cast<double>(RAND_MAX);
            const double x = min_pos_x + rx * span_x;
            const double y = min_pos_y + ry * span_y;
            const double z = min_pos_z + rz * span_z;
            bool ok = true;
            for (int q = 0; q < p; ++q) {
                const double dx = x - X[q], dy = y - Y[q], dz = z - Z[q];
                const double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
                if (dist - 2.0 * radius < radius / 3.0) {
                    counter++;
                    if (counter > 1000000)
                        return report(RBTP_E_PLACEMENT,
                                      "Can't distribute the particles, iter. limit=1000000");
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
    return RBTP_OK;
}

// Upload a full particle state (RBTP_P_COUNT arrays of np doubles) and build both geometry
// generations from its positions (build_initial_ibnode, pinit.cu:146-179, and build_links'
// marking, links.cu:143-191).
int part_load_state(PartDev &pd, const double *host, const Dims &dims, hipStream_t st) {
    const int np = pd.np;
    PHIP(hipMemsetAsync(pd.base, 0, (size_t)(RBTP_P_COUNT + OLD_COUNT) * np * sizeof(double), st));
    PHIP(copy_to_device(pd.base, host, (size_t)RBTP_P_COUNT * np * sizeof(double), st));
    PHIP(hipMemsetAsync(pd.owner[0], 0xFF, (size_t)dims.n * sizeof(int32_t), st));
    hipLaunchKernelGGL(k_geo_mark_init, np, PB, 0, st, pd.owner[0], pd.box[0], pd.p[RBTP_P_POS_X],
                       pd.p[RBTP_P_POS_Y], pd.p[RBTP_P_POS_Z], pd.radius, pd.radius_sq, dims);
    PHIP(hipGetLastError());
    PHIP(hipMemcpyAsync(pd.owner[1], pd.owner[0], (size_t)dims.n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    PHIP(hipMemcpyAsync(pd.box[1], pd.box[0], 6 * (size_t)np * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
    PHIP(hipMemsetAsync(pd.err, 0, ERR_COUNT * sizeof(int32_t), st));
    PHIP(hipMemsetAsync(pd.sp.count, 0, 2 * sizeof(int), st));
    PHIP(hipMemsetAsync(pd.sp.map, 0xFF, (size_t)dims.n * sizeof(int32_t), st));
    PHIP(hipStreamSynchronize(st));
    pd.geo_cur = 0;
    pd.special_pending = false;
    return RBTP_OK;
}

// init_particles (pinit.cu:183-295): zero state, positions (given or placed), temperatures.
int part_init(PartDev &pd, const rbtp_case &c, const rbtp_derived &d, const Dims &dims,
              const double *ppos, const double *ptemp, hipStream_t st) {
    const int np = pd.np;
    std::vector<double> host((size_t)RBTP_P_COUNT * np, 0.0);
    std::vector<double> X(np), Y(np), Z(np);
    if (ppos) {
        for (int p = 0; p < np; p++) {
            X[p] = ppos[p];
            Y[p] = ppos[np + p];
            Z[p] = ppos[2 * np + p];
        }
    } else {
        const int r = place_particles(c, d, X, Y, Z);
        if (r != RBTP_OK) return r;
    }
    for (int p = 0; p < np; p++) {
        host[(size_t)RBTP_P_POS_X * np + p] = X[p];
        host[(size_t)RBTP_P_POS_Y * np + p] = Y[p];
        host[(size_t)RBTP_P_POS_Z * np + p] = Z[p];
        host[(size_t)RBTP_P_TEMP * np + p] = ptemp ? ptemp[p] : c.particle_init_temp;
    }
    return part_load_state(pd, host.data(), dims, st);
}

int part_head_blocks(const PartDev &pd) {
    const long long head = (long long)pd.np * pd.lnk_stride;
    return (int)std::min(std::max((head + 127) / 128, 1LL), 16384LL);
}
int part_refill_blocks(const PartDev &pd) { return std::min(std::max(pd.np, 1), 4096); }

void part_tables_free(PartTables &t) {
    (void)hipFree(t.mem);
    t.mem = nullptr;
    t.bytes = 0;
    t.g[0] = t.g[1] = PartGrid{};
}

int part_tables_build(PartTables &t, const std::vector<PartSlot> (&slots)[2], hipStream_t st) {
    // kernels already queued on st may still read the old tables
    PHIP(hipStreamSynchronize(st));
    const int n = (int)slots[0].size();
    const size_t tab_bytes = ((sizeof(PartSlot) * (size_t)(n > 0 ? n : 1) + 255) / 256) * 256;
    const size_t pre_bytes = ((sizeof(int) * (size_t)(n + 1) + 255) / 256) * 256;
    const size_t need = 2 * tab_bytes + 5 * pre_bytes;
    if (t.bytes < need) {
        (void)hipFree(t.mem);
        t.mem = nullptr;
        t.bytes = 0;
        PHIP(hipMalloc(&t.mem, need));
        t.bytes = need;
    }
    std::vector<int> pre[5];
    for (auto &v : pre) v.assign(n + 1, 0);
    for (int j = 0; j < n; j++) {
        const PartDev &pd = slots[0][j].pd;
        pre[0][j + 1] = pre[0][j] + pd.np;
        pre[1][j + 1] = pre[1][j] + (pd.use_cells ? pd.ncell[0] * pd.ncell[1] * pd.ncell[2] : 0);
        pre[2][j + 1] = pre[2][j] + part_refill_blocks(pd);
        pre[3][j + 1] = pre[3][j] + part_head_blocks(pd);
        pre[4][j + 1] = pre[4][j] + pd.np * pd.lnk_stride;
    }
    std::vector<char> host(need, 0);
    for (int ph = 0; ph < 2; ph++)
        if (n > 0) std::memcpy(host.data() + ph * tab_bytes, slots[ph].data(), sizeof(PartSlot) * n);
    for (int k = 0; k < 5; k++)
        std::memcpy(host.data() + 2 * tab_bytes + k * pre_bytes, pre[k].data(), sizeof(int) * (n + 1));
    PHIP(hipMemcpyAsync(t.mem, host.data(), need, hipMemcpyHostToDevice, st));
    PHIP(hipStreamSynchronize(st));
    for (int ph = 0; ph < 2; ph++) {
        PartGrid &g = t.g[ph];
        g.tab = (const PartSlot *)(t.mem + ph * tab_bytes);
        g.pstart = (const int *)(t.mem + 2 * tab_bytes);
        g.cstart = (const int *)(t.mem + 2 * tab_bytes + pre_bytes);
        g.sstart = (const int *)(t.mem + 2 * tab_bytes + 2 * pre_bytes);
        g.lstart = (const int *)(t.mem + 2 * tab_bytes + 3 * pre_bytes);
        g.bstart = (const int *)(t.mem + 2 * tab_bytes + 4 * pre_bytes);
        g.nslots = n;
        g.total_p = pre[0][n];
        g.total_c = pre[1][n];
        g.total_s = pre[2][n];
        g.total_l = pre[3][n];
        g.total_b = pre[4][n];
    }
    return RBTP_OK;
}

// One particle phase (main.cu:52-84 minus the lattice passes) of every slot of g, enqueued on
// st: history shift and repulsion, hydrodynamic sums, motion, the new geometry, the special-node
// list and the special nodes' corrected post-stream populations.
int part_phase(const PartGrid &g, int prec, hipStream_t st) {
    if (g.nslots == 0 || g.total_p == 0) return RBTP_OK;
    const int tp = g.total_p;
    if (g.total_c > 0) {
        hipLaunchKernelGGL(k_cell_clear, std::min((g.total_c + 255) / 256, 1024), 256, 0, st, g);
        PHIP(hipGetLastError());
    }
    // the cell binning, and the previous list leaving the node map (its last reader, the lattice
    // kernel, ran before this phase)
    const int nbin = (tp + 255) / 256;
    hipLaunchKernelGGL(k_bin_unmap, nbin + (g.total_l + 1) / 2, 256, 0, st, g, nbin);
    PHIP(hipGetLastError());
    // contacts four particles per workgroup, in a multiple of 8 workgroups ahead of the link nodes
    const int nc = ((tp + 3) / 4 + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
    PHIP(dispatch_prec(prec, [&](auto *tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_contacts_links<S>), nc + (g.total_b + 255) / 256, 256, 0, st, g, nc);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        // the hydrodynamic sums and the particle step (update_particles) in one launch
        hipLaunchKernelGGL((k_link_reduce<0, 7, true>), (tp + 256 / RL - 1) / (256 / RL), 256, 0,
                           st, g);
        return hipGetLastError();
    }));
    hipLaunchKernelGGL(k_geo_mark, tp, PB, 0, st, g);
    PHIP(hipGetLastError());
    hipLaunchKernelGGL(k_refill_find, tp, LB, 0, st, g);
    PHIP(hipGetLastError());
    PHIP(dispatch_prec(prec, [&](auto *tag) {
        using S = std::remove_pointer_t<decltype(tag)>;
        hipLaunchKernelGGL((k_fixups<S>), g.total_l + g.total_s, 128, 0, st, g);
        return hipGetLastError();
    }));
    return RBTP_OK;
}

int part_stage(int stage, const PartGrid &g, double *fo, double *go, hipStream_t st) {
    if (g.nslots == 0 || g.total_p == 0) return RBTP_OK;
    const int tp = g.total_p;
    const int nb = (g.total_b + 255) / 256, np128 = (tp + 127) / 128;
    const int nred = (tp + 256 / RL - 1) / (256 / RL);
    switch (stage) {
        case RBTP_STAGE_IBB_F:
            hipLaunchKernelGGL(k_st_links<STL_IBB_F>, nb, 256, 0, st, g, fo, go);
            break;
        case RBTP_STAGE_IBB_G:
            hipLaunchKernelGGL(k_st_links<STL_IBB_G>, nb, 256, 0, st, g, fo, go);
            break;
        case RBTP_STAGE_PREPARE:
            hipLaunchKernelGGL(k_st_prepare, np128, 128, 0, st, g);
            break;
        case RBTP_STAGE_REPULSION:
            hipLaunchKernelGGL(k_st_repulsion, np128, 128, 0, st, g);
            break;
        case RBTP_STAGE_FORCES:
            hipLaunchKernelGGL(k_st_links<STL_FORCES>, nb, 256, 0, st, g, fo, go);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL((k_link_reduce<0, 6>), nred, 256, 0, st, g);
            break;
        case RBTP_STAGE_HEAT:
            hipLaunchKernelGGL(k_st_heat_reset, np128, 128, 0, st, g);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL(k_st_links<STL_HEAT>, nb, 256, 0, st, g, fo, go);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL((k_link_reduce<6, 1>), nred, 256, 0, st, g);
            break;
        case RBTP_STAGE_MOVE:
            hipLaunchKernelGGL(k_move, np128, 128, 0, st, g);
            break;
        case RBTP_STAGE_BUILD_LINKS:
            hipLaunchKernelGGL(k_geo_clear, tp, PB, 0, st, g);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL(k_geo_mark, tp, PB, 0, st, g);
            break;
        case RBTP_STAGE_REFILL:
            hipLaunchKernelGGL(k_st_special_reset, 1, 64, 0, st, g);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL(k_refill_find, tp, LB, 0, st, g);
            PHIP(hipGetLastError());
            hipLaunchKernelGGL((k_fixup_refill<double, true>), g.total_s, 128, 0, st, g, fo, go);
            break;
        default:
            return report(RBTP_E_ARG, "not a particle stage");
    }
    PHIP(hipGetLastError());
    return RBTP_OK;
}

hipError_t part_overwrite_special_blocked(PartDev &pd, double *f, double *g, const Dims &dims,
                                          hipStream_t st) {
    if (!pd.special_pending) return hipSuccess;
    hipLaunchKernelGGL((k_overwrite_special<double, true>), grid_for(pd.sp.cap, 256, 1024), 256, 0,
                       st, pd, f, g, dims);
    return hipGetLastError();
}

int part_check_errors(PartDev &pd) {
    int32_t h[ERR_COUNT];
    PHIP(hipMemcpy(h, pd.err, sizeof(h), hipMemcpyDeviceToHost));
    if (h[ERR_SPECIAL_CAP])
        return report(RBTP_E_CAPACITY, "special-node list overflowed its capacity");
    return RBTP_OK;
}

hipError_t part_overwrite_special(PartDev &pd, double *f, double *g, const Dims &dims,
                                  hipStream_t st) {
    if (!pd.special_pending) return hipSuccess;
    if (pd.sp_elem == sizeof(double))
        hipLaunchKernelGGL((k_overwrite_special<double>), grid_for(pd.sp.cap, 256, 1024), 256, 0, st,
                           pd, f, g, dims);
    else
        hipLaunchKernelGGL((k_overwrite_special<float>), grid_for(pd.sp.cap, 256, 1024), 256, 0, st,
                           pd, f, g, dims);
    return hipGetLastError();
}

int part_collect_links(PartDev &pd, int gen, bool snapshot, const Dims &dims, hipStream_t st,
                       std::vector<rbtp_link> &out) {
    int *count = nullptr;
    PHIP(hipMalloc(&count, sizeof(int)));
    int total = 0;
    hipError_t e = hipMemsetAsync(count, 0, sizeof(int), st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_collect_links, pd.np, PB, 0, st, pd, pd.owner[gen], pd.box[gen], dims,
                           snapshot ? 1 : 0, (rbtp_link *)nullptr, 0, count);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&total, count, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    rbtp_link *buf = nullptr;
    if (e == hipSuccess && total > 0) e = hipMalloc(&buf, (size_t)total * sizeof(rbtp_link));
    if (e == hipSuccess && total > 0) {
        e = hipMemsetAsync(count, 0, sizeof(int), st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_collect_links, pd.np, PB, 0, st, pd, pd.owner[gen], pd.box[gen],
                               dims, snapshot ? 1 : 0, buf, total, count);
            e = hipGetLastError();
        }
        out.resize(total);
        if (e == hipSuccess)
            e = copy_to_host(out.data(), buf, (size_t)total * sizeof(rbtp_link), st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    (void)hipFree(buf);
    (void)hipFree(count);
    if (e != hipSuccess) return report_hip(e, "collect links");
    return RBTP_OK;
}

int part_collect_new_fluid(PartDev &pd, hipStream_t st, std::vector<std::pair<int32_t, int32_t>> &out) {
    out.clear();
    if (!pd.special_pending) return RBTP_OK;
    // the refilled nodes are the list's tail, entries count[1] .. count[0]-1
    int cnt[2] = {0, 0};
    PHIP(hipMemcpyAsync(cnt, pd.sp.count, sizeof(cnt), hipMemcpyDeviceToHost, st));
    PHIP(hipStreamSynchronize(st));
    const int count = std::min(cnt[0], pd.sp.cap);
    const int first = std::min(cnt[1], count);
    const int m = count - first;
    std::vector<int32_t> nodes(m), kinds(m);
    if (m > 0) {
        PHIP(copy_to_host(nodes.data(), pd.sp.node + first, m * sizeof(int32_t), st));
        PHIP(copy_to_host(kinds.data(), pd.sp_kind + first, m * sizeof(int32_t), st));
    }
    for (int j = 0; j < m; j++)
        if (kinds[j] >= 0) out.emplace_back(nodes[j], kinds[j]);
    return RBTP_OK;
}

}  // namespace rbtp
