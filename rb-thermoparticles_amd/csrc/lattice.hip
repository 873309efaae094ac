// lattice.hip — fused stream + macrovar + forcing + BGK(f) + BGK(g) kernels for gfx950.
//
// State between time steps is the post-collision populations of both distributions (what the
// reference copies into d_f_collide / d_g_collide, main.cu:45,61).  One launch per time step
// pulls them across the lattice with the wall rules of streaming (streaming.cu:3-23) and
// streaming_scalar (streaming.cu:25-63), takes the moments of macrovar (macrovar.cu:3-34),
// applies forcing (macrovar.cu:36-47), collides f (collision.cu:3-33) and g
// (collision.cu:35-89), and writes the next post-collision state into the other buffer of a
// ping-pong pair.  Every population is read once and written once per step (one thread per
// node, 54 coalesced read and 54 coalesced write streams), which is the HBM roofline of the
// method: 864 B/node/step in FP64 plus the F_x auxiliary field.
//
// Two arithmetic forms:
//   * strict FP64 (collide_strict): the reference's operation order, true divisions by RT, tau
//     and tauc, no fused multiply-add (the translation unit is compiled with
//     -ffp-contract=off).  Terms whose velocity component is zero are skipped; adding or
//     multiplying an exact zero changes at most the sign of a zero, never a value.
//   * FP32 (collide_shifted): the same model in FP32 arithmetic with divisions replaced by
//     reciprocal multiplications and contraction allowed, on g shifted by the node's conduction
//     temperature and scaled by 1/w_i (the variables the FP32 and FP16/BF16 storage forms
//     hold).  Used by the FP32 and FP16/BF16-storage forms.
#include "lattice.h"
#include "../../include/rbtp.h"
#include "lattice_device.h"

#include <type_traits>
#include <vector>

namespace rbtp {

namespace {

constexpr int BLOCK = 256;

// Node outputs shared by the lattice kernel and the special-node kernel.
template <class S> struct NodeOut {
    S *fout;
    S *gout;
};

// Pieces of the strict FP64 collisions, shared by the fused kernel and the per-stage kernels.
// Every rounding is the one of the reference's expression (divisions via div_rn, which returns
// the IEEE quotient).
//   eu_i = (c_i . u) / RT and h_i = (eu_i^2 - u.u/RT)/2      (collision.cu:20-22, 64-66)
__device__ inline void strict_eq(int i, double ux, double uy, double uz, double uv,
                                 const FluidK &k, double &eu, double &h) {
    const double RT = 1.0 / 3.0;
    eu = (i == 0) ? 0.0 : div_rn(cdot<double>(i, ux, uy, uz), RT, k.rt_h, k.rt_l);
    h = 0.5 * (eu * eu - uv);
}
// collision_BGK for one population (collision.cu:16-32); X = (1 + eu) + h, which g's
// equilibrium uses too.
__device__ inline double strict_f_new(int i, double fpop, double rho, double ux, double eu,
                                      double h, double fx, const FluidK &k, double &X) {
    const double RT = 1.0 / 3.0;
    const double tp = weight(i);
    const double feq = tp * ((rho + eu) + h);
    double fchange = div_rn(-(fpop - feq), k.tau, k.tau_h, k.tau_l);
    X = (1.0 + eu) + h;
    const double feq_rho0 = tp * X;
    const double cuF = ((double)D3Q27::cx[i] - ux) * fx;
    fchange += div_rn(k.force_coeff * cuF, RT, k.rt_h, k.rt_l) * feq_rho0;
    return fpop + fchange;
}
// geqnew of collision_BGK_scalar (collision.cu:53-56) from X.
__device__ inline double strict_geqnew(int i, double phi, double rho, double X, const FluidK &k) {
    const double tp = weight(i);
    const double geq = tp * phi * X;
    return (i == 0) ? (geq - k.lambda0 * phi * rho) : (geq + tp * phi * rho);
}
// Pass 1 (collision.cu:57-70): G += c_i (g_i - geqnew_i), terms with c = 0 skipped.
__device__ inline void strict_g_moment(int i, double diff, double &gx, double &gy, double &gz) {
    if (D3Q27::cx[i] == 1) gx += diff;
    if (D3Q27::cx[i] == -1) gx -= diff;
    if (D3Q27::cy[i] == 1) gy += diff;
    if (D3Q27::cy[i] == -1) gy -= diff;
    if (D3Q27::cz[i] == 1) gz += diff;
    if (D3Q27::cz[i] == -1) gz -= diff;
}
// Pass 2 (collision.cu:72-87) for one population; gd = grad_denom with its reciprocal pair.
__device__ inline double strict_g_new(int i, double gpop, double geqnew, double gx, double gy,
                                      double gz, double fx, double phi, double rho, double gd,
                                      double gd_h, double gd_l, const FluidK &k) {
    const double tp = weight(i);
    double gchange;
    if (i == 0) {
        gchange = div_rn(-(gpop - geqnew), k.tauc, k.tauc_h, k.tauc_l);
    } else {
        const int cx = D3Q27::cx[i];
        const double gneq = cdot<double>(i, gx, gy, gz);
        double gradt, src;
        if (cx != 0) {
            const double eb = cx > 0 ? fx : -fx;
            gradt = div_rn(-(2.0 * gneq + eb * phi), gd, gd_h, gd_l);
            src = gradt * rho + 3.0 * phi * eb;
        } else {
            gradt = div_rn(-(2.0 * gneq), gd, gd_h, gd_l);
            src = gradt * rho;
        }
        const double Ralph = tp * k.ralph_coeff * src;
        gchange = div_rn(-(gpop - geqnew), k.tauc, k.tauc_h, k.tauc_l) + Ralph;
    }
    return gpop + gchange;
}
// grad_denom = rho/3 + 2 tauc/3 (collision.cu:73) and its double-word reciprocal.
__device__ inline double strict_grad_denom(double rho, const FluidK &k, double &h, double &l) {
    const double gd = div_rn(rho, 3.0, k.three_h, k.three_l) + k.c2t;
    h = 1.0 / gd;
    l = __builtin_fma(-gd, h, 1.0) * h;
    return gd;
}

// Strict FP64 macrovar + forcing + collision_BGK + collision_BGK_scalar for one node.
// Returns the new F_x; writes the post-collision populations.
template <class S>
__device__ inline double collide_strict(double (&f)[NPOP], double (&g)[NPOP], double fx_old,
                                        int idx, const NodeOut<S> &o, const FluidK &k,
                                        double &rho, double &ux, double &uy, double &uz,
                                        double &phi) {
    moments<double>(f, g, rho, ux, uy, uz, phi);
    ux = ux + 0.5 * fx_old;  // macrovar.cu:30; F_y = F_z = 0 always (macrovar.cu:45-46)
    const double fx = k.gb * (phi - k.t0);  // forcing, macrovar.cu:44
    const double RT = 1.0 / 3.0;
    const double u2 = ux * ux + uy * uy + uz * uz;
    const double uv = div_rn(u2, RT, k.rt_h, k.rt_l);
    double X[NPOP];  // (1 + eu) + 0.5 (eu^2 - uv), shared by f's Guo term and g's equilibrium
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        double eu, h;
        strict_eq(i, ux, uy, uz, uv, k, eu, h);
        Enc<S>::st(o.fout + pop_at(i, idx), strict_f_new(i, f[i], rho, ux, eu, h, fx, k, X[i]));
    }
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        X[i] = strict_geqnew(i, phi, rho, X[i], k);
        strict_g_moment(i, g[i] - X[i], gx, gy, gz);
    }
    double gd_h, gd_l;
    const double gd = strict_grad_denom(rho, k, gd_h, gd_l);
#pragma unroll
    for (int i = 0; i < NPOP; i++)
        Enc<S>::st(o.gout + pop_at(i, idx),
                   strict_g_new(i, g[i], X[i], gx, gy, gz, fx, phi, rho, gd, gd_h, gd_l, k));
    return fx;
}

// The same node update in FP32 arithmetic (reciprocals instead of divisions, contraction
// allowed), for the shifted encodings.  g enters and leaves in the shifted variables of
// gather_shifted, p_i = g_i / w_i - T_s; with phi' = sum w_i p_i (so phi = T_s + phi') and
// A_i = eu_i + (eu_i^2 - u.u/RT)/2 (so geq_i = w_i phi (1 + A_i)), the non-equilibrium part of
// collision_BGK_scalar (collision.cu:55-87) is
//   (g_i - geqnew_i) / w_i = p_i - phi' - phi (A_i + s_i rho),  s_0 = -(1 - w_0)/w_0, s_i = 1,
// and the update p_i <- p_i - that / tauc + (source)_i / w_i.  Every term is a deviation of
// size |T - T_s| or |u|, never the O(T) value of g itself.
template <class S, class T>
__device__ inline T collide_shifted(T (&f)[NPOP], T (&p)[NPOP], T fx_old, int idx, T ts,
                                    const NodeOut<S> &o, const FluidK &k, T &rho, T &ux, T &uy,
                                    T &uz, T &phi) {
#pragma clang fp contract(fast)
    T phis = T(0);
    rho = T(0); ux = T(0); uy = T(0); uz = T(0);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        rho += f[i];
        if (D3Q27::cx[i] == 1) ux += f[i];
        if (D3Q27::cx[i] == -1) ux -= f[i];
        if (D3Q27::cy[i] == 1) uy += f[i];
        if (D3Q27::cy[i] == -1) uy -= f[i];
        if (D3Q27::cz[i] == 1) uz += f[i];
        if (D3Q27::cz[i] == -1) uz -= f[i];
        phis += T(weight(i)) * p[i];
    }
    phi = ts + phis;
    ux = ux + T(0.5) * fx_old;
    const T fx = T(k.gb) * (phi - T(k.t0));
    const T inv_tau = T(k.inv_tau), inv_tauc = T(k.inv_tauc);
    const T uv = T(3) * (ux * ux + uy * uy + uz * uz);
    const T fc3 = T(3) * T(k.force_coeff) * fx;
    const T s0 = T(-k.lambda0 / weight(0));
    T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const T tp = T(weight(i));
        const T eu = (i == 0) ? T(0) : T(3) * cdot<T>(i, ux, uy, uz);
        const T h = T(0.5) * (eu * eu - uv);
        const T ai = eu + h;
        const T feq = tp * ((rho + eu) + h);
        const T fpop = f[i];
        const T fnew = fpop + (feq - fpop) * inv_tau + (T(D3Q27::cx[i]) - ux) * fc3 * tp * (T(1) + ai);
        enc_f<S, T>(o.fout + pop_at(i, idx), fnew, i);
        const T dneq = p[i] - phis - phi * (ai + ((i == 0) ? s0 : T(1)) * rho);
        f[i] = dneq;  // reuse the register for the non-equilibrium part
        const T wd = tp * dneq;
        if (D3Q27::cx[i] == 1) gx += wd;
        if (D3Q27::cx[i] == -1) gx -= wd;
        if (D3Q27::cy[i] == 1) gy += wd;
        if (D3Q27::cy[i] == -1) gy -= wd;
        if (D3Q27::cz[i] == 1) gz += wd;
        if (D3Q27::cz[i] == -1) gz -= wd;
    }
    const T inv_gd = T(1) / (rho * T(1.0 / 3.0) + T(k.c2t));
    const T rc = T(k.ralph_coeff);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        T pnew = p[i] - f[i] * inv_tauc;
        if (i != 0) {
            const int cx = D3Q27::cx[i];
            const T gneq = cdot<T>(i, gx, gy, gz);
            const T eb = cx == 0 ? T(0) : (cx > 0 ? fx : -fx);
            const T gradt = -(T(2) * gneq + eb * phi) * inv_gd;
            pnew += rc * (gradt * rho + T(3) * phi * eb);
        }
        Enc<S>::st(o.gout + pop_at(i, idx), float(pnew));
    }
    return fx;
}

template <class S, class T, bool STRICT>
__device__ inline T collide_node(T (&f)[NPOP], T (&g)[NPOP], T fx_old, int idx, T gshift,
                                 const NodeOut<S> &o, const FluidK &k, T &rho, T &ux, T &uy,
                                 T &uz, T &phi) {
    if constexpr (STRICT) {
        static_assert(std::is_same<T, double>::value, "strict form needs FP64 arithmetic");
        (void)gshift;
        return collide_strict<S>(f, g, fx_old, idx, o, k, rho, ux, uy, uz, phi);
    } else {
        return collide_shifted<S, T>(f, g, fx_old, idx, gshift, o, k, rho, ux, uy, uz, phi);
    }
}

// Writes shared by both kernels: F_x, wall rho, optional rho field and macro fields.
template <class S, class T>
__device__ inline void write_node_aux(const StepArgs<S> &a, int idx, int ix, int iy, int iz,
                                      T fx, T rho, T ux, T uy, T uz, T phi) {
    using A = typename Enc<S>::A;
    a.fx[idx] = A(fx);
    if (ix == 0) a.rho_wall[iz * a.d.ly + iy] = A(rho);
    if (ix == a.d.lx - 1) a.rho_wall[a.d.ly * a.d.lz + iz * a.d.ly + iy] = A(rho);
    if (a.rho_full) a.rho_full[idx] = A(rho);
    if (a.macro_rho) {
        a.macro_rho[idx] = double(rho);
        a.macro_ux[idx] = double(ux);
        a.macro_uy[idx] = double(uy);
        a.macro_uz[idx] = double(uz);
        a.macro_phi[idx] = double(phi);
    }
}

// One node's time step (shared by the single-lattice and batched kernels).  ROWS: the FP64 pull
// through gather_rows (rows of whole waves, full workgroups; edges is the workgroup's LDS).
template <class S, class T, bool STRICT, bool NOSTREAM, bool ROWS = false>
__device__ inline void node_step(const StepArgs<S> &a, int idx, RowEdges *edges = nullptr) {
    const int iz = idx / a.d.lxy;
    const int rem = idx - iz * a.d.lxy;
    const int iy = rem / a.d.lx;
    const int ix = rem - iy * a.d.lx;
    // The node's entry in the special list (IBB link node or refilled node), -1 elsewhere.  The
    // FP32-arithmetic forms read it before the pull so that its latency passes under the pull's
    // (512x512x256 with 5000 spheres in FP16 storage: 3.17 -> 2.85 ms together with the single
    // wait below); in FP64 that order measured 1 % slower, so it reads it after the pull.
    const bool special = !NOSTREAM && a.sp_map;
    int sj = (!STRICT && special) ? a.sp_map[idx] : -1;
    T rho_w = T(0);
    if (!NOSTREAM) {
        if (ix == 0) rho_w = T(a.rho_wall[iz * a.d.ly + iy]);
        else if (ix == a.d.lx - 1) rho_w = T(a.rho_wall[a.d.ly * a.d.lz + iz * a.d.ly + iy]);
    }
    // g: the post-stream value (strict FP64) or its shifted form (FP32 arithmetic)
    T f[NPOP], g[NPOP];
    T ts = T(0);
    if constexpr (ROWS) {
        static_assert(STRICT && !NOSTREAM && std::is_same<S, double>::value,
                      "the row pull is the FP64 streaming step");
        gather_rows<BLOCK / 64>(a.fin, a.gin, idx, ix, iy, iz, a.d, rho_w, a.k.thot, a.k.tcold,
                                edges, f, g);
    } else if constexpr (STRICT) {
        gather<S, T, NOSTREAM>(a.fin, a.gin, idx, ix, iy, iz, a.d, rho_w, a.gshift, T(a.k.thot),
                               T(a.k.tcold), f, g);
    } else {
        ts = T(a.gshift[ix]);
        const T dm = T(a.gshift[ix > 0 ? ix - 1 : ix]) - ts;
        const T dp = T(a.gshift[ix < a.d.lx - 1 ? ix + 1 : ix]) - ts;
        gather_shifted<S, T, NOSTREAM>(a.fin, a.gin, idx, ix, iy, iz, a.d, rho_w, ts, dm, dp,
                                       T(a.k.thot), T(a.k.tcold), f, g);
    }
    if (STRICT && special) sj = a.sp_map[idx];
    // At a special node (IBB link node or refilled node) the populations the particle phase
    // corrected replace the pulled ones; the first step after a (re)load has none.  The node's
    // values lie together in its list entry (one or two cache lines).  All of them are loaded
    // before any is used, so the node waits for them once: the shifted form of g is taken in a
    // second pass (arithmetic inside the first would wait for each load in turn).
    if (sj >= 0) {
        const uint32_t m = a.sp_mask[sj];
#pragma unroll
        for (int i = 0; i < NPOP; i++) {
            if (m & (1u << i)) {
                f[i] = T(a.sp_pop[sp_at(sj, i)]);
                g[i] = T(a.sp_pop[sp_at(sj, NPOP + i)]);
            }
        }
        if constexpr (!STRICT) {
#pragma unroll
            for (int i = 0; i < NPOP; i++)
                g[i] = (m & (1u << i)) ? g[i] * T(1.0 / weight(i)) - ts : g[i];
        }
    }
    const T fx_old = T(a.fx[idx]);
    NodeOut<S> o{a.fout, a.gout};
    T rho, ux, uy, uz, phi;
    const T fx = collide_node<S, T, STRICT>(f, g, fx_old, idx, ts, o, a.k, rho, ux, uy, uz, phi);
    write_node_aux<S, T>(a, idx, ix, iy, iz, fx, rho, ux, uy, uz, phi);
    if (a.watch) {
        // a NaN fails the comparison; an infinite rho is caught by the finiteness test
        const T u2 = ux * ux + uy * uy + uz * uz;
        if (!__builtin_isfinite(rho) || !(u2 <= T(a.watch_u2))) {
            atomicMin(a.watch, a.watch_step);
            atomicAdd(a.watch + 1, 1);
        }
    }
}

// One time step for every node.
template <class S, class T, bool STRICT, bool NOSTREAM>
__global__ __launch_bounds__(BLOCK) void k_lattice_step(const StepArgs<S> a) {
    const int nblocks = (a.d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    node_step<S, T, STRICT, NOSTREAM>(a, idx);
}

// The step of the FP32-arithmetic forms with the pull: five waves per SIMD (at most 96 VGPRs,
// no spills; the compiler's own choice, 97, allows four).  The kernel waits on memory most of
// the time, and at 512x512x256 with 5000 spheres in FP16 storage the fifth wave took its time
// from 3.66 to 3.36 ms; without spheres it is unchanged.
template <class S>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(5))) void k_lattice_step_lp(
    const StepArgs<S> a) {
    const int nblocks = (a.d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    node_step<S, float, false, false>(a, idx);
}

// The FP64 step with the row pull, for lx a multiple of 64 and n a multiple of BLOCK: every
// workgroup is full, so every wave reaches the barrier of gather_rows.  The pull's aligned loads
// touch 4 cache lines per wave and population instead of 5 for the 18 populations with c_x != 0.
__global__ __launch_bounds__(BLOCK) void k_lattice_step_rows(const StepArgs<double> a) {
    __shared__ RowEdges edges[BLOCK / 64];
    const int idx = xcd_block(a.d.n / BLOCK) * BLOCK + threadIdx.x;
    node_step<double, double, true, false, true>(a, idx, edges);
}

// Whether the FP64 step of lattice d runs with the row pull.
inline bool rows_pull(const Dims &d) { return d.lx % 64 == 0 && d.n % BLOCK == 0; }

// One time step for every node of several independent lattices (serving slots) in one launch.
// Tiles of BLK nodes are numbered across slots (tile_start[s] is slot s's first tile); a
// persistent grid walks them, so the launch shape is independent of the batch composition.
template <class S, class T, bool STRICT, int BLK>
__global__ __launch_bounds__(BLK) void k_lattice_batched(const StepArgs<S> *__restrict__ slots,
                                                         const int *__restrict__ tile_start,
                                                         int nslots, int total_tiles) {
    for (int b = blockIdx.x; b < total_tiles; b += gridDim.x) {
        int lo = 0, hi = nslots - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tile_start[mid] <= b) lo = mid;
            else hi = mid - 1;
        }
        const StepArgs<S> a = slots[lo];
        const int idx = (b - tile_start[lo]) * BLK + threadIdx.x;
        if (idx < a.d.n) node_step<S, T, STRICT, false>(a, idx);
    }
}

// Post-stream populations only (for the parity accessors and the fg restart writer), decoded to
// FP64.  Special nodes are overwritten afterwards from the special-node buffer.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_stream_only(const S *__restrict__ fin,
                                                       const S *__restrict__ gin,
                                                       const typename Enc<S>::A *rho_wall,
                                                       double *fout, double *gout, Dims d,
                                                       const float *gshift, double thot,
                                                       double tcold, int nostream) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    using T = typename std::conditional<std::is_same<S, double>::value, double, float>::type;
    T rho_w = T(0);
    if (ix == 0) rho_w = T(rho_wall[iz * d.ly + iy]);
    else if (ix == d.lx - 1) rho_w = T(rho_wall[d.ly * d.lz + iz * d.ly + iy]);
    T f[NPOP], g[NPOP];
    if (nostream)
        gather<S, T, true>(fin, gin, idx, ix, iy, iz, d, rho_w, gshift, T(thot), T(tcold), f, g);
    else
        gather<S, T, false>(fin, gin, idx, ix, iy, iz, d, rho_w, gshift, T(thot), T(tcold), f, g);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        fout[(size_t)i * d.n + idx] = double(f[i]);
        gout[(size_t)i * d.n + idx] = double(g[i]);
    }
}

// Decode a population buffer to FP64 (post-collision accessor).
template <class S>
__global__ __launch_bounds__(BLOCK) void k_decode(const S *__restrict__ in, double *out, size_t n,
                                                  int is_g, const float *gshift, int lx) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= (size_t)NPOP * n) return;
    const int i = (int)(j / n);
    using T = typename std::conditional<std::is_same<S, double>::value, double, float>::type;
    // weight(i) with a runtime index: evaluate the same constant expression.
    const double w = i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
    const T s = T(Enc<S>::ld(in + pop_at(i, (int)(j % n))));
    if constexpr (Enc<S>::scaled) {
        const double shift = double(gshift[(j % n) % (size_t)lx]);
        out[j] = is_g ? (double(s) + shift) * w : double(s) * w;
    } else {
        out[j] = double(s);
    }
}

// Encode FP64 populations into the storage form (restart load, set_populations).
template <class S>
__global__ __launch_bounds__(BLOCK) void k_encode(const double *__restrict__ in, S *out, size_t n,
                                                  int is_g, const float *gshift, int lx) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= (size_t)NPOP * n) return;
    const int i = (int)(j / n);
    const double w = i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
    const double v = in[j];
    if constexpr (Enc<S>::scaled) {
        const double shift = double(gshift[(j % n) % (size_t)lx]);
        Enc<S>::st(out + pop_at(i, (int)(j % n)), is_g ? float(v * (1.0 / w) - shift) : float(v * (1.0 / w)));
    } else {
        Enc<S>::st(out + pop_at(i, (int)(j % n)), (S)v);
    }
}

// init_f (initial.cu:335-360) and init_g (initial.cu:362-397).  temperature may be null
// (uniform t0, the reference) or a field (benchmark initial state).  Writes the post-stream
// state (f, g), phi, F_x and the wall rho; the rho field of init_f is RHO0.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_init(S *f, S *g, typename Enc<S>::A *fx_field,
                                                typename Enc<S>::A *rho_wall,
                                                const double *temperature, double rho0,
                                                double t0, double grav0, double beta,
                                                const float *gshift, Dims d,
                                                double *macro_rho, double *macro_ux,
                                                double *macro_uy, double *macro_uz,
                                                double *macro_phi) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    // init_f: r = u = v = w = 0, so eu = 0 and feq = w_i (0 + 0 + 0.5 (0 - 0)) = 0.
    const double RT = 1.0 / 3.0;
    const double uv = (0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0) / RT;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const double eu = 0.0 / RT;
        const double feq = weight(i) * (0.0 + eu + 0.5 * (eu * eu - uv));
        enc_f<S, double>(f + pop_at(i, idx), feq, i);
    }
    // init_g with r = rho = RHO0 (written by init_f, initial.cu:350) and u = 0.
    const double r = rho0;
    const double tempT = temperature ? temperature[idx] : t0;
    const double lambda0 = 1.0 - weight(0);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const double eu = 0.0 / RT;
        double geq = weight(i) * tempT * (1.0 + eu + 0.5 * (eu * eu - uv));
        if (i == 0) geq -= lambda0 * tempT * r;
        else geq += weight(i) * tempT * r;
        enc_g<S, double>(g + pop_at(i, idx), geq, i, Enc<S>::scaled ? double(gshift[ix]) : 0.0);
    }
    using A = typename Enc<S>::A;
    fx_field[idx] = A(grav0 * beta * (tempT - t0));
    if (ix == 0) rho_wall[iz * d.ly + iy] = A(rho0);
    if (ix == d.lx - 1) rho_wall[d.ly * d.lz + iz * d.ly + iy] = A(rho0);
    if (macro_rho) {
        macro_rho[idx] = rho0;
        macro_ux[idx] = 0.0;
        macro_uy[idx] = 0.0;
        macro_uz[idx] = 0.0;
        macro_phi[idx] = tempT;
    }
}

// forcing (macrovar.cu:36-47) on a loaded phi, and the wall rho from a loaded rho field
// (restart, initial.cu:206-209).
template <class S>
__global__ __launch_bounds__(BLOCK) void k_forcing_field(typename Enc<S>::A *fx,
                                                         typename Enc<S>::A *rho_wall,
                                                         const double *rho, const double *phi,
                                                         double gb, double t0, Dims d) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    using A = typename Enc<S>::A;
    fx[idx] = A(gb * (phi[idx] - t0));
    const int iz = idx / d.lxy, rem = idx - iz * d.lxy, iy = rem / d.lx, ix = rem - iy * d.lx;
    if (ix == 0) rho_wall[iz * d.ly + iy] = A(rho[idx]);
    if (ix == d.lx - 1) rho_wall[d.ly * d.lz + iz * d.ly + iy] = A(rho[idx]);
}

inline int blocks_for(size_t n) { return (int)((n + BLOCK - 1) / BLOCK); }

// Plane means of the temperature for the shift profile of the reduced-precision encodings, from
// the stored g (T = gshift[ix] + sum_i w_i p_i at every node), in a fixed order: workgroup z
// sums its lattice plane's rows iy = 0, 1, ... for each column ix (thread ix), then
// k_shift_update adds the z partials in increasing z.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_shift_partial(const S *__restrict__ g, Dims d,
                                                         double *__restrict__ partial) {
    const int iz = blockIdx.y;
    const int ix = blockIdx.x * BLOCK + threadIdx.x;
    if (ix >= d.lx) return;
    double acc = 0.0;
    for (int iy = 0; iy < d.ly; iy++) {
        const int idx = iz * d.lxy + iy * d.lx + ix;
        double t = 0.0;
#pragma unroll
        for (int i = 0; i < NPOP; i++) t += weight(i) * double(Enc<S>::ld(g + pop_at(i, idx)));
        acc += t;
    }
    partial[(size_t)iz * d.lx + ix] = acc;
}

// New shift profile = the plane means; delta[ix] = old - new, by which every stored g value of
// column ix changes (k_shift_reencode).
__global__ __launch_bounds__(BLOCK) void k_shift_update(const double *__restrict__ partial, Dims d,
                                                        float *gshift, float *delta) {
    const int ix = blockIdx.x * BLOCK + threadIdx.x;
    if (ix >= d.lx) return;
    double sum = 0.0;
    for (int iz = 0; iz < d.lz; iz++) sum += partial[(size_t)iz * d.lx + ix];
    const float old = gshift[ix];
    const float now = float(double(old) + sum / (double(d.ly) * double(d.lz)));
    gshift[ix] = now;
    delta[ix] = old - now;
}

// p_i <- p_i + delta[ix] for every stored g value: the same g under the new shift profile.
template <class S>
__global__ __launch_bounds__(BLOCK) void k_shift_reencode(S *__restrict__ g, Dims d,
                                                          const float *__restrict__ delta) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const float dl = delta[idx % d.lx];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        S *q = g + pop_at(i, idx);
        Enc<S>::st(q, float(Enc<S>::ld(q)) + dl);
    }
}

// ---- per-stage execution (rbtp_stage) ---------------------------------------------------------
//
// One reference stage per launch on the FP64 arrays of StageFluid, with the strict form's
// arithmetic (the functions above, shared with the fused kernel).

// macrovar (macrovar.cu:3-34): moments of the post-stream populations and the previous F_x.
__global__ __launch_bounds__(BLOCK) void k_st_macrovar(const StageFluid a) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    double f[NPOP], g[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        f[i] = a.f[pop_at(i, idx)];
        g[i] = a.g[pop_at(i, idx)];
    }
    double rho, ux, uy, uz, phi;
    moments<double>(f, g, rho, ux, uy, uz, phi);
    a.rho[idx] = rho;
    a.ux[idx] = ux + 0.5 * a.fx[idx];
    a.uy[idx] = uy;
    a.uz[idx] = uz;
    a.phi[idx] = phi;
}

// forcing (macrovar.cu:36-47).
__global__ __launch_bounds__(BLOCK) void k_st_forcing(const StageFluid a) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    a.fx[idx] = a.k.gb * (a.phi[idx] - a.k.t0);
}

// collision_BGK (collision.cu:3-33) in place on d_f.
__global__ __launch_bounds__(BLOCK) void k_st_collide_f(const StageFluid a) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    const double rho = a.rho[idx], ux = a.ux[idx], uy = a.uy[idx], uz = a.uz[idx];
    const double fx = a.fx[idx];
    const double RT = 1.0 / 3.0;
    const double uv = div_rn(ux * ux + uy * uy + uz * uz, RT, a.k.rt_h, a.k.rt_l);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        double eu, h, X;
        strict_eq(i, ux, uy, uz, uv, a.k, eu, h);
        double *p = a.f + pop_at(i, idx);
        *p = strict_f_new(i, *p, rho, ux, eu, h, fx, a.k, X);
    }
}

// collision_BGK_scalar (collision.cu:35-89) in place on d_g.
__global__ __launch_bounds__(BLOCK) void k_st_collide_g(const StageFluid a) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= a.d.n) return;
    const double rho = a.rho[idx], ux = a.ux[idx], uy = a.uy[idx], uz = a.uz[idx];
    const double phi = a.phi[idx], fx = a.fx[idx];
    const double RT = 1.0 / 3.0;
    const double uv = div_rn(ux * ux + uy * uy + uz * uz, RT, a.k.rt_h, a.k.rt_l);
    double g[NPOP], geqnew[NPOP];
    double gx = 0.0, gy = 0.0, gz = 0.0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        double eu, h;
        strict_eq(i, ux, uy, uz, uv, a.k, eu, h);
        g[i] = a.g[pop_at(i, idx)];
        geqnew[i] = strict_geqnew(i, phi, rho, (1.0 + eu) + h, a.k);
        strict_g_moment(i, g[i] - geqnew[i], gx, gy, gz);
    }
    double gd_h, gd_l;
    const double gd = strict_grad_denom(rho, a.k, gd_h, gd_l);
#pragma unroll
    for (int i = 0; i < NPOP; i++)
        a.g[pop_at(i, idx)] =
            strict_g_new(i, g[i], geqnew[i], gx, gy, gz, fx, phi, rho, gd, gd_h, gd_l, a.k);
}

// dst_i(x) <- src_i(x) for the 27 populations of one distribution (f and g of a pair buffer
// are interleaved by runs, so a distribution is not one contiguous range).
__global__ __launch_bounds__(BLOCK) void k_st_copy(const double *__restrict__ src,
                                                   double *__restrict__ dst, int n) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= n) return;
#pragma unroll
    for (int i = 0; i < NPOP; i++) dst[pop_at(i, idx)] = src[pop_at(i, idx)];
}

// streaming (streaming.cu:3-23, thermal = 0) and streaming_scalar (streaming.cu:25-63,
// thermal = 1): dst_i(x) = src_i(x - c_i), or the wall rule where x - c_i is outside (z, then y
// walls reflect; the x walls of g apply -src_opp(x) + 2 w_i T_wall (1 + rho(x)), with rho from
// the field of this step's macrovar, or from the wall rho of the state the staged mode started
// from).
__global__ __launch_bounds__(BLOCK) void k_st_pull(const double *__restrict__ src,
                                                   double *__restrict__ dst,
                                                   const double *__restrict__ rho_node,
                                                   const double *__restrict__ rho_wall,
                                                   int thermal, Dims d, double thot,
                                                   double tcold) {
    const int idx = blockIdx.x * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    double rw = 0.0;
    if (ix == 0 || ix == d.lx - 1) {
        if (rho_node) rw = rho_node[idx];
        else rw = rho_wall[(ix == 0 ? 0 : d.ly * d.lz) + iz * d.ly + iy];
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool outx = (cx == 1 && ix == 0) || (cx == -1 && ix == d.lx - 1);
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        const bool inside = !(outx || outy || outz);
        const int off = cx + cy * d.lx + cz * d.lxy;
        double v = inside ? src[pop_at(i, idx - off)] : src[pop_at(D3Q27::opp[i], idx)];
        if (thermal && outx && !outy && !outz) {
            const double twall = (cx == 1) ? thot : tcold;
            v = -v + 2.0 * weight(i) * twall * (1.0 + rw);
        }
        dst[pop_at(i, idx)] = v;
    }
}

// The FP64 lattice kernel's memory traffic without its arithmetic: the same launch shape, the
// same pull of 54 populations under the wall rules (gather, or gather_rows where the kernel uses
// the row pull), the same F_x read and write, and the 54 values stored unchanged into the other
// pair buffer.  Its time is the ceiling the access pattern itself reaches on the device.
template <bool ROWS>
__global__ __launch_bounds__(BLOCK) void k_stencil_probe(const double *__restrict__ fin,
                                                         const double *__restrict__ gin,
                                                         double *__restrict__ fout,
                                                         double *__restrict__ gout,
                                                         double *__restrict__ fx, Dims d) {
    __shared__ RowEdges edges[ROWS ? BLOCK / 64 : 1];
    const int nblocks = (d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (!ROWS && idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    double f[NPOP], g[NPOP];
    if constexpr (ROWS)
        gather_rows<BLOCK / 64>(fin, gin, idx, ix, iy, iz, d, 0.0, 0.0, 0.0, edges, f, g);
    else
        gather<double, double, false>(fin, gin, idx, ix, iy, iz, d, 0.0, nullptr, 0.0, 0.0, f, g);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        Enc<double>::st(fout + pop_at(i, idx), f[i]);
        Enc<double>::st(gout + pop_at(i, idx), g[i]);
    }
    fx[idx] = fx[idx] + f[0];
}

// The FP64 divisions and square roots of the strict kernels, one operand pair per thread, with
// the same inline functions the kernels use (lattice_arith_probe).
__global__ __launch_bounds__(BLOCK) void k_arith_probe(int op, const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ bh,
                                                       const double *__restrict__ bl, size_t n,
                                                       double *__restrict__ out) {
    const size_t j = (size_t)blockIdx.x * BLOCK + threadIdx.x;
    if (j >= n) return;
    const double x = a[j];
    double r = 0.0;
    if (op == RBTP_ARITH_DIV) {
        // the pair formed on the device, as collide_strict does for grad_denom
        const double d = b[j];
        const double h = 1.0 / d;
        const double l = __builtin_fma(-d, h, 1.0) * h;
        r = div_rn(x, d, h, l);
    } else if (op == RBTP_ARITH_DIV_HOSTRECIP) {
        r = div_rn(x, b[j], bh[j], bl[j]);
    } else if (op == RBTP_ARITH_DIV6) {
        r = div_rn(x, 6.0, RECIP6_HI, RECIP6_LO);
    } else if (op == RBTP_ARITH_DIV_RT) {
        r = div_rn(x, 1.0 / 3.0, RECIP_RT_HI, RECIP_RT_LO);
    } else {
        r = sqrt(x);
    }
    out[j] = r;
}

}  // namespace

int lattice_arith_probe(int op, const double *a, const double *b, size_t n, double *out) {
    if (op < RBTP_ARITH_DIV || op > RBTP_ARITH_SQRT)
        return report(RBTP_E_ARG, "arith probe: unknown operation");
    if (n == 0) return RBTP_OK;
    const bool needs_b = op == RBTP_ARITH_DIV || op == RBTP_ARITH_DIV_HOSTRECIP;
    if (!a || !out || (needs_b && !b)) return report(RBTP_E_ARG, "arith probe: null operand");
    const size_t bytes = n * sizeof(double);
    std::vector<double> hh, hl;
    if (op == RBTP_ARITH_DIV_HOSTRECIP) {
        hh.resize(n);
        hl.resize(n);
        for (size_t j = 0; j < n; j++) dw_recip(b[j], hh[j], hl[j]);
    }
    double *buf = nullptr;
    hipStream_t st = nullptr;
    hipError_t e = hipMalloc(&buf, 5 * bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    double *da = buf, *db = buf + n, *dh = buf + 2 * n, *dl = buf + 3 * n, *dout = buf + 4 * n;
    if (e == hipSuccess) e = copy_to_device(da, a, bytes, st);
    if (e == hipSuccess && needs_b) e = copy_to_device(db, b, bytes, st);
    if (e == hipSuccess && !hh.empty()) e = copy_to_device(dh, hh.data(), bytes, st);
    if (e == hipSuccess && !hl.empty()) e = copy_to_device(dl, hl.data(), bytes, st);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_arith_probe, (unsigned)((n + BLOCK - 1) / BLOCK), BLOCK, 0, st, op,
                           (const double *)da, (const double *)db, (const double *)dh,
                           (const double *)dl, n, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = copy_to_host(out, dout, bytes, st);
    if (st) (void)hipStreamDestroy(st);
    (void)hipFree(buf);
    if (e != hipSuccess) return report_hip(e, "arith probe");
    return RBTP_OK;
}

// The two steps of in-place (AA) streaming as access patterns, for comparison with the pull
// above: one buffer per distribution, 54 loads and 54 stores per node plus F_x, the same launch
// shape.  Even step: a node reads its own slots i and writes its own slots opp(i).  Odd step: a
// node reads slot i at x + c_i (its own slot opp(i) where x + c_i is outside) and writes the
// same locations; each location belongs to exactly one node, so both steps are race-free.  A
// value is stored with 1 added, so that no store can be dropped as rewriting what was read.
__global__ __launch_bounds__(BLOCK) void k_aa_even_probe(double *__restrict__ f,
                                                         double *__restrict__ g,
                                                         double *__restrict__ fx, Dims d) {
    const int nblocks = (d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    double fv[NPOP], gv[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        fv[i] = f[pop_at(i, idx)];
        gv[i] = g[pop_at(i, idx)];
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        Enc<double>::st(f + pop_at(D3Q27::opp[i], idx), fv[i] + 1.0);
        Enc<double>::st(g + pop_at(D3Q27::opp[i], idx), gv[i] + 1.0);
    }
    fx[idx] = fx[idx] + fv[0];
}

__global__ __launch_bounds__(BLOCK) void k_aa_odd_probe(double *__restrict__ f,
                                                        double *__restrict__ g,
                                                        double *__restrict__ fx, Dims d) {
    const int nblocks = (d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    double fv[NPOP], gv[NPOP];
    size_t loc[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool inside = ix + cx >= 0 && ix + cx < d.lx && iy + cy >= 0 && iy + cy < d.ly &&
                            iz + cz >= 0 && iz + cz < d.lz;
        const int off = cx + cy * d.lx + cz * d.lxy;
        loc[i] = inside ? pop_at(i, idx + off) : pop_at(D3Q27::opp[i], idx);
        fv[i] = f[loc[i]];
        gv[i] = g[loc[i]];
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        Enc<double>::st(f + loc[i], fv[D3Q27::opp[i]] + 1.0);
        Enc<double>::st(g + loc[i], gv[D3Q27::opp[i]] + 1.0);
    }
    fx[idx] = fx[idx] + fv[0];
}

// ---- 16-bit storage patterns (FP16/BF16 forms): 2-byte populations, F_x as float ----------
// Each stores what it pulled plus one (so that no store rewrites the value it read), like the
// FP64 probes; 216 B of populations and 8 B of F_x move per node.

// The pull of gather_shifted: one node per thread, every population one 2-byte load.
__global__ __launch_bounds__(BLOCK) void k_pull16_probe(const unsigned short *__restrict__ fin,
                                                        const unsigned short *__restrict__ gin,
                                                        unsigned short *__restrict__ fout,
                                                        unsigned short *__restrict__ gout,
                                                        float *__restrict__ fx, Dims d) {
    const int nblocks = (d.n + BLOCK - 1) / BLOCK;
    const int idx = xcd_block(nblocks) * BLOCK + threadIdx.x;
    if (idx >= d.n) return;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    unsigned f[NPOP], g[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool outx = (cx == 1 && ix == 0) || (cx == -1 && ix == d.lx - 1);
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        const bool inside = !(outx || outy || outz);
        const size_t src = inside ? pop_at(i, idx - (cx + cy * d.lx + cz * d.lxy))
                                  : pop_at(D3Q27::opp[i], idx);
        f[i] = fin[src];
        g[i] = gin[src];
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        __builtin_nontemporal_store((unsigned short)(f[i] + 1u), fout + pop_at(i, idx));
        __builtin_nontemporal_store((unsigned short)(g[i] + 1u), gout + pop_at(i, idx));
    }
    fx[idx] = fx[idx] + (float)f[0];
}

__device__ inline unsigned lane_shr1(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ inline unsigned lane_shl1(unsigned v) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xf, 0xf, false);
}

// The row pull of gather_rows on 2-byte populations (lx a multiple of 64, n of BLOCK): aligned
// loads, the +-1 x shift across lanes (DPP), the wave's edge values through LDS, or a load at
// the workgroup's edge, or the x-wall copy.
__global__ __launch_bounds__(BLOCK) void k_rows16_probe(const unsigned short *__restrict__ fin,
                                                        const unsigned short *__restrict__ gin,
                                                        unsigned short *__restrict__ fout,
                                                        unsigned short *__restrict__ gout,
                                                        float *__restrict__ fx, Dims d) {
    constexpr int NW = BLOCK / 64;
    __shared__ unsigned last[NW][NSHIFT], first[NW][NSHIFT];
    const int idx = xcd_block(d.n / BLOCK) * BLOCK + threadIdx.x;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    unsigned fa[NPOP], ga[NPOP];
    bool yz_out[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        yz_out[i] = outy || outz;
        const size_t a = yz_out[i] ? pop_at(D3Q27::opp[i], idx)
                                   : pop_at(i, idx - cy * d.lx - cz * d.lxy);
        fa[i] = fin[a];
        ga[i] = gin[a];
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        if (D3Q27::cx[i] == 0) continue;
        if (D3Q27::cx[i] == 1 && lane == 63) { last[w][k] = fa[i]; last[w][k + 1] = ga[i]; }
        if (D3Q27::cx[i] == -1 && lane == 0) { first[w][k] = fa[i]; first[w][k + 1] = ga[i]; }
        k += 2;
    }
    __syncthreads();
    unsigned f[NPOP], g[NPOP];
    k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        if (cx == 0 || yz_out[i]) {
            f[i] = fa[i];
            g[i] = ga[i];
            if (cx) k += 2;
            continue;
        }
        unsigned fv = cx == 1 ? lane_shr1(fa[i]) : lane_shl1(fa[i]);
        unsigned gv = cx == 1 ? lane_shr1(ga[i]) : lane_shl1(ga[i]);
        if (cx == 1 ? lane == 0 : lane == 63) {
            const int nw = cx == 1 ? w - 1 : w + 1;
            if (cx == 1 ? ix == 0 : ix == d.lx - 1) {
                fv = fin[pop_at(D3Q27::opp[i], idx)];
                gv = gin[pop_at(D3Q27::opp[i], idx)];
            } else if (nw >= 0 && nw < NW) {
                fv = cx == 1 ? last[nw][k] : first[nw][k];
                gv = cx == 1 ? last[nw][k + 1] : first[nw][k + 1];
            } else {
                const size_t a = pop_at(i, idx - (cx + cy * d.lx + cz * d.lxy));
                fv = fin[a];
                gv = gin[a];
            }
        }
        f[i] = fv;
        g[i] = gv;
        k += 2;
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        __builtin_nontemporal_store((unsigned short)(f[i] + 1u), fout + pop_at(i, idx));
        __builtin_nontemporal_store((unsigned short)(g[i] + 1u), gout + pop_at(i, idx));
    }
    fx[idx] = fx[idx] + (float)f[0];
}

// Two nodes of a row per thread (lx even, n a multiple of 2 BLOCK): each population one aligned
// 4-byte word; an x-shifted one takes the neighbouring lane's word (DPP, the edge lane's through
// LDS or a load) and byte-aligns the pair; a wall node copies its own opposite value.  Where two
// rows meet inside a wave, the lane holding a row's first (last) node is a wall lane, so no
// value crosses rows.
__device__ inline unsigned shifted_pair(unsigned own, unsigned nb, int cx) {
    // c_x = +1: nodes (2t, 2t+1) need elements (2t-1, 2t); c_x = -1: (2t+1, 2t+2)
    return cx == 1 ? __builtin_amdgcn_alignbyte(own, nb, 2) : __builtin_amdgcn_alignbyte(nb, own, 2);
}
__global__ __launch_bounds__(BLOCK) void k_pairs16_probe(const unsigned short *__restrict__ fin,
                                                         const unsigned short *__restrict__ gin,
                                                         unsigned short *__restrict__ fout,
                                                         unsigned short *__restrict__ gout,
                                                         float *__restrict__ fx, Dims d) {
    constexpr int NW = BLOCK / 64;
    __shared__ unsigned last[NW][NSHIFT], first[NW][NSHIFT];
    const int t = xcd_block(d.n / (2 * BLOCK)) * BLOCK + threadIdx.x;
    const int idx = 2 * t;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int iz = idx / d.lxy;
    const int rem = idx - iz * d.lxy;
    const int iy = rem / d.lx;
    const int ix = rem - iy * d.lx;
    unsigned fa[NPOP], ga[NPOP];
    bool yz_out[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        yz_out[i] = outy || outz;
        const size_t a = yz_out[i] ? pop_at(D3Q27::opp[i], idx)
                                   : pop_at(i, idx - cy * d.lx - cz * d.lxy);
        fa[i] = *(const unsigned *)(fin + a);
        ga[i] = *(const unsigned *)(gin + a);
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        if (D3Q27::cx[i] == 0) continue;
        if (D3Q27::cx[i] == 1 && lane == 63) { last[w][k] = fa[i]; last[w][k + 1] = ga[i]; }
        if (D3Q27::cx[i] == -1 && lane == 0) { first[w][k] = fa[i]; first[w][k + 1] = ga[i]; }
        k += 2;
    }
    __syncthreads();
    unsigned f[NPOP], g[NPOP];
    k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        if (cx == 0 || yz_out[i]) {
            f[i] = fa[i];
            g[i] = ga[i];
            if (cx) k += 2;
            continue;
        }
        unsigned fnb = cx == 1 ? lane_shr1(fa[i]) : lane_shl1(fa[i]);
        unsigned gnb = cx == 1 ? lane_shr1(ga[i]) : lane_shl1(ga[i]);
        // rows meet only at a pair holding a row's first or last node
        const bool wall = cx == 1 ? ix == 0 : ix + 1 == d.lx - 1;
        if ((cx == 1 ? lane == 0 : lane == 63) && !wall) {
            const int nw = cx == 1 ? w - 1 : w + 1;
            if (nw >= 0 && nw < NW) {
                fnb = cx == 1 ? last[nw][k] : first[nw][k];
                gnb = cx == 1 ? last[nw][k + 1] : first[nw][k + 1];
            } else {
                const size_t a = pop_at(i, idx - 2 * cx - cy * d.lx - cz * d.lxy);
                fnb = *(const unsigned *)(fin + a);
                gnb = *(const unsigned *)(gin + a);
            }
        }
        unsigned fv = shifted_pair(fa[i], fnb, cx), gv = shifted_pair(ga[i], gnb, cx);
        if (wall) {
            const int wid = cx == 1 ? idx : idx + 1;
            const unsigned fw = fin[pop_at(D3Q27::opp[i], wid)], gw = gin[pop_at(D3Q27::opp[i], wid)];
            fv = cx == 1 ? ((fv & 0xffff0000u) | fw) : ((fv & 0x0000ffffu) | (fw << 16));
            gv = cx == 1 ? ((gv & 0xffff0000u) | gw) : ((gv & 0x0000ffffu) | (gw << 16));
        }
        f[i] = fv;
        g[i] = gv;
        k += 2;
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        __builtin_nontemporal_store(f[i] + 0x00010001u, (unsigned *)(fout + pop_at(i, idx)));
        __builtin_nontemporal_store(g[i] + 0x00010001u, (unsigned *)(gout + pop_at(i, idx)));
    }
    float2 *fx2 = (float2 *)fx;
    float2 v = fx2[t];
    v.x += (float)(f[0] & 0xffffu);
    v.y += (float)(f[0] >> 16);
    fx2[t] = v;
}

int lattice_stencil_probe(int lx, int ly, int lz, int reps, double *gbs) {
    return lattice_streaming_probe(0, lx, ly, lz, reps, gbs);
}

int lattice_streaming_probe(int scheme, int lx, int ly, int lz, int reps, double *gbs) {
    if (lx < 2 || ly < 2 || lz < 2 || reps < 1 || (long long)lx * ly * lz > 0x7fffffffLL)
        return report(RBTP_E_ARG, "stencil probe: lattice of 2..2^31 nodes, reps >= 1");
    if (scheme < 0 || scheme > 6)
        return report(RBTP_E_ARG, "streaming probe: scheme 0 .. 6");
    Dims d;
    d.lx = lx; d.ly = ly; d.lz = lz; d.lxy = lx * ly; d.n = lx * ly * lz;
    if (scheme == 5 && (lx % 64 != 0 || d.n % BLOCK != 0))
        return report(RBTP_E_ARG, "streaming probe 5: lx a multiple of 64, nodes a multiple of 256");
    if (scheme == 6 && (lx % 2 != 0 || d.n % (2 * BLOCK) != 0))
        return report(RBTP_E_ARG, "streaming probe 6: lx even, nodes a multiple of 512");
    const bool p16 = scheme >= 4;
    // two pair buffers, as a sim holds them: f at the buffer, g G_OFFSET elements after it
    const size_t len = pop_buffer_len((size_t)d.n);
    const size_t elem = p16 ? sizeof(unsigned short) : sizeof(double);
    const size_t fx_elem = p16 ? sizeof(float) : sizeof(double);
    double *buf[2] = {nullptr, nullptr};
    double *fx = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSuccess;
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMalloc(&buf[k], len * elem);
    if (e == hipSuccess) e = hipMalloc(&fx, (size_t)d.n * fx_elem);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int k = 0; k < 2 && e == hipSuccess; k++) e = hipMemsetAsync(buf[k], 0, len * elem, st);
    if (e == hipSuccess) e = hipMemsetAsync(fx, 0, (size_t)d.n * fx_elem, st);
    const int nb = blocks_for((size_t)d.n);
    const bool rows = rows_pull(d);
    int cur = 0;
    auto launch = [&]() {
        double *f0 = buf[0], *g0 = buf[0] + G_OFFSET;
        if (scheme == 1) {
            hipLaunchKernelGGL(k_aa_even_probe, nb, BLOCK, 0, st, f0, g0, fx, d);
        } else if (scheme == 2) {
            hipLaunchKernelGGL(k_aa_odd_probe, nb, BLOCK, 0, st, f0, g0, fx, d);
        } else if (p16) {
            using U = unsigned short;
            U *fi = (U *)buf[cur], *fo = (U *)buf[1 - cur];
            if (scheme == 4)
                hipLaunchKernelGGL(k_pull16_probe, nb, BLOCK, 0, st, fi, fi + G_OFFSET, fo,
                                   fo + G_OFFSET, (float *)fx, d);
            else if (scheme == 5)
                hipLaunchKernelGGL(k_rows16_probe, d.n / BLOCK, BLOCK, 0, st, fi, fi + G_OFFSET,
                                   fo, fo + G_OFFSET, (float *)fx, d);
            else
                hipLaunchKernelGGL(k_pairs16_probe, d.n / (2 * BLOCK), BLOCK, 0, st, fi,
                                   fi + G_OFFSET, fo, fo + G_OFFSET, (float *)fx, d);
            cur = 1 - cur;
        } else {
            double *fi = buf[cur], *fo = buf[1 - cur];
            if (rows && scheme == 0)
                hipLaunchKernelGGL(k_stencil_probe<true>, nb, BLOCK, 0, st, fi, fi + G_OFFSET, fo,
                                   fo + G_OFFSET, fx, d);
            else
                hipLaunchKernelGGL(k_stencil_probe<false>, nb, BLOCK, 0, st, fi, fi + G_OFFSET,
                                   fo, fo + G_OFFSET, fx, d);
            cur = 1 - cur;
        }
        return hipGetLastError();
    };
    if (e == hipSuccess) e = launch();  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; r++) e = launch();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess) *gbs = (p16 ? 224.0 : 880.0) * (double)d.n * reps / (ms * 1e-3) / 1e9;
    for (double *p : buf) (void)hipFree(p);
    (void)hipFree(fx);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) return report_hip(e, "stencil probe");
    return RBTP_OK;
}

// ---- host launchers --------------------------------------------------------------------------

template <class S>
hipError_t launch_lattice_step(const StepArgs<S> &a, bool strict, bool nostream,
                               hipStream_t st) {
    const int nb = blocks_for((size_t)a.d.n);
    if constexpr (std::is_same<S, double>::value) {
        // FP64 storage has one arithmetic form: the strict one
        if (!strict) return hipErrorInvalidValue;
        if (nostream) hipLaunchKernelGGL((k_lattice_step<S, double, true, true>), nb, BLOCK, 0, st, a);
        else if (rows_pull(a.d)) hipLaunchKernelGGL(k_lattice_step_rows, nb, BLOCK, 0, st, a);
        else hipLaunchKernelGGL((k_lattice_step<S, double, true, false>), nb, BLOCK, 0, st, a);
    } else {
        (void)strict;
        if (nostream) hipLaunchKernelGGL((k_lattice_step<S, float, false, true>), nb, BLOCK, 0, st, a);
        else hipLaunchKernelGGL((k_lattice_step_lp<S>), nb, BLOCK, 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_stage_fluid(int stage, const StageFluid &a, hipStream_t st) {
    const int nb = blocks_for((size_t)a.d.n);
    switch (stage) {
        case RBTP_STAGE_MACROVAR:
            hipLaunchKernelGGL(k_st_macrovar, nb, BLOCK, 0, st, a);
            break;
        case RBTP_STAGE_FORCING:
            hipLaunchKernelGGL(k_st_forcing, nb, BLOCK, 0, st, a);
            break;
        case RBTP_STAGE_COLLIDE_F:
            hipLaunchKernelGGL(k_st_collide_f, nb, BLOCK, 0, st, a);
            break;
        case RBTP_STAGE_COLLIDE_G:
            hipLaunchKernelGGL(k_st_collide_g, nb, BLOCK, 0, st, a);
            break;
        case RBTP_STAGE_STREAM_F:
        case RBTP_STAGE_STREAM_G: {
            // main.cu:45,61 (the pre-stream copy), then the pull (main.cu:47-48,63-64)
            const bool g = stage == RBTP_STAGE_STREAM_G;
            hipLaunchKernelGGL(k_st_copy, nb, BLOCK, 0, st, (const double *)(g ? a.g : a.f),
                               g ? a.gc : a.fc, a.d.n);
            hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_st_pull, nb, BLOCK, 0, st, (const double *)(g ? a.gc : a.fc),
                               g ? a.g : a.f, (const double *)a.rho, (const double *)nullptr,
                               g ? 1 : 0, a.d, a.k.thot, a.k.tcold);
            break;
        }
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_stage_enter(const StageFluid &a, bool poststream, hipStream_t st) {
    const int nb = blocks_for((size_t)a.d.n);
    for (int which = 0; which < 2; which++) {
        const double *src = which ? a.gc : a.fc;
        double *dst = which ? a.g : a.f;
        if (poststream)
            hipLaunchKernelGGL(k_st_copy, nb, BLOCK, 0, st, src, dst, a.d.n);
        else
            hipLaunchKernelGGL(k_st_pull, nb, BLOCK, 0, st, src, dst, (const double *)nullptr,
                               a.rho_wall, which, a.d, a.k.thot, a.k.tcold);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

template <class S>
hipError_t launch_stream_only(const S *fin, const S *gin, const typename Enc<S>::A *rho_wall,
                              double *fout, double *gout, Dims d, const float *gshift,
                              double thot, double tcold, bool nostream, hipStream_t st) {
    hipLaunchKernelGGL((k_stream_only<S>), blocks_for((size_t)d.n), BLOCK, 0, st, fin, gin,
                       rho_wall, fout, gout, d, gshift, thot, tcold, nostream ? 1 : 0);
    return hipGetLastError();
}

template <class S>
hipError_t launch_decode(const S *in, double *out, size_t n, int is_g, const float *gshift,
                         int lx, hipStream_t st) {
    hipLaunchKernelGGL((k_decode<S>), blocks_for((size_t)NPOP * n), BLOCK, 0, st, in, out, n, is_g,
                       gshift, lx);
    return hipGetLastError();
}

template <class S>
hipError_t launch_encode(const double *in, S *out, size_t n, int is_g, const float *gshift,
                         int lx, hipStream_t st) {
    hipLaunchKernelGGL((k_encode<S>), blocks_for((size_t)NPOP * n), BLOCK, 0, st, in, out, n, is_g,
                       gshift, lx);
    return hipGetLastError();
}

template <class S>
hipError_t launch_init(S *f, S *g, typename Enc<S>::A *fx, typename Enc<S>::A *rho_wall,
                       const double *temperature, double rho0, double t0, double grav0,
                       double beta, const float *gshift, Dims d, double *mr, double *mux,
                       double *muy, double *muz, double *mphi, hipStream_t st) {
    hipLaunchKernelGGL((k_init<S>), blocks_for((size_t)d.n), BLOCK, 0, st, f, g, fx, rho_wall,
                       temperature, rho0, t0, grav0, beta, gshift, d, mr, mux, muy, muz, mphi);
    return hipGetLastError();
}

template <class S>
hipError_t launch_shift_refresh(S *g, Dims d, float *gshift, double *partial, float *delta,
                                hipStream_t st) {
    if constexpr (!Enc<S>::scaled) {
        return hipSuccess;  // FP64 storage is not shifted
    } else {
        hipLaunchKernelGGL((k_shift_partial<S>), dim3(blocks_for((size_t)d.lx), d.lz), BLOCK, 0, st,
                           (const S *)g, d, partial);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_shift_update, blocks_for((size_t)d.lx), BLOCK, 0, st,
                           (const double *)partial, d, gshift, delta);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_shift_reencode<S>), blocks_for((size_t)d.n), BLOCK, 0, st, g, d,
                           (const float *)delta);
        return hipGetLastError();
    }
}

template <class S>
hipError_t launch_forcing_field(typename Enc<S>::A *fx, typename Enc<S>::A *rho_wall,
                                const double *rho, const double *phi, double gb, double t0,
                                Dims d, hipStream_t st) {
    hipLaunchKernelGGL((k_forcing_field<S>), blocks_for((size_t)d.n), BLOCK, 0, st, fx, rho_wall,
                       rho, phi, gb, t0, d);
    return hipGetLastError();
}

template <class S>
hipError_t launch_lattice_batched(const StepArgs<S> *slots_dev, const int *tile_start_dev,
                                  int nslots, int total_tiles, int block, int grid, bool strict,
                                  hipStream_t st) {
    if (nslots <= 0 || total_tiles <= 0) return hipSuccess;
#define RBTP_BATCHED(T, STRICT_, BLK)                                                          \
    hipLaunchKernelGGL((k_lattice_batched<S, T, STRICT_, BLK>), grid, BLK, 0, st, slots_dev,   \
                       tile_start_dev, nslots, total_tiles)
    if constexpr (std::is_same<S, double>::value) {
        if (!strict) return hipErrorInvalidValue;
        switch (block) {
            case 64: RBTP_BATCHED(double, true, 64); break;
            case 128: RBTP_BATCHED(double, true, 128); break;
            default: RBTP_BATCHED(double, true, 256); break;
        }
    } else {
        (void)strict;
        switch (block) {
            case 64: RBTP_BATCHED(float, false, 64); break;
            case 128: RBTP_BATCHED(float, false, 128); break;
            default: RBTP_BATCHED(float, false, 256); break;
        }
    }
#undef RBTP_BATCHED
    return hipGetLastError();
}

#define RBTP_INSTANTIATE(S)                                                                     \
    template hipError_t launch_lattice_step<S>(const StepArgs<S> &, bool, bool, hipStream_t);   \
    template hipError_t launch_lattice_batched<S>(const StepArgs<S> *, const int *, int, int, int, \
                                                  int, bool, hipStream_t);                        \
    template hipError_t launch_stream_only<S>(const S *, const S *,                             \
                                              const typename Enc<S>::A *, double *, double *,   \
                                              Dims, const float *, double, double, bool,        \
                                              hipStream_t);                                      \
    template hipError_t launch_decode<S>(const S *, double *, size_t, int, const float *, int,   \
                                         hipStream_t);                                          \
    template hipError_t launch_encode<S>(const double *, S *, size_t, int, const float *, int,  \
                                         hipStream_t);                                          \
    template hipError_t launch_init<S>(S *, S *, typename Enc<S>::A *, typename Enc<S>::A *,    \
                                       const double *, double, double, double, double,          \
                                       const float *, Dims, double *, double *, double *,       \
                                       double *, double *, hipStream_t);                         \
    template hipError_t launch_shift_refresh<S>(S *, Dims, float *, double *, float *,          \
                                                hipStream_t);                                    \
    template hipError_t launch_forcing_field<S>(typename Enc<S>::A *, typename Enc<S>::A *,     \
                                                const double *, const double *, double, double, \
                                                Dims, hipStream_t);

RBTP_INSTANTIATE(double)
RBTP_INSTANTIATE(float)
RBTP_INSTANTIATE(__half)
RBTP_INSTANTIATE(__hip_bfloat16)

}  // namespace rbtp
