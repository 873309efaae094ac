// rbtp_internal.h — types shared by the HIP translation units of librbtp.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>
#include <cmath>
#include <cstdint>
#include <cstddef>

#include "d3q27.h"

namespace rbtp {

// Record an error for rbtp_last_error() and return its code (defined with the C-ABI).
int report(int code, const char *msg);
int report_hip(hipError_t e, const char *what);

// Double-word reciprocal of a divisor d for div_rn (lattice_device.h), formed on the host:
// hi = RN(1/d), lo = RN((1 - d*hi)/d); the residual 1 - d*hi is exact in one fused multiply-add.
inline void dw_recip(double d, double &hi, double &lo) {
    hi = 1.0 / d;
    lo = std::fma(-d, hi, 1.0) / d;
}
// The pairs of the two constant divisors of the particle kernels, as dw_recip forms them:
// 6 (compute_link_q's 2a for a diagonal direction, links.cu:21) and RT = RN(1/3) (the IBB
// momentum term, ibb.cu:48-49), whose RN(1/RT) is exactly 3.
constexpr double RECIP6_HI = 0x1.5555555555555p-3, RECIP6_LO = 0x1.5555555555555p-57;
constexpr double RECIP_RT_HI = 3.0, RECIP_RT_LO = 0x1.8p-53;

// Lattice shape.  n = lx*ly*lz must stay below 2^31 (checked by rbtp_create).
struct Dims {
    int lx, ly, lz, lxy;
    int n;
};

// Per-case fluid constants, evaluated on the host in the reference's operation order.
struct FluidK {
    double tau;          // 3*visc + 0.5                       (globals.cu:29)
    double tauc;         // 3*diff + 0.5                       (globals.cu:32)
    double gb;           // grav0 * beta, the first product of forcing (macrovar.cu:44)
    double t0;           // T_REF
    double thot, tcold;  // wall temperatures
    double force_coeff;  // 1.0 - 0.5/tau                      (collision.cu:18)
    double lambda0;      // 1.0 - w0                           (collision.cu:55)
    double ralph_coeff;  // 1.0 - 1.0/(2.0*tauc)               (collision.cu:72)
    double c2t;          // (2.0*tauc)/3.0, second term of grad_denom (collision.cu:73)
    double inv_tau, inv_tauc;
    // Double-word reciprocals (hi = RN(1/d), lo = RN((1 - d*hi)/d)) of the constant divisors of
    // the strict form: RT = 1.0/3.0, 3.0, tau, tauc.
    double rt_h, rt_l, three_h, three_l, tau_h, tau_l, tauc_h, tauc_l;
    // The conduction profile T(ix) = tl0 + tls*ix between the hot and cold walls at cell
    // centres: the initial shift profile of the reduced-precision g encoding.
    double tl0, tls;
};

// Storage encodings of the populations.
//   double:           the reference's value.
//   float/half/bf16:  value/w_i for f; value/w_i - gshift[ix] for g (shifted and scaled), where
//                     gshift is a per-column temperature profile: the linear conduction profile
//                     at first, then the plane-mean temperature, refreshed every few hundred
//                     steps.  The FP32-arithmetic collision works in the same shifted variables.
// Pointer into global memory.  Loads and stores through it compile to global_* instructions;
// through a generic pointer the compiler cannot place (one read from a table in memory) they
// become flat_* instructions, which also count against the scalar-load counter and so make
// every wait for a scalar load wait for them too.
template <class T> using gptr = __attribute__((address_space(1))) T *;
template <class T> __device__ inline gptr<T> gbl(T *p) { return (gptr<T>)p; }

// Population encodings.  Stores are nontemporal (streaming): a value written in one step is
// read only in the next, after the whole sweep, so it should not take L2 space from the lines
// that neighbouring workgroups of the same sweep are still pulling from.  This shortens the
// lattice kernel by 4-5 % in every form at 320x320x160 (2.74 -> 2.62 ms in FP64).
template <class S> struct Enc;
template <> struct Enc<double> {
    using A = double;  // auxiliary per-node field type (F_x, wall rho)
    static constexpr bool scaled = false;
    __device__ static inline double ld(const double *p) { return *p; }
    __device__ static inline double ld(gptr<const double> p) { return *p; }
    __device__ static inline void st(double *p, double v) { __builtin_nontemporal_store(v, p); }
};
template <> struct Enc<float> {
    using A = float;
    static constexpr bool scaled = true;
    __device__ static inline float ld(const float *p) { return *p; }
    __device__ static inline float ld(gptr<const float> p) { return *p; }
    __device__ static inline void st(float *p, float v) { __builtin_nontemporal_store(v, p); }
};
template <> struct Enc<__half> {
    using A = float;
    static constexpr bool scaled = true;
    __device__ static inline float ld(const __half *p) { return __half2float(*p); }
    __device__ static inline float ld(gptr<const __half> p) {
        return __half2float(__ushort_as_half(*(gptr<const unsigned short>)p));
    }
    __device__ static inline void st(__half *p, float v) {
        __builtin_nontemporal_store(__half_as_ushort(__float2half_rn(v)), (unsigned short *)p);
    }
};
template <> struct Enc<__hip_bfloat16> {
    using A = float;
    static constexpr bool scaled = true;
    __device__ static inline float ld(const __hip_bfloat16 *p) { return __bfloat162float(*p); }
    __device__ static inline float ld(gptr<const __hip_bfloat16> p) {
        return __bfloat162float(__ushort_as_bfloat16(*(gptr<const unsigned short>)p));
    }
    __device__ static inline void st(__hip_bfloat16 *p, float v) {
        __builtin_nontemporal_store(__bfloat16_as_ushort(__float2bfloat16(v)), (unsigned short *)p);
    }
};

// Arguments of the fused lattice kernels.  Populations are SoA: pop[i*n + idx].
template <class S> struct StepArgs {
    const S *fin;
    const S *gin;
    S *fout;
    S *gout;
    typename Enc<S>::A *fx;        // F_x of the previous forcing, updated in place
    typename Enc<S>::A *rho_wall;  // rho at ix=0 (first ly*lz) and ix=lx-1 (next ly*lz), in place
    typename Enc<S>::A *rho_full;  // rho of this step at every node (particle coupling), or null
    // Special nodes (particle coupling), or null: sp_map[idx] is the node's entry j in the
    // special list (-1 elsewhere); bit i of sp_mask[j] marks population i as replaced by its
    // corrected post-stream value (f: sp_pop[sp_at(j, i)], g: sp_pop[sp_at(j, 27 + i)], in the
    // auxiliary type).
    const int32_t *sp_map;
    const uint32_t *sp_mask;
    const typename Enc<S>::A *sp_pop;
    double *macro_rho, *macro_ux, *macro_uy, *macro_uz, *macro_phi;  // output fields, or null
    const float *gshift;  // [lx] shift profile of the reduced-precision g encoding
    // Divergence watchdog, or null: a node with a non-finite rho or u, or |u|^2 > watch_u2, takes
    // watch[0] down to watch_step (atomic min) and adds one to watch[1].
    int32_t *watch;
    int watch_step;
    double watch_u2;
    Dims d;
    FluidK k;
};

// Post-stream populations of the special nodes (IBB link nodes and refilled nodes) that the
// particle coupling replaces, computed by the particle phase from the post-collision state; the
// next lattice kernel uses them in place of its own pull of those populations.  A link node
// replaces the populations its links bounce back (f[opp][x], ibb.cu:58-71,135); a refilled
// node replaces all 27 (refill.cu:214-215).
struct SpecialBuf {
    int *node;       // lattice index, [cap]
    void *pop;       // [cap][54]: per entry f_0..f_26 then g_0..g_26 (post-stream,
                     // post-IBB/refill; sp_at), in the auxiliary type of the precision form
                     // (FP64 for the FP64 form, FP32 for the others).  Entry-major: an entry's
                     // values share one or two cache lines, which its writer and its reader
                     // (one lattice thread) each touch once.
    uint32_t *mask;  // [cap]: bit i set when population i (f and g) is replaced
    int *count;      // device counters: [0] list length, [1] IBB link nodes at its head
    int cap;
    int32_t *map;    // [n]: list entry of each node, -1 when it is not special
};

// Position of value v (f_i: v = i, g_i: v = 27 + i) of special-list entry j in SpecialBuf::pop.
__host__ __device__ inline size_t sp_at(int j, int v) { return (size_t)j * (2 * NPOP) + v; }

// Synchronous copies between caller-owned host memory and the device on stream st, staged
// through library-owned pinned buffers (xfer.hip).
hipError_t copy_to_device(void *dst, const void *src, size_t bytes, hipStream_t st);
hipError_t copy_to_host(void *dst, const void *src, size_t bytes, hipStream_t st);

}  // namespace rbtp
