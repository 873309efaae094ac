// lattice_device.h — device helpers shared by the lattice and particle kernels: population
// encodings, the pull gather with the wall rules of streaming/streaming_scalar, moments.
#pragma once

#include "rbtp_internal.h"

namespace rbtp {

// x / d, correctly rounded, from a double-word reciprocal yh + yl of d (yh = RN(1/d)).
// q0 = RN(x*yh + RN(x*yl)) is within half an ulp plus 2^-100 relative of x/d, hence a faithful
// rounding; by Markstein's theorem (yh within half an ulp of 1/d) the residual x - q0*d is exact
// in one FMA and RN(q0 + r*yh) is the correctly rounded quotient.  For x = +-0 the sum of signed
// zeros can lose the quotient's sign (+0 for 0 / negative d when RN(x*yl) and x*yh differ in
// sign), so v_div_fixup_f64 gives the result the IEEE sign (and the special cases), as the
// compiler's own division sequence does last.  Five FP64 operations replace the ~10-instruction
// IEEE division sequence with an identical result.
__device__ inline double div_rn(double x, double d, double yh, double yl) {
    const double q0 = __builtin_fma(x, yh, x * yl);
    const double r = __builtin_fma(-q0, d, x);
    return __builtin_amdgcn_div_fixup(__builtin_fma(r, yh, q0), d, x);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.  The remapped id
// gives each XCD one contiguous range of logical workgroups (a bijection of [0, nblocks)), so
// that neighbouring workgroups, which share cache lines, share an L2.  xcd_remap takes the
// workgroup's position bid in a range of nblocks workgroups that starts at a multiple of 8 of
// the grid (its XCD is then bid % 8).
constexpr int NUM_XCD = 8;
__device__ inline int xcd_remap(int bid, int nblocks) {
    const int q = nblocks / NUM_XCD, r = nblocks % NUM_XCD;
    const int xcd = bid % NUM_XCD, local = bid / NUM_XCD;
    return xcd * q + (xcd < r ? xcd : r) + local;
}
__device__ inline int xcd_block(int nblocks) { return xcd_remap(blockIdx.x, nblocks); }

// ---- populations in and out ------------------------------------------------------------------

template <class S, class T>
__device__ inline T dec_f(S s, int i) {
    if constexpr (Enc<S>::scaled) return T(s) * T(weight(i));
    else return T(s);
}
template <class S, class T>
__device__ inline T dec_g(S s, int i, T shift) {
    if constexpr (Enc<S>::scaled) return (T(s) + shift) * T(weight(i));
    else return T(s);
}
template <class S, class T>
__device__ inline void enc_f(S *p, T v, int i) {
    if constexpr (Enc<S>::scaled) Enc<S>::st(p, float(v * T(1.0 / weight(i))));
    else Enc<S>::st(p, v);
}
template <class S, class T>
__device__ inline void enc_g(S *p, T v, int i, T shift) {
    if constexpr (Enc<S>::scaled) Enc<S>::st(p, float(v * T(1.0 / weight(i)) - shift));
    else Enc<S>::st(p, v);
}

// Pull the post-stream populations of node (ix,iy,iz) from the post-collision state.
//   f: f_i(x) = f*_i(x - c_i), or f*_opp(x) when x - c_i is outside (streaming.cu:10-21).
//   g: as f inside; outside, the z then y walls reflect (adiabatic) and the x walls apply
//      -g*_opp(x) + 2 w_i T_wall (1 + rho(x)) (streaming.cu:37-61).
// gshift: the shift profile of the reduced-precision encodings (one value per x column; unused
// for FP64 storage).
template <class S, class T, bool NOSTREAM>
__device__ inline void gather(const S *__restrict__ fin, const S *__restrict__ gin, int idx,
                              int ix, int iy, int iz, const Dims &d, T rho_w,
                              const float *__restrict__ gshift, T thot, T tcold, T (&f)[NPOP],
                              T (&g)[NPOP]) {
    T own_shift = T(0), shift_m = T(0), shift_p = T(0);  // at ix, ix - 1, ix + 1
    if constexpr (Enc<S>::scaled) {
        own_shift = T(gshift[ix]);
        shift_m = T(gshift[ix > 0 ? ix - 1 : ix]);
        shift_p = T(gshift[ix < d.lx - 1 ? ix + 1 : ix]);
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const int op = D3Q27::opp[i];
        if (NOSTREAM) {
            f[i] = dec_f<S, T>(Enc<S>::ld(fin + pop_at(i, idx)), i);
            g[i] = dec_g<S, T>(Enc<S>::ld(gin + pop_at(i, idx)), i, own_shift);
            continue;
        }
        const bool outx = (cx == 1 && ix == 0) || (cx == -1 && ix == d.lx - 1);
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        const bool inside = !(outx || outy || outz);
        const int off = cx + cy * d.lx + cz * d.lxy;
        const size_t src = inside ? pop_at(i, idx - off) : pop_at(op, idx);
        const T fv = dec_f<S, T>(Enc<S>::ld(fin + src), i);
        const T src_shift = (inside && cx == 1) ? shift_m : ((inside && cx == -1) ? shift_p : own_shift);
        const T gv = dec_g<S, T>(Enc<S>::ld(gin + src), i, src_shift);
        f[i] = fv;
        if (inside || outz || outy) {
            g[i] = gv;
        } else {
            // x wall only: hot plate at x=0 (c_x=+1), cold plate at x=LX (c_x=-1).
            const T twall = (cx == 1) ? thot : tcold;
            g[i] = -gv + T(2.0) * T(weight(i)) * twall * (T(1.0) + rho_w);
        }
    }
}

// Exchange of one 64-bit value between neighbouring lanes of a wave (DPP wave_shr:1 and
// wave_shl:1): shr gives lane l the value of lane l - 1, shl that of lane l + 1; the lane that has
// no neighbour gets 0 and is filled in by the caller.
__device__ inline double wave_shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ inline double wave_shl1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Populations pulled with c_x != 0 (18 per distribution) and the two edge values a wave hands to
// its neighbours for each of them.
constexpr int NSHIFT = 36;
struct RowEdges {
    double last[NSHIFT];   // lane 63's aligned value (lane 0 of the next wave needs it, c_x = +1)
    double first[NSHIFT];  // lane 0's aligned value (lane 63 of the previous wave, c_x = -1)
};

// gather() of the FP64 form for lattices whose rows are whole waves (lx a multiple of 64, every
// workgroup full).  A pull with c_x = +-1 reads element idx - c_x of a row that is 64-aligned, so
// each wave loads its own aligned 512 bytes (4 cache lines instead of 5) and moves the values one
// lane over (DPP); the lane at the wave's edge takes its value from the neighbouring wave of the
// workgroup through LDS, or loads it (first or last wave of the workgroup), or applies the x-wall
// rule (a row's first and last node are the lanes at the edges).  The y and z walls cover whole
// rows, so they are decided per wave.  Same values as gather(), in the same order of operations.
template <int NW>
__device__ inline void gather_rows(const double *__restrict__ fin,
                                   const double *__restrict__ gin, int idx, int ix, int iy,
                                   int iz, const Dims &d, double rho_w, double thot,
                                   double tcold, RowEdges *edges, double (&f)[NPOP],
                                   double (&g)[NPOP]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    bool rowout[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        rowout[i] = outy || outz;
        const int aidx = rowout[i] ? idx : idx - cy * d.lx - cz * d.lxy;
        const int sp = rowout[i] ? D3Q27::opp[i] : i;
        f[i] = fin[pop_at(sp, aidx)];
        g[i] = gin[pop_at(sp, aidx)];
    }
    int k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        if (D3Q27::cx[i] == 0) continue;
        if (D3Q27::cx[i] == 1 && lane == 63) {
            edges[w].last[k] = f[i];
            edges[w].last[k + 1] = g[i];
        }
        if (D3Q27::cx[i] == -1 && lane == 0) {
            edges[w].first[k] = f[i];
            edges[w].first[k + 1] = g[i];
        }
        k += 2;
    }
    __syncthreads();
    k = 0;
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        if (cx == 0) continue;
        if (!rowout[i]) {
            double fv = cx == 1 ? wave_shr1(f[i]) : wave_shl1(f[i]);
            double gv = cx == 1 ? wave_shr1(g[i]) : wave_shl1(g[i]);
            if (cx == 1 ? lane == 0 : lane == 63) {
                const int nw = w - cx;
                if (cx == 1 ? ix == 0 : ix == d.lx - 1) {
                    // x wall only: bounce-back for f, hot/cold plate for g (streaming.cu:10-21,
                    // 45-61)
                    fv = fin[pop_at(D3Q27::opp[i], idx)];
                    const double twall = (cx == 1) ? thot : tcold;
                    gv = -gin[pop_at(D3Q27::opp[i], idx)] + 2.0 * weight(i) * twall * (1.0 + rho_w);
                } else if (nw >= 0 && nw < NW) {
                    fv = cx == 1 ? edges[nw].last[k] : edges[nw].first[k];
                    gv = cx == 1 ? edges[nw].last[k + 1] : edges[nw].first[k + 1];
                } else {
                    const int sidx = idx - (cx + cy * d.lx + cz * d.lxy);
                    fv = fin[pop_at(i, sidx)];
                    gv = gin[pop_at(i, sidx)];
                }
            }
            f[i] = fv;
            g[i] = gv;
        }
        k += 2;
    }
}

// gather() for the shifted encodings (float, half, bf16), with g returned in the variables the
// FP32 collision works in: p_i = g_i / w_i - T_s, T_s = gshift[ix] the shift of the node's own
// column.  A value pulled from x - c_i is stored relative to the shift of that node's column, so
// it gains dm = gshift[ix-1] - T_s (c_x = +1) or dp = gshift[ix+1] - T_s (c_x = -1); the wall
// rules of streaming_scalar become
//   z and y walls (adiabatic):  p_i = p*_opp(x)
//   x walls:                    p_i = -p*_opp(x) + 2 (T_wall - T_s) + 2 T_wall rho(x)
// so that no O(T) quantity enters the arithmetic on g, only its deviation from T_s.
template <class S, class T, bool NOSTREAM>
__device__ inline void gather_shifted(const S *__restrict__ fin, const S *__restrict__ gin,
                                      int idx, int ix, int iy, int iz, const Dims &d, T rho_w,
                                      T ts, T dm, T dp, T thot, T tcold, T (&f)[NPOP],
                                      T (&p)[NPOP]) {
    static_assert(Enc<S>::scaled, "shifted gather needs a shifted encoding");
    // all 54 loads first, then the arithmetic, so that they are in flight together
    T sf[NPOP], sg[NPOP];
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        size_t src = pop_at(i, idx);
        if (!NOSTREAM) {
            const bool outx = (cx == 1 && ix == 0) || (cx == -1 && ix == d.lx - 1);
            const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
            const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
            const bool inside = !(outx || outy || outz);
            const int off = cx + cy * d.lx + cz * d.lxy;
            src = inside ? pop_at(i, idx - off) : pop_at(D3Q27::opp[i], idx);
        }
        sf[i] = T(Enc<S>::ld(fin + src));
        sg[i] = T(Enc<S>::ld(gin + src));
    }
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
        f[i] = sf[i] * T(weight(i));
        if (NOSTREAM || cx == 0) {
            p[i] = sg[i];
            continue;
        }
        const bool outx = (cx == 1 && ix == 0) || (cx == -1 && ix == d.lx - 1);
        const bool outy = (cy == 1 && iy == 0) || (cy == -1 && iy == d.ly - 1);
        const bool outz = (cz == 1 && iz == 0) || (cz == -1 && iz == d.lz - 1);
        const bool inside = !(outx || outy || outz);
        const T sv = sg[i];
        const T pulled = cx > 0 ? sv + (inside ? dm : T(0)) : sv + (inside ? dp : T(0));
        // x wall only: hot plate at x=0 (c_x=+1), cold plate at x=LX (c_x=-1).
        const T twall = (cx == 1) ? thot : tcold;
        const T walled = -sv + T(2.0) * (twall - ts) + T(2.0) * twall * rho_w;
        p[i] = (outx && !outy && !outz) ? walled : pulled;
    }
}

// Moments of macrovar (macrovar.cu:18-33): sequential sums in population order; terms with a
// zero velocity component are skipped (they add an exact zero).
template <class T>
__device__ inline void moments(const T (&f)[NPOP], const T (&g)[NPOP], T &rho, T &ux, T &uy,
                               T &uz, T &phi) {
    rho = T(0); ux = T(0); uy = T(0); uz = T(0); phi = T(0);
#pragma unroll
    for (int i = 0; i < NPOP; i++) {
        rho += f[i];
        if (D3Q27::cx[i] == 1) ux += f[i];
        if (D3Q27::cx[i] == -1) ux -= f[i];
        if (D3Q27::cy[i] == 1) uy += f[i];
        if (D3Q27::cy[i] == -1) uy -= f[i];
        if (D3Q27::cz[i] == 1) uz += f[i];
        if (D3Q27::cz[i] == -1) uz -= f[i];
        phi += g[i];
    }
}

// c_i . v with the reference's (x + y) + z order and exact-zero terms skipped.
template <class T>
__device__ inline T cdot(int i, T vx, T vy, T vz) {
    const int cx = D3Q27::cx[i], cy = D3Q27::cy[i], cz = D3Q27::cz[i];
    T s = T(0);
    bool first = true;
    if (cx != 0) { s = cx > 0 ? vx : -vx; first = false; }
    if (cy != 0) { s = first ? (cy > 0 ? vy : -vy) : (cy > 0 ? s + vy : s - vy); first = false; }
    if (cz != 0) { s = first ? (cz > 0 ? vz : -vz) : (cz > 0 ? s + vz : s - vz); first = false; }
    return s;
}

}  // namespace rbtp
