// lattice.h — host launchers of lattice.hip.
#pragma once

#include "rbtp_internal.h"

namespace rbtp {

template <class S>
hipError_t launch_lattice_step(const StepArgs<S> &a, bool strict, bool nostream, hipStream_t st);
// Batched step over several lattices; tile_start has nslots entries (first tile of each slot,
// tiles of `block` nodes).  block is 64, 128 or 256; grid is the number of persistent
// workgroups.
template <class S>
hipError_t launch_lattice_batched(const StepArgs<S> *slots_dev, const int *tile_start_dev,
                                  int nslots, int total_tiles, int block, int grid, bool strict,
                                  hipStream_t st);
template <class S>
hipError_t launch_stream_only(const S *fin, const S *gin, const typename Enc<S>::A *rho_wall,
                              double *fout, double *gout, Dims d, const float *gshift,
                              double thot, double tcold, bool nostream, hipStream_t st);
template <class S>
hipError_t launch_decode(const S *in, double *out, size_t n, int is_g, const float *gshift,
                         int lx, hipStream_t st);
template <class S>
hipError_t launch_encode(const double *in, S *out, size_t n, int is_g, const float *gshift,
                         int lx, hipStream_t st);
template <class S>
hipError_t launch_init(S *f, S *g, typename Enc<S>::A *fx, typename Enc<S>::A *rho_wall,
                       const double *temperature, double rho0, double t0, double grav0,
                       double beta, const float *gshift, Dims d, double *mr, double *mux,
                       double *muy, double *muz, double *mphi, hipStream_t st);

// Arrays of the per-stage execution (rbtp_stage): FP64, populations in the blocked layout of
// pop_at.  f/g are d_f/d_g (post-stream, bounced back, refilled), fc/gc d_f_collide/d_g_collide.
struct StageFluid {
    double *f, *fc, *g, *gc;
    double *rho, *ux, *uy, *uz, *phi;
    double *fx;
    const double *rho_wall;  // wall rho of the state the staged mode starts from
    Dims d;
    FluidK k;
};
// One fluid stage (RBTP_STAGE_MACROVAR ... RBTP_STAGE_STREAM_G, not the IBB stages).
hipError_t launch_stage_fluid(int stage, const StageFluid &a, hipStream_t st);
// Start of the staged mode from the fused state: fc/gc hold the post-collision populations of
// the last step (or, when poststream, the post-stream ones); f/g receive the post-stream
// populations (pull with the wall rules, or a copy).
hipError_t launch_stage_enter(const StageFluid &a, bool poststream, hipStream_t st);

// rbtp_arith_probe: op RBTP_ARITH_* over n host operand pairs, result into host out.
int lattice_arith_probe(int op, const double *a, const double *b, size_t n, double *out);

// Time the FP64 lattice step's access pattern alone on an lx x ly x lz lattice (rbtp_stencil_probe).
int lattice_stencil_probe(int lx, int ly, int lz, int reps, double *gbs);
// scheme 0: the pull (as lattice_stencil_probe); 1 and 2: the even and odd steps of in-place
// AA streaming.  GB/s at 880 B per node.
int lattice_streaming_probe(int scheme, int lx, int ly, int lz, int reps, double *gbs);

// Refresh the shift profile of the reduced-precision g encoding to the plane-mean temperature
// and re-encode the stored g (buffer g, the state the next step reads) under it.  partial holds
// lx*lz doubles, delta lx floats.  No-op for FP64 storage.
template <class S>
hipError_t launch_shift_refresh(S *g, Dims d, float *gshift, double *partial, float *delta,
                                hipStream_t st);

template <class S>
hipError_t launch_forcing_field(typename Enc<S>::A *fx, typename Enc<S>::A *rho_wall,
                                const double *rho, const double *phi, double gb, double t0,
                                Dims d, hipStream_t st);

}  // namespace rbtp
