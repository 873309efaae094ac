// diag.hip — FP64 diagnostics on the device (output/output_statis.cu).
//
// Plane and wall sums are formed on the GPU with a fixed reduction shape (256 lane partials,
// pairwise tree), so results are reproducible run to run; the host only combines the LX plane
// values in the reference's order.  Nothing but a few kilobytes crosses to the host, instead of
// the reference's 2 x 8N + 4N bytes per call.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "particles.h"
#include "lattice_device.h"

namespace rbtp {

namespace {

constexpr int DB = 256;

template <int C>
__device__ inline void block_tree(double (&red)[C][DB], int t) {
    __syncthreads();
    for (int stride = DB / 2; stride >= 1; stride >>= 1) {
        if (t < stride)
            for (int c = 0; c < C; c++) red[c][t] = red[c][t] + red[c][t + stride];
        __syncthreads();
    }
}

// Per x-plane sums of ux*phi over fluid and particle nodes (output_statis.cu:138-156).
__global__ __launch_bounds__(DB) void k_nu_planes(const double *ux, const double *phi,
                                                  const int32_t *owner, Dims d, double *out) {
    __shared__ double red[4][DB];
    const int ix = blockIdx.x, t = threadIdx.x;
    double sf = 0.0, sp = 0.0, cf = 0.0, cp = 0.0;
    const int m = d.ly * d.lz;
    for (int j = t; j < m; j += DB) {
        const int iz = j / d.ly, iy = j - iz * d.ly;
        const int idx = iz * d.lxy + iy * d.lx + ix;
        const double v = ux[idx] * phi[idx];
        const bool fluid = owner ? owner[idx] == -1 : true;
        if (fluid) { sf += v; cf += 1.0; }
        else { sp += v; cp += 1.0; }
    }
    red[0][t] = sf; red[1][t] = sp; red[2][t] = cf; red[3][t] = cp;
    block_tree<4>(red, t);
    if (t == 0)
        for (int c = 0; c < 4; c++) out[ix * 4 + c] = red[c][0];
}

// Hot (ix=0) and cold (ix=LX-1) wall sums (output_statis.cu:176-203).
__global__ __launch_bounds__(DB) void k_nu_walls(const double *phi, const int32_t *owner, Dims d,
                                                 double thot, double tcold, double *out) {
    __shared__ double red[2][DB];
    const int w = blockIdx.x, t = threadIdx.x;
    const int ix = w == 0 ? 0 : d.lx - 1;
    double s = 0.0, c = 0.0;
    const int m = d.ly * d.lz;
    for (int j = t; j < m; j += DB) {
        const int iz = j / d.ly, iy = j - iz * d.ly;
        const int idx = iz * d.lxy + iy * d.lx + ix;
        const bool fluid = owner ? owner[idx] == -1 : true;
        if (!fluid) continue;
        s += w == 0 ? (thot - phi[idx]) / 0.5 : (phi[idx] - tcold) / 0.5;
        c += 1.0;
    }
    red[0][t] = s; red[1][t] = c;
    block_tree<2>(red, t);
    if (t == 0) { out[w * 2] = red[0][0]; out[w * 2 + 1] = red[1][0]; }
}

// Means of u and phi with particle nodes overwritten by the particle's rigid-body velocity and
// temperature (diag_flow after overwrite_particle_nodes, output_statis.cu:5-60,82-96).
__global__ __launch_bounds__(DB) void k_means(const double *ux, const double *uy,
                                              const double *uz, const double *phi,
                                              const int32_t *owner, PartDev pd, Dims d,
                                              double *out) {
    __shared__ double red[4][DB];
    const int t = threadIdx.x;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    for (int idx = blockIdx.x * DB + t; idx < d.n; idx += gridDim.x * DB) {
        double u = ux[idx], v = uy[idx], w = uz[idx], T = phi[idx];
        const int pid = owner ? owner[idx] : -1;
        if (pid >= 0 && pid < pd.np) {
            const int iz = idx / d.lxy, rem = idx - iz * d.lxy, iy = rem / d.lx, ix = rem - iy * d.lx;
            const double relx = (static_cast<double>(ix) + 0.5) - pd.p[RBTP_P_POS_X][pid];
            const double rely = (static_cast<double>(iy) + 0.5) - pd.p[RBTP_P_POS_Y][pid];
            const double relz = (static_cast<double>(iz) + 0.5) - pd.p[RBTP_P_POS_Z][pid];
            const double wx = pd.p[RBTP_P_OMEGA_X][pid], wy = pd.p[RBTP_P_OMEGA_Y][pid], wz = pd.p[RBTP_P_OMEGA_Z][pid];
            u = pd.p[RBTP_P_VEL_X][pid] + (relz * wy - rely * wz);
            v = pd.p[RBTP_P_VEL_Y][pid] + (relx * wz - relz * wx);
            w = pd.p[RBTP_P_VEL_Z][pid] + (rely * wx - relx * wy);
            T = pd.p[RBTP_P_TEMP][pid];
        }
        s[0] += u; s[1] += v; s[2] += w; s[3] += T;
    }
    for (int c = 0; c < 4; c++) red[c][t] = s[c];
    block_tree<4>(red, t);
    if (t == 0)
        for (int c = 0; c < 4; c++) out[blockIdx.x * 4 + c] = red[c][0];
}

// Kinetic energy over the nodes that are fluid in both the current and the previous geometry
// (every node without particles): per-block partial sums of 0.5|u|^2 and of the node count.
// Nodes refilled this step are left out: refill_new_fluid_kernel sets their u to momentum over
// the density deviation sum(f) (refill.cu:224-233), which is not a velocity.
__global__ __launch_bounds__(DB) void k_ke(const double *ux, const double *uy, const double *uz,
                                           const int32_t *owner, const int32_t *owner_prev,
                                           Dims d, double *out) {
    __shared__ double red[2][DB];
    const int t = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int idx = blockIdx.x * DB + t; idx < d.n; idx += gridDim.x * DB) {
        if (owner && (owner[idx] != -1 || owner_prev[idx] != -1)) continue;
        s += 0.5 * (ux[idx] * ux[idx] + uy[idx] * uy[idx] + uz[idx] * uz[idx]);
        c += 1.0;
    }
    red[0][t] = s;
    red[1][t] = c;
    block_tree<2>(red, t);
    if (t == 0) {
        out[blockIdx.x * 2] = red[0][0];
        out[blockIdx.x * 2 + 1] = red[1][0];
    }
}


// Output fields of %09d.dat (output_flow.cu:64-110): rho, ux, uy, uz, phi from the last
// macrovar (and refill), ted and Ked from the post-collision populations (compute_ted,
// collision.cu:154-198; compute_Ked, collision.cu:91-152), then overwrite_particle_nodes
// (output_flow.cu:5-60) at solid nodes.  out is 7*n doubles in file order.
template <class S>
__global__ __launch_bounds__(DB) void k_output_fields(const S *fc, const S *gc,
                                                      const typename Enc<S>::A *fxf,
                                                      const double *rho, const double *ux,
                                                      const double *uy, const double *uz,
                                                      const double *phi, const int32_t *owner,
                                                      PartDev pd, OutK k, Dims d, double *out) {
    const int idx = blockIdx.x * DB + threadIdx.x;
    if (idx >= d.n) return;
    const size_t n = (size_t)d.n;
    const double fx = double(fxf[idx]), fy = 0.0, fz = 0.0;
    const double rho_local = rho[idx], ux_local = ux[idx], uy_local = uy[idx], uz_local = uz[idx];
    const double phi_local = phi[idx];
    const double RT = 1.0 / 3.0;
    const double u2 = ux_local * ux_local + uy_local * uy_local + uz_local * uz_local;
    const double uv = u2 / RT;
    // compute_Ked
    const double bu = fx * ux_local + fy * uy_local + fz * uz_local;
    double txx = 0.0, txy = 0.0, txz = 0.0, tyy = 0.0, tyz = 0.0, tzz = 0.0;
    // compute_ted
    const double lambda0 = 1.0 - weight(0);
    double gneqx = 0.0, gneqy = 0.0, gneqz = 0.0;
#pragma unroll
    for (int ip = 0; ip < NPOP; ip++) {
        const int cx = D3Q27::cx[ip], cy = D3Q27::cy[ip], cz = D3Q27::cz[ip];
        const double tp = weight(ip);
        const double eu = (cx * ux_local + cy * uy_local + cz * uz_local) / RT;
        const double feq = tp * (rho_local + eu + 0.5 * (eu * eu - uv));
        const double fneq = dec_f<S, double>(Enc<S>::ld(fc + pop_at(ip, idx)), ip) - feq;
        const double cidotF = cx * fx + cy * fy + cz * fz;
        const double feq_term = 0.5 * (cidotF - bu) * feq / RT;
        txx += cx * cx * (fneq + feq_term);
        txy += cx * cy * (fneq + feq_term);
        txz += cx * cz * (fneq + feq_term);
        tyy += cy * cy * (fneq + feq_term);
        tyz += cy * cz * (fneq + feq_term);
        tzz += cz * cz * (fneq + feq_term);
        const double geq = tp * phi_local * (1.0 + eu + 0.5 * (eu * eu - uv));
        const double geqnew = (ip == 0) ? (geq - lambda0 * phi_local * rho_local)
                                        : (geq + tp * phi_local * rho_local);
        const double diff = dec_g<S, double>(Enc<S>::ld(gc + pop_at(ip, idx)), ip,
                                             Enc<S>::scaled ? double(k.gshift[idx % d.lx]) : 0.0) -
                            geqnew;
        gneqx += cx * diff;
        gneqy += cy * diff;
        gneqz += cz * diff;
    }
    const double coeff = -(1.0 - 0.5 / k.tau);
    txx *= coeff; txy *= coeff; txz *= coeff; tyy *= coeff; tyz *= coeff; tzz *= coeff;
    const double Sxx = txx / (2.0 * k.visc), Sxy = txy / (2.0 * k.visc), Sxz = txz / (2.0 * k.visc);
    const double Syy = tyy / (2.0 * k.visc), Syz = tyz / (2.0 * k.visc), Szz = tzz / (2.0 * k.visc);
    double ked = 0.5 * k.visc * (4.0 * (Sxx * Sxx + Syy * Syy + Szz * Szz) +
                                 8.0 * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
    const double p = rho_local * RT;
    const double denominator = 2.0 * k.tauc * RT + p;
    const double tx = -(2.0 * gneqx + phi_local * fx) / denominator;
    const double ty = -(2.0 * gneqy + phi_local * fy) / denominator;
    const double tz = -(2.0 * gneqz + phi_local * fz) / denominator;
    double ted = (k.diff > 0.0) ? k.diff * (tx * tx + ty * ty + tz * tz) : 0.0;
    double r = rho_local, u = ux_local, v = uy_local, w = uz_local, T = phi_local;
    const int pid = owner ? owner[idx] : -1;
    if (pid >= 0 && pid < pd.np) {
        const int iz = idx / d.lxy, rem = idx - iz * d.lxy, iy = rem / d.lx, ix = rem - iy * d.lx;
        const double relx = (static_cast<double>(ix) + 0.5) - pd.p[RBTP_P_POS_X][pid];
        const double rely = (static_cast<double>(iy) + 0.5) - pd.p[RBTP_P_POS_Y][pid];
        const double relz = (static_cast<double>(iz) + 0.5) - pd.p[RBTP_P_POS_Z][pid];
        const double wx = pd.p[RBTP_P_OMEGA_X][pid], wy = pd.p[RBTP_P_OMEGA_Y][pid], wz = pd.p[RBTP_P_OMEGA_Z][pid];
        r = k.rho0;
        u = pd.p[RBTP_P_VEL_X][pid] + (relz * wy - rely * wz);
        v = pd.p[RBTP_P_VEL_Y][pid] + (relx * wz - relz * wx);
        w = pd.p[RBTP_P_VEL_Z][pid] + (rely * wx - relx * wy);
        T = pd.p[RBTP_P_TEMP][pid];
        ted = 0.0;
        ked = 0.0;
    }
    out[idx] = r;
    out[n + idx] = u;
    out[2 * n + idx] = v;
    out[3 * n + idx] = w;
    out[4 * n + idx] = T;
    out[5 * n + idx] = ted;
    out[6 * n + idx] = ked;
}

// output_profile (output_statis.cu:225-285): per x-plane sums of ux, uy, uz, phi after
// overwrite_particle_nodes, from the seven output fields.
__global__ __launch_bounds__(DB) void k_profile(const double *fields, Dims d, double *out) {
    __shared__ double red[4][DB];
    const int ix = blockIdx.x, t = threadIdx.x;
    const size_t n = (size_t)d.n;
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int m = d.ly * d.lz;
    for (int j = t; j < m; j += DB) {
        const int iz = j / d.ly, iy = j - iz * d.ly;
        const size_t idx = (size_t)iz * d.lxy + (size_t)iy * d.lx + ix;
        for (int c = 0; c < 4; c++) s[c] += fields[(size_t)(c + 1) * n + idx];
    }
    for (int c = 0; c < 4; c++) red[c][t] = s[c];
    block_tree<4>(red, t);
    if (t == 0)
        for (int c = 0; c < 4; c++) out[ix * 4 + c] = red[c][0];
}


// Streaming copy of 16-byte vectors: the reference kernel for the achievable HBM bandwidth and
// for calibrating the memory counters (it reads and writes exactly `count` vectors).  Each
// thread moves U vectors 256 apart, all loads issued before the stores; one workgroup per 256*U
// vectors.  The probe runs U = 1, 2 and 4, with ordinary and with nontemporal stores (as the
// lattice kernels use), and keeps the fastest: which is best differs between devices.
using dvec2 = double __attribute__((ext_vector_type(2)));
template <int U, bool NT>
__global__ __launch_bounds__(256) void k_copy_probe(const dvec2 *__restrict__ src,
                                                    dvec2 *__restrict__ dst, size_t count) {
    const size_t base = (size_t)blockIdx.x * (256 * U) + threadIdx.x;
    dvec2 v[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < count) v[u] = src[i];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < count) {
            if constexpr (NT) __builtin_nontemporal_store(v[u], dst + i);
            else dst[i] = v[u];
        }
    }
}

constexpr int RED_BLOCKS = 1024;

// ---- allocation-free Nu (rbtp_nu_noalloc) -------------------------------------------------------
//
// The plane sums (4 per x-plane), the two wall sums and the seven results of output_nu live in a
// file-scope device array instead of a per-call allocation.  Every call on one device shares it
// and it keeps the last call's contents; each call writes every entry it reads before reading
// it, so nothing needs clearing.  Calls that overlap on one device race on it.
constexpr int NU_NOALLOC_MAX_LX = 8192;
constexpr size_t NU_SCRATCH_DOUBLES = 4 * (size_t)NU_NOALLOC_MAX_LX + 4 + 7;
static __device__ __align__(256) double scratch[NU_SCRATCH_DOUBLES];

// The host-side combination of diag_nu (output_statis.cu:133-219) on the device, one thread,
// same operations in the same order.
__global__ void k_nu_combine(const double *sums, int lx, int ly, int lz, double volume_fraction,
                             double cp_particle_ratio, double diff, double thot, double tcold,
                             int step, double *out) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double fluid_fraction = fmax(0.0, 1.0 - volume_fraction);
    double plane_avg_uT_fluid = 0.0, plane_avg_uT_part = 0.0;
    for (int ix = 0; ix < lx; ++ix) {
        const double sf = sums[ix * 4], sp = sums[ix * 4 + 1];
        const double cf = sums[ix * 4 + 2], cp = sums[ix * 4 + 3];
        if (cf > 0) plane_avg_uT_fluid += sf / cf;
        if (cp > 0) plane_avg_uT_part += sp / cp;
    }
    plane_avg_uT_fluid /= static_cast<double>(lx);
    plane_avg_uT_part /= static_cast<double>(lx);
    const double conductive_flux = diff * (thot - tcold) / static_cast<double>(lx);
    double Nu_bulk_fluid = fluid_fraction, Nu_bulk_part = 0.0;
    if (conductive_flux != 0.0) {
        Nu_bulk_fluid += fluid_fraction * (plane_avg_uT_fluid / conductive_flux);
        Nu_bulk_part = volume_fraction * cp_particle_ratio * (plane_avg_uT_part / conductive_flux);
    }
    const double Nu_bulk = Nu_bulk_fluid + Nu_bulk_part;
    const double bottom = sums[4 * lx], bottom_count = sums[4 * lx + 1];
    const double top = sums[4 * lx + 2], top_count = sums[4 * lx + 3];
    const double section_area = static_cast<double>(ly) * static_cast<double>(lz);
    const double wall_scale = static_cast<double>(lx) / (thot - tcold);
    double Tbottom = 0.0, Ttop = 0.0;
    if (bottom_count > 0) Tbottom = (bottom / section_area) * wall_scale;
    if (top_count > 0) Ttop = (top / section_area) * wall_scale;
    out[0] = (double)step;
    out[1] = volume_fraction;
    out[2] = Nu_bulk;
    out[3] = Ttop;
    out[4] = Tbottom;
    out[5] = Nu_bulk_fluid;
    out[6] = Nu_bulk_part;
}

// Device address of `scratch` on each device, looked up once per device with that device
// current (the array's name is a host-side stand-in, not a device address).
constexpr int NU_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[NU_MAX_DEVICES] = {};

#define DHIP(call)                                         \
    do {                                                   \
        hipError_t e_ = (call);                            \
        if (e_ != hipSuccess) return report_hip(e_, #call); \
    } while (0)

struct DevBuf {
    double *p = nullptr;
    ~DevBuf() { (void)hipFree(p); }
};

}  // namespace

int diag_nu(const rbtp_case &c, const rbtp_derived &d, const Dims &dims, const double *ux,
            const double *phi, const int32_t *owner, int step, hipStream_t st, double out[7]) {
    const int LX = dims.lx;
    DevBuf buf;
    DHIP(hipMalloc(&buf.p, (size_t)(4 * LX + 4) * sizeof(double)));
    hipLaunchKernelGGL(k_nu_planes, LX, DB, 0, st, ux, phi, owner, dims, buf.p);
    DHIP(hipGetLastError());
    hipLaunchKernelGGL(k_nu_walls, 2, DB, 0, st, phi, owner, dims, c.t_hot, c.t_cold, buf.p + 4 * LX);
    DHIP(hipGetLastError());
    std::vector<double> h(4 * LX + 4);
    DHIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    // output_statis.cu:133-219
    const double volume_fraction = c.particle_fraction;
    const double fluid_fraction = std::fmax(0.0, 1.0 - volume_fraction);
    double plane_avg_uT_fluid = 0.0, plane_avg_uT_part = 0.0;
    for (int ix = 0; ix < LX; ++ix) {
        const double sf = h[ix * 4], sp = h[ix * 4 + 1];
        const double cf = h[ix * 4 + 2], cp = h[ix * 4 + 3];
        if (cf > 0) plane_avg_uT_fluid += sf / cf;
        if (cp > 0) plane_avg_uT_part += sp / cp;
    }
    plane_avg_uT_fluid /= static_cast<double>(LX);
    plane_avg_uT_part /= static_cast<double>(LX);
    const double conductive_flux = d.diff * (c.t_hot - c.t_cold) / static_cast<double>(LX);
    double Nu_bulk_fluid = fluid_fraction, Nu_bulk_part = 0.0;
    if (conductive_flux != 0.0) {
        Nu_bulk_fluid += fluid_fraction * (plane_avg_uT_fluid / conductive_flux);
        Nu_bulk_part = volume_fraction * c.cp_particle_ratio * (plane_avg_uT_part / conductive_flux);
    }
    const double Nu_bulk = Nu_bulk_fluid + Nu_bulk_part;
    const double bottom = h[4 * LX], bottom_count = h[4 * LX + 1];
    const double top = h[4 * LX + 2], top_count = h[4 * LX + 3];
    const double section_area = static_cast<double>(dims.ly) * static_cast<double>(dims.lz);
    const double wall_scale = static_cast<double>(LX) / (c.t_hot - c.t_cold);
    double Tbottom = 0.0, Ttop = 0.0;
    if (bottom_count > 0) Tbottom = (bottom / section_area) * wall_scale;
    if (top_count > 0) Ttop = (top / section_area) * wall_scale;
    out[0] = (double)step;
    out[1] = volume_fraction;
    out[2] = Nu_bulk;
    out[3] = Ttop;
    out[4] = Tbottom;
    out[5] = Nu_bulk_fluid;
    out[6] = Nu_bulk_part;
    return RBTP_OK;
}

// rbtp_nu without any allocation: the kernels of diag_nu on the device's scratch array, the
// combination on the device, and the seven results copied into the caller's out.
int diag_nu_noalloc(const rbtp_case &c, const rbtp_derived &d, const Dims &dims, const double *ux,
                    const double *phi, const int32_t *owner, int step, hipStream_t st,
                    double out[7]) {
    const int LX = dims.lx;
    if (LX > NU_NOALLOC_MAX_LX) {
        char msg[160];
        snprintf(msg, sizeof(msg),
                 "rbtp_nu_noalloc: lx = %d exceeds the scratch limit of %d x-planes", LX,
                 NU_NOALLOC_MAX_LX);
        return report(RBTP_E_CAPACITY, msg);
    }
    int dev = 0;
    DHIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NU_MAX_DEVICES) return report(RBTP_E_ARG, "rbtp_nu_noalloc: device id out of range");
    double *buf = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (!g_scratch_dev[dev]) {
            void *p = nullptr;
            DHIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
            g_scratch_dev[dev] = (double *)p;
        }
        buf = g_scratch_dev[dev];
    }
    hipLaunchKernelGGL(k_nu_planes, LX, DB, 0, st, ux, phi, owner, dims, buf);
    DHIP(hipGetLastError());
    hipLaunchKernelGGL(k_nu_walls, 2, DB, 0, st, phi, owner, dims, c.t_hot, c.t_cold, buf + 4 * LX);
    DHIP(hipGetLastError());
    double *res = buf + 4 * (size_t)LX + 4;
    hipLaunchKernelGGL(k_nu_combine, 1, 1, 0, st, (const double *)buf, LX, dims.ly, dims.lz,
                       c.particle_fraction, c.cp_particle_ratio, d.diff, c.t_hot, c.t_cold, step,
                       res);
    DHIP(hipGetLastError());
    DHIP(hipMemcpyAsync(out, res, 7 * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    return RBTP_OK;
}

int diag_means(const Dims &dims, double *const *macro, const PartDev *pd, hipStream_t st,
               double out[4]) {
    DevBuf buf;
    DHIP(hipMalloc(&buf.p, (size_t)RED_BLOCKS * 4 * sizeof(double)));
    PartDev empty;
    std::memset(&empty, 0, sizeof(empty));
    const PartDev &p = pd ? *pd : empty;
    const int32_t *owner = pd ? pd->owner[pd->geo_cur] : nullptr;
    hipLaunchKernelGGL(k_means, RED_BLOCKS, DB, 0, st, macro[1], macro[2], macro[3], macro[4],
                       owner, p, dims, buf.p);
    DHIP(hipGetLastError());
    std::vector<double> h((size_t)RED_BLOCKS * 4);
    DHIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    double s[4] = {0, 0, 0, 0};
    for (int b = 0; b < RED_BLOCKS; b++)
        for (int c = 0; c < 4; c++) s[c] += h[b * 4 + c];
    for (int c = 0; c < 4; c++) out[c] = s[c] / (double)dims.n;
    return RBTP_OK;
}

int diag_kinetic_energy(const Dims &dims, double *const *macro, const int32_t *owner,
                        const int32_t *owner_prev, hipStream_t st, double *out) {
    DevBuf buf;
    DHIP(hipMalloc(&buf.p, (size_t)2 * RED_BLOCKS * sizeof(double)));
    hipLaunchKernelGGL(k_ke, RED_BLOCKS, DB, 0, st, macro[1], macro[2], macro[3], owner,
                       owner_prev, dims, buf.p);
    DHIP(hipGetLastError());
    std::vector<double> h(2 * RED_BLOCKS);
    DHIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    double s = 0.0, c = 0.0;
    for (int b = 0; b < RED_BLOCKS; b++) {
        s += h[2 * b];
        c += h[2 * b + 1];
    }
    *out = c > 0.0 ? s / c : 0.0;
    return RBTP_OK;
}


int diag_output_fields(int prec, const void *fc, const void *gc, const void *fx, double *const *macro,
                       const PartDev *pd, const OutK &k, const Dims &dims, hipStream_t st,
                       double *dev_out) {
    PartDev empty;
    std::memset(&empty, 0, sizeof(empty));
    const PartDev &p = pd ? *pd : empty;
    const int32_t *owner = pd ? pd->owner[pd->geo_cur] : nullptr;
    const int nb = (dims.n + DB - 1) / DB;
    hipError_t e;
    switch (prec) {
#define OUTF(S)                                                                                   \
    hipLaunchKernelGGL((k_output_fields<S>), nb, DB, 0, st, (const S *)fc, (const S *)gc,         \
                       (const typename Enc<S>::A *)fx, macro[0], macro[1], macro[2], macro[3],   \
                       macro[4], owner, p, k, dims, dev_out);                                     \
    e = hipGetLastError();                                                                        \
    break;
        case RBTP_PREC_FP64: OUTF(double)
        case RBTP_PREC_FP32: OUTF(float)
        case RBTP_PREC_FP16S: OUTF(__half)
        default: OUTF(__hip_bfloat16)
#undef OUTF
    }
    if (e != hipSuccess) return report_hip(e, "output fields");
    return RBTP_OK;
}

int diag_profile(const double *dev_fields, const Dims &dims, hipStream_t st, double *out) {
    DevBuf buf;
    DHIP(hipMalloc(&buf.p, (size_t)4 * dims.lx * sizeof(double)));
    hipLaunchKernelGGL(k_profile, dims.lx, DB, 0, st, dev_fields, dims, buf.p);
    DHIP(hipGetLastError());
    std::vector<double> h((size_t)4 * dims.lx);
    DHIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    const double count = (double)dims.ly * (double)dims.lz;
    const double inv_count = (count > 0) ? (1.0 / count) : 0.0;
    for (size_t i = 0; i < h.size(); i++) out[i] = h[i] * inv_count;
    return RBTP_OK;
}


int diag_copy_probe(size_t bytes, int reps, double *gbs) {
    const size_t count = bytes / sizeof(dvec2);
    if (count == 0 || reps < 1) return report(RBTP_E_ARG, "copy probe: bytes >= 16 and reps >= 1");
    if ((count + 255) / 256 > 0x7fffffffu) return report(RBTP_E_ARG, "copy probe: too many bytes");
    dvec2 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, count * sizeof(dvec2));
    if (e == hipSuccess) e = hipMalloc(&b, count * sizeof(dvec2));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, count * sizeof(dvec2), st);
    double best = 0.0;
    auto run = [&](auto kernel, int unroll) {
        const unsigned grid = (unsigned)((count + 256 * (size_t)unroll - 1) / (256 * (size_t)unroll));
        if (e != hipSuccess) return;
        hipLaunchKernelGGL(kernel, grid, 256, 0, st, a, b, count);  // warm-up
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(e0, st);
        for (int r = 0; r < reps && e == hipSuccess; r++) {
            hipLaunchKernelGGL(kernel, grid, 256, 0, st, a, b, count);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.f)
            best = std::max(best, 2.0 * (double)(count * sizeof(dvec2)) * reps / (ms * 1e-3) / 1e9);
    };
    run(k_copy_probe<1, false>, 1);
    run(k_copy_probe<2, false>, 2);
    run(k_copy_probe<4, false>, 4);
    run(k_copy_probe<1, true>, 1);
    run(k_copy_probe<2, true>, 2);
    run(k_copy_probe<4, true>, 4);
    if (e == hipSuccess) *gbs = best;
    (void)hipFree(a);
    (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) return report_hip(e, "copy probe");
    return RBTP_OK;
}

}  // namespace rbtp
