// diag.hip — FP64 diagnostics on the device (output/output_statis.cu).
//
// Plane and wall sums are formed on the GPU with a fixed reduction shape (256 lane partials,
// pairwise tree), so results are reproducible run to run; the host only combines the LX plane
// values in the reference's order.  Nothing but a few kilobytes crosses to the host, instead of
// the reference's 2 x 8N + 4N bytes per call.
#include <cmath>
#include <cstring>
#include <vector>

#include "particles.h"

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

__global__ __launch_bounds__(DB) void k_ke(const double *ux, const double *uy, const double *uz,
                                           Dims d, double *out) {
    __shared__ double red[1][DB];
    const int t = threadIdx.x;
    double s = 0.0;
    for (int idx = blockIdx.x * DB + t; idx < d.n; idx += gridDim.x * DB)
        s += 0.5 * (ux[idx] * ux[idx] + uy[idx] * uy[idx] + uz[idx] * uz[idx]);
    red[0][t] = s;
    block_tree<1>(red, t);
    if (t == 0) out[blockIdx.x] = red[0][0];
}

constexpr int RED_BLOCKS = 1024;

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

int diag_kinetic_energy(const Dims &dims, double *const *macro, hipStream_t st, double *out) {
    DevBuf buf;
    DHIP(hipMalloc(&buf.p, (size_t)RED_BLOCKS * sizeof(double)));
    hipLaunchKernelGGL(k_ke, RED_BLOCKS, DB, 0, st, macro[1], macro[2], macro[3], dims, buf.p);
    DHIP(hipGetLastError());
    std::vector<double> h(RED_BLOCKS);
    DHIP(hipMemcpyAsync(h.data(), buf.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    DHIP(hipStreamSynchronize(st));
    double s = 0.0;
    for (int b = 0; b < RED_BLOCKS; b++) s += h[b];
    *out = s / (double)dims.n;
    return RBTP_OK;
}

}  // namespace rbtp
