// Semi-Lagrangian heat scalar for gfx950 (reference behaviour: HEAT,
// src/engine/wgsl.ts:460-511).  One thread per cell: back-trace along the cell's
// velocity, trilinear gather of the previous temperature (z < 0 mirrored, x < 0
// cold), 6-neighbour smoothing, source relaxation and cooling.  The source test
// reads neighbour emit values only where the flag byte marks a hot neighbour.
#include "aero_internal.h"

namespace aero {

struct HeatArgs {
    const float4* fields;
    const float* emit;
    const uint8_t* flags;
    const float* tin;
    float* tout;
    uint32_t nx, ny, nz;
    uint32_t pad;
    float dt, cool, diffuse;
};

// Cell index with the reference's addressing rules: z < 0 mirrors to -z-1, then every
// coordinate is clamped into the grid (wgsl.ts:486-492).  32-bit: grids stay below 2^31
// cells.
__device__ __forceinline__ uint32_t heat_idx(const HeatArgs& h, int x, int y, int z) {
    if (z < 0) z = -z - 1;
    x = x < 0 ? 0 : (x > (int)h.nx - 1 ? (int)h.nx - 1 : x);
    y = y < 0 ? 0 : (y > (int)h.ny - 1 ? (int)h.ny - 1 : y);
    z = z > (int)h.nz - 1 ? (int)h.nz - 1 : z;
    return (uint32_t)x + h.nx * ((uint32_t)y + h.ny * (uint32_t)z);
}
__device__ __forceinline__ float heat_at(const HeatArgs& h, int x, int y, int z) {
    return x < 0 ? 0.0f : h.tin[heat_idx(h, x, y, z)];
}

// Each thread handles the cells (x, y, z) and (x, y, z + 1): the loads of both cells go
// out before either waits, so every wavefront keeps twice the memory requests in flight
// through the two dependent rounds (velocity, then the back-traced gather).  The
// arithmetic of each cell is unchanged.
constexpr int kHeatCells = 2;

__global__ void __launch_bounds__(256) k_heat(HeatArgs h) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z0 = blockIdx.z * kHeatCells;
    if (x >= (int)h.nx || y >= (int)h.ny) return;
    const int off[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    bool live[kHeatCells];
    uint32_t idx[kHeatCells], fl[kHeatCells];
    float nbv[kHeatCells][6];
    float4 fv[kHeatCells];
    // Loads that do not depend on the velocity go out first, together with it: the six
    // neighbour temperatures and the flag byte, for both cells.
#pragma unroll
    for (int c = 0; c < kHeatCells; c++) {
        const int z = z0 + c;
        live[c] = z < (int)h.nz;
        const int zs = live[c] ? z : z0;  // a dead second cell repeats the first one's loads
        idx[c] = (uint32_t)x + h.nx * ((uint32_t)y + h.ny * (uint32_t)zs);
#pragma unroll
        for (int k = 0; k < 6; k++) nbv[c][k] = heat_at(h, x + off[k][0], y + off[k][1], zs + off[k][2]);
        fl[c] = h.flags[idx[c]];
        fv[c] = h.fields[idx[c]];
    }
    float fx[kHeatCells], fy[kHeatCells], fz[kHeatCells], tv[kHeatCells][8];
#pragma unroll
    for (int c = 0; c < kHeatCells; c++) {
        const int zs = live[c] ? z0 + c : z0;
        const float gx = (float)x - fv[c].x * h.dt, gy = (float)y - fv[c].y * h.dt, gz = (float)zs - fv[c].z * h.dt;
        const float flx = __builtin_floorf(gx), fly = __builtin_floorf(gy), flz = __builtin_floorf(gz);
        const int ix = (int)flx, iy = (int)fly, iz = (int)flz;
        fx[c] = gx - (float)ix;
        fy[c] = gy - (float)iy;
        fz[c] = gz - (float)iz;
#pragma unroll
        for (int k = 0; k < 8; k++) tv[c][k] = heat_at(h, ix + (k & 1), iy + ((k >> 1) & 1), iz + ((k >> 2) & 1));
    }
#pragma unroll
    for (int c = 0; c < kHeatCells; c++) {
        if (!live[c]) continue;
        const int z = z0 + c;
        if (fv[c].w <= 0.0f) {
            __builtin_nontemporal_store(h.emit[idx[c]], h.tout + idx[c]);
            continue;
        }
        float t = 0.0f;
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
            const float w = (ox ? fx[c] : 1.0f - fx[c]) * (oy ? fy[c] : 1.0f - fy[c]) * (oz ? fz[c] : 1.0f - fz[c]);
            t = t + tv[c][k] * w;
        }
        float nb = 0.0f;
#pragma unroll
        for (int k = 0; k < 6; k++) nb = nb + nbv[c][k];
        float src = 0.0f;
        if (fl[c] & FL_HOTNB) {
#pragma unroll
            for (int k = 0; k < 6; k++) {
                const int cx = x + off[k][0], cy = y + off[k][1], cz = z + off[k][2];
                if (cx >= 0 && cy >= 0 && cx < (int)h.nx && cy < (int)h.ny && cz < (int)h.nz) {
                    const float e = h.emit[heat_idx(h, cx, cy, cz)];
                    src = src > e ? src : e;
                }
            }
        }
        t = t * (1.0f - h.diffuse) + (nb / 6.0f) * h.diffuse;
        if (src > 0.0f) t = t * (1.0f - 0.35f) + src * 0.35f;
        const float o = t * (1.0f - h.cool);
        __builtin_nontemporal_store(o > 0.0f ? o : 0.0f, h.tout + idx[c]);
    }
}

hipError_t launch_heat(aero_ctx* c, const aero_heat* hp) {
    HeatArgs h{};
    h.fields = c->fields;
    h.emit = c->emit;
    h.flags = c->flags;
    h.tin = c->temps[c->tcur];
    h.tout = c->temps[c->tcur ^ 1];
    h.nx = c->grid.nx; h.ny = c->grid.ny; h.nz = c->grid.nz;
    h.dt = hp->dt; h.cool = hp->cool; h.diffuse = hp->diffuse;
    const dim3 grid((c->grid.nx + 63) / 64, (c->grid.ny + 3) / 4, (c->grid.nz + kHeatCells - 1) / kHeatCells);
    hipLaunchKernelGGL(k_heat, grid, dim3(256), 0, c->stream, h);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) c->tcur ^= 1;
    return e;
}

}  // namespace aero
