// Distance-grid baker for gfx950: the W17 body and part grids that the voxelizer's
// kind 5 samples, computed from the car's triangles.
//
// Reference behaviour: bakeGrid of scripts/bake-car.ts:237-316 (closestOnTriangle
// :201-224, toHalf :226-235).  Four phases, each a data-parallel kernel:
//   k_bake_tris   one block per triangle: every grid point of the triangle's box grown
//                 by `band` gets its FP64 point-triangle distance, folded into an f32
//                 grid with an atomic minimum on the float's bits (non-negative floats
//                 order as unsigned integers, and rounding to f32 is monotonic, so the
//                 result is f32 of the minimum whatever the order of the triangles);
//   k_bake_flood  the closing flood: 6-connected, from every face but z = 0, through
//                 cells at distance >= closing; each thread sweeps one x row forward and
//                 back, passes repeat until one changes nothing (the flooded set is the
//                 connected component, so the order of updates does not matter);
//   k_bake_grow   the regrowth by ceil(closing / h) cells, one 26-neighbour level per
//                 launch (a cell joins at level s when a neighbour is at s - 1);
//   k_bake_out    the sign and shell, and the truncating f32 -> f16 encode.
// FP64 arithmetic in the reference's order (no FMA contraction), so the grids are
// bit-identical to the baker's own output files.
#include <string>

#include "aero_internal.h"

namespace aero {

struct BakeGeom {
    double ox, oy, oz, h;
    uint32_t nx, ny, nz, pad;
};

__device__ __forceinline__ double bk_sq(double x, double y, double z) { return __builtin_sqrt((x * x + y * y) + z * z); }

// closestOnTriangle (bake-car.ts:201-224)
__device__ double closest_on_triangle(double px, double py, double pz, const double* t) {
    const double ax = t[0], ay = t[1], az = t[2], bx = t[3], by = t[4], bz = t[5], cx = t[6], cy = t[7], cz = t[8];
    const double abx = bx - ax, aby = by - ay, abz = bz - az, acx = cx - ax, acy = cy - ay, acz = cz - az;
    const double apx = px - ax, apy = py - ay, apz = pz - az;
    const double d1 = (abx * apx + aby * apy) + abz * apz, d2 = (acx * apx + acy * apy) + acz * apz;
    if (d1 <= 0 && d2 <= 0) return bk_sq(apx, apy, apz);
    const double bpx = px - bx, bpy = py - by, bpz = pz - bz;
    const double d3 = (abx * bpx + aby * bpy) + abz * bpz, d4 = (acx * bpx + acy * bpy) + acz * bpz;
    if (d3 >= 0 && d4 <= d3) return bk_sq(bpx, bpy, bpz);
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0 && d1 >= 0 && d3 <= 0) {
        const double v = d1 / (d1 - d3);
        return bk_sq(apx - v * abx, apy - v * aby, apz - v * abz);
    }
    const double cpx = px - cx, cpy = py - cy, cpz = pz - cz;
    const double d5 = (abx * cpx + aby * cpy) + abz * cpz, d6 = (acx * cpx + acy * cpy) + acz * cpz;
    if (d6 >= 0 && d5 <= d6) return bk_sq(cpx, cpy, cpz);
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0 && d2 >= 0 && d6 <= 0) {
        const double w = d2 / (d2 - d6);
        return bk_sq(apx - w * acx, apy - w * acy, apz - w * acz);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        return bk_sq(px - (bx + w * (cx - bx)), py - (by + w * (cy - by)), pz - (bz + w * (cz - bz)));
    }
    const double denom = 1 / ((va + vb) + vc);
    const double v = vb * denom, w = vc * denom;
    return bk_sq((apx - abx * v) - acx * w, (apy - aby * v) - acy * w, (apz - abz * v) - acz * w);
}

__global__ void __launch_bounds__(256) k_bake_fill(uint32_t* dist, uint64_t n, uint32_t bits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dist[i] = bits;
}

__global__ void __launch_bounds__(256) k_bake_tris(BakeGeom g, const double* tris, uint32_t ntri, double band,
                                                   uint32_t* dist) {
    for (uint32_t f = blockIdx.x; f < ntri; f += gridDim.x) {
        double t[9];
#pragma unroll
        for (int k = 0; k < 9; k++) t[k] = tris[9 * (size_t)f + k];
        const double zmax = t[2] > t[5] ? (t[2] > t[8] ? t[2] : t[8]) : (t[5] > t[8] ? t[5] : t[8]);
        if (zmax < -band) continue;
        const double org[3] = {g.ox, g.oy, g.oz};
        const uint32_t dims[3] = {g.nx, g.ny, g.nz};
        int64_t lo[3], hi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double v0 = t[a], v1 = t[a + 3], v2 = t[a + 6];
            const double mn = v0 < v1 ? (v0 < v2 ? v0 : v2) : (v1 < v2 ? v1 : v2);
            const double mx = v0 > v1 ? (v0 > v2 ? v0 : v2) : (v1 > v2 ? v1 : v2);
            const double l = __builtin_floor((mn - band - org[a]) / g.h), u = __builtin_ceil((mx + band - org[a]) / g.h);
            lo[a] = l > 0 ? (int64_t)l : 0;
            hi[a] = u < (double)dims[a] - 1 ? (int64_t)u : (int64_t)dims[a] - 1;
        }
        if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) continue;
        const uint64_t bx = (uint64_t)(hi[0] - lo[0] + 1), by = (uint64_t)(hi[1] - lo[1] + 1),
                       bz = (uint64_t)(hi[2] - lo[2] + 1);
        const uint64_t nbox = bx * by * bz;
        for (uint64_t c = threadIdx.x; c < nbox; c += blockDim.x) {
            const int64_t i = lo[0] + (int64_t)(c % bx), j = lo[1] + (int64_t)((c / bx) % by),
                          k = lo[2] + (int64_t)(c / (bx * by));
            const double d = closest_on_triangle(g.ox + (double)i * g.h, g.oy + (double)j * g.h,
                                                 g.oz + (double)k * g.h, t);
            const float df = (float)d;
            uint32_t* slot = dist + (((uint64_t)k * g.ny + (uint64_t)j) * g.nx + (uint64_t)i);
            // Values only decrease: a stale read is an upper bound, so d >= it needs no
            // atomic.  The minimum itself is a system-scope atomic, performed at the
            // memory side and so coherent across the XCDs' separate L2 caches.
            if (d < (double)__builtin_bit_cast(float, __atomic_load_n(slot, __ATOMIC_RELAXED)))
                (void)__hip_atomic_fetch_min(slot, __builtin_bit_cast(uint32_t, df), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
}

__device__ __forceinline__ bool bk_open(const uint32_t* dist, uint64_t idx, double closing) {
    return !((double)__builtin_bit_cast(float, dist[idx]) < closing);
}

// Seeds: the open cells of every face but z = 0 (bake-car.ts:271-274).
__global__ void __launch_bounds__(256) k_bake_seed(BakeGeom g, const uint32_t* dist, double closing, uint8_t* reached) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
    if (idx >= n) return;
    const uint32_t i = (uint32_t)(idx % g.nx), j = (uint32_t)((idx / g.nx) % g.ny), k = (uint32_t)(idx / ((uint64_t)g.nx * g.ny));
    const bool face = i == 0 || i == g.nx - 1 || j == 0 || j == g.ny - 1 || k == g.nz - 1;
    reached[idx] = (face && bk_open(dist, idx, closing)) ? 1 : 0;
}

// One flood pass: thread (j, k) sweeps its x row forward then back.  Neighbour rows
// are read while other threads write them; writes only ever set 0 -> 1, and passes
// repeat until one sets nothing, so the racing reads only delay propagation.
__global__ void __launch_bounds__(256) k_bake_flood(BakeGeom g, const uint32_t* dist, double closing, uint8_t* reached,
                                                    uint32_t* changed) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= (uint64_t)g.ny * g.nz) return;
    const uint32_t j = (uint32_t)(r % g.ny), k = (uint32_t)(r / g.ny);
    const uint64_t row = (uint64_t)g.nx * r, sy = g.nx, sz = (uint64_t)g.nx * g.ny;
    volatile uint8_t* R = reached;
    bool any = false;
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t s = 0; s < g.nx; s++) {
            const uint32_t i = pass == 0 ? s : g.nx - 1 - s;
            const uint64_t idx = row + i;
            if (R[idx] || !bk_open(dist, idx, closing)) continue;
            const bool nb = (i > 0 && R[idx - 1]) || (i < g.nx - 1 && R[idx + 1]) || (j > 0 && R[idx - sy]) ||
                            (j < g.ny - 1 && R[idx + sy]) || (k > 0 && R[idx - sz]) || (k < g.nz - 1 && R[idx + sz]);
            if (nb) {
                R[idx] = 1;
                any = true;
            }
        }
    }
    if (any) *changed = 1u;
}

__global__ void __launch_bounds__(256) k_bake_grow_init(const uint8_t* reached, int16_t* grow, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) grow[i] = reached[i] ? (int16_t)0 : (int16_t)-1;
}

// Level s of the regrowth (bake-car.ts:291-306).
__global__ void __launch_bounds__(256) k_bake_grow(BakeGeom g, int16_t* grow, int16_t s) {
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t n = (uint64_t)g.nx * g.ny * g.nz;
    if (idx >= n || grow[idx] >= 0) return;
    const int64_t i = (int64_t)(idx % g.nx), j = (int64_t)((idx / g.nx) % g.ny), k = (int64_t)(idx / ((uint64_t)g.nx * g.ny));
    for (int dk = -1; dk <= 1; dk++)
        for (int dj = -1; dj <= 1; dj++)
            for (int di = -1; di <= 1; di++) {
                const int64_t a = i + di, b = j + dj, c = k + dk;
                if (a < 0 || b < 0 || c < 0 || a >= (int64_t)g.nx || b >= (int64_t)g.ny || c >= (int64_t)g.nz) continue;
                if (grow[((uint64_t)c * g.ny + (uint64_t)b) * g.nx + (uint64_t)a] == s - 1) {
                    grow[idx] = s;
                    return;
                }
            }
}

// toHalf (bake-car.ts:226-235) of f32(v): truncated mantissa, underflow to signed zero.
__device__ __forceinline__ uint16_t bk_to_half(double v) {
    const uint32_t x = __builtin_bit_cast(uint32_t, (float)v);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const int e = (int)((x >> 23) & 0xffu) - 127 + 15;
    const uint32_t mant = x & 0x7fffffu;
    if (e <= 0) return (uint16_t)sign;
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    return (uint16_t)(sign | ((uint32_t)e << 10) | (mant >> 13));
}

__global__ void __launch_bounds__(256) k_bake_out(const uint32_t* dist, const int16_t* grow, uint64_t n, double shell,
                                                  uint16_t* out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = (double)__builtin_bit_cast(float, dist[i]);
    out[i] = bk_to_half(grow[i] >= 0 ? d - shell : -(d > 0.001 ? d : 0.001) - shell);
}

}  // namespace aero

using namespace aero;

extern "C" int aero_bake_sdf(const double* tris, uint32_t ntri, const aero_sdf_grid* grid, double band, double closing,
                             double shell, uint16_t* out) {
    auto fail_msg = [](int code, const std::string& m) {
        report_global_error(m);
        return code;
    };
    if (!grid || !out || (ntri && !tris)) return fail_msg(AERO_E_INVALID, "aero_bake_sdf: null argument");
    const uint64_t n = (uint64_t)grid->dims[0] * grid->dims[1] * grid->dims[2];
    if (n == 0 || n > (1ull << 31)) return fail_msg(AERO_E_INVALID, "aero_bake_sdf: grid must hold 1 .. 2^31 cells");
    if (!(grid->h > 0.0) || !(band > 0.0) || !(closing >= 0.0))
        return fail_msg(AERO_E_INVALID, "aero_bake_sdf: needs h > 0, band > 0, closing >= 0");
    const double steps = __builtin_ceil(closing / grid->h);
    if (steps > 32767.0) return fail_msg(AERO_E_INVALID, "aero_bake_sdf: closing / h above 32767");
    BakeGeom g{grid->origin[0], grid->origin[1], grid->origin[2], grid->h, grid->dims[0], grid->dims[1], grid->dims[2], 0};
    hipStream_t st = nullptr;
    double* d_tris = nullptr;
    uint32_t *d_dist = nullptr, *d_changed = nullptr;
    uint8_t* d_reached = nullptr;
    int16_t* d_grow = nullptr;
    uint16_t* d_out = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess && ntri) e = hipMalloc(&d_tris, (size_t)ntri * 9 * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_dist, n * 4);
    if (e == hipSuccess) e = hipMalloc(&d_reached, n);
    if (e == hipSuccess) e = hipMalloc(&d_grow, n * 2);
    if (e == hipSuccess) e = hipMalloc(&d_out, n * 2);
    if (e == hipSuccess) e = hipMalloc(&d_changed, 4);
    if (e == hipSuccess && ntri)
        e = hipMemcpyAsync(d_tris, tris, (size_t)ntri * 9 * sizeof(double), hipMemcpyHostToDevice, st);
    const dim3 cells((unsigned)((n + 255) / 256));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bake_fill, cells, dim3(256), 0, st, d_dist, n, __builtin_bit_cast(uint32_t, (float)band));
        e = hipGetLastError();
    }
    if (e == hipSuccess && ntri) {
        const unsigned blocks = ntri < (1u << 20) ? ntri : (1u << 20);
        hipLaunchKernelGGL(k_bake_tris, dim3(blocks), dim3(256), 0, st, g, d_tris, ntri, band, d_dist);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bake_seed, cells, dim3(256), 0, st, g, d_dist, closing, d_reached);
        e = hipGetLastError();
    }
    // flood passes until one changes nothing; each pass that continues sets at least one
    // cell, so there are at most n + 1 of them
    const uint64_t rows = (uint64_t)g.ny * g.nz;
    for (uint64_t pass = 0; e == hipSuccess && pass <= n; pass++) {
        uint32_t changed = 0;
        e = hipMemsetAsync(d_changed, 0, 4, st);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_bake_flood, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st, g, d_dist, closing,
                               d_reached, d_changed);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(&changed, d_changed, 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess || !changed) break;
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bake_grow_init, cells, dim3(256), 0, st, d_reached, d_grow, n);
        e = hipGetLastError();
    }
    for (int s = 1; e == hipSuccess && s <= (int)steps; s++) {
        hipLaunchKernelGGL(k_bake_grow, cells, dim3(256), 0, st, g, d_grow, (int16_t)s);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_bake_out, cells, dim3(256), 0, st, d_dist, d_grow, n, shell, d_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * 2, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (void* p : {(void*)d_tris, (void*)d_dist, (void*)d_reached, (void*)d_grow, (void*)d_out, (void*)d_changed})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess)
        return fail_msg(e == hipErrorOutOfMemory ? AERO_E_NOMEM : AERO_E_HIP,
                        std::string("aero_bake_sdf: ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
    return AERO_OK;
}
