// Field consumers for gfx950: trilinear flow sampling, smoke/fog particle advection
// and streamline tracing over the `fields` (ux, uy, uz, rho) and `temps` buffers.
//
// Reference behaviour: sampleFlow / vorticityAt (src/engine/wgsl.ts:324-364), View and
// fieldValue (:430-457), PARTICLES_SIM (:517-595) and STREAM_SIM
// (src/engine/render-wgsl.ts:520-576).  One thread per point, particle or line, as
// there: every sample is 8 independent gathers of a 16-byte field cell and 8 of a
// temperature, issued together; the grids stay resident in HBM and the gathers of
// neighbouring particles share cache lines.  Vector expressions are evaluated
// component by component in the order the WGSL writes them (the library is built
// without FMA contraction), mix(a, b, t) = a*(1-t) + b*t and normalize(v) =
// v / length(v), so the results are bit-identical to the f32 CPU restatement
// (the sines and cosines of emitter-ring seeds aside).
#include "aero_internal.h"
#include "sdf.h"

namespace aero {

struct FlowArgs {
    const float4* fields;
    const float* temps;
    uint32_t nx, ny, nz;
    float ox, oy, oz, h;
};

struct FSample {
    float ux, uy, uz, rho, temp, fluid;
};

__device__ __forceinline__ float f_mix(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float f_len(float x, float y, float z) { return __builtin_sqrtf((x * x + y * y) + z * z); }
__device__ __forceinline__ float f_clamp(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }

// i32 of a floored value: saturating, NaN -> 0.
__device__ __forceinline__ int32_t f_to_i32(float v) {
    if (!(v == v)) return 0;
    if (v <= -2147483648.0f) return INT32_MIN;
    if (v >= 2147483520.0f) return 2147483520;
    return (int32_t)v;
}

// cellIdx (wgsl.ts:325-329): every coordinate clamped into the grid.
__device__ __forceinline__ uint32_t f_cell(const FlowArgs& a, int32_t x, int32_t y, int32_t z) {
    const int32_t dx = (int32_t)a.nx, dy = (int32_t)a.ny, dz = (int32_t)a.nz;
    x = x < 0 ? 0 : (x > dx - 1 ? dx - 1 : x);
    y = y < 0 ? 0 : (y > dy - 1 ? dy - 1 : y);
    z = z < 0 ? 0 : (z > dz - 1 ? dz - 1 : z);
    return (uint32_t)x + a.nx * ((uint32_t)y + a.ny * (uint32_t)z);
}

// sampleFlow (wgsl.ts:331-355): trilinear over the fluid corners of the cell-centred
// grid, renormalised by their weight; z < 0 samples the mirror image.
__device__ FSample sample_flow(const FlowArgs& a, float px, float py, float pz) {
    const float zs = pz < 0.0f ? -1.0f : 1.0f;
    const float qz = __builtin_fabsf(pz);
    const float gx = (px - a.ox) / a.h - 0.5f, gy = (py - a.oy) / a.h - 0.5f, gz = (qz - a.oz) / a.h - 0.5f;
    const int32_t ix = f_to_i32(__builtin_floorf(gx)), iy = f_to_i32(__builtin_floorf(gy)),
                  iz = f_to_i32(__builtin_floorf(gz));
    const float fx = gx - (float)ix, fy = gy - (float)iy, fz = gz - (float)iz;
    uint32_t id[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
        int32_t cz = iz + ((k >> 2) & 1);
        if (cz < 0) cz = -cz - 1;
        id[k] = f_cell(a, ix + (k & 1), iy + ((k >> 1) & 1), cz);
    }
    float4 v[8];
    float t[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = a.fields[id[k]];
#pragma unroll
    for (int k = 0; k < 8; k++) t[k] = a.temps[id[k]];
    float ax = 0.0f, ay = 0.0f, az = 0.0f, aw = 0.0f, tacc = 0.0f, wsum = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
        const float w = ((ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy)) * (oz ? fz : 1.0f - fz);
        if (v[k].w > 0.0f) {
            ax = ax + v[k].x * w;
            ay = ay + v[k].y * w;
            az = az + v[k].z * w;
            aw = aw + v[k].w * w;
            tacc = tacc + t[k] * w;
            wsum = wsum + w;
        }
    }
    FSample s;
    if (wsum < 1e-5f) {
        s.ux = s.uy = s.uz = 0.0f;
        s.rho = 1.0f;
        s.temp = 0.0f;
        s.fluid = 0.0f;
        return s;
    }
    s.ux = ax / wsum;
    s.uy = ay / wsum;
    s.uz = (az / wsum) * zs;
    s.rho = aw / wsum;
    s.temp = tacc / wsum;
    s.fluid = wsum;
    return s;
}

// vorticityAt (wgsl.ts:356-363): |curl u| / 2 by central differences over +-h.
__device__ float vorticity_at(const FlowArgs& a, float px, float py, float pz) {
    const float e = a.h;
    const FSample dx0 = sample_flow(a, px - e, py - 0.0f, pz - 0.0f);
    const FSample dx1 = sample_flow(a, px + e, py + 0.0f, pz + 0.0f);
    const FSample dy0 = sample_flow(a, px - 0.0f, py - e, pz - 0.0f);
    const FSample dy1 = sample_flow(a, px + 0.0f, py + e, pz + 0.0f);
    const FSample dz0 = sample_flow(a, px - 0.0f, py - 0.0f, pz - e);
    const FSample dz1 = sample_flow(a, px + 0.0f, py + 0.0f, pz + e);
    const float wx = ((dy1.uz - dy0.uz) - (dz1.uy - dz0.uy)) * 0.5f;
    const float wy = ((dz1.ux - dz0.ux) - (dx1.uz - dx0.uz)) * 0.5f;
    const float wz = ((dx1.uy - dx0.uy) - (dy1.ux - dy0.ux)) * 0.5f;
    return f_len(wx, wy, wz);
}

// rhoRef, pressureCoef, adiabaticDelta and fieldValue (wgsl.ts:435-457).
__device__ __forceinline__ float rho_ref(const FlowArgs& a) {
    const float r = a.fields[f_cell(a, 6, (int32_t)a.ny - 6, (int32_t)a.nz - 6)].w;
    return r > 0.5f ? r : 1.0f;
}
__device__ __forceinline__ float pressure_coef(const FlowArgs& a, const aero_view& v, const FSample& s) {
    return ((s.rho - rho_ref(a)) / 3.0f) / ((0.5f * v.u_in) * v.u_in);
}
__device__ float field_value(const FlowArgs& a, const aero_view& v, uint32_t mode, const FSample& s, float px,
                             float py, float pz) {
    switch (mode) {
    case 1u: return (pressure_coef(a, v, s) - v.cp_min) / (v.cp_max - v.cp_min);
    case 2u: {
        const float ad = (((pressure_coef(a, v, s) * 0.5f) * v.v_real) * v.v_real) / 1005.0f;
        const float t = s.temp + ad;
        return (t - v.t_min) / (v.t_max - v.t_min);
    }
    case 3u: return vorticity_at(a, px, py, pz) / (v.u_in * v.vort_max);
    case 4u: return f_clamp((1.0f - s.ux / v.u_in) / 0.9f, 0.0f, 1.0f);
    default: return f_len(s.ux, s.uy, s.uz) / (v.u_in * v.speed_max);
    }
}

// pcg and rand (wgsl.ts:537-544).
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    const uint32_t s = v * 747796405u + 2891336453u;
    const uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
__device__ __forceinline__ float rand01(uint32_t& seed) {
    seed = pcg(seed);
    return (float)seed / 4294967295.0f;
}

// spawn (wgsl.ts:547-567): rake (mode 1) or fog box; scatter spreads x over the tunnel.
__device__ void spawn(const FlowArgs& a, const aero_psim& ps, uint32_t i, uint32_t& seed, float4& pos, float4& vel) {
    float x, y, z;
    if (ps.spawn_mode == 1u) {
        const uint32_t lanes = ps.rake_rows * ps.rake_cols;
        const uint32_t lane = i % lanes;
        const uint32_t row = lane / ps.rake_cols, col = lane % ps.rake_cols;
        const float fy = ((float)row + 0.5f) / (float)ps.rake_rows;
        const float fz = ((float)col + 0.5f) / (float)ps.rake_cols;
        x = ps.spawn_min[0];
        y = f_mix(ps.spawn_min[1], ps.spawn_max[1], fy);
        z = f_mix(ps.spawn_min[2], ps.spawn_max[2], fz);
        const float r0 = rand01(seed);
        const float r1 = rand01(seed);
        const float r2 = rand01(seed);
        x = x + r0 * 0.12f;
        y = y + (r1 - 0.5f) * 0.012f;
        z = z + (r2 - 0.5f) * 0.012f;
    } else {
        const float r0 = rand01(seed);
        const float r1 = rand01(seed);
        const float r2 = rand01(seed);
        x = ps.spawn_min[0] + r0 * 0.15f;
        y = f_mix(ps.spawn_min[1], ps.spawn_max[1], r1);
        z = f_mix(ps.spawn_min[2], ps.spawn_max[2], r2);
    }
    if (ps.scatter != 0u) {
        const float dmx = a.ox + (float)a.nx * a.h;
        x = f_mix(ps.spawn_min[0], dmx - a.h * 2.0f, rand01(seed));
    }
    pos = make_float4(x, y, z, 0.0f);
    vel = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
}

// PARTICLES_SIM cs_main (wgsl.ts:569-594): midpoint advection over dt lattice steps,
// displacement capped at 2h, respawn when old, outside or in a solid.
__global__ void __launch_bounds__(256) k_particles(FlowArgs a, aero_psim ps, aero_view view, float4* parts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ps.count) return;
    uint32_t seed = pcg(i * 9781u + ps.frame * 6271u + 17u);
    float4 pos = parts[2 * (size_t)i], vel = parts[2 * (size_t)i + 1];
    const float max_life = ps.life * (0.6f + (0.4f * (float)(pcg(i * 31u) % 1000u)) / 1000.0f);
    const float hx = a.ox + (float)a.nx * a.h, hy = a.oy + (float)a.ny * a.h, hz = a.oz + (float)a.nz * a.h;
    if (ps.scatter != 0u || pos.w > max_life) spawn(a, ps, i, seed, pos, vel);
    const FSample s1 = sample_flow(a, pos.x, pos.y, pos.z);
    const float mx = pos.x + ((s1.ux * a.h) * ps.dt) * 0.5f;
    const float my = pos.y + ((s1.uy * a.h) * ps.dt) * 0.5f;
    const float mz = pos.z + ((s1.uz * a.h) * ps.dt) * 0.5f;
    const FSample s2 = sample_flow(a, mx, my, mz);
    float dx = (s2.ux * a.h) * ps.dt, dy = (s2.uy * a.h) * ps.dt, dz = (s2.uz * a.h) * ps.dt;
    const float dl = f_len(dx, dy, dz);
    if (dl > a.h * 2.0f) {
        const float sc = (a.h * 2.0f) / dl;
        dx = dx * sc; dy = dy * sc; dz = dz * sc;
    }
    float nx = pos.x + dx, ny = pos.y + dy, nz = pos.z + dz;
    const bool dead = nx > hx - a.h * 1.5f || ny < 0.0f || ny > hy || nz > hz || s2.fluid < 0.05f;
    if (dead) {
        spawn(a, ps, i, seed, pos, vel);
        pos.w = 0.0f;
        parts[2 * (size_t)i] = pos;
        parts[2 * (size_t)i + 1] = vel;
        return;
    }
    if (nz < 0.0f) nz = -nz;
    const FSample sv = sample_flow(a, nx, ny, nz);
    const float val = f_clamp(field_value(a, view, ps.color_mode, sv, nx, ny, nz), 0.0f, 1.0f);
    parts[2 * (size_t)i] = make_float4(nx, ny, nz, pos.w + 1.0f);
    parts[2 * (size_t)i + 1] = make_float4(dx, dy, dz, val);
}

// seedOf (render-wgsl.ts:536-550): emitter rings (mode 1) or a rows x cols rake.
__device__ void seed_of(const aero_stream& st, const float4* em, uint32_t line, float& x, float& y, float& z) {
    if (st.seed_mode == 1u) {
        const uint32_t last = st.emitters - 1u;
        uint32_t ei = line / st.per_emitter;
        ei = ei < last ? ei : last;
        const uint32_t k = line % st.per_emitter;
        const float4 e = em[ei];
        const float ang = (((float)k + 0.5f) / (float)st.per_emitter) * 6.2831853f + st.time * 0.6f;
        const float ring = 0.35f + (0.65f * (float)(k % 3u)) / 2.0f;
        float sa, ca;  // FP64-evaluated, as the oracle evaluates the ring angle
        sdf::sincos_r64(ang, &sa, &ca);
        x = e.x + (0.0f * e.w) * ring;
        y = e.y + (ca * e.w) * ring;
        z = e.z + (sa * e.w) * ring;
        return;
    }
    const uint32_t row = line / st.cols, col = line % st.cols;
    const float fy = ((float)row + 0.5f) / (float)st.rows;
    const float fz = ((float)col + 0.5f) / (float)st.cols;
    x = st.seed_min[0];
    y = f_mix(st.seed_min[1], st.seed_max[1], fy);
    z = f_mix(st.seed_min[2], st.seed_max[2], fz);
}

// STREAM_SIM cs_main (render-wgsl.ts:551-575): midpoint tracing of `steps` points of
// stepLen per line; a dead line repeats its last point with value -1.
__global__ void __launch_bounds__(64) k_streams(FlowArgs a, aero_stream st, aero_view view, const float4* em,
                                                float4* pts) {
    const uint32_t line = blockIdx.x * blockDim.x + threadIdx.x;
    if (line >= st.lines) return;
    float4* out = pts + (size_t)line * st.steps;
    float px, py, pz;
    seed_of(st, em, line, px, py, pz);
    bool alive = true;
    const float hx = a.ox + (float)a.nx * a.h, hy = a.oy + (float)a.ny * a.h, hz = a.oz + (float)a.nz * a.h;
    for (uint32_t k = 0; k < st.steps; k++) {
        if (!alive) { out[k] = make_float4(px, py, pz, -1.0f); continue; }
        const FSample s1 = sample_flow(a, px, py, pz);
        const float l1 = f_len(s1.ux, s1.uy, s1.uz);
        if (s1.fluid < 0.3f || l1 < view.u_in * 0.03f) {
            alive = false;
            out[k] = make_float4(px, py, pz, -1.0f);
            continue;
        }
        out[k] = make_float4(px, py, pz, f_clamp(field_value(a, view, st.color_mode, s1, px, py, pz), 0.0f, 1.0f));
        const float mx = px + ((s1.ux / l1) * st.step_len) * 0.5f;
        const float my = py + ((s1.uy / l1) * st.step_len) * 0.5f;
        const float mz = pz + ((s1.uz / l1) * st.step_len) * 0.5f;
        const FSample s2 = sample_flow(a, mx, my, mz);
        const bool use2 = s2.fluid > 0.3f;
        const float dx = use2 ? s2.ux : s1.ux, dy = use2 ? s2.uy : s1.uy, dz = use2 ? s2.uz : s1.uz;
        const float l2 = f_len(dx, dy, dz);
        px = px + (dx / l2) * st.step_len;
        py = py + (dy / l2) * st.step_len;
        pz = pz + (dz / l2) * st.step_len;
        if (px > hx - a.h * 3.0f || py < 0.01f || py > hy || __builtin_fabsf(pz) > hz) alive = false;
    }
}

// Point probes: (ux, uy, uz, rho, temp, fluid, vorticityAt, fieldValue(mode)) per point.
__global__ void __launch_bounds__(256) k_probe(FlowArgs a, aero_view view, uint32_t mode, const float* pts, uint32_t n,
                                               float* out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float px = pts[3 * (size_t)i], py = pts[3 * (size_t)i + 1], pz = pts[3 * (size_t)i + 2];
    const FSample s = sample_flow(a, px, py, pz);
    float* o = out + 8 * (size_t)i;
    o[0] = s.ux; o[1] = s.uy; o[2] = s.uz; o[3] = s.rho; o[4] = s.temp; o[5] = s.fluid;
    o[6] = vorticity_at(a, px, py, pz);
    o[7] = field_value(a, view, mode, s, px, py, pz);
}

namespace {
FlowArgs flow_args(const aero_ctx* c) {
    FlowArgs a{};
    a.fields = c->fields;
    a.temps = c->temps[c->tcur];
    a.nx = c->grid.nx; a.ny = c->grid.ny; a.nz = c->grid.nz;
    a.ox = c->grid.ox; a.oy = c->grid.oy; a.oz = c->grid.oz; a.h = c->grid.h;
    return a;
}
}  // namespace

hipError_t launch_probe(aero_ctx* c, const aero_view* v, uint32_t mode, const float* d_pts, uint32_t n, float* d_out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, c->stream, flow_args(c), *v, mode, d_pts, n, d_out);
    return hipGetLastError();
}

hipError_t launch_particles(aero_ctx* c, const aero_psim* ps, const aero_view* v) {
    if (ps->count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_particles, dim3((ps->count + 255) / 256), dim3(256), 0, c->stream, flow_args(c), *ps, *v,
                       (float4*)c->parts);
    return hipGetLastError();
}

hipError_t launch_streams(aero_ctx* c, const aero_stream* st, const aero_view* v, const float* d_em) {
    if (st->lines == 0 || st->steps == 0) return hipSuccess;
    hipLaunchKernelGGL(k_streams, dim3((st->lines + 63) / 64), dim3(64), 0, c->stream, flow_args(c), *st, *v,
                       (const float4*)d_em, (float4*)c->stream_pts);
    return hipGetLastError();
}

}  // namespace aero
