// Vortex-lattice wing loads for gfx950: many lattices (wing setups) solved in one set
// of launches, each lattice on its own workgroups.
//
// Reference behaviour: wingLoads of src/engine/vlm.ts:140-182 after buildLattice
// (segment :42-51, horseshoeVelocity :54-66, solve :115-138).  Four kernels:
//   k_vlm_matrix  one thread per (collocation point i, horseshoe j): the normal
//                 velocity A[i][j] of horseshoe j (unit circulation, with its ground
//                 image) at cp_i, and rhs[i] = -vinf . n_i;
//   k_vlm_solve   one 1024-thread workgroup per lattice: Gaussian elimination with
//                 partial pivoting on the L2-resident matrix, the pivot column kept in
//                 LDS and handed from one step to the next, then back substitution;
//   k_vlm_forces  one workgroup per bound vortex: the 10 n segment velocities at its
//                 midpoint computed in parallel into LDS, then added to the free stream
//                 in the reference's order by three lanes (one per component), and the
//                 Kutta-Joukowski force and moment terms of that vortex;
//   k_vlm_sum     one lane per lattice: the force, moment, area and centroid sums in
//                 index order and the four outputs.
// Every value is produced by the reference's FP64 operations in the reference's order
// (no FMA contraction, IEEE division and square root, Math.hypot as V8 evaluates it),
// so the outputs are bit-identical to a sequential evaluation.  The elimination's
// updates are independent per matrix entry within one pivot step, so running them in
// parallel changes no rounding; entries that the reference updates but never reads
// again (below the diagonal) are left alone.
#include <string>
#include <vector>

#include "aero_internal.h"

namespace aero {

constexpr double VLM_RHO = 1.225;
constexpr double VLM_CD0 = 0.018;
constexpr double VLM_CORE = 0.01;
constexpr double VLM_FAR = 60.0;
constexpr double VLM_PI = 3.141592653589793;  // Math.PI
constexpr int VLM_REC = 20;                     // doubles per horseshoe record
constexpr int VLM_SOLVE_THREADS = 1024;
constexpr int VLM_SOLVE_WAVES = VLM_SOLVE_THREADS / 64;
constexpr int VLM_FORCE_THREADS = 256;
constexpr int VLM_FORCE_CHUNK = 2048;  // segments per LDS pass of k_vlm_forces

struct VlmBatch {
    const double* hs;       // horseshoe records, all lattices back to back
    const uint32_t* off;    // count + 1 record offsets
    const uint64_t* moff;   // count matrix offsets (doubles)
    const double* speed;    // count free-stream speeds
    double* A;              // matrices, row-major n x n each
    double* rhs;            // off[count]
    double* gamma;          // off[count]
    double* terms;          // off[count] x 5: force x, y, z, moment x, moment y terms
    double* out;            // count x 4
    uint32_t count;
};

// Math.hypot(x, y, z) as V8 evaluates it: NaN and infinity first, then the magnitudes
// scaled by the largest and their squares Kahan-summed.
__device__ __forceinline__ double js_hypot3(double x, double y, double z) {
    const bool nx = x != x, ny = y != y, nz = z != z;
    const double ax = nx ? 0.0 : __builtin_fabs(x), ay = ny ? 0.0 : __builtin_fabs(y),
                 az = nz ? 0.0 : __builtin_fabs(z);
    double m = 0.0;
    if (ax > m) m = ax;
    if (ay > m) m = ay;
    if (az > m) m = az;
    if (m == __builtin_inf()) return __builtin_inf();
    if (nx || ny || nz) return __builtin_nan("");
    if (m == 0.0) return 0.0;
    double sum = 0.0, comp = 0.0;
    const double a[3] = {ax, ay, az};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const double n = a[i] / m;
        const double s = n * n - comp;
        const double p = sum + s;
        comp = (p - sum) - s;
        sum = p;
    }
    return __builtin_sqrt(sum) * m;
}

__device__ __forceinline__ double vdot(const double* a, const double* b) {
    return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// segment (vlm.ts:42-51): the velocity at p of the vortex segment a -> b of circulation
// gamma into o; false when p lies on an end point (the reference adds nothing then).
__device__ __forceinline__ bool vlm_segment(const double* p, const double* a, const double* b, double gamma,
                                            double* o) {
    const double r1[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
    const double r2[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
    const double c[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2],
                         r1[0] * r2[1] - r1[1] * r2[0]};
    const double r0[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    const double l1 = js_hypot3(r1[0], r1[1], r1[2]), l2 = js_hypot3(r2[0], r2[1], r2[2]);
    const double c2 = vdot(c, c) + (VLM_CORE * VLM_CORE) * vdot(r0, r0);
    if (l1 < 1e-6 || l2 < 1e-6) return false;
    const double k = (gamma / ((4.0 * VLM_PI) * c2)) * (vdot(r0, r1) / l1 - vdot(r0, r2) / l2);
    o[0] = c[0] * k;
    o[1] = c[1] * k;
    o[2] = c[2] * k;
    return true;
}

// Segment q (0..9) of horseshoeVelocity (vlm.ts:54-66): q < 5 the horseshoe itself,
// q >= 5 its ground image (y -> -y, circulation negated); within each, the legs
// far -> ta, ta -> a, the bound vortex a -> b, b -> tb, tb -> far.
__device__ __forceinline__ bool vlm_horseshoe_segment(const double* p, const double* h, double gamma, int q,
                                                      double* o) {
    const double sign = q < 5 ? 1.0 : -1.0;
    const int leg = q < 5 ? q : q - 5;
    // the chain afar, ta, a, b, tb, bfar: leg l runs from point l to point l + 1
    double e0[3], e1[3];
    auto point = [&](int k, double* d) {
        if (k == 0 || k == 5) {
            const double* t = h + (k == 0 ? 6 : 9);
            d[0] = t[0] + VLM_FAR;
            d[1] = t[1] * sign;
            d[2] = t[2];
        } else {
            const double* t = h + (k == 1 ? 6 : k == 2 ? 0 : k == 3 ? 3 : 9);
            d[0] = t[0];
            d[1] = t[1] * sign;
            d[2] = t[2];
        }
    };
    point(leg, e0);
    point(leg + 1, e1);
    return vlm_segment(p, e0, e1, gamma * sign, o);
}

__global__ __launch_bounds__(256) void k_vlm_matrix(VlmBatch B) {
    const uint32_t l = blockIdx.y;
    const uint32_t base = B.off[l], n = B.off[l + 1] - base;
    const uint64_t e = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (uint64_t)n * n) return;
    const uint32_t i = (uint32_t)(e / n), j = (uint32_t)(e % n);
    const double* hi = B.hs + (size_t)VLM_REC * (base + i);
    const double* hj = B.hs + (size_t)VLM_REC * (base + j);
    double v[3] = {0.0, 0.0, 0.0};
    for (int q = 0; q < 10; q++) {
        double o[3];
        if (vlm_horseshoe_segment(hi + 12, hj, 1.0, q, o)) {
            v[0] += o[0];
            v[1] += o[1];
            v[2] += o[2];
        }
    }
    B.A[B.moff[l] + e] = vdot(v, hi + 15);
    if (j == 0) {
        const double vinf[3] = {B.speed[l], 0.0, 0.0};
        B.rhs[base + i] = -vdot(vinf, hi + 15);
    }
}

// solve (vlm.ts:115-138) on one lattice per workgroup.
__global__ __launch_bounds__(VLM_SOLVE_THREADS) void k_vlm_solve(VlmBatch B) {
    __shared__ double col[2][AERO_VLM_MAX_HORSESHOES];
    __shared__ double row[AERO_VLM_MAX_HORSESHOES];
    __shared__ double rhs[AERO_VLM_MAX_HORSESHOES];
    __shared__ double x[AERO_VLM_MAX_HORSESHOES];
    __shared__ double red_v[VLM_SOLVE_WAVES];
    __shared__ int red_i[VLM_SOLVE_WAVES];
    __shared__ int piv_s;
    const uint32_t l = blockIdx.x;
    const uint32_t base = B.off[l];
    const int n = (int)(B.off[l + 1] - base);
    double* A = B.A + B.moff[l];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < n; i += VLM_SOLVE_THREADS) {
        rhs[i] = B.rhs[base + i];
        col[0][i] = A[(size_t)i * n];
    }
    __syncthreads();
    int cur = 0;
    for (int k = 0; k < n; k++) {
        // pivot: the first row of largest |A[i][k]|, NaNs never taken, and no move at
        // all when |A[k][k]| is NaN (the reference's strict > scan)
        double bv = -1.0;
        int bi = n;
        for (int i = k + tid; i < n; i += VLM_SOLVE_THREADS) {
            const double v = __builtin_fabs(col[cur][i]);
            if (v > bv) {
                bv = v;
                bi = i;
            }
        }
        for (int s = 32; s > 0; s >>= 1) {
            const double ov = __shfl_xor(bv, s);
            const int oi = __shfl_xor(bi, s);
            if (ov > bv || (ov == bv && oi < bi)) {
                bv = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            red_v[wave] = bv;
            red_i[wave] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            double v = red_v[0];
            int p = red_i[0];
            for (int w = 1; w < VLM_SOLVE_WAVES; w++)
                if (red_v[w] > v || (red_v[w] == v && red_i[w] < p)) {
                    v = red_v[w];
                    p = red_i[w];
                }
            const double akk = col[cur][k];
            piv_s = (akk != akk || p >= n) ? k : p;
        }
        __syncthreads();
        const int piv = piv_s;
        // row swap (columns below the diagonal are never read again) and the pivot row
        double* rk = A + (size_t)k * n;
        double* rp = A + (size_t)piv * n;
        for (int j = k + tid; j < n; j += VLM_SOLVE_THREADS) {
            const double b = rp[j];
            if (piv != k) rp[j] = rk[j];
            rk[j] = b;
            row[j] = b;
        }
        if (tid == 0 && piv != k) {
            const double t = rhs[k];
            rhs[k] = rhs[piv];
            rhs[piv] = t;
            const double c = col[cur][k];
            col[cur][k] = col[cur][piv];
            col[cur][piv] = c;
        }
        __syncthreads();
        double d = row[k];
        if (d == 0.0 || d != d) d = 1e-12;
        const int nxt = cur ^ 1;
        for (int i = k + 1 + wave; i < n; i += VLM_SOLVE_WAVES) {
            const double f = col[cur][i] / d;
            double* ri = A + (size_t)i * n;
            if (f == 0.0) {
                if (lane == 0 && k + 1 < n) col[nxt][i] = ri[k + 1];
                continue;
            }
            for (int j = k + 1 + lane; j < n; j += 64) {
                const double v = ri[j] - f * row[j];
                ri[j] = v;
                if (j == k + 1) col[nxt][i] = v;
            }
            if (lane == 0) rhs[i] = rhs[i] - f * rhs[k];
        }
        __syncthreads();
        cur = nxt;
    }
    // back substitution: products in parallel, the running difference in order
    for (int i = n - 1; i >= 0; i--) {
        const double* ri = A + (size_t)i * n;
        for (int j = i + 1 + tid; j < n; j += VLM_SOLVE_THREADS) row[j] = ri[j] * x[j];
        __syncthreads();
        if (tid == 0) {
            double s = rhs[i];
            for (int j = i + 1; j < n; j++) s -= row[j];
            double d = ri[i];
            if (d == 0.0 || d != d) d = 1e-12;
            x[i] = s / d;
        }
        __syncthreads();
    }
    for (int i = tid; i < n; i += VLM_SOLVE_THREADS) B.gamma[base + i] = x[i];
}

// The force of bound vortex i (vlm.ts:159-171): the velocity at its midpoint is the
// free stream plus the 10 n segment velocities, added in horseshoe and segment order.
__global__ __launch_bounds__(VLM_FORCE_THREADS) void k_vlm_forces(VlmBatch B) {
    __shared__ double seg[VLM_FORCE_CHUNK * 3];
    __shared__ uint8_t used[VLM_FORCE_CHUNK];
    __shared__ double vel[3];
    const uint32_t l = blockIdx.y;
    const uint32_t base = B.off[l], n = B.off[l + 1] - base;
    const uint32_t i = blockIdx.x;
    if (i >= n) return;
    const double* h = B.hs + (size_t)VLM_REC * (base + i);
    const double m[3] = {(h[0] + h[3]) / 2, (h[1] + h[4]) / 2, (h[2] + h[5]) / 2};
    const int tid = threadIdx.x;
    double acc = tid == 0 ? B.speed[l] : 0.0;  // lanes 0, 1, 2 hold vel x, y, z
    const uint32_t total = 10 * n;
    for (uint32_t s0 = 0; s0 < total; s0 += VLM_FORCE_CHUNK) {
        const uint32_t cnt = total - s0 < (uint32_t)VLM_FORCE_CHUNK ? total - s0 : (uint32_t)VLM_FORCE_CHUNK;
        for (uint32_t s = tid; s < cnt; s += VLM_FORCE_THREADS) {
            const uint32_t gs = s0 + s, j = gs / 10;
            double o[3] = {0.0, 0.0, 0.0};
            used[s] = vlm_horseshoe_segment(m, B.hs + (size_t)VLM_REC * (base + j), B.gamma[base + j],
                                            (int)(gs - 10 * j), o);
            seg[3 * s] = o[0];
            seg[3 * s + 1] = o[1];
            seg[3 * s + 2] = o[2];
        }
        __syncthreads();
        if (tid < 3)
            for (uint32_t s = 0; s < cnt; s++)
                if (used[s]) acc += seg[3 * s + tid];
        __syncthreads();
    }
    if (tid < 3) vel[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        const double g = B.gamma[base + i];
        const double ab[3] = {h[3] - h[0], h[4] - h[1], h[5] - h[2]};
        const double f[3] = {vel[1] * ab[2] - vel[2] * ab[1], vel[2] * ab[0] - vel[0] * ab[2],
                             vel[0] * ab[1] - vel[1] * ab[0]};
        double* t = B.terms + (size_t)5 * (base + i);
        t[0] = (VLM_RHO * g) * f[0];
        t[1] = (VLM_RHO * g) * f[1];
        t[2] = (VLM_RHO * g) * f[2];
        t[3] = (((-VLM_RHO) * g) * f[1]) * m[0];
        t[4] = ((VLM_RHO * g) * f[0]) * m[1];
    }
}

// The sums and outputs of vlm.ts:157-181, one lane per lattice.
__global__ __launch_bounds__(64) void k_vlm_sum(VlmBatch B) {
    const uint32_t l = blockIdx.x * 64 + threadIdx.x;
    if (l >= B.count) return;
    const uint32_t base = B.off[l], n = B.off[l + 1] - base;
    double fx = 0.0, fy = 0.0, area = 0.0, mx = 0.0, my = 0.0, cx = 0.0, cy = 0.0;
    for (uint32_t i = 0; i < n; i++) {
        const double* t = B.terms + (size_t)5 * (base + i);
        fx += t[0];
        fy += t[1];
        mx += t[3];
        my += t[4];
        area += B.hs[(size_t)VLM_REC * (base + i) + 18];
    }
    for (uint32_t i = 0; i < n; i++) cx = cx + B.hs[(size_t)VLM_REC * (base + i) + 12];
    for (uint32_t i = 0; i < n; i++) cy = cy + B.hs[(size_t)VLM_REC * (base + i) + 13];
    const double speed = B.speed[l];
    const double q = ((0.5 * VLM_RHO) * speed) * speed;
    const double down = -fy;
    const double nn = n > 1 ? (double)n : 1.0;
    double* o = B.out + (size_t)4 * l;
    o[0] = down;
    o[1] = fx + (VLM_CD0 * q) * area;
    o[2] = __builtin_fabs(down) > 1.0 ? mx / down : cx / nn;
    o[3] = __builtin_fabs(fx) > 1.0 ? my / fx : cy / nn;
}

}  // namespace aero

using namespace aero;

extern "C" int aero_vlm_loads(const double* hs, const uint32_t* offsets, uint32_t count, const double* speeds,
                              double* out, double* gamma_out) {
    auto fail_msg = [](int code, const std::string& m) {
        report_global_error(m);
        return code;
    };
    if (!offsets || !speeds || !out) return fail_msg(AERO_E_INVALID, "aero_vlm_loads: null argument");
    if (count == 0 || count > 65535) return fail_msg(AERO_E_INVALID, "aero_vlm_loads: count must be 1 .. 65535");
    if (offsets[0] != 0) return fail_msg(AERO_E_INVALID, "aero_vlm_loads: offsets[0] must be 0");
    std::vector<uint64_t> moff(count);
    uint64_t msize = 0;
    uint32_t nmax = 0;
    for (uint32_t l = 0; l < count; l++) {
        if (offsets[l + 1] < offsets[l] || offsets[l + 1] - offsets[l] > AERO_VLM_MAX_HORSESHOES)
            return fail_msg(AERO_E_INVALID, "aero_vlm_loads: lattice " + std::to_string(l) + " must hold 0 .. " +
                                                std::to_string(AERO_VLM_MAX_HORSESHOES) + " horseshoes");
        const uint64_t n = offsets[l + 1] - offsets[l];
        if (!(__builtin_isfinite(speeds[l])))
            return fail_msg(AERO_E_INVALID, "aero_vlm_loads: speed of lattice " + std::to_string(l) + " not finite");
        moff[l] = msize;
        msize += n * n;
        if (n > nmax) nmax = (uint32_t)n;
    }
    const uint64_t nh = offsets[count];
    if (nh && !hs) return fail_msg(AERO_E_INVALID, "aero_vlm_loads: null horseshoes");
    VlmBatch B{};
    B.count = count;
    hipStream_t st = nullptr;
    double *d_hs = nullptr, *d_speed = nullptr, *d_A = nullptr, *d_rhs = nullptr, *d_gamma = nullptr,
           *d_terms = nullptr, *d_out = nullptr;
    uint32_t* d_off = nullptr;
    uint64_t* d_moff = nullptr;
    const size_t hb = (size_t)nh * VLM_REC * sizeof(double), vb = (size_t)(nh ? nh : 1) * sizeof(double);
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess && nh) e = hipMalloc(&d_hs, hb);
    if (e == hipSuccess) e = hipMalloc(&d_off, (size_t)(count + 1) * 4);
    if (e == hipSuccess) e = hipMalloc(&d_moff, (size_t)count * 8);
    if (e == hipSuccess) e = hipMalloc(&d_speed, (size_t)count * 8);
    if (e == hipSuccess && msize) e = hipMalloc(&d_A, msize * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&d_rhs, vb);
    if (e == hipSuccess) e = hipMalloc(&d_gamma, vb);
    if (e == hipSuccess) e = hipMalloc(&d_terms, 5 * vb);
    if (e == hipSuccess) e = hipMalloc(&d_out, (size_t)count * 4 * sizeof(double));
    if (e == hipSuccess && nh) e = hipMemcpyAsync(d_hs, hs, hb, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_off, offsets, (size_t)(count + 1) * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_moff, moff.data(), (size_t)count * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_speed, speeds, (size_t)count * 8, hipMemcpyHostToDevice, st);
    B.hs = d_hs;
    B.off = d_off;
    B.moff = d_moff;
    B.speed = d_speed;
    B.A = d_A;
    B.rhs = d_rhs;
    B.gamma = d_gamma;
    B.terms = d_terms;
    B.out = d_out;
    if (e == hipSuccess && nmax) {
        const uint64_t cells = (uint64_t)nmax * nmax;
        hipLaunchKernelGGL(k_vlm_matrix, dim3((unsigned)((cells + 255) / 256), count), dim3(256), 0, st, B);
        e = hipGetLastError();
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_vlm_solve, dim3(count), dim3(VLM_SOLVE_THREADS), 0, st, B);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_vlm_forces, dim3(nmax, count), dim3(VLM_FORCE_THREADS), 0, st, B);
            e = hipGetLastError();
        }
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_vlm_sum, dim3((count + 63) / 64), dim3(64), 0, st, B);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)count * 4 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && gamma_out && nh)
        e = hipMemcpyAsync(gamma_out, d_gamma, (size_t)nh * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    for (void* p : {(void*)d_hs, (void*)d_off, (void*)d_moff, (void*)d_speed, (void*)d_A, (void*)d_rhs,
                    (void*)d_gamma, (void*)d_terms, (void*)d_out})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess)
        return fail_msg(e == hipErrorOutOfMemory ? AERO_E_NOMEM : AERO_E_HIP,
                        std::string("aero_vlm_loads: ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")");
    return AERO_OK;
}
