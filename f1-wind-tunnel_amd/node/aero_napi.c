/* N-API addon over libaero_lbm.so: the binding a TypeScript host loads to drive the
 * simulation half of the reference's AeroEngine (src/engine/engine.ts:159-769) on an
 * MI355X.  Every function mirrors one C-ABI entry point of include/aero_lbm.h; host
 * arrays cross as typed arrays and are only borrowed for the call.  A negative return
 * code becomes a thrown Error carrying the library's message, the way the reference
 * throws or rejects (engine.ts:760).
 *
 * Build: gcc -shared -fPIC -I/usr/include/node -Iinclude aero_napi.c -Llib -laero_lbm
 */
#define NAPI_VERSION 4
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "aero_lbm.h"

#define NAPI_CALL(env, call)                                                     \
    do {                                                                         \
        if ((call) != napi_ok) {                                                 \
            napi_throw_error((env), "AERO_NAPI", "N-API call failed: " #call);   \
            return NULL;                                                         \
        }                                                                        \
    } while (0)

static napi_value throw_rc(napi_env env, int rc, aero_ctx* c, const char* what) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s failed (%d): %s", what, rc, c ? aero_last_error(c) : aero_global_error());
    napi_throw_error(env, "AERO_E", buf);
    return NULL;
}

/* The JS handle: the context and its cell count (read buffers are sized from it, never
 * from an argument).  destroy() releases the context; the handle itself is freed by
 * the garbage collector. */
typedef struct {
    aero_ctx* c;
    uint32_t n;
    uint32_t nparts;  /* particles allocated by particlesAlloc */
} handle_t;

static void finalize_handle(napi_env env, void* data, void* hint) {
    (void)env;
    (void)hint;
    free(data);
}

/* Argument helpers. */
static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value* argv) {
    size_t argc = want;
    if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok || argc < want) {
        napi_throw_type_error(env, "AERO_ARGS", "too few arguments");
        return 0;
    }
    return 1;
}

static handle_t* get_handle(napi_env env, napi_value v) {
    void* p = NULL;
    napi_valuetype t;
    if (napi_typeof(env, v, &t) != napi_ok || t != napi_external || napi_get_value_external(env, v, &p) != napi_ok ||
        !p || !((handle_t*)p)->c) {
        napi_throw_type_error(env, "AERO_ARGS", "expected a live engine context");
        return NULL;
    }
    return (handle_t*)p;
}

static aero_ctx* get_ctx(napi_env env, napi_value v) {
    handle_t* h = get_handle(env, v);
    return h ? h->c : NULL;
}

/* typed array data pointer and element count; type checked */
static void* typed(napi_env env, napi_value v, napi_typedarray_type want, size_t* len) {
    bool is = false;
    if (napi_is_typedarray(env, v, &is) != napi_ok || !is) {
        napi_throw_type_error(env, "AERO_ARGS", "expected a typed array");
        return NULL;
    }
    napi_typedarray_type t;
    void* data = NULL;
    size_t n = 0;
    if (napi_get_typedarray_info(env, v, &t, &n, &data, NULL, NULL) != napi_ok || t != want) {
        napi_throw_type_error(env, "AERO_ARGS", "typed array of the wrong element type");
        return NULL;
    }
    if (len) *len = n;
    return data;
}

static int is_nullish(napi_env env, napi_value v) {
    napi_valuetype t;
    napi_typeof(env, v, &t);
    return t == napi_null || t == napi_undefined;
}

static napi_value make_typed(napi_env env, napi_typedarray_type type, size_t n, size_t esz, void** data) {
    napi_value ab, arr;
    if (napi_create_arraybuffer(env, n * esz, data, &ab) != napi_ok) return NULL;
    if (napi_create_typedarray(env, type, n, ab, 0, &arr) != napi_ok) return NULL;
    return arr;
}

/* create(grid: Float64Array[8] = nx,ny,nz,n,ox,oy,oz,h, sim: Float64Array[3] = uIn,tau0,smag,
 *        storage: number) -> context */
static napi_value js_create(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return NULL;
    size_t ng = 0, ns = 0;
    const double* g = (const double*)typed(env, a[0], napi_float64_array, &ng);
    if (!g) return NULL;
    const double* s = (const double*)typed(env, a[1], napi_float64_array, &ns);
    if (!s) return NULL;
    if (ng != 8 || ns < 3) {
        napi_throw_range_error(env, "AERO_ARGS", "grid needs 8 values, sim 3");
        return NULL;
    }
    int32_t storage = 0;
    NAPI_CALL(env, napi_get_value_int32(env, a[2], &storage));
    aero_grid grid = {(uint32_t)g[0], (uint32_t)g[1], (uint32_t)g[2], (uint32_t)g[3],
                      (float)g[4],    (float)g[5],    (float)g[6],    (float)g[7]};
    aero_sim sim = {(float)s[0], (float)s[1], (float)s[2], 0.0f};
    aero_ctx* ctx = NULL;
    const int rc = aero_create(&grid, &sim, storage, &ctx);
    if (rc) return throw_rc(env, rc, NULL, "aero_create");
    handle_t* h = (handle_t*)malloc(sizeof(handle_t));
    if (!h) {
        aero_destroy(ctx);
        napi_throw_error(env, "AERO_NAPI", "allocation failed");
        return NULL;
    }
    h->c = ctx;
    h->n = grid.n;
    h->nparts = 0;
    napi_value ext;
    if (napi_create_external(env, h, finalize_handle, NULL, &ext) != napi_ok) {
        aero_destroy(ctx);
        free(h);
        napi_throw_error(env, "AERO_NAPI", "napi_create_external failed");
        return NULL;
    }
    return ext;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    const int rc = aero_destroy(h->c);
    h->c = NULL;  /* later calls on this handle throw instead of touching freed memory */
    if (rc) return throw_rc(env, rc, NULL, "aero_destroy");
    return NULL;
}

static napi_value js_init(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    const int rc = aero_init(c);
    if (rc) return throw_rc(env, rc, c, "aero_init");
    return NULL;
}

/* voxelize(ctx, shape: Uint8Array[96] (the packed Shape uniform), profile: Float32Array|null,
 *          prims: Uint8Array (64 B each)|null) */
static napi_value js_voxelize(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!get_args(env, info, 4, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t n = 0;
    const uint8_t* shape = (const uint8_t*)typed(env, a[1], napi_uint8_array, &n);
    if (!shape) return NULL;
    if (n != sizeof(aero_shape)) {
        napi_throw_range_error(env, "AERO_ARGS", "shape must be 96 bytes");
        return NULL;
    }
    aero_shape sh;
    memcpy(&sh, shape, sizeof sh);
    const float* profile = NULL;
    if (!is_nullish(env, a[2])) {
        profile = (const float*)typed(env, a[2], napi_float32_array, &n);
        if (!profile) return NULL;
        if (n != 256 * 128) {
            napi_throw_range_error(env, "AERO_ARGS", "profile must hold 256 x 128 values");
            return NULL;
        }
    }
    const uint8_t* prims = NULL;
    size_t np = 0;
    if (!is_nullish(env, a[3])) {
        prims = (const uint8_t*)typed(env, a[3], napi_uint8_array, &n);
        if (!prims) return NULL;
        if (n % sizeof(aero_prim)) {
            napi_throw_range_error(env, "AERO_ARGS", "prims must be whole 64-byte records");
            return NULL;
        }
        np = n / sizeof(aero_prim);
    }
    const int rc = aero_voxelize(c, &sh, profile, (const aero_prim*)prims, (uint32_t)np);
    if (rc) return throw_rc(env, rc, c, "aero_voxelize");
    return NULL;
}

static napi_value js_set_flags(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t ns = 0, ne = 0;
    const uint32_t* solid = (const uint32_t*)typed(env, a[1], napi_uint32_array, &ns);
    if (!solid) return NULL;
    const float* emit = (const float*)typed(env, a[2], napi_float32_array, &ne);
    if (!emit) return NULL;
    if (ns != ne || ns != get_handle(env, a[0])->n) {
        napi_throw_range_error(env, "AERO_ARGS", "solid and emit must hold one value per cell");
        return NULL;
    }
    const int rc = aero_set_flags(c, solid, emit);
    if (rc) return throw_rc(env, rc, c, "aero_set_flags");
    return NULL;
}

static napi_value js_step(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    uint32_t n = 0;
    NAPI_CALL(env, napi_get_value_uint32(env, a[1], &n));
    const int rc = aero_step(c, n);
    if (rc) return throw_rc(env, rc, c, "aero_step");
    return NULL;
}

/* heatStep(ctx, heat: Float32Array[4] = dt, cool, diffuse, pad, n) */
static napi_value js_heat_step(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t n4 = 0;
    const float* h = (const float*)typed(env, a[1], napi_float32_array, &n4);
    if (!h) return NULL;
    if (n4 < 3) {
        napi_throw_range_error(env, "AERO_ARGS", "heat needs dt, cool, diffuse");
        return NULL;
    }
    uint32_t n = 0;
    NAPI_CALL(env, napi_get_value_uint32(env, a[2], &n));
    aero_heat heat = {h[0], h[1], h[2], 0.0f};
    const int rc = aero_heat_step(c, &heat, n);
    if (rc) return throw_rc(env, rc, c, "aero_heat_step");
    return NULL;
}

/* readForces(ctx, reset: boolean) -> { sums: Float64Array[9], quantized: Float64Array[9], steps } */
static napi_value js_read_forces(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    bool reset = false;
    NAPI_CALL(env, napi_get_value_bool(env, a[1], &reset));
    double d[9];
    int64_t q[9];
    uint64_t steps = 0;
    const int rc = aero_read_forces(c, d, q, &steps, reset ? 1 : 0);
    if (rc) return throw_rc(env, rc, c, "aero_read_forces");
    void* pd = NULL;
    void* pq = NULL;
    napi_value sums = make_typed(env, napi_float64_array, 9, 8, &pd);
    napi_value quant = make_typed(env, napi_float64_array, 9, 8, &pq);
    if (!sums || !quant) {
        napi_throw_error(env, "AERO_NAPI", "allocation failed");
        return NULL;
    }
    memcpy(pd, d, sizeof d);
    for (int k = 0; k < 9; k++) ((double*)pq)[k] = (double)q[k];  /* |q| < 2^53 for any run this size */
    napi_value obj, st;
    NAPI_CALL(env, napi_create_object(env, &obj));
    NAPI_CALL(env, napi_create_double(env, (double)steps, &st));
    NAPI_CALL(env, napi_set_named_property(env, obj, "sums", sums));
    NAPI_CALL(env, napi_set_named_property(env, obj, "quantized", quant));
    NAPI_CALL(env, napi_set_named_property(env, obj, "steps", st));
    return obj;
}

/* readFields(ctx) -> Float32Array[4n]; readTemps(ctx) -> Float32Array[n];
 * readPopulations(ctx) -> Float32Array[19n]; readFlags(ctx) -> {solid, emit};
 * n is the context's cell count. */
static napi_value js_read_fields(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    aero_ctx* c = h->c;
    const uint32_t n = h->n;
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, (size_t)n * 4, 4, &p);
    if (!arr) return NULL;
    const int rc = aero_read_fields(c, (float*)p);
    if (rc) return throw_rc(env, rc, c, "aero_read_fields");
    return arr;
}

static napi_value js_read_temps(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    aero_ctx* c = h->c;
    const uint32_t n = h->n;
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, n, 4, &p);
    if (!arr) return NULL;
    const int rc = aero_read_temps(c, (float*)p);
    if (rc) return throw_rc(env, rc, c, "aero_read_temps");
    return arr;
}

static napi_value js_read_populations(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    aero_ctx* c = h->c;
    const uint32_t n = h->n;
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, (size_t)n * 19, 4, &p);
    if (!arr) return NULL;
    const int rc = aero_read_populations(c, (float*)p);
    if (rc) return throw_rc(env, rc, c, "aero_read_populations");
    return arr;
}

static napi_value js_read_flags(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    aero_ctx* c = h->c;
    const uint32_t n = h->n;
    void* ps = NULL;
    void* pe = NULL;
    napi_value solid = make_typed(env, napi_uint32_array, n, 4, &ps);
    napi_value emit = make_typed(env, napi_float32_array, n, 4, &pe);
    if (!solid || !emit) return NULL;
    const int rc = aero_read_flags(c, (uint32_t*)ps, (float*)pe);
    if (rc) return throw_rc(env, rc, c, "aero_read_flags");
    napi_value obj;
    NAPI_CALL(env, napi_create_object(env, &obj));
    NAPI_CALL(env, napi_set_named_property(env, obj, "solid", solid));
    NAPI_CALL(env, napi_set_named_property(env, obj, "emit", emit));
    return obj;
}

static napi_value js_guard_count(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    uint64_t n = 0;
    const int rc = aero_guard_count(c, &n);
    if (rc) return throw_rc(env, rc, c, "aero_guard_count");
    napi_value v;
    NAPI_CALL(env, napi_create_double(env, (double)n, &v));
    return v;
}

static napi_value js_sync(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    const int rc = aero_sync(c);
    if (rc) return throw_rc(env, rc, c, "aero_sync");
    return NULL;
}

/* ---- field consumers (wgsl.ts:324-457,517-595; render-wgsl.ts:520-576) ---- */

static int get_view(napi_env env, napi_value v, aero_view* out) {
    size_t n = 0;
    const float* f = (const float*)typed(env, v, napi_float32_array, &n);
    if (!f) return 0;
    if (n != 8) {
        napi_throw_range_error(env, "AERO_ARGS", "view needs 8 values (the View uniform)");
        return 0;
    }
    memcpy(out, f, sizeof *out);
    return 1;
}

/* a packed 64-byte uniform (PSim or Stream) */
static int get_u64b(napi_env env, napi_value v, void* out) {
    size_t n = 0;
    const uint8_t* b = (const uint8_t*)typed(env, v, napi_uint8_array, &n);
    if (!b) return 0;
    if (n != 64) {
        napi_throw_range_error(env, "AERO_ARGS", "uniform must be 64 bytes");
        return 0;
    }
    memcpy(out, b, 64);
    return 1;
}

/* writeFields(ctx, fields: Float32Array[4n]) */
static napi_value js_write_fields(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    size_t n = 0;
    const float* f = (const float*)typed(env, a[1], napi_float32_array, &n);
    if (!f) return NULL;
    if (n != (size_t)h->n * 4) {
        napi_throw_range_error(env, "AERO_ARGS", "fields must hold 4 values per cell");
        return NULL;
    }
    const int rc = aero_write_fields(h->c, f);
    if (rc) return throw_rc(env, rc, h->c, "aero_write_fields");
    return NULL;
}

/* sampleFlow(ctx, pts: Float32Array[3k], view: Float32Array[8], colorMode) -> Float32Array[8k] */
static napi_value js_sample_flow(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!get_args(env, info, 4, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    size_t n = 0;
    const float* pts = (const float*)typed(env, a[1], napi_float32_array, &n);
    if (!pts) return NULL;
    if (n % 3 || n / 3 > (1u << 26)) {
        napi_throw_range_error(env, "AERO_ARGS", "points must be whole (x, y, z) triples, at most 2^26");
        return NULL;
    }
    aero_view view;
    if (!get_view(env, a[2], &view)) return NULL;
    uint32_t mode = 0;
    NAPI_CALL(env, napi_get_value_uint32(env, a[3], &mode));
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, n / 3 * 8, 4, &p);
    if (!arr) return NULL;
    const int rc = aero_sample_flow(c, pts, (uint32_t)(n / 3), &view, mode, (float*)p);
    if (rc) return throw_rc(env, rc, c, "aero_sample_flow");
    return arr;
}

/* particlesAlloc(ctx, count) */
static napi_value js_particles_alloc(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    uint32_t count = 0;
    NAPI_CALL(env, napi_get_value_uint32(env, a[1], &count));
    h->nparts = 0;
    const int rc = aero_particles_alloc(h->c, count);
    if (rc) return throw_rc(env, rc, h->c, "aero_particles_alloc");
    h->nparts = count;
    return NULL;
}

/* particlesStep(ctx, psim: Uint8Array[64], view: Float32Array[8]) */
static napi_value js_particles_step(napi_env env, napi_callback_info info) {
    napi_value a[3];
    if (!get_args(env, info, 3, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    aero_psim ps;
    aero_view view;
    if (!get_u64b(env, a[1], &ps) || !get_view(env, a[2], &view)) return NULL;
    const int rc = aero_particles_step(c, &ps, &view);
    if (rc) return throw_rc(env, rc, c, "aero_particles_step");
    return NULL;
}

/* readParticles(ctx) -> Float32Array[8 * count]; writeParticles(ctx, Float32Array[8 * count]) */
static napi_value js_read_particles(napi_env env, napi_callback_info info) {
    napi_value a[1];
    if (!get_args(env, info, 1, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, (size_t)h->nparts * 8, 4, &p);
    if (!arr) return NULL;
    if (h->nparts) {
        const int rc = aero_particles_read(h->c, (float*)p);
        if (rc) return throw_rc(env, rc, h->c, "aero_particles_read");
    }
    return arr;
}

static napi_value js_write_particles(napi_env env, napi_callback_info info) {
    napi_value a[2];
    if (!get_args(env, info, 2, a)) return NULL;
    handle_t* h = get_handle(env, a[0]);
    if (!h) return NULL;
    size_t n = 0;
    const float* f = (const float*)typed(env, a[1], napi_float32_array, &n);
    if (!f) return NULL;
    if (n != (size_t)h->nparts * 8) {
        napi_throw_range_error(env, "AERO_ARGS", "particles must hold 8 values per allocated particle");
        return NULL;
    }
    const int rc = aero_particles_write(h->c, f);
    if (rc) return throw_rc(env, rc, h->c, "aero_particles_write");
    return NULL;
}

/* streamlines(ctx, stream: Uint8Array[64], view: Float32Array[8], emitters: Float32Array[32]|null)
 *   -> Float32Array[4 * lines * steps] */
static napi_value js_streamlines(napi_env env, napi_callback_info info) {
    napi_value a[4];
    if (!get_args(env, info, 4, a)) return NULL;
    aero_ctx* c = get_ctx(env, a[0]);
    if (!c) return NULL;
    aero_stream st;
    aero_view view;
    if (!get_u64b(env, a[1], &st) || !get_view(env, a[2], &view)) return NULL;
    const float* em = NULL;
    if (!is_nullish(env, a[3])) {
        size_t n = 0;
        em = (const float*)typed(env, a[3], napi_float32_array, &n);
        if (!em) return NULL;
        if (n != 32) {
            napi_throw_range_error(env, "AERO_ARGS", "emitters must hold 8 x (x, y, z, r)");
            return NULL;
        }
    }
    const uint64_t npts = (uint64_t)st.lines * st.steps;
    if (npts > (1ull << 26)) {
        napi_throw_range_error(env, "AERO_ARGS", "more than 2^26 streamline points");
        return NULL;
    }
    void* p = NULL;
    napi_value arr = make_typed(env, napi_float32_array, (size_t)npts * 4, 4, &p);
    if (!arr) return NULL;
    const int rc = aero_streamlines(c, &st, &view, em, npts ? (float*)p : NULL);
    if (rc) return throw_rc(env, rc, c, "aero_streamlines");
    return arr;
}

static napi_value js_version(napi_env env, napi_callback_info info) {
    (void)info;
    napi_value v;
    NAPI_CALL(env, napi_create_uint32(env, aero_version(), &v));
    return v;
}

static napi_value init_module(napi_env env, napi_value exports) {
    const struct {
        const char* name;
        napi_callback fn;
    } fns[] = {
        {"create", js_create},          {"destroy", js_destroy},
        {"init", js_init},              {"voxelize", js_voxelize},
        {"setFlags", js_set_flags},     {"step", js_step},
        {"heatStep", js_heat_step},     {"readForces", js_read_forces},
        {"readFields", js_read_fields}, {"readTemps", js_read_temps},
        {"readFlags", js_read_flags},   {"readPopulations", js_read_populations},
        {"guardCount", js_guard_count}, {"sync", js_sync},
        {"version", js_version},        {"writeFields", js_write_fields},
        {"sampleFlow", js_sample_flow}, {"particlesAlloc", js_particles_alloc},
        {"particlesStep", js_particles_step}, {"readParticles", js_read_particles},
        {"writeParticles", js_write_particles}, {"streamlines", js_streamlines},
    };
    for (size_t i = 0; i < sizeof fns / sizeof fns[0]; i++) {
        napi_value f;
        NAPI_CALL(env, napi_create_function(env, fns[i].name, NAPI_AUTO_LENGTH, fns[i].fn, NULL, &f));
        NAPI_CALL(env, napi_set_named_property(env, exports, fns[i].name, f));
    }
    return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init_module)
