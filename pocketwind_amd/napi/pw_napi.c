/*
 * pw_napi.c -- Node-API binding of libpocketwind's solver entry points (include/pw.h), the
 * boundary a TypeScript host calls in place of GpuSolver's WebGPU work
 * (src/lbm/gpu-solver.ts:135-465, gpu-force.ts:175-210). Plain C against node_api.h; built by
 * pocketwind_amd/build.py into pw_napi.node next to this file.
 *
 * Errors map as the reference raises them: PW_EINVAL throws a RangeError with the library's
 * message (gpu-solver.ts:98-105, geometry-contract.ts), PW_ENOMEM returns
 * {kind: 'allocation-failure', message} (gpu-resources.ts:3-27), PW_EBUSY from readForce returns
 * 'busy' (gpu-force.ts:192-205), anything else throws an Error.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pw.h"

/* A solver handle as JS holds it: destroy() clears `s`, the finalizer destroys what is left. */
typedef struct {
  pw_solver *s;
  size_t cells;
} handle_box;

#define CHECK(call)                                                        \
  do {                                                                     \
    if ((call) != napi_ok) {                                               \
      napi_throw_error(env, NULL, "pocketwind addon: Node-API call failed"); \
      return NULL;                                                         \
    }                                                                      \
  } while (0)

static napi_value throw_status(napi_env env, pw_status st) {
  const char *msg = pw_last_error_message();
  if (!msg || !*msg) msg = "pocketwind call failed";
  if (st == PW_EINVAL)
    napi_throw_range_error(env, NULL, msg);
  else
    napi_throw_error(env, NULL, msg);
  return NULL;
}

static napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

static int get_args(napi_env env, napi_callback_info info, size_t want, napi_value *argv) {
  size_t argc = want;
  if (napi_get_cb_info(env, info, &argc, argv, NULL, NULL) != napi_ok) return 0;
  for (size_t k = argc; k < want; ++k) napi_get_undefined(env, &argv[k]);
  return 1;
}

/* A named number property; `fallback` when it is absent (undefined). Returns 0 on a wrong type. */
static int number_prop(napi_env env, napi_value obj, const char *name, double fallback, double *out) {
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_undefined) {
    *out = fallback;
    return 1;
  }
  if (t != napi_number) return 0;
  return napi_get_value_double(env, v, out) == napi_ok;
}

static int bool_prop(napi_env env, napi_value obj, const char *name, int *out) {
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_undefined) {
    *out = 0;
    return 1;
  }
  bool b = false;
  if (t != napi_boolean || napi_get_value_bool(env, v, &b) != napi_ok) return 0;
  *out = b;
  return 1;
}

/* A string property among `names` (index returned in *out); `fallback` when absent. */
static int choice_prop(napi_env env, napi_value obj, const char *name, const char *const *names, int n,
                       int fallback, uint32_t *out) {
  napi_value v;
  napi_valuetype t;
  if (napi_get_named_property(env, obj, name, &v) != napi_ok || napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_undefined) {
    *out = (uint32_t)fallback;
    return 1;
  }
  char buf[32];
  size_t len = 0;
  if (t != napi_string || napi_get_value_string_utf8(env, v, buf, sizeof buf, &len) != napi_ok) return 0;
  for (int k = 0; k < n; ++k)
    if (strcmp(buf, names[k]) == 0) {
      *out = (uint32_t)k;
      return 1;
    }
  return 0;
}

/* {nx, ny, nz} -> pw_layout; dimensions must be positive integers below 2^32. */
static int read_layout(napi_env env, napi_value obj, pw_layout *L) {
  double d[3];
  const char *names[3] = {"nx", "ny", "nz"};
  for (int k = 0; k < 3; ++k) {
    if (!number_prop(env, obj, names[k], -1, &d[k]) || !(d[k] >= 1) || d[k] > 4294967295.0 || d[k] != (double)(uint64_t)d[k])
      return 0;
  }
  L->nx = (uint32_t)d[0];
  L->ny = (uint32_t)d[1];
  L->nz = (uint32_t)d[2];
  return 1;
}

/* SolverParams (src/lbm/gpu-solver.ts:19-30) -> pw_params, with this library's extra fields. */
static int read_params(napi_env env, napi_value obj, pw_params *p) {
  static const char *const ops[] = {"trt", "reg"};
  static const char *const walls[] = {"bb", "bouzidi"};
  static const char *const precisions[] = {"f32", "f16", "bf16"};
  static const char *const models[] = {"law", "learned"};
  static const char *const sides[] = {"periodic", "free-slip"};
  static const char *const outlets[] = {"equilibrium", "zero-gradient"};
  memset(p, 0, sizeof *p);
  int les = 0;
  double ramp = 0, slip_floor = 0;
  if (!number_prop(env, obj, "tau0", -1, &p->tau0) || !number_prop(env, obj, "uIn", 0, &p->u_in) ||
      !bool_prop(env, obj, "les", &les) || !choice_prop(env, obj, "operator", ops, 2, 0, &p->op) ||
      !choice_prop(env, obj, "walls", walls, 2, 0, &p->walls) ||
      !choice_prop(env, obj, "precision", precisions, 3, 0, &p->storage) ||
      !choice_prop(env, obj, "wallModel", models, 2, 0, &p->wall_model) ||
      !choice_prop(env, obj, "sides", sides, 2, 0, &p->sides) ||
      !choice_prop(env, obj, "outlet", outlets, 2, 0, &p->outlet) || !number_prop(env, obj, "rampSteps", 0, &ramp) ||
      !number_prop(env, obj, "slipFloor", 0, &slip_floor) ||
      !(ramp >= 0) || ramp > 4294967295.0 || ramp != (double)(uint64_t)ramp)
    return 0;
  p->les = (uint32_t)les;
  p->ramp_steps = (uint32_t)ramp;
  p->slip_floor = (float)slip_floor;
  return 1;
}

static int u32_array(napi_env env, napi_value v, uint32_t **data, size_t *n) {
  bool is_typed = false;
  if (napi_is_typedarray(env, v, &is_typed) != napi_ok || !is_typed) return 0;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  void *ptr;
  if (napi_get_typedarray_info(env, v, &t, n, &ptr, &ab, &off) != napi_ok || t != napi_uint32_array) return 0;
  *data = (uint32_t *)ptr;
  return 1;
}

static pw_solver *solver_of(napi_env env, napi_value v) {
  handle_box *box = NULL;
  if (napi_get_value_external(env, v, (void **)&box) != napi_ok || !box) {
    napi_throw_type_error(env, NULL, "Expected a solver handle.");
    return NULL;
  }
  if (!box->s) {
    napi_throw_error(env, NULL, "The solver has been destroyed.");
    return NULL;
  }
  return box->s;
}

static void finalize_box(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  handle_box *box = (handle_box *)data;
  if (box->s) pw_solver_destroy(box->s);
  free(box);
}

static napi_value js_abi_version(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value r;
  CHECK(napi_create_int32(env, pw_abi_version(), &r));
  return r;
}

static napi_value js_device_count(napi_env env, napi_callback_info info) {
  (void)info;
  napi_value r;
  CHECK(napi_create_int32(env, pw_device_count(), &r));
  return r;
}

/* packParams(layout, params, writeMacro, step) -> ArrayBuffer(48) (gpu-solver.ts:80-116). */
static napi_value js_pack_params(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  pw_layout L;
  pw_params p;
  if (!read_layout(env, argv[0], &L)) {
    napi_throw_range_error(env, NULL, "Grid dimensions must be positive safe integers.");
    return NULL;
  }
  if (!read_params(env, argv[1], &p)) {
    napi_throw_range_error(env, NULL, "Invalid solver parameters.");
    return NULL;
  }
  bool macro = false;
  double step = 0;
  napi_valuetype t;
  if (napi_typeof(env, argv[2], &t) == napi_ok && t == napi_boolean) napi_get_value_bool(env, argv[2], &macro);
  if (napi_typeof(env, argv[3], &t) == napi_ok && t == napi_number) napi_get_value_double(env, argv[3], &step);
  if (!(step >= 0) || step > 9007199254740991.0 || step != (double)(uint64_t)step) {
    napi_throw_range_error(env, NULL, "Solver step must be a non-negative safe integer.");
    return NULL;
  }
  void *bytes = NULL;
  napi_value ab;
  CHECK(napi_create_arraybuffer(env, PW_PARAMS_BYTES, &bytes, &ab));
  pw_status st = pw_pack_params(&L, &p, macro, (uint64_t)step, bytes, PW_PARAMS_BYTES);
  if (st) return throw_status(env, st);
  return ab;
}

/* create(layout, geometry: Uint32Array, boundaryCells: Uint32Array, params) -> handle, or
 * {kind: 'allocation-failure', message} (gpu-solver.ts:158-335). */
static napi_value js_create(napi_env env, napi_callback_info info) {
  napi_value argv[4];
  if (!get_args(env, info, 4, argv)) return NULL;
  pw_layout L;
  pw_params p;
  uint32_t *geo = NULL, *bnd = NULL;
  size_t ng = 0, nb = 0;
  if (!read_layout(env, argv[0], &L)) {
    napi_throw_range_error(env, NULL, "Grid dimensions must be positive safe integers.");
    return NULL;
  }
  if (!u32_array(env, argv[1], &geo, &ng) || !u32_array(env, argv[2], &bnd, &nb)) {
    napi_throw_range_error(env, NULL, "Geometry and boundary cells must be Uint32Arrays.");
    return NULL;
  }
  if (!read_params(env, argv[3], &p)) {
    napi_throw_range_error(env, NULL, "Invalid solver parameters.");
    return NULL;
  }
  pw_solver *s = NULL;
  pw_status st = pw_solver_create(&L, geo, ng, bnd, nb, &p, &s);
  if (st == PW_ENOMEM) {
    napi_value r, kind, msg;
    CHECK(napi_create_object(env, &r));
    CHECK(napi_create_string_utf8(env, "allocation-failure", NAPI_AUTO_LENGTH, &kind));
    CHECK(napi_create_string_utf8(env, pw_last_error_message(), NAPI_AUTO_LENGTH, &msg));
    CHECK(napi_set_named_property(env, r, "kind", kind));
    CHECK(napi_set_named_property(env, r, "message", msg));
    return r;
  }
  if (st) return throw_status(env, st);
  handle_box *box = (handle_box *)malloc(sizeof *box);
  if (!box) {
    pw_solver_destroy(s);
    napi_throw_error(env, NULL, "Host allocation failed.");
    return NULL;
  }
  box->s = s;
  box->cells = (size_t)L.nx * L.ny * L.nz;
  napi_value ext;
  if (napi_create_external(env, box, finalize_box, NULL, &ext) != napi_ok) {
    pw_solver_destroy(s);
    free(box);
    napi_throw_error(env, NULL, "pocketwind addon: Node-API call failed");
    return NULL;
  }
  return ext;
}

/* step(handle, count, writeMacroOnLast) (gpu-solver.ts:378-395). */
static napi_value js_step(napi_env env, napi_callback_info info) {
  napi_value argv[3];
  if (!get_args(env, info, 3, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  double count = -1;
  bool macro = false;
  napi_valuetype t;
  if (napi_typeof(env, argv[1], &t) == napi_ok && t == napi_number) napi_get_value_double(env, argv[1], &count);
  if (napi_typeof(env, argv[2], &t) == napi_ok && t == napi_boolean) napi_get_value_bool(env, argv[2], &macro);
  if (!(count >= 0) || count > 4294967295.0 || count != (double)(uint64_t)count) {
    napi_throw_range_error(env, NULL, "Step count must be a non-negative safe integer.");
    return NULL;
  }
  pw_status st = pw_solver_step(s, (uint32_t)count, macro);
  if (st) return throw_status(env, st);
  return undefined(env);
}

static napi_value js_encode_force(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  pw_status st = pw_solver_encode_force(s);
  if (st) return throw_status(env, st);
  return undefined(env);
}

/* readForce(handle) -> {step, fx, fy, fz} | 'busy' (gpu-force.ts:192-205). */
static napi_value js_read_force(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  pw_force_sample f;
  pw_status st = pw_solver_read_force(s, &f);
  napi_value r;
  if (st == PW_EBUSY) {
    CHECK(napi_create_string_utf8(env, "busy", NAPI_AUTO_LENGTH, &r));
    return r;
  }
  if (st) return throw_status(env, st);
  napi_value step, fx, fy, fz;
  CHECK(napi_create_object(env, &r));
  CHECK(napi_create_double(env, (double)f.step, &step));
  CHECK(napi_create_double(env, f.fx, &fx));
  CHECK(napi_create_double(env, f.fy, &fy));
  CHECK(napi_create_double(env, f.fz, &fz));
  CHECK(napi_set_named_property(env, r, "step", step));
  CHECK(napi_set_named_property(env, r, "fx", fx));
  CHECK(napi_set_named_property(env, r, "fy", fy));
  CHECK(napi_set_named_property(env, r, "fz", fz));
  return r;
}

static napi_value js_reset(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  pw_status st = pw_solver_reset(s);
  if (st) return throw_status(env, st);
  return undefined(env);
}

static napi_value js_step_count(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  napi_value r;
  CHECK(napi_create_double(env, (double)pw_solver_step_count(s), &r));
  return r;
}

static napi_value js_destroy(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  handle_box *box = NULL;
  if (napi_get_value_external(env, argv[0], (void **)&box) != napi_ok || !box) {
    napi_throw_type_error(env, NULL, "Expected a solver handle.");
    return NULL;
  }
  if (box->s) pw_solver_destroy(box->s);
  box->s = NULL;
  return undefined(env);
}

/* readPopulations(handle) -> ArrayBuffer of the latest state's 19 planes in storage format. */
static napi_value js_read_populations(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  const size_t bytes = pw_solver_population_bytes(s);
  void *data = NULL;
  napi_value ab;
  CHECK(napi_create_arraybuffer(env, bytes, &data, &ab));
  pw_status st = pw_solver_read_populations(s, data, bytes);
  if (st) return throw_status(env, st);
  return ab;
}

/* writePopulations(handle, typedArray) -- the test-side writeBuffer (force.spec.ts:145-153). */
static napi_value js_write_populations(napi_env env, napi_callback_info info) {
  napi_value argv[2];
  if (!get_args(env, info, 2, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  bool is_typed = false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off, n;
  void *ptr;
  if (napi_is_typedarray(env, argv[1], &is_typed) != napi_ok || !is_typed ||
      napi_get_typedarray_info(env, argv[1], &t, &n, &ptr, &ab, &off) != napi_ok) {
    napi_throw_range_error(env, NULL, "Population input must be a typed array.");
    return NULL;
  }
  size_t elem = 1;
  switch (t) {
    case napi_int16_array:
    case napi_uint16_array: elem = 2; break;
    case napi_int32_array:
    case napi_uint32_array:
    case napi_float32_array: elem = 4; break;
    case napi_float64_array:
    case napi_bigint64_array:
    case napi_biguint64_array: elem = 8; break;
    default: elem = 1;
  }
  pw_status st = pw_solver_write_populations(s, ptr, n * elem);
  if (st) return throw_status(env, st);
  return undefined(env);
}

/* readMacro(handle) -> Float32Array of (rho, ux, uy, uz) per cell, as last written. */
static napi_value js_read_macro(napi_env env, napi_callback_info info) {
  napi_value argv[1];
  if (!get_args(env, info, 1, argv)) return NULL;
  pw_solver *s = solver_of(env, argv[0]);
  if (!s) return NULL;
  handle_box *box = NULL;
  napi_get_value_external(env, argv[0], (void **)&box);
  void *data = NULL;
  napi_value ab, arr;
  CHECK(napi_create_arraybuffer(env, box->cells * 16, &data, &ab));
  pw_status st = pw_solver_read_macro(s, (float *)data, box->cells * 4);
  if (st) return throw_status(env, st);
  CHECK(napi_create_typedarray(env, napi_float32_array, box->cells * 4, ab, 0, &arr));
  return arr;
}

static napi_value init(napi_env env, napi_value exports) {
  struct {
    const char *name;
    napi_callback fn;
  } fns[] = {{"abiVersion", js_abi_version},       {"deviceCount", js_device_count},
             {"packParams", js_pack_params},       {"create", js_create},
             {"step", js_step},                    {"encodeForce", js_encode_force},
             {"readForce", js_read_force},         {"reset", js_reset},
             {"stepCount", js_step_count},         {"destroy", js_destroy},
             {"readPopulations", js_read_populations}, {"writePopulations", js_write_populations},
             {"readMacro", js_read_macro}};
  for (size_t k = 0; k < sizeof fns / sizeof fns[0]; ++k) {
    napi_value f;
    if (napi_create_function(env, fns[k].name, NAPI_AUTO_LENGTH, fns[k].fn, NULL, &f) != napi_ok ||
        napi_set_named_property(env, exports, fns[k].name, f) != napi_ok)
      return NULL;
  }
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
