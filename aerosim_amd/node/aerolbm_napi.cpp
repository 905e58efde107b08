// Node N-API addon over the C ABI of include/aerolbm.h. The JavaScript host (lbmEngine.js) keeps the
// reference's LBMEngine / voxelizeMeshToLattice interface (src/solver/lbmEngine.ts, src/geometry/
// voxelizer.ts) and calls these bindings; every failure of the C ABI becomes a thrown JS Error carrying
// aero_last_error().
#define NAPI_VERSION 6
#include <node_api.h>

#include <cstring>
#include <string>
#include <vector>

#include "../../include/aerolbm.h"

namespace {

#define NAPI_CALL(env, call)                                      \
  do {                                                            \
    if ((call) != napi_ok) {                                      \
      napi_throw_error((env), nullptr, "N-API call failed: " #call); \
      return nullptr;                                             \
    }                                                             \
  } while (0)

napi_value throw_aero(napi_env env, aero_ctx* ctx, int rc, const char* what) {
  std::string msg = std::string(what) + " failed (" + std::to_string(rc) + "): " + aero_last_error(ctx);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

// What a JS context handle holds: the C context and its cell count, which sizes every per-cell array the
// bindings accept or return.
struct CtxBox {
  aero_ctx* c;
  int64_t ncells;
};

CtxBox* get_box(napi_env env, napi_value v) {
  void* p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok) return nullptr;
  return static_cast<CtxBox*>(p);
}

aero_ctx* get_ctx(napi_env env, napi_value v) {
  CtxBox* b = get_box(env, v);
  return b ? b->c : nullptr;
}

// Cell count of the handle (-1 if it is not one).
int64_t get_cells(napi_env env, napi_value v) {
  CtxBox* b = get_box(env, v);
  return b && b->c ? b->ncells : -1;
}

void finalize_box(napi_env, void* data, void*) {
  CtxBox* box = static_cast<CtxBox*>(data);
  if (box->c) aero_destroy(box->c);
  delete box;
}

// An optional cell-count argument must equal the context's own (the library reads and writes that many).
bool cells_arg_ok(napi_env env, size_t argc, napi_value* argv, size_t i, int64_t ncells) {
  if (argc <= i) return true;
  napi_valuetype t;
  napi_typeof(env, argv[i], &t);
  if (t == napi_undefined || t == napi_null) return true;
  int64_t n = -1;
  napi_get_value_int64(env, argv[i], &n);
  return n == ncells;
}

double num(napi_env env, napi_value obj, const char* key, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, key, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, key, &v);
  napi_valuetype t;
  napi_typeof(env, v, &t);
  if (t == napi_boolean) {
    bool b;
    napi_get_value_bool(env, v, &b);
    return b ? 1.0 : 0.0;
  }
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

template <typename T>
bool typed(napi_env env, napi_value v, napi_typedarray_type want, T** data, size_t* len) {
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) return false;
  napi_typedarray_type t;
  napi_value ab;
  size_t off;
  void* p;
  napi_get_typedarray_info(env, v, &t, len, &p, &ab, &off);
  if (t != want) return false;
  *data = static_cast<T*>(p);
  return true;
}

napi_value make_typed(napi_env env, napi_typedarray_type t, size_t n, size_t esz, void** data) {
  napi_value ab, arr;
  napi_create_arraybuffer(env, n * esz, data, &ab);
  napi_create_typedarray(env, t, n, ab, 0, &arr);
  return arr;
}

void set_num(napi_env env, napi_value o, const char* k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
}

// createContext(nx, ny, nz, storage = 0, device = 0)
napi_value CreateContext(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t v[5] = {0, 0, 0, 0, 0};
  for (size_t i = 0; i < argc && i < 5; ++i) napi_get_value_int32(env, argv[i], &v[i]);
  aero_grid g{v[0], v[1], v[2], v[3], v[4]};
  aero_ctx* c = nullptr;
  int rc = aero_create(&g, &c);
  if (rc) return throw_aero(env, nullptr, rc, "aero_create");
  CtxBox* box = new CtxBox{c, (int64_t)v[0] * v[1] * v[2]};
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, box, finalize_box, nullptr, &ext));
  return ext;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  void* p = nullptr;
  NAPI_CALL(env, napi_get_value_external(env, argv[0], &p));
  CtxBox* box = static_cast<CtxBox*>(p);
  if (box->c) aero_destroy(box->c);
  box->c = nullptr;
  return nullptr;
}

// setMesh(ctx, Float32Array positions, Float64Array|null world, {length, width, height})
napi_value SetMesh(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  float* pos;
  size_t n;
  if (!c || !typed(env, argv[1], napi_float32_array, &pos, &n)) {
    napi_throw_type_error(env, nullptr, "setMesh(ctx, Float32Array, world, tunnel)");
    return nullptr;
  }
  double* world = nullptr;
  size_t wn = 0;
  napi_valuetype wt;
  napi_typeof(env, argv[2], &wt);
  if (wt != napi_null && wt != napi_undefined && (!typed(env, argv[2], napi_float64_array, &world, &wn) || wn != 16)) {
    napi_throw_type_error(env, nullptr, "world must be a Float64Array(16) or null");
    return nullptr;
  }
  aero_tunnel t{num(env, argv[3], "length", 0), num(env, argv[3], "width", 0), num(env, argv[3], "height", 0)};
  int rc = aero_set_mesh(c, pos, (int32_t)(n / 9), world, &t);
  if (rc) return throw_aero(env, c, rc, "aero_set_mesh");
  return nullptr;
}

// setVoxels(ctx, Uint8Array flags, Int32Array bmap, Int32Array blist, Float32Array q)
napi_value SetVoxels(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  uint8_t* fl;
  int32_t *bm, *bl;
  float* q;
  size_t n1, n2, n3, n4;
  if (!c || !typed(env, argv[1], napi_uint8_array, &fl, &n1) || !typed(env, argv[2], napi_int32_array, &bm, &n2) ||
      !typed(env, argv[3], napi_int32_array, &bl, &n3) || !typed(env, argv[4], napi_float32_array, &q, &n4)) {
    napi_throw_type_error(env, nullptr, "setVoxels(ctx, Uint8Array, Int32Array, Int32Array, Float32Array)");
    return nullptr;
  }
  const int64_t ncells = get_cells(env, argv[0]);
  if ((int64_t)n1 != ncells || (int64_t)n2 != ncells || n3 > (size_t)INT32_MAX || n4 != 18 * n3) {
    napi_throw_type_error(env, nullptr,
                          "setVoxels: flags and bmap need one entry per cell, q 18 per boundary-list entry");
    return nullptr;
  }
  int rc = aero_set_voxels(c, fl, bm, bl, q, (int32_t)n3);
  if (rc) return throw_aero(env, c, rc, "aero_set_voxels");
  return nullptr;
}

// setParams(ctx, {uInlet, nu, smagorinskyCs, groundNoSlip, inflowProfile, z0, zRef, rampSteps, fp16Shift, uInit})
napi_value SetParams(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  aero_params p;
  std::memset(&p, 0, sizeof(p));
  p.u_inlet = num(env, argv[1], "uInlet", 0.08);
  p.nu = num(env, argv[1], "nu", 0.0005);
  p.smagorinsky_cs = num(env, argv[1], "smagorinskyCs", 0.12);
  p.ground_noslip = (int32_t)num(env, argv[1], "groundNoSlip", 0);
  p.inflow_profile = (int32_t)num(env, argv[1], "inflowProfile", 0);
  p.z0 = num(env, argv[1], "z0", 0.5);
  p.z_ref = num(env, argv[1], "zRef", 64.0);
  p.ramp_steps = (int32_t)num(env, argv[1], "rampSteps", 0);
  p.fp16_shift = (int32_t)num(env, argv[1], "fp16Shift", 0);
  p.u_init = num(env, argv[1], "uInit", -1.0);
  int rc = aero_set_params(c, &p);
  if (rc) return throw_aero(env, c, rc, "aero_set_params");
  return nullptr;
}

napi_value Reset(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  int64_t seed = 0;
  double amp = 0;
  if (argc > 1) napi_get_value_int64(env, argv[1], &seed);
  if (argc > 2) napi_get_value_double(env, argv[2], &amp);
  int rc = aero_reset(c, (uint64_t)seed, amp);
  if (rc) return throw_aero(env, c, rc, "aero_reset");
  return nullptr;
}

napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  int32_t n = 1;
  int32_t flags = 0;
  if (argc > 1) napi_get_value_int32(env, argv[1], &n);
  if (argc > 2) {
    // true: AERO_STEP_READOUT (macro fields and populations); a number: the AERO_STEP_* flags themselves
    napi_valuetype t;
    napi_typeof(env, argv[2], &t);
    if (t == napi_boolean) {
      bool ro = false;
      napi_get_value_bool(env, argv[2], &ro);
      flags = ro ? AERO_STEP_READOUT : 0;
    } else if (t == napi_number) {
      napi_get_value_int32(env, argv[2], &flags);
    }
  }
  int rc = aero_step(c, n, flags);
  if (!rc) rc = aero_sync(c);
  if (rc) return throw_aero(env, c, rc, "aero_step");
  int64_t s = 0;
  aero_step_index(c, &s);
  napi_value out;
  napi_create_int64(env, s, &out);
  return out;
}

// readForces(ctx, refs?) -> {fx, fy, fz, mx, my, mz, cutFaceLinks, cdRaw, cd, cl, cm, xcp}
napi_value ReadForces(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  aero_refs r;
  aero_refs* rp = nullptr;
  if (argc > 1) {
    napi_valuetype t;
    napi_typeof(env, argv[1], &t);
    if (t == napi_object) {
      r.frontal_area = num(env, argv[1], "frontalAreaM2", 1.0);
      r.char_length = num(env, argv[1], "characteristicLengthM", 1.0);
      r.tunnel_width = num(env, argv[1], "tunnelWidth", 1.0);
      r.tunnel_height = num(env, argv[1], "tunnelHeight", 1.0);
      r.dx = num(env, argv[1], "dx", 1.0);
      r.u_ref = num(env, argv[1], "uInletLattice", 0.08);
      rp = &r;
    }
  }
  aero_forces f;
  int rc = aero_read_forces(c, rp, &f);
  if (rc) return throw_aero(env, c, rc, "aero_read_forces");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "fx", f.fx); set_num(env, o, "fy", f.fy); set_num(env, o, "fz", f.fz);
  set_num(env, o, "mx", f.mx); set_num(env, o, "my", f.my); set_num(env, o, "mz", f.mz);
  set_num(env, o, "cutFaceLinks", (double)f.cut_face_links);
  set_num(env, o, "cdRaw", f.cd_raw); set_num(env, o, "cd", f.cd); set_num(env, o, "cl", f.cl);
  set_num(env, o, "cm", f.cm); set_num(env, o, "xcp", f.xcp);
  return o;
}

// readMacro(ctx, n?) -> {ux, uy, uz, rho}; n, when given, must be the context's cell count
napi_value ReadMacro(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  const int64_t n = get_cells(env, argv[0]);
  if (!c || !cells_arg_ok(env, argc, argv, 1, n)) {
    napi_throw_type_error(env, nullptr, "readMacro(ctx, n?): n must be the context's cell count");
    return nullptr;
  }
  void* p[4];
  napi_value a[4];
  for (int i = 0; i < 4; ++i) a[i] = make_typed(env, napi_float32_array, (size_t)n, 4, &p[i]);
  int rc = aero_read_macro(c, (float*)p[0], (float*)p[1], (float*)p[2], (float*)p[3]);
  if (rc) return throw_aero(env, c, rc, "aero_read_macro");
  napi_value o;
  napi_create_object(env, &o);
  const char* k[4] = {"ux", "uy", "uz", "rho"};
  for (int i = 0; i < 4; ++i) napi_set_named_property(env, o, k[i], a[i]);
  return o;
}

// readPopulations(ctx, n?) -> Float32Array(19 n); n, when given, must be the context's cell count
napi_value ReadPopulations(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  const int64_t n = get_cells(env, argv[0]);
  if (!c || !cells_arg_ok(env, argc, argv, 1, n)) {
    napi_throw_type_error(env, nullptr, "readPopulations(ctx, n?): n must be the context's cell count");
    return nullptr;
  }
  void* p;
  napi_value a = make_typed(env, napi_float32_array, (size_t)(19 * n), 4, &p);
  int rc = aero_read_populations(c, (float*)p);
  if (rc) return throw_aero(env, c, rc, "aero_read_populations");
  return a;
}

napi_value VoxelCounts(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  int32_t counts[3];
  int rc = aero_read_voxels(c, nullptr, nullptr, nullptr, nullptr, nullptr, counts);
  if (rc) return throw_aero(env, c, rc, "aero_read_voxels");
  double ms = 0;
  aero_voxelize_ms(c, &ms);
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "solidCellCount", counts[0]);
  set_num(env, o, "boundaryCellCount", counts[1]);
  set_num(env, o, "fluidCellCount", counts[2]);
  set_num(env, o, "voxelizationTimeMs", ms);
  return o;
}

// readVoxelArrays(ctx, N) -> {cellFlags, boundaryIndexMap, boundaryCellIndices, boundaryQValues, boundaryMinQ}
napi_value ReadVoxelArrays(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = get_ctx(env, argv[0]);
  const int64_t n = get_cells(env, argv[0]);
  if (!c || !cells_arg_ok(env, argc, argv, 1, n)) {
    napi_throw_type_error(env, nullptr, "readVoxelArrays(ctx, n?): n must be the context's cell count");
    return nullptr;
  }
  int32_t counts[3];
  int rc = aero_read_voxels(c, nullptr, nullptr, nullptr, nullptr, nullptr, counts);
  if (rc) return throw_aero(env, c, rc, "aero_read_voxels");
  const size_t nb = (size_t)counts[1];
  void *pf, *pm, *pl, *pq, *pn;
  napi_value af = make_typed(env, napi_uint8_array, (size_t)n, 1, &pf);
  napi_value am = make_typed(env, napi_int32_array, (size_t)n, 4, &pm);
  napi_value al = make_typed(env, napi_int32_array, nb, 4, &pl);
  napi_value aq = make_typed(env, napi_float32_array, nb * 18, 4, &pq);
  napi_value an = make_typed(env, napi_float32_array, nb, 4, &pn);
  rc = aero_read_voxels(c, (uint8_t*)pf, (int32_t*)pm, nb ? (int32_t*)pl : nullptr, nb ? (float*)pq : nullptr,
                        nb ? (float*)pn : nullptr, counts);
  if (rc) return throw_aero(env, c, rc, "aero_read_voxels");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "cellFlags", af);
  napi_set_named_property(env, o, "boundaryIndexMap", am);
  napi_set_named_property(env, o, "boundaryCellIndices", al);
  napi_set_named_property(env, o, "boundaryQValues", aq);
  napi_set_named_property(env, o, "boundaryMinQ", an);
  return o;
}

// meshIcosphere(level, radius, [cx, cy, cz]) -> Float32Array
napi_value MeshIcosphere(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  int32_t level = 3;
  double radius = 1.0, ctr[3] = {0, 0, 0};
  napi_get_value_int32(env, argv[0], &level);
  napi_get_value_double(env, argv[1], &radius);
  for (uint32_t i = 0; i < 3; ++i) {
    napi_value e;
    napi_get_element(env, argv[2], i, &e);
    napi_get_value_double(env, e, &ctr[i]);
  }
  int32_t nt = 0;
  int rc = aero_mesh_icosphere(level, radius, ctr, nullptr, 0, &nt);
  if (rc) return throw_aero(env, nullptr, rc, "aero_mesh_icosphere");
  void* p;
  napi_value a = make_typed(env, napi_float32_array, (size_t)nt * 9, 4, &p);
  aero_mesh_icosphere(level, radius, ctr, (float*)p, nt, &nt);
  return a;
}

napi_value MeshStats(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  float* pos;
  size_t n;
  if (!typed(env, argv[0], napi_float32_array, &pos, &n)) {
    napi_throw_type_error(env, nullptr, "meshStats(Float32Array, world?)");
    return nullptr;
  }
  double* world = nullptr;
  size_t wn = 0;
  if (argc > 1) typed(env, argv[1], napi_float64_array, &world, &wn);
  double out[6];
  int rc = aero_mesh_stats(pos, (int32_t)(n / 9), wn == 16 ? world : nullptr, out);
  if (rc) return throw_aero(env, nullptr, rc, "aero_mesh_stats");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "frontalArea", out[0]);
  set_num(env, o, "sizeX", out[1]); set_num(env, o, "sizeY", out[2]); set_num(env, o, "sizeZ", out[3]);
  set_num(env, o, "volume", out[4]);
  set_num(env, o, "openEdgeCount", out[5]);
  return o;
}

int32_t get_i32(napi_env env, napi_value v, int32_t d) {
  int32_t x = d;
  napi_get_value_int32(env, v, &x);
  return x;
}
double get_f64(napi_env env, napi_value v, double d) {
  double x = d;
  napi_get_value_double(env, v, &x);
  return x;
}

// readSliceRGBA(ctx, axis, fraction, field, rules, uRef) -> {width, height, rgba: Uint8Array}
napi_value ReadSliceRGBA(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = argc >= 6 ? get_ctx(env, argv[0]) : nullptr;
  if (!c) {
    napi_throw_type_error(env, nullptr, "readSliceRGBA(ctx, axis, fraction, field, rules, uRef)");
    return nullptr;
  }
  const int32_t axis = get_i32(env, argv[1], 0), field = get_i32(env, argv[3], 0), rules = get_i32(env, argv[4], 0);
  const double frac = get_f64(env, argv[2], 0.5), uref = get_f64(env, argv[5], -1.0);
  int32_t wh[2] = {0, 0};
  int rc = aero_read_slice_rgba(c, axis, frac, field, rules, uref, nullptr, wh);
  if (rc) return throw_aero(env, c, rc, "aero_read_slice_rgba");
  void* p;
  napi_value a = make_typed(env, napi_uint8_array, (size_t)wh[0] * wh[1] * 4, 1, &p);
  rc = aero_read_slice_rgba(c, axis, frac, field, rules, uref, (uint8_t*)p, wh);
  if (rc) return throw_aero(env, c, rc, "aero_read_slice_rgba");
  napi_value o;
  napi_create_object(env, &o);
  set_num(env, o, "width", wh[0]);
  set_num(env, o, "height", wh[1]);
  napi_set_named_property(env, o, "rgba", a);
  return o;
}

// qCriterion(ctx, n, dx) -> {qValues, omegaMag, omegaVec, maxQ, maxOmega} (computeQCriterion, vortexIso.ts:22)
napi_value QCriterion(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = argc >= 3 ? get_ctx(env, argv[0]) : nullptr;
  const int64_t n = c ? get_cells(env, argv[0]) : 0;
  if (!c || !cells_arg_ok(env, argc, argv, 1, n)) {
    napi_throw_type_error(env, nullptr, "qCriterion(ctx, n, dx): n must be the context's cell count");
    return nullptr;
  }
  const double dx = get_f64(env, argv[2], 1.0);
  void *pq, *po, *pv;
  napi_value q = make_typed(env, napi_float32_array, (size_t)n, 4, &pq);
  napi_value om = make_typed(env, napi_float32_array, (size_t)n, 4, &po);
  napi_value ov = make_typed(env, napi_float32_array, (size_t)n * 3, 4, &pv);
  double mx[2] = {0, 0};
  int rc = aero_read_qcriterion(c, dx, (float*)pq, (float*)po, (float*)pv, mx);
  if (rc) return throw_aero(env, c, rc, "aero_read_qcriterion");
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "qValues", q);
  napi_set_named_property(env, o, "omegaMag", om);
  napi_set_named_property(env, o, "omegaVec", ov);
  set_num(env, o, "maxQ", mx[0]);
  set_num(env, o, "maxOmega", mx[1]);
  return o;
}

// probe(ctx, Float64Array xyz, Uint8Array|null onSurface, {originX, originY, originZ, dx, windSpeedMps,
//       uInletLattice, fluidDensity}) -> Float64Array (13 per point)
napi_value Probe(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  aero_ctx* c = argc >= 4 ? get_ctx(env, argv[0]) : nullptr;
  double* xyz;
  size_t n3;
  if (!c || !typed(env, argv[1], napi_float64_array, &xyz, &n3) || n3 % 3) {
    napi_throw_type_error(env, nullptr, "probe(ctx, Float64Array xyz, onSurface, params)");
    return nullptr;
  }
  uint8_t* surf = nullptr;
  size_t ns = 0;
  if (typed(env, argv[2], napi_uint8_array, &surf, &ns) && ns != n3 / 3) {
    napi_throw_type_error(env, nullptr, "probe: onSurface needs one byte per point");
    return nullptr;
  }
  aero_probe_params pp;
  pp.origin[0] = num(env, argv[3], "originX", 0.0);
  pp.origin[1] = num(env, argv[3], "originY", 0.0);
  pp.origin[2] = num(env, argv[3], "originZ", 0.0);
  pp.dx = num(env, argv[3], "dx", 1.0);
  pp.wind_speed_mps = num(env, argv[3], "windSpeedMps", 25.0);
  pp.u_inlet_lattice = num(env, argv[3], "uInletLattice", 0.085);
  pp.fluid_density = num(env, argv[3], "fluidDensity", 1.204);
  void* p;
  napi_value a = make_typed(env, napi_float64_array, n3 / 3 * 13, 8, &p);
  int rc = aero_probe(c, &pp, xyz, surf, (int32_t)(n3 / 3), (double*)p);
  if (rc) return throw_aero(env, c, rc, "aero_probe");
  return a;
}

// parseSTL(ArrayBuffer | TypedArray | Buffer) -> Float32Array of triangle positions (stlLoader.ts:9-95)
napi_value ParseSTL(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  const uint8_t* bytes = nullptr;
  size_t len = 0;
  bool is_ab = false, is_ta = false;
  if (argc >= 1) {
    napi_is_arraybuffer(env, argv[0], &is_ab);
    napi_is_typedarray(env, argv[0], &is_ta);
  }
  if (is_ab) {
    void* d;
    NAPI_CALL(env, napi_get_arraybuffer_info(env, argv[0], &d, &len));
    bytes = (const uint8_t*)d;
  } else if (is_ta) {
    napi_typedarray_type t;
    size_t n, off;
    void* d;
    napi_value ab;
    NAPI_CALL(env, napi_get_typedarray_info(env, argv[0], &t, &n, &d, &ab, &off));
    size_t el = (t == napi_int8_array || t == napi_uint8_array || t == napi_uint8_clamped_array) ? 1
                : (t == napi_int16_array || t == napi_uint16_array)                                 ? 2
                : (t == napi_float64_array || t == napi_bigint64_array || t == napi_biguint64_array) ? 8
                                                                                                    : 4;
    bytes = (const uint8_t*)d;
    len = n * el;
  } else {
    napi_throw_type_error(env, nullptr, "parseSTL(ArrayBuffer | TypedArray)");
    return nullptr;
  }
  int64_t nt = 0;
  int rc = aero_stl_parse(bytes, (int64_t)len, nullptr, 0, &nt);
  if (rc) return throw_aero(env, nullptr, rc, "aero_stl_parse");
  void* p;
  napi_value a = make_typed(env, napi_float32_array, (size_t)nt * 9, 4, &p);
  if (nt) {
    rc = aero_stl_parse(bytes, (int64_t)len, (float*)p, nt, &nt);
    if (rc) return throw_aero(env, nullptr, rc, "aero_stl_parse");
  }
  return a;
}

// normalizeAndCenterGeometry(Float32Array) -> {position, autoScaledFromMm}; the array is changed in place
// (stlLoader.ts:101-127)
napi_value NormalizeSTL(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  NAPI_CALL(env, napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  float* pos;
  size_t n;
  if (argc < 1 || !typed(env, argv[0], napi_float32_array, &pos, &n)) {
    napi_throw_type_error(env, nullptr, "normalizeAndCenterGeometry(Float32Array)");
    return nullptr;
  }
  int32_t scaled = 0;
  int rc = aero_stl_normalize(pos, (int64_t)(n / 9), &scaled);
  if (rc) return throw_aero(env, nullptr, rc, "aero_stl_normalize");
  napi_value o, b;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "position", argv[0]);
  napi_get_boolean(env, scaled != 0, &b);
  napi_set_named_property(env, o, "autoScaledFromMm", b);
  return o;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char* name;
    napi_callback cb;
  } fns[] = {{"createContext", CreateContext}, {"destroy", Destroy},       {"setMesh", SetMesh},
             {"setVoxels", SetVoxels},         {"voxelCounts", VoxelCounts}, {"readVoxelArrays", ReadVoxelArrays},
             {"setParams", SetParams},         {"reset", Reset},           {"step", Step},
             {"readForces", ReadForces},       {"readMacro", ReadMacro},   {"readPopulations", ReadPopulations},
             {"meshIcosphere", MeshIcosphere}, {"meshStats", MeshStats},
             {"parseSTL", ParseSTL},           {"normalizeAndCenterGeometry", NormalizeSTL},
             {"readSliceRGBA", ReadSliceRGBA}, {"qCriterion", QCriterion}, {"probe", Probe}};
  for (auto& f : fns) {
    napi_value fn;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.cb, nullptr, &fn);
    napi_set_named_property(env, exports, f.name, fn);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
