// N-API (version 8) addon over the C ABI of include/vt.h: the native half of the JavaScript TwinRenderer in
// js/renderer.mjs. Blocking calls (sync, readback, stats, autotune, backward) run as napi_async_work and return
// Promises, as the reference's queue.onSubmittedWorkDone / mapAsync do (gpu/renderer.js:986-996, 1191-1231);
// enqueue-only calls (upload, setCamera, render, sensor, updateMaterials, updateLights) return synchronously and
// throw Error on a bad status. Calls into libvt.so are serialised by one lock (the pool's jobs included).
#define NAPI_VERSION 8
#include <node_api.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vt.h"

namespace {

// the C ABI wants one host thread per context at a time; the async jobs run on libuv's pool, so every call into
// libvt.so from this addon (main thread or pool) holds this lock
std::mutex g_mu;

#define CHECK_NAPI(call)                                                   \
  do {                                                                     \
    if ((call) != napi_ok) {                                               \
      napi_throw_error(env, nullptr, "N-API call failed: " #call);         \
      return nullptr;                                                      \
    }                                                                      \
  } while (0)

napi_value throw_vt(napi_env env, vt_ctx *c, int rc, const char *what) {
  std::string m = std::string(what) + " failed (" + std::to_string(rc) + "): " + (c ? vt_last_error(c) : "");
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

// What the JS external holds: the context and the async jobs queued on it. destroy() stops new calls at once and
// frees the context when the last queued job has completed (each job's promise still settles); a handle whose
// external is collected goes the same way. The flags and the count change on the main thread only (the calls, the
// jobs' completions and the finalizer run there); the pool's job_exec reads c under g_mu.
struct Handle {
  vt_ctx *c = nullptr;
  int pending = 0;          // queued jobs that will use c
  bool destroyed = false;   // destroy() was called
  bool collected = false;   // the JS external was garbage-collected
  uint32_t baked_nodes = 0; // nodes of the baked table last uploaded (bakedJacobi's size check)
};

// with g_mu held
void settle(Handle *h) {
  if ((h->destroyed || h->collected) && h->pending == 0 && h->c) {
    vt_destroy(h->c);
    h->c = nullptr;
  }
  if (h->collected && h->pending == 0) delete h;
}

void finalize_handle(napi_env, void *data, void *) {
  Handle *h = static_cast<Handle *>(data);
  std::lock_guard<std::mutex> lk(g_mu);
  h->collected = true;
  settle(h);
}

Handle *get_handle(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "first argument must be a vt context");
    return nullptr;
  }
  return static_cast<Handle *>(p);
}

vt_ctx *get_ctx(napi_env env, napi_value v) {
  Handle *h = get_handle(env, v);
  if (!h) return nullptr;
  if (h->destroyed || !h->c) {
    napi_throw_error(env, nullptr, "the vt context was destroyed");
    return nullptr;
  }
  return h->c;
}

bool get_prop(napi_env env, napi_value obj, const char *name, napi_value *out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

double num(napi_env env, napi_value obj, const char *name, double dflt) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return dflt;
  double d = dflt;
  napi_get_value_double(env, v, &d);
  return d;
}

// typed array view of a property: data pointer and element count (nullptr when absent)
template <typename T>
const T *typed(napi_env env, napi_value obj, const char *name, size_t *count, napi_typedarray_type want) {
  napi_value v;
  *count = 0;
  if (!get_prop(env, obj, name, &v)) return nullptr;
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    std::string m = std::string(name) + " must be a typed array";
    napi_throw_type_error(env, nullptr, m.c_str());
    return nullptr;
  }
  napi_typedarray_type tt;
  size_t len;
  void *data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &tt, &len, &data, &ab, &off);
  if (tt != want) {
    std::string m = std::string(name) + (want == napi_float32_array ? " must be a Float32Array"
                                         : want == napi_uint16_array ? " must be a Uint16Array" : " must be a Uint32Array");
    napi_throw_type_error(env, nullptr, m.c_str());
    return nullptr;
  }
  *count = len;
  return static_cast<const T *>(data);
}

bool float_list(napi_env env, napi_value obj, const char *name, float *dst, uint32_t n) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return false;
  for (uint32_t i = 0; i < n; i++) {
    napi_value e;
    double d = 0;
    if (napi_get_element(env, v, i, &e) != napi_ok || napi_get_value_double(env, e, &d) != napi_ok) return false;
    dst[i] = (float)d;
  }
  return true;
}

napi_value Create(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_create_info ci{0, 0};
  if (argc > 0) {
    int32_t dev = 0;
    napi_get_value_int32(env, argv[0], &dev);
    ci.device = dev;
  }
  vt_ctx *c = nullptr;
  int rc = vt_create(&ci, &c);
  if (rc) return throw_vt(env, nullptr, rc, "vt_create (no HIP device?)");
  Handle *h = new Handle();
  h->c = c;
  napi_value out;
  if (napi_create_external(env, h, finalize_handle, nullptr, &out) != napi_ok) {
    vt_destroy(c);
    delete h;
    napi_throw_error(env, nullptr, "napi_create_external failed");
    return nullptr;
  }
  return out;
}

napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  Handle *h = get_handle(env, argv[0]);
  if (!h) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  h->destroyed = true;      // a second destroy() is a no-op
  settle(h);
  return nullptr;
}

napi_value BuildBVH(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  bool a = false, b = false;
  napi_is_typedarray(env, argv[0], &a);
  napi_is_typedarray(env, argv[1], &b);
  if (argc < 2 || !a || !b) {
    napi_throw_type_error(env, nullptr, "buildBVH(pos: Float32Array, idx: Uint32Array)");
    return nullptr;
  }
  napi_typedarray_type t0, t1;
  size_t n0, n1, o0, o1;
  void *d0, *d1;
  napi_value ab0, ab1;
  napi_get_typedarray_info(env, argv[0], &t0, &n0, &d0, &ab0, &o0);
  napi_get_typedarray_info(env, argv[1], &t1, &n1, &d1, &ab1, &o1);
  if (t0 != napi_float32_array || t1 != napi_uint32_array) {
    napi_throw_type_error(env, nullptr, "buildBVH(pos: Float32Array, idx: Uint32Array)");
    return nullptr;
  }
  const uint32_t ntri = (uint32_t)(n1 / 3), cap = ntri + 2;
  std::vector<float> nodes((size_t)cap * 16);
  std::vector<uint32_t> order(ntri ? ntri : 1);
  uint32_t count = 0;
  int rc = vt_build_bvh((const float *)d0, (uint32_t)(n0 / 3), (const uint32_t *)d1, ntri, 4, 16, nodes.data(), cap,
                        &count, order.data());
  if (rc) return throw_vt(env, nullptr, rc, "vt_build_bvh");
  napi_value res, nab, nta, oab, ota, ncount;
  void *dst;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)count * 64, &dst, &nab));
  std::memcpy(dst, nodes.data(), (size_t)count * 64);
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, (size_t)count * 16, nab, 0, &nta));
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)ntri * 4, &dst, &oab));
  if (ntri) std::memcpy(dst, order.data(), (size_t)ntri * 4);
  CHECK_NAPI(napi_create_typedarray(env, napi_uint32_array, ntri, oab, 0, &ota));
  CHECK_NAPI(napi_create_object(env, &res));
  CHECK_NAPI(napi_set_named_property(env, res, "nodes", nta));
  CHECK_NAPI(napi_set_named_property(env, res, "order", ota));
  CHECK_NAPI(napi_create_uint32(env, count, &ncount));
  CHECK_NAPI(napi_set_named_property(env, res, "nodeCount", ncount));
  return res;
}

napi_value UploadScene(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  size_t n;
  vt_scene s{};
  s.nodes = typed<float>(env, argv[1], "nodes", &n, napi_float32_array);
  s.node_count = (uint32_t)(n / 16);
  s.tris = typed<float>(env, argv[1], "tris", &n, napi_float32_array);
  s.tri_count = (uint32_t)(n / 16);
  s.verts = typed<float>(env, argv[1], "verts", &n, napi_float32_array);
  s.vert_count = (uint32_t)(n / 8);
  s.meshes = typed<uint32_t>(env, argv[1], "meshes", &n, napi_uint32_array);
  s.mesh_count = (uint32_t)(n / 4);
  s.materials = typed<float>(env, argv[1], "materials", &n, napi_float32_array);
  s.material_count = (uint32_t)(n / 16);
  s.lights = typed<float>(env, argv[1], "lights", &n, napi_float32_array);
  s.light_count = (uint32_t)(n / 4);
  s.emit = typed<float>(env, argv[1], "emit", &n, napi_float32_array);
  s.emit_count = (uint32_t)(n / 16);
  s.emit_area = (float)num(env, argv[1], "emitArea", 0);
  s.tri_body = typed<uint32_t>(env, argv[1], "triBody", &n, napi_uint32_array);
  s.body_count = s.tri_body ? (uint32_t)num(env, argv[1], "bodyCount", 0) : 0;
  s.shadow_skip = (uint32_t)num(env, argv[1], "shadowSkip", 0);
  // emitBody: the body the light rides on (its frame holds the emit records); absent or negative: static
  const double eb = num(env, argv[1], "emitBody", -1);
  s.emit_posed = eb >= 0 ? 1u : 0u;
  s.emit_body = eb >= 0 ? (uint32_t)eb : 0u;
  if (s.tri_body && n != s.tri_count) {
    napi_throw_range_error(env, nullptr, "triBody needs one entry per triangle");
    return nullptr;
  }
  // uvs (optional): 2 floats per vertex, for VT_MESH_TEXTURED meshes
  s.uvs = typed<float>(env, argv[1], "uvs", &n, napi_float32_array);
  if (s.uvs && n != (size_t)s.vert_count * 2) {
    napi_throw_range_error(env, nullptr, "uvs needs two floats per vertex");
    return nullptr;
  }
  bool pending = false;
  napi_is_exception_pending(env, &pending);
  if (pending) return nullptr;
  int rc = vt_upload_scene(c, &s);
  if (rc) return throw_vt(env, c, rc, "vt_upload_scene");
  return nullptr;
}

napi_value SetCamera(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t slot = 0;
  napi_get_value_uint32(env, argv[1], &slot);
  vt_camera k{};
  if (!float_list(env, argv[2], "R", k.R, 9) || !float_list(env, argv[2], "pos", k.pos, 3) ||
      !float_list(env, argv[2], "lens", k.lens, 9)) {
    napi_throw_type_error(env, nullptr, "camera needs R[9], pos[3], lens[9]");
    return nullptr;
  }
  k.mode = (uint32_t)num(env, argv[2], "mode", 0);
  k.width = (uint32_t)num(env, argv[2], "width", 0);
  k.height = (uint32_t)num(env, argv[2], "height", 0);
  k.spp = (uint32_t)num(env, argv[2], "spp", 1);
  k.seed = (uint32_t)num(env, argv[2], "seed", 0);
  k.cam_index = (uint32_t)num(env, argv[2], "camIndex", slot);
  k.light_samples = (uint32_t)num(env, argv[2], "lightSamples", 1);
  k.max_depth = (uint32_t)num(env, argv[2], "maxDepth", 1);
  k.flags = (uint32_t)num(env, argv[2], "options", 0);
  k.scale = (float)num(env, argv[2], "scale", 1.0);
  k.skip_bodies = (uint32_t)num(env, argv[2], "skipBodies", 0);
  k.refine = (uint32_t)num(env, argv[2], "refine", 0);
  k.refine_light_samples = (uint32_t)num(env, argv[2], "refineLightSamples", 0);
  k.shutter[0] = 0.0f;
  k.shutter[1] = 1.0f;
  napi_value sh;
  if (get_prop(env, argv[2], "shutter", &sh) && !float_list(env, argv[2], "shutter", k.shutter, 2)) {
    napi_throw_type_error(env, nullptr, "camera shutter must be [t0, t1]");
    return nullptr;
  }
  int rc = vt_set_camera(c, slot, &k);
  if (rc) return throw_vt(env, c, rc, "vt_set_camera");
  return nullptr;
}

// setInstances(ctx, [{q0, q1, p0, p1} | null, ...]): one entry per body, null = hidden
napi_value SetInstances(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t len = 0;
  if (argc < 2 || napi_get_array_length(env, argv[1], &len) != napi_ok) {
    napi_throw_type_error(env, nullptr, "setInstances(ctx, poses: Array)");
    return nullptr;
  }
  std::vector<vt_instance> v(len ? len : 1);
  for (uint32_t i = 0; i < len; i++) {
    napi_value e;
    CHECK_NAPI(napi_get_element(env, argv[1], i, &e));
    napi_valuetype t;
    napi_typeof(env, e, &t);
    vt_instance &I = v[i];
    I = vt_instance{};
    I.q0[3] = I.q1[3] = 1.0f;
    if (t == napi_null || t == napi_undefined) {
      I.hidden = 1;
      continue;
    }
    if (!float_list(env, e, "q0", I.q0, 4) || !float_list(env, e, "p0", I.p0, 3)) {
      napi_throw_type_error(env, nullptr, "pose needs q0[4] and p0[3]");
      return nullptr;
    }
    if (!float_list(env, e, "q1", I.q1, 4)) std::memcpy(I.q1, I.q0, sizeof I.q0);
    if (!float_list(env, e, "p1", I.p1, 3)) std::memcpy(I.p1, I.p0, sizeof I.p0);
    // occlusion capsule (body frame) and effective reflectance, when the pose carries them
    if (float_list(env, e, "occA", I.occ_a, 3) && float_list(env, e, "occB", I.occ_b, 3)) {
      I.occ_r = (float)num(env, e, "occR", 0);
      I.rho_eff = (float)num(env, e, "rhoEff", 0);
    }
    // the cavity cut of the occlusion capsule: {origin: [3], axis: [3], s0, r}
    napi_value cav;
    bool has = false;
    napi_has_named_property(env, e, "cavity", &has);
    if (has && napi_get_named_property(env, e, "cavity", &cav) == napi_ok) {
      napi_valuetype t;
      napi_typeof(env, cav, &t);
      if (t == napi_object && float_list(env, cav, "origin", I.cavity_origin, 3) &&
          float_list(env, cav, "axis", I.cavity_axis, 3)) {
        I.cavity = 1;
        I.cavity_s0 = (float)num(env, cav, "s0", 0);
        I.cavity_r = (float)num(env, cav, "r", 0);
      }
    }
  }
  int rc = vt_set_instances(c, v.data(), len);
  if (rc) return throw_vt(env, c, rc, "vt_set_instances");
  return nullptr;
}

// setLightProfile(ctx, light, {axis: [3], theta: [..], p: [..]} | null)
// setBoardTexture(ctx, mask: Uint8Array (width * height, row 0 the image's top), width, height)
napi_value SetBoardTexture(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t w = 0, h = 0;
  bool ta = false;
  if (argc < 4 || napi_is_typedarray(env, argv[1], &ta) != napi_ok || !ta ||
      napi_get_value_uint32(env, argv[2], &w) != napi_ok || napi_get_value_uint32(env, argv[3], &h) != napi_ok) {
    napi_throw_type_error(env, nullptr, "setBoardTexture(ctx, Uint8Array mask, width, height)");
    return nullptr;
  }
  napi_typedarray_type type;
  size_t len = 0, off = 0;
  void *data = nullptr;
  napi_value ab;
  CHECK_NAPI(napi_get_typedarray_info(env, argv[1], &type, &len, &data, &ab, &off));
  if ((type != napi_uint8_array && type != napi_uint8_clamped_array) || len != (size_t)w * h) {
    napi_throw_range_error(env, nullptr, "the mask needs width * height bytes (Uint8Array)");
    return nullptr;
  }
  int rc = vt_set_board_texture(c, (const uint8_t *)data, w, h);
  if (rc) return throw_vt(env, c, rc, "vt_set_board_texture");
  return nullptr;
}

napi_value SetLightProfile(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t light = 0;
  napi_get_value_uint32(env, argv[1], &light);
  float axis[3] = {0.f, 0.f, 1.f};
  std::vector<float> th, p;
  napi_valuetype t = napi_undefined;
  if (argc > 2) napi_typeof(env, argv[2], &t);
  if (t == napi_object) {
    napi_value a, b;
    uint32_t na = 0, nb = 0;
    if (!float_list(env, argv[2], "axis", axis, 3) || !get_prop(env, argv[2], "theta", &a) ||
        !get_prop(env, argv[2], "p", &b) || napi_get_array_length(env, a, &na) != napi_ok ||
        napi_get_array_length(env, b, &nb) != napi_ok || na != nb || na > VT_MAX_KNOTS) {
      napi_throw_type_error(env, nullptr, "light profile needs axis[3] and equally long theta[], p[] (<= 64 knots)");
      return nullptr;
    }
    th.resize(na);
    p.resize(na);
    if (na && (!float_list(env, argv[2], "theta", th.data(), na) || !float_list(env, argv[2], "p", p.data(), na))) {
      napi_throw_type_error(env, nullptr, "light profile knots must be numbers");
      return nullptr;
    }
  }
  int rc = vt_set_light_profile(c, light, axis, th.data(), p.data(), (uint32_t)th.size());
  if (rc) return throw_vt(env, c, rc, "vt_set_light_profile");
  return nullptr;
}

// visible(ctx, cam, points: Float32Array (n x 3), tol, time) -> Uint8Array (1 = visible); synchronous, as the
// reference's CPU query is
napi_value Visible(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t cam = 0;
  double tol = 2e-4, time = 0.5;
  napi_get_value_uint32(env, argv[1], &cam);
  bool is = false;
  if (argc < 5 || napi_is_typedarray(env, argv[2], &is) != napi_ok || !is) {
    napi_throw_type_error(env, nullptr, "visible(ctx, cam, points: Float32Array, tol, time)");
    return nullptr;
  }
  napi_typedarray_type tt;
  size_t len, off;
  void *data;
  napi_value ab;
  napi_get_typedarray_info(env, argv[2], &tt, &len, &data, &ab, &off);
  if (tt != napi_float32_array || len % 3) {
    napi_throw_type_error(env, nullptr, "visible: points must be a Float32Array of n x 3");
    return nullptr;
  }
  napi_get_value_double(env, argv[3], &tol);
  napi_get_value_double(env, argv[4], &time);
  const uint32_t n = (uint32_t)(len / 3);
  napi_value outab, out;
  void *dst;
  CHECK_NAPI(napi_create_arraybuffer(env, n ? n : 1, &dst, &outab));
  int rc = vt_visible(c, cam, (const float *)data, n, (float)tol, (float)time, (uint8_t *)dst);
  if (rc) return throw_vt(env, c, rc, "vt_visible");
  CHECK_NAPI(napi_create_typedarray(env, napi_uint8_array, n, outab, 0, &out));
  return out;
}

napi_value Render(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t len = 0;
  napi_get_array_length(env, argv[1], &len);
  std::vector<uint32_t> slots(len);
  for (uint32_t i = 0; i < len; i++) {
    napi_value e;
    napi_get_element(env, argv[1], i, &e);
    napi_get_value_uint32(env, e, &slots[i]);
  }
  vt_render_opts o{};
  if (argc > 2) {
    o.mode = (uint32_t)num(env, argv[2], "mode", 0);
    o.spp_begin = (uint32_t)num(env, argv[2], "sppBegin", 0);
    o.spp_end = (uint32_t)num(env, argv[2], "sppEnd", 0);
    o.accumulate = (uint32_t)num(env, argv[2], "accumulate", 0);
    o.chunk = (uint32_t)num(env, argv[2], "chunk", 0);
    o.block = (uint32_t)num(env, argv[2], "block", 0);
    o.stack_lds = (uint32_t)num(env, argv[2], "stackLds", 0);
    o.queues = (uint32_t)num(env, argv[2], "queues", 0);
    o.graph = (uint32_t)num(env, argv[2], "graph", 0);
    o.batch = (uint32_t)num(env, argv[2], "batch", 0);
    o.sort_bits = (uint32_t)num(env, argv[2], "sortBits", 0);
  }
  int rc = vt_render(c, slots.data(), len, &o);
  if (rc) return throw_vt(env, c, rc, "vt_render");
  return nullptr;
}

napi_value Sensor(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t slot = 0;
  napi_get_value_uint32(env, argv[1], &slot);
  vt_sensor_params p{};
  napi_value o = argv[2];
  p.cfa = (uint32_t)num(env, o, "cfa", 0);
  p.frame = (uint32_t)num(env, o, "frame", 0);
  p.seed = (uint32_t)num(env, o, "seed", 0);
  p.optics_k = (float)num(env, o, "opticsK", 0.17671458676442586);
  p.esf_sigma = (float)num(env, o, "esfSigma", 0.6374);
  p.exposure = (float)num(env, o, "exposure", 1e-3);
  p.gain = (float)num(env, o, "gain", 1.0);
  if (!float_list(env, o, "m", p.m, 9)) {
    napi_throw_type_error(env, nullptr, "sensor params need m[9] (response matrix rows)");
    return nullptr;
  }
  p.read_noise = (float)num(env, o, "readNoise", 3.0);
  p.full_well = (float)num(env, o, "fullWell", 10000.0);
  p.dark = (float)num(env, o, "dark", 0.0);
  p.conv_gain = (float)num(env, o, "convGain", 0.33);
  p.max_dn = (float)num(env, o, "maxDn", 4095.0);
  p.black = (float)num(env, o, "black", 0.0);
  p.noise_on = (uint32_t)num(env, o, "noise", 1);
  p.pupil_exp = (float)num(env, o, "pupilExp", 1.0);
  p.glare_fraction = (float)num(env, o, "glareFraction", 0.0);
  p.glare_sigma = (float)num(env, o, "glareSigma", 1.0);
  p.glare_level = (uint32_t)num(env, o, "glareLevel", 16);
  {
    napi_value vt, vv;
    uint32_t nt = 0, nv = 0;
    if (get_prop(env, o, "vignettingThetaDeg", &vt) && get_prop(env, o, "vignetting", &vv)) {
      napi_get_array_length(env, vt, &nt);
      napi_get_array_length(env, vv, &nv);
      if (nt != nv || nt > VT_MAX_VIGNETTING) {
        napi_throw_range_error(env, nullptr, "vignettingThetaDeg and vignetting need the same length, at most 8");
        return nullptr;
      }
      if (!float_list(env, o, "vignettingThetaDeg", p.vignetting_theta_deg, nt) ||
          !float_list(env, o, "vignetting", p.vignetting, nt)) {
        napi_throw_type_error(env, nullptr, "vignetting tables must hold numbers");
        return nullptr;
      }
      p.vignetting_count = nt;
    }
  }
  int rc = vt_sensor(c, slot, &p);
  if (rc) return throw_vt(env, c, rc, "vt_sensor");
  return nullptr;
}

// isp(ctx, slot, {demosaic, nr, tone, chroma420, wb[3], ccm[9], gamma, sCurveC, blackOffset, nrSigmaPx,
// nrRangeScale, sharpenAmount, sharpenRadius}) (numeric codes; js/renderer.mjs maps the reference's strings)
napi_value Isp(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t slot = 0;
  napi_get_value_uint32(env, argv[1], &slot);
  vt_isp_params q{};
  napi_value o = argv[2];
  q.demosaic = (uint32_t)num(env, o, "demosaic", 0);
  q.nr = (uint32_t)num(env, o, "nr", 0);
  q.tone = (uint32_t)num(env, o, "tone", 0);
  q.chroma420 = (uint32_t)num(env, o, "chroma420", 1);
  if (!float_list(env, o, "wb", q.wb, 3)) q.wb[0] = q.wb[1] = q.wb[2] = 1.0f;
  if (!float_list(env, o, "ccm", q.ccm, 9)) {
    for (int k = 0; k < 9; k++) q.ccm[k] = (k % 4 == 0) ? 1.0f : 0.0f;
  }
  q.gamma = (float)num(env, o, "gamma", 2.2);
  q.s_curve_c = (float)num(env, o, "sCurveC", 1.0);
  q.black_offset = (float)num(env, o, "blackOffset", 0.0);
  q.nr_sigma_px = (float)num(env, o, "nrSigmaPx", 0.4);
  q.nr_range_scale = (float)num(env, o, "nrRangeScale", 2.0);
  q.sharpen_amount = (float)num(env, o, "sharpenAmount", 0.0);
  q.sharpen_radius = (float)num(env, o, "sharpenRadius", 1.0);
  int rc = vt_isp(c, slot, &q);
  if (rc) return throw_vt(env, c, rc, "vt_isp");
  return nullptr;
}

// ------------------------------------------------------------------ async: sync / readback / stats
struct Job {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  Handle *hd = nullptr;   // counted in hd->pending from queue_job to job_done
  vt_ctx *c = nullptr;
  int kind = -1;          // -1 sync, -2 stats, -3 autotune, -4 backward, >= 0 readback kind
  uint32_t slot = 0, w = 0, h = 0;
  std::vector<uint8_t> buf;
  vt_frame_stats st{};
  std::vector<uint32_t> slots;            // autotune, backward
  vt_render_opts opts{};
  uint32_t probe_stride = 0;
  vt_tune_result tune{};
  std::vector<std::vector<float>> adj;    // backward: copies of the adjoint images
  std::vector<float> gmat, glight, gcam;
  int rc = 0;
  std::string err;
};

void job_exec(napi_env, void *data) {
  Job *j = static_cast<Job *>(data);
  std::lock_guard<std::mutex> lk(g_mu);
  if (j->kind == -1) j->rc = vt_sync(j->c);
  else if (j->kind == -2) j->rc = vt_frame_stats_get(j->c, &j->st);
  else if (j->kind == -3)
    j->rc = vt_autotune(j->c, j->slots.data(), (uint32_t)j->slots.size(), &j->opts, j->probe_stride, &j->tune);
  else if (j->kind == -4) {
    std::vector<const float *> ptrs;
    for (auto &a : j->adj) ptrs.push_back(a.data());
    vt_grads g{j->gmat.data(), (uint32_t)(j->gmat.size() / 3), j->glight.data(), (uint32_t)(j->glight.size() / 3),
               j->gcam.empty() ? nullptr : j->gcam.data(), (uint32_t)(j->gcam.size() / VT_CAMERA_PARAMS)};
    j->rc = vt_backward(j->c, j->slots.data(), (uint32_t)j->slots.size(), ptrs.data(), &j->opts, &g);
  } else j->rc = vt_readback(j->c, j->slot, j->kind, j->buf.data(), j->buf.size());
  if (j->rc) j->err = vt_last_error(j->c);
}

void job_done(napi_env env, napi_status, void *data) {
  Job *j = static_cast<Job *>(data);
  napi_value res;
  if (j->rc) {
    napi_value msg, e;
    std::string m = "vt call failed (" + std::to_string(j->rc) + "): " + j->err;
    napi_create_string_utf8(env, m.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &e);
    napi_reject_deferred(env, j->deferred, e);
  } else if (j->kind == -1) {
    napi_get_undefined(env, &res);
    napi_resolve_deferred(env, j->deferred, res);
  } else if (j->kind == -2) {
    napi_create_object(env, &res);
    const char *names[] = {"rays", "samples", "msRender", "msSensor", "msKernelAvg", "launches", "stackOverflows",
                           "msBackward", "raysBackward", "graphUsed", "graphBuilds", "msIsp"};
    double vals[] = {j->st.rays, j->st.samples, j->st.ms_render, j->st.ms_sensor, j->st.ms_kernel_avg,
                     (double)j->st.launches, (double)j->st.stack_overflows, j->st.ms_backward, j->st.rays_backward,
                     (double)j->st.graph_used, (double)j->st.graph_builds, j->st.ms_isp};
    for (int k = 0; k < 12; k++) {
      napi_value v;
      napi_create_double(env, vals[k], &v);
      napi_set_named_property(env, res, names[k], v);
    }
    napi_resolve_deferred(env, j->deferred, res);
  } else if (j->kind == -3) {
    napi_create_object(env, &res);
    const char *names[] = {"block", "stackLds", "chunk", "queues", "candidates", "probeMraysBest",
                           "probeMraysDefault", "probeFraction", "mode", "wfBatch", "wfStackLds", "wfRefill",
                           "probeMraysWavefront"};
    double vals[] = {(double)j->tune.block, (double)j->tune.stack_lds, (double)j->tune.chunk, (double)j->tune.queues,
                     (double)j->tune.candidates, j->tune.probe_mrays_best, j->tune.probe_mrays_default,
                     j->tune.probe_fraction, (double)j->tune.mode, (double)j->tune.wf_batch,
                     (double)j->tune.wf_stack_lds, (double)j->tune.wf_refill, j->tune.probe_mrays_wavefront};
    for (int k = 0; k < 13; k++) {
      napi_value v;
      napi_create_double(env, vals[k], &v);
      napi_set_named_property(env, res, names[k], v);
    }
    napi_resolve_deferred(env, j->deferred, res);
  } else if (j->kind == -4) {
    napi_create_object(env, &res);
    std::vector<float> *parts[] = {&j->gmat, &j->glight, &j->gcam};
    const char *names[] = {"materials", "lights", "cameras"};
    for (int k = 0; k < 3; k++) {
      if (k == 2 && j->gcam.empty()) continue;
      napi_value ab, ta;
      void *dst;
      napi_create_arraybuffer(env, parts[k]->size() * 4, &dst, &ab);
      if (!parts[k]->empty()) std::memcpy(dst, parts[k]->data(), parts[k]->size() * 4);
      napi_create_typedarray(env, napi_float32_array, parts[k]->size(), ab, 0, &ta);
      napi_set_named_property(env, res, names[k], ta);
    }
    napi_resolve_deferred(env, j->deferred, res);
  } else {
    napi_value ab;
    void *dst;
    napi_create_arraybuffer(env, j->buf.size(), &dst, &ab);
    std::memcpy(dst, j->buf.data(), j->buf.size());
    const bool bytes = j->kind == VT_KIND_RGB8 || j->kind == VT_KIND_RGBA8 || j->kind == VT_KIND_YUV420;
    napi_typedarray_type t = j->kind == VT_KIND_RAW     ? napi_uint16_array
                             : bytes                    ? napi_uint8_array
                             : j->kind == VT_KIND_STATS ? napi_uint32_array
                                                        : napi_float32_array;
    size_t elems = j->kind == VT_KIND_RAW ? j->buf.size() / 2 : bytes ? j->buf.size() : j->buf.size() / 4;
    napi_create_typedarray(env, t, elems, ab, 0, &res);
    napi_resolve_deferred(env, j->deferred, res);
  }
  napi_delete_async_work(env, j->work);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    j->hd->pending--;
    settle(j->hd);
  }
  delete j;
}

napi_value queue_job(napi_env env, Job *j) {
  napi_value promise, name;
  if (napi_create_promise(env, &j->deferred, &promise) != napi_ok) {
    delete j;
    napi_throw_error(env, nullptr, "promise creation failed");
    return nullptr;
  }
  napi_create_string_utf8(env, "vt", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, job_exec, job_done, j, &j->work);
  j->hd->pending++;
  napi_queue_async_work(env, j->work);
  return promise;
}

napi_value Sync(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  Handle *hd = c ? get_handle(env, argv[0]) : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  Job *j = new Job();
  j->c = c;
  j->hd = hd;
  j->kind = -1;
  return queue_job(env, j);
}

napi_value Stats(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  Handle *hd = c ? get_handle(env, argv[0]) : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  Job *j = new Job();
  j->c = c;
  j->hd = hd;
  j->kind = -2;
  return queue_job(env, j);
}

napi_value Readback(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  Handle *hd = c ? get_handle(env, argv[0]) : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  Job *j = new Job();
  j->c = c;
  j->hd = hd;
  int32_t kind = 0;
  napi_get_value_uint32(env, argv[1], &j->slot);
  napi_get_value_int32(env, argv[2], &kind);
  napi_get_value_uint32(env, argv[3], &j->w);
  napi_get_value_uint32(env, argv[4], &j->h);
  j->kind = kind;
  const size_t px = (size_t)j->w * j->h;
  const size_t cpx = (size_t)((j->w + 1) / 2) * ((j->h + 1) / 2);
  j->buf.resize(kind == VT_KIND_LINEAR   ? px * 16
                : kind == VT_KIND_RAW    ? px * 2
                : kind == VT_KIND_RGB8   ? px * 3
                : kind == VT_KIND_YUV420 ? px + 2 * cpx
                : kind == VT_KIND_STATS  ? 16
                                         : px * 4);
  return queue_job(env, j);
}

bool slot_list(napi_env env, napi_value arr, std::vector<uint32_t> &out) {
  uint32_t len = 0;
  if (napi_get_array_length(env, arr, &len) != napi_ok) return false;
  out.resize(len);
  for (uint32_t i = 0; i < len; i++) {
    napi_value e;
    napi_get_element(env, arr, i, &e);
    napi_get_value_uint32(env, e, &out[i]);
  }
  return true;
}

// autotune(ctx, slots, {sppBegin, sppEnd, probeStride}) -> Promise<{block, stackLds, chunk, queues, ...}>
napi_value Autotune(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  Handle *hd = c ? get_handle(env, argv[0]) : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  Job *j = new Job();
  j->c = c;
  j->hd = hd;
  j->kind = -3;
  if (!slot_list(env, argv[1], j->slots)) {
    delete j;
    napi_throw_type_error(env, nullptr, "autotune: camera slots must be an array");
    return nullptr;
  }
  if (argc > 2) {
    j->opts.spp_begin = (uint32_t)num(env, argv[2], "sppBegin", 0);
    j->opts.spp_end = (uint32_t)num(env, argv[2], "sppEnd", 0);
    j->probe_stride = (uint32_t)num(env, argv[2], "probeStride", 0);
  }
  return queue_job(env, j);
}

// backward(ctx, slots, adjImages[], nMat, nLight, {sppBegin, sppEnd, cameras}) -> Promise<{materials, lights[, cameras]}>
napi_value Backward(napi_env env, napi_callback_info info) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  Handle *hd = c ? get_handle(env, argv[0]) : nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  Job *j = new Job();
  j->c = c;
  j->hd = hd;
  j->kind = -4;
  uint32_t na = 0, nm = 0, nl = 0;
  if (argc < 5 || !slot_list(env, argv[1], j->slots) || napi_get_array_length(env, argv[2], &na) != napi_ok ||
      na != j->slots.size()) {
    delete j;
    napi_throw_type_error(env, nullptr, "backward: (ctx, slots, adjImages, nMat, nLight) with one image per slot");
    return nullptr;
  }
  napi_get_value_uint32(env, argv[3], &nm);
  napi_get_value_uint32(env, argv[4], &nl);
  for (uint32_t i = 0; i < na; i++) {
    napi_value e;
    napi_get_element(env, argv[2], i, &e);
    bool is_ta = false;
    napi_is_typedarray(env, e, &is_ta);
    napi_typedarray_type t;
    size_t len = 0;
    void *data = nullptr;
    if (!is_ta || napi_get_typedarray_info(env, e, &t, &len, &data, nullptr, nullptr) != napi_ok ||
        t != napi_float32_array) {
      delete j;
      napi_throw_type_error(env, nullptr, "backward: adjoint images must be Float32Array (w*h*4)");
      return nullptr;
    }
    j->adj.emplace_back(static_cast<const float *>(data), static_cast<const float *>(data) + len);
  }
  j->gmat.assign(3 * (size_t)nm, 0.0f);
  j->glight.assign(3 * (size_t)nl, 0.0f);
  if (argc > 5) {
    j->opts.spp_begin = (uint32_t)num(env, argv[5], "sppBegin", 0);
    j->opts.spp_end = (uint32_t)num(env, argv[5], "sppEnd", 0);
    if (num(env, argv[5], "cameras", 0) != 0) j->gcam.assign((size_t)VT_CAMERA_PARAMS * na, 0.0f);
  }
  return queue_job(env, j);
}

// updateMaterials(ctx, Float32Array 16 per material) / updateLights(ctx, Float32Array 4 per light)
napi_value update_table(napi_env env, napi_callback_info info, bool mats) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  napi_typedarray_type t;
  size_t len = 0;
  void *data = nullptr;
  bool is_ta = false;
  napi_is_typedarray(env, argv[1], &is_ta);
  if (!is_ta || napi_get_typedarray_info(env, argv[1], &t, &len, &data, nullptr, nullptr) != napi_ok ||
      t != napi_float32_array) {
    napi_throw_type_error(env, nullptr, "expected a Float32Array");
    return nullptr;
  }
  const int rc = mats ? vt_update_materials(c, static_cast<const float *>(data), (uint32_t)(len / 16))
                      : vt_update_lights(c, static_cast<const float *>(data), (uint32_t)(len / 4));
  if (rc) return throw_vt(env, c, rc, mats ? "vt_update_materials" : "vt_update_lights");
  return nullptr;
}
napi_value UpdateMaterials(napi_env env, napi_callback_info info) { return update_table(env, info, true); }
napi_value UpdateLights(napi_env env, napi_callback_info info) { return update_table(env, info, false); }

// radiosityBuild(ctx, {gatherK, nodeSamples, nbrK, smoothRadius, smoothSigma}) -> {nodes, staticNodes, active, ms}
napi_value RadiosityBuild(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  vt_radiosity_opts o{64, 4, 16, 0.006f, 0.003f};
  napi_valuetype t = napi_undefined;
  if (argc > 1) napi_typeof(env, argv[1], &t);
  if (t == napi_object) {
    o.gather_k = (uint32_t)num(env, argv[1], "gatherK", o.gather_k);
    o.node_samples = (uint32_t)num(env, argv[1], "nodeSamples", o.node_samples);
    o.nbr_k = (uint32_t)num(env, argv[1], "nbrK", o.nbr_k);
    o.smooth_radius = (float)num(env, argv[1], "smoothRadius", o.smooth_radius);
    o.smooth_sigma = (float)num(env, argv[1], "smoothSigma", o.smooth_sigma);
  }
  vt_radiosity_info inf{};
  int rc = vt_radiosity_build(c, &o, &inf);
  if (rc) return throw_vt(env, c, rc, "vt_radiosity_build");
  napi_value out, v;
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_create_uint32(env, inf.nodes, &v));
  CHECK_NAPI(napi_set_named_property(env, out, "nodes", v));
  CHECK_NAPI(napi_create_uint32(env, inf.static_nodes, &v));
  CHECK_NAPI(napi_set_named_property(env, out, "staticNodes", v));
  CHECK_NAPI(napi_create_uint32(env, inf.active, &v));
  CHECK_NAPI(napi_set_named_property(env, out, "active", v));
  CHECK_NAPI(napi_create_double(env, inf.ms_build, &v));
  CHECK_NAPI(napi_set_named_property(env, out, "ms", v));
  return out;
}

// interiorBvh(ctx, enable: bool) -> static triangles the walks see (vt_interior_bvh)
napi_value InteriorBvh(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  bool enable = true;
  if (argc > 1) napi_get_value_bool(env, argv[1], &enable);
  uint32_t n = 0;
  int rc = vt_interior_bvh(c, enable ? 1u : 0u, &n);
  if (rc) return throw_vt(env, c, rc, "vt_interior_bvh");
  napi_value out;
  CHECK_NAPI(napi_create_uint32(env, n, &out));
  return out;
}

// radiosityUpdate(ctx, direct: bool, iters) (enqueue)
napi_value RadiosityUpdate(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  bool direct = true;
  uint32_t iters = 0;
  if (argc > 1) napi_get_value_bool(env, argv[1], &direct);
  if (argc > 2) napi_get_value_uint32(env, argv[2], &iters);
  int rc = vt_radiosity_update(c, direct ? 1u : 0u, iters);
  if (rc) return throw_vt(env, c, rc, "vt_radiosity_update");
  return nullptr;
}

// radiosityUpdateMoved(ctx, movedMask, iters) -> 'none' | 'incremental' | 'full' (enqueue)
napi_value RadiosityUpdateMoved(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t moved = 0, iters = 0, ran = 0;
  if (argc > 1) napi_get_value_uint32(env, argv[1], &moved);
  if (argc > 2) napi_get_value_uint32(env, argv[2], &iters);
  int rc = vt_radiosity_update_moved(c, moved, iters, &ran);
  if (rc) return throw_vt(env, c, rc, "vt_radiosity_update_moved");
  static const char *kinds[] = {"none", "incremental", "full"};
  napi_value out;
  CHECK_NAPI(napi_create_string_utf8(env, kinds[ran < 3 ? ran : 0], NAPI_AUTO_LENGTH, &out));
  return out;
}

// radiosityState(ctx, nodes) -> {eAvg: [3], state: Float32Array(nodes * 16)}; blocks until the queue is done
napi_value RadiosityState(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t nodes = 0;
  if (argc > 1) napi_get_value_uint32(env, argv[1], &nodes);
  napi_value ab, st, out, ea;
  void *dst = nullptr;
  CHECK_NAPI(napi_create_arraybuffer(env, (size_t)nodes * 64, &dst, &ab));
  float e[3] = {0, 0, 0};
  int rc = vt_radiosity_state(c, nodes ? (float *)dst : nullptr, (size_t)nodes * 64, e);
  if (rc) return throw_vt(env, c, rc, "vt_radiosity_state");
  CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, (size_t)nodes * 16, ab, 0, &st));
  CHECK_NAPI(napi_create_array_with_length(env, 3, &ea));
  for (uint32_t k = 0; k < 3; k++) {
    napi_value v;
    CHECK_NAPI(napi_create_double(env, e[k], &v));
    CHECK_NAPI(napi_set_element(env, ea, k, v));
  }
  CHECK_NAPI(napi_create_object(env, &out));
  CHECK_NAPI(napi_set_named_property(env, out, "eAvg", ea));
  CHECK_NAPI(napi_set_named_property(env, out, "state", st));
  return out;
}

// bakedUpload(ctx, {nodes, K, gather: Uint16Array, nodeA: Float32Array, nodeN: Float32Array, hh: Float32Array})
// (the lite renderer's baked node table; blocks)
napi_value BakedUpload(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c || argc < 2) return nullptr;
  vt_baked_table t{};
  size_t ng = 0, na = 0, nn = 0, nh = 0;
  t.nodes = (uint32_t)num(env, argv[1], "nodes", 0);
  t.K = (uint32_t)num(env, argv[1], "K", 0);
  t.gather = typed<uint16_t>(env, argv[1], "gather", &ng, napi_uint16_array);
  t.node_a = typed<float>(env, argv[1], "nodeA", &na, napi_float32_array);
  t.node_n = typed<float>(env, argv[1], "nodeN", &nn, napi_float32_array);
  t.hh = typed<float>(env, argv[1], "hh", &nh, napi_float32_array);
  t.meshes = (uint32_t)(nh / 4);
  if (!t.gather || !t.node_a || !t.node_n || !t.hh || ng != (size_t)t.nodes * t.K || na != 4 * (size_t)t.nodes ||
      nn != 4 * (size_t)t.nodes) {
    napi_throw_range_error(env, nullptr, "bakedUpload: gather nodes*K, nodeA and nodeN nodes*4 entries");
    return nullptr;
  }
  int rc = vt_baked_upload(c, &t);
  if (rc) return throw_vt(env, c, rc, "vt_baked_upload");
  get_handle(env, argv[0])->baked_nodes = t.nodes;
  return nullptr;
}

// bakedOcclusion(ctx, caps: Float32Array(8 per capsule: a.xyz, r, b.xyz, rho)) (enqueue)
// bakedJacobi(ctx, eDir: Float32Array(nodes * 4) | null, iters) (enqueue)
napi_value BakedOcclusion(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value argv[2];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c || argc < 2) return nullptr;
  napi_value o;
  CHECK_NAPI(napi_create_object(env, &o));
  CHECK_NAPI(napi_set_named_property(env, o, "caps", argv[1]));
  size_t n = 0;
  const float *caps = typed<float>(env, o, "caps", &n, napi_float32_array);
  if (!caps && n) return nullptr;
  int rc = vt_baked_occlusion(c, caps, (uint32_t)(n / 8));
  if (rc) return throw_vt(env, c, rc, "vt_baked_occlusion");
  return nullptr;
}

napi_value BakedJacobi(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  const float *ed = nullptr;
  if (argc > 1) {
    napi_valuetype vt;
    napi_typeof(env, argv[1], &vt);
    if (vt != napi_null && vt != napi_undefined) {
      napi_value o;
      CHECK_NAPI(napi_create_object(env, &o));
      CHECK_NAPI(napi_set_named_property(env, o, "eDir", argv[1]));
      size_t n = 0;
      ed = typed<float>(env, o, "eDir", &n, napi_float32_array);
      if (!ed) return nullptr;
      if (n != 4 * (size_t)get_handle(env, argv[0])->baked_nodes) {
        napi_throw_range_error(env, nullptr, "bakedJacobi: eDir needs nodes * 4 entries");
        return nullptr;
      }
    }
  }
  uint32_t iters = 1;
  if (argc > 2) napi_get_value_uint32(env, argv[2], &iters);
  int rc = vt_baked_jacobi(c, ed, iters);
  if (rc) return throw_vt(env, c, rc, "vt_baked_jacobi");
  return nullptr;
}

// bakedRead(ctx, which, nodes, K) -> Float32Array (B, E_ind: nodes * 4) | Uint8Array (bytes: nodes * K); blocks
napi_value BakedRead(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  uint32_t which = 0, nodes = 0, K = 0;
  if (argc > 1) napi_get_value_uint32(env, argv[1], &which);
  if (argc > 2) napi_get_value_uint32(env, argv[2], &nodes);
  if (argc > 3) napi_get_value_uint32(env, argv[3], &K);
  const size_t bytes = which == VT_BAKED_OCC ? (size_t)nodes * K : (size_t)nodes * 16;
  napi_value ab, out;
  void *dst = nullptr;
  CHECK_NAPI(napi_create_arraybuffer(env, std::max<size_t>(bytes, 1), &dst, &ab));
  int rc = vt_baked_read(c, which, dst, bytes);
  if (rc) return throw_vt(env, c, rc, "vt_baked_read");
  if (which == VT_BAKED_OCC) CHECK_NAPI(napi_create_typedarray(env, napi_uint8_array, bytes, ab, 0, &out));
  else CHECK_NAPI(napi_create_typedarray(env, napi_float32_array, bytes / 4, ab, 0, &out));
  return out;
}

napi_value LastError(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value argv[1], out;
  CHECK_NAPI(napi_get_cb_info(env, info, &argc, argv, nullptr, nullptr));
  vt_ctx *c = get_ctx(env, argv[0]);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!c) return nullptr;
  CHECK_NAPI(napi_create_string_utf8(env, vt_last_error(c), NAPI_AUTO_LENGTH, &out));
  return out;
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char *name;
    napi_callback fn;
  } fns[] = {{"create", Create}, {"destroy", Destroy}, {"buildBVH", BuildBVH}, {"uploadScene", UploadScene},
             {"setCamera", SetCamera}, {"setInstances", SetInstances}, {"visible", Visible}, {"setLightProfile", SetLightProfile}, {"setBoardTexture", SetBoardTexture}, {"render", Render}, {"sensor", Sensor}, {"isp", Isp}, {"sync", Sync}, {"stats", Stats},
             {"readback", Readback}, {"lastError", LastError}, {"autotune", Autotune}, {"backward", Backward},
             {"updateMaterials", UpdateMaterials}, {"updateLights", UpdateLights},
             {"radiosityBuild", RadiosityBuild}, {"interiorBvh", InteriorBvh}, {"radiosityUpdate", RadiosityUpdate},
             {"radiosityUpdateMoved", RadiosityUpdateMoved},
             {"bakedUpload", BakedUpload}, {"bakedOcclusion", BakedOcclusion}, {"bakedJacobi", BakedJacobi},
             {"bakedRead", BakedRead},
             {"radiosityState", RadiosityState}};
  for (auto &f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  napi_value kinds;
  napi_create_object(env, &kinds);
  const char *kn[] = {"linear", "raw", "ri"};
  for (int k = 0; k < 3; k++) {
    napi_value v;
    napi_create_int32(env, k, &v);
    napi_set_named_property(env, kinds, kn[k], v);
  }
  napi_set_named_property(env, exports, "KIND", kinds);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
