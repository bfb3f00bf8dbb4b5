// The C ABI, part 4: the copy-bandwidth probe, visible(), and the reference's test entry points
// (gpu/wgsl/kernels.wgsl.js:487-551) on caller-given inputs.
#include "vt_ctx.h"

using namespace vt;

// ====================================================================== bandwidth probe
extern "C" int vt_copy_bandwidth(vt_ctx *c, uint64_t bytes, uint32_t iters, double *gbs) {
  VT_LIVE(c);
  if (!gbs || bytes < 4096 || iters == 0) return fail(c, VT_ERR_ARG, "copy_bandwidth: bytes >= 4096, iters > 0");
  VT_HIP(c, hipSetDevice(c->device));
  DevBuf a, b;
  int rc = ensure(c, a, bytes);
  if (!rc) rc = ensure(c, b, bytes);
  struct Free {
    DevBuf &x, &y;
    ~Free() { release(x); release(y); }
  } fr{a, b};
  if (rc) return rc;
  const uint64_t n = bytes / 16;
  const uint32_t grid = (uint32_t)c->cu_count * 8;
  VT_HIP(c, launch_zero_f4((float4 *)a.p, n, c->stream));
  for (int w = 0; w < 3; w++) VT_HIP(c, launch_copy_f4((const float4 *)a.p, (float4 *)b.p, n, grid, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  for (uint32_t k = 0; k < iters; k++) VT_HIP(c, launch_copy_f4((const float4 *)a.p, (float4 *)b.p, n, grid, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  VT_HIP(c, hipEventSynchronize(c->ev_s1));
  float ms = 0;
  VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
  *gbs = 2.0 * 16.0 * (double)n * iters / (ms * 1e-3) / 1e9;
  c->sensor_timed = false;
  return VT_OK;
}

extern "C" int vt_read_bandwidth(vt_ctx *c, uint64_t bytes, uint32_t iters, double *gbs) {
  VT_LIVE(c);
  if (!gbs || bytes < 4096 || iters == 0) return fail(c, VT_ERR_ARG, "read_bandwidth: bytes >= 4096, iters > 0");
  VT_HIP(c, hipSetDevice(c->device));
  DevBuf a, sink;
  int rc = ensure(c, a, bytes);
  if (!rc) rc = ensure(c, sink, 256);
  struct Free {
    DevBuf &x, &y;
    ~Free() { release(x); release(y); }
  } fr{a, sink};
  if (rc) return rc;
  const uint64_t n = bytes / 16;
  const uint32_t grid = (uint32_t)c->cu_count * 16;
  VT_HIP(c, launch_zero_f4((float4 *)a.p, n, c->stream));   // zeros fold to 0, so the sink is never written
  for (int w = 0; w < 3; w++) VT_HIP(c, launch_read_u4((const uint4 *)a.p, n, grid, (uint32_t *)sink.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  for (uint32_t k = 0; k < iters; k++) VT_HIP(c, launch_read_u4((const uint4 *)a.p, n, grid, (uint32_t *)sink.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  VT_HIP(c, hipEventSynchronize(c->ev_s1));
  float ms = 0;
  VT_HIP(c, hipEventElapsedTime(&ms, c->ev_s0, c->ev_s1));
  *gbs = 16.0 * (double)n * iters / (ms * 1e-3) / 1e9;
  c->sensor_timed = false;
  return VT_OK;
}

// ====================================================================== test entry points
extern "C" int vt_trace_test(vt_ctx *c, const float *rays, uint32_t n, int any_hit, float *hits) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "trace_test: no scene");
  if (n && (!rays || !hits)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 32))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 16))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, rays, (size_t)n * 32, hipMemcpyHostToDevice));
  TestLaunch T{c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes,
               (uint32_t)(c->spill_lanes / DEFAULT_BLOCK)};
  VT_HIP(c, launch_trace_test(T, (const float *)c->test_in.p, n, any_hit, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(hits, c->test_out.p, (size_t)n * 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_visible(vt_ctx *c, uint32_t cam, const float *points, uint32_t n, float tol, float time,
                          uint8_t *out) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "visible: no scene");
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "visible: camera %u not set", cam);
  if (n && (!points || !out)) return fail(c, VT_ERR_ARG, "visible: null points or output");
  if (!(tol >= 0.0f) || !(time >= 0.0f && time <= 1.0f)) return fail(c, VT_ERR_ARG, "visible: tol >= 0, time in [0, 1]");
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 12))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n))) return rc;
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(c->test_in.p, points, (size_t)n * 12, hipMemcpyHostToDevice));
  TestLaunch T{c->S, (uint32_t *)c->overflow.p, (uint2 *)c->spill.p, c->spill_lanes,
               (uint32_t)(c->spill_lanes / DEFAULT_BLOCK)};
  T.S.root = 0;                                 // the whole static scene, as the reference's CPU walk
  const vt_camera &k = c->cams[cam].c;
  VT_HIP(c, launch_visible(T, k.pos, k.skip_bodies, tol, time, (const float *)c->test_in.p, n, (uint8_t *)c->test_out.p,
                           c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, n, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_lens_test(vt_ctx *c, const float lens[9], const float *uv, uint32_t n, float *rays) {
  VT_LIVE(c);
  if (!lens || (n && (!uv || !rays))) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 8 + 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 12))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, lens, 36, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((char *)c->test_in.p + 64, uv, (size_t)n * 8, hipMemcpyHostToDevice));
  VT_HIP(c, launch_lens_test((const float *)c->test_in.p, (const float *)((char *)c->test_in.p + 64), n,
                             (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(rays, c->test_out.p, (size_t)n * 12, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_camera_ray_test(vt_ctx *c, uint32_t cam, const uint32_t *xyi, uint32_t n, float *out) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "camera %u not set", cam);
  if (n && (!xyi || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  const vt_camera &k = c->cams[cam].c;
  DevCam d{};
  std::memcpy(d.R, k.R, sizeof d.R);
  std::memcpy(d.pos, k.pos, sizeof d.pos);
  d.mode = k.mode;
  std::memcpy(d.lens, k.lens, sizeof d.lens);
  d.width = k.width; d.height = k.height; d.spp = k.spp; d.seed = k.seed; d.cam_index = k.cam_index;
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 12 + 256))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 24))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, &d, sizeof d, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy((char *)c->test_in.p + 256, xyi, (size_t)n * 12, hipMemcpyHostToDevice));
  VT_HIP(c, launch_camray_test((const DevCam *)c->test_in.p, (const uint32_t *)((char *)c->test_in.p + 256), n,
                               (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 24, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_bsdf_test(vt_ctx *c, const float *cases, uint32_t n, float *out) {
  VT_LIVE(c);
  if (!c->has_scene) return fail(c, VT_ERR_STATE, "bsdf_test: no scene (materials)");
  if (n && (!cases || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 32))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, cases, (size_t)n * 64, hipMemcpyHostToDevice));
  VT_HIP(c, launch_bsdf_test(c->S.mats, (const float *)c->test_in.p, n, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 32, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_capsule_test(vt_ctx *c, const float *cases, uint32_t n, float *out) {
  VT_LIVE(c);
  if (n && (!cases || !out)) return VT_ERR_ARG;
  if (n == 0) return VT_OK;
  VT_HIP(c, hipSetDevice(c->device));
  int rc;
  if ((rc = ensure(c, c->test_in, (size_t)n * 64))) return rc;
  if ((rc = ensure(c, c->test_out, (size_t)n * 4))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, cases, (size_t)n * 64, hipMemcpyHostToDevice));
  VT_HIP(c, launch_capsule_test((const float *)c->test_in.p, n, (float *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(out, c->test_out.p, (size_t)n * 4, hipMemcpyDeviceToHost));
  return VT_OK;
}

extern "C" int vt_sensor_test(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, const float *radiance,
                              const float *ri, uint16_t *dn) {
  VT_LIVE(c);
  if (!radiance || !ri || !dn || w == 0 || h == 0) return VT_ERR_ARG;
  VT_HIP(c, hipSetDevice(c->device));
  SensorDev P;
  int rc = sensor_params(c, p, w, h, 0, 1.0f, P);
  if (rc) return rc;
  P.cam_index = 0;
  const size_t px = (size_t)w * h;
  // test_aux: the RI map, then the two glare levels (float4-aligned)
  const size_t gb = glare_bytes(w, h, P), ga = (px * 4 + 255) / 256 * 256;
  if ((rc = ensure(c, c->test_in, px * 16))) return rc;
  if ((rc = ensure(c, c->test_aux, ga + 2 * gb))) return rc;
  if ((rc = ensure(c, c->test_out, px * 2))) return rc;
  VT_HIP(c, hipMemcpy(c->test_in.p, radiance, px * 16, hipMemcpyHostToDevice));
  VT_HIP(c, hipMemcpy(c->test_aux.p, ri, px * 4, hipMemcpyHostToDevice));
  char *g = (char *)c->test_aux.p + ga;
  VT_HIP(c, launch_sensor(P, (const float4 *)c->test_in.p, (const float *)c->test_aux.p, (float4 *)g,
                          (float4 *)(g + gb), (uint16_t *)c->test_out.p, c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(dn, c->test_out.p, px * 2, hipMemcpyDeviceToHost));
  return VT_OK;
}

// a device buffer of ctx's device holding at least `bytes` from p (hipPointerGetAttributes, no allocation)
static int device_buffer(vt_ctx *c, const void *p, size_t bytes, const char *what) {
  hipPointerAttribute_t a;
  std::memset(&a, 0, sizeof a);
  const hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, VT_ERR_ARG, "sensor_raw_device: %s is not device memory", what);
  }
  if (a.type != hipMemoryTypeDevice && a.type != hipMemoryTypeManaged)
    return fail(c, VT_ERR_ARG, "sensor_raw_device: %s is not device memory", what);
  if (a.device != c->device) return fail(c, VT_ERR_ARG, "sensor_raw_device: %s is on device %d, not %d", what, a.device, c->device);
  // the allocation holding p must reach p + bytes (allocations whose range HIP cannot report pass unchecked)
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange(&base, &size, (hipDeviceptr_t)p) == hipSuccess && base) {
    const uintptr_t b = (uintptr_t)base, q = (uintptr_t)p;
    if (q < b || q - b + bytes > size)
      return fail(c, VT_ERR_SIZE, "sensor_raw_device: %s needs %zu bytes, its allocation holds %zu from there", what,
                  bytes, (size_t)(b + size > q ? b + size - q : 0));
  } else {
    (void)hipGetLastError();
  }
  return VT_OK;
}

extern "C" int vt_sensor_raw_device(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, const float *radiance,
                                    const float *ri, uint16_t *dn) {
  VT_LIVE(c);
  if (!radiance || !ri || !dn || w == 0 || h == 0) return fail(c, VT_ERR_ARG, "sensor_raw_device: null buffer or empty size");
  if (w > 16384 || h > 16384) return fail(c, VT_ERR_SIZE, "sensor_raw_device: %ux%u above 16384x16384", w, h);
  SensorDev P;
  int rc = sensor_params(c, p, w, h, 0, 1.0f, P);
  if (rc) return rc;
  P.cam_index = 0;
  if (P.glare_k > 0.0f) {
    const uint64_t F = P.glare_f, level = ((uint64_t)(w + F - 1) / F) * ((h + F - 1) / F);
    if (level > GLARE_LEVEL_MAX)
      return fail(c, VT_ERR_SIZE, "sensor_raw_device: glare level %llu > %u pixels (raise the level factor)",
                  (unsigned long long)level, GLARE_LEVEL_MAX);
  }
  VT_HIP(c, hipSetDevice(c->device));
  const size_t px = (size_t)w * h;
  if ((rc = device_buffer(c, radiance, px * 16, "radiance"))) return rc;
  if ((rc = device_buffer(c, ri, px * 4, "ri"))) return rc;
  if ((rc = device_buffer(c, dn, px * 2, "dn"))) return rc;
  VT_HIP(c, launch_sensor_noalloc(c->device, P, (const float4 *)radiance, ri, dn, c->stream));
  return VT_OK;
}

extern "C" int vt_isp_test(vt_ctx *c, const vt_sensor_params *sp, const vt_isp_params *q, uint32_t w, uint32_t h,
                           const uint16_t *dn, uint8_t *rgba, uint8_t *yuv, uint32_t *stats) {
  VT_LIVE(c);
  if (!sp || !dn || !rgba || !stats || w == 0 || h == 0) return fail(c, VT_ERR_ARG, "isp_test: null argument");
  if (sp->cfa > 3) return fail(c, VT_ERR_ARG, "isp_test: cfa code %u", sp->cfa);
  VT_HIP(c, hipSetDevice(c->device));
  IspDev P;
  int rc = isp_params(c, *sp, q, w, h, P);
  if (rc) return rc;
  const size_t px = (size_t)w * h, cpx = (size_t)((w + 1) / 2) * ((h + 1) / 2);
  // test_in: the RGB plane, then the partial sums; test_aux: rgba, yuv, stats; test_out: the DN
  if ((rc = ensure(c, c->test_in, px * 16 + (size_t)isp_partials(w, h) * 16))) return rc;
  const size_t o_rgba = 0, o_yuv = o_rgba + (px * 4 + 255) / 256 * 256,
               o_st = o_yuv + (px + 2 * cpx + 255) / 256 * 256;
  if ((rc = ensure(c, c->test_aux, o_st + 16))) return rc;
  if ((rc = ensure(c, c->test_out, px * 2))) return rc;
  char *aux = (char *)c->test_aux.p;
  VT_HIP(c, hipMemcpy(c->test_out.p, dn, px * 2, hipMemcpyHostToDevice));
  VT_HIP(c, launch_isp(P, (const uint16_t *)c->test_out.p, (float4 *)c->test_in.p,
                       (uint4 *)((float4 *)c->test_in.p + px), (uchar4 *)(aux + o_rgba), (uint8_t *)(aux + o_yuv),
                       (unsigned int *)(aux + o_st), c->stream));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(rgba, aux + o_rgba, px * 4, hipMemcpyDeviceToHost));
  if (yuv && P.chroma420) VT_HIP(c, hipMemcpy(yuv, aux + o_yuv, px + 2 * cpx, hipMemcpyDeviceToHost));
  VT_HIP(c, hipMemcpy(stats, aux + o_st, 16, hipMemcpyDeviceToHost));
  return VT_OK;
}

