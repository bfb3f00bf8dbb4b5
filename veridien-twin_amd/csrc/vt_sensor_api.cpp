// The C ABI, part 3: the sensor (gpu/renderer.js:1056-1099: the fused raw stage with glare, the ISP) and the
// readback kinds (gpu/renderer.js:1191-1231).
#include "vt_ctx.h"

using namespace vt;

// ====================================================================== sensor
int vt::sensor_params(vt_ctx *c, const vt_sensor_params *p, uint32_t w, uint32_t h, uint32_t cam_index,
                             float scale, SensorDev &P) {
  if (!p) return fail(c, VT_ERR_ARG, "sensor: null parameters");
  if (p->cfa > 3) return fail(c, VT_ERR_ARG, "sensor: cfa code %u", p->cfa);
  if (!(p->esf_sigma >= 0) || p->esf_sigma > 3.8f) return fail(c, VT_ERR_ARG, "sensor: esf_sigma outside [0, 3.8] px");
  P.w = w; P.h = h; P.cfa = p->cfa; P.frame = p->frame; P.seed = p->seed; P.cam_index = cam_index;
  P.optics_k = p->optics_k; P.esf_sigma = p->esf_sigma; P.exposure = p->exposure; P.gain = p->gain;
  std::memcpy(P.m, p->m, sizeof P.m);
  P.read_noise = p->read_noise; P.full_well = p->full_well; P.dark = p->dark; P.conv_gain = p->conv_gain;
  P.max_dn = p->max_dn; P.black = p->black; P.noise_on = p->noise_on ? 1u : 0u;
  P.scale = scale;
  P.glare_k = p->glare_fraction;
  P.glare_sigma = p->glare_sigma;
  P.glare_f = p->glare_level;
  if (!(p->glare_fraction >= 0.0f && p->glare_fraction <= 1.0f))
    return fail(c, VT_ERR_ARG, "sensor: glare fraction outside [0, 1]");
  if (p->glare_fraction > 0.0f && (p->glare_level < 1 || p->glare_level > 64))
    return fail(c, VT_ERR_ARG, "sensor: glare level factor %u outside 1..64", p->glare_level);
  if (p->vignetting_count > VT_MAX_VIGNETTING)
    return fail(c, VT_ERR_ARG, "sensor: at most %d vignetting knots", VT_MAX_VIGNETTING);
  for (uint32_t k = 0; k < p->vignetting_count; k++)
    if (!std::isfinite(p->vignetting_theta_deg[k]) || !(p->vignetting[k] >= 0.0f) ||
        (k && !(p->vignetting_theta_deg[k] > p->vignetting_theta_deg[k - 1])))
      return fail(c, VT_ERR_ARG, "sensor: vignetting knots need ascending angles and values >= 0 (knot %u)", k);
  return VT_OK;
}
size_t vt::glare_bytes(uint32_t w, uint32_t h, const SensorDev &P) {
  if (!(P.glare_k > 0.0f)) return 0;
  const size_t gw = (w + P.glare_f - 1) / P.glare_f, gh = (h + P.glare_f - 1) / P.glare_f;
  return gw * gh * sizeof(float4);
}

// the ISP's uniform words from the raw stage's parameters and the ISP parameters (gpu/sensor.js:169-176)
int vt::isp_params(vt_ctx *c, const vt_sensor_params &sp, const vt_isp_params *q, uint32_t w, uint32_t h,
                      IspDev &P) {
  if (!q) return fail(c, VT_ERR_ARG, "isp: null parameters");
  if (q->demosaic > 2) return fail(c, VT_ERR_ARG, "isp: demosaic mode %u", q->demosaic);
  if (q->tone > 3) return fail(c, VT_ERR_ARG, "isp: tone mode %u", q->tone);
  if (!(q->sharpen_amount >= 0.0f)) return fail(c, VT_ERR_ARG, "isp: negative sharpening amount");
  P.w = w; P.h = h; P.cfa = sp.cfa; P.demosaic = q->demosaic;
  P.nr = (q->nr && q->nr_sigma_px > 0.0f) ? 1u : 0u;
  P.tone = q->tone; P.chroma420 = q->chroma420 ? 1u : 0u;
  P.max_dn = sp.max_dn; P.black = sp.black; P.read_noise = sp.read_noise; P.conv_gain = sp.conv_gain;
  P.gain = sp.gain;
  std::memcpy(P.wb, q->wb, sizeof P.wb);
  std::memcpy(P.ccm, q->ccm, sizeof P.ccm);
  P.nr_sigma = q->nr_sigma_px; P.nr_scale = q->nr_range_scale; P.gamma = q->gamma; P.s_c = q->s_curve_c;
  P.black_off = q->black_offset; P.sharpen_amount = q->sharpen_amount;
  const float rr = std::nearbyint(q->sharpen_radius);
  P.sharpen_rad = rr != rr ? 1 : (int)std::min(4.0f, std::max(1.0f, rr));
  return VT_OK;
}
static int ensure_isp(vt_ctx *c, CamSlot &s, uint32_t w, uint32_t h) {
  const size_t px = (size_t)w * h, cpx = (size_t)((w + 1) / 2) * ((h + 1) / 2);
  int rc = ensure(c, s.isp_rgb, px * 16);
  if (!rc) rc = ensure(c, s.isp_part, (size_t)isp_partials(w, h) * 16);
  if (!rc) rc = ensure(c, s.rgba, px * 4);
  if (!rc) rc = ensure(c, s.yuv, px + 2 * cpx);
  if (!rc) rc = ensure(c, s.stats, 16);
  return rc;
}

extern "C" int vt_sensor(vt_ctx *c, uint32_t cam, const vt_sensor_params *p) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "sensor: camera %u not set", cam);
  CamSlot &s = c->cams[cam];
  if (!s.rendered || s.spp_total == 0) return fail(c, VT_ERR_STATE, "sensor: camera %u has no radiance", cam);
  VT_HIP(c, hipSetDevice(c->device));
  SensorDev P;
  const float inv = s.c.scale > 0 ? s.c.scale : 1.0f;     // over each pixel's own sample count
  int rc = sensor_params(c, p, s.c.width, s.c.height, s.c.cam_index, inv, P);
  if (rc) return rc;
  const uint32_t vn = p->vignetting_count;
  const bool same_vig = s.ri_vn == vn &&
                        std::memcmp(s.ri_vt, p->vignetting_theta_deg, vn * sizeof(float)) == 0 &&
                        std::memcmp(s.ri_vv, p->vignetting, vn * sizeof(float)) == 0;
  if (!s.ri_valid || s.ri_pupil != p->pupil_exp || !same_vig) {
    std::vector<float> ri;
    ri_map_host(s.c, p->pupil_exp, vn, p->vignetting_theta_deg, p->vignetting, ri);
    VT_HIP(c, hipMemcpy(s.ri.p, ri.data(), ri.size() * 4, hipMemcpyHostToDevice));
    s.ri_valid = true;
    s.ri_pupil = p->pupil_exp;
    s.ri_vn = vn;
    std::memcpy(s.ri_vt, p->vignetting_theta_deg, sizeof s.ri_vt);
    std::memcpy(s.ri_vv, p->vignetting, sizeof s.ri_vv);
  }
  if (const size_t gb = glare_bytes(s.c.width, s.c.height, P)) {
    if ((rc = ensure(c, s.glare_a, gb)) || (rc = ensure(c, s.glare_b, gb))) return rc;
  }
  VT_HIP(c, hipEventRecord(c->ev_s0, c->stream));
  VT_HIP(c, launch_sensor(P, (const float4 *)s.acc.p, (const float *)s.ri.p, (float4 *)s.glare_a.p,
                          (float4 *)s.glare_b.p, (uint16_t *)s.dn.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_s1, c->stream));
  c->sensor_timed = true;
  s.sensed = true;
  s.isp_done = false;
  s.sp = *p;
  return VT_OK;
}

extern "C" int vt_isp(vt_ctx *c, uint32_t cam, const vt_isp_params *q) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "isp: camera %u not set", cam);
  CamSlot &s = c->cams[cam];
  if (!s.sensed) return fail(c, VT_ERR_STATE, "isp: camera %u has no raw DN (vt_sensor first)", cam);
  VT_HIP(c, hipSetDevice(c->device));
  IspDev P;
  int rc = isp_params(c, s.sp, q, s.c.width, s.c.height, P);
  if (!rc) rc = ensure_isp(c, s, s.c.width, s.c.height);
  if (rc) return rc;
  VT_HIP(c, hipEventRecord(c->ev_i0, c->stream));
  VT_HIP(c, launch_isp(P, (const uint16_t *)s.dn.p, (float4 *)s.isp_rgb.p, (uint4 *)s.isp_part.p,
                       (uchar4 *)s.rgba.p, (uint8_t *)s.yuv.p, (unsigned int *)s.stats.p, c->stream));
  VT_HIP(c, hipEventRecord(c->ev_i1, c->stream));
  c->isp_timed = true;
  s.isp_done = true;
  s.isp420 = P.chroma420 != 0;
  return VT_OK;
}

// ====================================================================== readback
extern "C" int vt_readback(vt_ctx *c, uint32_t cam, int kind, void *dst, size_t bytes) {
  VT_LIVE(c);
  if (cam >= VT_MAX_CAMERAS || !c->cams[cam].set) return fail(c, VT_ERR_STATE, "readback: camera %u not set", cam);
  if (!dst) return fail(c, VT_ERR_ARG, "readback: null destination");
  CamSlot &s = c->cams[cam];
  const size_t px = (size_t)s.c.width * s.c.height;
  VT_HIP(c, hipSetDevice(c->device));
  if (kind == VT_KIND_LINEAR) {
    if (bytes != px * 16) return fail(c, VT_ERR_SIZE, "readback linear: %zu bytes, want %zu", bytes, px * 16);
    if (!s.rendered || s.spp_total == 0) return fail(c, VT_ERR_STATE, "readback: camera %u not rendered", cam);
    const float sc = s.c.scale > 0 ? s.c.scale : 1.0f;     // over each pixel's own sample count
    VT_HIP(c, launch_finalize_linear((const float4 *)s.acc.p, sc, px, (float4 *)s.lin.p, c->stream));
    VT_HIP(c, hipMemcpyAsync(dst, s.lin.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RAW) {
    if (bytes != px * 2) return fail(c, VT_ERR_SIZE, "readback raw: %zu bytes, want %zu", bytes, px * 2);
    if (!s.sensed) return fail(c, VT_ERR_STATE, "readback raw: camera %u has no sensor output", cam);
    VT_HIP(c, hipMemcpyAsync(dst, s.dn.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RI) {
    if (bytes != px * 4) return fail(c, VT_ERR_SIZE, "readback ri: %zu bytes, want %zu", bytes, px * 4);
    if (!s.ri_valid) return fail(c, VT_ERR_STATE, "readback ri: no sensor run yet");
    VT_HIP(c, hipMemcpyAsync(dst, s.ri.p, bytes, hipMemcpyDeviceToHost, c->stream));
  } else if (kind == VT_KIND_RGB8 || kind == VT_KIND_RGBA8 || kind == VT_KIND_YUV420 || kind == VT_KIND_STATS) {
    if (!s.isp_done) return fail(c, VT_ERR_STATE, "readback: camera %u has no ISP output (vt_isp first)", cam);
    const size_t cpx = (size_t)((s.c.width + 1) / 2) * ((s.c.height + 1) / 2);
    const size_t want = kind == VT_KIND_RGB8 ? px * 3 : kind == VT_KIND_RGBA8 ? px * 4
                        : kind == VT_KIND_YUV420 ? px + 2 * cpx : 16;
    if (bytes != want) return fail(c, VT_ERR_SIZE, "readback kind %d: %zu bytes, want %zu", kind, bytes, want);
    if (kind == VT_KIND_RGB8) {
      int rc = ensure(c, s.rgb8, px * 3);
      if (rc) return rc;
      VT_HIP(c, launch_pack_rgb8((const uchar4 *)s.rgba.p, px, (uint8_t *)s.rgb8.p, c->stream));
      VT_HIP(c, hipMemcpyAsync(dst, s.rgb8.p, bytes, hipMemcpyDeviceToHost, c->stream));
    } else if (kind == VT_KIND_YUV420) {
      if (!s.isp420) return fail(c, VT_ERR_STATE, "readback yuv420: the last ISP pass ran 4:4:4");
      VT_HIP(c, hipMemcpyAsync(dst, s.yuv.p, bytes, hipMemcpyDeviceToHost, c->stream));
    } else {
      VT_HIP(c, hipMemcpyAsync(dst, kind == VT_KIND_RGBA8 ? s.rgba.p : s.stats.p, bytes, hipMemcpyDeviceToHost,
                               c->stream));
    }
  } else {
    return fail(c, VT_ERR_ARG, "readback: unknown kind %d", kind);
  }
  VT_HIP(c, hipStreamSynchronize(c->stream));
  return VT_OK;
}

