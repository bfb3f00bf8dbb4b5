// The C ABI, part 6: the lite renderer's radiosity on its baked node table (lite/renderer.js:214-222, 486-517;
// lite/shaders.js:220-283): the table stays on the device; per pose change the gather occlusion by the posed
// capsules, then warm-started Jacobi steps.
#include "vt_ctx.h"

using namespace vt;

namespace {

BakedDev baked_dev(vt_ctx *c) {
  BakedDev B{};
  B.nodes = c->bk_nodes;
  B.K = c->bk_K;
  B.n_caps = c->bk_ncaps;
  B.gather = (const uint16_t *)c->bk_gather.p;
  B.node_a = (const float4 *)c->bk_a.p;
  B.node_n = (const float4 *)c->bk_n.p;
  B.hh = (const float4 *)c->bk_hh.p;
  B.caps = (const float4 *)c->bk_caps.p;
  B.occ = (uint32_t *)c->bk_occ.p;
  B.e_dir = (const float4 *)c->bk_edir.p;
  B.e_ind = (float4 *)c->bk_eind.p;
  return B;
}

}  // namespace

extern "C" int vt_baked_upload(vt_ctx *c, const vt_baked_table *t) {
  VT_LIVE(c);
  if (!t || !t->gather || !t->node_a || !t->node_n || !t->hh) return fail(c, VT_ERR_ARG, "baked: null table");
  if (t->nodes == 0 || t->nodes >= 0xffffu || t->K == 0 || t->K % 4 || t->K > VT_BAKED_MAX_K || t->meshes == 0)
    return fail(c, VT_ERR_ARG, "baked: 1..65534 nodes, K a multiple of 4 up to %d, meshes > 0", VT_BAKED_MAX_K);
  const size_t N = t->nodes, K = t->K;
  for (size_t k = 0; k < N * K; k++)
    if (t->gather[k] != 0xffffu && t->gather[k] >= N) return fail(c, VT_ERR_ARG, "baked: gather entry %zu out of range", k);
  for (size_t i = 0; i < N; i++) {
    uint32_t m;
    std::memcpy(&m, t->node_n + 4 * i + 3, 4);
    if (m >= t->meshes) return fail(c, VT_ERR_ARG, "baked: node %zu mesh out of range", i);
  }
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  int rc;
  if ((rc = upload(c, c->bk_gather, t->gather, N * K * 2))) return rc;
  if ((rc = upload(c, c->bk_a, t->node_a, N * 16))) return rc;
  if ((rc = upload(c, c->bk_n, t->node_n, N * 16))) return rc;
  if ((rc = upload(c, c->bk_hh, t->hh, (size_t)t->meshes * 16))) return rc;
  if ((rc = ensure(c, c->bk_occ, N * K))) return rc;
  if ((rc = ensure(c, c->bk_edir, N * 16))) return rc;
  if ((rc = ensure(c, c->bk_eind, N * 16))) return rc;
  for (DevBuf &b : c->bk_b)
    if ((rc = ensure(c, b, N * 16))) return rc;
  if ((rc = ensure(c, c->bk_caps, VT_BAKED_MAX_CAPS * 32))) return rc;
  c->bk_nodes = (uint32_t)N;
  c->bk_K = (uint32_t)K;
  c->bk_meshes = t->meshes;
  c->bk_ncaps = 0;
  c->bk_cur = 0;
  // every ray transmits until vt_baked_occlusion; no light yet
  if (N * K >= 16)
    VT_HIP(c, launch_fill_u4((uint4 *)c->bk_occ.p, N * K / 16, make_uint4(~0u, ~0u, ~0u, ~0u), c->stream));
  if ((N * K) % 16) {
    std::vector<uint8_t> tail((N * K) % 16, 255);
    VT_HIP(c, hipStreamSynchronize(c->stream));
    VT_HIP(c, hipMemcpy((uint8_t *)c->bk_occ.p + (N * K / 16) * 16, tail.data(), tail.size(), hipMemcpyHostToDevice));
  }
  for (DevBuf *b : {&c->bk_edir, &c->bk_eind, &c->bk_b[0], &c->bk_b[1]})
    VT_HIP(c, launch_zero_f4((float4 *)b->p, N, c->stream));
  return VT_OK;
}

extern "C" int vt_baked_occlusion(vt_ctx *c, const float *caps, uint32_t n_caps) {
  VT_LIVE(c);
  if (!c->bk_nodes) return fail(c, VT_ERR_STATE, "baked: no table uploaded");
  if (n_caps > VT_BAKED_MAX_CAPS || (n_caps && !caps))
    return fail(c, VT_ERR_ARG, "baked: at most %d capsules", VT_BAKED_MAX_CAPS);
  for (uint32_t i = 0; i < n_caps; i++)
    for (int k = 0; k < 8; k++)
      if (!std::isfinite(caps[8 * i + k])) return fail(c, VT_ERR_ARG, "baked: capsule %u not finite", i);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  if (n_caps) VT_HIP(c, hipMemcpy(c->bk_caps.p, caps, (size_t)n_caps * 32, hipMemcpyHostToDevice));
  c->bk_ncaps = n_caps;
  VT_HIP(c, launch_baked_occ(baked_dev(c), c->stream));
  return VT_OK;
}

extern "C" int vt_baked_jacobi(vt_ctx *c, const float *e_dir, uint32_t iters) {
  VT_LIVE(c);
  if (!c->bk_nodes) return fail(c, VT_ERR_STATE, "baked: no table uploaded");
  VT_HIP(c, hipSetDevice(c->device));
  if (e_dir) {
    VT_HIP(c, hipStreamSynchronize(c->stream));
    VT_HIP(c, hipMemcpy(c->bk_edir.p, e_dir, (size_t)c->bk_nodes * 16, hipMemcpyHostToDevice));
  }
  const BakedDev B = baked_dev(c);
  for (uint32_t it = 0; it < iters; it++) {
    VT_HIP(c, launch_baked_jacobi(B, (const float4 *)c->bk_b[c->bk_cur].p, (float4 *)c->bk_b[c->bk_cur ^ 1].p,
                                  c->stream));
    c->bk_cur ^= 1;
  }
  return VT_OK;
}

extern "C" int vt_baked_read(vt_ctx *c, uint32_t which, void *dst, size_t bytes) {
  VT_LIVE(c);
  if (!c->bk_nodes) return fail(c, VT_ERR_STATE, "baked: no table uploaded");
  const void *src = nullptr;
  size_t want = 0;
  switch (which) {
    case VT_BAKED_B: src = c->bk_b[c->bk_cur].p; want = (size_t)c->bk_nodes * 16; break;
    case VT_BAKED_EIND: src = c->bk_eind.p; want = (size_t)c->bk_nodes * 16; break;
    case VT_BAKED_OCC: src = c->bk_occ.p; want = (size_t)c->bk_nodes * c->bk_K; break;
    default: return fail(c, VT_ERR_ARG, "baked: unknown read kind %u", which);
  }
  if (!dst || bytes != want) return fail(c, VT_ERR_SIZE, "baked: read needs %zu bytes", want);
  VT_HIP(c, hipSetDevice(c->device));
  VT_HIP(c, hipStreamSynchronize(c->stream));
  VT_HIP(c, hipMemcpy(dst, src, want, hipMemcpyDeviceToHost));
  return VT_OK;
}
