// Static-mesh preparation of the reference's prepareScene (gpu/scene_gpu.js:66-143): weld coincident vertices, then
// split edges by conforming bisection until none is longer than the node spacing. Host code in FP64, as the
// reference's plain JS arrays are: positions and normals come in as float32, every midpoint and normal is formed
// in double, and the caller rounds to float32 when it packs the scene.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/vt.h"

namespace {

// ECMAScript Math.round: the nearest integer, ties toward +infinity (floor(x + 0.5) rounds some values up wrongly)
inline int64_t js_round(double x) {
  const double f = std::floor(x);
  return (int64_t)(x - f >= 0.5 ? f + 1.0 : f);
}

struct Key {
  int64_t k[6];
  bool operator==(const Key &o) const { return std::memcmp(k, o.k, sizeof k) == 0; }
};
struct KeyHash {
  size_t operator()(const Key &a) const {
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < 6; i++) {
      h ^= (uint64_t)a.k[i];
      h *= 1099511628211ull;
    }
    return (size_t)h;
  }
};

struct Mesh {
  std::vector<double> P, N;
  std::vector<uint32_t> I;
};

// weld (gpu/scene_gpu.js:66-84): one vertex per (position at 1e-7 m, normal at 1e-4), numbered in order of first
// appearance; triangles that collapse are dropped
void weld(const float *pos, const float *nrm, uint32_t nv, const uint32_t *idx, uint32_t nt, Mesh &m) {
  std::unordered_map<Key, uint32_t, KeyHash> map;
  map.reserve(nv * 2);
  std::vector<uint32_t> remap(nv);
  for (uint32_t v = 0; v < nv; v++) {
    Key key;
    for (int c = 0; c < 3; c++) {
      key.k[c] = js_round((double)pos[3 * v + c] * 1e7);
      key.k[3 + c] = js_round((double)nrm[3 * v + c] * 1e4);
    }
    auto it = map.find(key);
    uint32_t id;
    if (it == map.end()) {
      id = (uint32_t)(m.P.size() / 3);
      map.emplace(key, id);
      for (int c = 0; c < 3; c++) {
        m.P.push_back(pos[3 * v + c]);
        m.N.push_back(nrm[3 * v + c]);
      }
    } else {
      id = it->second;
    }
    remap[v] = id;
  }
  m.I.reserve((size_t)nt * 3);
  for (uint32_t t = 0; t < nt; t++) {
    const uint32_t a = remap[idx[3 * t]], b = remap[idx[3 * t + 1]], c = remap[idx[3 * t + 2]];
    if (a == b || b == c || a == c) continue;
    m.I.push_back(a);
    m.I.push_back(b);
    m.I.push_back(c);
  }
}

// conforming edge bisection (gpu/scene_gpu.js:86-143): a pass splits every edge longer than h at its midpoint (one
// shared vertex per edge) and re-triangulates each triangle by its pattern of split edges
void subdivide(Mesh &m, double h, uint32_t max_passes) {
  const double h2 = h * h;
  std::vector<double> &P = m.P, &N = m.N;
  auto len2 = [&](uint32_t a, uint32_t b) {
    const double x = P[3 * a] - P[3 * b], y = P[3 * a + 1] - P[3 * b + 1], z = P[3 * a + 2] - P[3 * b + 2];
    return x * x + y * y + z * z;
  };
  std::vector<uint32_t> tris = m.I, out;
  for (uint32_t pass = 0; pass < max_passes; pass++) {
    std::unordered_map<uint64_t, uint32_t> mid;
    bool any = false;
    auto midpoint = [&](uint32_t a, uint32_t b) {
      const uint64_t key = a < b ? ((uint64_t)a << 32) | b : ((uint64_t)b << 32) | a;
      auto it = mid.find(key);
      if (it != mid.end()) return it->second;
      const uint32_t id = (uint32_t)(P.size() / 3);
      const double px = 0.5 * (P[3 * a] + P[3 * b]), py = 0.5 * (P[3 * a + 1] + P[3 * b + 1]),
                   pz = 0.5 * (P[3 * a + 2] + P[3 * b + 2]);
      P.push_back(px);
      P.push_back(py);
      P.push_back(pz);
      double nx = N[3 * a] + N[3 * b], ny = N[3 * a + 1] + N[3 * b + 1], nz = N[3 * a + 2] + N[3 * b + 2];
      const double l = std::hypot(nx, ny, nz);
      if (l > 1e-6) {
        nx /= l;
        ny /= l;
        nz /= l;
      } else {
        nx = N[3 * a];
        ny = N[3 * a + 1];
        nz = N[3 * a + 2];
      }
      N.push_back(nx);
      N.push_back(ny);
      N.push_back(nz);
      mid.emplace(key, id);
      return id;
    };
    out.clear();
    out.reserve(tris.size() * 2);
    auto push3 = [&](uint32_t a, uint32_t b, uint32_t c) {
      out.push_back(a);
      out.push_back(b);
      out.push_back(c);
    };
    for (size_t k = 0; k < tris.size(); k += 3) {
      const uint32_t v[3] = {tris[k], tris[k + 1], tris[k + 2]};
      const bool mk[3] = {len2(v[0], v[1]) > h2, len2(v[1], v[2]) > h2, len2(v[2], v[0]) > h2};
      const int cnt = (int)mk[0] + (int)mk[1] + (int)mk[2];
      if (!cnt) {
        push3(v[0], v[1], v[2]);
        continue;
      }
      any = true;
      if (cnt == 3) {
        const uint32_t a = midpoint(v[0], v[1]);
        const uint32_t b = midpoint(v[1], v[2]);
        const uint32_t c = midpoint(v[2], v[0]);
        push3(v[0], a, c);
        push3(a, v[1], b);
        push3(c, b, v[2]);
        push3(a, b, c);
      } else if (cnt == 1) {
        const int e = mk[0] ? 0 : mk[1] ? 1 : 2;
        const uint32_t p = v[e], q = v[(e + 1) % 3], r = v[(e + 2) % 3];
        const uint32_t mm = midpoint(p, q);
        push3(p, mm, r);
        push3(mm, q, r);
      } else {
        // the unsplit edge is (A, B) = (v[e], v[e+1]); edges BC and CA are split
        const int e = !mk[0] ? 0 : !mk[1] ? 1 : 2;
        const uint32_t A = v[e], B = v[(e + 1) % 3], C = v[(e + 2) % 3];
        const uint32_t mBC = midpoint(B, C);
        const uint32_t mCA = midpoint(C, A);
        push3(mCA, mBC, C);
        // the quad A, B, mBC, mCA along its shorter diagonal
        if (len2(A, mBC) < len2(B, mCA)) {
          push3(A, B, mBC);
          push3(A, mBC, mCA);
        } else {
          push3(A, B, mCA);
          push3(B, mBC, mCA);
        }
      }
    }
    tris.swap(out);
    if (!any) break;
  }
  m.I.swap(tris);
}

}  // namespace

extern "C" int vt_weld_subdivide(const float *pos, const float *nrm, uint32_t nvert, const uint32_t *idx,
                                 uint32_t ntri, double max_edge, uint32_t max_passes, double *pos_out,
                                 double *nrm_out, uint32_t cap_vert, uint32_t *idx_out, uint32_t cap_tri,
                                 uint32_t *nvert_out, uint32_t *ntri_out) {
  if (!pos || !nrm || !idx || !nvert_out || !ntri_out) return VT_ERR_ARG;
  if (!(max_edge > 0.0)) return VT_ERR_ARG;
  for (uint64_t k = 0; k < (uint64_t)ntri * 3; k++)
    if (idx[k] >= nvert) return VT_ERR_ARG;
  Mesh m;
  weld(pos, nrm, nvert, idx, ntri, m);
  subdivide(m, max_edge, max_passes);
  const size_t nv = m.P.size() / 3, nt = m.I.size() / 3;
  if (nv > 0xffffffffu || nt > 0xffffffffu) return VT_ERR_SIZE;
  *nvert_out = (uint32_t)nv;
  *ntri_out = (uint32_t)nt;
  if (nv > cap_vert || nt > cap_tri || !pos_out || !nrm_out || !idx_out) return VT_ERR_SIZE;
  std::memcpy(pos_out, m.P.data(), nv * 3 * sizeof(double));
  std::memcpy(nrm_out, m.N.data(), nv * 3 * sizeof(double));
  std::memcpy(idx_out, m.I.data(), nt * 3 * sizeof(uint32_t));
  return VT_OK;
}
