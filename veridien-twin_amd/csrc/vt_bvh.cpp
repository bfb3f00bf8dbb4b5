// Host BVH work: the reference's binned-SAH BVH2 builder (native, for the 2M-triangle scenes) and the
// re-layout of a BVH2 into the device's 128-B four-wide nodes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "vt_internal.h"

namespace vt {
namespace {

// JS Math.min / Math.max on finite values (-0 below +0): the reference computes its boxes with them
inline double jmin(double a, double b) {
  if (a < b) return a;
  if (b < a) return b;
  return std::signbit(a) ? a : b;
}
inline double jmax(double a, double b) {
  if (a > b) return a;
  if (b > a) return b;
  return std::signbit(a) ? b : a;
}

struct Box3 {
  double lo[3], hi[3], clo[3], chi[3];
};

// The SAH split search and partition follow gpu/bvh.js:66-116 step for step (bin formula, the right-to-left
// then left-to-right sweeps, strict-less cost tie-break, the two-pointer swap partition, the middle split of
// coincident centroids) so the tree, the triangle order and the node numbering equal the reference's.
class Sah2Builder {
 public:
  Sah2Builder(const float *pos, const uint32_t *idx, uint32_t n, int leaf, int bins)
      : n_(n), leaf_(leaf), bins_(bins), lo_(3 * (size_t)n), hi_(3 * (size_t)n), c_(3 * (size_t)n),
        cnt_(bins), rcnt_(bins), blo_(3 * bins), bhi_(3 * bins), rarea_(bins) {
    for (uint32_t t = 0; t < n; t++) {
      const uint32_t v[3] = {idx[3 * t] * 3, idx[3 * t + 1] * 3, idx[3 * t + 2] * 3};
      for (int k = 0; k < 3; k++) {
        double a = pos[v[0] + k], b = pos[v[1] + k], c = pos[v[2] + k];
        double l = jmin(jmin(a, b), c), h = jmax(jmax(a, b), c);
        lo_[3 * t + k] = (float)l;
        hi_[3 * t + k] = (float)h;
        c_[3 * t + k] = (float)(0.5 * (l + h));
      }
    }
  }

  int64_t run(float *nodes, uint32_t cap, uint32_t *order) {
    order_ = order;
    for (uint32_t i = 0; i < n_; i++) order[i] = i;
    uint32_t count = 0;
    auto alloc = [&](uint32_t &out) {
      if (count >= cap) return false;
      out = count++;
      return true;
    };
    uint32_t root;
    if (!alloc(root)) return -1;
    Box3 R = bounds(0, n_);
    int64_t m0 = n_ ? split(0, n_, R) : -1;
    struct Task {
      uint32_t node;
      int64_t s, e, m;
    };
    std::vector<Task> stack;
    if (n_ == 0) {
      const double far[3] = {1e30, 1e30, 1e30};
      put(nodes, root, 0, far, far, 0, 1);
      put(nodes, root, 1, far, far, 0, 1);
    } else if (m0 < 0) {
      int64_t h = n_ > 1 ? n_ >> 1 : n_;
      Box3 A = bounds(0, h);
      put(nodes, root, 0, A.lo, A.hi, 0, (uint32_t)h);
      if (n_ > 1) {
        Box3 B = bounds(h, n_);
        put(nodes, root, 1, B.lo, B.hi, (uint32_t)h, (uint32_t)(n_ - h));
      } else {
        put(nodes, root, 1, R.lo, R.hi, 0, n_);
      }
    } else {
      stack.push_back({root, 0, (int64_t)n_, m0});
    }
    while (!stack.empty()) {
      Task t = stack.back();
      stack.pop_back();
      const int64_t a[2] = {t.s, t.m}, b[2] = {t.m, t.e};
      for (int slot = 0; slot < 2; slot++) {
        Box3 C = bounds(a[slot], b[slot]);
        int64_t mm = split(a[slot], b[slot], C);
        if (mm < 0) {
          put(nodes, t.node, slot, C.lo, C.hi, (uint32_t)a[slot], (uint32_t)(b[slot] - a[slot]));
        } else {
          uint32_t child;
          if (!alloc(child)) return -1;
          put(nodes, t.node, slot, C.lo, C.hi, child, 0);
          stack.push_back({child, a[slot], b[slot], mm});
        }
      }
    }
    return count;
  }

 private:
  static void put(float *nodes, uint32_t node, int slot, const double *lo, const double *hi, uint32_t ref,
                  uint32_t cnt) {
    float *o = nodes + (size_t)node * VT_NODE_WORDS_ + slot * 8;
    o[0] = (float)lo[0]; o[1] = (float)lo[1]; o[2] = (float)lo[2];
    o[4] = (float)hi[0]; o[5] = (float)hi[1]; o[6] = (float)hi[2];
    std::memcpy(o + 3, &ref, 4);
    std::memcpy(o + 7, &cnt, 4);
  }
  static constexpr int VT_NODE_WORDS_ = 16;

  Box3 bounds(int64_t s, int64_t e) const {
    Box3 B;
    for (int k = 0; k < 3; k++) { B.lo[k] = 1e30; B.hi[k] = -1e30; B.clo[k] = 1e30; B.chi[k] = -1e30; }
    for (int64_t i = s; i < e; i++) {
      const uint32_t t = order_[i];
      for (int k = 0; k < 3; k++) {
        double a = lo_[3 * t + k], b = hi_[3 * t + k], c = c_[3 * t + k];
        if (a < B.lo[k]) B.lo[k] = a;
        if (b > B.hi[k]) B.hi[k] = b;
        if (c < B.clo[k]) B.clo[k] = c;
        if (c > B.chi[k]) B.chi[k] = c;
      }
    }
    return B;
  }
  static double area(const double *lo, const double *hi) {
    double dx = jmax(0, hi[0] - lo[0]), dy = jmax(0, hi[1] - lo[1]), dz = jmax(0, hi[2] - lo[2]);
    return dx * dy + dy * dz + dz * dx;
  }
  int bin(double c, double lo, double sc) const {
    int b = (int)std::floor((c - lo) * sc);
    return b >= bins_ ? bins_ - 1 : b;
  }

  int64_t split(int64_t s, int64_t e, const Box3 &B) {
    const int64_t cnt = e - s;
    if (cnt <= 1) return -1;
    double best = INFINITY;
    int best_ax = -1, best_bin = -1;
    const double pA = area(B.lo, B.hi);
    for (int ax = 0; ax < 3; ax++) {
      const double lo = B.clo[ax], ext = B.chi[ax] - lo;
      if (!(ext > 0)) continue;
      std::fill(cnt_.begin(), cnt_.end(), 0u);
      std::fill(blo_.begin(), blo_.end(), 1e30);
      std::fill(bhi_.begin(), bhi_.end(), -1e30);
      const double sc = bins_ / ext;
      for (int64_t i = s; i < e; i++) {
        const uint32_t t = order_[i];
        const int b = bin(c_[3 * t + ax], lo, sc);
        cnt_[b]++;
        for (int k = 0; k < 3; k++) {
          if (lo_[3 * t + k] < blo_[3 * b + k]) blo_[3 * b + k] = lo_[3 * t + k];
          if (hi_[3 * t + k] > bhi_[3 * b + k]) bhi_[3 * b + k] = hi_[3 * t + k];
        }
      }
      double l3[3] = {1e30, 1e30, 1e30}, h3[3] = {-1e30, -1e30, -1e30};
      uint32_t c = 0;
      for (int b = bins_ - 1; b > 0; b--) {
        c += cnt_[b];
        for (int k = 0; k < 3; k++) { l3[k] = jmin(l3[k], blo_[3 * b + k]); h3[k] = jmax(h3[k], bhi_[3 * b + k]); }
        rcnt_[b] = c;
        rarea_[b] = c ? area(l3, h3) : 0;
      }
      for (int k = 0; k < 3; k++) { l3[k] = 1e30; h3[k] = -1e30; }
      c = 0;
      for (int b = 0; b < bins_ - 1; b++) {
        c += cnt_[b];
        for (int k = 0; k < 3; k++) { l3[k] = jmin(l3[k], blo_[3 * b + k]); h3[k] = jmax(h3[k], bhi_[3 * b + k]); }
        if (c == 0 || rcnt_[b + 1] == 0) continue;
        const double cost = 1.0 + 1.0 * ((double)c * area(l3, h3) + (double)rcnt_[b + 1] * rarea_[b + 1]) / pA;
        if (cost < best) { best = cost; best_ax = ax; best_bin = b; }
      }
    }
    if (best_ax < 0) return cnt > (int64_t)leaf_ * 4 ? s + (cnt >> 1) : -1;
    if (cnt <= leaf_ && best >= (double)cnt) return -1;
    const double lo = B.clo[best_ax], sc = bins_ / (B.chi[best_ax] - lo);
    int64_t i = s, j = e - 1;
    while (i <= j) {
      const uint32_t t = order_[i];
      if (bin(c_[3 * t + best_ax], lo, sc) <= best_bin) {
        i++;
      } else {
        order_[i] = order_[j];
        order_[j] = t;
        j--;
      }
    }
    if (i == s || i == e) return s + (cnt >> 1);
    return i;
  }

  uint32_t n_;
  int leaf_, bins_;
  std::vector<float> lo_, hi_, c_;
  std::vector<uint32_t> cnt_, rcnt_;
  std::vector<double> blo_, bhi_, rarea_;
  uint32_t *order_ = nullptr;
};

struct Child2 {
  float lo[3], hi[3];
  uint32_t ref, count;  // count > 0: leaf
  double area() const {
    double dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
    return dx < 0 ? 0 : dx * dy + dy * dz + dz * dx;
  }
};

Child2 child_of(const float *n2, uint32_t node, int slot) {
  const float *o = n2 + (size_t)node * 16 + slot * 8;
  Child2 c;
  for (int k = 0; k < 3; k++) { c.lo[k] = o[k]; c.hi[k] = o[4 + k]; }
  std::memcpy(&c.ref, o + 3, 4);
  std::memcpy(&c.count, o + 7, 4);
  return c;
}

}  // namespace

int64_t build_bvh2(const float *pos, const uint32_t *idx, uint32_t ntri, int leaf, int bins, float *nodes,
                   uint32_t cap_nodes, uint32_t *order) {
  Sah2Builder B(pos, idx, ntri, leaf, bins);
  return B.run(nodes, cap_nodes, order);
}

// Four-wide re-layout: a node's two children are opened, largest surface area first, until it has four
// children or only leaves; nodes are emitted depth-first so a subtree stays contiguous in memory.
std::vector<Node4> collapse_bvh4(const float *n2, uint32_t node_count, uint32_t tri_count) {
  std::vector<Node4> out;
  if (node_count == 0) throw std::runtime_error("BVH has no nodes");
  out.reserve(node_count / 2 + 1);
  struct Item {
    uint32_t node2;   // BVH2 node whose children seed this four-wide node
    uint32_t slot4;   // index in out
  };
  std::vector<Item> work;
  out.emplace_back();
  work.push_back({0, 0});
  std::vector<uint8_t> seen(node_count, 0);
  while (!work.empty()) {
    Item it = work.back();
    work.pop_back();
    if (it.node2 >= node_count || seen[it.node2]) throw std::runtime_error("malformed BVH (cycle or bad index)");
    seen[it.node2] = 1;
    Child2 ch[4];
    int nc = 2;
    ch[0] = child_of(n2, it.node2, 0);
    ch[1] = child_of(n2, it.node2, 1);
    while (nc < 4) {
      int best = -1;
      double ba = -1;
      for (int c = 0; c < nc; c++)
        if (ch[c].count == 0 && ch[c].area() > ba) { ba = ch[c].area(); best = c; }
      if (best < 0) break;
      uint32_t inner = ch[best].ref;
      if (inner >= node_count || inner <= it.node2 || seen[inner])
        throw std::runtime_error("malformed BVH (child not after its parent)");
      seen[inner] = 1;
      Child2 a = child_of(n2, inner, 0), b = child_of(n2, inner, 1);
      ch[best] = a;
      ch[nc++] = b;
    }
    Node4 nd;
    for (int c = 0; c < 4; c++) {
      nd.pad[c] = 0;
      if (c >= nc) {
        nd.lox[c] = nd.loy[c] = nd.loz[c] = INFINITY;
        nd.hix[c] = nd.hiy[c] = nd.hiz[c] = -INFINITY;
        nd.ref[c] = EMPTY_REF;
        continue;
      }
      nd.lox[c] = ch[c].lo[0]; nd.loy[c] = ch[c].lo[1]; nd.loz[c] = ch[c].lo[2];
      nd.hix[c] = ch[c].hi[0]; nd.hiy[c] = ch[c].hi[1]; nd.hiz[c] = ch[c].hi[2];
      if (ch[c].count > 0) {
        if (ch[c].count > 32) throw std::runtime_error("BVH leaf with more than 32 triangles");
        if ((uint64_t)ch[c].ref + ch[c].count > tri_count && tri_count > 0)
          throw std::runtime_error("BVH leaf outside the triangle range");
        if (ch[c].ref >= (1u << 26)) throw std::runtime_error("more than 2^26 triangles");
        nd.ref[c] = tri_count == 0 ? EMPTY_REF : (LEAF_BIT | ((ch[c].count - 1) << 26) | ch[c].ref);
        if (tri_count == 0) {
          nd.lox[c] = nd.loy[c] = nd.loz[c] = INFINITY;
          nd.hix[c] = nd.hiy[c] = nd.hiz[c] = -INFINITY;
        }
      } else {
        uint32_t slot = (uint32_t)out.size();
        out.emplace_back();
        nd.ref[c] = slot;
        work.push_back({ch[c].ref, slot});
      }
    }
    out[it.slot4] = nd;
  }
  return out;
}

}  // namespace vt
