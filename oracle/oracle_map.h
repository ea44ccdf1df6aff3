/*
 * CPU ORACLE (test infrastructure) — sparse octree, allocation, integration,
 * raycasting.  Restates lib/octree.hpp, lib/node.hpp, lib/octant_ops.hpp,
 * lib/utils/morton_utils.hpp, lib/algorithms/{unique,filter}.hpp,
 * lib/interpolation/interp_gather.hpp, lib/functors/projective_functor.hpp,
 * apps/kfusion/src/cpp/kfusion/{alloc_impl,mapping_impl,rendering_impl}.hpp
 * and rendering/rendering.cpp:9-181.
 */
#pragma once
#include "oracle_core.h"

namespace orc {

// Host fminf/fmaxf of utils/cutil_math.h:43-49 (a<b?a:b / a>b?a:b; these
// differ from C99 fminf/fmaxf when an argument is NaN).
static inline R fmin_r(R a, R b) { return a < b ? a : b; }
static inline R fmax_r(R a, R b) { return a > b ? a : b; }
static inline float fminf_r(float a, float b) { return a < b ? a : b; }
static inline float fmaxf_r(float a, float b) { return a > b ? a : b; }

// ---- Morton keys (morton_utils.hpp:38-73, octant_ops.hpp:51-55,
// octree_defines.h:38-80)
static const uint64_t SCALE_MASK = 0x1FF;
static const int MAX_BITS = 21;
static inline uint64_t mask_at(int i) {
  uint64_t m = 0x7000000000000000ULL, base = 0x7000000000000000ULL;
  for (int k = 1; k <= i; k++) m |= base >> (3 * k);
  return m;
}
static inline uint64_t expand(uint64_t v) {
  uint64_t x = v & 0x1fffff;
  x = (x | x << 32) & 0x1f00000000ffffULL;
  x = (x | x << 16) & 0x1f0000ff0000ffULL;
  x = (x | x << 8) & 0x100f00f00f00f00fULL;
  x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
  x = (x | x << 2) & 0x1249249249249249ULL;
  return x;
}
static inline uint64_t compact(uint64_t v) {
  uint64_t x = v & 0x1249249249249249ULL;
  x = (x | x >> 2) & 0x10c30c30c30c30c3ULL;
  x = (x | x >> 4) & 0x100f00f00f00f00fULL;
  x = (x | x >> 8) & 0x1f0000ff0000ffULL;
  x = (x | x >> 16) & 0x1f00000000ffffULL;
  x = (x | x >> 32) & 0x1fffffULL;
  return x;
}
static inline uint64_t morton(uint64_t x, uint64_t y, uint64_t z) { return expand(x) | (expand(y) << 1) | (expand(z) << 2); }
static inline void unpack(uint64_t c, int& x, int& y, int& z) { x = (int)compact(c); y = (int)compact(c >> 1); z = (int)compact(c >> 2); }
static inline uint64_t encode(int x, int y, int z, int level, int max_depth) {
  return (morton(x, y, z) & mask_at(MAX_BITS - max_depth + level - 1)) | (uint64_t)level;
}
static inline int child_id(uint64_t key, int level, int max_depth) {
  uint64_t c = (key & ~SCALE_MASK) >> (3 * (max_depth - level));
  return (int)(c & 7);
}
// unique_multiscale (unique.hpp:50-98); keys[0] is always kept.
static inline int unique_ms(uint64_t* keys, int n) {
  int end = 1;
  if (n < 2) return end;
  uint64_t pk = keys[0] & ~SCALE_MASK, pl = keys[0] & SCALE_MASK;
  for (int i = 1; i < n; i++) {
    uint64_t k = keys[i] & ~SCALE_MASK, l = keys[i] & SCALE_MASK;
    if (k != pk) keys[end++] = keys[i];
    else if (l > pl) keys[end - 1] = keys[i];
    pk = k; pl = l;
  }
  return end;
}
static inline int unique_ms_level(uint64_t* keys, int n, unsigned cur) {
  int end = 1;
  if (n < 2) return end;
  uint64_t pk = keys[0] & ~SCALE_MASK, pl = keys[0] & SCALE_MASK;
  for (int i = 1; i < n; i++) {
    uint64_t k = keys[i] & ~SCALE_MASK, l = keys[i] & SCALE_MASK;
    if (l >= cur) {
      if (k != pk) keys[end++] = keys[i];
      else if (l > pl) keys[end - 1] = keys[i];
      pk = k; pl = l;
    }
  }
  return end;
}

// ---- Voxel (continuous/volume_traits.hpp:26-40): tsdf x, weight y, rgb,
// colour votes c, foreground probability fg (double), fg witness w.
struct Voxel { R x, y, r, g, b; int c; double fg; int w; };
static inline Voxel voxel_init() { Voxel v; v.x = 1; v.y = 0; v.r = v.g = v.b = 0; v.c = 0; v.fg = 0; v.w = 0; return v; }

struct Node {
  Node* child[8];
  uint64_t code;
  unsigned side;
  bool leaf;
  // leaf payload
  int cx, cy, cz;
  bool active;
  Voxel* vox;   // 512 voxels, x + 8y + 64z (node.hpp:145-157)
  Node() : code(0), side(0), leaf(false), cx(0), cy(0), cz(0), active(false), vox(nullptr) {
    for (int i = 0; i < 8; i++) child[i] = nullptr;
  }
};

struct Octree {
  int size = 0;      // voxels per side
  R dim = 0;         // metres per side
  int max_level = 0;
  Node* root = nullptr;
  std::vector<Node*> blocks, nodes;

  void init(int s, R d) {
    size = s; dim = d; max_level = 0; while ((1 << max_level) < s) max_level++;
    root = new Node(); root->side = s;
  }
  void release() {
    for (Node* b : blocks) { delete[] b->vox; delete b; }
    for (Node* n : nodes) delete n;
    delete root; root = nullptr; blocks.clear(); nodes.clear();
  }
  int leaf_level() const { return max_level - 3; }
  // Octree::fetch (octree.hpp:439-457).  Coordinates outside [0,size) are
  // reported missing (the reference aliases them through its bit masks).
  Node* fetch(int x, int y, int z) const {
    if (x < 0 || y < 0 || z < 0 || x >= size || y >= size || z >= size) return nullptr;
    Node* n = root;
    for (unsigned edge = size / 2; edge >= 8; edge /= 2) {
      n = n->child[((x & edge) > 0u) + 2 * ((y & edge) > 0u) + 4 * ((z & edge) > 0u)];
      if (!n) return nullptr;
    }
    return n;
  }
  // Octree::get_fine (octree.hpp:359-379): init value if the block is missing.
  Voxel get_fine(int x, int y, int z) const {
    Node* b = fetch(x, y, z);
    if (!b) return voxel_init();
    return b->vox[(x - b->cx) + 8 * (y - b->cy) + 64 * (z - b->cz)];
  }
  // Octree::updateLevel (octree.hpp:826-863)
  void update_level(const uint64_t* keys, int n, int target) {
    const int leaves = leaf_level();
    for (int i = 0; i < n; i++) {
      Node** np = &root;
      uint64_t k = keys[i] & ~SCALE_MASK;
      int edge = size / 2;
      for (int level = 1; level <= target; ++level) {
        int idx = child_id(k, level, max_level);
        Node* parent = *np;
        np = &parent->child[idx];
        if (!*np) {
          Node* c = new Node();
          if (level == leaves) {
            c->leaf = true;
            int x, y, z; unpack(k, x, y, z);
            c->cx = x; c->cy = y; c->cz = z;
            c->active = true;
            c->code = k | (uint64_t)level;
            c->vox = new Voxel[512];
            for (int v = 0; v < 512; v++) c->vox[v] = voxel_init();
            blocks.push_back(c);
          } else {
            c->code = k | (uint64_t)level;
            c->side = edge;
            nodes.push_back(c);
          }
          *np = c;
        }
        edge /= 2;
      }
    }
  }
  // Octree::alloc_update (octree.hpp:761-786)
  void alloc_update(uint64_t* keys, int n) {
    std::sort(keys, keys + n);
    n = unique_ms(keys, n);
    std::vector<uint64_t> lvl((size_t)std::max(n, 1));
    const int leaves = leaf_level();
    const int shift = MAX_BITS - max_level - 1;
    for (int level = 1; level <= leaves; level++) {
      const uint64_t m = mask_at(level + shift) | SCALE_MASK;
      for (int i = 0; i < n; i++) lvl[i] = keys[i] & m;
      int last = unique_ms_level(lvl.data(), n, (unsigned)level);
      update_level(lvl.data(), last, level);
    }
  }
};

// ---- trilinear interpolation (octree.hpp:479-501, interp_gather.hpp:107-237).
// A missing block reads as tsdf 1, fg 0 in every gather case.
enum Field { F_TSDF = 0, F_FG = 1, F_R = 2, F_G = 3, F_B = 4 };
static inline R sel(const Voxel& v, int f) {
  switch (f) {
    case F_TSDF: return v.x;
    case F_FG: return (R)v.fg;
    case F_R: return v.r;
    case F_G: return v.g;
    default: return v.b;
  }
}
static inline R interp(const Octree& o, V3 pos, int f) {
  const int bx = (int)std::floor(pos.x), by = (int)std::floor(pos.y), bz = (int)std::floor(pos.z);
  const R fx = pos.x - std::floor(pos.x), fy = pos.y - std::floor(pos.y), fz = pos.z - std::floor(pos.z);
  const int lx = std::max(bx, 0), ly = std::max(by, 0), lz = std::max(bz, 0);
  R p[8];
  for (int i = 0; i < 8; i++) p[i] = sel(o.get_fine(lx + (i & 1), ly + ((i >> 1) & 1), lz + ((i >> 2) & 1)), f);
  return (((p[0] * (1 - fx) + p[1] * fx) * (1 - fy) + (p[2] * (1 - fx) + p[3] * fx) * fy) * (1 - fz) +
          ((p[4] * (1 - fx) + p[5] * fx) * (1 - fy) + (p[6] * (1 - fx) + p[7] * fx) * fy) * fz);
}
// Octree::grad(pos, select) (octree.hpp:593-680), tsdf field.
static inline V3 grad(const Octree& o, V3 pos) {
  const int b[3] = {(int)std::floor(pos.x), (int)std::floor(pos.y), (int)std::floor(pos.z)};
  const R f[3] = {pos.x - std::floor(pos.x), pos.y - std::floor(pos.y), pos.z - std::floor(pos.z)};
  int ll[3], lu[3], ul[3], uu[3];
  for (int a = 0; a < 3; a++) {
    ll[a] = std::max(b[a] - 1, 0); lu[a] = std::max(b[a], 0);
    ul[a] = std::min(b[a] + 1, o.size - 1); uu[a] = std::min(b[a] + 2, o.size - 1);
  }
  // lower = lower_upper, upper = upper_lower (octree.hpp:604-605)
  auto G = [&](int x, int y, int z) -> R { return (R)o.get_fine(x, y, z).x; };
  R g[3];
  for (int a = 0; a < 3; a++) {
    auto D = [&](int ix, int iy, int iz) -> R {
      int id[3] = {ix, iy, iz};
      int hi[3], lo[3];
      for (int c = 0; c < 3; c++) {
        if (c == a) { hi[c] = id[c] ? uu[c] : ul[c]; lo[c] = id[c] ? lu[c] : ll[c]; }
        else { hi[c] = lo[c] = id[c] ? ul[c] : lu[c]; }
      }
      return G(hi[0], hi[1], hi[2]) - G(lo[0], lo[1], lo[2]);
    };
    g[a] = (((D(0, 0, 0) * (1 - f[0]) + D(1, 0, 0) * f[0]) * (1 - f[1]) + (D(0, 1, 0) * (1 - f[0]) + D(1, 1, 0) * f[0]) * f[1]) * (1 - f[2]) +
            ((D(0, 0, 1) * (1 - f[0]) + D(1, 0, 1) * f[0]) * (1 - f[1]) + (D(0, 1, 1) * (1 - f[0]) + D(1, 1, 1) * f[0]) * f[1]) * f[2]);
  }
  const R s = o.dim / (R)o.size;
  return v3(g[0] * s * R(0.5), g[1] * s * R(0.5), g[2] * s * R(0.5));
}

// ---- ray_iterator (octree.hpp:919-1130): ESVO traversal of the octree,
// scaled to [1,2]^3, to the first allocated leaf block.  Always in float:
// the traversal relies on float bit patterns.
static inline int f2i(float f) { int i; std::memcpy(&i, &f, 4); return i; }
static inline float i2f(int i) { float f; std::memcpy(&f, &i, 4); return f; }
struct RayIter {
  static const int DEPTH = 23;   // CAST_STACK_DEPTH (octree_defines.h:40)
  struct Entry { int scale; const Node* parent; float t_max; };
  const Octree* map;
  Entry stack[DEPTH + 1];   // one spare: a pop past the root reads index DEPTH
  const Node* parent; const Node* child;
  int idx, scale, min_scale, octant_mask;
  float scale_exp2, t_min, t_max, tc_max, h;
  float pos[3], t_coef[3], t_bias[3], t_corner[3];

  RayIter(const Octree& m, const float o[3], const float d_in[3], float nearp, float farp, float eps) {
    map = &m;
    for (int i = 0; i <= DEPTH; i++) { stack[i].scale = 0; stack[i].parent = nullptr; stack[i].t_max = 0; }
    pos[0] = pos[1] = pos[2] = 1.0f;
    idx = 0; parent = m.root; child = nullptr; scale_exp2 = 0.5f; scale = DEPTH - 1;
    int lg = 0; while ((8 << lg) < m.size) lg++;
    min_scale = DEPTH - lg;
    float d[3];
    for (int a = 0; a < 3; a++) d[a] = std::fabs(d_in[a]) < eps ? std::copysign(eps, d_in[a]) : d_in[a];
    const float dimf = (float)m.dim;
    for (int a = 0; a < 3; a++) {
      float so = o[a] / dimf + 1.f;
      t_coef[a] = -1.f / std::fabs(d[a]);
      t_bias[a] = t_coef[a] * so;
    }
    octant_mask = 7;
    for (int a = 0; a < 3; a++)
      if (d[a] > 0.0f) { octant_mask ^= (1 << a); t_bias[a] = 3.0f * t_coef[a] - t_bias[a]; }
    t_min = fmaxf_r(fmaxf_r(2.0f * t_coef[0] - t_bias[0], 2.0f * t_coef[1] - t_bias[1]), 2.0f * t_coef[2] - t_bias[2]);
    t_max = fminf_r(fminf_r(t_coef[0] - t_bias[0], t_coef[1] - t_bias[1]), t_coef[2] - t_bias[2]);
    h = t_max;
    t_min = fmaxf_r(t_min, nearp / dimf);
    t_max = fminf_r(t_max, farp / dimf);
    for (int a = 0; a < 3; a++)
      if (1.5f * t_coef[a] - t_bias[a] > t_min) { idx ^= (1 << a); pos[a] = 1.5f; }
    tc_max = 0; h = h; t_corner[0] = t_corner[1] = t_corner[2] = 0;
  }
  void advance() {
    int step_mask = (t_corner[0] <= tc_max) | ((t_corner[1] <= tc_max) << 1) | ((t_corner[2] <= tc_max) << 2);
    for (int a = 0; a < 3; a++) pos[a] -= scale_exp2 * (float)(bool)(step_mask & (1 << a));
    t_min = tc_max;
    idx ^= step_mask;
    if ((idx & step_mask) != 0) {
      unsigned diff = 0;
      for (int a = 0; a < 3; a++)
        if (step_mask & (1 << a)) diff |= (unsigned)(f2i(pos[a]) ^ f2i(pos[a] + scale_exp2));
      scale = (f2i((float)diff) >> 23) - 127;
      scale_exp2 = i2f((scale - DEPTH + 127) << 23);
      if (scale < 0 || scale > DEPTH) { parent = nullptr; scale = DEPTH; return; }
      const Entry& e = stack[scale];
      parent = e.parent;
      t_max = e.t_max;
      int sh[3];
      for (int a = 0; a < 3; a++) { sh[a] = f2i(pos[a]) >> scale; pos[a] = i2f(sh[a] << scale); }
      idx = (sh[0] & 1) | ((sh[1] & 1) << 1) | ((sh[2] & 1) << 2);
      h = 0.0f;
      child = nullptr;
    }
  }
  void descend() {
    float tv_max = fminf_r(t_max, tc_max);
    float half = scale_exp2 * 0.5f;
    float tc[3];
    for (int a = 0; a < 3; a++) tc[a] = half * t_coef[a] + t_corner[a];
    if (tc_max < h) { stack[scale].scale = scale; stack[scale].parent = parent; stack[scale].t_max = t_max; }
    h = tc_max;
    parent = child;
    scale--;
    scale_exp2 = half;
    idx = (tc[0] > t_min) | ((tc[1] > t_min) << 1) | ((tc[2] > t_min) << 2);
    for (int a = 0; a < 3; a++) pos[a] += scale_exp2 * (float)(bool)(idx & (1 << a));
    t_max = tv_max;
    child = nullptr;
  }
  // First call of next(): voxel-unit distance to the first allocated leaf
  // block, or -1.
  float first() {
    const float vs = (float)map->dim / (float)map->size;
    for (int guard = 0; scale < DEPTH; guard++) {
      if (!parent || guard >= 8192) return -1.f;   // same walk bound as the HIP traversal
      for (int a = 0; a < 3; a++) t_corner[a] = pos[a] * t_coef[a] - t_bias[a];
      tc_max = fminf_r(fminf_r(t_corner[0], t_corner[1]), t_corner[2]);
      child = parent->child[idx ^ octant_mask ^ 7];
      if (scale == min_scale && child != nullptr) return t_min * (float)map->dim / vs;
      else if (child != nullptr && t_min <= t_max) { descend(); continue; }
      advance();
    }
    return -1.f;
  }
};

// ---- allocation list, Volume variant (kfusion/alloc_impl.hpp:121-191).
// Appends the key of every band sample whose block is missing (block keys
// repeat; the caller dedups).  `quirks` reproduces the reference's
// single-threaded list exactly: keys start at slot 1 and the count excludes
// the last one, so the caller's [0,count) range holds slot 0's stale value.
static inline size_t build_alloc_list(std::vector<uint64_t>& out, Octree& o, const M4& T_w_c, const M4& T_w_o,
                                      const M4& K, const R* depth, const R* mask, int w, int h, R voxel, R band) {
  const R inv_vox = R(1) / voxel;
  const int level = o.max_level - 3;
  const M4 invK = inverse(K);
  const M4 kPose = mul(T_w_c, invK);
  const M4 inv_T_w_o = inverse(T_w_o);
  const V3 wc = translation(T_w_c);
  const V3 oc = affine(inv_T_w_o, wc);
  const int num_steps = (int)std::ceil(band * inv_vox);
  const R size = (R)o.size;
  std::vector<std::vector<uint64_t>> rows((size_t)h);
#pragma omp parallel for schedule(dynamic, 8)
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const R d = depth[x + y * w];
      if (d == 0 || (mask && mask[x + y * w] < 0)) continue;
      const V3 wv = affine(kPose, v3((R(x) + R(0.5)) * d, (R(y) + R(0.5)) * d, d));
      const V3 ov = affine(inv_T_w_o, wv);
      const V3 dir = normalize(oc - ov);
      const V3 origin = ov - (band * R(0.5)) * dir;
      const V3 step = (dir * band) / (R)num_steps;
      V3 p = origin;
      for (int i = 0; i < num_steps; i++) {
        const V3 s = v3(std::floor(p.x * inv_vox), std::floor(p.y * inv_vox), std::floor(p.z * inv_vox));
        if (s.x < size && s.y < size && s.z < size && s.x >= 0 && s.y >= 0 && s.z >= 0) {
          const int vx = (int)s.x, vy = (int)s.y, vz = (int)s.z;
          Node* n = o.fetch(vx, vy, vz);
          if (!n) rows[y].push_back(encode(vx, vy, vz, level, o.max_level));
          else n->active = true;
        }
        p = p + step;
      }
    }
  size_t count = 0;
  for (auto& r : rows) { out.insert(out.end(), r.begin(), r.end()); count += r.size(); }
  return count;
}

// ---- integration: projective_functor::apply (projective_functor.hpp:54-156)
// with sdf_update (kfusion/mapping_impl.hpp:15-72).  update_node (I6) writes
// only coarse node values that nothing on this path reads, and is skipped.
static inline void integrate(Octree& o, const M4& Tco, const M4& K, const R* depth, const V3* rgb, const R* mask,
                             int w, int h, R mu, int maxweight) {
  const R vs = o.dim / (R)o.size;
  const M4 KT = mul(K, Tco);
  std::vector<Node*> active;
  for (Node* b : o.blocks) {
    bool keep = b->active;
    if (!keep) {   // algorithms::in_frustum (filter.hpp:42-51)
      const V3 bc = v3(R(b->cx) * vs, R(b->cy) * vs, R(b->cz) * vs);
      const V3 vc = affine(KT, bc);
      const int px = (int)(vc.x / vc.z), py = (int)(vc.y / vc.z);
      keep = px >= 0 && px < w && py >= 0 && py < h;
    }
    if (keep) active.push_back(b);
  }
  const V3 delta = rotate(Tco, v3(vs, 0, 0));
  const V3 cdelta = rotate(K, delta);
#pragma omp parallel for schedule(dynamic, 16)
  for (size_t bi = 0; bi < active.size(); bi++) {
    Node* b = active[bi];
    bool vis = false;
    for (int z = b->cz; z < b->cz + 8; ++z)
      for (int y = b->cy; y < b->cy + 8; ++y) {
        const V3 start = affine(Tco, v3(R(b->cx) * vs, R(y) * vs, R(z) * vs));
        const V3 cstart = affine(K, start);
        for (int x = 0; x < 8; ++x) {
          const V3 cv = cstart + R(x) * cdelta;
          const V3 pos = start + R(x) * delta;
          if (pos.z < R(0.0001)) continue;
          const R invd = R(1) / cv.z;
          const R pxf = cv.x * invd + R(0.5), pyf = cv.y * invd + R(0.5);
          if (pxf < R(0.5) || pxf > R(w) - R(1.5) || pyf < R(0.5) || pyf > R(h) - R(1.5)) continue;
          vis = true;
          // sdf_update
          const unsigned ux = (unsigned)pxf, uy = (unsigned)pyf;
          const R ds = depth[ux + w * uy];
          if (ds <= 0 || mask[ux + w * uy] < 0) continue;
          const R qx = pos.x / pos.z, qy = pos.y / pos.z;
          const R diff = (ds - pos.z) * std::sqrt(R(1) + qx * qx + qy * qy);
          if (diff > -mu) {
            const R sdf = fmin_r(R(1), diff / mu);
            Voxel& v = b->vox[x + 8 * (y - b->cy) + 64 * (z - b->cz)];
            v.x = fmax_r(R(-1), fmin_r((v.y * v.x + sdf) / (v.y + 1), R(1)));
            v.y = fmin_r(v.y + 1, (R)maxweight);
            const V3 c = rgb[ux + w * uy];
            v.r = (c.x + v.r * (R)v.c) / (R)(v.c + 1);
            v.g = (c.y + v.g * (R)v.c) / (R)(v.c + 1);
            v.b = (c.z + v.b * (R)v.c) / (R)(v.c + 1);
            v.c += 1;
            const R fgm = mask[ux + w * uy];
            v.fg = ((double)fgm + v.fg * (double)v.w) / (double)(v.w + 1);
            v.w++;
          }
        }
      }
    b->active = vis;
  }
}

// ---- raycast(volume, origin, dir, ...) (kfusion/rendering_impl.hpp:80-140).
// Returns (hit, t); t == 0 for no hit.
static inline void raycast_ray(const Octree& o, V3 origin, V3 dir, R nearp, R farp, R mu, R step, R largestep,
                               V3& hit, R& tout) {
  hit = v3(0, 0, 0); tout = 0;
  const R size = o.dim;
  const R inv_vs = (R)o.size / o.dim;
  const V3 invR = v3(R(1) / dir.x, R(1) / dir.y, R(1) / dir.z);
  const V3 tbot = cmul(R(-1) * invR, origin);
  const V3 ttop = cmul(invR, v3(size - origin.x, size - origin.y, size - origin.z));
  const V3 tmn = v3(fmin_r(ttop.x, tbot.x), fmin_r(ttop.y, tbot.y), fmin_r(ttop.z, tbot.z));
  const V3 tmx = v3(fmax_r(ttop.x, tbot.x), fmax_r(ttop.y, tbot.y), fmax_r(ttop.z, tbot.z));
  const R largest_tmin = fmax_r(fmax_r(tmn.x, tmn.y), fmax_r(tmn.x, tmn.z));
  const R smallest_tmax = fmin_r(fmin_r(tmx.x, tmx.y), fmin_r(tmx.x, tmx.z));
  const R tnear = fmax_r(largest_tmin, nearp);
  const R tfar = fmin_r(smallest_tmax, farp);
  if (!(tnear < tfar)) return;
  R t = tnear, stepsize = largestep;
  V3 p = origin + dir * t;
  R f_t = interp(o, p * inv_vs, F_TSDF);
  R f_tt = 0;
  if (f_t > 0) {
    for (; t < tfar; t += stepsize) {
      const Voxel v = o.get_fine((int)(p.x * inv_vs), (int)(p.y * inv_vs), (int)(p.z * inv_vs));
      if (v.y < 0) { stepsize = largestep; p = p + stepsize * dir; continue; }
      f_tt = v.x;
      if ((double)f_tt <= 0.1 && f_tt >= R(-0.5)) f_tt = interp(o, p * inv_vs, F_TSDF);
      if (f_tt < 0) break;
      stepsize = fmax_r(f_tt * mu, step);
      p = p + stepsize * dir;
      f_t = f_tt;
    }
    if (f_tt < 0) {
      t = t + stepsize * f_tt / (f_t - f_tt);
      hit = origin + dir * t; tout = t;
    }
  }
}

}  // namespace orc
