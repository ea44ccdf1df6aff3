/*
 * TEST INFRASTRUCTURE ONLY — a C-ABI shim over the reference's own
 * header-only octree core (lib/ of smartroboticslab/mid-fusion), compiled in
 * place from /root/reference by oracle/Makefile.ref into oracle/_ref/.
 * Nothing here is copied from the reference: the file includes the
 * reference headers where they lie and exposes their functions to the tests,
 * which use them to pin the oracle's restatement (Morton keys, multiscale
 * dedup, alloc_update, trilinear interp/gather, gradient, ray_iterator).
 * The application layer (tracking, preprocessing, mapping) needs TooN and
 * OpenCV, which are absent, and is not built.
 */
#include <octree.hpp>
#include <continuous/volume_traits.hpp>
#include <continuous/volume_instance.hpp>
#include <kfusion/rendering_impl.hpp>
#include <vector>
#include <algorithm>

typedef Octree<SDF> RefOctree;

extern "C" {

uint64_t ref_encode(int x, int y, int z, int level, int max_depth) {
  return octlib::keyops::encode(x, y, z, level, max_depth);
}
uint64_t ref_morton(uint32_t x, uint32_t y, uint32_t z) { return compute_morton(x, y, z); }
void ref_unpack(uint64_t code, uint32_t* xyz) {
  uint3 c = unpack_morton(code);
  xyz[0] = c.x; xyz[1] = c.y; xyz[2] = c.z;
}
uint64_t ref_parent(uint64_t key, int max_depth) { return parent(key, max_depth); }
int ref_child_id(uint64_t key, int level, int max_depth) { return child_id(key, level, max_depth); }
int ref_unique_ms(uint64_t* keys, int n, int level) {
  if (level < 0) return algorithms::unique_multiscale(keys, n, (uint64_t)SCALE_MASK);
  return algorithms::unique_multiscale(keys, n, (uint64_t)SCALE_MASK, (unsigned)level);
}

void* ref_octree_create(int size, float dim) {
  RefOctree* o = new RefOctree();
  o->init(size, dim);
  return o;
}
void ref_octree_alloc(void* h, uint64_t* keys, int n) { ((RefOctree*)h)->alloc_update(keys, n); }
uint64_t ref_octree_hash(void* h, int x, int y, int z, int scale) { return ((RefOctree*)h)->hash(x, y, z, scale); }
long ref_octree_keys(void* h, uint64_t* out, long cap) {
  RefOctree* o = (RefOctree*)h;
  auto& pool = o->getBlockBuffer();
  std::vector<uint64_t> k;
  for (size_t i = 0; i < pool.size(); i++) k.push_back(pool[i]->code);
  std::sort(k.begin(), k.end());
  for (long i = 0; i < (long)k.size() && i < cap; i++) out[i] = k[i];
  return (long)k.size();
}
long ref_octree_node_keys(void* h, uint64_t* out, long cap) {
  RefOctree* o = (RefOctree*)h;
  auto& pool = o->getNodesBuffer();
  std::vector<uint64_t> k;
  for (size_t i = 0; i < pool.size(); i++) k.push_back(pool[i]->code);
  std::sort(k.begin(), k.end());
  for (long i = 0; i < (long)k.size() && i < cap; i++) out[i] = k[i];
  return (long)k.size();
}
int ref_octree_set(void* h, int x, int y, int z, float tsdf, float fg) {
  RefOctree* o = (RefOctree*)h;
  VoxelBlock<SDF>* b = o->fetch(x, y, z);
  if (!b) return 0;
  auto v = b->data(make_int3(x, y, z));
  v.x = tsdf; v.fg = fg; v.y = 1;
  b->data(make_int3(x, y, z), v);
  return 1;
}
float ref_octree_interp(void* h, float x, float y, float z, int field) {
  RefOctree* o = (RefOctree*)h;
  if (field == 0) return o->interp(make_float3(x, y, z), [](const auto& v) { return v.x; });
  return o->interp(make_float3(x, y, z), [](const auto& v) { return v.fg; });
}
void ref_octree_grad(void* h, float x, float y, float z, float* g) {
  RefOctree* o = (RefOctree*)h;
  float3 r = o->grad(make_float3(x, y, z), [](const auto& v) { return v.x; });
  g[0] = r.x; g[1] = r.y; g[2] = r.z;
}
float ref_octree_ray_tmin(void* h, const float* org, const float* dir, float nearp, float farp) {
  RefOctree* o = (RefOctree*)h;
  ray_iterator<SDF> it(*o, make_float3(org[0], org[1], org[2]), make_float3(dir[0], dir[1], dir[2]), nearp, farp);
  return std::get<0>(it.next());
}
void ref_octree_destroy(void* h) { delete (RefOctree*)h; }

// Volume<SDF> + raycast(volume, origin, dir, near, far, mu, step, largestep)
// (kfusion/rendering_impl.hpp:80-140).
void* ref_volume_create(int res, float size) {
  Volume<SDF>* v = new Volume<SDF>();
  v->init(res, size);
  return v;
}
void ref_volume_alloc(void* h, uint64_t* keys, int n) { ((Volume<SDF>*)h)->_map_index.alloc_update(keys, n); }
int ref_volume_set(void* h, int x, int y, int z, float tsdf, float fg) {
  return ref_octree_set(&((Volume<SDF>*)h)->_map_index, x, y, z, tsdf, fg);
}
float ref_volume_raycast(void* h, const float* org, const float* dir, float nearp, float farp, float mu, float step,
                         float largestep, float* hit) {
  float4 r = raycast(*(Volume<SDF>*)h, make_float3(org[0], org[1], org[2]), make_float3(dir[0], dir[1], dir[2]),
                     nearp, farp, mu, step, largestep);
  hit[0] = r.x; hit[1] = r.y; hit[2] = r.z;
  return r.w;
}
void ref_volume_destroy(void* h) { delete (Volume<SDF>*)h; }

}  // extern "C"
