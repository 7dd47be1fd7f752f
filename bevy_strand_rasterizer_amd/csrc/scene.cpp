// scene.cpp -- host-side helpers of the strand runtime:
//   * the seeded scalp-hair generator used by the benchmark workloads
//     (packing follows pack_strands, src/plugin.rs:1315-1355: float4 vertices
//     with w = 1, identity index slab, StrandMeta{count, offset, material});
//   * the frustum-table builder of use_prepass_buffers (src/plugin.rs:386-458);
//   * camera / light matrices in the conventions of Bevy's projections
//     (reverse-Z infinite perspective for cameras, reverse-Z orthographic for
//     directional-light cascades: larger depth is nearer).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/strand_rt.h"

namespace {

struct Rng {  // splitmix64
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {  // Box-Muller
    double u1 = std::max(uniform(), 1e-300), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

struct D3 { double x, y, z; };
inline D3 add(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 sub(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 mul(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 norm(D3 a) { double l = std::sqrt(dot(a, a)); return l > 0 ? mul(a, 1.0 / l) : D3{0, 1, 0}; }

// Uniform Catmull-Rom segment between p1 and p2.
inline D3 catmull_rom(D3 p0, D3 p1, D3 p2, D3 p3, double t) {
  double t2 = t * t, t3 = t2 * t;
  auto c = [&](double a, double b, double cc, double d) {
    return 0.5 * ((2.0 * b) + (-a + cc) * t + (2.0 * a - 5.0 * b + 4.0 * cc - d) * t2 + (-a + 3.0 * b - 3.0 * cc + d) * t3);
  };
  return {c(p0.x, p1.x, p2.x, p3.x), c(p0.y, p1.y, p2.y, p3.y), c(p0.z, p1.z, p2.z, p3.z)};
}

void mat_mul(const double* a, const double* b, double* out) {  // column-major out = a * b
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a[k * 4 + r] * b[c * 4 + k];
      out[c * 4 + r] = s;
    }
}

void look_at_rh(D3 eye, D3 target, D3 up, double* m) {
  D3 f = norm(sub(target, eye));
  D3 s = norm(cross(f, up));
  D3 u = cross(s, f);
  double v[16] = {s.x, u.x, -f.x, 0, s.y, u.y, -f.y, 0, s.z, u.z, -f.z, 0, -dot(s, eye), -dot(u, eye), dot(f, eye), 1};
  std::memcpy(m, v, sizeof(v));
}

}  // namespace

extern "C" {

int sr_scalp_sizes(const sr_scalp_params* p, uint32_t* vertex_count, uint32_t* index_count, uint32_t* strand_count,
                   uint32_t* material_count) {
  if (!p || p->strand_count == 0 || p->points_per_strand < 2) return SR_ERR_INVALID_ARGUMENT;
  uint64_t v = (uint64_t)p->strand_count * p->points_per_strand;
  if (v > 0xFFFFFFFFull) return SR_ERR_INVALID_ARGUMENT;
  if (vertex_count) *vertex_count = (uint32_t)v;
  if (index_count) *index_count = (uint32_t)v;
  if (strand_count) *strand_count = p->strand_count;
  if (material_count) *material_count = p->strand_count;
  return SR_OK;
}

// Roots uniform on the upper hemisphere of radius scalp_radius (y up); each
// strand is a gravity-drooping Catmull-Rom curve of length
// U(length_min, length_max) with a helical curl of radius
// U(0, curl_radius_max); colour N(color_mean, color_sigma) clamped per channel.
int sr_scalp_generate(const sr_scalp_params* p, float* vertices, uint32_t* indices, sr_strand_meta* meta,
                      sr_strand_geo* geo, sr_strand_material* materials) {
  uint32_t nv = 0;
  int rc = sr_scalp_sizes(p, &nv, nullptr, nullptr, nullptr);
  if (rc != SR_OK) return rc;
  if (!vertices || !indices || !meta || !geo || !materials) return SR_ERR_INVALID_ARGUMENT;
  Rng rng{p->seed};
  const uint32_t P = p->points_per_strand;
  const D3 g{0.0, -1.0, 0.0};
  double bmin[3] = {1e30, 1e30, 1e30}, bmax[3] = {-1e30, -1e30, -1e30};
  for (uint32_t s = 0; s < p->strand_count; ++s) {
    double cy = rng.uniform();
    double phi = 6.283185307179586 * rng.uniform();
    double rxy = std::sqrt(std::max(0.0, 1.0 - cy * cy));
    D3 n{rxy * std::cos(phi), cy, rxy * std::sin(phi)};
    D3 root = mul(n, p->scalp_radius);
    double L = rng.uniform(p->length_min, p->length_max);
    double curl = rng.uniform(0.0, p->curl_radius_max);
    double pitch = rng.uniform(0.012, 0.03);
    double phase = rng.uniform(0.0, 6.283185307179586);
    double stiff = rng.uniform(0.25, 0.45);
    // guide points: out along the scalp normal, then droop under gravity
    D3 c1 = root;
    D3 c2 = add(root, add(mul(n, stiff * L), mul(g, 0.08 * L)));
    D3 c3 = add(root, add(mul(n, (stiff + 0.12) * L), mul(g, 0.5 * L)));
    D3 c4 = add(root, add(mul(n, (stiff + 0.18) * L), mul(g, 0.9 * L)));
    D3 c0 = sub(mul(c1, 2.0), c2);
    D3 c5 = sub(mul(c4, 2.0), c3);
    D3 guide[6] = {c0, c1, c2, c3, c4, c5};
    double turns = L / pitch;
    for (uint32_t k = 0; k < P; ++k) {
      double u = 3.0 * (double)k / (double)(P - 1);
      int span = std::min(2, (int)u);
      double t = u - span;
      D3 pos = catmull_rom(guide[span], guide[span + 1], guide[span + 2], guide[span + 3], t);
      double u2 = std::min(3.0, u + 1e-3);
      int span2 = std::min(2, (int)u2);
      D3 pos2 = catmull_rom(guide[span2], guide[span2 + 1], guide[span2 + 2], guide[span2 + 3], u2 - span2);
      D3 tan = norm(sub(pos2, pos));
      if (k == P - 1) {
        D3 prev = catmull_rom(guide[2], guide[3], guide[4], guide[5], 1.0 - 1e-3);
        tan = norm(sub(pos, prev));
      }
      D3 ref = std::fabs(tan.y) < 0.9 ? D3{0, 1, 0} : D3{1, 0, 0};
      D3 e1 = norm(cross(tan, ref));
      D3 e2 = cross(tan, e1);
      double sfrac = (double)k / (double)(P - 1);
      double ramp = std::min(1.0, sfrac * 4.0);
      double ang = 6.283185307179586 * turns * sfrac + phase;
      D3 off = add(mul(e1, curl * ramp * std::cos(ang)), mul(e2, curl * ramp * std::sin(ang)));
      D3 q = add(pos, off);
      uint32_t vi = s * P + k;
      vertices[4ull * vi + 0] = (float)q.x;
      vertices[4ull * vi + 1] = (float)q.y;
      vertices[4ull * vi + 2] = (float)q.z;
      vertices[4ull * vi + 3] = 1.0f;
      indices[vi] = vi;
      double qq[3] = {q.x, q.y, q.z};
      for (int a = 0; a < 3; ++a) { bmin[a] = std::min(bmin[a], (double)(float)qq[a]); bmax[a] = std::max(bmax[a], (double)(float)qq[a]); }
    }
    meta[s] = sr_strand_meta{P, s * P, s, 0};
    sr_strand_material& m = materials[s];
    for (int c = 0; c < 3; ++c) {
      double v = p->color_mean + p->color_sigma * rng.normal();
      m.absorption_color[c] = (float)std::min(1.0, std::max(0.0, v));
    }
    m.absorption_color[3] = 0.45f;
    m.specular_color[0] = 0.9f; m.specular_color[1] = 0.72f; m.specular_color[2] = 0.52f; m.specular_color[3] = 0.2f;
    m.ambient_factor = 0.15f; m.ao_factor = 0.1f; m.eta = 1.55f; m.beta = 0.85f; m.alpha = 0.35f; m.shift = 0.1f;
    m.min_radius_pixels = p->radius_pixels;
    m.max_radius_pixels = p->radius_pixels;
  }
  std::memset(geo, 0, sizeof(*geo));
  geo->strand_count = p->strand_count;
  geo->max_segments_in_strand = P;  // plugin.rs:1334 stores the max vertex count here
  geo->offset = 0;
  for (int a = 0; a < 3; ++a) { geo->aabb.min[a] = (float)bmin[a]; geo->aabb.max[a] = (float)bmax[a]; }
  return SR_OK;
}

// use_prepass_buffers frustum table, src/plugin.rs:386-458.
int sr_build_frustum_table(const sr_froxel_config* cfgs, const uint32_t* is_light, const uint32_t* cascade,
                           uint32_t count, sr_frustum_desc* out) {
  if (!cfgs || !out || count == 0) return SR_ERR_INVALID_ARGUMENT;
  uint32_t bucket_base = 0, coarse_base = 0, fine_base = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const sr_froxel_config& c = cfgs[i];
    if (c.screen_width == 0 || c.screen_height == 0 || c.froxel_size_x == 0 || c.froxel_size_y == 0)
      return SR_ERR_INVALID_ARGUMENT;
    uint32_t ftx = (c.screen_width + c.froxel_size_x - 1) / c.froxel_size_x;
    uint32_t fty = (c.screen_height + c.froxel_size_y - 1) / c.froxel_size_y;
    uint32_t buckets = ftx * fty * c.depth_slices;
    uint32_t ctx = std::max(1u, (ftx + SR_COARSE_FINE_TILE_EXTENT - 1) / SR_COARSE_FINE_TILE_EXTENT);
    uint32_t cty = std::max(1u, (fty + SR_COARSE_FINE_TILE_EXTENT - 1) / SR_COARSE_FINE_TILE_EXTENT);
    sr_frustum_desc d{};
    d.screen_width = c.screen_width;
    d.screen_height = c.screen_height;
    d.froxel_size_x = c.froxel_size_x;
    d.froxel_size_y = c.froxel_size_y;
    d.depth_slices = c.depth_slices;
    d.bucket_base = bucket_base;
    d.bucket_count = buckets;
    d.kind = is_light && is_light[i] ? 1u : 0u;
    d.coarse_depth_tile_base = coarse_base;
    d.coarse_depth_tile_count = ctx * cty;
    d.coarse_tiles_x = ctx;
    d.coarse_tiles_y = cty;
    d.cascade_index = cascade ? cascade[i] : 0u;
    d.fine_depth_tile_base = fine_base;
    out[i] = d;
    bucket_base += buckets;
    coarse_base += ctx * cty;
    fine_base += ftx * fty;
  }
  return SR_OK;
}

// Camera view: Bevy PerspectiveProjection (reverse-Z, infinite far plane).
int sr_camera_view(const float* eye, const float* target, const float* up, float fov_y, float aspect, float z_near,
                   sr_view* out) {
  if (!eye || !target || !up || !out || !(fov_y > 0) || !(aspect > 0) || !(z_near > 0)) return SR_ERR_INVALID_ARGUMENT;
  double view[16];
  look_at_rh({eye[0], eye[1], eye[2]}, {target[0], target[1], target[2]}, {up[0], up[1], up[2]}, view);
  double h = 1.0 / std::tan(0.5 * (double)fov_y), w = h / aspect;
  double proj[16] = {w, 0, 0, 0, 0, h, 0, 0, 0, 0, 0, -1, 0, 0, z_near, 0};
  double m[16];
  mat_mul(proj, view, m);
  for (int i = 0; i < 16; ++i) out->clip_from_world[i] = (float)m[i];
  out->world_position[0] = eye[0]; out->world_position[1] = eye[1]; out->world_position[2] = eye[2];
  out->world_position[3] = 1.0f;
  return SR_OK;
}

// Directional-light view: orthographic, reverse-Z (nearest to the light = 1),
// fitted to a bounding sphere (centre, radius).
int sr_light_view(const float* direction_to_light, const float* center, float radius, sr_view* out) {
  if (!direction_to_light || !center || !out || !(radius > 0)) return SR_ERR_INVALID_ARGUMENT;
  D3 d = norm({direction_to_light[0], direction_to_light[1], direction_to_light[2]});
  D3 c{center[0], center[1], center[2]};
  D3 eye = add(c, mul(d, 2.0 * radius));
  D3 up = std::fabs(d.y) < 0.95 ? D3{0, 1, 0} : D3{1, 0, 0};
  double view[16];
  look_at_rh(eye, c, up, view);
  double near_ = radius, far_ = 3.0 * radius;
  double proj[16] = {1.0 / radius, 0, 0, 0, 0, 1.0 / radius, 0, 0, 0, 0, 1.0 / (far_ - near_), 0,
                     0, 0, far_ / (far_ - near_), 1};
  double m[16];
  mat_mul(proj, view, m);
  for (int i = 0; i < 16; ++i) out->clip_from_world[i] = (float)m[i];
  out->world_position[0] = (float)eye.x; out->world_position[1] = (float)eye.y; out->world_position[2] = (float)eye.z;
  out->world_position[3] = 0.0f;
  return SR_OK;
}

}  // extern "C"
