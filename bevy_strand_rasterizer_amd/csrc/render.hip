// render.hip -- deep-shadow-map build, stampback, hair shading, strand
// raster and composite on MI355X (gfx950, wave64).
//
// Reference kernels restated here (paths relative to the reference root):
//   rasterize_strands [SHADOWS]   src/shaders/strand_rasterizer.wgsl:391-711
//   stampback fragment            src/shaders/shadow_dom_stampback.wgsl:74-123
//   shade_strands                 src/shaders/strand_shading.wgsl:608-695
//   rasterize_strands [LINEAR]    src/shaders/strand_rasterizer.wgsl:855-1044
//   composite fragment            src/shaders/strand_composite.wgsl:12-33
//
// Design: one 256-lane workgroup per tile run; every lane owns one pixel of
// the tile (8x8 or 16x16 froxels, all pixels rasterised) and all lanes stage
// a 256-reference batch of projected segments in LDS, so each reference is
// fetched and projected once per tile.  Work items arrive near to far; a
// workgroup vote ends the walk once every pixel is opaque (camera) or has
// its front depth (shadow pass 1).  Deep-shadow partial opacities live in
// LDS (pixel x layer), so no dynamic register indexing spills.  Shadow maps
// are dense per light (no virtual paging), outputs stay FP32.
#include "common.hpp"
#include "kernels.hpp"

namespace srk {

constexpr uint32_t kBatch = 256;

__global__ void k_fill_u32(uint32_t* dst, uint64_t n, uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = v;
}

void launch_fill_u32(uint32_t* dst, uint64_t count, uint32_t value, hipStream_t s) {
  if (count == 0) return;
  uint64_t g = (count + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_fill_u32, dim3((uint32_t)g), dim3(256), 0, s, dst, count, value);
}

__device__ __forceinline__ V3 world_v3(const InstRec& in, uint32_t vi) {
  float4 v = in.vertices[vi];
  V4 w = mat_vec(in.world_from_local, V4{v.x, v.y, v.z, 1.0f});
  return V3{w.x, w.y, w.z};
}

// --------------------------------------------------------- K25 shadow
struct ShadowBatch {
  float p0x[kBatch], p0y[kBatch], p0z[kBatch];
  float p1x[kBatch], p1y[kBatch], p1z[kBatch];
  float rmin[kBatch], rmax[kBatch], absorb[kBatch];
  uint32_t valid[kBatch];
};

__device__ __forceinline__ bool shadow_cover(const ShadowBatch& b, uint32_t k, float px, float py, float* pz,
                                             float* cov) {
  if (!b.valid[k]) return false;
  float x0 = b.p0x[k], y0 = b.p0y[k], z0 = b.p0z[k];
  float abx = b.p1x[k] - x0, aby = b.p1y[k] - y0;
  float l2 = abx * abx + aby * aby;
  float t = 0.0f;
  if (!(l2 == 0.0f)) t = ((px - x0) * abx + (py - y0) * aby) / l2;
  if (t < 0.0f || t > 1.0f) return false;
  float qx = wmix(x0, b.p1x[k], t), qy = wmix(y0, b.p1y[k], t), qz = wmix(z0, b.p1z[k], t);
  float r = wmix(b.rmin[k], b.rmax[k], wclamp(qz, 0.0f, 1.0f));
  float dx = px - qx, dy = py - qy;
  float c = wclamp(1.0f - sqrtf(dx * dx + dy * dy) / r, 0.0f, 1.0f);
  *pz = qz;
  *cov = c;
  return c > 0.0f;
}

__device__ void load_shadow_batch(const FrameParams& fp, const FrustumRec& fr, uint32_t fi, uint32_t screen_tile,
                                  const sr_raster_work_item& wi, uint32_t base, ShadowBatch& b) {
  uint32_t k = threadIdx.x;
  uint32_t ref_idx = base + k;
  uint32_t end = wi.seg_ref_base + wi.seg_ref_count;
  b.valid[k] = 0u;
  if (ref_idx < end && wi.frustum_id == fi && wi.screen_tile_id == screen_tile) {
    const sr_fine_seg_ref sr = fp.refs[ref_idx];
    const InstRec& in = fp.inst[sr.inst_id];
    V3 a = world_v3(in, in.indices[sr.seg_id]);
    V3 c = world_v3(in, in.indices[sr.seg_id + 1u]);
    float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
    V4 ra = world_to_screen_raw(V4{a.x, a.y, a.z, 1.0f}, fr.clip, vw, vh);
    V4 rc = world_to_screen_raw(V4{c.x, c.y, c.z, 1.0f}, fr.clip, vw, vh);
    const sr_strand_material& m = in.materials[sr.material_idx];
    float minr = fmaxf(m.min_radius_pixels, 1e-4f);
    b.p0x[k] = ra.x; b.p0y[k] = ra.y; b.p0z[k] = normalize_depth01(ra.z);
    b.p1x[k] = rc.x; b.p1y[k] = rc.y; b.p1z[k] = normalize_depth01(rc.z);
    b.rmin[k] = minr;
    b.rmax[k] = fmaxf(m.max_radius_pixels, minr);
    b.absorb[k] = m.absorption_color[3];
    b.valid[k] = 1u;
  }
}

__global__ __launch_bounds__(256) void k_shadow_raster(FrameParams fp) {
  __shared__ ShadowBatch b;
  __shared__ float part[256 * SR_MAX_DOM_LAYERS];
  const uint32_t R = fp.counters[C_SHADOW_RUNS];
  const uint32_t L = fp.dom_layers;
  for (uint32_t ri = blockIdx.x; ri < R; ri += gridDim.x) {
    const sr_raster_tile_run run = fp.shadow_runs[ri];
    const uint32_t fi = run.frustum_id;
    const FrustumRec& fr = fp.fr[fi];
    if (fr.light_layer == kInvalid || run.work_count == 0u) continue;
    if ((uint64_t)run.work_base + run.work_count > fp.work_cap) continue;  // emission overflowed
    const uint32_t fx = fr.d.froxel_size_x, fy = fr.d.froxel_size_y;
    const uint32_t tx = run.screen_tile_id % fr.fine_tiles_x, ty = run.screen_tile_id / fr.fine_tiles_x;
    const uint32_t pidx = threadIdx.x;
    const uint32_t pxu = tx * fx + pidx % fx, pyu = ty * fy + pidx / fx;
    const bool active = pidx < fx * fy && pxu < fr.d.screen_width && pyu < fr.d.screen_height;
    const float pfx = (float)pxu + 0.5f, pfy = (float)pyu + 0.5f;
    // pass 1: front depth from the first (nearest) work item with a hit
    float z0 = -1.0f;
    bool found = false;
    for (uint32_t w = 0; w < run.work_count; ++w) {
      const sr_raster_work_item wi = fp.work[run.work_base + w];
      bool hit = false;
      for (uint32_t base = wi.seg_ref_base; base < wi.seg_ref_base + wi.seg_ref_count; base += kBatch) {
        __syncthreads();
        load_shadow_batch(fp, fr, fi, run.screen_tile_id, wi, base, b);
        __syncthreads();
        if (active && !found) {
          uint32_t n = min(kBatch, wi.seg_ref_base + wi.seg_ref_count - base);
          for (uint32_t k = 0; k < n; ++k) {
            float pz, cov;
            if (shadow_cover(b, k, pfx, pfy, &pz, &cov)) { z0 = fmaxf(z0, pz); hit = true; }
          }
        }
      }
      if (hit) found = true;
      if (__syncthreads_count(active && !found) == 0) break;
    }
    const size_t plane = (size_t)fr.d.screen_width * fr.d.screen_height;
    const size_t pix = (size_t)pyu * fr.d.screen_width + pxu;
    float* part_px = part + threadIdx.x * SR_MAX_DOM_LAYERS;
    for (uint32_t i = 0; i < L; ++i) part_px[i] = 0.0f;
    const bool alpha_active = active && found;
    const float inv_span = 1.0f / fmaxf(1e-6f, alpha_active ? z0 : 1.0f);
    if (__syncthreads_count(alpha_active) != 0) {
      for (uint32_t w = 0; w < run.work_count; ++w) {
        const sr_raster_work_item wi = fp.work[run.work_base + w];
        for (uint32_t base = wi.seg_ref_base; base < wi.seg_ref_base + wi.seg_ref_count; base += kBatch) {
          __syncthreads();
          load_shadow_batch(fp, fr, fi, run.screen_tile_id, wi, base, b);
          __syncthreads();
          if (alpha_active) {
            uint32_t n = min(kBatch, wi.seg_ref_base + wi.seg_ref_count - base);
            for (uint32_t k = 0; k < n; ++k) {
              float pz, cov;
              if (!shadow_cover(b, k, pfx, pfy, &pz, &cov)) continue;
              float dzp = fmaxf(0.0f, z0 - pz) * inv_span;
              float u = wclamp(dzp, 0.0f, 1.0f);  // pow(u, DOM_GAMMA = 1)
              float tL = u * (float)L;
              uint32_t si = min(f2u(floorf(tL)), L - 1u);
              float wgt = tL - floorf(tL);
              float a = cov * b.absorb[k] * 0.5f;
              float a0 = part_px[si];
              part_px[si] = wclamp(a0 + (1.0f - a0) * (1.0f - wgt) * a, 0.0f, 1.0f);
              if (si + 1u < L) {
                float a1 = part_px[si + 1u];
                part_px[si + 1u] = wclamp(a1 + (1.0f - a1) * wgt * a, 0.0f, 1.0f);
              }
            }
          }
        }
      }
    }
    if (active) {
      float* dd = fp.dom_depth + fr.dom_depth_off;
      float* op = fp.dom_opacity + fr.dom_opacity_off;
      if (!found) {
        dd[pix] = 0.0f;
        for (uint32_t i = 0; i < L; ++i) op[i * plane + pix] = 0.0f;
      } else {
        dd[pix] = z0;
        float acc = 0.0f;
        for (uint32_t i = 0; i < L; ++i) {
          acc = acc + (1.0f - acc) * part_px[i];
          op[i * plane + pix] = acc;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------ K26 stampback
__global__ void k_stampback(FrameParams fp, uint32_t fi) {
  const FrustumRec& fr = fp.fr[fi];
  const uint32_t tw = fr.stamp_w, th = fr.stamp_h;
  const uint32_t W = fr.d.screen_width, H = fr.d.screen_height;
  const float* dd = fp.dom_depth + fr.dom_depth_off;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < tw * th; i += gridDim.x * blockDim.x) {
    uint32_t x = i % tw, y = i / tw;
    uint32_t x0 = min((x * W) / tw, W - 1u), y0 = min((y * H) / th, H - 1u);
    uint32_t x1 = min(((x + 1u) * W + tw - 1u) / tw, W), y1 = min(((y + 1u) * H + th - 1u) / th, H);
    float z = 0.0f;
    for (uint32_t yy = y0; yy < y1; ++yy)
      for (uint32_t xx = x0; xx < x1; ++xx) z = fmaxf(z, dd[(size_t)yy * W + xx]);
    fp.stamp[fr.stamp_off + i] = z;
  }
}

// --------------------------------------------------------- K27 shading
constexpr float kPiF = 3.14159265359f;
#define PI_HALF_LO ((float)(-3.14159265359 / 2.0 + 1e-6))
#define PI_HALF_HI ((float)(3.14159265359 / 2.0 - 1e-6))
#define TWO_PI_F ((float)(2.0 * 3.14159265359))
#define SQRT_2PI_F ((float)2.5066282746310002)
#define DELTA_TOL_F ((float)(0.1 * 3.14159265359 / 180.0))

__constant__ float kLegRoots[10] = {-0.9739065285171717f, -0.8650633666889844f, -0.6794095682990244f,
                                    -0.4333953941292472f, -0.14887433898163116f, 0.14887433898163116f,
                                    0.4333953941292472f, 0.6794095682990244f, 0.8650633666889844f,
                                    0.9739065285171717f};
__constant__ float kLegWeights[10] = {0.06667134430868714f, 0.14945134915058053f, 0.21908636251598224f,
                                      0.26926671930999674f, 0.2955242247147533f, 0.2955242247147533f,
                                      0.26926671930999674f, 0.21908636251598224f, 0.14945134915058053f,
                                      0.06667134430868714f};

__device__ __forceinline__ V3 v3sub(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float v3dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 v3scale(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 v3cross(V3 a, V3 b) {
  return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ V3 v3norm(V3 a) {
  float l = sqrtf(v3dot(a, a));
  return V3{a.x / l, a.y / l, a.z / l};
}
__device__ __forceinline__ float sacos(float x) { return acosf(wclamp(x, -1.0f, 1.0f)); }

__device__ float fresnel_reflectance(float eta_t, float c_in) {  // eta_i = 1
  float c = fabsf(c_in);
  float cc = wclamp(0.0f, 1.0f, c);
  float sin2 = 1.0f - cc * cc;
  float ratio = 1.0f / eta_t;
  float cos_t = sqrtf(1.0f - ratio * ratio * sin2);
  float rpar = ((eta_t * c) - (1.0f * cos_t)) / ((eta_t * c) + (1.0f * cos_t));
  float rper = ((1.0f * c) - (eta_t * cos_t)) / ((1.0f * c) + (eta_t * cos_t));
  return (rpar * rpar + rper * rper) * 0.5f;
}

__device__ float bessel5(float x) {
  // sum_{i<5} x^(2i) / (4^i (i!)^2)
  const float fact[5] = {1.0f, 1.0f, 2.0f, 6.0f, 24.0f};
  float y = 0.0f;
  for (int i = 0; i < 5; ++i) {
    float num = powf(x, 2.0f * (float)i);
    float den = powf(4.0f, (float)i) * fact[i] * fact[i];
    y = y + num / den;
  }
  return y;
}

__device__ float Mp(float v, float theta_i_, float theta_r, float shift) {
  float trs = theta_r + shift;
  float theta_i = wclamp(PI_HALF_LO, PI_HALF_HI, theta_i_);
  trs = wclamp(PI_HALF_LO, PI_HALF_HI, trs);
  float cti = cosf(-theta_i), ctr = cosf(trs);
  float sti = sinf(-theta_i), str = sinf(trs);
  if (v < 1e-6f) return fabsf(trs + theta_i) < DELTA_TOL_F ? 0.0f : 1e6f / DELTA_TOL_F;
  return bessel5((cti * ctr) / v) * expf((sti * str) / v);
}

__device__ V3 Np(uint32_t p, float phi, float theta_i, float theta_r, float eta, V3 mu, float beta) {
  float theta_d = (theta_r - theta_i) * 0.5f;
  float sti = sinf(theta_i);
  float ctd = cosf(theta_d);
  float eta_prime = sqrtf(fmaxf(0.0f, eta * eta - sti * sti)) / ctd;
  V3 mup{mu.x / ctd, mu.y / ctd, mu.z / ctd};
  float var = beta * beta;
  float norm = 1.0f / (SQRT_2PI_F * beta);
  float cos_td = cosf(theta_d);
  V3 acc{0.0f, 0.0f, 0.0f};
  for (int i = 0; i < 10; ++i) {
    float h = wclamp(-1.0f + 1e-6f, 1.0f - 1e-6f, kLegRoots[i]);
    // A(p, h)
    float gamma_t = asinf(h / eta_prime);
    float gamma_i = asinf(h);
    float f = fresnel_reflectance(eta, sacos(cos_td * cosf(gamma_i)));
    float e = 1.0f + cosf(2.0f * gamma_t);
    V3 T{expf(-2.0f * mup.x * e), expf(-2.0f * mup.y * e), expf(-2.0f * mup.z * e)};
    V3 att;
    if (p == 0u) {
      att = V3{f * T.x * T.x, f * T.y * T.y, f * T.z * T.z};
    } else {
      float f_ = 1.0f - f;
      float s = (f_ * f_) * powf(f, (float)(p - 1u));
      att = V3{s * powf(T.x, (float)p), s * powf(T.y, (float)p), s * powf(T.z, (float)p)};
    }
    // Phi(p, h)
    float gt2 = asinf(wclamp(h / eta_prime, -1.0f, 1.0f));
    float shift = 2.0f * (float)p * gt2 - 2.0f * asinf(h) + (float)p * kPiF;
    // gaussian_detector(beta, phi - shift, 10)
    float det;
    float pd = phi - shift;
    if (var < 1e-6f) {
      det = fabsf(fmodf(pd, TWO_PI_F)) < 1e-3f ? 0.0f : 1e6f;
    } else {
      float total = 0.0f;
      for (int k = -10; k <= 10; ++k) {
        float ang = pd - TWO_PI_F * (float)k;
        total = total + expf(-ang * ang / (2.0f * var));
      }
      det = norm * total;
    }
    acc.x = acc.x + att.x * det * kLegWeights[i];
    acc.y = acc.y + att.y * det * kLegWeights[i];
    acc.z = acc.z + att.z * det * kLegWeights[i];
  }
  return v3scale(acc, 0.5f);
}

__device__ V3 marschner(V3 dir, V3 view_n, V3 light_n, const sr_strand_material& m, float occlusion) {
  float theta_i = sacos(v3dot(dir, light_n));
  float theta_r = sacos(v3dot(dir, view_n));
  V3 lp = v3norm(v3sub(light_n, v3scale(dir, v3dot(light_n, dir))));
  V3 vp = v3norm(v3sub(view_n, v3scale(dir, v3dot(view_n, dir))));
  float sg = v3dot(v3cross(lp, vp), dir);
  float phi = sacos(v3dot(v3norm(lp), v3norm(vp))) * (sg > 0.0f ? 1.0f : (sg < 0.0f ? -1.0f : 0.0f));
  V3 sigma{1.0f - m.absorption_color[0], 1.0f - m.absorption_color[1], 1.0f - m.absorption_color[2]};
  float v_long = m.alpha * m.alpha, v_azim = m.beta * m.beta;
  float beta_azim = sqrtf(v_azim);
  float scat = wclamp(1.0f - occlusion, 0.0f, 1.0f);
  float direct = wmix(0.18f, 1.0f, scat);
  float mp = Mp(v_long, theta_i, theta_r, m.shift);
  V3 total{mp * m.specular_color[0] * m.specular_color[3] * direct, mp * m.specular_color[1] * m.specular_color[3] * direct,
           mp * m.specular_color[2] * m.specular_color[3] * direct};
  for (uint32_t p = 0; p < 3u; ++p) {
    V3 n = Np(p, phi, theta_i, theta_r, m.eta, sigma, beta_azim);
    float s = ((float)(p + 1u) * scat) / 3.0f;
    total.x = total.x + n.x * s;
    total.y = total.y + n.y * s;
    total.z = total.z + n.z * s;
  }
  return total;
}

__device__ float dom_visibility(const FrameParams& fp, V3 pw, uint32_t layer) {
  uint32_t fi = kInvalid, l = 0;
  for (uint32_t i = 0; i < fp.F; ++i) {
    if (fp.fr[i].d.kind != 1u) continue;
    if (l == layer) { fi = i; break; }
    ++l;
  }
  if (fi == kInvalid) return 1.0f;
  const FrustumRec& fr = fp.fr[fi];
  float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
  V4 raw = world_to_screen_raw(V4{pw.x, pw.y, pw.z, 1.0f}, fr.clip, vw, vh);
  if (raw.x < 0.0f || raw.y < 0.0f || raw.x >= vw || raw.y >= vh || raw.z < 0.0f || raw.z > 1.0f) return 1.0f;
  uint32_t px = f2u(raw.x), py = f2u(raw.y);
  size_t pix = (size_t)py * fr.d.screen_width + px;
  float z0 = fp.dom_depth[fr.dom_depth_off + pix];
  if (z0 <= 0.0f) return 1.0f;
  float z = normalize_depth01(raw.z);
  if (z > z0 - 1e-3f) return 1.0f;
  uint32_t L = fp.dom_layers;
  float u = wclamp(fmaxf(0.0f, z0 - z) / fmaxf(z0, 1e-6f), 0.0f, 1.0f);
  uint32_t s = min(f2u(floorf(u * (float)L)), L - 1u);
  float op = fp.dom_opacity[fr.dom_opacity_off + (size_t)s * fr.d.screen_width * fr.d.screen_height + pix];
  return wclamp(1.0f - op, 0.08f, 1.0f);
}

__global__ __launch_bounds__(128) void k_shading(FrameParams fp) {
  const uint32_t n = min(fp.counters[C_CAM_TASKS], fp.task_cap);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
    const uint32_t t = fp.shading_queue[q];
    const sr_binning_task tk = fp.tasks[t];
    const uint32_t fi = tk.packed_field >> SR_BINNING_FRUSTUM_SHIFT;
    const FrustumRec& fr = fp.fr[fi];
    const InstRec& in = fp.inst[tk.id_info];
    const uint32_t idx = tk.seg_idx;
    if (idx >= fp.atlas_texels) continue;
    const sr_strand_meta sm = in.meta[tk.chunk_id];
    V3 v0 = world_v3(in, in.indices[idx]);
    V3 v1 = world_v3(in, in.indices[idx + 1u]);
    V3 U = v3norm(v3sub(v1, v0));
    V3 cam_dir = v3norm(v3sub(V3{fr.eye[0], fr.eye[1], fr.eye[2]}, v0));
    V3 bin = v3norm(v3cross(U, cam_dir));
    V3 Vv = v3norm(v3cross(bin, U));
    const sr_strand_material& m = in.materials[sm.material_idx];
    V3 mid{wmix(v0.x, v1.x, 0.5f), wmix(v0.y, v1.y, 0.5f), wmix(v0.z, v1.z, 0.5f)};
    const uint32_t nl = min(fp.NL, 4u);
    float cur = 1.0f;
    if (nl > 0u) {
      float vis = 0.0f;
      for (uint32_t i = 0; i < nl; ++i) vis = vis + dom_visibility(fp, mid, i);
      cur = vis / (float)nl;
    }
    const size_t layer = (size_t)fr.cam_index * fp.NI + tk.id_info;
    const size_t texel = layer * fp.atlas_texels + idx;
    float prev = (fp.st.shadow_history_valid && fp.history_prev) ? fp.history_prev[texel].x : 0.0f;
    float scat_vis = prev > 0.0f ? wmix(cur, prev, 0.88f) : cur;
    fp.history[texel] = make_float2(scat_vis, cur);
    float occ = 1.0f - scat_vis;
    float r = 0.0f, g = 0.0f, bl = 0.0f;
    for (uint32_t j = 0; j < fp.NL; ++j) {
      const sr_directional_light lt = fp.lights[j];
      V3 L = v3norm(V3{lt.direction_to_light[0], lt.direction_to_light[1], lt.direction_to_light[2]});
      V3 bc = marschner(L, Vv, U, m, occ);
      float amb = m.ambient_factor;
      float cx = bc.x * (lt.color[0] / 255.0f), cy = bc.y * (lt.color[1] / 255.0f), cz = bc.z * (lt.color[2] / 255.0f);
      float ax = m.absorption_color[0] * amb + (fp.st.ambient_color[0] / 255.0f) * amb;
      float ay = m.absorption_color[1] * amb + (fp.st.ambient_color[1] / 255.0f) * amb;
      float az = m.absorption_color[2] * amb + (fp.st.ambient_color[2] / 255.0f) * amb;
      r += wmix(cx, ax, amb);
      g += wmix(cy, ay, amb);
      bl += wmix(cz, az, amb);
    }
    float4 col = make_float4(wclamp(r, 0.0f, 1.0f), wclamp(g, 0.0f, 1.0f), wclamp(bl, 0.0f, 1.0f),
                             wclamp(m.absorption_color[3], 0.0f, 1.0f));
    fp.atlas[texel] = col;
    // texel idx + 1 is also written by the segment starting there; the
    // segment that starts at a texel owns it
    uint32_t nxt = idx + 1u;
    if (nxt < fp.atlas_texels) {
      bool nxt_shaded = nxt < in.index_count && fp.shaded[(size_t)fr.cam_index * fp.slots + in.pos_base + nxt];
      if (!nxt_shaded) fp.atlas[texel + 1] = col;
    }
  }
}

// --------------------------------------------------- K28 camera raster
struct ColorBatch {
  float p0x[kBatch], p0y[kBatch], abx[kBatch], aby[kBatch], inv_l2[kBatch];
  float d0[kBatch], d1[kBatch], w0[kBatch], w1[kBatch];
  float rmin[kBatch], rmax[kBatch];
  float mnx[kBatch], mny[kBatch], mxx[kBatch], mxy[kBatch];
  float4 s0[kBatch], s1[kBatch];
  float vis[kBatch];
  uint32_t valid[kBatch];
};

__global__ __launch_bounds__(256) void k_camera_raster(FrameParams fp) {
  __shared__ ColorBatch b;
  const uint32_t R = fp.counters[C_CAM_RUNS];
  for (uint32_t ri = blockIdx.x; ri < R; ri += gridDim.x) {
    const sr_raster_tile_run run = fp.cam_runs[ri];
    const uint32_t fi = run.frustum_id;
    const FrustumRec& fr = fp.fr[fi];
    if (fr.d.kind != 0u || run.work_count == 0u) continue;
    if ((uint64_t)run.work_base + run.work_count > fp.work_cap) continue;  // emission overflowed
    const uint32_t fx = fr.d.froxel_size_x, fy = fr.d.froxel_size_y;
    const uint32_t tx = run.screen_tile_id % fr.fine_tiles_x, ty = run.screen_tile_id / fr.fine_tiles_x;
    const uint32_t pidx = threadIdx.x;
    const uint32_t pxu = tx * fx + pidx % fx, pyu = ty * fy + pidx / fx;
    const bool active = pidx < fx * fy && pxu < fr.d.screen_width && pyu < fr.d.screen_height;
    const float cx = (float)pxu + 0.5f, cy = (float)pyu + 0.5f;
    const float vw = (float)fr.d.screen_width, vh = (float)fr.d.screen_height;
    const size_t layer_base = (size_t)fr.cam_index * fp.NI;
    float fc0 = 0.0f, fc1 = 0.0f, fc2 = 0.0f, fc3 = 0.0f, gdepth = 0.0f;
    bool done = !active;
    for (uint32_t w = 0; w < run.work_count; ++w) {
      const sr_raster_work_item wi = fp.work[run.work_base + w];
      const uint32_t end = wi.seg_ref_base + wi.seg_ref_count;
      for (uint32_t base = wi.seg_ref_base; base < end; base += kBatch) {
        __syncthreads();
        {
          uint32_t k = threadIdx.x, ref_idx = base + k;
          b.valid[k] = 0u;
          if (ref_idx < end) {
            const sr_fine_seg_ref sr = fp.refs[ref_idx];
            const InstRec& in = fp.inst[sr.inst_id];
            V3 a3 = world_v3(in, in.indices[sr.seg_id]);
            V3 c3 = world_v3(in, in.indices[sr.seg_id + 1u]);
            V4 a = world_to_screen_raw(V4{a3.x, a3.y, a3.z, 1.0f}, fr.clip, vw, vh);
            V4 c = world_to_screen_raw(V4{c3.x, c3.y, c3.z, 1.0f}, fr.clip, vw, vh);
            if (!(a.x < 0.0f && c.x < 0.0f)) {
              const sr_strand_material& m = in.materials[sr.material_idx];
              float minr = fmaxf(m.min_radius_pixels, 1e-4f);
              float maxr = fmaxf(m.max_radius_pixels, minr);
              const size_t lt = (layer_base + sr.inst_id) * fp.atlas_texels;
              uint32_t i0 = sr.seg_id, i1 = min(i0 + 1u, fp.atlas_texels - 1u);
              float abx = c.x - a.x, aby = c.y - a.y;
              b.p0x[k] = a.x; b.p0y[k] = a.y; b.abx[k] = abx; b.aby[k] = aby;
              b.inv_l2[k] = 1.0f / (abx * abx + aby * aby);
              b.d0[k] = a.z; b.d1[k] = c.z; b.w0[k] = a.w; b.w1[k] = c.w;
              b.rmin[k] = minr; b.rmax[k] = maxr;
              b.mnx[k] = fminf(a.x, c.x) - maxr; b.mny[k] = fminf(a.y, c.y) - maxr;
              b.mxx[k] = fmaxf(a.x, c.x) + maxr; b.mxy[k] = fmaxf(a.y, c.y) + maxr;
              b.s0[k] = fp.atlas[lt + i0];
              b.s1[k] = fp.atlas[lt + i1];
              b.vis[k] = fp.history[lt + i0].x;
              b.valid[k] = 1u;
            }
          }
        }
        __syncthreads();
        if (!done) {
          uint32_t n = min(kBatch, end - base);
          for (uint32_t k = 0; k < n; ++k) {
            if (!b.valid[k]) continue;
            if (cx < b.mnx[k] || cy < b.mny[k] || cx > b.mxx[k] || cy > b.mxy[k]) continue;
            float t = ((cx - b.p0x[k]) * b.abx[k] + (cy - b.p0y[k]) * b.aby[k]) * b.inv_l2[k];
            if (!(t >= 0.0f && t <= 1.0f)) continue;
            float qx = b.p0x[k] + b.abx[k] * t, qy = b.p0y[k] + b.aby[k] * t;
            float qz = wmix(b.d0[k], b.d1[k], t);
            float minr = fmaxf(b.rmin[k], 1e-4f);
            float r = wmix(minr, fmaxf(b.rmax[k], minr), wclamp(normalize_depth01(qz), 0.0f, 1.0f));
            float dx = cx - qx, dy = cy - qy;
            float d2 = dx * dx + dy * dy;
            if (!(d2 < r * r)) continue;
            float cov = wclamp(1.0f - sqrtf(d2) / r, 0.0f, 1.0f);
            float aw0 = fmaxf(fabsf(b.w0[k]), 1e-6f), aw1 = fmaxf(fabsf(b.w1[k]), 1e-6f);
            float tw = (t * aw0) / ((1.0f - t) * aw1 + t * aw0);
            float4 s0 = b.s0[k], s1 = b.s1[k];
            float sa = wmix(s0.w, s1.w, tw);
            float alpha = sa * cov;
            float vis = b.vis[k];
            float om = 1.0f - fc3;
            fc0 = fc0 + (wmix(s0.x, s1.x, tw) * vis * alpha) * om;
            fc1 = fc1 + (wmix(s0.y, s1.y, tw) * vis * alpha) * om;
            fc2 = fc2 + (wmix(s0.z, s1.z, tw) * vis * alpha) * om;
            fc3 = fc3 + alpha * om;
            gdepth = fmaxf(gdepth, qz);
          }
        }
      }
      if (!done && fc3 > 0.98f) {
        fc0 = fc0 / fc3; fc1 = fc1 / fc3; fc2 = fc2 / fc3; fc3 = 1.0f;
        done = true;
      }
      if (__syncthreads_count(!done) == 0) break;
    }
    if (active) {
      float o0 = fc0, o1 = fc1, o2 = fc2, o3 = fc3;
      if (o3 > 1e-6f && o3 < 1.0f) { o0 = o0 / o3; o1 = o1 / o3; o2 = o2 / o3; }
      size_t o = fr.image_off + (size_t)pyu * fr.d.screen_width + pxu;
      fp.image[o] = make_float4(wclamp(o0, 0.0f, 1.0f), wclamp(o1, 0.0f, 1.0f), wclamp(o2, 0.0f, 1.0f),
                                wclamp(o3, 0.0f, 1.0f));
      fp.depth[o] = gdepth;
    }
  }
}

// ------------------------------------------------------- K29 composite
__global__ void k_composite(FrameParams fp, uint64_t pixels) {
  const float4 sc = make_float4(fp.st.scene_clear_color[0], fp.st.scene_clear_color[1], fp.st.scene_clear_color[2],
                                fp.st.scene_clear_color[3]);
  const float sd = fp.st.scene_clear_depth;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < pixels; i += (uint64_t)gridDim.x * blockDim.x) {
    float d = fp.depth[i];
    if (sd > d) { fp.composite[i] = sc; continue; }
    float4 im = fp.image[i];
    fp.composite[i] = make_float4(wmix(sc.x, im.x, im.w), wmix(sc.y, im.y, im.w), wmix(sc.z, im.z, im.w), sc.w);
  }
}

// ------------------------------------------------------------ launchers
void launch_shadow(const FrameParams& fp, hipStream_t s) {
  hipLaunchKernelGGL(k_shadow_raster, dim3(256u * 8u), dim3(256), 0, s, fp);
}

void launch_shading(const FrameParams& fp, hipStream_t s) {
  hipLaunchKernelGGL(k_shading, dim3(256u * 16u), dim3(128), 0, s, fp);
}

void launch_camera_raster(const FrameParams& fp, hipStream_t s) {
  hipLaunchKernelGGL(k_camera_raster, dim3(256u * 8u), dim3(256), 0, s, fp);
}

void launch_stampback(const FrameParams& fp, uint32_t fi, uint32_t texels, hipStream_t s) {
  if (texels == 0) return;
  uint32_t g = (texels + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_stampback, dim3(g), dim3(256), 0, s, fp, fi);
}

void launch_composite(const FrameParams& fp, uint64_t pixels, hipStream_t s) {
  if (pixels == 0) return;
  uint64_t g = (pixels + 255) / 256;
  if (g > 8192) g = 8192;
  hipLaunchKernelGGL(k_composite, dim3((uint32_t)g), dim3(256), 0, s, fp, pixels);
}

}  // namespace srk
