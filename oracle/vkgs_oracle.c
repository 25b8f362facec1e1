/*
 * vkgs_oracle.c -- CPU restatement of vkgs's splat path (TEST INFRASTRUCTURE).
 *
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg.  The product (vkgs_amd/) never calls into this file.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  With
 * -ffp-contract=off every a*b+c below rounds twice, exactly as written; fmaf()
 * is used only where the HIP kernels use a fused multiply-add on purpose.
 *
 * Conventions shared with the HIP kernels:
 *  - mat4/mat3/mat2 are column-major (m[col*R + row]) as in GLSL/glm.
 *  - A product C = A*B sums k = 0..R-1 left to right: ((a0*b0 + a1*b1) + a2*b2) ...
 *  - GLSL expressions are evaluated left-associatively as written in the shader.
 *
 * Parity status (what pins each function; tests/test_oracle.py, tests/test_golden.py):
 *  - oracle_parse_ply: rotation vs scipy Rotation.from_quat, closed forms for
 *    identity quaternions, fp16 against numpy float16, sigmoid closed form.
 *  - oracle_rank: NDC of the camera frustum corners/centre (closed forms),
 *    float64 numpy projection + cull with margin, boundary cases.
 *  - oracle_sort_kv: numpy argsort(kind="stable").
 *  - oracle_inverse: permutation invariants.
 *  - oracle_projection: cov2d against first-order propagation of the 3D
 *    covariance through the NDC map (numerical Jacobian in float64),
 *    rot_scale*rot_scale^T == cov2d, SH basis against scipy spherical
 *    harmonics, DC-only colour closed form.
 *  - oracle_draw / oracle_bin: brute force over all pixels, closed forms (empty
 *    scene, single splat at its centre, two-splat over operator, uniform stack).
 *
 * Deliberate restatement choices (documented in DESIGN.md):
 *  - projection.comp:130-132 computes theta = atan(sin2t, cos2t)/2 and then
 *    cos/sin of it.  Here cos(theta), sin(theta) come from the half-angle
 *    identities cos = sqrt((1+cos2t)/2), sin = sign(sin2t) sqrt((1-cos2t)/2),
 *    which are algebraically identical for theta in (-pi/2, pi/2] and need only
 *    correctly rounded sqrt, so CPU and GPU agree to the bit.
 *  - Coverage is analytic (pixel centre inside the +-3 sigma square), not the
 *    fixed-point triangle rasterizer of the driver.
 */
#include "vkgs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int oracle_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------- fp16 --- */

uint16_t oracle_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) { /* inf or nan */
    return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? (0x200u | ((ax >> 13) & 0x3ffu)) : 0u));
  }
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u); /* rounds to >= 65520 -> inf */
  if (ax < 0x38800000u) {                                   /* result is subnormal or zero */
    if (ax < 0x33000000u) return (uint16_t)sign;            /* < 2^-25: rounds to 0 */
    uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    int shift = 126 - (int)(ax >> 23); /* 14 .. 24 */
    uint32_t half = mant >> shift;
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t mid = 1u << (shift - 1);
    if (rem > mid || (rem == mid && (half & 1u))) half++;
    return (uint16_t)(sign | half);
  }
  uint32_t e = (ax >> 23) - 112u; /* rebias 127 -> 15 */
  uint32_t mant = ax & 0x7fffffu;
  uint32_t h = (e << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (h & 1u))) h++;
  return (uint16_t)(sign | h);
}

float oracle_f16_to_f32(uint16_t h) {
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu;
  uint32_t m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else { /* subnormal */
      int k = 0;
      while (!(m & 0x400u)) { m <<= 1; k++; }
      m &= 0x3ffu;
      x = sign | ((uint32_t)(113 - k) << 23) | (m << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

/* ------------------------------------------------------------ matrices --- */

static void m4_mul(const float* a, const float* b, float* c) {
  for (int col = 0; col < 4; ++col)
    for (int row = 0; row < 4; ++row) {
      float s = a[0 * 4 + row] * b[col * 4 + 0];
      s = s + a[1 * 4 + row] * b[col * 4 + 1];
      s = s + a[2 * 4 + row] * b[col * 4 + 2];
      s = s + a[3 * 4 + row] * b[col * 4 + 3];
      c[col * 4 + row] = s;
    }
}

static void m4_mulv(const float* m, const float* v, float* o) {
  for (int row = 0; row < 4; ++row) {
    float s = m[0 * 4 + row] * v[0];
    s = s + m[1 * 4 + row] * v[1];
    s = s + m[2 * 4 + row] * v[2];
    s = s + m[3 * 4 + row] * v[3];
    o[row] = s;
  }
}

static void m3_mul(const float* a, const float* b, float* c) {
  float t[9];
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) {
      float s = a[0 * 3 + row] * b[col * 3 + 0];
      s = s + a[1 * 3 + row] * b[col * 3 + 1];
      s = s + a[2 * 3 + row] * b[col * 3 + 2];
      t[col * 3 + row] = s;
    }
  memcpy(c, t, sizeof(t));
}

static void m3_transpose(const float* a, float* t) {
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) t[col * 3 + row] = a[row * 3 + col];
}

static void m3_from_m4(const float* m4, float* m3) {
  for (int col = 0; col < 3; ++col)
    for (int row = 0; row < 3; ++row) m3[col * 3 + row] = m4[col * 4 + row];
}

/* S <- (A*S)*A^T, the shader's `cov3d = A * cov3d * transpose(A)`. */
static void m3_sandwich(const float* a, float* s) {
  float as[9], at[9];
  m3_mul(a, s, as);
  m3_transpose(a, at);
  m3_mul(as, at, s);
}

static void m2_mul(const float* a, const float* b, float* c) {
  float t[4];
  for (int col = 0; col < 2; ++col)
    for (int row = 0; row < 2; ++row) {
      float s = a[0 * 2 + row] * b[col * 2 + 0];
      s = s + a[1 * 2 + row] * b[col * 2 + 1];
      t[col * 2 + row] = s;
    }
  memcpy(c, t, sizeof(t));
}

static uint32_t f2u(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

/* ----------------------------------------------------------- parse_ply --- */

void oracle_parse_ply(const float* ply, const uint32_t offsets[60], uint32_t n, float* position,
                      float* cov3d, float* opacity, uint16_t* sh) {
  const uint32_t stride = offsets[59];
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t id = (uint32_t)ii;
    const float* row = ply + (size_t)stride * id;
    /* parse_ply.comp:46-48 */
    float s[3] = {expf(row[offsets[3]]), expf(row[offsets[4]]), expf(row[offsets[5]])};
    /* :50-52, q = (rot_1, rot_2, rot_3, rot_0) = (x, y, z, w) */
    float q[4] = {row[offsets[6]], row[offsets[7]], row[offsets[8]], row[offsets[9]]};
    float len = sqrtf(((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3]);
    for (int k = 0; k < 4; ++k) q[k] = q[k] / len;
    /* :54-72 */
    float xx = q[0] * q[0], yy = q[1] * q[1], zz = q[2] * q[2];
    float xy = q[0] * q[1], xz = q[0] * q[2], yz = q[1] * q[2];
    float wx = q[3] * q[0], wy = q[3] * q[1], wz = q[3] * q[2];
    float rot[9];
    rot[0 * 3 + 0] = 1.f - 2.f * (yy + zz);
    rot[0 * 3 + 1] = 2.f * (xy + wz);
    rot[0 * 3 + 2] = 2.f * (xz - wy);
    rot[1 * 3 + 0] = 2.f * (xy - wz);
    rot[1 * 3 + 1] = 1.f - 2.f * (xx + zz);
    rot[1 * 3 + 2] = 2.f * (yz + wx);
    rot[2 * 3 + 0] = 2.f * (xz + wy);
    rot[2 * 3 + 1] = 2.f * (yz - wx);
    rot[2 * 3 + 2] = 1.f - 2.f * (xx + yy);
    /* :74-78, cov3d = rot * ss * transpose(rot) */
    float ss[9] = {0};
    ss[0] = s[0] * s[0];
    ss[4] = s[1] * s[1];
    ss[8] = s[2] * s[2];
    float rs[9], rt[9], cov[9];
    m3_mul(rot, ss, rs);
    m3_transpose(rot, rt);
    m3_mul(rs, rt, cov);
    /* :80-85 */
    float* c = cov3d + 6 * (size_t)id;
    c[0] = cov[0 * 3 + 0];
    c[1] = cov[1 * 3 + 0];
    c[2] = cov[2 * 3 + 0];
    c[3] = cov[1 * 3 + 1];
    c[4] = cov[2 * 3 + 1];
    c[5] = cov[2 * 3 + 2];
    /* :87-89 */
    position[3 * (size_t)id + 0] = row[offsets[0]];
    position[3 * (size_t)id + 1] = row[offsets[1]];
    position[3 * (size_t)id + 2] = row[offsets[2]];
    /* :91-94 */
    for (int i = 0; i < 48; ++i) sh[48 * (size_t)id + i] = oracle_f32_to_f16(row[offsets[10 + i]]);
    /* :30, :96 */
    opacity[id] = 1.f / (1.f + expf(-row[offsets[58]]));
  }
}

/* ---------------------------------------------------------------- rank --- */

void oracle_rank(const oracle_camera* cam, const float model[16], const float* position, uint32_t n,
                 uint32_t* visible_count, uint32_t* key, uint32_t* index) {
  /* rank.comp:32 `projection * view * model * pos` = ((P*V)*M)*pos */
  float pv[16], pvm[16];
  m4_mul(cam->projection, cam->view, pv);
  m4_mul(pv, model, pvm);
  uint32_t c = 0;
  for (uint32_t id = 0; id < n; ++id) {
    float v[4] = {position[3 * (size_t)id + 0], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2],
                  1.f};
    float p[4];
    m4_mulv(pvm, v, p);
    /* :33 pos = pos / pos.w */
    float x = p[0] / p[3], y = p[1] / p[3], z = p[2] / p[3];
    /* :37 */
    if (fabsf(x) <= 1.f && fabsf(y) <= 1.f && z >= 0.f && z <= 1.f) {
      key[c] = f2u(1.f - z); /* :39 */
      index[c] = id;         /* :40 */
      c++;
    }
  }
  *visible_count = c;
}

/* ---------------------------------------------------------------- sort --- */

typedef struct {
  uint32_t key, pos, value;
} kv_item;

static int kv_cmp(const void* a, const void* b) {
  const kv_item* x = (const kv_item*)a;
  const kv_item* y = (const kv_item*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->pos < y->pos ? -1 : (x->pos > y->pos);
}

void oracle_sort_kv(uint32_t count, uint32_t* key, uint32_t* value) {
  kv_item* t = (kv_item*)malloc(sizeof(kv_item) * (count ? count : 1));
  for (uint32_t i = 0; i < count; ++i) {
    t[i].key = key[i];
    t[i].pos = i;
    t[i].value = value[i];
  }
  qsort(t, count, sizeof(kv_item), kv_cmp);
  for (uint32_t i = 0; i < count; ++i) {
    key[i] = t[i].key;
    value[i] = t[i].value;
  }
  free(t);
}

/* ------------------------------------------------------------- inverse --- */

void oracle_inverse(uint32_t count, const uint32_t* index, uint32_t n, int32_t* inverse) {
  for (uint32_t i = 0; i < n; ++i) inverse[i] = -1; /* engine.cc:1226 */
  for (uint32_t k = 0; k < count; ++k) inverse[index[k]] = (int32_t)k;
}

/* ---------------------------------------------------------- projection --- */

/* GLSL max(x, y) = x < y ? y : x (NaN in x propagates). */
static float glsl_max(float x, float y) { return x < y ? y : x; }

void oracle_projection(const oracle_camera* cam, const float model[16], const float model_inv[16],
                       const float* position, const float* cov3d, const float* opacity,
                       const uint16_t* sh, const int32_t* inverse, uint32_t n, uint32_t visible_count,
                       float* instances, uint32_t draw_indirect[12]) {
  /* projection.comp:64-75 */
  draw_indirect[0] = 6u * visible_count;
  draw_indirect[1] = 1;
  draw_indirect[2] = 0;
  draw_indirect[3] = 0;
  draw_indirect[4] = 0;
  draw_indirect[5] = draw_indirect[6] = draw_indirect[7] = 0;
  draw_indirect[8] = visible_count;
  draw_indirect[9] = 1;
  draw_indirect[10] = 0;
  draw_indirect[11] = 0;

  const float* P = cam->projection;
  const float* V = cam->view;
  float model3[9], view3[9];
  m3_from_m4(model, model3);
  m3_from_m4(V, view3);
  /* :85-86 camera position in model space */
  float eye4[4] = {cam->camera_position[0], cam->camera_position[1], cam->camera_position[2], 1.f};
  float cm[4];
  m4_mulv(model_inv, eye4, cm);
  {
    const float w = cm[3];
    cm[0] = cm[0] / w;
    cm[1] = cm[1] / w;
    cm[2] = cm[2] / w;
  }
  const float inv_w2 = (1.f / (float)cam->screen_size[0]) / (float)cam->screen_size[0];
  const float inv_h2 = (1.f / (float)cam->screen_size[1]) / (float)cam->screen_size[1];

#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
    const uint32_t id = (uint32_t)ii;
    int32_t slot = inverse[id]; /* :77-78 */
    if (slot == -1) continue;
    const float* c = cov3d + 6 * (size_t)id;
    /* :92 mat3(v0, v0.y, v1.xy, v0.z, v1.yz) */
    float S[9] = {c[0], c[1], c[2], c[1], c[3], c[4], c[2], c[4], c[5]};
    float p4[4] = {position[3 * (size_t)id + 0], position[3 * (size_t)id + 1], position[3 * (size_t)id + 2],
                   1.f};
    /* :87 dir = normalize(pos.xyz - camera_model_position.xyz) */
    float d[3] = {p4[0] - cm[0], p4[1] - cm[1], p4[2] - cm[2]};
    float dl = sqrtf((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2]);
    float dir[3] = {d[0] / dl, d[1] / dl, d[2] / dl};
    /* :95-97 */
    float t4[4];
    m3_sandwich(model3, S);
    m4_mulv(model, p4, t4);
    memcpy(p4, t4, sizeof(t4));
    /* :100-102 */
    m3_sandwich(view3, S);
    m4_mulv(V, p4, t4);
    memcpy(p4, t4, sizeof(t4));
    /* :105-108 */
    float x = p4[0], y = p4[1], z = p4[2];
    float r = sqrtf((x * x + y * y) + z * z);
    float J[9] = {-1.f / z, 0.f, (-2.f * x) / r, 0.f, -1.f / z, (-2.f * y) / r, (x / z) / z, (y / z) / z,
                  (-2.f * z) / r};
    m3_sandwich(J, S);
    /* :111-112 cov2d = projection_scale * mat2(cov3d) * projection_scale */
    float Ps[4] = {P[0], P[1], P[4], P[5]};
    float S2[4] = {S[0], S[1], S[3], S[4]};
    float cov2[4];
    m2_mul(Ps, S2, cov2);
    m2_mul(cov2, Ps, cov2);
    /* :115-116 low-pass */
    cov2[0] = cov2[0] + inv_w2;
    cov2[3] = cov2[3] + inv_h2;
    /* :121-134 eigendecomposition */
    float a = cov2[0], b = cov2[3], cc = cov2[2];
    float D = sqrtf((a - b) * (a - b) + (4.f * cc) * cc);
    float s0 = sqrtf(0.5f * ((a + b) + D));
    float s1 = sqrtf(0.5f * ((a + b) - D));
    float sin2t = (2.f * cc) / D;
    float cos2t = (a - b) / D;
    float ct = (1.f + cos2t) * 0.5f;
    float st = (1.f - cos2t) * 0.5f;
    ct = ct < 0.f ? 0.f : ct;
    st = st < 0.f ? 0.f : st;
    float cos_t = sqrtf(ct);
    float sin_t = copysignf(sqrtf(st), sin2t);
    if (sin2t != sin2t || cos2t != cos2t) { /* atan(NaN) propagates */
      cos_t = NAN;
      sin_t = NAN;
    }
    float rs00 = s0 * cos_t, rs01 = s0 * sin_t, rs10 = (-s1) * sin_t, rs11 = s1 * cos_t;
    /* A NaN anywhere in the eigendecomposition (D = 0 for an isotropic cov2d,
     * or a + b < D by rounding for a needle) makes the whole rot_scale the
     * canonical quiet NaN, so its bits do not depend on how a platform
     * propagates NaN signs; the splat draws nothing either way. */
    if (rs00 != rs00 || rs01 != rs01 || rs10 != rs10 || rs11 != rs11) {
      union { uint32_t u; float f; } qnan = {0x7fc00000u};
      rs00 = rs01 = rs10 = rs11 = qnan.f;
    }
    /* :136-137 */
    m4_mulv(P, p4, t4);
    float ndc[3] = {t4[0] / t4[3], t4[1] / t4[3], t4[2] / t4[3]};
    /* :140-171 spherical harmonics */
    const float C0 = 0.28209479177387814f, C1 = 0.4886025119029199f, C20 = 1.0925484305920792f,
                C21 = 0.31539156525252005f, C22 = 0.5462742152960396f, C30 = 0.5900435899266435f,
                C31 = 2.890611442640554f, C32 = 0.4570457994644658f, C33 = 0.3731763325901154f,
                C34 = 1.445305721320277f;
    float dx = dir[0], dy = dir[1], dz = dir[2];
    float xx = dx * dx, yy = dy * dy, zz = dz * dz, xy = dx * dy, yz = dy * dz, xz = dx * dz;
    float basis[16];
    basis[0] = C0;
    basis[1] = -C1 * dy;
    basis[2] = C1 * dz;
    basis[3] = -C1 * dx;
    basis[4] = C20 * xy;
    basis[5] = -C20 * yz;
    basis[6] = C21 * ((2.f * zz - xx) - yy);
    basis[7] = -C20 * xz;
    basis[8] = C22 * (xx - yy);
    basis[9] = (-C30 * dy) * (3.f * xx - yy);
    basis[10] = (C31 * xy) * dz;
    basis[11] = (-C32 * dy) * ((4.f * zz - xx) - yy);
    basis[12] = (C33 * dz) * ((2.f * zz - 3.f * xx) - 3.f * yy);
    basis[13] = (-C32 * dx) * ((4.f * zz - xx) - yy);
    basis[14] = (C34 * dz) * (xx - yy);
    basis[15] = (-C30 * dx) * (xx - 3.f * yy);
    const uint16_t* h = sh + 48 * (size_t)id;
    float color[3];
    for (int ch = 0; ch < 3; ++ch) {
      float acc = 0.f;
      for (int j = 0; j < 4; ++j) {
        const uint16_t* q = h + ch * 16 + 4 * j;
        float dot = basis[4 * j + 0] * oracle_f16_to_f32(q[0]);
        dot = dot + basis[4 * j + 1] * oracle_f16_to_f32(q[1]);
        dot = dot + basis[4 * j + 2] * oracle_f16_to_f32(q[2]);
        dot = dot + basis[4 * j + 3] * oracle_f16_to_f32(q[3]);
        acc = (j == 0) ? dot : acc + dot;
      }
      color[ch] = glsl_max(acc + 0.5f, 0.f); /* :174 */
    }
    /* :177-179 */
    float* inst = instances + 12 * (size_t)slot;
    inst[0] = ndc[0];
    inst[1] = ndc[1];
    inst[2] = ndc[2];
    inst[3] = 0.f;
    inst[4] = rs00;
    inst[5] = rs01;
    inst[6] = rs10;
    inst[7] = rs11;
    inst[8] = color[0];
    inst[9] = color[1];
    inst[10] = color[2];
    inst[11] = opacity[id];
  }
}

/* -------------------------------------------------------------- raster --- */

/* Centre, inverse pixel-space rot_scale and colour; returns 0 when the depth
 * test rejects the quad. */
static int setup_core(const float* inst, uint32_t width, uint32_t height, oracle_splat2d* s, float* ex,
                      float* ey) {
  const float cx = inst[0], cy = inst[1], cz = inst[2];
  /* rot_scale = mat2(inst1.xy, inst1.zw) (splat.vert:19): col0 = (m00, m10), col1 = (m01, m11) */
  const float m00 = inst[4], m10 = inst[5], m01 = inst[6], m11 = inst[7];
  s->r = inst[8];
  s->g = inst[9];
  s->b = inst[10];
  s->o = inst[11];
  /* depth test LESS against the cleared 1.0 with depth range [0,1]
   * (engine.cc:298-299, graphics_pipeline.cc:77-81, viewport engine.cc:1419-1426) */
  if (!(cz >= 0.f && cz < 1.f)) return 0;
  /* viewport transform x_fb = (x_ndc + 1) * W/2 (engine.cc:1419-1426) */
  const float hw = 0.5f * (float)width, hh = 0.5f * (float)height;
  s->cx = (cx + 1.f) * hw;
  s->cy = (cy + 1.f) * hh;
  const float a00 = hw * m00, a01 = hw * m01, a10 = hh * m10, a11 = hh * m11;
  const float det = a00 * a11 - a01 * a10;
  const float inv = 1.f / det;
  s->i00 = a11 * inv;
  s->i01 = -(a01 * inv);
  s->i10 = -(a10 * inv);
  s->i11 = a00 * inv;
  /* half extents of the quad c + A*(+-3, +-3) (splat.vert:21-23), padded */
  float hx = 3.f * (fabsf(a00) + fabsf(a01));
  float hy = 3.f * (fabsf(a10) + fabsf(a11));
  *ex = hx * 1.00390625f + 0.015625f;
  *ey = hy * 1.00390625f + 0.015625f;
  return 1;
}

int oracle_setup_splat(const float* inst, uint32_t width, uint32_t height, oracle_splat2d* s) {
  float ex, ey;
  if (!setup_core(inst, width, height, s, &ex, &ey)) return 0;
  if (!(isfinite(s->cx) && isfinite(s->cy) && isfinite(s->i00) && isfinite(s->i01) && isfinite(s->i10) &&
        isfinite(s->i11) && isfinite(ex) && isfinite(ey)))
    return 0;
  float lox = ceilf(s->cx - ex - 0.5f), hix = floorf(s->cx + ex - 0.5f);
  float loy = ceilf(s->cy - ey - 0.5f), hiy = floorf(s->cy + ey - 0.5f);
  if (lox < 0.f) lox = 0.f;
  if (loy < 0.f) loy = 0.f;
  if (hix > (float)(width - 1)) hix = (float)(width - 1);
  if (hiy > (float)(height - 1)) hiy = (float)(height - 1);
  if (!(lox <= hix && loy <= hiy)) return 0;
  s->x0 = (int32_t)lox;
  s->x1 = (int32_t)hix;
  s->y0 = (int32_t)loy;
  s->y1 = (int32_t)hiy;
  return 1;
}

/* Coverage of pixel (px, py) by the +-3 sigma square (splat.vert:21-25) in the
 * eigenframe scaled by ORACLE_FRAME_SCALE = sqrt(log2(e) / 2), u' = s u, so
 * the square is max(|u'_x|, |u'_y|) <= fl(3 s) and exp(-|u|^2 / 2) =
 * exp2(-|u'|^2).  u' is formed from the corner (X0, Y0) of the pixel's 64 x 64
 * block with the operations of the HIP raster, so both decide coverage
 * identically:
 *   i' = s * (row of the inverse rot_scale), o' = fma(i'0, X0 - cx, i'1 * (Y0 - cy)),
 *   u' = fma(i'0, px + 0.5 - X0, fma(i'1, py + 0.5 - Y0, o')). */
#define ORACLE_FRAME_SCALE 0.849321800288019f
#define ORACLE_COVER_SCALED (3.0f * ORACLE_FRAME_SCALE)

int oracle_eval_pixel(const oracle_splat2d* s, int px, int py, float* alpha) {
  const int bx = px & ~63, by = py & ~63;
  const float X0 = (float)bx, Y0 = (float)by;
  const float i00 = s->i00 * ORACLE_FRAME_SCALE, i01 = s->i01 * ORACLE_FRAME_SCALE;
  const float i10 = s->i10 * ORACLE_FRAME_SCALE, i11 = s->i11 * ORACLE_FRAME_SCALE;
  const float dx = X0 - s->cx, dy = Y0 - s->cy;
  const float ox = fmaf(i00, dx, i01 * dy), oy = fmaf(i10, dx, i11 * dy);
  const float lx = (float)(px - bx) + 0.5f, ly = (float)(py - by) + 0.5f;
  const float ux = fmaf(i00, lx, fmaf(i01, ly, ox));
  const float uy = fmaf(i10, lx, fmaf(i11, ly, oy));
  if (!(fmaxf(fabsf(ux), fabsf(uy)) <= ORACLE_COVER_SCALED)) return 0;
  /* splat.frag:9-10: alpha = opacity * exp(-|u|^2 / 2) */
  *alpha = s->o * exp2f(-fmaf(ux, ux, uy * uy));
  return 1;
}

static void clear_image(float* rgba, size_t npix) {
  for (size_t i = 0; i < npix; ++i) {
    rgba[4 * i + 0] = 0.f;
    rgba[4 * i + 1] = 0.f;
    rgba[4 * i + 2] = 0.f;
    rgba[4 * i + 3] = 1.f;
  }
}

/* Fixed-function blend: SRC_ALPHA, ONE_MINUS_SRC_ALPHA for colour and alpha. */
static void blend(float* px, const oracle_splat2d* s, float a) {
  const float om = 1.f - a;
  px[0] = s->r * a + px[0] * om;
  px[1] = s->g * a + px[1] * om;
  px[2] = s->b * a + px[2] * om;
  px[3] = a * a + px[3] * om;
}

void oracle_draw(const float* instances, uint32_t count, uint32_t width, uint32_t height, float* rgba) {
  clear_image(rgba, (size_t)width * height);
  oracle_splat2d* set = (oracle_splat2d*)malloc(sizeof(oracle_splat2d) * (count ? count : 1));
  uint8_t* ok = (uint8_t*)malloc(count ? count : 1);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)count; ++k) ok[k] = (uint8_t)oracle_setup_splat(instances + 12 * k, width, height, &set[k]);
  /* each thread owns a band of rows and walks every splat in draw order */
#pragma omp parallel
  {
    int nt = 1, t = 0;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    t = omp_get_thread_num();
#endif
    int band = ((int)height + nt - 1) / nt;
    int r0 = t * band, r1 = r0 + band;
    if (r1 > (int)height) r1 = (int)height;
    for (uint32_t k = 0; k < count; ++k) {
      if (!ok[k]) continue;
      const oracle_splat2d* s = &set[k];
      int y0 = s->y0 > r0 ? s->y0 : r0;
      int y1 = s->y1 < r1 - 1 ? s->y1 : r1 - 1;
      for (int py = y0; py <= y1; ++py)
        for (int px = s->x0; px <= s->x1; ++px) {
          float a;
          if (oracle_eval_pixel(s, px, py, &a)) blend(rgba + 4 * ((size_t)py * width + px), s, a);
        }
    }
  }
  free(set);
  free(ok);
}

void oracle_draw_bruteforce(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                            float* rgba) {
  clear_image(rgba, (size_t)width * height);
  for (uint32_t k = 0; k < count; ++k) {
    oracle_splat2d s;
    float ex, ey;
    if (!setup_core(instances + 12 * (size_t)k, width, height, &s, &ex, &ey)) continue;
    for (uint32_t py = 0; py < height; ++py)
      for (uint32_t px = 0; px < width; ++px) {
        float a;
        if (oracle_eval_pixel(&s, (int)px, (int)py, &a)) blend(rgba + 4 * ((size_t)py * width + px), &s, a);
      }
  }
}

/* ------------------------------------------------- triangle rasterization --- */

/* The reference's draw as the fixed-function pipeline executes it: each
 * instance is a quad of two triangles, index pattern 4i + {0,1,2, 2,1,3}
 * (engine.cc:607-616); vertex k sits at ndc + (rot_scale * q_k) * 3 with
 * q_k = (2 (k / 2) - 1, 2 (k % 2) - 1) and carries the varying p = 3 q_k
 * (splat.vert:12-25); the viewport maps x_fb = W/2 x_ndc + W/2 (engine.cc:
 * 1419-1426); vertices are snapped to 1 / 2^subpixel_bits of a pixel (round
 * to nearest, Vulkan's subPixelPrecisionBits, 8 on common implementations);
 * back faces are culled with CCW front faces (graphics_pipeline.cc:72-73),
 * using Vulkan's signed area a = -1/2 sum(x_i y_i+1 - x_i+1 y_i) in
 * framebuffer coordinates (y down); a pixel centre is covered when it is
 * inside all three edges, a centre exactly on an edge only for a top or left
 * edge (the tie rule), so the shared diagonal is covered once; p is
 * interpolated linearly from the snapped vertices (w = 1, so perspective
 * correction is the identity) and alpha = opacity * exp(-p.p / 2)
 * (splat.frag:9-11); the depth test is LESS against 1.0 with z_fb = z_ndc,
 * and z < 0 is clipped.  The driver's exact fixed-point and interpolation
 * arithmetic are not public; this restatement uses exact integer edge
 * functions and double-precision barycentrics. */
static int64_t snap_coord(float v, int bits) { return (int64_t)llrint((double)v * (double)(1 << bits)); }

static void draw_triangle(const int64_t X[3], const int64_t Y[3], const float P[3][2], const float* inst,
                          uint32_t width, uint32_t height, int bits, float* rgba) {
  /* Vulkan's signed area in framebuffer coordinates (y down): a > 0 is CCW, front */
  const int64_t cross = (X[1] - X[0]) * (Y[2] - Y[0]) - (Y[1] - Y[0]) * (X[2] - X[0]);
  /* a = -1/2 (sum x_i y_i+1 - x_i+1 y_i) = -cross / 2 */
  if (!(-cross > 0)) return; /* back-facing (culled) or degenerate */
  /* interior-positive edge functions: with cross < 0, E_i(P) = -((B - A) x (P - A)) > 0 inside */
  int64_t xmin = X[0], xmax = X[0], ymin = Y[0], ymax = Y[0];
  for (int i = 1; i < 3; ++i) {
    xmin = X[i] < xmin ? X[i] : xmin;
    xmax = X[i] > xmax ? X[i] : xmax;
    ymin = Y[i] < ymin ? Y[i] : ymin;
    ymax = Y[i] > ymax ? Y[i] : ymax;
  }
  const int64_t one = (int64_t)1 << bits, half = one >> 1;
  int64_t px0 = (xmin - half + one - 1) >> bits, px1 = (xmax - half) >> bits;
  int64_t py0 = (ymin - half + one - 1) >> bits, py1 = (ymax - half) >> bits;
  if (px0 < 0) px0 = 0;
  if (py0 < 0) py0 = 0;
  if (px1 > (int64_t)width - 1) px1 = (int64_t)width - 1;
  if (py1 > (int64_t)height - 1) py1 = (int64_t)height - 1;
  const double area = (double)(-cross);
  for (int64_t py = py0; py <= py1; ++py)
    for (int64_t px = px0; px <= px1; ++px) {
      const int64_t sx = px * one + half, sy = py * one + half;
      int64_t e[3];
      int inside = 1;
      for (int i = 0; i < 3 && inside; ++i) {
        const int a = i, b = (i + 1) % 3;
        const int64_t dx = X[b] - X[a], dy = Y[b] - Y[a];
        e[i] = -(dx * (sy - Y[a]) - dy * (sx - X[a]));
        /* interior-positive orientation: a left edge has dy > 0 here, a top edge dy == 0, dx < 0
         * (the negation above mirrors the orientation of the y-down cross product) */
        const int top_left = (dy > 0) || (dy == 0 && dx < 0);
        if (e[i] < 0 || (e[i] == 0 && !top_left)) inside = 0;
      }
      if (!inside) continue;
      /* barycentric weight of vertex c is the edge function of the edge opposite c */
      const double l0 = (double)e[1] / area, l1 = (double)e[2] / area, l2 = (double)e[0] / area;
      const double p0 = l0 * P[0][0] + l1 * P[1][0] + l2 * P[2][0];
      const double p1 = l0 * P[0][1] + l1 * P[1][1] + l2 * P[2][1];
      const float g = expf(-0.5f * (float)(p0 * p0 + p1 * p1));
      const float a = inst[11] * g;
      float* dst = rgba + 4 * ((size_t)py * width + (size_t)px);
      const float om = 1.f - a;
      dst[0] = inst[8] * a + dst[0] * om;
      dst[1] = inst[9] * a + dst[1] * om;
      dst[2] = inst[10] * a + dst[2] * om;
      dst[3] = a * a + dst[3] * om;
    }
}

void oracle_draw_triangles(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                           uint32_t subpixel_bits, float* rgba) {
  clear_image(rgba, (size_t)width * height);
  const int bits = (int)subpixel_bits;
  const float hw = 0.5f * (float)width, hh = 0.5f * (float)height;
#pragma omp parallel
  {
    int nt = 1, t = 0;
#ifdef _OPENMP
    nt = omp_get_num_threads();
    t = omp_get_thread_num();
#endif
    /* each thread draws every splat into its own band of rows, in draw order */
    const int band = ((int)height + nt - 1) / nt;
    const int r0 = t * band, r1 = r0 + band < (int)height ? r0 + band : (int)height;
    for (uint32_t k = 0; k < count && r0 < r1; ++k) {
      const float* inst = instances + 12 * (size_t)k;
      const float z = inst[2];
      if (!(z >= 0.f && z < 1.f)) continue;
      const float m00 = inst[4], m10 = inst[5], m01 = inst[6], m11 = inst[7];
      int64_t X[4], Y[4];
      float P[4][2];
      int ok = 1;
      for (int v = 0; v < 4 && ok; ++v) {
        const float qx = (float)(v / 2) * 2.f - 1.f, qy = (float)(v % 2) * 2.f - 1.f;
        /* (rot_scale * position) * confidence_radius, splat.vert:19-23 */
        const float rx = (m00 * qx + m01 * qy) * 3.f, ry = (m10 * qx + m11 * qy) * 3.f;
        const float xn = inst[0] + rx, yn = inst[1] + ry;
        const float xf = xn * hw + hw, yf = yn * hh + hh;
        if (!(isfinite(xf) && isfinite(yf)) || fabsf(xf) > 1e6f || fabsf(yf) > 1e6f) ok = 0;
        X[v] = ok ? snap_coord(xf, bits) : 0;
        Y[v] = ok ? snap_coord(yf, bits) : 0;
        P[v][0] = qx * 3.f;
        P[v][1] = qy * 3.f;
      }
      if (!ok) continue;
      /* restrict to this thread's rows: the triangle loop clamps to the image, so clamp the band here */
      const int64_t tri[2][3] = {{0, 1, 2}, {2, 1, 3}};
      for (int ti = 0; ti < 2; ++ti) {
        int64_t tx[3], ty[3];
        float tp[3][2];
        for (int j = 0; j < 3; ++j) {
          tx[j] = X[tri[ti][j]];
          ty[j] = Y[tri[ti][j]];
          tp[j][0] = P[tri[ti][j]][0];
          tp[j][1] = P[tri[ti][j]][1];
        }
        /* draw into rows [r0, r1) only: offset the image and its height */
        const int64_t off = (int64_t)r0 << bits;
        int64_t ty_b[3] = {ty[0] - off, ty[1] - off, ty[2] - off};
        draw_triangle(tx, ty_b, tp, inst, width, (uint32_t)(r1 - r0), bits, rgba + 4 * (size_t)r0 * width);
      }
    }
  }
}

uint64_t oracle_bin(const float* instances, uint32_t count, uint32_t width, uint32_t height,
                    uint32_t tile_w, uint32_t tile_h, uint32_t* ranges, uint32_t* list) {
  const uint32_t tx_n = (width + tile_w - 1) / tile_w, ty_n = (height + tile_h - 1) / tile_h;
  const uint32_t nt = tx_n * ty_n;
  uint64_t* cnt = (uint64_t*)calloc(nt + 1, sizeof(uint64_t));
  for (uint32_t k = 0; k < count; ++k) {
    oracle_splat2d s;
    if (!oracle_setup_splat(instances + 12 * (size_t)k, width, height, &s)) continue;
    for (int ty = s.y0 / (int)tile_h; ty <= s.y1 / (int)tile_h; ++ty)
      for (int tx = s.x0 / (int)tile_w; tx <= s.x1 / (int)tile_w; ++tx) cnt[(uint32_t)ty * tx_n + tx]++;
  }
  uint64_t total = 0;
  for (uint32_t t = 0; t < nt; ++t) {
    uint64_t c = cnt[t];
    if (ranges) {
      ranges[2 * t + 0] = (uint32_t)total;
      ranges[2 * t + 1] = (uint32_t)(total + c);
    }
    cnt[t] = total;
    total += c;
  }
  if (list) {
    for (uint32_t k = 0; k < count; ++k) {
      oracle_splat2d s;
      if (!oracle_setup_splat(instances + 12 * (size_t)k, width, height, &s)) continue;
      for (int ty = s.y0 / (int)tile_h; ty <= s.y1 / (int)tile_h; ++ty)
        for (int tx = s.x0 / (int)tile_w; tx <= s.x1 / (int)tile_w; ++tx) list[cnt[(uint32_t)ty * tx_n + tx]++] = k;
    }
  }
  free(cnt);
  return total;
}
