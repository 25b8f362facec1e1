// Host-side pieces of the C ABI: status strings, camera producer, 4x4 inverse,
// frame constants and the checked PLY reader.  No device work happens here.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <unordered_map>

#include "common.h"

namespace vkgs {

void make_frame_const(const vkgs_camera* cam, const float* model, FrameConst* fc) {
  static const float kIdentity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  if (model == nullptr) model = kIdentity;
  memset(fc, 0, sizeof(*fc));
  memcpy(fc->P, cam->projection, sizeof(fc->P));
  memcpy(fc->V, cam->view, sizeof(fc->V));
  memcpy(fc->M, model, sizeof(fc->M));
  float pv[16];
  m4_mul(fc->P, fc->V, pv);  // rank.comp:32, ((P*V)*M)
  m4_mul(pv, fc->M, fc->PVM);
  float minv[16];
  if (vkgs_mat4_inverse(model, minv) != VKGS_OK) memcpy(minv, kIdentity, sizeof(minv));
  const float eye4[4] = {cam->camera_position[0], cam->camera_position[1], cam->camera_position[2], 1.f};
  float cm[4];
  m4_mulv(minv, eye4, cm);  // projection.comp:85-86
  const float w = cm[3];
  fc->eye_model[0] = cm[0] / w;
  fc->eye_model[1] = cm[1] / w;
  fc->eye_model[2] = cm[2] / w;
  fc->eye_model[3] = 1.f;
  const float W = (float)cam->screen_size[0], H = (float)cam->screen_size[1];
  fc->inv_w2 = (1.f / W) / W;  // projection.comp:115-116
  fc->inv_h2 = (1.f / H) / H;
  fc->width = cam->screen_size[0];
  fc->height = cam->screen_size[1];
}

}  // namespace vkgs

using namespace vkgs;

extern "C" {

const char* vkgs_status_string(int status) {
  switch (status) {
    case VKGS_OK: return "ok";
    case VKGS_ERR_INVALID_ARG: return "invalid argument";
    case VKGS_ERR_CAPACITY: return "capacity exceeded";
    case VKGS_ERR_HIP: return "HIP runtime error";
    case VKGS_ERR_IO: return "file not found or unreadable";
    case VKGS_ERR_FORMAT: return "malformed or unsupported PLY";
    case VKGS_ERR_OVERFLOW: return "device list overflow";
    case VKGS_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

int vkgs_abi_version(void) { return VKGS_ABI_VERSION; }

void vkgs_raster_config_default(vkgs_raster_config* cfg) {
  if (!cfg) return;
  memset(cfg, 0, sizeof(*cfg));
  cfg->tile_w = 8;
  cfg->tile_h = 8;
  cfg->t_min = 1.f / 8192.f;
}

// glm::perspective / glm::lookAt restated from their defining formulas
// (gluPerspective, gluLookAt), then camera.cc:29-34's conversion to Vulkan
// clip space.  Eye position per camera.cc:39-45.
int vkgs_camera_orbit(float fovy, uint32_t width, uint32_t height, float znear, float zfar, const float center[3],
                      float r, float phi, float theta, vkgs_camera* out) {
  if (!out || !center || width == 0 || height == 0 || !(zfar > znear) || !(znear > 0.f) || !(fovy > 0.f) ||
      !(fovy < 3.14159265f))
    return VKGS_ERR_INVALID_ARG;
  memset(out, 0, sizeof(*out));
  const float aspect = (float)width / (float)height;  // camera.cc:26
  const float t = tanf(fovy / 2.f);
  float P[16] = {0};
  P[0] = 1.f / (aspect * t);
  P[5] = 1.f / t;
  P[10] = -(zfar + znear) / (zfar - znear);
  P[11] = -1.f;
  P[14] = -(2.f * zfar * znear) / (zfar - znear);
  float C[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
  C[5] = -1.f;   // conversion[1][1]
  C[10] = 0.5f;  // conversion[2][2]
  C[14] = 0.5f;  // conversion[3][2]
  m4_mul(C, P, out->projection);

  const float sp = sinf(phi), cp = cosf(phi), st = sinf(theta), ct = cosf(theta);
  const float eye[3] = {center[0] + r * (sp * st), center[1] + r * cp, center[2] + r * (sp * ct)};
  float f[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
  float fl = sqrtf((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]);
  for (int i = 0; i < 3; ++i) f[i] = f[i] / fl;
  const float up[3] = {0.f, 1.f, 0.f};
  float s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
  float sl = sqrtf((s[0] * s[0] + s[1] * s[1]) + s[2] * s[2]);
  if (!(sl > 0.f)) return VKGS_ERR_INVALID_ARG;  // looking straight along +-y
  for (int i = 0; i < 3; ++i) s[i] = s[i] / sl;
  const float u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
  float* V = out->view;
  V[0] = s[0];
  V[4] = s[1];
  V[8] = s[2];
  V[1] = u[0];
  V[5] = u[1];
  V[9] = u[2];
  V[2] = -f[0];
  V[6] = -f[1];
  V[10] = -f[2];
  V[12] = -((s[0] * eye[0] + s[1] * eye[1]) + s[2] * eye[2]);
  V[13] = -((u[0] * eye[0] + u[1] * eye[1]) + u[2] * eye[2]);
  V[14] = (f[0] * eye[0] + f[1] * eye[1]) + f[2] * eye[2];
  V[15] = 1.f;
  out->camera_position[0] = eye[0];
  out->camera_position[1] = eye[1];
  out->camera_position[2] = eye[2];
  out->screen_size[0] = width;
  out->screen_size[1] = height;
  return VKGS_OK;
}

int vkgs_mat4_inverse(const float m[16], float out[16]) {
  if (!m || !out) return VKGS_ERR_INVALID_ARG;
  // cofactor expansion in double, rounded once to float
  double a[16], inv[16];
  for (int i = 0; i < 16; ++i) a[i] = m[i];
  inv[0] = a[5] * a[10] * a[15] - a[5] * a[11] * a[14] - a[9] * a[6] * a[15] + a[9] * a[7] * a[14] +
           a[13] * a[6] * a[11] - a[13] * a[7] * a[10];
  inv[4] = -a[4] * a[10] * a[15] + a[4] * a[11] * a[14] + a[8] * a[6] * a[15] - a[8] * a[7] * a[14] -
           a[12] * a[6] * a[11] + a[12] * a[7] * a[10];
  inv[8] = a[4] * a[9] * a[15] - a[4] * a[11] * a[13] - a[8] * a[5] * a[15] + a[8] * a[7] * a[13] +
           a[12] * a[5] * a[11] - a[12] * a[7] * a[9];
  inv[12] = -a[4] * a[9] * a[14] + a[4] * a[10] * a[13] + a[8] * a[5] * a[14] - a[8] * a[6] * a[13] -
            a[12] * a[5] * a[10] + a[12] * a[6] * a[9];
  inv[1] = -a[1] * a[10] * a[15] + a[1] * a[11] * a[14] + a[9] * a[2] * a[15] - a[9] * a[3] * a[14] -
           a[13] * a[2] * a[11] + a[13] * a[3] * a[10];
  inv[5] = a[0] * a[10] * a[15] - a[0] * a[11] * a[14] - a[8] * a[2] * a[15] + a[8] * a[3] * a[14] +
           a[12] * a[2] * a[11] - a[12] * a[3] * a[10];
  inv[9] = -a[0] * a[9] * a[15] + a[0] * a[11] * a[13] + a[8] * a[1] * a[15] - a[8] * a[3] * a[13] -
           a[12] * a[1] * a[11] + a[12] * a[3] * a[9];
  inv[13] = a[0] * a[9] * a[14] - a[0] * a[10] * a[13] - a[8] * a[1] * a[14] + a[8] * a[2] * a[13] +
            a[12] * a[1] * a[10] - a[12] * a[2] * a[9];
  inv[2] = a[1] * a[6] * a[15] - a[1] * a[7] * a[14] - a[5] * a[2] * a[15] + a[5] * a[3] * a[14] +
           a[13] * a[2] * a[7] - a[13] * a[3] * a[6];
  inv[6] = -a[0] * a[6] * a[15] + a[0] * a[7] * a[14] + a[4] * a[2] * a[15] - a[4] * a[3] * a[14] -
           a[12] * a[2] * a[7] + a[12] * a[3] * a[6];
  inv[10] = a[0] * a[5] * a[15] - a[0] * a[7] * a[13] - a[4] * a[1] * a[15] + a[4] * a[3] * a[13] +
            a[12] * a[1] * a[7] - a[12] * a[3] * a[5];
  inv[14] = -a[0] * a[5] * a[14] + a[0] * a[6] * a[13] + a[4] * a[1] * a[14] - a[4] * a[2] * a[13] -
            a[12] * a[1] * a[6] + a[12] * a[2] * a[5];
  inv[3] = -a[1] * a[6] * a[11] + a[1] * a[7] * a[10] + a[5] * a[2] * a[11] - a[5] * a[3] * a[10] -
           a[9] * a[2] * a[7] + a[9] * a[3] * a[6];
  inv[7] = a[0] * a[6] * a[11] - a[0] * a[7] * a[10] - a[4] * a[2] * a[11] + a[4] * a[3] * a[10] +
           a[8] * a[2] * a[7] - a[8] * a[3] * a[6];
  inv[11] = -a[0] * a[5] * a[11] + a[0] * a[7] * a[9] + a[4] * a[1] * a[11] - a[4] * a[3] * a[9] -
            a[8] * a[1] * a[7] + a[8] * a[3] * a[5];
  inv[15] = a[0] * a[5] * a[10] - a[0] * a[6] * a[9] - a[4] * a[1] * a[10] + a[4] * a[2] * a[9] +
            a[8] * a[1] * a[6] - a[8] * a[2] * a[5];
  const double det = a[0] * inv[0] + a[1] * inv[4] + a[2] * inv[8] + a[3] * inv[12];
  if (!(fabs(det) > 0.0) || !isfinite(det)) return VKGS_ERR_INVALID_ARG;
  for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
  return VKGS_OK;
}

// ------------------------------------------------------------------ PLY ---

static int ply_type_size(const std::string& t) {
  if (t == "char" || t == "uchar" || t == "int8" || t == "uint8") return 1;
  if (t == "short" || t == "ushort" || t == "int16" || t == "uint16") return 2;
  if (t == "int" || t == "uint" || t == "float" || t == "int32" || t == "uint32" || t == "float32") return 4;
  if (t == "double" || t == "float64") return 8;
  return -1;
}

int vkgs_ply_read_header(const char* path, vkgs_ply_info* info) {
  if (!path || !info) return VKGS_ERR_INVALID_ARG;
  FILE* f = fopen(path, "rb");
  if (!f) return VKGS_ERR_IO;
  memset(info, 0, sizeof(*info));
  std::unordered_map<std::string, std::pair<int, std::string>> props;  // name -> (byte offset, type)
  std::string line;
  int c;
  int line_no = 0;
  bool in_vertex = false, saw_vertex = false, saw_format = false, ok_end = false;
  uint64_t row_bytes = 0;
  uint64_t header_bytes = 0;
  int status = VKGS_OK;
  while (status == VKGS_OK) {
    line.clear();
    while ((c = fgetc(f)) != EOF && c != '\n') {
      line.push_back((char)c);
      if (line.size() > 4096) break;
    }
    header_bytes = (uint64_t)ftell(f);
    if (c == EOF || line.size() > 4096 || header_bytes > (1u << 20)) {
      status = VKGS_ERR_FORMAT;
      break;
    }
    if (!line.empty() && line.back() == '\r') line.pop_back();
    ++line_no;
    char w0[64] = {0}, w1[64] = {0}, w2[64] = {0}, w3[64] = {0};
    const int nw = sscanf(line.c_str(), "%63s %63s %63s %63s", w0, w1, w2, w3);
    if (line_no == 1) {
      if (strcmp(line.c_str(), "ply") != 0) status = VKGS_ERR_FORMAT;
      continue;
    }
    if (nw <= 0) continue;
    const std::string word = w0;
    if (word == "end_header") {
      ok_end = true;
      break;
    } else if (word == "format") {
      if (nw < 2 || strcmp(w1, "binary_little_endian") != 0) status = VKGS_ERR_FORMAT;
      saw_format = true;
    } else if (word == "element") {
      if (nw < 3) {
        status = VKGS_ERR_FORMAT;
        break;
      }
      if (strcmp(w1, "vertex") == 0) {
        char* endp = nullptr;
        const unsigned long long n = strtoull(w2, &endp, 10);
        if (!endp || *endp != '\0') status = VKGS_ERR_FORMAT;
        info->point_count = n;
        in_vertex = true;
        saw_vertex = true;
      } else {
        if (!saw_vertex) status = VKGS_ERR_FORMAT;  // data of another element would precede the vertices
        in_vertex = false;
      }
    } else if (word == "property") {
      if (!in_vertex) continue;
      if (nw < 3 || strcmp(w1, "list") == 0) {
        status = VKGS_ERR_FORMAT;
        break;
      }
      const int sz = ply_type_size(w1);
      if (sz != 4) {  // rows are read as 4-byte words (splat_load_thread.cc:91)
        status = VKGS_ERR_FORMAT;
        break;
      }
      props[w2] = {(int)row_bytes, w1};
      row_bytes += (uint64_t)sz;
    }
  }
  if (status == VKGS_OK && (!ok_end || !saw_format || !saw_vertex || row_bytes == 0)) status = VKGS_ERR_FORMAT;
  if (status == VKGS_OK) {
    // offset table, splat_load_thread.cc:115-135
    auto need = [&](const std::string& name, uint32_t* dst) {
      auto it = props.find(name);
      if (it == props.end()) return false;
      const std::string& t = it->second.second;
      if (t != "float" && t != "float32") return false;
      *dst = (uint32_t)(it->second.first / 4);
      return true;
    };
    bool ok = need("x", &info->offsets[0]) && need("y", &info->offsets[1]) && need("z", &info->offsets[2]) &&
              need("scale_0", &info->offsets[3]) && need("scale_1", &info->offsets[4]) &&
              need("scale_2", &info->offsets[5]) && need("rot_1", &info->offsets[6]) &&
              need("rot_2", &info->offsets[7]) && need("rot_3", &info->offsets[8]) &&
              need("rot_0", &info->offsets[9]) && need("f_dc_0", &info->offsets[10 + 0]) &&
              need("f_dc_1", &info->offsets[10 + 16]) && need("f_dc_2", &info->offsets[10 + 32]) &&
              need("opacity", &info->offsets[58]);
    for (int i = 0; ok && i < 15; ++i) {
      ok = need("f_rest_" + std::to_string(i), &info->offsets[10 + 1 + i]) &&
           need("f_rest_" + std::to_string(15 + i), &info->offsets[10 + 17 + i]) &&
           need("f_rest_" + std::to_string(30 + i), &info->offsets[10 + 33 + i]);
    }
    if (!ok) status = VKGS_ERR_FORMAT;
    info->stride = (uint32_t)(row_bytes / 4);
    info->offsets[59] = info->stride;
    info->header_bytes = header_bytes;
  }
  if (status == VKGS_OK) {
    // truncation check
    if (fseek(f, 0, SEEK_END) != 0) status = VKGS_ERR_IO;
    const long size = ftell(f);
    const unsigned __int128 need_bytes =
        (unsigned __int128)info->header_bytes + (unsigned __int128)info->point_count * row_bytes;
    if (status == VKGS_OK && (size < 0 || (unsigned __int128)size < need_bytes)) status = VKGS_ERR_FORMAT;
  }
  fclose(f);
  return status;
}

uint64_t vkgs_ply_buffer_bytes(const vkgs_ply_info* info) {
  if (!info) return 0;
  return 60ull * 4ull + info->point_count * (uint64_t)info->stride * 4ull;
}

int vkgs_ply_read(const char* path, const vkgs_ply_info* info, void* dst, uint64_t dst_bytes) {
  if (!path || !info || !dst) return VKGS_ERR_INVALID_ARG;
  if (dst_bytes < vkgs_ply_buffer_bytes(info)) return VKGS_ERR_CAPACITY;
  FILE* f = fopen(path, "rb");
  if (!f) return VKGS_ERR_IO;
  memcpy(dst, info->offsets, 60 * 4);
  int status = VKGS_OK;
  if (fseek(f, (long)info->header_bytes, SEEK_SET) != 0) status = VKGS_ERR_IO;
  const uint64_t bytes = info->point_count * (uint64_t)info->stride * 4ull;
  char* p = static_cast<char*>(dst) + 240;
  uint64_t done = 0;
  while (status == VKGS_OK && done < bytes) {
    const uint64_t chunk = bytes - done < (64ull << 20) ? bytes - done : (64ull << 20);
    const size_t got = fread(p + done, 1, (size_t)chunk, f);
    if (got != chunk) status = VKGS_ERR_FORMAT;  // truncated
    done += got;
  }
  fclose(f);
  return status;
}

}  // extern "C"
