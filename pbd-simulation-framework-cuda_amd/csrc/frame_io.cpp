// frame_io.cpp - headless frame output: the arrays the reference's renderer
// draws, written to files instead of a GL surface.
//
// The reference draws particle i from texel (i % width, i / width) of the
// positions4 / colors4 textures (textureWidth x textureWidth;
// src/app/rendering/Particles.cc:266-274, shader/particles/
// particleSpriteGeom.vert:35-41) and shades it
// mix((0.95, 0.97, 0.97), (0, 0, 1), colors4.x^(1/6)).
//
// Formats (PBD_OUT_*):
//   RAW  32-byte header {"PBDF", version 1, n, texture width, frame, 0, 0, 0}
//        (u32 little-endian), then n float4 positions4, then n float4 colors4,
//        in buffer index order: texel (i % width, i / width) holds entry i.
//   PLY  ASCII or binary little-endian PLY, one vertex per particle in buffer
//        index order: x y z (float) and the renderer's shade as red green blue
//        (uchar).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "pbd/pbd.h"

namespace pbd {

namespace {

struct File {
  FILE* f = nullptr;
  explicit File(const char* path) : f(std::fopen(path, "wb")) {}
  ~File() {
    if (f) std::fclose(f);
  }
  bool write(const void* p, size_t bytes) { return bytes == 0 || std::fwrite(p, 1, bytes, f) == bytes; }
  bool close() {
    const bool ok = std::fclose(f) == 0;
    f = nullptr;
    return ok;
  }
};

// The renderer's shade of a particle from its colour (particleSpriteGeom.vert:40).
void shade(const float* c, unsigned char rgb[3]) {
  const float x = c[0] > 0.0f ? c[0] : 0.0f;
  const float t = std::pow(x, 1.0f / 6.0f);
  const float base[3] = {0.95f, 0.97f, 0.97f}, blue[3] = {0.0f, 0.0f, 1.0f};
  for (int k = 0; k < 3; ++k) {
    float v = base[k] + (blue[k] - base[k]) * t;
    v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    rgb[k] = (unsigned char)std::lround(v * 255.0f);
  }
}

}  // namespace

int write_frame_arrays(const char* path, int format, uint32_t frame, uint32_t n, uint32_t width,
                       const float* pos4, const float* col4, std::string* err) {
  if (!path || (n && !pos4)) {
    if (err) *err = "write_frame: path and positions are required";
    return PBD_ERR_INVALID;
  }
  if (format != PBD_OUT_RAW && format != PBD_OUT_PLY_ASCII && format != PBD_OUT_PLY_BINARY) {
    if (err) *err = "write_frame: unknown format";
    return PBD_ERR_INVALID;
  }
  std::vector<float> white;
  if (!col4) {
    white.assign((size_t)n * 4, 1.0f);
    col4 = white.data();
  }
  File out(path);
  if (!out.f) {
    if (err) *err = std::string("write_frame: cannot open ") + path;
    return PBD_ERR_IO;
  }
  bool ok = true;
  if (format == PBD_OUT_RAW) {
    const uint32_t hdr[8] = {0x46444250u /* "PBDF" */, 1u, n, width, frame, 0u, 0u, 0u};
    ok = out.write(hdr, sizeof(hdr)) && out.write(pos4, (size_t)n * 16) && out.write(col4, (size_t)n * 16);
  } else {
    const bool binary = format == PBD_OUT_PLY_BINARY;
    char head[512];
    const int len = std::snprintf(head, sizeof(head),
                                  "ply\nformat %s 1.0\ncomment pbd frame %u, texture width %u, vertex i = texel "
                                  "(i %% width, i / width)\nelement vertex %u\nproperty float x\nproperty float y\n"
                                  "property float z\nproperty uchar red\nproperty uchar green\nproperty uchar blue\n"
                                  "end_header\n",
                                  binary ? "binary_little_endian" : "ascii", frame, width, n);
    ok = out.write(head, (size_t)len);
    std::vector<unsigned char> buf;
    buf.reserve(binary ? (size_t)n * 15 : 0);
    char line[96];
    for (uint32_t i = 0; i < n && ok; ++i) {
      unsigned char rgb[3];
      shade(col4 + 4 * (size_t)i, rgb);
      const float* p = pos4 + 4 * (size_t)i;
      if (binary) {
        const unsigned char* b = reinterpret_cast<const unsigned char*>(p);
        buf.insert(buf.end(), b, b + 12);
        buf.insert(buf.end(), rgb, rgb + 3);
      } else {
        const int l = std::snprintf(line, sizeof(line), "%.9g %.9g %.9g %u %u %u\n", p[0], p[1], p[2], rgb[0], rgb[1],
                                    rgb[2]);
        ok = out.write(line, (size_t)l);
      }
    }
    if (ok && binary) ok = out.write(buf.data(), buf.size());
  }
  ok = out.close() && ok;
  if (!ok) {
    if (err) *err = std::string("write_frame: write to ") + path + " failed";
    return PBD_ERR_IO;
  }
  return PBD_OK;
}

}  // namespace pbd
