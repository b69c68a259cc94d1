// pbd_run.cpp - headless driver: the reference's engine loop without the
// window.  Engine::run (src/app/Engine.cu:66-72) calls Simulation::step and
// then renders; here each frame is one Simulation::step (substeps computes,
// replayed as a captured graph) and every k-th frame is written to a file
// instead of drawn (pbd_write_frame).
//
//   pbd_run <config.txt> <frames> [--out DIR] [--every K] [--format raw|ply|plyb]
//
// Prints the particle count, then one line per written frame and the
// throughput at the end.  Uses only the C-ABI (include/pbd/pbd.h).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pbd/pbd.h"

static int usage() {
  std::fprintf(stderr, "usage: pbd_run <config.txt> <frames> [--out DIR] [--every K] [--format raw|ply|plyb]\n");
  return 2;
}

int main(int argc, char** argv) {
  if (argc < 3) return usage();
  const char* config = argv[1];
  const long frames = std::strtol(argv[2], nullptr, 10);
  std::string out;
  long every = 1;
  int format = PBD_OUT_RAW;
  for (int i = 3; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--out") && i + 1 < argc) {
      out = argv[++i];
    } else if (!std::strcmp(argv[i], "--every") && i + 1 < argc) {
      every = std::strtol(argv[++i], nullptr, 10);
    } else if (!std::strcmp(argv[i], "--format") && i + 1 < argc) {
      const char* f = argv[++i];
      if (!std::strcmp(f, "raw")) format = PBD_OUT_RAW;
      else if (!std::strcmp(f, "ply")) format = PBD_OUT_PLY_ASCII;
      else if (!std::strcmp(f, "plyb")) format = PBD_OUT_PLY_BINARY;
      else return usage();
    } else {
      return usage();
    }
  }
  if (frames < 0 || every < 1) return usage();

  std::vector<char> params(pbd_params_size());
  pbd_params* p = reinterpret_cast<pbd_params*>(params.data());
  char err[512];
  if (pbd_params_load(config, p, err, sizeof err) != PBD_OK) {
    std::fprintf(stderr, "pbd_run: %s\n", err);
    return 1;
  }
  pbd_ctx* ctx = nullptr;
  if (pbd_create(p, 0, &ctx) != PBD_OK) {
    std::fprintf(stderr, "pbd_run: %s\n", pbd_last_error(nullptr));
    return 1;
  }
  uint32_t n = 0;
  pbd_num_particles(ctx, &n);
  std::printf("pbd_run: %u particles, %ld frames\n", n, frames);
  const char* ext = format == PBD_OUT_RAW ? "pbdf" : "ply";
  const auto t0 = std::chrono::steady_clock::now();
  int rc = PBD_OK;
  for (long f = 1; f <= frames && rc == PBD_OK; ++f) {
    rc = pbd_step_graph(ctx, 1);
    if (rc == PBD_OK && !out.empty() && f % every == 0) {
      char path[4096];
      std::snprintf(path, sizeof path, "%s/frame_%06ld.%s", out.c_str(), f, ext);
      rc = pbd_write_frame(ctx, path, format, (uint32_t)f);
      if (rc == PBD_OK) std::printf("frame %ld -> %s\n", f, path);
    }
  }
  if (rc == PBD_OK) rc = pbd_sync(ctx);
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (rc != PBD_OK) {
    std::fprintf(stderr, "pbd_run: %s\n", pbd_last_error(ctx));
    pbd_destroy(ctx);
    return 1;
  }
  uint32_t sub = 1;
  {
    char v[64];
    if (pbd_params_get(p, "Application.Simulation.Forces.substeps", v, sizeof v) == PBD_OK) sub = (uint32_t)std::atoi(v);
  }
  std::printf("pbd_run: %.3f s, %.3e particle-steps/s\n", s, s > 0 ? (double)n * frames * sub / s : 0.0);
  pbd_destroy(ctx);
  return 0;
}
