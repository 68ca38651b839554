// scene_xml_tests.cpp -- the scene-file front end (include/scx_scene_xml.hpp) as a small
// command-line driver in the manner of ball2d_cli / rigidbody3d_cli.
//   scene_xml_tests parse <scene.xml>              prints the parsed scene as JSON (host only)
//   scene_xml_tests run <scene.xml> <steps> <out>  steps it on the GPU through the C ABI and
//                                                  writes the final q, v (and omega) as raw doubles
// Exit status: 0 ok, 2 the scene was refused (message on stderr), 1 a device error.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "scx_scene_xml.hpp"

static void print_list(const char* key, const std::vector<double>& v, bool last = false) {
  std::printf("\"%s\": [", key);
  for (size_t k = 0; k < v.size(); ++k) std::printf("%s%.17g", k ? ", " : "", v[k]);
  std::printf("]%s\n", last ? "" : ",");
}

static int parse(const std::string& path) {
  const scx_scisim::SceneFile sc = scx_scisim::parseSceneFile(path);
  const scx_step_params p = sc.params();
  std::printf("{\n\"dim\": %d,\n", sc.dim);
  print_list("q", sc.q);
  print_list("v", sc.v);
  print_list("omega", sc.omega);
  print_list("r", sc.r);
  print_list("m", sc.m);
  print_list("inertia", sc.inertia);
  print_list("plane_x", sc.plane_x);
  print_list("plane_n", sc.plane_n);
  print_list("g", {sc.g[0], sc.g[1], sc.g[2]});
  std::printf("\"integrator\": \"%s\", \"dt_string\": \"%s\", \"dt\": %.17g, ", sc.integrator.c_str(),
              sc.dt_string.c_str(), sc.dt);
  if (std::isinf(sc.end_time))
    std::printf("\"end_time\": null,\n");
  else
    std::printf("\"end_time\": %.17g,\n", sc.end_time);
  std::printf("\"mu\": %.17g, \"CoR\": %.17g, \"tol\": %.17g, \"max_iters\": %u, \"eval_every\": %u,\n", sc.mu, sc.CoR,
              sc.tol, sc.max_iters, sc.eval_every);
  std::printf("\"staggering\": \"%s\", \"cache_impulses\": \"%s\", \"stabilization\": %d, "
              "\"penetration_threshold\": %.17g,\n",
              sc.staggering.c_str(), sc.cache_impulses.c_str(), sc.stabilization ? 1 : 0, sc.penetration_threshold);
  std::printf("\"params\": {\"integrator\": %d, \"staggering\": %d, \"stabilize\": %d, \"warm_start\": %d, "
              "\"precision\": %d}\n}\n",
              p.integrator, p.staggering, p.stabilize, p.warm_start, p.precision);
  return 0;
}

static int run(const std::string& path, int steps, const std::string& out) {
  const scx_scisim::SceneFile sc = scx_scisim::parseSceneFile(path);
  scx_config cfg{};
  cfg.dim = sc.dim;
  cfg.max_bodies = sc.nbodies();
  cfg.max_planes = sc.nplanes() > 0 ? sc.nplanes() : 1;
  scx_ctx* c = nullptr;
  auto check = [&](int rc, const char* what) {
    if (rc != SCX_OK) {
      std::cerr << what << ": " << (c ? scx_last_error(c) : "invalid configuration") << std::endl;
      if (c) scx_destroy(c);
      std::exit(EXIT_FAILURE);
    }
  };
  check(scx_create(&cfg, &c), "scx_create");
  check(scx_set_planes(c, sc.nplanes(), sc.plane_x.data(), sc.plane_n.data()), "scx_set_planes");
  check(scx_set_bodies(c, sc.nbodies(), sc.q.data(), sc.v.data(), sc.dim == 3 ? sc.omega.data() : nullptr,
                       sc.r.data(), sc.m.data(), sc.dim == 3 ? sc.inertia.data() : nullptr),
        "scx_set_bodies");
  unsigned failures = 0, most = 0;
  for (int it = 1; it <= steps; ++it) {
    const scx_step_params p = sc.params((uint32_t)it);
    scx_step_stats st{};
    check(scx_step(c, &p, &st), "scx_step");
    if (!st.succeeded) ++failures;  // the reference warns and continues
    if (st.sweeps > most) most = st.sweeps;
  }
  std::vector<double> q(sc.q.size()), v(sc.v.size()), w(sc.omega.size());
  check(scx_get_state(c, q.data(), v.data(), sc.dim == 3 ? w.data() : nullptr), "scx_get_state");
  scx_destroy(c);
  std::ofstream f(out, std::ios::binary);
  f.write(reinterpret_cast<const char*>(q.data()), (std::streamsize)(sizeof(double) * q.size()));
  f.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(sizeof(double) * v.size()));
  f.write(reinterpret_cast<const char*>(w.data()), (std::streamsize)(sizeof(double) * w.size()));
  std::printf("steps %d bodies %u planes %u most sweeps %u unconverged steps %u\n", steps, sc.nbodies(), sc.nplanes(),
              most, failures);
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::cerr << "Usage: " << argv[0] << " parse <scene.xml> | run <scene.xml> <steps> <out.bin>" << std::endl;
    return EXIT_FAILURE;
  }
  const std::string mode = argv[1];
  try {
    if (mode == "parse") return parse(argv[2]);
    if (mode == "run" && argc == 5) return run(argv[2], std::atoi(argv[3]), argv[4]);
  } catch (const std::exception& e) {
    std::cerr << "scene refused: " << e.what() << std::endl;
    return 2;
  }
  std::cerr << "bad arguments" << std::endl;
  return EXIT_FAILURE;
}
