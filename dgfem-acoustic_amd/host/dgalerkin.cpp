// dgalerkin mesh.msh config.conf
//
// Host driver with the reference's command line and run semantics
// (src/dgalerkin.cpp:11-62, src/solver.cpp:171-292): reads the Gmsh mesh and the
// config, sets the Gaussian initial conditions, then runs the time loop whose steps
// are executed by libdgk on the GPU. Output views (Pressure, Density, Velocity as
// ElementNodeData) are written to saveFile at the steps where the reference saves
// them: step 0 and whenever the accumulated display time reaches timeRate, before the
// step is taken.
#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../include/dgk.h"
#include "../../include/dgk_host.h"
#include "config.hpp"
#include "meshio.hpp"

namespace {

void log(const std::string& s) { printf("Info    : %s\n", s.c_str()); fflush(stdout); }

int die(const std::string& s) {
  fprintf(stderr, "Error   : %s\n", s.c_str());
  return EXIT_FAILURE;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return E2BIG;  // src/dgalerkin.cpp:20
  const std::string msh_name = argv[1], config_name = argv[2];
  dgh::MshModel model;
  dgh::Config cfg;
  try {
    model = dgh::read_msh(msh_name);
    cfg = dgh::parse_config(config_name);
    dgh::resolve_bcs(cfg, model);
  } catch (std::exception& e) {
    return die(e.what());
  }
  std::remove(cfg.saveFile.c_str());  // src/configParser.cpp:115
  log("Config loaded : " + config_name);
  dgh::TopMesh top;
  try {
    top = dgh::extract_top(model);
  } catch (std::exception& e) {
    return die(e.what());
  }
  const int64_t K = top.K, Np = top.Np, KN = K * Np;
  log("==================================================");
  log("Number of Elements : " + std::to_string(K));
  log("Element dimension : " + std::to_string(top.dim));
  log("Element Order : " + std::to_string(top.order));
  log("Element Nbr Nodes : " + std::to_string(Np));
  log("Time step : " + std::to_string(cfg.timeStep));
  log("Final time : " + std::to_string(cfg.timeEnd));
  log("Mean flow velocity : (" + std::to_string(cfg.v0[0]) + "," + std::to_string(cfg.v0[1]) + "," +
      std::to_string(cfg.v0[2]) + ")");
  log("Mean density : " + std::to_string(cfg.rho0));
  log("Speed of sound : " + std::to_string(cfg.c0));
  log("Solver : " + cfg.timeIntMethod);

  // boundary code per element node (the face rule reads it at the face's first node)
  std::vector<uint8_t> node_bc(KN, 0);
  {
    int bcdim = dgh::model_dimension(model) - 1;
    for (auto& kv : cfg.physBCs) {
      auto nodes = dgh::nodes_for_physical_group(model, bcdim, kv.first);
      uint8_t v = kv.second.first == "Reflecting" ? 1 : 0;
      for (int64_t i = 0; i < KN; ++i)
        if (std::binary_search(nodes.begin(), nodes.end(), top.node_tags[i])) node_bc[i] = v;
    }
  }
  // initial condition p += amp*exp(-r^2/size) (src/dgalerkin.cpp:36-50)
  std::vector<double> u(4 * KN, 0.0);
  for (auto& ic : cfg.initConditions)
    for (int64_t n = 0; n < KN; ++n) {
      const double* X = &top.node_xyz[3 * n];
      u[n] += ic[5] * std::exp(-((X[0] - ic[1]) * (X[0] - ic[1]) + (X[1] - ic[2]) * (X[1] - ic[2]) +
                                 (X[2] - ic[3]) * (X[2] - ic[3])) / ic[4]);
    }
  if (cfg.timeIntMethod != "Runge-Kutta" && cfg.timeIntMethod != "Euler1") {
    log("Unknown timeIntMethod '" + cfg.timeIntMethod + "': nothing to integrate");  // src/dgalerkin.cpp:55-58
    return EXIT_SUCCESS;
  }
  dgk_mesh m{top.dim, top.order, K, top.node_xyz.data(), top.node_tags.data(), node_bc.data()};
  dgk_phys ph{cfg.c0, cfg.rho0, {cfg.v0[0], cfg.v0[1], cfg.v0[2]}};
  dgk_opts op{};
  op.precision = cfg.precision == "fp32" ? DGK_FP32 : DGK_FP64;
  op.flux_sign = cfg.fluxSign == "stable" ? DGK_SIGN_STABLE : DGK_SIGN_REFERENCE;
  op.device = cfg.device;
  // launch settings tuned for this machine: $DGK_TUNING, else tuning.txt next to the executable
  {
    std::string tpath;
    if (const char* e = std::getenv("DGK_TUNING")) tpath = e;
    else {
      char buf[4096];
      const ssize_t len = readlink("/proc/self/exe", buf, sizeof(buf) - 1);
      if (len > 0) {
        std::string exe(buf, (size_t)len);
        tpath = exe.substr(0, exe.find_last_of('/') + 1) + "tuning.txt";
      }
    }
    if (!tpath.empty() && access(tpath.c_str(), R_OK) == 0 && dgk_load_tuning(tpath.c_str()) != DGK_OK)
      log("ignoring tuning table " + tpath + ": " + dgk_last_error(nullptr));
  }
  dgk_ctx* ctx = nullptr;
  if (dgk_create(&m, &ph, &op, &ctx) != DGK_OK) return die(std::string("libdgk: ") + dgk_last_error(nullptr));
  auto check = [&](dgk_status st) {
    if (st != DGK_OK) {
      fprintf(stderr, "Error   : libdgk: %s\n", dgk_last_error(ctx));
      dgk_destroy(ctx);
      exit(EXIT_FAILURE);
    }
  };
  dgk_info info{};
  check(dgk_info_get(ctx, &info));
  log("Number of Faces : " + std::to_string(info.num_faces) + " (" + std::to_string(info.num_boundary_faces) +
      " on the boundary)");
  check(dgk_set_state(ctx, u.data()));
  if (!cfg.sources.empty()) {
    std::vector<double> src;
    for (auto& s : cfg.sources) src.insert(src.end(), s.begin(), s.end());
    check(dgk_set_sources(ctx, (int)cfg.sources.size(), src.data(), top.node_xyz.data()));
  }
  // receivers: pressure interpolated at each point after every step, recorded on the device
  const int nrec = (int)cfg.receivers.size();
  long total_steps = 0;
  for (double tt = cfg.timeStart; tt <= cfg.timeEnd; tt += cfg.timeStep) ++total_steps;  // the loop below
  if (nrec) {
    std::vector<double> rxyz;
    for (auto& r : cfg.receivers) rxyz.insert(rxyz.end(), r.begin(), r.begin() + 3);
    check(dgk_set_receivers(ctx, nrec, rxyz.data(), top.node_xyz.data(), (int)std::max(1L, total_steps)));
  }
  dgh::ViewWriter views;
  views.path = cfg.saveFile;
  bool write_views = !cfg.saveFile.empty();
  if (write_views) {
    try { views.begin(model); } catch (std::exception& e) { return die(e.what()); }
  }
  const bool rk = cfg.timeIntMethod == "Runge-Kutta";
  auto start = std::chrono::system_clock::now();
  std::vector<double> pending;  // start times of steps not yet run
  std::vector<double> g_p(KN), g_rho(KN), g_v(3 * KN);
  auto flush = [&]() {
    if (pending.empty()) return;
    check(rk ? dgk_rk4(ctx, (int)pending.size(), cfg.timeStep, pending.data())
             : dgk_euler(ctx, (int)pending.size(), cfg.timeStep, pending.data()));
    pending.clear();
  };
  double t = cfg.timeStart, tDisplay = 0;
  long step = 0;
  for (; t <= cfg.timeEnd; t += cfg.timeStep, tDisplay += cfg.timeStep, ++step) {  // src/solver.cpp:216-217
    if (tDisplay >= cfg.timeRate || step == 0) {
      tDisplay = 0;
      flush();
      check(dgk_get_state(ctx, u.data()));
      if (write_views) {
        for (int64_t i = 0; i < KN; ++i) {
          g_p[i] = u[i];
          g_rho[i] = u[i] / (cfg.c0 * cfg.c0);
          for (int c = 0; c < 3; ++c) g_v[3 * i + c] = u[(1 + c) * KN + i];
        }
        views.add_step("Pressure", (int)step, t, 1, top.elem_tags, (int)Np, g_p.data());
        views.add_step("Density", (int)step, t, 1, top.elem_tags, (int)Np, g_rho.data());
        views.add_step("Velocity", (int)step, t, 3, top.elem_tags, (int)Np, g_v.data());
      }
      auto el = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - start);
      log("[" + std::to_string(t) + "/" + std::to_string(cfg.timeEnd) + "s] Step number : " + std::to_string(step) +
          ", Elapsed time: " + std::to_string(el.count()) + "s");
    }
    pending.push_back(t);
  }
  flush();
  check(dgk_sync(ctx));
  if (dgk_check_finite(ctx) != DGK_OK) log("warning: the final state holds NaN or Inf");
  if (nrec) {
    int rows = 0;
    check(dgk_get_receivers(ctx, nullptr, 0, &rows));
    std::vector<double> series((size_t)rows * nrec);
    check(dgk_get_receivers(ctx, series.data(), rows, &rows));
    std::string path = cfg.receiverFile.empty() ? cfg.saveFile + ".receivers.csv" : cfg.receiverFile;
    FILE* f = fopen(path.c_str(), "w");
    if (f) {
      fprintf(f, "step");
      for (int r = 0; r < nrec; ++r) fprintf(f, ",p%d", r);
      fprintf(f, "\n");
      for (int i = 0; i < rows; ++i) {
        fprintf(f, "%d", i + 1);
        for (int r = 0; r < nrec; ++r) fprintf(f, ",%.17g", series[(size_t)i * nrec + r]);
        fprintf(f, "\n");
      }
      fclose(f);
    }
  }
  dgk_destroy(ctx);
  return EXIT_SUCCESS;
}
