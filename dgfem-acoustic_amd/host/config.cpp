#include "config.hpp"

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "meshio.hpp"

namespace dgh {

namespace {

std::vector<std::string> split(const std::string& s, char delim) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string tok;
  while (std::getline(ss, tok, delim)) out.push_back(tok);
  return out;
}

double to_double(const std::map<std::string, std::string>& m, const std::string& key) {
  auto it = m.find(key);
  std::string v = it == m.end() ? "" : it->second;
  try {
    size_t used = 0;
    double d = std::stod(v, &used);
    return d;
  } catch (...) {
    throw std::runtime_error("config: required numeric key '" + key + "' is missing or not a number ('" + v + "')");
  }
}

double field(const std::vector<std::string>& sep, size_t i, const std::string& key) {
  if (i >= sep.size()) throw std::runtime_error("config: key '" + key + "' has too few comma-separated fields");
  try {
    return std::stod(sep[i]);
  } catch (...) {
    throw std::runtime_error("config: key '" + key + "' field " + std::to_string(i) + " is not a number");
  }
}

}  // namespace

Config parse_config_text(const std::string& text) {
  Config c;
  std::map<std::string, std::string> cm;
  std::stringstream in(text);
  std::string line;
  while (std::getline(in, line)) {
    std::string l;
    for (char ch : line)
      if (!isspace((unsigned char)ch)) l.push_back(ch);
    if (l.empty() || l[0] == '#') continue;
    size_t eq = l.find('=');
    std::string name = l.substr(0, eq);
    std::string value = eq == std::string::npos ? l : l.substr(eq + 1);
    cm[name] = value;
  }
  c.raw = cm;
  c.timeStart = to_double(cm, "timeStart");
  c.timeEnd = to_double(cm, "timeEnd");
  c.timeStep = to_double(cm, "timeStep");
  c.timeRate = to_double(cm, "timeRate");
  c.elementType = cm.count("elementType") ? cm["elementType"] : "";
  c.timeIntMethod = cm.count("timeIntMethod") ? cm["timeIntMethod"] : "";
  c.saveFile = cm.count("saveFile") ? cm["saveFile"] : "";
  {
    auto it = cm.find("numThreads");
    std::string v = it == cm.end() ? "" : it->second;
    try { c.numThreads = std::stoi(v); } catch (...) {
      throw std::runtime_error("config: required numeric key 'numThreads' is missing or not a number ('" + v + "')");
    }
    if (c.numThreads == 1) c.numThreads = 0;  // src/configParser.cpp:58
  }
  c.v0[0] = to_double(cm, "v0_x");
  c.v0[1] = to_double(cm, "v0_y");
  c.v0[2] = to_double(cm, "v0_z");
  c.rho0 = to_double(cm, "rho0");
  c.c0 = to_double(cm, "c0");
  const double pi = M_PI;
  for (auto& kv : cm) {  // ascending key order, as the reference's std::map walk
    const std::string& key = kv.first;
    if (key.find("source") == 0) {
      auto sep = split(kv.second, ',');
      double x = field(sep, 1, key), y = field(sep, 2, key), z = field(sep, 3, key), size = field(sep, 4, key);
      double amp = field(sep, 5, key), freq = field(sep, 6, key), phase = field(sep, 7, key), dur = field(sep, 8, key);
      if (sep[0] == "dipole") {
        c.sources.push_back({1, x - size, y, z, size / 2., amp, freq, phase, dur});
        c.sources.push_back({1, x + size, y, z, size / 2., amp, freq, phase + pi, dur});
      } else if (sep[0] == "quadrupole") {
        c.sources.push_back({2, x - size, y, z, size / 2., amp, freq, phase, dur});
        c.sources.push_back({2, x + size, y, z, size / 2., amp, freq, phase, dur});
        c.sources.push_back({2, x, y - size, z, size / 2., amp, freq, phase + pi, dur});
        c.sources.push_back({2, x, y + size, z, size / 2., amp, freq, phase + pi, dur});
      } else {
        c.sources.push_back({0, x, y, z, size, amp, freq, phase, dur});
      }
    } else if (key.find("initialCondtition") == 0) {
      auto sep = split(kv.second, ',');
      c.initConditions.push_back({0, field(sep, 1, key), field(sep, 2, key), field(sep, 3, key), field(sep, 4, key),
                                  field(sep, 5, key)});
    } else if (key.find("receiver") == 0) {
      auto sep = split(kv.second, ',');
      c.receivers.push_back({field(sep, 0, key), field(sep, 1, key), field(sep, 2, key)});
    }
  }
  if (cm.count("gpu.precision")) c.precision = cm["gpu.precision"];
  if (cm.count("gpu.fluxSign")) c.fluxSign = cm["gpu.fluxSign"];
  if (cm.count("gpu.device")) c.device = std::stoi(cm["gpu.device"]);
  if (cm.count("gpu.receiverFile")) c.receiverFile = cm["gpu.receiverFile"];
  if (c.precision != "fp64" && c.precision != "fp32")
    throw std::runtime_error("config: gpu.precision must be fp64 or fp32");
  if (c.fluxSign != "reference" && c.fluxSign != "stable")
    throw std::runtime_error("config: gpu.fluxSign must be reference or stable");
  return c;
}

Config parse_config(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("config: cannot open '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  return parse_config_text(ss.str());
}

void resolve_bcs(Config& c, const MshModel& m) {
  int bcdim = model_dimension(m) - 1;
  for (auto& g : physical_groups(m, bcdim)) {
    auto it = c.raw.find(g.second);
    std::string v = it == c.raw.end() ? "" : it->second;
    if (v.find("Absorbing") == 0) c.physBCs[g.first] = {"Absorbing", 0};
    else if (v.find("Reflecting") == 0) c.physBCs[g.first] = {"Reflecting", 0};
  }
}

}  // namespace dgh
