// dgalerkin configuration file: the reference's key=value format and semantics
// (src/configParser.cpp:30-150, include/configParser.h:7-41), plus keys under
// the "gpu." prefix and "receiver*" that the reference never reads.
#pragma once
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace dgh {

struct MshModel;

struct Config {
  // defaults as in include/configParser.h:9-40
  double timeStart = 0, timeEnd = 1, timeStep = 0.1, timeRate = 0.1;
  std::string elementType = "Lagrange";
  std::string timeIntMethod = "Euler1";
  std::map<int, std::pair<std::string, double>> physBCs;  // physical tag -> (type, value), ascending tag
  double v0[3] = {0, 0, 0};
  double rho0 = 1, c0 = 1;
  int numThreads = 1;
  std::vector<std::vector<double>> sources;         // {pole, x, y, z, radius, amp, freq, phase, duration}
  std::vector<std::vector<double>> initConditions;  // {0, x, y, z, size, amp}
  std::string saveFile = "results.msh";
  // keys the reference does not read
  std::string precision = "fp64";        // gpu.precision = fp64 | fp32
  std::string fluxSign = "reference";    // gpu.fluxSign = reference | stable
  int device = 0;                        // gpu.device
  std::vector<std::vector<double>> receivers;  // receiverN = x,y,z
  std::string receiverFile;              // gpu.receiverFile: CSV of recorded pressure
  std::map<std::string, std::string> raw;  // whitespace-stripped key -> value
};

// Parses the file with the reference's rules: all whitespace removed from each
// line, lines starting with '#' skipped, first '=' splits key and value,
// later duplicates win. Throws std::runtime_error when the file cannot be read
// or a required numeric key is missing or malformed (the reference terminates
// the process in both cases). Does not touch the filesystem besides reading.
Config parse_config_text(const std::string& text);
Config parse_config(const std::string& path);

// Resolve boundary conditions against the mesh's physical groups of dimension
// model_dim-1 (src/configParser.cpp:117-132): a group is honoured when the
// config value for its name starts with "Absorbing" or "Reflecting".
void resolve_bcs(Config& c, const MshModel& m);

}  // namespace dgh
