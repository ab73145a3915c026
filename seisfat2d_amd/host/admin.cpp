#include "admin.hpp"

#include <algorithm>
#include <cctype>
#include <fstream>
#include <iostream>
#include <sstream>

namespace seisfat {

namespace {

bool find_parameter(const std::string& target, const std::string& file, std::string& value) {
  std::ifstream in(file);
  if (!in.is_open()) throw std::invalid_argument("Error: " + file + " could not be opened!");
  std::string line;
  while (std::getline(in, line)) {
    if (line.empty() || line.front() == '#' || line.front() == ' ') continue;
    if (line.find(target) != 0) continue;
    const size_t eq = line.find('=');
    std::string v;
    if (eq != std::string::npos) {
      for (size_t i = eq + 2; i < line.size(); ++i) {
        if (line[i] == '#') break;
        v += line[i];
      }
    }
    v.erase(std::remove(v.begin(), v.end(), ' '), v.end());
    v.erase(std::remove(v.begin(), v.end(), '\r'), v.end());
    value = v;
    return true;
  }
  return false;
}

}  // namespace

std::string catch_parameter(const std::string& target, const std::string& file) {
  std::string v;
  if (!find_parameter(target, file, v)) throw std::invalid_argument("Error: parameter " + target + " not found in " + file);
  return v;
}

std::string catch_parameter_or(const std::string& target, const std::string& file, const std::string& fallback) {
  std::string v;
  return find_parameter(target, file, v) ? v : fallback;
}

bool str2bool(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return (char)std::tolower(c); });
  if (s == "true") return true;
  if (s == "false") return false;
  throw std::invalid_argument("Error: '" + s + "' is not a boolean");
}

void import_binary_float(const std::string& path, float* array, size_t n) {
  std::ifstream file(path, std::ios::in | std::ios::binary);
  if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
  file.read(reinterpret_cast<char*>(array), (std::streamsize)(n * sizeof(float)));
  if ((size_t)file.gcount() != n * sizeof(float))
    throw std::invalid_argument("Error: " + path + " holds fewer than " + std::to_string(n) + " float32 values");
}

void export_binary_float(const std::string& path, const float* array, size_t n) {
  std::ofstream file(path, std::ios::out | std::ios::binary);
  if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
  file.write(reinterpret_cast<const char*>(array), (std::streamsize)(n * sizeof(float)));
  if (!file) throw std::invalid_argument("Error: writing " + path + " failed");
}

void import_text_file(const std::string& path, std::vector<std::string>& elements) {
  std::ifstream file(path, std::ios::in);
  if (!file.is_open()) throw std::invalid_argument("Error: " + path + " could not be opened!");
  std::string line;
  while (std::getline(file, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    elements.push_back(line);
  }
}

std::vector<std::string> split(const std::string& s, char delimiter) {
  std::vector<std::string> tokens;
  std::string token;
  std::istringstream ts(s);
  while (std::getline(ts, token, delimiter)) tokens.push_back(token);
  return tokens;
}

float cubic1d(const float P[4], float dx) {
  return P[1] + 0.5f * dx *
                    (P[2] - P[0] +
                     dx * (2.0f * P[0] - 5.0f * P[1] + 4.0f * P[2] - P[3] + dx * (3.0f * (P[1] - P[2]) + P[3] - P[0])));
}

float cubic2d(const float P[4][4], float dx, float dy) {
  float p[4];
  for (int k = 0; k < 4; ++k) p[k] = cubic1d(P[k], dy);
  return cubic1d(p, dx);
}

}  // namespace seisfat
