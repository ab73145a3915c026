// Parameter file, binary and text I/O helpers of the modeling driver.
//
// Same contract as SeisFAT2D's admin module (src/admin/admin.hpp:17-33):
// key = value parameter files with '#' comments, raw float32 binaries and
// "x, z" text geometry files.  Differences, all towards failing loudly:
// empty lines are skipped instead of dereferenced, a missing key throws
// instead of returning "", and a short binary file throws instead of leaving
// the tail of the buffer unread.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

namespace seisfat {

// Value of the first line whose text starts with `target` (prefix match, as
// src/admin/admin.cpp:82); the value starts two characters after '=', ends at
// '#', and has every space removed (src/admin/admin.cpp:84-97).  Lines that
// start with '#' or ' ' are skipped.  Throws std::invalid_argument when the
// file cannot be opened or holds no such key.
std::string catch_parameter(const std::string& target, const std::string& file);

// Same, returning `fallback` when the key is absent.
std::string catch_parameter_or(const std::string& target, const std::string& file, const std::string& fallback);

// "true"/"false" in any case (src/admin/admin.cpp:15-23).
bool str2bool(std::string s);

// Reads exactly n float32 values (throws on a missing or short file).
void import_binary_float(const std::string& path, float* array, size_t n);
void export_binary_float(const std::string& path, const float* array, size_t n);

// Lines of a text file that do not start with '#' (src/admin/admin.cpp:51-63);
// empty lines are skipped.
void import_text_file(const std::string& path, std::vector<std::string>& elements);

std::vector<std::string> split(const std::string& s, char delimiter);

// Catmull-Rom interpolation (src/admin/admin.cpp:121-134).
float cubic1d(const float P[4], float dx);
float cubic2d(const float P[4][4], float dx, float dy);

}  // namespace seisfat
