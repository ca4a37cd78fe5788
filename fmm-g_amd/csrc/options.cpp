#include "options.hpp"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <vector>

namespace kifmm {

static std::string lower(std::string s) {
  for (auto& c : s) c = char(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

// A token is an option name if it starts with '-' followed by a letter or '_'
// (so "-1" or "-.5" are values).
static bool is_name(const std::string& t) {
  return t.size() > 1 && t[0] == '-' && (std::isalpha(static_cast<unsigned char>(t[1])) || t[1] == '_');
}

int Options::load(const std::string& path, std::string* err) {
  std::ifstream in(path);
  if (!in) {
    if (err) *err = "cannot open options file " + path;
    return -4;
  }
  std::vector<std::string> toks;
  std::string line;
  while (std::getline(in, line)) {
    size_t cut = line.find_first_of("%#");
    if (cut != std::string::npos) line.resize(cut);
    std::istringstream ls(line);
    std::string tok;
    while (ls >> tok) toks.push_back(tok);
  }
  for (size_t i = 0; i < toks.size(); i++) {
    if (!is_name(toks[i])) continue;  // stray value: ignored as PETSc does
    std::string name = lower(toks[i]);
    std::string value;
    if (i + 1 < toks.size() && !is_name(toks[i + 1])) value = toks[++i];
    kv_[name] = value;
  }
  return 0;
}

bool Options::has(const std::string& name) const { return kv_.count(lower(name)) > 0; }

bool Options::get(const std::string& name, std::string* value) const {
  auto it = kv_.find(lower(name));
  if (it == kv_.end()) return false;
  if (value) *value = it->second;
  return true;
}

bool Options::get_int(const std::string& name, long* value) const {
  std::string s;
  if (!get(name, &s) || s.empty()) return false;
  char* end = nullptr;
  long v = std::strtol(s.c_str(), &end, 10);
  if (end == s.c_str() || *end) return false;
  *value = v;
  return true;
}

bool Options::get_double(const std::string& name, double* value) const {
  std::string s;
  if (!get(name, &s) || s.empty()) return false;
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end == s.c_str() || *end) return false;
  *value = v;
  return true;
}

}  // namespace kifmm
