// PETSc-style options files as the reference reads them (PetscOptionsGetInt /
// HasName / GetString with -options_file; see fmm3d_mpi/options_*): tokens are
// whitespace separated, '%' and '#' start a comment that runs to the end of the
// line, an option is a token starting with '-' and takes the next token as its
// value unless that token is itself an option.  Names compare
// case-insensitively (the code asks for -ptsMax, the files say -ptsmax).
#pragma once
#include <map>
#include <string>

namespace kifmm {

class Options {
 public:
  int load(const std::string& path, std::string* err);
  bool has(const std::string& name) const;
  bool get(const std::string& name, std::string* value) const;
  bool get_int(const std::string& name, long* value) const;
  bool get_double(const std::string& name, double* value) const;
 private:
  std::map<std::string, std::string> kv_;
};

}  // namespace kifmm
