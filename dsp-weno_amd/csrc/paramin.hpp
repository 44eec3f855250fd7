// paramin.hpp — the reference's input contract on the host: the param.in
// reader (tokens, comments and sections of enflo/src/reader.h:49-217), an
// expression evaluator for the initial/exact-solution formulas (the subset of
// fparser 4.5.1 syntax the examples use: + - * / % ^, unary -, !, comparisons
// = != < <= > >= yielding 1/0, & and |, parentheses, functions, named
// constants, variables x y z t), and Parameter::read (enflo/src/parameter.cc,
// enflo_scalar/src_scalar/parameter.cc) into a tecno_cfg per mesh level.
#pragma once
#include <cmath>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tecno.h"

namespace tecno {
namespace cli {

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Reader: whitespace, ',' and ';' separate tokens; '#', '%', '//' comment to
// the end of the line, '/* */' block comments
class Reader {
 public:
  explicit Reader(const std::string& path) : f_(path) {
    if (!f_) throw InputError("cannot open " + path);
  }
  std::string token();
  double number();
  int integer();
  std::string line();  // rest of the current line after leading blanks
  bool eos();          // next token is "}" (consumed) or not (untouched)
  void entry(const std::string& name);
  void begin_section(const std::string& name) {
    entry(name);
    entry("{");
  }
  void end_section() { entry("}"); }
  std::string peek();

 private:
  void skip();
  std::ifstream f_;
};

// compiled expression: postfix code over a value stack
class Expr {
 public:
  Expr() = default;
  Expr(const std::string& text, const std::map<std::string, double>& constants);
  double operator()(double x, double y, double z, double t) const;
  const std::string& text() const { return text_; }
  bool empty() const { return code_.empty(); }

  struct Op {
    int kind;
    double v;
  };

 private:
  std::string text_;
  std::vector<Op> code_;
};

struct Param {
  bool scalar = false;
  tecno_cfg base{};                 // mesh-independent fields
  std::vector<int> Nx, Ny, Nz;      // mesh levels
  double xmin = 0, xmax = 1, ymin = 0, ymax = 1, zmin = 0, zmax = 1;
  std::string reconstruction;       // the reconstruct keyword (SOLN_FILES_<name>)
  std::string net_file;
  double final_time = 0;
  int max_iter = 0;
  std::map<std::string, double> constants;
  std::vector<Expr> ic;             // scalar: 1; systems: rho, u, v, w, p
  bool exact_available = false;
  std::vector<Expr> exact;
  bool ooc = false;
  int pdim = 1;
  int t_stamps = 1;
  std::string output_path = ".";
  int mesh_levels = 1;
};

Param read_param(const std::string& path);

}  // namespace cli
}  // namespace tecno
