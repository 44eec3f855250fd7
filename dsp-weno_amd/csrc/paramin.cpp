// paramin.cpp — param.in reader, expression evaluator, Parameter::read.
#include "paramin.hpp"

#include <cctype>
#include <cstdlib>
#include <sstream>

namespace tecno {
namespace cli {

// ================================================================ Reader
void Reader::skip() {
  // Reader::skipComment (reader.h): blanks, ',' ';' and comments
  for (;;) {
    int c = f_.get();
    while (c != EOF && (std::isspace(c) || c == ',' || c == ';')) c = f_.get();
    if (c == '#' || c == '%' || (c == '/' && f_.peek() == '/')) {
      std::string dummy;
      std::getline(f_, dummy);
      continue;
    }
    if (c == '/' && f_.peek() == '*') {
      f_.get();
      int last = 0;
      while (f_.good()) {
        c = f_.get();
        if (c == EOF || (c == '/' && last == '*')) break;
        last = c;
      }
      continue;
    }
    if (c != EOF) f_.putback((char)c);
    return;
  }
}

std::string Reader::token() {
  skip();
  std::string s;
  if (!(f_ >> s)) throw InputError("unexpected end of param.in");
  return s;
}

std::string Reader::peek() {
  std::streampos pos = f_.tellg();
  std::string s;
  skip();
  f_ >> s;
  f_.clear();
  f_.seekg(pos);
  return s;
}

double Reader::number() {
  skip();
  double d;
  if (!(f_ >> d)) throw InputError("expected a number in param.in");
  return d;
}

int Reader::integer() {
  skip();
  int i;
  if (!(f_ >> i)) throw InputError("expected an integer in param.in");
  return i;
}

std::string Reader::line() {
  skip();
  std::string s;
  std::getline(f_, s);
  return s;
}

bool Reader::eos() {
  std::streampos pos = f_.tellg();
  skip();
  std::string s;
  f_ >> s;
  if (s == "}") return true;
  f_.clear();
  f_.seekg(pos);
  return false;
}

void Reader::entry(const std::string& name) {
  const std::string s = token();
  if (s != name) throw InputError("param.in: expecting " + name + ", but found " + s);
}

// ================================================================ Expr
namespace {
enum OpKind {
  OP_NUM = 0, OP_X, OP_Y, OP_Z, OP_T,
  OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_POW, OP_NEG, OP_NOT,
  OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE, OP_AND, OP_OR,
  OP_F1,  // unary function, index in v
  OP_F2,  // binary function
  OP_IF
};

typedef double (*F1)(double);
typedef double (*F2)(double, double);

const std::pair<const char*, F1> kF1[] = {
    {"sin", [](double a) { return std::sin(a); }},    {"cos", [](double a) { return std::cos(a); }},
    {"tan", [](double a) { return std::tan(a); }},    {"asin", [](double a) { return std::asin(a); }},
    {"acos", [](double a) { return std::acos(a); }},  {"atan", [](double a) { return std::atan(a); }},
    {"sinh", [](double a) { return std::sinh(a); }},  {"cosh", [](double a) { return std::cosh(a); }},
    {"tanh", [](double a) { return std::tanh(a); }},  {"asinh", [](double a) { return std::asinh(a); }},
    {"acosh", [](double a) { return std::acosh(a); }}, {"atanh", [](double a) { return std::atanh(a); }},
    {"exp", [](double a) { return std::exp(a); }},    {"exp2", [](double a) { return std::exp2(a); }},
    {"log", [](double a) { return std::log(a); }},    {"log2", [](double a) { return std::log2(a); }},
    {"log10", [](double a) { return std::log10(a); }}, {"sqrt", [](double a) { return std::sqrt(a); }},
    {"cbrt", [](double a) { return std::cbrt(a); }},  {"abs", [](double a) { return std::fabs(a); }},
    {"floor", [](double a) { return std::floor(a); }}, {"ceil", [](double a) { return std::ceil(a); }},
    {"trunc", [](double a) { return std::trunc(a); }}, {"int", [](double a) { return std::round(a); }},
    {"sec", [](double a) { return 1.0 / std::cos(a); }}, {"csc", [](double a) { return 1.0 / std::sin(a); }},
    {"cot", [](double a) { return 1.0 / std::tan(a); }}};
const std::pair<const char*, F2> kF2[] = {
    {"pow", [](double a, double b) { return std::pow(a, b); }},
    {"atan2", [](double a, double b) { return std::atan2(a, b); }},
    {"hypot", [](double a, double b) { return std::hypot(a, b); }},
    {"min", [](double a, double b) { return a < b ? a : b; }},
    {"max", [](double a, double b) { return a > b ? a : b; }}};

// recursive-descent parser (fparser precedence, lowest first):
//   |   &   comparisons   + -   * / %   unary - !   ^ (right)   primary
class Parser {
 public:
  Parser(const std::string& s, const std::map<std::string, double>& k) : s_(s), k_(k) {}
  std::vector<Expr::Op> parse() {
    orx();
    ws();
    if (i_ != s_.size()) fail("unexpected '" + s_.substr(i_) + "'");
    return code_;
  }

 private:
  void fail(const std::string& m) { throw InputError("expression \"" + s_ + "\": " + m); }
  void ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool eat(const char* t) {
    ws();
    const size_t n = std::char_traits<char>::length(t);
    if (s_.compare(i_, n, t) == 0) {
      i_ += n;
      return true;
    }
    return false;
  }
  void emit(int k, double v = 0) { code_.push_back(Expr::Op{k, v}); }
  void orx() {
    andx();
    while (eat("|")) {
      andx();
      emit(OP_OR);
    }
  }
  void andx() {
    cmp();
    while (eat("&")) {
      cmp();
      emit(OP_AND);
    }
  }
  void cmp() {
    add();
    for (;;) {
      if (eat("!=")) { add(); emit(OP_NE); }
      else if (eat("<=")) { add(); emit(OP_LE); }
      else if (eat(">=")) { add(); emit(OP_GE); }
      else if (eat("==") || eat("=")) { add(); emit(OP_EQ); }
      else if (eat("<")) { add(); emit(OP_LT); }
      else if (eat(">")) { add(); emit(OP_GT); }
      else return;
    }
  }
  void add() {
    mul();
    for (;;) {
      if (eat("+")) { mul(); emit(OP_ADD); }
      else if (eat("-")) { mul(); emit(OP_SUB); }
      else return;
    }
  }
  void mul() {
    unary();
    for (;;) {
      if (eat("*")) { unary(); emit(OP_MUL); }
      else if (eat("/")) { unary(); emit(OP_DIV); }
      else if (eat("%")) { unary(); emit(OP_MOD); }
      else return;
    }
  }
  void unary() {
    if (eat("-")) { unary(); emit(OP_NEG); return; }
    if (eat("!")) { unary(); emit(OP_NOT); return; }
    if (eat("+")) { unary(); return; }
    power();
  }
  void power() {
    primary();
    if (eat("^")) {
      unary();  // right associative, binds tighter than the left operand's sign
      emit(OP_POW);
    }
  }
  void primary() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    const char c = s_[i_];
    if (c == '(') {
      ++i_;
      orx();
      if (!eat(")")) fail("missing )");
      return;
    }
    if (std::isdigit((unsigned char)c) || c == '.') {
      const char* b = s_.c_str() + i_;
      char* e = nullptr;
      const double v = std::strtod(b, &e);
      if (e == b) fail("bad number");
      i_ += (size_t)(e - b);
      emit(OP_NUM, v);
      return;
    }
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i_;
      while (j < s_.size() && (std::isalnum((unsigned char)s_[j]) || s_[j] == '_')) ++j;
      const std::string name = s_.substr(i_, j - i_);
      i_ = j;
      ws();
      if (i_ < s_.size() && s_[i_] == '(') {
        ++i_;
        if (name == "if") {
          orx();
          if (!eat(",")) fail("if(c, a, b)");
          orx();
          if (!eat(",")) fail("if(c, a, b)");
          orx();
          if (!eat(")")) fail("missing )");
          emit(OP_IF);
          return;
        }
        for (size_t q = 0; q < sizeof(kF1) / sizeof(kF1[0]); ++q)
          if (name == kF1[q].first) {
            orx();
            if (!eat(")")) fail("missing )");
            emit(OP_F1, (double)q);
            return;
          }
        for (size_t q = 0; q < sizeof(kF2) / sizeof(kF2[0]); ++q)
          if (name == kF2[q].first) {
            orx();
            if (!eat(",")) fail(name + " takes two arguments");
            orx();
            if (!eat(")")) fail("missing )");
            emit(OP_F2, (double)q);
            return;
          }
        fail("unknown function " + name);
      }
      if (name == "x") { emit(OP_X); return; }
      if (name == "y") { emit(OP_Y); return; }
      if (name == "z") { emit(OP_Z); return; }
      if (name == "t") { emit(OP_T); return; }
      auto it = k_.find(name);
      if (it != k_.end()) { emit(OP_NUM, it->second); return; }
      if (name == "pi") { emit(OP_NUM, M_PI); return; }
      fail("unknown name " + name);
    }
    fail(std::string("unexpected '") + c + "'");
  }

  const std::string& s_;
  const std::map<std::string, double>& k_;
  size_t i_ = 0;
  std::vector<Expr::Op> code_;
};
}  // namespace

Expr::Expr(const std::string& text, const std::map<std::string, double>& constants)
    : text_(text) {
  Parser p(text_, constants);
  code_ = p.parse();
}

double Expr::operator()(double x, double y, double z, double t) const {
  double st[64];
  int n = 0;
  for (const Op& o : code_) {
    if (n >= 62) throw InputError("expression too deep: " + text_);
    switch (o.kind) {
      case OP_NUM: st[n++] = o.v; break;
      case OP_X: st[n++] = x; break;
      case OP_Y: st[n++] = y; break;
      case OP_Z: st[n++] = z; break;
      case OP_T: st[n++] = t; break;
      case OP_NEG: st[n - 1] = -st[n - 1]; break;
      case OP_NOT: st[n - 1] = (st[n - 1] == 0.0) ? 1.0 : 0.0; break;
      case OP_F1: st[n - 1] = kF1[(int)o.v].second(st[n - 1]); break;
      case OP_F2: st[n - 2] = kF2[(int)o.v].second(st[n - 2], st[n - 1]); --n; break;
      case OP_IF: st[n - 3] = (st[n - 3] != 0.0) ? st[n - 2] : st[n - 1]; n -= 2; break;
      default: {
        const double a = st[n - 2], b = st[n - 1];
        double r = 0;
        switch (o.kind) {
          case OP_ADD: r = a + b; break;
          case OP_SUB: r = a - b; break;
          case OP_MUL: r = a * b; break;
          case OP_DIV: r = a / b; break;
          case OP_MOD: r = std::fmod(a, b); break;
          case OP_POW: r = std::pow(a, b); break;
          case OP_EQ: r = (a == b); break;
          case OP_NE: r = (a != b); break;
          case OP_LT: r = (a < b); break;
          case OP_LE: r = (a <= b); break;
          case OP_GT: r = (a > b); break;
          case OP_GE: r = (a >= b); break;
          case OP_AND: r = (a != 0.0 && b != 0.0); break;
          case OP_OR: r = (a != 0.0 || b != 0.0); break;
        }
        st[n - 2] = r;
        --n;
      }
    }
  }
  return n ? st[n - 1] : 0.0;
}

// ================================================================ Parameter
namespace {
double const_expr(const std::string& s, const std::map<std::string, double>& k) {
  return Expr(s, k)(0, 0, 0, 0);
}

void read_grid(Reader& r, Param& p) {
  // grid expressions see only pi: constants are read later (parameter.cc:34-98)
  const std::map<std::string, double> none;
  r.begin_section("grid");
  const char* ax[3] = {"x", "y", "z"};
  std::vector<int>* N[3] = {&p.Nx, &p.Ny, &p.Nz};
  double* lo[3] = {&p.xmin, &p.ymin, &p.zmin};
  double* hi[3] = {&p.xmax, &p.ymax, &p.zmax};
  for (int d = 0; d < 3; ++d) {
    r.entry(std::string(ax[d]) + "_min");
    *lo[d] = const_expr(r.line(), none);
    r.entry(std::string(ax[d]) + "_max");
    *hi[d] = const_expr(r.line(), none);
    r.begin_section(std::string("N") + ax[d] + "_cell");
    while (!r.eos()) {
      const int n = r.integer();
      if (n <= 0) throw InputError("cell counts must be positive");
      N[d]->push_back(n);
    }
  }
  r.end_section();
  if (!(p.xmin < p.xmax && p.ymin < p.ymax && p.zmin < p.zmax)) throw InputError("empty domain");
  if (p.Nx.size() != p.Ny.size() || p.Nx.size() != p.Nz.size() || p.Nx.empty())
    throw InputError("Nx_cell, Ny_cell and Nz_cell need the same number of levels");
}

void read_numeric(Reader& r, Param& p) {
  tecno_cfg& c = p.base;
  r.begin_section("numeric");
  r.entry("time_scheme");
  const std::string ts = r.token();
  if (ts == "ssprk3") c.time_scheme = TECNO_TS_SSPRK3;
  else if (ts == "rk1") c.time_scheme = TECNO_TS_RK1;
  else if (ts == "jameson_rk4") c.time_scheme = TECNO_TS_JAMESON_RK4;
  else if (ts == "rk4") c.time_scheme = TECNO_TS_RK4;
  else throw InputError("unknown time_scheme " + ts);
  r.entry("cfl");
  c.cfl = r.number();
  r.entry("max_iter");
  p.max_iter = r.integer();
  r.entry("final_time");
  p.final_time = r.number();
  r.entry("min_residue");
  (void)r.number();  // parsed but unused by the reference
  r.entry("reconstruct");
  const std::string rec = r.token();
  p.reconstruction = rec;
  c.spweno_corr = 0;
  c.eno_level = 2;
  if (rec == "sp_weno") c.recon = TECNO_REC_SP_WENO;
  else if (rec == "sp_wenoc") { c.recon = TECNO_REC_SP_WENO; c.spweno_corr = 1; }
  else if (rec == "sp_weno_corr2") { c.recon = TECNO_REC_SP_WENO; c.spweno_corr = 2; }
  else if (rec == "sp_weno_dl") c.recon = TECNO_REC_SP_WENO_DL;
  else if (rec == "first") c.recon = TECNO_REC_FIRST;
  else if (rec == "second") c.recon = TECNO_REC_SECOND;
  else if (rec == "minmod") c.recon = TECNO_REC_MINMOD;
  else if (rec == "minabs") c.recon = TECNO_REC_MINABS;
  else if (rec == "eno") {
    c.recon = TECNO_REC_ENO;
    c.eno_level = r.integer();
  } else
    throw InputError("unknown reconstruct " + rec);
  r.entry("net_file");
  p.net_file = r.token();
  r.end_section();
}

void read_material(Reader& r, Param& p) {
  tecno_cfg& c = p.base;
  r.begin_section("material");
  const std::string first = r.peek();
  if (first == "model") {  // enflo_scalar parameter.cc:230-320
    p.scalar = true;
    c.system = TECNO_SYS_SCALAR;
    r.entry("model");
    const std::string m = r.token();
    if (m == "linadv") {
      c.model = TECNO_MODEL_LINADV;
      r.begin_section("vel");
      for (int d = 0; d < 3; ++d) c.vel[d] = r.number();
      r.end_section();
    } else if (m == "burgers") {
      c.model = TECNO_MODEL_BURGERS;
    } else
      throw InputError("unknown scalar model " + m);
    r.entry("flux");
    const std::string f = r.token();
    const char* names[8] = {"lin_ec", "lin_ec4", "lin_tecno", "lin_tecno4",
                            "bur_ec", "bur_ec4", "bur_tecno", "bur_tecno4"};
    c.flux = -1;
    for (int q = 0; q < 8; ++q)
      if (f == names[q]) c.flux = q;
    if (c.flux < 0) throw InputError("unknown scalar flux " + f);
  } else {  // enflo/src/parameter.cc:222-409
    c.system = TECNO_SYS_EULER;
    r.entry("gamma");
    c.gamma = r.number();
    r.entry("gas_const");
    c.gas_const = r.number();
    r.begin_section("viscosity");
    r.entry("model");
    const std::string vm = r.token();
    if (vm == "constant") {
      c.mu_model = TECNO_MU_CONSTANT;
      r.entry("mu_ref");
      c.mu_ref = r.number();
    } else if (vm == "sutherland") {
      c.mu_model = TECNO_MU_SUTHERLAND;
      r.entry("mu_ref");
      c.mu_ref = r.number();
      r.entry("T_ref");
      c.T_ref = r.number();
      r.entry("T_0");
      c.T_0 = r.number();
    } else if (vm == "power") {
      c.mu_model = TECNO_MU_POWER;
      r.entry("mu_ref");
      c.mu_ref = r.number();
      r.entry("T_ref");
      c.T_ref = r.number();
      r.entry("omega");
      c.omega = r.number();
    } else
      throw InputError("unknown viscosity type " + vm);
    r.end_section();
    r.entry("prandtl");
    c.prandtl = r.number();
    r.entry("model");
    const std::string m = r.token();
    if (m == "euler") c.model = TECNO_MODEL_EULER;
    else if (m == "ns_kep") c.model = TECNO_MODEL_NS_KEP;
    else if (m == "ns_es") c.model = TECNO_MODEL_NS_ES;
    else throw InputError("unknown model " + m);
    r.entry("flux");
    const std::string f = r.token();
    const std::pair<const char*, int> names[] = {
        {"kepes_tecno_roe", 100}, {"kepes_tecno4_roe", 101}, {"kepes_tecno_rusanov", 102},
        {"kepes_tecno4_rusanov", 103}, {"roe_tecno_roe", 104}, {"roe_tecno4_roe", 105},
        {"kep_tecno_roe", 106}, {"kepec", 120}, {"kepec4", 121}, {"roe_ec", 122},
        {"roe_ec4", 123}, {"kep", 124}, {"roe", 125}, {"kep4", 126}, {"keps", 127},
        {"keps4", 128}};
    c.flux = -1;
    for (auto& nm : names)
      if (f == nm.first) c.flux = nm.second;
    if (c.flux < 0) throw InputError("unknown flux " + f);
    r.entry("EC1_fix");
    const std::string e = r.token();
    if (e == "yes") c.ec1_fix = 1;
    else if (e == "no") c.ec1_fix = 0;
    else throw InputError("EC1_fix must be yes or no");
  }
  r.end_section();
}

void read_constants(Reader& r, Param& p) {
  r.begin_section("constants");
  while (!r.eos()) {
    const std::string name = r.token();
    p.constants[name] = r.number();
  }
}

std::vector<Expr> read_fields(Reader& r, Param& p) {
  std::vector<Expr> out;
  if (p.scalar) {
    out.emplace_back(r.line(), p.constants);
  } else {
    const char* names[5] = {"density", "xvelocity", "yvelocity", "zvelocity", "pressure"};
    for (const char* n : names) {
      r.entry(n);
      out.emplace_back(r.line(), p.constants);
    }
  }
  return out;
}

void read_initial(Reader& r, Param& p) {
  r.begin_section("initial_condition");
  p.ic = read_fields(r, p);
  r.end_section();
}

void read_exact(Reader& r, Param& p) {
  r.begin_section("exact_soln");
  r.entry("available");
  const std::string a = r.token();
  if (a == "yes") {
    p.exact_available = true;
    p.exact = read_fields(r, p);
    r.end_section();
  } else if (a == "no") {
    while (!r.eos()) (void)r.token();
  } else
    throw InputError("exact_soln available must be yes or no");
}

void read_bc(Reader& r, Param& p) {
  r.begin_section("boundary_condition");
  const char* names[6] = {"xleft", "xright", "yleft", "yright", "zleft", "zright"};
  for (int i = 0; i < 6; ++i) {
    r.begin_section(names[i]);
    r.entry("type");
    const std::string t = r.token();
    if (t == "periodic") p.base.bc[i] = TECNO_BC_PERIODIC;
    else if (t == "neumann") p.base.bc[i] = TECNO_BC_NEUMANN;
    else if (t == "wall") p.base.bc[i] = TECNO_BC_WALL;
    else if (t == "dirichlet")
      throw InputError("dirichlet boundaries are rejected (as parameter.cc:558-562)");
    else
      throw InputError("unknown boundary type " + t);
    r.end_section();
  }
  r.end_section();
  for (int d = 0; d < 3; ++d)
    if ((p.base.bc[2 * d] == TECNO_BC_PERIODIC) != (p.base.bc[2 * d + 1] == TECNO_BC_PERIODIC))
      throw InputError("periodic boundaries must be paired");
}

void read_output(Reader& r, Param& p) {
  r.begin_section("output");
  r.entry("OOC_study");
  const std::string o = r.token();
  if (o == "yes") p.ooc = true;
  else if (o == "no") p.ooc = false;
  else throw InputError("OOC_study must be yes or no");
  r.entry("problem_dim");
  p.pdim = r.integer();
  if (p.pdim < 1 || p.pdim > 3) throw InputError("problem_dim must be 1, 2 or 3");
  if (p.ooc && p.Nx.size() < 2)
    throw InputError("Cannot perform OOC study without at least two mesh levels");
  if (p.ooc && !p.exact_available)
    throw InputError("Cannot perform OOC study as exact solution not available");
  r.entry("time_stamps");
  p.t_stamps = r.integer();
  if (p.t_stamps < 1) throw InputError("time_stamps must be >= 1");
  r.entry("output_path");
  p.output_path = r.token();
  r.end_section();
  p.mesh_levels = p.ooc ? (int)p.Nx.size() : 1;
}
}  // namespace

Param read_param(const std::string& path) {
  Reader r(path);
  Param p;
  tecno_cfg_default(&p.base);
  read_grid(r, p);
  read_numeric(r, p);
  read_material(r, p);
  read_constants(r, p);
  read_initial(r, p);
  read_exact(r, p);
  read_bc(r, p);
  read_output(r, p);
  return p;
}

}  // namespace cli
}  // namespace tecno
