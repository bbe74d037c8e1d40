// Host-side inputs: condition table reader, UDF expression evaluator (libmatheval's C entry
// points, as bound by src/libfosolvers/UDF/UDF.f90:24-46), UDF table reader, error state.
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/fosolvers_amd.h"
#include "common.h"

namespace fos {
static thread_local std::string g_err;
void set_error(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}
}  // namespace fos

namespace {

// ---- expression evaluator: recursive descent to a postfix program ----
enum Op : int { PUSH, VAR, ADD, SUB, MUL, DIV, POW, NEG, FUN };
enum Fn : int { F_EXP, F_LOG, F_SQRT, F_SIN, F_COS, F_TAN, F_ASIN, F_ACOS, F_ATAN, F_SINH, F_COSH,
                F_TANH, F_ABS, F_STEP, F_DELTA, F_NANDELTA, F_ERF };
struct Ins { int op; int arg; double val; };
struct Expr {
  std::vector<Ins> code;
  std::vector<std::string> vars;
};

struct Parser {
  const char *s;
  Expr *ex;
  bool ok = true;
  void ws() { while (*s && std::isspace((unsigned char)*s)) s++; }
  int var_index(const std::string &n) {
    for (size_t i = 0; i < ex->vars.size(); i++) if (ex->vars[i] == n) return (int)i;
    ex->vars.push_back(n);
    return (int)ex->vars.size() - 1;
  }
  // grammar (libmatheval precedence): expr = term {(+|-) term}; term = unary {(*|/) unary};
  // unary = -unary | power; power = primary {^ primary} (left associative)
  void expr() {
    term();
    for (;;) {
      ws();
      if (*s == '+') { s++; term(); ex->code.push_back({ADD, 0, 0}); }
      else if (*s == '-') { s++; term(); ex->code.push_back({SUB, 0, 0}); }
      else break;
    }
  }
  void term() {
    unary();
    for (;;) {
      ws();
      if (*s == '*') { s++; unary(); ex->code.push_back({MUL, 0, 0}); }
      else if (*s == '/') { s++; unary(); ex->code.push_back({DIV, 0, 0}); }
      else break;
    }
  }
  void unary() {
    ws();
    if (*s == '-') { s++; unary(); ex->code.push_back({NEG, 0, 0}); return; }
    if (*s == '+') { s++; unary(); return; }
    power();
  }
  void power() {
    primary();
    for (;;) {
      ws();
      if (*s == '^') { s++; primary(); ex->code.push_back({POW, 0, 0}); }
      else break;
    }
  }
  void primary() {
    ws();
    if (*s == '(') {
      s++;
      expr();
      ws();
      if (*s != ')') { ok = false; return; }
      s++;
      return;
    }
    if (std::isdigit((unsigned char)*s) || *s == '.') {
      char *end;
      double v = std::strtod(s, &end);
      if (end == s) { ok = false; return; }
      s = end;
      ex->code.push_back({PUSH, 0, v});
      return;
    }
    if (std::isalpha((unsigned char)*s) || *s == '_') {
      std::string name;
      while (std::isalnum((unsigned char)*s) || *s == '_') name += *s++;
      ws();
      static const char *fnames[] = {"exp", "log", "sqrt", "sin", "cos", "tan", "asin", "acos", "atan",
                                     "sinh", "cosh", "tanh", "abs", "step", "delta", "nandelta", "erf"};
      for (int f = 0; f < (int)(sizeof(fnames) / sizeof(fnames[0])); f++) {
        if (name == fnames[f]) {
          if (*s != '(') { ok = false; return; }
          s++;
          expr();
          ws();
          if (*s != ')') { ok = false; return; }
          s++;
          ex->code.push_back({FUN, f, 0});
          return;
        }
      }
      if (name == "e") { ex->code.push_back({PUSH, 0, 2.718281828459045235}); return; }
      if (name == "pi") { ex->code.push_back({PUSH, 0, 3.141592653589793238}); return; }
      ex->code.push_back({VAR, var_index(name), 0});
      return;
    }
    ok = false;
  }
};

double run(const Expr &ex, const double *vals) {
  double st[256];
  int sp = 0;
  for (const Ins &in : ex.code) {
    switch (in.op) {
      case PUSH: st[sp++] = in.val; break;
      case VAR: st[sp++] = vals[in.arg]; break;
      case ADD: sp--; st[sp - 1] += st[sp]; break;
      case SUB: sp--; st[sp - 1] -= st[sp]; break;
      case MUL: sp--; st[sp - 1] *= st[sp]; break;
      case DIV: sp--; st[sp - 1] /= st[sp]; break;
      case POW: sp--; st[sp - 1] = std::pow(st[sp - 1], st[sp]); break;
      case NEG: st[sp - 1] = -st[sp - 1]; break;
      case FUN: {
        double &a = st[sp - 1];
        switch (in.arg) {
          case F_EXP: a = std::exp(a); break;
          case F_LOG: a = std::log(a); break;
          case F_SQRT: a = std::sqrt(a); break;
          case F_SIN: a = std::sin(a); break;
          case F_COS: a = std::cos(a); break;
          case F_TAN: a = std::tan(a); break;
          case F_ASIN: a = std::asin(a); break;
          case F_ACOS: a = std::acos(a); break;
          case F_ATAN: a = std::atan(a); break;
          case F_SINH: a = std::sinh(a); break;
          case F_COSH: a = std::cosh(a); break;
          case F_TANH: a = std::tanh(a); break;
          case F_ABS: a = std::fabs(a); break;
          case F_STEP: a = a < 0.0 ? 0.0 : 1.0; break;
          case F_DELTA: a = a == 0.0 ? 1.0 : 0.0; break;
          case F_NANDELTA: a = a == 0.0 ? NAN : 0.0; break;
          case F_ERF: a = std::erf(a); break;
        }
        break;
      }
    }
  }
  return sp == 1 ? st[0] : NAN;
}

// strip a trailing comment ("# ...") and whitespace from a line of an input file
std::string strip(const std::string &l) {
  std::string s = l.substr(0, l.find('#'));
  size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

}  // namespace

extern "C" {

const char *fos_last_error(void) { return fos::g_err.c_str(); }
const char *fos_version(void) { return "fosolvers_amd 0.1 (gfx950)"; }

void *evaluator_create(char *string) {
  auto ex = std::make_unique<Expr>();
  Parser p{string, ex.get()};
  p.expr();
  p.ws();
  if (!p.ok || *p.s != '\0' || ex->code.empty()) {
    fos::set_error("evaluator_create: cannot parse \"%s\"", string);
    return nullptr;
  }
  // a postfix program of depth > 256 is refused rather than overflowing the stack
  int depth = 0, maxd = 0;
  for (auto &in : ex->code) {
    depth += (in.op == PUSH || in.op == VAR) ? 1 : (in.op == NEG || in.op == FUN) ? 0 : -1;
    maxd = depth > maxd ? depth : maxd;
  }
  if (maxd > 256) { fos::set_error("evaluator_create: expression too deep"); return nullptr; }
  return ex.release();
}

void evaluator_destroy(void *evaluator) { delete static_cast<Expr *>(evaluator); }

double evaluator_evaluate(void *evaluator, int count, char **names, double *values) {
  const Expr *ex = static_cast<const Expr *>(evaluator);
  if (!ex) return NAN;
  std::vector<double> v(ex->vars.size(), 0.0);
  for (size_t i = 0; i < ex->vars.size(); i++)
    for (int k = 0; k < count; k++)
      if (ex->vars[i] == names[k]) { v[i] = values[k]; break; }
  return run(*ex, v.data());
}

int fos_read_udf_tab(const char *path, void **handles, int maxHandles) {
  std::ifstream in(path);
  if (!in) { fos::set_error("readUDFTab: cannot open %s", path); return -1; }
  std::string line;
  std::getline(in, line);
  int n = std::atoi(strip(line).c_str());
  if (n < 0 || n > maxHandles) { fos::set_error("readUDFTab: %d UDFs do not fit %d handles", n, maxHandles); return -1; }
  for (int i = 0; i < n; i++) {
    if (!std::getline(in, line)) { fos::set_error("readUDFTab: missing line %d", i + 2); return -1; }
    std::string s = line;
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    s = a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
    handles[i] = evaluator_create(const_cast<char *>(s.c_str()));
    if (!handles[i]) return -1;
  }
  return n;
}

int fos_read_cond_tab(const char *path, fos_cond_tab *out) {
  // condTab.f90:4-21: n, m, then per condition: separator line, gid, type, k, k parameters
  std::ifstream in(path);
  if (!in) { fos::set_error("readCondTab: cannot open %s", path); return -1; }
  std::vector<std::string> lines;
  std::string l;
  while (std::getline(in, l)) lines.push_back(l);
  size_t pos = 0;
  // list-directed reads skip empty records until a value is found; a bare read(fid,*) consumes one
  auto next = [&](std::string &s) -> bool {
    while (pos < lines.size()) {
      s = strip(lines[pos++]);
      if (!s.empty()) return true;
    }
    return false;
  };
  auto skip = [&]() -> bool { return pos++ < lines.size(); };
  std::string s;
  if (!next(s)) return -1;
  int n = std::atoi(s.c_str());
  if (!next(s)) return -1;
  int m = std::atoi(s.c_str());
  out->n = n;
  out->maxPar = m;
  out->gid = fos::xalloc<int>(n);
  out->t = fos::xalloc<int>(n);
  out->p = fos::xalloc<double>((size_t)n * m);
  std::fill(out->p, out->p + (size_t)n * m, 0.0);
  for (int i = 0; i < n; i++) {
    bool ok = skip();  // the separator line (read(fid,*) with no items)
    ok = ok && next(s);
    out->gid[i] = std::atoi(s.c_str());
    ok = ok && next(s);
    out->t[i] = std::atoi(s.c_str());
    ok = ok && next(s);
    int k = std::atoi(s.c_str());
    if (!ok || k > m) { fos::set_error("readCondTab: malformed block %d", i + 1); fos_cond_tab_free(out); return -1; }
    for (int j = 0; j < k; j++) {
      if (!next(s)) { fos::set_error("readCondTab: missing parameter"); fos_cond_tab_free(out); return -1; }
      out->p[(size_t)m * i + j] = std::strtod(s.c_str(), nullptr);
    }
  }
  return 0;
}

void fos_cond_tab_free(fos_cond_tab *c) {
  if (!c) return;
  std::free(c->gid);
  std::free(c->t);
  std::free(c->p);
  std::memset(c, 0, sizeof(*c));
}

}  // extern "C"
