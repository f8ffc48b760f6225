// Expressions of x and y, as the reference's input files write them
// (VLP2D_Data/InputData/*.txt: funcRHS, funcBoundary, funcAnalytic, funcLeft..
// funcBottom, xMax, yMax), compiled to a postfix (RPN) program that the host
// evaluator below and the GPU interpreter (mg_setup.hpp) run identically.
//
// Grammar, after the reference's operator table
// (VisualLaplacePoisson2D/Substitutor/OperationsRegistry.cs:64-84):
//   ^ (power, left-associative, binds tightest), then unary - and +
//   (right-associative), then * and / (left), then + and - (left), then the
//   conditional c ? a : b (right-associative; c is true when nonzero).
// Operands: numbers (invariant culture, optional exponent), the variables x
// and y, parenthesised expressions, and the functions sin cos tan asin acos
// atan sinh cosh tanh exp ln log (natural) log10 sqrt abs, and pi(a) = pi*a
// (the reference's custom function, Common/Utils.cs:352-364).
// Subexpressions without variables are folded to constants at compile time.
// a^n with an integer constant n, |n| <= 64, is evaluated by a fixed sequence of
// multiplications (x^2 = x*x), so host and GPU agree bit for bit on
// polynomial expressions; other powers and the functions use each side's
// libm (the GPU's may differ from the host's in the last place).
#pragma once
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <string>
#include <vector>

#if defined(__HIPCC__)
#define VLP2D_HD __host__ __device__
#else
#define VLP2D_HD
#endif

namespace vlp2d {

enum ExprCode : int {
  EX_CONST = 0,  // push val
  EX_X,          // push x
  EX_Y,          // push y
  EX_ADD,
  EX_SUB,
  EX_MUL,
  EX_DIV,
  EX_POW,   // a ^ b, b not a small integer constant
  EX_POWI,  // a ^ arg (integer constant exponent)
  EX_NEG,
  EX_FN,    // unary function arg
  EX_COND,  // c ? a : b
};

enum ExprFn : int {
  FN_SIN = 0,
  FN_COS,
  FN_TAN,
  FN_ASIN,
  FN_ACOS,
  FN_ATAN,
  FN_SINH,
  FN_COSH,
  FN_TANH,
  FN_EXP,
  FN_LN,
  FN_LOG10,
  FN_SQRT,
  FN_ABS,
  FN_PI,  // pi * a
  FN_COUNT
};

struct ExprOp {
  int code;
  int arg;
  double val;
};

constexpr int kExprMaxDepth = 16;  // evaluation stack of the GPU interpreter
constexpr int kExprMaxOps = 256;

// a^n by square-and-multiply from the least significant bit; the same
// sequence on host and device
template <typename F> VLP2D_HD inline F expr_powi(F a, int n) {
  const bool inv = n < 0;
  unsigned m = (unsigned)(inv ? -n : n);
  F r = F(1), b = a;
  bool first = true;
  while (m) {
    if (m & 1u) {
      r = first ? b : r * b;
      first = false;
    }
    m >>= 1;
    if (m) b = b * b;
  }
  return inv ? F(1) / r : r;
}

inline double expr_fn_host(int fn, double a) {
  switch (fn) {
    case FN_SIN: return std::sin(a);
    case FN_COS: return std::cos(a);
    case FN_TAN: return std::tan(a);
    case FN_ASIN: return std::asin(a);
    case FN_ACOS: return std::acos(a);
    case FN_ATAN: return std::atan(a);
    case FN_SINH: return std::sinh(a);
    case FN_COSH: return std::cosh(a);
    case FN_TANH: return std::tanh(a);
    case FN_EXP: return std::exp(a);
    case FN_LN: return std::log(a);
    case FN_LOG10: return std::log10(a);
    case FN_SQRT: return std::sqrt(a);
    case FN_ABS: return std::fabs(a);
    case FN_PI: return 3.141592653589793 * a;
    default: return std::nan("");
  }
}

// Evaluates a compiled program (host).
inline double expr_eval(const ExprOp *ops, int n, double x, double y) {
  double st[kExprMaxDepth];
  int sp = 0;
  for (int k = 0; k < n; ++k) {
    const ExprOp &o = ops[k];
    switch (o.code) {
      case EX_CONST: st[sp++] = o.val; break;
      case EX_X: st[sp++] = x; break;
      case EX_Y: st[sp++] = y; break;
      case EX_ADD: --sp; st[sp - 1] = st[sp - 1] + st[sp]; break;
      case EX_SUB: --sp; st[sp - 1] = st[sp - 1] - st[sp]; break;
      case EX_MUL: --sp; st[sp - 1] = st[sp - 1] * st[sp]; break;
      case EX_DIV: --sp; st[sp - 1] = st[sp - 1] / st[sp]; break;
      case EX_POW: --sp; st[sp - 1] = std::pow(st[sp - 1], st[sp]); break;
      case EX_POWI: st[sp - 1] = expr_powi(st[sp - 1], o.arg); break;
      case EX_NEG: st[sp - 1] = -st[sp - 1]; break;
      case EX_FN: st[sp - 1] = expr_fn_host(o.arg, st[sp - 1]); break;
      case EX_COND: sp -= 2; st[sp - 1] = st[sp - 1] != 0.0 ? st[sp] : st[sp + 1]; break;
      default: return std::nan("");
    }
  }
  return sp == 1 ? st[0] : std::nan("");
}

struct Expr {
  std::vector<ExprOp> ops;
  int depth = 0;
  bool uses_x = false, uses_y = false;
  bool empty() const { return ops.empty(); }
  double eval(double x, double y) const { return expr_eval(ops.data(), (int)ops.size(), x, y); }
};

namespace detail {

struct Node {
  int code;
  int arg;
  double val;
  int a, b, c;  // children (-1 = none)
};

class ExprParser {
 public:
  ExprParser(const std::string &s) : s_(s) {}

  bool parse(std::vector<Node> &nodes, int &root, std::string &err) {
    nodes_ = &nodes;
    root = cond();
    skip();
    if (err_.empty() && p_ != s_.size()) fail("unexpected '" + std::string(1, s_[p_]) + "'");
    err = err_;
    return err_.empty();
  }

 private:
  const std::string &s_;
  size_t p_ = 0;
  std::string err_;
  std::vector<Node> *nodes_ = nullptr;

  void fail(const std::string &m) {
    if (err_.empty()) err_ = m + " at position " + std::to_string(p_ + 1);
  }
  void skip() {
    while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
  }
  bool eat(char c) {
    skip();
    if (p_ < s_.size() && s_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  int node(int code, int arg, double val, int a = -1, int b = -1, int c = -1) {
    nodes_->push_back(Node{code, arg, val, a, b, c});
    return (int)nodes_->size() - 1;
  }
  int cond() {
    const int c = sum();
    if (!err_.empty()) return -1;
    if (eat('?')) {
      const int a = cond();
      if (!eat(':')) {
        fail("expected ':'");
        return -1;
      }
      const int b = cond();
      return node(EX_COND, 0, 0.0, c, a, b);
    }
    return c;
  }
  int sum() {
    int l = prod();
    while (err_.empty()) {
      if (eat('+')) l = node(EX_ADD, 0, 0.0, l, prod());
      else if (eat('-')) l = node(EX_SUB, 0, 0.0, l, prod());
      else break;
    }
    return l;
  }
  int prod() {
    int l = unary();
    while (err_.empty()) {
      if (eat('*')) l = node(EX_MUL, 0, 0.0, l, unary());
      else if (eat('/')) l = node(EX_DIV, 0, 0.0, l, unary());
      else break;
    }
    return l;
  }
  int unary() {
    if (eat('-')) return node(EX_NEG, 0, 0.0, unary());
    if (eat('+')) return unary();
    return power();
  }
  // ^ binds tighter than unary minus and is left-associative; its right
  // operand may carry signs (2^-x)
  int power() {
    int l = primary();
    while (err_.empty() && eat('^')) {
      int r;
      if (eat('-')) r = node(EX_NEG, 0, 0.0, signed_primary());
      else if (eat('+')) r = signed_primary();
      else r = primary();
      l = node(EX_POW, 0, 0.0, l, r);
    }
    return l;
  }
  int signed_primary() {
    if (eat('-')) return node(EX_NEG, 0, 0.0, signed_primary());
    if (eat('+')) return signed_primary();
    return primary();
  }
  int primary() {
    skip();
    if (p_ >= s_.size()) {
      fail("unexpected end of expression");
      return -1;
    }
    const char ch = s_[p_];
    if (ch == '(') {
      ++p_;
      const int e = cond();
      if (!eat(')')) fail("expected ')'");
      return e;
    }
    if (std::isdigit((unsigned char)ch) || ch == '.') {
      const char *b = s_.c_str() + p_;
      char *e = nullptr;
      const double v = std::strtod(b, &e);
      if (e == b) {
        fail("bad number");
        return -1;
      }
      p_ += (size_t)(e - b);
      return node(EX_CONST, 0, v);
    }
    if (std::isalpha((unsigned char)ch) || ch == '_') {
      size_t q = p_;
      while (q < s_.size() && (std::isalnum((unsigned char)s_[q]) || s_[q] == '_')) ++q;
      const std::string name = s_.substr(p_, q - p_);
      p_ = q;
      if (name == "x") return node(EX_X, 0, 0.0);
      if (name == "y") return node(EX_Y, 0, 0.0);
      static const char *const names[FN_COUNT] = {"sin",  "cos", "tan", "asin", "acos",  "atan", "sinh", "cosh",
                                                   "tanh", "exp", "ln",  "log10", "sqrt", "abs",  "pi"};
      int fn = -1;
      for (int k = 0; k < FN_COUNT; ++k)
        if (name == names[k]) fn = k;
      if (name == "log") fn = FN_LN;
      if (fn < 0) {
        fail("unknown name '" + name + "'");
        return -1;
      }
      if (!eat('(')) {
        fail("expected '(' after " + name);
        return -1;
      }
      const int a = cond();
      if (!eat(')')) fail("expected ')'");
      return node(EX_FN, fn, 0.0, a);
    }
    fail("unexpected '" + std::string(1, ch) + "'");
    return -1;
  }
};

inline bool is_const(const std::vector<Node> &n, int k) { return k >= 0 && n[k].code == EX_CONST; }

// constant folding and integer powers, bottom-up
inline void fold(std::vector<Node> &n, int k) {
  if (k < 0) return;
  fold(n, n[k].a);
  fold(n, n[k].b);
  fold(n, n[k].c);
  Node &z = n[k];
  const bool ca = is_const(n, z.a), cb = is_const(n, z.b), cc = is_const(n, z.c);
  auto va = [&] { return n[z.a].val; };
  auto vb = [&] { return n[z.b].val; };
  auto set = [&](double v) {
    z.code = EX_CONST;
    z.val = v;
    z.a = z.b = z.c = -1;
  };
  switch (z.code) {
    case EX_ADD: if (ca && cb) set(va() + vb()); break;
    case EX_SUB: if (ca && cb) set(va() - vb()); break;
    case EX_MUL: if (ca && cb) set(va() * vb()); break;
    case EX_DIV: if (ca && cb) set(va() / vb()); break;
    case EX_NEG: if (ca) set(-va()); break;
    case EX_FN: if (ca) set(expr_fn_host(z.arg, va())); break;
    case EX_COND: if (ca && cb && cc) set(va() != 0.0 ? vb() : n[z.c].val); break;
    case EX_POW:
      if (cb && std::fabs(vb()) <= 64.0 && vb() == std::floor(vb())) {
        const int e = (int)vb();
        if (ca) set(expr_powi(va(), e));
        else {
          z.code = EX_POWI;
          z.arg = e;
          z.b = -1;
        }
      } else if (ca && cb) {
        set(std::pow(va(), vb()));
      }
      break;
    default: break;
  }
}

inline void emit(const std::vector<Node> &n, int k, Expr &out, int &depth, int &maxd) {
  const Node &z = n[k];
  for (int ch : {z.a, z.b, z.c})
    if (ch >= 0) emit(n, ch, out, depth, maxd);
  const int pops = (z.a >= 0) + (z.b >= 0) + (z.c >= 0);
  depth += 1 - pops;
  if (depth > maxd) maxd = depth;
  if (z.code == EX_X) out.uses_x = true;
  if (z.code == EX_Y) out.uses_y = true;
  out.ops.push_back(ExprOp{z.code, z.arg, z.val});
}

}  // namespace detail

// Compiles text into out.  An empty or all-blank text gives an empty
// program (the caller's "absent function").  Returns false with a message on
// a syntax error, an unknown name, or a program too deep or too long for the
// GPU interpreter.
inline bool expr_compile(const std::string &text, Expr &out, std::string &err) {
  out = Expr();
  err.clear();
  bool blank = true;
  for (char ch : text)
    if (!std::isspace((unsigned char)ch)) blank = false;
  if (blank) return true;
  std::vector<detail::Node> nodes;
  int root = -1;
  detail::ExprParser p(text);
  if (!p.parse(nodes, root, err)) return false;
  detail::fold(nodes, root);
  int depth = 0, maxd = 0;
  detail::emit(nodes, root, out, depth, maxd);
  out.depth = maxd;
  if (maxd > kExprMaxDepth) {
    err = "expression nests deeper than " + std::to_string(kExprMaxDepth);
    out = Expr();
    return false;
  }
  if ((int)out.ops.size() > kExprMaxOps) {
    err = "expression longer than " + std::to_string(kExprMaxOps) + " operations";
    out = Expr();
    return false;
  }
  return true;
}

}  // namespace vlp2d
