// xacc_lite.cpp -- XASM front end and accelerator lookup of the XACC API subset.
#include "xacc_lite.hpp"

#include <cctype>

namespace xacc {

namespace {

struct Lexer {
  std::string s;
  size_t p = 0;
  void ws() {
    while (p < s.size()) {
      if (isspace((unsigned char)s[p])) { ++p; continue; }
      if (s.compare(p, 2, "//") == 0) { while (p < s.size() && s[p] != '\n') ++p; continue; }
      break;
    }
  }
  bool eat(const std::string& t) {
    ws();
    if (s.compare(p, t.size(), t) == 0) { p += t.size(); return true; }
    return false;
  }
  void expect(const std::string& t) {
    if (!eat(t)) error("xasm: expected '" + t + "' at offset " + std::to_string(p));
  }
  std::string ident() {
    ws();
    size_t b = p;
    while (p < s.size() && (isalnum((unsigned char)s[p]) || s[p] == '_')) ++p;
    if (b == p) error("xasm: expected identifier at offset " + std::to_string(p));
    return s.substr(b, p - b);
  }
  char peek() { ws(); return p < s.size() ? s[p] : '\0'; }
};

// expression value: numeric, or a symbolic "[-][k*]var[/d]" form bound later
struct Val {
  bool sym = false;
  double v = 0;
  std::string var;
  double mul = 1.0, div = 1.0;
};

struct Parser {
  Lexer L;
  std::vector<std::string> vars;               // declared double parameters
  std::map<std::string, long> ints;            // loop variables in scope

  Val primary() {
    char c = L.peek();
    if (c == '(') { L.eat("("); Val v = expr(); L.expect(")"); return v; }
    if (c == '-') { L.eat("-"); Val v = primary(); if (v.sym) v.mul = -v.mul; else v.v = -v.v; return v; }
    if (isdigit((unsigned char)c) || c == '.') {
      size_t b = L.p;
      while (L.p < L.s.size() && (isdigit((unsigned char)L.s[L.p]) || L.s[L.p] == '.' || L.s[L.p] == 'e' ||
                                  ((L.s[L.p] == '-' || L.s[L.p] == '+') && (L.s[L.p - 1] == 'e'))))
        ++L.p;
      Val v;
      v.v = std::stod(L.s.substr(b, L.p - b));
      return v;
    }
    std::string id = L.ident();
    Val v;
    if (id == "pi") { v.v = constants::pi; return v; }
    auto it = ints.find(id);
    if (it != ints.end()) { v.v = (double)it->second; return v; }
    if (std::find(vars.begin(), vars.end(), id) == vars.end()) error("xasm: unknown symbol " + id);
    v.sym = true;
    v.var = id;
    return v;
  }
  Val term() {
    Val a = primary();
    for (;;) {
      if (L.eat("*")) {
        Val b = primary();
        if (a.sym && b.sym) error("xasm: product of two variables");
        if (a.sym) a.mul *= b.v;
        else if (b.sym) { b.mul *= a.v; a = b; }
        else a.v *= b.v;
      } else if (L.peek() == '/') {
        L.eat("/");
        Val b = primary();
        if (b.sym) error("xasm: division by a variable");
        if (a.sym) a.div *= b.v; else a.v /= b.v;
      } else {
        return a;
      }
    }
  }
  Val expr() {
    Val a = term();
    for (;;) {
      if (L.peek() == '+' ) { L.eat("+"); Val b = term(); if (a.sym || b.sym) error("xasm: sums with variables unsupported"); a.v += b.v; }
      else if (L.peek() == '-') { L.eat("-"); Val b = term(); if (a.sym || b.sym) error("xasm: sums with variables unsupported"); a.v -= b.v; }
      else return a;
    }
  }
  long int_expr() {
    Val v = expr();
    if (v.sym) error("xasm: qubit index must be an integer expression");
    return (long)std::llround(v.v);
  }
  InstructionParameter to_param(const Val& v) {
    if (!v.sym) return InstructionParameter(v.v);
    std::ostringstream o;
    if (v.mul < 0) o << "-";
    double m = std::fabs(v.mul);
    if (m != 1.0) o << m << "*";
    o << v.var;
    if (v.div != 1.0) o << "/" << v.div;
    return InstructionParameter(o.str());
  }

  void statements(std::shared_ptr<CompositeInstruction>& out, const std::string& qreg) {
    while (L.peek() != '}' && L.peek() != '\0') statement(out, qreg);
  }
  void statement(std::shared_ptr<CompositeInstruction>& out, const std::string& qreg) {
    if (L.eat("for")) {
      // for (int i = A; i < B; i++) { ... }
      L.expect("(");
      L.eat("int");
      std::string v = L.ident();
      L.expect("=");
      long a = int_expr();
      L.expect(";");
      std::string v2 = L.ident();
      if (v2 != v) error("xasm: loop variable mismatch");
      L.expect("<");
      long b = int_expr();
      L.expect(";");
      std::string v3 = L.ident();
      if (v3 != v) error("xasm: loop variable mismatch");
      L.expect("++");
      L.expect(")");
      L.expect("{");
      size_t body = L.p;
      size_t end = body;
      for (long i = a; i < b; ++i) {
        ints[v] = i;
        L.p = body;
        statements(out, qreg);
        end = L.p;
      }
      if (a >= b) {  // skip an empty loop body
        int depth = 1;
        while (L.p < L.s.size() && depth) { if (L.s[L.p] == '{') ++depth; if (L.s[L.p] == '}') --depth; ++L.p; }
        --L.p;
        end = L.p;
      }
      L.p = end;
      L.expect("}");
      ints.erase(v);
      return;
    }
    std::string g = L.ident();
    L.expect("(");
    std::vector<size_t> bits;
    std::vector<InstructionParameter> params;
    bool first = true;
    while (!L.eat(")")) {
      if (!first) L.expect(",");
      first = false;
      L.ws();
      size_t save = L.p;
      std::string id;
      if (isalpha((unsigned char)L.peek())) id = L.ident();
      if (id == qreg && L.eat("[")) {
        long q = int_expr();
        L.expect("]");
        bits.push_back((size_t)q);
      } else {
        L.p = save;
        params.push_back(to_param(expr()));
      }
    }
    L.expect(";");
    out->addInstruction(createGate(g, bits, params));
  }
};

}  // namespace

std::shared_ptr<IR> XasmCompiler::compile(const std::string& src, std::shared_ptr<Accelerator>) {
  auto ir = std::make_shared<IR>();
  Parser P;
  P.L.s = src;
  while (P.L.peek() != '\0') {
    P.L.expect("__qpu__");
    P.L.expect("void");
    std::string name = P.L.ident();
    P.L.expect("(");
    std::string qreg = "q";
    P.vars.clear();
    bool first = true;
    while (!P.L.eat(")")) {
      if (!first) P.L.expect(",");
      first = false;
      std::string type = P.L.ident();
      std::string var = P.L.ident();
      if (type == "qbit" || type == "qreg") qreg = var;
      else P.vars.push_back(var);
    }
    P.L.expect("{");
    auto c = std::make_shared<CompositeInstruction>(name);
    for (auto& v : P.vars) c->addVariable(v);
    P.statements(c, qreg);
    P.L.expect("}");
    ir->composites.push_back(c);
  }
  return ir;
}

std::shared_ptr<Accelerator> getAccelerator(const std::string& name, const HeterogeneousMap& params) {
  auto acc = getService<Accelerator>(name);
  acc->initialize(params);
  return acc;
}

}  // namespace xacc
