// udf.cpp -- reader for the UDF subset KAPSEL's input uses (include/spm_udf.h).
//
// A UDF file holds a schema and data.  The schema sections (global_def and
// def) declare classes (`class Vector3d:{x: float, ...}`) and entries
// `name: type "description"`, with types int, double, float, string,
// `select {'a','b'}`, a braced record of entries, or a class name; `name[]`
// declares an array (one [] per dimension).  The data section gives
// `name: value` for top-level entries, each value laid out as its schema
// says: a record's member values in braces (commas optional), an array's
// elements in brackets, a select as the chosen quoted string.  A record
// member that is itself a record (the bodies of a select's alternatives, as
// in constitutive_eq) appears in braces in schema order whether chosen or
// not.  Comments run from // to the end of the line.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/spm_udf.h"
#include "mt19937.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

// ---------------------------------------------------------------------
// tokens
// ---------------------------------------------------------------------
enum Tok { T_END, T_LBRACE, T_RBRACE, T_LBRACK, T_RBRACK, T_COMMA, T_COLON, T_STRING, T_WORD, T_DIRECTIVE };

struct Token {
  Tok t;
  std::string s;  // string contents, word, or directive "begin:data" / "end:data" / "include:path"
  int line;
};

bool lex(const std::string &src, std::vector<Token> &out, std::string &err) {
  size_t i = 0;
  int line = 1;
  const size_t n = src.size();
  while (i < n) {
    const char ch = src[i];
    if (ch == '\n') {
      ++line;
      ++i;
    } else if (isspace((unsigned char)ch)) {
      ++i;
    } else if (ch == '/' && i + 1 < n && src[i + 1] == '/') {
      while (i < n && src[i] != '\n') ++i;
    } else if (ch == '{' || ch == '}' || ch == '[' || ch == ']' || ch == ',' || ch == ':') {
      const Tok t = ch == '{' ? T_LBRACE : ch == '}' ? T_RBRACE : ch == '[' ? T_LBRACK : ch == ']' ? T_RBRACK
                                                                                                    : ch == ',' ? T_COMMA
                                                                                                                : T_COLON;
      out.push_back({t, std::string(1, ch), line});
      ++i;
    } else if (ch == '"' || ch == '\'') {
      size_t j = i + 1;
      while (j < n && src[j] != ch) {
        if (src[j] == '\n') ++line;
        ++j;
      }
      if (j >= n) {
        err = "unterminated string at line " + std::to_string(line);
        return false;
      }
      out.push_back({T_STRING, src.substr(i + 1, j - i - 1), line});
      i = j + 1;
    } else if (ch == '\\') {
      // \begin{x}, \end{x}, \include{"file"}
      size_t j = i + 1;
      while (j < n && isalpha((unsigned char)src[j])) ++j;
      const std::string cmd = src.substr(i + 1, j - i - 1);
      if (j >= n || src[j] != '{') {
        err = "malformed directive at line " + std::to_string(line);
        return false;
      }
      const size_t k = src.find('}', j);
      if (k == std::string::npos) {
        err = "malformed directive at line " + std::to_string(line);
        return false;
      }
      std::string arg = src.substr(j + 1, k - j - 1);
      if (!arg.empty() && (arg.front() == '"' || arg.front() == '\'')) arg = arg.substr(1, arg.size() - 2);
      out.push_back({T_DIRECTIVE, cmd + ":" + arg, line});
      i = k + 1;
    } else {
      size_t j = i;
      while (j < n && !isspace((unsigned char)src[j]) && strchr("{}[],:\"'", src[j]) == nullptr &&
             !(src[j] == '/' && j + 1 < n && src[j + 1] == '/'))
        ++j;
      out.push_back({T_WORD, src.substr(i, j - i), line});
      i = j;
    }
  }
  out.push_back({T_END, "", line});
  return true;
}

// ---------------------------------------------------------------------
// schema and values
// ---------------------------------------------------------------------
struct Type;
using TypeP = std::shared_ptr<Type>;
enum Kind { K_INT, K_REAL, K_STRING, K_SELECT, K_RECORD, K_ARRAY, K_REF };

struct Type {
  Kind kind = K_REAL;
  std::vector<std::pair<std::string, TypeP>> fields;  // K_RECORD
  std::vector<std::string> options;                   // K_SELECT
  TypeP elem;                                         // K_ARRAY
  std::string ref;                                    // K_REF: class name
};

struct Value {
  Kind kind = K_REAL;
  double num = 0.0;
  std::string str;
  std::vector<std::pair<std::string, Value>> fields;  // K_RECORD (names from the schema)
  std::vector<Value> items;                           // K_ARRAY
};

struct Doc {
  std::map<std::string, TypeP> classes;
  std::vector<std::pair<std::string, TypeP>> entries;  // top-level schema entries
  std::map<std::string, Value> data;
  std::vector<std::map<std::string, Value>> records;  // \begin{record} blocks in file order
};

class Parser {
 public:
  Parser(Doc &d, std::vector<Token> &t) : doc(d), tok(t) {}

  bool run(std::string &err) {
    try {
      while (peek().t != T_END) {
        const Token &d = peek();
        if (d.t != T_DIRECTIVE) throw std::runtime_error("expected a section at line " + std::to_string(d.line));
        const std::string s = next().s;
        if (s == "begin:global_def" || s == "begin:def")
          schema_section();
        else if (s == "begin:data")
          data_section(doc.data, "end:data");
        else if (s == "begin:record") {
          doc.records.emplace_back();
          data_section(doc.records.back(), "end:record");
        }
        else if (s == "begin:header")
          skip_to("end:header");
        else if (s.rfind("begin:", 0) == 0)
          skip_to("end:" + s.substr(6));
        else
          throw std::runtime_error("unexpected " + s + " at line " + std::to_string(d.line));
      }
    } catch (const std::exception &e) {
      err = e.what();
      return false;
    }
    return true;
  }

 private:
  Doc &doc;
  std::vector<Token> &tok;
  size_t pos = 0;

  // The token list always ends with one T_END; the cursor never moves past
  // it, so a file cut anywhere reads as T_END from there on.
  const Token &peek() const { return tok[pos]; }
  void adv() {
    if (tok[pos].t != T_END) ++pos;
  }
  const Token &next() {
    const Token &t = tok[pos];
    adv();
    return t;
  }
  void expect(Tok t, const char *what) {
    if (peek().t != t) throw std::runtime_error(std::string("expected ") + what + " at line " + std::to_string(peek().line));
    adv();
  }
  void skip_to(const std::string &dir) {
    while (peek().t != T_END && !(peek().t == T_DIRECTIVE && peek().s == dir)) adv();
    if (peek().t == T_END) throw std::runtime_error("missing \\" + dir);
    adv();
  }
  void skip_commas() {
    while (peek().t == T_COMMA) adv();
  }

  // name with optional [] dimensions; returns the dimension count
  std::string entry_name(int &dims) {
    const Token &t = next();
    if (t.t != T_WORD) throw std::runtime_error("expected a name at line " + std::to_string(t.line));
    std::string name = t.s;
    dims = 0;
    while (peek().t == T_LBRACK) {
      adv();
      expect(T_RBRACK, "]");
      ++dims;
    }
    return name;
  }

  TypeP type_expr() {
    auto ty = std::make_shared<Type>();
    const Token &t = next();
    if (t.t == T_LBRACE) {
      ty->kind = K_RECORD;
      skip_commas();
      while (peek().t != T_RBRACE) {
        int dims = 0;
        const std::string nm = entry_name(dims);
        expect(T_COLON, ":");
        TypeP ft = type_expr();
        for (int k = 0; k < dims; ++k) {
          auto a = std::make_shared<Type>();
          a->kind = K_ARRAY;
          a->elem = ft;
          ft = a;
        }
        ty->fields.push_back({nm, ft});
        skip_commas();
      }
      adv();
    } else if (t.t == T_WORD) {
      if (t.s == "int") {
        ty->kind = K_INT;
      } else if (t.s == "double" || t.s == "float") {
        ty->kind = K_REAL;
      } else if (t.s == "string") {
        ty->kind = K_STRING;
      } else if (t.s == "select") {
        ty->kind = K_SELECT;
        expect(T_LBRACE, "{");
        while (peek().t != T_RBRACE) {
          if (peek().t == T_STRING || peek().t == T_WORD) ty->options.push_back(next().s);
          else if (peek().t == T_COMMA) adv();
          else throw std::runtime_error("bad select at line " + std::to_string(peek().line));
        }
        adv();
      } else {
        ty->kind = K_REF;
        ty->ref = t.s;
      }
    } else {
      throw std::runtime_error("expected a type at line " + std::to_string(t.line));
    }
    if (peek().t == T_STRING) adv();  // description
    return ty;
  }

  void schema_section() {
    while (!(peek().t == T_DIRECTIVE && peek().s.rfind("end:", 0) == 0)) {
      if (peek().t == T_END) throw std::runtime_error("unterminated schema section");
      if (peek().t == T_WORD && peek().s == "class") {
        adv();
        const std::string nm = next().s;
        expect(T_COLON, ":");
        doc.classes[nm] = type_expr();
        continue;
      }
      int dims = 0;
      const std::string nm = entry_name(dims);
      expect(T_COLON, ":");
      TypeP ty = type_expr();
      for (int k = 0; k < dims; ++k) {
        auto a = std::make_shared<Type>();
        a->kind = K_ARRAY;
        a->elem = ty;
        ty = a;
      }
      doc.entries.push_back({nm, ty});
      skip_commas();
    }
    adv();
  }

  TypeP resolve(TypeP t) {
    int guard = 0;
    while (t->kind == K_REF) {
      auto it = doc.classes.find(t->ref);
      if (it == doc.classes.end() || ++guard > 64) throw std::runtime_error("unknown class " + t->ref);
      t = it->second;
    }
    return t;
  }

  Value value(TypeP ty) {
    ty = resolve(ty);
    Value v;
    v.kind = ty->kind;
    switch (ty->kind) {
      case K_INT:
      case K_REAL: {
        const Token &t = next();
        if (t.t != T_WORD) throw std::runtime_error("expected a number at line " + std::to_string(t.line));
        char *end = nullptr;
        v.num = std::strtod(t.s.c_str(), &end);
        if (end == t.s.c_str() || *end != '\0')
          throw std::runtime_error("bad number '" + t.s + "' at line " + std::to_string(t.line));
        break;
      }
      case K_STRING:
      case K_SELECT: {
        const Token &t = next();
        if (t.t != T_STRING && t.t != T_WORD)
          throw std::runtime_error("expected a string at line " + std::to_string(t.line));
        v.str = t.s;
        break;
      }
      case K_RECORD: {
        expect(T_LBRACE, "{");
        for (auto &f : ty->fields) {
          skip_commas();
          v.fields.push_back({f.first, value(f.second)});
        }
        skip_commas();
        expect(T_RBRACE, "}");
        break;
      }
      case K_ARRAY: {
        expect(T_LBRACK, "[");
        skip_commas();
        while (peek().t != T_RBRACK) {
          v.items.push_back(value(ty->elem));
          skip_commas();
        }
        adv();
        break;
      }
      default:
        throw std::runtime_error("unresolved type");
    }
    return v;
  }

  void data_section(std::map<std::string, Value> &out, const std::string &end) {
    while (!(peek().t == T_DIRECTIVE && peek().s == end)) {
      if (peek().t == T_END) throw std::runtime_error("missing \\" + end);
      int dims = 0;
      const int line = peek().line;
      const std::string nm = entry_name(dims);
      expect(T_COLON, ":");
      TypeP ty;
      for (auto &e : doc.entries)
        if (e.first == nm) ty = e.second;
      if (!ty) throw std::runtime_error("no schema for '" + nm + "' at line " + std::to_string(line));
      out[nm] = value(ty);
      skip_commas();
    }
    adv();
  }
};

bool read_file(const std::string &path, std::string &out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

// tokens of `path` with \include{...} replaced by the included file's tokens
bool lex_file(const std::string &path, std::vector<Token> &out, std::string &err, int depth = 0) {
  if (depth > 8) {
    err = "include nesting too deep";
    return false;
  }
  std::string src;
  if (!read_file(path, src)) {
    err = "cannot read " + path;
    return false;
  }
  std::vector<Token> t;
  if (!lex(src, t, err)) {
    err = path + ": " + err;
    return false;
  }
  t.pop_back();  // T_END
  const std::string dir = path.find('/') == std::string::npos ? "" : path.substr(0, path.rfind('/') + 1);
  for (auto &x : t) {
    if (x.t == T_DIRECTIVE && x.s.rfind("include:", 0) == 0) {
      std::string inc = x.s.substr(8);
      if (!inc.empty() && inc[0] != '/') inc = dir + inc;
      if (!lex_file(inc, out, err, depth + 1)) return false;
    } else {
      out.push_back(x);
    }
  }
  return true;
}

// ---------------------------------------------------------------------
// locations
// ---------------------------------------------------------------------
// A location is a top-level name followed by .member and [i] steps; an
// empty [] stands for every element of that array.
struct Step {
  enum { MEMBER, INDEX, ALL } kind;
  std::string name;
  long index;
};

bool split_location(const std::string &loc, std::string &head, std::vector<Step> &steps) {
  size_t i = 0;
  while (i < loc.size() && loc[i] != '.' && loc[i] != '[') ++i;
  head = loc.substr(0, i);
  if (head.empty()) return false;
  while (i < loc.size()) {
    if (loc[i] == '.') {
      size_t j = i + 1;
      while (j < loc.size() && loc[j] != '.' && loc[j] != '[') ++j;
      if (j == i + 1) return false;
      steps.push_back({Step::MEMBER, loc.substr(i + 1, j - i - 1), 0});
      i = j;
    } else {  // '['
      const size_t k = loc.find(']', i);
      if (k == std::string::npos) return false;
      const std::string idx = loc.substr(i + 1, k - i - 1);
      if (idx.empty()) {
        steps.push_back({Step::ALL, "", 0});
      } else {
        if (idx.find_first_not_of("0123456789") != std::string::npos || idx.size() > 9) return false;
        steps.push_back({Step::INDEX, "", std::strtol(idx.c_str(), nullptr, 10)});
      }
      i = k + 1;
    }
  }
  return true;
}

// Appends the values `steps[s..]` reaches from v; false if a step fails.
bool walk(const Value *v, const std::vector<Step> &steps, size_t s, std::vector<const Value *> &out) {
  for (; s < steps.size(); ++s) {
    const Step &st = steps[s];
    if (st.kind == Step::MEMBER) {
      if (v->kind != K_RECORD) return false;
      const Value *nxt = nullptr;
      for (auto &f : v->fields)
        if (f.first == st.name) nxt = &f.second;
      if (!nxt) return false;
      v = nxt;
    } else if (st.kind == Step::INDEX) {
      if (v->kind != K_ARRAY || st.index >= (long)v->items.size()) return false;
      v = &v->items[st.index];
    } else {
      if (v->kind != K_ARRAY) return false;
      if (s + 1 == steps.size()) break;  // trailing []: the array itself
      for (auto &it : v->items)
        if (!walk(&it, steps, s + 1, out)) return false;
      return true;
    }
  }
  out.push_back(v);
  return true;
}

// Values at `loc`: the current record's entries first, then the data section.
bool lookup(const Doc &d, int record, const std::string &loc, std::vector<const Value *> &out) {
  std::string head;
  std::vector<Step> steps;
  if (!split_location(loc, head, steps)) return false;
  const Value *root = nullptr;
  if (record >= 0) {
    auto it = d.records[record].find(head);
    if (it != d.records[record].end()) root = &it->second;
  }
  if (!root) {
    auto it = d.data.find(head);
    if (it == d.data.end()) return false;
    root = &it->second;
  }
  return walk(root, steps, 0, out);
}

// The single value at `loc` (no [] inside it); a trailing [] gives the array.
const Value *find(const Doc &d, int record, const std::string &loc) {
  std::vector<const Value *> v;
  const size_t inner = loc.find("[]");
  if (inner != std::string::npos && inner + 2 != loc.size()) return nullptr;
  if (!lookup(d, record, loc, v) || v.size() != 1) return nullptr;
  return v[0];
}

}  // namespace

struct spm_udf {
  Doc doc;
  int record = -1;  // record the getters read first (spm_udf_jump), -1 for none
};

extern "C" {

const char *spm_udf_last_error(void) { return g_err.c_str(); }

int spm_udf_open(const char *path, spm_udf **out) {
  if (!path || !out) return fail(SPM_ERR_INVALID, "null argument");
  *out = nullptr;
  std::vector<Token> tok;
  std::string err;
  if (!lex_file(path, tok, err)) return fail(SPM_ERR_INVALID, err);
  tok.push_back({T_END, "", 0});
  auto u = std::make_unique<spm_udf>();
  Parser p(u->doc, tok);
  if (!p.run(err)) return fail(SPM_ERR_INVALID, std::string(path) + ": " + err);
  *out = u.release();
  return SPM_OK;
}

void spm_udf_close(spm_udf *udf) { delete udf; }

int spm_udf_get_double(const spm_udf *u, const char *loc, double *v) {
  if (!u || !loc || !v) return fail(SPM_ERR_INVALID, "null argument");
  const Value *x = std::strstr(loc, "[]") ? nullptr : find(u->doc, u->record, loc);
  if (!x || (x->kind != K_REAL && x->kind != K_INT)) return fail(SPM_ERR_INVALID, std::string("no number at ") + loc);
  *v = x->num;
  return SPM_OK;
}

int spm_udf_get_int(const spm_udf *u, const char *loc, int *v) {
  double d = 0.0;
  const int rc = spm_udf_get_double(u, loc, &d);
  if (rc) return rc;
  if (d != std::floor(d)) return fail(SPM_ERR_INVALID, std::string("not an integer at ") + loc);
  *v = (int)d;
  return SPM_OK;
}

int spm_udf_get_string(const spm_udf *u, const char *loc, char *buf, int len) {
  if (!u || !loc || !buf || len <= 0) return fail(SPM_ERR_INVALID, "null argument");
  const Value *x = std::strstr(loc, "[]") ? nullptr : find(u->doc, u->record, loc);
  if (!x || (x->kind != K_STRING && x->kind != K_SELECT)) return fail(SPM_ERR_INVALID, std::string("no string at ") + loc);
  if ((int)x->str.size() >= len) return fail(SPM_ERR_INVALID, "buffer too small");
  std::memcpy(buf, x->str.c_str(), x->str.size() + 1);
  return SPM_OK;
}

int spm_udf_size(const spm_udf *u, const char *loc, int *n) {
  if (!u || !loc || !n) return fail(SPM_ERR_INVALID, "null argument");
  std::string l(loc);
  if (l.size() < 2 || l.compare(l.size() - 2, 2, "[]") != 0) l += "[]";
  const Value *x = find(u->doc, u->record, l);
  if (!x || x->kind != K_ARRAY) return fail(SPM_ERR_INVALID, std::string("no array at ") + loc);
  *n = (int)x->items.size();
  return SPM_OK;
}

int spm_udf_get_doubles(const spm_udf *u, const char *loc, double *out, int cap, int *n) {
  if (!u || !loc || !n || (cap > 0 && !out)) return fail(SPM_ERR_INVALID, "null argument");
  std::vector<const Value *> v;
  if (!lookup(u->doc, u->record, loc, v)) return fail(SPM_ERR_INVALID, std::string("no values at ") + loc);
  for (auto *x : v)
    if (x->kind != K_REAL && x->kind != K_INT) return fail(SPM_ERR_INVALID, std::string("not numbers at ") + loc);
  *n = (int)v.size();
  if ((int)v.size() > cap) return fail(SPM_ERR_INVALID, std::string("more values than room at ") + loc);
  for (size_t k = 0; k < v.size(); ++k) out[k] = v[k]->num;
  return SPM_OK;
}

int spm_udf_num_records(const spm_udf *u, int *n) {
  if (!u || !n) return fail(SPM_ERR_INVALID, "null argument");
  *n = (int)u->doc.records.size();
  return SPM_OK;
}

int spm_udf_jump(spm_udf *u, int record) {
  if (!u) return fail(SPM_ERR_INVALID, "null argument");
  if (record < -1 || record >= (int)u->doc.records.size()) return fail(SPM_ERR_INVALID, "no such record");
  u->record = record;
  return SPM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------
// output file: the input's text with includes inlined, then one
// \begin{record} block per snapshot
// ---------------------------------------------------------------------
namespace {

bool inline_includes(const std::string &path, std::string &out, std::string &err, int depth = 0) {
  if (depth > 8) {
    err = "include nesting too deep";
    return false;
  }
  std::string src;
  if (!read_file(path, src)) {
    err = "cannot read " + path;
    return false;
  }
  const std::string dir = path.find('/') == std::string::npos ? "" : path.substr(0, path.rfind('/') + 1);
  std::istringstream in(src);
  std::string line;
  while (std::getline(in, line)) {
    const size_t b = line.find_first_not_of(" \t");
    if (b != std::string::npos && line.compare(b, 9, "\\include{") == 0) {
      const size_t e = line.find('}', b);
      std::string inc = e == std::string::npos ? "" : line.substr(b + 9, e - b - 9);
      if (!inc.empty() && (inc.front() == '"' || inc.front() == '\'')) inc = inc.substr(1, inc.size() - 2);
      if (inc.empty()) {
        err = "malformed include in " + path;
        return false;
      }
      if (inc[0] != '/') inc = dir + inc;
      if (!inline_includes(inc, out, err, depth + 1)) return false;
      if (!out.empty() && out.back() != '\n') out += '\n';
    } else {
      out += line;
      out += '\n';
    }
  }
  return true;
}

void put3(std::string &o, const double *v) {
  char b[96];
  std::snprintf(b, sizeof(b), "{%.17g,%.17g,%.17g}", v[0], v[1], v[2]);
  o += b;
}

}  // namespace

extern "C" int spm_udf_create_output(const char *in_path, const char *out_path) {
  if (!in_path || !out_path) return fail(SPM_ERR_INVALID, "null argument");
  spm_udf *probe = nullptr;
  const int rc = spm_udf_open(in_path, &probe);
  if (rc) return rc;
  spm_udf_close(probe);
  std::string text, err;
  if (!inline_includes(in_path, text, err)) return fail(SPM_ERR_INVALID, err);
  std::ofstream f(out_path, std::ios::binary | std::ios::trunc);
  if (!f) return fail(SPM_ERR_INVALID, std::string("cannot write ") + out_path);
  f << text;
  return f.good() ? SPM_OK : fail(SPM_ERR_INVALID, std::string("write failed on ") + out_path);
}

extern "C" int spm_udf_append_record(const char *out_path, int ts, const spm_particle *p, int n) {
  if (!out_path || (!p && n > 0) || n < 0) return fail(SPM_ERR_INVALID, "null argument");
  // Output_udf (avs_output.cxx:343-396): E = 1.0, t = ts, and per particle
  // R, R_raw, v, q, omega, then the previous-step hydrodynamic, LJ and slip
  // forces and torques (torque_r written as zero)
  std::string o = "\\begin{record}\nE:1.0\n";
  o += "t:" + std::to_string(ts) + "\nParticles[]:[\n";
  const double zero[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < n; ++k) {
    const spm_particle &q = p[k];
    const double qq[4] = {q.q.s, q.q.v[0], q.q.v[1], q.q.v[2]};
    char b[128];
    o += "{";
    put3(o, q.x);
    put3(o, q.x_nopbc);
    put3(o, q.v);
    std::snprintf(b, sizeof(b), "{%.17g,%.17g,%.17g,%.17g}", qq[0], qq[1], qq[2], qq[3]);
    o += b;
    put3(o, q.omega);
    put3(o, q.f_hydro_previous);
    put3(o, q.torque_hydro_previous);
    put3(o, q.fr_previous);
    put3(o, zero);
    put3(o, q.f_slip_previous);
    put3(o, q.torque_slip_previous);
    o += "}\n";
  }
  o += "]\nRigidParticles[]:[]\n\\end{record}\n";
  std::ofstream f(out_path, std::ios::binary | std::ios::app);
  if (!f) return fail(SPM_ERR_INVALID, std::string("cannot append to ") + out_path);
  f << o;
  return f.good() ? SPM_OK : fail(SPM_ERR_INVALID, std::string("write failed on ") + out_path);
}

// ---------------------------------------------------------------------
// input.cxx mapping
// ---------------------------------------------------------------------
namespace {

struct Reader {
  const spm_udf *u;
  int rc = SPM_OK;
  double d(const std::string &loc) {
    double v = 0.0;
    if (rc == SPM_OK) rc = spm_udf_get_double(u, loc.c_str(), &v);
    return v;
  }
  int i(const std::string &loc) {
    int v = 0;
    if (rc == SPM_OK) rc = spm_udf_get_int(u, loc.c_str(), &v);
    return v;
  }
  std::string s(const std::string &loc) {
    char buf[256] = {0};
    if (rc == SPM_OK) rc = spm_udf_get_string(u, loc.c_str(), buf, (int)sizeof(buf));
    return buf;
  }
  int size(const std::string &loc) {
    int v = 0;
    if (rc == SPM_OK) rc = spm_udf_size(u, loc.c_str(), &v);
    return v;
  }
};

int unsupported(const std::string &what) { return fail(SPM_ERR_UNSUPPORTED, what + " is not run by this library"); }
int invalid(const std::string &what) { return fail(SPM_ERR_INVALID, "invalid " + what); }

}  // namespace

extern "C" int spm_udf_params(const spm_udf *u, spm_params *prm, spm_electrolyte *el, spm_udf_run *run) {
  if (!u || !prm || !run) return fail(SPM_ERR_INVALID, "null argument");
  Reader r{u};
  spm_params p;
  std::memset(&p, 0, sizeof(p));
  spm_udf_run rn;
  std::memset(&rn, 0, sizeof(rn));
  // constitutive_eq (input.cxx:438-790)
  const std::string eq = r.s("constitutive_eq.type");
  if (r.rc) return r.rc;
  std::string base;
  if (eq == "Navier_Stokes") {
    rn.equation = SPM_UDF_NAVIER_STOKES;
    base = "constitutive_eq.Navier_Stokes.";
  } else if (eq == "Shear_Navier_Stokes_Lees_Edwards") {
    rn.equation = SPM_UDF_LEES_EDWARDS;
    base = "constitutive_eq.Shear_Navier_Stokes_Lees_Edwards.";
  } else if (eq == "Electrolyte") {
    rn.equation = SPM_UDF_ELECTROLYTE;
    base = "constitutive_eq.Electrolyte.";
    if (!el) return fail(SPM_ERR_INVALID, "Electrolyte input needs an spm_electrolyte");
  } else {
    return unsupported("constitutive_eq " + eq);
  }
  p.dx = r.d(base + "DX");
  p.rho = r.d(base + "RHO");
  p.eta = r.d(base + "ETA");
  p.kbt = r.d(base + "kBT");
  if (rn.equation != SPM_UDF_ELECTROLYTE) {
    // the Electrolyte branch does not read alpha_v / alpha_o, which keep
    // their zero initial values (input.cxx:267, 660-667)
    p.alpha_v = r.d(base + "alpha_v");
    p.alpha_o = r.d(base + "alpha_o");
  }
  double frequency = 0.0;
  if (rn.equation == SPM_UDF_LEES_EDWARDS) {
    // input.cxx:634-653: DC and AC both read Shear_rate (the AC branch as a
    // steady shear); the time step is the Navier-Stokes one (:323-340)
    const std::string t = r.s(base + "External_field.type");
    if (t != "DC" && t != "AC" && r.rc == SPM_OK) return invalid("External_field type " + t);
    rn.shear_rate = r.d(base + "External_field." + t + ".Shear_rate");
  }
  if (rn.equation == SPM_UDF_ELECTROLYTE) {
    std::memset(el, 0, sizeof(*el));
    el->elementary_charge = 1.0;  // Elementary_charge (input.cxx:247)
    el->dielectric = r.d(base + "Dielectric_cst");
    const std::string init = r.s(base + "INIT_profile");
    rn.poisson_boltzmann = init == "Poisson_Boltzmann";
    if (!rn.poisson_boltzmann && init != "Uniform" && r.rc == SPM_OK) return invalid("INIT_profile " + init);
    const std::string salt = r.s(base + "Add_salt.type");
    if (salt == "saltfree") {
      el->n_spec = 1;
      el->valency[0] = r.d(base + "Add_salt.saltfree.Valency_counterion");
      el->onsager[0] = r.d(base + "Add_salt.saltfree.Onsager_coeff_counterion");
    } else if (salt == "salt") {
      el->n_spec = 2;
      el->valency[0] = r.d(base + "Add_salt.salt.Valency_positive_ion");
      el->valency[1] = r.d(base + "Add_salt.salt.Valency_negative_ion");
      el->onsager[0] = r.d(base + "Add_salt.salt.Onsager_coeff_positive_ion");
      el->onsager[1] = r.d(base + "Add_salt.salt.Onsager_coeff_negative_ion");
      el->debye_length = r.d(base + "Add_salt.salt.Debye_length");
    } else if (r.rc == SPM_OK) {
      return invalid("Add_salt " + salt);
    }
    const std::string ef = r.s(base + "Electric_field.type");
    if (ef == "ON") {
      el->external_field = 1;
      const std::string t = r.s(base + "Electric_field.ON.type");
      const std::string b = base + "Electric_field.ON." + t + ".";
      if (t == "AC") {
        el->ac = 1;
        frequency = r.d(b + "Frequency");
        el->angular_frequency = 2.0 * M_PI * frequency;  // PI2 * Frequency, input.cxx:356
      } else if (t != "DC" && r.rc == SPM_OK) {
        return invalid("Electric_field type " + t);
      }
      el->e_ext[0] = r.d(b + "Ex");
      el->e_ext[1] = r.d(b + "Ey");
      el->e_ext[2] = r.d(b + "Ez");
    } else if (ef != "OFF" && r.rc == SPM_OK) {
      return invalid("Electric_field " + ef);
    }
  }
  if (r.rc) return r.rc;
  // object_type (input.cxx:800-1330)
  const std::string obj = r.s("object_type.type");
  if (r.rc) return r.rc;
  if (obj != "spherical_particle") return unsupported("object_type " + obj);
  const int nspec = r.size("object_type.spherical_particle.Particle_spec[]");
  if (r.rc) return r.rc;
  if (nspec < 1 || nspec > SPM_MAX_SPECIES) return invalid("Particle_spec count");
  p.n_species = nspec;
  for (int s = 0; s < nspec; ++s) {
    const std::string b = "object_type.spherical_particle.Particle_spec[" + std::to_string(s) + "].";
    p.n_particles += r.i(b + "Particle_number");
    p.mass_ratio[s] = r.d(b + "MASS_RATIO");
    const double q = r.d(b + "Surface_charge");
    if (el && rn.equation == SPM_UDF_ELECTROLYTE) el->surface_charge[s] = q;
    const std::string jp = r.s(b + "janus_propulsion");
    if (r.rc) return r.rc;
    if (jp != "OFF") return unsupported("janus_propulsion " + jp);
  }
  // geometry and forces (input.cxx:1332-1405)
  p.a_xi = r.d("A_XI");
  p.a = r.d("A");
  p.g = r.d("gravity.G");
  const std::string gd = r.s("gravity.G_direction");
  p.epsilon = r.d("EPSILON");
  const std::string ljp = r.s("LJ_powers");
  int np[3] = {r.i("mesh.NPX"), r.i("mesh.NPY"), r.i("mesh.NPZ")};
  if (r.rc) return r.rc;
  if (gd == "-X") p.g_direction = 0;
  else if (gd == "-Y") p.g_direction = 1;
  else if (gd == "-Z") p.g_direction = 2;
  else return invalid("G_direction " + gd);
  if (ljp == "12:6") p.lj_powers = SPM_LJ_12_6;
  else if (ljp == "24:12") p.lj_powers = SPM_LJ_24_12;
  else if (ljp == "36:18") p.lj_powers = SPM_LJ_36_18;
  else return invalid("LJ_powers " + ljp);
  for (int d = 0; d < 3; ++d)
    if (np[d] < 0 || np[d] > 12) return invalid("mesh");
  p.nx = 1 << np[0];
  p.ny = 1 << np[1];
  p.nz = 1 << np[2];
  // time_increment (input.cxx:1406-1426)
  const std::string ti = r.s("time_increment.type");
  if (ti == "auto") {
    rn.auto_dt = 1;
    rn.dt_factor = r.d("time_increment.auto.factor");
  } else if (ti == "manual") {
    p.dt = r.d("time_increment.manual.delta_t");
  } else if (r.rc == SPM_OK) {
    return invalid("time_increment " + ti);
  }
  // switch (input.cxx:1427-1630)
  const std::string rot = r.s("switch.ROTATION");
  const std::string ljt = r.s("switch.LJ_truncate");
  const std::string dist = r.s("switch.INIT_distribution.type");
  if (r.rc) return r.rc;
  if (rot == "ON") p.rotation = 1;
  else if (rot == "OFF") p.rotation = 0;
  else return invalid("ROTATION " + rot);
  if (ljt == "ON") p.lj_truncate = SPM_LJ_WCA;
  else if (ljt == "OFF") p.lj_truncate = SPM_LJ_FULL;
  else if (ljt == "NONE") p.lj_truncate = SPM_LJ_NONE;
  else return invalid("LJ_truncate " + ljt);
  static const char *dists[] = {"NONE", "uniform_random", "random_walk", "FCC", "BCC", "user_specify"};
  rn.distribution = -1;
  for (int k = 0; k < 6; ++k)
    if (dist == dists[k]) rn.distribution = k;
  if (rn.distribution < 0) return invalid("INIT_distribution " + dist);
  const char *xyz[3] = {"x", "y", "z"};
  for (int d = 0; d < 3; ++d) {
    const std::string fc = r.s(std::string("switch.FIX_CELL.") + xyz[d]);
    if (r.rc) return r.rc;
    if (fc == "ON") p.fix_cell[d] = 1;
    else if (fc == "OFF") p.fix_cell[d] = 0;
    else return invalid(std::string("FIX_CELL") + xyz[d]);
  }
  {
    char pin[16] = {0};
    if (spm_udf_get_string(u, "switch.pin.type", pin, sizeof(pin)) == SPM_OK && std::strcmp(pin, "YES") == 0)
      return unsupported("pinning");
  }
  // output (input.cxx:1645-1660)
  rn.gts = r.i("output.GTS");
  rn.num_snap = r.i("output.Num_snap");
  if (r.rc) return r.rc;
  // derived time step, Set_global_parameters (input.cxx:312-381)
  if (rn.auto_dt) {
    const double nu = p.eta * (1. / p.rho);  // NU = ETA * IRHO (input.cxx:288-289)
    const int trn[3] = {(p.nx + 2) / 3, (p.ny + 2) / 3, (p.nz + 2) / 3};
    const double L[3] = {p.nx * p.dx, p.ny * p.dx, p.nz * p.dx};
    const double pi2 = 2. * M_PI;
    const double kmax2 = (pi2 / L[0] * trn[0]) * (pi2 / L[0] * trn[0]) + (pi2 / L[1] * trn[1]) * (pi2 / L[1] * trn[1]) +
                         (pi2 / L[2] * trn[2]) * (pi2 / L[2] * trn[2]);
    double tdump = 1. / (nu * kmax2);
    if (rn.equation == SPM_UDF_ELECTROLYTE) {
      const double onsager = el->n_spec == 1 ? el->onsager[0] : std::fmax(el->onsager[0], el->onsager[1]);
      const double diffusion_time = 1. / (p.kbt * onsager * kmax2);
      tdump = std::fmin(tdump, diffusion_time);
      if (el->external_field && el->ac) tdump = std::fmin(tdump, 1.e-2 * (1. / frequency));
    }
    p.dt = rn.dt_factor * tdump;
  }
  p.precision = SPM_FP64;
  *prm = p;
  *run = rn;
  return SPM_OK;
}

namespace {

// glibc's rand() (the TYPE_3 additive feedback generator of random_r:
// r[i] = r[i-31] + r[i-3] mod 2^32 over a state seeded by the Park-Miller
// recurrence 16807 r mod (2^31 - 1), the first 310 outputs discarded, each
// output r >> 1).  KAPSEL draws initial positions and orientations with it
// (RA, RAx, macro.h:75-80); a process that has not called srand runs as
// srand(1).
class GlibcRand {
 public:
  explicit GlibcRand(uint32_t seed = 1) { srand(seed); }
  void srand(uint32_t seed) {
    if (seed == 0) seed = 1;
    r_[0] = (int32_t)seed;
    for (int i = 1; i < 31; ++i) {
      const int64_t hi = r_[i - 1] / 127773, lo = r_[i - 1] % 127773;
      int64_t w = 16807 * lo - 2836 * hi;
      if (w < 0) w += 2147483647;
      r_[i] = (int32_t)w;
    }
    for (int i = 31; i < 34; ++i) r_[i] = r_[i - 31];
    k_ = 0;  // r[34] goes to slot 0
    for (int i = 0; i < 310; ++i) next();
  }
  int rand() { return (int)(next() >> 1); }
  static constexpr double kRandMax = 2147483647.0;
  double RA() { return 2.0 * ((double)rand() - kRandMax * 0.5) / kRandMax; }  // uniform in [-1, 1]
  double RAx(double x) { return (double)rand() / kRandMax * x; }             // uniform in [0, x)

 private:
  int32_t r_[34];  // the last 34 values, r[i] in slot i % 34
  int k_ = 0;      // slot of the next value
  uint32_t next() {
    const uint32_t v = (uint32_t)r_[(k_ + 3) % 34] + (uint32_t)r_[(k_ + 31) % 34];  // r[i-31] + r[i-3]
    r_[k_] = (int32_t)v;
    k_ = (k_ + 1) % 34;
    return v;
  }
};

// RA_on_sphere4D (macro.h:113-123) from two RA_circle draws (macro.h:82-91)
void random_rqtn(GlibcRand &g, double q[4]) {
  double u[4];
  for (int h = 0; h < 2; ++h) {
    do {
      u[2 * h] = g.RA();
      u[2 * h + 1] = g.RA();
    } while (!(u[2 * h] * u[2 * h] + u[2 * h + 1] * u[2 * h + 1] <= 1));
  }
  const double scale = std::sqrt((1.0 - (u[0] * u[0] + u[1] * u[1])) / (u[2] * u[2] + u[3] * u[3]));
  q[0] = u[0];
  q[1] = u[1];
  q[2] = u[2] * scale;
  q[3] = u[3] * scale;
}

int nint(double a) { return a >= 0.0 ? int(a + .5) : int(a - .5); }  // Nint (macro.h:38-46)

// MT19937 (mt19937ar.c: init_genrand, genrand_int32, genrand_real3) and
// Gauss2 (fluct.h:48-60), the generator the reference seeds with
// MT_seed(GIVEN_SEED, 0) (sp_3d_ns.cxx:477) before Init_Particle; FCC and
// random_walk draw their walks from it and the thermostat goes on from the
// state they leave.
class Mt {
 public:
  explicit Mt(uint32_t s) { spm::mt_init(mt_, mti_, s); }
  uint32_t next() { return spm::mt_next(mt_, mti_); }
  double real3() { return spm::mt_real3(mt_, mti_); }
  void gauss2(double *out) { spm::mt_gauss2(mt_, mti_, out); }
  void state(uint32_t *mt, int *mti) const {
    std::memcpy(mt, mt_, sizeof(mt_));
    *mti = mti_;
  }

 private:
  uint32_t mt_[SPM_MT_N];
  int mti_ = 0;
};

// The LJ model of the initial relaxation: Lennard_Jones_f
// (interaction.h:156-181) with LJ_dia and the cutoff A_R_cutoff LJ_dia.
struct InitLj {
  int powers;
  double epsilon, dia, cutoff;
  double f(double x) const {
    double dmy = dia / x;
    double coeff = 24. * epsilon;
    if (powers == 0) {
      dmy = dmy * dmy * (dmy * dmy) * (dmy * dmy);
    } else if (powers == 1) {
      const double d2 = dmy * dmy;
      dmy = d2 * d2 * d2 * d2 * d2 * d2;
      coeff = 48. * epsilon;
    } else {
      const double d2 = dmy * dmy;
      dmy = d2 * d2 * d2 * d2 * d2 * d2 * d2 * d2 * d2;
      coeff = 72. * epsilon;
    }
    return coeff / (x * x) * (2.0 * (dmy * dmy) - dmy);
  }
};

// Calc_f_Lennard_Jones_shear_cap_primitive_lnk (md_force.cxx:23-134) with
// Distance0 and no cap, as Force_random_walk calls it: cells of 16 grid
// spacings, the 27 neighbours of each cell with periodic wrap, pairs i > j
// in the order the cell lists hold them.  Where fewer than 3 cells span an
// axis the reference visits a neighbour cell twice and counts those pairs
// twice; so does this, since it decides where the particles start.
void init_lj_lnk(std::vector<spm_particle> &p, const int N[3], double dx, const double L[3], const InitLj &lj) {
  const int cell_length = 16;
  const int lc[3] = {N[0] / cell_length, N[1] / cell_length, N[2] / cell_length};
  const double lc_r = cell_length * dx;
  const int lcyz = lc[1] * lc[2], lcxyz = lc[0] * lcyz;
  const int n = (int)p.size();
  std::vector<int> head((size_t)lcxyz, -1), lscl((size_t)std::max(n, 1));
  for (int i = 0; i < n; ++i) {
    int mc[3];
    for (int d = 0; d < 3; ++d) mc[d] = int(p[i].x[d] / lc_r);
    const int cn = mc[0] * lcyz + mc[1] * lc[2] + mc[2];
    lscl[i] = head[cn];
    head[cn] = i;
  }
  int ic[3], icl[3], icls[3];
  for (ic[0] = 0; ic[0] < lc[0]; ic[0]++)
    for (ic[1] = 0; ic[1] < lc[1]; ic[1]++)
      for (ic[2] = 0; ic[2] < lc[2]; ic[2]++) {
        const int cn = ic[0] * lcyz + ic[1] * lc[2] + ic[2];
        for (icl[0] = ic[0] - 1; icl[0] <= ic[0] + 1; icl[0]++)
          for (icl[1] = ic[1] - 1; icl[1] <= ic[1] + 1; icl[1]++)
            for (icl[2] = ic[2] - 1; icl[2] <= ic[2] + 1; icl[2]++) {
              for (int d = 0; d < 3; ++d) {
                icls[d] = icl[d];
                if (icls[d] < 0) icls[d] = lc[d] - 1;
                if (icls[d] > lc[d] - 1) icls[d] = 0;
              }
              const int cl = ((icls[0] + lc[0]) % lc[0]) * lcyz + ((icls[1] + lc[1]) % lc[1]) * lc[2] +
                             ((icls[2] + lc[2]) % lc[2]);
              for (int i = head[cn]; i != -1; i = lscl[i])
                for (int j = head[cl]; j != -1; j = lscl[j]) {
                  if (!(i > j)) continue;
                  double x12[3], r2 = 0.0;
                  for (int d = 0; d < 3; ++d) {  // Distance0 (interaction.h:110-123)
                    x12[d] = p[j].x[d] - p[i].x[d];
                    x12[d] -= (double)nint(x12[d] / L[d]) * L[d];
                    r2 += x12[d] * x12[d];
                  }
                  const double r = std::sqrt(r2);
                  if (r < lj.cutoff) {
                    const double dmy = std::min(DBL_MAX / r, lj.f(r));
                    for (int d = 0; d < 3; ++d) {
                      const double dmy1 = dmy * -x12[d];
                      p[i].fr[d] += dmy1;
                      p[j].fr[d] -= dmy1;
                    }
                  }
                }
            }
      }
}

// Random_Walk and Steepest_descent (fluct.h:66-120): every particle moves
// by RADIUS dr_factor, along a Gaussian direction or along its force (fr
// accumulates over the calls, as in the reference), wrapped into the box.
void random_walk(std::vector<spm_particle> &p, Mt &mt, double radius, double dr_factor, const double L[3]) {
  const double dr = radius * dr_factor;
  for (auto &q : p) {
    double dmy[4];
    mt.gauss2(dmy);
    mt.gauss2(dmy + 2);
    double h = 0.;
    for (int d = 0; d < 3; ++d) h += dmy[d] * dmy[d];
    h = std::sqrt(h);
    if (h > 0.) h = dr / h;
    for (int d = 0; d < 3; ++d) {
      q.x[d] += h * dmy[d];
      q.x[d] -= std::floor(q.x[d] * (1. / L[d])) * L[d];
    }
  }
}

void steepest_descent(std::vector<spm_particle> &p, double radius, const int N[3], double dx, const double L[3],
                      const InitLj &lj) {
  const double dr = radius * 0.5e-1;
  init_lj_lnk(p, N, dx, L, lj);
  for (auto &q : p) {
    double h = 0.;
    for (int d = 0; d < 3; ++d) h += q.fr[d] * q.fr[d];
    h = std::sqrt(h);
    if (h > 0.) h = dr / h;
    for (int d = 0; d < 3; ++d) {
      q.x[d] += h * q.fr[d];
      q.x[d] -= std::floor(q.x[d] * (1. / L[d])) * L[d];
    }
  }
}

// FCC placement (init_particle.cxx:57-150): nn cells per axis chosen for
// the smallest anisotropy, four particles per cell, a last partial layer
// stretched along z.
int fcc_lattice(std::vector<spm_particle> &p, const double L[3], double sigma) {
  const int n = (int)p.size();
  double lratio[3];
  const double min_l = std::min(L[0], std::min(L[1], L[2]));
  for (int d = 0; d < 3; ++d) lratio[d] = L[d] / min_l;
  const double dmy = std::pow((double)n / (4. * lratio[0] * lratio[1] * lratio[2]), 1. / 3);
  int nn_base = (int)std::ceil(dmy);
  const int nn_up = (int)std::ceil(dmy), nn_low = (int)dmy;
  const int nxny_up = 2 * nn_up * nn_up * (int)lratio[0] * (int)lratio[1];
  const int nxny_low = 2 * nn_low * nn_low * (int)lratio[0] * (int)lratio[1];
  const int nz_up = nxny_up > 0 ? (int)std::ceil((double)n / nxny_up) : 0;
  const int nz_low = nxny_low > 0 ? (int)std::ceil((double)n / nxny_low) : 0;
  const double density_xy_up = std::sqrt((double)nxny_up / (L[0] * L[1]));
  const double density_xy_low = std::sqrt((double)nxny_low / (L[0] * L[1]));
  const double density_z_up = nz_up / L[2], density_z_low = nz_low / L[2];
  const double skew_up = std::fabs(density_xy_up / density_z_up - 1.);
  const double skew_low = std::fabs(density_xy_low / density_z_low - 1.);
  if (nxny_low > 0 && nz_low > 0 && skew_low <= skew_up)
    nn_base = nn_low;
  else if (nxny_up > 0 && nz_up > 0)
    nn_base = nn_up;
  int nn[3];
  for (int d = 0; d < 3; ++d) nn[d] = (int)(nn_base * std::ceil(lratio[d]));
  const bool just_packed = n - 4 * nn[0] * nn[1] * nn[2] == 0;
  int nz = 0;
  if (!just_packed) nz = (int)std::ceil((double)n / (2 * nn[0] * nn[1]));
  double lattice[3], origin[3];
  for (int d = 0; d < 3; ++d) {
    lattice[d] = L[d] / nn[d];
    if (!just_packed && d == 2) lattice[d] = L[d] / (nz * .5);
    origin[d] = lattice[d] * .25;
    if (lattice[d] < sigma * std::sqrt(2.))
      return fail(SPM_ERR_INVALID, "FCC: beyond closed packing along axis " + std::to_string(d));
  }
  const int zlayer = 2 * nn[0] * nn[1];
  for (int i = 0; i < n; ++i) {
    const int ix = i % nn[0], iy = (i % zlayer) / nn[0], iz = i / zlayer;
    p[i].x[0] = origin[0] + (double)ix * lattice[0] + (lattice[0] / 2.0) * ((iy + iz) % 2);
    p[i].x[1] = origin[1] + (double)iy * lattice[1] / 2.0;
    p[i].x[2] = origin[2] + (double)iz * lattice[2] / 2.0;
  }
  return SPM_OK;
}

// Init_Particle (init_particle.cxx:12-393) for spherical particles; mt and
// mti (when given) receive the MT19937 state it leaves.
int init_particles(const spm_udf *u, spm_particle *p, int n, bool user_only, uint32_t *mt_out = nullptr,
                   int *mti_out = nullptr) {
  if (!u || (!p && n > 0)) return fail(SPM_ERR_INVALID, "null argument");
  Reader r{u};
  const std::string dist = r.s("switch.INIT_distribution.type");
  const std::string rot = r.s("switch.ROTATION");
  const std::string orient = r.s("switch.INIT_orientation");
  const int nspec = r.size("object_type.spherical_particle.Particle_spec[]");
  const double A = r.d("A");
  int np[3] = {r.i("mesh.NPX"), r.i("mesh.NPY"), r.i("mesh.NPZ")};
  double DX = 0.0;
  {
    char eq[64] = {0};
    if (spm_udf_get_string(u, "constitutive_eq.type", eq, sizeof(eq)) == SPM_OK)
      DX = r.d(std::string("constitutive_eq.") + eq + ".DX");
  }
  if (r.rc) return r.rc;
  if (user_only && dist != "user_specify")
    return fail(SPM_ERR_INVALID, "INIT_distribution is " + dist + ", not user_specify");
  if (orient != "user_specify" && orient != "random" && orient != "space_align")
    return fail(SPM_ERR_INVALID, "invalid INIT_orientation " + orient);
  // species by Particle_spec counts in order (init_particle.cxx:351-357)
  std::vector<int> spec;
  for (int s = 0; s < nspec; ++s) {
    const int c = r.i("object_type.spherical_particle.Particle_spec[" + std::to_string(s) + "].Particle_number");
    if (r.rc) return r.rc;
    if (c < 0) return fail(SPM_ERR_INVALID, "negative Particle_number");
    spec.insert(spec.end(), c, s);
  }
  if ((int)spec.size() != n) return fail(SPM_ERR_INVALID, "n differs from the Particle_number total");
  double L[3];
  for (int d = 0; d < 3; ++d) {
    if (np[d] < 0 || np[d] > 12) return fail(SPM_ERR_INVALID, "invalid mesh");
    L[d] = (1 << np[d]) * DX;  // L_particle = L = N * DX (input.cxx:304-311)
  }
  const double sigma = 2.0 * (A * DX);  // SIGMA = 2 RADIUS (input.cxx:1339-1340)
  GlibcRand g;
  std::vector<spm_particle> out((size_t)n);
  std::memset(out.data(), 0, sizeof(spm_particle) * out.size());
  for (auto &q : out) q.q.s = 1.0;
  const bool rotation = rot == "ON";
  if (dist == "user_specify") {  // init_particle.cxx:226-288
    const std::string b0 = "switch.INIT_distribution.user_specify.Particles";
    const int m = r.size(b0 + "[]");
    if (r.rc) return r.rc;
    if (m != n) return fail(SPM_ERR_INVALID, "Particles[] holds " + std::to_string(m) + " entries");
    for (int k = 0; k < n; ++k) {
      const std::string b = b0 + "[" + std::to_string(k) + "].";
      const char *c[3] = {"x", "y", "z"};
      for (int d = 0; d < 3; ++d) {
        out[k].x[d] = r.d(b + "R." + c[d]);
        // v and omega are read, then set to 0.0*RA() below
        (void)r.d(b + "v." + c[d]);
        (void)r.d(b + "omega." + c[d]);
      }
      double q[4] = {r.d(b + "q.q0"), r.d(b + "q.q1"), r.d(b + "q.q2"), r.d(b + "q.q3")};
      if (r.rc) return r.rc;
      if (orient == "user_specify") {  // qtn_normalize (quaternion.h:237-245)
        const double norm = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        if (!(norm > 0.0)) return fail(SPM_ERR_INVALID, "zero quaternion in Particles[" + std::to_string(k) + "]");
        const double inv = 1.0 / norm;
        for (double &x : q) x *= inv;
        out[k].q.s = q[0];
        for (int d = 0; d < 3; ++d) out[k].q.v[d] = q[d + 1];
      }
    }
  } else if (dist == "uniform_random") {  // init_particle.cxx:27-48
    const double overlap = sigma * 1.05;
    const long budget = 1000000L + 1000L * n;  // draws before giving up
    long draws = 0;
    for (int i = 0; i < n; ++i) {
      for (;;) {
        if (++draws > budget) return fail(SPM_ERR_INVALID, "uniform_random: no room for the particles");
        for (int d = 0; d < 3; ++d) out[i].x[d] = g.RAx(L[d]);
        int j = 0;
        for (; j < i; ++j) {  // Distance (interaction.h:110-135)
          double s2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            double x12 = out[j].x[d] - out[i].x[d];
            x12 -= (double)nint(x12 / L[d]) * L[d];
            s2 += x12 * x12;
          }
          if (std::sqrt(s2) <= overlap) break;
        }
        if (j >= i) break;
      }
    }
  } else if (dist == "BCC") {  // init_particle.cxx:291-316
    const int nn = (int)std::ceil(std::pow((double)n / 2., 1. / 3));
    const double ax = L[0] / (double)nn, ay = L[1] / (double)nn, az = L[2] / (double)nn;
    for (int i = 0; i < n; ++i) {
      const int ix = i % nn, iy = (i % (nn * nn)) / nn, iz = i / (nn * nn);
      out[i].x[0] = ax / 4.0 + (double)ix * ax;
      out[i].x[1] = ay / 4.0 + (double)iy * ay;
      out[i].x[2] = az / 4.0 + (double)iz * az / 2.0;
      if (iz % 2 == 1) {
        out[i].x[0] = out[i].x[0] + ax / 2.0;
        out[i].x[1] = out[i].x[1] + ax / 2.0;  // ax, as the reference has it
      }
    }
  } else if (dist == "FCC" || dist == "random_walk") {  // init_particle.cxx:50-224
    // the LJ model the relaxation uses: LJ_dia = SIGMA, or under shear
    // min((SIGMA + XI) / 2^(1/p), SIGMA) (input.cxx:389-403)
    const std::string ljp = r.s("LJ_powers");
    const double eps = r.d("EPSILON"), a_xi = r.d("A_XI");
    int iters = 0;
    if (dist == "random_walk") iters = r.i("switch.INIT_distribution.random_walk.iteration");
    char eq[64] = {0};
    (void)spm_udf_get_string(u, "constitutive_eq.type", eq, sizeof(eq));
    if (r.rc) return r.rc;
    InitLj lj;
    lj.powers = ljp == "12:6" ? 0 : ljp == "24:12" ? 1 : ljp == "36:18" ? 2 : -1;
    if (lj.powers < 0) return fail(SPM_ERR_INVALID, "invalid LJ_powers " + ljp);
    if (iters < 0) return fail(SPM_ERR_INVALID, "negative random_walk iteration");
    const int N[3] = {1 << np[0], 1 << np[1], 1 << np[2]};
    if (N[0] < 16 || N[1] < 16 || N[2] < 16)
      return unsupported("FCC / random_walk on a box under 16 grid points per axis (its LJ cells)");
    const double radius = A * DX;
    const double pw[3] = {std::pow(2., 1. / 6.), std::pow(2., 1. / 12.), std::pow(2., 1. / 18.)};
    lj.epsilon = eps;
    lj.dia = sigma;
    if (std::string(eq) == "Shear_Navier_Stokes" || std::string(eq) == "Shear_Navier_Stokes_Lees_Edwards")
      lj.dia = std::min((sigma + a_xi * DX) / pw[lj.powers], sigma);
    if (int rc = fcc_lattice(out, L, sigma)) return rc;
    for (auto &q : out)
      for (int d = 0; d < 3; ++d) q.fr[d] = q.fr_previous[d] = 0.0;
    Mt mt(0);
    lj.cutoff = pw[lj.powers] * lj.dia;  // A_R_cutoff by LJ_powers while the lattice settles
    random_walk(out, mt, radius, 1.e-3, L);
    steepest_descent(out, radius, N, DX, L, lj);
    // (the reference repeats the pair while a particle lies outside [0, L)
    // in y or z; both moves wrap into the box, so it runs once)
    if (dist == "random_walk") {
      for (auto &q : out)
        for (int d = 0; d < 3; ++d) q.fr[d] = q.fr_previous[d] = 0.0;
      lj.cutoff = pw[0] * lj.dia;  // A_R_cutoff = 2^(1/6)
      for (int it = 0; it < iters; ++it) {
        random_walk(out, mt, radius, 0.5e-1 * .5, L);
        steepest_descent(out, radius, N, DX, L, lj);
      }
    }
    for (auto &q : out)
      for (int d = 0; d < 3; ++d) q.fr[d] = q.fr_previous[d] = 0.0;  // init_particle.cxx:369-370
    if (mt_out) mt.state(mt_out, mti_out);
  } else {
    return unsupported("INIT_distribution " + dist);
  }
  if (mt_out && dist != "FCC" && dist != "random_walk") Mt(0).state(mt_out, mti_out);
  // orientation (init_particle.cxx:318-345)
  for (auto &q : out) {
    if (rotation && orient == "random") {
      double v[4];
      random_rqtn(g, v);
      q.q.s = v[0];
      for (int d = 0; d < 3; ++d) q.q.v[d] = v[d + 1];
    } else if (!(rotation && orient == "user_specify" && dist == "user_specify")) {
      q.q.s = 1.0;
      q.q.v[0] = q.q.v[1] = q.q.v[2] = 0.0;
    }
    q.q_old = q.q;
  }
  // species, velocity 0*RA(), angular velocity 0*RA() (init_particle.cxx:349-393)
  for (int k = 0; k < n; ++k) {
    out[k].spec = spec[k];
    for (int d = 0; d < 3; ++d) out[k].x_nopbc[d] = out[k].x[d];
  }
  std::copy(out.begin(), out.end(), p);
  return SPM_OK;
}

}  // namespace

extern "C" int spm_udf_user_particles(const spm_udf *u, spm_particle *p, int n) {
  return init_particles(u, p, n, true);
}

extern "C" int spm_udf_init_particles(const spm_udf *u, spm_particle *p, int n) {
  return init_particles(u, p, n, false);
}

extern "C" int spm_udf_init_particles_rng(const spm_udf *u, spm_particle *p, int n, uint32_t mt[SPM_MT_N],
                                          int *mti) {
  if (!mt || !mti) return fail(SPM_ERR_INVALID, "null argument");
  return init_particles(u, p, n, false, mt, mti);
}

// ---------------------------------------------------------------------
// restart file: Save_Restart_udf / Force_restore_parameters
// (resume.cxx:10-78, 274-364; Read_Particle_udf and Save_Particle_udf)
// ---------------------------------------------------------------------
namespace {

void put_num(std::string &o, double v) {
  char b[32];
  std::snprintf(b, sizeof(b), "%.17g", v);
  o += b;
}

void serialize(std::string &o, const Value &v) {
  switch (v.kind) {
    case K_INT: {
      char b[32];
      std::snprintf(b, sizeof(b), "%.0f", v.num);
      o += b;
      break;
    }
    case K_REAL:
      put_num(o, v.num);
      break;
    case K_STRING:
    case K_SELECT:
      o += '"';
      o += v.str;
      o += '"';
      break;
    case K_RECORD:
      o += '{';
      for (size_t k = 0; k < v.fields.size(); ++k) {
        if (k) o += ',';
        serialize(o, v.fields[k].second);
      }
      o += '}';
      break;
    case K_ARRAY:
      o += '[';
      for (size_t k = 0; k < v.items.size(); ++k) {
        if (k) o += ',';
        serialize(o, v.items[k]);
      }
      o += ']';
      break;
    default:
      break;
  }
}

// sParticle member order (Udf/define.udf restart schema)
const char *const kSaved3[] = {"R", "v", "v_old", "f_hydro", "f_hydro_previous", "f_hydro1", "f_slip",
                               "f_slip_previous", "fr", "fr_previous", "omega", "omega_old", "torque_hydro",
                               "torque_hydro_previous", "torque_hydro1", "torque_slip", "torque_slip_previous"};

const double *saved3(const spm_particle &p, int k) {
  const double *const f[] = {p.x, p.v, p.v_old, p.f_hydro, p.f_hydro_previous, p.f_hydro1, p.f_slip,
                             p.f_slip_previous, p.fr, p.fr_previous, p.omega, p.omega_old, p.torque_hydro,
                             p.torque_hydro_previous, p.torque_hydro1, p.torque_slip, p.torque_slip_previous};
  return f[k];
}

double *saved3(spm_particle &p, int k) { return const_cast<double *>(saved3(static_cast<const spm_particle &>(p), k)); }

int check_restart(const spm_restart *r) {
  if (!r || r->nx <= 0 || r->ny <= 0 || r->nz <= 0 || r->n_particles < 0 || r->n_ion < 0 || r->n_ion > 2)
    return fail(SPM_ERR_INVALID, "invalid spm_restart sizes");
  if (!r->zeta0 || !r->zeta1 || (r->n_particles > 0 && !r->particles) || (r->n_ion > 0 && !r->concentration))
    return fail(SPM_ERR_INVALID, "spm_restart buffer missing");
  return SPM_OK;
}

const Value *member(const Value *v, const char *name) {
  if (!v || v->kind != K_RECORD) return nullptr;
  for (auto &f : v->fields)
    if (f.first == name) return &f.second;
  return nullptr;
}

bool num3(const Value *v, double *out) {
  const char *c[3] = {"x", "y", "z"};
  for (int d = 0; d < 3; ++d) {
    const Value *x = member(v, c[d]);
    if (!x || (x->kind != K_REAL && x->kind != K_INT)) return false;
    out[d] = x->num;
  }
  return true;
}

bool quat(const Value *v, double *out) {
  const char *c[4] = {"q0", "q1", "q2", "q3"};
  for (int d = 0; d < 4; ++d) {
    const Value *x = member(v, c[d]);
    if (!x || (x->kind != K_REAL && x->kind != K_INT)) return false;
    out[d] = x->num;
  }
  return true;
}

}  // namespace

extern "C" int spm_udf_write_restart(const char *in_path, const char *out_path, const spm_restart *r) {
  if (!in_path || !out_path) return fail(SPM_ERR_INVALID, "null argument");
  if (int rc = check_restart(r)) return rc;
  spm_udf *u = nullptr;
  if (int rc = spm_udf_open(in_path, &u)) return rc;
  std::unique_ptr<spm_udf> hold(u);
  bool has_resume = false;
  for (auto &e : u->doc.entries) has_resume |= e.first == "resume";
  if (!has_resume) return fail(SPM_ERR_INVALID, "the schema has no resume entry");
  std::string text, err;
  if (!inline_includes(in_path, text, err)) return fail(SPM_ERR_INVALID, err);
  // the data section proper, not the one inside \begin{header}
  const size_t hdr = text.find("\\end{header}");
  const size_t b = text.find("\\begin{data}", hdr == std::string::npos ? 0 : hdr);
  if (b == std::string::npos) return fail(SPM_ERR_INVALID, "no data section");
  std::string o = text.substr(0, b);
  o += "\\begin{data}\n";
  for (auto &e : u->doc.entries) {
    if (e.first == "resume") continue;
    auto it = u->doc.data.find(e.first);
    if (it == u->doc.data.end()) continue;
    o += e.first + ":";
    serialize(o, it->second);
    o += '\n';
  }
  // resume:{"CONTINUE",{Saved_Data:{jikan, Particles[], GR_body[], Zeta[][][],
  //         uk_dc, Concentration[][][][], oblique}}}
  o += "resume:{\"CONTINUE\",{{\n{" + std::to_string(r->ts) + ",";
  put_num(o, r->time);
  o += "}\n[";
  for (int j = 0; j < r->n_particles; ++j) {
    const spm_particle &p = r->particles[j];
    o += j ? ",\n{" : "\n{";
    for (int k = 0; k < 17; ++k) put3(o, saved3(p, k));
    char qb[128];
    std::snprintf(qb, sizeof(qb), "{%.17g,%.17g,%.17g,%.17g}", p.q.s, p.q.v[0], p.q.v[1], p.q.v[2]);
    o += qb;
    std::snprintf(qb, sizeof(qb), "{%.17g,%.17g,%.17g,%.17g}", p.q_old.s, p.q_old.v[0], p.q_old.v[1],
                  p.q_old.v[2]);
    o += qb;
    o += '}';
  }
  o += "]\n[]\n[";
  const int nz_ = 2 * (r->nz / 2 + 1);
  for (int i = 0; i < r->nx; ++i) {
    o += i ? ",\n[" : "\n[";
    for (int j = 0; j < r->ny; ++j) {
      o += j ? ",[" : "[";
      for (int k = 0; k < nz_; ++k) {
        const size_t im = ((size_t)i * r->ny + j) * nz_ + k;
        o += k ? ",{" : "{";
        put_num(o, r->zeta0[im]);
        o += ',';
        put_num(o, r->zeta1[im]);
        o += '}';
      }
      o += ']';
    }
    o += ']';
  }
  o += "]\n";
  put3(o, r->uk_dc);
  o += "\n[";
  const size_t nf = (size_t)r->nx * r->ny * nz_;
  for (int n = 0; n < r->n_ion; ++n) {
    o += n ? ",[" : "[";
    for (int i = 0; i < r->nx; ++i) {
      o += i ? ",\n[" : "\n[";
      for (int j = 0; j < r->ny; ++j) {
        o += j ? ",[" : "[";
        for (int k = 0; k < nz_; ++k) {
          o += k ? ",{" : "{";
          put_num(o, r->concentration[n * nf + ((size_t)i * r->ny + j) * nz_ + k]);
          o += '}';
        }
        o += ']';
      }
      o += ']';
    }
    o += ']';
  }
  o += "]\n{";
  put_num(o, r->degree_oblique);
  o += "}\n}}}\n\\end{data}\n";
  const std::string tmp = std::string(out_path) + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary | std::ios::trunc);
    if (!f) return fail(SPM_ERR_INVALID, "cannot write " + tmp);
    f << o;
    f.close();
    if (!f) return fail(SPM_ERR_INVALID, "write failed on " + tmp);
  }
  if (std::rename(tmp.c_str(), out_path) != 0) return fail(SPM_ERR_INVALID, std::string("cannot rename to ") + out_path);
  return SPM_OK;
}

extern "C" int spm_udf_read_restart(const spm_udf *u, spm_restart *r) {
  if (!u) return fail(SPM_ERR_INVALID, "null argument");
  if (int rc = check_restart(r)) return rc;
  char calc[16] = {0};
  if (spm_udf_get_string(u, "resume.Calculation", calc, sizeof(calc)) != SPM_OK || std::strcmp(calc, "CONTINUE"))
    return fail(SPM_ERR_INVALID, "resume.Calculation is not CONTINUE");
  const Value *sd = find(u->doc, -1, "resume.CONTINUE.Saved_Data");
  const Value *jikan = member(sd, "jikan");
  const Value *parts = member(sd, "Particles");
  const Value *zeta = member(sd, "Zeta");
  const Value *ukdc = member(sd, "uk_dc");
  const Value *conc = member(sd, "Concentration");
  const Value *obl = member(sd, "oblique");
  const Value *ts = member(jikan, "ts"), *tm = member(jikan, "time"), *deg = member(obl, "degree_oblique");
  if (!ts || !tm || !parts || !zeta || !ukdc || !conc || !deg)
    return fail(SPM_ERR_INVALID, "resume.CONTINUE.Saved_Data is incomplete");
  const int nz_ = 2 * (r->nz / 2 + 1);
  // sizes first, so that a mismatch leaves the caller's buffers untouched
  if ((int)parts->items.size() != r->n_particles) return fail(SPM_ERR_INVALID, "Particles[] size differs");
  if ((int)zeta->items.size() != r->nx) return fail(SPM_ERR_INVALID, "Zeta[] size differs from nx");
  for (auto &a : zeta->items) {
    if ((int)a.items.size() != r->ny) return fail(SPM_ERR_INVALID, "Zeta[][] size differs from ny");
    for (auto &b : a.items)
      if ((int)b.items.size() != nz_) return fail(SPM_ERR_INVALID, "Zeta[][][] size differs from 2*(nz/2+1)");
  }
  if ((int)conc->items.size() != r->n_ion) return fail(SPM_ERR_INVALID, "Concentration[] size differs");
  for (auto &n : conc->items) {
    if ((int)n.items.size() != r->nx) return fail(SPM_ERR_INVALID, "Concentration[][] size differs");
    for (auto &a : n.items) {
      if ((int)a.items.size() != r->ny) return fail(SPM_ERR_INVALID, "Concentration[][][] size differs");
      for (auto &b : a.items)
        if ((int)b.items.size() != nz_) return fail(SPM_ERR_INVALID, "Concentration[][][][] size differs");
    }
  }
  // particles (Read_Particle_udf, resume.cxx:366-460)
  std::vector<spm_particle> ps(r->particles, r->particles + r->n_particles);
  for (int j = 0; j < r->n_particles; ++j) {
    const Value *pv = &parts->items[j];
    spm_particle &p = ps[j];
    for (int k = 0; k < 17; ++k)
      if (!num3(member(pv, kSaved3[k]), saved3(p, k))) return fail(SPM_ERR_INVALID, "bad Particles[] entry");
    double q[4], qo[4];
    if (!quat(member(pv, "q"), q) || !quat(member(pv, "q_old"), qo))
      return fail(SPM_ERR_INVALID, "bad Particles[] quaternion");
    for (double *a : {q, qo}) {  // qtn_normalize (quaternion.h:237-245)
      const double norm = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
      if (!(norm > 0.0)) return fail(SPM_ERR_INVALID, "zero quaternion in Particles[]");
      const double inv = 1.0 / norm;
      for (int d = 0; d < 4; ++d) a[d] *= inv;
    }
    p.q.s = q[0];
    p.q_old.s = qo[0];
    for (int d = 0; d < 3; ++d) {
      p.q.v[d] = q[d + 1];
      p.q_old.v[d] = qo[d + 1];
      p.x_previous[d] = p.x[d];
      p.x_nopbc[d] = p.x[d];
      p.v_slip[d] = 0.0;
      p.omega_slip[d] = 0.0;
      p.mass_center[d] = p.surface_mass_center[d] = p.surface_dv[d] = p.surface_dw[d] = 0.0;
      for (int l = 0; l < 3; ++l) p.inertia[d][l] = p.surface_inertia[d][l] = 0.0;
    }
    p.mass = p.surface_mass = 0.0;
  }
  double dc[3];
  if (!num3(ukdc, dc)) return fail(SPM_ERR_INVALID, "bad uk_dc");
  // fields (resume.cxx:280-301)
  const size_t nf = (size_t)r->nx * r->ny * nz_;
  std::vector<double> z0(nf), z1(nf), cc((size_t)r->n_ion * nf);
  for (int i = 0; i < r->nx; ++i)
    for (int j = 0; j < r->ny; ++j)
      for (int k = 0; k < nz_; ++k) {
        const Value &e = zeta->items[i].items[j].items[k];
        const Value *a = member(&e, "zeta0"), *b = member(&e, "zeta1");
        if (!a || !b || a->kind != K_REAL || b->kind != K_REAL) return fail(SPM_ERR_INVALID, "bad Zeta entry");
        const size_t im = ((size_t)i * r->ny + j) * nz_ + k;
        z0[im] = a->num;
        z1[im] = b->num;
        for (int n = 0; n < r->n_ion; ++n) {
          const Value *ck = member(&conc->items[n].items[i].items[j].items[k], "ck");
          if (!ck || ck->kind != K_REAL) return fail(SPM_ERR_INVALID, "bad Concentration entry");
          cc[n * nf + im] = ck->num;
        }
      }
  std::copy(ps.begin(), ps.end(), r->particles);
  std::copy(z0.begin(), z0.end(), r->zeta0);
  std::copy(z1.begin(), z1.end(), r->zeta1);
  if (r->n_ion > 0) std::copy(cc.begin(), cc.end(), r->concentration);
  for (int d = 0; d < 3; ++d) r->uk_dc[d] = dc[d];
  r->ts = (int)ts->num;
  r->time = tm->num;
  r->degree_oblique = deg->num;
  return SPM_OK;
}
