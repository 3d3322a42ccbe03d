// msimio.cpp -- the msim input and restart formats on the host side of the
// nonbonded path, so that msim-style input drives the engine:
//   * the input preprocessor (PreprocessedStream, prepro.cpp:23-318):
//     '#' comments, $name substitution, !define, !include / <file, !path,
//     !foreach, !repeat, !ignore ... !end;
//   * ElecLJ's parameters in the object syntax that the io generator gives
//     every class (io:94-312): "{ field value ... }", with the fields of
//     eleclj.hpp:310-327 and the constructor's defaults (eleclj.cpp:12-40);
//     a deck holds them as "potential { eleclj { ... } }" inside the
//     program's braces (msimprog.cpp:40-43, 665-674);
//   * the per-type site parameters ElecLJ::types_changed derives from
//     fixed_charge, sigma, epsilon and general_type (eleclj.cpp:440-500);
//   * xyz coordinate / restart files (MSys::read_as_xyz,
//     msys.cpp:823-868; write_as_xyz, msys.cpp:916-932; Atom::write_as_xyz,
//     atom.cpp:129-134).
// The molecular system's own block (atoms, types, bonds) is not read here:
// callers give atom types and bonds (mm_derive_sites) themselves.
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <map>
#include <string>
#include <unistd.h>
#include <vector>

#include "molmodel_amd.h"

namespace mmamd {
void set_last_error(const std::string &s);
}
using mmamd::set_last_error;

namespace {

bool is_directory(const std::string &f) {
  DIR *d = opendir(f.c_str());
  if (d) closedir(d);
  return d != nullptr;
}

// ---------------------------------------------------------------- preprocessor

struct Prepro {
  std::map<std::string, std::string> var;
  std::vector<std::string> path;
  std::string out;
  int depth = 0;

  // the file as given, else in each search-path directory (out.cpp:33-52)
  bool open_lines(const std::string &fname, std::vector<std::string> &lines, std::string &err) {
    std::vector<std::string> cand{fname};
    for (auto &d : path) cand.push_back(d + "/" + fname);
    for (auto &c : cand) {
      FILE *f = fopen(c.c_str(), "r");
      if (!f) continue;
      if (is_directory(c)) { fclose(f); continue; }
      char buf[1024];
      lines.clear();
      while (fgets(buf, sizeof buf, f)) lines.emplace_back(buf);
      fclose(f);
      return true;
    }
    err = "cannot open " + fname;
    return false;
  }

  bool read_fname(const std::string &fname, std::string &err) {
    std::vector<std::string> lines;
    if (fname.empty()) { err = "expecting a filename"; return false; }
    if (!open_lines(fname, lines, err)) return false;
    if (++depth > 64) { err = fname + ": files included too deeply"; return false; }
    size_t next;
    const bool ok = read_lines(lines, 0, fname, false, next, err);
    depth--;
    return ok;
  }

  static std::string where(const std::string &f, size_t i) {
    return f + ":" + std::to_string(i + 1) + ": ";
  }

  // lines [start, ...) up to the end of the file, or through the matching
  // !end when expecting_end; next = the line after it
  bool read_lines(const std::vector<std::string> &lines, size_t start, const std::string &fname,
                  bool expecting_end, size_t &next, std::string &err) {
    bool ignore = false;
    for (size_t i = start; i < lines.size(); i++) {
      std::string ln = lines[i];
      if (ln.size() >= 1023 && ln.back() != '\n') { err = where(fname, i) + "line too long"; return false; }
      const size_t hash = ln.find('#');
      if (hash != std::string::npos) ln = ln.substr(0, hash) + "\n";   // comment
      size_t c = 0;
      while (c < ln.size() && isspace((unsigned char)ln[c])) c++;
      auto at = [&](const char *w) { return ln.compare(c, strlen(w), w) == 0; };
      if (at("!ignore")) ignore = true;
      if (ignore) {
        if (!at("!end")) continue;
        // the rest of the line after the !end is an ordinary line
        ignore = false;
        for (size_t k = 0; k < c + 4; k++) ln[k] = ' ';
        if (!read_line(ln, fname, i, err)) return false;
        continue;
      }
      if (c >= ln.size() || ln[c] != '!' || at("!define") || at("!parallel_define") ||
          at("!include") || at("!path")) {
        if (!read_line(ln, fname, i, err)) return false;   // ordinary line, define, include
        continue;
      }
      if (at("!end")) {
        if (!expecting_end) { err = where(fname, i) + "unexpected !end"; return false; }
        next = i + 1;
        return true;
      }
      if (at("!foreach")) {
        // !foreach v { values ... } (a single value needs no braces)
        std::vector<std::string> tok;
        if (!words(ln.substr(c + 8), tok) || tok.size() < 2) {
          err = where(fname, i) + "expecting !foreach var { values ... }";
          return false;
        }
        const std::string v = tok[0];
        std::vector<std::string> vals;
        if (tok[1] == "{") {
          size_t k = 2;
          for (; k < tok.size() && tok[k] != "}"; k++) vals.push_back(tok[k]);
          if (k == tok.size()) vals.clear();
        } else {
          vals.push_back(tok[1]);
        }
        if (vals.empty()) { err = where(fname, i) + "expecting !foreach var { values ... }"; return false; }
        size_t end = i + 1;
        for (auto &x : vals) {
          var[v] = x;
          if (!read_lines(lines, i + 1, fname, true, end, err)) return false;
        }
        var[v] = "";
        i = end - 1;
      } else if (at("!repeat")) {
        char *e = nullptr;
        const std::string arg = ln.substr(c + 7);
        const long n = strtol(arg.c_str(), &e, 10);
        if (e == arg.c_str() || n <= 0) { err = where(fname, i) + "expecting !repeat n where n > 0"; return false; }
        size_t end = i + 1;
        for (long r = 0; r < n; r++)
          if (!read_lines(lines, i + 1, fname, true, end, err)) return false;
        i = end - 1;
      } else {
        err = where(fname, i) + "unknown command: " + ln;
        return false;
      }
    }
    if (expecting_end) { err = fname + ": expecting !end, reached end of file"; return false; }
    next = lines.size();
    return true;
  }

  // whitespace-separated words, with '{' and '}' words of their own
  static bool words(const std::string &s, std::vector<std::string> &w) {
    std::string cur;
    for (char ch : s) {
      if (isspace((unsigned char)ch) || ch == '{' || ch == '}') {
        if (!cur.empty()) w.push_back(cur), cur.clear();
        if (ch == '{' || ch == '}') w.push_back(std::string(1, ch));
      } else {
        cur += ch;
      }
    }
    if (!cur.empty()) w.push_back(cur);
    return true;
  }

  bool read_line(const std::string &ln0, const std::string &fname, size_t nline, std::string &err) {
    std::string ln = ln0;
    if (depth > 1000) { err = where(fname, nline) + "variable substitution does not end"; return false; }
    // $name: defined variables first, then the environment; the line is
    // processed again after each substitution
    size_t d = ln.find('$');
    if (d != std::string::npos) {
      size_t e = d + 1;
      while (e < ln.size() && (isalnum((unsigned char)ln[e]) || ln[e] == '_')) e++;
      const std::string name = ln.substr(d + 1, e - d - 1);
      if (name.empty()) { err = where(fname, nline) + "no variable name after '$'"; return false; }
      if (name.size() >= 256) { err = where(fname, nline) + "variable name too long"; return false; }
      std::string val;
      auto it = var.find(name);
      if (it != var.end() && !it->second.empty()) val = it->second;
      else if (const char *env = getenv(name.c_str())) val = env;
      else { err = where(fname, nline) + "undefined variable: '" + name + "'"; return false; }
      depth++;
      const bool ok = read_line(ln.substr(0, d) + val + ln.substr(e), fname, nline, err);
      depth--;
      return ok;
    }
    size_t p = ln.find("!path");
    if (p != std::string::npos) {
      if (!read_line(ln.substr(0, p) + "\n", fname, nline, err)) return false;
      // the rest of the line, without surrounding whitespace, is the
      // directory (the reference keeps the whitespace in the name)
      std::string dir = ln.substr(p + 5);
      size_t a = 0, b = dir.size();
      while (a < b && isspace((unsigned char)dir[a])) a++;
      while (b > a && isspace((unsigned char)dir[b - 1])) b--;
      path.push_back(dir.substr(a, b - a));
      return true;
    }
    p = ln.find("!define");
    if (p != std::string::npos) {
      if (!read_line(ln.substr(0, p) + "\n", fname, nline, err)) return false;
      size_t a = p + 7;
      while (a < ln.size() && (ln[a] == ' ' || ln[a] == '\t' || ln[a] == '\n')) a++;
      size_t b = a;
      while (b < ln.size() && !(ln[b] == ' ' || ln[b] == '\t' || ln[b] == '\n')) b++;
      const std::string name = ln.substr(a, b - a);
      if (name.empty()) { err = where(fname, nline) + "no variable name after !define"; return false; }
      for (char ch : name)
        if (!(isalnum((unsigned char)ch) || ch == '_')) {
          err = where(fname, nline) + "illegal variable name: " + name;
          return false;
        }
      // the value starts after the one character that ends the name
      std::string val = b < ln.size() ? ln.substr(b + 1) : "";
      while (!val.empty() && isspace((unsigned char)val.back())) val.pop_back();
      var[name] = val;
      return true;
    }
    if (ln.find("!parallel_define") != std::string::npos) {
      err = where(fname, nline) + "parallel_define needs MPI ranks, which this library does not have";
      return false;
    }
    // !include file, or <file, outside double quotes
    bool quote = false;
    for (size_t c = 0; c < ln.size(); c++) {
      if (ln[c] == '"') quote = !quote;
      if (quote) continue;
      const bool inc = ln.compare(c, 9, "!include ") == 0;
      if (!inc && ln[c] != '<') continue;
      size_t a = c + (inc ? 9 : 1);
      while (a < ln.size() && isspace((unsigned char)ln[a])) a++;
      size_t b = a;
      while (b < ln.size() && !isspace((unsigned char)ln[b]) && ln[b] != '{' && ln[b] != '}') b++;
      if (!read_line(ln.substr(0, c), fname, nline, err)) return false;
      if (!read_line(" ", fname, nline, err)) return false;
      const std::string f = ln.substr(a, b - a);
      if (f.empty()) { err = where(fname, nline) + "expecting filename after '!include' or '<'"; return false; }
      if (!read_fname(f, err)) return false;
      return read_line(ln.substr(b), fname, nline, err);
    }
    out += ln;
    return true;
  }
};

// ---------------------------------------------------------------- object syntax

// words of a preprocessed deck: '{' and '}' are words of their own, a
// double-quoted string is one word (Str's reader, str.hpp:71-100)
struct Tokens {
  std::vector<std::string> w;
  explicit Tokens(const char *s) {
    const size_t n = strlen(s);
    size_t i = 0;
    while (i < n) {
      if (isspace((unsigned char)s[i])) { i++; continue; }
      if (s[i] == '{' || s[i] == '}') { w.emplace_back(1, s[i++]); continue; }
      std::string cur;
      if (s[i] == '"') {
        i++;
        while (i < n && s[i] != '"') cur += s[i++];
        i++;
      } else {
        while (i < n && !isspace((unsigned char)s[i]) && s[i] != '{' && s[i] != '}') cur += s[i++];
      }
      w.push_back(cur);
    }
  }
};

bool parse_double(const std::string &t, double &v) {
  errno = 0;
  char *e = nullptr;
  v = strtod(t.c_str(), &e);
  return !t.empty() && e && *e == 0 && errno == 0;
}

bool parse_int(const std::string &t, int &v) {
  errno = 0;
  char *e = nullptr;
  const long x = strtol(t.c_str(), &e, 10);
  v = (int)x;
  return !t.empty() && e && *e == 0 && errno == 0;
}

// Boolean's reader (boo.hpp:26-41)
bool parse_bool(const std::string &t, int &v) {
  const char *c = t.c_str();
  if (!strcasecmp(c, "y") || !strcasecmp(c, "yes") || !strcasecmp(c, "t") ||
      !strcasecmp(c, "true") || !strcmp(c, "1")) { v = 1; return true; }
  if (!strcasecmp(c, "n") || !strcasecmp(c, "no") || !strcasecmp(c, "f") ||
      !strcasecmp(c, "false") || !strcmp(c, "0")) { v = 0; return true; }
  return false;
}

// index just past the balanced group that opens at w[i] == "{"
bool skip_group(const std::vector<std::string> &w, size_t &i) {
  int depth = 0;
  for (; i < w.size(); i++) {
    if (w[i] == "{") depth++;
    else if (w[i] == "}" && --depth == 0) { i++; return true; }
  }
  return false;
}

}  // namespace

// ElecLJ's io parameters (eleclj.hpp:310-327) with the constructor's
// defaults and its pre-set table entries (eleclj.cpp:12-40)
struct mm_msim_eleclj {
  std::map<std::string, double> dbl{{"elec_1_4_scale_factor", 1},
                                    {"lj_1_4_scale_factor", 1},
                                    {"rspace_cutoff", 0},
                                    {"smoothing_width", 1},
                                    {"neighbor_list_width", 0},
                                    {"kspace_cutoff", 5},
                                    {"ewald_screening", -1},
                                    {"tolerance", 1e-10},
                                    {"bond_charge_increment_frequency", 1500},
                                    {"p3m_grid_spacing", 0.5},
                                    {"p3m_grid_density", -1}};
  std::map<std::string, int> num{{"verbose", 1},
                                 {"max_iterations", 1000},
                                 {"p3m_alias_sum_limit", 2},
                                 {"p3m_assignment_order", 3}};
  std::map<std::string, int> flag{{"reorder_sites", 0},
                                  {"geometric_combining_for_sigma", 0},
                                  {"include_lj_correction", 1},
                                  {"include_rspace", 1},
                                  {"include_kspace", 1},
                                  {"include_self", 1},
                                  {"use_p3m", 0},
                                  {"p3m_ik_differentiate", 0}};
  std::map<std::string, std::map<std::string, double>> dtab{
      {"fixed_charge", {{"dummy", 0.0}}},
      {"screening_radius", {{"dummy", 2.0}}},
      {"sigma", {{"Lj", 1.0}, {"dummy", 1.0}}},
      {"epsilon", {{"Lj", 1.0}, {"dummy", 0.0}}}};
  std::map<std::string, std::string> general_type;
  // tables of parameter kinds the device path does not evaluate (virtual
  // sites, bond charge increments, 1-3 interactions): whether each was given
  std::map<std::string, int> other{{"msite", 0}, {"sp3", 0}, {"sp2", 0},
                                   {"bond_charge_increment", 0},
                                   {"one_three_interaction", 0}};
};

namespace {

std::string field_names(const mm_msim_eleclj &p) {
  std::string s;
  auto add = [&](const std::string &n) { s += (s.empty() ? "" : ", ") + ("'" + n + "'"); };
  for (auto &x : p.other) add(x.first);
  for (auto &x : p.dtab) add(x.first);
  add("general_type");
  for (auto &x : p.dbl) add(x.first);
  for (auto &x : p.num) add(x.first);
  for (auto &x : p.flag) add(x.first);
  return s;
}

// one "{ field value ... }" block at w[i] (the generated operator>>,
// io:219-245; HTab's reader, htab.hpp:222-247); i ends past its '}'
bool read_eleclj(const std::vector<std::string> &w, size_t &i, mm_msim_eleclj &p, std::string &err) {
  if (i >= w.size() || w[i] != "{") { err = "expecting '{' while reading class 'ElecLJ'"; return false; }
  i++;
  while (i < w.size()) {
    const std::string f = w[i++];
    if (f == "}") return true;
    if (i >= w.size()) break;
    const std::string &v = w[i];
    if (p.dbl.count(f)) {
      if (!parse_double(v, p.dbl[f])) { err = "ElecLJ: expecting a number for '" + f + "', read '" + v + "'"; return false; }
      i++;
    } else if (p.num.count(f)) {
      if (!parse_int(v, p.num[f])) { err = "ElecLJ: expecting an integer for '" + f + "', read '" + v + "'"; return false; }
      i++;
    } else if (p.flag.count(f)) {
      if (!parse_bool(v, p.flag[f])) {
        err = "Boolean: expecting 'yes'/'true' or 'no'/'false': read " + v;
        return false;
      }
      i++;
    } else if (p.dtab.count(f) || f == "general_type") {
      if (v != "{") {
        err = "ElecLJ: table '" + f + "' read from file '" + v + "': give its entries inline";
        return false;
      }
      i++;
      while (i < w.size() && w[i] != "}") {
        if (i + 1 >= w.size()) break;
        const std::string key = w[i], val = w[i + 1];
        if (f == "general_type") {
          p.general_type[key] = val;
        } else if (!parse_double(val, p.dtab[f][key])) {
          err = "ElecLJ: expecting a number for " + f + " '" + key + "', read '" + val + "'";
          return false;
        }
        i += 2;
      }
      if (i >= w.size()) { err = "HTab::operator>>: EOF reached before reading '}'"; return false; }
      i++;
    } else if (p.other.count(f)) {
      if (v == "{") {
        const size_t a = i;
        if (!skip_group(w, i)) { err = "HTab::operator>>: EOF reached before reading '}'"; return false; }
        if (i - a > 2) p.other[f] = 1;
      } else {
        p.other[f] = 1;   // a table file
        i++;
      }
    } else {
      err = "unknown field name: '" + f + "' while reading class 'ElecLJ'. Must be one of: " +
            field_names(p);
      return false;
    }
  }
  err = "EOF or error reading class 'ElecLJ'";
  return false;
}

}  // namespace

extern "C" {

void mm_free(void *p) { free(p); }

int mm_msim_preprocess(int nargs, const char *const *args, char **out) {
  if (nargs < 0 || (nargs > 0 && !args) || !out) { set_last_error("mm_msim_preprocess: bad arguments"); return MM_ERR_ARG; }
  *out = nullptr;
  Prepro pp;
  // predefined variables (prepro.cpp:26-36): the working directory's base
  // name and the process id
  char buf[1024];
  if (getcwd(buf, sizeof buf)) {
    const char *b = strrchr(buf, '/');
    pp.var["pwd"] = b ? b + 1 : buf;
  }
  pp.var["pid"] = std::to_string((long)getpid());
  // msim's parameter directory (msimprog.cpp:33-39)
  if (const char *dir = getenv("MSIM_DIR"))
    if (*dir) pp.path.push_back(strstr(dir, "/param") ? std::string(dir) : std::string(dir) + "/param");
  std::string err;
  bool any = false;
  for (int k = 0; k < nargs; k++) {
    const std::string a = args[k] ? args[k] : "";
    if (a.size() > 1 && a[0] == '-') {
      if (k + 1 >= nargs) {
        set_last_error("PreprocessedStream: expecting variable definition for '" + a + "'");
        return MM_ERR_ARG;
      }
      pp.var[a.substr(1)] = args[++k];
      continue;
    }
    any = true;
    std::string base = a;
    const size_t dot = base.rfind('.');
    if (dot != std::string::npos && dot > 0) base = base.substr(0, dot);
    pp.var["base"] = base;
    if (!pp.read_fname(a, err)) { set_last_error("PreprocessedStream: " + err); return MM_ERR_ARG; }
  }
  if (!any) { set_last_error("mm_msim_preprocess: no input file"); return MM_ERR_ARG; }
  *out = (char *)malloc(pp.out.size() + 1);
  if (!*out) { set_last_error("mm_msim_preprocess: out of memory"); return MM_ERR_ARG; }
  memcpy(*out, pp.out.c_str(), pp.out.size() + 1);
  return MM_OK;
}

int mm_msim_eleclj_read(const char *deck, mm_msim_eleclj **out) {
  if (!deck || !out) { set_last_error("mm_msim_eleclj_read: bad arguments"); return MM_ERR_ARG; }
  *out = nullptr;
  Tokens t(deck);
  const auto &w = t.w;
  std::string err;
  auto *p = new mm_msim_eleclj;
  size_t i = 0;
  if (!w.empty() && w[0] == "{" && !(w.size() > 1 && w[1] == "}")) {
    // a whole program: find potential { ... eleclj { ... } ... } at depth 1
    int depth = 0;
    bool found = false, potential = false;
    for (i = 0; i < w.size();) {
      if (w[i] == "{") { depth++; i++; continue; }
      if (w[i] == "}") { depth--; i++; continue; }
      if (depth == 1 && w[i] == "potential" && i + 1 < w.size() && w[i + 1] == "{") {
        potential = true;
        i += 2;
        int d2 = 1;
        while (i < w.size() && d2 > 0) {
          if (w[i] == "{") { d2++; i++; continue; }
          if (w[i] == "}") { d2--; i++; continue; }
          if (d2 == 1 && w[i] == "eleclj") {
            found = true;
            i++;
            // "eleclj" alone adds the potential with its defaults
            if (i < w.size() && w[i] == "{" && !read_eleclj(w, i, *p, err)) {
              delete p;
              set_last_error(err);
              return MM_ERR_ARG;
            }
            continue;
          }
          i++;
        }
        continue;
      }
      i++;
    }
    if (found) { *out = p; return MM_OK; }
    if (potential) {
      delete p;
      set_last_error("mm_msim_eleclj_read: the deck's potential block has no eleclj");
      return MM_ERR_ARG;
    }
    // else: the block itself
    i = 0;
  }
  if (!read_eleclj(w, i, *p, err)) {
    delete p;
    set_last_error(err);
    return MM_ERR_ARG;
  }
  *out = p;
  return MM_OK;
}

void mm_msim_eleclj_free(mm_msim_eleclj *p) { delete p; }

int mm_msim_eleclj_get(const mm_msim_eleclj *p, const char *field, double *value) {
  if (!p || !field || !value) { set_last_error("mm_msim_eleclj_get: bad arguments"); return MM_ERR_ARG; }
  if (p->dbl.count(field)) { *value = p->dbl.at(field); return MM_OK; }
  if (p->num.count(field)) { *value = p->num.at(field); return MM_OK; }
  if (p->flag.count(field)) { *value = p->flag.at(field); return MM_OK; }
  if (p->other.count(field)) { *value = p->other.at(field); return MM_OK; }
  set_last_error(std::string("mm_msim_eleclj_get: no scalar field '") + field + "'");
  return MM_ERR_ARG;
}

int mm_msim_eleclj_table(const mm_msim_eleclj *p, const char *table, const char *key,
                         double *value) {
  if (!p || !table || !key || !value) { set_last_error("mm_msim_eleclj_table: bad arguments"); return MM_ERR_ARG; }
  auto t = p->dtab.find(table);
  if (t == p->dtab.end()) { set_last_error(std::string("mm_msim_eleclj_table: no table '") + table + "'"); return MM_ERR_ARG; }
  auto v = t->second.find(key);
  if (v == t->second.end()) return 1;   // no entry
  *value = v->second;
  return MM_OK;
}

int mm_msim_eleclj_desc(const mm_msim_eleclj *p, mm_system_desc *d) {
  if (!p || !d) { set_last_error("mm_msim_eleclj_desc: bad arguments"); return MM_ERR_ARG; }
  for (auto &x : p->other)
    if (x.second) {
      set_last_error("ElecLJ: '" + x.first + "' parameters are not evaluated on the device");
      return MM_ERR_ARG;
    }
  for (const char *f : {"include_rspace", "include_kspace", "include_self"})
    if (!p->flag.at(f)) {
      set_last_error(std::string("ElecLJ: ") + f + " no: the device path evaluates every term");
      return MM_ERR_ARG;
    }
  if (p->dbl.at("elec_1_4_scale_factor") != 1 || p->dbl.at("lj_1_4_scale_factor") != 1) {
    set_last_error("ElecLJ: 1-4 scale factors other than 1 are not evaluated on the device");
    return MM_ERR_ARG;
  }
  const double boxl = d->boxl;
  mm_desc_init(d);
  d->boxl = boxl;
  d->rc = p->dbl.at("rspace_cutoff");
  d->smoothing_width = p->dbl.at("smoothing_width");
  d->skin = p->dbl.at("neighbor_list_width");
  d->eta = p->dbl.at("ewald_screening");
  d->kc = p->dbl.at("kspace_cutoff");
  d->kmode = p->flag.at("use_p3m") ? 1 : 0;
  d->order = p->num.at("p3m_assignment_order");
  d->nalias = p->num.at("p3m_alias_sum_limit");
  d->ik = p->flag.at("p3m_ik_differentiate") ? 1 : 0;
  d->include_lj_correction = p->flag.at("include_lj_correction");
  d->geometric_sigma = p->flag.at("geometric_combining_for_sigma");
  return mm_desc_resolve(d, p->dbl.at("p3m_grid_spacing"), p->dbl.at("p3m_grid_density"));
}

int mm_msim_site_params(const mm_msim_eleclj *p, int natom, const char *const *types, double *q,
                        double *sigma, double *epsilon, int *mask) {
  if (!p || natom < 0 || (natom > 0 && (!types || !q || !sigma || !epsilon || !mask))) {
    set_last_error("mm_msim_site_params: bad arguments");
    return MM_ERR_ARG;
  }
  // units.h:23-26, 91-94
  const double charge_unit = sqrt(0.5291772108 * 4.35974417 * 6.0221415 * 100 / 4.184);
  const double small = 64 * 2.220446049250313e-16;   // fns.c:22-29
  const auto &fc = p->dtab.at("fixed_charge"), &sr = p->dtab.at("screening_radius");
  const auto &sg = p->dtab.at("sigma"), &ep = p->dtab.at("epsilon");
  for (int i = 0; i < natom; i++) {
    const std::string t = types[i] ? types[i] : "";
    q[i] = 0;
    sigma[i] = epsilon[i] = 0;
    mask[i] = 0;
    auto c = fc.find(t);
    if (c != fc.end() && !(fabs(c->second) < small)) {
      q[i] = c->second * charge_unit;
      mask[i] |= MM_SITE_HAS_Q;
    }
    auto r = sr.find(t);
    if (r != sr.end()) {
      if (r->second < 0) {
        set_last_error("ElecLJ::types_changed: screening radius for type " + t + " is < 0");
        return MM_ERR_ARG;
      }
      // screening acts on charges (rspace.c:104-156): a charged site with a
      // screening radius is beyond the device path
      if (r->second > 0 && (mask[i] & MM_SITE_HAS_Q)) {
        set_last_error("ElecLJ: type " + t + " has a charge and a screening radius; the device "
                       "evaluates unscreened charges only");
        return MM_ERR_ARG;
      }
    }
    auto g = p->general_type.find(t);
    const std::string lt = g != p->general_type.end() ? g->second : t;
    auto s = sg.find(lt), e = ep.find(lt);
    if (s != sg.end() && e != ep.end() && !(fabs(e->second) < small)) {
      if (e->second < 0) { set_last_error("ElecLJ: epsilon must be >= 0 for type " + lt); return MM_ERR_ARG; }
      if (s->second <= 0) { set_last_error("ElecLJ: sigma must be > 0 for type " + lt); return MM_ERR_ARG; }
      mask[i] |= MM_SITE_HAS_LJ;
      sigma[i] = s->second;
      epsilon[i] = e->second;
    }
  }
  return MM_OK;
}

int mm_xyz_read(const char *fname, int natom, char *symbols, char *types, double *r, double *v,
                int *has_velocities, char *bc, int bc_len) {
  if (!fname) { set_last_error("mm_xyz_read: bad arguments"); return MM_ERR_ARG; }
  FILE *f = fopen(fname, "r");
  if (!f) { set_last_error(std::string("mm_xyz_read: cannot open ") + fname); return MM_ERR_ARG; }
  char buf[256];
  auto fail = [&](const std::string &m) {
    fclose(f);
    set_last_error("MSys::read_as_xyz: " + m);
    return MM_ERR_ARG;
  };
  auto line = [&]() { return fgets(buf, sizeof buf, f) && buf[strlen(buf) - 1] == '\n'; };
  int n;
  if (!line()) return fail("error reading line 1 of xyz file");
  if (sscanf(buf, "%d", &n) != 1) return fail("could not read number of atoms");
  if (natom <= 0) {   // count only
    fclose(f);
    return n;
  }
  if (n != natom)
    return fail("number of atoms in molsys (" + std::to_string(natom) +
                ") does not match number of atoms in xyz file (" + std::to_string(n) + ")");
  if (!line()) return fail("error reading line 2 of xyz file");
  if (bc && bc_len > 0) {
    bc[0] = 0;
    const char *b = strstr(buf, "boundary_conditions");
    if (b) {
      // the type word and its dimensions (bcond.cpp:233-276)
      b += strlen("boundary_conditions");
      char word[32] = {0};
      double x[6];
      int used = 0;
      if (sscanf(b, "%31s%n", word, &used) == 1) {
        int nd = !strcmp(word, "non_periodic") ? 0 : !strcmp(word, "orthorhombic") ? 3
                 : !strcmp(word, "triclinic") ? 6 : 1;
        if (nd == 1 && strcmp(word, "cubic") && strcmp(word, "bcc") && strcmp(word, "fcc"))
          return fail(std::string("bcondtype_t: must be one of 'non_periodic','cubic',"
                                  "'orthorhombic','bcc','fcc','triclinic': read ") + word);
        const char *c = b + used;
        std::string s = word;
        for (int k = 0; k < nd; k++) {
          char num[64];
          int u = 0;
          if (sscanf(c, "%63s%n", num, &u) != 1 || !parse_double(num, x[k]))
            return fail("boundary_conditions: expecting " + std::to_string(nd) + " dimensions for " + word);
          c += u;
          s += std::string(" ") + num;
        }
        snprintf(bc, bc_len, "%s", s.c_str());
      }
    }
  }
  int vel = 0;
  for (int i = 0; i < n; i++) {
    if (!line()) return fail("error reading line " + std::to_string(i + 2) + " of xyz file");
    char sym[256], type[256];
    double p[3], u[3];
    const int m = sscanf(buf, "%255s%lf%lf%lf%255s%lf%lf%lf", sym, &p[0], &p[1], &p[2], type, &u[0],
                         &u[1], &u[2]);
    if (!(m == 4 || m == 5 || m == 8)) return fail("error reading line " + std::to_string(i + 2) + " of xyz file");
    if (symbols) snprintf(symbols + 16 * i, 16, "%s", sym);
    if (types) snprintf(types + 16 * i, 16, "%s", m >= 5 ? type : "*");
    if (r)
      for (int k = 0; k < 3; k++) r[3 * i + k] = p[k];
    if (m == 8) {
      vel++;
      if (v)
        for (int k = 0; k < 3; k++) v[3 * i + k] = u[k];
    }
  }
  fclose(f);
  if (has_velocities) *has_velocities = (vel == n);
  return MM_OK;
}

int mm_xyz_write(const char *fname, int natom, const char *symbols, const char *types,
                 const double *r, const double *v, const char *bc, const char *comment,
                 int precision) {
  if (!fname || natom < 0 || (natom > 0 && (!symbols || !r)) || !bc) {
    set_last_error("mm_xyz_write: bad arguments");
    return MM_ERR_ARG;
  }
  FILE *f = fopen(fname, "w");
  if (!f) { set_last_error(std::string("mm_xyz_write: cannot open ") + fname); return MM_ERR_ARG; }
  // ostream's default: 6 significant digits
  const int pr = precision > 0 ? precision : 6;
  auto num = [&](double x) {
    char b[64];
    snprintf(b, sizeof b, "%.*g", pr, x);
    return std::string(b);
  };
  fprintf(f, "%d\nboundary_conditions %s", natom, bc);
  if (comment && *comment) fprintf(f, "  %s", comment);
  fprintf(f, "\n");
  for (int i = 0; i < natom; i++) {
    const char *t = types && types[16 * i] ? types + 16 * i : "*";
    const double z[3] = {0, 0, 0};
    const double *u = v ? v + 3 * i : z;
    fprintf(f, "%s  %s %s %s %s  %s %s %s\n", symbols + 16 * i, num(r[3 * i]).c_str(),
            num(r[3 * i + 1]).c_str(), num(r[3 * i + 2]).c_str(), t, num(u[0]).c_str(),
            num(u[1]).c_str(), num(u[2]).c_str());
  }
  const bool ok = fclose(f) == 0;
  if (!ok) { set_last_error(std::string("mm_xyz_write: error writing ") + fname); return MM_ERR_ARG; }
  return MM_OK;
}

}  // extern "C"
