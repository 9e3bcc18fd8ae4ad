// Structure and curve ingest (src/structure.rs, src/io.rs): PDB and mmCIF
// coordinate files to atom records, and the two/three-column SAXS curve
// format.  Host code: the work is text parsing, bound by reading the file.
//
// The reference reads coordinates through pdbtbx 0.12 (a Cargo dependency).
// This file reads the fixed PDB columns and the mmCIF _atom_site loop itself
// and then applies the reference's own selection code
// (structure.rs:231-325): residue filters, element check, grouping of
// alternate locations by (chain, residue number, insertion code, atom name),
// the altloc and occupancy policies and the final (chain, residue, name)
// sort.  Records are grouped as they appear in the file.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <string_view>
#include <unordered_map>
#include <vector>

#include "../../include/crabsaxs.h"
#include "../../include/crabsaxs_tables.h"
#include "api_range.hpp"

namespace csx {
csx_status fail(csx_status st, const std::string& msg);
}

struct csx_atoms {
  std::vector<csx_atom_record> atoms;
};

namespace {

using csx::fail;

std::string_view trim(std::string_view s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

std::string upper(std::string_view v) {
  std::string s(v);
  for (char& c : s) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
  return s;
}

std::string lower(std::string_view v) {
  std::string s(v);
  for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

// Rust's f64 grammar (str::parse::<f64>): optional sign, then inf,
// infinity or nan (any case), or digits with an optional fraction and
// exponent; nothing else (no hexadecimal, no surrounding space).  The value
// is the correctly rounded one, as strtod gives.
bool parse_f64(std::string_view t, double& out) {
  size_t i = 0;
  if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
  if (i < t.size() && !std::isdigit(static_cast<unsigned char>(t[i])) && t[i] != '.') {
    const std::string rest = lower(t.substr(i));
    if (rest != "inf" && rest != "infinity" && rest != "nan") return false;
    const bool neg = t[0] == '-';
    out = rest == "nan" ? std::nan("") : (neg ? -HUGE_VAL : HUGE_VAL);
    return true;
  }
  size_t digits = 0;
  while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++digits;
  if (i < t.size() && t[i] == '.') {
    ++i;
    while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++digits;
  }
  if (digits == 0) return false;
  if (i < t.size() && (t[i] == 'e' || t[i] == 'E')) {
    ++i;
    if (i < t.size() && (t[i] == '+' || t[i] == '-')) ++i;
    size_t ed = 0;
    while (i < t.size() && std::isdigit(static_cast<unsigned char>(t[i]))) ++i, ++ed;
    if (ed == 0) return false;
  }
  if (i != t.size()) return false;
  const size_t skip = t[0] == '+' ? 1 : 0;  // from_chars takes no '+'
  const auto r = std::from_chars(t.data() + skip, t.data() + t.size(), out);
  if (r.ec == std::errc::result_out_of_range) {  // overflow to infinity, underflow to zero
    out = std::strtod(std::string(t).c_str(), nullptr);
    return true;
  }
  return r.ec == std::errc() && r.ptr == t.data() + t.size();
}

bool parse_int(std::string_view t, long long& out) {
  if (t.empty()) return false;
  const size_t skip = t[0] == '+' ? 1 : 0;
  const auto r = std::from_chars(t.data() + skip, t.data() + t.size(), out);
  return r.ec == std::errc() && r.ptr == t.data() + t.size();
}

bool read_file(const char* path, std::string& text) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  text = ss.str();
  return true;
}

// Lines without their "\n" or "\r\n" ending (BufRead::lines).
std::vector<std::string_view> split_lines(std::string_view text) {
  std::vector<std::string_view> lines;
  size_t start = 0;
  while (start < text.size()) {
    size_t end = text.find('\n', start);
    if (end == std::string_view::npos) end = text.size();
    size_t stop = end;
    if (stop > start && text[stop - 1] == '\r') --stop;
    lines.push_back(text.substr(start, stop - start));
    start = end + 1;
  }
  return lines;
}

// Element::from_symbol (structure.rs:109-128).
int element_from_symbol(std::string_view symbol) {
  const std::string s = upper(trim(symbol));
  static const std::map<std::string, int> table = {
      {"H", 0},  {"D", 0},   {"C", 1},   {"N", 2},   {"O", 3},   {"S", 4},  {"P", 5}, {"NA", 6},
      {"MG", 7}, {"CL", 8},  {"CA", 9},  {"FE", 10}, {"ZN", 11}, {"K", 12}, {"I", 13}};
  const auto it = table.find(s);
  return it == table.end() ? 14 : it->second;
}

// One coordinate record as read from the file.
struct Raw {
  double x = 0, y = 0, z = 0, occupancy = 1.0;
  std::string name, residue, chain, icode, element;
  bool has_element = false, hetero = false;
  long long seq = 0;
};

using Model = std::vector<Raw>;

csx_status parse_error(const std::string& msg) { return fail(CSX_PARSE_ERROR, msg); }

std::string_view column(std::string_view line, size_t first, size_t last) {  // 1-based, inclusive
  if (line.size() < first) return std::string_view();
  return line.substr(first - 1, std::min(last, line.size()) - (first - 1));
}

// PDB ATOM/HETATM records (fixed columns of the wwPDB format), split into
// models at MODEL records; reading stops at END.
csx_status read_pdb(const std::string& text, std::vector<Model>& models) {
  const std::vector<std::string_view> lines = split_lines(text);
  for (size_t ln = 0; ln < lines.size(); ++ln) {
    const std::string_view line = lines[ln];
    const std::string_view rec = trim(column(line, 1, 6));
    auto where = [ln] { return " on line " + std::to_string(ln + 1); };
    if (rec == "MODEL") {
      models.emplace_back();
      continue;
    }
    if (rec == "END") break;
    if (rec != "ATOM" && rec != "HETATM") continue;
    if (line.size() < 54) return parse_error("coordinate record is too short" + where());
    if (models.empty()) models.emplace_back();  // a file without MODEL records
    Raw a;
    a.hetero = rec == "HETATM";
    a.name = trim(column(line, 13, 16));
    a.residue = trim(column(line, 18, 20));
    a.chain = trim(column(line, 22, 22));
    if (!parse_int(trim(column(line, 23, 26)), a.seq))
      return parse_error("invalid residue number" + where());
    a.icode = trim(column(line, 27, 27));
    if (!parse_f64(trim(column(line, 31, 38)), a.x) || !parse_f64(trim(column(line, 39, 46)), a.y) ||
        !parse_f64(trim(column(line, 47, 54)), a.z))
      return parse_error("invalid coordinate" + where());
    const std::string_view occ = trim(column(line, 55, 60));
    if (!occ.empty() && !parse_f64(occ, a.occupancy))
      return parse_error("invalid occupancy" + where());
    a.element = trim(column(line, 77, 78));
    a.has_element = !a.element.empty();
    models.back().push_back(std::move(a));
  }
  if (models.empty()) models.emplace_back();
  return CSX_OK;
}

// mmCIF tokens: whitespace-separated words, '...' / "..." strings (closed by
// the quote followed by whitespace), ;-delimited text fields and # comments.
struct CifToken {
  std::string text;
  bool quoted = false;
};

std::vector<CifToken> cif_tokens(const std::string& text) {
  std::vector<CifToken> out;
  size_t i = 0;
  const size_t n = text.size();
  bool line_start = true;
  while (i < n) {
    const char c = text[i];
    if (c == '\n') {
      line_start = true;
      ++i;
      continue;
    }
    if (std::isspace(static_cast<unsigned char>(c))) {
      ++i;
      continue;
    }
    if (line_start && c == ';') {
      size_t end = text.find("\n;", i);
      if (end == std::string::npos) end = n;
      out.push_back({text.substr(i + 1, end - i - 1), true});
      i = end == n ? n : end + 2;
      line_start = false;
      continue;
    }
    line_start = false;
    if (c == '#') {
      while (i < n && text[i] != '\n') ++i;
      continue;
    }
    if (c == '\'' || c == '"') {
      size_t j = i + 1;
      while (j < n && !(text[j] == c && (j + 1 == n || std::isspace(static_cast<unsigned char>(text[j + 1])))))
        ++j;
      out.push_back({text.substr(i + 1, j - i - 1), true});
      i = j + 1;
      continue;
    }
    size_t j = i;
    while (j < n && !std::isspace(static_cast<unsigned char>(text[j]))) ++j;
    out.push_back({text.substr(i, j - i), false});
    i = j;
  }
  return out;
}

bool cif_missing(const CifToken& t) { return !t.quoted && (t.text == "." || t.text == "?"); }

// mmCIF number, allowing a standard-uncertainty suffix such as 1.234(5).
bool cif_number(const CifToken& t, double& v) {
  std::string s = t.text;
  const size_t p = s.find('(');
  if (p != std::string::npos && !s.empty() && s.back() == ')') s = s.substr(0, p);
  return parse_f64(s, v);
}

// The _atom_site loop of the first data block, split into models by
// pdbx_PDB_model_num in order of first appearance.  Author fields (auth_*)
// are used where present, else the label_* ones.
csx_status read_mmcif(const std::string& text, std::vector<Model>& models) {
  const std::vector<CifToken> tok = cif_tokens(text);
  size_t i = 0;
  bool seen_block = false;
  while (i < tok.size()) {
    const CifToken& t = tok[i];
    if (!t.quoted && t.text.compare(0, 5, "data_") == 0) {
      if (seen_block) break;
      seen_block = true;
      ++i;
      continue;
    }
    if (t.quoted || lower(t.text) != "loop_") {
      ++i;
      continue;
    }
    ++i;
    std::vector<std::string> tags;
    while (i < tok.size() && !tok[i].quoted && !tok[i].text.empty() && tok[i].text[0] == '_')
      tags.push_back(lower(tok[i++].text));
    if (tags.empty() || tags[0].compare(0, 11, "_atom_site.") != 0) continue;
    auto col = [&](const char* name) -> int {
      for (size_t k = 0; k < tags.size(); ++k)
        if (tags[k] == std::string("_atom_site.") + name) return static_cast<int>(k);
      return -1;
    };
    auto either = [&](const char* a, const char* b) { return col(a) >= 0 ? col(a) : col(b); };
    const int c_group = col("group_pdb"), c_type = col("type_symbol");
    const int c_name = either("auth_atom_id", "label_atom_id");
    const int c_res = either("auth_comp_id", "label_comp_id");
    const int c_chain = either("auth_asym_id", "label_asym_id");
    const int c_seq = either("auth_seq_id", "label_seq_id");
    const int c_ins = col("pdbx_pdb_ins_code"), c_occ = col("occupancy");
    const int c_x = col("cartn_x"), c_y = col("cartn_y"), c_z = col("cartn_z");
    const int c_model = col("pdbx_pdb_model_num");
    if (c_x < 0 || c_y < 0 || c_z < 0 || c_name < 0)
      return parse_error("_atom_site lacks Cartn_x/y/z or an atom id");
    const size_t w = tags.size();
    std::vector<std::string> model_ids;
    while (i < tok.size()) {
      const CifToken& first = tok[i];
      if (!first.quoted && (first.text[0] == '_' || lower(first.text) == "loop_" ||
                            first.text.compare(0, 5, "data_") == 0))
        break;
      if (i + w > tok.size()) return parse_error("_atom_site row is incomplete");
      const CifToken* r = &tok[i];
      i += w;
      Raw a;
      auto text_of = [&](int c) { return c >= 0 && !cif_missing(r[c]) ? r[c].text : std::string(); };
      a.hetero = upper(text_of(c_group)) == "HETATM";
      a.name = text_of(c_name);
      a.residue = text_of(c_res);
      a.chain = text_of(c_chain);
      a.icode = text_of(c_ins);
      a.element = text_of(c_type);
      a.has_element = !a.element.empty();
      const std::string seq = text_of(c_seq);
      if (!seq.empty() && !parse_int(seq, a.seq)) return parse_error("invalid residue number " + seq);
      if (!cif_number(r[c_x], a.x) || !cif_number(r[c_y], a.y) || !cif_number(r[c_z], a.z))
        return parse_error("invalid coordinate in _atom_site row " + text_of(col("id")));
      if (c_occ >= 0 && !cif_missing(r[c_occ]) && !cif_number(r[c_occ], a.occupancy))
        return parse_error("invalid occupancy in _atom_site row " + text_of(col("id")));
      const std::string mid = text_of(c_model);
      size_t m = 0;
      while (m < model_ids.size() && model_ids[m] != mid) ++m;
      if (m == model_ids.size()) {
        model_ids.push_back(mid);
        models.emplace_back();
      }
      models[m].push_back(std::move(a));
    }
    return CSX_OK;
  }
  return parse_error("no _atom_site loop found");
}

int residue_kind(const std::string& name) {
  const std::string code = upper(trim(name));
  if (code == "HOH" || code == "WAT" || code == "DOD") return CSX_RESIDUE_WATER;
  static const char* glycans[] = {
      "NAG", "NDG", "BMA", "MAN", "GLC", "GAL", "FUC", "SIA", "NAN", "NEU", "XYS", "XYP", "ARA", "ARB",
      "RIB", "FRU", "G6D", "G1P", "BGC", "AFD", "AGS", "AMU", "API", "DGA", "DGL", "DGU", "GCU", "GTR",
      "IDR", "LBT", "MMA", "RM4", "SUC", "TRE", "MAL", "LMT", "KDO", "KDN", "DAN", "XUL"};
  for (const char* g : glycans)
    if (code == g) return CSX_RESIDUE_GLYCAN;
  static const char* polymer[] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
                                  "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL",
                                  "A",   "C",   "G",   "U",   "T",   "DA",  "DC",  "DG",  "DT"};
  for (const char* p : polymer)
    if (code == p) return CSX_RESIDUE_POLYMER;
  return CSX_RESIDUE_OTHER;
}

bool copy_name(const std::string& s, char (&dst)[16]) {
  if (s.size() >= sizeof dst) return false;
  std::memcpy(dst, s.c_str(), s.size() + 1);
  return true;
}

// The selection of structure.rs:231-325 applied to one model.
csx_status select_atoms(const Model& model, const csx_parse_options& o, csx_atoms& out) {
  std::vector<std::string> chains;
  if (o.chains) {
    std::stringstream ss(o.chains);
    std::string c;
    while (std::getline(ss, c, ',')) chains.push_back(c);
  }
  // key (chain, residue number, insertion code, name), joined with a byte
  // that does not occur in PDB / mmCIF fields
  std::unordered_map<std::string, size_t> group_of;
  group_of.reserve(model.size());
  std::string key;
  // residue kinds, elements and united-atom groups repeat over few distinct
  // names: computed once per name
  std::unordered_map<std::string, int> kinds, element_of, ua_group;
  auto memo = [](std::unordered_map<std::string, int>& m, const std::string& k, auto&& f) {
    auto it = m.find(k);
    if (it == m.end()) it = m.emplace(k, f()).first;
    return it->second;
  };
  std::vector<std::vector<const Raw*>> groups;
  std::vector<int> elements(model.size());
  for (size_t k = 0; k < model.size(); ++k) {
    const Raw& a = model[k];
    const std::string residue = a.residue.empty() ? std::string("UNK") : a.residue;
    const int kind = memo(kinds, residue, [&] { return residue_kind(residue); });
    if (o.chains && std::find(chains.begin(), chains.end(), a.chain) == chains.end()) continue;
    if (a.hetero) {
      if (kind == CSX_RESIDUE_WATER && !o.include_waters) continue;
      if (!o.include_hetatm && !(o.include_glycans && kind == CSX_RESIDUE_GLYCAN)) continue;
    }
    const std::string& symbol = a.has_element ? a.element : a.name;
    const int element = memo(element_of, symbol, [&] { return element_from_symbol(symbol); });
    if (element == 14 && !o.allow_unknown_elements)
      return fail(CSX_UNSUPPORTED_ELEMENT, "unsupported element: " + a.name);
    elements[k] = element;
    key.assign(a.chain);
    key += '\x1f';
    char num[24];
    key.append(num, std::to_chars(num, num + sizeof num, a.seq).ptr);
    key += '\x1f';
    key += a.icode;
    key += '\x1f';
    key += a.name;
    auto it = group_of.find(key);
    if (it == group_of.end()) {
      it = group_of.emplace(key, groups.size()).first;
      groups.emplace_back();
    }
    groups[it->second].push_back(&a);
  }
  auto record = [&](const Raw& a, double occupancy, csx_atom_record& r) -> bool {
    std::memset(&r, 0, sizeof r);
    r.x = a.x;
    r.y = a.y;
    r.z = a.z;
    r.occupancy = occupancy;
    r.element = elements[static_cast<size_t>(&a - model.data())];
    r.is_hetatm = a.hetero ? 1 : 0;
    r.residue_id = a.seq;
    const std::string residue = a.residue.empty() ? std::string("UNK") : a.residue;
    if (!copy_name(a.name, r.atom_name) || !copy_name(residue, r.residue_name) ||
        !copy_name(a.chain, r.chain_id))
      return false;
    key.assign(residue);
    key += '\x1f';
    key += a.name;
    r.group = memo(ua_group, key,
                   [&] { return csx_united_atom_group(r.residue_name, r.atom_name); });
    return true;
  };
  std::vector<csx_atom_record> atoms;
  for (const auto& alts : groups) {
    if (o.altloc_policy == CSX_ALTLOC_ALL) {
      for (const Raw* a : alts) {
        atoms.emplace_back();
        if (!record(*a, std::fmax(a->occupancy, 0.0), atoms.back()))
          return parse_error("name longer than 15 characters: " + a->name);
      }
      continue;
    }
    const size_t count = alts.size();
    const bool has_alternatives = count > 1;
    bool same_position = has_alternatives;
    for (size_t k = 1; k < count && same_position; ++k)
      same_position = alts[k]->x == alts[0]->x && alts[k]->y == alts[0]->y && alts[k]->z == alts[0]->z;
    // Occupancies are max(reported, 0) (structure.rs:273); Highest takes the
    // last of the equally largest (Iterator::max_by).
    const Raw* selected = alts[0];
    if (o.altloc_policy == CSX_ALTLOC_HIGHEST)
      for (const Raw* a : alts)
        if (std::fmax(a->occupancy, 0.0) >= std::fmax(selected->occupancy, 0.0)) selected = a;
    double occupancy = std::fmax(selected->occupancy, 0.0);
    if (o.occupancy_policy == CSX_OCCUPANCY_SELECTED_FULL && has_alternatives)
      occupancy = 1.0;
    else if (o.occupancy_policy == CSX_OCCUPANCY_REPORTED && same_position)
      occupancy = std::min(occupancy * static_cast<double>(count + 1), 1.0);
    atoms.emplace_back();
    if (!record(*selected, occupancy, atoms.back()))
      return parse_error("name longer than 15 characters: " + selected->name);
  }
  std::stable_sort(atoms.begin(), atoms.end(), [](const csx_atom_record& l, const csx_atom_record& r) {
    const int c = std::strcmp(l.chain_id, r.chain_id);
    if (c != 0) return c < 0;
    if (l.residue_id != r.residue_id) return l.residue_id < r.residue_id;
    return std::strcmp(l.atom_name, r.atom_name) < 0;
  });
  out.atoms = std::move(atoms);
  return CSX_OK;
}

bool ends_with(const std::string& s, const char* suffix) {
  const size_t n = std::strlen(suffix);
  return s.size() >= n && lower(s.substr(s.size() - n)) == suffix;
}

bool looks_like_mmcif(const std::string& text) {
  for (const CifToken& t : cif_tokens(text.substr(0, std::min<size_t>(text.size(), 4096))))
    return !t.quoted && t.text.compare(0, 5, "data_") == 0;
  return false;
}

csx_status parse_structure(const std::string& text, int format, const csx_parse_options* opts,
                           csx_atoms** out) {
  if (!out) return fail(CSX_INVALID_INPUT, "output pointer is required");
  *out = nullptr;
  csx_parse_options o;
  csx_parse_options_default(&o);
  if (opts) o = *opts;
  if (o.altloc_policy < 0 || o.altloc_policy > 2 || o.occupancy_policy < 0 || o.occupancy_policy > 1)
    return fail(CSX_INVALID_INPUT, "unknown altloc or occupancy policy");
  if (format == CSX_FORMAT_AUTO) format = looks_like_mmcif(text) ? CSX_FORMAT_MMCIF : CSX_FORMAT_PDB;
  std::vector<Model> models;
  csx_status st = format == CSX_FORMAT_MMCIF ? read_mmcif(text, models) : read_pdb(text, models);
  if (st != CSX_OK) return st;
  if (o.model_index >= models.size())
    return parse_error("model " + std::to_string(o.model_index) + " is not present");
  std::unique_ptr<csx_atoms> a(new csx_atoms);
  st = select_atoms(models[o.model_index], o, *a);
  if (st != CSX_OK) return st;
  *out = a.release();
  return CSX_OK;
}

// ExperimentalCurve::from_dat_file (io.rs:19-112), line for line.
csx_status parse_dat(const std::string& text, std::vector<double>& q, std::vector<double>& in,
                     std::vector<double>& sg, int& has_errors) {
  has_errors = 1;
  double q_scale = 1.0;
  bool saw_numeric = false;
  const std::vector<std::string_view> lines = split_lines(text);
  for (size_t ln = 0; ln < lines.size(); ++ln) {
    const std::string trimmed(trim(lines[ln]));
    const std::string low = lower(trimmed);
    const std::string where = " on line " + std::to_string(ln + 1);
    if (low.find("nm") != std::string::npos &&
        (low.find('q') != std::string::npos || low.find("angstrom") != std::string::npos))
      q_scale = 0.1;
    if (trimmed.empty() || trimmed[0] == '#') continue;
    std::string norm = trimmed;
    std::replace(norm.begin(), norm.end(), ',', ' ');
    std::vector<std::string> fields;
    {
      std::istringstream ss(norm);
      std::string f;
      while (ss >> f) fields.push_back(f);
    }
    double v0;
    if (fields.empty() || !parse_f64(fields[0], v0)) continue;
    if (fields.size() < 2) return fail(CSX_PARSE_ERROR, "expected q and intensity" + where);
    double values[3];
    const size_t nv = std::min<size_t>(fields.size(), 3);
    for (size_t k = 0; k < nv; ++k)
      if (!parse_f64(fields[k], values[k])) return fail(CSX_PARSE_ERROR, "invalid numeric value" + where);
    if (!saw_numeric && q_scale == 1.0 && values[0] > 1.0) q_scale = 0.1;
    saw_numeric = true;
    const double qq = values[0] * q_scale, ii = values[1], ss = nv > 2 ? values[2] : 1.0;
    if (ii == 0.0 && ss == 0.0) continue;
    if (!std::isfinite(qq) || qq < 0.0 || !std::isfinite(ii) || !std::isfinite(ss) || ss <= 0.0)
      return fail(CSX_PARSE_ERROR, "q and sigma must be positive and all values finite" + where);
    if (!q.empty() && qq < q.back())
      return fail(CSX_PARSE_ERROR, "q values must be non-decreasing" + where);
    q.push_back(qq);
    in.push_back(ii);
    sg.push_back(ss);
    if (fields.size() < 3) has_errors = 0;
  }
  if (q.empty()) return fail(CSX_PARSE_ERROR, "no data rows found");
  return CSX_OK;
}

csx_status dat_out(csx_status st, const std::vector<double>& q, const std::vector<double>& in,
                   const std::vector<double>& sg, int has, double* q_out, double* i_out,
                   double* s_out, size_t capacity, size_t* count, int* has_errors) {
  if (st != CSX_OK) return st;
  if (count) *count = q.size();
  if (has_errors) *has_errors = has;
  const size_t m = std::min(capacity, q.size());
  if (q_out) std::copy(q.begin(), q.begin() + m, q_out);
  if (i_out) std::copy(in.begin(), in.begin() + m, i_out);
  if (s_out) std::copy(sg.begin(), sg.begin() + m, s_out);
  return CSX_OK;
}

}  // namespace

extern "C" {

void csx_parse_options_default(csx_parse_options* out) {
  if (!out) return;
  out->include_hetatm = 0;
  out->include_waters = 0;
  out->include_glycans = 1;
  out->model_index = 0;
  out->allow_unknown_elements = 0;
  out->occupancy_policy = CSX_OCCUPANCY_REPORTED;
  out->altloc_policy = CSX_ALTLOC_HIGHEST;
  out->chains = nullptr;
}

int csx_residue_kind(const char* residue_name) {
  return residue_kind(residue_name ? residue_name : "");
}

csx_status csx_parse_structure(const char* text, size_t len, int format,
                               const csx_parse_options* options, csx_atoms** out) {
  CSX_API_RANGE();
  if (!text && len) return fail(CSX_INVALID_INPUT, "text is required");
  return parse_structure(std::string(text ? text : "", len), format, options, out);
}

csx_status csx_read_structure_file(const char* path, const csx_parse_options* options,
                                   csx_atoms** out) {
  CSX_API_RANGE();
  if (!path) return fail(CSX_INVALID_INPUT, "path is required");
  std::string text;
  if (!read_file(path, text)) return fail(CSX_IO_ERROR, std::string("cannot read ") + path);
  const std::string p(path);
  if (ends_with(p, ".gz")) return fail(CSX_PARSE_ERROR, "compressed structure files are not supported");
  const int format = ends_with(p, ".cif") || ends_with(p, ".mmcif") ? CSX_FORMAT_MMCIF
                     : ends_with(p, ".pdb") || ends_with(p, ".ent") ? CSX_FORMAT_PDB
                                                                    : CSX_FORMAT_AUTO;
  return parse_structure(text, format, options, out);
}

size_t csx_atoms_count(const csx_atoms* a) { return a ? a->atoms.size() : 0; }

csx_status csx_atoms_get(const csx_atoms* a, csx_atom_record* out) {
  if (!a || (!out && !a->atoms.empty())) return fail(CSX_INVALID_INPUT, "atoms and output are required");
  std::copy(a->atoms.begin(), a->atoms.end(), out);
  return CSX_OK;
}

csx_status csx_atoms_arrays(const csx_atoms* a, double* xyz, uint8_t* element, int8_t* group,
                            double* occupancy) {
  if (!a) return fail(CSX_INVALID_INPUT, "atoms are required");
  for (size_t i = 0; i < a->atoms.size(); ++i) {
    const csx_atom_record& r = a->atoms[i];
    if (xyz) {
      xyz[3 * i] = r.x;
      xyz[3 * i + 1] = r.y;
      xyz[3 * i + 2] = r.z;
    }
    if (element) element[i] = static_cast<uint8_t>(r.element);
    if (group) group[i] = static_cast<int8_t>(r.group);
    if (occupancy) occupancy[i] = r.occupancy;
  }
  return CSX_OK;
}

void csx_atoms_destroy(csx_atoms* a) { delete a; }

csx_status csx_parse_dat(const char* text, size_t len, double* q, double* intensity, double* sigma,
                         size_t capacity, size_t* count, int* has_errors) {
  CSX_API_RANGE();
  if (!text && len) return fail(CSX_INVALID_INPUT, "text is required");
  std::vector<double> qq, ii, ss;
  int has = 1;
  const csx_status st = parse_dat(std::string(text ? text : "", len), qq, ii, ss, has);
  return dat_out(st, qq, ii, ss, has, q, intensity, sigma, capacity, count, has_errors);
}

csx_status csx_read_dat_file(const char* path, double* q, double* intensity, double* sigma,
                             size_t capacity, size_t* count, int* has_errors) {
  CSX_API_RANGE();
  if (!path) return fail(CSX_INVALID_INPUT, "path is required");
  std::string text;
  if (!read_file(path, text)) return fail(CSX_IO_ERROR, std::string("cannot read ") + path);
  std::vector<double> qq, ii, ss;
  int has = 1;
  const csx_status st = parse_dat(text, qq, ii, ss, has);
  return dat_out(st, qq, ii, ss, has, q, intensity, sigma, capacity, count, has_errors);
}

}  // extern "C"
