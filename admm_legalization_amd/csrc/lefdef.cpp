// LEF / DEF reader and DEF writer (include/skyline_lefdef.h).
//
// A token stream over the whole file (words, quoted strings, ';', '(' and
// ')' as tokens of their own, '#' comments dropped) and two recursive-descent
// readers of the statements the legalization flow uses; unknown statements
// are skipped to their ';' and unknown blocks to their END.  Geometry in LEF
// microns becomes dbu as dbTech::getDbuLength does (micron * dbu, truncated
// to int, dbTech.cpp:98-102).
#include <algorithm>
#include <array>
#include <cctype>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <regex>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/skyline_lefdef.h"

namespace {

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct Tok {
  std::string s;
  bool quoted = false;
  size_t pos = 0;  // offset of the token in the file
};

// Tokens of a LEF / DEF text.
std::vector<Tok> tokenize(const std::string& t) {
  std::vector<Tok> out;
  size_t i = 0, n = t.size();
  while (i < n) {
    const char c = t[i];
    if (std::isspace((unsigned char)c)) { i++; continue; }
    if (c == '#') {
      while (i < n && t[i] != '\n') i++;
      continue;
    }
    if (c == '"') {
      size_t j = i + 1;
      while (j < n && t[j] != '"') j++;
      out.push_back(Tok{t.substr(i + 1, j - i - 1), true, i});
      i = j + 1;
      continue;
    }
    if (c == ';' || c == '(' || c == ')') {
      out.push_back(Tok{std::string(1, c), false, i});
      i++;
      continue;
    }
    size_t j = i;
    while (j < n && !std::isspace((unsigned char)t[j]) && t[j] != ';' && t[j] != '(' && t[j] != ')' && t[j] != '"')
      j++;
    out.push_back(Tok{t.substr(i, j - i), false, i});
    i = j;
  }
  return out;
}

struct Stream {
  const std::vector<Tok>& v;
  size_t i = 0;
  std::string err;
  explicit Stream(const std::vector<Tok>& t) : v(t) {}
  bool eof() const { return i >= v.size(); }
  const std::string& peek(size_t k = 0) const {
    static const std::string empty;
    return i + k < v.size() ? v[i + k].s : empty;
  }
  std::string next() { return i < v.size() ? v[i++].s : std::string(); }
  size_t pos() const { return i < v.size() ? v[i].pos : std::string::npos; }
  // past the next ';'
  void skip_statement() {
    while (!eof() && next() != ";") {}
  }
  // past "END <name>" (or a lone END when name is empty)
  void skip_block(const std::string& name) {
    while (!eof()) {
      if (next() == "END" && (name.empty() || peek() == name)) {
        if (!name.empty()) next();
        return;
      }
    }
  }
  bool num(double& d) {
    const std::string s = next();
    char* e = nullptr;
    d = std::strtod(s.c_str(), &e);
    if (s.empty() || *e) {
      err = "number expected, got '" + s + "'";
      return false;
    }
    return true;
  }
  bool integer(long long& v) {
    const std::string s = next();
    char* e = nullptr;
    v = std::strtoll(s.c_str(), &e, 10);
    if (s.empty() || *e) {
      err = "integer expected, got '" + s + "'";
      return false;
    }
    return true;
  }
};

enum Use { USE_SIGNAL = 0, USE_POWER = 1, USE_GROUND = 2 };

struct Port {
  int layer = -1;  // index into LdDb::layers
  std::vector<std::pair<int, int>> pts;
};

struct MPin {
  std::string name;
  int use = USE_SIGNAL;
  std::vector<Port> ports;
  int avg_x = 0, avg_y = 0;  // dbMTerm::setBoundary's average
};

struct Master {
  std::string name;
  std::string cls;
  int sx = 0, sy = 0;
  int ledge = 0, redge = 0;
  std::vector<MPin> pins;
  std::unordered_map<std::string, int> pin_idx;
};

struct Layer {
  std::string name;
  bool routing = false;
};

struct Comp {
  std::string name;
  int master = -1;
  std::string status;  // PLACED, FIXED, COVER, UNPLACED or ""
  int x = 0, y = 0;
  std::string orient = "N";
  std::string source;
};

struct IoPin {
  std::string name, net;
  int layer = -1;
  int rx1 = 0, ry1 = 0, rx2 = 0, ry2 = 0;  // port box from the placement point
  int x = 0, y = 0;
  std::string orient = "N";
  bool placed = false;
};

struct Row {
  int x = 0, y = 0, nx = 1, ny = 1, sx = 0, sy = 0, site_w = 0, site_h = 0;
  std::string orient = "N";
};

struct Wire {
  int layer, x1, y1, x2, y2, width;
};

int orient_code(const std::string& o) {
  static const char* names[] = {"N", "S", "W", "E", "FN", "FS", "FW", "FE"};
  for (int k = 0; k < 8; k++)
    if (o == names[k]) return k;
  return -1;
}

}  // namespace

struct LdDb {
  // technology
  int lef_dbu = 0;
  std::vector<Layer> layers;
  std::unordered_map<std::string, int> layer_idx;
  std::unordered_map<std::string, std::pair<int, int>> sites;
  std::vector<Master> masters;
  std::unordered_map<std::string, int> master_idx;
  std::vector<int> es_t1, es_t2;
  std::vector<double> es_sp;
  // design
  bool have_def = false;
  std::string design, def_text;
  int def_dbu = 0;
  std::vector<Row> rows;
  std::vector<Comp> comps;
  std::unordered_map<std::string, int> comp_idx;
  std::vector<IoPin> io;
  std::unordered_map<std::string, int> io_idx;
  std::vector<std::vector<std::pair<int, std::string>>> nets;  // (component or -1, pin name)
  std::vector<Wire> wires;
  struct Region {
    std::string name, type;
    std::vector<std::array<int, 4>> boxes;
  };
  std::vector<Region> regions;
  std::unordered_map<std::string, int> region_idx;
  struct Group {
    std::string name, region;
    std::vector<std::string> members;
  };
  std::vector<Group> groups;
  size_t comp_begin = 0, comp_end = 0;  // offsets of the COMPONENTS section in def_text
  // LgDesign export
  std::vector<int32_t> row_lx, row_ly, row_dx, row_dy, row_sw, row_or;
  std::vector<int32_t> lx, ly, dx, dy, group, inst_master;
  std::vector<uint8_t> fixed, macro, gnd, m_block;
  std::vector<int32_t> box_group, box_lx, box_ly, box_ux, box_uy;
  std::vector<int32_t> net_start, pin_inst, pin_dx, pin_dy;
  std::vector<const char*> name_ptrs;
  std::vector<int32_t> m_led, m_red, m_pin_start, p_layer, p_lx, p_ly, p_ux, p_uy;
  std::vector<int32_t> w_layer, w_x1, w_y1, w_x2, w_y2, w_w;
  std::vector<int32_t> routing_index;  // layer -> routing-layer index or -1
};

namespace {

bool read_file(const char* path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

int dbu_len(const LdDb* db, double micron) { return (int)(micron * (double)db->lef_dbu); }

// ---------------------------------------------------------------------------
// LEF
// ---------------------------------------------------------------------------
bool lef_layer(LdDb* db, Stream& s) {
  Layer l;
  l.name = s.next();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      if (s.peek() == l.name) s.next();
      break;
    }
    if (t == "TYPE") {
      l.routing = s.next() == "ROUTING";
      s.skip_statement();
    } else if (t == "PROPERTY" || t == "SPACINGTABLE" || t == "ACCURRENTDENSITY" || t == "DCCURRENTDENSITY" ||
               t == "ANTENNAAREARATIO") {
      s.skip_statement();
    } else if (t != ";") {
      s.skip_statement();
    }
  }
  if (!db->layer_idx.count(l.name)) {
    db->layer_idx[l.name] = (int)db->layers.size();
    db->layers.push_back(l);
  }
  return true;
}

bool lef_site(LdDb* db, Stream& s) {
  const std::string name = s.next();
  int w = 0, h = 0;
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      if (s.peek() == name) s.next();
      break;
    }
    if (t == "SIZE") {
      double a, b;
      if (!s.num(a)) return false;
      if (s.next() != "BY") return (s.err = "SITE SIZE: BY expected", false);
      if (!s.num(b)) return false;
      w = dbu_len(db, a);
      h = dbu_len(db, b);
    }
    if (t != ";") s.skip_statement();
  }
  db->sites[name] = {w, h};
  return true;
}

// one PORT (up to its END) or an OBS block: LAYER / RECT / POLYGON
bool lef_geometry(LdDb* db, Stream& s, std::vector<Port>* ports) {
  int layer = -1;
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") return true;
    if (t == "LAYER") {
      const std::string ln = s.next();
      auto it = db->layer_idx.find(ln);
      layer = it == db->layer_idx.end() ? -1 : it->second;
      s.skip_statement();
    } else if (t == "RECT" || t == "POLYGON") {
      std::vector<double> v;
      while (!s.eof() && s.peek() != ";") {
        const std::string w = s.next();
        if (w == "MASK" || w == "ITERATE") { s.next(); continue; }
        char* e = nullptr;
        const double d = std::strtod(w.c_str(), &e);
        if (*e) return (s.err = "coordinate expected in " + t + ", got '" + w + "'", false);
        v.push_back(d);
      }
      s.next();
      if (!ports) continue;
      Port p;
      p.layer = layer;
      if (t == "RECT") {
        if (v.size() != 4) return (s.err = "RECT takes four coordinates", false);
        const int x1 = dbu_len(db, v[0]), y1 = dbu_len(db, v[1]), x2 = dbu_len(db, v[2]), y2 = dbu_len(db, v[3]);
        // the four corners, as dbTech::addPinToMacro stores a rectangle
        p.pts = {{x1, y1}, {x2, y1}, {x2, y2}, {x1, y2}};
      } else {
        if (v.size() < 6 || v.size() % 2) return (s.err = "POLYGON takes coordinate pairs", false);
        for (size_t k = 0; k < v.size(); k += 2) p.pts.push_back({dbu_len(db, v[k]), dbu_len(db, v[k + 1])});
        // dbTech::addPinToMacro keeps no layer for a polygon port
        p.layer = -1;
      }
      ports->push_back(std::move(p));
    } else if (t != ";") {
      s.skip_statement();
    }
  }
  return true;
}

bool lef_pin(LdDb* db, Stream& s, Master& m) {
  MPin p;
  p.name = s.next();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      if (s.peek() == p.name) s.next();
      break;
    }
    if (t == "USE") {
      const std::string u = s.next();
      p.use = u == "POWER" ? USE_POWER : (u == "GROUND" ? USE_GROUND : USE_SIGNAL);
      s.skip_statement();
    } else if (t == "PORT") {
      if (!lef_geometry(db, s, &p.ports)) return false;
    } else if (t != ";") {
      s.skip_statement();
    }
  }
  // dbMTerm::setBoundary (dbMTerm.cpp:20-58): per port the integer average
  // of its points, then the integer average over the ports
  if (!p.ports.empty()) {
    int sx = 0, sy = 0;
    for (const Port& q : p.ports) {
      int ax = 0, ay = 0;
      for (auto [x, y] : q.pts) {
        ax += x;
        ay += y;
      }
      const int np = (int)q.pts.size();
      sx += ax / np;
      sy += ay / np;
    }
    p.avg_x = sx / (int)p.ports.size();
    p.avg_y = sy / (int)p.ports.size();
  }
  m.pin_idx[p.name] = (int)m.pins.size();
  m.pins.push_back(std::move(p));
  return true;
}

// LEF58_EDGETYPE "EDGETYPE LEFT 1 ; EDGETYPE RIGHT 2 ;" (dbTech.cpp:604-626)
void edge_types(const std::string& v, int& left, int& right) {
  std::smatch m;
  static const std::regex rl("LEFT (\\d+) ;"), rr("RIGHT (\\d+) ;");
  if (std::regex_search(v, m, rl)) left = std::atoi(m.str(1).c_str());
  if (std::regex_search(v, m, rr)) right = std::atoi(m.str(1).c_str());
}

bool lef_macro(LdDb* db, Stream& s) {
  Master m;
  m.name = s.next();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      if (s.peek() == m.name) {
        s.next();
        break;
      }
      continue;
    }
    if (t == "CLASS") {
      m.cls = s.next();
      s.skip_statement();
    } else if (t == "SIZE") {
      double a, b;
      if (!s.num(a)) return false;
      if (s.next() != "BY") return (s.err = "MACRO SIZE: BY expected", false);
      if (!s.num(b)) return false;
      m.sx = dbu_len(db, a);
      m.sy = dbu_len(db, b);
      s.skip_statement();
    } else if (t == "PIN") {
      if (!lef_pin(db, s, m)) return false;
    } else if (t == "OBS") {
      if (!lef_geometry(db, s, nullptr)) return false;
    } else if (t == "PROPERTY") {
      const std::string name = s.next();
      std::string val;
      while (!s.eof() && s.peek() != ";") val += s.next() + " ";
      s.next();
      if (name == "LEF58_EDGETYPE") edge_types(val, m.ledge, m.redge);
    } else if (t != ";") {
      s.skip_statement();
    }
  }
  auto it = db->master_idx.find(m.name);
  if (it != db->master_idx.end()) {
    db->masters[it->second] = std::move(m);
  } else {
    db->master_idx[m.name] = (int)db->masters.size();
    db->masters.push_back(std::move(m));
  }
  return true;
}

// LEF58_CELLEDGESPACINGTABLE "... EDGETYPE t1 t2 spacing ..." (dbLefReader.cpp:180-212)
void edge_spacing(LdDb* db, const std::string& v) {
  static const std::regex r("EDGETYPE (\\d+) (\\d+) ([0-9.]+)");
  for (auto it = std::sregex_iterator(v.begin(), v.end(), r); it != std::sregex_iterator(); ++it) {
    db->es_t1.push_back(std::atoi((*it).str(1).c_str()));
    db->es_t2.push_back(std::atoi((*it).str(2).c_str()));
    db->es_sp.push_back(std::strtod((*it).str(3).c_str(), nullptr));
  }
}

int read_lef(LdDb* db, const std::string& text) {
  const auto toks = tokenize(text);
  Stream s(toks);
  while (!s.eof()) {
    const std::string t = s.next();
    bool ok = true;
    if (t == "UNITS") {
      while (!s.eof()) {
        const std::string u = s.next();
        if (u == "END") {
          s.next();
          break;
        }
        if (u == "DATABASE" && s.peek() == "MICRONS") {
          s.next();
          long long v;
          if (!s.integer(v)) return fail(1, "LEF UNITS: " + s.err);
          db->lef_dbu = (int)v;
        }
        if (u != ";") s.skip_statement();
      }
    } else if (t == "LAYER") {
      ok = lef_layer(db, s);
    } else if (t == "SITE") {
      ok = lef_site(db, s);
    } else if (t == "MACRO") {
      if (db->lef_dbu <= 0) return fail(1, "LEF: MACRO before UNITS DATABASE MICRONS");
      ok = lef_macro(db, s);
    } else if (t == "PROPERTY") {
      const std::string name = s.next();
      std::string val;
      while (!s.eof() && s.peek() != ";") val += s.next() + " ";
      s.next();
      if (name == "LEF58_CELLEDGESPACINGTABLE") edge_spacing(db, val);
    } else if (t == "VIA" || t == "VIARULE" || t == "NONDEFAULTRULE") {
      s.skip_block(s.next());
    } else if (t == "PROPERTYDEFINITIONS" || t == "SPACING" || t == "IRDROP" || t == "NOISETABLE" ||
               t == "CORRECTIONTABLE" || t == "ARRAY") {
      s.skip_block(t);
    } else if (t == "BEGINEXT") {
      while (!s.eof() && s.next() != "ENDEXT") {}
    } else if (t == "END") {
      if (s.peek() == "LIBRARY") break;
      s.next();
    } else if (t != ";") {
      s.skip_statement();
    }
    if (!ok) return fail(1, "LEF: " + s.err);
  }
  if (db->lef_dbu <= 0) return fail(1, "LEF: no UNITS DATABASE MICRONS");
  return 0;
}

// ---------------------------------------------------------------------------
// DEF
// ---------------------------------------------------------------------------
bool def_point(Stream& s, long long& x, long long& y, long long px, long long py) {
  if (s.next() != "(") return (s.err = "'(' expected", false);
  std::string a = s.next(), b = s.next();
  x = a == "*" ? px : std::strtoll(a.c_str(), nullptr, 10);
  y = b == "*" ? py : std::strtoll(b.c_str(), nullptr, 10);
  while (!s.eof() && s.peek() != ")") s.next();  // an extension value
  s.next();
  return true;
}

bool def_components(LdDb* db, Stream& s) {
  s.skip_statement();  // COMPONENTS n ;
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();  // COMPONENTS
      return true;
    }
    if (t != "-") return (s.err = "COMPONENTS: '-' expected, got '" + t + "'", false);
    Comp c;
    c.name = s.next();
    const std::string mname = s.next();
    auto it = db->master_idx.find(mname);
    if (it == db->master_idx.end()) return (s.err = "component " + c.name + ": unknown master " + mname, false);
    c.master = it->second;
    while (!s.eof() && s.peek() != ";") {
      const std::string w = s.next();
      if (w != "+") continue;
      const std::string k = s.next();
      if (k == "PLACED" || k == "FIXED" || k == "COVER") {
        c.status = k;
        long long x, y;
        if (!def_point(s, x, y, 0, 0)) return false;
        c.x = (int)x;
        c.y = (int)y;
        c.orient = s.next();
        if (orient_code(c.orient) < 0) return (s.err = "component " + c.name + ": orientation " + c.orient, false);
      } else if (k == "UNPLACED") {
        c.status = k;
      } else if (k == "SOURCE") {
        c.source = s.next();
      }
    }
    s.next();
    db->comp_idx[c.name] = (int)db->comps.size();
    db->comps.push_back(std::move(c));
  }
  return (s.err = "COMPONENTS without END", false);
}

bool def_pins(LdDb* db, Stream& s) {
  s.skip_statement();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();
      return true;
    }
    if (t != "-") return (s.err = "PINS: '-' expected", false);
    IoPin p;
    p.name = s.next();
    while (!s.eof() && s.peek() != ";") {
      const std::string w = s.next();
      if (w != "+") continue;
      const std::string k = s.next();
      if (k == "NET") {
        p.net = s.next();
      } else if (k == "LAYER") {
        auto it = db->layer_idx.find(s.next());
        p.layer = it == db->layer_idx.end() ? -1 : it->second;
        while (s.peek() == "MASK" || s.peek() == "SPACING" || s.peek() == "DESIGNRULEWIDTH") {
          s.next();
          s.next();
        }
        long long x1, y1, x2, y2;
        if (!def_point(s, x1, y1, 0, 0) || !def_point(s, x2, y2, x1, y1)) return false;
        p.rx1 = (int)x1; p.ry1 = (int)y1; p.rx2 = (int)x2; p.ry2 = (int)y2;
      } else if (k == "PLACED" || k == "FIXED" || k == "COVER") {
        long long x, y;
        if (!def_point(s, x, y, 0, 0)) return false;
        p.x = (int)x;
        p.y = (int)y;
        p.orient = s.next();
        p.placed = true;
      }
    }
    s.next();
    db->io_idx[p.name] = (int)db->io.size();
    db->io.push_back(std::move(p));
  }
  return (s.err = "PINS without END", false);
}

bool def_nets(LdDb* db, Stream& s) {
  s.skip_statement();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();
      return true;
    }
    if (t != "-") return (s.err = "NETS: '-' expected", false);
    s.next();  // net name
    std::vector<std::pair<int, std::string>> pins;
    bool in_opts = false;
    while (!s.eof() && s.peek() != ";") {
      const std::string w = s.next();
      if (w == "+") in_opts = true;
      if (in_opts || w != "(") continue;
      const std::string comp = s.next(), pin = s.next();
      while (!s.eof() && s.next() != ")") {}
      if (comp == "PIN") {
        pins.push_back({-1, pin});
      } else if (comp == "*") {
        continue;  // every component with that pin: not a placement pin
      } else {
        auto it = db->comp_idx.find(comp);
        if (it == db->comp_idx.end()) return (s.err = "net pin on unknown component " + comp, false);
        pins.push_back({it->second, pin});
      }
    }
    s.next();
    db->nets.push_back(std::move(pins));
  }
  return (s.err = "NETS without END", false);
}

// SPECIALNETS: the segments of every ROUTED / FIXED / COVER / SHIELD wire
bool def_specialnets(LdDb* db, Stream& s) {
  s.skip_statement();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();
      return true;
    }
    if (t != "-") return (s.err = "SPECIALNETS: '-' expected", false);
    s.next();  // net name
    int layer = -1;
    long long width = 0, px = 0, py = 0;
    bool in_wire = false, have_pt = false;
    while (!s.eof() && s.peek() != ";") {
      const std::string w = s.next();
      if (w == "+") {
        const std::string k = s.next();
        if (k == "ROUTED" || k == "FIXED" || k == "COVER" || k == "SHIELD") {
          if (k == "SHIELD") s.next();  // shielded net name
          auto it = db->layer_idx.find(s.next());
          layer = it == db->layer_idx.end() ? -1 : it->second;
          long long wd;
          if (!s.integer(wd)) return false;
          width = wd;
          in_wire = true;
          have_pt = false;
        } else if (k == "SHAPE" || k == "MASK" || k == "STYLE") {
          s.next();  // an attribute of the current wire, which goes on
        } else {
          in_wire = false;  // + USE / + RECT / + POLYGON / ...: no more points of a wire
        }
        continue;
      }
      if (w == "NEW") {
        auto it = db->layer_idx.find(s.next());
        layer = it == db->layer_idx.end() ? -1 : it->second;
        long long wd;
        if (!s.integer(wd)) return false;
        width = wd;
        in_wire = true;
        have_pt = false;
        continue;
      }
      if (in_wire && w == "(") {
        const std::string a = s.next(), b = s.next();
        while (!s.eof() && s.peek() != ")") s.next();
        s.next();
        const long long x = a == "*" ? px : std::strtoll(a.c_str(), nullptr, 10);
        const long long y = b == "*" ? py : std::strtoll(b.c_str(), nullptr, 10);
        if (have_pt && layer >= 0)
          db->wires.push_back(Wire{layer, (int)px, (int)py, (int)x, (int)y, (int)width});
        px = x;
        py = y;
        have_pt = true;
      }
    }
    s.next();
  }
  return (s.err = "SPECIALNETS without END", false);
}

bool def_regions(LdDb* db, Stream& s) {
  s.skip_statement();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();
      return true;
    }
    if (t != "-") return (s.err = "REGIONS: '-' expected", false);
    LdDb::Region r;
    r.name = s.next();
    while (!s.eof() && s.peek() != ";") {
      if (s.peek() == "(") {
        long long x1, y1, x2, y2;
        if (!def_point(s, x1, y1, 0, 0) || !def_point(s, x2, y2, x1, y1)) return false;
        r.boxes.push_back({(int)std::min(x1, x2), (int)std::min(y1, y2), (int)std::max(x1, x2), (int)std::max(y1, y2)});
        continue;
      }
      const std::string w = s.next();
      if (w == "+" && s.peek() == "TYPE") {
        s.next();
        r.type = s.next();
      }
    }
    s.next();
    db->region_idx[r.name] = (int)db->regions.size();
    db->regions.push_back(std::move(r));
  }
  return (s.err = "REGIONS without END", false);
}

bool def_groups(LdDb* db, Stream& s) {
  s.skip_statement();
  while (!s.eof()) {
    const std::string t = s.next();
    if (t == "END") {
      s.next();
      return true;
    }
    if (t != "-") return (s.err = "GROUPS: '-' expected", false);
    LdDb::Group g;
    g.name = s.next();
    bool opts = false;
    while (!s.eof() && s.peek() != ";") {
      const std::string w = s.next();
      if (w == "+") {
        opts = true;
        if (s.peek() == "REGION") {
          s.next();
          g.region = s.next();
        }
        continue;
      }
      if (!opts) g.members.push_back(w);
    }
    s.next();
    db->groups.push_back(std::move(g));
  }
  return (s.err = "GROUPS without END", false);
}

int read_def(LdDb* db, std::string text) {
  db->def_text = std::move(text);
  const auto toks = tokenize(db->def_text);
  Stream s(toks);
  bool have_comp = false;
  while (!s.eof()) {
    const size_t at = s.pos();
    const std::string t = s.next();
    bool ok = true;
    if (t == "DESIGN") {
      db->design = s.next();
      s.skip_statement();
    } else if (t == "UNITS") {
      s.next();  // DISTANCE
      s.next();  // MICRONS
      long long v;
      if (!s.integer(v)) return fail(1, "DEF UNITS: " + s.err);
      db->def_dbu = (int)v;
      s.skip_statement();
    } else if (t == "ROW") {
      Row r;
      s.next();  // row name
      const std::string site = s.next();
      auto it = db->sites.find(site);
      if (it == db->sites.end()) return fail(1, "DEF ROW: unknown site " + site);
      r.site_w = it->second.first;
      r.site_h = it->second.second;
      long long v;
      if (!s.integer(v)) return fail(1, "DEF ROW: " + s.err);
      r.x = (int)v;
      if (!s.integer(v)) return fail(1, "DEF ROW: " + s.err);
      r.y = (int)v;
      r.orient = s.next();
      while (!s.eof() && s.peek() != ";") {
        const std::string w = s.next();
        long long a, b;
        if (w == "DO") {
          if (!s.integer(a) || s.next() != "BY" || !s.integer(b)) return fail(1, "DEF ROW: bad DO");
          r.nx = (int)a;
          r.ny = (int)b;
        } else if (w == "STEP") {
          if (!s.integer(a) || !s.integer(b)) return fail(1, "DEF ROW: bad STEP");
          r.sx = (int)a;
          r.sy = (int)b;
        }
      }
      s.next();
      db->rows.push_back(r);
    } else if (t == "COMPONENTS") {
      db->comp_begin = at;
      ok = def_components(db, s);
      db->comp_end = s.pos() == std::string::npos ? db->def_text.size() : s.pos();
      have_comp = true;
    } else if (t == "PINS") {
      ok = def_pins(db, s);
    } else if (t == "NETS") {
      ok = def_nets(db, s);
    } else if (t == "SPECIALNETS") {
      ok = def_specialnets(db, s);
    } else if (t == "REGIONS") {
      ok = def_regions(db, s);
    } else if (t == "GROUPS") {
      ok = def_groups(db, s);
    } else if (t == "VIAS" || t == "NONDEFAULTRULES" || t == "BLOCKAGES" || t == "FILLS" || t == "SCANCHAINS" ||
               t == "PROPERTYDEFINITIONS" || t == "STYLES" || t == "SLOTS") {
      while (!s.eof()) {
        if (s.next() == "END" && s.peek() == t) {
          s.next();
          break;
        }
      }
    } else if (t == "END") {
      if (s.peek() == "DESIGN") break;
      s.next();
    } else if (t != ";") {
      s.skip_statement();
    }
    if (!ok) return fail(1, "DEF: " + s.err);
  }
  if (!have_comp) return fail(1, "DEF: no COMPONENTS section");
  if (db->rows.empty()) return fail(1, "DEF: no ROW");
  if (db->lef_dbu > 0 && db->def_dbu > 0 && db->lef_dbu != db->def_dbu)
    return fail(1, "DEF UNITS (" + std::to_string(db->def_dbu) + ") differ from the LEF's (" +
                       std::to_string(db->lef_dbu) + ")");
  db->have_def = true;
  return 0;
}

// The LgDesign arrays of the design read.
void export_design(LdDb* db) {
  const int nr = (int)db->rows.size(), n = (int)db->comps.size();
  db->row_lx.resize(nr); db->row_ly.resize(nr); db->row_dx.resize(nr); db->row_dy.resize(nr);
  db->row_sw.resize(nr); db->row_or.resize(nr);
  for (int i = 0; i < nr; i++) {  // dbRow lx / ux / dx / dy (dbRow.h:25-31)
    const Row& r = db->rows[i];
    db->row_lx[i] = r.x;
    db->row_ly[i] = r.y;
    db->row_dx[i] = r.sx * (r.nx - 1) + r.site_w;
    db->row_dy[i] = r.sy * (r.ny - 1) + r.site_h;
    db->row_sw[i] = r.site_w;
    db->row_or[i] = std::max(orient_code(r.orient), 0);
  }
  // routing-layer index: ROUTING layers in LEF order (Legalizer::importLayers)
  db->routing_index.assign(db->layers.size(), -1);
  for (int k = 0, r = 0; k < (int)db->layers.size(); k++)
    if (db->layers[k].routing) db->routing_index[k] = r++;
  auto rindex = [&](int layer) { return layer >= 0 ? db->routing_index[layer] : -1; };
  db->lx.resize(n); db->ly.resize(n); db->dx.resize(n); db->dy.resize(n);
  db->fixed.resize(n); db->macro.resize(n); db->gnd.resize(n); db->group.assign(n, -1);
  db->inst_master.resize(n);
  db->name_ptrs.resize(n);
  for (int i = 0; i < n; i++) {
    const Comp& c = db->comps[i];
    const Master& m = db->masters[c.master];
    const int o = orient_code(c.orient);
    const bool turn = o == 2 || o == 3 || o == 6 || o == 7;  // W, E, FW, FE
    db->lx[i] = c.x;
    db->ly[i] = c.y;
    db->dx[i] = turn ? m.sy : m.sx;
    db->dy[i] = turn ? m.sx : m.sy;
    db->fixed[i] = (c.status == "FIXED" || c.status == "COVER") ? 1 : 0;
    db->macro[i] = m.cls == "BLOCK" ? 1 : 0;
    // LGCellTechInfo (LGCellTechInfo.cpp:9-37): the lowest point of the
    // POWER pins against that of the GROUND pins
    int vdd = INT_MAX, vss = INT_MAX;
    for (const MPin& p : m.pins)
      for (const Port& q : p.ports)
        for (auto [x, y] : q.pts) {
          if (p.use == USE_POWER) vdd = std::min(vdd, y);
          else if (p.use == USE_GROUND) vss = std::min(vss, y);
        }
    db->gnd[i] = vss > vdd ? 0 : 1;
    db->inst_master[i] = c.master;
    db->name_ptrs[i] = c.name.c_str();
  }
  // groups in DEF order; a member is an instance name or "prefix/*", which
  // takes every instance whose name up to its last '/' is the prefix
  std::unordered_map<std::string, int> pattern;
  for (int g = 0; g < (int)db->groups.size(); g++)
    for (const std::string& mem : db->groups[g].members) {
      auto it = db->comp_idx.find(mem);
      if (it != db->comp_idx.end()) {
        db->group[it->second] = g;
      } else if (mem.size() >= 2 && mem.compare(mem.size() - 2, 2, "/*") == 0) {
        pattern[mem.substr(0, mem.size() - 2)] = g;
      }
    }
  if (!pattern.empty())
    for (int i = 0; i < n; i++) {
      const std::string& nm = db->comps[i].name;
      const size_t cut = nm.rfind('/');
      if (cut == std::string::npos) continue;
      auto it = pattern.find(nm.substr(0, cut));
      if (it != pattern.end()) db->group[i] = it->second;
    }
  // fence boxes: the regions of groups, region by region (importGroupRegions)
  db->box_group.clear(); db->box_lx.clear(); db->box_ly.clear(); db->box_ux.clear(); db->box_uy.clear();
  std::vector<int> region_group(db->regions.size(), -1);
  for (int g = 0; g < (int)db->groups.size(); g++) {
    auto it = db->region_idx.find(db->groups[g].region);
    if (it != db->region_idx.end()) region_group[it->second] = g;
  }
  for (int r = 0; r < (int)db->regions.size(); r++) {
    if (region_group[r] < 0) continue;
    for (const auto& b : db->regions[r].boxes) {
      db->box_group.push_back(region_group[r]);
      db->box_lx.push_back(b[0]);
      db->box_ly.push_back(b[1]);
      db->box_ux.push_back(b[2]);
      db->box_uy.push_back(b[3]);
    }
  }
  // nets: component pins at the master pin's average offset (orientation N),
  // IO pins at the centre of their box (dbBTerm cx / cy)
  db->net_start.assign(1, 0);
  db->pin_inst.clear(); db->pin_dx.clear(); db->pin_dy.clear();
  for (const auto& net : db->nets) {
    for (const auto& [ci, pname] : net) {
      if (ci < 0) {
        auto it = db->io_idx.find(pname);
        if (it == db->io_idx.end()) continue;
        const IoPin& p = db->io[it->second];
        db->pin_inst.push_back(-1);
        db->pin_dx.push_back((p.x + p.rx1 + p.x + p.rx2) / 2);
        db->pin_dy.push_back((p.y + p.ry1 + p.y + p.ry2) / 2);
      } else {
        const Master& m = db->masters[db->comps[ci].master];
        auto it = m.pin_idx.find(pname);
        if (it == m.pin_idx.end()) continue;
        db->pin_inst.push_back(ci);
        db->pin_dx.push_back(m.pins[it->second].avg_x);
        db->pin_dy.push_back(m.pins[it->second].avg_y);
      }
    }
    db->net_start.push_back((int32_t)db->pin_inst.size());
  }
  // technology: per master the edge types and the signal pins' layer and box
  const int M = (int)db->masters.size();
  db->m_led.resize(M); db->m_red.resize(M); db->m_block.resize(M);
  db->m_pin_start.assign(1, 0);
  db->p_layer.clear(); db->p_lx.clear(); db->p_ly.clear(); db->p_ux.clear(); db->p_uy.clear();
  for (int k = 0; k < M; k++) {
    const Master& m = db->masters[k];
    db->m_led[k] = m.ledge;
    db->m_red[k] = m.redge;
    db->m_block[k] = m.cls == "BLOCK" ? 1 : 0;
    for (const MPin& p : m.pins) {
      // signal pins with ports (DatabaseIO.cpp:465-493)
      if (p.use == USE_POWER || p.use == USE_GROUND || p.ports.empty()) continue;
      int bx0 = INT_MAX, by0 = INT_MAX, bx1 = 0, by1 = 0, layer = -1;
      for (const Port& q : p.ports) {
        for (auto [x, y] : q.pts) {
          bx0 = std::min(bx0, x);
          by0 = std::min(by0, y);
          bx1 = std::max(bx1, x);
          by1 = std::max(by1, y);
        }
        // one layer per pin, the last port's (DatabaseIO.cpp:488-489); a port
        // on no routing layer (a polygon keeps none) reads index 0 there, the
        // default of the layer map
        const int r = rindex(q.layer);
        layer = r >= 0 ? r : 0;
      }
      db->p_layer.push_back(layer);
      db->p_lx.push_back(bx0);
      db->p_ly.push_back(by0);
      db->p_ux.push_back(bx1);
      db->p_uy.push_back(by1);
    }
    db->m_pin_start.push_back((int32_t)db->p_layer.size());
  }
  db->w_layer.clear(); db->w_x1.clear(); db->w_y1.clear(); db->w_x2.clear(); db->w_y2.clear(); db->w_w.clear();
  for (const Wire& w : db->wires) {
    db->w_layer.push_back(rindex(w.layer));
    db->w_x1.push_back(w.x1);
    db->w_y1.push_back(w.y1);
    db->w_x2.push_back(w.x2);
    db->w_y2.push_back(w.y2);
    db->w_w.push_back(w.width);
  }
}

}  // namespace

extern "C" {

const char* ld_last_error(void) { return g_err.c_str(); }

int ld_create(LdDb** out) {
  if (!out) return fail(1, "null argument");
  *out = new LdDb();
  return 0;
}

int ld_read_lef(LdDb* db, const char* path) {
  if (!db || !path) return fail(1, "null argument");
  std::string text;
  if (!read_file(path, text)) return fail(1, std::string("cannot read ") + path);
  return read_lef(db, text);
}

int ld_read_def(LdDb* db, const char* path) {
  if (!db || !path) return fail(1, "null argument");
  if (db->masters.empty()) return fail(1, "read_def before read_lef");
  std::string text;
  if (!read_file(path, text)) return fail(1, std::string("cannot read ") + path);
  db->rows.clear(); db->comps.clear(); db->comp_idx.clear(); db->io.clear(); db->io_idx.clear();
  db->nets.clear(); db->wires.clear(); db->regions.clear(); db->region_idx.clear(); db->groups.clear();
  db->have_def = false;
  const int st = read_def(db, std::move(text));
  if (st) return st;
  export_design(db);
  return 0;
}

int ld_get_design(LdDb* db, LgDesign* d) {
  if (!db || !d) return fail(1, "null argument");
  if (!db->have_def) return fail(5, "no DEF read");
  std::memset(d, 0, sizeof(*d));
  d->dbu_per_micron = db->def_dbu > 0 ? db->def_dbu : db->lef_dbu;
  // core: bounding box of the rows (dbDesign.cpp:325-328)
  int cl = INT_MAX, cb = INT_MAX, cr = INT_MIN, ct = INT_MIN;
  for (size_t i = 0; i < db->rows.size(); i++) {
    cl = std::min(cl, db->row_lx[i]);
    cb = std::min(cb, db->row_ly[i]);
    cr = std::max(cr, db->row_lx[i] + db->row_dx[i]);
    ct = std::max(ct, db->row_ly[i] + db->row_dy[i]);
  }
  d->core_lx = cl; d->core_ly = cb; d->core_ux = cr; d->core_uy = ct;
  d->num_rows = (int32_t)db->rows.size();
  d->row_lx = db->row_lx.data(); d->row_ly = db->row_ly.data();
  d->row_dx = db->row_dx.data(); d->row_dy = db->row_dy.data();
  d->row_site_width = db->row_sw.data(); d->row_orient = db->row_or.data();
  d->num_insts = (int32_t)db->comps.size();
  d->inst_lx = db->lx.data(); d->inst_ly = db->ly.data();
  d->inst_dx = db->dx.data(); d->inst_dy = db->dy.data();
  d->inst_fixed = db->fixed.data(); d->inst_macro = db->macro.data();
  d->inst_group = db->group.data(); d->inst_ground_at_bottom = db->gnd.data();
  d->num_groups = (int32_t)db->groups.size();
  d->num_region_boxes = (int32_t)db->box_group.size();
  d->box_group = db->box_group.data(); d->box_lx = db->box_lx.data(); d->box_ly = db->box_ly.data();
  d->box_ux = db->box_ux.data(); d->box_uy = db->box_uy.data();
  d->num_nets = (int32_t)db->nets.size();
  d->net_start = db->net_start.data();
  d->pin_inst = db->pin_inst.data(); d->pin_dx = db->pin_dx.data(); d->pin_dy = db->pin_dy.data();
  d->inst_name = db->name_ptrs.data();
  d->num_masters = (int32_t)db->masters.size();
  d->inst_master = db->inst_master.data();
  d->master_ledge = db->m_led.data(); d->master_redge = db->m_red.data();
  d->master_is_block = db->m_block.data();
  d->master_pin_start = db->m_pin_start.data();
  d->mpin_layer = db->p_layer.data();
  d->mpin_lx = db->p_lx.data(); d->mpin_ly = db->p_ly.data(); d->mpin_ux = db->p_ux.data(); d->mpin_uy = db->p_uy.data();
  d->num_pwires = (int32_t)db->wires.size();
  d->pwire_layer = db->w_layer.data();
  d->pwire_x1 = db->w_x1.data(); d->pwire_y1 = db->w_y1.data();
  d->pwire_x2 = db->w_x2.data(); d->pwire_y2 = db->w_y2.data(); d->pwire_width = db->w_w.data();
  d->num_edge_spacings = (int32_t)db->es_t1.size();
  d->es_type1 = db->es_t1.data(); d->es_type2 = db->es_t2.data(); d->es_spacing_um = db->es_sp.data();
  return 0;
}

int32_t ld_num_insts(const LdDb* db) { return db ? (int32_t)db->comps.size() : 0; }

const char* ld_inst_name(const LdDb* db, int32_t i) {
  if (!db || i < 0 || i >= (int32_t)db->comps.size()) return nullptr;
  return db->comps[i].name.c_str();
}

int ld_write_def(LdDb* db, const char* path, const int32_t* inst_lx, const int32_t* inst_ly, const uint8_t* flipped) {
  if (!db || !path) return fail(1, "null argument");
  if (!db->have_def) return fail(5, "no DEF read");
  std::ofstream f(path, std::ios::binary);
  if (!f) return fail(1, std::string("cannot write ") + path);
  f.write(db->def_text.data(), (std::streamsize)db->comp_begin);
  // dbDesign::writeDef's COMPONENTS section (dbDesign.cpp:1314-1352)
  f << "COMPONENTS " << db->comps.size() << " ;\n";
  for (size_t i = 0; i < db->comps.size(); i++) {
    const Comp& c = db->comps[i];
    const bool fx = db->fixed[i] != 0;
    f << "    - " << c.name << " " << db->masters[c.master].name << " ";
    if (!c.source.empty()) f << "+ SOURCE " << c.source << " ";
    const int x = (!fx && inst_lx) ? inst_lx[i] : c.x;
    const int y = (!fx && inst_ly) ? inst_ly[i] : c.y;
    const std::string o = fx ? c.orient : (flipped ? (flipped[i] ? "FS" : "N") : c.orient);
    f << "+ " << (fx ? (c.status.empty() ? "FIXED" : c.status) : "PLACED") << " ( " << x << " " << y << " ) " << o
      << " ;\n";
  }
  f << "END COMPONENTS\n";
  f.write(db->def_text.data() + db->comp_end, (std::streamsize)(db->def_text.size() - db->comp_end));
  if (!f) return fail(1, std::string("write failed: ") + path);
  return 0;
}

int ld_destroy(LdDb* db) {
  delete db;
  return 0;
}

}  // extern "C"
