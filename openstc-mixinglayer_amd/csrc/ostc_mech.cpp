// ostc_mech.cpp -- host reader of the finite-rate mixture tables
// (include/openstc/mech.h): CHEMKIN mechanism (ELEMENTS / SPECIES /
// REACTIONS), NASA-7 thermodynamic data and Lennard-Jones transport data, in
// the file formats and with the derived coefficients of the reference's
// Species and Reaction constructors (src/ChemData.cpp:11-948,
// src/Element.cpp:13-167, src/Constants.h:17-34).
//
// The files are read line by line; the column positions of therm.dat and the
// keyword handling of the REACTIONS section follow the reference, so that a
// file the reference accepts yields the same tables here.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "openstc/mech.h"

namespace {

constexpr double kRc = 1.98720425864083;          // cal/(mol K), src/Constants.h:21
constexpr double kBoltzmann = 1.380649e-23;       // J/K, src/Constants.h:30

// Standard atomic weights of the elements a combustion mechanism names
// (the values of src/Element.cpp:13-...).
double atomic_weight(const std::string& e) {
  static const std::map<std::string, double> w = {
      {"H", 1.008},     {"HE", 4.002602}, {"LI", 6.94},       {"BE", 9.0121831}, {"B", 10.81},
      {"C", 12.011},    {"N", 14.007},    {"O", 15.999},      {"F", 18.998403163}, {"NE", 20.1797},
      {"NA", 22.98976928}, {"MG", 24.305}, {"AL", 26.9815384}, {"SI", 28.085},   {"P", 30.973761998},
      {"S", 32.06},     {"CL", 35.45},    {"AR", 39.95},      {"K", 39.0983},    {"CA", 40.078}};
  auto it = w.find(e);
  return it == w.end() ? 0.0 : it->second;
}

std::string upper(std::string s) {
  for (auto& ch : s) ch = (char)std::toupper((unsigned char)ch);
  return s;
}

std::string ltrim(const std::string& s) {
  const auto p = s.find_first_not_of(" \t\n\v\f\r");
  return p == std::string::npos ? std::string() : s.substr(p);
}

std::string field(const std::string& s, size_t pos, size_t len) {
  return pos >= s.size() ? std::string() : s.substr(pos, len);
}

struct Reader {
  ostc_mech_tables& T;
  std::map<std::string, int> spec;   // name -> index (mechanism order)
  std::map<std::string, int> elem;   // name -> index (mechanism order)
  std::string err;

  bool fail(const std::string& m) {
    err = m;
    return false;
  }

  // Species::Species, src/ChemData.cpp:11-76: ELEMENTS and SPECIES sections,
  // whitespace-separated tokens, '!' starts a comment.
  bool read_species(std::ifstream& f, bool* has_therm) {
    std::vector<std::string> tok;
    std::string line;
    while (std::getline(f, line)) {
      std::istringstream ls(line);
      std::string t;
      while (ls >> t) {
        if (t[0] == '!') break;
        tok.push_back(t);
      }
    }
    size_t p = 0;
    while (p < tok.size() && tok[p] != "ELEMENTS" && tok[p] != "ELEM") ++p;
    if (p == tok.size()) return fail("mechanism: no ELEMENTS section");
    ++p;
    for (; p < tok.size(); ++p) {
      const std::string t = upper(tok[p]);
      if (t == "END") continue;
      if (t == "SPECIES" || t == "SPEC") break;
      if (!elem.count(t)) {
        if ((int)elem.size() >= OSTC_MAX_ELEM) return fail("mechanism: too many elements");
        const int n = (int)elem.size();
        elem.emplace(t, n);
      }
    }
    if (p == tok.size()) return fail("mechanism: no SPECIES section");
    ++p;
    *has_therm = false;
    for (; p < tok.size(); ++p) {
      const std::string t = upper(tok[p]);
      if (t == "END") continue;
      if (t == "REACTIONS" || t == "REAC") break;
      if (t == "THERMO") {
        *has_therm = true;
        break;
      }
      if (!spec.count(t)) {
        if ((int)spec.size() >= OSTC_MAX_SPEC) return fail("mechanism: more than OSTC_MAX_SPEC species");
        const int n = (int)spec.size();
        spec.emplace(t, n);
      }
    }
    T.n_spec = (int)spec.size();
    T.n_elem = (int)elem.size();
    for (auto& [n, i] : spec) std::snprintf(T.spec_name[i], sizeof(T.spec_name[i]), "%s", n.c_str());
    for (auto& [n, i] : elem) std::snprintf(T.elem_name[i], sizeof(T.elem_name[i]), "%s", n.c_str());
    for (int i = 0; i < T.n_spec; ++i) {  // defaults of Species::set_nspec, src/ChemData.cpp:200-222
      T.t_low[i] = 300;
      T.t_mid[i] = 1000;
      T.t_high[i] = 5000;
    }
    return T.n_spec > 0 || fail("mechanism: no species");
  }

  // Species::read_therm (Combustion2Part), src/ChemData.cpp:224-542.  Card 1:
  // name in columns 1-18, element counts in four 5-column fields from column
  // 25, T_low / T_high / T_mid in 10-column fields from column 46; cards 2-4:
  // 15-column coefficients, high range a1..a7 then low range a1..a7.
  bool read_therm(std::ifstream& f, bool positioned) {
    std::string line;
    if (!positioned) {
      bool found = false;
      while (std::getline(f, line)) {
        if (upper(ltrim(line)).rfind("THERMO", 0) == 0) {
          found = true;
          break;
        }
      }
      if (!found) return fail("therm: THERMO not found");
    }
    while (std::getline(f, line)) {  // default temperature ranges (unused: src/ChemData.cpp:240-244)
      if (line.empty() || line[0] == '!') continue;
      break;
    }
    std::vector<int> have(T.n_spec, 0);
    int n_read = 0;
    auto next_upper = [&](std::string& s) {
      if (!std::getline(f, s)) return false;
      s = upper(ltrim(s));
      return true;
    };
    std::string card;
    while (next_upper(card)) {
      if (card.empty() || card[0] == '!') continue;
      std::istringstream ks(card);
      std::string key;
      ks >> key;
      if (key == "END") continue;
      if (key == "REACTIONS" || key == "TRANSPORT" || key == "TRAN") break;
      if (n_read >= T.n_spec) break;
      std::istringstream ns(field(card, 0, 18));
      std::string name;
      ns >> name;
      auto it = spec.find(name);
      if (it == spec.end()) {
        for (int r = 0; r < 3; ++r) std::getline(f, line);
        continue;
      }
      const int s = it->second;
      if (have[s]) {
        // a second entry of an already-read species: its three cards and the
        // next line are skipped (src/ChemData.cpp:340-347)
        for (int r = 0; r < 4; ++r) std::getline(f, line);
        continue;
      }
      T.t_low[s] = std::stod(field(card, 45, 10));
      T.t_high[s] = std::stod(field(card, 55, 10));
      {
        std::istringstream ms(field(card, 65, 10));
        std::string tm;
        if (ms >> tm) T.t_mid[s] = std::stod(tm);
      }
      for (int r = 0; r < 4; ++r) {
        std::string c = ltrim(field(card, 24 + r * 5, 5));
        if (c.empty() || c[0] == '0') break;
        std::istringstream cs(c);
        std::string e;
        int n = 0;
        cs >> e;
        if (!elem.count(e)) continue;
        cs >> n;
        T.elem_comp[s][elem[e]] = n;
      }
      double w = 0;
      for (auto& [e, idx] : elem) w += atomic_weight(e) * T.elem_comp[s][idx];  // element-name order
      T.mw[s] = w;
      std::string c2, c3, c4;
      if (!std::getline(f, c2) || !std::getline(f, c3) || !std::getline(f, c4))
        return fail("therm: truncated entry for " + name);
      auto num = [](const std::string& l, int k) { return std::stod(field(l, 15 * k, 15)); };
      for (int k = 0; k < 5; ++k) T.high_coeff[s][k] = num(c2, k);
      T.high_coeff[s][5] = num(c3, 0);
      T.high_coeff[s][6] = num(c3, 1);
      T.low_coeff[s][0] = num(c3, 2);
      T.low_coeff[s][1] = num(c3, 3);
      T.low_coeff[s][2] = num(c3, 4);
      for (int k = 0; k < 4; ++k) T.low_coeff[s][3 + k] = num(c4, k);
      have[s] = 1;
      ++n_read;
    }
    for (int s = 0; s < T.n_spec; ++s)
      if (!have[s]) return fail(std::string("therm: no data for ") + T.spec_name[s]);
    return true;
  }

  // Species::read_tran + the pair tables, src/ChemData.cpp:544-669.
  bool read_tran(std::ifstream& f) {
    std::vector<double> geom(T.n_spec), eps(T.n_spec), sig(T.n_spec), dip(T.n_spec), pol(T.n_spec);
    std::vector<int> have(T.n_spec, 0);
    int n_read = 0;
    std::string line;
    while (std::getline(f, line)) {
      line = upper(ltrim(line));
      if (line.empty() || line[0] == '!') continue;
      std::istringstream ls(line);
      std::string key;
      ls >> key;
      if (key.rfind("END", 0) == 0 || key.rfind("REAC", 0) == 0) break;
      auto it = spec.find(key);
      if (it == spec.end()) continue;
      if (n_read >= T.n_spec) break;
      const int s = it->second;
      if (have[s]) continue;
      double zrot = 0;
      ls >> geom[s] >> eps[s] >> sig[s] >> dip[s] >> pol[s] >> zrot;
      T.lj_potent_inv[s] = 1.0 / eps[s];
      T.vis_coeff[s] = 2.6693e-6 * std::sqrt(T.mw[s]) / (sig[s] * sig[s]);
      have[s] = 1;
      ++n_read;
    }
    const int n = T.n_spec;
    for (int s = 0; s < n; ++s)
      if (!have[s]) return fail(std::string("tran: no data for ") + T.spec_name[s]);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        T.wj_div_wi_to_one4th[i][j] = std::pow(T.mw[j] / T.mw[i], 0.25);
        T.sqrt_wi_div_wj_pl1_mul8[i][j] = 1.0 / std::sqrt(8 * (1 + T.mw[i] / T.mw[j]));
      }
    // reduced dipole moment mu* = mu / sqrt(eps sigma^3), Debye and Angstrom
    // converted to SI (src/ChemData.cpp:657-668)
    auto mu_star = [&](int i) {
      if (dip[i] == 0) return 0.0;
      const double debye = 1e-18 * std::sqrt(1e-7);
      const double angstrom = std::pow(1e+8, 1.5);
      const double conv = debye * angstrom / std::sqrt(kBoltzmann);
      return conv * dip[i] / std::sqrt(eps[i] * sig[i] * sig[i] * sig[i]);
    };
    for (int j = 0; j < n; ++j) {
      const double wjj = T.mw[j] * 0.5;
      T.binary_diffusivity_coeff[j][j] = 0.0188324 / (std::sqrt(wjj) * sig[j] * sig[j]);
      for (int k = j + 1; k < n; ++k) {
        const double wjk = T.mw[j] * T.mw[k] / (T.mw[j] + T.mw[k]);
        double xi = 1;
        if ((dip[j] != 0) != (dip[k] != 0)) {  // one polar, one non-polar: src/ChemData.cpp:641-655
          const int p = dip[j] != 0 ? j : k, q = dip[j] != 0 ? k : j;
          const double a_star = pol[q] / std::pow(sig[q], 3);
          const double m_star = mu_star(p);
          xi = 1.0 + 0.25 * a_star * m_star * m_star * std::sqrt(eps[p] / eps[q]);
        }
        const double sjk = 0.5 * (sig[j] + sig[k]) * std::pow(xi, -1.0 / 6.0);
        const double bd = 0.0188324 / (std::sqrt(wjk) * sjk * sjk);
        T.binary_diffusivity_coeff[j][k] = T.binary_diffusivity_coeff[k][j] = bd;
        const double ejk = xi * xi * std::sqrt(eps[j] * eps[k]);
        T.kb_over_eps_jk[j][k] = T.kb_over_eps_jk[k][j] = 1.0 / ejk;
      }
    }
    return true;
  }

  // Reaction::read_reaction_line, src/ChemData.cpp:790-886.
  bool reaction_line(std::string in, int r) {
    T.label[r] = OSTC_REAC_REVERSIBLE;
    const auto cat = in.find("+M");
    if (cat != std::string::npos) {
      T.label[r] = OSTC_REAC_THIRD_BODY;
      const auto lp = in.find("(+");
      if (lp != std::string::npos) {
        T.label[r] = OSTC_REAC_LINDEMANN;
        in.erase(lp, in.find(')', lp) - lp + 1);
        const auto lp2 = in.find("(+");
        if (lp2 != std::string::npos) in.erase(lp2, in.find_last_of(')') - lp2 + 1);
      } else {
        in.erase(cat, 2);
        const auto c2 = in.find("+M");
        if (c2 != std::string::npos) in.erase(c2, 2);
      }
    }
    std::istringstream ls(in);
    std::string eq;
    ls >> eq >> T.A[r] >> T.b[r] >> T.Ea[r];
    if (!ls) return fail("reaction: cannot read '" + in + "'");
    T.Ea[r] /= kRc;
    const auto e = eq.find('=');
    if (e == std::string::npos) return fail("reaction: no '=' in '" + eq + "'");
    size_t lhs_end = e, rhs_begin = e + 1;
    if (e > 0 && eq[e - 1] == '<') lhs_end = e - 1;
    if (e + 1 < eq.size() && eq[e + 1] == '>') {
      rhs_begin = e + 2;
      if (lhs_end == e) T.label[r] = OSTC_REAC_IRREVERSIBLE;  // "=>"
    }
    auto side = [&](std::string s, int* stoi, int sign) {
      while (!s.empty()) {
        size_t plus = s.find('+');
        while (plus != std::string::npos && plus + 1 < s.size() && s[plus + 1] == '+') ++plus;  // "X+" names
        std::string term = s.substr(0, plus);
        s = plus == std::string::npos ? std::string() : s.substr(plus + 1);
        int nu = 1;
        if (!term.empty() && std::isdigit((unsigned char)term[0])) {
          nu = term[0] - '0';
          term.erase(0, 1);
        }
        auto it = spec.find(term);
        if (it == spec.end()) return fail("reaction: unknown species '" + term + "'");
        stoi[it->second] += nu;
        T.order[r] += sign * nu;
      }
      return true;
    };
    if (!side(eq.substr(0, lhs_end), T.stoi_f[r], -1)) return false;
    return side(eq.substr(rhs_begin), T.stoi_b[r], +1);
  }

  // Reaction::Reaction + get_auxi_info, src/ChemData.cpp:670-767, :888-947.
  bool read_reactions(std::ifstream& f) {
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
      const std::string u = ltrim(line);
      if (u.rfind("REACTIONS", 0) == 0 || u.rfind("REAC", 0) == 0) {
        found = true;
        break;
      }
    }
    if (!found) return fail("mechanism: REACTIONS not found");
    auto next = [&](std::string& s) {
      if (!std::getline(f, s)) return false;
      s = upper(ltrim(s));
      return true;
    };
    std::vector<int> dups;
    int n = 0;
    std::string in;
    if (!next(in)) return fail("mechanism: empty REACTIONS section");
    while (in != "END") {
      if (in.empty() || in[0] == '!') {
        if (!next(in)) break;
        continue;
      }
      if (n >= OSTC_MAX_REAC) return fail("mechanism: more than OSTC_MAX_REAC reactions");
      // a reused row (after a merged duplicate) starts clean
      std::memset(T.stoi_f[n], 0, sizeof(T.stoi_f[n]));
      std::memset(T.stoi_b[n], 0, sizeof(T.stoi_b[n]));
      T.order[n] = 0;
      if (!reaction_line(in, n)) return false;
      bool dup = false;
      in = "END";
      std::string aux;
      while (next(aux)) {
        std::replace(aux.begin(), aux.end(), '/', ' ');
        if (aux.empty() || aux[0] == '!') continue;
        std::istringstream ls(aux);
        std::string key;
        ls >> key;
        if (key == "END") break;
        if (aux.find('=') != std::string::npos) {
          in = aux;
          break;
        }
        if (key == "PLOG") continue;
        if (key == "DUPLICATE" || key == "DUP") {
          dup = true;
          T.label[n] = OSTC_REAC_DUPLICATE;
        } else if (key == "REV") {
          ls >> T.A2[n] >> T.b2[n] >> T.Ea2[n];
          T.Ea2[n] /= kRc;
          T.rev_type[n] = 1;
        } else if (key == "LOW") {
          ls >> T.A2[n] >> T.b2[n] >> T.Ea2[n];
          T.Ea2[n] /= kRc;
        } else if (key == "TROE") {
          T.label[n] = OSTC_REAC_TROE3;
          ls >> T.troe_alpha[n] >> T.troe_t3[n] >> T.troe_t1[n];
          std::string t2;
          if (ls >> t2) {
            T.troe_t2[n] = std::stod(t2);
            T.label[n] = OSTC_REAC_TROE4;
          }
        } else {  // third-body efficiencies "SPECIES value" pairs
          std::istringstream ps(aux);
          std::string sp;
          double v;
          while (ps >> sp >> v) {
            auto it = spec.find(sp);
            if (it == spec.end()) return fail("reaction: unknown third body '" + sp + "'");
            T.third_body_coeff[n][it->second] = v;
          }
        }
      }
      if (dup) {
        // the second of a duplicate pair (same stoichiometry) becomes the
        // second Arrhenius set of the first (src/ChemData.cpp:710-741)
        int match = -1;
        for (size_t d = 0; d < dups.size() && match < 0; ++d) {
          const int r = dups[d];
          if (std::equal(T.stoi_f[r], T.stoi_f[r] + T.n_spec, T.stoi_f[n]) &&
              std::equal(T.stoi_b[r], T.stoi_b[r] + T.n_spec, T.stoi_b[n]))
            match = (int)d;
        }
        if (match >= 0) {
          const int r = dups[match];
          T.A2[r] = T.A[n];
          T.b2[r] = T.b[n];
          T.Ea2[r] = T.Ea[n];
          dups.erase(dups.begin() + match);
          // the merged line leaves no row behind (its row is reused)
          std::memset(T.stoi_f[n], 0, sizeof(T.stoi_f[n]));
          std::memset(T.stoi_b[n], 0, sizeof(T.stoi_b[n]));
          T.label[n] = OSTC_REAC_REVERSIBLE;
          T.order[n] = 0;
          T.A[n] = T.b[n] = T.Ea[n] = 0;
          continue;
        }
        dups.push_back(n);
      }
      ++n;
    }
    T.n_reac = n;
    for (int r = n; r < OSTC_MAX_REAC; ++r) T.label[r] = 0;
    return true;
  }
};

}  // namespace

extern "C" int ostc_mech_load(const char* mech_file, const char* therm_file, const char* tran_file, int reaction,
                              ostc_mech_tables* out, char* err, int err_cap) {
  auto report = [&](const std::string& m) {
    if (err && err_cap > 0) std::snprintf(err, (size_t)err_cap, "%s", m.c_str());
    return -1;
  };
  if (!mech_file || !out) return report("ostc_mech_load: null argument");
  std::memset(out, 0, sizeof(*out));
  for (int r = 0; r < OSTC_MAX_REAC; ++r)
    for (int s = 0; s < OSTC_MAX_SPEC; ++s) out->third_body_coeff[r][s] = 1;  // src/ChemData.cpp:784
  Reader R{*out, {}, {}, {}};
  std::ifstream mf(mech_file);
  if (!mf) return report(std::string("cannot open ") + mech_file);
  bool has_therm = false;
  if (!R.read_species(mf, &has_therm)) return report(R.err);
  mf.close();
  // thermodynamic data: after THERMO in the mechanism, or the therm file
  std::ifstream tf;
  if (has_therm) {
    tf.open(mech_file);
    std::string line;
    while (std::getline(tf, line))
      if (upper(ltrim(line)).rfind("THERMO", 0) == 0) break;
  } else {
    if (!therm_file) return report("no thermodynamic data file");
    tf.open(therm_file);
    if (!tf) return report(std::string("cannot open ") + therm_file);
  }
  if (!R.read_therm(tf, has_therm)) return report(R.err);
  if (!tran_file) return report("no transport data file");
  std::ifstream rf(tran_file);
  if (!rf) return report(std::string("cannot open ") + tran_file);
  if (!R.read_tran(rf)) return report(R.err);
  if (reaction == 1) {
    std::ifstream mf2(mech_file);
    if (!R.read_reactions(mf2)) return report(R.err);
  }
  return 0;
}
