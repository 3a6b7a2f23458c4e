// On-disk formats of the reference's drivers (host code only):
//   * the multigrid parameter file: Grid's XML serialisation of MGTestParams
//     (include/MultigridParams.h:73-185; read by src/Test_wilsonclover_mg.cc:51-58
//     with XmlReader, written by the commented-out XmlWriter at :60-64);
//   * gauge configurations in the NERSC format and in ILDG (LIME records)
//     (src/Test_wilsonclover_mg.cc:94-101, IldgReader; Grid's NerscIO).
// Gauge fields are host arrays [mu][site][3][3] of complex doubles, sites
// lexicographic with x fastest, as the rest of the library.
#include "common.hpp"

#include <cctype>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

namespace mgc {
namespace io {

// ---------------------------------------------------------------------------
// a small XML tree: elements with text, attributes ignored
// ---------------------------------------------------------------------------
struct XNode {
  std::string name, text;
  std::vector<std::unique_ptr<XNode>> kids;
  const XNode* child(const std::string& n) const {
    for (auto& k : kids)
      if (k->name == n) return k.get();
    return nullptr;
  }
};

class XmlParser {
 public:
  explicit XmlParser(const std::string& s) : s_(s) {}
  std::unique_ptr<XNode> parse() {
    auto root = std::make_unique<XNode>();
    root->name = "#document";
    parse_children(*root, "");
    return root;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  [[noreturn]] void fail(const std::string& m) {
    throw Error(MGC_ERR_ARG, "XML parse error at byte " + std::to_string(i_) + ": " + m);
  }
  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_to(const char* p) {
    const size_t j = s_.find(p, i_);
    if (j == std::string::npos) fail(std::string("missing ") + p);
    i_ = j + std::strlen(p);
  }
  static std::string unescape(const std::string& t) {
    std::string o;
    for (size_t k = 0; k < t.size(); ++k) {
      if (t[k] != '&') {
        o += t[k];
        continue;
      }
      const size_t e = t.find(';', k);
      const std::string ent = e == std::string::npos ? "" : t.substr(k + 1, e - k - 1);
      if (ent == "lt") o += '<';
      else if (ent == "gt") o += '>';
      else if (ent == "amp") o += '&';
      else if (ent == "quot") o += '"';
      else if (ent == "apos") o += '\'';
      else {
        o += '&';
        continue;
      }
      k = e;
    }
    return o;
  }
  // children of `parent` up to its closing tag </close> (empty: end of input)
  void parse_children(XNode& parent, const std::string& close) {
    std::string text;
    while (i_ < s_.size()) {
      if (s_[i_] != '<') {
        text += s_[i_++];
        continue;
      }
      if (starts("<?")) {
        skip_to("?>");
      } else if (starts("<!--")) {
        skip_to("-->");
      } else if (starts("<!")) {
        skip_to(">");
      } else if (starts("</")) {
        i_ += 2;
        size_t j = i_;
        while (j < s_.size() && s_[j] != '>') ++j;
        std::string name = s_.substr(i_, j - i_);
        while (!name.empty() && std::isspace((unsigned char)name.back())) name.pop_back();
        if (name != close) fail("unexpected </" + name + ">");
        i_ = j + 1;
        parent.text = unescape(text);
        return;
      } else {
        ++i_;
        size_t j = i_;
        while (j < s_.size() && !std::isspace((unsigned char)s_[j]) && s_[j] != '>' &&
               s_[j] != '/')
          ++j;
        auto node = std::make_unique<XNode>();
        node->name = s_.substr(i_, j - i_);
        if (node->name.empty()) fail("empty element name");
        const size_t gt = s_.find('>', j);
        if (gt == std::string::npos) fail("unterminated tag");
        const bool empty = s_[gt - 1] == '/';
        i_ = gt + 1;
        if (!empty) parse_children(*node, node->name);
        parent.kids.push_back(std::move(node));
      }
    }
    if (!close.empty()) fail("missing </" + close + ">");
    parent.text = unescape(text);
  }
};

static std::string trim(const std::string& t) {
  size_t a = 0, b = t.size();
  while (a < b && std::isspace((unsigned char)t[a])) ++a;
  while (b > a && std::isspace((unsigned char)t[b - 1])) --b;
  return t.substr(a, b - a);
}

static const XNode& need(const XNode& n, const std::string& name) {
  const XNode* c = n.child(name);
  MGC_CHECK(c, MGC_ERR_ARG, "XML: <" + n.name + "> has no <" + name + ">");
  return *c;
}
static double as_double(const XNode& n) {
  const std::string t = trim(n.text);
  char* end = nullptr;
  const double v = std::strtod(t.c_str(), &end);
  MGC_CHECK(!t.empty() && end && *end == 0, MGC_ERR_ARG,
            "XML: <" + n.name + "> is not a number: '" + t + "'");
  return v;
}
static int as_int(const XNode& n) {
  const double v = as_double(n);
  MGC_CHECK(v == std::floor(v) && std::fabs(v) < 2e9, MGC_ERR_ARG,
            "XML: <" + n.name + "> is not an integer");
  return (int)v;
}
static int as_bool(const XNode& n) {
  const std::string t = trim(n.text);
  if (t == "true" || t == "1") return 1;
  if (t == "false" || t == "0") return 0;
  throw Error(MGC_ERR_ARG, "XML: <" + n.name + "> is not a boolean: '" + t + "'");
}
// Grid writes std::vector members as a list of <elem> children
static std::vector<const XNode*> elems(const XNode& n) {
  std::vector<const XNode*> out;
  for (auto& k : n.kids)
    if (k->name == "elem") out.push_back(k.get());
  return out;
}

// enum names of include/MultigridParams.h:36-49
static int solver_type(const XNode& n) {
  const std::string t = trim(n.text);
  if (t == "MR") return MGC_SOLVER_MR;
  if (t == "GMRES") return MGC_SOLVER_GMRES;
  if (t == "FGMRES") return MGC_SOLVER_FGMRES;
  if (t == "BiCGSTAB") return MGC_SOLVER_BICGSTAB;
  throw Error(MGC_ERR_ARG, "XML: unknown solver type '" + t + "' in <" + n.name + ">");
}
static const char* solver_name(int t) {
  switch (t) {
    case MGC_SOLVER_MR: return "MR";
    case MGC_SOLVER_GMRES: return "GMRES";
    case MGC_SOLVER_FGMRES: return "FGMRES";
    case MGC_SOLVER_BICGSTAB: return "BiCGSTAB";
  }
  throw Error(MGC_ERR_ARG, "unknown solver type " + std::to_string(t));
}

// One value per level in the file; this build takes one value for all levels
template <typename T, typename F>
static T same_on_levels(const XNode& n, F&& get, int nlev) {
  const auto e = elems(n);
  MGC_CHECK((int)e.size() == nlev, MGC_ERR_ARG,
            "XML: <" + n.name + "> needs nLevels - 1 = " + std::to_string(nlev) + " entries");
  T v = get(*e[0]);
  for (size_t i = 1; i < e.size(); ++i)
    MGC_CHECK(get(*e[i]) == v, MGC_ERR_ARG,
              "XML: <" + n.name + "> differs between levels; this build takes one value");
  return v;
}

struct SolverParams {
  double tol;
  int maxit, restart, rb;
  bool operator==(const SolverParams& o) const {
    return tol == o.tol && maxit == o.maxit && restart == o.restart && rb == o.rb;
  }
};
static SolverParams solver_params(const XNode& n) {
  SolverParams p{as_double(need(n, "tolerance")), as_int(need(n, "maxIter")),
                 as_int(need(n, "restartLength")), as_bool(need(n, "useRBPrec"))};
  MGC_CHECK(!p.rb, MGC_ERR_ARG, "XML: useRBPrec = true is not supported");
  return p;
}

// <mg> (MultiGridParams) into p; members missing from the file keep the defaults
static void read_mg(const XNode& mg, mgc_mg_params* p) {
  p->nlevels = as_int(need(mg, "nLevels"));
  MGC_CHECK(p->nlevels >= 2 && p->nlevels <= 4, MGC_ERR_ARG, "XML: nLevels must be 2..4");
  const int nl = p->nlevels - 1;
  const auto blocks = elems(need(mg, "blockSizes"));
  MGC_CHECK((int)blocks.size() == nl, MGC_ERR_ARG, "XML: blockSizes needs nLevels - 1 entries");
  for (int l = 0; l < nl; ++l) {
    const auto b = elems(*blocks[l]);
    MGC_CHECK(b.size() == 4, MGC_ERR_ARG, "XML: each block size has 4 extents");
    for (int d = 0; d < 4; ++d) p->block[l][d] = as_int(*b[d]);
  }
  if (const XNode* n = mg.child("setupInitialIter"))
    p->setup_iters = same_on_levels<int>(*n, as_int, nl);
  if (const XNode* n = mg.child("setupRefinementIter"))
    p->setup_refine_iters = same_on_levels<int>(*n, as_int, nl);
  if (const XNode* n = mg.child("setupBasisVectorsType")) {
    const std::string t = trim(n->text);
    MGC_CHECK(t == "testVectors" || t == "nullVectors", MGC_ERR_ARG,
              "XML: setupBasisVectorsType must be testVectors or nullVectors");
    p->setup_basis_type = t == "testVectors" ? MGC_BASIS_TEST_VECTORS : MGC_BASIS_NULL_VECTORS;
  }
  if (const XNode* n = mg.child("setupPreOrthonormalise")) p->setup_pre_orthonormalise = as_bool(*n);
  if (const XNode* n = mg.child("setupPostOrthonormalise"))
    p->setup_post_orthonormalise = as_bool(*n);
  if (const XNode* n = mg.child("setupGenerateOnAllLevels"))
    MGC_CHECK(as_bool(*n), MGC_ERR_ARG,
              "XML: setupGenerateOnAllLevels = false (projected basis vectors) is not supported");
  if (const XNode* n = mg.child("setupRecursiveUpdates")) p->setup_recursive_updates = as_bool(*n);
  if (const XNode* n = mg.child("cyclingStrategy")) {
    const std::string c = same_on_levels<std::string>(
        *n, [](const XNode& e) { return trim(e.text); }, nl);
    MGC_CHECK(c == "KCYCLE", MGC_ERR_ARG, "XML: the cycling strategy must be KCYCLE");
  }
  if (const XNode* n = mg.child("setupSolverType"))
    p->setup_solver = same_on_levels<int>(*n, solver_type, nl);
  if (const XNode* n = mg.child("setupSolverParams")) {
    const SolverParams sp = same_on_levels<SolverParams>(*n, solver_params, nl);
    p->setup_tol = sp.tol;
    p->setup_solve_iters = sp.maxit;
    p->setup_restart = sp.restart;
  }
  if (const XNode* n = mg.child("wrapperSolverType"))
    MGC_CHECK(same_on_levels<int>(*n, solver_type, nl) == MGC_SOLVER_FGMRES, MGC_ERR_ARG,
              "XML: the wrapper solver must be FGMRES");
  if (const XNode* n = mg.child("wrapperSolverParams")) {
    const SolverParams sp = same_on_levels<SolverParams>(*n, solver_params, nl);
    p->wrapper_tol = sp.tol;
    p->wrapper_iters = sp.maxit;
    p->wrapper_restart = sp.restart;
  }
  if (const XNode* n = mg.child("smootherSolverType"))
    p->smoother_solver = same_on_levels<int>(*n, solver_type, nl);
  if (const XNode* n = mg.child("smootherSolverParams")) {
    const SolverParams sp = same_on_levels<SolverParams>(*n, solver_params, nl);
    p->smoother_tol = sp.tol;
    p->smoother_iters = sp.maxit;
    p->smoother_restart = sp.restart;
  }
  if (const XNode* n = mg.child("coarsestSolverType")) p->coarsest_solver = solver_type(*n);
  if (const XNode* n = mg.child("coarsestSolverParams")) {
    const SolverParams sp = solver_params(*n);
    p->coarsest_tol = sp.tol;
    p->coarsest_iters = sp.maxit;
    p->coarsest_restart = sp.restart;
  }
}

static void read_test(const XNode& t, mgc_test_params* p) {
  if (const XNode* n = t.child("outerSolverTol")) p->outer_tol = as_double(*n);
  if (const XNode* n = t.child("outerSolverMaxIter")) p->outer_maxit = as_int(*n);
  if (const XNode* n = t.child("outerSolverRestartLength")) p->outer_restart = as_int(*n);
  if (const XNode* n = t.child("massSetup")) p->mass_setup = as_double(*n);
  if (const XNode* n = t.child("massSolve")) p->mass_solve = as_double(*n);
  if (const XNode* n = t.child("csw")) p->csw = as_double(*n);
  if (const XNode* n = t.child("config")) {
    const std::string c = trim(n->text);
    MGC_CHECK(c.size() < sizeof(p->config), MGC_ERR_ARG, "XML: config path too long");
    std::snprintf(p->config, sizeof(p->config), "%s", c.c_str());
  }
  if (const XNode* n = t.child("sourceType")) {
    const std::string c = trim(n->text);
    MGC_CHECK(c == "ones" || c == "random" || c == "gaussian", MGC_ERR_ARG,
              "XML: sourceType must be ones, random or gaussian");
    std::snprintf(p->source_type, sizeof(p->source_type), "%s", c.c_str());
  }
  if (const XNode* n = t.child("useAntiPeriodicBC")) p->antiperiodic_t = as_bool(*n);
}

static std::string fmt(double v) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", v);
  return b;
}

static void write_xml(std::ostream& o, const mgc_mg_params& p, const mgc_test_params& t) {
  const int nl = p.nlevels - 1;
  auto per_level = [&](const std::string& name, const std::string& v) {
    o << "    <" << name << ">\n";
    for (int l = 0; l < nl; ++l) o << "      <elem>" << v << "</elem>\n";
    o << "    </" << name << ">\n";
  };
  auto sp = [&](double tol, int maxit, int restart) {
    return "\n        <tolerance>" + fmt(tol) + "</tolerance>\n        <maxIter>" +
           std::to_string(maxit) + "</maxIter>\n        <restartLength>" +
           std::to_string(restart) + "</restartLength>\n        <useRBPrec>false</useRBPrec>\n      ";
  };
  auto b = [](int v) { return std::string(v ? "true" : "false"); };
  o << "<?xml version=\"1.0\"?>\n<grid>\n  <Params>\n  <mg>\n";
  o << "    <nLevels>" << p.nlevels << "</nLevels>\n    <blockSizes>\n";
  for (int l = 0; l < nl; ++l) {
    o << "      <elem>";
    for (int d = 0; d < 4; ++d) o << "\n        <elem>" << p.block[l][d] << "</elem>";
    o << "\n      </elem>\n";
  }
  o << "    </blockSizes>\n";
  per_level("setupInitialIter", std::to_string(p.setup_iters));
  per_level("setupRefinementIter", std::to_string(p.setup_refine_iters));
  o << "    <setupBasisVectorsType>"
    << (p.setup_basis_type == MGC_BASIS_NULL_VECTORS ? "nullVectors" : "testVectors")
    << "</setupBasisVectorsType>\n";
  o << "    <setupPreOrthonormalise>" << b(p.setup_pre_orthonormalise)
    << "</setupPreOrthonormalise>\n";
  o << "    <setupPostOrthonormalise>" << b(p.setup_post_orthonormalise)
    << "</setupPostOrthonormalise>\n";
  o << "    <setupGenerateOnAllLevels>true</setupGenerateOnAllLevels>\n";
  o << "    <setupRecursiveUpdates>" << b(p.setup_recursive_updates) << "</setupRecursiveUpdates>\n";
  per_level("cyclingStrategy", "KCYCLE");
  per_level("setupSolverType", solver_name(p.setup_solver));
  per_level("setupSolverParams", sp(p.setup_tol, p.setup_solve_iters, p.setup_restart));
  per_level("wrapperSolverType", "FGMRES");
  per_level("wrapperSolverParams", sp(p.wrapper_tol, p.wrapper_iters, p.wrapper_restart));
  per_level("smootherSolverType", solver_name(p.smoother_solver));
  per_level("smootherSolverParams", sp(p.smoother_tol, p.smoother_iters, p.smoother_restart));
  o << "    <coarsestSolverType>" << solver_name(p.coarsest_solver) << "</coarsestSolverType>\n";
  o << "    <coarsestSolverParams>" << sp(p.coarsest_tol, p.coarsest_iters, p.coarsest_restart)
    << "</coarsestSolverParams>\n";
  o << "  </mg>\n  <test>\n";
  o << "    <outerSolverTol>" << fmt(t.outer_tol) << "</outerSolverTol>\n";
  o << "    <outerSolverMaxIter>" << t.outer_maxit << "</outerSolverMaxIter>\n";
  o << "    <outerSolverRestartLength>" << t.outer_restart << "</outerSolverRestartLength>\n";
  o << "    <massSetup>" << fmt(t.mass_setup) << "</massSetup>\n";
  o << "    <massSolve>" << fmt(t.mass_solve) << "</massSolve>\n";
  o << "    <csw>" << fmt(t.csw) << "</csw>\n";
  o << "    <config>" << t.config << "</config>\n";
  o << "    <sourceType>" << t.source_type << "</sourceType>\n";
  o << "    <useAntiPeriodicBC>" << b(t.antiperiodic_t) << "</useAntiPeriodicBC>\n";
  o << "  </test>\n  </Params>\n</grid>\n";
}

// ---------------------------------------------------------------------------
// gauge files
// ---------------------------------------------------------------------------
typedef std::complex<double> C;

static bool host_big_endian() {
  const uint16_t x = 1;
  return *reinterpret_cast<const uint8_t*>(&x) == 0;
}
static void byteswap(void* p, size_t word, size_t n) {
  uint8_t* b = static_cast<uint8_t*>(p);
  for (size_t i = 0; i < n; ++i, b += word)
    for (size_t k = 0; k < word / 2; ++k) std::swap(b[k], b[word - 1 - k]);
}

// average plaquette Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag) / 3
// over x and mu < nu, and the average link trace Re tr(U) / 3 (the NERSC
// header values)
static void plaquette(const C* U, const Lat4& lat, double* plaq, double* ltrace) {
  const int V = lat.volume();
  auto link = [&](int mu, int x) { return U + ((size_t)mu * V + x) * 9; };
  double sp = 0.0, st = 0.0;
  for (int x = 0; x < V; ++x)
    for (int mu = 0; mu < 4; ++mu) {
      const C* a = link(mu, x);
      st += (a[0] + a[4] + a[8]).real();
      for (int nu = mu + 1; nu < 4; ++nu) {
        const C* b = link(nu, lat.shift(x, mu, 1));
        const C* c = link(mu, lat.shift(x, nu, 1));
        const C* d = link(nu, x);
        C ab[9], abc[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            C s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[i * 3 + k] * b[k * 3 + j];
            ab[i * 3 + j] = s;
          }
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            C s = 0.0;
            for (int k = 0; k < 3; ++k) s += ab[i * 3 + k] * std::conj(c[j * 3 + k]);
            abc[i * 3 + j] = s;
          }
        C tr = 0.0;
        for (int i = 0; i < 3; ++i)
          for (int k = 0; k < 3; ++k) tr += abc[i * 3 + k] * std::conj(d[i * 3 + k]);
        sp += tr.real();
      }
    }
  *plaq = sp / (3.0 * 6.0 * V);
  *ltrace = st / (3.0 * 4.0 * V);
}

// NERSC checksum: the sum of the 32-bit words of every site's four 3x3
// double-precision complex matrices in little-endian order
static uint32_t nersc_checksum(const C* U, int V) {
  uint32_t sum = 0;
  const bool be = host_big_endian();
  for (int x = 0; x < V; ++x)
    for (int mu = 0; mu < 4; ++mu) {
      double m[18];
      std::memcpy(m, U + ((size_t)mu * V + x) * 9, sizeof(m));
      if (be) byteswap(m, 8, 18);
      const uint32_t* w = reinterpret_cast<const uint32_t*>(m);
      for (int k = 0; k < 36; ++k) sum += w[k];
    }
  return sum;
}

struct GaugeFile {
  int dims[4] = {0, 0, 0, 0};
  std::vector<C> U;  // [mu][site][3][3]
  bool has_plaq = false, has_trace = false, has_sum = false;
  double plaq = 0.0, ltrace = 0.0;
  uint32_t checksum = 0;
};

static std::string read_all(const char* path) {
  std::ifstream f(path, std::ios::binary);
  MGC_CHECK(f.good(), MGC_ERR_ARG, std::string("cannot open ") + path);
  std::ostringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// file site order (x fastest, then y, z, t; per site mu = 0..3) of a field
// with `rows` stored rows per link in precision `bytes` (4 or 8), byte order
// big (be) or little
static void unpack_links(const char* data, size_t avail, GaugeFile& g, int rows, int bytes,
                         bool be) {
  const Lat4 lat{{g.dims[0], g.dims[1], g.dims[2], g.dims[3]}};
  const int V = lat.volume();
  const size_t per_link = (size_t)rows * 3 * 2 * bytes;
  MGC_CHECK(avail >= per_link * 4 * V, MGC_ERR_ARG, "gauge file: binary data too short");
  g.U.assign((size_t)4 * V * 9, C(0.0, 0.0));
  std::vector<uint8_t> buf(per_link);
  const bool swap = be != host_big_endian();
  for (int x = 0; x < V; ++x)
    for (int mu = 0; mu < 4; ++mu) {
      std::memcpy(buf.data(), data + ((size_t)x * 4 + mu) * per_link, per_link);
      if (swap) byteswap(buf.data(), bytes, per_link / bytes);
      C* u = g.U.data() + ((size_t)mu * V + x) * 9;
      for (int k = 0; k < rows * 3; ++k) {
        double re, im;
        if (bytes == 8) {
          std::memcpy(&re, buf.data() + (2 * k) * 8, 8);
          std::memcpy(&im, buf.data() + (2 * k + 1) * 8, 8);
        } else {
          float fr, fi;
          std::memcpy(&fr, buf.data() + (2 * k) * 4, 4);
          std::memcpy(&fi, buf.data() + (2 * k + 1) * 4, 4);
          re = fr;
          im = fi;
        }
        u[k] = C(re, im);
      }
      if (rows == 2)  // third row = conj(row0 x row1)
        for (int j = 0; j < 3; ++j) {
          const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
          u[6 + j] = std::conj(u[j1] * u[3 + j2] - u[j2] * u[3 + j1]);
        }
    }
}

static void pack_links(const GaugeFile& g, int rows, int bytes, bool be, std::string& out) {
  const Lat4 lat{{g.dims[0], g.dims[1], g.dims[2], g.dims[3]}};
  const int V = lat.volume();
  const size_t per_link = (size_t)rows * 3 * 2 * bytes;
  out.resize(per_link * 4 * V);
  const bool swap = be != host_big_endian();
  std::vector<uint8_t> buf(per_link);
  for (int x = 0; x < V; ++x)
    for (int mu = 0; mu < 4; ++mu) {
      const C* u = g.U.data() + ((size_t)mu * V + x) * 9;
      for (int k = 0; k < rows * 3; ++k) {
        if (bytes == 8) {
          const double re = u[k].real(), im = u[k].imag();
          std::memcpy(buf.data() + (2 * k) * 8, &re, 8);
          std::memcpy(buf.data() + (2 * k + 1) * 8, &im, 8);
        } else {
          const float re = (float)u[k].real(), im = (float)u[k].imag();
          std::memcpy(buf.data() + (2 * k) * 4, &re, 4);
          std::memcpy(buf.data() + (2 * k + 1) * 4, &im, 4);
        }
      }
      if (swap) byteswap(buf.data(), bytes, per_link / bytes);
      std::memcpy(&out[((size_t)x * 4 + mu) * per_link], buf.data(), per_link);
    }
}

static bool read_nersc(const std::string& s, GaugeFile& g) {
  if (s.compare(0, 12, "BEGIN_HEADER") != 0) return false;
  const size_t he = s.find("END_HEADER");
  MGC_CHECK(he != std::string::npos, MGC_ERR_ARG, "NERSC: no END_HEADER");
  const size_t data = s.find('\n', he);
  MGC_CHECK(data != std::string::npos, MGC_ERR_ARG, "NERSC: no data after the header");
  std::map<std::string, std::string> kv;
  std::istringstream hs(s.substr(0, he));
  std::string line;
  while (std::getline(hs, line)) {
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    kv[trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
  }
  for (int d = 0; d < 4; ++d) {
    const auto it = kv.find("DIMENSION_" + std::to_string(d + 1));
    MGC_CHECK(it != kv.end(), MGC_ERR_ARG, "NERSC: no DIMENSION_" + std::to_string(d + 1));
    g.dims[d] = std::atoi(it->second.c_str());
    MGC_CHECK(g.dims[d] >= 1, MGC_ERR_ARG, "NERSC: bad dimension");
  }
  const std::string dt = kv.count("DATATYPE") ? kv["DATATYPE"] : "4D_SU3_GAUGE";
  int rows;
  if (dt == "4D_SU3_GAUGE") rows = 2;
  else if (dt == "4D_SU3_GAUGE_3x3") rows = 3;
  else throw Error(MGC_ERR_ARG, "NERSC: unsupported DATATYPE " + dt);
  const std::string fp = kv.count("FLOATING_POINT") ? kv["FLOATING_POINT"] : "IEEE64BIG";
  int bytes;
  bool be;
  if (fp == "IEEE64BIG") bytes = 8, be = true;
  else if (fp == "IEEE32BIG" || fp == "IEEE32") bytes = 4, be = true;
  else if (fp == "IEEE64LITTLE") bytes = 8, be = false;
  else if (fp == "IEEE32LITTLE") bytes = 4, be = false;
  else throw Error(MGC_ERR_ARG, "NERSC: unsupported FLOATING_POINT " + fp);
  unpack_links(s.data() + data + 1, s.size() - data - 1, g, rows, bytes, be);
  if (kv.count("PLAQUETTE")) g.has_plaq = true, g.plaq = std::atof(kv["PLAQUETTE"].c_str());
  if (kv.count("LINK_TRACE")) g.has_trace = true, g.ltrace = std::atof(kv["LINK_TRACE"].c_str());
  if (kv.count("CHECKSUM")) {
    g.has_sum = true;
    g.checksum = (uint32_t)std::strtoul(kv["CHECKSUM"].c_str(), nullptr, 16);
  }
  return true;
}

// LIME records: 144-byte big-endian header (magic 0x456789ab, version,
// flags, 64-bit length, 128-byte type string), data padded to 8 bytes
static bool read_lime(const std::string& s, GaugeFile& g) {
  auto be32 = [&](size_t o) {
    return (uint32_t)(uint8_t)s[o] << 24 | (uint32_t)(uint8_t)s[o + 1] << 16 |
           (uint32_t)(uint8_t)s[o + 2] << 8 | (uint32_t)(uint8_t)s[o + 3];
  };
  if (s.size() < 144 || be32(0) != 0x456789abu) return false;
  std::string fmt_xml;
  size_t bin_off = 0, bin_len = 0;
  bool have_bin = false;
  for (size_t o = 0; o + 144 <= s.size();) {
    MGC_CHECK(be32(o) == 0x456789abu, MGC_ERR_ARG, "ILDG: bad LIME record magic");
    const uint64_t len = (uint64_t)be32(o + 8) << 32 | be32(o + 12);
    const std::string type(s.c_str() + o + 16, strnlen(s.c_str() + o + 16, 128));
    const size_t d = o + 144;
    MGC_CHECK(d + len <= s.size(), MGC_ERR_ARG, "ILDG: record past the end of the file");
    if (type == "ildg-format") fmt_xml = s.substr(d, len);
    if (type == "ildg-binary-data") {
      bin_off = d;
      bin_len = len;
      have_bin = true;
    }
    o = d + (len + 7) / 8 * 8;
  }
  MGC_CHECK(!fmt_xml.empty() && have_bin, MGC_ERR_ARG,
            "ILDG: needs ildg-format and ildg-binary-data records");
  XmlParser xp(fmt_xml);
  auto doc = xp.parse();
  const XNode* f = doc->child("ildgFormat");
  MGC_CHECK(f, MGC_ERR_ARG, "ILDG: no <ildgFormat>");
  const std::string field = f->child("field") ? trim(f->child("field")->text) : "su3gauge";
  MGC_CHECK(field == "su3gauge", MGC_ERR_ARG, "ILDG: field must be su3gauge");
  const int prec = as_int(need(*f, "precision"));
  MGC_CHECK(prec == 32 || prec == 64, MGC_ERR_ARG, "ILDG: precision must be 32 or 64");
  const char* names[4] = {"lx", "ly", "lz", "lt"};
  for (int d = 0; d < 4; ++d) g.dims[d] = as_int(need(*f, names[d]));
  unpack_links(s.data() + bin_off, bin_len, g, 3, prec / 8, true);
  return true;
}

static void lime_record(std::string& out, const std::string& type, const std::string& data,
                        bool mb, bool me) {
  std::string h(144, '\0');
  auto put32 = [&](size_t o, uint32_t v) {
    for (int k = 0; k < 4; ++k) h[o + k] = (char)(v >> (24 - 8 * k));
  };
  put32(0, 0x456789abu);
  h[4] = 0;
  h[5] = 1;  // version 1
  h[6] = (char)((mb ? 0x80 : 0) | (me ? 0x40 : 0));
  put32(8, (uint32_t)((uint64_t)data.size() >> 32));
  put32(12, (uint32_t)data.size());
  std::memcpy(&h[16], type.c_str(), std::min<size_t>(type.size(), 127));
  out += h;
  out += data;
  out.append((8 - data.size() % 8) % 8, '\0');
}

static void load(const char* path, GaugeFile& g) {
  const std::string s = read_all(path);
  if (read_nersc(s, g)) return;
  if (read_lime(s, g)) return;
  throw Error(MGC_ERR_ARG, std::string("not a NERSC or ILDG gauge file: ") + path);
}

static void write_file(const char* path, const std::string& s) {
  std::ofstream f(path, std::ios::binary);
  MGC_CHECK(f.good(), MGC_ERR_ARG, std::string("cannot write ") + path);
  f.write(s.data(), (std::streamsize)s.size());
  MGC_CHECK(f.good(), MGC_ERR_ARG, std::string("write failed: ") + path);
}

}  // namespace io
}  // namespace mgc

using namespace mgc;
using namespace mgc::io;

extern "C" {

int mgc_test_params_default(mgc_test_params* t) {
  return guarded([&] {
    MGC_CHECK(t, MGC_ERR_ARG, "null params");
    // include/MultigridParams.h:154-176
    t->outer_tol = 1e-12;
    t->outer_maxit = 100;
    t->outer_restart = 20;
    t->mass_setup = 0.1;
    t->mass_solve = 0.1;
    t->csw = 1.0;
    std::snprintf(t->config, sizeof(t->config), "foo");
    std::snprintf(t->source_type, sizeof(t->source_type), "random");
    t->antiperiodic_t = 0;
  });
}

int mgc_params_read_xml(const char* path, mgc_mg_params* mg, mgc_test_params* test) {
  return guarded([&] {
    MGC_CHECK(path && (mg || test), MGC_ERR_ARG, "null argument");
    const std::string s = read_all(path);
    XmlParser xp(s);
    auto doc = xp.parse();
    // <grid><Params><mg/><test/></Params></grid> as Grid's XmlWriter writes
    // MGTestParams, a bare <Params>, or a bare <Multigrid> (MultiGridParams)
    const XNode* top = doc.get();
    if (const XNode* gr = top->child("grid")) top = gr;
    const XNode* mgn = nullptr;
    const XNode* tn = nullptr;
    if (const XNode* pr = top->child("Params")) {
      mgn = pr->child("mg");
      tn = pr->child("test");
    } else {
      mgn = top->child("Multigrid");
    }
    MGC_CHECK(mgn || tn, MGC_ERR_ARG, "XML: no <Params> or <Multigrid> element");
    if (mg) {
      MGC_CHECK(mgn, MGC_ERR_ARG, "XML: no multigrid parameters");
      MGC_CHECK(mgc_mg_params_default(mg) == MGC_OK, MGC_ERR_ARG, "defaults");
      read_mg(*mgn, mg);
    }
    if (test) {
      MGC_CHECK(mgc_test_params_default(test) == MGC_OK, MGC_ERR_ARG, "defaults");
      if (tn) read_test(*tn, test);
    }
  });
}

int mgc_params_write_xml(const char* path, const mgc_mg_params* mg, const mgc_test_params* test) {
  return guarded([&] {
    MGC_CHECK(path && mg, MGC_ERR_ARG, "null argument");
    MGC_CHECK(mg->nlevels >= 2 && mg->nlevels <= 4, MGC_ERR_ARG, "nlevels must be 2..4");
    mgc_test_params t;
    if (test) t = *test;
    else mgc_test_params_default(&t);
    std::ostringstream o;
    write_xml(o, *mg, t);
    write_file(path, o.str());
  });
}

int mgc_gauge_read(const char* path, int dims[4], void* gauge_host_f64, size_t capacity_bytes,
                   double* plaq_out) {
  return guarded([&] {
    MGC_CHECK(path && dims, MGC_ERR_ARG, "null argument");
    GaugeFile g;
    load(path, g);
    for (int d = 0; d < 4; ++d) dims[d] = g.dims[d];
    const Lat4 lat{{g.dims[0], g.dims[1], g.dims[2], g.dims[3]}};
    double plaq = 0.0, ltr = 0.0;
    plaquette(g.U.data(), lat, &plaq, &ltr);
    if (g.has_plaq)
      MGC_CHECK(std::fabs(plaq - g.plaq) <= 1e-6 * std::max(1.0, std::fabs(g.plaq)), MGC_ERR_ARG,
                "gauge file: plaquette " + std::to_string(plaq) + " does not match the header's " +
                    std::to_string(g.plaq));
    if (g.has_trace)
      MGC_CHECK(std::fabs(ltr - g.ltrace) <= 1e-6 * std::max(1.0, std::fabs(g.ltrace)),
                MGC_ERR_ARG, "gauge file: link trace does not match the header");
    if (g.has_sum)
      MGC_CHECK(nersc_checksum(g.U.data(), lat.volume()) == g.checksum, MGC_ERR_ARG,
                "gauge file: checksum does not match the header");
    if (plaq_out) *plaq_out = plaq;
    if (gauge_host_f64) {
      const size_t need = g.U.size() * sizeof(C);
      MGC_CHECK(capacity_bytes >= need, MGC_ERR_SHAPE, "gauge buffer too small");
      std::memcpy(gauge_host_f64, g.U.data(), need);
    }
  });
}

int mgc_gauge_write(const char* path, int format, const int dims[4], const void* gauge_host_f64,
                    int precision_bits) {
  return guarded([&] {
    MGC_CHECK(path && dims && gauge_host_f64, MGC_ERR_ARG, "null argument");
    MGC_CHECK(precision_bits == 32 || precision_bits == 64, MGC_ERR_ARG,
              "precision must be 32 or 64 bits");
    GaugeFile g;
    for (int d = 0; d < 4; ++d) {
      MGC_CHECK(dims[d] >= 1, MGC_ERR_SHAPE, "lattice extent < 1");
      g.dims[d] = dims[d];
    }
    const Lat4 lat{{dims[0], dims[1], dims[2], dims[3]}};
    const int V = lat.volume();
    g.U.resize((size_t)4 * V * 9);
    std::memcpy(g.U.data(), gauge_host_f64, g.U.size() * sizeof(C));
    std::string bin;
    if (format == MGC_GAUGE_NERSC) {
      pack_links(g, 3, precision_bits / 8, true, bin);
      double plaq = 0.0, ltr = 0.0;
      plaquette(g.U.data(), lat, &plaq, &ltr);
      // the checksum is that of the field as a reader reconstructs it
      GaugeFile back;
      for (int d = 0; d < 4; ++d) back.dims[d] = dims[d];
      unpack_links(bin.data(), bin.size(), back, 3, precision_bits / 8, true);
      char head[1024];
      std::snprintf(head, sizeof(head),
                    "BEGIN_HEADER\nHDR_VERSION = 1.0\nDATATYPE = 4D_SU3_GAUGE_3x3\n"
                    "STORAGE_FORMAT = 1.0\nDIMENSION_1 = %d\nDIMENSION_2 = %d\n"
                    "DIMENSION_3 = %d\nDIMENSION_4 = %d\nLINK_TRACE = %.12g\n"
                    "PLAQUETTE = %.12g\nBOUNDARY_1 = PERIODIC\nBOUNDARY_2 = PERIODIC\n"
                    "BOUNDARY_3 = PERIODIC\nBOUNDARY_4 = PERIODIC\nCHECKSUM = %x\n"
                    "ENSEMBLE_ID = libmgc\nSEQUENCE_NUMBER = 1\nFLOATING_POINT = %s\n"
                    "END_HEADER\n",
                    dims[0], dims[1], dims[2], dims[3], ltr, plaq,
                    nersc_checksum(back.U.data(), V),
                    precision_bits == 64 ? "IEEE64BIG" : "IEEE32BIG");
      write_file(path, std::string(head) + bin);
    } else if (format == MGC_GAUGE_ILDG) {
      pack_links(g, 3, precision_bits / 8, true, bin);
      char xml[512];
      std::snprintf(xml, sizeof(xml),
                    "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat>"
                    "<version>1.0</version><field>su3gauge</field><precision>%d</precision>"
                    "<lx>%d</lx><ly>%d</ly><lz>%d</lz><lt>%d</lt></ildgFormat>",
                    precision_bits, dims[0], dims[1], dims[2], dims[3]);
      std::string out;
      lime_record(out, "ildg-format", xml, true, false);
      lime_record(out, "ildg-binary-data", bin, false, false);
      lime_record(out, "ildg-data-lfn", "libmgc", false, true);
      write_file(path, out);
    } else {
      throw Error(MGC_ERR_ARG, "format must be MGC_GAUGE_NERSC or MGC_GAUGE_ILDG");
    }
  });
}

}  // extern "C"
