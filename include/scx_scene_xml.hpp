// scx_scene_xml.hpp -- scene-file front end: reads the reference's XML scene files
// (<ball2d_scene>, ball2dutils/Ball2DSceneParser.cpp:1486-1651; <rigidbody3d_scene> of
// sphere bodies, rigidbody3dutils/RigidBody3DSceneParser.cpp:126-484,1152-1210,1742-1965,
// 2202-...) into the arrays and step parameters of the scx C ABI, so that reference assets
// run unchanged on the MI355X pipeline. Header-only C++17 with its own small XML reader
// (elements, attributes, comments, declarations; the reference uses RapidXML).
//
// Supported, with the reference's attribute names, required attributes and value checks:
//   ball2d:      <integrator type="symplectic_euler|verlet" dt>, <sobogus_friction_solver>,
//                <gravity fx fy> (several add up), <static_plane x n>, <ball x y vx vy m r fixed>,
//                <end_time t>, <camera> (ignored)
//   rigidbody3d: <integrator type="split_ham|dmv" dt> (their translational parts are the same
//                velocity Verlet), <sobogus_friction_solver>, <near_earth_gravity f>,
//                <static_plane x n>, <geometry type="sphere" r>,
//                <rigid_body_with_density x v omega rho fixed geo_idx [R]>, <end_time t>,
//                <camera_perspective>, <camera_orthographic>, <static_plane_renderer> (ignored)
//   <sobogus_friction_solver mu CoR max_iters tol eval_every staggering
//                [cache_impulses] [stabilization penetration_threshold]>
// Anything else the reference accepts but this pipeline does not run (fixed bodies, drums,
// portals, scripting, penalty forces, other solvers and geometries, the exponential Euler
// map) is refused with an Error naming it, rather than silently dropped.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "scx.h"

namespace scx_scisim {
namespace xml {

class ParseError : public std::runtime_error {
 public:
  explicit ParseError(const std::string& what) : std::runtime_error(what) {}
};

struct Element {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  std::vector<Element> children;

  const std::string* attr(const std::string& key) const {
    for (const auto& a : attrs)
      if (a.first == key) return &a.second;
    return nullptr;
  }
  std::vector<const Element*> all(const std::string& child) const {
    std::vector<const Element*> out;
    for (const Element& e : children)
      if (e.name == child) out.push_back(&e);
    return out;
  }
  const Element* first(const std::string& child) const {
    for (const Element& e : children)
      if (e.name == child) return &e;
    return nullptr;
  }
};

// Recursive-descent reader of one XML document: prolog (declaration, comments), one root
// element; character data is skipped, entities in attribute values are decoded.
class Reader {
 public:
  explicit Reader(const std::string& text) : s_(text) {}

  Element document() {
    skip_misc();
    if (!at("<")) fail("expected the root element");
    Element root = element();
    skip_misc();
    if (p_ != s_.size()) fail("content after the root element");
    return root;
  }

 private:
  const std::string& s_;
  size_t p_ = 0;

  [[noreturn]] void fail(const std::string& what) const {
    size_t line = 1;
    for (size_t k = 0; k < p_ && k < s_.size(); ++k) line += s_[k] == '\n';
    throw ParseError("XML: " + what + " (line " + std::to_string(line) + ")");
  }
  bool at(const char* lit) const { return s_.compare(p_, std::strlen(lit), lit) == 0; }
  static bool space(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r'; }
  void skip_space() {
    while (p_ < s_.size() && space(s_[p_])) ++p_;
  }
  void skip_until(const char* end) {
    const size_t q = s_.find(end, p_);
    if (q == std::string::npos) fail(std::string("unterminated construct, missing '") + end + "'");
    p_ = q + std::strlen(end);
  }
  // whitespace, comments, declarations and processing instructions between elements
  void skip_misc() {
    for (;;) {
      skip_space();
      if (at("<!--"))
        skip_until("-->");
      else if (at("<?"))
        skip_until("?>");
      else if (at("<!"))
        skip_until(">");
      else
        return;
    }
  }
  std::string name() {
    const size_t b = p_;
    while (p_ < s_.size() && !space(s_[p_]) && s_[p_] != '=' && s_[p_] != '>' && s_[p_] != '/' && s_[p_] != '<')
      ++p_;
    if (p_ == b) fail("expected a name");
    return s_.substr(b, p_ - b);
  }
  static std::string decode(const std::string& v) {
    std::string out;
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] != '&') {
        out += v[k];
        continue;
      }
      static const char* names[] = {"&lt;", "&gt;", "&amp;", "&quot;", "&apos;"};
      static const char chars[] = {'<', '>', '&', '"', '\''};
      bool hit = false;
      for (int e = 0; e < 5 && !hit; ++e)
        if (v.compare(k, std::strlen(names[e]), names[e]) == 0) {
          out += chars[e];
          k += std::strlen(names[e]) - 1;
          hit = true;
        }
      if (!hit) out += '&';
    }
    return out;
  }
  Element element() {
    ++p_;  // '<'
    Element e;
    e.name = name();
    for (;;) {
      skip_space();
      if (p_ >= s_.size()) fail("unterminated start tag <" + e.name + ">");
      if (at("/>")) {
        p_ += 2;
        return e;
      }
      if (s_[p_] == '>') {
        ++p_;
        break;
      }
      std::string key = name();
      skip_space();
      if (!at("=")) fail("attribute " + key + " of <" + e.name + "> has no value");
      ++p_;
      skip_space();
      if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\'')) fail("attribute " + key + " is not quoted");
      const char q = s_[p_++];
      const size_t end = s_.find(q, p_);
      if (end == std::string::npos) fail("unterminated value of attribute " + key);
      if (e.attr(key)) fail("attribute " + key + " repeated in <" + e.name + ">");
      e.attrs.emplace_back(key, decode(s_.substr(p_, end - p_)));
      p_ = end + 1;
    }
    // content: child elements, comments and character data, up to the end tag
    for (;;) {
      while (p_ < s_.size() && s_[p_] != '<') ++p_;
      if (p_ >= s_.size()) fail("missing end tag </" + e.name + ">");
      if (at("<!--")) {
        skip_until("-->");
      } else if (at("<![CDATA[")) {
        skip_until("]]>");
      } else if (at("<?")) {
        skip_until("?>");
      } else if (at("</")) {
        p_ += 2;
        const std::string closing = name();
        if (closing != e.name) fail("end tag </" + closing + "> does not close <" + e.name + ">");
        skip_space();
        if (!at(">")) fail("malformed end tag </" + closing + ">");
        ++p_;
        return e;
      } else {
        e.children.push_back(element());
      }
    }
  }
};

inline Element parse(const std::string& text) { return Reader(text).document(); }

}  // namespace xml

// A scene file's contents for the scx C ABI.
struct SceneFile {
  int dim = 2;
  // bodies, in file order: dim per body for q and v; 3 per body for omega (3D)
  std::vector<double> q, v, omega, r, m, inertia;
  std::vector<double> plane_x, plane_n;  // dim per static plane, as written (normalised by the library)
  double g[3] = {0.0, 0.0, 0.0};         // sum of the gravity forces
  std::string integrator;                // as written
  std::string dt_string;
  double dt = 0.0;
  double end_time = std::numeric_limits<double>::infinity();
  // <sobogus_friction_solver>
  double mu = 0.0, CoR = 0.0, tol = 0.0, penetration_threshold = 0.0;
  unsigned max_iters = 0, eval_every = 0;
  std::string staggering, cache_impulses;
  bool stabilization = false;

  // The scx step parameters of this scene (FP64 sweeps, as the reference solves).
  scx_step_params params(uint32_t iteration = 1) const {
    scx_step_params p{};
    p.dt = dt;
    for (int a = 0; a < 3; ++a) p.g[a] = g[a];
    p.mu = mu;
    p.CoR = CoR;
    p.tol = tol;
    p.max_iters = max_iters;
    p.eval_every = eval_every;
    p.integrator = (integrator == "symplectic_euler") ? SCX_INTEGRATOR_SYMPLECTIC_EULER : SCX_INTEGRATOR_VERLET;
    p.precision = SCX_PRECISION_FP64;
    p.iteration = iteration;
    p.staggering = staggering == "symplectic_euler" ? SCX_STAGGER_SYMPLECTIC_EULER
                   : staggering == "stabilized"     ? SCX_STAGGER_STABILIZED
                                                    : SCX_STAGGER_GEOMETRIC;
    p.stabilize = (p.staggering == SCX_STAGGER_SYMPLECTIC_EULER && stabilization) ? 1 : 0;
    p.penetration_threshold = penetration_threshold;
    // ImpulsesToCache: the 3D geometric map never caches
    // (RigidBody3DSceneParser.cpp:1877-1880); the 3D symplectic Euler map's
    // normal_and_friction cache is the transported one (SymplecticEulerImpactFrictionMap.cpp:87-127)
    p.warm_start = SCX_WARM_NONE;
    if (cache_impulses == "normal" && !(dim == 3 && p.staggering == SCX_STAGGER_GEOMETRIC))
      p.warm_start = SCX_WARM_NORMAL;
    if (cache_impulses == "normal_and_friction" && !(dim == 3 && p.staggering == SCX_STAGGER_GEOMETRIC))
      p.warm_start = (dim == 2) ? SCX_WARM_CACHE : SCX_WARM_TRANSPORT;
    return p;
  }
  uint32_t nbodies() const { return (uint32_t)r.size(); }
  uint32_t nplanes() const { return (uint32_t)(plane_x.size() / (size_t)dim); }
};

namespace scene_detail {

class SceneError : public std::runtime_error {
 public:
  explicit SceneError(const std::string& what) : std::runtime_error(what) {}
};

[[noreturn]] inline void refuse(const std::string& what) { throw SceneError(what); }

inline const std::string& need(const xml::Element& e, const char* key) {
  const std::string* v = e.attr(key);
  if (!v) refuse("could not locate " + std::string(key) + " for " + e.name);
  return *v;
}

// StringUtilities::extractFromString: a stream extraction of the whole leading token
inline bool to_double(const std::string& s, double& out) {
  std::istringstream is(s);
  is >> out;
  return !is.fail();
}
inline bool to_int(const std::string& s, long long& out) {
  std::istringstream is(s);
  is >> out;
  return !is.fail();
}
inline bool to_bool(const std::string& s, bool& out) {
  std::istringstream is(s);
  is >> out;  // "0" / "1", as a stream of bool reads them
  return !is.fail();
}
inline double number(const xml::Element& e, const char* key) {
  double x;
  if (!to_double(need(e, key), x)) refuse("could not load " + std::string(key) + " value for " + e.name);
  return x;
}
// whitespace-separated scalars, exactly `count` of them (readScalarList / the stream loops)
inline std::vector<double> list(const xml::Element& e, const char* key, size_t count) {
  std::istringstream is(need(e, key));
  std::vector<double> out;
  double x;
  while (is >> x) out.push_back(x);
  if (out.size() != count)
    refuse("invalid number of components for " + e.name + " " + key + ": " + std::to_string(count) + " required");
  return out;
}

// Rational dt (scisim/Math/Rational.h:204-258): "a/b", or a decimal read as an exact
// fraction over a power of ten, reduced, then converted as double(num) / double(den).
inline double rational(const std::string& text) {
  std::string s;
  for (char c : text)
    if (c != ' ' && c != '\t' && c != '\n' && c != '\r') s += c;
  long long num = 0, den = 1;
  const size_t slash = s.rfind('/');
  if (slash != std::string::npos) {
    if (s.find('.') != std::string::npos || s.find('/') != slash) refuse("invalid rational dt '" + text + "'");
    if (!to_int(s.substr(0, slash), num) || !to_int(s.substr(slash + 1), den)) refuse("invalid dt '" + text + "'");
  } else {
    const size_t dot = s.rfind('.');
    std::string left = (dot == std::string::npos) ? s : s.substr(0, dot);
    std::string right = (dot == std::string::npos) ? std::string() : s.substr(dot + 1);
    while (!right.empty() && right.back() == '0') right.pop_back();
    if (left.empty() || left == "+" || left == "-") left += "0";
    long long l = 0, rr = 0;
    if (!to_int(left, l) || (!right.empty() && !to_int(right, rr))) refuse("invalid dt '" + text + "'");
    if (right.size() > 18) refuse("dt '" + text + "' has too many decimals");
    for (size_t k = 0; k < right.size(); ++k) den *= 10;
    num = l * den + (l < 0 ? -1 : 1) * rr;
  }
  if (den == 0) refuse("dt '" + text + "' has a zero denominator");
  if (den < 0) {
    num = -num;
    den = -den;
  }
  long long a = num < 0 ? -num : num, b = den;
  while (b != 0) {
    const long long t = a % b;
    a = b;
    b = t;
  }
  if (a > 1) {
    num /= a;
    den /= a;
  }
  return double(num) / double(den);
}

inline void load_solver(const xml::Element& root, SceneFile& sc) {
  for (const char* other : {"impact_operator", "staggered_projections_friction_solver", "grr_friction_solver"})
    if (root.first(other)) refuse(std::string("<") + other + "> is not supported (only <sobogus_friction_solver>)");
  const xml::Element* s = root.first("sobogus_friction_solver");
  if (!s) refuse("the scene has no <sobogus_friction_solver>");
  sc.mu = number(*s, "mu");
  if (sc.mu < 0.0) refuse("value of mu must be a nonnegative scalar");
  sc.CoR = number(*s, "CoR");
  if (sc.CoR < 0.0 || sc.CoR > 1.0) refuse("value of CoR must lie in [0, 1]");
  long long mi, ee;
  if (!to_int(need(*s, "max_iters"), mi) || mi <= 0) refuse("value of max_iters must be a positive integer");
  if (!to_int(need(*s, "eval_every"), ee) || ee <= 0 || ee > mi)
    refuse("value of eval_every must be a positive integer no larger than max_iters");
  sc.max_iters = (unsigned)mi;
  sc.eval_every = (unsigned)ee;
  sc.tol = number(*s, "tol");
  if (sc.tol < 0.0) refuse("value of tol must be a nonnegative scalar");
  sc.staggering = need(*s, "staggering");
  if (sc.staggering != "geometric" && sc.staggering != "symplectic_euler" && sc.staggering != "stabilized")
    refuse("invalid staggering attribute '" + sc.staggering + "'");
  // ball2d reads cache_impulses for every staggering (Ball2DSceneParser.cpp:1319-1345);
  // rigidbody3d only for symplectic_euler (RigidBody3DSceneParser.cpp:1926-1951)
  const bool want_cache = sc.dim == 2 || sc.staggering == "symplectic_euler";
  sc.cache_impulses = "none";
  if (want_cache) {
    sc.cache_impulses = need(*s, "cache_impulses");
    if (sc.cache_impulses != "none" && sc.cache_impulses != "normal" && sc.cache_impulses != "normal_and_friction")
      refuse("invalid option specified for cache_impulses: '" + sc.cache_impulses + "'");
  }
  if (sc.staggering == "symplectic_euler") {
    if (sc.CoR != 0.0) refuse("CoR must be set to 0 for the symplectic_euler staggering");
    if (!to_bool(need(*s, "stabilization"), sc.stabilization)) refuse("value of stabilization must be a boolean");
    sc.penetration_threshold = number(*s, "penetration_threshold");
    if (sc.penetration_threshold < 0.0) refuse("value of penetration_threshold must be a non-negative scalar");
  }
  if (sc.staggering == "stabilized" && sc.cache_impulses != "none")
    refuse("constraint cache not supported for the stabilized impact friction map");
}

inline void load_common(const xml::Element& root, SceneFile& sc) {
  if (const xml::Element* s = root.first("scripting")) {
    const std::string* cb = s->attr("callback");
    refuse("scripting callbacks are not supported (callback '" + (cb ? *cb : std::string()) + "')");
  }
  if (const xml::Element* e = root.first("end_time")) {
    sc.end_time = number(*e, "t");
    if (!(sc.end_time > 0.0)) refuse("end_time must be a positive scalar");
  }
  const xml::Element* integ = root.first("integrator");
  if (!integ) refuse("failed to locate integrator node");
  sc.dt_string = need(*integ, "dt");
  sc.dt = rational(sc.dt_string);
  if (!(sc.dt > 0.0)) refuse("dt must be positive");
  sc.integrator = need(*integ, "type");
  for (const xml::Element* p : root.all("static_plane")) {
    const std::vector<double> x = list(*p, "x", (size_t)sc.dim), n = list(*p, "n", (size_t)sc.dim);
    sc.plane_x.insert(sc.plane_x.end(), x.begin(), x.end());
    sc.plane_n.insert(sc.plane_n.end(), n.begin(), n.end());
  }
  for (const char* other : {"planar_portal", "lees_edwards_portal", "static_drum", "static_cylinder", "penalty_force",
                            "hertzian_penalty", "simulation_boundary", "kinematic_script"})
    if (root.first(other)) refuse(std::string("<") + other + "> is not supported");
  load_solver(root, sc);
}

inline SceneFile load_ball2d(const xml::Element& root) {
  SceneFile sc;
  sc.dim = 2;
  load_common(root, sc);
  if (sc.integrator != "symplectic_euler" && sc.integrator != "verlet")
    refuse("invalid integrator type '" + sc.integrator + "' (options: verlet, symplectic_euler)");
  for (const xml::Element* gr : root.all("gravity")) {
    sc.g[0] += number(*gr, "fx");
    sc.g[1] += number(*gr, "fy");
  }
  for (const xml::Element* b : root.all("ball")) {
    bool fixed;
    if (!to_bool(need(*b, "fixed"), fixed)) refuse("could not load fixed for ball");
    if (fixed) refuse("fixed balls are not supported");
    sc.q.push_back(number(*b, "x"));
    sc.q.push_back(number(*b, "y"));
    sc.v.push_back(number(*b, "vx"));
    sc.v.push_back(number(*b, "vy"));
    sc.m.push_back(number(*b, "m"));
    sc.r.push_back(number(*b, "r"));
  }
  return sc;
}

inline SceneFile load_rigidbody3d(const xml::Element& root) {
  SceneFile sc;
  sc.dim = 3;
  load_common(root, sc);
  if (sc.integrator == "exponential_euler") refuse("the exponential_euler integrator is not supported");
  if (sc.integrator != "split_ham" && sc.integrator != "dmv")
    refuse("invalid integrator type '" + sc.integrator + "' (options: split_ham, dmv)");
  for (const xml::Element* gr : root.all("near_earth_gravity")) {
    const std::vector<double> f = list(*gr, "f", 3);
    for (int a = 0; a < 3; ++a) sc.g[a] += f[a];
  }
  // geometry list: spheres only (others are refused when a body uses them)
  std::vector<double> radius;  // < 0: not a sphere
  std::vector<std::string> kind;
  for (const xml::Element* ge : root.all("geometry")) {
    const std::string& t = need(*ge, "type");
    kind.push_back(t);
    if (t == "sphere") {
      const double r = number(*ge, "r");
      if (!(r > 0.0)) refuse("sphere radius must be positive");
      radius.push_back(r);
    } else {
      radius.push_back(-1.0);
    }
  }
  for (const xml::Element* b : root.all("rigid_body_with_density")) {
    const std::vector<double> x = list(*b, "x", 3), v = list(*b, "v", 3), w = list(*b, "omega", 3);
    const double rho = number(*b, "rho");
    bool fixed;
    if (!to_bool(need(*b, "fixed"), fixed)) refuse("could not load fixed for rigid_body_with_density");
    long long gi;
    if (!to_int(need(*b, "geo_idx"), gi)) refuse("could not load geo_idx");
    const long long ng = (long long)radius.size();
    if (gi < 0) {
      if (-gi > ng) refuse("geo_idx out of range");
      gi += ng;
    } else if (gi >= ng) {
      refuse("geo_idx out of range");
    }
    if (radius[(size_t)gi] < 0.0) refuse("geometry type '" + kind[(size_t)gi] + "' is not supported (spheres only)");
    if (fixed) refuse("fixed bodies are not supported");
    if (b->attr("R")) list(*b, "R", 3);  // orientation: checked, irrelevant for a sphere
    const double r = radius[(size_t)gi];
    // RigidBodySphere::computeMassAndInertia (rigidbody3d/Geometry/RigidBodySphere.cpp:10-18)
    const double pi = 3.14159265358979323846264338327950288;
    double M = 4.0 * pi * r * r * r / 3.0;
    double I = 2.0 * M * r * r / 5.0;
    M *= rho;
    I *= rho;
    for (int a = 0; a < 3; ++a) {
      sc.q.push_back(x[a]);
      sc.v.push_back(v[a]);
      sc.omega.push_back(w[a]);
    }
    sc.r.push_back(r);
    sc.m.push_back(M);
    sc.inertia.push_back(I);
  }
  return sc;
}

}  // namespace scene_detail

// Parses a scene document (the text of a .xml file). Throws scene_detail::SceneError or
// xml::ParseError with the reason.
inline SceneFile parseSceneText(const std::string& text) {
  const xml::Element root = xml::parse(text);
  if (root.name == "ball2d_scene") return scene_detail::load_ball2d(root);
  if (root.name == "rigidbody3d_scene") return scene_detail::load_rigidbody3d(root);
  scene_detail::refuse("unknown scene root <" + root.name + "> (ball2d_scene or rigidbody3d_scene)");
}

inline SceneFile parseSceneFile(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) scene_detail::refuse("failed to open scene file " + path);
  std::ostringstream ss;
  ss << in.rdbuf();
  return parseSceneText(ss.str());
}

}  // namespace scx_scisim
