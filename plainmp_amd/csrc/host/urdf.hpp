// Minimal URDF reader for the kinematic part of a robot description:
// links, and joints with type, parent, child, origin, axis and limits.
// Numerics follow urdfdom as vendored by plainmp: origin rpy converted to a
// normalised quaternion (third/urdf_parser/include/urdf_model/pose.h:202-250),
// joint axis kept as written with default (1, 0, 0) (src/joint.cpp:427-452),
// missing limit bounds read as 0 (src/joint.cpp:101-130).
#pragma once

#include <cctype>
#include <cmath>
#include <cstring>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace pmp_host {

struct XmlNode {
  std::string tag;
  std::map<std::string, std::string> attr;
  std::vector<XmlNode> children;
  const XmlNode* child(const std::string& t) const {
    for (auto& c : children)
      if (c.tag == t) return &c;
    return nullptr;
  }
  const std::string* get(const std::string& k) const {
    auto it = attr.find(k);
    return it == attr.end() ? nullptr : &it->second;
  }
};

class XmlReader {
 public:
  explicit XmlReader(const std::string& s) : s_(s) {}
  XmlNode parse_document() {
    skip_misc();
    XmlNode root = parse_element();
    return root;
  }

 private:
  const std::string& s_;
  size_t i_ = 0;
  [[noreturn]] void fail(const std::string& m) const {
    throw std::runtime_error("URDF parse error at offset " + std::to_string(i_) + ": " + m);
  }
  void skip_ws() {
    while (i_ < s_.size() && std::isspace((unsigned char)s_[i_])) ++i_;
  }
  bool starts(const char* p) const { return s_.compare(i_, std::strlen(p), p) == 0; }
  void skip_misc() {  // whitespace, <?...?>, <!-- ... -->, <!DOCTYPE ...>
    for (;;) {
      skip_ws();
      if (starts("<?")) {
        size_t e = s_.find("?>", i_);
        if (e == std::string::npos) fail("unterminated processing instruction");
        i_ = e + 2;
      } else if (starts("<!--")) {
        size_t e = s_.find("-->", i_);
        if (e == std::string::npos) fail("unterminated comment");
        i_ = e + 3;
      } else if (starts("<!")) {
        size_t e = s_.find('>', i_);
        if (e == std::string::npos) fail("unterminated declaration");
        i_ = e + 1;
      } else {
        return;
      }
    }
  }
  std::string name() {
    size_t b = i_;
    while (i_ < s_.size() && (std::isalnum((unsigned char)s_[i_]) || s_[i_] == '_' || s_[i_] == ':' ||
                              s_[i_] == '-' || s_[i_] == '.'))
      ++i_;
    if (b == i_) fail("expected a name");
    return s_.substr(b, i_ - b);
  }
  XmlNode parse_element() {
    if (i_ >= s_.size() || s_[i_] != '<') fail("expected '<'");
    ++i_;
    XmlNode n;
    n.tag = name();
    for (;;) {
      skip_ws();
      if (starts("/>")) {
        i_ += 2;
        return n;
      }
      if (starts(">")) {
        ++i_;
        break;
      }
      std::string k = name();
      skip_ws();
      if (i_ >= s_.size() || s_[i_] != '=') fail("expected '='");
      ++i_;
      skip_ws();
      if (i_ >= s_.size() || (s_[i_] != '"' && s_[i_] != '\'')) fail("expected a quoted value");
      char q = s_[i_++];
      size_t e = s_.find(q, i_);
      if (e == std::string::npos) fail("unterminated attribute");
      n.attr[k] = s_.substr(i_, e - i_);
      i_ = e + 1;
    }
    for (;;) {  // content
      size_t lt = s_.find('<', i_);
      if (lt == std::string::npos) fail("unterminated element <" + n.tag + ">");
      i_ = lt;
      if (starts("<!--")) {
        size_t e = s_.find("-->", i_);
        if (e == std::string::npos) fail("unterminated comment");
        i_ = e + 3;
      } else if (starts("</")) {
        i_ += 2;
        std::string t = name();
        if (t != n.tag) fail("mismatched closing tag </" + t + "> for <" + n.tag + ">");
        skip_ws();
        if (i_ >= s_.size() || s_[i_] != '>') fail("expected '>'");
        ++i_;
        return n;
      } else if (starts("<![CDATA[")) {
        size_t e = s_.find("]]>", i_);
        if (e == std::string::npos) fail("unterminated CDATA");
        i_ = e + 3;
      } else {
        n.children.push_back(parse_element());
      }
    }
  }
};

struct UJoint {
  std::string name, type, parent, child;
  double xyz[3] = {0, 0, 0};
  double quat[4] = {0, 0, 0, 1};  // x, y, z, w (normalised)
  double axis[3] = {1, 0, 0};
  double lower = 0, upper = 0;
};

struct UModel {
  std::vector<std::string> links;
  std::vector<UJoint> joints;
};

inline void parse_vec3(const std::string& s, double out[3]) {
  std::istringstream is(s);
  for (int k = 0; k < 3; ++k)
    if (!(is >> out[k])) throw std::runtime_error("malformed vector '" + s + "'");
}

// urdf::Rotation::setFromRPY followed by normalize().
inline void urdf_rpy_to_quat(double r, double p, double y, double q[4]) {
  double phi = r / 2.0, the = p / 2.0, psi = y / 2.0;
  double x = std::sin(phi) * std::cos(the) * std::cos(psi) - std::cos(phi) * std::sin(the) * std::sin(psi);
  double yy = std::cos(phi) * std::sin(the) * std::cos(psi) + std::sin(phi) * std::cos(the) * std::sin(psi);
  double z = std::cos(phi) * std::cos(the) * std::sin(psi) - std::sin(phi) * std::sin(the) * std::cos(psi);
  double w = std::cos(phi) * std::cos(the) * std::cos(psi) + std::sin(phi) * std::sin(the) * std::sin(psi);
  double s = std::sqrt(x * x + yy * yy + z * z + w * w);
  if (s == 0.0) {
    q[0] = q[1] = q[2] = 0.0;
    q[3] = 1.0;
  } else {
    q[0] = x / s;
    q[1] = yy / s;
    q[2] = z / s;
    q[3] = w / s;
  }
}

inline UModel parse_urdf(const std::string& xml) {
  if (xml.empty()) throw std::runtime_error("xml string is empty");
  XmlReader rd(xml);
  XmlNode root = rd.parse_document();
  if (root.tag != "robot") throw std::runtime_error("URDF root element is not <robot>");
  UModel m;
  for (auto& c : root.children) {
    if (c.tag == "link") {
      auto* n = c.get("name");
      if (!n) throw std::runtime_error("link without a name");
      m.links.push_back(*n);
    } else if (c.tag == "joint") {
      UJoint j;
      auto* n = c.get("name");
      auto* t = c.get("type");
      if (!n || !t) throw std::runtime_error("joint without name or type");
      j.name = *n;
      j.type = *t;
      auto* pa = c.child("parent");
      auto* ch = c.child("child");
      if (!pa || !ch || !pa->get("link") || !ch->get("link"))
        throw std::runtime_error("Joint [" + j.name + "] is missing a parent and/or child link specification.");
      j.parent = *pa->get("link");
      j.child = *ch->get("link");
      if (auto* o = c.child("origin")) {
        if (auto* x = o->get("xyz")) parse_vec3(*x, j.xyz);
        double rpy[3] = {0, 0, 0};
        if (auto* r = o->get("rpy")) parse_vec3(*r, rpy);
        urdf_rpy_to_quat(rpy[0], rpy[1], rpy[2], j.quat);
      }
      if (j.type != "fixed" && j.type != "floating") {
        if (auto* a = c.child("axis"))
          if (auto* x = a->get("xyz")) parse_vec3(*x, j.axis);
      }
      if (auto* l = c.child("limit")) {
        if (auto* lo = l->get("lower")) j.lower = std::stod(*lo);
        if (auto* up = l->get("upper")) j.upper = std::stod(*up);
      } else if (j.type == "revolute" || j.type == "prismatic") {
        throw std::runtime_error("joint [" + j.name + "] of type " + j.type + " has no <limit>");
      }
      m.joints.push_back(j);
    }
  }
  return m;
}

}  // namespace pmp_host
