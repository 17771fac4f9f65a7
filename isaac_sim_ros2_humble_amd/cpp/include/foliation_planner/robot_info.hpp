// RobotInfo without ROS or MoveIt: the flat robot arrays the planner API takes
// (SingleArmSpace / CPRRTC::RobotSolver), read from the files MoveIt builds its robot model from:
// a URDF, an SRDF (planning group, disabled collision pairs) and the collision-sphere YAML.  The
// reference builds the same arrays from a MoveIt RobotModel
// (src/foliation_planner/include/foliation_planner/robot_info.hpp:13-310); this class keeps its
// getters and reproduces the model semantics it inherits from MoveIt and urdfdom:
//   * links in depth-first pre-order from the URDF root, a link's children in the order of their
//     joints' names (urdfdom keeps joints in a name-keyed map);
//   * one joint per link; the root gets type 0, identity origin, zero axis;
//   * joint types as MoveIt numbers them: 1 revolute (URDF revolute and continuous), 2 prismatic,
//     5 fixed;
//   * origins: translation, then the rotation of the roll-pitch-yaw angles via a unit quaternion
//     (urdf::Rotation::setFromRPY, then Eigen's quaternion-to-matrix);
//   * axes normalised (MoveIt's joint models); URDF's default axis is (1, 0, 0);
//   * bounds of the group's active joints from the URDF limits, continuous joints [-pi, pi], other
//     joints 0; default values: the default-value map's entry, else 0 when 0 is within the
//     bounds, else the bounds' midpoint (MoveIt's default positions); fixed joints 0;
//   * collision spheres in YAML order, per link; spheres of links the model lacks are skipped;
//   * self-collision enable map all true, with the SRDF's disabled pairs cleared symmetrically.
// Only what these files need is parsed: XML elements and attributes, and the sphere YAML's
// "collision_spheres:" block of link keys with "- center: [x, y, z]" / "radius: r" items.
#pragma once

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../mp_eigen.h"

namespace robot_info_detail {

struct XmlNode {
    std::string tag;
    std::map<std::string, std::string> attr;
    std::vector<std::unique_ptr<XmlNode>> children;
    std::string get(const std::string& k, const std::string& def = "") const {
        auto it = attr.find(k);
        return it == attr.end() ? def : it->second;
    }
    const XmlNode* child(const std::string& t) const {
        for (const auto& c : children)
            if (c->tag == t) return c.get();
        return nullptr;
    }
};

// Element tree of an XML document: declarations, comments, DOCTYPE and text are skipped.
class XmlReader {
  public:
    explicit XmlReader(const std::string& text) : s_(text) {}
    std::unique_ptr<XmlNode> parse() {
        auto root = std::make_unique<XmlNode>();
        root->tag = "#document";
        std::vector<XmlNode*> stack{root.get()};
        while (true) {
            const size_t lt = s_.find('<', p_);
            if (lt == std::string::npos) break;
            p_ = lt;
            if (s_.compare(p_, 4, "<!--") == 0) {
                p_ = need(s_.find("-->", p_), "comment") + 3;
            } else if (s_.compare(p_, 2, "<?") == 0) {
                p_ = need(s_.find("?>", p_), "declaration") + 2;
            } else if (s_.compare(p_, 2, "<!") == 0) {
                p_ = need(s_.find('>', p_), "declaration") + 1;
            } else if (s_.compare(p_, 2, "</") == 0) {
                const size_t gt = need(s_.find('>', p_), "closing tag");
                std::string name = trim(s_.substr(p_ + 2, gt - p_ - 2));
                if (stack.size() < 2 || stack.back()->tag != name) throw std::runtime_error("XML: unbalanced </" + name + ">");
                stack.pop_back();
                p_ = gt + 1;
            } else {
                ++p_;
                auto node = std::make_unique<XmlNode>();
                node->tag = name_at();
                bool self_closing = false;
                while (true) {
                    skip_space();
                    if (p_ >= s_.size()) throw std::runtime_error("XML: unterminated <" + node->tag + ">");
                    if (s_[p_] == '/') {
                        self_closing = true;
                        p_ = need(s_.find('>', p_), "tag") + 1;
                        break;
                    }
                    if (s_[p_] == '>') {
                        ++p_;
                        break;
                    }
                    const std::string key = name_at();
                    skip_space();
                    if (p_ >= s_.size() || s_[p_] != '=') throw std::runtime_error("XML: attribute " + key + " without value");
                    ++p_;
                    skip_space();
                    const char q = s_[p_];
                    if (q != '"' && q != '\'') throw std::runtime_error("XML: unquoted attribute " + key);
                    const size_t end = need(s_.find(q, p_ + 1), "attribute value");
                    node->attr[key] = s_.substr(p_ + 1, end - p_ - 1);
                    p_ = end + 1;
                }
                XmlNode* raw = node.get();
                stack.back()->children.push_back(std::move(node));
                if (!self_closing) stack.push_back(raw);
            }
        }
        if (stack.size() != 1) throw std::runtime_error("XML: unclosed <" + stack.back()->tag + ">");
        return root;
    }

  private:
    const std::string& s_;
    size_t p_ = 0;
    size_t need(size_t pos, const char* what) const {
        if (pos == std::string::npos) throw std::runtime_error(std::string("XML: unterminated ") + what);
        return pos;
    }
    void skip_space() {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) ++p_;
    }
    std::string name_at() {
        const size_t b = p_;
        while (p_ < s_.size() && !std::isspace((unsigned char)s_[p_]) && s_[p_] != '>' && s_[p_] != '/' && s_[p_] != '=') ++p_;
        return s_.substr(b, p_ - b);
    }
    static std::string trim(const std::string& x) {
        size_t b = 0, e = x.size();
        while (b < e && std::isspace((unsigned char)x[b])) ++b;
        while (e > b && std::isspace((unsigned char)x[e - 1])) --e;
        return x.substr(b, e - b);
    }
};

inline std::string read_file(const std::string& path) {
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

inline std::vector<double> numbers(const std::string& s, size_t n, const std::vector<double>& def) {
    if (s.empty()) return def;
    std::istringstream is(s);
    std::vector<double> v;
    double x;
    while (is >> x) v.push_back(x);
    if (v.size() != n) throw std::runtime_error("expected " + std::to_string(n) + " numbers in '" + s + "'");
    return v;
}

// Row-major rotation of roll-pitch-yaw angles through their unit quaternion.
inline void rpy_matrix(double roll, double pitch, double yaw, double R[9]) {
    const double phi = roll / 2.0, the = pitch / 2.0, psi = yaw / 2.0;
    double x = std::sin(phi) * std::cos(the) * std::cos(psi) - std::cos(phi) * std::sin(the) * std::sin(psi);
    double y = std::cos(phi) * std::sin(the) * std::cos(psi) + std::sin(phi) * std::cos(the) * std::sin(psi);
    double z = std::cos(phi) * std::cos(the) * std::sin(psi) - std::sin(phi) * std::sin(the) * std::cos(psi);
    double w = std::cos(phi) * std::cos(the) * std::cos(psi) + std::sin(phi) * std::sin(the) * std::sin(psi);
    const double n = std::sqrt(x * x + y * y + z * z + w * w);
    x /= n;
    y /= n;
    z /= n;
    w /= n;
    const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
    const double twx = tx * w, twy = ty * w, twz = tz * w;
    const double txx = tx * x, txy = ty * x, txz = tz * x, tyy = ty * y, tyz = tz * y, tzz = tz * z;
    const double M[9] = {1 - (tyy + tzz), txy - twz, txz + twy, txy + twz, 1 - (txx + tzz),
                         tyz - twx,       txz - twy, tyz + twx, 1 - (txx + tyy)};
    std::copy(M, M + 9, R);
}

struct UrdfJoint {
    std::string name, type, parent, child;
    double xyz[3] = {0, 0, 0}, rpy[3] = {0, 0, 0}, axis[3] = {1, 0, 0};
    double lower = 0, upper = 0;
};

// The "collision_spheres:" block: link name -> spheres (centre x, y, z, radius) in file order.
inline std::vector<std::pair<std::string, std::vector<std::array<float, 4>>>> read_sphere_yaml(const std::string& path) {
    std::istringstream in(read_file(path));
    std::vector<std::pair<std::string, std::vector<std::array<float, 4>>>> out;
    std::string line;
    int block = -1;  // indent of "collision_spheres:", -1 before it
    auto indent_of = [](const std::string& l) {
        size_t i = 0;
        while (i < l.size() && l[i] == ' ') ++i;
        return (int)i;
    };
    auto unquote = [](std::string k) {
        while (!k.empty() && std::isspace((unsigned char)k.back())) k.pop_back();
        if (k.size() >= 2 && (k[0] == '"' || k[0] == '\'') && k.back() == k[0]) k = k.substr(1, k.size() - 2);
        return k;
    };
    std::array<float, 4>* cur = nullptr;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        const int ind = indent_of(line);
        std::string body = line.substr(ind);
        while (!body.empty() && (body.back() == '\r' || body.back() == ' ')) body.pop_back();
        if (block < 0) {
            if (body == "collision_spheres:") block = ind;
            continue;
        }
        if (ind <= block) break;  // the block ended
        bool item = false;
        if (body[0] == '-') {  // a new sphere of the current link
            if (out.empty()) throw std::runtime_error(path + ": sphere before its link");
            out.back().second.push_back({0.0f, 0.0f, 0.0f, 0.0f});
            cur = &out.back().second.back();
            body = body.substr(1);
            body = body.substr(std::min(body.find_first_not_of(' '), body.size()));
            item = true;
        }
        const size_t colon = body.find(':');
        if (colon == std::string::npos) throw std::runtime_error(path + ": cannot read '" + line + "'");
        const std::string key = unquote(body.substr(0, colon));
        std::string val = body.substr(colon + 1);
        val = val.substr(std::min(val.find_first_not_of(' '), val.size()));
        if (!item && val.empty()) {  // a link key
            out.push_back({key, {}});
            cur = nullptr;
            continue;
        }
        if (!cur) throw std::runtime_error(path + ": value outside a sphere item: '" + line + "'");
        if (key == "center") {
            const size_t a = val.find('['), b = val.find(']');
            if (a == std::string::npos || b == std::string::npos) throw std::runtime_error(path + ": center is not a list");
            std::string inner = val.substr(a + 1, b - a - 1);
            std::replace(inner.begin(), inner.end(), ',', ' ');
            std::istringstream is(inner);
            for (int j = 0; j < 3; j++)
                if (!(is >> (*cur)[j])) throw std::runtime_error(path + ": center needs 3 numbers");
        } else if (key == "radius") {
            (*cur)[3] = std::strtof(val.c_str(), nullptr);
        }
    }
    if (block < 0) throw std::runtime_error(path + ": no collision_spheres block");
    return out;
}

}  // namespace robot_info_detail

class RobotInfo {
  public:
    RobotInfo(const std::string& urdf_path, const std::string& srdf_path, const std::string& group_name,
              const std::string& collision_spheres_file_path,
              const std::map<std::string, float>& default_joint_values_map = std::map<std::string, float>()) {
        namespace d = robot_info_detail;
        const std::string urdf_text = d::read_file(urdf_path);
        auto urdf = d::XmlReader(urdf_text).parse();
        const d::XmlNode* robot = urdf->child("robot");
        if (!robot) throw std::runtime_error(urdf_path + ": no <robot>");
        std::vector<std::string> links;
        std::map<std::string, d::UrdfJoint> joints;  // by name: urdfdom's order of a link's children
        for (const auto& c : robot->children) {
            if (c->tag == "link") links.push_back(c->get("name"));
            if (c->tag != "joint") continue;
            d::UrdfJoint j;
            j.name = c->get("name");
            j.type = c->get("type");
            const d::XmlNode* par = c->child("parent");
            const d::XmlNode* chi = c->child("child");
            if (!par || !chi) throw std::runtime_error("joint " + j.name + " lacks its parent or child");
            j.parent = par->get("link");
            j.child = chi->get("link");
            if (const d::XmlNode* o = c->child("origin")) {
                auto x = d::numbers(o->get("xyz"), 3, {0, 0, 0}), r = d::numbers(o->get("rpy"), 3, {0, 0, 0});
                std::copy(x.begin(), x.end(), j.xyz);
                std::copy(r.begin(), r.end(), j.rpy);
            }
            if (const d::XmlNode* a = c->child("axis")) {
                auto x = d::numbers(a->get("xyz"), 3, {1, 0, 0});
                std::copy(x.begin(), x.end(), j.axis);
            }
            if (const d::XmlNode* l = c->child("limit")) {
                j.lower = std::strtod(l->get("lower", "0").c_str(), nullptr);
                j.upper = std::strtod(l->get("upper", "0").c_str(), nullptr);
            }
            joints[j.name] = j;
        }
        std::map<std::string, const d::UrdfJoint*> parent_joint;
        std::map<std::string, std::vector<std::string>> children;
        for (const auto& kv : joints) {
            parent_joint[kv.second.child] = &kv.second;
            children[kv.second.parent].push_back(kv.second.child);
        }
        std::vector<std::string> roots;
        for (const auto& l : links)
            if (!parent_joint.count(l)) roots.push_back(l);
        if (roots.size() != 1) throw std::runtime_error(urdf_path + ": the links must form one tree");
        std::vector<std::string> stack{roots[0]};
        while (!stack.empty()) {  // depth-first pre-order
            const std::string l = stack.back();
            stack.pop_back();
            link_names.push_back(l);
            const auto& ch = children[l];
            for (auto it = ch.rbegin(); it != ch.rend(); ++it) stack.push_back(*it);
        }
        if (link_names.size() != links.size()) throw std::runtime_error(urdf_path + ": the links must form one tree");

        // the planning group's joints (SRDF <group>: <joint> entries, or <chain> base -> tip)
        const std::string srdf_text = d::read_file(srdf_path);
        auto srdf = d::XmlReader(srdf_text).parse();
        const d::XmlNode* srobot = srdf->child("robot");
        if (!srobot) throw std::runtime_error(srdf_path + ": no <robot>");
        std::vector<std::string> group;
        bool found = false;
        for (const auto& c : srobot->children) {
            if (c->tag != "group" || c->get("name") != group_name) continue;
            found = true;
            for (const auto& e : c->children) {
                if (e->tag == "joint") group.push_back(e->get("name"));
                if (e->tag == "chain") {
                    std::vector<std::string> chain;
                    for (std::string l = e->get("tip_link"); l != e->get("base_link");) {
                        auto it = parent_joint.find(l);
                        if (it == parent_joint.end()) throw std::runtime_error("chain tip is not below its base");
                        chain.push_back(it->second->name);
                        l = it->second->parent;
                    }
                    group.insert(group.end(), chain.rbegin(), chain.rend());
                }
            }
        }
        if (!found) throw std::runtime_error(srdf_path + ": no group " + group_name);

        const double kPi = 3.14159265358979323846;
        for (size_t i = 0; i < link_names.size(); i++) {
            auto it = parent_joint.find(link_names[i]);
            if (it == parent_joint.end()) {
                joint_name_to_parent_link.push_back("");
                joint_types.push_back(0);
                joint_poses.push_back(mpx::Isometry3d::Identity());
                joint_axes.push_back(mpx::Vector3d::Zero());
                link_maps.push_back(-1);
                default_joint_values.push_back(0.0f);
                lower_bounds.push_back(0.0f);
                upper_bounds.push_back(0.0f);
                active_joint_map.push_back(false);
                continue;
            }
            const d::UrdfJoint& j = *it->second;
            joint_name_to_parent_link.push_back(j.name);
            link_maps.push_back((int)(std::find(link_names.begin(), link_names.end(), j.parent) - link_names.begin()));
            mpx::Isometry3d T = mpx::Isometry3d::Identity();
            double R[9];
            d::rpy_matrix(j.rpy[0], j.rpy[1], j.rpy[2], R);
            for (int r = 0; r < 3; r++) {
                for (int c = 0; c < 3; c++) T.matrix()(r, c) = R[3 * r + c];
                T.matrix()(r, 3) = j.xyz[r];
            }
            joint_poses.push_back(T);
            double lo = 0, hi = 0;
            if (j.type == "revolute" || j.type == "continuous" || j.type == "prismatic") {
                joint_types.push_back(j.type == "prismatic" ? 2 : 1);
                const double n = std::sqrt(j.axis[0] * j.axis[0] + j.axis[1] * j.axis[1] + j.axis[2] * j.axis[2]);
                joint_axes.push_back(mpx::Vector3d(j.axis[0] / n, j.axis[1] / n, j.axis[2] / n));
                lo = j.type == "continuous" ? -kPi : j.lower;
                hi = j.type == "continuous" ? kPi : j.upper;
                auto dv = default_joint_values_map.find(j.name);
                default_joint_values.push_back(dv != default_joint_values_map.end()
                                                   ? dv->second
                                                   : (float)(lo <= 0.0 && 0.0 <= hi ? 0.0 : (lo + hi) / 2.0));
            } else if (j.type == "fixed") {
                joint_types.push_back(5);
                joint_axes.push_back(mpx::Vector3d::Zero());
                default_joint_values.push_back(0.0f);
            } else {
                throw std::runtime_error("joint " + j.name + ": unsupported type " + j.type);
            }
            const bool active = std::find(group.begin(), group.end(), j.name) != group.end() && joint_types.back() != 5;
            active_joint_map.push_back(active);
            dimension += active ? 1 : 0;
            lower_bounds.push_back(active ? (float)lo : 0.0f);
            upper_bounds.push_back(active ? (float)hi : 0.0f);
        }

        // collision spheres in file order, per link of the model
        for (const auto& kv : d::read_sphere_yaml(collision_spheres_file_path)) {
            const auto it = std::find(link_names.begin(), link_names.end(), kv.first);
            if (it == link_names.end()) continue;
            for (const auto& s : kv.second) {
                collision_spheres_map.push_back((int)(it - link_names.begin()));
                collision_spheres_pos.push_back({s[0], s[1], s[2]});
                collision_spheres_radius.push_back(s[3]);
            }
        }
        // self-collision enable map: SRDF disabled pairs cleared
        const size_t L = link_names.size();
        self_collision_enabled_map.assign(L, std::vector<bool>(L, true));
        for (const auto& c : srobot->children) {
            if (c->tag != "disable_collisions") continue;
            const auto a = std::find(link_names.begin(), link_names.end(), c->get("link1"));
            const auto b = std::find(link_names.begin(), link_names.end(), c->get("link2"));
            if (a == link_names.end() || b == link_names.end()) continue;
            self_collision_enabled_map[a - link_names.begin()][b - link_names.begin()] = false;
            self_collision_enabled_map[b - link_names.begin()][a - link_names.begin()] = false;
        }
    }

    std::vector<int> getJointTypes() const { return joint_types; }
    std::vector<mpx::Isometry3d> getJointPoses() const { return joint_poses; }
    std::vector<mpx::Vector3d> getJointAxes() const { return joint_axes; }
    std::vector<int> getLinkMaps() const { return link_maps; }
    std::vector<std::string> getLinkNames() const { return link_names; }
    std::vector<int> getCollisionSpheresMap() const { return collision_spheres_map; }
    std::vector<std::vector<float>> getCollisionSpheresPos() const { return collision_spheres_pos; }
    std::vector<mpx::Vector3d> getCollisionSpheresPosEigen() const {
        std::vector<mpx::Vector3d> v;
        for (const auto& p : collision_spheres_pos) v.push_back(mpx::Vector3d(p[0], p[1], p[2]));
        return v;
    }
    std::vector<float> getCollisionSpheresRadius() const { return collision_spheres_radius; }
    std::vector<std::vector<bool>> getSelfCollisionEnabledMap() const { return self_collision_enabled_map; }
    std::vector<bool> getActiveJointMap() const { return active_joint_map; }
    int getDimension() const { return dimension; }
    std::vector<float> getUpperBounds() const { return upper_bounds; }
    std::vector<float> getLowerBounds() const { return lower_bounds; }
    std::vector<float> getDefaultJointValues() const { return default_joint_values; }
    std::vector<std::string> getJointNames() const { return joint_name_to_parent_link; }

  private:
    std::vector<int> joint_types, link_maps, collision_spheres_map;
    std::vector<mpx::Isometry3d> joint_poses;
    std::vector<mpx::Vector3d> joint_axes;
    std::vector<std::string> link_names, joint_name_to_parent_link;
    std::vector<std::vector<float>> collision_spheres_pos;
    std::vector<float> collision_spheres_radius, upper_bounds, lower_bounds, default_joint_values;
    std::vector<std::vector<bool>> self_collision_enabled_map;
    std::vector<bool> active_joint_map;
    int dimension = 0;
};
